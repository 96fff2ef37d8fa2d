// Block substitution (trsm(Hierarchical, Dense), src/operations/BLAS/trsm.cpp:40-90)
// as one persistent launch per sweep.
//
// A sweep solves x_p = op(Tinv_p) (b_p - sum_{q before p} M_pq x_q) over the
// blocks p in sweep order, every M_pq low rank (U S V, strided: transposed
// sweeps pass the transposed views).  The launch is a dataflow over tasks
// claimed from one ticket counter in a host-built order; a task waits only for
// tasks earlier in that order, so the launch completes whatever number of its
// workgroups is resident:
//   D(p, g)  rows [16 g, 16 g + 16) of x_p.  Before any wait it loads its rows
//            of op(Tinv_p) and forms W = op(Tinv_p)[rows] U S V for the update
//            from the block just before p (16 x n), then
//            pre = op(Tinv_p)[rows] (b_p - older updates); once x_{p-1} is
//            complete, x[rows] = pre - W x_{p-1}.
//   Y(q, c)  y_pq = S_pq (V_pq x_q) for up to four of the blocks p three or
//            more positions after q, one wave per update (the updates into
//            the next two blocks are applied by those blocks' own tasks).
//   A(p, h)  rows [64 h, 64 h + 64) of bt_p = b_p - sum of U_pq S_pq V_pq x_q
//            over the updates of p from blocks two or more positions before p:
//            U_pq y_pq for q three or more positions back (the bases staged
//            before the wait, then one round of loads brings every y_pq), and
//            the update from two positions back formed from x_q itself once
//            that block is complete.
// The critical chain is one D step per block (wait, load x_{p-1}, 16 rows of
// W x_{p-1}, store); every other load that does not depend on the chain is
// issued before the task waits.
// Waits poll relaxed agent-scope atomics and take one acquire fence; a wait
// past the spin limit marks the launch failed and every task then exits.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

namespace frk {

namespace {

constexpr int kDfGroup = 16;   // threads per row in a D task (256 threads = 16 rows x 16)
constexpr int kDfAcols = 128;  // A task: basis columns staged per round
constexpr int kDfAtiles = 256; // A task: updates described per round
constexpr int kDfSmax = 16;    // Y task: S staged in LDS up to this rank
constexpr int kDfCuLd = 65;    // A task: padded row length of the staged bases

// LDS of one workgroup, shared by the three task types.
template <typename T, int NR>
struct DfLds {
  static constexpr size_t xs = sizeof(T) * 512 * NR;
  static constexpr size_t d = xs + sizeof(T) * kDfRows * 64;
  static constexpr size_t y = xs + sizeof(T) * kDfYtiles * 64 * NR + sizeof(T) * kDfYtiles * kDfSmax * kDfSmax;
  // A: the older updates' staging; then x of the block two back (over the
  // staging) with V x, S V x and S beyond it
  static constexpr size_t a0 = sizeof(T) * kDfAcols * (kDfCuLd + NR) + 8 * 3 * kDfAtiles + 4 * (2 * kDfAtiles + 4);
  static constexpr size_t a = a0 + sizeof(T) * (2 * 64 * NR + kDfSmax * kDfSmax);
  static_assert(xs <= sizeof(T) * kDfAcols * kDfCuLd, "x of the A task fits over the staged bases");
  static constexpr size_t bytes = d > y ? (d > a ? d : a) : (y > a ? y : a);
};

__device__ __forceinline__ float sum16(float v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 8);
  return v;
}
__device__ __forceinline__ double sum16(double v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 8);
  return v;
}
template <typename T>
__device__ __forceinline__ T sum64(T v) {
  v = sum16(v);
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

// Poll until the n flags f[0 .. n) are all set (wave 0, a lane per flag, n <= 64),
// then acquire.  A task's completion is its own flag, never a shared counter:
// the tasks of a block finish at nearly the same time, and read-modify-writes
// of one word from many workgroups serialise.
__device__ __forceinline__ bool df_wait_all(const int32_t* f, int n, int32_t* err, int64_t limit, int* bail_sh) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    bool ok = lane >= n;
    int64_t spins = 0;
    int bail = 0;
    for (;;) {
      if (!ok) ok = __hip_atomic_load(&f[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
      if (__all(ok)) break;
      if (++spins > limit) {
        if (lane == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bail = 1;
        break;
      }
      if ((spins & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        bail = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (lane == 0) *bail_sh = bail;
  }
  __syncthreads();
  return *bail_sh == 0;
}

// Publish: every thread's stores precede the release (barrier + cumulative release).
__device__ __forceinline__ void df_set(int32_t* flag) {
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// x (n x nrhs, row-major) into xs[512][NR], zero past n and nrhs: every load of
// a thread issued before the first store (one memory round trip).
template <typename T, int NR>
__device__ __forceinline__ void df_stage(T (*xs)[NR], const T* __restrict__ src, int n, int nrhs) {
  constexpr int E = 512 * NR / 256;
  T v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int e = threadIdx.x + 256 * i;
    const int k = e / NR, c = e % NR;
    const bool in = k < n && c < nrhs;
    v[i] = src[in ? (int64_t)k * nrhs + c : 0];
  }
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int e = threadIdx.x + 256 * i;
    const int k = e / NR, c = e % NR;
    xs[k][c] = (k < n && c < nrhs) ? v[i] : T(0);
  }
}

// ---------------------------------------------------------------- D(p, g)
template <typename T, int NR>
__device__ bool df_task_d(const DfSweep& sw, int4 tk, unsigned char* lds, int* bail_sh) {
  T (*xs)[NR] = reinterpret_cast<T (*)[NR]>(lds);
  T (*zs)[64] = reinterpret_cast<T (*)[64]>(lds + DfLds<T, NR>::xs);
  const DfBlock* blks = static_cast<const DfBlock*>(sw.blocks);
  const DfTile* tiles = static_cast<const DfTile*>(sw.tiles);
  int32_t* err = sw.flags + 1;
  int32_t* xflag = sw.flags + 2;
  int32_t* aflag = xflag + kDfXflags * sw.nb;
  const int nrhs = sw.nrhs;
  const int tid = threadIdx.x;
  const DfBlock B = blks[tk.y];
  const int n = B.n, r0 = kDfRows * tk.z;
  const int i = tid / kDfGroup, sub = tid % kDfGroup;
  const int row = r0 + i;
  const T* Ti = static_cast<const T*>(B.tinv);
  // this thread's part of row `row` of op(Tinv): columns sub + 16 j
  T ta[512 / kDfGroup];
#pragma unroll
  for (int j = 0; j < 512 / kDfGroup; ++j) {
    const int k = sub + kDfGroup * j;
    const bool in = row < n && k < n;
    const int64_t off = B.trans ? (int64_t)k * B.ld + row : (int64_t)row * B.ld + k;
    ta[j] = Ti[in ? off : 0];
  }
#pragma unroll
  for (int j = 0; j < 512 / kDfGroup; ++j) {
    const int k = sub + kDfGroup * j;
    ta[j] = (row < n && k < n) ? ta[j] : T(0);
  }
  // W = op(Tinv)[rows] U S V of the update from the previous block, this
  // thread's columns sub + 16 j of row `row` (the update is a factor: no wait)
  T wa[512 / kDfGroup];
#pragma unroll
  for (int j = 0; j < 512 / kDfGroup; ++j) wa[j] = T(0);
  int nprev = 0;
  if (B.last >= 0) {
    const DfTile L = tiles[B.last];
    const T* U = static_cast<const T*>(L.U);
    const T* S = static_cast<const T*>(L.S);
    const T* V = static_cast<const T*>(L.V);
    nprev = blks[tk.y - 1].n;
    const int rl = L.r;
    // zs[i][c] = (op(Tinv)[rows] U)[i][c]
    for (int c = 0; c < rl; ++c) {
      T u[512 / kDfGroup];
#pragma unroll
      for (int j = 0; j < 512 / kDfGroup; ++j) {
        const int k = sub + kDfGroup * j;
        u[j] = U[k < n ? k * L.u_rs + c * L.u_cs : 0];
      }
      T acc = T(0);
#pragma unroll
      for (int j = 0; j < 512 / kDfGroup; ++j) acc += ta[j] * u[j];  // ta is zero past n
      acc = sum16(acc);
      if (sub == 0) zs[i][c] = acc;
    }
    __syncthreads();
    // zs <- zs S (row by row: thread sub of row i forms entries sub, sub + 16, ...)
    T zrow[64 / kDfGroup];
#pragma unroll
    for (int h = 0; h < 64 / kDfGroup; ++h) {
      const int c = sub + kDfGroup * h;
      T acc = T(0);
      if (c < rl)
        for (int s2 = 0; s2 < rl; ++s2) acc += zs[i][s2] * S[s2 * L.s_rs + c * L.s_cs];
      zrow[h] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 64 / kDfGroup; ++h) {
      const int c = sub + kDfGroup * h;
      if (c < rl) zs[i][c] = zrow[h];
    }
    __syncthreads();
    // W = (zs) V: columns k = sub + 16 j of the previous block
    for (int c = 0; c < rl; ++c) {
      const T z = zs[i][c];
      T v[512 / kDfGroup];
#pragma unroll
      for (int j = 0; j < 512 / kDfGroup; ++j) {
        const int k = sub + kDfGroup * j;
        v[j] = V[k < nprev ? c * L.v_rs + k * L.v_cs : 0];
      }
#pragma unroll
      for (int j = 0; j < 512 / kDfGroup; ++j) {
        const int k = sub + kDfGroup * j;
        wa[j] += z * (k < nprev ? v[j] : T(0));
      }
    }
  }
  // pre = op(Tinv)[rows] (b - older updates)
  const T* src = static_cast<const T*>(B.b);
  if (B.has_bt) {
    if (!df_wait_all(aflag + kDfAflags * tk.y, B.na, err, sw.spin_limit, bail_sh)) return false;
    src = static_cast<const T*>(B.bt);
  }
  df_stage<T, NR>(xs, src, n, nrhs);
  __syncthreads();
  T pre[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) pre[c] = T(0);
#pragma unroll
  for (int j = 0; j < 512 / kDfGroup; ++j) {
    const int k = sub + kDfGroup * j;
#pragma unroll
    for (int c = 0; c < NR; ++c) pre[c] += ta[j] * xs[k][c];  // xs and ta are zero past n
  }
#pragma unroll
  for (int c = 0; c < NR; ++c) pre[c] = sum16(pre[c]);
  if (B.last >= 0) {
    // the previous block complete: x[rows] = pre - W x_{p-1}
    if (!df_wait_all(xflag + kDfXflags * (tk.y - 1), blks[tk.y - 1].nd, err, sw.spin_limit, bail_sh))
      return false;
    df_stage<T, NR>(xs, static_cast<const T*>(blks[tk.y - 1].x), nprev, nrhs);
    __syncthreads();
    T d[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) d[c] = T(0);
#pragma unroll
    for (int j = 0; j < 512 / kDfGroup; ++j) {
      const int k = sub + kDfGroup * j;
#pragma unroll
      for (int c = 0; c < NR; ++c) d[c] += wa[j] * xs[k][c];  // wa is zero past nprev
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) pre[c] -= sum16(d[c]);
  }
  T* x = static_cast<T*>(B.x);
  if (row < n && sub < nrhs) {
    T v = pre[0];
#pragma unroll
    for (int c = 1; c < NR; ++c)
      if (c == sub) v = pre[c];
    x[(int64_t)row * nrhs + sub] = v;
  }
  df_set(xflag + kDfXflags * tk.y + tk.z);
  return true;
}

// ---------------------------------------------------------------- Y(q, c)
// Wave w forms y for update c0 + w of block q's outgoing list (c0 + w < c1).
template <typename T, int NR>
__device__ bool df_task_y(const DfSweep& sw, int4 tk, unsigned char* lds, int* bail_sh, int* yti_sh) {
  T (*xs)[NR] = reinterpret_cast<T (*)[NR]>(lds);
  T* tsb = reinterpret_cast<T*>(lds + DfLds<T, NR>::xs);
  T* ssb = tsb + kDfYtiles * 64 * NR;
  const DfBlock* blks = static_cast<const DfBlock*>(sw.blocks);
  const DfTile* tiles = static_cast<const DfTile*>(sw.tiles);
  int32_t* err = sw.flags + 1;
  int32_t* xflag = sw.flags + 2;
  int32_t* ydone = xflag + (kDfXflags + kDfAflags) * sw.nb;
  T* ybuf = static_cast<T*>(sw.y);
  const int nrhs = sw.nrhs;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const DfBlock B = blks[tk.y];
  const int n = B.n;
  const int e = tk.z + w;
  const bool has = e < tk.w;
  int ti = 0;
  DfTile L{};
  if (has) {
    ti = sw.by_src[B.src0 + e];
    L = tiles[ti];
  }
  const int r = has ? L.r : 0;
  if (lane == 0) yti_sh[w] = has ? ti : -1;
  T (*ts)[NR] = reinterpret_cast<T (*)[NR]>(tsb + w * 64 * NR);
  T (*ss)[kDfSmax] = reinterpret_cast<T (*)[kDfSmax]>(ssb + w * kDfSmax * kDfSmax);
  const T* V = static_cast<const T*>(L.V);
  const T* S = static_cast<const T*>(L.S);
  constexpr int KM = 512 / 64;
  // before the wait: rows [0, 4) of V at this lane's columns, and S (r <= 16)
  T vp[4][KM];
  if (r > 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int m = 0; m < KM; ++m) {
        const int k = lane + 64 * m;
        vp[i][m] = V[(i < r && k < n) ? (int64_t)i * L.v_rs + (int64_t)k * L.v_cs : 0];
      }
    if (r <= kDfSmax) {
      T sv[kDfSmax * kDfSmax / 64];
#pragma unroll
      for (int j = 0; j < kDfSmax * kDfSmax / 64; ++j) {
        const int o = lane + 64 * j, a = o / kDfSmax, b = o % kDfSmax;
        sv[j] = S[(a < r && b < r) ? (int64_t)a * L.s_rs + (int64_t)b * L.s_cs : 0];
      }
#pragma unroll
      for (int j = 0; j < kDfSmax * kDfSmax / 64; ++j) {
        const int o = lane + 64 * j, a = o / kDfSmax, b = o % kDfSmax;
        ss[a][b] = (a < r && b < r) ? sv[j] : T(0);
      }
    }
  }
  if (!df_wait_all(xflag + kDfXflags * tk.y, B.nd, err, sw.spin_limit, bail_sh)) return false;
  df_stage<T, NR>(xs, static_cast<const T*>(B.x), n, nrhs);
  __syncthreads();
  // ts = V x (r x NR), four rows of V at a time, lanes over columns
  for (int i0 = 0; i0 < r; i0 += 4) {
    if (i0 > 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int m = 0; m < KM; ++m) {
          const int k = lane + 64 * m;
          vp[i][m] = V[(i0 + i < r && k < n) ? (int64_t)(i0 + i) * L.v_rs + (int64_t)k * L.v_cs : 0];
        }
    }
    T acc[4][NR];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[i][c] = T(0);
#pragma unroll
    for (int m = 0; m < KM; ++m) {
      const int k = lane + 64 * m;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const T v = (i0 + i < r && k < n) ? vp[i][m] : T(0);
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[i][c] += v * xs[k][c];
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const T s = sum64(acc[i][c]);
        if (lane == 0 && i0 + i < r) ts[i0 + i][c] = s;
      }
  }
  __syncthreads();
  // y = S ts (r x NR)
  if (r > 0) {
    T* y = ybuf + L.yoff;
    for (int o = lane; o < r * NR; o += 64) {
      const int rho = o / NR, c = o % NR;
      T acc = T(0);
      if (r <= kDfSmax) {
        for (int s2 = 0; s2 < r; ++s2) acc += ss[rho][s2] * ts[s2][c];
      } else {
        for (int s2 = 0; s2 < r; ++s2) acc += S[(int64_t)rho * L.s_rs + (int64_t)s2 * L.s_cs] * ts[s2][c];
      }
      y[o] = acc;
    }
  }
  __syncthreads();
  if (tid == 0)
    for (int j = 0; j < kDfYtiles; ++j)
      if (yti_sh[j] >= 0) __hip_atomic_store(&ydone[yti_sh[j]], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// ---------------------------------------------------------------- A(p, h)
template <typename T, int NR>
__device__ bool df_task_a(const DfSweep& sw, int4 tk, unsigned char* lds, int* bail_sh, int* wsum) {
  T (*cu)[kDfCuLd] = reinterpret_cast<T (*)[kDfCuLd]>(lds);
  T (*ys)[NR] = reinterpret_cast<T (*)[NR]>(lds + sizeof(T) * kDfAcols * kDfCuLd);
  unsigned char* q8 = lds + sizeof(T) * kDfAcols * (kDfCuLd + NR);
  const T** tU = reinterpret_cast<const T**>(q8);
  int64_t* tus = reinterpret_cast<int64_t*>(q8 + 8 * kDfAtiles);
  int64_t* tuc = reinterpret_cast<int64_t*>(q8 + 16 * kDfAtiles);
  int32_t* tyo = reinterpret_cast<int32_t*>(q8 + 24 * kDfAtiles);
  int32_t* pref = tyo + kDfAtiles;  // kDfAtiles + 1 entries
  // the update from two positions back (after the older ones are summed)
  T (*xs)[NR] = reinterpret_cast<T (*)[NR]>(lds);
  T (*ts)[NR] = reinterpret_cast<T (*)[NR]>(lds + DfLds<T, NR>::a0);
  T (*yn)[NR] = ts + 64;
  T (*ss)[kDfSmax] = reinterpret_cast<T (*)[kDfSmax]>(yn + 64);
  const DfBlock* blks = static_cast<const DfBlock*>(sw.blocks);
  const DfTile* tiles = static_cast<const DfTile*>(sw.tiles);
  int32_t* err = sw.flags + 1;
  int32_t* xflag = sw.flags + 2;
  int32_t* aflag = xflag + kDfXflags * sw.nb;
  int32_t* ydone = aflag + kDfAflags * sw.nb;
  const T* ybuf = static_cast<const T*>(sw.y);
  const int nrhs = sw.nrhs;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const DfBlock B = blks[tk.y];
  const int n = B.n, r0 = kDfArows * tk.z;
  constexpr int NC = NR / 4;  // right-hand sides per thread in the product
  T acc[NC];
#pragma unroll
  for (int m = 0; m < NC; ++m) acc[m] = T(0);
  bool waited = false;
  for (int tb = 0; tb < B.ntile; tb += kDfAtiles) {
    const int nbt = min(kDfAtiles, B.ntile - tb);
    // this round's update descriptors and the prefix sum of their ranks
    int rr = 0;
    if (tid < nbt) {
      const int ti = B.tile0 + tb + tid;
      if (ti != B.last && ti != B.near) {
        const DfTile L = tiles[ti];
        rr = L.r;
        tU[tid] = static_cast<const T*>(L.U);
        tus[tid] = L.u_rs;
        tuc[tid] = L.u_cs;
        tyo[tid] = L.yoff;
      }
    }
    int v = rr;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int u = __shfl_up(v, d);
      if (lane >= d) v += u;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    int base = 0;
    for (int j = 0; j < w; ++j) base += wsum[j];
    pref[tid + 1] = base + v;
    if (tid == 0) pref[0] = 0;
    __syncthreads();
    const int J = pref[kDfAtiles];
    for (int jb = 0; jb < J; jb += kDfAcols) {
      // thread (column jj, half) stages rows [32 half, 32 half + 32) of basis column jb + jj
      const int jj = tid % kDfAcols, half = tid / kDfAcols;
      const int j = jb + jj;
      const bool colok = j < J;
      int e = 0, rho = 0;
      if (colok) {
        int lo = 0, hi = nbt;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (pref[mid + 1] <= j)
            lo = mid + 1;
          else
            hi = mid;
        }
        e = lo;
        rho = j - pref[e];
      }
      if (colok) {
        const T* U = tU[e];
        const int64_t urs = tus[e], ucs = tuc[e];
        T uv[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) {
          const int row = r0 + 32 * half + i;
          uv[i] = U[(int64_t)(row < n ? row : 0) * urs + (int64_t)rho * ucs];
        }
#pragma unroll
        for (int i = 0; i < 32; ++i) cu[jj][32 * half + i] = (r0 + 32 * half + i < n) ? uv[i] : T(0);
      } else {
#pragma unroll
        for (int i = 0; i < 32; ++i) cu[jj][32 * half + i] = T(0);
      }
      if (!waited) {
        // every older update of the block at once: thread e' polls the flag of update e'
        int bail = 0;
        for (int e2 = tid; e2 < B.ntile && !bail; e2 += 256) {
          const int ti = B.tile0 + e2;
          if (ti == B.last || ti == B.near) continue;
          int64_t spins = 0;
          while (__hip_atomic_load(&ydone[ti], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < 1) {
            if (++spins > sw.spin_limit) {
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              bail = 1;
              break;
            }
            if ((spins & 63) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
              bail = 1;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (__syncthreads_or(bail)) return false;
        waited = true;
      }
      // y rows of the staged columns: thread (jj, half) loads right-hand sides half + 2 m
      {
        T yv[NR / 2];
        const T* yb = ybuf + (colok ? (int64_t)tyo[e] + (int64_t)rho * NR : 0);
#pragma unroll
        for (int m = 0; m < NR / 2; ++m) yv[m] = yb[half + 2 * m];
#pragma unroll
        for (int m = 0; m < NR / 2; ++m) ys[jj][half + 2 * m] = colok ? yv[m] : T(0);
      }
      __syncthreads();
      const int ncol = min(kDfAcols, J - jb);
      for (int c2 = 0; c2 < ncol; ++c2) {
        const T u = cu[c2][lane];
#pragma unroll
        for (int m = 0; m < NC; ++m) acc[m] += u * ys[c2][w + 4 * m];
      }
      __syncthreads();
    }
  }
  if (B.near >= 0) {
    // the update from two positions back: acc += U (S (V x_{p-2})) on rows r0 + lane
    const DfTile L = tiles[B.near];
    const DfBlock Q = blks[tk.y - 2];
    const int r = L.r, nq = Q.n;
    const T* U = static_cast<const T*>(L.U);
    const T* S = static_cast<const T*>(L.S);
    const T* V = static_cast<const T*>(L.V);
    const int row = r0 + lane;
    constexpr int KM = 512 / 64;
    // before the wait: wave w's first row of V (row w), S (r <= 16) and this row of U (r <= 8)
    T vp[KM], up[8];
#pragma unroll
    for (int m = 0; m < KM; ++m) {
      const int k = lane + 64 * m;
      vp[m] = V[(w < r && k < nq) ? (int64_t)w * L.v_rs + (int64_t)k * L.v_cs : 0];
    }
#pragma unroll
    for (int h = 0; h < 8; ++h) up[h] = U[(h < r && row < n) ? (int64_t)row * L.u_rs + (int64_t)h * L.u_cs : 0];
    if (r <= kDfSmax) {
      T sv[kDfSmax * kDfSmax / 256];
#pragma unroll
      for (int j = 0; j < kDfSmax * kDfSmax / 256; ++j) {
        const int o = tid + 256 * j, a = o / kDfSmax, b = o % kDfSmax;
        sv[j] = S[(a < r && b < r) ? (int64_t)a * L.s_rs + (int64_t)b * L.s_cs : 0];
      }
#pragma unroll
      for (int j = 0; j < kDfSmax * kDfSmax / 256; ++j) {
        const int o = tid + 256 * j, a = o / kDfSmax, b = o % kDfSmax;
        ss[a][b] = (a < r && b < r) ? sv[j] : T(0);
      }
    }
    if (!df_wait_all(xflag + kDfXflags * (tk.y - 2), Q.nd, err, sw.spin_limit, bail_sh)) return false;
    df_stage<T, NR>(xs, static_cast<const T*>(Q.x), nq, nrhs);
    __syncthreads();
    // ts = V x: wave w forms rows w, w + 4, ... (lanes over columns)
    for (int rho = w; rho < r; rho += 4) {
      if (rho != w) {
#pragma unroll
        for (int m = 0; m < KM; ++m) {
          const int k = lane + 64 * m;
          vp[m] = V[k < nq ? (int64_t)rho * L.v_rs + (int64_t)k * L.v_cs : 0];
        }
      }
      T t[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) t[c] = T(0);
#pragma unroll
      for (int m = 0; m < KM; ++m) {
        const int k = lane + 64 * m;
        const T vv = k < nq ? vp[m] : T(0);
#pragma unroll
        for (int c = 0; c < NR; ++c) t[c] += vv * xs[k][c];
      }
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        const T sm = sum64(t[c]);
        if (lane == 0) ts[rho][c] = sm;
      }
    }
    __syncthreads();
    // yn = S ts (r x NR)
    for (int o = tid; o < r * NR; o += 256) {
      const int rho = o / NR, c = o % NR;
      T a2 = T(0);
      if (r <= kDfSmax) {
        for (int s2 = 0; s2 < r; ++s2) a2 += ss[rho][s2] * ts[s2][c];
      } else {
        for (int s2 = 0; s2 < r; ++s2) a2 += S[(int64_t)rho * L.s_rs + (int64_t)s2 * L.s_cs] * ts[s2][c];
      }
      yn[rho][c] = a2;
    }
    __syncthreads();
    // acc += U[row] yn
#pragma unroll
    for (int h = 0; h < 8; ++h)
      if (h < r) {
        const T uu = row < n ? up[h] : T(0);
#pragma unroll
        for (int m = 0; m < NC; ++m) acc[m] += uu * yn[h][w + 4 * m];
      }
    for (int h = 8; h < r; ++h) {
      const T uu = row < n ? U[(int64_t)row * L.u_rs + (int64_t)h * L.u_cs] : T(0);
#pragma unroll
      for (int m = 0; m < NC; ++m) acc[m] += uu * yn[h][w + 4 * m];
    }
  }
  // bt = b - acc: row r0 + lane, right-hand sides w + 4 m
  const int row = r0 + lane;
  const T* b = static_cast<const T*>(B.b);
  T* bt = static_cast<T*>(B.bt);
  if (row < n)
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const int c = w + 4 * m;
      if (c < nrhs) bt[(int64_t)row * nrhs + c] = b[(int64_t)row * nrhs + c] - acc[m];
    }
  df_set(aflag + kDfAflags * tk.y + tk.z);
  return true;
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void solve_df_kernel(DfSweep sw) {
  __shared__ __align__(16) unsigned char lds[DfLds<T, NR>::bytes];
  __shared__ int task_sh, bail_sh;
  __shared__ int misc_sh[kDfYtiles > 4 ? kDfYtiles : 4];
  const int4* tasks = static_cast<const int4*>(sw.tasks);
  int32_t* ticket = sw.flags;
  for (;;) {
    if (threadIdx.x == 0) {
      task_sh = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      bail_sh = 0;
    }
    __syncthreads();
    const int t = task_sh;
    if (t >= sw.ntasks) return;
    const int4 tk = tasks[t];
    __syncthreads();
    bool ok;
    if (tk.x == 0)
      ok = df_task_d<T, NR>(sw, tk, lds, &bail_sh);
    else if (tk.x == 1)
      ok = df_task_y<T, NR>(sw, tk, lds, &bail_sh, misc_sh);
    else
      ok = df_task_a<T, NR>(sw, tk, lds, &bail_sh, misc_sh);
    if (!ok) return;
    __syncthreads();
  }
}

__global__ void df_clear_kernel(int32_t* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

}  // namespace

template <typename T>
void launch_solve_df(const DfSweep& sw, int nflags, int nrhs, int grid, hipStream_t s) {
  if (nrhs <= 0 || nrhs > 16) throw std::runtime_error("dataflow solve: nrhs must be in [1, 16]");
  hipLaunchKernelGGL(df_clear_kernel, dim3((nflags + 255) / 256), dim3(256), 0, s, sw.flags, nflags);
  FRK_HIP_CHECK(hipGetLastError());
  if (nrhs > 8)
    hipLaunchKernelGGL((solve_df_kernel<T, 16>), dim3(grid), dim3(256), 0, s, sw);
  else
    hipLaunchKernelGGL((solve_df_kernel<T, 8>), dim3(grid), dim3(256), 0, s, sw);
  FRK_HIP_CHECK(hipGetLastError());
}
template void launch_solve_df<double>(const DfSweep&, int, int, int, hipStream_t);
template void launch_solve_df<float>(const DfSweep&, int, int, int, hipStream_t);

}  // namespace frk

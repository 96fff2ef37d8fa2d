// gfx950 kernels for the BLR factorization hot path.
//
// Design notes (CDNA4, wave64):
//  * Small dense factorizations (Householder QR, reflector application) keep
//    the whole m x K operand in VGPRs: wave w owns columns c = q*W + w and
//    lane l holds rows l + 64*t.  Column reductions are DPP / swizzle /
//    permlane butterflies (bitwise identical in every lane); the only workgroup
//    barrier per Householder step publishes the reflector through LDS.
//  * Jacobi SVD of the K x K core: for K <= 16 one wave keeps the core in
//    VGPRs (a lane group per column pair, columns moved between pairs by DPP
//    wave shifts); larger cores run in LDS, QR-preconditioned, with a
//    round-robin pairing so K/2 rotations proceed in parallel.
//  * Column-pivoted QR streams the tile through L2 with one thread per column
//    (row-major tiles => every row sweep is coalesced).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "kernels.hpp"
#include "diag64.hpp"

namespace frk {

template <typename T>
struct Eps;
template <>
struct Eps<double> {
  static constexpr double v = 2.220446049250313080847e-16;
  static constexpr double half = 1.1102230246251565404e-16;  // dlamch('e')
};
template <>
struct Eps<float> {
  static constexpr float v = 1.1920928955078125e-07f;
  static constexpr float half = 5.9604644775390625e-08f;
};

// Lane exchanges of the 64-wide wavefront without the LDS crossbar: DPP for
// partners inside a row of 16 lanes, ds_swizzle for xor 16 inside a half
// wave, v_permlane32_swap (gfx950) across the two halves.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  return __hiloint2double(dpp_i<CTRL>(__double2hiint(v)), dpp_i<CTRL>(__double2loint(v)));
}
__device__ __forceinline__ int swz16_i(int v) { return __builtin_amdgcn_ds_swizzle(v, 0x401F); }
__device__ __forceinline__ float xor16(float v) { return __int_as_float(swz16_i(__float_as_int(v))); }
__device__ __forceinline__ double xor16(double v) {
  return __hiloint2double(swz16_i(__double2hiint(v)), swz16_i(__double2loint(v)));
}
// v[l] + v[l ^ 32] with the operands in the same order in both halves
__device__ __forceinline__ float half_sum(float v) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(v), __float_as_int(v), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
__device__ __forceinline__ double half_sum(double v) {
  const auto h = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
  const auto l = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
  return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}

// Sum over aligned groups of g lanes (g a power of two <= 64, uniform over
// the wave).  Every step adds a value that is uniform over one sub-group to
// the value of its partner sub-group, so all lanes of a group end with the
// same bits.
template <typename T>
__device__ __forceinline__ T group_sum(T v, int g) {
  if (g >= 2) v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  if (g >= 4) v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  if (g >= 8) v += dpp_mov<0x141>(v);  // row_half_mirror
  if (g >= 16) v += dpp_mov<0x140>(v); // row_mirror
  if (g >= 32) v += xor16(v);
  if (g >= 64) v = half_sum(v);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  v += dpp_mov<0xB1>(v);
  v += dpp_mov<0x4E>(v);
  v += dpp_mov<0x141>(v);
  v += dpp_mov<0x140>(v);
  v += xor16(v);
  return half_sum(v);
}

// v of lane l (uniform l) in every lane
__device__ __forceinline__ float wave_read(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double wave_read(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}
// 1 / x: the hardware reciprocal in FP32 (about 1 ulp), correctly rounded in FP64
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ===========================================================================
// Strided batched GEMM.  One TM x TN output tile per workgroup (256 threads,
// RM x RN outputs each), K staged through LDS in chunks of TK with loads
// mapped to whichever operand stride is unit (coalesced).  Variants:
//   64 x 64 x 16 (square-ish), 64 x 16 x 32 (tall-skinny: m large, n <= 16),
//   16 x 64 x 32 (short-wide), 16 x 16 x 64 (tiny output, long K).
// ===========================================================================
template <typename T, int TM, int TN, int TK, int RM, int RN>
__global__ __launch_bounds__(256) void gemm_kernel(const GemmItem* __restrict__ items) {
  static_assert((TM / RM) * (TN / RN) == 256, "256 threads");
  constexpr int TX = TN / RN;  // threads along n
  const GemmItem it = items[blockIdx.y];
  const int tiles_n = (it.n + TN - 1) / TN;
  const int tiles_m = (it.m + TM - 1) / TM;
  if ((int)blockIdx.x >= tiles_m * tiles_n) return;
  const int row0 = (blockIdx.x / tiles_n) * TM;
  const int col0 = (blockIdx.x % tiles_n) * TN;
  __shared__ T As[TK][TM + 1];
  __shared__ T Bs[TK][TN + 1];
  const int tid = threadIdx.x;
  const int tx = tid % TX, ty = tid / TX;
  T acc[RM][RN];
#pragma unroll
  for (int a = 0; a < RM; ++a)
#pragma unroll
    for (int b = 0; b < RN; ++b) acc[a][b] = T(0);
  const T* __restrict__ A = static_cast<const T*>(it.A);
  const T* __restrict__ B = static_cast<const T*>(it.B);
  T* C = static_cast<T*>(it.C);
  const T alpha = static_cast<T>(it.alpha), beta = static_cast<T>(it.beta);
  const bool a_kfast = (it.a_cs == 1 && it.a_rs != 1);
  const bool b_kfast = (it.b_rs == 1 && it.b_cs != 1);
  constexpr int AE = (TM * TK + 255) / 256, BE = (TN * TK + 255) / 256;
  // chunk loads go to registers first: the next chunk is in flight while the
  // current one is multiplied out of LDS
  T ra[AE], rb[BE];
  auto load_chunk = [&](int k0) {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      const int idx = tid + e * 256;
      ra[e] = T(0);
      if (idx < TM * TK) {
        int i, kk;
        if (a_kfast) {
          kk = idx % TK;
          i = idx / TK;
        } else {
          i = idx % TM;
          kk = idx / TM;
        }
        const int gi = row0 + i, gk = k0 + kk;
        if (gi < it.m && gk < it.k) ra[e] = A[(int64_t)gi * it.a_rs + (int64_t)gk * it.a_cs];
      }
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      const int idx = tid + e * 256;
      rb[e] = T(0);
      if (idx < TN * TK) {
        int j, kk;
        if (b_kfast) {
          kk = idx % TK;
          j = idx / TK;
        } else {
          j = idx % TN;
          kk = idx / TN;
        }
        const int gj = col0 + j, gk = k0 + kk;
        if (gj < it.n && gk < it.k) rb[e] = B[(int64_t)gk * it.b_rs + (int64_t)gj * it.b_cs];
      }
    }
  };
  auto store_chunk = [&]() {
#pragma unroll
    for (int e = 0; e < AE; ++e) {
      const int idx = tid + e * 256;
      if (idx < TM * TK) {
        const int i = a_kfast ? idx / TK : idx % TM, kk = a_kfast ? idx % TK : idx / TM;
        As[kk][i] = ra[e];
      }
    }
#pragma unroll
    for (int e = 0; e < BE; ++e) {
      const int idx = tid + e * 256;
      if (idx < TN * TK) {
        const int j = b_kfast ? idx / TK : idx % TN, kk = b_kfast ? idx % TK : idx / TN;
        Bs[kk][j] = rb[e];
      }
    }
  };
  // the C tile is read early so its latency overlaps the K loop
  T cv[RM][RN];
#pragma unroll
  for (int a = 0; a < RM; ++a)
#pragma unroll
    for (int b = 0; b < RN; ++b) {
      cv[a][b] = T(0);
      const int i = row0 + ty + (TM / RM) * a, j = col0 + tx + TX * b;
      if (beta != T(0) && i < it.m && j < it.n) cv[a][b] = C[(int64_t)i * it.c_rs + (int64_t)j * it.c_cs];
    }
  if (it.k > 0) load_chunk(0);
  for (int k0 = 0; k0 < it.k; k0 += TK) {
    store_chunk();
    __syncthreads();
    if (k0 + TK < it.k) load_chunk(k0 + TK);
#pragma unroll 8
    for (int kk = 0; kk < TK; ++kk) {
      T av[RM], bv[RN];
#pragma unroll
      for (int a = 0; a < RM; ++a) av[a] = As[kk][ty + (TM / RM) * a];
#pragma unroll
      for (int b = 0; b < RN; ++b) bv[b] = Bs[kk][tx + TX * b];
#pragma unroll
      for (int a = 0; a < RM; ++a)
#pragma unroll
        for (int b = 0; b < RN; ++b) acc[a][b] += av[a] * bv[b];
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < RM; ++a) {
    const int i = row0 + ty + (TM / RM) * a;
    if (i >= it.m) continue;
#pragma unroll
    for (int b = 0; b < RN; ++b) {
      const int j = col0 + tx + TX * b;
      if (j >= it.n) continue;
      const int64_t idx = (int64_t)i * it.c_rs + (int64_t)j * it.c_cs;
      T v = alpha * acc[a][b];
      if (beta != T(0)) v += beta * cv[a][b];
      C[idx] = v;
    }
  }
}

template <typename T>
void launch_gemm(const GemmItem* d_items, int n_items, int max_m, int max_n, hipStream_t s) {
  if (n_items <= 0 || max_m <= 0 || max_n <= 0) return;
  if (max_m <= 16 && max_n <= 16) {
    dim3 grid(1, n_items);
    hipLaunchKernelGGL((gemm_kernel<T, 16, 16, 64, 1, 1>), grid, dim3(256), 0, s, d_items);
  } else if (max_n <= 16) {
    dim3 grid(cdiv(max_m, 64), n_items);
    hipLaunchKernelGGL((gemm_kernel<T, 64, 16, 32, 4, 1>), grid, dim3(256), 0, s, d_items);
  } else if (max_m <= 16) {
    dim3 grid(cdiv(max_n, 64), n_items);
    hipLaunchKernelGGL((gemm_kernel<T, 16, 64, 32, 1, 4>), grid, dim3(256), 0, s, d_items);
  } else {
    dim3 grid(cdiv(max_m, 64) * cdiv(max_n, 64), n_items);
    hipLaunchKernelGGL((gemm_kernel<T, 64, 64, 16, 4, 4>), grid, dim3(256), 0, s, d_items);
  }
  FRK_HIP_CHECK(hipGetLastError());
}

// ===========================================================================
// Batched left triangular solve, in place, blocked by 64 x 64 diagonal blocks
// whose inverses are given:  X_k = inv(T_kk) (B_k - sum_j T_kj X_j).
// One workgroup per (item, strip of 16 right-hand sides).  The strip lives in
// LDS; every T block is staged through LDS with coalesced loads, and the next
// block is prefetched into registers while the current one is consumed.
// ===========================================================================
constexpr int kTrsmMaxN = 512;
constexpr int kTrsmW = 16;

template <typename T>
__device__ __forceinline__ void trsm_fetch(T (&pf)[16], const T* __restrict__ A, int64_t ars, int64_t acs, bool acol,
                                           int r0, int nk, int q0, int nj, int tid) {
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = tid + u * 256;
    int a, b;
    if (acol) {
      a = e >> 6;
      b = e & 63;
    } else {
      b = e >> 6;
      a = e & 63;
    }
    pf[u] = (a < nk && b < nj) ? A[(int64_t)(r0 + a) * ars + (int64_t)(q0 + b) * acs] : T(0);
  }
}

template <typename T>
__device__ __forceinline__ void trsm_store(const T (&pf)[16], T (*Tb)[kTriBlk + 1], bool acol, int tid) {
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int e = tid + u * 256;
    int a, b;
    if (acol) {
      a = e >> 6;
      b = e & 63;
    } else {
      b = e >> 6;
      a = e & 63;
    }
    Tb[a][b] = pf[u];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void trsm_kernel(const TrsmItem* __restrict__ items) {
  const TrsmItem it = items[blockIdx.y];
  const int c0 = blockIdx.x * kTrsmW;
  if (c0 >= it.nrhs) return;
  const int nc = min(kTrsmW, it.nrhs - c0);
  const int n = it.n;
  __shared__ T X[kTrsmMaxN][kTrsmW + 1];
  __shared__ T Tb[kTriBlk][kTriBlk + 1];
  __shared__ T Yb[kTriBlk][kTrsmW + 1];
  const T* __restrict__ A = static_cast<const T*>(it.A);
  const T* __restrict__ Dv = static_cast<const T*>(it.dinv);
  T* B = static_cast<T*>(it.B);
  const int tid = threadIdx.x;
  const bool rfast = (it.b_rs == 1);
  for (int e = tid; e < n * kTrsmW; e += 256) {
    int r, c;
    if (rfast) {
      r = e % n;
      c = e / n;
    } else {
      r = e / kTrsmW;
      c = e % kTrsmW;
    }
    X[r][c] = (c < nc) ? B[(int64_t)r * it.b_rs + (int64_t)(c0 + c) * it.b_cs] : T(0);
  }
  const int row = tid & 63;
  const int cg = (tid >> 6) * 4;
  const int nblk = (n + 63) / 64;
  const bool acol = (it.a_cs == 1);
  const bool dcol = (it.d_cs == 1);
  T pf[16];
  for (int bb = 0; bb < nblk; ++bb) {
    const int k = it.lower ? bb : nblk - 1 - bb;
    const int r0 = k * 64, nk = min(64, n - r0);
    const int jlo = it.lower ? 0 : k + 1, jhi = it.lower ? k : nblk;
    __syncthreads();  // X complete from the previous block
    T acc[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = (row < nk) ? X[r0 + row][cg + c] : T(0);
    if (jlo < jhi) trsm_fetch(pf, A, it.a_rs, it.a_cs, acol, r0, nk, jlo * 64, min(64, n - jlo * 64), tid);
    for (int j = jlo; j < jhi; ++j) {
      const int q0 = j * 64, nj = min(64, n - q0);
      __syncthreads();  // previous Tb consumed
      trsm_store(pf, Tb, acol, tid);
      __syncthreads();
      if (j + 1 < jhi) trsm_fetch(pf, A, it.a_rs, it.a_cs, acol, r0, nk, q0 + 64, min(64, n - q0 - 64), tid);
      for (int b2 = 0; b2 < nj; ++b2) {
        const T t = Tb[row][b2];
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c] -= t * X[q0 + b2][cg + c];
      }
    }
    // inverse of the diagonal block
    const T* D = Dv + (int64_t)k * (kTriBlk * kTriBlk);
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int e = tid + u * 256;
      int a, b2;
      if (dcol) {
        a = e >> 6;
        b2 = e & 63;
      } else {
        b2 = e >> 6;
        a = e & 63;
      }
      pf[u] = (a < nk && b2 < nk) ? D[(int64_t)a * it.d_rs + (int64_t)b2 * it.d_cs] : T(0);
    }
    __syncthreads();  // Tb / Yb free
#pragma unroll
    for (int c = 0; c < 4; ++c) Yb[row][cg + c] = acc[c];
    trsm_store(pf, Tb, dcol, tid);
    __syncthreads();
    T out[4] = {T(0), T(0), T(0), T(0)};
    for (int b2 = 0; b2 < nk; ++b2) {
      const T d = Tb[row][b2];
#pragma unroll
      for (int c = 0; c < 4; ++c) out[c] += d * Yb[b2][cg + c];
    }
    if (row < nk)
#pragma unroll
      for (int c = 0; c < 4; ++c) X[r0 + row][cg + c] = out[c];
  }
  __syncthreads();
  for (int e = tid; e < n * kTrsmW; e += 256) {
    int r, c;
    if (rfast) {
      r = e % n;
      c = e / n;
    } else {
      r = e / kTrsmW;
      c = e % kTrsmW;
    }
    if (c < nc) B[(int64_t)r * it.b_rs + (int64_t)(c0 + c) * it.b_cs] = X[r][c];
  }
}

template <typename T>
void launch_trsm(const TrsmItem* d_items, int n_items, int max_n, int max_nrhs, hipStream_t s) {
  if (n_items <= 0 || max_nrhs <= 0 || max_n <= 0) return;
  if (max_n > kTrsmMaxN) throw std::runtime_error("trsm: n exceeds 512");
  dim3 grid(cdiv(max_nrhs, kTrsmW), n_items);
  hipLaunchKernelGGL(trsm_kernel<T>, grid, dim3(256), 0, s, d_items);
  FRK_HIP_CHECK(hipGetLastError());
}

// Inverses of the 64 x 64 diagonal blocks of a strided triangle: lane j
// solves for column j by substitution against the block staged in LDS.
template <typename T>
__global__ __launch_bounds__(64) void tri_inv_kernel(const TriInvItem* __restrict__ items) {
  const TriInvItem it = items[blockIdx.x];
  const T* A = static_cast<const T*>(it.a);
  T* out = static_cast<T*>(it.out);
  const int n = it.n, lane = threadIdx.x;
  __shared__ T Bk[64][65];
  for (int i = 0; i < 64; ++i)
    Bk[i][lane] = (i < n && lane < n) ? A[(int64_t)i * it.rs + (int64_t)lane * it.cs] : T(i == lane ? 1 : 0);
  __syncthreads();
  T x[64];
  if (it.lower) {
#pragma unroll
    for (int i = 0; i < 64; ++i) {
      T sacc = (i == lane) ? T(1) : T(0);
#pragma unroll
      for (int kk = 0; kk < i; ++kk) sacc -= Bk[i][kk] * x[kk];
      x[i] = (it.unit || i >= n) ? sacc : sacc / Bk[i][i];
    }
  } else {
#pragma unroll
    for (int i = 63; i >= 0; --i) {
      T sacc = (i == lane) ? T(1) : T(0);
#pragma unroll
      for (int kk = i + 1; kk < 64; ++kk) sacc -= Bk[i][kk] * x[kk];
      x[i] = (it.unit || i >= n) ? sacc : sacc / Bk[i][i];
    }
  }
#pragma unroll
  for (int i = 0; i < 64; ++i) out[i * 64 + lane] = x[i];
}

template <typename T>
void launch_tri_inv(const TriInvItem* d_items, int n_items, hipStream_t s) {
  if (n_items <= 0) return;
  hipLaunchKernelGGL(tri_inv_kernel<T>, dim3(n_items), dim3(64), 0, s, d_items);
  FRK_HIP_CHECK(hipGetLastError());
}

// ===========================================================================
// No-pivot LU of a <= 64 x 64 diagonal block by 256 threads (thread t owns row
// t / 4 and columns (t % 4) + 4 q), two barriers per pivot; then the inverses
// of the unit-lower and upper factors by recursive doubling in LDS: 16 x 16
// diagonal blocks by substitution, combined 16 -> 32 -> 64 with small products
// ([[A,0],[C,D]]^-1 has off-diagonal block -D^-1 C A^-1; the upper case is
// symmetric).  Padding beyond n is the identity.
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(256) void getrf_small_kernel(const LuItem* __restrict__ items,
                                                         int32_t* __restrict__ info) {
  const LuItem it = items[blockIdx.x];
  T* A = static_cast<T*>(it.a);
  const int n = it.n;
  const int64_t rs = it.rs, cs = it.cs;
  const int tid = threadIdx.x;
  const int row = tid >> 2, cq = tid & 3;
  __shared__ T rowk[2][64];
  __shared__ T colk[2][64];
  __shared__ T M[64][65];
  __shared__ T Li[64][65];
  __shared__ T Ui[64][65];
  __shared__ T W[2][32][33];
  __shared__ int bad_sh;
  T a[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int j = cq + 4 * q;
    a[q] = (row < n && j < n) ? A[row * rs + j * cs] : T(row == j ? 1 : 0);
  }
  if (tid == 0) bad_sh = 0;
#pragma unroll
  for (int k = 0; k < 64; ++k) {
    const int buf = k & 1;
    if (row == k) {
#pragma unroll
      for (int q = 0; q < 16; ++q) rowk[buf][cq + 4 * q] = a[q];
    }
    __syncthreads();
    const T piv = rowk[buf][k];
    if (cq == (k & 3) && row > k) {
      const T l = a[k >> 2] / piv;
      a[k >> 2] = l;
      colk[buf][row] = l;
    }
    if (tid == 0 && piv == T(0) && k < n && bad_sh == 0) bad_sh = k + 1;
    __syncthreads();
    if (row > k) {
      const T l = colk[buf][row];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int j = cq + 4 * q;
        if (j > k) a[q] -= l * rowk[buf][j];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int j = cq + 4 * q;
    if (row < n && j < n) A[row * rs + j * cs] = a[q];
    M[row][j] = a[q];
    Li[row][j] = T(0);
    Ui[row][j] = T(0);
  }
  __syncthreads();
  if (tid == 0 && info != nullptr) info[blockIdx.x] = bad_sh;
  if (it.linv == nullptr && it.uinv == nullptr) return;
  // level 0: eight 16 x 16 substitutions (4 unit-lower, 4 upper), 16 columns each
  if (tid < 128) {
    const int prob = tid >> 4, col = tid & 15;
    const int o = 16 * (prob & 3);
    T x[16];
    if (prob < 4) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        T sacc = (r == col) ? T(1) : T(0);
#pragma unroll
        for (int kk = 0; kk < r; ++kk) sacc -= M[o + r][o + kk] * x[kk];
        x[r] = sacc;
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) Li[o + r][o + col] = x[r];
    } else {
#pragma unroll
      for (int r = 15; r >= 0; --r) {
        T sacc = (r == col) ? T(1) : T(0);
#pragma unroll
        for (int kk = r + 1; kk < 16; ++kk) sacc -= M[o + r][o + kk] * x[kk];
        x[r] = sacc / M[o + r][o + r];
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) Ui[o + r][o + col] = x[r];
    }
  }
  __syncthreads();
  // levels 1 (s = 16, two pairs) and 2 (s = 32, one pair)
#pragma unroll
  for (int lev = 0; lev < 2; ++lev) {
    const int sz = 16 << lev;
    const int npair = 2 >> lev;  // pairs per factor
    // W = C A^-1 (lower) and B D^-1 (upper), sz x sz each, for every pair
    for (int e = tid; e < 2 * npair * sz * sz; e += 256) {
      const int f = e / (npair * sz * sz);         // 0: L, 1: U
      const int pr = (e / (sz * sz)) % npair;
      const int r = (e / sz) % sz, c = e % sz;
      const int o0 = 2 * sz * pr, o1 = o0 + sz;
      T sacc = T(0);
      if (f == 0) {
        for (int t = 0; t < sz; ++t) sacc += M[o1 + r][o0 + t] * Li[o0 + t][o0 + c];
      } else {
        for (int t = 0; t < sz; ++t) sacc += M[o0 + r][o1 + t] * Ui[o1 + t][o1 + c];
      }
      W[f][pr * sz + r][c] = sacc;
    }
    __syncthreads();
    for (int e = tid; e < 2 * npair * sz * sz; e += 256) {
      const int f = e / (npair * sz * sz);
      const int pr = (e / (sz * sz)) % npair;
      const int r = (e / sz) % sz, c = e % sz;
      const int o0 = 2 * sz * pr, o1 = o0 + sz;
      T sacc = T(0);
      if (f == 0) {
        for (int t = 0; t < sz; ++t) sacc += Li[o1 + r][o1 + t] * W[0][pr * sz + t][c];
        Li[o1 + r][o0 + c] = -sacc;
      } else {
        for (int t = 0; t < sz; ++t) sacc += Ui[o0 + r][o0 + t] * W[1][pr * sz + t][c];
        Ui[o0 + r][o1 + c] = -sacc;
      }
    }
    __syncthreads();
  }
  T* lo = static_cast<T*>(it.linv);
  T* uo = static_cast<T*>(it.uinv);
  const int64_t ld = it.ld_inv;
  for (int e = tid; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    if (r < n && c < n) {
      if (lo) lo[r * ld + c] = Li[r][c];
      if (uo) uo[r * ld + c] = Ui[r][c];
    }
  }
}

// ===========================================================================
// Dataflow LU of one tile (n <= 512): workgroup (bi, bj) owns the 64 x 64
// block (bi, bj), held in LDS.  It waits for the blocks it needs through
// per-block flags, applies its Schur updates (FP64 MFMA), factorizes
// (diagonal block: 16-blocked LU plus both 64 x 64 triangular inverses) or
// solves (panel blocks: products with the diagonal inverses), and finally
// computes its block of the full inverses of L and U:
//   Linv_ij = -Linv_ii sum_{k=j}^{i-1} L_ik Linv_kj        (i > j)
//   Uinv_ij = -(sum_{k=i}^{j-1} Uinv_ik U_kj) Uinv_jj      (i < j)
// Flags: F(i,j) = factor block final, I(i,j) = inverse block final; both
// hold the launch epoch once set.
// ===========================================================================
// Cross-workgroup handoff.  The waiting thread polls with relaxed agent-scope
// loads (no cache maintenance per poll) and issues one acquire fence once the
// flag is seen; the producer's workgroup barrier followed by one release store
// publishes every wave's stores (release is cumulative through the barrier).
__device__ __forceinline__ bool tile_wait(const int32_t* flag, int32_t target, int64_t limit, int32_t* err,
                                          int* bail_sh) {
  // err holds the latest epoch whose factorization timed out: a wait gives up
  // once its own epoch is marked failed, so one timeout ends every waiter.
  if (threadIdx.x == 0) {
    int64_t spins = 0;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++spins > limit) {
        *bail_sh = 1;
        __hip_atomic_fetch_max(err, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      if ((spins & 31) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) {
        *bail_sh = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return *bail_sh == 0;
}

__device__ __forceinline__ void tile_signal(int32_t* flag, int32_t value) {
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(flag, value, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// 64 x 64 blocks live in LDS with row stride kTileLd (the padding spreads the
// MFMA operand reads over the banks); rows / columns beyond the matrix edge
// are zero (identity on a diagonal block).
constexpr int kTileLd = 68;

// Block loads issue all 16 of a thread's loads before the first use: each load
// is unconditional, at the element itself or (outside the block) at element
// (0, 0), and zero is selected afterwards, so no branch separates the loads and
// one memory round trip serves the whole block.
__device__ __forceinline__ void blk_load(double* dst, const double* src, int64_t ld, int nr, int nc) {
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = threadIdx.x + 256 * q, r = e >> 6, c = e & 63;
    const bool in = r < nr && c < nc;
    v[q] = src[in ? (int64_t)r * ld + c : 0];
    v[q] = in ? v[q] : 0.0;
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = threadIdx.x + 256 * q, r = e >> 6, c = e & 63;
    dst[r * kTileLd + c] = v[q];
  }
}
__device__ __forceinline__ void blk_store(const double* src, double* dst, int64_t ld, int nr, int nc) {
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int r = e >> 6, c = e & 63;
    if (r < nr && c < nc) dst[(int64_t)r * ld + c] = src[r * kTileLd + c];
  }
}
// Store of a factor64 inverse: the 16 x 16 tiles on the far side of the block
// diagonal (above it for the lower factor, below it for the upper) hold no
// data in LDS and are written as zeros.
__device__ __forceinline__ void blk_store_tri(const double* src, double* dst, int64_t ld, int nr, int nc, bool lower) {
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int r = e >> 6, c = e & 63;
    const bool keep = lower ? ((c >> 4) <= (r >> 4)) : ((c >> 4) >= (r >> 4));
    if (r < nr && c < nc) dst[(int64_t)r * ld + c] = keep ? src[r * kTileLd + c] : 0.0;
  }
}
__device__ __forceinline__ void blk_zero(double* dst, int64_t ld, int nr, int nc) {
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int r = e >> 6, c = e & 63;
    if (r < nr && c < nc) dst[(int64_t)r * ld + c] = 0.0;
  }
}
// FP32 input / output variants (the mixed-precision path keeps its tiles in FP32)
__device__ __forceinline__ void blk_load32(double* dst, const float* src, int64_t ld, int nr, int nc) {
  float v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = threadIdx.x + 256 * q, r = e >> 6, c = e & 63;
    const bool in = r < nr && c < nc;
    v[q] = src[in ? (int64_t)r * ld + c : 0];
    v[q] = in ? v[q] : 0.f;
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = threadIdx.x + 256 * q, r = e >> 6, c = e & 63;
    dst[r * kTileLd + c] = (double)v[q];
  }
}
__device__ __forceinline__ void blk_store32(const double* src, float* dst, int64_t ld, int nr, int nc) {
  if (!dst) return;
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int r = e >> 6, c = e & 63;
    if (r < nr && c < nc) dst[(int64_t)r * ld + c] = (float)src[r * kTileLd + c];
  }
}
__device__ __forceinline__ void blk_store_tri32(const double* src, float* dst, int64_t ld, int nr, int nc,
                                                bool lower) {
  if (!dst) return;
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int r = e >> 6, c = e & 63;
    const bool keep = lower ? ((c >> 4) <= (r >> 4)) : ((c >> 4) >= (r >> 4));
    if (r < nr && c < nc) dst[(int64_t)r * ld + c] = keep ? (float)src[r * kTileLd + c] : 0.f;
  }
}
__device__ __forceinline__ void blk_zero32(float* dst, int64_t ld, int nr, int nc) {
  if (!dst) return;
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int r = e >> 6, c = e & 63;
    if (r < nr && c < nc) dst[(int64_t)r * ld + c] = 0.f;
  }
}

__global__ __launch_bounds__(256) void tile_lu_kernel(TileLuItem it) {
  using d64::mma64;
  const int n = it.n;
  const int nbk = (n + 63) / 64;
  const int bi = blockIdx.x / nbk, bj = blockIdx.x % nbk;
  const int tid = threadIdx.x;
  const int64_t ld = it.ld, ldi = it.ldi;
  double* A = static_cast<double*>(it.a);
  double* Li = static_cast<double*>(it.linv);
  double* Ui = static_cast<double*>(it.uinv);
  int32_t* F = it.flags;
  int32_t* I = it.flags + nbk * nbk;
  const int32_t ep = it.epoch;
  const int64_t lim = it.spin_limit;
  auto rows = [&](int b) { return min(64, n - 64 * b); };
  constexpr int LD = kTileLd;
  __shared__ __align__(16) double M[64 * LD];
  __shared__ __align__(16) double Ls[64 * LD];
  __shared__ __align__(16) double Us[64 * LD];
  __shared__ int bail_sh, zero_sh;
  if (tid == 0) bail_sh = 0;
  auto stamp = [&](int slot) {
    if (it.trace && tid == 0) it.trace[4 * blockIdx.x + slot] = wall_clock64();
  };
  stamp(0);
  const int mi = rows(bi), nj = rows(bj);
  double* Ablk = A + (int64_t)(64 * bi) * ld + 64 * bj;
  const int64_t off_a = (int64_t)(64 * bi) * ld + 64 * bj, off_i = (int64_t)(64 * bi) * ldi + 64 * bj;
  if (it.a32_in)
    blk_load32(M, it.a32_in + off_a, ld, mi, nj);
  else
    blk_load(M, Ablk, ld, mi, nj);
  if (bi == bj)
    for (int r = mi + tid; r < 64; r += 256) M[r * LD + r] = 1.0;  // identity padding of the last diagonal block
  if (it.pre_core != nullptr) {
    // 0. the low-rank update of the tile, on this block: M += U_blk (core V_blk)
    __syncthreads();
    const int kc = it.pre_kcap;
    const int r1 = min(kc, max(0, *it.pre_r1)), r2 = min(kc, max(0, *it.pre_r2));
    if (r1 > 0 && r2 > 0) {
      auto ld_el = [&](const void* p, int64_t e) -> double {
        return it.pre_f32 ? (double)static_cast<const float*>(p)[e] : static_cast<const double*>(p)[e];
      };
      for (int e = tid; e < 64 * 64; e += 256) {
        const int a = e >> 6, b = e & 63;
        Ls[a * LD + b] = (a < r1 && b < r2) ? ld_el(it.pre_core, (int64_t)a * it.pre_ldk + b) : 0.0;
        Us[a * LD + b] = (a < r2 && b < nj) ? ld_el(it.pre_v, (int64_t)a * it.pre_v_rs + 64 * bj + b) : 0.0;
      }
      __syncthreads();
      double w[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int e = tid + 256 * q, a = e >> 6, c = e & 63;
        double acc = 0.0;
        if (a < r1)
          for (int b = 0; b < r2; ++b) acc += Ls[a * LD + b] * Us[b * LD + c];
        w[q] = acc;
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int e = tid + 256 * q;
        Us[(e >> 6) * LD + (e & 63)] = w[q];
      }
      for (int e = tid; e < 64 * 64; e += 256) {
        const int row = e >> 6, a = e & 63;
        Ls[row * LD + a] = (row < mi && a < r1) ? ld_el(it.pre_u, (int64_t)(64 * bi + row) + a * it.pre_u_cs) : 0.0;
      }
      __syncthreads();
      mma64(M, LD, Ls, LD, Us, LD, 64, 1.0, true, false);
    }
  }
  __syncthreads();
  // 1. Schur updates from the earlier block columns / rows
  const int pmax = min(bi, bj);
  for (int p = 0; p < pmax; ++p) {
    if (!tile_wait(&F[bi * nbk + p], ep, lim, it.err, &bail_sh)) return;
    if (!tile_wait(&F[p * nbk + bj], ep, lim, it.err, &bail_sh)) return;
    blk_load(Ls, A + (int64_t)(64 * bi) * ld + 64 * p, ld, mi, 64);
    blk_load(Us, A + (int64_t)(64 * p) * ld + 64 * bj, ld, 64, nj);
    __syncthreads();
    mma64(M, LD, Ls, LD, Us, LD, 64, -1.0, true, false);
    __syncthreads();
  }
  stamp(1);
  if (bi == bj) {
    // 2a. LU of the diagonal block and the full inverses of both factors
    const int z = d64::factor64(M, Ls, Us, LD, &zero_sh);
    if (tid == 0 && z >= 0 && z < mi) atomicCAS(it.info, 0, 64 * bi + z + 1);
    // the inverses go first, each with its own flag (the L panels of this block
    // column need only U^-1, the U panels of its row only L^-1): I(k, k) marks
    // U^-1, F(k, k) marks L^-1; nobody waits for the packed L\U block
    blk_store_tri(Us, Ui + (int64_t)(64 * bi) * ldi + 64 * bi, ldi, mi, mi, false);
    tile_signal(&I[bi * nbk + bj], ep);
    stamp(2);
    blk_store_tri(Ls, Li + (int64_t)(64 * bi) * ldi + 64 * bi, ldi, mi, mi, true);
    tile_signal(&F[bi * nbk + bj], ep);
    stamp(3);
    blk_store(M, Ablk, ld, mi, nj);
    blk_store32(M, it.a32_out ? it.a32_out + off_a : nullptr, ld, mi, nj);
    blk_store_tri32(Ls, it.linv32 ? it.linv32 + off_i : nullptr, ldi, mi, mi, true);
    blk_store_tri32(Us, it.uinv32 ? it.uinv32 + off_i : nullptr, ldi, mi, mi, false);
    return;
  }
  // 2b. panel blocks: L_ij = A_ij Uinv_jj (i > j), U_ij = Linv_ii A_ij (i < j)
  const int dgb = (bi > bj) ? bj : bi;
  if (!tile_wait(bi > bj ? &I[dgb * nbk + dgb] : &F[dgb * nbk + dgb], ep, lim, it.err, &bail_sh)) return;
  if (bi > bj) {
    blk_load(Us, Ui + (int64_t)(64 * bj) * ldi + 64 * bj, ldi, nj, nj);
    __syncthreads();
    mma64(M, LD, M, LD, Us, LD, 64, 1.0, false, false);
  } else {
    blk_load(Ls, Li + (int64_t)(64 * bi) * ldi + 64 * bi, ldi, mi, mi);
    __syncthreads();
    mma64(M, LD, Ls, LD, M, LD, 64, 1.0, false, true);
  }
  __syncthreads();
  blk_store(M, Ablk, ld, mi, nj);
  tile_signal(&F[bi * nbk + bj], ep);
  stamp(2);
  blk_store32(M, it.a32_out ? it.a32_out + off_a : nullptr, ld, mi, nj);
  // 3. block (bi, bj) of the full inverse of L (below the diagonal) or U (above)
  if (bi > bj) {
    // M <- L_ij Linv_jj + sum_{k=j+1}^{i-1} L_ik Linv_kj, then Linv_ij = -Linv_ii M
    if (!tile_wait(&F[bj * nbk + bj], ep, lim, it.err, &bail_sh)) return;
    blk_load(Us, Li + (int64_t)(64 * bj) * ldi + 64 * bj, ldi, nj, nj);
    __syncthreads();
    mma64(M, LD, M, LD, Us, LD, 64, 1.0, false, false);
    __syncthreads();
    for (int k = bj + 1; k < bi; ++k) {
      if (!tile_wait(&F[bi * nbk + k], ep, lim, it.err, &bail_sh)) return;
      if (!tile_wait(&I[k * nbk + bj], ep, lim, it.err, &bail_sh)) return;
      blk_load(Ls, A + (int64_t)(64 * bi) * ld + 64 * k, ld, mi, 64);
      blk_load(Us, Li + (int64_t)(64 * k) * ldi + 64 * bj, ldi, 64, nj);
      __syncthreads();
      mma64(M, LD, Ls, LD, Us, LD, 64, 1.0, true, false);
      __syncthreads();
    }
    if (!tile_wait(&F[bi * nbk + bi], ep, lim, it.err, &bail_sh)) return;
    blk_load(Ls, Li + (int64_t)(64 * bi) * ldi + 64 * bi, ldi, mi, mi);
    __syncthreads();
    mma64(M, LD, Ls, LD, M, LD, 64, -1.0, false, true);
    __syncthreads();
    blk_store(M, Li + (int64_t)(64 * bi) * ldi + 64 * bj, ldi, mi, nj);
    blk_zero(Ui + (int64_t)(64 * bi) * ldi + 64 * bj, ldi, mi, nj);
    blk_store32(M, it.linv32 ? it.linv32 + off_i : nullptr, ldi, mi, nj);
    blk_zero32(it.uinv32 ? it.uinv32 + off_i : nullptr, ldi, mi, nj);
  } else {
    // M <- Uinv_ii U_ij + sum_{k=i+1}^{j-1} Uinv_ik U_kj, then Uinv_ij = -M Uinv_jj
    if (!tile_wait(&I[bi * nbk + bi], ep, lim, it.err, &bail_sh)) return;
    blk_load(Ls, Ui + (int64_t)(64 * bi) * ldi + 64 * bi, ldi, mi, mi);
    __syncthreads();
    mma64(M, LD, Ls, LD, M, LD, 64, 1.0, false, true);
    __syncthreads();
    for (int k = bi + 1; k < bj; ++k) {
      if (!tile_wait(&I[bi * nbk + k], ep, lim, it.err, &bail_sh)) return;
      if (!tile_wait(&F[k * nbk + bj], ep, lim, it.err, &bail_sh)) return;
      blk_load(Ls, Ui + (int64_t)(64 * bi) * ldi + 64 * k, ldi, mi, 64);
      blk_load(Us, A + (int64_t)(64 * k) * ld + 64 * bj, ld, 64, nj);
      __syncthreads();
      mma64(M, LD, Ls, LD, Us, LD, 64, 1.0, true, false);
      __syncthreads();
    }
    if (!tile_wait(&I[bj * nbk + bj], ep, lim, it.err, &bail_sh)) return;
    blk_load(Us, Ui + (int64_t)(64 * bj) * ldi + 64 * bj, ldi, nj, nj);
    __syncthreads();
    mma64(M, LD, M, LD, Us, LD, 64, -1.0, false, false);
    __syncthreads();
    blk_store(M, Ui + (int64_t)(64 * bi) * ldi + 64 * bj, ldi, mi, nj);
    blk_zero(Li + (int64_t)(64 * bi) * ldi + 64 * bj, ldi, mi, nj);
    blk_store32(M, it.uinv32 ? it.uinv32 + off_i : nullptr, ldi, mi, nj);
    blk_zero32(it.linv32 ? it.linv32 + off_i : nullptr, ldi, mi, nj);
  }
  tile_signal(&I[bi * nbk + bj], ep);
  stamp(3);
}

void launch_tile_lu(const TileLuItem& it, hipStream_t s) {
  if (it.n <= 0 || it.n > 512) throw std::runtime_error("tile LU: n must be in [1, 512]");
  const int nbk = (it.n + 63) / 64;
  hipLaunchKernelGGL(tile_lu_kernel, dim3(nbk * nbk), dim3(256), 0, s, it);
  FRK_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_getrf_small(const LuItem* d_items, int n_items, int32_t* d_info, hipStream_t s) {
  if (n_items <= 0) return;
  hipLaunchKernelGGL(getrf_small_kernel<T>, dim3(n_items), dim3(256), 0, s, d_items, d_info);
  FRK_HIP_CHECK(hipGetLastError());
}

// Batched strided copy: C <- A (m x n), GemmItem fields A/a_rs/a_cs, C/c_rs/c_cs, m, n.
template <typename T>
__global__ __launch_bounds__(256) void copy_kernel(const GemmItem* __restrict__ items) {
  const GemmItem it = items[blockIdx.y];
  const T* src = static_cast<const T*>(it.A);
  T* dst = static_cast<T*>(it.C);
  const int64_t total = (int64_t)it.m * it.n;
  const bool rfast = (it.c_rs == 1);
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    int64_t i, j;
    if (rfast) {
      i = e % it.m;
      j = e / it.m;
    } else {
      i = e / it.n;
      j = e % it.n;
    }
    dst[i * it.c_rs + j * it.c_cs] = src[i * it.a_rs + j * it.a_cs];
  }
}

template <typename T>
void launch_copy(const GemmItem* d_items, int n_items, int64_t max_elems, hipStream_t s) {
  if (n_items <= 0 || max_elems <= 0) return;
  const int gx = (int)std::min<int64_t>((max_elems + 255) / 256, 64);
  hipLaunchKernelGGL(copy_kernel<T>, dim3(gx, n_items), dim3(256), 0, s, d_items);
  FRK_HIP_CHECK(hipGetLastError());
}

// ===========================================================================
// Item sources of the batched QR / SVD / basis-update kernels: an uploaded item
// array, or the items of a batch of low-rank additions derived on the device
// from their descriptors (the same construction the host would do, so only the
// additions' descriptors cross PCIe).  Addition t owns QR items 2t (U side) and
// 2t+1 (V side), SVD item t, basis-update items 2t and 2t+1.
// ===========================================================================
template <class Item>
struct ItemArray {
  const Item* p;
  __device__ Item get(int b) const { return p[b]; }
};

template <typename T>
__device__ __forceinline__ bool lr_fast(const LrAddDesc<T>& x, const LrBatch<T>& b) {
  return b.lra == 2 && x.eps == 0.0;
}
// the core size of a rounded addition: K = kc + kx2, or less when K exceeds a
// tile dimension (the QR factors then have min(m, K) / min(n, K) nonzero rows)
template <typename T>
__device__ __forceinline__ int lr_core_dim(const LrAddDesc<T>& x, const LrBatch<T>& b) {
  const int Kt = x.kc + x.kx2;
  if (b.Mw == nullptr || lr_fast(x, b)) return Kt;
  return max(min(x.m, Kt), min(x.n, Kt));
}

template <typename T>
struct LrQrItems {
  const LrAddDesc<T>* d;
  LrBatch<T> b;
  __device__ QrItem get(int w) const {
    const int t = w >> 1;
    const LrAddDesc<T> x = d[t];
    const int K = b.K;
    QrItem q{};
    if ((w & 1) == 0) {
      q.Y = b.Yu + (size_t)t * b.mmax * K;
      q.y_cs = x.m;
      q.tau = b.tu + (size_t)t * K;
      q.R = b.Ru + (size_t)t * K * K;
      q.m = x.m;
    } else {
      q.Y = b.Yv + (size_t)t * b.nmax * K;
      q.y_cs = x.n;
      q.tau = b.tv + (size_t)t * K;
      q.R = b.Rv + (size_t)t * K * K;
      q.m = x.n;
    }
    q.y_rs = 1;
    q.ldr = K;
    const bool fast = lr_fast(x, b);
    if ((w & 1) == 0) {
      if (!fast) {  // Uc = [Uc Sc | Ux Sx] (formed inside the QR kernel)
        q.A1 = x.cu;
        q.a1_rs = x.cu_rs;
        q.a1_cs = x.cu_cs;
        q.S1 = x.cs;
        q.s1_ld = x.lds;
        q.k1a = q.k1 = x.kc;
        q.A2 = x.xu;
        q.a2_rs = x.xu_rs;
        q.a2_cs = x.xu_cs;
        q.S2 = x.xs;
        q.s2_ld = x.xs_ld;
        q.k2a = x.kx1;
        q.k2 = x.kx2;
      } else {  // project X's U basis on C's, QR the remainder
        q.A1 = x.xu;
        q.a1_rs = x.xu_rs;
        q.a1_cs = x.xu_cs;
        q.S1 = x.xs;
        q.s1_ld = x.xs_ld;
        q.k1a = x.kx1;
        q.k1 = x.kx2;
        q.k2 = 0;
        q.P = x.cu;
        q.p_rs = x.cu_rs;
        q.p_cs = x.cu_cs;
        q.kp = x.kc;
      }
    } else {
      if (!fast) {  // Vc^T = [Vc^T | Vx^T]
        q.A1 = x.cv;
        q.a1_rs = x.cv_cs;
        q.a1_cs = x.cv_rs;
        q.k1a = q.k1 = x.kc;
        q.A2 = x.xv;
        q.a2_rs = x.xv_cs;
        q.a2_cs = x.xv_rs;
        q.k2a = q.k2 = x.kx2;
      } else {
        q.A1 = x.xv;
        q.a1_rs = x.xv_cs;
        q.a1_cs = x.xv_rs;
        q.k1a = q.k1 = x.kx2;
        q.k2 = 0;
        q.P = x.cv;
        q.p_rs = x.cv_cs;
        q.p_cs = x.cv_rs;
        q.kp = x.kc;
      }
    }
    return q;
  }
};

template <typename T>
struct LrSvdItems {
  const LrAddDesc<T>* d;
  LrBatch<T> b;
  int t0;  // first item of this launch (workspace slots are indexed by the item)
  __device__ SvdItem get(int tl) const {
    const int t = t0 + tl;
    const LrAddDesc<T> x = d[t];
    const int K = b.K;
    const int Kt = x.kc + x.kx2;
    const bool fast = lr_fast(x, b);
    SvdItem sv{};
    sv.mode = 1;
    sv.Ru = b.Ru + (size_t)t * K * K;
    sv.Rv = b.Rv + (size_t)t * K * K;
    sv.ldr = K;
    sv.K = Kt;
    const int Ke = lr_core_dim(x, b);
    if (Ke < Kt) {
      // the core M = Ru Rv^T was formed into Mw (lr_core_prod_kernel)
      sv.mode = 0;
      sv.M = b.Mw + (size_t)t * K * K;
      sv.m_rs = K;
      sv.m_cs = 1;
      sv.K = Ke;
    } else if (!fast) {
      sv.kx = Kt;
      sv.ka = 0;
    } else {
      sv.kx = x.kx2;
      sv.ka = x.kc;
      sv.Sa = x.cs;
      sv.lds_in = x.lds;
    }
    sv.fixed_rank = (x.eps != 0.0) ? -1 : x.kc;
    sv.kdim = min(min(x.m, Kt), min(x.n, Kt));
    sv.eps = x.eps;
    sv.S = x.cs;
    sv.lds = x.lds;
    sv.Um = b.Um + (size_t)t * K * K;
    sv.ldu = K;
    sv.Vm = b.Vm + (size_t)t * K * K;
    sv.ldv = K;
    sv.rank_out = x.d_rank;
    sv.cap = x.cap;
    sv.cap_trunc = x.cap_trunc ? 1 : 0;
    sv.stats = b.stats;
    sv.vwork = b.Vw ? b.Vw + (size_t)t * K * K : nullptr;
    return sv;
  }
};

template <typename T>
struct LrApplyItems {
  const LrAddDesc<T>* d;
  LrBatch<T> b;
  __device__ ApplyQItem get(int w) const {
    const int t = w >> 1;
    const LrAddDesc<T> x = d[t];
    const int K = b.K;
    const int Kt = x.kc + x.kx2;
    const bool fast = lr_fast(x, b);
    T* um = b.Um + (size_t)t * K * K;
    T* vm = b.Vm + (size_t)t * K * K;
    ApplyQItem a{};
    a.y_rs = 1;
    a.ncol_ptr = x.d_rank;
    a.ncol = Kt;
    a.trans = 0;
    a.x_rs = K;
    a.x_cs = 1;
    if ((w & 1) == 0) {
      a.Y = b.Yu + (size_t)t * b.mmax * K;
      a.y_cs = x.m;
      a.tau = b.tu + (size_t)t * K;
      a.m = x.m;
      a.out = x.cu;
      a.out_rs = x.cu_rs;
      a.out_cs = x.cu_cs;
      if (!fast) {
        a.nref = min(x.m, Kt);
        a.X = um;
        a.kx = Kt;
      } else {
        a.nref = min(x.m, x.kx2);
        a.X = um + (size_t)x.kc * K;
        a.kx = x.kx2;
        a.Base = x.cu;
        a.base_rs = x.cu_rs;
        a.base_cs = x.cu_cs;
        a.Xb = um;
        a.xb_rs = K;
        a.xb_cs = 1;
        a.kb = x.kc;
      }
    } else {
      a.Y = b.Yv + (size_t)t * b.nmax * K;
      a.y_cs = x.n;
      a.tau = b.tv + (size_t)t * K;
      a.m = x.n;
      a.out = x.cv;
      a.out_rs = x.cv_cs;
      a.out_cs = x.cv_rs;
      if (!fast) {
        a.nref = min(x.n, Kt);
        a.X = vm;
        a.kx = Kt;
      } else {
        a.nref = min(x.n, x.kx2);
        a.X = vm + (size_t)x.kc * K;
        a.kx = x.kx2;
        a.Base = x.cv;
        a.base_rs = x.cv_cs;
        a.base_cs = x.cv_rs;
        a.Xb = vm;
        a.xb_rs = K;
        a.xb_cs = 1;
        a.kb = x.kc;
      }
    }
    return a;
  }
};

// ===========================================================================
// Householder QR of [A1 S1 | A2 S2] (m x K, m <= 64 * MR, K <= W * CPW).
// Optional MGS projection against an orthonormal basis P first; its
// coefficients occupy rows [0, kp) of the R buffer, R itself rows [kp, kp+K).
// ===========================================================================
template <typename T, int MR, int W, int CPW, class Src>
__global__ __launch_bounds__(W * 64) void qr_kernel(Src items) {
  const QrItem it = items.get(blockIdx.x);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int m = it.m;
  const int K = it.k1 + it.k2;
  const int nref = min(m, K);
  __shared__ T vbuf[2][MR * 64];
  __shared__ T taus[W * CPW];
  T a[CPW][MR];
#pragma unroll
  for (int q = 0; q < CPW; ++q)
#pragma unroll
    for (int t = 0; t < MR; ++t) a[q][t] = T(0);
  constexpr int CS = 8;
  __shared__ T stage[CS][MR * 64];
  __shared__ T sstage[CS][W * CPW];
  for (int blk = 0; blk < 2; ++blk) {
    const T* Ab = static_cast<const T*>(blk ? it.A2 : it.A1);
    const T* S = static_cast<const T*>(blk ? it.S2 : it.S1);
    const int64_t rs = blk ? it.a2_rs : it.a1_rs, cs = blk ? it.a2_cs : it.a1_cs;
    const int sld = blk ? it.s2_ld : it.s1_ld;
    const int kc = blk ? it.k2 : it.k1, ka = blk ? it.k2a : it.k1a;
    const int cbase = blk ? it.k1 : 0;
    if (kc <= 0) continue;
    if (S == nullptr) {
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        const int c = q * W + w;
        if (c >= cbase && c < cbase + kc) {
#pragma unroll
          for (int t = 0; t < MR; ++t) {
            const int row = lane + 64 * t;
            a[q][t] = (row < m) ? Ab[row * rs + (c - cbase) * cs] : T(0);
          }
        }
      }
    } else {
      // column c of A S = sum_s A[:, s] S[s, c]: stage CS columns of A and the
      // matching CS rows of S at a time (coalesced loads, LDS broadcasts)
      for (int s0 = 0; s0 < ka; s0 += CS) {
        const int cnt = min(CS, ka - s0);
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * MR * 64; e += W * 64) {
          const int si = e / (MR * 64), row = e % (MR * 64);
          stage[si][row] = (row < m) ? Ab[row * rs + (int64_t)(s0 + si) * cs] : T(0);
        }
        for (int e = threadIdx.x; e < cnt * kc; e += W * 64) {
          const int si = e / kc, cc = e % kc;
          sstage[si][cc] = S[(s0 + si) * sld + cc];
        }
        __syncthreads();
        for (int si = 0; si < cnt; ++si) {
          T av[MR];
#pragma unroll
          for (int t = 0; t < MR; ++t) av[t] = stage[si][lane + 64 * t];
#pragma unroll
          for (int q = 0; q < CPW; ++q) {
            const int c = q * W + w;
            if (c >= cbase && c < cbase + kc) {
              const T coef = sstage[si][c - cbase];
#pragma unroll
              for (int t = 0; t < MR; ++t) a[q][t] += av[t] * coef;
            }
          }
        }
      }
    }
  }
  T* Rb = static_cast<T*>(it.R);
  const int ldr = it.ldr;
  if (it.P != nullptr && it.kp > 0) {
    const T* P = static_cast<const T*>(it.P);
    for (int s = 0; s < it.kp; ++s) {
      T pv[MR];
#pragma unroll
      for (int t = 0; t < MR; ++t) {
        const int row = lane + 64 * t;
        pv[t] = (row < m) ? P[row * it.p_rs + s * it.p_cs] : T(0);
      }
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        const int c = q * W + w;
        if (c < K) {
          T d = T(0);
#pragma unroll
          for (int t = 0; t < MR; ++t) d += pv[t] * a[q][t];
          d = wave_sum(d);
#pragma unroll
          for (int t = 0; t < MR; ++t) a[q][t] -= d * pv[t];
          if (lane == 0) Rb[s * ldr + c] = d;
        }
      }
    }
  }
  const int kp = (it.P != nullptr) ? it.kp : 0;
  // Step j = qj W + wj reflects column j, a[qj][] of wave wj: qj is unrolled, so the
  // column is addressed in registers directly (no selection over the array).
  static_assert(W * CPW <= 128, "row j < 128 lies in a[][0] or a[][1]");
#pragma unroll
  for (int qj = 0; qj < CPW; ++qj) {
    for (int wj = 0; wj < W; ++wj) {
      const int j = qj * W + wj;
      if (j >= nref) break;
      const int buf = j & 1;
      if (w == wj) {
        T x2 = T(0);
#pragma unroll
        for (int t = 0; t < MR; ++t) {
          const int row = lane + 64 * t;
          if (row > j) x2 += a[qj][t] * a[qj][t];
        }
        x2 = wave_sum(x2);
        T alpha;
        if constexpr (MR == 1)
          alpha = wave_read(a[qj][0], j);  // nref <= m <= 64
        else
          alpha = (j < 64) ? wave_read(a[qj][0], j) : wave_read(a[qj][1], j - 64);
        T beta, tau, scal;
        if (x2 == T(0)) {
          tau = T(0);
          beta = alpha;
          scal = T(0);
        } else {
          const T nrm = sqrt(alpha * alpha + x2);
          beta = (alpha >= T(0)) ? -nrm : nrm;
          tau = (beta - alpha) * recip(beta);
          scal = recip(alpha - beta);
        }
#pragma unroll
        for (int t = 0; t < MR; ++t) {
          const int row = lane + 64 * t;
          if (row > j) a[qj][t] *= scal;
          if (row == j) a[qj][t] = beta;
          vbuf[buf][lane + 64 * t] = (row == j) ? T(1) : ((row > j && row < m) ? a[qj][t] : T(0));
        }
        if (lane == 0) taus[j] = tau;
      }
      __syncthreads();
      const T tj = taus[j];
      if (tj != T(0)) {
        T v[MR];
#pragma unroll
        for (int t = 0; t < MR; ++t) v[t] = vbuf[buf][lane + 64 * t];
#pragma unroll
        for (int q = qj; q < CPW; ++q) {
          const int c = q * W + w;
          if ((q > qj || w > wj) && c < K) {
            T d = T(0);
#pragma unroll
            for (int t = 0; t < MR; ++t) d += v[t] * a[q][t];
            d = wave_sum(d) * tj;
#pragma unroll
            for (int t = 0; t < MR; ++t) a[q][t] -= d * v[t];
          }
        }
      }
    }
  }
  T* Y = static_cast<T*>(it.Y);
  T* tau_out = static_cast<T*>(it.tau);
#pragma unroll
  for (int q = 0; q < CPW; ++q) {
    const int c = q * W + w;
    if (c < K) {
#pragma unroll
      for (int t = 0; t < MR; ++t) {
        const int row = lane + 64 * t;
        if (row < m && row > c && c < nref) Y[row * it.y_rs + c * it.y_cs] = a[q][t];
        if (it.r_upper_only) {
          if (row <= c && row < nref) Rb[(int64_t)(kp + row) * ldr + c] = a[q][t];
        } else if (row < nref) {
          Rb[(int64_t)(kp + row) * ldr + c] = (row <= c) ? a[q][t] : T(0);
        } else if (row < K) {
          Rb[(int64_t)(kp + row) * ldr + c] = T(0);
        }
      }
      if (lane == 0 && c < nref) tau_out[c] = taus[c];
    }
  }
}

template <typename T, int MR, class Src>
static void launch_qr_mr(Src d_items, int n_items, int max_k, hipStream_t s) {
  if (max_k <= 16) {
    hipLaunchKernelGGL((qr_kernel<T, MR, 4, 4, Src>), dim3(n_items), dim3(256), 0, s, d_items);
  } else if (max_k <= 32) {
    hipLaunchKernelGGL((qr_kernel<T, MR, 4, 8, Src>), dim3(n_items), dim3(256), 0, s, d_items);
  } else if (max_k <= 64) {
    // 16 waves x 4 columns: 32 register values per lane stay within the 1024-thread budget
    hipLaunchKernelGGL((qr_kernel<T, MR, 16, 4, Src>), dim3(n_items), dim3(1024), 0, s, d_items);
  } else if (max_k <= 128) {
    hipLaunchKernelGGL((qr_kernel<T, MR, 16, 8, Src>), dim3(n_items), dim3(1024), 0, s, d_items);
  } else {
    throw std::runtime_error("qr: K exceeds 128");
  }
  FRK_HIP_CHECK(hipGetLastError());
}

// max_m: the most rows of any item; a lane holds ceil(max_m / 64) rows (1, 4 or 8)
template <typename T, class Src>
static void launch_qr_src(Src d_items, int n_items, int max_k, hipStream_t s, int max_m = 512) {
  if (n_items <= 0) return;
  if (max_m > 512) throw std::runtime_error("qr: more than 512 rows");
  if (max_m <= 64)
    launch_qr_mr<T, 1>(d_items, n_items, max_k, s);
  else if (max_m <= 256)
    launch_qr_mr<T, 4>(d_items, n_items, max_k, s);
  else
    launch_qr_mr<T, 8>(d_items, n_items, max_k, s);
}

template <typename T>
void launch_qr(const QrItem* d_items, int n_items, int max_k, hipStream_t s) {
  launch_qr_src<T>(ItemArray<QrItem>{d_items}, n_items, max_k, s);
}

// ===========================================================================
// out = Q [X; 0] (+ Base Xb), Q = H_0 ... H_{nref-1} (trans: Q^T).
// Columns are independent: no workgroup barrier except before the write when
// a base term may alias the output.
// ===========================================================================
template <typename T, int MR, int W, int CPW, class Src>
__global__ __launch_bounds__(W * 64) void apply_q_kernel(Src items) {
  ApplyQItem it = items.get(blockIdx.x);
  if (it.ncol_ptr != nullptr) it.ncol = *it.ncol_ptr;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int m = it.m;
  const T* Y = static_cast<const T*>(it.Y);
  const T* tau = static_cast<const T*>(it.tau);
  const T* X = static_cast<const T*>(it.X);
  T a[CPW][MR];
#pragma unroll
  for (int q = 0; q < CPW; ++q) {
    const int c = q * W + w;
#pragma unroll
    for (int t = 0; t < MR; ++t) {
      const int row = lane + 64 * t;
      a[q][t] = (X != nullptr && c < it.ncol && row < it.kx && row < m)
                    ? X[row * it.x_rs + c * it.x_cs]
                    : T(0);
    }
  }
  constexpr int CH = 8;
  __shared__ T vs[CH][MR * 64];
  __shared__ T ts[CH];
  const int nref = it.nref;
  for (int base = 0; base < nref; base += CH) {
    const int cnt = min(CH, nref - base);
    __syncthreads();  // previous chunk consumed
    for (int e = threadIdx.x; e < cnt * MR * 64; e += W * 64) {
      const int ci = e / (MR * 64), row = e % (MR * 64);
      const int jj = base + ci;
      const int j = it.trans ? jj : nref - 1 - jj;
      vs[ci][row] = (row == j) ? T(1) : ((row > j && row < m) ? Y[row * it.y_rs + j * it.y_cs] : T(0));
    }
    if ((int)threadIdx.x < cnt) {
      const int jj = base + threadIdx.x;
      ts[threadIdx.x] = tau[it.trans ? jj : nref - 1 - jj];
    }
    __syncthreads();
    for (int ci = 0; ci < cnt; ++ci) {
      const T tj = ts[ci];
      if (tj == T(0)) continue;
      T v[MR];
#pragma unroll
      for (int t = 0; t < MR; ++t) v[t] = vs[ci][lane + 64 * t];
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        const int c = q * W + w;
        if (c < it.ncol) {
          T d = T(0);
#pragma unroll
          for (int t = 0; t < MR; ++t) d += v[t] * a[q][t];
          d = wave_sum(d) * tj;
#pragma unroll
          for (int t = 0; t < MR; ++t) a[q][t] -= d * v[t];
        }
      }
    }
  }
  if (it.Base != nullptr && it.kb > 0) {
    const T* Bs = static_cast<const T*>(it.Base);
    const T* Xb = static_cast<const T*>(it.Xb);
    for (int s = 0; s < it.kb; ++s) {
      T bv[MR];
#pragma unroll
      for (int t = 0; t < MR; ++t) {
        const int row = lane + 64 * t;
        bv[t] = (row < m) ? Bs[row * it.base_rs + s * it.base_cs] : T(0);
      }
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        const int c = q * W + w;
        if (c < it.ncol) {
          const T coef = Xb[s * it.xb_rs + c * it.xb_cs];
#pragma unroll
          for (int t = 0; t < MR; ++t) a[q][t] += bv[t] * coef;
        }
      }
    }
    __syncthreads();  // the base may alias the output
  }
  T* out = static_cast<T*>(it.out);
#pragma unroll
  for (int q = 0; q < CPW; ++q) {
    const int c = q * W + w;
    if (c < it.ncol) {
#pragma unroll
      for (int t = 0; t < MR; ++t) {
        const int row = lane + 64 * t;
        if (row < m) out[row * it.out_rs + c * it.out_cs] = a[q][t];
      }
    }
  }
}

template <typename T, int MR, class Src>
static void launch_apply_q_mr(Src d_items, int n_items, int max_ncol, hipStream_t s) {
  if (max_ncol <= 16) {
    hipLaunchKernelGGL((apply_q_kernel<T, MR, 4, 4, Src>), dim3(n_items), dim3(256), 0, s, d_items);
  } else if (max_ncol <= 32) {
    hipLaunchKernelGGL((apply_q_kernel<T, MR, 4, 8, Src>), dim3(n_items), dim3(256), 0, s, d_items);
  } else if (max_ncol <= 64) {
    hipLaunchKernelGGL((apply_q_kernel<T, MR, 8, 8, Src>), dim3(n_items), dim3(512), 0, s, d_items);
  } else if (max_ncol <= 128) {
    hipLaunchKernelGGL((apply_q_kernel<T, MR, 16, 8, Src>), dim3(n_items), dim3(1024), 0, s, d_items);
  } else {
    throw std::runtime_error("apply_q: more than 128 columns");
  }
  FRK_HIP_CHECK(hipGetLastError());
}

// max_m: the most rows of any item; a lane holds ceil(max_m / 64) rows (1, 4 or 8)
template <typename T, class Src>
static void launch_apply_q_src(Src d_items, int n_items, int max_ncol, hipStream_t s, int max_m = 512) {
  if (n_items <= 0) return;
  if (max_m > 512) throw std::runtime_error("apply_q: more than 512 rows");
  if (max_m <= 64)
    launch_apply_q_mr<T, 1>(d_items, n_items, max_ncol, s);
  else if (max_m <= 256)
    launch_apply_q_mr<T, 4>(d_items, n_items, max_ncol, s);
  else
    launch_apply_q_mr<T, 8>(d_items, n_items, max_ncol, s);
}

template <typename T>
void launch_apply_q(const ApplyQItem* d_items, int n_items, int max_ncol, hipStream_t s) {
  launch_apply_q_src<T>(ItemArray<ApplyQItem>{d_items}, n_items, max_ncol, s);
}

// ===========================================================================
// Pentagonal Householder QR of [Rt; B] (LAPACK tpqrt with l = 0): reflector j
// is [e_j; v_j] with v_j in B's rows, so step j touches only row j of the top
// triangle.  B (m <= 64 * MR rows, K <= W * CPW columns) lives in VGPRs
// (wave w owns columns q * W + w); rows of Rt are staged through LDS RC at a
// time and written back once final.
// ===========================================================================
template <typename T, int MR, int W, int CPW>
__global__ __launch_bounds__(W * 64) void tpqrt_kernel(const TpqrtItem* __restrict__ items) {
  const TpqrtItem it = items[blockIdx.x];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int m = it.m, K = it.K;
  constexpr int KC = W * CPW;
  constexpr int RC = 8;
  __shared__ T vbuf[2][MR * 64];
  __shared__ T taus[KC];
  __shared__ T Rst[RC][KC];
  T* R = static_cast<T*>(it.R);
  T* B = static_cast<T*>(it.B);
  T a[CPW][MR];
#pragma unroll
  for (int q = 0; q < CPW; ++q) {
    const int c = q * W + w;
#pragma unroll
    for (int t = 0; t < MR; ++t) {
      const int row = lane + 64 * t;
      a[q][t] = (row < m && c < K) ? B[row * it.b_rs + c * it.b_cs] : T(0);
    }
  }
  int j0 = -RC;  // first row of the staged chunk
  for (int j = 0; j < K; ++j) {
    const int qj = j / W, wj = j - qj * W;
    if (j - j0 >= RC) {
      // write back the finished chunk, stage the next RC rows (columns >= their row)
      __syncthreads();
      if (j0 >= 0)
        for (int e = threadIdx.x; e < RC * K; e += W * 64) {
          const int rr = e / K, c = e % K, row = j0 + rr;
          if (row < K && c >= row) R[row * it.r_rs + c * it.r_cs] = Rst[rr][c];
        }
      j0 = j;
      __syncthreads();
      for (int e = threadIdx.x; e < RC * K; e += W * 64) {
        const int rr = e / K, c = e % K, row = j0 + rr;
        Rst[rr][c] = (row < K && c >= row) ? R[row * it.r_rs + c * it.r_cs] : T(0);
      }
      // the pivot Rst[jj][j] read below is staged by another wave
      __syncthreads();
    }
    const int jj = j - j0;
    const int buf = j & 1;
    if (w == wj) {
      // column j by predicated selection (keeps a[][] in registers)
      T col[MR];
#pragma unroll
      for (int t = 0; t < MR; ++t) col[t] = a[0][t];
#pragma unroll
      for (int q = 1; q < CPW; ++q)
#pragma unroll
        for (int t = 0; t < MR; ++t) col[t] = (q == qj) ? a[q][t] : col[t];
      T x2 = T(0);
#pragma unroll
      for (int t = 0; t < MR; ++t) x2 += col[t] * col[t];
      x2 = wave_sum(x2);
      const T alpha = Rst[jj][j];
      T beta, tau, scal;
      if (x2 == T(0)) {
        tau = T(0);
        beta = alpha;
        scal = T(1);
      } else {
        const T nrm = sqrt(alpha * alpha + x2);
        beta = (alpha >= T(0)) ? -nrm : nrm;
        tau = (beta - alpha) / beta;
        scal = T(1) / (alpha - beta);
      }
#pragma unroll
      for (int t = 0; t < MR; ++t) {
        col[t] *= scal;
        vbuf[buf][lane + 64 * t] = col[t];
      }
#pragma unroll
      for (int q = 0; q < CPW; ++q)
#pragma unroll
        for (int t = 0; t < MR; ++t) a[q][t] = (q == qj) ? col[t] : a[q][t];
      // every lane has read alpha (wave-uniform) before lane 0 overwrites it
      if (lane == 0) {
        Rst[jj][j] = beta;
        taus[j] = tau;
      }
    }
    __syncthreads();
    const T tj = taus[j];
    if (tj != T(0)) {
      T v[MR];
#pragma unroll
      for (int t = 0; t < MR; ++t) v[t] = vbuf[buf][lane + 64 * t];
#pragma unroll
      for (int q = 0; q < CPW; ++q) {
        const int c = q * W + w;
        if (c > j && c < K) {
          T d = T(0);
#pragma unroll
          for (int t = 0; t < MR; ++t) d += v[t] * a[q][t];
          const T rjc = Rst[jj][c];
          d = (rjc + wave_sum(d)) * tj;
#pragma unroll
          for (int t = 0; t < MR; ++t) a[q][t] -= d * v[t];
          if (lane == 0) Rst[jj][c] = rjc - d;
        }
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < RC * K; e += W * 64) {
    const int rr = e / K, c = e % K, row = j0 + rr;
    if (row < K && c >= row) R[row * it.r_rs + c * it.r_cs] = Rst[rr][c];
  }
  T* tau_out = static_cast<T*>(it.tau);
#pragma unroll
  for (int q = 0; q < CPW; ++q) {
    const int c = q * W + w;
    if (c < K) {
#pragma unroll
      for (int t = 0; t < MR; ++t) {
        const int row = lane + 64 * t;
        if (row < m) B[row * it.b_rs + c * it.b_cs] = a[q][t];
      }
      if (lane == 0) tau_out[c] = taus[c];
    }
  }
}

template <typename T>
void launch_tpqrt(const TpqrtItem* d_items, int n_items, int max_m, int max_k, hipStream_t s) {
  if (n_items <= 0) return;
  if (max_m <= 64 && max_k <= 512) {
    hipLaunchKernelGGL((tpqrt_kernel<T, 1, 16, 32>), dim3(n_items), dim3(1024), 0, s, d_items);
  } else if (max_m <= 512 && max_k <= 64) {
    hipLaunchKernelGGL((tpqrt_kernel<T, 8, 8, 8>), dim3(n_items), dim3(512), 0, s, d_items);
  } else {
    throw std::runtime_error("tpqrt: block exceeds 64 rows x 512 columns or 512 rows x 64 columns");
  }
  FRK_HIP_CHECK(hipGetLastError());
}

// LAPACK larft (forward, columnwise) for K <= 64 reflectors, one wave per item.
template <typename T>
__global__ __launch_bounds__(64) void larft_kernel(const LarftItem* __restrict__ items) {
  const LarftItem it = items[blockIdx.x];
  const int K = it.K, r = threadIdx.x;
  const T* G = static_cast<const T*>(it.G);
  const T* tau = static_cast<const T*>(it.tau);
  __shared__ T Ts[64][65];
  __shared__ T gcol[64];
  for (int j = 0; j < K; ++j) {
    if (r < j) gcol[r] = G[r * it.g_rs + j * it.g_cs];
    __syncthreads();
    const T tj = tau[j];
    if (r < j) {
      T sacc = T(0);
      for (int c = r; c < j; ++c) sacc += Ts[r][c] * gcol[c];
      Ts[r][j] = -tj * sacc;
    } else if (r == j) {
      Ts[j][j] = tj;
    }
    __syncthreads();
  }
  T* Tout = static_cast<T*>(it.T);
  for (int c = 0; c < K; ++c)
    if (r < K) Tout[r * it.t_rs + c * it.t_cs] = (c >= r) ? Ts[r][c] : T(0);
}

template <typename T>
void launch_larft(const LarftItem* d_items, int n_items, hipStream_t s) {
  if (n_items <= 0) return;
  hipLaunchKernelGGL(larft_kernel<T>, dim3(n_items), dim3(64), 0, s, d_items);
  FRK_HIP_CHECK(hipGetLastError());
}

// Copy of a triangle of a strided n x n view into a dense row-major n x n
// buffer: mode 0 = unit lower (strict lower part, ones on the diagonal,
// zeros above), mode 1 = upper with diagonal (zeros below).
template <typename T>
__global__ __launch_bounds__(256) void tri_extract_kernel(const T* __restrict__ src, int64_t rs, int64_t cs,
                                                          T* __restrict__ dst, int64_t ld, int n, int mode) {
  const int64_t total = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e / n), j = (int)(e % n);
    T v;
    if (mode == 0)
      v = (i > j) ? src[i * rs + j * cs] : (i == j ? T(1) : T(0));
    else
      v = (i <= j) ? src[i * rs + j * cs] : T(0);
    dst[(int64_t)i * ld + j] = v;
  }
}

template <typename T>
void launch_tri_extract(const T* src, int64_t rs, int64_t cs, T* dst, int64_t ld, int n, int mode, hipStream_t s) {
  if (n <= 0) return;
  const int g = (int)std::min<int64_t>(((int64_t)n * n + 255) / 256, 1024);
  hipLaunchKernelGGL(tri_extract_kernel<T>, dim3(g), dim3(256), 0, s, src, rs, cs, dst, ld, n, mode);
  FRK_HIP_CHECK(hipGetLastError());
}

// ===========================================================================
// One-sided (Hestenes) Jacobi SVD of the K x K core with truncation.
// ===========================================================================
// Rotation (c, s) that zeroes the off-diagonal of [[al, ga], [ga, be]]
// (Rutishauser's formulas: t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
// zeta = (be - al) / (2 ga), c = 1 / sqrt(1 + t^2), s = c t).  FP32 uses the
// hardware reciprocal, square root and reciprocal square root (about 1 ulp: the rotation
// stays orthogonal to rounding level, only its angle moves by an ulp); FP64
// keeps the correctly rounded operations.
__device__ __forceinline__ void jacobi_rotation(float al, float be, float ga, float& c, float& s) {
  const float zeta = (be - al) * __builtin_amdgcn_rcpf(2.0f * ga);
  const float t = copysignf(1.0f, zeta) * __builtin_amdgcn_rcpf(fabsf(zeta) + __builtin_amdgcn_sqrtf(1.0f + zeta * zeta));
  c = __builtin_amdgcn_rsqf(1.0f + t * t);
  s = c * t;
}
__device__ __forceinline__ void jacobi_rotation(double al, double be, double ga, double& c, double& s) {
  const double zeta = (be - al) / (2.0 * ga);
  const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
  c = 1.0 / sqrt(1.0 + t * t);
  s = c * t;
}

__device__ __forceinline__ int rr_pos(int pos, int r, int k2) {
  return pos == 0 ? 0 : ((pos - 1 + r) % (k2 - 1)) + 1;
}

template <typename T, int KMAX, int NT, bool VG, class Src>
__global__ __launch_bounds__(NT) void svd_kernel(Src items) {
  const SvdItem it = items.get(blockIdx.x);
  const int K = it.K;
  __shared__ T G[KMAX][KMAX + 1];
  // right vectors: LDS, or a K x K global scratch (column j at vg + j * K) when they do not fit
  __shared__ T Vs[VG ? 1 : KMAX][KMAX + 1];
  __shared__ T coef[KMAX];
  __shared__ int accept_sh;
  T* vg = static_cast<T*>(it.vwork);
  auto VA = [&](int j, int row) -> T& {
    if constexpr (VG)
      return vg[(int64_t)j * K + row];
    else
      return Vs[j][row];
  };
  __shared__ T sig[KMAX];
  __shared__ int order[KMAX];
  __shared__ int rank_sh;
  // QR preconditioning (Drmac-Veselic): M P = Q R by column-pivoted Householder,
  // then one-sided Jacobi on R^T, whose graded rows converge in few sweeps.  Q is
  // formed in the right-vector array before the rotations, which then turn it
  // into the left vectors, unless the right vectors live in global memory (VG:
  // the left vectors are then formed as M v / sigma).  Every variant
  // preconditions (the flag marks the preconditioning code).
  constexpr bool PRE = true;
  __shared__ int perm[KMAX];
  __shared__ T tauq[KMAX];
  __shared__ T cnorm[KMAX];
  __shared__ int piv_sh;
  const int tid = threadIdx.x;
  if (K <= 0) {
    if (tid == 0 && it.rank_out) *it.rank_out = 0;
    return;
  }
  if constexpr (VG) {
    if (vg == nullptr) {  // missing scratch: report an impossible rank instead of touching memory
      if (tid == 0 && it.rank_out) *it.rank_out = -(1 << 20);
      return;
    }
  }
  if (it.mode == 0) {
    for (int e = tid; e < K * K; e += NT) {
      const int i = e / K, j = e % K;
      G[j][i] = static_cast<const T*>(it.M)[i * it.m_rs + j * it.m_cs];
      VA(j, i) = (i == j) ? T(1) : T(0);
    }
  } else {
    // stage Ru (K x J) into G and Rv (K x J) into Vs with independent loads
    const int J = it.kx;
    const T* Ru = static_cast<const T*>(it.Ru);
    const T* Rv = static_cast<const T*>(it.Rv);
    for (int e = tid; e < K * J; e += NT) {
      const int i = e / J, t = e % J;
      G[t][i] = Ru[i * it.ldr + t];
      VA(t, i) = Rv[i * it.ldr + t];
    }
    __syncthreads();
    constexpr int PER = (KMAX * KMAX + NT - 1) / NT;
    T mv[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + u * NT;
      T sacc = T(0);
      if (e < K * K) {
        const int i = e / K, j = e % K;
        for (int t = 0; t < J; ++t) sacc += G[t][i] * VA(t, j);
        if (i < it.ka && j < it.ka) sacc += static_cast<const T*>(it.Sa)[i * it.lds_in + j];
      }
      mv[u] = sacc;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = tid + u * NT;
      if (e < K * K) {
        const int i = e / K, j = e % K;
        G[j][i] = mv[u];
        VA(j, i) = (i == j) ? T(1) : T(0);
      }
    }
  }
  __syncthreads();
  // Scale by a power of two so that max |M_ij| lies in [1, 2): exact, and it
  // keeps squared column norms clear of underflow in the rotation tests.
  using UB = typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned int>::type;
  __shared__ UB amax_bits;
  __shared__ T tiny_sh;
  if (tid == 0) amax_bits = 0;
  __syncthreads();
  {
    T mx = T(0);
    for (int e = tid; e < K * K; e += NT) mx = fmax(mx, fabs(G[e / K][e % K]));
    UB bits;
    memcpy(&bits, &mx, sizeof(T));
    atomicMax(&amax_bits, bits);
  }
  __syncthreads();
  T amax;
  {
    const UB bits = amax_bits;
    memcpy(&amax, &bits, sizeof(T));
  }
  int sexp = 0;
  if (amax > T(0) && isfinite(amax)) sexp = ilogb(amax);
  const T scale = ldexp(T(1), -sexp), unscale = ldexp(T(1), sexp);
  for (int e = tid; e < K * K; e += NT) G[e / K][e % K] *= scale;
  __syncthreads();
  if constexpr (VG) {
    // the rotations' vectors do not fit in LDS: the sweeps rotate only G, and
    // the kept left vectors are formed afterwards as M P g_c / sigma_c from this
    // copy of the (scaled) matrix, column j at vg + j K
    for (int e = tid; e < K * K; e += NT) vg[(int64_t)(e / K) * K + e % K] = G[e / K][e % K];
    __syncthreads();
  }
  if constexpr (PRE) {
    for (int c = tid; c < K; c += NT) perm[c] = c;
    __syncthreads();
    // gq lanes of one wave per column (rows split between them)
    int gq = 64;
    while (gq > 1 && K * gq > NT) gq >>= 1;
    const int qc = tid / gq, qg = tid % gq;
    for (int j = 0; j < K; ++j) {
      // pivot: the column of largest norm over rows j..K-1
      for (int c0 = j; c0 < K; c0 += NT / gq) {
        const int c = c0 + qc;
        T q = T(0);
        if (c < K)
          for (int row = j + qg; row < K; row += gq) q += G[c][row] * G[c][row];
        q = group_sum(q, gq);
        if (c < K && qg == 0) cnorm[c] = q;
      }
      __syncthreads();
      if (tid == 0) {
        int pv = j;
        for (int c = j + 1; c < K; ++c)
          if (cnorm[c] > cnorm[pv]) pv = c;
        piv_sh = pv;
        if (pv != j) {
          const int t0 = perm[j];
          perm[j] = perm[pv];
          perm[pv] = t0;
        }
      }
      __syncthreads();
      const int pv = piv_sh;
      if (pv != j)
        for (int row = tid; row < K; row += NT) {
          const T a = G[j][row];
          G[j][row] = G[pv][row];
          G[pv][row] = a;
        }
      __syncthreads();
      // Householder reflector of column j below its diagonal (LAPACK dlarfg)
      if (tid < 64) {
        T q = T(0);
        for (int row = j + 1 + tid; row < K; row += 64) q += G[j][row] * G[j][row];
        q = wave_sum(q);
        if (tid == 0) {
          const T alpha = G[j][j];
          T tau = T(0), beta = alpha, inv = T(0);
          if (q > T(0)) {
            beta = (alpha >= T(0) ? T(-1) : T(1)) * sqrt(alpha * alpha + q);
            tau = (beta - alpha) / beta;
            inv = T(1) / (alpha - beta);
          }
          tauq[j] = tau;
          cnorm[j] = inv;  // the scaling of v below the diagonal
          G[j][j] = beta;
        }
      }
      __syncthreads();
      const T inv = cnorm[j], tau = tauq[j];
      for (int row = j + 1 + tid; row < K; row += NT) G[j][row] *= inv;
      __syncthreads();
      // apply H_j to the trailing columns: x -= tau v (v^T x), v = [1; G[j][j+1..]]
      if (tau != T(0))
        for (int c0 = j + 1; c0 < K; c0 += NT / gq) {
          const int c = c0 + qc;
          T w = T(0);
          if (c < K) {
            if (qg == 0) w = G[c][j];
            for (int row = j + 1 + qg; row < K; row += gq) w += G[j][row] * G[c][row];
          }
          w = group_sum(w, gq) * tau;
          if (c < K) {
            if (qg == 0) G[c][j] -= w;
            for (int row = j + 1 + qg; row < K; row += gq) G[c][row] -= w * G[j][row];
          }
        }
      __syncthreads();
    }
    if constexpr (!VG) {
      // Q = H_0 H_1 ... H_{K-1} formed in the right-vector array (identity so
      // far), from the reflectors kept below G's diagonal, last one first: the
      // rotations then carry Q along, so the array ends as Q V_x, the left
      // vectors, with no back-transform (and no LDS for the reflectors)
      for (int j = K - 1; j >= 0; --j) {
        const T tau = tauq[j];
        if (tau != T(0))
          for (int c0 = j; c0 < K; c0 += NT / gq) {
            const int c = c0 + qc;
            T w = T(0);
            if (c < K) {
              if (qg == 0) w = VA(c, j);
              for (int row = j + 1 + qg; row < K; row += gq) w += G[j][row] * VA(c, row);
            }
            w = group_sum(w, gq) * tau;
            if (c < K) {
              if (qg == 0) VA(c, j) -= w;
              for (int row = j + 1 + qg; row < K; row += gq) VA(c, row) -= w * G[j][row];
            }
          }
        __syncthreads();
      }
    }
    // G = R^T (in place): column j of G becomes row j of R, and the reflectors
    // below R's diagonal give way to zeros
    for (int e = tid; e < K * K; e += NT) {
      const int c = e / K, row = e % K;
      if (row > c) {
        G[c][row] = G[row][c];
        G[row][c] = T(0);
      }
    }
    __syncthreads();
  }
  for (int j = tid; j < K; j += NT) {
    T c2 = T(0);
    for (int row = 0; row < K; ++row) c2 += G[j][row] * G[j][row];
    sig[j] = c2;
  }
  __syncthreads();
  if (tid == 0) {
    T fro2 = T(0);
    for (int j = 0; j < K; ++j) fro2 += sig[j];
    // columns below u ||M||_F (u the unit roundoff) are numerically zero: rotating
    // them only churns rounding noise.  The floor also stays a factor 16 under the
    // truncation threshold eps ||sigma|| / sqrt(K) of an eps-rank decision, so in
    // FP32 at eps near 1e-6 the columns the rank is decided on are still rotated.
    T fl = Eps<T>::v;
    if (it.fixed_rank < 0 && it.eps > 0.0) fl = fmin(fl, static_cast<T>(it.eps / (16.0 * sqrt((double)K))));
    tiny_sh = fl * fl * fro2;
  }
  __syncthreads();
  const T tiny = tiny_sh;
  const int K2 = K + (K & 1);
  const int P = K2 / 2;
  int gl = 64;
  while (gl > 1 && P * gl > NT) gl >>= 1;
  const int p = tid / gl, sub = tid % gl;
  // rotation threshold: above the rounding noise of a length-K dot product
  const T tol = T(8 * max(K, 8)) * Eps<T>::v;
  const T itol = T(1) / tol;
  // The rounds run the round-robin schedule cyclically and stop once K2 - 1
  // consecutive rounds (every pair once, from any starting round) rotated
  // nothing: then no column changed while every pair was tested, as after a
  // rotation-free sweep.
  int rounds = 0;
  if (K >= 2) {
    int quiet = 0;
    for (int r = 0; quiet < K2 - 1 && rounds < 60 * (K2 - 1); r = (r + 1 == K2 - 1) ? 0 : r + 1) {
      ++rounds;
      int rotated = 0;
      if (p < P) {
        int i = rr_pos(p, r, K2), j = rr_pos(K2 - 1 - p, r, K2);
        if (i > j) {
          const int tmp = i;
          i = j;
          j = tmp;
        }
        if (j < K) {
          T al = T(0), be = T(0), ga = T(0);
          for (int row = sub; row < K; row += gl) {
            const T x = G[i][row], y = G[j][row];
            al += x * x;
            be += y * y;
            ga += x * y;
          }
          al = group_sum(al, gl);
          be = group_sum(be, gl);
          ga = group_sum(ga, gl);
          const T gt = ga * itol;  // |ga| > tol sqrt(al be), squared with no underflow (al, be > tiny)
          if (al > tiny && be > tiny && gt * gt > al * be) {
            T cth, sth;
            jacobi_rotation(al, be, ga, cth, sth);
            for (int row = sub; row < K; row += gl) {
              const T x = G[i][row], y = G[j][row];
              G[i][row] = cth * x - sth * y;
              G[j][row] = sth * x + cth * y;
              if constexpr (!VG) {
                const T vx = VA(i, row), vy = VA(j, row);
                VA(i, row) = cth * vx - sth * vy;
                VA(j, row) = sth * vx + cth * vy;
              }
            }
            rotated = 1;
          }
        }
      }
      quiet = __syncthreads_or(rotated) ? 0 : quiet + 1;
    }
  }
  const int sweeps = (rounds + K2 - 2) / max(1, K2 - 1);
  if (tid == 0 && it.stats != nullptr) {
    // this workgroup's copy of the counters (a line of its own: no contention
    // between the items of a launch)
    int32_t* st = it.stats + (blockIdx.x % kCounterCopies) * kCounterStride;
    atomicAdd(st, 1);
    atomicAdd(st + 1, sweeps);
    atomicMax(st + 2, sweeps);
  }
  for (int j = tid; j < K; j += NT) {
    T s = T(0);
    for (int row = 0; row < K; ++row) s += G[j][row] * G[j][row];
    sig[j] = sqrt(s);
  }
  __syncthreads();
  for (int j = tid; j < K; j += NT) {
    int rk = 0;
    const T sj = sig[j];
    for (int i = 0; i < K; ++i)
      if (sig[i] > sj || (sig[i] == sj && i < j)) ++rk;
    order[rk] = j;
  }
  __syncthreads();
  if (tid == 0) {
    int r;
    if (it.fixed_rank >= 0) {
      r = min(it.fixed_rank, K);
    } else {
      const int md = (it.kdim > 0) ? min(it.kdim, K) : K;
      T ss = T(0);
      for (int i = 0; i < md; ++i) {
        const T v = sig[order[i]];
        ss += v * v;
      }
      const T threshold = static_cast<T>(it.eps) * sqrt(ss);
      r = 0;
      T err;
      do {
        ++r;
        err = T(0);
        if (r < md) {
          const T sr = sig[order[r]];
          for (int jj = r; jj < md; ++jj) err += sr * sr;
        }
      } while (r < md && sqrt(err) > threshold);
    }
    if (r > it.cap) {
      if (it.rank_out) *it.rank_out = it.cap_trunc ? it.cap : -r;
      r = it.cap;
    } else if (it.rank_out) {
      *it.rank_out = r;
    }
    rank_sh = r;
  }
  __syncthreads();
  const int r = rank_sh;
  // normalise left vectors in place
  for (int e = tid; e < K * K; e += NT) {
    const int c = e / K, row = e % K;
    const T s = sig[c];
    if (s > T(0)) G[c][row] = G[c][row] / s;
  }
  __syncthreads();
  // Orthonormal completion of kept left vectors whose singular value is
  // numerically zero (below eps * ||M||_F: Jacobi leaves those columns
  // unrotated): two passes of classical Gram-Schmidt of e_u against the
  // earlier kept vectors, accepted when the remainder keeps norm^2 > thr.
  // Candidates are scanned from u = K - 1 down (the pivoted leading coordinates
  // carry the kept vectors) and the scan resumes where the previous column's
  // stopped: a rejected remainder only shrinks as the span grows, and an
  // accepted e_u lies in it.  The remainders' squared norms sum to K - c, so a
  // second scan with thr = 1/(4K) always accepts one.
  const T sig_floor = sqrt(tiny);
  int u = K - 1;
  T thr = T(0.25);
  for (int c = 0; c < r; ++c) {
    const int src = order[c];
    if (sig[src] > sig_floor) continue;
    for (int tries = 0; tries < 2 * K; ++tries, --u) {
      if (u < 0) {
        u = K - 1;
        thr = T(0.25) / T(K);
      }
      for (int row = tid; row < K; row += NT) G[src][row] = (row == u) ? T(1) : T(0);
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass) {
        for (int pp = tid; pp < c; pp += NT) {
          const int ps = order[pp];
          T d = T(0);
          for (int row = 0; row < K; ++row) d += G[ps][row] * G[src][row];
          coef[pp] = d;
        }
        __syncthreads();
        for (int row = tid; row < K; row += NT) {
          T x = G[src][row];
          for (int pp = 0; pp < c; ++pp) x -= coef[pp] * G[order[pp]][row];
          G[src][row] = x;
        }
        __syncthreads();
      }
      if (tid == 0) {
        T nn = T(0);
        for (int row = 0; row < K; ++row) nn += G[src][row] * G[src][row];
        accept_sh = nn > thr;
        coef[0] = (nn > T(0)) ? T(1) / sqrt(nn) : T(0);
      }
      __syncthreads();
      const int ok = accept_sh;
      const T inv = coef[0];
      __syncthreads();
      if (ok) {
        for (int row = tid; row < K; row += NT) G[src][row] *= inv;
        __syncthreads();
        --u;
        break;
      }
    }
  }
  T* S = static_cast<T*>(it.S);
  T* Um = static_cast<T*>(it.Um);
  T* Vm = static_cast<T*>(it.Vm);
  if (S != nullptr)
    for (int e = tid; e < r * r; e += NT) {
      const int i = e / r, j = e % r;
      S[i * it.lds + j] = (i == j) ? sig[order[i]] * unscale : T(0);
    }
  if constexpr (PRE) {
    // M = (Q V_x) Sigma (P U_x)^T: right vectors from G (rows permuted back),
    // left vectors Q V_x (the rotated Q), or, in VG mode, M v / sigma
    for (int e = tid; e < K * r; e += NT) {
      const int row = e / r, c = e % r;
      if (Vm != nullptr) Vm[perm[row] * it.ldv + c] = G[order[c]][row];
    }
    if constexpr (!VG) {
      for (int e = tid; e < K * r; e += NT) {
        const int row = e / r, c = e % r;
        if (Um != nullptr) Um[row * it.ldu + c] = VA(order[c], row);
      }
    } else if (Um != nullptr) {
      // u_c = M P g_c / sigma_c: column perm[j] of M against entry j of g_c, a wave per output
      const int lane = tid & 63, wv = tid >> 6;
      for (int o = wv; o < K * r; o += NT / 64) {
        const int row = o / r, c = o % r;
        const int src = order[c];
        const T sc = sig[src];
        T acc = T(0);
        for (int j2 = lane; j2 < K; j2 += 64) acc += vg[(int64_t)perm[j2] * K + row] * G[src][j2];
        acc = wave_sum(acc);
        if (lane == 0) Um[row * it.ldu + c] = (sc > sig_floor) ? acc / sc : T(0);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Register-resident one-sided Jacobi SVD for K <= KMAX (16 or 32): one wave per
// core and no LDS traffic inside the rounds.  The round-robin schedule has K2
// positions (K rounded up to even; a zero column pads odd K) in NS slots of two:
// the top of slot s is position s, its bottom position K2-1-s.  Lane l serves
// slot s = l % NS together with the other NL lanes of that slot, holding rows
// g + NL q (g = l / NS) of the top column in x[], of the bottom column in y[] and
// of their right vectors in vx[], vy[].  A round rotates every slot's pair at
// once; the columns then move one place round the circle (top 0 fixed, top k-1
// -> top k, bottom 0 -> top 1, bottom k+1 -> bottom k, top P-1 -> bottom P-1).
// Neighbouring slots are neighbouring lanes, so a move is one DPP wave shift per
// register; the column sums are a DPP / swizzle / permlane butterfly.
// ---------------------------------------------------------------------------
// (lane 0, resp. 63, receives zero: bound_ctrl writes every lane, so no old value is needed)
template <int CTRL>
__device__ __forceinline__ int dpp_bound_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, true);
}
__device__ __forceinline__ float wave_shr1(float v) {  // lane l receives lane l - 1
  return __int_as_float(dpp_bound_i<0x138>(__float_as_int(v)));
}
__device__ __forceinline__ double wave_shr1(double v) {
  return __hiloint2double(dpp_bound_i<0x138>(__double2hiint(v)), dpp_bound_i<0x138>(__double2loint(v)));
}
__device__ __forceinline__ float wave_shl1(float v) {  // lane l receives lane l + 1
  return __int_as_float(dpp_bound_i<0x130>(__float_as_int(v)));
}
__device__ __forceinline__ double wave_shl1(double v) {
  return __hiloint2double(dpp_bound_i<0x130>(__double2hiint(v)), dpp_bound_i<0x130>(__double2loint(v)));
}
__device__ __forceinline__ float half_max(float v) {
  const auto r = __builtin_amdgcn_permlane32_swap(__float_as_int(v), __float_as_int(v), false, false);
  return fmaxf(__int_as_float(r[0]), __int_as_float(r[1]));
}
__device__ __forceinline__ double half_max(double v) {
  const auto h = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
  const auto l = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
  return fmax(__hiloint2double(h[0], l[0]), __hiloint2double(h[1], l[1]));
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  v = fmax(v, dpp_mov<0xB1>(v));
  v = fmax(v, dpp_mov<0x4E>(v));
  v = fmax(v, dpp_mov<0x141>(v));
  v = fmax(v, dpp_mov<0x140>(v));
  v = fmax(v, xor16(v));
  return half_max(v);
}
// v[l] + v[l ^ 16] through v_permlane16_swap (gfx950), operands in the same order in every lane
__device__ __forceinline__ float pair16_sum(float v) {
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_int(v), __float_as_int(v), false, false);
  return __int_as_float(r[0]) + __int_as_float(r[1]);
}
__device__ __forceinline__ double pair16_sum(double v) {
  const auto h = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
  const auto l = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
  return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}
// sum over the NL lanes of a slot (lanes s + NS t), the same bits in each of them
template <int NL, typename T>
__device__ __forceinline__ T slot_sum(T v) {
  static_assert(NL == 4 || NL == 8, "slot of 4 or 8 lanes");
  if constexpr (NL == 8) v += dpp_mov<0x128>(v);  // row_ror:8 = lane l ^ 8
  return half_sum(pair16_sum(v));
}

template <typename T, int KMAX, class Src>
__global__ __launch_bounds__(64) void svd_reg_kernel(Src items) {
  constexpr int NS = KMAX / 2, NL = 64 / NS, NR = KMAX / NL;
  const SvdItem it = items.get(blockIdx.x);
  const int K = it.K;
  const int lane = threadIdx.x;
  const int s = lane % NS, g = lane / NS;
  // staging (Ru, or M in mode 0; Rv); the left vectors on the completion path
  __shared__ T A1[KMAX][KMAX + 1];
  __shared__ T A2[KMAX][KMAX + 1];
  __shared__ T sig[KMAX];   // by position
  __shared__ T sigs[KMAX];  // descending
  __shared__ int order[KMAX];
  __shared__ int rk_of[KMAX];
  __shared__ T coef[KMAX];
  if (K <= 0) {
    if (lane == 0 && it.rank_out) *it.rank_out = 0;
    return;
  }
  const int K2 = K + (K & 1), P = K2 / 2;
  const int pt = s, pb = K2 - 1 - s;
  const bool act = s < P;
  if (it.mode == 0) {
    const T* M = static_cast<const T*>(it.M);
#pragma unroll
    for (int k = 0; k < KMAX * KMAX / 64; ++k) {
      const int e = lane + 64 * k;
      const int i = e / KMAX, j = e % KMAX;
      A1[i][j] = (i < K && j < K) ? M[i * it.m_rs + j * it.m_cs] : T(0);
    }
  } else {
    const int J = it.kx;
    const T* Ru = static_cast<const T*>(it.Ru);
    const T* Rv = static_cast<const T*>(it.Rv);
#pragma unroll
    for (int k = 0; k < KMAX * KMAX / 64; ++k) {
      const int e = lane + 64 * k;
      const int i = e / KMAX, t = e % KMAX;
      const bool in = i < K && t < J;
      A1[i][t] = in ? Ru[i * it.ldr + t] : T(0);
      A2[i][t] = in ? Rv[i * it.ldr + t] : T(0);
    }
  }
  __syncthreads();
  T x[NR], y[NR], vx[NR], vy[NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int ri = g + NL * q;
    x[q] = T(0);
    y[q] = T(0);
    vx[q] = (act && ri == pt && ri < K) ? T(1) : T(0);
    vy[q] = (act && ri == pb && ri < K) ? T(1) : T(0);
  }
  if (act) {
    if (it.mode == 0) {
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        x[q] = A1[g + NL * q][pt];
        y[q] = A1[g + NL * q][pb];
      }
    } else {
      // M(i, j) = sum_t Ru(i, t) Rv(j, t) (+ Sa(i, j) inside the leading ka x ka block)
      const int J = it.kx;
      for (int t = 0; t < J; ++t) {
        const T a = A2[pt][t], b = A2[pb][t];
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const T u = A1[g + NL * q][t];
          x[q] += u * a;
          y[q] += u * b;
        }
      }
      if (it.ka > 0) {
        const T* Sa = static_cast<const T*>(it.Sa);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const int ri = g + NL * q;
          if (ri < it.ka) {
            if (pt < it.ka) x[q] += Sa[ri * it.lds_in + pt];
            if (pb < it.ka) y[q] += Sa[ri * it.lds_in + pb];
          }
        }
      }
    }
  }
  // exact power-of-two scaling: max |M_ij| in [1, 2)
  T mx = T(0);
#pragma unroll
  for (int q = 0; q < NR; ++q) mx = fmax(mx, fmax(fabs(x[q]), fabs(y[q])));
  const T amax = wave_max(mx);
  int sexp = 0;
  if (amax > T(0) && isfinite(amax)) sexp = ilogb(amax);
  const T scale = ldexp(T(1), -sexp), unscale = ldexp(T(1), sexp);
  T f2 = T(0);
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    x[q] *= scale;
    y[q] *= scale;
    f2 += x[q] * x[q] + y[q] * y[q];
  }
  const T fro2 = wave_sum(f2);
  // numerically zero columns (below u ||M||_F, and 16x under an eps-rank
  // threshold) are not rotated
  T fl = Eps<T>::v;
  if (it.fixed_rank < 0 && it.eps > 0.0) fl = fmin(fl, static_cast<T>(it.eps / (16.0 * sqrt((double)K))));
  const T tiny = fl * fl * fro2;
  const T tol = T(8 * max(K, 8)) * Eps<T>::v;
  const T itol = T(1) / tol;
  // cyclic rounds until K2 - 1 consecutive rounds rotate nothing
  int rounds = 0;
  if (K >= 2) {
    int quiet = 0;
    while (quiet < K2 - 1 && rounds < 60 * (K2 - 1)) {
      ++rounds;
      T al = T(0), be = T(0), ga = T(0);
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        al += x[q] * x[q];
        be += y[q] * y[q];
        ga += x[q] * y[q];
      }
      al = slot_sum<NL>(al);
      be = slot_sum<NL>(be);
      ga = slot_sum<NL>(ga);
      const T gt = ga * itol;
      const bool rot = act && al > tiny && be > tiny && gt * gt > al * be;
      if (rot) {
        T cth, sth;
        jacobi_rotation(al, be, ga, cth, sth);
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const T a = x[q], b = y[q];
          x[q] = cth * a - sth * b;
          y[q] = sth * a + cth * b;
          const T va = vx[q], vb = vy[q];
          vx[q] = cth * va - sth * vb;
          vy[q] = sth * va + cth * vb;
        }
      }
      quiet = __any(rot) ? 0 : quiet + 1;
      if (P > 1) {
#pragma unroll
        for (int q = 0; q < NR; ++q) {
          const T nt = wave_shr1((s == 0) ? y[q] : x[q]);
          const T ntv = wave_shr1((s == 0) ? vy[q] : vx[q]);
          const T nb = wave_shl1(y[q]);
          const T nbv = wave_shl1(vy[q]);
          if (s == P - 1) {
            y[q] = x[q];
            vy[q] = vx[q];
          } else {
            y[q] = nb;
            vy[q] = nbv;
          }
          if (s != 0) {
            x[q] = nt;
            vx[q] = ntv;
          }
        }
      }
    }
  }
  const int sweeps = (rounds + K2 - 2) / max(1, K2 - 1);
  if (lane == 0 && it.stats != nullptr) {
    int32_t* st = it.stats + (blockIdx.x % kCounterCopies) * kCounterStride;
    atomicAdd(st, 1);
    atomicAdd(st + 1, sweeps);
    atomicMax(st + 2, sweeps);
  }
  T nx = T(0), ny = T(0);
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    nx += x[q] * x[q];
    ny += y[q] * y[q];
  }
  nx = sqrt(slot_sum<NL>(nx));
  ny = sqrt(slot_sum<NL>(ny));
  // the padding column of odd K started at position K2 - 1 and moved one place
  // per round along the circle of positions 1 .. K2 - 1
  const int padpos = (K & 1) ? ((K2 - 2 + rounds) % (K2 - 1)) + 1 : -1;
  if (g == 0 && act) {
    sig[pt] = (pt == padpos) ? T(-1) : nx;
    sig[pb] = (pb == padpos) ? T(-1) : ny;
  }
  __syncthreads();
  if (lane < K2) {
    const T sj = sig[lane];
    int rk = 0;
    for (int i = 0; i < K2; ++i) {
      const T si = sig[i];
      rk += (si > sj || (si == sj && i < lane)) ? 1 : 0;
    }
    order[rk] = lane;
    sigs[rk] = sj;
    rk_of[lane] = rk;
  }
  __syncthreads();
  int r;
  if (it.fixed_rank >= 0) {
    r = min(it.fixed_rank, K);
  } else {
    // src/operations/misc/misc.cpp:133-146: the smallest r >= 1 with r = md or
    // sqrt((md - r) sigma_r^2) <= eps ||sigma||, the error summed term by term
    const int md = (it.kdim > 0) ? min(it.kdim, K) : K;
    T ss = T(0);
    for (int i = 0; i < md; ++i) ss += sigs[i] * sigs[i];
    const T threshold = static_cast<T>(it.eps) * sqrt(ss);
    bool stop = lane >= 1;
    if (lane >= 1 && lane < md) {
      const T sr = sigs[lane];
      T err = T(0);
      for (int jj = lane; jj < md; ++jj) err += sr * sr;
      stop = !(sqrt(err) > threshold);
    }
    r = __ffsll((unsigned long long)__ballot(stop)) - 1;
  }
  int rout = r;
  if (r > it.cap) {
    rout = it.cap_trunc ? it.cap : -r;
    r = it.cap;
  }
  if (lane == 0 && it.rank_out) *it.rank_out = rout;
  if (nx > T(0)) {
#pragma unroll
    for (int q = 0; q < NR; ++q) x[q] /= nx;
  }
  if (ny > T(0)) {
#pragma unroll
    for (int q = 0; q < NR; ++q) y[q] /= ny;
  }
  // Orthonormal completion of kept left vectors whose singular value is
  // numerically zero (Jacobi leaves those columns unrotated): two passes of
  // classical Gram-Schmidt of e_u against the earlier kept vectors, accepted
  // when the remainder keeps norm^2 > thr, with the resumed descending scan of
  // svd_kernel (a rejected remainder only shrinks as the span grows; a second
  // scan with thr = 1/(4K) always accepts one).  Rare: it runs on an LDS copy.
  const T sig_floor = sqrt(tiny);
  bool need = false;
  for (int c = 0; c < r; ++c) need = need || !(sigs[c] > sig_floor);
  if (need) {
    if (act) {
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        const int ri = g + NL * q;
        A1[pt][ri] = x[q];
        A1[pb][ri] = y[q];
      }
    }
    __syncthreads();
    __shared__ int accept_sh;
    int u = K - 1;
    T thr = T(0.25);
    for (int c = 0; c < r; ++c) {
      const int src = order[c];
      if (sigs[c] > sig_floor) continue;
      for (int tries = 0; tries < 2 * K; ++tries, --u) {
        if (u < 0) {
          u = K - 1;
          thr = T(0.25) / T(K);
        }
        for (int row = lane; row < K; row += 64) A1[src][row] = (row == u) ? T(1) : T(0);
        __syncthreads();
        for (int pass = 0; pass < 2; ++pass) {
          for (int pp = lane; pp < c; pp += 64) {
            const int ps = order[pp];
            T d = T(0);
            for (int row = 0; row < K; ++row) d += A1[ps][row] * A1[src][row];
            coef[pp] = d;
          }
          __syncthreads();
          for (int row = lane; row < K; row += 64) {
            T v = A1[src][row];
            for (int pp = 0; pp < c; ++pp) v -= coef[pp] * A1[order[pp]][row];
            A1[src][row] = v;
          }
          __syncthreads();
        }
        if (lane == 0) {
          T nn = T(0);
          for (int row = 0; row < K; ++row) nn += A1[src][row] * A1[src][row];
          accept_sh = nn > thr;
          coef[0] = (nn > T(0)) ? T(1) / sqrt(nn) : T(0);
        }
        __syncthreads();
        const int ok = accept_sh;
        const T inv = coef[0];
        __syncthreads();
        if (ok) {
          for (int row = lane; row < K; row += 64) A1[src][row] *= inv;
          __syncthreads();
          --u;
          break;
        }
      }
    }
    if (act) {
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        const int ri = g + NL * q;
        x[q] = A1[pt][ri];
        y[q] = A1[pb][ri];
      }
    }
  }
  T* S = static_cast<T*>(it.S);
  T* Um = static_cast<T*>(it.Um);
  T* Vm = static_cast<T*>(it.Vm);
  if (S != nullptr) {
#pragma unroll
    for (int k = 0; k < KMAX * KMAX / 64; ++k) {
      const int e = lane + 64 * k;
      const int i = e / KMAX, j = e % KMAX;
      if (i < r && j < r) S[i * it.lds + j] = (i == j) ? sigs[i] * unscale : T(0);
    }
  }
  const int ct = act ? rk_of[pt] : KMAX, cb = act ? rk_of[pb] : KMAX;
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int ri = g + NL * q;
    if (ri < K) {
      if (ct < r) {
        if (Um != nullptr) Um[ri * it.ldu + ct] = x[q];
        if (Vm != nullptr) Vm[ri * it.ldv + ct] = vx[q];
      }
      if (cb < r) {
        if (Um != nullptr) Um[ri * it.ldu + cb] = y[q];
        if (Vm != nullptr) Vm[ri * it.ldv + cb] = vy[q];
      }
    }
  }
}

template <typename T, class Src>
static void launch_svd_src(Src d_items, int n_items, int max_k, hipStream_t s) {
  if (n_items <= 0) return;
  if (max_k <= 16) {
    hipLaunchKernelGGL((svd_reg_kernel<T, 16, Src>), dim3(n_items), dim3(64), 0, s, d_items);
  } else if (max_k <= 32 && sizeof(T) == 8) {
    hipLaunchKernelGGL((svd_reg_kernel<T, 32, Src>), dim3(n_items), dim3(64), 0, s, d_items);
  } else if (max_k <= 32) {
    hipLaunchKernelGGL((svd_kernel<T, 32, 64, false, Src>), dim3(n_items), dim3(64), 0, s, d_items);
  } else if (max_k <= 64) {
    hipLaunchKernelGGL((svd_kernel<T, 64, 256, false, Src>), dim3(n_items), dim3(256), 0, s, d_items);
  } else if (max_k <= 128) {
    if constexpr (sizeof(T) == 4) {
      hipLaunchKernelGGL((svd_kernel<T, 128, 1024, false, Src>), dim3(n_items), dim3(1024), 0, s, d_items);
    } else {
      // every item must carry a K x K scratch (SvdItem::vwork) for its right vectors
      hipLaunchKernelGGL((svd_kernel<T, 128, 1024, true, Src>), dim3(n_items), dim3(1024), 0, s, d_items);
    }
  } else {
    throw std::runtime_error("svd: K above 128");
  }
  FRK_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_svd(const SvdItem* d_items, int n_items, int max_k, hipStream_t s) {
  launch_svd_src<T>(ItemArray<SvdItem>{d_items}, n_items, max_k, s);
}

template <typename T>
void launch_lr_add_qr(const LrAddDesc<T>* d, int nd, const LrBatch<T>& b, int max_k, hipStream_t s) {
  launch_qr_src<T>(LrQrItems<T>{d, b}, 2 * nd, max_k, s, std::max(b.mmax, b.nmax));
}
// M = Ru Rv^T (Ke x Ke, row stride K) of the additions whose core is smaller
// than K = kc + kx2 (a workgroup each; the others return at once): Ru and Rv
// staged through LDS in chunks of 32 of the K columns.
template <typename T>
__global__ __launch_bounds__(256) void lr_core_prod_kernel(const LrAddDesc<T>* __restrict__ d, LrBatch<T> b) {
  const int t = blockIdx.x;
  const LrAddDesc<T> x = d[t];
  const int K = b.K, Kt = x.kc + x.kx2, Ke = lr_core_dim(x, b);
  if (Ke >= Kt) return;
  constexpr int QC = 32, NO = 128 * 128 / 256;
  __shared__ T Us[128][QC + 1];
  __shared__ T Vs[128][QC + 1];
  const T* Ru = b.Ru + (size_t)t * K * K;
  const T* Rv = b.Rv + (size_t)t * K * K;
  T* M = b.Mw + (size_t)t * K * K;
  const int tid = threadIdx.x;
  T acc[NO];
#pragma unroll
  for (int k = 0; k < NO; ++k) acc[k] = T(0);
  for (int q0 = 0; q0 < Kt; q0 += QC) {
    __syncthreads();
    for (int e = tid; e < Ke * QC; e += 256) {
      const int i = e / QC, q = e % QC;
      const bool in = q0 + q < Kt;
      Us[i][q] = in ? Ru[(int64_t)i * K + q0 + q] : T(0);
      Vs[i][q] = in ? Rv[(int64_t)i * K + q0 + q] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NO; ++k) {
      const int e = tid + 256 * k;
      if (e < Ke * Ke) {
        const int i = e / Ke, j = e - (e / Ke) * Ke;
        T a = acc[k];
#pragma unroll 8
        for (int q = 0; q < QC; ++q) a += Us[i][q] * Vs[j][q];
        acc[k] = a;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    const int e = tid + 256 * k;
    if (e < Ke * Ke) M[(int64_t)(e / Ke) * K + e % Ke] = acc[k];
  }
}

template <typename T>
void launch_lr_add_core(const LrAddDesc<T>* d, int nd, const LrBatch<T>& b, hipStream_t s) {
  if (nd <= 0 || b.Mw == nullptr) return;
  if (b.K > 128) throw std::runtime_error("lr core product: K above 128");
  hipLaunchKernelGGL(lr_core_prod_kernel<T>, dim3(nd), dim3(256), 0, s, d, b);
  FRK_HIP_CHECK(hipGetLastError());
}

template <typename T>
void launch_lr_add_svd(const LrAddDesc<T>* d, int t0, int n, int max_k, const LrBatch<T>& b, hipStream_t s) {
  launch_svd_src<T>(LrSvdItems<T>{d, b, t0}, n, max_k, s);
}
template <typename T>
void launch_lr_add_apply(const LrAddDesc<T>* d, int nd, const LrBatch<T>& b, int max_ncol, hipStream_t s) {
  launch_apply_q_src<T>(LrApplyItems<T>{d, b}, 2 * nd, max_ncol, s, std::max(b.mmax, b.nmax));
}

// ===========================================================================
// Truncated column-pivoted Householder QR (geqp3 with LAWN-176 downdating),
// one 512-thread workgroup per row-major tile; thread-per-column sweeps.
// ===========================================================================
template <typename T>
__device__ T block_sum512(T v, T* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = 256; s > 0; s >>= 1) {
    if (tid < s) red[tid] = red[tid] + red[tid + s];
    __syncthreads();
  }
  const T out = red[0];
  __syncthreads();
  return out;
}

template <typename T>
__global__ __launch_bounds__(512) void geqp3_kernel(const Geqp3Item* __restrict__ items) {
  const Geqp3Item it = items[blockIdx.x];
  T* A = static_cast<T*>(it.A);
  const int m = it.m, n = it.n;
  const int64_t lda = it.lda;
  const int tid = threadIdx.x;
  __shared__ T cn[512], pcn[512], vv[512], red[512];
  __shared__ int piv[512], redi[512];
  const T tol = Eps<T>::half;
  const T tol3z = sqrt(tol);
  const int min_dim = min(m, n);
  T part = T(0);
  for (int c = tid; c < n; c += 512) {
    T s = T(0);
    for (int i = 0; i < m; ++i) {
      const T x = A[i * lda + c];
      s += x * x;
    }
    part += s;
    cn[c] = sqrt(s);
    pcn[c] = cn[c];
    piv[c] = c;
  }
  const T total = block_sum512(part, red);
  const T threshold = static_cast<T>(it.eps) * sqrt(total);
  T* tau_out = static_cast<T*>(it.tau);
  // initial max
  auto block_argmax = [&](int from, T& best, int& besti) {
    T bv = -T(1);
    int bi = 0x7fffffff;
    for (int c = from + tid; c < n; c += 512) {
      const T x = cn[c];
      if (x > bv || (x == bv && c < bi)) {
        bv = x;
        bi = c;
      }
    }
    red[tid] = bv;
    redi[tid] = bi;
    __syncthreads();
    for (int s = 256; s > 0; s >>= 1) {
      if (tid < s) {
        const T o = red[tid + s];
        const int oi = redi[tid + s];
        if (o > red[tid] || (o == red[tid] && oi < redi[tid])) {
          red[tid] = o;
          redi[tid] = oi;
        }
      }
      __syncthreads();
    }
    best = red[0];
    besti = redi[0];
    __syncthreads();
  };
  T maxv;
  int kmax;
  block_argmax(0, maxv, kmax);
  if (maxv <= tol) {
    // zero matrix: Q = e1, R = 0 (rank 1)
    for (int c = tid; c < n; c += 512) {
      A[c] = T(0);
      it.piv[c] = c;
    }
    if (tid == 0) {
      tau_out[0] = T(0);
      *it.rank_out = 1;
    }
    return;
  }
  // eps < 0: exactly min(cap, min_dim) pivot steps (dgeqp3 prefix, used by the ID)
  const bool fixed = it.eps < 0.0;
  const int steps = (fixed || it.cap_trunc) ? min(it.cap, min_dim) : min_dim;
  int r = 0;
  while (r < steps && (fixed || maxv > threshold)) {
    // pivot: first maximum of cn[r:] (computed at the end of the previous step)
    const int k = kmax;
    if (k != r) {
      for (int i = tid; i < m; i += 512) {
        const T x = A[i * lda + r];
        A[i * lda + r] = A[i * lda + k];
        A[i * lda + k] = x;
      }
    }
    __syncthreads();
    if (tid == 0 && k != r) {
      T x = cn[r];
      cn[r] = cn[k];
      cn[k] = x;
      x = pcn[r];
      pcn[r] = pcn[k];
      pcn[k] = x;
      const int pi = piv[r];
      piv[r] = piv[k];
      piv[k] = pi;
    }
    // reflector from column r
    T x2p = T(0);
    for (int i = r + 1 + tid; i < m; i += 512) {
      const T x = A[i * lda + r];
      x2p += x * x;
    }
    const T x2 = (r < m - 1) ? block_sum512(x2p, red) : T(0);
    const T alpha = A[r * lda + r];
    T beta, tau, scal;
    if (x2 == T(0)) {
      tau = T(0);
      beta = alpha;
      scal = T(0);
    } else {
      const T nrm = sqrt(alpha * alpha + x2);
      beta = (alpha >= T(0)) ? -nrm : nrm;
      tau = (beta - alpha) / beta;
      scal = T(1) / (alpha - beta);
    }
    for (int i = r + 1 + tid; i < m; i += 512) {
      const T x = A[i * lda + r] * scal;
      A[i * lda + r] = x;
      vv[i] = x;
    }
    if (tid == 0) {
      vv[r] = T(1);
      tau_out[r] = tau;
    }
    __syncthreads();
    if (tid == 0) A[r * lda + r] = beta;
    if (r < min_dim - 1 && tau != T(0)) {
      for (int c = r + 1 + tid; c < n; c += 512) {
        T wsum = T(0);
        for (int i = r; i < m; ++i) wsum += vv[i] * A[i * lda + c];
        wsum *= tau;
        for (int i = r; i < m; ++i) A[i * lda + c] -= wsum * vv[i];
      }
    }
    __syncthreads();
    for (int c = r + 1 + tid; c < n; c += 512) {
      if (cn[c] != T(0)) {
        T temp = fabs(A[r * lda + c] / cn[c]);
        temp = fmax(T(0), (T(1) - temp) * (T(1) + temp));
        const T ratio = cn[c] / pcn[c];
        const T temp2 = temp * ratio * ratio;
        if (temp2 > tol3z) {
          cn[c] = cn[c] * sqrt(temp);
        } else if (r < m - 1) {
          T s = T(0);
          for (int i = r + 1; i < m; ++i) {
            const T x = A[i * lda + c];
            s += x * x;
          }
          cn[c] = sqrt(s);
          pcn[c] = cn[c];
        } else {
          cn[c] = T(0);
          pcn[c] = T(0);
        }
      }
    }
    __syncthreads();
    ++r;
    if (r < n) {
      block_argmax(r, maxv, kmax);
    } else {
      maxv = T(0);
    }
  }
  for (int c = tid; c < n; c += 512) it.piv[c] = piv[c];
  if (tid == 0) *it.rank_out = (!fixed && !it.cap_trunc && r > it.cap) ? -r : r;
}

template <typename T>
void launch_geqp3(const Geqp3Item* d_items, int n_items, hipStream_t s) {
  if (n_items <= 0) return;
  hipLaunchKernelGGL(geqp3_kernel<T>, dim3(n_items), dim3(512), 0, s, d_items);
  FRK_HIP_CHECK(hipGetLastError());
}

// R (first `rank` rows, upper part) scattered back to original column order.
template <typename T>
__global__ __launch_bounds__(256) void permute_r_kernel(const PermuteItem* __restrict__ items) {
  const PermuteItem it = items[blockIdx.y];
  const int rk = *it.rank;
  const int i = blockIdx.x;
  if (i >= rk) return;
  const T* A = static_cast<const T*>(it.A);
  T* out = static_cast<T*>(it.out);
  for (int j = threadIdx.x; j < it.n; j += 256) {
    const T v = (j >= i) ? A[(int64_t)i * it.lda + j] : T(0);
    out[(int64_t)i * it.ldo + it.piv[j]] = v;
  }
}

template <typename T>
void launch_permute_r(const PermuteItem* d_items, int n_items, int max_n, int max_rows, hipStream_t s) {
  if (n_items <= 0 || max_rows <= 0) return;
  (void)max_n;
  hipLaunchKernelGGL(permute_r_kernel<T>, dim3(max_rows, n_items), dim3(256), 0, s, d_items);
  FRK_HIP_CHECK(hipGetLastError());
}

template <typename T>
__global__ __launch_bounds__(256) void flip_r_kernel(const FlipItem* __restrict__ items) {
  const FlipItem it = items[blockIdx.x];
  const int r = *it.rank;
  const T* R = static_cast<const T*>(it.R);
  T* S = static_cast<T*>(it.S);
  for (int e = threadIdx.x; e < r * r; e += 256) {
    const int i = e / r, j = e % r;
    S[i * it.lds + j] = (i <= j) ? R[(r - 1 - j) * it.ldr + (r - 1 - i)] : T(0);
  }
}

template <typename T>
void launch_flip_r(const FlipItem* d_items, int n_items, int max_r, hipStream_t s) {
  if (n_items <= 0 || max_r <= 0) return;
  hipLaunchKernelGGL(flip_r_kernel<T>, dim3(n_items), dim3(256), 0, s, d_items);
  FRK_HIP_CHECK(hipGetLastError());
}

// ===========================================================================
// Kernel-matrix tiles (same operation order as the CPU generators: the
// squared distance is accumulated without contraction).
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(256) void gen_kernel(KernelSpec ks, const GenItem* __restrict__ items) {
#pragma clang fp contract(off)
  const GenItem it = items[blockIdx.y];
  const int64_t total = (int64_t)it.nr * it.nc;
  T* out = static_cast<T*>(it.out);
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int i = (int)(e / it.nc), j = (int)(e % it.nc);
    const int64_t gi = it.r0 + i, gj = it.c0 + j;
    double rij = 0.0;
    for (int d = 0; d < ks.dim; ++d) {
      const double dd = ks.pts[d * ks.n_pts + gi] - ks.pts[d * ks.n_pts + gj];
      rij = rij + dd * dd;
    }
    double v;
    if (ks.kind == 0) {
      v = 1.0 / (sqrt(rij) + 1e-3);
    } else if (ks.kind == 1) {
      v = exp(-rij / ((2.0 * ks.ell) * ks.ell));
    } else if (ks.kind == 2) {  // identity (src/functions.cpp:25-34)
      v = (gi == gj) ? 1.0 : 0.0;
    } else {  // d/d ell of the gaussian kernel
      v = exp(-rij / ((2.0 * ks.ell) * ks.ell)) * (rij / (ks.ell * ks.ell * ks.ell));
    }
    if (gi == gj) v = v + ks.shift;
    out[(int64_t)i * it.ld + j] = static_cast<T>(v);
  }
}

template <typename T>
void launch_gen(const KernelSpec& ks, const GenItem* d_items, int n_items, int max_nr, int max_nc,
                hipStream_t s) {
  if (n_items <= 0) return;
  const int64_t total = (int64_t)max_nr * max_nc;
  int gx = (int)std::min<int64_t>((total + 255) / 256, 1024);
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(gen_kernel<T>, dim3(gx, n_items), dim3(256), 0, s, ks, d_items);
  FRK_HIP_CHECK(hipGetLastError());
}

__global__ void set_int_kernel(const SetIntItem* __restrict__ items, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) *items[i].p = items[i].v;
}

void launch_set_int(const SetIntItem* d_items, int n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(set_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, d_items, n);
  FRK_HIP_CHECK(hipGetLastError());
}

template <typename T>
__global__ void fill_identity_kernel(T* p, int64_t rows, int64_t cols, int64_t ld) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int64_t i = e / cols, j = e % cols;
    p[i * ld + j] = (i == j) ? T(1) : T(0);
  }
}

template <typename T>
void launch_fill_identity(T* p, int64_t rows, int64_t cols, int64_t ld, hipStream_t s) {
  const int64_t total = rows * cols;
  if (total <= 0) return;
  int gx = (int)std::min<int64_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_identity_kernel<T>, dim3(gx), dim3(256), 0, s, p, rows, cols, ld);
  FRK_HIP_CHECK(hipGetLastError());
}

template <typename TD, typename TS>
__global__ void convert_kernel(TD* dst, const TS* src, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
    dst[e] = static_cast<TD>(src[e]);
}

template <typename TD, typename TS>
void launch_convert(TD* dst, const TS* src, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int gx = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL((convert_kernel<TD, TS>), dim3(gx), dim3(256), 0, s, dst, src, n);
  FRK_HIP_CHECK(hipGetLastError());
}

// ===========================================================================
// Matrix-free residual R = B - A X (FP64) for the kernel matrices.
// Row block of 256 per workgroup, columns split over gridDim.y partial sums
// reduced in a fixed order by a second kernel (deterministic).
// ===========================================================================

// ===========================================================================
// Block substitution kernels.  diag_apply: w = op(Tinv) b for one diagonal
// block; a 32-lane group per output row (rows of Tinv read contiguously), or
// for the transpose a lane per output row with the four waves splitting k.
// solve_update: one workgroup per target block, b -= U (S (V x)) with x
// staged in LDS, or b -= D x for a dense block.
// ===========================================================================
// One wave per output row: 64 lanes over k, all loads of a lane issued before
// the products, then a wave reduction per right-hand side.
template <typename T, int NR>
__global__ __launch_bounds__(256) void diag_apply_kernel(const T* __restrict__ Tinv, int64_t ld,
                                                         const T* __restrict__ b, T* __restrict__ w, int n,
                                                         int nrhs) {
  const int lane = threadIdx.x & 63;
  const int row = 4 * blockIdx.x + (threadIdx.x >> 6);
  if (row >= n) return;
  T acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = T(0);
  const T* tr = Tinv + (int64_t)row * ld;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int k = lane + 64 * j;
    if (k < n) {
      const T a = tr[k];
      const T* bk = b + (int64_t)k * nrhs;
#pragma unroll
      for (int c = 0; c < NR; ++c)
        if (c < nrhs) acc[c] += a * bk[c];
    }
  }
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = wave_sum(acc[c]);
  if (lane < nrhs) {
    T v = acc[0];
#pragma unroll
    for (int c = 1; c < NR; ++c)
      if (lane == c) v = acc[c];
    w[(int64_t)row * nrhs + lane] = v;
  }
}

// Transpose: a workgroup per 16 output rows (columns of Tinv), 16 groups of
// 16 lanes over k, partial sums combined in LDS.
template <typename T, int NR>
__global__ __launch_bounds__(256) void diag_apply_t_kernel(const T* __restrict__ Tinv, int64_t ld,
                                                           const T* __restrict__ b, T* __restrict__ w, int n,
                                                           int nrhs) {
  __shared__ T red[16][16][NR + 1];
  const int tid = threadIdx.x, cc = tid & 15, kq = tid >> 4;
  const int row = 16 * blockIdx.x + cc;
  T acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = T(0);
  if (row < n) {
#pragma unroll 4
    for (int k = kq; k < n; k += 16) {
      const T a = Tinv[(int64_t)k * ld + row];
      const T* bk = b + (int64_t)k * nrhs;
#pragma unroll
      for (int c = 0; c < NR; ++c)
        if (c < nrhs) acc[c] += a * bk[c];
    }
  }
#pragma unroll
  for (int c = 0; c < NR; ++c) red[kq][cc][c] = acc[c];
  __syncthreads();
  if (tid < 16 * NR) {
    const int r16 = tid / NR, c = tid % NR;
    const int orow = 16 * blockIdx.x + r16;
    if (orow < n && c < nrhs) {
      T sum = T(0);
#pragma unroll
      for (int q = 0; q < 16; ++q) sum += red[q][r16][c];
      w[(int64_t)orow * nrhs + c] = sum;
    }
  }
}

template <typename T>
void launch_diag_apply(const T* Tinv, int64_t ld, bool trans, const T* b, T* w, int n, int nrhs, hipStream_t s) {
  if (n <= 0 || n > 512 || nrhs <= 0 || nrhs > 16) throw std::runtime_error("diag apply: n <= 512, nrhs <= 16");
  const bool wide = nrhs > 8;
  if (!trans) {
    const dim3 grid((n + 3) / 4);
    if (wide)
      hipLaunchKernelGGL((diag_apply_kernel<T, 16>), grid, dim3(256), 0, s, Tinv, ld, b, w, n, nrhs);
    else
      hipLaunchKernelGGL((diag_apply_kernel<T, 8>), grid, dim3(256), 0, s, Tinv, ld, b, w, n, nrhs);
  } else {
    const dim3 grid((n + 15) / 16);
    if (wide)
      hipLaunchKernelGGL((diag_apply_t_kernel<T, 16>), grid, dim3(256), 0, s, Tinv, ld, b, w, n, nrhs);
    else
      hipLaunchKernelGGL((diag_apply_t_kernel<T, 8>), grid, dim3(256), 0, s, Tinv, ld, b, w, n, nrhs);
  }
  FRK_HIP_CHECK(hipGetLastError());
}

// Rows [0, n) of a row-major n x nrhs block into LDS rows of NR entries
// (entries past nrhs zero), n <= 512, in two halves: load_rows issues every
// load of the thread (out-of-range entries load element 0 and are zeroed),
// store_rows writes the registers to LDS, so other loads can go out between.
template <typename T, int NR>
struct RowsReg {
  T v[512 * NR / 256];
};
template <typename T, int NR>
__device__ __forceinline__ void load_rows(RowsReg<T, NR>& g, const T* __restrict__ src, int n, int nrhs) {
#pragma unroll
  for (int q = 0; q < 2 * NR; ++q) {
    const int e = threadIdx.x + 256 * q, k = e / NR, c = e % NR;
    const bool in = k < n && c < nrhs;
    g.v[q] = src[in ? (int64_t)k * nrhs + c : 0];
    g.v[q] = in ? g.v[q] : T(0);
  }
}
template <typename T, int NR>
__device__ __forceinline__ void store_rows(T (*dst)[NR], const RowsReg<T, NR>& g, int n) {
#pragma unroll
  for (int q = 0; q < 2 * NR; ++q) {
    const int e = threadIdx.x + 256 * q, k = e / NR, c = e % NR;
    if (k < n) dst[k][c] = g.v[q];
  }
}
template <typename T, int NR>
__device__ __forceinline__ void stage_rows(T (*dst)[NR], const T* __restrict__ src, int n, int nrhs) {
  RowsReg<T, NR> g;
  load_rows<T, NR>(g, src, n, nrhs);
  store_rows<T, NR>(dst, g, n);
}

// b -= U (S (V x)) (low-rank) or b -= D x (dense) for one target block; the
// source x is staged in xs.  With bl != nullptr the target is the LDS copy bl
// (rows x NR) instead of the global b.  Loads are issued unconditionally at a
// clamped index (element 0 of every operand exists) and the out-of-range
// values zeroed after, so the loads of one batch are in flight together.
template <typename T, int NR>
__device__ __forceinline__ void solve_update_body(const SolveUpdItem& it, int nrhs, T (*xs)[NR], T (*t1)[NR],
                                                  T (*t2)[NR], T (*bl)[NR], const RowsReg<T, NR>* bpend = nullptr,
                                                  int bn = 0) {
  const int tid = threadIdx.x;
  const T* x = static_cast<const T*>(it.x);
  T* b = static_cast<T*>(it.b);
  if (it.D) {
    stage_rows<T, NR>(xs, x, it.n, nrhs);
    if (bpend) store_rows<T, NR>(bl, *bpend, bn);
    __syncthreads();
    const T* D = static_cast<const T*>(it.D);
    for (int row = tid; row < it.m; row += 256) {
      T acc[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = T(0);
      for (int k0 = 0; k0 < it.n; k0 += 16) {
        T a[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const bool in = k0 + j < it.n;
          a[j] = D[row * it.d_rs + (in ? k0 + j : 0) * it.d_cs];
          a[j] = in ? a[j] : T(0);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j)
          if (k0 + j < it.n)
#pragma unroll
            for (int c = 0; c < NR; ++c) acc[c] += a[j] * xs[k0 + j][c];
      }
#pragma unroll
      for (int c = 0; c < NR; ++c) {
        if (bl)
          bl[row][c] -= acc[c];
        else if (c < nrhs)
          b[(int64_t)row * nrhs + c] -= acc[c];
      }
    }
    return;
  }
  const T* U = static_cast<const T*>(it.U);
  const T* S = static_cast<const T*>(it.S);
  const T* V = static_cast<const T*>(it.V);
  const int r = it.r, n = it.n;
  if (r <= 0) {
    if (bpend) store_rows<T, NR>(bl, *bpend, bn);
    return;
  }
  const int wv = tid >> 6, lane = tid & 63;
  // t1 = V x: one wave per row of V (lanes over k), two rows per round, wave
  // reductions; the first round's loads go out together with the staging of x
  T a[2][8];
  auto load_v = [&](int rho0) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int rho = rho0 + 4 * h;
      const T* vr = V + (rho < r ? rho : 0) * it.v_rs;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = lane + 64 * j;
        const bool in = rho < r && k < n;
        a[h][j] = vr[(in ? k : 0) * it.v_cs];
        a[h][j] = in ? a[h][j] : T(0);
      }
    }
  };
  load_v(wv);
  {
    RowsReg<T, NR> xr;
    load_rows<T, NR>(xr, x, n, nrhs);
    if (bpend) store_rows<T, NR>(bl, *bpend, bn);
    store_rows<T, NR>(xs, xr, n);
  }
  __syncthreads();
  for (int rho0 = wv; rho0 < r; rho0 += 8) {
    if (rho0 != wv) load_v(rho0);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      T acc[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = T(0);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int k = lane + 64 * j;
        if (k < n)
#pragma unroll
          for (int c = 0; c < NR; ++c) acc[c] += a[h][j] * xs[k][c];
      }
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = wave_sum(acc[c]);
      const int rho = rho0 + 4 * h;
      if (lane == 0 && rho < r)
#pragma unroll
        for (int c = 0; c < NR; ++c) t1[rho][c] = acc[c];
    }
  }
  __syncthreads();
  for (int e = tid; e < r * NR; e += 256) {  // t2 = S t1
    const int rho = e / NR, c = e % NR;
    T acc = T(0);
    for (int q0 = 0; q0 < r; q0 += 16) {
      T sv[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const bool in = q0 + j < r;
        sv[j] = S[rho * it.s_rs + (in ? q0 + j : 0) * it.s_cs];
        sv[j] = in ? sv[j] : T(0);
      }
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (q0 + j < r) acc += sv[j] * t1[q0 + j][c];
    }
    t2[rho][c] = acc;
  }
  __syncthreads();
  // b -= U t2: two rows per thread (row, row + 256), sixteen columns of U per round
  for (int row0 = tid; row0 < it.m; row0 += 512) {
    T acc[2][NR];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[h][c] = T(0);
    for (int q0 = 0; q0 < r; q0 += 16) {
      T uv[2][16];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int row = row0 + 256 * h;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const bool in = row < it.m && q0 + j < r;
          uv[h][j] = U[in ? row * it.u_rs + (q0 + j) * it.u_cs : 0];
          uv[h][j] = in ? uv[h][j] : T(0);
        }
      }
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (q0 + j < r)
#pragma unroll
          for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int c = 0; c < NR; ++c) acc[h][c] += uv[h][j] * t2[q0 + j][c];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int row = row0 + 256 * h;
      if (row < it.m)
#pragma unroll
        for (int c = 0; c < NR; ++c) {
          if (bl)
            bl[row][c] -= acc[h][c];
          else if (c < nrhs)
            b[(int64_t)row * nrhs + c] -= acc[h][c];
        }
    }
  }
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void solve_update_kernel(const SolveUpdItem* __restrict__ items, int nrhs) {
  __shared__ T xs[512][NR];
  __shared__ T t1[64][NR];
  __shared__ T t2[64][NR];
  const SolveUpdItem it = items[blockIdx.x];
  solve_update_body<T, NR>(it, nrhs, xs, t1, t2, nullptr);
}

// One block step of a substitution sweep in one launch: workgroups
// [0, n_upd) apply the updates from the block solved last to the blocks two or
// more positions ahead; the remaining ceil(n / kStepRows) workgroups each take
// the next block's right-hand side, apply the one update it still lacks
// (items[n_upd], when present) in LDS, and compute kStepRows rows of
// x = op(Tinv) b.  The two roles
// do not depend on each other within the launch.
constexpr int kStepRows = 16;  // rows of x per diagonal workgroup of a fused step

template <typename T, int NR>
__global__ __launch_bounds__(256) void solve_step_kernel(const SolveUpdItem* __restrict__ items, int n_upd,
                                                         StepDiag sd, int nrhs) {
  __shared__ T xs[512][NR];
  __shared__ T t1[64][NR];
  __shared__ T t2[64][NR];
  __shared__ T bs[512][NR];
  const int tid = threadIdx.x;
  if ((int)blockIdx.x < n_upd) {
    const SolveUpdItem it = items[blockIdx.x];
    solve_update_body<T, NR>(it, nrhs, xs, t1, t2, nullptr);
    return;
  }
  const int g = blockIdx.x - n_upd, n = sd.n, r0 = kStepRows * g;
  const T* Ti = static_cast<const T*>(sd.tinv);
  const int lane = tid & 63, wv = tid >> 6;
  // the rows of Tinv this workgroup applies depend on nothing computed here:
  // their loads go out first, beside those of b (and of the last update)
  constexpr int RPW = kStepRows / 4;  // rows per wave
  T ta[RPW][8];
  if (!sd.trans) {
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const int row = r0 + wv + 4 * i;
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int k = lane + 64 * jj;
        ta[i][jj] = Ti[(row < n && k < n) ? (int64_t)row * sd.ld + k : 0];
      }
    }
  }
  {
    RowsReg<T, NR> br;
    load_rows<T, NR>(br, static_cast<const T*>(sd.b), n, nrhs);
    if (sd.has_last) {  // the update stores b to LDS once its own loads are out
      const SolveUpdItem it = items[n_upd];
      solve_update_body<T, NR>(it, nrhs, xs, t1, t2, bs, &br, n);
    } else {
      store_rows<T, NR>(bs, br, n);
    }
  }
  __syncthreads();
  T* x = static_cast<T*>(sd.x);
  if (!sd.trans) {
    // x[r] = sum_k Tinv[r][k] b[k]: a wave per row, lanes over k
#pragma unroll
    for (int i = 0; i < RPW; ++i) {
      const int rr = wv + 4 * i;
      if (r0 + rr >= n) break;
      T acc[NR];
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = T(0);
      const T* a = ta[i];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
        const int k = lane + 64 * jj;
        if (k < n)
#pragma unroll
          for (int c = 0; c < NR; ++c) acc[c] += a[jj] * bs[k][c];
      }
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = wave_sum(acc[c]);
      if (lane < nrhs) {
        T v = acc[0];
#pragma unroll
        for (int c = 1; c < NR; ++c)
          if (lane == c) v = acc[c];
        x[(int64_t)(r0 + rr) * nrhs + lane] = v;
      }
    }
  } else {
    // x[r] = sum_k Tinv[k][r] b[k]: a lane per row, the four waves split k,
    // partial sums meet in LDS (xs is free once the update is applied)
    T (*red)[NR + 1] = reinterpret_cast<T (*)[NR + 1]>(&xs[0][0]);
    const int r = r0 + lane;
    T acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = T(0);
    if (lane < kStepRows && r < n)
      for (int k = wv; k < n; k += 4) {
        const T a = Ti[(int64_t)k * sd.ld + r];
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[c] += a * bs[k][c];
      }
#pragma unroll
    for (int c = 0; c < NR; ++c) red[wv * 64 + lane][c] = acc[c];
    __syncthreads();
    for (int e = tid; e < kStepRows * NR; e += 256) {
      const int rr = e / NR, c = e % NR;
      if (r0 + rr < n && c < nrhs)
        x[(int64_t)(r0 + rr) * nrhs + c] =
            red[rr][c] + red[64 + rr][c] + red[128 + rr][c] + red[192 + rr][c];
    }
  }
}

template <typename T>
void launch_solve_step(const SolveUpdItem* d_items, int n_upd, const StepDiag& sd, int nrhs, hipStream_t s) {
  if (nrhs <= 0 || nrhs > 16 || sd.n <= 0 || sd.n > 512) throw std::runtime_error("solve step: nrhs <= 16, n <= 512");
  const dim3 grid(n_upd + (sd.n + kStepRows - 1) / kStepRows);
  if (nrhs > 8)
    hipLaunchKernelGGL((solve_step_kernel<T, 16>), grid, dim3(256), 0, s, d_items, n_upd, sd, nrhs);
  else
    hipLaunchKernelGGL((solve_step_kernel<T, 8>), grid, dim3(256), 0, s, d_items, n_upd, sd, nrhs);
  FRK_HIP_CHECK(hipGetLastError());
}
template void launch_solve_step<double>(const SolveUpdItem*, int, const StepDiag&, int, hipStream_t);
template void launch_solve_step<float>(const SolveUpdItem*, int, const StepDiag&, int, hipStream_t);

template <typename T>
void launch_solve_update(const SolveUpdItem* d_items, int n_items, int nrhs, hipStream_t s) {
  if (n_items <= 0) return;
  if (nrhs <= 0 || nrhs > 16) throw std::runtime_error("solve update: nrhs <= 16");
  if (nrhs > 8)
    hipLaunchKernelGGL((solve_update_kernel<T, 16>), dim3(n_items), dim3(256), 0, s, d_items, nrhs);
  else
    hipLaunchKernelGGL((solve_update_kernel<T, 8>), dim3(n_items), dim3(256), 0, s, d_items, nrhs);
  FRK_HIP_CHECK(hipGetLastError());
}
template void launch_diag_apply<double>(const double*, int64_t, bool, const double*, double*, int, int, hipStream_t);
template void launch_diag_apply<float>(const float*, int64_t, bool, const float*, float*, int, int, hipStream_t);
template void launch_solve_update<double>(const SolveUpdItem*, int, int, hipStream_t);
template void launch_solve_update<float>(const SolveUpdItem*, int, int, hipStream_t);


// ===========================================================================
// Stacked-row products: the low-rank panel solves of one elimination step
// (V <- V U^-1 for every tile of the block column, U^T <- U^T L^-T for every
// tile of the block row) as one GEMM over the concatenated rows.  A 16 x 64
// output tile per workgroup, K staged through LDS in chunks of 64.
// ===========================================================================
// The 16-row slab is staged in LDS once; the kRowsWaves waves split K, every
// lane owns one output column (so the reads of B are contiguous in the
// row-major case and cache-line reuse serves the transposed one), and the
// partial sums of the waves meet in LDS.  The kernel is bound by the chain of
// B reads of one wave: eight waves halve it against four.
template <typename T>
constexpr int rows_waves() { return sizeof(T) == 4 ? 16 : 8; }
template <typename T, class RowFn>
__device__ __forceinline__ void rows_gemm_body(RowFn row_of, int R, const int32_t* __restrict__ d_count,
                                               const T* __restrict__ B, int64_t b_rs, int64_t b_cs,
                                               T* __restrict__ W, int N, int K) {
  constexpr int kRowsWaves = rows_waves<T>();
  constexpr int NT = 64 * kRowsWaves, KH = (512 + NT - 1) / NT;
  __shared__ T As[16][512 + 1];
  __shared__ T red[kRowsWaves][16][64];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int r0 = 16 * blockIdx.x, col = 64 * blockIdx.y + lane;
  if (r0 >= R || 64 * (int)blockIdx.y >= N) return;
  {
    // stage the 16-row slab with every load issued up front: rows past R and
    // columns past K read row r0, column 0 (always present) and are zeroed after
    const T* rp[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) rp[r] = row_of(r0 + r < R ? r0 + r : r0);
    T v[16][KH];
#pragma unroll
    for (int r = 0; r < 16; ++r)
#pragma unroll
      for (int h = 0; h < KH; ++h) {
        const int k = tid + NT * h;
        v[r][h] = rp[r][k < K ? k : 0];
      }
    // the device count (rows actually present, <= R) is read only now, so its
    // load overlaps the slab's; rows of the table past it are in bounds
    if (d_count) R = min(R, max(0, *d_count));
    if (r0 >= R) return;
#pragma unroll
    for (int r = 0; r < 16; ++r)
#pragma unroll
      for (int h = 0; h < KH; ++h) {
        const int k = tid + NT * h;
        if (k < K) As[r][k] = (r0 + r < R) ? v[r][h] : T(0);
      }
  }
  __syncthreads();
  T acc[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = T(0);
  const int kq = (K + kRowsWaves - 1) / kRowsWaves, k0 = wv * kq, k1 = min(K, k0 + kq);
  if (col < N) {
    const T* bc = B + (int64_t)col * b_cs;
#pragma unroll 4
    for (int k = k0; k < k1; ++k) {
      const T bv = bc[(int64_t)k * b_rs];
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] += As[r][k] * bv;
    }
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) red[wv][r][lane] = acc[r];
  __syncthreads();
  for (int e = tid; e < 16 * 64; e += NT) {
    const int r = e >> 6, c = e & 63;
    const int oc = 64 * blockIdx.y + c;
    if (r0 + r < R && oc < N) {
      T sum = T(0);
#pragma unroll
      for (int w2 = 0; w2 < kRowsWaves; ++w2) sum += red[w2][r][c];
      W[(int64_t)(r0 + r) * N + oc] = sum;
    }
  }
}

// d_count (optional): the number of rows actually present, read on the device
// (the table then holds R >= count entries); workgroups past it exit.
template <typename T>
__global__ __launch_bounds__(64 * rows_waves<T>()) void rows_gemm_kernel(const RowRef* __restrict__ rows, int R,
                                                                    const T* __restrict__ B, int64_t b_rs,
                                                                    int64_t b_cs, T* __restrict__ W, int N, int K,
                                                                    const int32_t* __restrict__ d_count) {
  rows_gemm_body<T>([&](int r) { return static_cast<const T*>(rows[r].a); }, R, d_count, B, b_rs, b_cs, W, N, K);
}

template <typename T>
__global__ __launch_bounds__(64 * rows_waves<T>()) void rows_pair_kernel(RowsPair p) {
  const RowsJob j = blockIdx.z ? p.job[1] : p.job[0];
  const T* base = static_cast<const T*>(j.base);
  const int64_t st = j.stride;
  rows_gemm_body<T>([&](int r) { return base + (int64_t)r * st; }, j.R, j.d_count, static_cast<const T*>(j.b),
                    j.b_rs, j.b_cs, static_cast<T*>(j.w), j.N, j.K);
}

template <typename T>
__global__ __launch_bounds__(256) void rows_scatter_kernel(const RowRef* __restrict__ rows, int R,
                                                           const T* __restrict__ W, int N,
                                                           const int32_t* __restrict__ d_count) {
  const int r = blockIdx.x;
  if (d_count) R = min(R, max(0, *d_count));
  if (r >= R) return;
  T* c = static_cast<T*>(rows[r].c);
  for (int j = threadIdx.x; j < N; j += 256) c[j] = W[(int64_t)r * N + j];
}

template <typename T>
void launch_rows_gemm(const RowRef* d_rows, int R, const T* B, int64_t b_rs, int64_t b_cs, T* W, int N, int K,
                      hipStream_t s, const int32_t* d_count) {
  if (R <= 0) return;
  if (N <= 0 || K <= 0 || K > 512) throw std::runtime_error("rows gemm: N > 0, 0 < K <= 512");
  hipLaunchKernelGGL(rows_gemm_kernel<T>, dim3((R + 15) / 16, (N + 63) / 64), dim3(64 * rows_waves<T>()), 0, s, d_rows, R, B, b_rs,
                     b_cs, W, N, K, d_count);
  FRK_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_rows_scatter(const RowRef* d_rows, int R, const T* W, int N, hipStream_t s, const int32_t* d_count) {
  if (R <= 0) return;
  hipLaunchKernelGGL(rows_scatter_kernel<T>, dim3(R), dim3(256), 0, s, d_rows, R, W, N, d_count);
  FRK_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_rows_pair(const RowsPair& p, hipStream_t s) {
  int R = 0, N = 0;
  for (const auto& j : p.job) {
    if (j.R <= 0) continue;
    if (j.N <= 0 || j.K <= 0 || j.K > 512) throw std::runtime_error("rows pair: N > 0, 0 < K <= 512");
    R = std::max(R, (int)j.R);
    N = std::max(N, (int)j.N);
  }
  if (R <= 0) return;
  hipLaunchKernelGGL(rows_pair_kernel<T>, dim3((R + 15) / 16, (N + 63) / 64, 2), dim3(64 * rows_waves<T>()), 0, s, p);
  FRK_HIP_CHECK(hipGetLastError());
}
template void launch_rows_pair<double>(const RowsPair&, hipStream_t);
template void launch_rows_pair<float>(const RowsPair&, hipStream_t);

template <typename T>
__global__ __launch_bounds__(64 * rows_waves<T>()) void rows_gemm2_kernel(RowsList2 p) {
  const RowsListJob j = blockIdx.z ? p.job[1] : p.job[0];
  const RowRef* rows = j.rows;
  rows_gemm_body<T>([&](int r) { return static_cast<const T*>(rows[r].a); }, j.R, nullptr,
                    static_cast<const T*>(j.b), j.b_rs, j.b_cs, static_cast<T*>(j.w), j.N, j.K);
}

template <typename T>
void launch_rows_gemm2(const RowsList2& p, hipStream_t s) {
  int R = 0, N = 0;
  for (const auto& j : p.job) {
    if (j.R <= 0) continue;
    if (j.N <= 0 || j.K <= 0 || j.K > 512) throw std::runtime_error("rows gemm2: N > 0, 0 < K <= 512");
    R = std::max(R, (int)j.R);
    N = std::max(N, (int)j.N);
  }
  if (R <= 0) return;
  hipLaunchKernelGGL(rows_gemm2_kernel<T>, dim3((R + 15) / 16, (N + 63) / 64, 2), dim3(64 * rows_waves<T>()), 0, s, p);
  FRK_HIP_CHECK(hipGetLastError());
}
template void launch_rows_gemm2<double>(const RowsList2&, hipStream_t);
template void launch_rows_gemm2<float>(const RowsList2&, hipStream_t);

// Panel tiles after their solve, a workgroup per (tile, 64 positions): the
// r solved rows are staged through LDS (and written back to the tile), then
// each of the r outputs of the stacked factor is formed from them.
template <typename T>
__global__ __launch_bounds__(256) void panel_finish_kernel(const PanelFin* __restrict__ items) {
  const PanelFin it = items[blockIdx.y];
  const int x0 = 64 * blockIdx.x;
  if (x0 >= it.n || it.r <= 0) return;
  __shared__ T Ws[64][65];
  __shared__ T Sc[64][65];
  const int tid = threadIdx.x, lx = tid & 63, g = tid >> 6;
  const int r = min(it.r, 64), x = x0 + lx;
  const T* __restrict__ src = static_cast<const T*>(it.src);
  T* __restrict__ dst = static_cast<T*>(it.dst);
  const T* __restrict__ S = static_cast<const T*>(it.s);
  T* __restrict__ pq = static_cast<T*>(it.pq);
  for (int e = tid; e < r * r; e += 256) {
    const int o = e / r, sg = e - o * r;
    Sc[o][sg] = S[o * it.s_rs + sg * it.s_cs];
  }
  for (int sg = g; sg < r; sg += 4) {
    const T v = (x < it.n) ? src[sg * it.src_ld + x] : T(0);
    Ws[sg][lx] = v;
    if (dst != nullptr && x < it.n) dst[sg * it.dst_ld + x] = v;
  }
  __syncthreads();
  if (x < it.n)
    for (int o = g; o < r; o += 4) {
      T acc = T(0);
      for (int sg = 0; sg < r; ++sg) acc += Sc[o][sg] * Ws[sg][lx];
      pq[(int64_t)o * it.n + x] = static_cast<T>(it.alpha) * acc;
    }
}

template <typename T>
void launch_panel_finish(const PanelFin* d_items, int n_items, int max_n, hipStream_t s) {
  if (n_items <= 0 || max_n <= 0) return;
  hipLaunchKernelGGL(panel_finish_kernel<T>, dim3((max_n + 63) / 64, n_items), dim3(256), 0, s, d_items);
  FRK_HIP_CHECK(hipGetLastError());
}
template void launch_panel_finish<double>(const PanelFin*, int, int, hipStream_t);
template void launch_panel_finish<float>(const PanelFin*, int, int, hipStream_t);
template void launch_rows_gemm<double>(const RowRef*, int, const double*, int64_t, int64_t, double*, int, int,
                                       hipStream_t, const int32_t*);
template void launch_rows_gemm<float>(const RowRef*, int, const float*, int64_t, int64_t, float*, int, int,
                                      hipStream_t, const int32_t*);
template void launch_rows_scatter<double>(const RowRef*, int, const double*, int, hipStream_t, const int32_t*);
template void launch_rows_scatter<float>(const RowRef*, int, const float*, int, hipStream_t, const int32_t*);

// ===========================================================================
// Schur update of a dense tile by a product of two low-rank tiles whose ranks
// are read on the device, so the launch can be queued before they are known:
// C += alpha U_l (S_l (V_l U_b) S_b) V_b.  lr_core_kernel (one workgroup)
// forms the rl x rb core; lr_update_kernel applies it to 64 x 64 tiles of C.
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(256) void lr_core_kernel(LrCoreItem it) {
  constexpr int KC = sizeof(T) == 8 ? 64 : 256;  // K chunk staged in LDS (FP64 has less room)
  __shared__ T M1[64][65];
  __shared__ T T1[64][65];
  __shared__ T Sls[64][65];
  __shared__ T Sbs[64][65];
  __shared__ T Va[16][KC + 1];
  __shared__ T Ub[KC][17];
  const int tid = threadIdx.x;
  const int kc = it.kcap;
  const T* __restrict__ V = static_cast<const T*>(it.vl);
  const T* __restrict__ U = static_cast<const T*>(it.ub);
  const T* __restrict__ Sl = static_cast<const T*>(it.sl);
  const T* __restrict__ Sb = static_cast<const T*>(it.sb);
  T* __restrict__ core = static_cast<T*>(it.core);
  const int K = it.k;
  const int ti = tid >> 4, tj = tid & 15;
  // Loads are issued unconditionally at a clamped index (element 0 of every
  // operand exists: rl, rb, K >= 1) and out-of-range entries zeroed after, so
  // a thread's loads of one round are in flight together.
  // the S cores into LDS (one round of loads, at the capacity kc, so that it
  // overlaps the load of the device ranks)
  {
    T sl[16], sb[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = tid + 256 * q, r = e >> 6, c = e & 63;
      const bool in = r < kc && c < kc;
      sl[q] = Sl[in ? (int64_t)r * it.lds + c : 0];
      sb[q] = Sb[in ? (int64_t)r * it.lds + c : 0];
    }
    // lr_core is the only writer of the solved panel rows (vl_dst / ub_dst), so
    // it returns here only when both panels are empty; with one rank zero the
    // other panel is still written back below and only the core is skipped
    const int rl0 = min(kc, max(0, *it.rl)), rb0 = min(kc, max(0, *it.rb));
    if (rl0 == 0 && rb0 == 0) return;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = tid + 256 * q, r = e >> 6, c = e & 63;
      sl[q] = (r < rl0 && c < rl0) ? sl[q] : T(0);
      sb[q] = (r < rb0 && c < rb0) ? sb[q] : T(0);
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int e = tid + 256 * q, r = e >> 6, c = e & 63;
      Sls[r][c] = sl[q];
      Sbs[r][c] = sb[q];
    }
  }
  const int rl = min(kc, max(0, *it.rl)), rb = min(kc, max(0, *it.rb));
  // the panel rows written back: rows of V_l (columns of U_b), eight per round
  if (it.vl_dst) {
    T* __restrict__ d = static_cast<T*>(it.vl_dst);
    for (int j0 = 0; j0 < it.nv; j0 += 256)
      for (int r0 = 0; r0 < rl; r0 += 8) {
        const int j = j0 + tid;
        T v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const bool in = r0 + q < rl && j < it.nv;
          v[q] = V[in ? (int64_t)(r0 + q) * it.v_rs + (int64_t)j * it.v_cs : 0];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (r0 + q < rl && j < it.nv) d[(int64_t)(r0 + q) * it.vl_dst_ld + j] = v[q];
      }
  }
  if (it.ub_dst) {
    T* __restrict__ d = static_cast<T*>(it.ub_dst);
    for (int j0 = 0; j0 < it.nu; j0 += 256)
      for (int c0 = 0; c0 < rb; c0 += 8) {
        const int j = j0 + tid;
        T v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const bool in = c0 + q < rb && j < it.nu;
          v[q] = U[in ? (int64_t)j * it.u_rs + (int64_t)(c0 + q) * it.u_cs : 0];
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (c0 + q < rb && j < it.nu) d[(int64_t)(c0 + q) * it.ub_dst_ld + j] = v[q];
      }
  }
  // an empty product leaves the core unused (lr_update_kernel skips the target)
  if (rl == 0 || rb == 0) return;
  // M1 = V_l U_b (rl x rb), 16 x 16 output blocks, K in chunks of KC
  constexpr int NQ = 16 * KC / 256;
  if (rl <= 16 && rb <= 16) {
    // one output block: FP64 MFMA 16 x 16 x 4 products, each wave a quarter of
    // every chunk, the four partial blocks summed through T1
    const int lane = tid & 63, w = tid >> 6, li = lane & 15, lk = lane >> 4;
    constexpr int KW = KC / 4;
    d64::d4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += KC) {
      T va[NQ], ub[NQ];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int e = tid + 256 * q;
        const int r = e / KC, kk = e % KC, kb = e >> 4, c = e & 15;
        const bool iv = r < rl && k0 + kk < K, iu = c < rb && k0 + kb < K;
        va[q] = V[iv ? (int64_t)r * it.v_rs + (int64_t)(k0 + kk) * it.v_cs : 0];
        ub[q] = U[iu ? (int64_t)(k0 + kb) * it.u_rs + (int64_t)c * it.u_cs : 0];
        va[q] = iv ? va[q] : T(0);
        ub[q] = iu ? ub[q] : T(0);
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int e = tid + 256 * q;
        Va[e / KC][e % KC] = va[q];
        Ub[e >> 4][e & 15] = ub[q];
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = w * KW; kk < w * KW + KW; kk += 4) {
        const double a = static_cast<double>(Va[li][kk + lk]);
        const double b = static_cast<double>(Ub[kk + lk][li]);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
      }
    }
    // lane (li, lk) holds rows 4 v + lk, column li of the wave's partial block
#pragma unroll
    for (int v = 0; v < 4; ++v) T1[16 * w + 4 * v + lk][li] = static_cast<T>(acc[v]);
    __syncthreads();
    {
      const int i = tid >> 4, j = tid & 15;
      if (i < rl && j < rb) M1[i][j] = T1[i][j] + T1[16 + i][j] + T1[32 + i][j] + T1[48 + i][j];
    }
  } else {
    for (int bi = 0; bi < rl; bi += 16)
      for (int bj = 0; bj < rb; bj += 16) {
        T acc = T(0);
        for (int k0 = 0; k0 < K; k0 += KC) {
          T va[NQ], ub[NQ];
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const int e = tid + 256 * q;
            const int r = e / KC, kk = e % KC, kb = e >> 4, c = e & 15;
            const bool iv = bi + r < rl && k0 + kk < K, iu = bj + c < rb && k0 + kb < K;
            va[q] = V[iv ? (int64_t)(bi + r) * it.v_rs + (int64_t)(k0 + kk) * it.v_cs : 0];
            ub[q] = U[iu ? (int64_t)(k0 + kb) * it.u_rs + (int64_t)(bj + c) * it.u_cs : 0];
            va[q] = iv ? va[q] : T(0);
            ub[q] = iu ? ub[q] : T(0);
          }
          __syncthreads();
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const int e = tid + 256 * q;
            Va[e / KC][e % KC] = va[q];
            Ub[e >> 4][e & 15] = ub[q];
          }
          __syncthreads();
#pragma unroll 8
          for (int kk = 0; kk < KC; ++kk) acc += Va[ti][kk] * Ub[kk][tj];
        }
        if (bi + ti < rl && bj + tj < rb) M1[bi + ti][bj + tj] = acc;
      }
  }
  __syncthreads();
  // T1 = M1 S_b, core = alpha S_l T1 (all operands in LDS)
  for (int e = tid; e < rl * rb; e += 256) {
    const int i = e / rb, j = e % rb;
    T sacc = T(0);
    for (int q = 0; q < rb; ++q) sacc += M1[i][q] * Sbs[q][j];
    T1[i][j] = sacc;
  }
  __syncthreads();
  for (int e = tid; e < rl * rb; e += 256) {
    const int i = e / rb, j = e % rb;
    T sacc = T(0);
    for (int q = 0; q < rl; ++q) sacc += Sls[i][q] * T1[q][j];
    core[(int64_t)i * it.ldk + j] = T(it.alpha) * sacc;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void lr_update_kernel(LrUpdItem it) {
  __shared__ T Ks[64][65];
  __shared__ T Vt[64][65];
  __shared__ T Wt[64][65];
  __shared__ T Ut[64][65];
  const int tid = threadIdx.x;
  const int kc = it.kcap;
  const int c0 = 64 * blockIdx.x, i0 = 64 * blockIdx.y;
  const T* core = static_cast<const T*>(it.core);
  const T* V = static_cast<const T*>(it.v);
  const T* U = static_cast<const T*>(it.u);
  T* C = static_cast<T*>(it.c);
  // every load of the thread issued before the first use (unconditional, at a
  // clamped index), the low-rank operands at their capacity kc (always
  // allocated), so that none waits for the device ranks; entries outside the
  // operands or past the ranks are zeroed afterwards
  T cv[16], kv[16], vv[16], uv[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = tid + 256 * q, a = e >> 6, b = e & 63;
    const bool ic = i0 + a < it.m && c0 + b < it.n, ik = a < kc && b < kc, iv = a < kc && c0 + b < it.n,
               iu = i0 + a < it.m && b < kc;
    cv[q] = C[ic ? (int64_t)(i0 + a) * it.ldc + c0 + b : 0];
    kv[q] = core[ik ? (int64_t)a * it.ldk + b : 0];
    vv[q] = V[iv ? (int64_t)a * it.v_rs + (int64_t)(c0 + b) * it.v_cs : 0];
    uv[q] = U[iu ? (int64_t)(i0 + a) * it.u_rs + (int64_t)b * it.u_cs : 0];
  }
  const int r1 = min(kc, max(0, *it.r1)), r2 = min(kc, max(0, *it.r2));
  if (r1 == 0 || r2 == 0) return;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = tid + 256 * q, a = e >> 6, b = e & 63;
    cv[q] = (i0 + a < it.m && c0 + b < it.n) ? cv[q] : T(0);
    kv[q] = (a < r1 && b < r2) ? kv[q] : T(0);
    vv[q] = (a < r2 && c0 + b < it.n) ? vv[q] : T(0);
    uv[q] = (i0 + a < it.m && b < r1) ? uv[q] : T(0);
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int e = tid + 256 * q, a = e >> 6, b = e & 63;
    Ks[a][b] = kv[q];
    Vt[a][b] = vv[q];
    Ut[a][b] = uv[q];
  }
  __syncthreads();
  // W = core V_b(:, c0:c0+64)  (r1 x 64)
  for (int e = tid; e < r1 * 64; e += 256) {
    const int a = e >> 6, b = e & 63;
    T sacc = T(0);
    for (int q = 0; q < r2; ++q) sacc += Ks[a][q] * Vt[q][b];
    Wt[a][b] = sacc;
  }
  __syncthreads();
#pragma unroll
  for (int q2 = 0; q2 < 16; ++q2) {
    const int e = tid + 256 * q2, a = e >> 6, b = e & 63;
    T sacc = T(0);
    for (int q = 0; q < r1; ++q) sacc += Ut[a][q] * Wt[q][b];
    if (i0 + a < it.m && c0 + b < it.n) C[(int64_t)(i0 + a) * it.ldc + c0 + b] = cv[q2] + sacc;
  }
}

template <typename T>
void launch_lr_core(const LrCoreItem& it, hipStream_t s) {
  if (it.kcap <= 0 || it.kcap > 64 || it.k <= 0) throw std::runtime_error("lr core: 0 < kcap <= 64, k > 0");
  hipLaunchKernelGGL(lr_core_kernel<T>, dim3(1), dim3(256), 0, s, it);
  FRK_HIP_CHECK(hipGetLastError());
}
template <typename T>
void launch_lr_update(const LrUpdItem& it, hipStream_t s) {
  if (it.kcap <= 0 || it.kcap > 64 || it.m <= 0 || it.n <= 0) throw std::runtime_error("lr update: 0 < kcap <= 64");
  hipLaunchKernelGGL(lr_update_kernel<T>, dim3((it.n + 63) / 64, (it.m + 63) / 64), dim3(256), 0, s, it);
  FRK_HIP_CHECK(hipGetLastError());
}
template void launch_lr_core<double>(const LrCoreItem&, hipStream_t);
template void launch_lr_core<float>(const LrCoreItem&, hipStream_t);
template void launch_lr_update<double>(const LrUpdItem&, hipStream_t);
template void launch_lr_update<float>(const LrUpdItem&, hipStream_t);

constexpr int kResMaxRhs = 16;
constexpr int kResSplits = 16;

__global__ __launch_bounds__(256) void residual_partial_kernel(KernelSpec ks, const double* __restrict__ X,
                                                               double* __restrict__ part, int64_t n,
                                                               int nrhs) {
#pragma clang fp contract(off)
  __shared__ double xs[3][256];
  __shared__ double vs[256][kResMaxRhs + 1];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int split = blockIdx.y;
  const int64_t chunk = (n + kResSplits - 1) / kResSplits;
  const int64_t j0 = split * chunk, j1 = std::min<int64_t>(n, j0 + chunk);
  double acc[kResMaxRhs];
  for (int c = 0; c < kResMaxRhs; ++c) acc[c] = 0.0;
  double xi[3] = {0.0, 0.0, 0.0};
  if (i < n)
    for (int d = 0; d < ks.dim; ++d) xi[d] = ks.pts[d * ks.n_pts + i];
  for (int64_t jb = j0; jb < j1; jb += 256) {
    const int64_t j = jb + threadIdx.x;
    __syncthreads();
    if (j < j1) {
      for (int d = 0; d < ks.dim; ++d) xs[d][threadIdx.x] = ks.pts[d * ks.n_pts + j];
      for (int c = 0; c < nrhs; ++c) vs[threadIdx.x][c] = X[j * nrhs + c];
    }
    __syncthreads();
    const int cnt = (int)std::min<int64_t>(256, j1 - jb);
    if (i < n) {
      for (int jj = 0; jj < cnt; ++jj) {
        double rij = 0.0;
        for (int d = 0; d < ks.dim; ++d) {
          const double dd = xi[d] - xs[d][jj];
          rij = rij + dd * dd;
        }
        double a;
        if (ks.kind == 0) {
          a = 1.0 / (sqrt(rij) + 1e-3);
        } else if (ks.kind == 1) {
          a = exp(-rij / ((2.0 * ks.ell) * ks.ell));
        } else if (ks.kind == 2) {
          a = (jb + jj == i) ? 1.0 : 0.0;
        } else {  // d/d ell of the gaussian kernel: exp(-r^2 / (2 ell^2)) r^2 / ell^3
          a = exp(-rij / ((2.0 * ks.ell) * ks.ell)) * (rij / (ks.ell * ks.ell * ks.ell));
        }
        if (jb + jj == i) a = a + ks.shift;
        for (int c = 0; c < nrhs; ++c) acc[c] = fma(a, vs[jj][c], acc[c]);
      }
    }
  }
  if (i < n)
    for (int c = 0; c < nrhs; ++c) part[((int64_t)split * n + i) * nrhs + c] = acc[c];
}


// Laplace kernel 1 / (|x_i - x_j| + 1e-3) (+ shift on the diagonal), with the
// dimension and the (padded) right-hand-side count fixed at compile time and
// the reciprocal from the hardware estimate refined by two Newton steps.
template <int DIM, int NR>
__global__ __launch_bounds__(256) void residual_laplace_kernel(KernelSpec ks, const double* __restrict__ X,
                                                               double* __restrict__ part, int64_t n, int nrhs) {
#pragma clang fp contract(off)
  __shared__ double xs[DIM][256];
  __shared__ __align__(16) double vs[256][NR];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int split = blockIdx.y;
  const int64_t chunk = (n + kResSplits - 1) / kResSplits;
  const int64_t j0 = split * chunk, j1 = std::min<int64_t>(n, j0 + chunk);
  double acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) acc[c] = 0.0;
  double xi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) xi[d] = (i < n) ? ks.pts[d * ks.n_pts + i] : 0.0;
  for (int64_t jb = j0; jb < j1; jb += 256) {
    const int64_t j = jb + threadIdx.x;
    __syncthreads();
    if (j < j1) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) xs[d][threadIdx.x] = ks.pts[d * ks.n_pts + j];
#pragma unroll
      for (int c = 0; c < NR; ++c) vs[threadIdx.x][c] = (c < nrhs) ? X[j * nrhs + c] : 0.0;
    }
    __syncthreads();
    const int cnt = (int)std::min<int64_t>(256, j1 - jb);
    for (int jj = 0; jj < cnt; ++jj) {
      double r;
      if (DIM == 1) {
        r = fabs(xi[0] - xs[0][jj]);  // equals sqrt(d * d) exactly
      } else {
        double rij = 0.0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const double dd = xi[d] - xs[d][jj];
          rij = rij + dd * dd;
        }
        r = sqrt(rij);
      }
      const double a = d64::rcp2(r + 1e-3);
#pragma unroll
      for (int c = 0; c < NR; ++c) acc[c] = fma(a, vs[jj][c], acc[c]);
    }
  }
  // the shift on the diagonal, once, by the split whose column range holds i
  // (kept out of the pair loop: no compare and select per entry)
  if (i < n && i >= j0 && i < j1 && ks.shift != 0.0)
#pragma unroll
    for (int c = 0; c < NR; ++c)
      if (c < nrhs) acc[c] = fma(ks.shift, X[i * nrhs + c], acc[c]);
  if (i < n)
#pragma unroll
    for (int c = 0; c < NR; ++c)
      if (c < nrhs) part[((int64_t)split * n + i) * nrhs + c] = acc[c];
}

template <int DIM>
static void launch_residual_laplace(dim3 grid, hipStream_t s, const KernelSpec& ks, const double* X, double* part,
                                    int64_t n, int nrhs) {
  if (nrhs <= 1)
    hipLaunchKernelGGL((residual_laplace_kernel<DIM, 1>), grid, dim3(256), 0, s, ks, X, part, n, nrhs);
  else if (nrhs <= 2)
    hipLaunchKernelGGL((residual_laplace_kernel<DIM, 2>), grid, dim3(256), 0, s, ks, X, part, n, nrhs);
  else if (nrhs <= 4)
    hipLaunchKernelGGL((residual_laplace_kernel<DIM, 4>), grid, dim3(256), 0, s, ks, X, part, n, nrhs);
  else if (nrhs <= 8)
    hipLaunchKernelGGL((residual_laplace_kernel<DIM, 8>), grid, dim3(256), 0, s, ks, X, part, n, nrhs);
  else
    hipLaunchKernelGGL((residual_laplace_kernel<DIM, 16>), grid, dim3(256), 0, s, ks, X, part, n, nrhs);
}

__global__ void residual_reduce_kernel(const double* __restrict__ part, const double* __restrict__ B,
                                       double* __restrict__ R, int64_t n, int nrhs) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= n * nrhs) return;
  double s = 0.0;
  for (int sp = 0; sp < kResSplits; ++sp) s += part[(int64_t)sp * n * nrhs + e];
  R[e] = (B != nullptr) ? B[e] - s : s;
}

size_t kernel_residual_scratch(int64_t n, int nrhs) { return (size_t)kResSplits * n * nrhs; }

void launch_kernel_residual(const KernelSpec& ks, const double* X, const double* B, double* R, int64_t n,
                            int nrhs, double* part, hipStream_t s) {
  if (nrhs > kResMaxRhs) throw std::runtime_error("residual: at most 16 right-hand sides");
  if (part == nullptr) throw std::invalid_argument("residual: missing partial-sum scratch");
  dim3 grid((unsigned)((n + 255) / 256), kResSplits);
  if (ks.kind == 0 && ks.dim == 1)
    launch_residual_laplace<1>(grid, s, ks, X, part, n, nrhs);
  else if (ks.kind == 0 && ks.dim == 2)
    launch_residual_laplace<2>(grid, s, ks, X, part, n, nrhs);
  else if (ks.kind == 0 && ks.dim == 3)
    launch_residual_laplace<3>(grid, s, ks, X, part, n, nrhs);
  else
    hipLaunchKernelGGL(residual_partial_kernel, grid, dim3(256), 0, s, ks, X, part, n, nrhs);
  FRK_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(residual_reduce_kernel, dim3((unsigned)((n * nrhs + 255) / 256)), dim3(256), 0, s, part,
                     B, R, n, nrhs);
  FRK_HIP_CHECK(hipGetLastError());
}

__global__ __launch_bounds__(256) void colmax_kernel(const double* __restrict__ a, int64_t n, int nrhs,
                                                     double* __restrict__ out) {
  __shared__ double red[256];
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = blockIdx.x * chunk, r1 = std::min<int64_t>(n, r0 + chunk);
  for (int c = 0; c < nrhs; ++c) {
    double m = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) m = fmax(m, fabs(a[i * nrhs + c]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
      __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x * nrhs + c] = red[0];
    __syncthreads();
  }
}

void launch_colmax_abs(const double* a, int64_t n, int nrhs, double* out, hipStream_t s) {
  hipLaunchKernelGGL(colmax_kernel, dim3(kColmaxBlocks), dim3(256), 0, s, a, n, nrhs, out);
  FRK_HIP_CHECK(hipGetLastError());
}

__global__ void fill_value_kernel(double* p, int64_t n, double v) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) p[e] = v;
}

void launch_fill_value(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_value_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, s, p,
                     n, v);
  FRK_HIP_CHECK(hipGetLastError());
}

template <typename T>
__global__ void axpy_convert_kernel(double* x, const T* d, int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
    x[e] += static_cast<double>(d[e]);
}

template <typename T>
void launch_axpy_convert(double* x, const T* d, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpy_convert_kernel<T>, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0,
                     s, x, d, n);
  FRK_HIP_CHECK(hipGetLastError());
}
template void launch_axpy_convert<double>(double*, const double*, int64_t, hipStream_t);
template void launch_axpy_convert<float>(double*, const float*, int64_t, hipStream_t);

// ---------------------------------------------------------------------------
// Explicit instantiations
// ---------------------------------------------------------------------------
#define FRK_INST(T)                                                                              \
  template void launch_gemm<T>(const GemmItem*, int, int, int, hipStream_t);                     \
  template void launch_trsm<T>(const TrsmItem*, int, int, int, hipStream_t);                     \
  template void launch_getrf_small<T>(const LuItem*, int, int32_t*, hipStream_t);                \
  template void launch_tri_inv<T>(const TriInvItem*, int, hipStream_t);                          \
  template void launch_copy<T>(const GemmItem*, int, int64_t, hipStream_t);                      \
  template void launch_qr<T>(const QrItem*, int, int, hipStream_t);                              \
  template void launch_tpqrt<T>(const TpqrtItem*, int, int, int, hipStream_t);                   \
  template void launch_larft<T>(const LarftItem*, int, hipStream_t);                             \
  template void launch_tri_extract<T>(const T*, int64_t, int64_t, T*, int64_t, int, int, hipStream_t); \
  template void launch_apply_q<T>(const ApplyQItem*, int, int, hipStream_t);                     \
  template void launch_svd<T>(const SvdItem*, int, int, hipStream_t);                            \
  template void launch_lr_add_qr<T>(const LrAddDesc<T>*, int, const LrBatch<T>&, int, hipStream_t);  \
  template void launch_lr_add_svd<T>(const LrAddDesc<T>*, int, int, int, const LrBatch<T>&, hipStream_t); \
  template void launch_lr_add_core<T>(const LrAddDesc<T>*, int, const LrBatch<T>&, hipStream_t);     \
  template void launch_lr_add_apply<T>(const LrAddDesc<T>*, int, const LrBatch<T>&, int, hipStream_t); \
  template void launch_geqp3<T>(const Geqp3Item*, int, hipStream_t);                             \
  template void launch_permute_r<T>(const PermuteItem*, int, int, int, hipStream_t);             \
  template void launch_flip_r<T>(const FlipItem*, int, int, hipStream_t);                        \
  template void launch_gen<T>(const KernelSpec&, const GenItem*, int, int, int, hipStream_t);    \
  template void launch_fill_identity<T>(T*, int64_t, int64_t, int64_t, hipStream_t);

// Descriptor upload: an in-stream kernel reading the pinned staging ring (no
// copy engine, so no cross-engine wait before the launches that use it).
__global__ __launch_bounds__(256) void stage_copy_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                         size_t n16) {
  constexpr int U = 4;
  const size_t base = (size_t)blockIdx.x * (256 * U) + threadIdx.x;
  uint4 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t e = base + (size_t)u * 256;
    if (e < n16) v[u] = src[e];
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const size_t e = base + (size_t)u * 256;
    if (e < n16) dst[e] = v[u];
  }
}

__global__ __launch_bounds__(256) void snapshot_kernel(int32_t* __restrict__ dst, const int32_t* __restrict__ src,
                                                       int n, const int32_t* __restrict__ tail) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e < n)
    dst[e] = src[e];
  else if (e == n)
    dst[e] = *tail;
}

void launch_snapshot(int32_t* dst_host, const int32_t* src, int n, const int32_t* tail, hipStream_t s) {
  hipLaunchKernelGGL(snapshot_kernel, dim3((n + 256) / 256), dim3(256), 0, s, dst_host, src, n, tail);
  FRK_HIP_CHECK(hipGetLastError());
}

void launch_stage_copy(void* dst, const void* src_host, size_t n16, hipStream_t s) {
  if (n16 == 0) return;
  const size_t g = (n16 + 1023) / 1024;
  hipLaunchKernelGGL(stage_copy_kernel, dim3((unsigned)g), dim3(256), 0, s, static_cast<uint4*>(dst),
                     static_cast<const uint4*>(src_host), n16);
  FRK_HIP_CHECK(hipGetLastError());
}

FRK_INST(double)
FRK_INST(float)
template void launch_convert<float, double>(float*, const double*, int64_t, hipStream_t);
template void launch_convert<double, double>(double*, const double*, int64_t, hipStream_t);
template void launch_convert<double, float>(double*, const float*, int64_t, hipStream_t);
template void launch_convert<float, float>(float*, const float*, int64_t, hipStream_t);

}  // namespace frk

// Wave-level building blocks for the FP64 tile LU (gfx950, wave64):
//   * 16 x 16 LU and triangular inverses held in one wave's registers, with
//     DPP row_newbcast for broadcasts inside a 16-lane row and ds_bpermute
//     (the LDS crossbar, no LDS storage) for broadcasts across rows (no barrier);
//   * 16 x 16 output tiles of FP64 MFMA products (v_mfma_f64_16x16x4f64)
//     with operands in LDS;
//   * factor64: the LU of a 64 x 64 block in LDS plus the full inverses of
//     both factors, blocked by 16 (four waves, three barriers per block step).
//
// Register layout of a 16 x 16 tile in one wave: lane (r, q), r = lane & 15,
// q = lane >> 4 holds row r, columns 4q .. 4q + 3.
// MFMA 16x16x4 layouts (checked on the device): A[i][k] at lane i + 16 k,
// B[k][j] at lane j + 16 k, D[i][j] at lane j + 16 (i % 4), element i / 4.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <utility>

namespace frk {
namespace d64 {

typedef double d4 __attribute__((ext_vector_type(4)));

// (every lane of a DPP row broadcast reads a valid lane, so no old value is
// needed and each half is one v_mov_b32_dpp)
template <int CTRL>
__device__ __forceinline__ double dppd(double v) {
  return __hiloint2double(__builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true),
                          __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true));
}
// every lane receives lane K of its own 16-lane row (DPP row_newbcast)
template <int K>
__device__ __forceinline__ double row_bcast(double v) {
  return dppd<0x150 + K>(v);
}
// every lane receives the value at its own position within row S: lane
// (r, q) reads lane (r, S) through the LDS crossbar (ds_bpermute, no LDS
// storage), one permute per 32-bit half
template <int S>
__device__ __forceinline__ double xrow_bcast(double v) {
  const int addr = ((threadIdx.x & 15) + 16 * S) * 4;
  return __hiloint2double(__builtin_amdgcn_ds_bpermute(addr, __double2hiint(v)),
                          __builtin_amdgcn_ds_bpermute(addr, __double2loint(v)));
}
// 1 / x from the hardware estimate and two Newton steps (the result equals the
// rounded quotient on the inputs tested; LAPACK's getf2 scales by 1 / pivot too)
__device__ __forceinline__ double rcp2(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}

// ---------------------------------------------------------------------------
// 16 x 16 LU without pivoting (one wave, tile in registers) with both
// triangular inverses built in the same rank-1 steps: E <- M_K E (row
// operations, E ends as L^-1) and X <- X C_K (column operations reducing U
// to I, X ends as U^-1).  Step K uses U's row K, the multipliers, E's row K
// and X's column K.
// ---------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ void lu16_step(double (&a)[4], double (&e)[4], double (&x)[4], int r, int q,
                                          int* zero_row) {
  constexpr int S = K >> 2, J = K & 3;
  const double dkk = row_bcast<K>(a[J]);  // a[K][K] in the lanes of row S
  if (q == S && r == K && dkk == 0.0) *zero_row = K;
  const double rp = rcp2(dkk);            // valid in row S
  const double lb = xrow_bcast<S>(a[J] * rp);  // multiplier a[r][K] / a[K][K]
  const double xs = xrow_bcast<S>(x[J] * rp);  // X[r][K] / a[K][K]
  const double l = (r > K) ? lb : 0.0;
  double u[4], ek[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const double v = row_bcast<K>(a[j]);  // a[K][4q + j]
    u[j] = (4 * q + j > K) ? v : 0.0;
    ek[j] = row_bcast<K>(e[j]);           // E[K][4q + j]
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    a[j] -= l * u[j];
    e[j] -= l * ek[j];
    x[j] -= xs * u[j];
  }
  if (q == S) {
    if (r > K) a[J] = lb;
    x[J] = xs;
  }
}
template <int... K>
__device__ __forceinline__ void lu16_all(double (&a)[4], double (&e)[4], double (&x)[4], int r, int q,
                                         int* zero_row, std::integer_sequence<int, K...>) {
  (lu16_step<K>(a, e, x, r, q, zero_row), ...);
}

__device__ __forceinline__ void load16(double (&a)[4], const double* t, int ld) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) a[j] = t[r * ld + 4 * q + j];
}
__device__ __forceinline__ void store16(const double (&a)[4], double* t, int ld) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
#pragma unroll
  for (int j = 0; j < 4; ++j) t[r * ld + 4 * q + j] = a[j];
}

// In-place LU of the 16 x 16 LDS tile t, inverses of its factors into li and
// ui; returns the first zero pivot (0-based) or -1.
__device__ __forceinline__ int lu16(double* t, double* li, double* ui, int ld) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  double a[4], e[4], x[4];
  load16(a, t, ld);
#pragma unroll
  for (int j = 0; j < 4; ++j) e[j] = x[j] = (r == 4 * q + j) ? 1.0 : 0.0;
  int zr = 16;
  lu16_all(a, e, x, r, q, &zr, std::make_integer_sequence<int, 16>{});
  store16(a, t, ld);
  store16(e, li, ld);
  store16(x, ui, ld);
  int z = zr;
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) z = min(z, __shfl_xor(z, s));
  return z < 16 ? z : -1;
}

// ---------------------------------------------------------------------------
// MFMA tiles (one wave): C (16 x 16) = beta C + sign A (16 x K) B (K x 16),
// all in LDS, K a multiple of 4.
// ---------------------------------------------------------------------------
__device__ __forceinline__ d4 mma16(const double* A, int lda, const double* B, int ldb, int K, double sign, d4 acc) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += 4) {
    const double a = sign * A[li * lda + k0 + lk];
    const double b = B[(k0 + lk) * ldb + li];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  }
  return acc;
}
__device__ __forceinline__ d4 ld_tile(const double* C, int ldc) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  d4 acc;
#pragma unroll
  for (int v = 0; v < 4; ++v) acc[v] = C[(4 * v + lk) * ldc + li];
  return acc;
}
__device__ __forceinline__ void st_tile(const d4& acc, double* C, int ldc) {
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int v = 0; v < 4; ++v) C[(4 * v + lk) * ldc + li] = acc[v];
}

// C (64 x 64) = beta C + sign A (64 x K) B (K x 64), four waves, wave w owns
// output rows 16w .. 16w + 15.  With sync_before_store the results are held
// in registers across a workgroup barrier, so C may alias B (every wave reads
// all of B); C may always alias A (each wave reads only its own rows of A).
// Every thread of the workgroup must call it.
__device__ __forceinline__ void mma64(double* C, int ldc, const double* A, int lda, const double* B, int ldb, int K,
                                      double sign, bool accumulate, bool sync_before_store) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, li = lane & 15, lk = lane >> 4;
  d4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = accumulate ? ld_tile(C + 16 * w * ldc + 16 * t, ldc) : d4{0, 0, 0, 0};
  const double* Aw = A + (16 * w + li) * lda + lk;
  for (int k0 = 0; k0 < K; k0 += 4) {
    const double a = sign * Aw[k0];
    const double* Bk = B + (k0 + lk) * ldb + li;
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Bk[16 * t], acc[t], 0, 0, 0);
  }
  if (sync_before_store) __syncthreads();
#pragma unroll
  for (int t = 0; t < 4; ++t) st_tile(acc[t], C + 16 * w * ldc + 16 * t, ldc);
}

// ---------------------------------------------------------------------------
// LU of the 64 x 64 block M (LDS, ld) in place, blocked by 16, plus the full
// inverses of its unit-lower and upper factors into Li and Ui (LDS, ld).  The
// block form of the rank-1 scheme: at block step kb
//   panel:    L(ib,kb) = M(ib,kb) U(kb,kb)^-1,  U(kb,jb) = L(kb,kb)^-1 M(kb,jb),
//             E(kb,j) = L(kb,kb)^-1 E(kb,j) (j < kb),  X(i,kb) = X(i,kb) U(kb,kb)^-1 (i < kb)
//   trailing: M(ib,jb) -= L(ib,kb) U(kb,jb),  E(ib,j) -= L(ib,kb) E(kb,j) (j <= kb),
//             X(i,jb) -= X(i,kb) U(kb,jb) (i <= kb)
// with E = Li, X = Ui.  Only the triangles of Li / Ui are written (the
// 16 x 16 tiles above Li's / below Ui's block diagonal are left untouched).
// Returns (in every thread) the first zero pivot row, or -1.  256 threads.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int factor64(double* M, double* Li, double* Ui, int ld, int* zero_sh) {
  const int tid = threadIdx.x, w = tid >> 6;
  if (tid == 0) *zero_sh = -1;
  // the LU of the diagonal tile kb + 1 runs in wave 0 during the trailing update of
  // step kb (lookahead): wave 0 updates that tile first and factors it, waves 1-3
  // update the other tiles, which never touch it
  auto diag_lu = [&](int kb) {
    double* Mkk = M + 16 * kb * ld + 16 * kb;
    const int z = lu16(Mkk, Li + 16 * kb * ld + 16 * kb, Ui + 16 * kb * ld + 16 * kb, ld);
    if (z >= 0 && (threadIdx.x & 63) == 0 && *zero_sh < 0) *zero_sh = 16 * kb + z;
  };
  if (w == 0) diag_lu(0);
  __syncthreads();
  for (int kb = 0; kb < 4; ++kb) {
    const double* Lkk = Li + 16 * kb * ld + 16 * kb;
    const double* Ukk = Ui + 16 * kb * ld + 16 * kb;
    const int rest = 3 - kb;
    for (int p = w; p < 6; p += 4) {
      if (p < rest) {  // L(ib, kb)
        double* T = M + 16 * (kb + 1 + p) * ld + 16 * kb;
        st_tile(mma16(T, ld, Ukk, ld, 16, 1.0, d4{0, 0, 0, 0}), T, ld);
      } else if (p < 2 * rest) {  // U(kb, jb)
        double* T = M + 16 * kb * ld + 16 * (kb + 1 + p - rest);
        st_tile(mma16(Lkk, ld, T, ld, 16, 1.0, d4{0, 0, 0, 0}), T, ld);
      } else if (p < 2 * rest + kb) {  // E(kb, j), j < kb
        double* T = Li + 16 * kb * ld + 16 * (p - 2 * rest);
        st_tile(mma16(Lkk, ld, T, ld, 16, 1.0, d4{0, 0, 0, 0}), T, ld);
      } else {  // X(i, kb), i < kb
        double* T = Ui + 16 * (p - 2 * rest - kb) * ld + 16 * kb;
        st_tile(mma16(T, ld, Ukk, ld, 16, 1.0, d4{0, 0, 0, 0}), T, ld);
      }
    }
    __syncthreads();
    if (rest == 0) break;
    const int nm = rest * rest, ne = rest * (kb + 1);
    // tile 0 of the trailing list is (kb + 1, kb + 1): wave 0's, then its LU
    auto trailing = [&](int t) {
      if (t < nm) {
        const int ib = kb + 1 + t / rest, jb = kb + 1 + t % rest;
        double* C = M + 16 * ib * ld + 16 * jb;
        st_tile(mma16(M + 16 * ib * ld + 16 * kb, ld, M + 16 * kb * ld + 16 * jb, ld, 16, -1.0, ld_tile(C, ld)), C,
                ld);
      } else if (t < nm + ne) {  // E(ib, j) -= L(ib, kb) E(kb, j)
        const int u = t - nm, ib = kb + 1 + u / (kb + 1), j = u % (kb + 1);
        double* C = Li + 16 * ib * ld + 16 * j;
        const d4 c0 = (j < kb) ? ld_tile(C, ld) : d4{0, 0, 0, 0};
        st_tile(mma16(M + 16 * ib * ld + 16 * kb, ld, Li + 16 * kb * ld + 16 * j, ld, 16, -1.0, c0), C, ld);
      } else {  // X(i, jb) -= X(i, kb) U(kb, jb)
        const int u = t - nm - ne, i = u % (kb + 1), jb = kb + 1 + u / (kb + 1);
        double* C = Ui + 16 * i * ld + 16 * jb;
        const d4 c0 = (i < kb) ? ld_tile(C, ld) : d4{0, 0, 0, 0};
        st_tile(mma16(Ui + 16 * i * ld + 16 * kb, ld, M + 16 * kb * ld + 16 * jb, ld, 16, -1.0, c0), C, ld);
      }
    };
    if (w == 0) {
      trailing(0);
      diag_lu(kb + 1);
    } else {
      for (int t = w; t < nm + 2 * ne; t += 3) trailing(t);
    }
    __syncthreads();
  }
  return *zero_sh;
}

}  // namespace d64
}  // namespace frk

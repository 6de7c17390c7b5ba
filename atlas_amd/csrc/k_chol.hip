/*
 * k_chol.hip -- Cholesky of one diagonal block (n <= 128) in LDS.
 *
 * Replaces the leaves of ATLAS's recursive Cholesky: the closed-form 4x4,
 * 3x3, 2x2 kernels (src/lapack/ATL_potrfL.c:104-202, ATL_potrfU.c:98-197)
 * and the small-N recursion above them.  The whole block is held in the
 * registers of one 256-thread workgroup and factored right-looking, one
 * column at a time: rank-1 update of the trailing triangle with the
 * unscaled column over the pivot; the square roots and the column scaling
 * are applied once, when the block is written back.  A non-positive (or NaN) pivot stops the
 * factorization and records the 1-based global column in *info, as
 * ATL_potrfL.c:238-240 does; later kernels of the same Cholesky see the
 * nonzero info and do nothing.  Only the referenced triangle is read or
 * written.
 */
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace {

constexpr int CNB = ATLMX_CHOL_NB;
constexpr int NT = 256;

template <typename T>
__device__ inline T dsqrt(T x);
template <>
__device__ inline double dsqrt<double>(double x) { return sqrt(x); }
template <>
__device__ inline float dsqrt<float>(float x) { return sqrtf(x); }

/* 1/x for the rank-1 multipliers: the hardware reciprocal refined by
 * Newton steps (two for double, one for float), within an ulp or so of
 * the rounded quotient and a short dependent chain on every column step. */
__device__ inline double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}
__device__ inline float rcp_nr(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

/* The block lives in registers: thread (ti, tj) of a 16 x 16 grid holds the
 * 8 x 8 tile of rows ti*8.., columns tj*8.. (tiles above the diagonal stay
 * idle).  Step k: the owners of column k publish it as it stands (below the
 * pivot, zeros above) and the pivot d_k through LDS, and every tile right
 * of column k subtracts a(i,k) * a(j,k) / d_k in registers.  Column k stays
 * unscaled: L(i,k) = a(i,k) / sqrt(d_k) and L(k,k) = sqrt(d_k) are formed
 * for the whole block at write-back, so no square root or division by it
 * sits on the step-to-step chain.  One barrier per step: the column buffers
 * alternate with the parity of k, so a step's writes never meet the
 * previous step's reads, and every thread reads the same d_k, so a failed
 * pivot stops all of them at once.  The eight steps of a tile column are
 * unrolled, so every register index is fixed at compile time. */
template <typename T>
__global__ __launch_bounds__(NT) void potf2_kernel(int lower, int n, T *A, long lda,
                                                   int info_base, int *info) {
  if (*info != 0) return;
  __shared__ T col[2][CNB];
  __shared__ T piv[CNB];
  const int tid = threadIdx.x;
  const int ti = tid >> 4, tj = tid & 15;
  const bool own = ti >= tj;
  T a[8][8];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int i = ti * 8 + r, j = tj * 8 + c;
      a[r][c] = (own && i < n && j < n && i >= j)
                    ? (lower ? A[(size_t)j * lda + i] : A[(size_t)i * lda + j])
                    : T(0);
    }
  int done = n;
  const int nkt = (n + 7) / 8;
  for (int kt = 0; kt < nkt && done == n; ++kt) {
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) {
      const int k = kt * 8 + kc;
      if (k >= n) break;
      T *ck = col[k & 1];
      if (tj == kt && own) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const int i = ti * 8 + r;
          ck[i] = (i > k && i < n) ? a[r][kc] : T(0);
        }
        if (ti == kt) piv[k] = a[kc][kc];
      }
      __syncthreads();
      const T d = piv[k];
      if (!__builtin_amdgcn_readfirstlane((int)(d > T(0)))) {
        if (tid == 0) atomicCAS(info, 0, info_base + k + 1);
        done = k;
        break;
      }
      if (own && tj >= kt) {
        const T rd = rcp_nr(d);
        T x[8], y[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) x[r] = ck[ti * 8 + r];
#pragma unroll
        for (int c = 0; c < 8; ++c) y[c] = ck[tj * 8 + c] * rd;
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
          for (int c = 0; c < 8; ++c) a[r][c] -= x[r] * y[c];
      }
    }
  }
  /* write back the factored columns (the whole block when it succeeded; on
   * failure the columns before the failing one, as the reference leaves
   * them factored), each scaled by its pivot's square root */
  if (own) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int j = tj * 8 + c;
      if (j >= done) continue;
      const T sd = dsqrt(piv[j]), rs = T(1) / sd;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int i = ti * 8 + r;
        if (i < n && i >= j) {
          const T v = (i == j) ? sd : a[r][c] * rs;
          if (lower) A[(size_t)j * lda + i] = v;
          else A[(size_t)i * lda + j] = v;
        }
      }
    }
  }
}

template <typename T>
hipError_t launch_potf2(int uplo, int n, T *A, long lda, int info_base, int *info,
                        hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n > CNB || lda < n || !A || !info) return hipErrorInvalidValue;
  hipLaunchKernelGGL(potf2_kernel<T>, dim3(1), dim3(NT), 0, s, uplo == KX_LOWER, n, A, lda,
                     info_base, info);
  return hipGetLastError();
}

}  // namespace

extern "C" hipError_t atlmx_k_dpotf2(int uplo, int n, double *A, long lda, int info_base,
                                     int *info, hipStream_t s) {
  return launch_potf2<double>(uplo, n, A, lda, info_base, info, s);
}
extern "C" hipError_t atlmx_k_spotf2(int uplo, int n, float *A, long lda, int info_base,
                                     int *info, hipStream_t s) {
  return launch_potf2<float>(uplo, n, A, lda, info_base, info, s);
}

/*
 * k_chol.hip -- Cholesky of one diagonal block (n <= 128) in LDS.
 *
 * Replaces the leaves of ATLAS's recursive Cholesky: the closed-form 4x4,
 * 3x3, 2x2 kernels (src/lapack/ATL_potrfL.c:104-202, ATL_potrfU.c:98-197)
 * and the small-N recursion above them.  The whole block is held in LDS by
 * one 256-thread workgroup and factored right-looking, one column at a
 * time: sqrt of the pivot, scale the column by its reciprocal, rank-1
 * update of the trailing triangle.  A non-positive (or NaN) pivot stops the
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

template <typename T>
__global__ __launch_bounds__(NT) void potf2_kernel(int lower, int n, T *A, long lda,
                                                   int info_base, int *info) {
  if (*info != 0) return;
  /* Lt[i][j] holds the lower-triangular view: L = A (lower) or U^T (upper) */
  __shared__ T Lt[CNB][CNB + 1];
  __shared__ int s_stop;
  const int tid = threadIdx.x;
  for (int idx = tid; idx < n * n; idx += NT) {
    const int j = idx / n, i = idx - j * n;
    if (i >= j) Lt[i][j] = lower ? A[(size_t)j * lda + i] : A[(size_t)i * lda + j];
  }
  if (tid == 0) s_stop = 0;
  __syncthreads();
  int done = n;
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      const T d = Lt[k][k];
      if (!(d > T(0))) {
        s_stop = 1;
        atomicCAS(info, 0, info_base + k + 1);
      } else {
        Lt[k][k] = dsqrt(d);
      }
    }
    __syncthreads();
    if (s_stop) { done = k; break; }
    const T rd = T(1) / Lt[k][k];
    for (int i = k + 1 + tid; i < n; i += NT) Lt[i][k] *= rd;
    __syncthreads();
    /* trailing update of the lower triangle, columns k+1..n-1 */
    const int m = n - k - 1;
    for (int idx = tid; idx < m * m; idx += NT) {
      const int jj = idx / m, ii = idx - jj * m;
      if (ii >= jj) {
        const int i = k + 1 + ii, j = k + 1 + jj;
        Lt[i][j] -= Lt[i][k] * Lt[j][k];
      }
    }
    __syncthreads();
  }
  /* write back the factored columns (the whole block when it succeeded; on
   * failure the columns before the failing one, as the reference leaves
   * them factored) */
  const int wcols = done;
  for (int idx = tid; idx < n * n; idx += NT) {
    const int j = idx / n, i = idx - j * n;
    if (i >= j && j < wcols) {
      if (lower) A[(size_t)j * lda + i] = Lt[i][j];
      else A[(size_t)i * lda + j] = Lt[i][j];
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

/*
 * k_aux.hip -- the small kernels around the GEMM: triangular diagonal-block
 * inversion (TRSM), 2-D scaling, iamax, fills.
 */
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>
#include "kernels.h"

namespace {

constexpr int TNB = ATLMX_TRI_NB;

/* ------------------------------------------------------------------------
 * Diagonal-block inversion for TRSM.  ATLAS's TRSM nano-kernels store the
 * reciprocal of each diagonal element and multiply by it
 * (tune/blas/gemm/trsm-nano.base:225,297); here the whole 64x64 diagonal
 * block is inverted once, so the TRSM becomes a chain of MFMA GEMMs
 * (X_k = inv(A_kk) * B_k, then the trailing update).
 *
 * One wavefront per 64x64 block, four blocks per workgroup.  Lane j builds
 * column j of the inverse by substitution with the column held in
 * registers (the loops are fully unrolled); the triangle is read from LDS
 * as a broadcast.  Lanes j > i compute x_i = 0 automatically, so all lanes
 * run the same instruction stream.
 * ---------------------------------------------------------------------- */
template <typename T>
__global__ __launch_bounds__(256) void tri_inv_kernel(int lower, int unit, int n,
                                                      const T *__restrict__ A, long lda,
                                                      T *__restrict__ inv,
                                                      const int *abort_flag) {
  if (abort_flag && *abort_flag) return;
  __shared__ T L[4][TNB][TNB + 1];
  const int wv = threadIdx.x >> 6, j = threadIdx.x & 63;
  const int nblk = (n + TNB - 1) / TNB;
  const int blk = blockIdx.x * 4 + wv;
  const int b0 = blk * TNB;
  const int nb = (blk < nblk) ? min(TNB, n - b0) : 0;
  /* coalesced load: lane = row */
  for (int c = 0; c < TNB; ++c) {
    const int i = j;
    T v = T(0);
    const bool in = lower ? (i >= c) : (i <= c);
    if (i < nb && c < nb) {
      if (in) v = A[(size_t)(b0 + c) * lda + (b0 + i)];
      if (i == c && unit) v = T(1);
    } else if (i == c) {
      v = T(1);
    }
    L[wv][i][c] = in ? v : T(0);
  }
  __syncthreads();
  if (blk >= nblk) return;
  T x[TNB];
  if (lower) {
#pragma unroll
    for (int i = 0; i < TNB; ++i) {
      T s = (i == j) ? T(1) : T(0);
#pragma unroll
      for (int k = 0; k < i; ++k) s -= L[wv][i][k] * x[k];
      x[i] = s / L[wv][i][i];
    }
  } else {
#pragma unroll
    for (int i = TNB - 1; i >= 0; --i) {
      T s = (i == j) ? T(1) : T(0);
#pragma unroll
      for (int k = i + 1; k < TNB; ++k) s -= L[wv][i][k] * x[k];
      x[i] = s / L[wv][i][i];
    }
  }
  T *out = inv + (size_t)blk * TNB * TNB + (size_t)j * TNB;
#pragma unroll
  for (int i = 0; i < TNB; ++i) out[i] = x[i];
}

template <typename T>
__global__ void scal2d_kernel(int M, int N, T alpha, T *B, long ldb) {
  const long total = (long)M * N;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long j = idx / M, i = idx - j * M;
    T *p = B + (size_t)j * ldb + i;
    *p = (alpha == T(0)) ? T(0) : *p * alpha;
  }
}

/* (value, index) arg-max with cblas_i?amax semantics: larger |x| wins, ties
 * go to the lower index, NaN never wins (its |x| is mapped to -1). */
template <typename T>
__device__ inline void amax_combine(T &v, int &i, T v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

template <typename T>
__global__ __launch_bounds__(1024) void iamax_kernel(int n, const T *x, int incx, int *out) {
  __shared__ T sv[16];
  __shared__ int si[16];
  T bv = T(-1);
  int bi = 0x7fffffff;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    T a = x[(size_t)i * incx];
    a = a < T(0) ? -a : a;
    if (!(a == a)) a = T(-1);
    amax_combine(bv, bi, a, i);
  }
  for (int off = 32; off > 0; off >>= 1) {
    const T v2 = __shfl_xor(bv, off);
    const int i2 = __shfl_xor(bi, off);
    amax_combine(bv, bi, v2, i2);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sv[w] = bv; si[w] = bi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    T v = sv[0];
    int i = si[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) amax_combine(v, i, sv[k], si[k]);
    *out = (v > T(0)) ? i : 0;   /* all zero / NaN: index 0, like iamax */
  }
}

__global__ void fill_i32_kernel(int *p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
hipError_t launch_tri_inv(int uplo, int diag, int n, const T *A, long lda, T *inv,
                          const int *abort_flag, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (lda < n || !A || !inv) return hipErrorInvalidValue;
  const int nblk = (n + TNB - 1) / TNB;
  hipLaunchKernelGGL(tri_inv_kernel<T>, dim3((nblk + 3) / 4), dim3(256), 0, s,
                     uplo == KX_LOWER, diag == KX_UNIT, n, A, lda, inv, abort_flag);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scal2d(int M, int N, T alpha, T *B, long ldb, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (ldb < M) return hipErrorInvalidValue;
  long total = (long)M * N;
  long blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(scal2d_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, M, N,
                     alpha, B, ldb);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_iamax(int n, const T *x, int incx, int *out, hipStream_t s) {
  if (n <= 0 || incx <= 0) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, s, out, 1, 0);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(iamax_kernel<T>, dim3(1), dim3(1024), 0, s, n, x, incx, out);
  return hipGetLastError();
}

}  // namespace

extern "C" {

hipError_t atlmx_k_dtri_inv(int uplo, int diag, int n, const double *A, long lda,
                            double *inv, const int *abort_flag, hipStream_t s) {
  return launch_tri_inv<double>(uplo, diag, n, A, lda, inv, abort_flag, s);
}
hipError_t atlmx_k_stri_inv(int uplo, int diag, int n, const float *A, long lda,
                            float *inv, const int *abort_flag, hipStream_t s) {
  return launch_tri_inv<float>(uplo, diag, n, A, lda, inv, abort_flag, s);
}
hipError_t atlmx_k_dscal2d(int M, int N, double alpha, double *B, long ldb, hipStream_t s) {
  return launch_scal2d<double>(M, N, alpha, B, ldb, s);
}
hipError_t atlmx_k_sscal2d(int M, int N, float alpha, float *B, long ldb, hipStream_t s) {
  return launch_scal2d<float>(M, N, alpha, B, ldb, s);
}
hipError_t atlmx_k_idamax(int n, const double *x, int incx, int *out, hipStream_t s) {
  return launch_iamax<double>(n, x, incx, out, s);
}
hipError_t atlmx_k_isamax(int n, const float *x, int incx, int *out, hipStream_t s) {
  return launch_iamax<float>(n, x, incx, out, s);
}
hipError_t atlmx_k_fill_i32(int *p, int n, int v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // extern "C"

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
 * One 256-thread workgroup per 64x64 block.  An upper block is inverted as
 * the transpose of the inverse of its (lower) transpose.  The four 16x16
 * diagonal sub-blocks are inverted by substitution (lane = column), then
 * two doubling levels fill in the off-diagonal parts,
 *     inv([A 0; C B]) = [inv(A) 0; -inv(B) C inv(A) inv(B)],
 * as two small products each, spread over all threads.  LDS holds the
 * factor's strict lower triangle, the inverse's strict lower triangle
 * transposed into the strict upper half of the same array, the inverse's
 * diagonal, and the product's intermediate; every product loop runs over
 * the nonzero (triangular) range only.
 * ---------------------------------------------------------------------- */
template <typename T>
__global__ __launch_bounds__(256) void tri_inv_kernel(int lower, int unit, int n,
                                                      const T *__restrict__ A, long lda,
                                                      T *__restrict__ inv,
                                                      const int *abort_flag) {
  static_assert(TNB == 64, "tri_inv_kernel is written for 64x64 blocks");
  if (abort_flag && *abort_flag) return;
  __shared__ T Ls[TNB][TNB + 1];   /* (i > j): L(i,j) at [i][j], X(i,j) at [j][i] */
  __shared__ T dL[TNB], dX[TNB];
  __shared__ T Tb[32][33];
  const int tid = threadIdx.x;
  const int blk = blockIdx.x;
  const int b0 = blk * TNB;
  const int nb = min(TNB, n - b0);
  /* load: the block as the lower triangle L (L = U^T for an upper block);
   * outside the n x n matrix the block is the identity */
  for (int e = tid; e < TNB * TNB; e += 256) {
    const int r = e & (TNB - 1), c = e >> 6;           /* element (r, c) of the block */
    const int i = lower ? r : c, j = lower ? c : r;    /* its place in L */
    if (i < j) continue;
    T v;
    if (r < nb && c < nb) v = A[(size_t)(b0 + c) * lda + (b0 + r)];
    else v = (i == j) ? T(1) : T(0);
    if (i == j) dL[i] = (unit && i < nb) ? T(1) : v;
    else Ls[i][j] = v;
  }
  __syncthreads();
  /* the 16x16 diagonal sub-blocks, by substitution: lane (q, c) builds
   * column c of inv(L_qq) */
  if (tid < 64) {
    const int q = tid >> 4, c = tid & 15, o = 16 * q;
    T x[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      T s = (i == c) ? T(1) : T(0);
#pragma unroll
      for (int k = 0; k < i; ++k) s -= Ls[o + i][o + k] * x[k];   /* x[k] = 0 for k < c */
      x[i] = (i >= c) ? s / dL[o + i] : T(0);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      if (i == c) dX[o + c] = x[i];
      else if (i > c) Ls[o + c][o + i] = x[i];
    }
  }
  __syncthreads();
  /* doubling: blocks of 2h from two blocks of h */
#pragma unroll
  for (int h = 16; h < TNB; h *= 2) {
    const int nout = (TNB / (2 * h)) * h * h;   /* 512, then 1024 */
    /* T = L21 * X11, L21 = L(r0.., c0..), X11 = X(c0.., c0..) lower */
    for (int e = tid; e < nout; e += 256) {
      const int p = e / (h * h), i = (e / h) % h, j = e % h;
      const int c0 = 2 * h * p, r0 = c0 + h;
      T s = Ls[r0 + i][c0 + j] * dX[c0 + j];
      for (int k = j + 1; k < h; ++k) s += Ls[r0 + i][c0 + k] * Ls[c0 + j][c0 + k];
      Tb[p * h + i][j] = s;
    }
    __syncthreads();
    /* X21 = -X22 * T, X22 = X(r0.., r0..) lower */
    for (int e = tid; e < nout; e += 256) {
      const int p = e / (h * h), i = (e / h) % h, j = e % h;
      const int c0 = 2 * h * p, r0 = c0 + h;
      T s = dX[r0 + i] * Tb[p * h + i][j];
      for (int k = 0; k < i; ++k) s += Ls[r0 + k][r0 + i] * Tb[p * h + k][j];
      Ls[c0 + j][r0 + i] = -s;
    }
    __syncthreads();
  }
  /* store inv(A) column-major with leading dimension 64, zeros outside
   * the triangle */
  T *out = inv + (size_t)blk * TNB * TNB;
  for (int e = tid; e < TNB * TNB; e += 256) {
    const int r = e & (TNB - 1), c = e >> 6;
    const int i = lower ? r : c, j = lower ? c : r;    /* inv(U) = inv(L)^T */
    T v = T(0);
    if (i == j) v = dX[i];
    else if (i > j) v = Ls[j][i];
    out[(size_t)c * TNB + r] = v;
  }
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

/* ------------------------------------------------------------------------
 * Triangle copies for the routines that run on the dense GEMM (TRMM, SYMM,
 * LAUUM) or on a dense workspace (TRTRI, GETRI).  One thread per element of
 * the n x n (or m x n) destination, 64-wide column strips per workgroup row
 * so the column-major accesses coalesce.
 *   mode TC_EXPAND_TRI: W = triangle of A, zero outside (unit diag -> 1)
 *   mode TC_EXPAND_SYM: W = the symmetric matrix held in A's triangle
 *   mode TC_STORE_TRI : A's triangle = W's (diagonal left alone if unit)
 *   mode TC_IDENTITY  : W = I
 *   mode TC_COPY      : A = W (m x n)
 *   mode TC_TRANSPOSE : W (n x m) = A^T (A m x n)
 * ---------------------------------------------------------------------- */
template <typename T>
__global__ __launch_bounds__(256) void tricopy_kernel(int mode, int lower, int unit, int m,
                                                      int n, T *A, long lda, T *W, long ldw) {
  const int i = blockIdx.y * 64 + (threadIdx.x & 63);
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= m || j >= n) return;
  const bool in = lower ? (i >= j) : (i <= j);
  switch (mode) {
    case ATLMX_TC_EXPAND_TRI: {
      T v = in ? A[(size_t)j * lda + i] : T(0);
      if (i == j && unit) v = T(1);
      W[(size_t)j * ldw + i] = v;
      break;
    }
    case ATLMX_TC_EXPAND_SYM:
      W[(size_t)j * ldw + i] = in ? A[(size_t)j * lda + i] : A[(size_t)i * lda + j];
      break;
    case ATLMX_TC_STORE_TRI:
      if (in && !(i == j && unit)) A[(size_t)j * lda + i] = W[(size_t)j * ldw + i];
      break;
    case ATLMX_TC_IDENTITY:
      W[(size_t)j * ldw + i] = (i == j) ? T(1) : T(0);
      break;
    case ATLMX_TC_TRANSPOSE:
      W[(size_t)i * ldw + j] = A[(size_t)j * lda + i];
      break;
    default:
      A[(size_t)j * lda + i] = W[(size_t)j * ldw + i];
      break;
  }
}

/* *info = 1 + the first i < n with A(i,i) == 0, else 0 (info starts at
 * INT_MAX; a second pass turns INT_MAX into 0) */
template <typename T>
__global__ __launch_bounds__(256) void diag_zero_kernel(int n, const T *A, long lda, int *info) {
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    if (A[(size_t)i * lda + i] == T(0)) atomicMin(info, i + 1);
}

__global__ void info_fix_kernel(int *info) {
  if (*info == 0x7fffffff) *info = 0;
}

template <typename T>
hipError_t launch_tricopy(int mode, int uplo, int diag, int m, int n, T *A, long lda, T *W,
                          long ldw, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (mode < ATLMX_TC_EXPAND_TRI || mode > ATLMX_TC_TRANSPOSE) return hipErrorInvalidValue;
  if (mode <= ATLMX_TC_STORE_TRI && m != n) return hipErrorInvalidValue;
  if ((mode != ATLMX_TC_IDENTITY && (!A || lda < m)) || !W ||
      ldw < (mode == ATLMX_TC_TRANSPOSE ? n : m))
    return hipErrorInvalidValue;
  const int gy = (m + 63) / 64;            /* row strips (grid y), column quads (grid x) */
  if (gy > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(tricopy_kernel<T>, dim3((n + 3) / 4, gy), dim3(256), 0, s,
                     mode, uplo == KX_LOWER, diag == KX_UNIT, m, n, A, lda, W, ldw);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_diag_zero(int n, const T *A, long lda, int *info, hipStream_t s) {
  hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, s, info, 1, 0x7fffffff);
  if (n > 0) {
    int g = (n + 255) / 256;
    if (g > 1024) g = 1024;
    hipLaunchKernelGGL(diag_zero_kernel<T>, dim3(g), dim3(256), 0, s, n, A, lda, info);
  }
  hipLaunchKernelGGL(info_fix_kernel, dim3(1), dim3(1), 0, s, info);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_tri_inv(int uplo, int diag, int n, const T *A, long lda, T *inv,
                          const int *abort_flag, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (lda < n || !A || !inv) return hipErrorInvalidValue;
  const int nblk = (n + TNB - 1) / TNB;
  hipLaunchKernelGGL(tri_inv_kernel<T>, dim3(nblk), dim3(256), 0, s,
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
hipError_t atlmx_k_dtricopy(int mode, int uplo, int diag, int m, int n, double *A, long lda,
                            double *W, long ldw, hipStream_t s) {
  return launch_tricopy<double>(mode, uplo, diag, m, n, A, lda, W, ldw, s);
}
hipError_t atlmx_k_stricopy(int mode, int uplo, int diag, int m, int n, float *A, long lda,
                            float *W, long ldw, hipStream_t s) {
  return launch_tricopy<float>(mode, uplo, diag, m, n, A, lda, W, ldw, s);
}
hipError_t atlmx_k_ddiag_zero(int n, const double *A, long lda, int *info, hipStream_t s) {
  return launch_diag_zero<double>(n, A, lda, info, s);
}
hipError_t atlmx_k_sdiag_zero(int n, const float *A, long lda, int *info, hipStream_t s) {
  return launch_diag_zero<float>(n, A, lda, info, s);
}
hipError_t atlmx_k_fill_i32(int *p, int n, int v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // extern "C"

/*
 * k_mp.hip -- memory-bound kernels of the mixed-precision solver: precision
 * conversion, residual GEMV in FP64 (R = B - A*X for a few right-hand
 * sides), infinity norms, and the solution update.  None has an ATLAS
 * counterpart: ATLAS has no iterative-refinement solver.
 *
 * All are HBM-bound: one pass over the n x n operand at most.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "kernels.h"
#include "kernels_mp.h"

namespace {

template <typename TO, typename TI>
__global__ void convert_kernel(int M, int N, const TI *__restrict__ src, long lds,
                               TO *__restrict__ dst, long ldd) {
  const long total = (long)M * N;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long j = idx / M, i = idx - j * M;
    dst[(size_t)j * ldd + i] = (TO)src[(size_t)j * lds + i];
  }
}

/* R(:, 0:nrhs) = B - A*X, column-major, nrhs <= MP_MAXRHS.  Grid:
 * (ceil(M/512), ksplit); each thread takes two consecutive rows (one
 * 16-byte load per column when A's columns are 16-byte aligned) over its
 * block's K slice, with eight columns' loads in flight, and writes the two
 * partial sums to part[ks][nrhs][M]; a second pass adds the slices in fixed
 * order, so the result is deterministic.  NR = 1 is the one-right-hand-side
 * case without per-column predicates. */
constexpr int MAXRHS = MP_MAXRHS;
template <int NR>
__global__ __launch_bounds__(256) void gemv_part_kernel(int M, int K, int nrhs, int kchunk,
                                                        const double *__restrict__ A, long lda,
                                                        const double *__restrict__ X, long ldx,
                                                        double *__restrict__ part) {
  const int i = (blockIdx.x * 256 + threadIdx.x) * 2;
  const int ks = blockIdx.y;
  const int k0 = ks * kchunk, k1 = min(K, k0 + kchunk);
  const bool vec = ((lda & 1) == 0) && (((uintptr_t)A & 15) == 0) && i + 1 < M;
  __shared__ double xs[64][NR];
  double acc0[NR], acc1[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc0[r] = acc1[r] = 0.0;
  for (int kb = k0; kb < k1; kb += 64) {
    const int kn = min(64, k1 - kb);
    __syncthreads();
    for (int t = threadIdx.x; t < 64 * NR; t += 256) {
      const int kk = t / NR, r = t - kk * NR;
      xs[kk][r] = (kk < kn && r < nrhs) ? X[(size_t)r * ldx + kb + kk] : 0.0;
    }
    __syncthreads();
    if (i >= M) continue;
    const double *p = A + (size_t)kb * lda + i;
    if (vec && kn == 64) {
#pragma unroll 8
      for (int kk = 0; kk < 64; ++kk) {
        const double2 a = *reinterpret_cast<const double2 *>(p + (size_t)kk * lda);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          acc0[r] += a.x * xs[kk][r];
          acc1[r] += a.y * xs[kk][r];
        }
      }
    } else {
      for (int kk = 0; kk < kn; ++kk) {
        const double a0 = p[(size_t)kk * lda];
        const double a1 = (i + 1 < M) ? p[(size_t)kk * lda + 1] : 0.0;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          acc0[r] += a0 * xs[kk][r];
          acc1[r] += a1 * xs[kk][r];
        }
      }
    }
  }
  if (i < M)
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < nrhs) {
        part[((size_t)ks * nrhs + r) * M + i] = acc0[r];
        if (i + 1 < M) part[((size_t)ks * nrhs + r) * M + i + 1] = acc1[r];
      }
}

__global__ void gemv_reduce_kernel(int M, int nrhs, int ksplit, const double *__restrict__ part,
                                   const double *__restrict__ B, long ldb, double *R,
                                   long ldr) {
  const long total = (long)M * nrhs;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const int r = (int)(idx / M), i = (int)(idx - (long)r * M);
    double s = 0.0;
    for (int ks = 0; ks < ksplit; ++ks) s += part[((size_t)ks * nrhs + r) * M + i];
    R[(size_t)r * ldr + i] = B[(size_t)r * ldb + i] - s;
  }
}

/* out[c] = max_i |X(i,c)| for c < N (one block per column) */
__global__ __launch_bounds__(256) void colmax_kernel(int M, const double *X, long ldx,
                                                     double *out) {
  const int c = blockIdx.x;
  double m = 0.0;
  for (int i = threadIdx.x; i < M; i += 256) {
    double v = fabs(X[(size_t)c * ldx + i]);
    if (!(v <= m)) m = v;   /* NaN propagates */
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(m, off);
    if (!(o <= m)) m = o;
  }
  __shared__ double s[4];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = s[0];
    for (int k = 1; k < 4; ++k) if (!(s[k] <= v)) v = s[k];
    out[c] = v;
  }
}

/* row abs-sums of A: rowsum[i] = sum_j |A(i,j)| (||A||_inf = max_i).
 * Grid (ceil(M/256), MP_KSPLIT): column slices summed into part[ks][M],
 * then added in fixed order. */
__global__ __launch_bounds__(256) void rowsum_part_kernel(int M, int N, int kchunk,
                                                          const double *A, long lda,
                                                          double *part) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int ks = blockIdx.y;
  if (i >= M) return;
  const int j0 = ks * kchunk, j1 = min(N, j0 + kchunk);
  double s = 0.0;
  for (int j = j0; j < j1; ++j) s += fabs(A[(size_t)j * lda + i]);
  part[(size_t)ks * M + i] = s;
}

__global__ void rowsum_reduce_kernel(int M, int ksplit, const double *part, double *rowsum) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  double s = 0.0;
  for (int ks = 0; ks < ksplit; ++ks) s += part[(size_t)ks * M + i];
  rowsum[i] = s;
}

__global__ void axpy_conv_kernel(int M, int N, const float *__restrict__ D, long ldd,
                                 double *__restrict__ X, long ldx) {
  const long total = (long)M * N;
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < total;
       idx += (long)gridDim.x * blockDim.x) {
    const long j = idx / M, i = idx - j * M;
    X[(size_t)j * ldx + i] += (double)D[(size_t)j * ldd + i];
  }
}

unsigned grid_for(long total) {
  long b = (total + 255) / 256;
  if (b > 16384) b = 16384;
  if (b < 1) b = 1;
  return (unsigned)b;
}

/* scale[j] = 2^-e with mx[j] = f 2^e, f in [0.5, 1): mx[j] scale[j] in
 * [0.5, 1), exactly (1 where mx[j] is zero or not finite) */
__global__ void pow2_scale_kernel(int n, const double *mx, double *scale) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const double m = mx[j];
  int e = 0;
  if (m > 0.0 && m <= 1.7976931348623157e308) frexp(m, &e);
  scale[j] = ldexp(1.0, -e);
}

/* X(i, :) (+)= scale[i] * D(i, :), D in FP32 */
__global__ void scale_conv_kernel(int M, int N, const float *D, long ldd, const double *scale,
                                  double *X, long ldx, int accumulate) {
  const long total = (long)M * N;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const long j = t / M, i = t - j * M;
    const double v = scale[i] * (double)D[(size_t)j * ldd + i];
    double *x = X + (size_t)j * ldx + i;
    *x = accumulate ? *x + v : v;
  }
}

/* Dot products out[j] = sum_i V(i, j) w(i), j < kc (the Gram-Schmidt
 * coefficients of GMRES).  Grid (ceil(M / GT_ROWS), ceil(kc / GT_CH)): each
 * block sums GT_CH columns over its GT_ROWS rows (w read once per block and
 * kept in a register across the columns), reduces across its 256 threads
 * and writes part[j][bx]; a second pass adds the row blocks in fixed order,
 * so the result is deterministic. */
constexpr int GT_ROWS = 1024;
constexpr int GT_CH = 8;
__global__ __launch_bounds__(256) void gemv_t_part_kernel(int M, int kc, int nbx,
                                                          const double *__restrict__ V,
                                                          long ldv,
                                                          const double *__restrict__ w,
                                                          double *__restrict__ part) {
  const int bx = blockIdx.x, j0 = blockIdx.y * GT_CH;
  const int r0 = bx * GT_ROWS, r1 = min(M, r0 + GT_ROWS);
  double acc[GT_CH];
#pragma unroll
  for (int c = 0; c < GT_CH; ++c) acc[c] = 0.0;
  for (int i = r0 + (int)threadIdx.x; i < r1; i += 256) {
    const double wi = w[i];
#pragma unroll
    for (int c = 0; c < GT_CH; ++c)
      if (j0 + c < kc) acc[c] += V[(size_t)(j0 + c) * ldv + i] * wi;
  }
  __shared__ double red[GT_CH][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < GT_CH; ++c) {
    double v = acc[c];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (lane == 0) red[c][wave] = v;
  }
  __syncthreads();
  const int c = threadIdx.x;
  if (c < GT_CH && j0 + c < kc)
    part[(size_t)(j0 + c) * nbx + bx] = (red[c][0] + red[c][1]) + (red[c][2] + red[c][3]);
}

__global__ void gemv_t_reduce_kernel(int kc, int nbx, const double *__restrict__ part,
                                     double *__restrict__ out) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= kc) return;
  double s = 0.0;
  for (int b = 0; b < nbx; ++b) s += part[(size_t)j * nbx + b];
  out[j] = s;
}

}  // namespace

extern "C" {

size_t atlmx_k_dots_ws_bytes(int M, int kc) {
  const size_t nbx = (size_t)(M > 0 ? (M + GT_ROWS - 1) / GT_ROWS : 1);
  return nbx * (size_t)(kc > 0 ? kc : 1) * sizeof(double) + 256;
}

hipError_t atlmx_k_ddots(int M, int kc, const double *V, long ldv, const double *w,
                         double *out, void *ws, hipStream_t s) {
  if (kc <= 0) return hipSuccess;
  if (M < 0 || !ws || (M > 0 && ldv < M)) return hipErrorInvalidValue;
  const int nbx = M > 0 ? (M + GT_ROWS - 1) / GT_ROWS : 0;
  double *part = (double *)ws;
  if (nbx > 0) {
    hipLaunchKernelGGL(gemv_t_part_kernel, dim3(nbx, (kc + GT_CH - 1) / GT_CH), dim3(256), 0, s,
                       M, kc, nbx, V, ldv, w, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(gemv_t_reduce_kernel, dim3((kc + 255) / 256), dim3(256), 0, s, kc, nbx,
                     part, out);
  return hipGetLastError();
}

hipError_t atlmx_k_pow2_scale(int n, const double *mx, double *scale, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(pow2_scale_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, mx, scale);
  return hipGetLastError();
}

hipError_t atlmx_k_scale_s2d(int M, int N, const float *D, long ldd, const double *scale,
                             double *X, long ldx, int accumulate, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL(scale_conv_kernel, dim3(grid_for((long)M * N)), dim3(256), 0, s, M, N, D,
                     ldd, scale, X, ldx, accumulate);
  return hipGetLastError();
}


hipError_t atlmx_k_d2s(int M, int N, const double *A, long lda, float *S, long lds,
                       hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL((convert_kernel<float, double>), dim3(grid_for((long)M * N)), dim3(256),
                     0, s, M, N, A, lda, S, lds);
  return hipGetLastError();
}

hipError_t atlmx_k_s2d(int M, int N, const float *S, long lds, double *A, long lda,
                       hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL((convert_kernel<double, float>), dim3(grid_for((long)M * N)), dim3(256),
                     0, s, M, N, S, lds, A, lda);
  return hipGetLastError();
}

size_t atlmx_k_resid_ws_bytes(int M, int nrhs) {
  return (size_t)MP_KSPLIT * (size_t)nrhs * (size_t)M * sizeof(double) + 256;
}

hipError_t atlmx_k_dresid(int M, int K, int nrhs, const double *A, long lda, const double *X,
                          long ldx, const double *B, long ldb, double *R, long ldr, void *ws,
                          hipStream_t s) {
  if (M <= 0 || nrhs <= 0) return hipSuccess;
  if (nrhs > MAXRHS || !ws) return hipErrorInvalidValue;
  const int kchunk = (K + MP_KSPLIT - 1) / MP_KSPLIT;
  const int ksplit = K > 0 ? (K + kchunk - 1) / kchunk : 1;
  double *part = (double *)ws;
  const dim3 grid((M + 511) / 512, ksplit);
  if (nrhs == 1)
    hipLaunchKernelGGL(gemv_part_kernel<1>, grid, dim3(256), 0, s, M, K, nrhs,
                       kchunk > 0 ? kchunk : 1, A, lda, X, ldx, part);
  else
    hipLaunchKernelGGL(gemv_part_kernel<MAXRHS>, grid, dim3(256), 0, s, M, K, nrhs,
                       kchunk > 0 ? kchunk : 1, A, lda, X, ldx, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(gemv_reduce_kernel, dim3(grid_for((long)M * nrhs)), dim3(256), 0, s, M,
                     nrhs, K > 0 ? ksplit : 0, part, B, ldb, R, ldr);
  return hipGetLastError();
}

hipError_t atlmx_k_dcolmax(int M, int N, const double *X, long ldx, double *out,
                           hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL(colmax_kernel, dim3(N), dim3(256), 0, s, M, X, ldx, out);
  return hipGetLastError();
}

hipError_t atlmx_k_drowsum(int M, int N, const double *A, long lda, double *rowsum,
                           void *ws, hipStream_t s) {
  if (M <= 0) return hipSuccess;
  if (!ws) return hipErrorInvalidValue;
  const int kchunk = N > 0 ? (N + MP_KSPLIT - 1) / MP_KSPLIT : 1;
  const int ksplit = N > 0 ? (N + kchunk - 1) / kchunk : 0;
  double *part = (double *)ws;
  if (ksplit > 0) {
    hipLaunchKernelGGL(rowsum_part_kernel, dim3((M + 255) / 256, ksplit), dim3(256), 0, s, M,
                       N, kchunk, A, lda, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(rowsum_reduce_kernel, dim3((M + 255) / 256), dim3(256), 0, s, M, ksplit,
                     part, rowsum);
  return hipGetLastError();
}

hipError_t atlmx_k_axpy_s2d(int M, int N, const float *D, long ldd, double *X, long ldx,
                            hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  hipLaunchKernelGGL(axpy_conv_kernel, dim3(grid_for((long)M * N)), dim3(256), 0, s, M, N, D,
                     ldd, X, ldx);
  return hipGetLastError();
}

}  // extern "C"

/*
 * k_qr.hip -- Householder QR kernels: the panel factorization one column at
 * a time (LAPACK's geqr2 with larfg, as src/lapack/ATL_geqr2.c and
 * ATL_larfg.c do it on the CPU) and the block reflector's triangular factor
 * (ATL_larft.c); the trailing update runs on the MFMA GEMM.
 *
 * Per panel column i (column 0 of P = A(i:, i:), w panel columns left):
 *   partial  -- every workgroup sums, over its rows r >= 1, x_r^2 and
 *               A(r, c) x_r for the panel's other columns c, with its
 *               largest |x_r| (the norm is taken scaled by it);
 *   finish   -- one workgroup adds the partials in a fixed order, forms
 *               larfg's beta, tau and the scale 1/(alpha - beta) of x, and
 *               w_c = A(0, c) + scale * sum_r A(r, c) x_r = (v^T A)_c;
 *   apply    -- v = [1; scale * x] written over the column, and every
 *               row r of the other columns gets A(r, c) -= tau v_r w_c.
 * Three launches per column; HBM traffic per column is the (m - i) x
 * (w - i) panel slice read twice and written once.
 */
#include <hip/hip_runtime.h>
#include <float.h>
#include "kernels.h"

namespace {

constexpr int NT = 256;
constexpr int PW = ATLMX_QR_NB;          /* widest panel */
constexpr int NV = PW + 2;               /* partial values: ss, max, dots */
constexpr int MAXG = 512;

template <typename T> struct QLim;
template <> struct QLim<double> {
  __device__ static double safmin() { return DBL_MIN; }
  __device__ static double eps() { return DBL_EPSILON * 0.5; }
};
template <> struct QLim<float> {
  __device__ static double safmin() { return FLT_MIN; }
  __device__ static double eps() { return FLT_EPSILON * 0.5; }
};

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

/* sums in double for both precisions; the norm is accumulated as
 * sum (x_r / s)^2 with s = the workgroup's largest |x_r|, found first */
template <typename T>
__global__ __launch_bounds__(NT) void qr_partial_kernel(int m, int w, const T *P, long ld,
                                                        double *part) {
  __shared__ double red[NT / 64][NV];
  __shared__ double s_scale;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long stride = (long)gridDim.x * NT;
  double mx = 0.0;
  for (long r = 1 + (long)blockIdx.x * NT + tid; r < m; r += stride)
    mx = fmax(mx, fabs((double)P[r]));
  mx = wave_max(mx);
  if (lane == 0) red[wave][0] = mx;
  __syncthreads();
  if (tid == 0) {
    double v = red[0][0];
    for (int k = 1; k < NT / 64; ++k) v = fmax(v, red[k][0]);
    s_scale = v;
  }
  __syncthreads();
  const double sc = s_scale;
  const double inv = sc > 0.0 ? 1.0 / sc : 0.0;
  double acc[NV - 1];
#pragma unroll
  for (int c = 0; c < NV - 1; ++c) acc[c] = 0.0;
  for (long r = 1 + (long)blockIdx.x * NT + tid; r < m; r += stride) {
    const double x = (double)P[r];
    const double xs = x * inv;
    acc[0] += xs * xs;
#pragma unroll
    for (int c = 1; c < PW; ++c)
      if (c < w) acc[c] += (double)P[(size_t)c * ld + r] * x;
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NV - 1; ++c) {
    const double v = wave_sum(acc[c]);
    if (lane == 0) red[wave][c] = v;
  }
  __syncthreads();
  if (tid < NV - 1) {
    double v = 0.0;
    for (int k = 0; k < NT / 64; ++k) v += red[k][tid];
    /* slot 0: sum (x/s)^2, slot PW: s (the dots keep slots 1..w-1) */
    part[(size_t)blockIdx.x * NV + tid] = v;
  }
  if (tid == 0) part[(size_t)blockIdx.x * NV + NV - 1] = sc;
}

/* coef: [0] tau, [1] scale of x, [2 + c] w_c for c = 1..w-1 */
template <typename T>
__global__ __launch_bounds__(64) void qr_finish_kernel(int G, int w, T *P, long ld,
                                                       const double *part, T *tau,
                                                       double *coef) {
  __shared__ double s_d[PW];
  const int lane = threadIdx.x;
  /* largest workgroup scale, then the scaled sum of squares re-based to it */
  double smax = 0.0;
  for (int g = 0; g < G; ++g) smax = fmax(smax, part[(size_t)g * NV + NV - 1]);
  if (lane < PW) {
    double v = 0.0;
    if (lane == 0) {
      for (int g = 0; g < G; ++g) {
        const double sg = part[(size_t)g * NV + NV - 1];
        const double f = smax > 0.0 ? sg / smax : 0.0;
        v += part[(size_t)g * NV] * f * f;
      }
    } else if (lane < w) {
      for (int g = 0; g < G; ++g) v += part[(size_t)g * NV + lane];
    }
    s_d[lane] = v;
  }
  __syncthreads();
  if (lane == 0) {
    const double xnorm = smax * sqrt(s_d[0]);
    double alpha = (double)P[0];
    double t = 0.0, scale = 0.0;
    if (xnorm != 0.0) {
      /* src/lapack/ATL_larfg.c: beta = -sign(alpha) hypot(alpha, xnorm);
       * tiny beta: rescale by 1/safmin (up to 20 times) first */
      const double safmin = QLim<T>::safmin() / QLim<T>::eps();
      double betap = hypot(alpha, xnorm);
      double beta = alpha >= 0.0 ? -betap : betap;
      double xn = xnorm, rs = 1.0;
      int knt = 0;
      if (betap < safmin) {
        const double rsafmn = 1.0 / safmin;
        do {
          ++knt;
          rs *= rsafmn;
          beta *= rsafmn;
          alpha *= rsafmn;
          betap = fabs(beta);
        } while (betap < safmin && knt < 20);
        xn = xnorm * rs;
        beta = hypot(alpha, xn);
        if (alpha > 0.0) beta = -beta;
      }
      t = (beta - alpha) / beta;
      scale = rs / (alpha - beta);
      for (int k = 0; k < knt; ++k) beta *= QLim<T>::safmin();
      P[0] = (T)beta;
    }
    *tau = (T)t;
    coef[0] = t;
    coef[1] = scale;
  }
  __syncthreads();
  if (lane >= 1 && lane < w) {
    /* (v^T A)_c with v_0 = 1: A(0, c) + scale * sum_r A(r, c) x_r */
    coef[2 + lane] = (double)P[(size_t)lane * ld] + coef[1] * s_d[lane];
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void qr_apply_kernel(int m, int w, T *P, long ld,
                                                      const double *coef) {
  const double t = coef[0];
  if (t == 0.0) return;
  const double scale = coef[1];
  for (long r = (long)blockIdx.x * NT + threadIdx.x; r < m; r += (long)gridDim.x * NT) {
    double v = 1.0;
    if (r > 0) {
      v = (double)P[r] * scale;
      P[r] = (T)v;
    }
    const double tv = t * v;
    for (int c = 1; c < w; ++c) {
      T *a = P + (size_t)c * ld + r;
      *a = (T)((double)*a - tv * coef[2 + c]);
    }
  }
}

/* V (m x w, ld m) = the panel's unit lower trapezoid */
template <typename T>
__global__ __launch_bounds__(NT) void qr_vcopy_kernel(int m, int w, const T *P, long ld, T *V) {
  const long total = (long)m * w;
  for (long e = (long)blockIdx.x * NT + threadIdx.x; e < total; e += (long)gridDim.x * NT) {
    const int c = (int)(e / m), r = (int)(e - (long)c * m);
    V[e] = r > c ? P[(size_t)c * ld + r] : (r == c ? T(1) : T(0));
  }
}

/* T (w x w, ld w) upper triangular with T(i,i) = tau_i and
 * T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i), G = V^T V (ld w) */
template <typename T>
__global__ __launch_bounds__(64) void qr_larft_kernel(int w, const T *G, const T *tau, T *Tm) {
  __shared__ double s_t[PW * PW];
  __shared__ double s_g[PW];
  const int lane = threadIdx.x;
  for (int e = lane; e < PW * PW; e += 64) s_t[e] = 0.0;
  __syncthreads();
  for (int i = 0; i < w; ++i) {
    const double ti = (double)tau[i];
    if (lane < i) s_g[lane] = (double)G[(size_t)i * w + lane];
    __syncthreads();
    if (lane < i) {
      double acc = 0.0;
      for (int k = lane; k < i; ++k) acc += s_t[k * PW + lane] * s_g[k];
      s_t[i * PW + lane] = -ti * acc;
    }
    if (lane == 0) s_t[i * PW + i] = ti;
    __syncthreads();
  }
  for (int e = lane; e < w * w; e += 64) {
    const int c = e / w, r = e - c * w;
    Tm[e] = (T)s_t[c * PW + r];
  }
}

template <typename T>
hipError_t launch_panel(int m, int w, T *P, long ld, T *tau, double *work, hipStream_t s) {
  if (m <= 0 || w <= 0) return hipSuccess;
  if (w > PW || ld < m || !P || !tau || !work) return hipErrorInvalidValue;
  double *part = work, *coef = work + (size_t)MAXG * NV;
  const int kmax = m < w ? m : w;
  for (int i = 0; i < kmax; ++i) {
    T *Pi = P + (size_t)i * ld + i;
    const int mi = m - i, wi = w - i;
    int G = (mi + NT - 1) / NT;
    if (G > MAXG) G = MAXG;
    hipLaunchKernelGGL(qr_partial_kernel<T>, dim3(G), dim3(NT), 0, s, mi, wi, (const T *)Pi,
                       ld, part);
    hipLaunchKernelGGL(qr_finish_kernel<T>, dim3(1), dim3(64), 0, s, G, wi, Pi, ld,
                       (const double *)part, tau + i, coef);
    hipLaunchKernelGGL(qr_apply_kernel<T>, dim3(G), dim3(NT), 0, s, mi, wi, Pi, ld,
                       (const double *)coef);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_vcopy(int m, int w, const T *P, long ld, T *V, hipStream_t s) {
  if (m <= 0 || w <= 0) return hipSuccess;
  if (ld < m || w > PW) return hipErrorInvalidValue;
  long blocks = ((long)m * w + NT - 1) / NT;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(qr_vcopy_kernel<T>, dim3((unsigned)blocks), dim3(NT), 0, s, m, w, P, ld,
                     V);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_larft(int w, const T *G, const T *tau, T *Tm, hipStream_t s) {
  if (w <= 0) return hipSuccess;
  if (w > PW) return hipErrorInvalidValue;
  hipLaunchKernelGGL(qr_larft_kernel<T>, dim3(1), dim3(64), 0, s, w, G, tau, Tm);
  return hipGetLastError();
}

}  // namespace

extern "C" {
size_t atlmx_k_qr_work_doubles(void) { return (size_t)MAXG * NV + NV + 2; }
hipError_t atlmx_k_dgeqr2_panel(int m, int w, double *P, long ld, double *tau, double *work,
                                hipStream_t s) {
  return launch_panel<double>(m, w, P, ld, tau, work, s);
}
hipError_t atlmx_k_sgeqr2_panel(int m, int w, float *P, long ld, float *tau, double *work,
                                hipStream_t s) {
  return launch_panel<float>(m, w, P, ld, tau, work, s);
}
hipError_t atlmx_k_dqr_vcopy(int m, int w, const double *P, long ld, double *V, hipStream_t s) {
  return launch_vcopy<double>(m, w, P, ld, V, s);
}
hipError_t atlmx_k_sqr_vcopy(int m, int w, const float *P, long ld, float *V, hipStream_t s) {
  return launch_vcopy<float>(m, w, P, ld, V, s);
}
hipError_t atlmx_k_dlarft(int w, const double *G, const double *tau, double *Tm, hipStream_t s) {
  return launch_larft<double>(w, G, tau, Tm, s);
}
hipError_t atlmx_k_slarft(int w, const float *G, const float *tau, float *Tm, hipStream_t s) {
  return launch_larft<float>(w, G, tau, Tm, s);
}
}  // extern "C"

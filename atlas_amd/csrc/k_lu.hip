/*
 * k_lu.hip -- the LU panel: partial-pivoting factorization of an m x w
 * column panel (w <= 32) in ONE cooperative launch.
 *
 * Replaces ATLAS's panel leaves -- getf2/LU1 (src/lapack/ATL_getf2.c:10-95),
 * LU2 (src/lapack/ATL_getrfC.c:46-136) and the threaded tgetf2 with its
 * cache-line combine of iamax results (src/threads/lapack/ATL_tgetf2.c,
 * src/threads/cbc/ATL_comb_iamax.c:15-106).  Per column it does what LU1
 * does: pivot = first index of max |a| (cblas_iamax: strict '>', ties to
 * the lower index, NaN never wins), swap the whole panel row, scale by the
 * reciprocal of the pivot when |p| >= SAFMIN (else divide), rank-1 update.
 *
 * Layout: the panel rows are spread over G workgroups of 256 threads; each
 * thread keeps RPT whole panel rows (w values each) in registers for the
 * entire factorization, so HBM is touched once to load and once to store.
 * Per column: a wavefront-level (shuffle) arg-max, a workgroup arg-max in
 * LDS, then each workgroup publishes its candidate (value, row, and the
 * candidate row itself) plus -- from the owner -- the current diagonal
 * row; one grid barrier; every workgroup then reduces the G candidates in
 * one wavefront and applies swap/scale/update locally.  One grid barrier
 * per column is the whole cross-CU cost.
 */
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>
#include "kernels.h"

namespace {

constexpr int PW = 32;      /* panel width held in registers */
constexpr int NT = 256;
constexpr int MAXG = 256;   /* cooperative workgroups at most (1 per CU) */

template <typename T> struct Lim;
template <> struct Lim<double> { __device__ static double safmin() { return DBL_MIN; } };
template <> struct Lim<float> { __device__ static float safmin() { return FLT_MIN; } };

template <typename T>
__device__ inline void amax_combine(T &v, int &i, T v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

/* scratch layout (after the 256-byte header holding the barrier counter) */
template <typename T>
struct PanelScratch {
  T *cval;      /* [2][G]      candidate |value| per workgroup  */
  int *cidx;    /* [2][G]      candidate panel row               */
  T *crow;      /* [2][G][PW]  candidate row values              */
  T *drow;      /* [2][PW]     current diagonal row values       */
};

template <typename T>
__device__ inline PanelScratch<T> scratch_view(void *base, int G) {
  PanelScratch<T> s;
  char *p = (char *)base + 256;
  s.cval = (T *)p;                 p += sizeof(T) * 2 * MAXG;
  s.cidx = (int *)p;               p += sizeof(int) * 2 * MAXG;
  s.crow = (T *)p;                 p += sizeof(T) * 2 * MAXG * PW;
  s.drow = (T *)p;
  (void)G;
  return s;
}

__device__ inline void grid_barrier(unsigned *cnt, unsigned target) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target)
      __builtin_amdgcn_s_sleep(1);
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

template <typename T, int RPT>
__global__ __launch_bounds__(NT) void getf2_panel_kernel(int m, int w, T *P, long ld,
                                                         int row0, int *ipiv, int *info,
                                                         void *scratch) {
  const int G = gridDim.x;
  const int g = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  unsigned *cnt = (unsigned *)scratch;
  PanelScratch<T> S = scratch_view<T>(scratch, G);

  __shared__ T s_wv[NT / 64];
  __shared__ int s_wi[NT / 64];
  __shared__ T s_prow[PW], s_drow[PW];
  __shared__ int s_piv;
  __shared__ T s_bv;

  /* rows owned by this thread: r = (g*RPT + q)*NT + tid */
  T a[RPT][PW];
  int rr[RPT];
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    rr[q] = (g * RPT + q) * NT + tid;
#pragma unroll
    for (int k = 0; k < PW; ++k)
      a[q][k] = (rr[q] < m && k < w) ? P[(size_t)k * ld + rr[q]] : T(0);
  }

#pragma unroll 1
  for (int c = 0; c < w; ++c) {
    const int slot = c & 1;
    /* extract column c from the register rows (c is not a compile-time
     * index here, so select with an unrolled compare) */
    T colv[RPT];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      T v = T(0);
#pragma unroll
      for (int k = 0; k < PW; ++k) if (k == c) v = a[q][k];
      colv[q] = v;
    }
    /* ---- local arg-max over rows >= c ---- */
    T bv = T(-1);
    int bi = 0x7fffffff;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      if (rr[q] < m && rr[q] >= c) {
        T v = colv[q] < T(0) ? -colv[q] : colv[q];
        if (!(v == v)) v = T(-1);
        amax_combine(bv, bi, v, rr[q]);
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const T v2 = __shfl_xor(bv, off);
      const int i2 = __shfl_xor(bi, off);
      amax_combine(bv, bi, v2, i2);
    }
    if (lane == 0) { s_wv[wave] = bv; s_wi[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      T v = s_wv[0];
      int i = s_wi[0];
#pragma unroll
      for (int k = 1; k < NT / 64; ++k) amax_combine(v, i, s_wv[k], s_wi[k]);
      S.cval[slot * MAXG + g] = v;
      S.cidx[slot * MAXG + g] = i;
      s_piv = i;
    }
    __syncthreads();
    /* publish the candidate row and (from its owner) the diagonal row */
    {
      const int wi = s_piv;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        if (rr[q] == wi) {
#pragma unroll
          for (int k = 0; k < PW; ++k) S.crow[((size_t)slot * MAXG + g) * PW + k] = a[q][k];
        }
        if (rr[q] == c) {
#pragma unroll
          for (int k = 0; k < PW; ++k) S.drow[slot * PW + k] = a[q][k];
        }
      }
    }
    grid_barrier(cnt, (unsigned)(G * (c + 1)));

    /* ---- global arg-max over the G candidates (wavefront 0) ---- */
    if (wave == 0) {
      T v = T(-1);
      int i = 0x7fffffff, gi = 0;
      for (int k = lane; k < G; k += 64) {
        const T v2 = S.cval[slot * MAXG + k];
        const int i2 = S.cidx[slot * MAXG + k];
        if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; gi = k; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const T v2 = __shfl_xor(v, off);
        const int i2 = __shfl_xor(i, off);
        const int g2 = __shfl_xor(gi, off);
        if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; gi = g2; }
      }
      /* all |a| zero or NaN: iamax answers the first row, i.e. c */
      const int p = (v > T(0)) ? i : c;
      if (lane < PW) {
        s_drow[lane] = S.drow[slot * PW + lane];
        s_prow[lane] = (p == c) ? S.drow[slot * PW + lane]
                                : S.crow[((size_t)slot * MAXG + gi) * PW + lane];
      }
      if (lane == 0) {
        s_piv = p;
        s_bv = v;
        if (g == 0) ipiv[c] = row0 + p;
      }
    }
    __syncthreads();

    const int p = s_piv;
    T pv = T(0);
#pragma unroll
    for (int k = 0; k < PW; ++k) if (k == c) pv = s_prow[k];
    if (pv != T(0)) {
      const bool recip = (pv < T(0) ? -pv : pv) >= Lim<T>::safmin();
      const T rcp = T(1) / pv;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        if (rr[q] >= m) continue;
        if (p != c && rr[q] == p) {
#pragma unroll
          for (int k = 0; k < PW; ++k) a[q][k] = s_drow[k];
        } else if (p != c && rr[q] == c) {
#pragma unroll
          for (int k = 0; k < PW; ++k) a[q][k] = s_prow[k];
        }
        if (rr[q] > c) {
          T l = T(0);
#pragma unroll
          for (int k = 0; k < PW; ++k) if (k == c) l = a[q][k];
          l = recip ? l * rcp : l / pv;
#pragma unroll
          for (int k = 0; k < PW; ++k) {
            if (k == c) a[q][k] = l;
            else if (k > c) a[q][k] -= l * s_prow[k];
          }
        }
      }
    } else if (g == 0 && tid == 0) {
      atomicCAS(info, 0, row0 + c + 1);
    }
    (void)s_bv;
    /* the next column's LDS writes must wait until everyone has read */
    __syncthreads();
  }

#pragma unroll
  for (int q = 0; q < RPT; ++q)
    if (rr[q] < m)
#pragma unroll
      for (int k = 0; k < PW; ++k)
        if (k < w) P[(size_t)k * ld + rr[q]] = a[q][k];
}

template <typename T>
hipError_t launch_panel(int m, int w, T *P, long ld, int row0, int *ipiv, int *info,
                        void *scratch, hipStream_t s) {
  if (m <= 0 || w <= 0) return hipSuccess;
  if (w > PW || ld < m || !P || !ipiv || !info || !scratch) return hipErrorInvalidValue;
  if (w > m) return hipErrorInvalidValue;          /* panel must be tall */
  int rpt = 1;
  while ((long)rpt * NT * MAXG < m) rpt *= 2;
  if (rpt > 4) return hipErrorInvalidValue;
  const int G = (int)((m + (long)rpt * NT - 1) / ((long)rpt * NT));
  void *args[] = {&m, &w, &P, &ld, &row0, &ipiv, &info, &scratch};
  hipError_t e;
  if (rpt == 1)
    e = hipLaunchCooperativeKernel((const void *)getf2_panel_kernel<T, 1>, dim3(G), dim3(NT),
                                   args, 0, s);
  else if (rpt == 2)
    e = hipLaunchCooperativeKernel((const void *)getf2_panel_kernel<T, 2>, dim3(G), dim3(NT),
                                   args, 0, s);
  else
    e = hipLaunchCooperativeKernel((const void *)getf2_panel_kernel<T, 4>, dim3(G), dim3(NT),
                                   args, 0, s);
  return e;
}

}  // namespace

extern "C" size_t atlmx_k_panel_scratch_bytes(int m, int w) {
  (void)m; (void)w;
  return 256 + (sizeof(double) + sizeof(int)) * 2 * MAXG + sizeof(double) * 2 * MAXG * PW +
         sizeof(double) * 2 * PW + 256;
}

extern "C" hipError_t atlmx_k_dgetf2_panel(int m, int w, double *P, long ld, int row0,
                                           int *ipiv, int *info, void *scratch,
                                           hipStream_t s) {
  return launch_panel<double>(m, w, P, ld, row0, ipiv, info, scratch, s);
}

extern "C" hipError_t atlmx_k_sgetf2_panel(int m, int w, float *P, long ld, int row0,
                                           int *ipiv, int *info, void *scratch,
                                           hipStream_t s) {
  return launch_panel<float>(m, w, P, ld, row0, ipiv, info, scratch, s);
}

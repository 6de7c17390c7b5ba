/*
 * k_lu.hip -- the LU panel: partial-pivoting factorization of an m x w
 * column panel (w <= 64) in ONE cooperative launch, and the row-interchange
 * kernels around it.
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
 * whole factorization, so HBM is touched once to load and once to store.
 *
 * Cross-workgroup exchange, per column: every workgroup reduces its rows to
 * one candidate (wavefront shuffles, then LDS across the 4 wavefronts) and
 * publishes it in its record for the column's parity: the row index, |value|
 * and the candidate row, every 32 bits of payload in a 64-bit word whose
 * upper half is a tag unique to (launch, column).  A word is valid when its
 * tag is the current one, so no store has to wait for an earlier one to
 * land and no slot is ever reset.  Readers poll the G records' headers at
 * once (one lane per workgroup), reduce them in one wavefront, and then
 * fetch the one row they need: the winning candidate row, or, in the
 * winner's own workgroup, the row it swaps with.  Two record sets
 * (column parity) suffice: a workgroup can publish column c + 2 only after
 * reading every record of column c + 1, which each writer publishes only
 * after reading all of column c.  All exchanged words move with
 * device-coherent (sc1) accesses, so no L2 writeback/invalidate fence is
 * needed.
 */
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>
#include "kernels.h"
#include "wave_argmax.h"

namespace {

constexpr int NT = 256;
constexpr int MAXG = 256;   /* cooperative workgroups at most (1 per CU) */
constexpr int MAXW = ATLMX_PANEL_MAXW;
constexpr int TALL_PARTS = 1024;   /* partial arg-maxes of the tall path */

template <typename T> struct Lim;
template <> struct Lim<double> { __device__ static double safmin() { return DBL_MIN; } };
template <> struct Lim<float> { __device__ static float safmin() { return FLT_MIN; } };

template <typename T>
__device__ inline void amax_combine(T &v, int &i, T v2, int i2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
}

/* Record of one workgroup for one column parity: 64-bit words tag << 32 |
 * payload, [0] the candidate's row index, [1], [2] the low and high halves
 * of its |value| as a double (-1 for NaN, -2 when the workgroup has no row
 * left), [RH + 2j], [RH + 2j + 1] the halves of entry j of the candidate
 * row.  One more record per parity holds the current row c (published by
 * workgroup 0, which holds rows 0..R-1 of the panel). */
constexpr int RH = 4;
constexpr int RECW = RH + 2 * MAXW;
typedef unsigned long long u64;

struct PanelSlots {
  u64 *rec;          /* [2 parity][MAXG + 1][RECW]; record MAXG: the current row */
  double *part_v;    /* tall path: per-workgroup arg-max partials */
  int *part_i;
};

__host__ __device__ inline size_t slots_bytes() {
  return (size_t)2 * (MAXG + 1) * RECW * 8 + (size_t)TALL_PARTS * (8 + 4);
}

__device__ inline PanelSlots slots_view(void *base) {
  PanelSlots s;
  char *p = (char *)base;
  s.rec = (u64 *)p;        p += (size_t)2 * (MAXG + 1) * RECW * 8;
  s.part_v = (double *)p;  p += (size_t)TALL_PARTS * 8;
  s.part_i = (int *)p;
  return s;
}

__device__ inline void put_u32(u64 *p, unsigned tag, unsigned v) {
  __hip_atomic_store(p, ((u64)tag << 32) | v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void put_f64(u64 *p, unsigned tag, double v) {
  const u64 b = (u64)__double_as_longlong(v);
  put_u32(p, tag, (unsigned)b);
  put_u32(p + 1, tag, (unsigned)(b >> 32));
}
/* the 32-bit payload of word p once it carries tag */
__device__ inline unsigned take_u32(u64 *p, unsigned tag) {
  u64 w;
  while (((w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 32) != tag)
    __builtin_amdgcn_s_sleep(1);
  return (unsigned)w;
}
__device__ inline double take_f64(u64 *p, unsigned tag) {
  const unsigned lo = take_u32(p, tag), hi = take_u32(p + 1, tag);
  return __longlong_as_double((long long)(((u64)hi << 32) | lo));
}

/* The panel rows of one workgroup live in LDS, row-major with a padded row
 * stride (W+1): a thread's own row is contiguous (its rank-1 update streams
 * through it), and column c of all rows is a stride-(W+1) gather. */
template <typename T, int W, int R>
__global__ __launch_bounds__(NT) void getf2_panel_lds_kernel(int m, int w, T *P, long ld,
                                                         int row0, int *ipiv, int *info,
                                                         void *scratch, int seq) {
  constexpr int RPT = R / NT;
  constexpr int LDW = W + 1;   /* odd stride: lanes hit distinct LDS banks */
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  T *rows = (T *)smem;                         /* [R][LDW] */
  __shared__ double s_wv[NT / 64];
  __shared__ int s_wi[NT / 64];
  __shared__ T s_prow[W];
  __shared__ int s_piv;
  const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int rbase = g * R;                     /* first panel row of this workgroup */
  PanelSlots S = slots_view(scratch);

  for (int k = 0; k < w; ++k)
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = rbase + q * NT + tid;
      rows[(q * NT + tid) * LDW + k] = (r < m) ? P[(size_t)k * ld + r] : T(0);
    }
  __syncthreads();

  for (int c = 0; c < w; ++c) {
    const int par = c & 1;
    /* ---- local arg-max over my rows >= c ---- */
    double bv = -2.0;
    int bi = 0x7fffffff, bz = 0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = rbase + q * NT + tid;
      if (r < m && r >= c) {
        double v = (double)rows[(q * NT + tid) * LDW + c];
        v = v < 0.0 ? -v : v;
        if (!(v == v)) v = -1.0;
        atlmx::argmax_combine(bv, bi, bz, v, r, 0);
      }
    }
    atlmx::wave_argmax(bv, bi, bz);
    if (lane == 0) { s_wv[wave] = bv; s_wi[wave] = bi; }
    __syncthreads();

    if (wave == 0) {
      double wv = s_wv[0];
      int wi = s_wi[0];
#pragma unroll
      for (int k = 1; k < NT / 64; ++k) atlmx::argmax_combine(wv, wi, bz, s_wv[k], s_wi[k], 0);
      if (wv < 0.0) wi = 0x7fffffff;    /* only NaN or no rows: nothing to offer */
      /* ---- publish this workgroup's record (and workgroup 0 the current
       * row c), each word tagged; no store waits for another ---- */
      const unsigned tag = (unsigned)seq * MAXW + (unsigned)c + 1u;
      u64 *recs = S.rec + (size_t)par * (MAXG + 1) * RECW;
      u64 *mine = recs + (size_t)g * RECW, *cur = recs + (size_t)MAXG * RECW;
      if (lane < w) {
        if (wi != 0x7fffffff) put_f64(mine + RH + 2 * lane, tag, (double)rows[(wi - rbase) * LDW + lane]);
        if (g == 0) put_f64(cur + RH + 2 * lane, tag, (double)rows[c * LDW + lane]);
      }
      if (lane == 0) {
        put_u32(mine, tag, (unsigned)wi);
        put_f64(mine + 1, tag, wv);
      }
      /* ---- read the G headers (one lane per workgroup), reduce ---- */
      double v = -2.0;
      int i = 0x7fffffff, gi = 0;
      for (int k0 = 0; k0 < G; k0 += 64) {
        const int k = k0 + lane;
        if (k < G) {
          u64 *rk = recs + (size_t)k * RECW;
          const int ik = (int)take_u32(rk, tag);
          atlmx::argmax_combine(v, i, gi, take_f64(rk + 1, tag), ik, k);
        }
      }
      atlmx::wave_argmax(v, i, gi);
      /* all |a| zero or NaN: iamax answers the first row, i.e. c */
      const int p = (v > 0.0) ? i : c;
      if (lane < w) {
        /* the pivot row: the winner's candidate (its own workgroup has it in
         * LDS); the current row c: needed only where row p lives, i.e. in
         * the winner's workgroup, which also has it in LDS when g == 0 */
        const bool own = (p != c) && gi == g;
        const T pr = (p == c) ? T(0)
                              : (own ? rows[(p - rbase) * LDW + lane]
                                     : (T)take_f64(recs + (size_t)gi * RECW + RH + 2 * lane, tag));
        if (p != c) {
          s_prow[lane] = pr;
          /* swap the whole panel rows c and p where this workgroup holds
           * them, one lane per column (p != c implies a nonzero pivot) */
          if (p >= rbase && p < rbase + R) {
            const T dr = (c >= rbase && c < rbase + R)
                             ? rows[(c - rbase) * LDW + lane]
                             : (T)take_f64(cur + RH + 2 * lane, tag);
            rows[(p - rbase) * LDW + lane] = dr;
          }
          if (c >= rbase && c < rbase + R) rows[(c - rbase) * LDW + lane] = pr;
        } else {
          /* no interchange: the pivot row is row c itself, held by
           * workgroup 0, published as the current row */
          s_prow[lane] = (c >= rbase && c < rbase + R) ? rows[(c - rbase) * LDW + lane]
                                                       : (T)take_f64(cur + RH + 2 * lane, tag);
        }
      }
      if (lane == 0) {
        s_piv = p;
        if (g == 0) ipiv[c] = row0 + p;
      }
    }
    __syncthreads();

    const int p = s_piv;
    const T pv = s_prow[c];
    if (pv != T(0)) {
      const bool recip = (pv < T(0) ? -pv : pv) >= Lim<T>::safmin();
      const T rcp = T(1) / pv;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int r = rbase + q * NT + tid;
        T *row = rows + (q * NT + tid) * LDW;
        if (r < m && r > c) {
          const T l = recip ? row[c] * rcp : row[c] / pv;
          row[c] = l;
#pragma unroll 8
          for (int k = c + 1; k < w; ++k) row[k] -= l * s_prow[k];
        }
      }
    } else if (g == 0 && tid == 0) {
      atomicCAS(info, 0, row0 + c + 1);
    }
    /* LDS rows are rewritten next column: wait until all have read */
    __syncthreads();
  }

  for (int k = 0; k < w; ++k)
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = rbase + q * NT + tid;
      if (r < m) P[(size_t)k * ld + r] = rows[(q * NT + tid) * LDW + k];
    }
}

__global__ void slots_init_kernel(u64 *h, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    h[i] = 0ull;                      /* tag 0: never a valid word */
}

/* ------------------------------------------------------------------------
 * Panels taller than the cooperative kernel holds (more than 2*NT*MAXG
 * rows): the same column steps as LU1, three launches per column on the
 * panel in HBM -- a partial arg-max per workgroup, a one-wavefront kernel
 * that reduces the partials (same tie and NaN rules), swaps the two panel
 * rows and records the pivot, and the scaling + rank-1 update of the rows
 * below.  The partials have their own area of the scratch; the tagged
 * records of the cooperative kernel are not touched, so the two paths can
 * alternate.
 * ---------------------------------------------------------------------- */
constexpr int TALL_G = TALL_PARTS;

template <typename T>
__global__ __launch_bounds__(NT) void tall_amax_kernel(int m, int c, const T *P, long ld,
                                                       double *pv, int *pi) {
  __shared__ double s_v[NT / 64];
  __shared__ int s_i[NT / 64];
  const T *col = P + (size_t)c * ld;
  double bv = -2.0;
  int bi = 0x7fffffff, bz = 0;
  for (long r = c + (long)blockIdx.x * NT + threadIdx.x; r < m; r += (long)gridDim.x * NT) {
    double v = (double)col[r];
    v = v < 0.0 ? -v : v;
    if (!(v == v)) v = -1.0;
    atlmx::argmax_combine(bv, bi, bz, v, (int)r, 0);
  }
  atlmx::wave_argmax(bv, bi, bz);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) { s_v[wave] = bv; s_i[wave] = bi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    bv = s_v[0];
    bi = s_i[0];
    for (int k = 1; k < NT / 64; ++k) atlmx::argmax_combine(bv, bi, bz, s_v[k], s_i[k], 0);
    pv[blockIdx.x] = bv;
    pi[blockIdx.x] = bi;
  }
}

template <typename T>
__global__ __launch_bounds__(64) void tall_swap_kernel(int w, int c, T *P, long ld, int row0,
                                                       int *ipiv, int *info, const double *pv,
                                                       const int *pi, int G) {
  const int lane = threadIdx.x;
  double v = -2.0;
  int i = 0x7fffffff, z = 0;
  for (int k = lane; k < G; k += 64) atlmx::argmax_combine(v, i, z, pv[k], pi[k], 0);
  atlmx::wave_argmax(v, i, z);
  const int p = (v > 0.0) ? i : c;      /* all zero or NaN: iamax answers c */
  const T pivval = P[(size_t)c * ld + p];
  if (p != c)
    for (int k = lane; k < w; k += 64) {
      const T a = P[(size_t)k * ld + c], b = P[(size_t)k * ld + p];
      P[(size_t)k * ld + c] = b;
      P[(size_t)k * ld + p] = a;
    }
  if (lane == 0) {
    ipiv[c] = row0 + p;
    if (pivval == T(0)) atomicCAS(info, 0, row0 + c + 1);
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void tall_update_kernel(int m, int w, int c, T *P, long ld) {
  const T pv = P[(size_t)c * ld + c];
  if (pv == T(0)) return;
  const bool recip = (pv < T(0) ? -pv : pv) >= Lim<T>::safmin();
  const T rcp = T(1) / pv;
  for (long r = c + 1 + (long)blockIdx.x * NT + threadIdx.x; r < m; r += (long)gridDim.x * NT) {
    const T x = P[(size_t)c * ld + r];
    const T l = recip ? x * rcp : x / pv;
    P[(size_t)c * ld + r] = l;
    for (int k = c + 1; k < w; ++k) P[(size_t)k * ld + r] -= l * P[(size_t)k * ld + c];
  }
}

template <typename T>
hipError_t launch_tall(int m, int w, T *P, long ld, int row0, int *ipiv, int *info,
                       void *scratch, hipStream_t s) {
  char *b = (char *)scratch + (size_t)2 * (MAXG + 1) * RECW * 8;   /* past the records */
  double *pv = (double *)b;                  /* TALL_G doubles, then TALL_G ints */
  int *pi = (int *)(pv + TALL_G);
  for (int c = 0; c < w; ++c) {
    long rows = (long)m - c;
    int G = (int)((rows + NT - 1) / NT);
    if (G > TALL_G) G = TALL_G;
    hipLaunchKernelGGL(tall_amax_kernel<T>, dim3(G), dim3(NT), 0, s, m, c, (const T *)P, ld,
                       pv, pi);
    hipLaunchKernelGGL(tall_swap_kernel<T>, dim3(1), dim3(64), 0, s, w, c, P, ld, row0, ipiv,
                       info, (const double *)pv, (const int *)pi, G);
    hipLaunchKernelGGL(tall_update_kernel<T>, dim3(G), dim3(NT), 0, s, m, w, c, P, ld);
  }
  return hipGetLastError();
}

template <typename T, int W, int R>
hipError_t launch_wr(int G, void **args, hipStream_t s) {
  static int attr_set = 0;
  const int lds = R * (W + 1) * (int)sizeof(T);
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void *)getf2_panel_lds_kernel<T, W, R>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    attr_set = 1;
  }
  return hipLaunchCooperativeKernel((const void *)getf2_panel_lds_kernel<T, W, R>, dim3(G),
                                    dim3(NT), args, (unsigned)lds, s);
}

template <typename T>
hipError_t launch_panel(int m, int w, T *P, long ld, int row0, int *ipiv, int *info,
                        void *scratch, int seq, hipStream_t s) {
  if (m <= 0 || w <= 0) return hipSuccess;
  if (w > MAXW || ld < m || !P || !ipiv || !info || !scratch) return hipErrorInvalidValue;
  if (w > m || seq < 0 || seq >= ATLMX_PANEL_SEQ_MAX) return hipErrorInvalidValue;
  void *args[] = {&m, &w, &P, &ld, &row0, &ipiv, &info, &scratch, &seq};
  /* rows per workgroup: 256, or 512 for taller panels (then w <= 32 in FP64
   * so the LDS tile stays within 160 KB) */
  if ((long)m <= (long)NT * MAXG) {
    const int G = (m + NT - 1) / NT;
    return launch_wr<T, MAXW, NT>(G, args, s);
  }
  if ((long)m > (long)2 * NT * MAXG) return launch_tall<T>(m, w, P, ld, row0, ipiv, info,
                                                          scratch, s);
  if (w > MAXW / 2) return hipErrorInvalidValue;
  const int G = (m + 2 * NT - 1) / (2 * NT);
  return launch_wr<T, MAXW / 2, 2 * NT>(G, args, s);
}

/* ------------------------------------------------------------------------
 * Row interchanges.  laswp with the swaps k1..k2 of one panel is applied to
 * the rest of the matrix as a gather/scatter: a one-workgroup plan kernel
 * replays the swaps on an index map in LDS and lists the rows that moved
 * (dst <- src); the apply kernel then moves those rows for many columns at
 * once with independent loads (src/lapack/ATL_laswp.c:7-110 does the same
 * swaps one pair at a time over 32-column chunks).
 * ---------------------------------------------------------------------- */
constexpr int PLAN_MAXROWS = ATLMX_LASWP_PLAN_MAXROWS;

__global__ __launch_bounds__(1024) void laswp_plan_kernel(int k1, int k2, int rows_end,
                                                          const int *ipiv, int *plan) {
  extern __shared__ int idx[];              /* idx[r - k1] = source row of row r */
  __shared__ int s_cnt;
  const int n = rows_end - k1;
  for (int r = threadIdx.x; r < n; r += blockDim.x) idx[r] = k1 + r;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = k1; k < k2; ++k) {
      const int p = ipiv[k];
      if (p != k) { const int t = idx[k - k1]; idx[k - k1] = idx[p - k1]; idx[p - k1] = t; }
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    if (idx[r] != k1 + r) {
      const int slot = atomicAdd(&s_cnt, 1);
      plan[2 + 2 * slot] = k1 + r;        /* dst */
      plan[3 + 2 * slot] = idx[r];        /* src */
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) plan[0] = s_cnt;
}

/* Apply a plan to columns [c0, c1) except [s0, s1): 32 columns per block. */
template <typename T>
__global__ __launch_bounds__(256) void laswp_apply_kernel(T *A, long lda, int c0, int c1,
                                                          int s0, int s1, const int *plan) {
  __shared__ T buf[2 * ATLMX_PANEL_MAXW * 2][33];
  __shared__ int s_dst[4 * ATLMX_PANEL_MAXW], s_src[4 * ATLMX_PANEL_MAXW];
  const int np = plan[0];
  const int cb = c0 + blockIdx.x * 32;
  for (int t = threadIdx.x; t < np; t += 256) {
    s_dst[t] = plan[2 + 2 * t];
    s_src[t] = plan[3 + 2 * t];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < np * 32; e += 256) {
    const int pi = e / 32, cc = e - pi * 32;   /* consecutive threads: columns */
    const int col = cb + cc;
    if (col < c1 && (col < s0 || col >= s1)) buf[pi][cc] = A[(size_t)col * lda + s_src[pi]];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < np * 32; e += 256) {
    const int pi = e / 32, cc = e - pi * 32;
    const int col = cb + cc;
    if (col < c1 && (col < s0 || col >= s1)) A[(size_t)col * lda + s_dst[pi]] = buf[pi][cc];
  }
}

/* Plain per-column swaps, used for right-hand sides and as the fallback for
 * row ranges beyond the plan kernel's LDS index map. */
template <typename T>
__global__ void laswp_kernel(int N, T *A, long lda, int k1, int k2, const int *ipiv, int inc) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  T *col = A + (size_t)c * lda;
  if (inc > 0) {
    for (int k = k1; k < k2; ++k) {
      const int p = ipiv[k];
      if (p != k) { const T t = col[k]; col[k] = col[p]; col[p] = t; }
    }
  } else {
    for (int k = k2 - 1; k >= k1; --k) {
      const int p = ipiv[k];
      if (p != k) { const T t = col[k]; col[k] = col[p]; col[p] = t; }
    }
  }
}

/* Few-column variant (the right-hand sides of getrs): one workgroup per
 * column holds rows [k1, rows_end) of the column in LDS and replays the
 * swaps there, so each swap costs LDS latency instead of two dependent
 * HBM round trips. */
constexpr int LDS_SWAP_ROWS = 16384;
template <typename T>
__global__ __launch_bounds__(256) void laswp_lds_kernel(T *A, long lda, int k1, int k2,
                                                        int rows_end, const int *ipiv,
                                                        int inc) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  T *col = (T *)smem;
  __shared__ int pv[1024];
  const int n = rows_end - k1;
  T *g = A + (size_t)blockIdx.x * lda + k1;
  for (int i = threadIdx.x; i < n; i += 256) col[i] = g[i];
  for (int base = 0; base < k2 - k1; base += 1024) {
    const int cnt = min(1024, k2 - k1 - base);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += 256) {
      const int k = (inc > 0) ? k1 + base + t : k2 - 1 - base - t;
      pv[t] = ipiv[k] - k1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int t = 0; t < cnt; ++t) {
        const int k = ((inc > 0) ? base + t : k2 - k1 - 1 - base - t);
        const int p = pv[t];
        if (p != k) { const T x = col[k]; col[k] = col[p]; col[p] = x; }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 256) g[i] = col[i];
}

/* The interchanges of getrs as one permutation (for right-hand sides: the
 * sequential replay above costs one LDS round trip per interchange).  With
 * p_k >= k (as getrf leaves them), interchange k is the last to touch row
 * k, so after the forward sequence row j holds what row p_j held just
 * before interchange j.  Let c(j) be what row j holds just before its own
 * interchange: the row last moved there, i.e. c(pred(j)) with pred(j) the
 * last k < j with p_k = j, or j itself.  Then the gather map is
 *   src[j] = c(j)                                  if j has no interchange
 *   src[j] = c(q) (q = the last k < j with p_k = p_j), else p_j.
 * One workgroup: the pairs (p_k, k) are sorted in LDS (bitonic), which
 * gives pred and q from neighbours, and c follows by pointer jumping
 * (log2 n rounds).  The reverse sequence (inc < 0) is the inverse map.  Any
 * p_k < k or out of range falls back to replaying the interchanges on row
 * labels in LDS (the map is then exact for whatever the sequence does,
 * entries out of range skipped). */
constexpr int PERM_NT = 1024;
__global__ __launch_bounds__(PERM_NT) void laswp_perm_kernel(int m, int nr, int npow,
                                                             const int *ipiv, int k1, int inc,
                                                             int *src, int *fwd) {
  extern __shared__ unsigned s_dyn[];
  unsigned *key = s_dyn;                          /* npow sorted (p << 16 | k) */
  int *r = reinterpret_cast<int *>(s_dyn + npow); /* nr: row labels, then c() */
  __shared__ int s_bad;
  const int tid = threadIdx.x;
  if (tid == 0) s_bad = 0;
  __syncthreads();
  for (int j = tid; j < nr; j += PERM_NT) r[j] = j;
  for (int k = tid; k < npow; k += PERM_NT) {
    unsigned v = 0xffffffffu;
    if (k < m) {
      const int p = ipiv[k1 + k] - k1;
      if (p < k || p >= nr) s_bad = 1;
      else if (p != k) v = ((unsigned)p << 16) | (unsigned)k;
    }
    key[k] = v;
  }
  __syncthreads();
  if (s_bad) {
    if (tid == 0) {
      for (int t = 0; t < m; ++t) {
        const int k = inc > 0 ? t : m - 1 - t;
        const int p = ipiv[k1 + k] - k1;
        if (p >= 0 && p < nr && p != k) { const int x = r[k]; r[k] = r[p]; r[p] = x; }
      }
    }
    __syncthreads();
    for (int j = tid; j < nr; j += PERM_NT) src[j] = r[j];
    return;
  }
  for (int size = 2; size <= npow; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = tid; i < npow / 2; i += PERM_NT) {
        const int lo = 2 * stride * (i / stride) + (i % stride), hi = lo + stride;
        const bool up = (lo & size) == 0;
        const unsigned a = key[lo], b = key[hi];
        if ((a > b) == up) { key[lo] = b; key[hi] = a; }
      }
      __syncthreads();
    }
  /* pred(p) = the largest k of p's run */
  for (int s = tid; s < m; s += PERM_NT) {
    const unsigned v = key[s];
    if (v == 0xffffffffu) continue;
    const unsigned nx = s + 1 < m ? key[s + 1] : 0xffffffffu;
    if (nx == 0xffffffffu || (nx >> 16) != (v >> 16)) r[v >> 16] = (int)(v & 0xffffu);
  }
  __syncthreads();
  /* r[j] = pred(j) or j; follow to the root: c(j) */
  for (int len = 1; len < nr; len <<= 1) {
    for (int j = tid; j < nr; j += PERM_NT) r[j] = r[r[j]];
    __syncthreads();
  }
  int *out = inc > 0 ? src : fwd;
  for (int j = tid; j < nr; j += PERM_NT) {
    const bool moved = j < m && (ipiv[k1 + j] - k1) != j;
    if (!moved) out[j] = r[j];
  }
  for (int s = tid; s < m; s += PERM_NT) {
    const unsigned v = key[s];
    if (v == 0xffffffffu) continue;
    const int p = (int)(v >> 16), k = (int)(v & 0xffffu);
    const unsigned pv = s > 0 ? key[s - 1] : 0xffffffffu;
    out[k] = (pv != 0xffffffffu && (int)(pv >> 16) == p) ? r[pv & 0xffffu] : p;
  }
  if (inc < 0) {
    __syncthreads();
    for (int j = tid; j < nr; j += PERM_NT) src[fwd[j]] = j;
  }
}

/* tmp(i, c) = A(k1 + src[i], c), then A(k1 + i, c) = tmp(i, c) */
template <typename T>
__global__ void gather_rows_kernel(int nr, const T *A, long lda, int k1, const int *src,
                                   T *tmp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (i < nr) tmp[(size_t)c * nr + i] = A[(size_t)c * lda + k1 + src[i]];
}
template <typename T>
__global__ void copy_rows_kernel(int nr, T *A, long lda, int k1, const T *tmp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
  if (i < nr) A[(size_t)c * lda + k1 + i] = tmp[(size_t)c * nr + i];
}

template <typename T>
hipError_t launch_permute_rows(int N, T *A, long lda, int k1, int nr, const int *src, T *tmp,
                               hipStream_t s) {
  if (N <= 0 || nr <= 0) return hipSuccess;
  if (N > 65535 || lda < k1 + nr || !src || !tmp) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((nr + 255) / 256), (unsigned)N);
  hipLaunchKernelGGL(gather_rows_kernel<T>, grid, dim3(256), 0, s, nr, A, lda, k1, src, tmp);
  hipLaunchKernelGGL(copy_rows_kernel<T>, grid, dim3(256), 0, s, nr, A, lda, k1, tmp);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_laswp(int N, T *A, long lda, int k1, int k2, const int *ipiv, int inc,
                        int rows_end, hipStream_t s) {
  static int attr_set = 0;
  if (N <= 0 || k2 <= k1) return hipSuccess;
  if (k1 < 0 || rows_end < k2 || lda < rows_end) return hipErrorInvalidValue;
  /* rows touched: [k1, rows_end); the LDS variant needs them to fit */
  if (N <= 256 && rows_end - k1 <= LDS_SWAP_ROWS) {
    if (!attr_set) {
      hipError_t e = hipFuncSetAttribute((const void *)laswp_lds_kernel<T>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         LDS_SWAP_ROWS * (int)sizeof(T));
      if (e != hipSuccess) return e;
      attr_set = 1;
    }
    hipLaunchKernelGGL(laswp_lds_kernel<T>, dim3(N), dim3(256),
                       (size_t)(rows_end - k1) * sizeof(T), s, A, lda, k1, k2, rows_end, ipiv,
                       inc);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(laswp_kernel<T>, dim3((N + 255) / 256), dim3(256), 0, s, N, A, lda, k1,
                     k2, ipiv, inc);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_laswp_apply(T *A, long lda, int c0, int c1, int s0, int s1, const int *plan,
                              hipStream_t s) {
  if (c1 <= c0) return hipSuccess;
  hipLaunchKernelGGL(laswp_apply_kernel<T>, dim3((c1 - c0 + 31) / 32), dim3(256), 0, s, A,
                     lda, c0, c1, s0, s1, plan);
  return hipGetLastError();
}

}  // namespace

extern "C" size_t atlmx_k_panel_scratch_bytes(int m, int w) {
  (void)m; (void)w;
  return slots_bytes() + 256;
}

extern "C" hipError_t atlmx_k_panel_scratch_init(void *scratch, hipStream_t s) {
  const size_t n = (size_t)2 * (MAXG + 1) * RECW;  /* the records of both parities */
  hipLaunchKernelGGL(slots_init_kernel, dim3(256), dim3(256), 0, s, (u64 *)scratch, n);
  return hipGetLastError();
}

extern "C" hipError_t atlmx_k_dgetf2_panel(int m, int w, double *P, long ld, int row0,
                                           int *ipiv, int *info, void *scratch, int seq,
                                           hipStream_t s) {
  return launch_panel<double>(m, w, P, ld, row0, ipiv, info, scratch, seq, s);
}

extern "C" hipError_t atlmx_k_sgetf2_panel(int m, int w, float *P, long ld, int row0,
                                           int *ipiv, int *info, void *scratch, int seq,
                                           hipStream_t s) {
  return launch_panel<float>(m, w, P, ld, row0, ipiv, info, scratch, seq, s);
}

extern "C" hipError_t atlmx_k_laswp_plan(int k1, int k2, int rows_end, const int *ipiv,
                                         int *plan, hipStream_t s) {
  static int attr_set = 0;
  if (k2 <= k1) return hipSuccess;
  if (k2 - k1 > 2 * ATLMX_PANEL_MAXW || rows_end < k2) return hipErrorInvalidValue;
  if (rows_end - k1 > PLAN_MAXROWS) return hipErrorInvalidValue;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void *)laswp_plan_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       PLAN_MAXROWS * (int)sizeof(int));
    if (e != hipSuccess) return e;
    attr_set = 1;
  }
  hipLaunchKernelGGL(laswp_plan_kernel, dim3(1), dim3(1024),
                     (size_t)(rows_end - k1) * sizeof(int), s, k1, k2, rows_end, ipiv, plan);
  return hipGetLastError();
}

extern "C" hipError_t atlmx_k_dlaswp_apply(double *A, long lda, int c0, int c1, int s0, int s1,
                                           const int *plan, hipStream_t s) {
  return launch_laswp_apply<double>(A, lda, c0, c1, s0, s1, plan, s);
}
extern "C" hipError_t atlmx_k_slaswp_apply(float *A, long lda, int c0, int c1, int s0, int s1,
                                           const int *plan, hipStream_t s) {
  return launch_laswp_apply<float>(A, lda, c0, c1, s0, s1, plan, s);
}
extern "C" hipError_t atlmx_k_dlaswp(int N, double *A, long lda, int k1, int k2,
                                     const int *ipiv, int inc, int rows_end, hipStream_t s) {
  return launch_laswp<double>(N, A, lda, k1, k2, ipiv, inc, rows_end, s);
}
extern "C" hipError_t atlmx_k_slaswp(int N, float *A, long lda, int k1, int k2,
                                     const int *ipiv, int inc, int rows_end, hipStream_t s) {
  return launch_laswp<float>(N, A, lda, k1, k2, ipiv, inc, rows_end, s);
}

extern "C" hipError_t atlmx_k_laswp_perm(int k1, int k2, int rows_end, const int *ipiv, int inc,
                                         int *src, int *tmp, hipStream_t s) {
  static int attr_set = 0;
  const int m = k2 - k1, nr = rows_end - k1;
  if (nr <= 0) return hipSuccess;
  if (k1 < 0 || m < 0 || m > nr || nr > ATLMX_LASWP_PERM_MAXROWS || !ipiv || !src || !tmp)
    return hipErrorInvalidValue;
  int npow = 1;
  while (npow < (m > 1 ? m : 1)) npow <<= 1;
  const size_t lds = (size_t)(npow + nr) * sizeof(int);
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute((const void *)laswp_perm_kernel,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       2 * ATLMX_LASWP_PERM_MAXROWS * (int)sizeof(int));
    if (e != hipSuccess) return e;
    attr_set = 1;
  }
  hipLaunchKernelGGL(laswp_perm_kernel, dim3(1), dim3(PERM_NT), lds, s, m, nr, npow, ipiv, k1,
                     inc, src, tmp);
  return hipGetLastError();
}
extern "C" hipError_t atlmx_k_dpermute_rows(int N, double *A, long lda, int k1, int nr,
                                            const int *src, double *tmp, hipStream_t s) {
  return launch_permute_rows<double>(N, A, lda, k1, nr, src, tmp, s);
}
extern "C" hipError_t atlmx_k_spermute_rows(int N, float *A, long lda, int k1, int nr,
                                            const int *src, float *tmp, hipStream_t s) {
  return launch_permute_rows<float>(N, A, lda, k1, nr, src, tmp, s);
}

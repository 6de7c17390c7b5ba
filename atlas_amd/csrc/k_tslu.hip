/*
 * k_tslu.hip -- kernels of the FP32 LU with tournament pivoting that the
 * mixed-precision solver uses (no ATLAS counterpart: ATLAS's LU uses
 * partial pivoting, src/lapack/ATL_getrfC.c:46-136, whose pivot search is a
 * global reduction for every column).
 *
 * Tournament pivoting (the TSLU panel of communication-avoiding LU:
 * Grigori, Demmel, Xiang, "CALU: a communication optimal LU factorization
 * algorithm", SIAM J. Matrix Anal. Appl. 32(4), 2011) picks the W pivot
 * rows of a panel with a reduction tree: every leaf runs Gaussian
 * elimination with partial pivoting on 256 rows of the panel and keeps the
 * W rows it chose; every tree node does the same on the ORIGINAL values of
 * its children's chosen rows (up to 4 x W).  The root's choice, in its
 * pivot order, becomes the panel's pivot rows; the panel is then factored
 * without further interchanges.  One reduction tree per panel replaces one
 * device-wide exchange per column.
 *
 * The FP32 matrix is stored transposed, S = A^T (column-major), so a row of
 * A -- what pivoting moves -- is a contiguous column of S.
 *
 * Tree synchronisation needs no waiting: a node is merged by whichever of
 * its children arrives last (an atomic counter per node), after every child
 * has published its candidate rows with device-coherent stores, so no
 * workgroup ever spins and no co-residency is assumed.  The merge order
 * within a node is fixed (child 0..3), so the result does not depend on
 * timing.
 */
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>
#include "kernels.h"

namespace {

constexpr int TW = ATLMX_TSLU_W;   /* 64 */
constexpr int TR = 256;            /* rows per elimination = threads */
constexpr int FAN = 4;

__device__ inline int ldc_i(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void stc_i(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int CTRL, int RM>
__device__ __forceinline__ float dpp_max(float v) {
  const int o = __builtin_amdgcn_update_dpp(__float_as_int(-2.0f), __float_as_int(v), CTRL, RM,
                                            0xf, false);
  return fmaxf(v, __int_as_float(o));
}

/* maximum over the wavefront, broadcast (values >= -2) */
__device__ __forceinline__ float wave_max(float v) {
  v = dpp_max<0x111, 0xf>(v);
  v = dpp_max<0x112, 0xf>(v);
  v = dpp_max<0x114, 0xf>(v);
  v = dpp_max<0x118, 0xf>(v);
  v = dpp_max<0x142, 0xa>(v);
  v = dpp_max<0x143, 0xc>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

typedef float f4 __attribute__((ext_vector_type(4)));

struct TsluArgs {
  const float *S;   /* A^T, column-major: A(i, j) = S[j + i*lds] */
  long lds;
  int k, n, w;      /* panel: rows k..n-1 of A, columns k..k+w-1 (w <= TW) */
  int nleaves;
  int *cand;        /* TW row indices per tree node */
  int *cnt;         /* one arrival counter per tree node (zero on entry) */
  int *plan;        /* out: move list, plan[0] = count, then (dst, src) rows */
  int *perm;        /* in/out: row i of the permuted A is row perm[i] of A */
  float *lu11;      /* out: the diagonal block's L\U, TW x TW column-major */
  float *invl, *invu; /* out: inv(L11) (unit lower), inv(U11), TW x TW */
  int *info;        /* zero pivot: 1 + its column; nonzero on entry = skip */
};

/* row `row` of A, columns k..k+w-1 (zero beyond w, or for row < 0) */
__device__ inline void load_row(f4 (&r)[TW / 4], const TsluArgs &a, int row) {
  const float *src = a.S + a.k + (size_t)(row < 0 ? 0 : row) * a.lds;
  if (row >= 0 && a.w == TW) {
#pragma unroll
    for (int q = 0; q < TW / 4; ++q) r[q] = *reinterpret_cast<const f4 *>(src + 4 * q);
  } else {
#pragma unroll
    for (int q = 0; q < TW / 4; ++q)
#pragma unroll
      for (int e = 0; e < 4; ++e) r[q][e] = (row >= 0 && 4 * q + e < a.w) ? src[4 * q + e] : 0.0f;
  }
}

/* Gaussian elimination with partial pivoting on the rows held by the
 * workgroup (one per thread, in registers as float4), recording in s_sel the
 * rows chosen as pivots, in order.  Ties go to the lower thread.  The
 * multipliers are kept in place, so a row retired at column c holds its L
 * entries in r[0..c) and its U entries in r[c..TW); *mypos = c.  All TW
 * columns are eliminated: columns past the panel width hold zeros, so
 * their steps change nothing that is used (their choices are ignored), and
 * the loop body has no branch that would make the compiler copy r. */
__device__ void gepp_select(f4 (&r)[TW / 4], int myrow, f4 (*s_row)[4][TW / 4],
                            float (*s_val)[4], int (*s_thr)[4], int *s_sel, int *mypos) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  bool act = myrow >= 0;
  int pos = -1;
#pragma unroll
  for (int c = 0; c < TW; ++c) {
    const int par = c & 1;
    const float rc = r[c / 4][c % 4];
    float v = act ? fabsf(rc) : -1.0f;
    if (!(v == v)) v = 0.0f;                      /* NaN: never better than a number */
    const float wm = wave_max(v);
    const unsigned long long bal = __ballot(v == wm);
    const int wl = __ffsll((long long)bal) - 1;
    if (lane == wl) {                             /* stores only */
      s_val[par][wave] = wm;
      s_thr[par][wave] = tid;
#pragma unroll
      for (int q = c / 4; q < TW / 4; ++q) s_row[par][wave][q] = r[q];
    }
    __syncthreads();
    float best = s_val[par][0];
    int bw = 0;
#pragma unroll
    for (int q = 1; q < 4; ++q)
      if (s_val[par][q] > best) { best = s_val[par][q]; bw = q; }
    const bool any = best >= 0.0f;                /* some row was still active */
    const bool me = any && tid == s_thr[par][bw];
    if (me) s_sel[c] = myrow;
    if (!any && tid == 0) s_sel[c] = -1;
    pos = me ? c : pos;
    act = act && !me;
    const float piv = s_row[par][bw][c / 4][c % 4];
    const float l = (act && piv != 0.0f) ? rc / piv : 0.0f;
    r[c / 4][c % 4] = act ? l : rc;
#pragma unroll
    for (int q = c / 4; q < TW / 4; ++q) {
      const f4 pr = s_row[par][bw][q];
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (4 * q + e > c) r[q][e] -= l * pr[e];
    }
  }
  *mypos = pos;
  __syncthreads();
}

/* inverses of the unit lower and the upper triangle of the w x w factor in
 * s_lu (row-major, ld TW+1) by recursive doubling: for b = 1, 2, .., 32 each
 * 2b x 2b diagonal block gets its off-diagonal block of the inverse,
 * X21 = -X22 L21 X11 (lower) and X12 = -X11 U12 X22 (upper), from the
 * inverses of its two b x b halves.  Outside w the factor is the identity. */
__device__ void tri_inverses(const float (*s_lu)[TW + 1], int w, float (*s_xl)[TW + 1],
                             float (*s_xu)[TW + 1], float *s_t) {
  const int tid = threadIdx.x;
  for (int e = tid; e < TW * TW; e += TR) {
    const int i = e / TW, j = e % TW;
    s_xl[i][j] = (i == j) ? 1.0f : 0.0f;
    s_xu[i][j] = (i == j) ? ((i < w) ? 1.0f / s_lu[i][i] : 1.0f) : 0.0f;
  }
  __syncthreads();
  for (int b = 1; b < TW; b *= 2) {
    const int nblk = TW / (2 * b), per = b * b;
    /* T = L21 X11 (lower) into s_t[0..], T' = U12 X22 (upper) into s_t[TW*TW/2..] */
    for (int e = tid; e < 2 * nblk * per; e += TR) {
      const int up = e >= nblk * per, f = e - (up ? nblk * per : 0);
      const int q = f / per, i = (f % per) / b, j = f % b, o = q * 2 * b;
      float acc = 0.0f;
      if (!up) {
        for (int l = j; l < b; ++l) {
          const int gr = o + b + i, gc = o + l;
          const float lv = (gr < w && gc < w) ? s_lu[gr][gc] : 0.0f;
          acc += lv * s_xl[o + l][o + j];
        }
        s_t[f] = acc;
      } else {
        for (int l = 0; l <= j; ++l) {
          const int gr = o + i, gc = o + b + l;
          const float uv = (gr < w && gc < w) ? s_lu[gr][gc] : 0.0f;
          acc += uv * s_xu[o + b + l][o + b + j];
        }
        s_t[TW * TW / 2 + f] = acc;
      }
    }
    __syncthreads();
    for (int e = tid; e < 2 * nblk * per; e += TR) {
      const int up = e >= nblk * per, f = e - (up ? nblk * per : 0);
      const int q = f / per, i = (f % per) / b, j = f % b, o = q * 2 * b;
      float acc = 0.0f;
      if (!up) {
        for (int l = 0; l <= i; ++l) acc += s_xl[o + b + i][o + b + l] * s_t[q * per + l * b + j];
        s_xl[o + b + i][o + j] = -acc;
      } else {
        for (int l = i; l < b; ++l)
          acc += s_xu[o + i][o + l] * s_t[TW * TW / 2 + q * per + l * b + j];
        s_xu[o + i][o + b + j] = -acc;
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(TR) void tslu_select_kernel(TsluArgs a) {
  __shared__ __attribute__((aligned(16))) f4 s_row[2][4][TW / 4];
  __shared__ float s_val[2][4];
  __shared__ int s_thr[2][4];
  __shared__ int s_sel[TW];
  __shared__ int s_go;
  __shared__ int s_aux[2 * TW];
  if (__hip_atomic_load(a.info, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
  const int tid = threadIdx.x;
  const int g = blockIdx.x;
  const int m = a.n - a.k;

  /* the tree: level 0 = leaves, each level FAN times smaller; nodes are
   * numbered level by level.  lsz/loff describe the current level. */
  int lsz = a.nleaves, loff = 0;

  /* ---- leaf: rows k + g*TR + tid ---- */
  f4 r[TW / 4];
  int myrow = a.k + g * TR + tid;
  if (myrow >= a.n) myrow = -1;
  load_row(r, a, myrow);
  int node = g, mypos = -1;
  for (;;) {
    gepp_select(r, myrow, s_row, s_val, s_thr, s_sel, &mypos);
    if (lsz == 1) break;                          /* root: s_sel is the answer */
    /* publish this node's choice, then count in at the parent */
    if (tid < TW) stc_i(&a.cand[(size_t)(loff + node) * TW + tid], tid < a.w ? s_sel[tid] : -1);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    const int parent = node / FAN;
    const int nchild = min(FAN, lsz - parent * FAN);
    const int poff = loff + lsz;
    if (tid == 0) {
      const int old = atomicAdd(&a.cnt[poff + parent], 1);
      s_go = (old == nchild - 1);
    }
    __syncthreads();
    if (!s_go) return;                            /* a later sibling merges */
    /* ---- merge: the original rows of the children's choices ---- */
    const int child = tid / TW, slot = tid % TW;
    myrow = -1;
    if (child < nchild)
      myrow = ldc_i(&a.cand[(size_t)(loff + parent * FAN + child) * TW + slot]);
    node = parent;
    loff = poff;
    lsz = (lsz + FAN - 1) / FAN;
    load_row(r, a, myrow);
  }

  /* ---- root: the diagonal block's factor and inverses ---- */
  const int ws = min(a.w, m);                     /* pivots actually chosen */
  __shared__ float s_lu[TW][TW + 1], s_xl[TW][TW + 1], s_xu[TW][TW + 1];
  __shared__ float s_t[TW * TW];
  for (int e = tid; e < TW * TW; e += TR) s_lu[e / TW][e % TW] = (e / TW == e % TW) ? 1.0f : 0.0f;
  __syncthreads();
  if (mypos >= 0 && mypos < ws) {
#pragma unroll
    for (int q = 0; q < TW / 4; ++q)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (4 * q + e < ws) s_lu[mypos][4 * q + e] = r[q][e];
  }
  if (tid == 0) s_aux[0] = 0x7fffffff;
  __syncthreads();
  if (tid < ws) {
    const float d = s_lu[tid][tid];
    if (!(fabsf(d) > 0.0f && fabsf(d) <= FLT_MAX)) atomicMin(&s_aux[0], tid);
  }
  __syncthreads();
  if (tid == 0 && s_aux[0] < ws) atomicCAS(a.info, 0, a.k + s_aux[0] + 1);
  tri_inverses(s_lu, ws, s_xl, s_xu, s_t);
  for (int e = tid; e < TW * TW; e += TR) {
    const int i = e % TW, j = e / TW;             /* column-major (i, j) */
    a.lu11[e] = s_lu[i][j];
    a.invl[e] = s_xl[i][j];
    a.invu[e] = s_xu[i][j];
  }

  /* ---- the move list that brings the chosen rows to k..k+ws-1 ---- */
  int *s_mark = s_aux;                            /* [TW] */
  __syncthreads();
  if (tid < TW) s_mark[tid] = 0;
  __syncthreads();
  const int sj = (tid < ws) ? s_sel[tid] : -1;
  if (sj >= a.k && sj < a.k + ws) s_mark[sj - a.k] = 1;
  __syncthreads();
  if (tid < 64) {
    /* displaced occupants D (unchosen rows of the block) pair with the
     * vacated rows V (chosen rows outside the block), both in order */
    const bool inD = (tid < ws) && !s_mark[tid];
    const bool inV = (tid < ws) && (sj < a.k || sj >= a.k + ws);
    const unsigned long long bd = __ballot(inD), bv = __ballot(inV);
    const unsigned long long below = (1ULL << tid) - 1;
    const int rd = __popcll(bd & below), rv = __popcll(bv & below);
    /* dst k+tid <- src sj for chosen rows not already in place */
    const bool mv = (tid < ws) && (sj != a.k + tid);
    const unsigned long long bm = __ballot(mv);
    const int nm = __popcll(bm), rm = __popcll(bm & below);
    if (mv) { a.plan[2 + 2 * rm] = a.k + tid; a.plan[3 + 2 * rm] = sj; }
    if (inV) s_aux[TW + rv] = sj;                 /* V in order */
    /* s_aux[TW..] is written and read by this one wavefront */
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (inD) { a.plan[2 + 2 * (nm + rd)] = s_aux[TW + rd]; a.plan[3 + 2 * (nm + rd)] = a.k + tid; }
    if (tid == 0) a.plan[0] = nm + __popcll(bd);
  }
  __syncthreads();
  /* perm follows the same moves (gather: read every source, then write) */
  const int np = a.plan[0];
  int pdst = -1, pval = 0;
  if (tid < np) {
    pdst = a.plan[2 + 2 * tid];
    pval = a.perm[a.plan[3 + 2 * tid]];
  }
  __syncthreads();
  if (pdst >= 0) a.perm[pdst] = pval;
  /* every child has counted in at every node by now: reset for the next panel
   * (loff is the root's number = the count of nodes below it) */
  for (int i = tid; i <= loff; i += TR) a.cnt[i] = 0;
}

/* S columns dst <- src (rows of A) for every move, over the entries
 * [ebeg, nrows) of each column: each workgroup moves a 64-entry slice of all
 * listed columns through LDS, so sources are read before any destination is
 * written.  A slice inside the diagonal block (rows and entries k..k+w-1)
 * then receives the block's factor lu11. */
__global__ __launch_bounds__(256) void tslu_swap_kernel(float *S, long lds, int nrows,
                                                        const int *plan, int k, int w,
                                                        const float *lu11, const int *info,
                                                        int ebeg) {
  __shared__ float buf[2 * TW][64];
  __shared__ int s_dst[2 * TW], s_src[2 * TW];
  if (*info) return;
  const int np = plan[0];
  const int e0 = ebeg + blockIdx.x * 64;
  for (int t = threadIdx.x; t < np; t += 256) {
    s_dst[t] = plan[2 + 2 * t];
    s_src[t] = plan[3 + 2 * t];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < np * 64; q += 256) {
    const int mv = q >> 6, e = e0 + (q & 63);
    if (e < nrows) buf[mv][q & 63] = S[(size_t)s_src[mv] * lds + e];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < np * 64; q += 256) {
    const int mv = q >> 6, e = e0 + (q & 63);
    if (e < nrows) S[(size_t)s_dst[mv] * lds + e] = buf[mv][q & 63];
  }
  if (e0 + 64 > k && e0 < k + w) {
    __syncthreads();
    for (int q = threadIdx.x; q < w * 64; q += 256) {
      const int r = q >> 6, e = e0 + (q & 63);       /* row k+r of A, entry e */
      if (e >= k && e < k + w) S[(size_t)(k + r) * lds + e] = lu11[(size_t)(e - k) * TW + r];
    }
  }
}

/* S (n x n, FP32) = A^T for the n x n FP64 A, through 64 x 64 LDS tiles;
 * also perm = 0..n-1 */
__global__ __launch_bounds__(256) void tslu_convert_kernel(int n, const double *A, long lda,
                                                           float *S, long lds, int *perm) {
  __shared__ float t[64][65];
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int jj = ty; jj < 64; jj += 4) {
    const int i = i0 + tx, j = j0 + jj;
    t[jj][tx] = (i < n && j < n) ? (float)A[(size_t)j * lda + i] : 0.0f;
  }
  __syncthreads();
  for (int ii = ty; ii < 64; ii += 4) {
    const int i = i0 + ii, j = j0 + tx;
    if (i < n && j < n) S[(size_t)i * lds + j] = t[tx][ii];
  }
  if (blockIdx.y == 0 && ty == 0 && i0 + tx < n) perm[i0 + tx] = i0 + tx;
}

/* Y (n x nrhs, FP32) = rows perm of X (FP64) */
__global__ __launch_bounds__(256) void tslu_gather_kernel(int n, int nrhs, const double *X,
                                                          long ldx, const int *perm, float *Y,
                                                          long ldy) {
  const int i = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y;
  if (i < n && k < nrhs) Y[(size_t)k * ldy + i] = (float)X[(size_t)k * ldx + perm[i]];
}

}  // namespace

extern "C" size_t atlmx_k_tslu_ws_ints(int n) {
  /* candidates for every tree node + counters, for the largest panel */
  const int leaves = (n + TR - 1) / TR;
  int nodes = leaves, sz = leaves;
  while (sz > 1) { sz = (sz + FAN - 1) / FAN; nodes += sz; }
  return (size_t)nodes * TW + (size_t)nodes + 2 + 4 * TW + 64;
}

extern "C" hipError_t atlmx_k_tslu_select(const float *S, long lds, int n, int k, int w,
                                          int *ws, int *plan, int *perm, float *lu11,
                                          float *invl, float *invu, int *info, hipStream_t s) {
  if (k >= n || w <= 0) return hipSuccess;
  if (w > TW || lds < n || (k & 3) || (lds & 3) || ((uintptr_t)S & 15) || !ws || !plan ||
      !perm || !lu11 || !invl || !invu || !info)
    return hipErrorInvalidValue;
  TsluArgs a;
  a.lu11 = lu11; a.invl = invl; a.invu = invu; a.info = info;
  a.S = S; a.lds = lds; a.k = k; a.n = n; a.w = min(w, n - k);
  a.nleaves = (n - k + TR - 1) / TR;
  {
    /* counters first; the tree of the first panel (k = 0) is the largest */
    const int leaves_max = (n + TR - 1) / TR;
    int nmax = leaves_max, szm = leaves_max;
    while (szm > 1) { szm = (szm + FAN - 1) / FAN; nmax += szm; }
    a.cnt = ws;
    a.cand = ws + nmax + 2;
  }
  a.plan = plan; a.perm = perm;
  hipLaunchKernelGGL(tslu_select_kernel, dim3(a.nleaves), dim3(TR), 0, s, a);
  return hipGetLastError();
}

extern "C" hipError_t atlmx_k_tslu_swap(float *S, long lds, int ebeg, int eend, const int *plan,
                                        int k, int w, const float *lu11, const int *info,
                                        hipStream_t s) {
  if (eend <= ebeg) return hipSuccess;
  if (ebeg < 0 || w < 0 || w > TW || !info) return hipErrorInvalidValue;
  hipLaunchKernelGGL(tslu_swap_kernel, dim3((eend - ebeg + 63) / 64), dim3(256), 0, s, S, lds,
                     eend, plan, k, w, lu11, info, ebeg);
  return hipGetLastError();
}

extern "C" hipError_t atlmx_k_tslu_convert(int n, const double *A, long lda, float *S, long lds,
                                           int *perm, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (lda < n || lds < n) return hipErrorInvalidValue;
  const int nb = (n + 63) / 64;
  hipLaunchKernelGGL(tslu_convert_kernel, dim3(nb, nb), dim3(256), 0, s, n, A, lda, S, lds,
                     perm);
  return hipGetLastError();
}

extern "C" hipError_t atlmx_k_tslu_gather(int n, int nrhs, const double *X, long ldx,
                                          const int *perm, float *Y, long ldy, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (ldx < n || ldy < n || nrhs > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(tslu_gather_kernel, dim3((n + 255) / 256, nrhs), dim3(256), 0, s, n, nrhs,
                     X, ldx, perm, Y, ldy);
  return hipGetLastError();
}

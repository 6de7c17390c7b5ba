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
 * elimination with partial pivoting on 256 rows of the panel and keeps
 * the W rows it chose; every tree node does the same on the ORIGINAL values
 * of its children's chosen rows (up to 4 x W).  The root's choice, in its
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
constexpr int NWV = TR / 64;       /* wavefronts per workgroup */
constexpr int FAN = TR / TW;       /* children per tree node: 4 */

__device__ inline int ldc_i(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void stc_i(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int CTRL, int RM>
__device__ __forceinline__ unsigned dpp_max(unsigned v) {
  const unsigned o = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, RM, 0xf, false);
  return v > o ? v : o;
}

/* maximum over the wavefront, broadcast */
__device__ __forceinline__ unsigned wave_max(unsigned v) {
  v = dpp_max<0x111, 0xf>(v);
  v = dpp_max<0x112, 0xf>(v);
  v = dpp_max<0x114, 0xf>(v);
  v = dpp_max<0x118, 0xf>(v);
  v = dpp_max<0x142, 0xa>(v);
  v = dpp_max<0x143, 0xc>(v);
  return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
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
  const int *prev;  /* the previous panel's move list, or null */
  int *plan2;       /* out (with prev): prev's moves followed by these, as one list */
  float *Sw;        /* with sw_e1 > sw_e0: apply the moves to S's entries */
  int sw_e0, sw_e1; /*   [sw_e0, sw_e1) and write the factored block, at the root */
  int u12_r1;       /* > k + w: S rows [k + w, u12_r1) get U12 = Ut inv(L11)^T */
};

/* row `row` of A, columns k..k+w-1 (zero beyond w, or for row < 0) */
__device__ inline void load_row(f4 (&r)[TW / 4], const TsluArgs &a, int row) {
  const float *src = a.S + a.k + (size_t)(row < 0 ? 0 : row) * a.lds;
  if (row >= 0 && a.w == TW) {
#pragma unroll
    for (int q = 0; q < TW / 4; ++q) r[q] = *reinterpret_cast<const f4 *>(src + 4 * q);
  } else {
    /* the bound as a per-lane value, so that its 64 tests stay here as
     * vector compares instead of being hoisted into scalar registers */
    int ww;
    asm volatile("v_mov_b32 %0, %1" : "=v"(ww) : "v"(row >= 0 ? a.w : 0));
#pragma unroll
    for (int q = 0; q < TW / 4; ++q)
#pragma unroll
      for (int e = 0; e < 4; ++e) r[q][e] = (4 * q + e < ww) ? src[4 * q + e] : 0.0f;
  }
}

/* Gaussian elimination with partial pivoting on the rows held by the
 * workgroup (one per thread, in registers as float4), recording in s_sel the
 * rows chosen as pivots, in order.  Each step ranks a row by the integer
 * bits(|entry|) + 1 (0 for a row already chosen, 1 for NaN: never better
 * than a number), which orders like |entry|; every wavefront takes the
 * maximum (ties: lowest lane), its owner stores the key
 * rank << 32 | (TR - 1 - tid) and its row in the wavefront's slot, and
 * after one barrier every thread takes the largest key -- the largest
 * entry, ties to the lower thread -- and reads the winner's row.  The
 * multiplier is the entry times the hardware reciprocal of the pivot
 * (within an ulp of the quotient), which keeps the IEEE division's
 * dependent chain off every step.  Each
 * thread notes the step that chose its row; s_sel is written from those
 * after the loop (-1 where no row was left).  The multipliers are kept in
 * place, so a row retired at column c holds its L entries in r[0..c) and its
 * U entries in r[c..TW); *mypos = c.  All TW columns are eliminated:
 * columns past the panel width hold zeros, so their steps change nothing
 * that is used (their choices are ignored), and the loop body has no branch
 * that would make the compiler copy r. */
__device__ void gepp_select(f4 (&r)[TW / 4], int myrow, f4 (*s_row)[NWV][TW / 4],
                            unsigned long long (*s_key)[NWV], int *s_sel, int *mypos) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  bool act = myrow >= 0;
  int pos = -1;
  if (tid < TW) s_sel[tid] = -1;
#pragma unroll
  for (int c = 0; c < TW; ++c) {
    const int par = c & 1;
    const float rc = r[c / 4][c % 4];
    const float av = fabsf(rc);
    const unsigned u = act ? ((av == av) ? __float_as_uint(av) + 1u : 1u) : 0u;
    const unsigned wm = wave_max(u);
    const unsigned long long bal = __ballot(u == wm);
    const int wl = __ffsll((long long)bal) - 1;
    if (lane == wl) {                             /* stores only */
      s_key[par][wave] = wm ? ((unsigned long long)wm << 32) | (unsigned)(TR - 1 - tid) : 0ull;
#pragma unroll
      for (int q = c / 4; q < TW / 4; ++q) s_row[par][wave][q] = r[q];
    }
    __syncthreads();
    unsigned long long kv[NWV];
#pragma unroll
    for (int q = 0; q < NWV; ++q) kv[q] = s_key[par][q];
#pragma unroll
    for (int h = NWV / 2; h > 0; h /= 2)
#pragma unroll
      for (int q = 0; q < h; ++q) kv[q] = kv[q] > kv[q + h] ? kv[q] : kv[q + h];
    const unsigned long long key = kv[0];
    const bool any = key != 0ull;                 /* some row was still active */
    const int wt = TR - 1 - (int)(unsigned)(key & 0xffffffffull);
    const bool me = any && tid == wt;
    pos = me ? c : pos;
    act = act && !me;
    const int bw = any ? (wt >> 6) : 0;
    const float piv = s_row[par][bw][c / 4][c % 4];
    const float l = (act && piv != 0.0f) ? rc * __builtin_amdgcn_rcpf(piv) : 0.0f;
    r[c / 4][c % 4] = act ? l : rc;
#pragma unroll
    for (int q = c / 4; q < TW / 4; ++q) {
      const f4 pr = s_row[par][bw][q];
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (4 * q + e > c) r[q][e] -= l * pr[e];
    }
  }
  if (pos >= 0) s_sel[pos] = myrow;
  *mypos = pos;
  __syncthreads();
}

/* inverses of the unit lower and the upper triangle of the factor in s_lu
 * (row-major, ld TW+1, the identity outside the leading w x w block) by
 * recursive doubling: for b = 16, 32 each 2b x 2b diagonal block gets
 * its off-diagonal block of the inverse, X21 = -X22 L21 X11 (lower) and
 * X12 = -X11 U12 X22 (upper), from the inverses of its two b x b halves.
 * The sums run over all b terms (the triangles' zeros make up the rest), so
 * every loop has a fixed trip count. */
template <int B>
__device__ __forceinline__ void tri_level(const float (*s_lu)[TW + 1], float (*s_xl)[TW + 1],
                                          float (*s_xu)[TW + 1], float *s_t) {
  constexpr int nblk = TW / (2 * B), per = B * B, tot = 2 * nblk * per;
  const int tid = threadIdx.x;
  /* T = L21 X11 (lower) into s_t[0..], T' = U12 X22 (upper) into s_t[TW*TW/2..] */
  for (int e = tid; e < tot; e += TR) {
    const bool up = e >= nblk * per;
    const int f = e - (up ? nblk * per : 0);
    const int q = f / per, i = (f % per) / B, j = f % B, o = q * 2 * B;
    float acc = 0.0f;
    if (!up) {
#pragma unroll
      for (int l = 0; l < B; ++l) acc += s_lu[o + B + i][o + l] * s_xl[o + l][o + j];
      s_t[f] = acc;
    } else {
#pragma unroll
      for (int l = 0; l < B; ++l) acc += s_lu[o + i][o + B + l] * s_xu[o + B + l][o + B + j];
      s_t[TW * TW / 2 + f] = acc;
    }
  }
  __syncthreads();
  for (int e = tid; e < tot; e += TR) {
    const bool up = e >= nblk * per;
    const int f = e - (up ? nblk * per : 0);
    const int q = f / per, i = (f % per) / B, j = f % B, o = q * 2 * B;
    float acc = 0.0f;
    if (!up) {
#pragma unroll
      for (int l = 0; l < B; ++l) acc += s_xl[o + B + i][o + B + l] * s_t[q * per + l * B + j];
      s_xl[o + B + i][o + j] = -acc;
    } else {
#pragma unroll
      for (int l = 0; l < B; ++l)
        acc += s_xu[o + i][o + l] * s_t[TW * TW / 2 + q * per + l * B + j];
      s_xu[o + i][o + B + j] = -acc;
    }
  }
  __syncthreads();
}

/* The 16 x 16 diagonal blocks first, by substitution (lane = one column of
 * one block: threads 0..63 the unit lower blocks, 64..127 the upper ones,
 * the upper diagonal applied as its reciprocal), then the doubling levels
 * 16 and 32 above them. */
__device__ void tri_inverses(const float (*s_lu)[TW + 1], int w, float (*s_xl)[TW + 1],
                             float (*s_xu)[TW + 1], float *s_t) {
  const int tid = threadIdx.x;
  for (int e = tid; e < TW * TW; e += TR) {
    const int i = e / TW, j = e % TW;
    s_xl[i][j] = 0.0f;
    s_xu[i][j] = 0.0f;
  }
  __syncthreads();
  if (tid < 128) {
    const int q = (tid >> 4) & 3, cc = tid & 15, o = 16 * q;
    float x[16];
    if (tid < 64) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        float acc = (i == cc) ? 1.0f : 0.0f;
#pragma unroll
        for (int kk = 0; kk < i; ++kk) acc -= s_lu[o + i][o + kk] * x[kk];
        x[i] = (i >= cc) ? acc : 0.0f;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) s_xl[o + i][o + cc] = x[i];
    } else {
#pragma unroll
      for (int i = 15; i >= 0; --i) {
        const float d = (o + i < w) ? 1.0f / s_lu[o + i][o + i] : 1.0f;
        float acc = (i == cc) ? 1.0f : 0.0f;
#pragma unroll
        for (int kk = i + 1; kk < 16; ++kk) acc -= s_lu[o + i][o + kk] * x[kk];
        x[i] = (i <= cc) ? acc * d : 0.0f;
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) s_xu[o + i][o + cc] = x[i];
    }
  }
  __syncthreads();
  tri_level<16>(s_lu, s_xl, s_xu, s_t);
  tri_level<32>(s_lu, s_xl, s_xu, s_t);
}

__global__ __launch_bounds__(TR) void tslu_select_kernel(TsluArgs a) {
  __shared__ f4 s_row[2][NWV][TW / 4];
  __shared__ unsigned long long s_key[2][NWV];
  __shared__ int s_sel[TW];
  __shared__ int s_go;
  __shared__ int s_aux[2 * TW];
  __shared__ float s_lu[TW][TW + 1], s_xl[TW][TW + 1], s_xu[TW][TW + 1];
  __shared__ float s_t[TW * TW];
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
    gepp_select(r, myrow, s_row, s_key, s_sel, &mypos);
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
  /* the chosen rows (all TW columns, staged in s_xl) -> s_lu, the identity
   * outside the leading ws x ws block */
  if (mypos >= 0) {
#pragma unroll
    for (int q = 0; q < TW / 4; ++q)
#pragma unroll
      for (int e = 0; e < 4; ++e) s_xl[mypos][4 * q + e] = r[q][e];
  }
  __syncthreads();
  for (int e = tid; e < TW * TW; e += TR) {
    const int i = e / TW, j = e % TW;
    s_lu[i][j] = (i < ws && j < ws) ? s_xl[i][j] : (i == j ? 1.0f : 0.0f);
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
  if (a.prev) {
    /* one list for the previous moves (P) followed by these (Q): a column Q
     * writes takes Q's source as it stood after P, i.e. P's source where P
     * wrote it; a column only P writes keeps P's move */
    int *qd = reinterpret_cast<int *>(s_t), *qs = qd + 2 * TW;
    int *pd = qs + 2 * TW, *ps = pd + 2 * TW;
    const int npp = a.prev[0];
    if (tid < np) { qd[tid] = a.plan[2 + 2 * tid]; qs[tid] = a.plan[3 + 2 * tid]; }
    if (tid < npp) { pd[tid] = a.prev[2 + 2 * tid]; ps[tid] = a.prev[3 + 2 * tid]; }
    if (tid == 0) s_aux[0] = np;
    __syncthreads();
    if (tid < np) {
      const int s0 = qs[tid];
      int src = s0;
      for (int j = 0; j < npp; ++j) src = (pd[j] == s0) ? ps[j] : src;
      a.plan2[2 + 2 * tid] = qd[tid];
      a.plan2[3 + 2 * tid] = src;
    }
    if (tid < npp) {
      bool inq = false;
      for (int j = 0; j < np; ++j) inq = inq || (qd[j] == pd[tid]);
      if (!inq) {
        const int slot = atomicAdd(&s_aux[0], 1);
        a.plan2[2 + 2 * slot] = pd[tid];
        a.plan2[3 + 2 * slot] = ps[tid];
      }
    }
    __syncthreads();
    if (tid == 0) a.plan2[0] = s_aux[0];
  }
  if (a.sw_e1 > a.sw_e0) {
    /* the moves on S's entries [sw_e0, sw_e1) (at most 2 TW), in two halves
     * of 64 entries (entries are independent): each thread holds 32
     * (move, entry) values in registers, all loads of a half are issued
     * together and complete before the barrier, then every destination is
     * written.  The factored block then overwrites rows k..k+ws-1 at
     * entries k..k+ws-1 (as tslu_swap_kernel with lu11). */
    int *s_mv = reinterpret_cast<int *>(&s_xu[0][0]);   /* inverses already stored */
    __syncthreads();
    for (int t = tid; t < np; t += TR) {
      s_mv[2 * t] = a.plan[2 + 2 * t];
      s_mv[2 * t + 1] = a.plan[3 + 2 * t];
    }
    __syncthreads();
    const int ne = a.sw_e1 - a.sw_e0;
    for (int h = 0; h < ne; h += 64) {
      float v[32];
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const int idx = tid + TR * i, mv = idx >> 6, e = h + (idx & 63);
        v[i] = (mv < np && e < ne) ? a.Sw[(size_t)s_mv[2 * mv + 1] * a.lds + a.sw_e0 + e]
                                   : 0.0f;
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const int idx = tid + TR * i, mv = idx >> 6, e = h + (idx & 63);
        if (mv < np && e < ne) a.Sw[(size_t)s_mv[2 * mv] * a.lds + a.sw_e0 + e] = v[i];
      }
    }
    __syncthreads();
    for (int t = tid; t < ws * TW; t += TR) {
      const int r = t / TW, e = t % TW;            /* row k+r of A, entry k+e */
      if (e < ws) a.Sw[(size_t)(a.k + r) * a.lds + a.k + e] = s_lu[r][e];
    }
    if (a.u12_r1 > a.k + ws) {
      /* U12 on the next panel's S rows [k + ws, u12_r1) (moved above):
       * Ut(i, j) <- sum_t Ut(i, t) invL(j, t), a 4 x 4 tile per thread */
      const int r0 = a.k + ws, nr = a.u12_r1 - r0;
      float (*ut)[TW + 1] = s_lu;                  /* the factor is written out */
      __syncthreads();
      for (int t = tid; t < TW * TW; t += TR) {
        const int i = t % TW, j = t / TW;          /* S row r0+i, column k+j */
        ut[i][j] = (i < nr && j < ws) ? a.Sw[(size_t)(a.k + j) * a.lds + r0 + i] : 0.0f;
      }
      __syncthreads();
      const int ti = tid >> 4, tj = tid & 15;
      float acc[4][4];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = 0.0f;
#pragma unroll 4
      for (int t = 0; t < TW; ++t) {
        float u[4], v[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) u[x] = ut[4 * ti + x][t];
#pragma unroll
        for (int y = 0; y < 4; ++y) v[y] = s_xl[4 * tj + y][t];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y) acc[x][y] += u[x] * v[y];
      }
#pragma unroll
      for (int y = 0; y < 4; ++y)
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          const int i = 4 * ti + x, j = 4 * tj + y;
          if (i < nr && j < ws) a.Sw[(size_t)(a.k + j) * a.lds + r0 + i] = acc[x][y];
        }
    }
  }
}

/* A value as the 16-bit update operand it becomes (round to nearest even):
 * lp = 1 bfloat16, lp = 2 IEEE half, else unchanged. */
__device__ inline float as_lp(float x, int lp) {
  if (lp == 1) {
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return __uint_as_float((u | 0x400000u) & 0xffff0000u);
    return __uint_as_float(((u + 0x7fffu + ((u >> 16) & 1u)) >> 16) << 16);
  }
  if (lp == 2) return (float)(_Float16)x;
  return x;
}

/* After the pair's first panel P (S rows k..k+63) is chosen and its block
 * factored: for the S columns [kq, n) (A's rows below P's block), 64 per
 * workgroup, L21 = inv(U11)^T Lt on S rows k..k+63, then the second panel's
 * S rows [kq, kq + wq) -= U12 L21 (U12 already formed on those rows at S
 * columns k..k+63), both 64 x 64 x 64 products in registers (a 4 x 4 tile
 * per thread).  A workgroup touches only its own columns, so the two steps
 * need no ordering between workgroups.  lp != 0 rounds the update's
 * operands to 16 bits as the low-precision update kernel does. */
__global__ __launch_bounds__(256) void tslu_l21up_kernel(float *S, long lds, int n, int k,
                                                         int wq, const float *invu, int lp,
                                                         const int *info) {
  __shared__ float s_iu[TW][TW + 1];   /* inv(U11)(t, r) at [t][r] */
  __shared__ float s_b[TW][TW + 1];    /* Lt, then L21: [t][column] */
  __shared__ float s_u[TW][TW + 1];    /* U12 on the second panel's rows: [i][t] */
  if (*info) return;
  const int kq = k + TW;
  const int j0 = kq + blockIdx.x * 64, nc = min(64, n - j0);
  const int tid = threadIdx.x, ti = tid >> 4, tj = tid & 15;
  for (int e = tid; e < TW * TW; e += 256) {
    const int t = e % TW, c = e / TW;
    s_iu[t][c] = invu[(size_t)c * TW + t];
    s_b[t][c] = (c < nc) ? S[(size_t)(j0 + c) * lds + k + t] : 0.0f;
    s_u[c][t] = (c < wq) ? as_lp(S[(size_t)(k + t) * lds + kq + c], lp) : 0.0f;
  }
  __syncthreads();
  float acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = 0.0f;
#pragma unroll 4
  for (int t = 0; t < TW; ++t) {
    float u[4], v[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) u[x] = s_iu[t][4 * ti + x];
#pragma unroll
    for (int y = 0; y < 4; ++y) v[y] = s_b[t][4 * tj + y];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y) acc[x][y] += u[x] * v[y];
  }
  __syncthreads();                       /* every Lt value read: s_b takes L21 */
#pragma unroll
  for (int y = 0; y < 4; ++y)
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int r = 4 * ti + x, c = 4 * tj + y;
      if (c < nc) S[(size_t)(j0 + c) * lds + k + r] = acc[x][y];
      s_b[r][c] = as_lp(acc[x][y], lp);
    }
  __syncthreads();
  float cv[4][4];
#pragma unroll
  for (int y = 0; y < 4; ++y)
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int i = 4 * ti + x, c = 4 * tj + y;
      cv[x][y] = (i < wq && c < nc) ? S[(size_t)(j0 + c) * lds + kq + i] : 0.0f;
      acc[x][y] = 0.0f;
    }
#pragma unroll 4
  for (int t = 0; t < TW; ++t) {
    float u[4], v[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) u[x] = s_u[4 * ti + x][t];
#pragma unroll
    for (int y = 0; y < 4; ++y) v[y] = s_b[t][4 * tj + y];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y) acc[x][y] += u[x] * v[y];
  }
#pragma unroll
  for (int y = 0; y < 4; ++y)
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int i = 4 * ti + x, c = 4 * tj + y;
      if (i < wq && c < nc) S[(size_t)(j0 + c) * lds + kq + i] = cv[x][y] - acc[x][y];
    }
}

/* S columns dst <- src (rows of A) for every move, over the entries
 * [ebeg, eend) of each column except the gap [gbeg, gbeg + glen): each
 * workgroup moves a 64-entry slice of all listed columns through LDS, so
 * sources are read before any destination is written.  With lu11 given, the
 * part of the diagonal block (rows and entries k..k+w-1) in the slice then
 * receives the block's factor. */
__global__ __launch_bounds__(256) void tslu_swap_kernel(float *S, long lds, int ebeg, int eend,
                                                        int gbeg, int glen, const int *plan,
                                                        int k, int w, const float *lu11,
                                                        const int *info) {
  __shared__ float buf[2 * TW][64];
  __shared__ int s_dst[2 * TW], s_src[2 * TW];
  if (*info) return;
  const int np = plan[0];
  const int i0 = ebeg + blockIdx.x * 64;
  const int j = threadIdx.x & 63;
  /* entry of slice position j (positions past the gap shift over it) */
  int e = i0 + j;
  if (e >= gbeg) e += glen;
  const bool live = e < eend;
  for (int t = threadIdx.x; t < np; t += 256) {
    s_dst[t] = plan[2 + 2 * t];
    s_src[t] = plan[3 + 2 * t];
  }
  __syncthreads();
  for (int m0 = threadIdx.x >> 6; m0 < np; m0 += 4 * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int mv = m0 + 4 * u;
      v[u] = (live && mv < np) ? S[(size_t)s_src[mv] * lds + e] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (m0 + 4 * u < np) buf[m0 + 4 * u][j] = v[u];
  }
  __syncthreads();
  for (int mv = threadIdx.x >> 6; mv < np; mv += 4)
    if (live) S[(size_t)s_dst[mv] * lds + e] = buf[mv][j];
  /* the block's factor replaces what the moves put into rows k..k+w-1 */
  __syncthreads();
  if (lu11 && live && e >= k && e < k + w)
    for (int r = threadIdx.x >> 6; r < w; r += 4)   /* row k+r of A, entry e */
      S[(size_t)(k + r) * lds + e] = lu11[(size_t)(e - k) * TW + r];
}

/* The rest of a panel pair P (columns k..k+TW-1 of A) and Q (the next TW)
 * once both were factored within the pair's own rows of S, for the S rows
 * (columns of A) outside the pair, [0, k) and [k2, n), k2 = k + 2 TW, in
 * blocks of 64: the combined row moves of P and Q (plan, at most 4 TW
 * moves); then, for rows from k2, with U = those rows' entries in S
 * columns k..k2-1 (A's U12 block of the pair, transposed):
 *   U_P <- U_P inv(L_P)^T, U_Q <- U_Q - U_P B, U_Q <- U_Q inv(L_Q)^T
 * with B = S[k.., kq..] (P's L21 restricted to Q's rows).  Everything is
 * per S row, so the blocks are independent.  The moved values of columns
 * k..k2-1 are taken from LDS, and those columns are written once, at the
 * end. */
struct TailArgs {
  float *S;
  long lds;
  int n, k, nlo;           /* nlo = blocks below k */
  const int *plan;
  const float *invlp, *invlq;  /* TW x TW column-major */
  const int *info;
};

__global__ __launch_bounds__(256) void tslu_tail_kernel(TailArgs t) {
  constexpr int NB = 64;                          /* S rows per block */
  __shared__ float buf[4 * TW][NB];               /* moved values, then operands */
  __shared__ float U[NB][2 * TW + 1];
  __shared__ int s_dst[4 * TW], s_src[4 * TW], s_map[2 * TW];
  if (*t.info) return;
  const int tid = threadIdx.x;
  const int k = t.k, kq = k + TW, k2 = k + 2 * TW, n = t.n;
  const bool hi = (int)blockIdx.x >= t.nlo;
  const int r0 = hi ? k2 + ((int)blockIdx.x - t.nlo) * NB : (int)blockIdx.x * NB;
  const int rend = hi ? n : k;
  const int nr = min(NB, rend - r0);
  float *S = t.S;
  const long lds = t.lds;
  const int np = t.plan[0];
  for (int q = tid; q < np; q += 256) {
    s_dst[q] = t.plan[2 + 2 * q];
    s_src[q] = t.plan[3 + 2 * q];
  }
  if (tid < 2 * TW) s_map[tid] = -1;
  __syncthreads();
  if (hi)
    for (int q = tid; q < np; q += 256)
      if (s_dst[q] >= k && s_dst[q] < k2) s_map[s_dst[q] - k] = q;
  /* every source is read before any destination is written; eight loads
   * per thread are in flight before their LDS stores */
  for (int q0 = tid; q0 < np * NB; q0 += 256 * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = q0 + 256 * u, mv = q / NB, i = q % NB;
      v[u] = (q < np * NB && i < nr) ? S[(size_t)s_src[mv] * lds + r0 + i] : 0.0f;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = q0 + 256 * u;
      if (q < np * NB) buf[q / NB][q % NB] = v[u];
    }
  }
  __syncthreads();
  for (int q = tid; q < np * NB; q += 256) {
    const int mv = q / NB, i = q % NB, d = s_dst[mv];
    if (i < nr && !(hi && d >= k && d < k2)) S[(size_t)d * lds + r0 + i] = buf[mv][i];
  }
  if (!hi) return;
  /* U: the pair's columns of these rows, after the moves */
  for (int q0 = tid; q0 < 2 * TW * NB; q0 += 256 * 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = q0 + 256 * u, j = q / NB, i = q % NB, mv = s_map[j];
      v[u] = (i >= nr) ? 0.0f : (mv >= 0 ? buf[mv][i] : S[(size_t)(k + j) * lds + r0 + i]);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = q0 + 256 * u;
      U[q % NB][q / NB] = v[u];
    }
  }
  __syncthreads();
  /* operands: inv(L_P) as [j][m] = invlp(j, m), B as [m][j], inv(L_Q) */
  float (*iLP)[TW + 1] = reinterpret_cast<float (*)[TW + 1]>(&buf[0][0]);
  float (*Bm)[TW + 1] = iLP + TW;
  float (*iLQ)[TW + 1] = Bm + TW;
  for (int q0 = tid; q0 < TW * TW; q0 += 256 * 4) {
    float x[4], y[4], z[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = q0 + 256 * u, a = q % TW, b = q / TW;   /* column-major (a, b) */
      x[u] = t.invlp[q];
      y[u] = t.invlq[q];
      z[u] = S[(size_t)(kq + b) * lds + k + a];             /* B(a, b) = S[k+a, kq+b] */
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int q = q0 + 256 * u, a = q % TW, b = q / TW;
      iLP[a][b] = x[u];
      iLQ[a][b] = y[u];
      Bm[a][b] = z[u];
    }
  }
  __syncthreads();
  /* thread: rows i0..i0+3, columns j0..j0+3 of each 64 x 64 product */
  const int i0 = (tid & 15) * 4, j0 = (tid >> 4) * 4;
  float acc[4][4];
  /* U_P <- U_P inv(L_P)^T */
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0f;
  for (int m = 0; m < TW; ++m) {
    float x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) x[a] = U[i0 + a][m];
#pragma unroll
    for (int b = 0; b < 4; ++b) y[b] = iLP[j0 + b][m];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += x[a] * y[b];
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) U[i0 + a][j0 + b] = acc[a][b];
  __syncthreads();
  /* U_Q <- U_Q - U_P B */
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = U[i0 + a][TW + j0 + b];
  for (int m = 0; m < TW; ++m) {
    float x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) x[a] = U[i0 + a][m];
#pragma unroll
    for (int b = 0; b < 4; ++b) y[b] = Bm[m][j0 + b];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] -= x[a] * y[b];
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) U[i0 + a][TW + j0 + b] = acc[a][b];
  __syncthreads();
  /* U_Q <- U_Q inv(L_Q)^T */
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0f;
  for (int m = 0; m < TW; ++m) {
    float x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) x[a] = U[i0 + a][TW + m];
#pragma unroll
    for (int b = 0; b < 4; ++b) y[b] = iLQ[j0 + b][m];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += x[a] * y[b];
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) U[i0 + a][TW + j0 + b] = acc[a][b];
  __syncthreads();
  for (int q = tid; q < 2 * TW * NB; q += 256) {
    const int j = q / NB, i = q % NB;
    if (i < nr) S[(size_t)(k + j) * lds + r0 + i] = U[i][j];
  }
}

/* S (n x n, FP32) = A^T for the n x n FP64 A (column j of A times
 * colscale[j] when given), through 64 x 64 LDS tiles; also perm = 0..n-1 */
__global__ __launch_bounds__(256) void tslu_convert_kernel(int n, const double *A, long lda,
                                                           float *S, long lds, int *perm,
                                                           const double *colscale) {
  __shared__ float t[64][65];
  const int i0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int jj = ty; jj < 64; jj += 4) {
    const int i = i0 + tx, j = j0 + jj;
    t[jj][tx] = (i < n && j < n)
                    ? (float)(colscale ? A[(size_t)j * lda + i] * colscale[j] : A[(size_t)j * lda + i])
                    : 0.0f;
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

extern "C" hipError_t atlmx_k_tslu_select_swap(float *S, long lds, int n, int k, int w,
                                               int *ws, int *plan, int *perm, float *lu11,
                                               float *invl, float *invu, int *info,
                                               const int *prev, int *plan2, int sw_e0,
                                               int sw_e1, int u12_r1, hipStream_t s) {
  if (k >= n || w <= 0) return hipSuccess;
  const int wk = min(w, n - k);
  if (sw_e1 > sw_e0 && (sw_e0 < 0 || sw_e1 > n || sw_e0 > k || sw_e1 - sw_e0 > 2 * TW ||
                        k + wk > sw_e1))
    return hipErrorInvalidValue;
  if (u12_r1 > k + wk && (sw_e1 <= sw_e0 || u12_r1 > sw_e1 || u12_r1 - (k + wk) > TW))
    return hipErrorInvalidValue;
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
  if (prev && !plan2) return hipErrorInvalidValue;
  a.plan = plan; a.perm = perm; a.prev = prev; a.plan2 = plan2;
  a.Sw = S; a.sw_e0 = sw_e0; a.sw_e1 = sw_e1; a.u12_r1 = u12_r1;
  hipLaunchKernelGGL(tslu_select_kernel, dim3(a.nleaves), dim3(TR), 0, s, a);
  return hipGetLastError();
}

extern "C" hipError_t atlmx_k_tslu_select(const float *S, long lds, int n, int k, int w,
                                          int *ws, int *plan, int *perm, float *lu11,
                                          float *invl, float *invu, int *info,
                                          const int *prev, int *plan2, hipStream_t s) {
  return atlmx_k_tslu_select_swap(const_cast<float *>(S), lds, n, k, w, ws, plan, perm, lu11,
                                  invl, invu, info, prev, plan2, 0, 0, 0, s);
}

extern "C" hipError_t atlmx_k_tslu_swap(float *S, long lds, int ebeg, int eend, int gbeg,
                                        int gend, const int *plan, int k, int w,
                                        const float *lu11, const int *info, hipStream_t s) {
  if (gend <= gbeg) gbeg = gend = eend;
  if (ebeg < 0 || gbeg < ebeg || gend > eend || w < 0 || w > TW || !info || !plan)
    return hipErrorInvalidValue;
  const int cnt = eend - ebeg - (gend - gbeg);
  if (cnt <= 0) return hipSuccess;
  hipLaunchKernelGGL(tslu_swap_kernel, dim3((cnt + 63) / 64), dim3(256), 0, s, S, lds, ebeg,
                     eend, gbeg, gend - gbeg, plan, k, w, lu11, info);
  return hipGetLastError();
}

extern "C" hipError_t atlmx_k_tslu_tail(float *S, long lds, int n, int k, const int *plan,
                                        const float *invlp, const float *invlq, const int *info,
                                        hipStream_t s) {
  const int k2 = k + 2 * TW;
  if (k < 0 || k2 > n || !plan || !invlp || !invlq || !info) return hipErrorInvalidValue;
  if ((k | (int)lds) < 0 || lds < n) return hipErrorInvalidValue;
  TailArgs t;
  t.S = S; t.lds = lds; t.n = n; t.k = k;
  t.nlo = (k + 63) / 64;
  t.plan = plan; t.invlp = invlp; t.invlq = invlq; t.info = info;
  const int nb = t.nlo + (n - k2 + 63) / 64;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(tslu_tail_kernel, dim3(nb), dim3(256), 0, s, t);
  return hipGetLastError();
}

extern "C" hipError_t atlmx_k_tslu_convert(int n, const double *A, long lda, float *S, long lds,
                                           int *perm, const double *colscale, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (lda < n || lds < n) return hipErrorInvalidValue;
  const int nb = (n + 63) / 64;
  hipLaunchKernelGGL(tslu_convert_kernel, dim3(nb, nb), dim3(256), 0, s, n, A, lda, S, lds,
                     perm, colscale);
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

extern "C" hipError_t atlmx_k_tslu_l21up(float *S, long lds, int n, int k, int wq,
                                         const float *invu, int lp, const int *info,
                                         hipStream_t s) {
  const int kq = k + TW;
  if (kq >= n || wq <= 0) return hipSuccess;
  if (wq > TW || kq + wq > n || lds < n || !S || !invu || !info) return hipErrorInvalidValue;
  hipLaunchKernelGGL(tslu_l21up_kernel, dim3((n - kq + 63) / 64), dim3(256), 0, s, S, lds, n, k,
                     wq, invu, lp, info);
  return hipGetLastError();
}

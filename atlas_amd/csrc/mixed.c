/*
 * mixed.c -- FP64-accurate dense solve from a low-precision LU plus FP64
 * iterative refinement (no ATLAS counterpart).
 *
 * The LU is done in FP32 on the MFMA GEMM; only the residual r = b - A*x,
 * the update x += d and the stopping test are in FP64.  The stopping test is
 * the normwise backward error ||r||_inf / (||A||_inf ||x||_inf + ||b||_inf)
 * <= n * 2^-52 for every right-hand side (or, on request, LAPACK dsgesv's
 * ||r||_inf <= ||x||_inf ||A||_inf 2^-53 sqrt(n)).
 *
 * First choice: the FP32 LU with tournament pivoting of S = A^T
 * (k_tslu.hip), whose panels need one reduction tree instead of one
 * device-wide exchange per column and whose row interchanges are contiguous
 * column moves of S.  If its plain refinement stalls, GMRES-IR with the same
 * factors; if that fails or a pivot is zero, the FP32 LU with partial
 * pivoting of A itself (plain, then GMRES), and failing that the FP64 LU.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "rt.h"
#include "kernels_mp.h"
#include "../../include/atlmx.h"

int dgetrf_dev(atlmx_ctx *c, int M, int N, double *A, long lda, int *d_ipiv, const char *rout);
int sgetrf_dev(atlmx_ctx *c, int M, int N, float *A, long lda, int *d_ipiv, const char *rout);
void dgetrs_dev(atlmx_ctx *c, int trans, int N, int NRHS, const double *A, long lda,
                const int *d_ipiv, double *B, long ldb, const char *rout);
void sgetrs_dev(atlmx_ctx *c, int trans, int N, int NRHS, const float *A, long lda,
                const int *d_ipiv, float *B, long ldb, const char *rout);
void sgemm_dev(atlmx_ctx *c, int ta, int tb, int M, int N, int K, float alpha,
               const float *A, long lda, const float *B, long ldb, float beta, float *C,
               long ldc, int tri, const int *abort_flag, const char *rout);
void strsm_dev(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N,
               float alpha, const float *A, long lda, float *B, long ldb,
               const int *abort_flag, const char *rout);
int spotrf_dev(atlmx_ctx *c, int uplo, int N, float *A, long lda, const char *rout);
void spotrs_dev(atlmx_ctx *c, int uplo, int N, int NRHS, const float *A, long lda, float *B,
                long ldb, const char *rout);
int dpotrf_dev(atlmx_ctx *c, int uplo, int N, double *A, long lda, const char *rout);
void dpotrs_dev(atlmx_ctx *c, int uplo, int N, int NRHS, const double *A, long lda, double *B,
                long ldb, const char *rout);
void dgemm_dev(atlmx_ctx *c, int ta, int tb, int M, int N, int K, double alpha,
               const double *A, long lda, const double *B, long ldb, double beta, double *C,
               long ldc, int tri, const int *abort_flag, const char *rout);

enum { MP_AUTO = 0, MP_TSLU = 1, MP_GEPP = 2, MP_FP64 = 3, MP_CHOL = 4, MP_BF16 = 5, MP_FP16 = 6 };
static int mp_method = MP_AUTO;
static int mp_stop_lapack = 0;

/* Stopping test of the refinement: 0 (default) the normwise backward error
 * ||r|| / (||A|| ||x|| + ||b||) <= n * 2^-52 (the accuracy the solver
 * promises); 1 LAPACK dsgesv's ||r|| <= ||x|| ||A|| 2^-53 sqrt(n). */
void atlmx_set_refine_stop(int lapack)
{
  mp_stop_lapack = (lapack != 0);
}

/* converged: residual norm rn of a column with solution norm xn, right-hand
 * side norm bn; cte = ||A|| 2^-53 sqrt(n) */
static int ir_done(double rn, double xn, double bn, double anrm, double cte, int n)
{
  if (mp_stop_lapack) return rn <= xn * cte;
  return rn <= (double)n * 2.220446049250313e-16 * (anrm * xn + bn);
}

/* 0: tournament-pivoting FP32 LU, then partial-pivoting FP32 LU, then
 * FP64; 1: tournament only (then FP64); 2: partial pivoting only (then FP64). */
void atlmx_set_mixed_method(int m)
{
  mp_method = (m == MP_TSLU || m == MP_GEPP || m == MP_BF16 || m == MP_FP16) ? m : MP_AUTO;
}

static double host_max(atlmx_ctx *c, const double *d, int n, const char *rout)
{
  double *h = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double m = 0.0;
  int i;
  RT_CHECK(rout, hipMemcpyAsync(h, d, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost,
                                c->stream));
  rt_sync(c, rout);
  for (i = 0; i < n; ++i) if (!(h[i] <= m)) m = h[i];
  free(h);
  return m;
}

/* S22 -= Ut Lt on the rows [r0, r0+M) x columns [c0, c0+N) of S with the
 * K = columns k0.. of Ut (S rows, A's U12^T) and rows k0.. of Lt (A's
 * L21^T): the operands are S's own blocks, so this is one short-K update */
static void tslu_update(atlmx_ctx *c, float *S, long lds, int r0, int M, int c0, int N,
                        int k0, int K, int lp, const int *d_info, hipStream_t st,
                        const char *rout)
{
  int pid;
  if (M <= 0 || N <= 0 || K <= 0) return;
  pid = rt_prof_begin2(c, PROF_UPDATE, 2.0 * M * N * K,
                       ((double)M * K + (double)K * N + 2.0 * M * N) * sizeof(float));
  if (lp)
    RT_CHECK(rout, atlmx_k_supdate_lp(lp, M, N, K, -1.0f, S + r0 + (size_t)k0 * lds, lds,
                                      S + k0 + (size_t)c0 * lds, lds,
                                      S + r0 + (size_t)c0 * lds, lds, d_info, st));
  else
    RT_CHECK(rout, atlmx_k_supdate(M, N, K, -1.0f, S + r0 + (size_t)k0 * lds, lds,
                                   S + k0 + (size_t)c0 * lds, lds, S + r0 + (size_t)c0 * lds,
                                   lds, d_info, st));
  rt_prof_end(c, pid);
}

/* The trailing update that runs beside a pair head, as two launches when it
 * is large: the first takes only as many 128 x 128 tiles as leave one
 * workgroup slot free for each of the head's tournament leaves (the update
 * kernel fits two workgroups per compute unit and then fills every
 * register; a tournament workgroup fits beside one), so the tournament
 * starts at once instead of waiting for update workgroups to retire; the
 * second launch takes the remaining rows. */
static void tslu_update_beside(atlmx_ctx *c, float *S, long lds, int r0, int M, int c0, int N,
                               int k0, int K, int lp, int leaves, const int *d_info,
                               hipStream_t st, const char *rout)
{
  const int tiles_n = (N + 127) / 128;
  const int cap = 2 * c->ncu - leaves;
  const int m1 = (tiles_n > 0 && cap > 0) ? (cap / tiles_n) * 128 : 0;
  if (m1 > 0 && M > m1) {
    tslu_update(c, S, lds, r0, m1, c0, N, k0, K, lp, d_info, st, rout);
    tslu_update(c, S, lds, r0 + m1, M - m1, c0, N, k0, K, lp, d_info, st, rout);
  } else {
    tslu_update(c, S, lds, r0, M, c0, N, k0, K, lp, d_info, st, rout);
  }
}

/* what a panel's tournament leaves for the later steps: the row moves, the
 * factored diagonal block and the inverses of its triangles */
typedef struct {
  int *plan;
  float *lu11, *invL, *invU;
} tslu_panel;

/* choose the pivot rows of the panel at k (width w); with prev (the
 * previous panel), plan2 receives prev's moves followed by these.  The
 * moves on S's rows [e0, e1) (A's columns) and the factored block are
 * applied by the same launch (e1 > e0). */
static void tslu_select(atlmx_ctx *c, float *S, long lds, int n, int k, int w, int *tws,
                        const tslu_panel *p, int *perm, const tslu_panel *prev, int *plan2,
                        int e0, int e1, int u12_r1, const int *d_info, const char *rout)
{
  int pid = rt_prof_begin(c, PROF_PANEL, (double)(n - k) * w * sizeof(float));
  RT_CHECK(rout, atlmx_k_tslu_select_swap(S, lds, n, k, w, tws, p->plan, perm, p->lu11,
                                          p->invL, p->invU, (int *)d_info,
                                          prev ? prev->plan : NULL, plan2, e0, e1, u12_r1,
                                          c->stream));
  rt_prof_end(c, pid);
}

/* the panel's row moves over the S rows [e0, e1) (columns of A) except
 * [g0, g1), with the factored diagonal block when overlay is set */
static void tslu_swap(atlmx_ctx *c, float *S, long lds, int e0, int e1, int g0, int g1, int k,
                      int w, const tslu_panel *p, int overlay, const int *d_info,
                      const char *rout)
{
  int pid = rt_prof_begin(c, PROF_LASWP, 0.0);
  RT_CHECK(rout, atlmx_k_tslu_swap(S, lds, e0, e1, g0, g1, p->plan, k, w,
                                   overlay ? p->lu11 : NULL, d_info, c->stream));
  rt_prof_end(c, pid);
}

/* L21 = A21 inv(U11) for the panel at k: in S, the rows k..k+w-1 right of
 * the block, Lt = inv(U11)^T Lt (an in-place GEMM covering all w rows in
 * one tile row, so every workgroup reads its operand in full before writing
 * it) */
static void tslu_l21(atlmx_ctx *c, float *S, long lds, int n, int k, int w,
                     const tslu_panel *p, const int *d_info, const char *rout)
{
  const int m = n - k - w;
  float *Lt = S + k + (size_t)(k + w) * lds;
  if (m > 0)
    sgemm_dev(c, KX_TRANS, KX_NOTRANS, w, m, w, 1.0f, p->invU, ATLMX_TSLU_W, Lt, lds, 0.0f,
              Lt, lds, 0, d_info, rout);
}

/* P's L21 on S rows k..k+W-1 and P's update of Q's S rows [k+W, k+W+wq),
 * over every S column from k+W, in one launch */
static void tslu_l21up(atlmx_ctx *c, float *S, long lds, int n, int k, int wq,
                       const tslu_panel *p, int lp, const int *d_info, const char *rout)
{
  const double m = (double)(n - k - ATLMX_TSLU_W);
  int pid = rt_prof_begin2(c, PROF_GEMM, 2.0 * m * ATLMX_TSLU_W * (ATLMX_TSLU_W + wq),
                           (m * (2.0 * ATLMX_TSLU_W + 2.0 * wq)) * sizeof(float));
  RT_CHECK(rout, atlmx_k_tslu_l21up(S, lds, n, k, wq, p->invU, lp, d_info, c->stream));
  rt_prof_end(c, pid);
}

/* the pair of panels P at k and Q after it, confined to the pair's own S
 * rows [k, k2) (A's columns): P's pivots, its moves and factored block on
 * those rows, its L21 and its U12 on Q's rows; Q's columns get P's update;
 * then Q's pivots, moves, block and L21.  Nothing outside rows [k, k2) of S
 * is read or written, so this runs beside the trailing update of the rows
 * below. */
static void tslu_pair_head(atlmx_ctx *c, float *S, long lds, int n, int k, int *tws,
                           const tslu_panel *P, const tslu_panel *Q, int *planPQ, int *perm,
                           int lp, hipEvent_t evq, const int *d_info, const char *rout)
{
  const int W = ATLMX_TSLU_W;
  const int wp = (n - k < W) ? n - k : W, kq = k + wp;
  const int wq = (n - kq < W) ? n - kq : W, k2 = kq + wq;
  /* P's tournament also moves the pair's rows and forms U12 on Q's rows */
  tslu_select(c, S, lds, n, k, wp, tws, P, perm, NULL, NULL, k, k2, k2, d_info, rout);
  if (wq == 0) {
    tslu_l21(c, S, lds, n, k, wp, P, d_info, rout);
    if (evq) RT_CHECK(rout, hipEventRecord(evq, c->stream));
    return;
  }
  tslu_l21up(c, S, lds, n, k, wq, P, lp, d_info, rout);   /* L21(P), Q's rows updated */
  tslu_select(c, S, lds, n, kq, wq, tws, Q, perm, P, planPQ, k, k2, 0, d_info, rout);
  /* what the pair's tail needs is ready here; Q's L21 is needed only by the
   * next update */
  if (evq) RT_CHECK(rout, hipEventRecord(evq, c->stream));
  tslu_l21(c, S, lds, n, kq, wq, Q, d_info, rout);
}

/* the rest of the pair at k, for the S rows outside [k, k2): P's and then
 * Q's moves (one combined list), P's U12 on the rows from k2, those rows' Q
 * columns updated by P (after Q's moves, so each moved row carries the
 * updates of the row it replaces), then Q's U12 -- one launch, as every
 * step is per S row.  With no Q (the last panel) there are no rows past
 * the pair and only P's moves remain. */
static void tslu_pair_tail(atlmx_ctx *c, float *S, long lds, int n, int k,
                           const tslu_panel *P, const tslu_panel *Q, const int *planPQ,
                           const int *d_info, const char *rout)
{
  const int W = ATLMX_TSLU_W;
  const int wp = (n - k < W) ? n - k : W, kq = k + wp;
  const int wq = (n - kq < W) ? n - kq : W, k2 = kq + wq;
  int pid;
  if (wq < W) {                                   /* the pair ends the matrix */
    tslu_swap(c, S, lds, 0, k, 0, 0, k, wp, P, 0, d_info, rout);
    if (wq > 0) tslu_swap(c, S, lds, 0, k, 0, 0, kq, wq, Q, 0, d_info, rout);
    return;
  }
  pid = rt_prof_begin2(c, PROF_LASWP, 2.0 * (n - k2) * 2 * W * 2 * W,
                       ((double)n * 4 * W + 2.0 * (n - k2) * 2 * W) * 2 * sizeof(float));
  RT_CHECK(rout, atlmx_k_tslu_tail(S, lds, n, k, planPQ, P->invL, Q->invL, d_info, c->stream));
  rt_prof_end(c, pid);
  (void)k2;
}

/* FP32 LU with tournament pivoting of A, held as S = A^T (column-major,
 * leading dimension lds), in pairs of W = ATLMX_TSLU_W column panels P, Q
 * whose row interchanges reach the rest of the matrix late:
 *   head (tslu_pair_head): both panels factored within the pair's own rows
 *     of S (A's columns k..k2-1): tournament pivots, moves and factored
 *     block, L21, U12 where Q needs it, P's update of Q's columns;
 *   tail (tslu_pair_tail): the moves on the other rows, U12 there, Q's
 *     columns of those rows updated by P.
 * The pair's tail waits for the head only through Q's tournament (Q's
 * L21, which only the next update reads, overlaps it).
 * The next pair's rows then get this pair's update (K = 2W), and its head
 * runs on a second stream while the rest of the trailing matrix gets the
 * same update on the first: the tournaments leave the critical path.  In
 * S, A22 -= L21 U12 reads S22 -= U12^T L21^T.  On return PA = LU with row
 * i of PA = row perm[i] of A; L (unit) is the strict upper triangle of S
 * transposed, U the lower triangle of S transposed.  Returns 0 or 1 + the
 * first column with a zero/non-finite pivot (every launch after that
 * returns at once). */
static int sgetrf_tslu_dev(atlmx_ctx *c, int n, float *S, long lds, int *perm, int lp,
                           const char *rout)
{
  const int W = ATLMX_TSLU_W;
  const size_t wsi = atlmx_k_tslu_ws_ints(n);
  float *blk = (float *)rt_ws(c, WS_NPINV, sizeof(float) * 6 * W * W + 256, rout);
  int *d_info = (int *)(blk + (size_t)6 * W * W);
  int *tws = (int *)rt_ws(c, WS_TSLU, sizeof(int) * (wsi + 2 * (2 + 4 * W) + 2 + 8 * W), rout);
  tslu_panel P = {tws + wsi, blk, blk + (size_t)W * W, blk + (size_t)2 * W * W};
  tslu_panel Q = {tws + wsi + 2 + 4 * W, blk + (size_t)3 * W * W, blk + (size_t)4 * W * W,
                  blk + (size_t)5 * W * W};
  int *planPQ = tws + wsi + 2 * (2 + 4 * W);
  hipStream_t main = c->stream, aux = rt_aux(c, rout);
  int info = 0, k = 0;
  RT_CHECK(rout, atlmx_k_fill_i32(d_info, 1, 0, main));
  RT_CHECK(rout, atlmx_k_fill_i32(tws, (int)wsi, 0, main));
  tslu_pair_head(c, S, lds, n, 0, tws, &P, &Q, planPQ, perm, lp, NULL, d_info, rout);
  for (;;) {
    const int k2 = (n - k <= 2 * W) ? n : k + 2 * W;
    const int R = (n - k2 < 2 * W) ? n - k2 : 2 * W;
    tslu_pair_tail(c, S, lds, n, k, &P, &Q, planPQ, d_info, rout);
    if (k2 >= n) break;
    /* the head of this pair ran on aux (but the first): its last launch,
     * Q's L21, overlapped the tail above; the update below reads it */
    if (k > 0) RT_CHECK(rout, hipStreamWaitEvent(main, c->ev[1], 0));
    tslu_update(c, S, lds, k2, R, k2, n - k2, k, k2 - k, lp, d_info, main, rout);
    RT_CHECK(rout, hipEventRecord(c->ev[0], main));
    RT_CHECK(rout, hipStreamWaitEvent(aux, c->ev[0], 0));
    rt_on_aux(c, main, 1);
    tslu_pair_head(c, S, lds, n, k2, tws, &P, &Q, planPQ, perm, lp, c->ev[2], d_info, rout);
    rt_on_aux(c, main, 0);
    RT_CHECK(rout, hipEventRecord(c->ev[1], aux));
    tslu_update_beside(c, S, lds, k2 + R, n - k2 - R, k2, n - k2, k, k2 - k, lp,
                       (n - k2 + 255) / 256, d_info, main, rout);
    RT_CHECK(rout, hipStreamWaitEvent(main, c->ev[2], 0));   /* head through Q's choice */
    k = k2;
  }
  if (k > 0) RT_CHECK(rout, hipStreamWaitEvent(main, c->ev[1], 0));
  RT_CHECK(rout, hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, main));
  rt_sync(c, rout);
  return info;
}

/* a low-precision factorization and how to apply its inverse */
typedef struct {
  int kind;                 /* MP_TSLU, MP_GEPP or MP_CHOL */
  int n;
  float *As;                /* FP32 factor: S = A^T (MP_TSLU) or A, leading dim lda */
  long lda;
  const int *ipiv;          /* MP_GEPP: LAPACK interchanges */
  const int *perm;          /* MP_TSLU: row permutation */
  float *Rs;                /* FP32 right-hand-side staging, n x MP_MAXRHS */
  const double *colscale;   /* MP_TSLU: the factor is of A diag(colscale), or NULL */
} lp_fact;

/* X (=, or += when accumulate) M^-1 R for nb <= MP_MAXRHS columns, with the
 * transforms in FP64 and the triangular solves in FP32 */
static void lp_solve(atlmx_ctx *c, const lp_fact *f, int nb, const double *R, long ldr,
                     double *X, long ldx, int accumulate, const char *rout)
{
  if (f->kind == MP_TSLU) {
    /* PA = LU, L = (strict upper of S)^T + I, U = (lower of S)^T */
    RT_CHECK(rout, atlmx_k_tslu_gather(f->n, nb, R, ldr, f->perm, f->Rs, f->n, c->stream));
    strsm_dev(c, KX_LEFT, KX_UPPER, KX_TRANS, KX_UNIT, f->n, nb, 1.0f, f->As, f->lda,
              f->Rs, f->n, NULL, rout);
    strsm_dev(c, KX_LEFT, KX_LOWER, KX_TRANS, KX_NONUNIT, f->n, nb, 1.0f, f->As, f->lda,
              f->Rs, f->n, NULL, rout);
    if (f->colscale)   /* (A D)^-1 = D^-1 A^-1: A^-1 r = D (A D)^-1 r */
      RT_CHECK(rout, atlmx_k_scale_s2d(f->n, nb, f->Rs, f->n, f->colscale, X, ldx, accumulate,
                                       c->stream));
    else if (accumulate)
      RT_CHECK(rout, atlmx_k_axpy_s2d(f->n, nb, f->Rs, f->n, X, ldx, c->stream));
    else
      RT_CHECK(rout, atlmx_k_s2d(f->n, nb, f->Rs, f->n, X, ldx, c->stream));
  } else if (f->kind == MP_CHOL) {
    RT_CHECK(rout, atlmx_k_d2s(f->n, nb, R, ldr, f->Rs, f->n, c->stream));
    spotrs_dev(c, KX_LOWER, f->n, nb, f->As, f->lda, f->Rs, f->n, rout);
    if (accumulate)
      RT_CHECK(rout, atlmx_k_axpy_s2d(f->n, nb, f->Rs, f->n, X, ldx, c->stream));
    else
      RT_CHECK(rout, atlmx_k_s2d(f->n, nb, f->Rs, f->n, X, ldx, c->stream));
  } else {
    RT_CHECK(rout, atlmx_k_d2s(f->n, nb, R, ldr, f->Rs, f->n, c->stream));
    sgetrs_dev(c, KX_NOTRANS, f->n, nb, f->As, f->lda, f->ipiv, f->Rs, f->n, rout);
    if (accumulate)
      RT_CHECK(rout, atlmx_k_axpy_s2d(f->n, nb, f->Rs, f->n, X, ldx, c->stream));
    else
      RT_CHECK(rout, atlmx_k_s2d(f->n, nb, f->Rs, f->n, X, ldx, c->stream));
  }
}

/* FP64 iterative refinement of every right-hand side with one low-precision
 * factorization.  Stops a group when the stopping test holds for all its
 * columns, or when a step fails to halve some column's residual (the
 * factorization is then not good enough).  Returns 1 when all converged. */
static int refine(atlmx_ctx *c, const lp_fact *f, int N, int NRHS, const double *A, int lda,
                  const double *B, int ldb, double *X, int ldx, int iter_max, double anrm,
                  double cte, double *R, double *nrm, void *ws, atlmx_ir_info *st,
                  const char *rout)
{
  double xn[MP_MAXRHS], rn[MP_MAXRHS], bn[MP_MAXRHS], rprev[MP_MAXRHS];
  int r0, nb, it, j, conv = 0;
  st->iterations = 0;
  st->backward_error = 0.0;
  for (r0 = 0; r0 < NRHS; r0 += MP_MAXRHS) {
    const double *Bg = B + (size_t)r0 * ldb;
    double *Xg = X + (size_t)r0 * ldx;
    nb = (NRHS - r0 < MP_MAXRHS) ? NRHS - r0 : MP_MAXRHS;
    RT_CHECK(rout, atlmx_k_dcolmax(N, nb, Bg, ldb, nrm + 64, c->stream));
    RT_CHECK(rout, hipMemcpyAsync(bn, nrm + 64, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                  c->stream));
    lp_solve(c, f, nb, Bg, ldb, Xg, ldx, 0, rout);
    conv = 0;
    for (it = 0; it <= iter_max; ++it) {
      int stalled = 0;
      RT_CHECK(rout, atlmx_k_dresid(N, N, nb, A, lda, Xg, ldx, Bg, ldb, R, N, ws, c->stream));
      RT_CHECK(rout, atlmx_k_dcolmax(N, nb, Xg, ldx, nrm, c->stream));
      RT_CHECK(rout, atlmx_k_dcolmax(N, nb, R, N, nrm + 32, c->stream));
      RT_CHECK(rout, hipMemcpyAsync(xn, nrm, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                    c->stream));
      RT_CHECK(rout, hipMemcpyAsync(rn, nrm + 32, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                    c->stream));
      rt_sync(c, rout);
      conv = 1;
      for (j = 0; j < nb; ++j) {
        const int done = ir_done(rn[j], xn[j], bn[j], anrm, cte, N);
        if (!done) conv = 0;
        if (it > 0 && !(rn[j] <= 0.5 * rprev[j]) && !done) stalled = 1;
        rprev[j] = rn[j];
      }
      if (conv || stalled || it == iter_max) break;
      lp_solve(c, f, nb, R, N, Xg, ldx, 1, rout);
    }
    if (it > st->iterations) st->iterations = it;
    for (j = 0; j < nb; ++j) {
      const double be = rn[j] / (anrm * xn[j] + bn[j]);
      if (!(be <= st->backward_error)) st->backward_error = be;
    }
    if (!conv) return 0;
  }
  return 1;
}

/* <u, v> of two device vectors, split over row blocks and summed in a fixed
 * order (synchronizes) */
static double ddot_dev(atlmx_ctx *c, int n, const double *u, const double *v, double *tmp,
                       void *ws, const char *rout)
{
  double h = 0.0;
  RT_CHECK(rout, atlmx_k_ddots(n, 1, u, n, v, tmp, ws, c->stream));
  RT_CHECK(rout, hipMemcpyAsync(&h, tmp, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  rt_sync(c, rout);
  return h;
}

#define GM_RESTART 100

/* GMRES-based refinement (Carson and Higham, "Accelerating the solution of
 * linear systems by iterative refinement in three precisions", SIAM J. Sci.
 * Comput. 40(2), 2018): each correction d of A d = r is found by GMRES in
 * FP64 on the left-preconditioned system M^{-1} A d = M^{-1} r, M^{-1}
 * applied with the low-precision factorization (classical Gram-Schmidt with
 * one reorthogonalisation, Givens rotations on the host).  It converges for
 * condition numbers well beyond plain refinement's kappa * u_factor < 1.
 * Same stopping test as refine(); returns 1 when every RHS converged. */
static int gmres_refine(atlmx_ctx *c, const lp_fact *f, int N, int NRHS, const double *A,
                        int lda, const double *B, int ldb, double *X, int ldx, int outer_max,
                        double anrm, double cte, double *R, double *nrm, void *ws,
                        atlmx_ir_info *st, const char *rout)
{
  const int m = GM_RESTART;
  double *V = (double *)rt_ws(c, WS_GMRES, sizeof(double) * ((size_t)N * (m + 3) + 4 * m + 64),
                              rout);
  /* hd: the two Gram-Schmidt passes' coefficients and the new vector's
   * squared norm, [0, m+1), [m+1, 2m+2), 2m+2; yd: the correction's weights */
  double *T = V + (size_t)N * (m + 1), *Z = T + N, *hd = Z + N, *yd = hd + 2 * m + 4;
  double H[GM_RESTART + 1][GM_RESTART], g[GM_RESTART + 1], cs[GM_RESTART], sn[GM_RESTART];
  double y[GM_RESTART], hh[2 * GM_RESTART + 3];
  int j, it, k, i, conv = 1;
  RT_CHECK(rout, atlmx_k_fill_i32((int *)Z, 2 * N, 0, c->stream));      /* Z = 0.0 */
  st->backward_error = 0.0;
  for (j = 0; j < NRHS; ++j) {
    const double *b = B + (size_t)j * ldb;
    double *x = X + (size_t)j * ldx;
    double bn = 0.0, xn = 0.0, rn = 0.0;
    int cj = 0, stalled = 0;
    RT_CHECK(rout, atlmx_k_dcolmax(N, 1, b, ldb, nrm + 64, c->stream));
    RT_CHECK(rout, hipMemcpyAsync(&bn, nrm + 64, sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
    for (it = 0; it <= outer_max; ++it) {
      double beta;
      int kk = 0;
      RT_CHECK(rout, atlmx_k_dresid(N, N, 1, A, lda, x, ldx, b, ldb, R, N, ws, c->stream));
      RT_CHECK(rout, atlmx_k_dcolmax(N, 1, x, ldx, nrm, c->stream));
      RT_CHECK(rout, atlmx_k_dcolmax(N, 1, R, N, nrm + 32, c->stream));
      RT_CHECK(rout, hipMemcpyAsync(&xn, nrm, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      RT_CHECK(rout, hipMemcpyAsync(&rn, nrm + 32, sizeof(double), hipMemcpyDeviceToHost,
                                    c->stream));
      rt_sync(c, rout);
      if (ir_done(rn, xn, bn, anrm, cte, N)) { cj = 1; break; }
      if (it == outer_max || stalled) break;
      /* GMRES on M^{-1} A d = M^{-1} r from d = 0 */
      lp_solve(c, f, 1, R, N, V, N, 0, rout);
      beta = sqrt(ddot_dev(c, N, V, V, hd, ws, rout));
      if (!(beta > 0.0) || !(beta < 1e300)) break;
      RT_CHECK(rout, atlmx_k_dscal2d(N, 1, 1.0 / beta, V, N, c->stream));
      memset(g, 0, sizeof(g));
      g[0] = beta;
      for (k = 0; k < m; ++k) {
        double *vk = V + (size_t)k * N, *w = V + (size_t)(k + 1) * N, hn, t;
        /* w = M^{-1} A v_k (T = 0 - A v_k, then negated after the solve) */
        RT_CHECK(rout, atlmx_k_dresid(N, N, 1, A, lda, vk, N, Z, N, T, N, ws, c->stream));
        lp_solve(c, f, 1, T, N, w, N, 0, rout);
        RT_CHECK(rout, atlmx_k_dscal2d(N, 1, -1.0, w, N, c->stream));
        /* classical Gram-Schmidt, twice: h = V^T w (split dot products),
         * then w -= V h (the residual GEMV with w as both B and R: every
         * entry is read and written by one thread, and V's columns 0..k do
         * not include w); one host synchronization per step */
        RT_CHECK(rout, atlmx_k_ddots(N, k + 1, V, N, w, hd, ws, c->stream));
        RT_CHECK(rout, atlmx_k_dresid(N, k + 1, 1, V, N, hd, k + 1, w, N, w, N, ws, c->stream));
        RT_CHECK(rout, atlmx_k_ddots(N, k + 1, V, N, w, hd + (m + 1), ws, c->stream));
        RT_CHECK(rout, atlmx_k_dresid(N, k + 1, 1, V, N, hd + (m + 1), k + 1, w, N, w, N, ws,
                                      c->stream));
        RT_CHECK(rout, atlmx_k_ddots(N, 1, w, N, w, hd + 2 * (m + 1), ws, c->stream));
        RT_CHECK(rout, hipMemcpyAsync(hh, hd, sizeof(double) * (2 * (m + 1) + 1),
                                      hipMemcpyDeviceToHost, c->stream));
        rt_sync(c, rout);
        hn = sqrt(hh[2 * (m + 1)]);
        for (i = 0; i <= k; ++i) H[i][k] = hh[i] + hh[m + 1 + i];
        H[k + 1][k] = hn;
        for (i = 0; i < k; ++i) {                           /* previous rotations */
          t = cs[i] * H[i][k] + sn[i] * H[i + 1][k];
          H[i + 1][k] = -sn[i] * H[i][k] + cs[i] * H[i + 1][k];
          H[i][k] = t;
        }
        t = hypot(H[k][k], H[k + 1][k]);
        cs[k] = (t > 0.0) ? H[k][k] / t : 1.0;
        sn[k] = (t > 0.0) ? H[k + 1][k] / t : 0.0;
        H[k][k] = t;
        H[k + 1][k] = 0.0;
        g[k + 1] = -sn[k] * g[k];
        g[k] = cs[k] * g[k];
        kk = k + 1;
        if (hn > 0.0) RT_CHECK(rout, atlmx_k_dscal2d(N, 1, 1.0 / hn, w, N, c->stream));
        if (!(fabs(g[k + 1]) > 1e-12 * beta) || !(hn > 0.0)) break;
      }
      /* d = V y with H y = g (upper triangular, kk x kk); x += d */
      for (i = kk - 1; i >= 0; --i) {
        double v = g[i];
        for (k = i + 1; k < kk; ++k) v -= H[i][k] * y[k];
        y[i] = v / H[i][i];
      }
      RT_CHECK(rout, hipMemcpyAsync(yd, y, sizeof(double) * kk, hipMemcpyHostToDevice,
                                    c->stream));
      dgemm_dev(c, KX_NOTRANS, KX_NOTRANS, N, 1, kk, 1.0, V, N, yd, kk, 1.0, x, ldx, 0, NULL,
                rout);
      rt_sync(c, rout);                                     /* y lives on the host stack */
      st->inner_iterations += kk;
      /* a whole restart cycle that did not halve the preconditioned
       * residual: the factors are too weak a preconditioner to go on */
      if (kk == m && !(fabs(g[kk]) <= 0.5 * beta)) stalled = 1;
    }
    if (it > st->iterations) st->iterations = it;
    {
      const double be = rn / (anrm * xn + bn);
      if (!(be <= st->backward_error)) st->backward_error = be;
    }
    if (!cj) { conv = 0; break; }
  }
  return conv;
}

int atlmx_dsgesv(int N, int NRHS, const double *A, int lda, const double *B, int ldb,
                 double *X, int ldx, int iter_max, atlmx_ir_info *st)
{
  static const char *rout = "atlmx_dsgesv";
  atlmx_ctx *c = rt_ctx();
  const double eps = 1.1102230246251565e-16;   /* 2^-53, LAPACK dlamch('E') */
  double anrm, cte;
  double xn[MP_MAXRHS], rn[MP_MAXRHS], bn[MP_MAXRHS];
  double *R, *nrm, *ws;
  int info = 0, j, conv = 0, nb, r0;
  lp_fact f;
  atlmx_ir_info loc;
  if (!st) st = &loc;
  memset(st, 0, sizeof(*st));
  if (N <= 0 || NRHS <= 0) return 0;
  if (lda < N || ldb < N || ldx < N) return -1;

  ws = (double *)rt_ws(c, WS_MP3, atlmx_k_resid_ws_bytes(N, MP_MAXRHS), rout);
  nrm = (double *)rt_ws(c, WS_MP4, sizeof(double) * ((size_t)N + 128), rout);
  R = (double *)rt_ws(c, WS_MP2, sizeof(double) * (size_t)N * MP_MAXRHS, rout);
  RT_CHECK(rout, atlmx_k_drowsum(N, N, A, lda, nrm, ws, c->stream));
  anrm = host_max(c, nrm, N, rout);
  cte = anrm * eps * sqrt((double)N);

  memset(&f, 0, sizeof(f));
  f.n = N;
  if (mp_method == MP_BF16 || mp_method == MP_FP16) {
    /* tournament-pivoting LU of S = (A D)^T, D the power-of-two column
     * scaling that brings every column's largest entry into [0.5, 1), with
     * the trailing updates' operands rounded to 16 bits */
    int *perm = (int *)rt_ws(c, WS_IPIV, sizeof(int) * (size_t)N, rout);
    double *csc = (double *)rt_ws(c, WS_MPSCALE, sizeof(double) * (size_t)N, rout);
    f.kind = MP_TSLU;
    f.lda = (N + 3) & ~3;
    f.perm = perm;
    f.colscale = csc;
    f.As = (float *)rt_ws(c, WS_MP0, sizeof(float) * (size_t)f.lda * N, rout);
    f.Rs = (float *)rt_ws(c, WS_MP1, sizeof(float) * (size_t)N * MP_MAXRHS, rout);
    RT_CHECK(rout, atlmx_k_dcolmax(N, N, A, lda, csc, c->stream));
    RT_CHECK(rout, atlmx_k_pow2_scale(N, csc, csc, c->stream));
    RT_CHECK(rout, atlmx_k_tslu_convert(N, A, lda, f.As, f.lda, perm, csc, c->stream));
    info = sgetrf_tslu_dev(c, N, f.As, f.lda, perm, mp_method == MP_BF16 ? 1 : 2, rout);
    st->factor_info = info;
    st->method = mp_method;
    if (!info)
      conv = refine(c, &f, N, NRHS, A, lda, B, ldb, X, ldx, iter_max, anrm, cte, R, nrm, ws,
                    st, rout);
    if (!info && !conv) {
      st->gmres = 1;
      conv = gmres_refine(c, &f, N, NRHS, A, lda, B, ldb, X, ldx, iter_max, anrm, cte, R, nrm,
                          ws, st, rout);
    }
    f.colscale = NULL;
  }
  if (!conv && mp_method != MP_GEPP) {
    /* tournament-pivoting LU of S = A^T */
    int *perm = (int *)rt_ws(c, WS_IPIV, sizeof(int) * (size_t)N, rout);
    f.kind = MP_TSLU;
    f.lda = (N + 3) & ~3;
    f.perm = perm;
    f.As = (float *)rt_ws(c, WS_MP0, sizeof(float) * (size_t)f.lda * N, rout);
    f.Rs = (float *)rt_ws(c, WS_MP1, sizeof(float) * (size_t)N * MP_MAXRHS, rout);
    st->gmres = 0;
    RT_CHECK(rout, atlmx_k_tslu_convert(N, A, lda, f.As, f.lda, perm, NULL, c->stream));
    info = sgetrf_tslu_dev(c, N, f.As, f.lda, perm, 0, rout);
    st->factor_info = info;
    st->method = MP_TSLU;
    if (!info)
      conv = refine(c, &f, N, NRHS, A, lda, B, ldb, X, ldx, iter_max, anrm, cte, R, nrm, ws,
                    st, rout);
    if (!info && !conv) {
      st->gmres = 1;
      conv = gmres_refine(c, &f, N, NRHS, A, lda, B, ldb, X, ldx, iter_max, anrm, cte, R, nrm,
                          ws, st, rout);
    }
  }
  if (!conv && mp_method != MP_TSLU) {
    /* partial-pivoting FP32 LU of A itself */
    int *ipiv = (int *)rt_ws(c, WS_IPIV, sizeof(int) * (size_t)N, rout);
    f.kind = MP_GEPP;
    f.lda = N;
    f.ipiv = ipiv;
    f.As = (float *)rt_ws(c, WS_MP0, sizeof(float) * (size_t)N * N, rout);
    f.Rs = (float *)rt_ws(c, WS_MP1, sizeof(float) * (size_t)N * MP_MAXRHS, rout);
    RT_CHECK(rout, atlmx_k_d2s(N, N, A, lda, f.As, N, c->stream));
    info = sgetrf_dev(c, N, N, f.As, N, ipiv, rout);
    st->factor_info = info;
    st->method = MP_GEPP;
    st->gmres = 0;
    if (!info)
      conv = refine(c, &f, N, NRHS, A, lda, B, ldb, X, ldx, iter_max, anrm, cte, R, nrm, ws,
                    st, rout);
    if (!info && !conv) {
      st->gmres = 1;
      conv = gmres_refine(c, &f, N, NRHS, A, lda, B, ldb, X, ldx, iter_max, anrm, cte, R, nrm,
                          ws, st, rout);
    }
  }
  st->converged = conv;
  if (!conv) {
    /* FP64 fallback on a copy of A (the input stays untouched) */
    double *Aw = (double *)rt_ws(c, WS_MP0, sizeof(double) * (size_t)N * N, rout);
    int *ipiv = (int *)rt_ws(c, WS_IPIV, sizeof(int) * (size_t)N, rout);
    st->fallback = 1;
    st->method = MP_FP64;
    RT_CHECK(rout, hipMemcpy2DAsync(Aw, sizeof(double) * (size_t)N, A, sizeof(double) * lda,
                                    sizeof(double) * (size_t)N, (size_t)N,
                                    hipMemcpyDeviceToDevice, c->stream));
    RT_CHECK(rout, hipMemcpy2DAsync(X, sizeof(double) * ldx, B, sizeof(double) * ldb,
                                    sizeof(double) * (size_t)N, (size_t)NRHS,
                                    hipMemcpyDeviceToDevice, c->stream));
    info = dgetrf_dev(c, N, N, Aw, N, ipiv, rout);
    if (info) return info;
    dgetrs_dev(c, KX_NOTRANS, N, NRHS, Aw, N, ipiv, X, ldx, rout);
    st->backward_error = 0.0;
    for (r0 = 0; r0 < NRHS; r0 += MP_MAXRHS) {
      nb = (NRHS - r0 < MP_MAXRHS) ? NRHS - r0 : MP_MAXRHS;
      RT_CHECK(rout, atlmx_k_dresid(N, N, nb, A, lda, X + (size_t)r0 * ldx, ldx,
                                    B + (size_t)r0 * ldb, ldb, R, N, ws, c->stream));
      RT_CHECK(rout, atlmx_k_dcolmax(N, nb, X + (size_t)r0 * ldx, ldx, nrm, c->stream));
      RT_CHECK(rout, atlmx_k_dcolmax(N, nb, R, N, nrm + 32, c->stream));
      RT_CHECK(rout, atlmx_k_dcolmax(N, nb, B + (size_t)r0 * ldb, ldb, nrm + 64, c->stream));
      RT_CHECK(rout, hipMemcpyAsync(xn, nrm, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                    c->stream));
      RT_CHECK(rout, hipMemcpyAsync(rn, nrm + 32, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                    c->stream));
      RT_CHECK(rout, hipMemcpyAsync(bn, nrm + 64, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                    c->stream));
      rt_sync(c, rout);
      for (j = 0; j < nb; ++j) {
        const double be = rn[j] / (anrm * xn[j] + bn[j]);
        if (!(be <= st->backward_error)) st->backward_error = be;
      }
    }
  }
  rt_sync(c, rout);
  return 0;
}

/* FP64-accurate SPD solve: FP32 Cholesky (lower) + FP64 refinement, the
 * FP64 Cholesky if that does not converge.  A holds the full symmetric
 * matrix (the residual uses all of it; the factorizations read the lower
 * triangle). */
int atlmx_dsposv(int N, int NRHS, const double *A, int lda, const double *B, int ldb,
                 double *X, int ldx, int iter_max, atlmx_ir_info *st)
{
  static const char *rout = "atlmx_dsposv";
  atlmx_ctx *c = rt_ctx();
  const double eps = 1.1102230246251565e-16;   /* 2^-53, LAPACK dlamch('E') */
  double anrm, cte;
  double *R, *nrm, *ws;
  int info = 0, conv = 0;
  lp_fact f;
  atlmx_ir_info loc;
  if (!st) st = &loc;
  memset(st, 0, sizeof(*st));
  if (N <= 0 || NRHS <= 0) return 0;
  if (lda < N || ldb < N || ldx < N) return -1;
  ws = (double *)rt_ws(c, WS_MP3, atlmx_k_resid_ws_bytes(N, MP_MAXRHS), rout);
  nrm = (double *)rt_ws(c, WS_MP4, sizeof(double) * ((size_t)N + 128), rout);
  R = (double *)rt_ws(c, WS_MP2, sizeof(double) * (size_t)N * MP_MAXRHS, rout);
  RT_CHECK(rout, atlmx_k_drowsum(N, N, A, lda, nrm, ws, c->stream));
  anrm = host_max(c, nrm, N, rout);
  cte = anrm * eps * sqrt((double)N);

  memset(&f, 0, sizeof(f));
  f.kind = MP_CHOL;
  f.n = N;
  f.lda = N;
  f.As = (float *)rt_ws(c, WS_MP0, sizeof(float) * (size_t)N * N, rout);
  f.Rs = (float *)rt_ws(c, WS_MP1, sizeof(float) * (size_t)N * MP_MAXRHS, rout);
  RT_CHECK(rout, atlmx_k_d2s(N, N, A, lda, f.As, N, c->stream));
  info = spotrf_dev(c, KX_LOWER, N, f.As, N, rout);
  st->factor_info = info;
  st->method = MP_CHOL;
  if (!info)
    conv = refine(c, &f, N, NRHS, A, lda, B, ldb, X, ldx, iter_max, anrm, cte, R, nrm, ws, st,
                  rout);
  if (!info && !conv) {
    st->gmres = 1;
    conv = gmres_refine(c, &f, N, NRHS, A, lda, B, ldb, X, ldx, iter_max, anrm, cte, R, nrm, ws,
                        st, rout);
  }
  st->converged = conv;
  if (!conv) {
    double *Aw = (double *)rt_ws(c, WS_MP0, sizeof(double) * (size_t)N * N, rout);
    double xn[MP_MAXRHS], rn[MP_MAXRHS], bn[MP_MAXRHS];
    int r0, nb, j;
    st->fallback = 1;
    st->method = MP_FP64;
    RT_CHECK(rout, hipMemcpy2DAsync(Aw, sizeof(double) * (size_t)N, A, sizeof(double) * lda,
                                    sizeof(double) * (size_t)N, (size_t)N,
                                    hipMemcpyDeviceToDevice, c->stream));
    RT_CHECK(rout, hipMemcpy2DAsync(X, sizeof(double) * ldx, B, sizeof(double) * ldb,
                                    sizeof(double) * (size_t)N, (size_t)NRHS,
                                    hipMemcpyDeviceToDevice, c->stream));
    info = dpotrf_dev(c, KX_LOWER, N, Aw, N, rout);
    if (info) return info;
    dpotrs_dev(c, KX_LOWER, N, NRHS, Aw, N, X, ldx, rout);
    st->backward_error = 0.0;
    for (r0 = 0; r0 < NRHS; r0 += MP_MAXRHS) {
      nb = (NRHS - r0 < MP_MAXRHS) ? NRHS - r0 : MP_MAXRHS;
      RT_CHECK(rout, atlmx_k_dresid(N, N, nb, A, lda, X + (size_t)r0 * ldx, ldx,
                                    B + (size_t)r0 * ldb, ldb, R, N, ws, c->stream));
      RT_CHECK(rout, atlmx_k_dcolmax(N, nb, X + (size_t)r0 * ldx, ldx, nrm, c->stream));
      RT_CHECK(rout, atlmx_k_dcolmax(N, nb, R, N, nrm + 32, c->stream));
      RT_CHECK(rout, atlmx_k_dcolmax(N, nb, B + (size_t)r0 * ldb, ldb, nrm + 64, c->stream));
      RT_CHECK(rout, hipMemcpyAsync(xn, nrm, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                    c->stream));
      RT_CHECK(rout, hipMemcpyAsync(rn, nrm + 32, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                    c->stream));
      RT_CHECK(rout, hipMemcpyAsync(bn, nrm + 64, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                    c->stream));
      rt_sync(c, rout);
      for (j = 0; j < nb; ++j) {
        const double be = rn[j] / (anrm * xn[j] + bn[j]);
        if (!(be <= st->backward_error)) st->backward_error = be;
      }
    }
  }
  rt_sync(c, rout);
  return 0;
}

/* Normwise backward error max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf
 * + ||b_j||_inf) of a device-resident solution, in FP64 (synchronizes). */
double atlmx_dbackward_error(int N, int NRHS, const double *A, int lda, const double *B,
                             int ldb, const double *X, int ldx)
{
  static const char *rout = "atlmx_dbackward_error";
  atlmx_ctx *c = rt_ctx();
  double *ws, *nrm, *R, anrm, be = 0.0;
  double xn[MP_MAXRHS], rn[MP_MAXRHS], bn[MP_MAXRHS];
  int r0, nb, j;
  if (N <= 0 || NRHS <= 0) return 0.0;
  ws = (double *)rt_ws(c, WS_MP3, atlmx_k_resid_ws_bytes(N, MP_MAXRHS), rout);
  nrm = (double *)rt_ws(c, WS_MP4, sizeof(double) * ((size_t)N + 128), rout);
  R = (double *)rt_ws(c, WS_MP2, sizeof(double) * (size_t)N * MP_MAXRHS, rout);
  RT_CHECK(rout, atlmx_k_drowsum(N, N, A, lda, nrm, ws, c->stream));
  anrm = host_max(c, nrm, N, rout);
  for (r0 = 0; r0 < NRHS; r0 += MP_MAXRHS) {
    nb = (NRHS - r0 < MP_MAXRHS) ? NRHS - r0 : MP_MAXRHS;
    RT_CHECK(rout, atlmx_k_dresid(N, N, nb, A, lda, X + (size_t)r0 * ldx, ldx,
                                  B + (size_t)r0 * ldb, ldb, R, N, ws, c->stream));
    RT_CHECK(rout, atlmx_k_dcolmax(N, nb, X + (size_t)r0 * ldx, ldx, nrm, c->stream));
    RT_CHECK(rout, atlmx_k_dcolmax(N, nb, R, N, nrm + 32, c->stream));
    RT_CHECK(rout, atlmx_k_dcolmax(N, nb, B + (size_t)r0 * ldb, ldb, nrm + 64, c->stream));
    RT_CHECK(rout, hipMemcpyAsync(xn, nrm, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                  c->stream));
    RT_CHECK(rout, hipMemcpyAsync(rn, nrm + 32, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                  c->stream));
    RT_CHECK(rout, hipMemcpyAsync(bn, nrm + 64, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                  c->stream));
    rt_sync(c, rout);
    for (j = 0; j < nb; ++j) {
      const double e = rn[j] / (anrm * xn[j] + bn[j]);
      if (!(e <= be)) be = e;
    }
  }
  return be;
}

/*
 * mixed.c -- FP64-accurate dense solve from a low-precision LU plus FP64
 * iterative refinement (no ATLAS counterpart).
 *
 * The LU is done in FP32 on the MFMA GEMM (sgetrf of this library); only
 * the residual r = b - A*x, the update x += d and the stopping test are in
 * FP64.  The stopping test is LAPACK dsgesv's: ||r||_inf <= ||x||_inf *
 * ||A||_inf * eps * sqrt(n) for every right-hand side.  If it is not met
 * within iter_max steps (or the FP32 factor is singular) the solve falls
 * back to the FP64 LU.
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

int atlmx_dsgesv(int N, int NRHS, const double *A, int lda, const double *B, int ldb,
                 double *X, int ldx, int iter_max, atlmx_ir_info *st)
{
  static const char *rout = "atlmx_dsgesv";
  atlmx_ctx *c = rt_ctx();
  const double eps = 1.1102230246251565e-16;   /* 2^-53, LAPACK dlamch('E') */
  double anrm, cte;
  double xn[MP_MAXRHS], rn[MP_MAXRHS], bn[MP_MAXRHS];
  float *As, *Rs;
  double *R, *nrm, *ws;
  int *ipiv;
  int info, it, j, conv = 0, nb, r0;
  atlmx_ir_info loc;
  if (!st) st = &loc;
  memset(st, 0, sizeof(*st));
  if (N <= 0 || NRHS <= 0) return 0;
  if (lda < N || ldb < N || ldx < N) return -1;

  ws = (double *)rt_ws(c, WS_MP3, atlmx_k_resid_ws_bytes(N, MP_MAXRHS), rout);
  nrm = (double *)rt_ws(c, WS_MP4, sizeof(double) * ((size_t)N + 128), rout);
  RT_CHECK(rout, atlmx_k_drowsum(N, N, A, lda, nrm, ws, c->stream));
  anrm = host_max(c, nrm, N, rout);
  cte = anrm * eps * sqrt((double)N);

  As = (float *)rt_ws(c, WS_MP0, sizeof(float) * (size_t)N * N, rout);
  ipiv = (int *)rt_ws(c, WS_IPIV, sizeof(int) * (size_t)N, rout);
  Rs = (float *)rt_ws(c, WS_MP1, sizeof(float) * (size_t)N * MP_MAXRHS, rout);
  R = (double *)rt_ws(c, WS_MP2, sizeof(double) * (size_t)N * MP_MAXRHS, rout);

  RT_CHECK(rout, atlmx_k_d2s(N, N, A, lda, As, N, c->stream));
  info = sgetrf_dev(c, N, N, As, N, ipiv, rout);
  st->factor_info = info;

  if (!info) {
    /* the right-hand sides go through in groups of MP_MAXRHS */
    for (r0 = 0; r0 < NRHS; r0 += MP_MAXRHS) {
      const double *Bg = B + (size_t)r0 * ldb;
      double *Xg = X + (size_t)r0 * ldx;
      nb = (NRHS - r0 < MP_MAXRHS) ? NRHS - r0 : MP_MAXRHS;
      RT_CHECK(rout, atlmx_k_dcolmax(N, nb, Bg, ldb, nrm, c->stream));
      RT_CHECK(rout, hipMemcpyAsync(bn, nrm, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                    c->stream));
      RT_CHECK(rout, atlmx_k_d2s(N, nb, Bg, ldb, Rs, N, c->stream));
      sgetrs_dev(c, KX_NOTRANS, N, nb, As, N, ipiv, Rs, N, rout);
      RT_CHECK(rout, atlmx_k_s2d(N, nb, Rs, N, Xg, ldx, c->stream));
      conv = 0;
      for (it = 0; it <= iter_max; ++it) {
        RT_CHECK(rout, atlmx_k_dresid(N, N, nb, A, lda, Xg, ldx, Bg, ldb, R, N, ws,
                                      c->stream));
        RT_CHECK(rout, atlmx_k_dcolmax(N, nb, Xg, ldx, nrm, c->stream));
        RT_CHECK(rout, atlmx_k_dcolmax(N, nb, R, N, nrm + 32, c->stream));
        RT_CHECK(rout, hipMemcpyAsync(xn, nrm, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                      c->stream));
        RT_CHECK(rout, hipMemcpyAsync(rn, nrm + 32, sizeof(double) * nb,
                                      hipMemcpyDeviceToHost, c->stream));
        rt_sync(c, rout);
        conv = 1;
        for (j = 0; j < nb; ++j) if (!(rn[j] <= xn[j] * cte)) conv = 0;
        if (conv || it == iter_max) break;
        RT_CHECK(rout, atlmx_k_d2s(N, nb, R, N, Rs, N, c->stream));
        sgetrs_dev(c, KX_NOTRANS, N, nb, As, N, ipiv, Rs, N, rout);
        RT_CHECK(rout, atlmx_k_axpy_s2d(N, nb, Rs, N, Xg, ldx, c->stream));
      }
      if (it > st->iterations) st->iterations = it;
      for (j = 0; j < nb; ++j) {
        const double be = rn[j] / (anrm * xn[j] + bn[j]);
        if (!(be <= st->backward_error)) st->backward_error = be;
      }
      if (!conv) break;
    }
    st->converged = conv;
  }
  if (info || !conv) {
    /* FP64 fallback on a copy of A (the input stays untouched) */
    double *Aw = (double *)rt_ws(c, WS_MP0, sizeof(double) * (size_t)N * N, rout);
    st->fallback = 1;
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

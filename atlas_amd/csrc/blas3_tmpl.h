/*
 * blas3_tmpl.h -- precision template of the Level-3 BLAS layer (included by
 * blas3.c with T=double/PFX=d and T=float/PFX=s, the way ATLAS builds each
 * routine per precision, makes/Make.ul3:7-23).
 *
 *   ?gemm_dev / ?trsm_dev / ?syrk_dev : stream-ordered drivers on device
 *                                        pointers (column-major)
 *   atlmx_?gemm / _?trsm / _?syrk      : public device-pointer API
 *   cblas_?gemm / _?trsm / _?syrk / cblas_i?amax : the ATLAS C ABI
 */
#define NM_(a, b) a##b
#define NM(a, b) NM_(a, b)
#define NM3_(a, b, c) a##b##c
#define NM3(a, b, c) NM3_(a, b, c)

#define GEMM_DEV NM(PFX, gemm_dev)
#define TRSM_DEV NM(PFX, trsm_dev)
#define SYRK_DEV NM(PFX, syrk_dev)

/* ------------------------------------------------------------------------
 * GEMM: one launch of the MFMA kernel; the shape router of
 * src/blas/ammm/ATL_ammm.c:12-247 (GER/GEMV/rank-K special cases) is not
 * needed because one tile kernel covers every shape, K=0 included.
 * ---------------------------------------------------------------------- */
void GEMM_DEV(atlmx_ctx *c, int ta, int tb, int M, int N, int K, T alpha, const T *A,
              long lda, const T *B, long ldb, T beta, T *C, long ldc, int tri,
              const int *abort_flag, const char *rout)
{
  int pid;
  if (M <= 0 || N <= 0) return;
  if ((alpha == (T)0 || K == 0) && beta == (T)1) return;
  if (alpha == (T)0) K = 0;   /* ATL_drefgemm.c: alpha == 0 only scales C */
  pid = rt_prof_begin2(c, PROF_GEMM, 2.0 * (double)M * N * K,
                      ((double)M * K + (double)K * N +
                       (double)M * N * (beta != (T)0 ? 2 : 1)) * sizeof(T));
  RT_CHECK(rout, K_GEMM(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tri,
                        abort_flag, c->stream));
  rt_prof_end(c, pid);
}

/* C += alpha * A * op(B) with A not transposed (beta = 1): the short-K
 * trailing updates of the factorizations.  The FP64 instantiation has a
 * kernel for this form that loads C behind the MFMAs (K_UPDATE); the other
 * runs the GEMM.  tri as GEMM_DEV. */
void NM(PFX, update_dev)(atlmx_ctx *c, int tb, int M, int N, int K, T alpha, const T *A,
                         long lda, const T *B, long ldb, T *C, long ldc, int tri,
                         const int *abort_flag, const char *rout)
{
#ifdef K_UPDATE
  int pid;
  if (M <= 0 || N <= 0 || K <= 0 || alpha == (T)0) return;
  pid = rt_prof_begin2(c, PROF_GEMM, 2.0 * (double)M * N * K,
                      ((double)M * K + (double)K * N + 2.0 * M * N) * sizeof(T));
  RT_CHECK(rout, K_UPDATE(tb, M, N, K, alpha, A, lda, B, ldb, C, ldc, tri, abort_flag,
                          c->stream));
  rt_prof_end(c, pid);
#else
  GEMM_DEV(c, KX_NOTRANS, tb, M, N, K, alpha, A, lda, B, ldb, (T)1, C, ldc, tri, abort_flag,
           rout);
#endif
}

/* ------------------------------------------------------------------------
 * TRSM: op(A) X = alpha B (Left) or X op(A) = alpha B (Right), in place.
 * Replaces src/blas/ulevel3/ATL_trsm.c:13-228 (recursive halving with GEMM
 * updates) and the amm outer loops ATL_trsmL_LN_UT.c:11-110 that alternate
 * utrsm diagonal solves with GEMM updates.  Here the 64x64 diagonal blocks
 * are inverted in one launch, then per diagonal block one in-place GEMM
 * (X_k = op(inv A_kk) * B_k, a single 128-row tile, so it reads B_k before
 * writing it) and one GEMM update of the rest of B.  Left-side solves with
 * one column (or a few, within one diagonal block) run as one flag-chained
 * launch, and with up to ATLMX_TRSM_TAG_MAXRHS as one launch of work items
 * of two 64-row blocks by 16 columns chained the same way (matrix-core
 * products).
 * ---------------------------------------------------------------------- */
/* inv_ws / ctl_ws: NULL, or caller-provided workspaces of at least
 * nblk*64*64 elements / atlmx_k_trsv_ctl_ints(na, T is double) ints (the
 * allocation-free entry points pass slices of their static array). */
void NM(PFX, trsm_dev_ws)(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N,
                          T alpha, const T *A, long lda, T *B, long ldb,
                          const int *abort_flag, T *inv_ws, int *ctl_ws, const char *rout)
{
  const int NB = ATLMX_TRI_NB;
  const int na = (side == KX_LEFT) ? M : N;
  const int nblk = (na + NB - 1) / NB;
  const int tr = (ta != KX_NOTRANS);
  const int opA_lower = ((uplo == KX_LOWER) != tr);
  const int tinv = tr ? KX_TRANS : KX_NOTRANS;
  T *inv;
  int step, pid;
  if (M <= 0 || N <= 0) return;
  if (alpha != (T)1) {
    RT_CHECK(rout, K_SCAL(M, N, alpha, B, ldb, c->stream));
    if (alpha == (T)0) return;
  }
  inv = inv_ws ? inv_ws
               : (T *)rt_ws(c, WS_TRI_INV, (size_t)nblk * NB * NB * sizeof(T), rout);
  pid = rt_prof_begin(c, PROF_TRSM_INV, (double)na * NB * sizeof(T));
  RT_CHECK(rout, K_TRI_INV(uplo, diag, na, A, lda, inv, abort_flag, c->stream));
  rt_prof_end(c, pid);

  if (side == KX_LEFT && (N == 1 || (N <= ATLMX_TRSV_MAXRHS && nblk < 2)) && !abort_flag) {
    /* few right-hand sides: the solve is bound by reading the factor once,
     * so one flag-chained launch replaces 2*nblk small GEMMs */
    const size_t nctl = atlmx_k_trsv_ctl_ints(na, sizeof(T) == sizeof(double));
    int *ctl = ctl_ws ? ctl_ws : (int *)rt_ws(c, WS_TRSV, nctl * sizeof(int), rout);
    RT_CHECK(rout, atlmx_k_fill_i32(ctl, (int)nctl, 0, c->stream));
    pid = rt_prof_begin(c, PROF_TRSV, (double)na * na / 2 * sizeof(T));
    RT_CHECK(rout, K_TRSV(uplo, tinv, M, N, A, lda, inv, B, ldb, ctl, c->stream));
    rt_prof_end(c, pid);
  } else if (side == KX_LEFT && N <= ATLMX_TRSM_TAG_MAXRHS && nblk >= 2 && !abort_flag) {
    /* two to a few hundred right-hand sides: still one launch, 64x16 blocks
     * of the solution chained through tagged words; with a caller's workspace
     * (sized for ATLMX_TRSM_TAG_CHUNK columns) one launch per that many
     * columns, which gives the same result since columns are independent */
    const int chunk = ctl_ws ? ATLMX_TRSM_TAG_CHUNK : N;
    const size_t nctl =
        atlmx_k_trsm_tag_ctl_ints(na, chunk < N ? chunk : N, sizeof(T) == sizeof(double));
    int *ctl = ctl_ws ? ctl_ws : (int *)rt_ws(c, WS_TRSV, nctl * sizeof(int), rout);
    int q;
    for (q = 0; q < N; q += chunk) {
      const int nq = (N - q < chunk) ? N - q : chunk;
      RT_CHECK(rout, atlmx_k_fill_i32(ctl, (int)atlmx_k_trsm_tag_ctl_ints(
                                               na, nq, sizeof(T) == sizeof(double)),
                                      0, c->stream));
      pid = rt_prof_begin(c, PROF_TRSV, (double)na * na / 2 * sizeof(T));
      RT_CHECK(rout, K_TRSM_TAG(uplo, tinv, M, nq, A, lda, inv, B + (size_t)q * ldb, ldb, ctl,
                                c->stream));
      rt_prof_end(c, pid);
    }
  } else if (side == KX_LEFT) {
    const int fwd = opA_lower;
    for (step = 0; step < nblk; ++step) {
      const int k = fwd ? step : nblk - 1 - step;
      const int r0 = k * NB, rb = (M - r0 < NB) ? M - r0 : NB;
      GEMM_DEV(c, tinv, KX_NOTRANS, rb, N, rb, (T)1, inv + (size_t)k * NB * NB, NB,
               B + r0, ldb, (T)0, B + r0, ldb, 0, abort_flag, rout);
      if (fwd && r0 + rb < M) {
        const int R0 = r0 + rb;
        const T *pa = tr ? A + r0 + (size_t)R0 * lda : A + R0 + (size_t)r0 * lda;
        GEMM_DEV(c, ta, KX_NOTRANS, M - R0, N, rb, (T)-1, pa, lda, B + r0, ldb, (T)1,
                 B + R0, ldb, 0, abort_flag, rout);
      } else if (!fwd && r0 > 0) {
        const T *pa = tr ? A + r0 : A + (size_t)r0 * lda;
        GEMM_DEV(c, ta, KX_NOTRANS, r0, N, rb, (T)-1, pa, lda, B + r0, ldb, (T)1, B, ldb,
                 0, abort_flag, rout);
      }
    }
  } else {
    const int fwd = !opA_lower;     /* X*U: columns left to right */
    for (step = 0; step < nblk; ++step) {
      const int k = fwd ? step : nblk - 1 - step;
      const int c0 = k * NB, cb = (N - c0 < NB) ? N - c0 : NB;
      T *Bk = B + (size_t)c0 * ldb;
      GEMM_DEV(c, KX_NOTRANS, tinv, M, cb, cb, (T)1, Bk, ldb, inv + (size_t)k * NB * NB,
               NB, (T)0, Bk, ldb, 0, abort_flag, rout);
      if (fwd && c0 + cb < N) {
        const int C0 = c0 + cb;
        const T *pa = tr ? A + C0 + (size_t)c0 * lda : A + c0 + (size_t)C0 * lda;
        GEMM_DEV(c, KX_NOTRANS, ta, M, N - C0, cb, (T)-1, Bk, ldb, pa, lda, (T)1,
                 B + (size_t)C0 * ldb, ldb, 0, abort_flag, rout);
      } else if (!fwd && c0 > 0) {
        const T *pa = tr ? A + (size_t)c0 * lda : A + c0;
        GEMM_DEV(c, KX_NOTRANS, ta, M, c0, cb, (T)-1, Bk, ldb, pa, lda, (T)1, B, ldb, 0,
                 abort_flag, rout);
      }
    }
  }
}

void TRSM_DEV(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N, T alpha,
              const T *A, long lda, T *B, long ldb, const int *abort_flag,
              const char *rout)
{
  NM(PFX, trsm_dev_ws)(c, side, uplo, ta, diag, M, N, alpha, A, lda, B, ldb, abort_flag, NULL,
                       NULL, rout);
}

/* ------------------------------------------------------------------------
 * SYRK: C = alpha*op(A)*op(A)^T + beta*C on one triangle, as the GEMM
 * kernel with the triangle mask (tiles wholly outside the triangle exit at
 * once).  Replaces src/blas/ulevel3/ATL_syrk.c:186-289.
 * ---------------------------------------------------------------------- */
void SYRK_DEV(atlmx_ctx *c, int uplo, int trans, int N, int K, T alpha, const T *A,
              long lda, T beta, T *C, long ldc, const int *abort_flag, const char *rout)
{
  if (trans == KX_NOTRANS)
    GEMM_DEV(c, KX_NOTRANS, KX_TRANS, N, N, K, alpha, A, lda, A, lda, beta, C, ldc, uplo,
             abort_flag, rout);
  else
    GEMM_DEV(c, KX_TRANS, KX_NOTRANS, N, N, K, alpha, A, lda, A, lda, beta, C, ldc, uplo,
             abort_flag, rout);
}

/* ------------------------------------------------------------------------
 * TRMM: B = alpha*op(A)*B (Left) or alpha*B*op(A) (Right), A triangular, in
 * place.  Replaces src/blas/ulevel3/ATL_trmm.c:149 (recursive halving on
 * GEMM).  Here A's triangle is expanded into a dense workspace (zeros
 * outside, ones on a unit diagonal) and B copied aside, so the product is
 * one launch of the MFMA GEMM over the whole of B whose tiles read only the
 * K slices that meet the triangle's nonzeros (the K-band instance).
 * ---------------------------------------------------------------------- */
void NM(PFX, trmm_dev)(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N,
                       T alpha, const T *A, long lda, T *B, long ldb, const char *rout)
{
  const int na = (side == KX_LEFT) ? M : N;
  int lower_op;
  T *Tm, *W;
  if (M <= 0 || N <= 0) return;
  if (alpha == (T)0) {                  /* ATL_dreftrmm.c:182: B = 0 */
    RT_CHECK(rout, K_SCAL(M, N, (T)0, B, ldb, c->stream));
    return;
  }
  Tm = (T *)rt_ws(c, WS_L3A, (size_t)na * na * sizeof(T), rout);
  W = (T *)rt_ws(c, WS_L3B, (size_t)M * N * sizeof(T), rout);
  RT_CHECK(rout, K_TRICOPY(ATLMX_TC_EXPAND_TRI, uplo, diag, na, na, (T *)A, lda, Tm, na,
                           c->stream));
  RT_CHECK(rout, K_TRICOPY(ATLMX_TC_COPY, 0, 0, M, N, W, M, B, ldb, c->stream));
  /* op(Tm) is lower when exactly one of "lower" and "transposed" holds;
   * the GEMM's tiles skip the K slices that meet only its zeros */
  lower_op = ((uplo == KX_LOWER) != (ta != KX_NOTRANS));
  if (side == KX_LEFT)
    GEMM_DEV(c, ta, KX_NOTRANS, M, N, M, alpha, Tm, na, W, M, (T)0, B, ldb,
             lower_op ? KX_KBAND_AL : KX_KBAND_AU, NULL, rout);
  else
    GEMM_DEV(c, KX_NOTRANS, ta, M, N, N, alpha, W, M, Tm, na, (T)0, B, ldb,
             lower_op ? KX_KBAND_RBL : KX_KBAND_RBU, NULL, rout);
}

/* SYMM: C = alpha*A*B + beta*C (Left) or alpha*B*A + beta*C (Right), A
 * symmetric in one triangle.  Replaces src/blas/ulevel3/ATL_symmL.c /
 * ATL_symmR.c: the symmetric matrix is expanded into a dense workspace and
 * the product is one GEMM launch. */
void NM(PFX, symm_dev)(atlmx_ctx *c, int side, int uplo, int M, int N, T alpha, const T *A,
                       long lda, const T *B, long ldb, T beta, T *C, long ldc,
                       const char *rout)
{
  const int na = (side == KX_LEFT) ? M : N;
  T *S;
  if (M <= 0 || N <= 0 || (alpha == (T)0 && beta == (T)1)) return;
  if (alpha == (T)0) {
    GEMM_DEV(c, KX_NOTRANS, KX_NOTRANS, M, N, 0, (T)0, B, ldb, B, ldb, beta, C, ldc, 0, NULL,
             rout);
    return;
  }
  S = (T *)rt_ws(c, WS_L3A, (size_t)na * na * sizeof(T), rout);
  RT_CHECK(rout, K_TRICOPY(ATLMX_TC_EXPAND_SYM, uplo, KX_NONUNIT, na, na, (T *)A, lda, S, na,
                           c->stream));
  if (side == KX_LEFT)
    GEMM_DEV(c, KX_NOTRANS, KX_NOTRANS, M, N, M, alpha, S, na, B, ldb, beta, C, ldc, 0, NULL,
             rout);
  else
    GEMM_DEV(c, KX_NOTRANS, KX_NOTRANS, M, N, N, alpha, B, ldb, S, na, beta, C, ldc, 0, NULL,
             rout);
}

/* SYR2K: C = alpha*(op(A) op(B)^T + op(B) op(A)^T) + beta*C on one
 * triangle.  Replaces src/blas/ulevel3/ATL_syr2k.c:255: two launches of the
 * triangle-masked GEMM (tiles outside the triangle exit at once), the
 * second accumulating onto the first. */
void NM(PFX, syr2k_dev)(atlmx_ctx *c, int uplo, int trans, int N, int K, T alpha,
                        const T *A, long lda, const T *B, long ldb, T beta, T *C, long ldc,
                        const char *rout)
{
  const int t1 = (trans == KX_NOTRANS) ? KX_NOTRANS : KX_TRANS;
  const int t2 = (trans == KX_NOTRANS) ? KX_TRANS : KX_NOTRANS;
  if (N <= 0) return;
  GEMM_DEV(c, t1, t2, N, N, K, alpha, A, lda, B, ldb, beta, C, ldc, uplo, NULL, rout);
  GEMM_DEV(c, t1, t2, N, N, K, alpha, B, ldb, A, lda, (T)1, C, ldc, uplo, NULL, rout);
}

/* ====================================================================== */
/* public device-pointer API (stream-ordered, column-major)                 */
/* ====================================================================== */
void NM3(atlmx_, PFX, gemm)(int ta, int tb, int M, int N, int K, T alpha, const T *A,
                            int lda, const T *B, int ldb, T beta, T *C, int ldc)
{
  GEMM_DEV(rt_ctx(), ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 0, NULL,
           "atlmx_gemm");
}

void NM3(atlmx_, PFX, trsm)(int side, int uplo, int ta, int diag, int M, int N, T alpha,
                            const T *A, int lda, T *B, int ldb)
{
  TRSM_DEV(rt_ctx(), side, uplo, ta, diag, M, N, alpha, A, lda, B, ldb, NULL, "atlmx_trsm");
}

void NM3(atlmx_, PFX, syrk)(int uplo, int trans, int N, int K, T alpha, const T *A,
                            int lda, T beta, T *C, int ldc)
{
  SYRK_DEV(rt_ctx(), uplo, trans, N, K, alpha, A, lda, beta, C, ldc, NULL, "atlmx_syrk");
}

/* ====================================================================== */
/* ATLAS C ABI                                                              */
/* ====================================================================== */

/* interfaces/blas/C/src/cblas_dgemm.c:16-134 */
void NM3(cblas_, PFX, gemm)(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TA,
                            const enum CBLAS_TRANSPOSE TB, const int M, const int N,
                            const int K, const T alpha, const T *A, const int lda,
                            const T *B, const int ldb, const T beta, T *C, const int ldc)
{
  static const char *rout = "cblas_" STR(PFX) "gemm";
  const int info = atlmx_gemm_check(Order, TA, TB, M, N, K, lda, ldb, ldc);
  atlmx_ctx *c;
  int ta, tb, m, n, ar, ac, br, bc;
  const T *pa, *pb;
  long la, lb, lc, dla, dlb, dlc;
  T *dA, *dB, *dC;

  if (info != 2000) { cblas_xerbla(info, rout, ""); return; }

  if (M == 0 || N == 0 || ((alpha == (T)0 || K == 0) && beta == (T)1)) return;
  /* RowMajor: C^T = op(B)^T op(A)^T, i.e. swap the operands (:133) */
  if (Order == CblasColMajor) {
    ta = (TA == CblasNoTrans) ? KX_NOTRANS : KX_TRANS;
    tb = (TB == CblasNoTrans) ? KX_NOTRANS : KX_TRANS;
    m = M; n = N; pa = A; pb = B; la = lda; lb = ldb;
  } else {
    ta = (TB == CblasNoTrans) ? KX_NOTRANS : KX_TRANS;
    tb = (TA == CblasNoTrans) ? KX_NOTRANS : KX_TRANS;
    m = N; n = M; pa = B; pb = A; la = ldb; lb = lda;
  }
  lc = ldc;
  ar = (ta == KX_NOTRANS) ? m : K; ac = (ta == KX_NOTRANS) ? K : m;
  br = (tb == KX_NOTRANS) ? K : n; bc = (tb == KX_NOTRANS) ? n : K;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, pa, ar, ac, la, sizeof(T), 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, pb, br, bc, lb, sizeof(T), 1, &dlb, rout);
  dC = (T *)rt_stage_in(c, WS_STAGE_C, C, m, n, lc, sizeof(T), beta != (T)0, &dlc, rout);
  GEMM_DEV(c, ta, tb, m, n, K, alpha, dA, dla, dB, dlb, beta, dC, dlc, 0, NULL, rout);
  rt_stage_out(c, C, dC, m, n, lc, dlc, sizeof(T), rout);
  rt_sync(c, rout);
}

/* interfaces/blas/C/src/cblas_dtrsm.c:13-106 */
void NM3(cblas_, PFX, trsm)(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                            const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TA,
                            const enum CBLAS_DIAG Diag, const int M, const int N,
                            const T alpha, const T *A, const int lda, T *B, const int ldb)
{
  static const char *rout = "cblas_" STR(PFX) "trsm";
  const int info = atlmx_trxm_check(Order, Side, Uplo, TA, Diag, M, N, lda, ldb);
  int side, uplo, m, n, na;
  atlmx_ctx *c;
  long dla, dlb;
  T *dA, *dB;

  if (info != 2000) { cblas_xerbla(info, rout, ""); return; }
  if (M == 0 || N == 0) return;
  /* RowMajor: flip Side and Uplo, swap M and N (:100-105) */
  if (Order == CblasColMajor) {
    side = (Side == CblasLeft) ? KX_LEFT : KX_RIGHT;
    uplo = (Uplo == CblasUpper) ? KX_UPPER : KX_LOWER;
    m = M; n = N;
  } else {
    side = (Side == CblasLeft) ? KX_RIGHT : KX_LEFT;
    uplo = (Uplo == CblasUpper) ? KX_LOWER : KX_UPPER;
    m = N; n = M;
  }
  na = (side == KX_LEFT) ? m : n;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, na, na, lda, sizeof(T), 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, m, n, ldb, sizeof(T), alpha != (T)0, &dlb, rout);
  TRSM_DEV(c, side, uplo, TA == CblasNoTrans ? KX_NOTRANS : KX_TRANS,
           Diag == CblasUnit ? KX_UNIT : KX_NONUNIT, m, n, alpha, dA, dla, dB, dlb, NULL,
           rout);
  rt_stage_out(c, B, dB, m, n, ldb, dlb, sizeof(T), rout);
  rt_sync(c, rout);
}

/* interfaces/blas/C/src/cblas_dsyrk.c */
void NM3(cblas_, PFX, syrk)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                            const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                            const T alpha, const T *A, const int lda, const T beta, T *C,
                            const int ldc)
{
  static const char *rout = "cblas_" STR(PFX) "syrk";
  int info = 2000;
  int uplo, trans, ar, ac;
  atlmx_ctx *c;
  long dla, dlc;
  T *dA, *dC;
  if (Uplo != CblasUpper && Uplo != CblasLower)
    info = cblas_errprn(2, info, "UPLO must be %d or %d, but is set to %d", CblasUpper,
                        CblasLower, Uplo);
  if (N < 0) info = cblas_errprn(4, info, "N cannot be less than zero; it is set to %d.", N);
  if (K < 0) info = cblas_errprn(5, info, "K cannot be less than zero; it is set to %d.", K);
  if (Order == CblasColMajor || Order == CblasRowMajor) {
    const int nt = (Order == CblasColMajor) == (Trans == CblasNoTrans);
    if (Trans != CblasNoTrans && Trans != CblasTrans && Trans != CblasConjTrans)
      info = cblas_errprn(3, info, "Trans must be %d, %d or %d, but is set to %d",
                          CblasNoTrans, CblasTrans, CblasConjTrans, Trans);
    if (nt) {
      if (lda < N || lda < 1)
        info = cblas_errprn(8, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
    } else {
      if (lda < K || lda < 1)
        info = cblas_errprn(8, info, "lda must be >= MAX(K,1): lda=%d K=%d", lda, K);
    }
  } else
    info = cblas_errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                        CblasColMajor, Order);
  if (ldc < N || ldc < 1)
    info = cblas_errprn(11, info, "ldc must be >= MAX(N,1): ldc=%d N=%d", ldc, N);
  if (info != 2000) { cblas_xerbla(info, rout, ""); return; }
  if (N == 0 || ((alpha == (T)0 || K == 0) && beta == (T)1)) return;
  /* RowMajor holds C^T and A^T: flip Uplo and Trans on the same memory */
  if (Order == CblasColMajor) {
    uplo = (Uplo == CblasUpper) ? KX_UPPER : KX_LOWER;
    trans = (Trans == CblasNoTrans) ? KX_NOTRANS : KX_TRANS;
  } else {
    uplo = (Uplo == CblasUpper) ? KX_LOWER : KX_UPPER;
    trans = (Trans == CblasNoTrans) ? KX_TRANS : KX_NOTRANS;
  }
  ar = (trans == KX_NOTRANS) ? N : K;
  ac = (trans == KX_NOTRANS) ? K : N;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, ar, ac, lda, sizeof(T), 1, &dla, rout);
  dC = (T *)rt_stage_in(c, WS_STAGE_C, C, N, N, ldc, sizeof(T), 1, &dlc, rout);
  SYRK_DEV(c, uplo, trans, N, K, alpha, dA, dla, beta, dC, dlc, NULL, rout);
  rt_stage_out(c, C, dC, N, N, ldc, dlc, sizeof(T), rout);
  rt_sync(c, rout);
}

/* interfaces/blas/C/src/cblas_dtrmm.c (RowMajor flips Side and Uplo and
 * swaps M and N, as for TRSM) */
void NM3(cblas_, PFX, trmm)(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                            const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TA,
                            const enum CBLAS_DIAG Diag, const int M, const int N,
                            const T alpha, const T *A, const int lda, T *B, const int ldb)
{
  static const char *rout = "cblas_" STR(PFX) "trmm";
  const int info = atlmx_trxm_check(Order, Side, Uplo, TA, Diag, M, N, lda, ldb);
  int side, uplo, m, n, na;
  atlmx_ctx *c;
  long dla, dlb;
  T *dA, *dB;
  if (info != 2000) { cblas_xerbla(info, rout, ""); return; }
  if (M == 0 || N == 0) return;
  if (Order == CblasColMajor) {
    side = (Side == CblasLeft) ? KX_LEFT : KX_RIGHT;
    uplo = (Uplo == CblasUpper) ? KX_UPPER : KX_LOWER;
    m = M; n = N;
  } else {
    side = (Side == CblasLeft) ? KX_RIGHT : KX_LEFT;
    uplo = (Uplo == CblasUpper) ? KX_LOWER : KX_UPPER;
    m = N; n = M;
  }
  na = (side == KX_LEFT) ? m : n;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, na, na, lda, sizeof(T), 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, m, n, ldb, sizeof(T), alpha != (T)0, &dlb, rout);
  NM(PFX, trmm_dev)(c, side, uplo, TA == CblasNoTrans ? KX_NOTRANS : KX_TRANS,
                    Diag == CblasUnit ? KX_UNIT : KX_NONUNIT, m, n, alpha, dA, dla, dB, dlb,
                    rout);
  rt_stage_out(c, B, dB, m, n, ldb, dlb, sizeof(T), rout);
  rt_sync(c, rout);
}

/* interfaces/blas/C/src/cblas_dsymm.c (RowMajor flips Side and Uplo and
 * swaps M and N) */
void NM3(cblas_, PFX, symm)(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                            const enum CBLAS_UPLO Uplo, const int M, const int N,
                            const T alpha, const T *A, const int lda, const T *B,
                            const int ldb, const T beta, T *C, const int ldc)
{
  static const char *rout = "cblas_" STR(PFX) "symm";
  const int info = atlmx_symm_check(Order, Side, Uplo, M, N, lda, ldb, ldc);
  int side, uplo, m, n, na;
  atlmx_ctx *c;
  long dla, dlb, dlc;
  T *dA, *dB, *dC;
  if (info != 2000) { cblas_xerbla(info, rout, ""); return; }
  if (M == 0 || N == 0 || (alpha == (T)0 && beta == (T)1)) return;
  if (Order == CblasColMajor) {
    side = (Side == CblasLeft) ? KX_LEFT : KX_RIGHT;
    uplo = (Uplo == CblasUpper) ? KX_UPPER : KX_LOWER;
    m = M; n = N;
  } else {
    side = (Side == CblasLeft) ? KX_RIGHT : KX_LEFT;
    uplo = (Uplo == CblasUpper) ? KX_LOWER : KX_UPPER;
    m = N; n = M;
  }
  na = (side == KX_LEFT) ? m : n;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, na, na, lda, sizeof(T), 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, m, n, ldb, sizeof(T), 1, &dlb, rout);
  dC = (T *)rt_stage_in(c, WS_STAGE_C, C, m, n, ldc, sizeof(T), beta != (T)0, &dlc, rout);
  NM(PFX, symm_dev)(c, side, uplo, m, n, alpha, dA, dla, dB, dlb, beta, dC, dlc, rout);
  rt_stage_out(c, C, dC, m, n, ldc, dlc, sizeof(T), rout);
  rt_sync(c, rout);
}

/* interfaces/blas/C/src/cblas_dsyr2k.c (RowMajor flips Uplo and Trans on
 * the same memory, as for SYRK) */
void NM3(cblas_, PFX, syr2k)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                             const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                             const T alpha, const T *A, const int lda, const T *B,
                             const int ldb, const T beta, T *C, const int ldc)
{
  static const char *rout = "cblas_" STR(PFX) "syr2k";
  int info = 2000, uplo, trans, ar, ac;
  atlmx_ctx *c;
  long dla, dlb, dlc;
  T *dA, *dB, *dC;
  if (Uplo != CblasUpper && Uplo != CblasLower)
    info = cblas_errprn(2, info, "UPLO must be %d or %d, but is set to %d", CblasUpper,
                        CblasLower, Uplo);
  if (N < 0) info = cblas_errprn(4, info, "N cannot be less than zero; it is set to %d.", N);
  if (K < 0) info = cblas_errprn(5, info, "K cannot be less than zero; it is set to %d.", K);
  if (Order == CblasColMajor || Order == CblasRowMajor) {
    const int nt = (Order == CblasColMajor) == (Trans == CblasNoTrans);
    if (Trans != CblasNoTrans && Trans != CblasTrans && Trans != CblasConjTrans)
      info = cblas_errprn(3, info, "Trans must be %d or %d, but is set to %d", CblasNoTrans,
                          CblasTrans, Trans);
    if (nt) {
      if (lda < N || lda < 1)
        info = cblas_errprn(8, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
      if (ldb < N || ldb < 1)
        info = cblas_errprn(10, info, "ldb must be >= MAX(N,1): ldb=%d N=%d", ldb, N);
    } else {
      if (lda < K || lda < 1)
        info = cblas_errprn(8, info, "lda must be >= MAX(K,1): lda=%d K=%d", lda, K);
      if (ldb < K || ldb < 1)
        info = cblas_errprn(10, info, "ldb must be >= MAX(K,1): ldb=%d K=%d", ldb, K);
    }
  } else
    info = cblas_errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                        CblasColMajor, Order);
  if (ldc < N || ldc < 1)
    info = cblas_errprn(13, info, "ldc must be >= MAX(N,1): ldc=%d N=%d", ldc, N);
  if (info != 2000) { cblas_xerbla(info, rout, ""); return; }
  if (N == 0 || ((alpha == (T)0 || K == 0) && beta == (T)1)) return;
  if (Order == CblasColMajor) {
    uplo = (Uplo == CblasUpper) ? KX_UPPER : KX_LOWER;
    trans = (Trans == CblasNoTrans) ? KX_NOTRANS : KX_TRANS;
  } else {
    uplo = (Uplo == CblasUpper) ? KX_LOWER : KX_UPPER;
    trans = (Trans == CblasNoTrans) ? KX_TRANS : KX_NOTRANS;
  }
  ar = (trans == KX_NOTRANS) ? N : K;
  ac = (trans == KX_NOTRANS) ? K : N;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, ar, ac, lda, sizeof(T), 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, ar, ac, ldb, sizeof(T), 1, &dlb, rout);
  dC = (T *)rt_stage_in(c, WS_STAGE_C, C, N, N, ldc, sizeof(T), 1, &dlc, rout);
  NM(PFX, syr2k_dev)(c, uplo, trans, N, K, alpha, dA, dla, dB, dlb, beta, dC, dlc, rout);
  rt_stage_out(c, C, dC, N, N, ldc, dlc, sizeof(T), rout);
  rt_sync(c, rout);
}

/* cblas_i?amax (ATLAS include/cblas.h:66-67; semantics of
 * src/blas/reference/level1/ATL_idrefamax.c:112) */
CBLAS_INDEX NM3(cblas_i, PFX, amax)(const int N, const T *X, const int incX)
{
  static const char *rout = "cblas_i" STR(PFX) "amax";
  atlmx_ctx *c;
  long dl;
  const T *dX;
  int *dout, out = 0;
  if (N < 1 || incX < 1) return 0;
  c = rt_ctx();
  dX = (const T *)rt_stage_in(c, WS_STAGE_A, X, 1, N, incX, sizeof(T), 1, &dl, rout);
  dout = (int *)rt_ws(c, WS_INFO, 256, rout);
  RT_CHECK(rout, K_IAMAX(N, dX, (int)dl, dout, c->stream));
  RT_CHECK(rout, hipMemcpyAsync(&out, dout, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  rt_sync(c, rout);
  return out;
}

#undef GEMM_DEV
#undef TRSM_DEV
#undef SYRK_DEV

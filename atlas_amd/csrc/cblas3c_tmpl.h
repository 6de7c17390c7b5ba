/*
 * cblas3c_tmpl.h -- precision template of the complex Level-3 BLAS (included
 * by cblas3c.c with T=double/PFX=z and T=float/PFX=c).
 *
 *   ?gemm_dev / ?trsm_dev : stream-ordered drivers on device pointers
 *   atlmx_?gemm / _?trsm  : public device-pointer API
 *   cblas_?gemm / _?trsm  : the ATLAS C ABI (alpha/beta passed by pointer)
 *
 * ATLAS runs a complex GEMM as four real access-major products on the split
 * real and imaginary parts (src/blas/ammm/ATL_iploops.c:257-347,
 * ATL_cammm_rk2.c).  Here it is one real MFMA GEMM: the interleaved complex
 * column-major B and C are already real matrices of twice the rows, and the
 * left operand alpha*op(A) is expanded into 2x2 real blocks
 * [Re -Im; Im Re] (k_cplx.hip), so C_real = Ahat * B_real (+ beta C) does
 * the 8mnk flops of the complex product in one launch of the real kernel.
 * ---------------------------------------------------------------------- */
#define NM_(a, b) a##b
#define NM(a, b) NM_(a, b)
#define NM3_(a, b, c) a##b##c
#define NM3(a, b, c) NM3_(a, b, c)

/* C = beta*C, C complex m x n on the device */
static void NM(PFX, scal_dev)(atlmx_ctx *c, int m, int n, const T *beta, T *C, long ldc,
                              const char *rout)
{
  RT_CHECK(rout, K_CSCAL(m, n, beta[0], beta[1], C, ldc, c->stream));
}

/* ------------------------------------------------------------------------
 * GEMM: C = alpha op(A) op(B) + beta C, op = N / T / C (conjugate
 * transpose), quick returns and alpha = 0 as src/blas/reference/level3/
 * ATL_zrefgemm.c.  A beta with an imaginary part scales C first (one
 * complex pass); a real beta goes to the real kernel.  op(B) other than N
 * is gathered into a K x N complex workspace.  The expanded left operand
 * (4*rows*K reals) is built and used in row chunks of at most
 * ATLMX_CPLX_WS_BYTES.  tri = KX_LOWER / KX_UPPER updates only C's lower /
 * upper triangle, diagonal included (the HERK shape; one chunk then);
 * tri = KX_KBAND_* (TRMM: op(A) or op(B) triangular with zeros stored)
 * skips the K slices that meet only the triangle's zeros (one chunk).
 * ---------------------------------------------------------------------- */
void NM(PFX, gemm_dev)(atlmx_ctx *c, int ta, int tb, int M, int N, int K, const T *alpha,
                       const T *A, long lda, const T *B, long ldb, const T *beta, T *C,
                       long ldc, int tri, const char *rout)
{
  const int a_zero = (alpha[0] == (T)0 && alpha[1] == (T)0);
  const int b_one = (beta[0] == (T)1 && beta[1] == (T)0);
  const T *Bv;
  long ldbv, ldh, mc;
  T breal, *H;
  int i0, pid;
  if (M <= 0 || N <= 0) return;
  if ((a_zero || K == 0) && b_one) return;
  if (a_zero || K == 0) { NM(PFX, scal_dev)(c, M, N, beta, C, ldc, rout); return; }
  if (beta[1] != (T)0) {
    NM(PFX, scal_dev)(c, M, N, beta, C, ldc, rout);
    breal = (T)1;
  } else
    breal = beta[0];
  if (tb == KX_NOTRANS) {
    Bv = B; ldbv = 2 * ldb;
  } else {
    const long ldk = (K + 1) & ~1L;         /* even: 16-byte aligned real columns */
    T *Bt = (T *)rt_ws(c, WS_L3B, (size_t)ldk * N * 2 * sizeof(T), rout);
    RT_CHECK(rout, K_CGATHER(K, N, B, ldb, 1, tb == KX_CONJTRANS, (T)1, (T)0, 0, Bt, ldk, 0, 0,
                             c->stream));
    Bv = Bt; ldbv = 2 * ldk;
  }
  mc = (long)(ATLMX_CPLX_WS_BYTES / (4.0 * K * sizeof(T)));
  mc = (mc / 64) * 64;
  if (mc < 64) mc = 64;
  if (mc > M || tri) mc = M;
  ldh = (2 * mc + 7) & ~7L;
  H = (T *)rt_ws(c, WS_L3A, (size_t)ldh * 2 * K * sizeof(T), rout);
  for (i0 = 0; i0 < M; i0 += (int)mc) {
    const int mi = (M - i0 < mc) ? M - i0 : (int)mc;
    const T *Ai = (ta == KX_NOTRANS) ? A + 2 * (size_t)i0 : A + 2 * (size_t)i0 * lda;
    RT_CHECK(rout, K_CGATHER(mi, K, Ai, lda, ta != KX_NOTRANS, ta == KX_CONJTRANS, alpha[0],
                             alpha[1], 1, H, ldh, 0, 0, c->stream));
    pid = rt_prof_begin2(c, PROF_GEMM, 8.0 * (double)mi * N * K,
                         2.0 * ((double)mi * K + (double)K * N +
                                (double)mi * N * (breal != (T)0 ? 2 : 1)) * sizeof(T));
    RT_CHECK(rout, K_GEMM(KX_NOTRANS, KX_NOTRANS, 2 * mi, N, 2 * K, (T)1, H, ldh, Bv, ldbv,
                          breal, C + 2 * (size_t)i0, 2 * ldc,
                          tri == KX_LOWER ? KX_CLOWER : (tri == KX_UPPER ? KX_CUPPER : tri),
                          NULL, c->stream));
    rt_prof_end(c, pid);
  }
}

/* ------------------------------------------------------------------------
 * TRSM: op(A) X = alpha B (Left) or X op(A) = alpha B (Right), in place.
 * Replaces src/blas/ulevel3/ATL_trsm.c's complex instantiation.  The
 * effective left operand E (op(A) for Left; op(A)^T for Right, whose
 * right-hand sides are transposed into a workspace and back) is expanded
 * with its triangle mask and unit diagonal, the inverses of its 32x32
 * complex diagonal blocks are formed expanded, and the solve is the real
 * blocked substitution on the expanded system: per block one in-place real
 * GEMM X_k = inv(E_kk) X_k (at most 64 real rows, one tile, so it reads X_k
 * before writing it) and one real GEMM update of the remaining rows.
 * ---------------------------------------------------------------------- */
void NM(PFX, trsm_dev)(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N,
                       const T *alpha, const T *A, long lda, T *B, long ldb, const char *rout)
{
  const int NB = ATLMX_CTRI_NB;
  const int na = (side == KX_LEFT) ? M : N, nrhs = (side == KX_LEFT) ? N : M;
  const int trans_e = (side == KX_LEFT) ? (ta != KX_NOTRANS) : (ta == KX_NOTRANS);
  const int conj_e = (ta == KX_CONJTRANS);
  const int lower_e = ((uplo == KX_LOWER) != trans_e);
  const int nblk = (na + NB - 1) / NB;
  long ldh, ldx;
  T *H, *inv, *X;
  int step, pid;
  if (M <= 0 || N <= 0) return;
  if (!(alpha[0] == (T)1 && alpha[1] == (T)0)) {
    NM(PFX, scal_dev)(c, M, N, alpha, B, ldb, rout);
    if (alpha[0] == (T)0 && alpha[1] == (T)0) return;
  }
  if (side == KX_LEFT) {
    X = B; ldx = ldb;
  } else {
    ldx = (N + 1) & ~1L;
    X = (T *)rt_ws(c, WS_L3B, (size_t)ldx * M * 2 * sizeof(T), rout);
    RT_CHECK(rout, K_CGATHER(N, M, B, ldb, 1, 0, (T)1, (T)0, 0, X, ldx, 0, 0, c->stream));
  }
  ldh = (2L * na + 7) & ~7L;
  H = (T *)rt_ws(c, WS_L3A, (size_t)ldh * 2 * na * sizeof(T), rout);
  RT_CHECK(rout, K_CGATHER(na, na, A, lda, trans_e, conj_e, (T)1, (T)0, 1, H, ldh,
                           lower_e ? KX_LOWER : KX_UPPER, diag == KX_UNIT, c->stream));
  inv = (T *)rt_ws(c, WS_TRI_INV, (size_t)nblk * 64 * 64 * sizeof(T), rout);
  pid = rt_prof_begin(c, PROF_TRSM_INV, 2.0 * na * NB * sizeof(T));
  RT_CHECK(rout, K_CTRI_INV(lower_e, na, H, ldh, inv, c->stream));
  rt_prof_end(c, pid);
  /* Up to a few hundred right-hand sides: the real one-launch chained
   * solves on the expanded system.  They read the matrix only off its 64x64
   * diagonal blocks and take those blocks' inverses from inv, so the 2x2
   * blocks on the expanded diagonal (not triangular in the real sense) are
   * no obstacle. */
  if (nrhs == 1 || (nrhs <= ATLMX_TRSV_MAXRHS && nblk < 2)) {
    const size_t nctl = atlmx_k_trsv_ctl_ints(2 * na, sizeof(T) == sizeof(double));
    int *ctl = (int *)rt_ws(c, WS_TRSV, nctl * sizeof(int), rout);
    RT_CHECK(rout, atlmx_k_fill_i32(ctl, (int)nctl, 0, c->stream));
    pid = rt_prof_begin(c, PROF_TRSV, 2.0 * na * na * sizeof(T));
    RT_CHECK(rout, K_RTRSV(lower_e ? KX_LOWER : KX_UPPER, KX_NOTRANS, 2 * na, nrhs, H, ldh, inv,
                           X, 2 * ldx, ctl, c->stream));
    rt_prof_end(c, pid);
  } else if (nrhs <= ATLMX_TRSM_TAG_MAXRHS && nblk >= 2) {
    const size_t nctl = atlmx_k_trsm_tag_ctl_ints(2 * na, nrhs, sizeof(T) == sizeof(double));
    int *ctl = (int *)rt_ws(c, WS_TRSV, nctl * sizeof(int), rout);
    RT_CHECK(rout, atlmx_k_fill_i32(ctl, (int)nctl, 0, c->stream));
    pid = rt_prof_begin(c, PROF_TRSV, 2.0 * na * na * sizeof(T));
    RT_CHECK(rout, K_RTRSM_TAG(lower_e ? KX_LOWER : KX_UPPER, KX_NOTRANS, 2 * na, nrhs, H, ldh,
                               inv, X, 2 * ldx, ctl, c->stream));
    rt_prof_end(c, pid);
  } else {
    for (step = 0; step < nblk; ++step) {
      const int k = lower_e ? step : nblk - 1 - step;
      const int r0 = k * NB, rb = (na - r0 < NB) ? na - r0 : NB;
      T *Xk = X + 2 * (size_t)r0;
      pid = rt_prof_begin2(c, PROF_GEMM, 8.0 * rb * rb * nrhs,
                           2.0 * ((double)rb * rb + 2.0 * rb * nrhs) * sizeof(T));
      RT_CHECK(rout, K_GEMM(KX_NOTRANS, KX_NOTRANS, 2 * rb, nrhs, 2 * rb, (T)1,
                            inv + (size_t)k * 64 * 64, 64, Xk, 2 * ldx, (T)0, Xk, 2 * ldx, 0, NULL,
                            c->stream));
      rt_prof_end(c, pid);
      if (lower_e && r0 + rb < na) {
        const int R0 = r0 + rb;
        pid = rt_prof_begin2(c, PROF_GEMM, 8.0 * (na - R0) * rb * nrhs,
                             2.0 * ((double)(na - R0) * rb + rb * nrhs + 2.0 * (na - R0) * nrhs) *
                                 sizeof(T));
        RT_CHECK(rout, K_GEMM(KX_NOTRANS, KX_NOTRANS, 2 * (na - R0), nrhs, 2 * rb, (T)-1,
                              H + 2 * (size_t)R0 + (size_t)(2 * r0) * ldh, ldh, Xk, 2 * ldx, (T)1,
                              X + 2 * (size_t)R0, 2 * ldx, 0, NULL, c->stream));
        rt_prof_end(c, pid);
      } else if (!lower_e && r0 > 0) {
        pid = rt_prof_begin2(c, PROF_GEMM, 8.0 * r0 * rb * nrhs,
                             2.0 * ((double)r0 * rb + rb * nrhs + 2.0 * r0 * nrhs) * sizeof(T));
        RT_CHECK(rout, K_GEMM(KX_NOTRANS, KX_NOTRANS, 2 * r0, nrhs, 2 * rb, (T)-1,
                              H + (size_t)(2 * r0) * ldh, ldh, Xk, 2 * ldx, (T)1, X, 2 * ldx, 0,
                              NULL, c->stream));
        rt_prof_end(c, pid);
      }
    }
  }
  if (side == KX_RIGHT)
    RT_CHECK(rout, K_CGATHER(M, N, X, ldx, 1, 0, (T)1, (T)0, 0, B, ldb, 0, 0, c->stream));
}

/* ====================================================================== */
/* public device-pointer API (stream-ordered, column-major)                 */
/* ====================================================================== */
void NM3(atlmx_, PFX, gemm)(int ta, int tb, int M, int N, int K, const T *alpha, const T *A,
                            int lda, const T *B, int ldb, const T *beta, T *C, int ldc)
{
  NM(PFX, gemm_dev)(rt_ctx(), ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, 0,
                    "atlmx_gemm");
}

void NM3(atlmx_, PFX, trsm)(int side, int uplo, int ta, int diag, int M, int N,
                            const T *alpha, const T *A, int lda, T *B, int ldb)
{
  NM(PFX, trsm_dev)(rt_ctx(), side, uplo, ta, diag, M, N, alpha, A, lda, B, ldb,
                    "atlmx_trsm");
}

/* ====================================================================== */
/* ATLAS C ABI                                                              */
/* ====================================================================== */
static int NM(PFX, op)(int t)
{
  return t == CblasNoTrans ? KX_NOTRANS : (t == CblasTrans ? KX_TRANS : KX_CONJTRANS);
}

/* interfaces/blas/C/src/cblas_zgemm.c:16-136 */
void NM3(cblas_, PFX, gemm)(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TA,
                            const enum CBLAS_TRANSPOSE TB, const int M, const int N,
                            const int K, const void *alpha, const void *A, const int lda,
                            const void *B, const int ldb, const void *beta, void *C,
                            const int ldc)
{
  static const char *rout = "cblas_" STR(PFX) "gemm";
  const int info = atlmx_gemm_check(Order, TA, TB, M, N, K, lda, ldb, ldc);
  const T *al = (const T *)alpha, *be = (const T *)beta;
  const size_t esz = 2 * sizeof(T);
  atlmx_ctx *c;
  int ta, tb, m, n, ar, ac, br, bc, b_zero;
  const void *pa, *pb;
  long la, lb, dla, dlb, dlc;
  T *dA, *dB, *dC;

  if (info != 2000) { cblas_xerbla(info, rout, ""); return; }
  b_zero = (be[0] == (T)0 && be[1] == (T)0);
  if (M == 0 || N == 0 ||
      (((al[0] == (T)0 && al[1] == (T)0) || K == 0) && be[0] == (T)1 && be[1] == (T)0))
    return;
  /* RowMajor: C^T = op(B)^T op(A)^T with the same op kinds (:133-135) */
  if (Order == CblasColMajor) {
    ta = NM(PFX, op)(TA); tb = NM(PFX, op)(TB);
    m = M; n = N; pa = A; pb = B; la = lda; lb = ldb;
  } else {
    ta = NM(PFX, op)(TB); tb = NM(PFX, op)(TA);
    m = N; n = M; pa = B; pb = A; la = ldb; lb = lda;
  }
  ar = (ta == KX_NOTRANS) ? m : K; ac = (ta == KX_NOTRANS) ? K : m;
  br = (tb == KX_NOTRANS) ? K : n; bc = (tb == KX_NOTRANS) ? n : K;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, pa, ar, ac, la, esz, 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, pb, br, bc, lb, esz, 1, &dlb, rout);
  dC = (T *)rt_stage_in(c, WS_STAGE_C, C, m, n, ldc, esz, !b_zero, &dlc, rout);
  NM(PFX, gemm_dev)(c, ta, tb, m, n, K, al, dA, dla, dB, dlb, be, dC, dlc, 0, rout);
  rt_stage_out(c, C, dC, m, n, ldc, dlc, esz, rout);
  rt_sync(c, rout);
}

/* interfaces/blas/C/src/cblas_ztrsm.c:13-106 */
void NM3(cblas_, PFX, trsm)(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                            const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TA,
                            const enum CBLAS_DIAG Diag, const int M, const int N,
                            const void *alpha, const void *A, const int lda, void *B,
                            const int ldb)
{
  static const char *rout = "cblas_" STR(PFX) "trsm";
  const int info = atlmx_trxm_check(Order, Side, Uplo, TA, Diag, M, N, lda, ldb);
  const T *al = (const T *)alpha;
  const size_t esz = 2 * sizeof(T);
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
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, na, na, lda, esz, 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, m, n, ldb, esz,
                        !(al[0] == (T)0 && al[1] == (T)0), &dlb, rout);
  NM(PFX, trsm_dev)(c, side, uplo, NM(PFX, op)(TA), Diag == CblasUnit ? KX_UNIT : KX_NONUNIT,
                    m, n, al, dA, dla, dB, dlb, rout);
  rt_stage_out(c, B, dB, m, n, ldb, dlb, esz, rout);
  rt_sync(c, rout);
}

/* ------------------------------------------------------------------------
 * The rest of the complex Level-3 BLAS on the complex GEMM:
 *   rank-k / rank-2k (HERK, SYRK, HER2K, SYR2K): one or two GEMMs masked to
 *     the referenced triangle (the complex-view mask of the real kernel)
 *     after beta has scaled that triangle, as src/blas/reference/level3/
 *     ATL_zrefherk.c / _zrefsyrk.c / _zrefher2k.c / _zrefsyr2k.c define
 *     them; the Hermitian forms keep the diagonal real;
 *   HEMM / SYMM: A's triangle expanded to the full matrix, then one GEMM;
 *   TRMM: op(A)'s triangle gathered (zeros outside, ones on a unit
 *     diagonal) and B copied aside, then one GEMM into B whose tiles skip
 *     the K slices that meet only the triangle's zeros (about half).
 * ---------------------------------------------------------------------- */
/* herm: Hermitian forms (beta's imaginary part is ignored, conj(alpha) on
 * the second product of the 2k form); B NULL: rank-k.  trans: KX_NOTRANS,
 * or KX_CONJTRANS (herm) / KX_TRANS (symmetric). */
static void NM(PFX, rk_dev)(atlmx_ctx *c, int herm, int uplo, int trans, int N, int K,
                            const T *alpha, const T *A, long lda, const T *B, long ldb,
                            const T *beta, T *C, long ldc, const char *rout)
{
  static const T one[2] = {(T)1, (T)0};
  const int lower = (uplo == KX_LOWER);
  const int op = herm ? KX_CONJTRANS : KX_TRANS;
  const T br = beta[0], bi = herm ? (T)0 : beta[1];
  const int a_zero = (alpha[0] == (T)0 && alpha[1] == (T)0);
  if (N <= 0) return;
  if (!(br == (T)1 && bi == (T)0) || herm)
    RT_CHECK(rout, K_CTRI_SCALE(N, lower, br, bi, herm, C, ldc, c->stream));
  if (a_zero || K == 0) return;
  {
    const T a2[2] = {alpha[0], herm ? -alpha[1] : alpha[1]};
    const T *X = B ? B : A;
    const long ldx = B ? ldb : lda;
    if (trans == KX_NOTRANS) {
      NM(PFX, gemm_dev)(c, KX_NOTRANS, op, N, N, K, alpha, A, lda, X, ldx, one, C, ldc, uplo,
                        rout);
      if (B)
        NM(PFX, gemm_dev)(c, KX_NOTRANS, op, N, N, K, a2, B, ldb, A, lda, one, C, ldc, uplo,
                          rout);
    } else {
      NM(PFX, gemm_dev)(c, op, KX_NOTRANS, N, N, K, alpha, A, lda, X, ldx, one, C, ldc, uplo,
                        rout);
      if (B)
        NM(PFX, gemm_dev)(c, op, KX_NOTRANS, N, N, K, a2, B, ldb, A, lda, one, C, ldc, uplo,
                          rout);
    }
  }
  if (herm) RT_CHECK(rout, K_CTRI_SCALE(N, lower, (T)1, (T)0, 1, C, ldc, c->stream));
}

/* C = alpha A B + beta C (Left) or alpha B A + beta C (Right), A Hermitian
 * (herm) or symmetric in its uplo triangle */
static void NM(PFX, hemm_dev)(atlmx_ctx *c, int herm, int side, int uplo, int M, int N,
                              const T *alpha, const T *A, long lda, const T *B, long ldb,
                              const T *beta, T *C, long ldc, const char *rout)
{
  const int na = (side == KX_LEFT) ? M : N;
  T *Ad;
  if (M <= 0 || N <= 0) return;
  Ad = (T *)rt_ws(c, WS_CPLX_X, (size_t)na * na * 2 * sizeof(T), rout);
  RT_CHECK(rout, K_CSYM_EXPAND(na, uplo == KX_LOWER, herm, A, lda, Ad, na, c->stream));
  if (side == KX_LEFT)
    NM(PFX, gemm_dev)(c, KX_NOTRANS, KX_NOTRANS, M, N, M, alpha, Ad, na, B, ldb, beta, C, ldc,
                      0, rout);
  else
    NM(PFX, gemm_dev)(c, KX_NOTRANS, KX_NOTRANS, M, N, N, alpha, B, ldb, Ad, na, beta, C, ldc,
                      0, rout);
}

/* B = alpha op(A) B (Left) or alpha B op(A) (Right), A triangular */
static void NM(PFX, trmm_dev)(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N,
                              const T *alpha, const T *A, long lda, T *B, long ldb,
                              const char *rout)
{
  static const T zero[2] = {(T)0, (T)0};
  const int na = (side == KX_LEFT) ? M : N;
  const int tr = (ta != KX_NOTRANS);
  const int lower_op = ((uplo == KX_LOWER) != tr);
  T *Td, *Bc;
  if (M <= 0 || N <= 0) return;
  Td = (T *)rt_ws(c, WS_CPLX_X, (size_t)na * na * 2 * sizeof(T), rout);
  Bc = (T *)rt_ws(c, WS_CPLX_Y, (size_t)M * N * 2 * sizeof(T), rout);
  RT_CHECK(rout, K_CGATHER(na, na, A, lda, tr, ta == KX_CONJTRANS, (T)1, (T)0, 0, Td, na,
                           lower_op ? KX_LOWER : KX_UPPER, diag == KX_UNIT, c->stream));
  RT_CHECK(rout, K_CGATHER(M, N, B, ldb, 0, 0, (T)1, (T)0, 0, Bc, M, 0, 0, c->stream));
  if (side == KX_LEFT)
    NM(PFX, gemm_dev)(c, KX_NOTRANS, KX_NOTRANS, M, N, M, alpha, Td, na, Bc, M, zero, B, ldb,
                      lower_op ? KX_KBAND_AL : KX_KBAND_AU, rout);
  else
    NM(PFX, gemm_dev)(c, KX_NOTRANS, KX_NOTRANS, M, N, N, alpha, Bc, M, Td, na, zero, B, ldb,
                      lower_op ? KX_KBAND_BL : KX_KBAND_BU, rout);
}

/* the rank-k / rank-2k C ABI: RowMajor holds C^T, so Uplo and Trans flip
 * (interfaces/blas/C/src/cblas_zherk.c, _zher2k.c with conj(alpha),
 * _zsyrk.c, _zsyr2k.c) */
static void NM(PFX, rk_abi)(const char *rout, int herm, int two, int Order, int Uplo, int Trans,
                            int N, int K, const T *alpha, const void *A, int lda, const void *B,
                            int ldb, const T *beta, void *C, int ldc)
{
  const int info = crk_check(Order, Uplo, Trans, N, K, lda, ldb, ldc, herm, two);
  const size_t esz = 2 * sizeof(T);
  const int alt = herm ? KX_CONJTRANS : KX_TRANS;
  int uplo, trans, ar, ac;
  T al[2];
  atlmx_ctx *c;
  long dla, dlb = 0, dlc;
  T *dA, *dB = NULL, *dC;
  if (info != 2000) { cblas_xerbla(info, rout, ""); return; }
  if (N == 0 || (((alpha[0] == (T)0 && alpha[1] == (T)0) || K == 0) && beta[0] == (T)1 &&
                 (herm || beta[1] == (T)0)))
    return;
  al[0] = alpha[0]; al[1] = alpha[1];
  if (Order == CblasColMajor) {
    uplo = (Uplo == CblasUpper) ? KX_UPPER : KX_LOWER;
    trans = (Trans == CblasNoTrans) ? KX_NOTRANS : alt;
  } else {
    uplo = (Uplo == CblasUpper) ? KX_LOWER : KX_UPPER;
    trans = (Trans == CblasNoTrans) ? alt : KX_NOTRANS;
    if (herm && two) al[1] = -al[1];
  }
  ar = (trans == KX_NOTRANS) ? N : K;
  ac = (trans == KX_NOTRANS) ? K : N;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, ar, ac, lda, esz, 1, &dla, rout);
  if (two) dB = (T *)rt_stage_in(c, WS_STAGE_B, B, ar, ac, ldb, esz, 1, &dlb, rout);
  dC = (T *)rt_stage_in(c, WS_STAGE_C, C, N, N, ldc, esz, 1, &dlc, rout);
  NM(PFX, rk_dev)(c, herm, uplo, trans, N, K, al, dA, dla, dB, dlb, beta, dC, dlc, rout);
  rt_stage_out(c, C, dC, N, N, ldc, dlc, esz, rout);
  rt_sync(c, rout);
}

/* interfaces/blas/C/src/cblas_zherk.c: alpha and beta real */
void NM3(cblas_, PFX, herk)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                            const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                            const T alpha, const void *A, const int lda, const T beta, void *C,
                            const int ldc)
{
  const T al[2] = {alpha, (T)0}, be[2] = {beta, (T)0};
  NM(PFX, rk_abi)("cblas_" STR(PFX) "herk", 1, 0, Order, Uplo, Trans, N, K, al, A, lda, NULL, 0,
                  be, C, ldc);
}

/* interfaces/blas/C/src/cblas_zher2k.c: beta real */
void NM3(cblas_, PFX, her2k)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                             const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                             const void *alpha, const void *A, const int lda, const void *B,
                             const int ldb, const T beta, void *C, const int ldc)
{
  const T be[2] = {beta, (T)0};
  NM(PFX, rk_abi)("cblas_" STR(PFX) "her2k", 1, 1, Order, Uplo, Trans, N, K, (const T *)alpha, A,
                  lda, B, ldb, be, C, ldc);
}

/* interfaces/blas/C/src/cblas_zsyrk.c */
void NM3(cblas_, PFX, syrk)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                            const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                            const void *alpha, const void *A, const int lda, const void *beta,
                            void *C, const int ldc)
{
  NM(PFX, rk_abi)("cblas_" STR(PFX) "syrk", 0, 0, Order, Uplo, Trans, N, K, (const T *)alpha, A,
                  lda, NULL, 0, (const T *)beta, C, ldc);
}

/* interfaces/blas/C/src/cblas_zsyr2k.c */
void NM3(cblas_, PFX, syr2k)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                             const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                             const void *alpha, const void *A, const int lda, const void *B,
                             const int ldb, const void *beta, void *C, const int ldc)
{
  NM(PFX, rk_abi)("cblas_" STR(PFX) "syr2k", 0, 1, Order, Uplo, Trans, N, K, (const T *)alpha, A,
                  lda, B, ldb, (const T *)beta, C, ldc);
}

/* HEMM / SYMM C ABI: RowMajor flips Side and Uplo and swaps M and N
 * (interfaces/blas/C/src/cblas_zhemm.c, cblas_zsymm.c) */
static void NM(PFX, hemm_abi)(const char *rout, int herm, int Order, int Side, int Uplo, int M,
                              int N, const T *alpha, const void *A, int lda, const void *B,
                              int ldb, const T *beta, void *C, int ldc)
{
  const int info = atlmx_symm_check(Order, Side, Uplo, M, N, lda, ldb, ldc);
  const size_t esz = 2 * sizeof(T);
  int side, uplo, m, n, na;
  atlmx_ctx *c;
  long dla, dlb, dlc;
  T *dA, *dB, *dC;
  if (info != 2000) { cblas_xerbla(info, rout, ""); return; }
  if (M == 0 || N == 0 ||
      (alpha[0] == (T)0 && alpha[1] == (T)0 && beta[0] == (T)1 && beta[1] == (T)0))
    return;
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
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, na, na, lda, esz, 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, m, n, ldb, esz, 1, &dlb, rout);
  dC = (T *)rt_stage_in(c, WS_STAGE_C, C, m, n, ldc, esz,
                        !(beta[0] == (T)0 && beta[1] == (T)0), &dlc, rout);
  NM(PFX, hemm_dev)(c, herm, side, uplo, m, n, alpha, dA, dla, dB, dlb, beta, dC, dlc, rout);
  rt_stage_out(c, C, dC, m, n, ldc, dlc, esz, rout);
  rt_sync(c, rout);
}

void NM3(cblas_, PFX, hemm)(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                            const enum CBLAS_UPLO Uplo, const int M, const int N,
                            const void *alpha, const void *A, const int lda, const void *B,
                            const int ldb, const void *beta, void *C, const int ldc)
{
  NM(PFX, hemm_abi)("cblas_" STR(PFX) "hemm", 1, Order, Side, Uplo, M, N, (const T *)alpha, A,
                    lda, B, ldb, (const T *)beta, C, ldc);
}

void NM3(cblas_, PFX, symm)(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                            const enum CBLAS_UPLO Uplo, const int M, const int N,
                            const void *alpha, const void *A, const int lda, const void *B,
                            const int ldb, const void *beta, void *C, const int ldc)
{
  NM(PFX, hemm_abi)("cblas_" STR(PFX) "symm", 0, Order, Side, Uplo, M, N, (const T *)alpha, A,
                    lda, B, ldb, (const T *)beta, C, ldc);
}

/* interfaces/blas/C/src/cblas_ztrmm.c (RowMajor flips Side and Uplo and
 * swaps M and N, as for TRSM) */
void NM3(cblas_, PFX, trmm)(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                            const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TA,
                            const enum CBLAS_DIAG Diag, const int M, const int N,
                            const void *alpha, const void *A, const int lda, void *B,
                            const int ldb)
{
  static const char *rout = "cblas_" STR(PFX) "trmm";
  const int info = atlmx_trxm_check(Order, Side, Uplo, TA, Diag, M, N, lda, ldb);
  const T *al = (const T *)alpha;
  const size_t esz = 2 * sizeof(T);
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
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, na, na, lda, esz, 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, m, n, ldb, esz, 1, &dlb, rout);
  NM(PFX, trmm_dev)(c, side, uplo, NM(PFX, op)(TA), Diag == CblasUnit ? KX_UNIT : KX_NONUNIT, m,
                    n, al, dA, dla, dB, dlb, rout);
  rt_stage_out(c, B, dB, m, n, ldb, dlb, esz, rout);
  rt_sync(c, rout);
}

/*
 * lapackc_tmpl.h -- precision template of the complex LU and Cholesky
 * families (included by cblas3c.c after cblas3c_tmpl.h, T=double/PFX=z and
 * T=float/PFX=c).
 *
 *   ?getrf_dev / ?getrs_dev / ?potrf_dev / ?potrs_dev : stream-ordered
 *                                       drivers on device pointers
 *   clapack_?getrf / ?getrs / ?gesv / ?potrf / ?potrs / ?posv : the ATLAS
 *                                       C ABI
 *
 * The LU is right-looking over panels of CLU_NB columns; each panel is
 * factored by the recursive halving of src/lapack/ATL_getrfC.c:242-361
 * (left half, its interchanges and TRSM/GEMM on the right half, right half,
 * its interchanges on the left half) down to leaves of at most CLU_LEAF
 * columns, which one workgroup factors unblocked with the rows in registers
 * (k_cplx.hip, the column loop of ATL_getf2.c).  The unit-lower TRSMs of
 * the recursion and the panel's U12 are one LDS kernel each; the GEMM
 * updates are the complex GEMM of cblas3c_tmpl.h (the real MFMA GEMM on an
 * expanded operand).  Row interchanges of a complex matrix are those of its
 * interleaved real view with each pivot doubled (rows 2k, 2k+1 <-> 2p,
 * 2p+1; the leaves write both forms): the real laswp kernels apply them,
 * and a panel's interchanges reach the rest of the matrix as one move list
 * (plan) applied to every other column in one launch.
 * ---------------------------------------------------------------------- */
#define CLU_NB 64
#define CLU_LEAF 8

/* the interchanges ipiv[k1..k2) (forward for inc > 0, reverse for inc < 0)
 * on ncols columns of the complex matrix A, rows below rows_end; ipiv2
 * holds the doubled pivots for [2 k1, 2 k2) already when d_ipiv is NULL,
 * else they are formed from d_ipiv first */
static void NM(PFX, laswp_dev)(atlmx_ctx *c, int ncols, T *A, long lda, int k1, int k2,
                               const int *d_ipiv, int inc, int rows_end, int *ipiv2,
                               const char *rout)
{
  if (ncols <= 0 || k2 <= k1) return;
  if (d_ipiv) RT_CHECK(rout, atlmx_k_cpiv2(k1, k2, d_ipiv, ipiv2, c->stream));
  RT_CHECK(rout, K_RLASWP(ncols, A, 2 * lda, 2 * k1, 2 * k2, ipiv2, inc, 2 * rows_end,
                          c->stream));
}

/* columns [c0, c0 + w) of the M-row matrix A, rows [c0, M); the columns
 * left of c0 in the panel are already factored and their interchanges
 * applied to these columns */
static void NM(PFX, getrf_rec)(atlmx_ctx *c, int M, int c0, int w, T *A, long lda,
                               int *d_ipiv, int *ipiv2, int *d_info, const char *rout)
{
  static const T one[2] = {(T)1, (T)0}, mone[2] = {(T)-1, (T)0};
  T *Acc = A + 2 * ((size_t)c0 * lda + c0);
  /* leaves whose rows fit the registers of one workgroup: 8 columns up to
   * 2048 rows, 4 up to 4096 (wider leaves above that run from memory) */
  const int leaf = (M - c0 > 2048 && M - c0 <= 4096) ? CLU_LEAF / 2 : CLU_LEAF;
  int w1;
  T *A12, *A22;
  if (w <= leaf) {
    RT_CHECK(rout, K_CGETF2(M - c0, w, Acc, lda, c0, d_ipiv + c0, ipiv2 + 2 * (size_t)c0,
                            d_info, c->stream));
    return;
  }
  w1 = w / 2;
  A12 = A + 2 * ((size_t)(c0 + w1) * lda + c0);
  A22 = A12 + 2 * (size_t)w1;
  NM(PFX, getrf_rec)(c, M, c0, w1, A, lda, d_ipiv, ipiv2, d_info, rout);
  NM(PFX, laswp_dev)(c, w - w1, A + 2 * (size_t)(c0 + w1) * lda, lda, c0, c0 + w1, NULL, 1,
                     M, ipiv2, rout);
  RT_CHECK(rout, K_CTRSM_LLNU(w1, w - w1, Acc, lda, A12, lda, c->stream));
  if (M - c0 - w1 > 0)
    NM(PFX, gemm_dev)(c, KX_NOTRANS, KX_NOTRANS, M - c0 - w1, w - w1, w1, mone,
                      Acc + 2 * (size_t)w1, lda, A12, lda, one, A22, lda, 0, rout);
  NM(PFX, getrf_rec)(c, M, c0 + w1, w - w1, A, lda, d_ipiv, ipiv2, d_info, rout);
  NM(PFX, laswp_dev)(c, w1, A + 2 * (size_t)c0 * lda, lda, c0 + w1, c0 + w, NULL, 1, M,
                     ipiv2, rout);
}

/* A = P L U in place (L unit lower, U upper), 0-based pivots in d_ipiv
 * (min(M,N) of them); returns ATLAS's info: the 1-based index of the first
 * exactly zero pivot, else 0. */
int NM(PFX, getrf_dev)(atlmx_ctx *c, int M, int N, T *A, long lda, int *d_ipiv,
                       const char *rout)
{
  static const T one[2] = {(T)1, (T)0}, mone[2] = {(T)-1, (T)0};
  const int mn = M < N ? M : N;
  int *d_info, *ipiv2, *plan, info = 0, k;
  if (mn <= 0) return 0;
  d_info = (int *)rt_ws(c, WS_INFO, sizeof(int), rout);
  ipiv2 = (int *)rt_ws(c, WS_CPIV, (size_t)2 * mn * sizeof(int), rout);
  plan = (int *)rt_ws(c, WS_PLAN, ATLMX_LASWP_PLAN_INTS * sizeof(int), rout);
  RT_CHECK(rout, atlmx_k_fill_i32(d_info, 1, 0, c->stream));
  for (k = 0; k < mn; k += CLU_NB) {
    const int jb = (mn - k < CLU_NB) ? mn - k : CLU_NB;
    T *Akk = A + 2 * ((size_t)k * lda + k);
    NM(PFX, getrf_rec)(c, M, k, jb, A, lda, d_ipiv, ipiv2, d_info, rout);
    /* the panel's interchanges on every other column */
    if (2 * jb <= 2 * ATLMX_PANEL_MAXW && 2L * (M - k) <= ATLMX_LASWP_PLAN_MAXROWS) {
      RT_CHECK(rout, atlmx_k_laswp_plan(2 * k, 2 * (k + jb), 2 * M, ipiv2, plan, c->stream));
      RT_CHECK(rout, K_RLASWP_APPLY(A, 2 * lda, 0, N, k, k + jb, plan, c->stream));
    } else {
      NM(PFX, laswp_dev)(c, k, A, lda, k, k + jb, NULL, 1, M, ipiv2, rout);
      if (k + jb < N)
        NM(PFX, laswp_dev)(c, N - k - jb, A + 2 * (size_t)(k + jb) * lda, lda, k, k + jb, NULL,
                           1, M, ipiv2, rout);
    }
    if (k + jb < N) {
      T *A12 = A + 2 * ((size_t)(k + jb) * lda + k);
      RT_CHECK(rout, K_CTRSM_LLNU(jb, N - k - jb, Akk, lda, A12, lda, c->stream));
      if (k + jb < M)
        NM(PFX, gemm_dev)(c, KX_NOTRANS, KX_NOTRANS, M - k - jb, N - k - jb, jb, mone,
                          Akk + 2 * (size_t)jb, lda, A12, lda, one, A12 + 2 * (size_t)jb, lda,
                          0, rout);
    }
  }
  RT_CHECK(rout, hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  rt_sync(c, rout);
  return info;
}

/* op(A) X = B with A = P L U from getrf_dev (src/lapack/ATL_getrs.c:17-45,
 * column-major): NoTrans = interchanges, L (unit), U; Trans / ConjTrans =
 * op(U), op(L), reverse interchanges. */
void NM(PFX, getrs_dev)(atlmx_ctx *c, int trans, int N, int NRHS, const T *A, long lda,
                        const int *d_ipiv, T *B, long ldb, const char *rout)
{
  static const T one[2] = {(T)1, (T)0};
  int *ipiv2;
  if (N <= 0 || NRHS <= 0) return;
  ipiv2 = (int *)rt_ws(c, WS_CPIV, (size_t)2 * N * sizeof(int), rout);
  if (trans == KX_NOTRANS) {
    NM(PFX, laswp_dev)(c, NRHS, B, ldb, 0, N, d_ipiv, 1, N, ipiv2, rout);
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_LOWER, KX_NOTRANS, KX_UNIT, N, NRHS, one, A, lda, B, ldb,
                      rout);
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_UPPER, KX_NOTRANS, KX_NONUNIT, N, NRHS, one, A, lda, B,
                      ldb, rout);
  } else {
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_UPPER, trans, KX_NONUNIT, N, NRHS, one, A, lda, B, ldb,
                      rout);
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_LOWER, trans, KX_UNIT, N, NRHS, one, A, lda, B, ldb,
                      rout);
    NM(PFX, laswp_dev)(c, NRHS, B, ldb, 0, N, d_ipiv, -1, N, ipiv2, rout);
  }
}

/* ====================================================================== */
/* ATLAS C ABI                                                              */
/* ====================================================================== */

/* interfaces/lapack/C/src/clapack_zgetrf.c; RowMajor factors the stored
 * A^T column-major (the real routines do the same) */
int NM3(clapack_, PFX, getrf)(const enum CBLAS_ORDER Order, const int M, const int N,
                              void *A, const int lda, int *ipiv)
{
  static const char *rout = "clapack_" STR(PFX) "getrf";
  const int ierr = atlmx_getrf_check(rout, Order, M, N, lda);
  const size_t esz = 2 * sizeof(T);
  int m, n, mn, info;
  atlmx_ctx *c;
  long dla;
  T *dA;
  int *dp;
  if (ierr) return ierr;
  m = (Order == CblasColMajor) ? M : N;
  n = (Order == CblasColMajor) ? N : M;
  mn = m < n ? m : n;
  if (mn == 0) return 0;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, m, n, lda, esz, 1, &dla, rout);
  dp = rt_is_device(ipiv) ? ipiv : (int *)rt_ws(c, WS_IPIV, (size_t)mn * sizeof(int), rout);
  info = NM(PFX, getrf_dev)(c, m, n, dA, dla, dp, rout);
  rt_stage_out(c, A, dA, m, n, lda, dla, esz, rout);
  if (dp != ipiv)
    RT_CHECK(rout, hipMemcpyAsync(ipiv, dp, (size_t)mn * sizeof(int), hipMemcpyDeviceToHost,
                                  c->stream));
  rt_sync(c, rout);
  return info;
}

/* interfaces/lapack/C/src/clapack_zgetrs.c.  RowMajor holds the factor of
 * A^T, so NoTrans is the column-major Trans solve and Trans the NoTrans
 * one; ConjTrans goes the way of Trans, as src/lapack/ATL_getrs.c:46-69
 * does (its RowMajor branches test only Trans == CblasNoTrans). */
int NM3(clapack_, PFX, getrs)(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE Trans,
                              const int N, const int NRHS, const void *A, const int lda,
                              const int *ipiv, void *B, const int ldb)
{
  static const char *rout = "clapack_" STR(PFX) "getrs";
  const int ierr = atlmx_getrs_check(rout, Order, Trans, N, NRHS, lda, ldb);
  const size_t esz = 2 * sizeof(T);
  int t;
  atlmx_ctx *c;
  long dla, dlb;
  T *dA, *dB;
  const int *dp;
  if (ierr) return ierr;
  if (N == 0 || NRHS == 0) return 0;
  if (Order == CblasColMajor)
    t = (Trans == CblasNoTrans) ? KX_NOTRANS : (Trans == CblasTrans ? KX_TRANS : KX_CONJTRANS);
  else
    t = (Trans == CblasNoTrans) ? KX_TRANS : KX_NOTRANS;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, esz, 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, N, NRHS, ldb, esz, 1, &dlb, rout);
  if (rt_is_device(ipiv)) {
    dp = ipiv;
  } else {
    int *d = (int *)rt_ws(c, WS_IPIV, (size_t)N * sizeof(int), rout);
    RT_CHECK(rout, hipMemcpyAsync(d, ipiv, (size_t)N * sizeof(int), hipMemcpyHostToDevice,
                                  c->stream));
    dp = d;
  }
  NM(PFX, getrs_dev)(c, t, N, NRHS, dA, dla, dp, dB, dlb, rout);
  rt_stage_out(c, B, dB, N, NRHS, ldb, dlb, esz, rout);
  rt_sync(c, rout);
  return 0;
}

/* interfaces/lapack/C/src/clapack_zgesv.c: getrf, then (info = 0) getrs */
int NM3(clapack_, PFX, gesv)(const enum CBLAS_ORDER Order, const int N, const int NRHS,
                             void *A, const int lda, int *ipiv, void *B, const int ldb)
{
  static const char *rout = "clapack_" STR(PFX) "gesv";
  const int ierr = atlmx_gesv_check(rout, Order, N, NRHS, lda, ldb);
  const size_t esz = 2 * sizeof(T);
  int info;
  atlmx_ctx *c;
  long dla, dlb;
  T *dA, *dB;
  int *dp;
  if (ierr) return ierr;
  if (N == 0) return 0;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, esz, 1, &dla, rout);
  dp = rt_is_device(ipiv) ? ipiv : (int *)rt_ws(c, WS_IPIV, (size_t)N * sizeof(int), rout);
  info = NM(PFX, getrf_dev)(c, N, N, dA, dla, dp, rout);
  if (!info && NRHS > 0) {
    dB = (T *)rt_stage_in(c, WS_STAGE_B, B, N, NRHS, ldb, esz, 1, &dlb, rout);
    NM(PFX, getrs_dev)(c, (Order == CblasRowMajor) ? KX_TRANS : KX_NOTRANS, N, NRHS, dA, dla,
                       dp, dB, dlb, rout);
    rt_stage_out(c, B, dB, N, NRHS, ldb, dlb, esz, rout);
  }
  rt_stage_out(c, A, dA, N, N, lda, dla, esz, rout);
  if (dp != ipiv)
    RT_CHECK(rout, hipMemcpyAsync(ipiv, dp, (size_t)N * sizeof(int), hipMemcpyDeviceToHost,
                                  c->stream));
  rt_sync(c, rout);
  return info;
}

/* ------------------------------------------------------------------------
 * Cholesky, A = L L^H (Lower) or U^H U (Upper), right-looking over blocks of
 * CCHOL_NB: the diagonal block by one workgroup in LDS (k_cplx.hip
 * cpotf2), the panel by the complex TRSM (L21 = A21 L11^-H, or U12 =
 * U11^-H A12), and the trailing Hermitian update as the complex GEMM masked
 * to the referenced triangle (the real MFMA kernel with the complex-view
 * triangle mask): the blocked right-looking form of src/lapack/ATL_potrfL.c
 * / ATL_potrfU.c.  The real part of the diagonal is the pivot; info is the
 * 1-based column of the first non-positive one, after which the
 * factorization stops (each block's info is read back before the next).
 * ---------------------------------------------------------------------- */
#define CCHOL_NB 64
int NM(PFX, potrf_dev)(atlmx_ctx *c, int uplo, int N, T *A, long lda, const char *rout)
{
  static const T one[2] = {(T)1, (T)0}, mone[2] = {(T)-1, (T)0};
  const int upper = (uplo == KX_UPPER);
  int *d_info, info = 0, k;
  if (N <= 0) return 0;
  d_info = (int *)rt_ws(c, WS_INFO, sizeof(int), rout);
  RT_CHECK(rout, atlmx_k_fill_i32(d_info, 1, 0, c->stream));
  for (k = 0; k < N; k += CCHOL_NB) {
    const int jb = (N - k < CCHOL_NB) ? N - k : CCHOL_NB, nr = N - k - jb;
    T *Akk = A + 2 * ((size_t)k * lda + k);
    RT_CHECK(rout, K_CPOTF2(upper, jb, Akk, lda, k, d_info, c->stream));
    RT_CHECK(rout, hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost,
                                  c->stream));
    rt_sync(c, rout);
    if (info) return info;
    if (nr <= 0) break;
    if (!upper) {
      T *L21 = Akk + 2 * (size_t)jb;
      NM(PFX, trsm_dev)(c, KX_RIGHT, KX_LOWER, KX_CONJTRANS, KX_NONUNIT, nr, jb, one, Akk, lda,
                        L21, lda, rout);
      NM(PFX, gemm_dev)(c, KX_NOTRANS, KX_CONJTRANS, nr, nr, jb, mone, L21, lda, L21, lda, one,
                        L21 + 2 * (size_t)jb * lda, lda, KX_LOWER, rout);
    } else {
      T *U12 = Akk + 2 * (size_t)jb * lda;
      NM(PFX, trsm_dev)(c, KX_LEFT, KX_UPPER, KX_CONJTRANS, KX_NONUNIT, jb, nr, one, Akk, lda,
                        U12, lda, rout);
      NM(PFX, gemm_dev)(c, KX_CONJTRANS, KX_NOTRANS, nr, nr, jb, mone, U12, lda, U12, lda, one,
                        U12 + 2 * (size_t)jb, lda, KX_UPPER, rout);
    }
  }
  return 0;
}

/* A X = B from potrf_dev's factor (src/lapack/ATL_potrs.c, column-major):
 * L (L^H X) = B, or U^H (U X) = B */
void NM(PFX, potrs_dev)(atlmx_ctx *c, int uplo, int N, int NRHS, const T *A, long lda, T *B,
                        long ldb, const char *rout)
{
  static const T one[2] = {(T)1, (T)0};
  if (N <= 0 || NRHS <= 0) return;
  if (uplo == KX_LOWER) {
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_LOWER, KX_NOTRANS, KX_NONUNIT, N, NRHS, one, A, lda, B,
                      ldb, rout);
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_LOWER, KX_CONJTRANS, KX_NONUNIT, N, NRHS, one, A, lda, B,
                      ldb, rout);
  } else {
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_UPPER, KX_CONJTRANS, KX_NONUNIT, N, NRHS, one, A, lda, B,
                      ldb, rout);
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_UPPER, KX_NOTRANS, KX_NONUNIT, N, NRHS, one, A, lda, B,
                      ldb, rout);
  }
}

/* interfaces/lapack/C/src/clapack_zpotrf.c; RowMajor holds A^T = conj(A)
 * column-major, whose Upper factor is the RowMajor Lower one (and the
 * other way round), as for the real routines */
int NM3(clapack_, PFX, potrf)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                              const int N, void *A, const int lda)
{
  static const char *rout = "clapack_" STR(PFX) "potrf";
  const int ierr = atlmx_potrf_check(rout, Order, Uplo, N, lda);
  const size_t esz = 2 * sizeof(T);
  int low, info;
  atlmx_ctx *c;
  long dla;
  T *dA;
  if (ierr) return ierr;
  if (N == 0) return 0;
  low = (Uplo == CblasLower);
  if (Order == CblasRowMajor) low = !low;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, esz, 1, &dla, rout);
  info = NM(PFX, potrf_dev)(c, low ? KX_LOWER : KX_UPPER, N, dA, dla, rout);
  rt_stage_out(c, A, dA, N, N, lda, dla, esz, rout);
  rt_sync(c, rout);
  return info;
}

/* the solve for clapack_?potrs / ?posv: RowMajor solves conj(A) conj(X) =
 * conj(B) with the column-major factor, so B is conjugated before and after,
 * as src/lapack/ATL_potrs.c:46-69 does */
static void NM(PFX, potrs_order)(atlmx_ctx *c, int rowmajor, int low, int N, int NRHS,
                                 const T *dA, long dla, T *dB, long dlb, const char *rout)
{
  if (rowmajor) RT_CHECK(rout, K_CCONJ(N, NRHS, dB, dlb, c->stream));
  NM(PFX, potrs_dev)(c, low ? KX_LOWER : KX_UPPER, N, NRHS, dA, dla, dB, dlb, rout);
  if (rowmajor) RT_CHECK(rout, K_CCONJ(N, NRHS, dB, dlb, c->stream));
}

/* interfaces/lapack/C/src/clapack_zpotrs.c */
int NM3(clapack_, PFX, potrs)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                              const int N, const int NRHS, const void *A, const int lda, void *B,
                              const int ldb)
{
  static const char *rout = "clapack_" STR(PFX) "potrs";
  const int ierr = atlmx_posv_check(rout, Order, Uplo, N, NRHS, lda, ldb);
  const size_t esz = 2 * sizeof(T);
  int low;
  atlmx_ctx *c;
  long dla, dlb;
  T *dA, *dB;
  if (ierr) return ierr;
  if (N == 0 || NRHS == 0) return 0;
  low = (Uplo == CblasLower);
  if (Order == CblasRowMajor) low = !low;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, esz, 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, N, NRHS, ldb, esz, 1, &dlb, rout);
  NM(PFX, potrs_order)(c, Order == CblasRowMajor, low, N, NRHS, dA, dla, dB, dlb, rout);
  rt_stage_out(c, B, dB, N, NRHS, ldb, dlb, esz, rout);
  rt_sync(c, rout);
  return 0;
}

/* interfaces/lapack/C/src/clapack_zposv.c: potrf, then (info = 0) potrs */
int NM3(clapack_, PFX, posv)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                             const int N, const int NRHS, void *A, const int lda, void *B,
                             const int ldb)
{
  static const char *rout = "clapack_" STR(PFX) "posv";
  const int ierr = atlmx_posv_check(rout, Order, Uplo, N, NRHS, lda, ldb);
  const size_t esz = 2 * sizeof(T);
  int info, low;
  atlmx_ctx *c;
  long dla, dlb;
  T *dA, *dB;
  if (ierr) return ierr;
  if (N == 0) return 0;
  low = (Uplo == CblasLower);
  if (Order == CblasRowMajor) low = !low;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, esz, 1, &dla, rout);
  info = NM(PFX, potrf_dev)(c, low ? KX_LOWER : KX_UPPER, N, dA, dla, rout);
  if (!info && NRHS > 0) {
    dB = (T *)rt_stage_in(c, WS_STAGE_B, B, N, NRHS, ldb, esz, 1, &dlb, rout);
    NM(PFX, potrs_order)(c, Order == CblasRowMajor, low, N, NRHS, dA, dla, dB, dlb, rout);
    rt_stage_out(c, B, dB, N, NRHS, ldb, dlb, esz, rout);
  }
  rt_stage_out(c, A, dA, N, N, lda, dla, esz, rout);
  rt_sync(c, rout);
  return info;
}

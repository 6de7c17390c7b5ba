/*
 * f77_tmpl.h -- precision template of the Fortran 77 entry points
 * (included by f77.c with T=double/P=d and T=float/P=s, and with
 * F77_COMPLEX defined for T=double/P=z and T=float/P=c: then the GEMM and
 * Level-3 BLAS take alpha/beta as {real, imaginary} pairs, HEMM/HERK/HER2K
 * are added and the LAPACK inverses and I?AMAX are left out).  Check order and
 * XERBLA positions follow interfaces/blas/F77/src/dgemm.f, dtrsm.f,
 * dtrmm.f, dsyrk.f, dsyr2k.f, dsymm.f, idamax.f and
 * interfaces/lapack/F77/src/dgetrf.f, dgetrs.f, dgesv.f, dpotrf.f,
 * dpotrs.f, dposv.f, dgetri.f, dpotri.f, dtrtri.f, dlauum.f.
 */
#define FN_(a, b, c) a##b##c
#define FN(a, b, c) FN_(a, b, c)
#define F77(name) FN(P, name, _)
#define CB(name) FN(cblas_, P, name)
#define CL(name) FN(clapack_, P, name)
/* a BLAS scalar as the C ABI takes it: by address for complex types */
#undef SC
#ifdef F77_COMPLEX
#define SC(x) (x)
#else
#define SC(x) (*(x))
#endif

void F77(gemm)(const char *ta, const char *tb, const int *m, const int *n, const int *k,
               const T *alpha, const T *a, const int *lda, const T *b, const int *ldb,
               const T *beta, T *c, const int *ldc, size_t lta, size_t ltb)
{
  const int ita = tr_code(ta), itb = tr_code(tb);
  const int nra = (ita == CblasNoTrans) ? *m : *k, nrb = (itb == CblasNoTrans) ? *k : *n;
  int info = 0;
  (void)lta; (void)ltb;
  if (!ita) info = 1;
  else if (!itb) info = 2;
  else if (*m < 0) info = 3;
  else if (*n < 0) info = 4;
  else if (*k < 0) info = 5;
  else if (*lda < imax1(nra)) info = 8;
  else if (*ldb < imax1(nrb)) info = 10;
  else if (*ldc < imax1(*m)) info = 13;
  if (info) { f77_err(PU "GEMM ", info); return; }
#ifdef F77_COMPLEX
  CB(gemm)(CblasColMajor, ita, itb, *m, *n, *k, alpha, a, *lda, b, *ldb, beta, c, *ldc);
#else
  CB(gemm)(CblasColMajor, ita, itb, *m, *n, *k, *alpha, a, *lda, b, *ldb, *beta, c, *ldc);
#endif
}

/* TRSM and TRMM share their checks (dtrsm.f / dtrmm.f) */
static int FN(P, trxm_f77, _check)(const char *side, const char *uplo, const char *ta,
                                   const char *diag, const int *m, const int *n,
                                   const int *lda, const int *ldb)
{
  const int left = lsame(side, 'L');
  const int nra = left ? *m : *n;
  if (!left && !lsame(side, 'R')) return 1;
  if (!lsame(uplo, 'U') && !lsame(uplo, 'L')) return 2;
  if (!tr_code(ta)) return 3;
  if (!lsame(diag, 'N') && !lsame(diag, 'U')) return 4;
  if (*m < 0) return 5;
  if (*n < 0) return 6;
  if (*lda < imax1(nra)) return 9;
  if (*ldb < imax1(*m)) return 11;
  return 0;
}

void F77(trsm)(const char *side, const char *uplo, const char *ta, const char *diag,
               const int *m, const int *n, const T *alpha, const T *a, const int *lda, T *b,
               const int *ldb, size_t l1, size_t l2, size_t l3, size_t l4)
{
  const int info = FN(P, trxm_f77, _check)(side, uplo, ta, diag, m, n, lda, ldb);
  (void)l1; (void)l2; (void)l3; (void)l4;
  if (info) { f77_err(PU "TRSM ", info); return; }
  CB(trsm)(CblasColMajor, lsame(side, 'L') ? CblasLeft : CblasRight,
           lsame(uplo, 'U') ? CblasUpper : CblasLower, tr_code(ta),
           lsame(diag, 'U') ? CblasUnit : CblasNonUnit, *m, *n,
#ifdef F77_COMPLEX
           alpha,
#else
           *alpha,
#endif
           a, *lda, b, *ldb);
}

void F77(trmm)(const char *side, const char *uplo, const char *ta, const char *diag,
               const int *m, const int *n, const T *alpha, const T *a, const int *lda, T *b,
               const int *ldb, size_t l1, size_t l2, size_t l3, size_t l4)
{
  const int info = FN(P, trxm_f77, _check)(side, uplo, ta, diag, m, n, lda, ldb);
  (void)l1; (void)l2; (void)l3; (void)l4;
  if (info) { f77_err(PU "TRMM ", info); return; }
  CB(trmm)(CblasColMajor, lsame(side, 'L') ? CblasLeft : CblasRight,
           lsame(uplo, 'U') ? CblasUpper : CblasLower, tr_code(ta),
           lsame(diag, 'U') ? CblasUnit : CblasNonUnit, *m, *n, SC(alpha), a, *lda, b, *ldb);
}

/* TRANS of ?SYRK / ?SYR2K: N, T or C for real types; N or T for complex
 * types (zsyrk.f, zsyr2k.f) */
static int FN(P, syrk_f77, _tr)(const char *tr)
{
  const int itr = tr_code(tr);
#ifdef F77_COMPLEX
  if (itr == CblasConjTrans) return 0;
#endif
  return itr;
}

void F77(syrk)(const char *uplo, const char *tr, const int *n, const int *k, const T *alpha,
               const T *a, const int *lda, const T *beta, T *c, const int *ldc, size_t l1,
               size_t l2)
{
  const int itr = FN(P, syrk_f77, _tr)(tr);
  const int nra = (itr == CblasNoTrans) ? *n : *k;
  int info = 0;
  (void)l1; (void)l2;
  if (!lsame(uplo, 'L') && !lsame(uplo, 'U')) info = 1;
  else if (!itr) info = 2;
  else if (*n < 0) info = 3;
  else if (*k < 0) info = 4;
  else if (*lda < imax1(nra)) info = 7;
  else if (*ldc < imax1(*n)) info = 10;
  if (info) { f77_err(PU "SYRK ", info); return; }
  CB(syrk)(CblasColMajor, lsame(uplo, 'U') ? CblasUpper : CblasLower, itr, *n, *k, SC(alpha),
           a, *lda, SC(beta), c, *ldc);
}

void F77(syr2k)(const char *uplo, const char *tr, const int *n, const int *k, const T *alpha,
                const T *a, const int *lda, const T *b, const int *ldb, const T *beta, T *c,
                const int *ldc, size_t l1, size_t l2)
{
  const int itr = FN(P, syrk_f77, _tr)(tr);
  const int nra = (itr == CblasNoTrans) ? *n : *k;
  int info = 0;
  (void)l1; (void)l2;
  if (!lsame(uplo, 'L') && !lsame(uplo, 'U')) info = 1;
  else if (!itr) info = 2;
  else if (*n < 0) info = 3;
  else if (*k < 0) info = 4;
  else if (*lda < imax1(nra)) info = 7;
  else if (*ldb < imax1(nra)) info = 9;
  else if (*ldc < imax1(*n)) info = 12;
  if (info) { f77_err(PU "SYR2K", info); return; }
  CB(syr2k)(CblasColMajor, lsame(uplo, 'U') ? CblasUpper : CblasLower, itr, *n, *k, SC(alpha),
            a, *lda, b, *ldb, SC(beta), c, *ldc);
}

/* SYMM and the complex HEMM share their checks (dsymm.f, zhemm.f) */
static int FN(P, symm_f77, _check)(const char *side, const char *uplo, const int *m,
                                   const int *n, const int *lda, const int *ldb, const int *ldc)
{
  const int left = lsame(side, 'L');
  const int nra = left ? *m : *n;
  if (!left && !lsame(side, 'R')) return 1;
  if (!lsame(uplo, 'U') && !lsame(uplo, 'L')) return 2;
  if (*m < 0) return 3;
  if (*n < 0) return 4;
  if (*lda < imax1(nra)) return 7;
  if (*ldb < imax1(*m)) return 9;
  if (*ldc < imax1(*m)) return 12;
  return 0;
}

void F77(symm)(const char *side, const char *uplo, const int *m, const int *n, const T *alpha,
               const T *a, const int *lda, const T *b, const int *ldb, const T *beta, T *c,
               const int *ldc, size_t l1, size_t l2)
{
  const int info = FN(P, symm_f77, _check)(side, uplo, m, n, lda, ldb, ldc);
  (void)l1; (void)l2;
  if (info) { f77_err(PU "SYMM ", info); return; }
  CB(symm)(CblasColMajor, lsame(side, 'L') ? CblasLeft : CblasRight,
           lsame(uplo, 'U') ? CblasUpper : CblasLower, *m, *n, SC(alpha), a, *lda, b, *ldb,
           SC(beta), c, *ldc);
}

#ifdef F77_COMPLEX
/* zhemm.f */
void F77(hemm)(const char *side, const char *uplo, const int *m, const int *n, const T *alpha,
               const T *a, const int *lda, const T *b, const int *ldb, const T *beta, T *c,
               const int *ldc, size_t l1, size_t l2)
{
  const int info = FN(P, symm_f77, _check)(side, uplo, m, n, lda, ldb, ldc);
  (void)l1; (void)l2;
  if (info) { f77_err(PU "HEMM ", info); return; }
  CB(hemm)(CblasColMajor, lsame(side, 'L') ? CblasLeft : CblasRight,
           lsame(uplo, 'U') ? CblasUpper : CblasLower, *m, *n, alpha, a, *lda, b, *ldb, beta,
           c, *ldc);
}

/* zherk.f / zher2k.f: TRANS N or C; HERK's alpha and beta and HER2K's
 * beta are real */
void F77(herk)(const char *uplo, const char *tr, const int *n, const int *k, const T *alpha,
               const T *a, const int *lda, const T *beta, T *c, const int *ldc, size_t l1,
               size_t l2)
{
  const int itr = lsame(tr, 'N') ? CblasNoTrans : lsame(tr, 'C') ? CblasConjTrans : 0;
  const int nra = (itr == CblasNoTrans) ? *n : *k;
  int info = 0;
  (void)l1; (void)l2;
  if (!lsame(uplo, 'L') && !lsame(uplo, 'U')) info = 1;
  else if (!itr) info = 2;
  else if (*n < 0) info = 3;
  else if (*k < 0) info = 4;
  else if (*lda < imax1(nra)) info = 7;
  else if (*ldc < imax1(*n)) info = 10;
  if (info) { f77_err(PU "HERK ", info); return; }
  CB(herk)(CblasColMajor, lsame(uplo, 'U') ? CblasUpper : CblasLower, itr, *n, *k, *alpha, a,
           *lda, *beta, c, *ldc);
}

void F77(her2k)(const char *uplo, const char *tr, const int *n, const int *k, const T *alpha,
                const T *a, const int *lda, const T *b, const int *ldb, const T *beta, T *c,
                const int *ldc, size_t l1, size_t l2)
{
  const int itr = lsame(tr, 'N') ? CblasNoTrans : lsame(tr, 'C') ? CblasConjTrans : 0;
  const int nra = (itr == CblasNoTrans) ? *n : *k;
  int info = 0;
  (void)l1; (void)l2;
  if (!lsame(uplo, 'L') && !lsame(uplo, 'U')) info = 1;
  else if (!itr) info = 2;
  else if (*n < 0) info = 3;
  else if (*k < 0) info = 4;
  else if (*lda < imax1(nra)) info = 7;
  else if (*ldb < imax1(nra)) info = 9;
  else if (*ldc < imax1(*n)) info = 12;
  if (info) { f77_err(PU "HER2K", info); return; }
  CB(her2k)(CblasColMajor, lsame(uplo, 'U') ? CblasUpper : CblasLower, itr, *n, *k, alpha, a,
            *lda, b, *ldb, *beta, c, *ldc);
}
#endif /* F77_COMPLEX: HEMM, HERK, HER2K */

#ifndef F77_COMPLEX

/* idamax.f: 0 when N < 1 or INCX < 1, else the 1-based index */
int FN(i, P, amax_)(const int *n, const T *x, const int *incx)
{
  if (*n < 1 || *incx < 1) return 0;
  return (int)FN(cblas_i, P, amax)(*n, x, *incx) + 1;
}

/* ---- LAPACK ---- */
#endif /* !F77_COMPLEX: I?AMAX */

void F77(getrf)(const int *m, const int *n, T *a, const int *lda, int *ipiv, int *info)
{
  int i, mn;
  *info = 0;
  if (*m < 0) *info = -1;
  else if (*n < 0) *info = -2;
  else if (*lda < imax1(*m)) *info = -4;
  if (*info) { f77_err(PU "GETRF", -*info); return; }
  *info = CL(getrf)(CblasColMajor, *m, *n, a, *lda, ipiv);
  mn = *m < *n ? *m : *n;
  for (i = 0; i < mn; ++i) ipiv[i] += 1;      /* ATL_f77wrap_getrf.c:27 */
}

/* 1-based ipiv -> a 0-based copy for the C routines */
static int *FN(P, piv0, _)(int n, const int *ipiv)
{
  int i, *p = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  if (!p) { fprintf(stderr, "atlas_mi355x: out of host memory\n"); exit(-1); }
  for (i = 0; i < n; ++i) p[i] = ipiv[i] - 1;
  return p;
}

void F77(getrs)(const char *tr, const int *n, const int *nrhs, const T *a, const int *lda,
                const int *ipiv, T *b, const int *ldb, int *info, size_t l1)
{
  const int itr = tr_code(tr);
  int *p0;
  (void)l1;
  *info = 0;
  if (!itr) *info = -1;
  else if (*n < 0) *info = -2;
  else if (*nrhs < 0) *info = -3;
  else if (*lda < imax1(*n)) *info = -5;
  else if (*ldb < imax1(*n)) *info = -8;
  if (*info) { f77_err(PU "GETRS", -*info); return; }
  if (*n == 0 || *nrhs == 0) return;
  p0 = FN(P, piv0, _)(*n, ipiv);
  CL(getrs)(CblasColMajor, itr, *n, *nrhs, a, *lda, p0, b, *ldb);
  free(p0);
}

void F77(gesv)(const int *n, const int *nrhs, T *a, const int *lda, int *ipiv, T *b,
               const int *ldb, int *info)
{
  int i;
  *info = 0;
  if (*n < 0) *info = -1;
  else if (*nrhs < 0) *info = -2;
  else if (*lda < imax1(*n)) *info = -4;
  else if (*ldb < imax1(*n)) *info = -7;
  if (*info) { f77_err(PU "GESV ", -*info); return; }
  *info = CL(gesv)(CblasColMajor, *n, *nrhs, a, *lda, ipiv, b, *ldb);
  for (i = 0; i < *n; ++i) ipiv[i] += 1;
}

static int FN(P, uplo_f77, _)(const char *u)
{
  return lsame(u, 'U') ? CblasUpper : (lsame(u, 'L') ? CblasLower : 0);
}

void F77(potrf)(const char *uplo, const int *n, T *a, const int *lda, int *info, size_t l1)
{
  const int u = FN(P, uplo_f77, _)(uplo);
  (void)l1;
  *info = 0;
  if (!u) *info = -1;
  else if (*n < 0) *info = -2;
  else if (*lda < imax1(*n)) *info = -4;
  if (*info) { f77_err(PU "POTRF", -*info); return; }
  *info = CL(potrf)(CblasColMajor, u, *n, a, *lda);
}

void F77(potrs)(const char *uplo, const int *n, const int *nrhs, const T *a, const int *lda,
                T *b, const int *ldb, int *info, size_t l1)
{
  const int u = FN(P, uplo_f77, _)(uplo);
  (void)l1;
  *info = 0;
  if (!u) *info = -1;
  else if (*n < 0) *info = -2;
  else if (*nrhs < 0) *info = -3;
  else if (*lda < imax1(*n)) *info = -5;
  else if (*ldb < imax1(*n)) *info = -7;
  if (*info) { f77_err(PU "POTRS", -*info); return; }
  CL(potrs)(CblasColMajor, u, *n, *nrhs, a, *lda, b, *ldb);
}

void F77(posv)(const char *uplo, const int *n, const int *nrhs, T *a, const int *lda, T *b,
               const int *ldb, int *info, size_t l1)
{
  const int u = FN(P, uplo_f77, _)(uplo);
  (void)l1;
  *info = 0;
  if (!u) *info = -1;
  else if (*n < 0) *info = -2;
  else if (*nrhs < 0) *info = -3;
  else if (*lda < imax1(*n)) *info = -5;
  else if (*ldb < imax1(*n)) *info = -7;
  if (*info) { f77_err(PU "POSV ", -*info); return; }
  *info = CL(posv)(CblasColMajor, u, *n, *nrhs, a, *lda, b, *ldb);
}

#ifndef F77_COMPLEX
/* dgetri.f: LWORK = -1 asks for the workspace size (WORK(1) = N*64); the
 * GPU routine keeps its own workspace, so WORK is otherwise unused */
void F77(getri)(const int *n, T *a, const int *lda, const int *ipiv, T *work, const int *lwork,
                int *info)
{
  const int lquery = (*lwork == -1);
  int *p0;
  *info = 0;
  if (*n < 0) *info = -1;
  else if (*lda < imax1(*n)) *info = -3;
  else if (*lwork < imax1(*n) && !lquery) *info = -6;
  if (*info) { f77_err(PU "GETRI", -*info); return; }
  if (lquery) { work[0] = (T)((*n > 0 ? *n : 1) * 64); return; }
  if (*n == 0) return;
  p0 = FN(P, piv0, _)(*n, ipiv);
  *info = CL(getri)(CblasColMajor, *n, a, *lda, p0);
  free(p0);
}

void F77(potri)(const char *uplo, const int *n, T *a, const int *lda, int *info, size_t l1)
{
  const int u = FN(P, uplo_f77, _)(uplo);
  (void)l1;
  *info = 0;
  if (!u) *info = -1;
  else if (*n < 0) *info = -2;
  else if (*lda < imax1(*n)) *info = -4;
  if (*info) { f77_err(PU "POTRI", -*info); return; }
  *info = CL(potri)(CblasColMajor, u, *n, a, *lda);
}

void F77(lauum)(const char *uplo, const int *n, T *a, const int *lda, int *info, size_t l1)
{
  const int u = FN(P, uplo_f77, _)(uplo);
  (void)l1;
  *info = 0;
  if (!u) *info = -1;
  else if (*n < 0) *info = -2;
  else if (*lda < imax1(*n)) *info = -4;
  if (*info) { f77_err(PU "LAUUM", -*info); return; }
  *info = CL(lauum)(CblasColMajor, u, *n, a, *lda);
}

void F77(trtri)(const char *uplo, const char *diag, const int *n, T *a, const int *lda,
                int *info, size_t l1, size_t l2)
{
  const int u = FN(P, uplo_f77, _)(uplo);
  (void)l1; (void)l2;
  *info = 0;
  if (!u) *info = -1;
  else if (!lsame(diag, 'N') && !lsame(diag, 'U')) *info = -2;
  else if (*n < 0) *info = -3;
  else if (*lda < imax1(*n)) *info = -5;
  if (*info) { f77_err(PU "TRTRI", -*info); return; }
  *info = CL(trtri)(CblasColMajor, u, lsame(diag, 'U') ? CblasUnit : CblasNonUnit, *n, a,
                    *lda);
}

#undef FN_
#undef FN
#undef F77
#undef CB
#undef CL
#endif /* !F77_COMPLEX: GETRI, POTRI, LAUUM, TRTRI */

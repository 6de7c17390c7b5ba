/*
 * atlas_f77.h -- the Fortran 77 ABI of libatlas_mi355x, for programs that
 * link ATLAS's -lf77blas / -llapack (gfortran convention: every argument
 * by reference, one hidden size_t length per CHARACTER argument at the
 * end, 1-based pivots).  Each entry replaces the same-named routine of
 * ATLAS's interfaces/blas/F77/src/<name>.f or
 * interfaces/lapack/F77/src/<name>.f (argument checks and XERBLA positions
 * as there; the pivot shift of
 * interfaces/lapack/F77/src/f77wrap/ATL_f77wrap_getrf.c:8-29).
 */
#ifndef ATLAS_F77_H
#define ATLAS_F77_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* interfaces/blas/F77/src/xerbla.f: message and STOP; weak, replaceable */
void xerbla_(const char *srname, const int *info, size_t len);

/* DGEMM, DTRSM, DTRMM, DSYRK, DSYR2K, DSYMM, IDAMAX
 * (interfaces/blas/F77/src/dgemm.f, dtrsm.f, dtrmm.f, dsyrk.f,
 * dsyr2k.f, dsymm.f, idamax.f) */
void dgemm_(const char *ta, const char *tb, const int *m, const int *n, const int *k,
            const double *alpha, const double *a, const int *lda, const double *b, const int *ldb,
            const double *beta, double *c, const int *ldc, size_t lta, size_t ltb);
void dtrsm_(const char *side, const char *uplo, const char *ta, const char *diag,
            const int *m, const int *n, const double *alpha, const double *a, const int *lda,
            double *b, const int *ldb, size_t l1, size_t l2, size_t l3, size_t l4);
void dtrmm_(const char *side, const char *uplo, const char *ta, const char *diag,
            const int *m, const int *n, const double *alpha, const double *a, const int *lda,
            double *b, const int *ldb, size_t l1, size_t l2, size_t l3, size_t l4);
void dsyrk_(const char *uplo, const char *tr, const int *n, const int *k, const double *alpha,
            const double *a, const int *lda, const double *beta, double *c, const int *ldc, size_t l1,
            size_t l2);
void dsyr2k_(const char *uplo, const char *tr, const int *n, const int *k,
             const double *alpha, const double *a, const int *lda, const double *b, const int *ldb,
             const double *beta, double *c, const int *ldc, size_t l1, size_t l2);
void dsymm_(const char *side, const char *uplo, const int *m, const int *n,
            const double *alpha, const double *a, const int *lda, const double *b, const int *ldb,
            const double *beta, double *c, const int *ldc, size_t l1, size_t l2);
int idamax_(const int *n, const double *x, const int *incx);

/* DGETRF, DGETRS, DGESV, DPOTRF, DPOTRS, DPOSV, DGETRI, DPOTRI,
 * DLAUUM, DTRTRI (interfaces/lapack/F77/src/dgetrf.f ... dtrtri.f) */
void dgetrf_(const int *m, const int *n, double *a, const int *lda, int *ipiv, int *info);
void dgetrs_(const char *tr, const int *n, const int *nrhs, const double *a, const int *lda,
             const int *ipiv, double *b, const int *ldb, int *info, size_t l1);
void dgesv_(const int *n, const int *nrhs, double *a, const int *lda, int *ipiv, double *b,
            const int *ldb, int *info);
void dpotrf_(const char *uplo, const int *n, double *a, const int *lda, int *info, size_t l1);
void dpotrs_(const char *uplo, const int *n, const int *nrhs, const double *a, const int *lda,
             double *b, const int *ldb, int *info, size_t l1);
void dposv_(const char *uplo, const int *n, const int *nrhs, double *a, const int *lda,
            double *b, const int *ldb, int *info, size_t l1);
void dgetri_(const int *n, double *a, const int *lda, const int *ipiv, double *work,
             const int *lwork, int *info);
void dpotri_(const char *uplo, const int *n, double *a, const int *lda, int *info, size_t l1);
void dlauum_(const char *uplo, const int *n, double *a, const int *lda, int *info, size_t l1);
void dtrtri_(const char *uplo, const char *diag, const int *n, double *a, const int *lda,
             int *info, size_t l1, size_t l2);

/* SGEMM, STRSM, STRMM, SSYRK, SSYR2K, SSYMM, ISAMAX
 * (interfaces/blas/F77/src/sgemm.f, strsm.f, strmm.f, ssyrk.f,
 * ssyr2k.f, ssymm.f, isamax.f) */
void sgemm_(const char *ta, const char *tb, const int *m, const int *n, const int *k,
            const float *alpha, const float *a, const int *lda, const float *b, const int *ldb,
            const float *beta, float *c, const int *ldc, size_t lta, size_t ltb);
void strsm_(const char *side, const char *uplo, const char *ta, const char *diag,
            const int *m, const int *n, const float *alpha, const float *a, const int *lda,
            float *b, const int *ldb, size_t l1, size_t l2, size_t l3, size_t l4);
void strmm_(const char *side, const char *uplo, const char *ta, const char *diag,
            const int *m, const int *n, const float *alpha, const float *a, const int *lda,
            float *b, const int *ldb, size_t l1, size_t l2, size_t l3, size_t l4);
void ssyrk_(const char *uplo, const char *tr, const int *n, const int *k, const float *alpha,
            const float *a, const int *lda, const float *beta, float *c, const int *ldc, size_t l1,
            size_t l2);
void ssyr2k_(const char *uplo, const char *tr, const int *n, const int *k,
             const float *alpha, const float *a, const int *lda, const float *b, const int *ldb,
             const float *beta, float *c, const int *ldc, size_t l1, size_t l2);
void ssymm_(const char *side, const char *uplo, const int *m, const int *n,
            const float *alpha, const float *a, const int *lda, const float *b, const int *ldb,
            const float *beta, float *c, const int *ldc, size_t l1, size_t l2);
int isamax_(const int *n, const float *x, const int *incx);

/* SGETRF, SGETRS, SGESV, SPOTRF, SPOTRS, SPOSV, SGETRI, SPOTRI,
 * SLAUUM, STRTRI (interfaces/lapack/F77/src/sgetrf.f ... strtri.f) */
void sgetrf_(const int *m, const int *n, float *a, const int *lda, int *ipiv, int *info);
void sgetrs_(const char *tr, const int *n, const int *nrhs, const float *a, const int *lda,
             const int *ipiv, float *b, const int *ldb, int *info, size_t l1);
void sgesv_(const int *n, const int *nrhs, float *a, const int *lda, int *ipiv, float *b,
            const int *ldb, int *info);
void spotrf_(const char *uplo, const int *n, float *a, const int *lda, int *info, size_t l1);
void spotrs_(const char *uplo, const int *n, const int *nrhs, const float *a, const int *lda,
             float *b, const int *ldb, int *info, size_t l1);
void sposv_(const char *uplo, const int *n, const int *nrhs, float *a, const int *lda,
            float *b, const int *ldb, int *info, size_t l1);
void sgetri_(const int *n, float *a, const int *lda, const int *ipiv, float *work,
             const int *lwork, int *info);
void spotri_(const char *uplo, const int *n, float *a, const int *lda, int *info, size_t l1);
void slauum_(const char *uplo, const int *n, float *a, const int *lda, int *info, size_t l1);
void strtri_(const char *uplo, const char *diag, const int *n, float *a, const int *lda,
             int *info, size_t l1, size_t l2);

/* Complex (z: COMPLEX*16, c: COMPLEX; arrays and alpha/beta as interleaved
 * {real, imaginary} pairs): ZGEMM, ZTRSM (interfaces/blas/F77/src/zgemm.f,
 * ztrsm.f), ZGETRF, ZGETRS, ZGESV, ZPOTRF, ZPOTRS, ZPOSV
 * (interfaces/lapack/F77/src/zgetrf.f ... zposv.f), and the c twins;
 * TRANSA/TRANSB/TRANS may be 'C'. */
void zgemm_(const char *ta, const char *tb, const int *m, const int *n, const int *k,
            const double *alpha, const double *a, const int *lda, const double *b, const int *ldb,
            const double *beta, double *c, const int *ldc, size_t lta, size_t ltb);
void ztrsm_(const char *side, const char *uplo, const char *ta, const char *diag,
            const int *m, const int *n, const double *alpha, const double *a, const int *lda,
            double *b, const int *ldb, size_t l1, size_t l2, size_t l3, size_t l4);
void zgetrf_(const int *m, const int *n, double *a, const int *lda, int *ipiv, int *info);
void zgetrs_(const char *tr, const int *n, const int *nrhs, const double *a, const int *lda,
             const int *ipiv, double *b, const int *ldb, int *info, size_t l1);
void zgesv_(const int *n, const int *nrhs, double *a, const int *lda, int *ipiv, double *b,
            const int *ldb, int *info);
void zpotrf_(const char *uplo, const int *n, double *a, const int *lda, int *info, size_t l1);
void zpotrs_(const char *uplo, const int *n, const int *nrhs, const double *a, const int *lda,
             double *b, const int *ldb, int *info, size_t l1);
void zposv_(const char *uplo, const int *n, const int *nrhs, double *a, const int *lda,
            double *b, const int *ldb, int *info, size_t l1);
void cgemm_(const char *ta, const char *tb, const int *m, const int *n, const int *k,
            const float *alpha, const float *a, const int *lda, const float *b, const int *ldb,
            const float *beta, float *c, const int *ldc, size_t lta, size_t ltb);
void ctrsm_(const char *side, const char *uplo, const char *ta, const char *diag,
            const int *m, const int *n, const float *alpha, const float *a, const int *lda,
            float *b, const int *ldb, size_t l1, size_t l2, size_t l3, size_t l4);
void cgetrf_(const int *m, const int *n, float *a, const int *lda, int *ipiv, int *info);
void cgetrs_(const char *tr, const int *n, const int *nrhs, const float *a, const int *lda,
             const int *ipiv, float *b, const int *ldb, int *info, size_t l1);
void cgesv_(const int *n, const int *nrhs, float *a, const int *lda, int *ipiv, float *b,
            const int *ldb, int *info);
void cpotrf_(const char *uplo, const int *n, float *a, const int *lda, int *info, size_t l1);
void cpotrs_(const char *uplo, const int *n, const int *nrhs, const float *a, const int *lda,
             float *b, const int *ldb, int *info, size_t l1);
void cposv_(const char *uplo, const int *n, const int *nrhs, float *a, const int *lda,
            float *b, const int *ldb, int *info, size_t l1);
/* complex TRMM, SYRK, SYR2K, SYMM, HEMM, HERK, HER2K (interfaces/blas/F77/src/
 * ztrmm.f, zsyrk.f, zsyr2k.f, zsymm.f, zhemm.f, zherk.f, zher2k.f) and the c
 * twins: ?SYRK/?SYR2K take TRANS N or T, ?HERK/?HER2K N or C; HERK's alpha
 * and beta and HER2K's beta are real scalars (one double / float). */
void ztrmm_(const char *side, const char *uplo, const char *ta, const char *diag,
            const int *m, const int *n, const double *alpha, const double *a, const int *lda,
            double *b, const int *ldb, size_t l1, size_t l2, size_t l3, size_t l4);
void zsyrk_(const char *uplo, const char *tr, const int *n, const int *k, const double *alpha,
            const double *a, const int *lda, const double *beta, double *c, const int *ldc, size_t l1,
            size_t l2);
void zsyr2k_(const char *uplo, const char *tr, const int *n, const int *k, const double *alpha,
             const double *a, const int *lda, const double *b, const int *ldb, const double *beta,
             double *c, const int *ldc, size_t l1, size_t l2);
void zsymm_(const char *side, const char *uplo, const int *m, const int *n, const double *alpha,
            const double *a, const int *lda, const double *b, const int *ldb, const double *beta,
            double *c, const int *ldc, size_t l1, size_t l2);
void zhemm_(const char *side, const char *uplo, const int *m, const int *n, const double *alpha,
            const double *a, const int *lda, const double *b, const int *ldb, const double *beta,
            double *c, const int *ldc, size_t l1, size_t l2);
void zherk_(const char *uplo, const char *tr, const int *n, const int *k, const double *alpha,
            const double *a, const int *lda, const double *beta, double *c, const int *ldc, size_t l1,
            size_t l2);
void zher2k_(const char *uplo, const char *tr, const int *n, const int *k, const double *alpha,
             const double *a, const int *lda, const double *b, const int *ldb, const double *beta,
             double *c, const int *ldc, size_t l1, size_t l2);
void ctrmm_(const char *side, const char *uplo, const char *ta, const char *diag,
            const int *m, const int *n, const float *alpha, const float *a, const int *lda,
            float *b, const int *ldb, size_t l1, size_t l2, size_t l3, size_t l4);
void csyrk_(const char *uplo, const char *tr, const int *n, const int *k, const float *alpha,
            const float *a, const int *lda, const float *beta, float *c, const int *ldc, size_t l1,
            size_t l2);
void csyr2k_(const char *uplo, const char *tr, const int *n, const int *k, const float *alpha,
             const float *a, const int *lda, const float *b, const int *ldb, const float *beta,
             float *c, const int *ldc, size_t l1, size_t l2);
void csymm_(const char *side, const char *uplo, const int *m, const int *n, const float *alpha,
            const float *a, const int *lda, const float *b, const int *ldb, const float *beta,
            float *c, const int *ldc, size_t l1, size_t l2);
void chemm_(const char *side, const char *uplo, const int *m, const int *n, const float *alpha,
            const float *a, const int *lda, const float *b, const int *ldb, const float *beta,
            float *c, const int *ldc, size_t l1, size_t l2);
void cherk_(const char *uplo, const char *tr, const int *n, const int *k, const float *alpha,
            const float *a, const int *lda, const float *beta, float *c, const int *ldc, size_t l1,
            size_t l2);
void cher2k_(const char *uplo, const char *tr, const int *n, const int *k, const float *alpha,
             const float *a, const int *lda, const float *b, const int *ldb, const float *beta,
             float *c, const int *ldc, size_t l1, size_t l2);

#ifdef __cplusplus
}
#endif
#endif

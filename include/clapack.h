/*
 * clapack.h -- ATLAS-style C LAPACK interface of libatlas_mi355x.
 *
 * Same names, argument meaning and return conventions as ATLAS's
 * include/clapack.h: ipiv is 0-based (src/lapack/ATL_getrfC.c:299,325);
 * getrf's positive info is the 1-based index of the first exactly-zero
 * pivot and the factorization completes (clapack_dgetrf.c:35-45); potrf's
 * positive info is the first non-positive leading minor and it stops
 * (src/lapack/ATL_potrfL.c:238-240); illegal arguments go through
 * cblas_xerbla, which exits.
 */
#ifndef CLAPACK_H
#define CLAPACK_H

#include "cblas.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ATLAS include/clapack.h:21-37 (single) and :60-76 (double). */
int clapack_sgesv(const enum CBLAS_ORDER Order, const int N, const int NRHS,
                  float *A, const int lda, int *ipiv, float *B, const int ldb);
int clapack_sgetrf(const enum CBLAS_ORDER Order, const int M, const int N,
                   float *A, const int lda, int *ipiv);
int clapack_sgetrs(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE Trans,
                   const int N, const int NRHS, const float *A, const int lda,
                   const int *ipiv, float *B, const int ldb);
int clapack_sposv(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const int N, const int NRHS, float *A, const int lda,
                  float *B, const int ldb);
int clapack_spotrf(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const int N, float *A, const int lda);
int clapack_spotrs(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const int N, const int NRHS, const float *A, const int lda,
                   float *B, const int ldb);

int clapack_dgesv(const enum CBLAS_ORDER Order, const int N, const int NRHS,
                  double *A, const int lda, int *ipiv, double *B, const int ldb);
int clapack_dgetrf(const enum CBLAS_ORDER Order, const int M, const int N,
                   double *A, const int lda, int *ipiv);
int clapack_dgetrs(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE Trans,
                   const int N, const int NRHS, const double *A, const int lda,
                   const int *ipiv, double *B, const int ldb);
int clapack_dposv(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const int N, const int NRHS, double *A, const int lda,
                  double *B, const int ldb);
int clapack_dpotrf(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const int N, double *A, const int lda);
int clapack_dpotrs(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const int N, const int NRHS, const double *A, const int lda,
                   double *B, const int ldb);

/* The inverses: ATLAS include/clapack.h:30,40,42,44 (single) and
 * :69,79,81,83 (double); interfaces/lapack/C/src/clapack_dgetri.c ->
 * src/lapack/ATL_getri.c:8, clapack_dpotri.c, clapack_dlauum.c ->
 * ATL_lauum.c:6, clapack_dtrtri.c -> ATL_trtri.c:8.  Positive info: the
 * 1-based index of the first exactly-zero diagonal entry of U (getri) or of
 * A (trtri, potri, non-unit); A is left unchanged then. */
int clapack_sgetri(const enum CBLAS_ORDER Order, const int N, float *A, const int lda,
                   const int *ipiv);
int clapack_spotri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   float *A, const int lda);
int clapack_slauum(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   float *A, const int lda);
int clapack_strtri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const enum CBLAS_DIAG Diag, const int N, float *A, const int lda);
int clapack_dgetri(const enum CBLAS_ORDER Order, const int N, double *A, const int lda,
                   const int *ipiv);
int clapack_dpotri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   double *A, const int lda);
int clapack_dlauum(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   double *A, const int lda);
int clapack_dtrtri(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                   const enum CBLAS_DIAG Diag, const int N, double *A, const int lda);

/* QR: ATLAS include/clapack.h:57 / :96, interfaces/lapack/C/src/
 * clapack_dgeqrf.c -> src/lapack/ATL_geqrf.c:29.  LAPACK's storage (R on
 * and above the diagonal, Householder vectors below, TAU[min(M,N)]);
 * returns 0. */
int clapack_sgeqrf(const enum CBLAS_ORDER Order, const int M, const int N, float *A,
                   const int lda, float *TAU);
int clapack_dgeqrf(const enum CBLAS_ORDER Order, const int M, const int N, double *A,
                   const int lda, double *TAU);

/* ---- complex LU family (c: float, z: double; interleaved real/imaginary
 * parts, leading dimensions in complex elements, 0-based pivots; pivot =
 * first row of largest |Re| + |Im|).  Replace ATLAS include/clapack.h:99-106
 * (c) and :138-145 (z), interfaces/lapack/C/src/clapack_zgetrf.c,
 * clapack_zgetrs.c, clapack_zgesv.c. */
int clapack_cgesv(const enum CBLAS_ORDER Order, const int N, const int NRHS, void *A,
                  const int lda, int *ipiv, void *B, const int ldb);
int clapack_cgetrf(const enum CBLAS_ORDER Order, const int M, const int N, void *A,
                   const int lda, int *ipiv);
int clapack_cgetrs(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE Trans,
                   const int N, const int NRHS, const void *A, const int lda,
                   const int *ipiv, void *B, const int ldb);
int clapack_zgesv(const enum CBLAS_ORDER Order, const int N, const int NRHS, void *A,
                  const int lda, int *ipiv, void *B, const int ldb);
int clapack_zgetrf(const enum CBLAS_ORDER Order, const int M, const int N, void *A,
                   const int lda, int *ipiv);
int clapack_zgetrs(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE Trans,
                   const int N, const int NRHS, const void *A, const int lda,
                   const int *ipiv, void *B, const int ldb);
/* Complex Cholesky family (Hermitian positive definite: A = L L^H or U^H U;
 * the real part of the diagonal is the pivot).  Replace ATLAS
 * include/clapack.h:107-115 (c) and :146-154 (z),
 * interfaces/lapack/C/src/clapack_zpotrf.c, clapack_zpotrs.c,
 * clapack_zposv.c. */
int clapack_cposv(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                  const int NRHS, void *A, const int lda, void *B, const int ldb);
int clapack_cpotrf(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   void *A, const int lda);
int clapack_cpotrs(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   const int NRHS, const void *A, const int lda, void *B, const int ldb);
int clapack_zposv(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                  const int NRHS, void *A, const int lda, void *B, const int ldb);
int clapack_zpotrf(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   void *A, const int lda);
int clapack_zpotrs(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo, const int N,
                   const int NRHS, const void *A, const int lda, void *B, const int ldb);

#ifdef __cplusplus
}
#endif
#endif

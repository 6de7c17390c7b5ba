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

#ifdef __cplusplus
}
#endif
#endif

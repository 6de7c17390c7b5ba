/*
 * blas3.c -- Level-3 BLAS layer (C host code) of libatlas_mi355x:
 * cblas_{s,d}{gemm,trsm,syrk}, cblas_i{s,d}amax and their device-pointer
 * twins, instantiated from blas3_tmpl.h per precision.
 */
#include <stdio.h>
#include <stdlib.h>
#include "rt.h"
#include "drivers.h"
#include "../../include/cblas.h"
#include "../../include/atlmx.h"

#define STR_(x) #x
#define STR(x) STR_(x)

/* Argument checks of cblas_?gemm for every type (real and complex), in
 * ATLAS's order, positions and messages: interfaces/blas/C/src/
 * cblas_dgemm.c:25-117 (cblas_zgemm.c:27-115 is the same).  2000 = valid. */
int atlmx_gemm_check(int Order, int TA, int TB, int M, int N, int K, int lda, int ldb,
                     int ldc)
{
  int info = 2000;
  if (M < 0) info = cblas_errprn(4, info, "M cannot be less than zero 0,; is set to %d.", M);
  if (N < 0) info = cblas_errprn(5, info, "N cannot be less than zero 0,; is set to %d.", N);
  if (K < 0) info = cblas_errprn(6, info, "K cannot be less than zero 0,; is set to %d.", K);
  if (Order == CblasRowMajor) {
    if (TA == CblasNoTrans) {
      if (lda < K || lda < 1)
        info = cblas_errprn(9, info, "lda must be >= MAX(K,1): lda=%d K=%d", lda, K);
    } else {
      if (TA != CblasConjTrans && TA != CblasTrans)
        info = cblas_errprn(2, info, "TransA must be %d, %d or %d, but is set to %d",
                            CblasNoTrans, CblasTrans, CblasConjTrans, TA);
      if (lda < M || lda < 1)
        info = cblas_errprn(9, info, "lda must be >= MAX(M,1): lda=%d M=%d", lda, M);
    }
    if (TB == CblasNoTrans) {
      if (ldb < N || ldb < 1)
        info = cblas_errprn(11, info, "ldb must be >= MAX(N,1): ldb=%d N=%d", ldb, N);
    } else {
      if (TB != CblasConjTrans && TB != CblasTrans)
        info = cblas_errprn(3, info, "TransB must be %d, %d or %d, but is set to %d",
                            CblasNoTrans, CblasTrans, CblasConjTrans, TB);
      if (ldb < K || ldb < 1)
        info = cblas_errprn(11, info, "ldb must be >= MAX(K,1): ldb=%d K=%d", ldb, K);
    }
    if (ldc < N || ldc < 1)
      info = cblas_errprn(14, info, "ldc must be >= MAX(N,1): ldc=%d N=%d", ldc, N);
  } else if (Order == CblasColMajor) {
    if (TA == CblasNoTrans) {
      if (lda < M || lda < 1)
        info = cblas_errprn(9, info, "lda must be >= MAX(M,1): lda=%d M=%d", lda, M);
    } else {
      if (TA != CblasConjTrans && TA != CblasTrans)
        info = cblas_errprn(2, info, "TransA must be %d, %d or %d, but is set to %d",
                            CblasNoTrans, CblasTrans, CblasConjTrans, TA);
      if (lda < K || lda < 1)
        info = cblas_errprn(9, info, "lda must be >= MAX(K,1): lda=%d K=%d", lda, K);
    }
    if (TB == CblasNoTrans) {
      if (ldb < K || ldb < 1)
        info = cblas_errprn(11, info, "ldb must be >= MAX(K,1): ldb=%d K=%d", ldb, K);
    } else {
      if (TB != CblasConjTrans && TB != CblasTrans)
        info = cblas_errprn(3, info, "TransB must be %d, %d or %d, but is set to %d",
                            CblasNoTrans, CblasTrans, CblasConjTrans, TB);
      if (ldb < N || ldb < 1)
        info = cblas_errprn(11, info, "ldb must be >= MAX(N,1): ldb=%d N=%d", ldb, N);
    }
    if (ldc < M || ldc < 1)
      info = cblas_errprn(14, info, "ldc must be >= MAX(M,1): ldc=%d M=%d", ldc, M);
  } else
    info = cblas_errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                        CblasColMajor, Order);
  return info;
}

/* Argument checks shared by cblas_?trsm and cblas_?trmm of every type,
 * whose checks are the same (interfaces/blas/C/src/cblas_dtrsm.c:23-75,
 * cblas_dtrmm.c, cblas_ztrsm.c).  2000 = valid. */
int atlmx_trxm_check(int Order, int Side, int Uplo, int TA, int Diag, int M, int N, int lda,
                     int ldb)
{
  int info = 2000;
  if (Order == CblasColMajor) {
    if (Side == CblasLeft) {
      if (lda < M || lda < 1)
        info = cblas_errprn(10, info, "lda must be >= MAX(M,1): lda=%d M=%d", lda, M);
    } else if (Side == CblasRight) {
      if (lda < N || lda < 1)
        info = cblas_errprn(10, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
    } else
      info = cblas_errprn(2, info, "SIDE must be %d or %d, but is set to %d", CblasRight,
                          CblasLeft, Side);
    if (ldb < M || ldb < 1)
      info = cblas_errprn(12, info, "ldb must be >= MAX(M,1): ldb=%d M=%d", ldb, M);
  } else if (Order == CblasRowMajor) {
    if (Side == CblasLeft) {
      if (lda < M || lda < 1)
        info = cblas_errprn(10, info, "lda must be >= MAX(M,1): lda=%d M=%d", lda, M);
    } else if (Side == CblasRight) {
      if (lda < N || lda < 1)
        info = cblas_errprn(10, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
    } else
      info = cblas_errprn(2, info, "SIDE must be %d or %d, but is set to %d", CblasRight,
                          CblasLeft, Side);
    if (ldb < N || ldb < 1)
      info = cblas_errprn(12, info, "ldb must be >= MAX(N,1): ldb=%d N=%d", ldb, N);
  } else
    info = cblas_errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                        CblasColMajor, Order);
  if (Uplo != CblasUpper && Uplo != CblasLower)
    info = cblas_errprn(3, info, "UPLO must be %d or %d, but is set to %d", CblasUpper,
                        CblasLower, Uplo);
  if (TA != CblasNoTrans && TA != CblasTrans && TA != CblasConjTrans)
    info = cblas_errprn(4, info, "TransA must be %d, %d or %d, but is set to %d",
                        CblasNoTrans, CblasTrans, CblasConjTrans, TA);
  if (Diag != CblasUnit && Diag != CblasNonUnit)
    info = cblas_errprn(5, info, "UPLO must be %d or %d, but is set to %d", CblasUnit,
                        CblasNonUnit, Diag);
  if (M < 0) info = cblas_errprn(6, info, "M cannot be less than zero; it is set to %d.", M);
  if (N < 0) info = cblas_errprn(7, info, "N cannot be less than zero; it is set to %d.", N);
  return info;
}

/* Argument checks shared by cblas_?symm and the complex cblas_?symm /
 * cblas_?hemm, whose checks are the same (interfaces/blas/C/src/
 * cblas_dsymm.c, cblas_zsymm.c, cblas_zhemm.c:25-72).  2000 = valid. */
int atlmx_symm_check(int Order, int Side, int Uplo, int M, int N, int lda, int ldb, int ldc)
{
  int info = 2000;
  if (Order == CblasColMajor || Order == CblasRowMajor) {
    const int rows = (Order == CblasColMajor) ? M : N;
    const char *rn = (Order == CblasColMajor) ? "M" : "N";
    if (Side == CblasLeft) {
      if (lda < M || lda < 1)
        info = cblas_errprn(8, info, "lda must be >= MAX(M,1): lda=%d M=%d", lda, M);
    } else if (Side == CblasRight) {
      if (lda < N || lda < 1)
        info = cblas_errprn(8, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
    } else
      info = cblas_errprn(2, info, "SIDE must be %d or %d, but is set to %d", CblasRight,
                          CblasLeft, Side);
    if (ldb < rows || ldb < 1)
      info = cblas_errprn(10, info, "ldb must be >= MAX(%s,1): ldb=%d %s=%d", rn, ldb, rn,
                          rows);
    if (ldc < rows || ldc < 1)
      info = cblas_errprn(13, info, "ldc must be >= MAX(%s,1): ldc=%d %s=%d", rn, ldc, rn,
                          rows);
  } else
    info = cblas_errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                        CblasColMajor, Order);
  if (Uplo != CblasUpper && Uplo != CblasLower)
    info = cblas_errprn(3, info, "UPLO must be %d or %d, but is set to %d", CblasUpper,
                        CblasLower, Uplo);
  if (M < 0) info = cblas_errprn(4, info, "M cannot be less than zero; it is set to %d.", M);
  if (N < 0) info = cblas_errprn(5, info, "N cannot be less than zero; it is set to %d.", N);
  return info;
}

#define T double
#define PFX d
#define K_GEMM atlmx_k_dgemm
#define K_TRI_INV atlmx_k_dtri_inv
#define K_SCAL atlmx_k_dscal2d
#define K_IAMAX atlmx_k_idamax
#define K_TRSV atlmx_k_dtrsv
#define K_TRSM_TAG atlmx_k_dtrsm_tag
#define K_TRICOPY atlmx_k_dtricopy
#define K_UPDATE atlmx_k_dupdate
#include "blas3_tmpl.h"
#undef T
#undef PFX
#undef K_UPDATE
#undef K_GEMM
#undef K_TRI_INV
#undef K_SCAL
#undef K_IAMAX
#undef K_TRSV
#undef K_TRSM_TAG
#undef K_TRICOPY

#define T float
#define PFX s
#define K_GEMM atlmx_k_sgemm
#define K_TRI_INV atlmx_k_stri_inv
#define K_SCAL atlmx_k_sscal2d
#define K_IAMAX atlmx_k_isamax
#define K_TRSV atlmx_k_strsv
#define K_TRSM_TAG atlmx_k_strsm_tag
#define K_TRICOPY atlmx_k_stricopy
#include "blas3_tmpl.h"

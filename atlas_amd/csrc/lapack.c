/*
 * lapack.c -- LAPACK layer (C host code) of libatlas_mi355x:
 * clapack_{s,d}{getrf,getrs,potrf,potrs,gesv,posv} and their
 * device-pointer twins, instantiated from lapack_tmpl.h per precision.
 */
#include <stdio.h>
#include <stdlib.h>
#include "rt.h"
#include "../../include/clapack.h"
#include "../../include/atlmx.h"

#define STR_(x) #x
#define STR(x) STR_(x)
#define NM_(a, b) a##b
#define NM(a, b) NM_(a, b)
#define NM3_(a, b, c) a##b##c
#define NM3(a, b, c) NM3_(a, b, c)

#include "drivers.h"

/* Argument checks of clapack_?getrf / ?getrs / ?gesv for every type (real
 * and complex), in ATLAS's order, positions and messages
 * (interfaces/lapack/C/src/clapack_dgetrf.c:48-86, clapack_dgetrs.c,
 * clapack_dgesv.c; the z/c files are the same).  Each returns 0 or the
 * negated position of the last bad argument, after xerbla for each. */
int atlmx_getrf_check(const char *rout, int Order, int M, int N, int lda)
{
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, rout, "Order must be %d or %d, but is set to %d\n", CblasRowMajor,
                 CblasColMajor, Order);
  }
  if (M < 0) {
    ierr = -2;
    cblas_xerbla(2, rout, "M cannot be less than zero 0,; is set to %d.\n", M);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, rout, "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (Order == CblasColMajor) {
    if (lda < M || lda < 1) {
      ierr = -6;
      cblas_xerbla(6, rout, "lda must be >= MAX(M,1): lda=%d M=%d\n", lda, M);
    }
  } else if (lda < N || lda < 1) {
    ierr = -6;
    cblas_xerbla(6, rout, "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  return ierr;
}

int atlmx_getrs_check(const char *rout, int Order, int Trans, int N, int NRHS, int lda,
                      int ldb)
{
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, rout, "Order must be %d or %d, but is set to %d\n", CblasRowMajor,
                 CblasColMajor, Order);
  }
  if (Trans != CblasNoTrans && Trans != CblasTrans && Trans != CblasConjTrans) {
    ierr = -2;
    cblas_xerbla(2, rout, "Trans must be %d, %d, or %d, but is set to %d\n", CblasNoTrans,
                 CblasTrans, CblasConjTrans, Trans);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, rout, "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (NRHS < 0) {
    ierr = -4;
    cblas_xerbla(4, rout, "NRHS cannot be less than zero 0,; is set to %d.\n", NRHS);
  }
  if (lda < N || lda < 1) {
    ierr = -6;
    cblas_xerbla(6, rout, "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (ldb < N || ldb < 1) {
    ierr = -9;
    cblas_xerbla(9, rout, "ldb must be >= MAX(N,1): ldb=%d N=%d\n", ldb, N);
  }
  return ierr;
}

int atlmx_gesv_check(const char *rout, int Order, int N, int NRHS, int lda, int ldb)
{
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, rout, "Order must be %d or %d, but is set to %d\n", CblasRowMajor,
                 CblasColMajor, Order);
  }
  if (N < 0) {
    ierr = -2;
    cblas_xerbla(2, rout, "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (NRHS < 0) {
    ierr = -3;
    cblas_xerbla(3, rout, "NRHS cannot be less than zero 0,; is set to %d.\n", NRHS);
  }
  if (lda < N || lda < 1) {
    ierr = -5;
    cblas_xerbla(5, rout, "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (ldb < N || ldb < 1) {
    ierr = -8;
    cblas_xerbla(8, rout, "ldb must be >= MAX(N,1): ldb=%d N=%d\n", ldb, N);
  }
  return ierr;
}

/* clapack_?potrf (interfaces/lapack/C/src/clapack_dpotrf.c) */
int atlmx_potrf_check(const char *rout, int Order, int Uplo, int N, int lda)
{
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, rout, "Order must be %d or %d, but is set to %d\n", CblasRowMajor,
                 CblasColMajor, Order);
  }
  if (Uplo != CblasUpper && Uplo != CblasLower) {
    ierr = -2;
    cblas_xerbla(2, rout, "Uplo must be %d or %d, but is set to %d\n", CblasUpper,
                 CblasLower, Uplo);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, rout, "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (lda < N || lda < 1) {
    ierr = -5;
    cblas_xerbla(5, rout, "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  return ierr;
}

/* clapack_?potrs and clapack_?posv, whose checks are the same
 * (clapack_dpotrs.c, clapack_dposv.c:19-58) */
int atlmx_posv_check(const char *rout, int Order, int Uplo, int N, int NRHS, int lda, int ldb)
{
  int ierr = 0;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, rout, "Order must be %d or %d, but is set to %d\n", CblasRowMajor,
                 CblasColMajor, Order);
  }
  if (Uplo != CblasUpper && Uplo != CblasLower) {
    ierr = -2;
    cblas_xerbla(2, rout, "Uplo must be %d or %d, but is set to %d\n", CblasUpper,
                 CblasLower, Uplo);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, rout, "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (NRHS < 0) {
    ierr = -4;
    cblas_xerbla(4, rout, "NRHS cannot be less than zero 0,; is set to %d.\n", NRHS);
  }
  if (lda < N || lda < 1) {
    ierr = -6;
    cblas_xerbla(6, rout, "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (ldb < N || ldb < 1) {
    ierr = -8;
    cblas_xerbla(8, rout, "ldb must be >= MAX(N,1): ldb=%d N=%d\n", ldb, N);
  }
  return ierr;
}

#define T double
#define PFX d
#define K_PANEL atlmx_k_dgetf2_panel
#define K_LASWP atlmx_k_dlaswp
#define K_PERMUTE_ROWS atlmx_k_dpermute_rows
#define K_LASWP_APPLY atlmx_k_dlaswp_apply
#define K_POTF2 atlmx_k_dpotf2
#define K_TRI_INV atlmx_k_dtri_inv
#define K_CHOL_PANEL atlmx_k_dchol_panel
#define K_TRICOPY atlmx_k_dtricopy
#define K_DIAG_ZERO atlmx_k_ddiag_zero
#define K_GEQR2 atlmx_k_dgeqr2_panel
#define K_QR_VCOPY atlmx_k_dqr_vcopy
#define K_LARFT atlmx_k_dlarft
#include "lapack_tmpl.h"
#undef T
#undef PFX
#undef K_PANEL
#undef K_LASWP
#undef K_PERMUTE_ROWS
#undef K_LASWP_APPLY
#undef K_POTF2
#undef K_TRI_INV
#undef K_CHOL_PANEL
#undef K_TRICOPY
#undef K_DIAG_ZERO
#undef K_GEQR2
#undef K_QR_VCOPY
#undef K_LARFT

#define T float
#define PFX s
#define K_PANEL atlmx_k_sgetf2_panel
#define K_LASWP atlmx_k_slaswp
#define K_PERMUTE_ROWS atlmx_k_spermute_rows
#define K_LASWP_APPLY atlmx_k_slaswp_apply
#define K_POTF2 atlmx_k_spotf2
#define K_TRI_INV atlmx_k_stri_inv
#define K_CHOL_PANEL atlmx_k_schol_panel
#define K_TRICOPY atlmx_k_stricopy
#define K_DIAG_ZERO atlmx_k_sdiag_zero
#define K_GEQR2 atlmx_k_sgeqr2_panel
#define K_QR_VCOPY atlmx_k_sqr_vcopy
#define K_LARFT atlmx_k_slarft
#include "lapack_tmpl.h"

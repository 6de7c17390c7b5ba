/*
 * cblas3c.c -- complex Level-3 BLAS and LU (C host code) of libatlas_mi355x:
 * cblas_{c,z}{gemm,trsm,trmm,herk,syrk,her2k,syr2k,hemm,symm},
 * clapack_{c,z}{getrf,getrs,gesv,potrf,potrs,posv} and their
 * device-pointer twins, instantiated from cblas3c_tmpl.h and lapackc_tmpl.h
 * per precision over the real MFMA GEMM kernels.
 */
#include <stdio.h>
#include <stdlib.h>
#include "rt.h"
#include "drivers.h"
#include "../../include/cblas.h"
#include "../../include/clapack.h"
#include "../../include/atlmx.h"

#define STR_(x) #x
#define STR(x) STR_(x)

/* Argument checks of cblas_?herk / ?her2k (herm: Trans N or C) and
 * cblas_?syrk / ?syr2k (Trans N or T) for c and z, in ATLAS's order,
 * positions and messages (interfaces/blas/C/src/cblas_zherk.c:20-75,
 * cblas_zsyrk.c, cblas_zher2k.c, cblas_zsyr2k.c; two_k adds ldb at
 * position 10 and moves ldc to 13).  2000 = valid. */
static int crk_check(int Order, int Uplo, int Trans, int N, int K, int lda, int ldb, int ldc,
                     int herm, int two_k)
{
  int info = 2000;
  if (Uplo != CblasUpper && Uplo != CblasLower)
    info = cblas_errprn(2, info, "UPLO must be %d or %d, but is set to %d", CblasUpper,
                        CblasLower, Uplo);
  if (N < 0) info = cblas_errprn(4, info, "N cannot be less than zero; it is set to %d.", N);
  if (K < 0) info = cblas_errprn(5, info, "K cannot be less than zero; it is set to %d.", K);
  if (Order == CblasColMajor || Order == CblasRowMajor) {
    const int nt = (Order == CblasColMajor) == (Trans == CblasNoTrans);
    if (Trans != CblasNoTrans && Trans != (herm ? CblasConjTrans : CblasTrans)) {
      if (herm)
        info = cblas_errprn(3, info, "Trans must be %d or %d, but is set to %d", CblasNoTrans,
                            CblasConjTrans, Trans);
      else
        info = cblas_errprn(3, info, "Trans must be %d, %d or %d, but is set to %d",
                            CblasNoTrans, CblasTrans, CblasConjTrans, Trans);
    }
    if (nt) {
      if (lda < N || lda < 1)
        info = cblas_errprn(8, info, "lda must be >= MAX(N,1): lda=%d N=%d", lda, N);
      if (two_k && (ldb < N || ldb < 1))
        info = cblas_errprn(10, info, "ldb must be >= MAX(N,1): ldb=%d N=%d", ldb, N);
    } else {
      if (lda < K || lda < 1)
        info = cblas_errprn(8, info, "lda must be >= MAX(K,1): lda=%d K=%d", lda, K);
      if (two_k && (ldb < K || ldb < 1))
        info = cblas_errprn(10, info, "ldb must be >= MAX(K,1): ldb=%d K=%d", ldb, K);
    }
  } else
    info = cblas_errprn(1, info, "Order must be %d or %d, but is set to %d", CblasRowMajor,
                        CblasColMajor, Order);
  if (ldc < N || ldc < 1)
    info = cblas_errprn(two_k ? 13 : 11, info, "ldc must be >= MAX(N,1): ldc=%d N=%d", ldc, N);
  return info;
}

#define T double
#define PFX z
#define K_GEMM atlmx_k_dgemm
#define K_CGATHER atlmx_k_zgather
#define K_CSCAL atlmx_k_zscal2d
#define K_CTRI_INV atlmx_k_ztri_inv
#define K_CGETF2 atlmx_k_zgetf2
#define K_RLASWP atlmx_k_dlaswp
#define K_RLASWP_APPLY atlmx_k_dlaswp_apply
#define K_CTRSM_LLNU atlmx_k_ztrsm_llnu
#define K_CPOTF2 atlmx_k_zpotf2
#define K_CCONJ atlmx_k_zconj
#define K_RTRSV atlmx_k_dtrsv
#define K_RTRSM_TAG atlmx_k_dtrsm_tag
#define K_CTRI_SCALE atlmx_k_ztri_scale
#define K_CSYM_EXPAND atlmx_k_zsym_expand
#include "cblas3c_tmpl.h"
#include "lapackc_tmpl.h"
#undef T
#undef PFX
#undef K_GEMM
#undef K_CGATHER
#undef K_CSCAL
#undef K_CTRI_INV
#undef K_CGETF2
#undef K_RLASWP
#undef K_RLASWP_APPLY
#undef K_CTRSM_LLNU
#undef K_CPOTF2
#undef K_CCONJ
#undef K_RTRSV
#undef K_RTRSM_TAG
#undef K_CTRI_SCALE
#undef K_CSYM_EXPAND

#define T float
#define PFX c
#define K_GEMM atlmx_k_sgemm
#define K_CGATHER atlmx_k_cgather
#define K_CSCAL atlmx_k_cscal2d
#define K_CTRI_INV atlmx_k_ctri_inv
#define K_CGETF2 atlmx_k_cgetf2
#define K_RLASWP atlmx_k_slaswp
#define K_RLASWP_APPLY atlmx_k_slaswp_apply
#define K_CTRSM_LLNU atlmx_k_ctrsm_llnu
#define K_CPOTF2 atlmx_k_cpotf2
#define K_CCONJ atlmx_k_cconj
#define K_RTRSV atlmx_k_strsv
#define K_RTRSM_TAG atlmx_k_strsm_tag
#define K_CTRI_SCALE atlmx_k_ctri_scale
#define K_CSYM_EXPAND atlmx_k_csym_expand
#include "cblas3c_tmpl.h"
#include "lapackc_tmpl.h"

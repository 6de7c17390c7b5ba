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

/* drivers defined in blas3.c */
void dgemm_dev(atlmx_ctx *c, int ta, int tb, int M, int N, int K, double alpha,
               const double *A, long lda, const double *B, long ldb, double beta, double *C,
               long ldc, int tri, const int *abort_flag, const char *rout);
void sgemm_dev(atlmx_ctx *c, int ta, int tb, int M, int N, int K, float alpha,
               const float *A, long lda, const float *B, long ldb, float beta, float *C,
               long ldc, int tri, const int *abort_flag, const char *rout);
void dtrsm_dev(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N,
               double alpha, const double *A, long lda, double *B, long ldb,
               const int *abort_flag, const char *rout);
void strsm_dev(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N,
               float alpha, const float *A, long lda, float *B, long ldb,
               const int *abort_flag, const char *rout);
void dsyrk_dev(atlmx_ctx *c, int uplo, int trans, int N, int K, double alpha,
               const double *A, long lda, double beta, double *C, long ldc,
               const int *abort_flag, const char *rout);
void ssyrk_dev(atlmx_ctx *c, int uplo, int trans, int N, int K, float alpha,
               const float *A, long lda, float beta, float *C, long ldc,
               const int *abort_flag, const char *rout);

#define T double
#define PFX d
#define K_PANEL atlmx_k_dgetf2_panel
#define K_LASWP atlmx_k_dlaswp
#define K_LASWP_APPLY atlmx_k_dlaswp_apply
#define K_POTF2 atlmx_k_dpotf2
#include "lapack_tmpl.h"
#undef T
#undef PFX
#undef K_PANEL
#undef K_LASWP
#undef K_LASWP_APPLY
#undef K_POTF2

#define T float
#define PFX s
#define K_PANEL atlmx_k_sgetf2_panel
#define K_LASWP atlmx_k_slaswp
#define K_LASWP_APPLY atlmx_k_slaswp_apply
#define K_POTF2 atlmx_k_spotf2
#include "lapack_tmpl.h"

/*
 * blas3.c -- Level-3 BLAS layer (C host code) of libatlas_mi355x:
 * cblas_{s,d}{gemm,trsm,syrk}, cblas_i{s,d}amax and their device-pointer
 * twins, instantiated from blas3_tmpl.h per precision.
 */
#include <stdio.h>
#include <stdlib.h>
#include "rt.h"
#include "../../include/cblas.h"
#include "../../include/atlmx.h"

#define STR_(x) #x
#define STR(x) STR_(x)

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

/*
 * cblas3c.c -- complex Level-3 BLAS and LU (C host code) of libatlas_mi355x:
 * cblas_{c,z}{gemm,trsm}, clapack_{c,z}{getrf,getrs,gesv,potrf,potrs,posv} and their
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
#include "cblas3c_tmpl.h"
#include "lapackc_tmpl.h"

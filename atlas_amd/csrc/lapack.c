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

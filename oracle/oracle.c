/*
 * ORACLE -- test infrastructure only (see oracle.h).  Not linked into, and
 * never called by, the MI355X library.
 */
#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* ======================================================================= */
/* Portable 31-bit LCG: src/testing/ATL_rand.c                              */
/* ======================================================================= */
/* A 31-bit number X is held as two ints: X[0] = low 16 bits, X[1] = high
 * 15 bits.  ATLAS does the arithmetic in C int, relying on two's-complement
 * wrap of intermediate products (signed overflow); the wraps are spelled out
 * here with uint32_t so the restatement is defined behaviour. */
static int32_t wrap32(uint32_t x) { return (int32_t)x; }

/* ATL_rand.c:41-54 (ATL_ladd): I = J + K modulo 2^31, no carry out. */
static void lcg_add(const int *J, const int *K, int *I)
{
   int t0 = K[0] + J[0], t1;
   t1 = t0 >> 16;
   I[0] = t0 - (t1 << 16);
   t0 = t1 + K[1] + J[1];
   I[1] = t0 - ((t0 >> 15) << 15);
}

/* ATL_rand.c:84-106 (ATL_lmul): I = K * J modulo 2^31.  A negative
 * intermediate has 2^31 added (ATLAS adds 2^30 twice). */
static void lcg_mul(const int *K, const int *J, int *I)
{
   const int32_t p30 = (int32_t)1 << 30;
   int32_t kt, lt;
   kt = wrap32((uint32_t)K[0] * (uint32_t)J[0]);
   if (kt < 0) kt = wrap32((uint32_t)(kt + p30) + (uint32_t)p30);
   I[0] = kt - ((kt >> 16) << 16);
   lt = wrap32((uint32_t)K[0] * (uint32_t)J[1] + (uint32_t)K[1] * (uint32_t)J[0]);
   if (lt < 0) lt = wrap32((uint32_t)(lt + p30) + (uint32_t)p30);
   kt = wrap32((uint32_t)(kt >> 16) + (uint32_t)lt);
   if (kt < 0) kt = wrap32((uint32_t)(kt + p30) + (uint32_t)p30);
   I[1] = kt - ((kt >> 15) << 15);
}

/* ATL_rand.c:248-257 (ATL_srand): jseed = {iseed, iseed>>16} -- note the low
 * word is NOT masked to 16 bits (:254) -- then a jump of one step with
 * a = {20077, 16838} (= 1103515245) and c = {12345, 0} gives X(1); the
 * generator state becomes X(1), a, c (ATL_xjumpm with JUMPM=1, :194-244). */
void orc_lcg_seed(orc_lcg_t *g, int iseed)
{
   const int mult[2] = {20077, 16838}, iadd[2] = {12345, 0};
   int jseed[2], j[2];
   jseed[0] = iseed;
   jseed[1] = iseed >> 16;
   lcg_mul(jseed, mult, j);              /* IRANM = IRANN*IAM + ICM */
   lcg_add(j, iadd, g->ir);
   g->ia[0] = mult[0]; g->ia[1] = mult[1];
   g->ic[0] = iadd[0]; g->ic[1] = iadd[1];
}

/* ATL_rand.c:259-265 + ATL_setran OPTION 3 (:143-148): return the current
 * value j[0] + (j[1] << 16), then advance X = a*X + c. */
int orc_lcg_next(orc_lcg_t *g)
{
   int j[2], v;
   v = g->ir[0] + (g->ir[1] << 16);
   lcg_mul(g->ir, g->ia, j);
   lcg_add(j, g->ic, g->ir);
   return v;
}

static orc_lcg_t orc_g_;
void orc_srand(int iseed) { orc_lcg_seed(&orc_g_, iseed); }
int orc_rand(void) { return orc_lcg_next(&orc_g_); }
/* include/atlas_tst.h:18 */
double orc_dumb_rand(void) { return 0.5 - ((double)orc_rand()) / 2147483648.0; }

void orc_rand_stream(int iseed, int n, int *out)
{
   orc_lcg_t g;
   int i;
   orc_lcg_seed(&g, iseed);
   for (i = 0; i < n; i++) out[i] = orc_lcg_next(&g);
}

/* ======================================================================= */
/* Knobs                                                                    */
/* ======================================================================= */
#define ORC_MAXTHR 256
static int orc_l1c_elts_ = 4096, orc_lcm_ = 4, orc_nthr_ = 1;
static double *orc_gaps_ = NULL;
static int orc_ngaps_ = 0;

void orc_set_recursion_params(int l1c_elts, int lcm)
{
   orc_l1c_elts_ = l1c_elts;
   orc_lcm_ = lcm;
}
void orc_set_threads(int n)
{
   orc_nthr_ = n < 1 ? 1 : (n > ORC_MAXTHR ? ORC_MAXTHR : n);
}
int orc_get_threads(void) { return orc_nthr_; }
void orc_track_pivot_gaps(double *gaps, int n)
{
   orc_gaps_ = gaps;
   orc_ngaps_ = gaps ? n : 0;
}

#define ORC_FILLCONST (-2560000000.0)   /* src/testing/ATL_gegen.c:6 */
#define CAT2_(a, b) a##b
#define CAT_(a, b) CAT2_(a, b)

/* ---- double ---- */
#define T double
#define P orc_d
#define ORC_ABS(x) fabs(x)
#define ORC_SQRT(x) sqrt(x)
#define ORC_SAFMIN DBL_MIN          /* LAPACK dlamch('S') */
#define ORC_GEMMTST_EPS 1.0e-15     /* include/atlas_misc.h:122 */
#include "orc_tmpl.h"
#undef T
#undef P
#undef ORC_ABS
#undef ORC_SQRT
#undef ORC_SAFMIN
#undef ORC_GEMMTST_EPS

/* ---- float ---- */
#define T float
#define P orc_s
#define ORC_ABS(x) fabsf(x)
#define ORC_SQRT(x) sqrtf(x)
#define ORC_SAFMIN FLT_MIN          /* LAPACK slamch('S') */
#define ORC_GEMMTST_EPS 5.0e-7      /* include/atlas_misc.h:107 */
#include "orc_tmpl.h"
#undef T
#undef P
#undef ORC_ABS
#undef ORC_SQRT
#undef ORC_SAFMIN
#undef ORC_GEMMTST_EPS

/* ---- complex double / complex float ---- */
#define T double
#define P orc_z
#define ORC_ABS(x) fabs(x)
#define ORC_SAFMIN DBL_MIN
#define ORC_SQRT(x) sqrt(x)
#include "orc_cplx.h"
#undef T
#undef P
#undef ORC_ABS
#undef ORC_SAFMIN
#undef ORC_SQRT
#define T float
#define P orc_c
#define ORC_ABS(x) fabsf(x)
#define ORC_SAFMIN FLT_MIN
#define ORC_SQRT(x) sqrtf(x)
#include "orc_cplx.h"

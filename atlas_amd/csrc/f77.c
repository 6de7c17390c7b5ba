/*
 * f77.c -- the Fortran 77 ABI of libatlas_mi355x (no Fortran compiler
 * involved): dgemm_, dgetrf_ and the rest with the gfortran calling
 * convention -- every argument by reference, a hidden string length per
 * CHARACTER argument after the others, 1-based pivots.
 *
 * Replaces ATLAS's interfaces/blas/F77/src/?*.f and
 * interfaces/lapack/F77/src/?*.f, which check their arguments (same order,
 * same positions, XERBLA on the first bad one) and call the C routines
 * through the f77wrap directories.  Here each entry point does the
 * same checks and calls this library's cblas_ / clapack_ routine; pivots
 * are shifted to 1-based on the way out
 * (interfaces/lapack/F77/src/f77wrap/ATL_f77wrap_getrf.c:27) and back to
 * 0-based on the way in.
 */
#include <ctype.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include "../../include/cblas.h"
#include "../../include/clapack.h"
#include "../../include/atlas_f77.h"

/* interfaces/blas/F77/src/xerbla.f: message, then STOP (exit status 0).
 * Weak, so a program may supply its own. */
__attribute__((weak)) void xerbla_(const char *srname, const int *info, size_t len)
{
  int n = (int)(len > 6 ? 6 : len);
  fprintf(stdout, " ** On entry to %-6.*s parameter number %2d had an illegal value\n", n,
          srname, *info);
  fflush(stdout);
  exit(0);
}

static void f77_err(const char *name, int info)
{
  xerbla_(name, &info, 6);
}

/* LSAME: first character, case-insensitive */
static int lsame(const char *c, char want) { return toupper((unsigned char)*c) == want; }

static int tr_code(const char *t)
{
  if (lsame(t, 'N')) return CblasNoTrans;
  if (lsame(t, 'T')) return CblasTrans;
  if (lsame(t, 'C')) return CblasConjTrans;
  return 0;
}

static int imax1(int a) { return a > 1 ? a : 1; }

#define T double
#define P d
#define PU "D"
#include "f77_tmpl.h"
#undef T
#undef P
#undef PU

#define T float
#define P s
#define PU "S"
#include "f77_tmpl.h"
#undef T
#undef P
#undef PU

/* complex: zgemm_, ztrsm_, zgetrf_ ... (interfaces/blas/F77/src/zgemm.f,
 * ztrsm.f, interfaces/lapack/F77/src/zgetrf.f, ... : the same checks) */
#define F77_COMPLEX
#define T double
#define P z
#define PU "Z"
#include "f77_tmpl.h"
#undef T
#undef P
#undef PU

#define T float
#define P c
#define PU "C"
#include "f77_tmpl.h"

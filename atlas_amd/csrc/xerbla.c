/*
 * xerbla.c -- error reporting of the C BLAS boundary.
 *
 * interfaces/blas/C/src/cblas_xerbla.c:13-28: print "Parameter %d to
 * routine %s was incorrect" (when p != 0) and the formatted message to
 * stderr, then exit(-1).  Weak, so a program may supply its own.
 * interfaces/blas/C/src/cblas_errprn.c:13-25: print the message and return
 * min(ierr, info), so the lowest failing parameter position is reported.
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include "../../include/cblas.h"

__attribute__((weak)) void cblas_xerbla(int p, const char *rout, const char *form, ...)
{
  va_list ap;
  va_start(ap, form);
  if (p) fprintf(stderr, "Parameter %d to routine %s was incorrect\n", p, rout);
  vfprintf(stderr, form, ap);
  va_end(ap);
  exit(-1);
}

int cblas_errprn(int ierr, int info, char *form, ...)
{
  va_list ap;
  va_start(ap, form);
  vfprintf(stderr, form, ap);
  va_end(ap);
  return ierr < info ? ierr : info;
}

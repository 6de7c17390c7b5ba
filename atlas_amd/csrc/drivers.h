/*
 * drivers.h -- stream-ordered drivers on device pointers shared between the
 * host C files (defined in blas3.c and lapack.c from their precision
 * templates).  Column-major, 64-bit leading dimensions, work ordered on the
 * context's stream.
 */
#ifndef ATLMX_DRIVERS_H
#define ATLMX_DRIVERS_H
#include "rt.h"

#define ATLMX_DRV_DECLARE(P, T)                                                      \
void P##gemm_dev(atlmx_ctx *c, int ta, int tb, int M, int N, int K, T alpha,          \
                 const T *A, long lda, const T *B, long ldb, T beta, T *C, long ldc,  \
                 int tri, const int *abort_flag, const char *rout);                   \
void P##update_dev(atlmx_ctx *c, int tb, int M, int N, int K, T alpha, const T *A,     \
                   long lda, const T *B, long ldb, T *C, long ldc, int tri,           \
                   const int *abort_flag, const char *rout);                          \
void P##trsm_dev(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N,    \
                 T alpha, const T *A, long lda, T *B, long ldb,                       \
                 const int *abort_flag, const char *rout);                            \
void P##trsm_dev_ws(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N, \
                    T alpha, const T *A, long lda, T *B, long ldb,                    \
                    const int *abort_flag, T *inv_ws, int *ctl_ws, const char *rout); \
void P##syrk_dev(atlmx_ctx *c, int uplo, int trans, int N, int K, T alpha,            \
                 const T *A, long lda, T beta, T *C, long ldc,                        \
                 const int *abort_flag, const char *rout);                            \
int  P##getrf_dev(atlmx_ctx *c, int M, int N, T *A, long lda, int *d_ipiv,            \
                  const char *rout);                                                  \
void P##getrs_dev(atlmx_ctx *c, int trans, int N, int NRHS, const T *A, long lda,     \
                  const int *d_ipiv, T *B, long ldb, const char *rout);               \
void P##getrs_dev_ws(atlmx_ctx *c, int trans, int N, int NRHS, const T *A, long lda,  \
                     const int *d_ipiv, T *B, long ldb, T *inv, int *ctl,             \
                     const char *rout);                                               \
int  P##potrf_dev(atlmx_ctx *c, int uplo, int N, T *A, long lda, const char *rout);   \
void P##potrs_dev(atlmx_ctx *c, int uplo, int N, int NRHS, const T *A, long lda,      \
                  T *B, long ldb, const char *rout);                                  \
void P##trmm_dev(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N,    \
                 T alpha, const T *A, long lda, T *B, long ldb, const char *rout);    \
void P##symm_dev(atlmx_ctx *c, int side, int uplo, int M, int N, T alpha,             \
                 const T *A, long lda, const T *B, long ldb, T beta, T *C, long ldc,  \
                 const char *rout);                                                   \
void P##syr2k_dev(atlmx_ctx *c, int uplo, int trans, int N, int K, T alpha,           \
                  const T *A, long lda, const T *B, long ldb, T beta, T *C, long ldc, \
                  const char *rout);                                                  \
int  P##trtri_dev(atlmx_ctx *c, int uplo, int diag, int N, T *A, long lda,            \
                  const char *rout);                                                  \
void P##lauum_dev(atlmx_ctx *c, int uplo, int N, T *A, long lda, const char *rout);   \
int  P##getri_dev(atlmx_ctx *c, int N, T *A, long lda, const int *d_ipiv,             \
                  const char *rout);                                                  \
void P##geqrf_dev(atlmx_ctx *c, int M, int N, T *A, long lda, T *d_tau,               \
                  const char *rout);

ATLMX_DRV_DECLARE(d, double)
ATLMX_DRV_DECLARE(s, float)

/* cblas argument checks shared by the real and complex routines (defined in
 * blas3.c); each returns 2000 when the arguments are valid, else the
 * position xerbla reports, after printing ATLAS's message. */
int atlmx_gemm_check(int Order, int TA, int TB, int M, int N, int K, int lda, int ldb,
                     int ldc);
int atlmx_trxm_check(int Order, int Side, int Uplo, int TA, int Diag, int M, int N, int lda,
                     int ldb);
int atlmx_symm_check(int Order, int Side, int Uplo, int M, int N, int lda, int ldb, int ldc);
/* clapack argument checks shared by the real and complex routines (defined
 * in lapack.c): 0, or the negated position of the last bad argument */
int atlmx_getrf_check(const char *rout, int Order, int M, int N, int lda);
int atlmx_getrs_check(const char *rout, int Order, int Trans, int N, int NRHS, int lda,
                      int ldb);
int atlmx_gesv_check(const char *rout, int Order, int N, int NRHS, int lda, int ldb);
int atlmx_potrf_check(const char *rout, int Order, int Uplo, int N, int lda);
int atlmx_posv_check(const char *rout, int Order, int Uplo, int N, int NRHS, int lda, int ldb);

/* complex drivers (defined in cblas3c.c); alpha and beta point at
 * {real, imaginary}; matrices interleaved, leading dimensions in complex
 * elements */
#define ATLMX_CDRV_DECLARE(P, T)                                                      \
void P##gemm_dev(atlmx_ctx *c, int ta, int tb, int M, int N, int K, const T *alpha,   \
                 const T *A, long lda, const T *B, long ldb, const T *beta, T *C,     \
                 long ldc, int tri, const char *rout);                                \
void P##trsm_dev(atlmx_ctx *c, int side, int uplo, int ta, int diag, int M, int N,    \
                 const T *alpha, const T *A, long lda, T *B, long ldb,                \
                 const char *rout);                                                   \
int  P##getrf_dev(atlmx_ctx *c, int M, int N, T *A, long lda, int *d_ipiv,            \
                  const char *rout);                                                  \
void P##getrs_dev(atlmx_ctx *c, int trans, int N, int NRHS, const T *A, long lda,     \
                  const int *d_ipiv, T *B, long ldb, const char *rout);               \
int  P##potrf_dev(atlmx_ctx *c, int uplo, int N, T *A, long lda, const char *rout);   \
void P##potrs_dev(atlmx_ctx *c, int uplo, int N, int NRHS, const T *A, long lda,      \
                  T *B, long ldb, const char *rout);
ATLMX_CDRV_DECLARE(z, double)
ATLMX_CDRV_DECLARE(c, float)

#endif

/*
 * cblas.h -- C BLAS interface of the MI355X Level-3 library (libatlas_mi355x).
 *
 * Drop-in for the subset of ATLAS's include/cblas.h that the Level-3 hot path
 * needs.  Enum values, argument order and meaning, CBLAS_INDEX = int, and the
 * error behaviour (cblas_xerbla prints "Parameter %d to routine %s was
 * incorrect" and calls exit(-1)) are ATLAS's.  Every routine accepts host
 * pointers (data is staged through the GPU) or device pointers (hipMalloc'd;
 * no staging), and returns when the result is in place.
 */
#ifndef CBLAS_H
#define CBLAS_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef CBLAS_ENUM_DEFINED_H
#define CBLAS_ENUM_DEFINED_H
/* ATLAS include/cblas.h:5-10 */
enum CBLAS_ORDER {CblasRowMajor=101, CblasColMajor=102};
enum CBLAS_TRANSPOSE {CblasNoTrans=111, CblasTrans=112, CblasConjTrans=113,
                      AtlasConj=114};
enum CBLAS_UPLO {CblasUpper=121, CblasLower=122};
enum CBLAS_DIAG {CblasNonUnit=131, CblasUnit=132};
enum CBLAS_SIDE {CblasLeft=141, CblasRight=142};
#endif

#define CBLAS_INDEX int   /* ATLAS include/cblas.h:15 */

/* ATLAS include/cblas.h:17-18; interfaces/blas/C/src/cblas_errprn.c:13-25 and
 * cblas_xerbla.c:13-28.  cblas_xerbla is a weak symbol: a program may
 * supply its own. */
int  cblas_errprn(int ierr, int info, char *form, ...);
void cblas_xerbla(int p, const char *rout, const char *form, ...);

/* Replaces ATLAS include/cblas.h:66-67 (cblas_i?amax), used by the LU panel.
 * First index of max |x_i|; NaN never wins (src/blas/reference/level1/
 * ATL_idrefamax.c:112). */
CBLAS_INDEX cblas_isamax(const int N, const float *X, const int incX);
CBLAS_INDEX cblas_idamax(const int N, const double *X, const int incX);

/* Replaces ATLAS include/cblas.h:441 / :471 and
 * interfaces/blas/C/src/cblas_dgemm.c:16-134. */
void cblas_sgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_TRANSPOSE TransB, const int M, const int N,
                 const int K, const float alpha, const float *A, const int lda,
                 const float *B, const int ldb, const float beta, float *C,
                 const int ldc);
void cblas_dgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_TRANSPOSE TransB, const int M, const int N,
                 const int K, const double alpha, const double *A, const int lda,
                 const double *B, const int ldb, const double beta, double *C,
                 const int ldc);

/* Replaces ATLAS include/cblas.h:451 / :481 (cblas_?syrk); the Cholesky
 * trailing update of src/lapack/ATL_potrfL.c:235. */
void cblas_ssyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                 const float alpha, const float *A, const int lda,
                 const float beta, float *C, const int ldc);
void cblas_dsyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                 const double alpha, const double *A, const int lda,
                 const double beta, double *C, const int ldc);

/* Replaces ATLAS include/cblas.h:465 / :495 and
 * interfaces/blas/C/src/cblas_dtrsm.c:13-106. */
void cblas_strsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N,
                 const float alpha, const float *A, const int lda, float *B,
                 const int ldb);
void cblas_dtrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N,
                 const double alpha, const double *A, const int lda, double *B,
                 const int ldb);

/* Replaces ATLAS include/cblas.h:460 / :490 (cblas_?trmm,
 * interfaces/blas/C/src/cblas_dtrmm.c -> src/blas/ulevel3/ATL_trmm.c:149). */
void cblas_strmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N,
                 const float alpha, const float *A, const int lda, float *B,
                 const int ldb);
void cblas_dtrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N,
                 const double alpha, const double *A, const int lda, double *B,
                 const int ldb);

/* Replaces ATLAS include/cblas.h:446 / :476 (cblas_?symm,
 * interfaces/blas/C/src/cblas_dsymm.c -> src/blas/ulevel3/ATL_symmL.c). */
void cblas_ssymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N,
                 const float alpha, const float *A, const int lda, const float *B,
                 const int ldb, const float beta, float *C, const int ldc);
void cblas_dsymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N,
                 const double alpha, const double *A, const int lda, const double *B,
                 const int ldb, const double beta, double *C, const int ldc);

/* Replaces ATLAS include/cblas.h:455 / :485 (cblas_?syr2k,
 * interfaces/blas/C/src/cblas_dsyr2k.c -> src/blas/ulevel3/ATL_syr2k.c:255). */
void cblas_ssyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                  const float alpha, const float *A, const int lda, const float *B,
                  const int ldb, const float beta, float *C, const int ldc);
void cblas_dsyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K,
                  const double alpha, const double *A, const int lda, const double *B,
                  const int ldb, const double beta, double *C, const int ldc);

/* ---- complex (c: float, z: double; interleaved real/imaginary parts,
 * alpha and beta point at {real, imaginary}; TransA/TransB may be
 * CblasConjTrans) ---- */
/* Replaces ATLAS include/cblas.h:501 / :531 and
 * interfaces/blas/C/src/cblas_zgemm.c:16-136 (ATL_zgemm). */
void cblas_cgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_TRANSPOSE TransB, const int M, const int N,
                 const int K, const void *alpha, const void *A, const int lda,
                 const void *B, const int ldb, const void *beta, void *C, const int ldc);
void cblas_zgemm(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_TRANSPOSE TransB, const int M, const int N,
                 const int K, const void *alpha, const void *A, const int lda,
                 const void *B, const int ldb, const void *beta, void *C, const int ldc);
/* Replaces ATLAS include/cblas.h:525 / :555 and
 * interfaces/blas/C/src/cblas_ztrsm.c:13-106 (ATL_ztrsm). */
void cblas_ctrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                 const void *A, const int lda, void *B, const int ldb);
void cblas_ztrsm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                 const void *A, const int lda, void *B, const int ldb);
/* The rest of the complex Level-3 BLAS, ATLAS include/cblas.h:506-520 and
 * :536-550 (symm, syrk, syr2k, trmm), :565-592 (hemm, herk, her2k):
 * interfaces/blas/C/src/cblas_zherk.c (alpha, beta real),
 * cblas_zher2k.c (beta real), cblas_zsyrk.c, cblas_zsyr2k.c,
 * cblas_zhemm.c, cblas_zsymm.c, cblas_ztrmm.c. */
void cblas_cherk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const float alpha,
                 const void *A, const int lda, const float beta, void *C, const int ldc);
void cblas_zherk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const double alpha,
                 const void *A, const int lda, const double beta, void *C, const int ldc);
void cblas_cher2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                  const void *A, const int lda, const void *B, const int ldb, const float beta,
                  void *C, const int ldc);
void cblas_zher2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                  const void *A, const int lda, const void *B, const int ldb, const double beta,
                  void *C, const int ldc);
void cblas_csyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                 const void *A, const int lda, const void *beta, void *C, const int ldc);
void cblas_zsyrk(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                 const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                 const void *A, const int lda, const void *beta, void *C, const int ldc);
void cblas_csyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                  const void *A, const int lda, const void *B, const int ldb, const void *beta,
                  void *C, const int ldc);
void cblas_zsyr2k(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                  const enum CBLAS_TRANSPOSE Trans, const int N, const int K, const void *alpha,
                  const void *A, const int lda, const void *B, const int ldb, const void *beta,
                  void *C, const int ldc);
void cblas_chemm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                 const void *A, const int lda, const void *B, const int ldb, const void *beta,
                 void *C, const int ldc);
void cblas_zhemm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                 const void *A, const int lda, const void *B, const int ldb, const void *beta,
                 void *C, const int ldc);
void cblas_csymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                 const void *A, const int lda, const void *B, const int ldb, const void *beta,
                 void *C, const int ldc);
void cblas_zsymm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const int M, const int N, const void *alpha,
                 const void *A, const int lda, const void *B, const int ldb, const void *beta,
                 void *C, const int ldc);
void cblas_ctrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                 const void *A, const int lda, void *B, const int ldb);
void cblas_ztrmm(const enum CBLAS_ORDER Order, const enum CBLAS_SIDE Side,
                 const enum CBLAS_UPLO Uplo, const enum CBLAS_TRANSPOSE TransA,
                 const enum CBLAS_DIAG Diag, const int M, const int N, const void *alpha,
                 const void *A, const int lda, void *B, const int ldb);

#ifdef __cplusplus
}
#endif
#endif

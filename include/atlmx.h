/*
 * atlmx.h -- extensions of libatlas_mi355x beyond ATLAS's C ABI.
 *
 * None of these has a counterpart in ATLAS.  They serve callers whose data
 * already lives in GPU memory (HBM-resident pipelines, PyTorch, the
 * benchmarks): every atlmx_* routine takes DEVICE pointers, column-major
 * storage, ATLAS's enum values (include/cblas.h), and is ordered on the
 * calling thread's stream (atlmx_set_stream), returning before the work
 * finishes unless it says "synchronizes".
 */
#ifndef ATLMX_H
#define ATLMX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ------------------------------------------------------------ */
int   atlmx_init(void);                 /* optional; creates this thread's stream */
void  atlmx_set_stream(void *hip_stream); /* NULL: the legacy default stream (the default) */
void *atlmx_get_stream(void);
void  atlmx_synchronize(void);

/* ---- profiler: hipEvent timing of every launch of the enabled classes on
 * this thread's stream.  Classes: 0 GEMM (work = 2*M*N*K flops), 1 LU panel
 * (work = bytes loaded+stored), 2 TRSM diagonal inversion, 3 laswp,
 * 4 Cholesky diagonal block, 5 other, 6 few-right-hand-side triangular
 * solve (work = bytes of the factor read), 7 short-K trailing update of the
 * mixed solver's FP32 LU (work = flops).  mask bit c enables class c (0 = off,
 * -1 = all); enabling clears earlier records. ---- */
void atlmx_prof_enable(int mask);
void atlmx_prof_reset(void);
int  atlmx_prof_query(int cls, double *ms_total, long *launches, double *work_total);
/* ... plus summed algorithmic HBM bytes (each operand read once, each output
 * written once) for the classes whose launchers state them (0 and 7). */
int  atlmx_prof_query2(int cls, double *ms_total, long *launches, double *work_total,
                       double *bytes_total);

/* ---- BLAS-3 on device pointers ----------------------------------------- */
void atlmx_dgemm(int TA, int TB, int M, int N, int K, double alpha, const double *A,
                 int lda, const double *B, int ldb, double beta, double *C, int ldc);
void atlmx_sgemm(int TA, int TB, int M, int N, int K, float alpha, const float *A,
                 int lda, const float *B, int ldb, float beta, float *C, int ldc);
void atlmx_dtrsm(int Side, int Uplo, int TA, int Diag, int M, int N, double alpha,
                 const double *A, int lda, double *B, int ldb);
void atlmx_strsm(int Side, int Uplo, int TA, int Diag, int M, int N, float alpha,
                 const float *A, int lda, float *B, int ldb);
/* complex twins (interleaved storage, leading dimensions in complex
 * elements, alpha/beta point at host {real, imaginary}; TA/TB may be
 * ConjTrans = 113) */
void atlmx_zgemm(int TA, int TB, int M, int N, int K, const double *alpha, const double *A,
                 int lda, const double *B, int ldb, const double *beta, double *C, int ldc);
void atlmx_cgemm(int TA, int TB, int M, int N, int K, const float *alpha, const float *A,
                 int lda, const float *B, int ldb, const float *beta, float *C, int ldc);
void atlmx_ztrsm(int Side, int Uplo, int TA, int Diag, int M, int N, const double *alpha,
                 const double *A, int lda, double *B, int ldb);
void atlmx_ctrsm(int Side, int Uplo, int TA, int Diag, int M, int N, const float *alpha,
                 const float *A, int lda, float *B, int ldb);
void atlmx_dsyrk(int Uplo, int Trans, int N, int K, double alpha, const double *A,
                 int lda, double beta, double *C, int ldc);
void atlmx_ssyrk(int Uplo, int Trans, int N, int K, float alpha, const float *A, int lda,
                 float beta, float *C, int ldc);

/* ---- LAPACK on device pointers (ipiv: device int array, 0-based) -------- */
int  atlmx_dgetrf(int M, int N, double *A, int lda, int *d_ipiv);    /* synchronizes */
int  atlmx_sgetrf(int M, int N, float *A, int lda, int *d_ipiv);     /* synchronizes */
void atlmx_dgetrs(int Trans, int N, int NRHS, const double *A, int lda, const int *d_ipiv,
                  double *B, int ldb);
void atlmx_sgetrs(int Trans, int N, int NRHS, const float *A, int lda, const int *d_ipiv,
                  float *B, int ldb);
int  atlmx_dpotrf(int Uplo, int N, double *A, int lda);              /* synchronizes */
int  atlmx_spotrf(int Uplo, int N, float *A, int lda);               /* synchronizes */
void atlmx_dpotrs(int Uplo, int N, int NRHS, const double *A, int lda, double *B, int ldb);
void atlmx_spotrs(int Uplo, int N, int NRHS, const float *A, int lda, float *B, int ldb);

/* ---- batched: thousands of small, uneven-size problems ------------------
 * A plan is made once per batch shape: sizes and leading dimensions come as
 * HOST arrays; matrices as DEVICE arrays of device pointers (column-major).
 * The plan orders the instances largest first, autotunes the launch (panel
 * width, persistent workgroups) on a synthetic replica of a sample of the
 * batch (cached per shape class for the process), and captures one
 * persistent launch into a hipGraph.  atlmx_batch_run replays it on this
 * thread's stream (asynchronous); the pointer arrays and the data they
 * point to are read at every run.
 *   posv: A[i] = L L^T (lower triangle, in place; the upper is not
 *         touched), then B[i] <- A[i]^{-1} B[i]; info[i] as clapack_?potrf
 *         (the instance stops there).
 *   gesv: A[i] = P L U (partial pivoting, ipiv[i] 0-based as clapack_?getrf),
 *         then B[i] <- A[i]^{-1} B[i]; info[i] as clapack_?getrf (no solve
 *         when nonzero).
 *   gemm: C[i] = alpha op(A[i]) op(B[i]) + beta C[i].
 * nrhs <= 8.  Orders are bounded by what one workgroup holds in LDS:
 * atlmx_batch_max_order(kind, dbl) (kind 0 posv, 1 gesv): 2270 / 768 in
 * FP64, 4830 / 1621 in FP32.  A plan with a larger order is refused: a
 * NULL plan and a message on stderr. */
int atlmx_batch_max_order(int kind, int dbl);
typedef struct atlmx_batch atlmx_batch;
atlmx_batch *atlmx_dposv_batch_plan(int count, const int *h_n, double *const *d_A,
                                    const int *h_lda, int nrhs, double *const *d_B,
                                    const int *h_ldb, int *d_info);
atlmx_batch *atlmx_sposv_batch_plan(int count, const int *h_n, float *const *d_A,
                                    const int *h_lda, int nrhs, float *const *d_B,
                                    const int *h_ldb, int *d_info);
atlmx_batch *atlmx_dgesv_batch_plan(int count, const int *h_n, double *const *d_A,
                                    const int *h_lda, int *const *d_ipiv, int nrhs,
                                    double *const *d_B, const int *h_ldb, int *d_info);
atlmx_batch *atlmx_sgesv_batch_plan(int count, const int *h_n, float *const *d_A,
                                    const int *h_lda, int *const *d_ipiv, int nrhs,
                                    float *const *d_B, const int *h_ldb, int *d_info);
atlmx_batch *atlmx_dgemm_batch_plan(int TA, int TB, int count, const int *h_m, const int *h_n,
                                    const int *h_k, double alpha, const double *const *d_A,
                                    const int *h_lda, const double *const *d_B,
                                    const int *h_ldb, double beta, double *const *d_C,
                                    const int *h_ldc);
atlmx_batch *atlmx_sgemm_batch_plan(int TA, int TB, int count, const int *h_m, const int *h_n,
                                    const int *h_k, float alpha, const float *const *d_A,
                                    const int *h_lda, const float *const *d_B, const int *h_ldb,
                                    float beta, float *const *d_C, const int *h_ldc);
/* adjoints: gesv_t solves A[i]^T X = B[i] after the same LU (b-bar =
 * A^{-T} x-bar); posv_adjoint takes L[i] (potrf's lower triangle, its strict
 * upper triangle is zeroed), X[i] = A^{-1} B and Bb[i] = A^{-1} Xb and
 * writes G[i] = the gradient of <Xb, X> with respect to the lower triangle
 * of A[i], through Phi(L^T L-bar) and two triangular solves; W[i]: n x n
 * scratch (leading dimension n). */
atlmx_batch *atlmx_dgesv_t_batch_plan(int count, const int *h_n, double *const *d_A,
                                      const int *h_lda, int *const *d_ipiv, int nrhs,
                                      double *const *d_B, const int *h_ldb, int *d_info);
atlmx_batch *atlmx_sgesv_t_batch_plan(int count, const int *h_n, float *const *d_A,
                                      const int *h_lda, int *const *d_ipiv, int nrhs,
                                      float *const *d_B, const int *h_ldb, int *d_info);
atlmx_batch *atlmx_dposv_adjoint_batch_plan(int count, const int *h_n, double *const *d_L,
                                            const int *h_lda, int nrhs, double *const *d_X,
                                            double *const *d_Bb, const int *h_ldx,
                                            double *const *d_G, const int *h_ldg,
                                            double *const *d_W);
atlmx_batch *atlmx_sposv_adjoint_batch_plan(int count, const int *h_n, float *const *d_L,
                                            const int *h_lda, int nrhs, float *const *d_X,
                                            float *const *d_Bb, const int *h_ldx,
                                            float *const *d_G, const int *h_ldg,
                                            float *const *d_W);
int  atlmx_batch_run(atlmx_batch *plan);
void atlmx_batch_destroy(atlmx_batch *plan);                       /* synchronizes */
void atlmx_batch_config(const atlmx_batch *plan, int *nb, int *grid);  /* the tuned launch */

/* ---- FP64-accurate solve from a low-precision LU + FP64 refinement ------
 * A (N x N), B (N x NRHS), X (N x NRHS) are FP64 device arrays; A and B are
 * not modified.  The LU is done in FP32; the residual, the update and the
 * stopping test (atlmx_set_refine_stop) are FP64.  First the FP32 LU with tournament pivoting (pivot rows of each
 * 64-column panel chosen by a reduction tree of partial-pivoting
 * eliminations) with plain refinement; if that stalls, GMRES-IR (GMRES in
 * FP64 preconditioned by the same FP32 factors) reaches condition numbers
 * far beyond 1/eps32; if a pivot is zero or that fails too, the FP32 LU with
 * partial pivoting (plain, then GMRES), then the FP64 LU.  Synchronizes;
 * returns 0 or the FP64 LU's info. */
typedef struct {
  int iterations;          /* refinement steps taken (max over RHS groups)  */
  int converged;           /* stopping test met without fallback            */
  int fallback;            /* solved by the FP64 LU instead                 */
  int factor_info;         /* info of the last low-precision factorization  */
  double backward_error;   /* max_j ||b-Ax||inf / (||A||inf ||x||inf + ||b||inf) */
  int method;              /* 1 tournament-pivoting FP32, 2 partial-pivoting
                              FP32, 3 FP64, 4 FP32 Cholesky, 5 (6) tournament
                              with bfloat16 (half) trailing updates */
  int gmres;               /* 1: the corrections came from GMRES-IR (plain
                              refinement with these factors stalled) */
  int inner_iterations;    /* GMRES iterations in total */
} atlmx_ir_info;
/* SPD twin: FP32 Cholesky (lower triangle) + FP64 refinement, then the FP64
 * Cholesky.  A must hold the full symmetric matrix (the FP64 residual uses
 * both triangles).  method: 4 FP32 Cholesky, 3 FP64. */
int atlmx_dsposv(int N, int NRHS, const double *A, int lda, const double *B, int ldb,
                 double *X, int ldx, int iter_max, atlmx_ir_info *info);
/* Refinement stopping test: 0 (default) normwise backward error
 * ||b - Ax||inf / (||A||inf ||x||inf + ||b||inf) <= N * 2^-52; 1 LAPACK
 * dsgesv's ||r||inf <= ||x||inf ||A||inf 2^-53 sqrt(N).  Process-wide. */
void atlmx_set_refine_stop(int lapack);
/* 0 (default): tournament, then partial pivoting, then FP64; 1: tournament
 * then FP64; 2: partial pivoting then FP64; 5 (6): tournament with the
 * trailing updates' operands rounded to bfloat16 (IEEE half), A's columns
 * scaled by powers of two, then as 0.  The method that produced the result
 * is reported in atlmx_ir_info.method.  Process-wide. */
void atlmx_set_mixed_method(int method);
int atlmx_dsgesv(int N, int NRHS, const double *A, int lda, const double *B, int ldb,
                 double *X, int ldx, int iter_max, atlmx_ir_info *info);
double atlmx_dbackward_error(int N, int NRHS, const double *A, int lda, const double *B,
                             int ldb, const double *X, int ldx);   /* synchronizes */

/* ---- allocation-free LU solve --------------------------------------------
 * No counterpart in ATLAS (nor a reference interface): it is for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  The same solve as
 * atlmx_dgetrs (same kernels, same order, bit-identical results), but its
 * scratch (the inverted 64 x 64 diagonal blocks of the factor and the
 * solve's published words) lives in one static device array of
 * atlmx_noalloc_scratch_bytes() bytes per GPU, held until the process ends;
 * no hipMalloc, hipFree or other allocation is made.  N <= ATLMX_NOALLOC_MAX_N:
 * a larger N (or another illegal argument) returns minus its position, with
 * a message on stderr, and does nothing.  Asynchronous on the calling
 * thread's stream; returns 0.  Calls that overlap on one GPU (other threads
 * or streams) race on the shared array. */
#define ATLMX_NOALLOC_MAX_N 16384
int atlmx_dgetrs_noalloc(int Trans, int N, int NRHS, const double *A, int lda,
                         const int *d_ipiv, double *B, int ldb);
size_t atlmx_noalloc_scratch_bytes(void);

#ifdef __cplusplus
}
#endif
#endif

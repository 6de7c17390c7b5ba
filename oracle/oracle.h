/*
 * ORACLE -- test infrastructure only.
 *
 * A CPU restatement, in plain C, of the ATLAS v3.11.41 algorithms on the
 * Level-3 hot path (GEMM, TRSM, recursive LU and Cholesky, getrs/potrs) and of
 * the ATLAS testers that judge them (gemmtst, lutst, llttst, slvtst,
 * l3blastst).  It is the checker for the MI355X library in atlas_amd/ and the
 * CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * Every function cites the reference file:line it restates.  How each one is
 * pinned (against oracle/_ref built from the reference's own C reference BLAS,
 * against glibc/numpy/scipy routines, closed forms, or brute force) is
 * recorded in the tests under tests/test_oracle_*.py.
 *
 * Enum values are those of ATLAS's include/cblas.h:5-10.
 */
#ifndef ORACLE_H
#define ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_ROW = 101, ORC_COL = 102 };
enum { ORC_NOTRANS = 111, ORC_TRANS = 112, ORC_CONJTRANS = 113 };
enum { ORC_UPPER = 121, ORC_LOWER = 122 };
enum { ORC_NONUNIT = 131, ORC_UNIT = 132 };
enum { ORC_LEFT = 141, ORC_RIGHT = 142 };

/* ---- portable 31-bit LCG (src/testing/ATL_rand.c) ---------------------- */
typedef struct { int ir[2], ia[2], ic[2]; } orc_lcg_t;
void   orc_lcg_seed(orc_lcg_t *g, int iseed);
int    orc_lcg_next(orc_lcg_t *g);
void   orc_srand(int iseed);              /* global stream, like ATL_srand */
int    orc_rand(void);                    /* like ATL_rand                 */
double orc_dumb_rand(void);               /* include/atlas_tst.h:18        */
void   orc_rand_stream(int iseed, int n, int *out);

/* ---- configuration knobs ------------------------------------------------ */
/* ATLAS's recursion cut-offs are tuning products: L1C_ELTS (getf2 base case,
 * src/lapack/ATL_getrfC.c:264) and ATL_VWipgen_100LCMMN (Nleft rounding,
 * ATL_getrfC.c:280-283, ATL_potrfL.c:223-226).  Defaults: 4096 and 4. */
void orc_set_recursion_params(int l1c_elts, int lcm);
void orc_set_threads(int nthreads);       /* threads for orc_?gemm_fast   */
int  orc_get_threads(void);
/* When gaps != NULL, every pivot search of orc_?getrf records, for global
 * column j, (|a_max| - |a_second|)/|a_max| of the searched column. */
void orc_track_pivot_gaps(double *gaps, int n);

#define ORC_DECLARE(P, T)                                                     \
void P##gegen(int M, int N, T *A, int lda, int seed);                         \
void P##gefillgap(int M, int N, T *A, int lda);                               \
int  P##gechkgap(int M, int N, const T *A, int lda);                          \
void P##matgen(int M, int N, T *A, int lda, int seed);                        \
T    P##epsilon(void);                                                        \
T    P##genrm1(int M, int N, const T *A, int lda);                            \
T    P##gediffnrm1(int M, int N, const T *A, int lda, const T *B, int ldb);   \
T    P##infnrm(int N, const T *X, int incX);                                  \
int  P##iamax(int N, const T *X, int incX);                                   \
void P##gemm_ref(int TA, int TB, int M, int N, int K, T alpha,                \
                 const T *A, int lda, const T *B, int ldb, T beta,            \
                 T *C, int ldc);                                              \
void P##gemm_fast(int TA, int TB, int M, int N, int K, T alpha,               \
                  const T *A, int lda, const T *B, int ldb, T beta,           \
                  T *C, int ldc);                                             \
void P##gemm(int order, int TA, int TB, int M, int N, int K, T alpha,         \
             const T *A, int lda, const T *B, int ldb, T beta, T *C, int ldc);\
void P##trsm_ref(int side, int uplo, int ta, int diag, int M, int N, T alpha, \
                 const T *A, int lda, T *B, int ldb);                         \
void P##trsm(int order, int side, int uplo, int ta, int diag, int M, int N,   \
             T alpha, const T *A, int lda, T *B, int ldb);                    \
void P##syrk_ref(int uplo, int trans, int N, int K, T alpha, const T *A,      \
                 int lda, T beta, T *C, int ldc);                             \
void P##trmm_ref(int side, int uplo, int ta, int diag, int M, int N, T alpha, \
                 const T *A, int lda, T *B, int ldb);                         \
void P##trmm_fast(int side, int uplo, int ta, int diag, int M, int N,         \
                  T alpha, const T *A, int lda, T *B, int ldb);               \
void P##symm_ref(int side, int uplo, int M, int N, T alpha, const T *A,       \
                 int lda, const T *B, int ldb, T beta, T *C, int ldc);        \
void P##syr2k_ref(int uplo, int trans, int N, int K, T alpha, const T *A,     \
                  int lda, const T *B, int ldb, T beta, T *C, int ldc);       \
void P##laswp(int N, T *A, int lda, int K1, int K2, const int *piv, int inci);\
int  P##getf2(int M, int N, T *A, int lda, int *ipiv);                        \
int  P##getrf(int order, int M, int N, T *A, int lda, int *ipiv);             \
void P##getrs(int order, int trans, int N, int NRHS, const T *A, int lda,     \
              const int *ipiv, T *B, int ldb);                                \
int  P##potrf(int order, int uplo, int N, T *A, int lda);                     \
void P##potrs(int order, int uplo, int N, int NRHS, const T *A, int lda,      \
              T *B, int ldb);                                                 \
int  P##gesv(int order, int N, int NRHS, T *A, int lda, int *ipiv, T *B,      \
             int ldb);                                                        \
int  P##posv(int order, int uplo, int N, int NRHS, T *A, int lda, T *B,       \
             int ldb);                                                        \
T    P##lu_resid(int M, int N, const T *LU, int lda, const int *ipiv,         \
                 const T *A0, int lda0);                                      \
T    P##lutest(int M, int N, int lda, int *npiv);                             \
void P##lltgen(int uplo, int N, T *A, int lda, int seed);                     \
T    P##lltnrm1(int uplo, int N, const T *A, int lda);                        \
T    P##llt_resid(int uplo, int N, const T *L, int lda, const T *A0,          \
                  int lda0);                                                  \
T    P##llttest(int uplo, int N, int lda);                                    \
T    P##slv_resid(int N, int NRHS, T nrmA, const T *XC, int ldxc,             \
                  const T *X, int ldx);                                       \
int  P##gemm_ferr_check(int M, int N, int K, T alpha, T beta, const T *C,     \
                        int ldc, const T *D, int ldd, double *maxdiff,        \
                        double *ferr);                                        \
T    P##trsm_resid(int side, int uplo, int ta, int diag, int M, int N,        \
                   T alpha, const T *A, int lda, const T *X, int ldx,         \
                   const T *B0, int ldb0);                                   \
T    P##trinrm1(int uplo, int diag, int N, const T *A, int lda);              \
void P##trigen(int uplo, int diag, int N, T *A, int lda, T padval, int seed); \
T    P##trtri_resid(int uplo, int diag, int N, const T *Ainv, int ldi,        \
                    const T *A0, int lda0);                                   \
T    P##inv_resid(int N, const T *A, int lda, const T *AI, int ldi);          \
T    P##poinv_resid(int uplo, int N, const T *A, int lda, const T *AI,        \
                    int ldi);                                                 \
T    P##lauum_resid(int uplo, int N, const T *R, int ldr, const T *T0,        \
                    int ld0);                                                 \
void P##geqr2(int M, int N, T *A, int lda, T *tau);                           \
T    P##qr_resid(int M, int N, const T *QR, int lda, const T *tau,            \
                 const T *A0, int lda0);

ORC_DECLARE(orc_d, double)
ORC_DECLARE(orc_s, float)

/* complex (z: double, c: float): interleaved storage, leading dimensions in
 * complex elements, alpha/beta point at {real, imaginary} */
#define ORC_CDECLARE(P, T)                                                    \
void P##gemm_ref(int TA, int TB, int M, int N, int K, const T *alpha,        \
                  const T *A, int lda, const T *B, int ldb, const T *beta,    \
                  T *C, int ldc);                                             \
void P##trsm_ref(int side, int uplo, int TA, int diag, int M, int N,         \
                  const T *alpha, const T *A, int lda, T *B, int ldb);        \
int  P##getf2(int M, int N, T *A, int lda, int *ipiv, T *mingap);           \
int  P##potrf(int uplo, int N, T *A, int lda);

ORC_CDECLARE(orc_z, double)
ORC_CDECLARE(orc_c, float)

#ifdef __cplusplus
}
#endif
#endif

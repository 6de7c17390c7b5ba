/*
 * kernels.h -- the thin C-ABI between the C host layer and the gfx950 HIP
 * kernels of libatlas_mi355x.  Every launcher takes device pointers, 64-bit
 * leading dimensions and the stream to launch on, checks the launch shape
 * against what its kernel assumes, and returns a hipError_t (0 = success).
 * Matrices are column-major.
 */
#ifndef ATLMX_KERNELS_H
#define ATLMX_KERNELS_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Operand codes used by the launchers (values of ATLAS include/cblas.h:5-10). */
#define KX_NOTRANS 111
#define KX_TRANS   112
#define KX_CONJTRANS 113
#define KX_UPPER   121
#define KX_LOWER   122
#define KX_NONUNIT 131
#define KX_UNIT    132
#define KX_LEFT    141
#define KX_RIGHT   142
#define KX_CUPPER  1121
#define KX_CLOWER  1122
/* K bands of a triangular operand (TRMM, gemm_kernel only): the K slices
 * that meet only the triangle's zeros are skipped, C is whole.  AL / AU:
 * op(A) is a real lower / upper triangle or the 2x2-block expansion of a
 * complex one; BL / BU: op(B) is the interleaved real view of a complex
 * lower / upper triangle; RBL / RBU: op(B) is a real lower / upper one. */
#define KX_KBAND_AL 1131
#define KX_KBAND_AU 1132
#define KX_KBAND_BL 1133
#define KX_KBAND_BU 1134
#define KX_KBAND_RBL 1135
#define KX_KBAND_RBU 1136

/* C = alpha*op(A)*op(B) + beta*C on the MFMA GEMM kernel.
 * tri: 0 = whole C; KX_LOWER / KX_UPPER = only elements on or below / above
 * the diagonal of C are read and written (the SYRK shape); KX_CLOWER /
 * KX_CUPPER = the same for C the interleaved real view of a complex matrix
 * (real row r is complex row r / 2: the HERK shape; gemm_kernel only);
 * KX_KBAND_* = whole C, a triangular operand's zero K slices skipped.
 * abort_flag: NULL, or a device int; the kernel does nothing when it is
 * nonzero (lets a stopped Cholesky skip the rest of its queued work). */
hipError_t atlmx_k_dgemm(int ta, int tb, int M, int N, int K, double alpha,
                         const double *A, long lda, const double *B, long ldb,
                         double beta, double *C, long ldc, int tri,
                         const int *abort_flag, hipStream_t s);
hipError_t atlmx_k_sgemm(int ta, int tb, int M, int N, int K, float alpha,
                         const float *A, long lda, const float *B, long ldb,
                         float beta, float *C, long ldc, int tri,
                         const int *abort_flag, hipStream_t s);

/* C += alpha A op(B) in FP64 (A not transposed, beta = 1) with the C tile
 * loaded behind the MFMAs: the short-K trailing updates of getrf / potrf.
 * tb: KX_NOTRANS or KX_TRANS; tri as atlmx_k_dgemm. */
hipError_t atlmx_k_dupdate(int tb, int M, int N, int K, double alpha, const double *A,
                           long lda, const double *B, long ldb, double *C, long ldc, int tri,
                           const int *abort_flag, hipStream_t s);

/* Largest row count one in-place GEMM tile holds: an in-place product
 * B := inv(T)*B with M <= ATLMX_GEMM_BM rows (or B := B*inv(T) with
 * N <= ATLMX_GEMM_BN columns) reads all of its input before any write. */
#define ATLMX_GEMM_BM 128
#define ATLMX_GEMM_BN 128

/* Invert the nb x nb diagonal blocks of a triangular n x n matrix A into
 * inv (nblk blocks of nb*nb, column-major, ld = nb, zero outside the
 * triangle; the last block is padded with the identity).  nb = 64. */
#define ATLMX_TRI_NB 64
hipError_t atlmx_k_dtri_inv(int uplo, int diag, int n, const double *A,
                            long lda, double *inv, const int *abort_flag,
                            hipStream_t s);
hipError_t atlmx_k_stri_inv(int uplo, int diag, int n, const float *A,
                            long lda, float *inv, const int *abort_flag,
                            hipStream_t s);

/* op(A) X = B for a few right-hand sides (nrhs <= ATLMX_TRSV_MAXRHS) in
 * one launch; X holds B on entry.  inv: the inverted 64x64 diagonal blocks
 * of A from atlmx_k_?tri_inv (with the same uplo/diag).  ctl: device ints,
 * atlmx_k_trsv_ctl_ints(n, is_double) of them (8-byte aligned), zero on
 * entry. */
size_t atlmx_k_trsv_ctl_ints(int n, int dbl);
#define ATLMX_TRSV_MAXRHS 4
hipError_t atlmx_k_dtrsv(int uplo, int trans, int n, int nrhs, const double *A,
                         long lda, const double *inv, double *X, long ldx,
                         int *ctl, hipStream_t s);
hipError_t atlmx_k_strsv(int uplo, int trans, int n, int nrhs, const float *A,
                         long lda, const float *inv, float *X, long ldx,
                         int *ctl, hipStream_t s);

/* The same for up to ATLMX_TRSM_TAG_MAXRHS right-hand sides (any number of
 * 64-row blocks; work item = one block of 16 columns, products on the matrix
 * cores).  ctl: atlmx_k_trsm_tag_ctl_ints(n, nrhs, is_double) ints, 8-byte
 * aligned, zero on entry.  The columns are independent, so a solve split
 * into launches over column ranges gives the same result. */
#define ATLMX_TRSM_TAG_MAXRHS 256
/* columns per launch when the caller brings the workspace */
#define ATLMX_TRSM_TAG_CHUNK 64
size_t atlmx_k_trsm_tag_ctl_ints(int n, int nrhs, int dbl);
hipError_t atlmx_k_dtrsm_tag(int uplo, int trans, int n, int nrhs, const double *A,
                             long lda, const double *inv, double *X, long ldx,
                             int *ctl, hipStream_t s);
hipError_t atlmx_k_strsm_tag(int uplo, int trans, int n, int nrhs, const float *A,
                             long lda, const float *inv, float *X, long ldx,
                             int *ctl, hipStream_t s);

/* Tournament pivoting (FP32) on S = A^T (column-major, A(i,j) = S[j+i*lds],
 * lds and k multiples of 4, S 16-byte aligned): chooses the w <=
 * ATLMX_TSLU_W pivot rows of the panel rows k..n-1, columns k..k+w-1,
 * writes the move list that brings them to rows k.. in pivot order
 * (plan[0] = count <= 2*ATLMX_TSLU_W, then dst, src pairs), applies those
 * moves to perm, and writes the w x w diagonal block's factor L\U (lu11)
 * and inv(L11) (unit lower), inv(U11) (invl, invu), all ATLMX_TSLU_W x
 * ATLMX_TSLU_W column-major, identity beyond w.  A zero or non-finite U11
 * pivot sets *info = k + j + 1 (if 0); with *info nonzero on entry nothing
 * is done.  ws: atlmx_k_tslu_ws_ints(n) ints, zero before the first call
 * (each call leaves its counters zero).  With prev (the previous panel's
 * plan) given, plan2 (2 + 8*ATLMX_TSLU_W ints) receives one move list equal
 * to prev's moves followed by this panel's. */
#define ATLMX_TSLU_W 64
size_t atlmx_k_tslu_ws_ints(int n);
hipError_t atlmx_k_tslu_select(const float *S, long lds, int n, int k, int w, int *ws,
                               int *plan, int *perm, float *lu11, float *invl, float *invu,
                               int *info, const int *prev, int *plan2, hipStream_t s);
/* the same, and then, with sw_e1 > sw_e0 (sw_e0 <= k, k + w <= sw_e1,
 * at most 2*ATLMX_TSLU_W entries), what atlmx_k_tslu_swap(S, lds, sw_e0,
 * sw_e1, 0, 0, plan, k, w, lu11, ...) does: the moves on S's entries
 * [sw_e0, sw_e1) and the factored block written over rows k..k+w-1 --
 * done by the last workgroup, which holds the plan and the block.  With
 * u12_r1 > k + w (<= sw_e1, at most ATLMX_TSLU_W rows) it also forms U12
 * on S rows [k + w, u12_r1): Ut = Ut inv(L11)^T over S columns k..k+w-1,
 * as the in-place GEMM with invl would. */
hipError_t atlmx_k_tslu_select_swap(float *S, long lds, int n, int k, int w, int *ws,
                                    int *plan, int *perm, float *lu11, float *invl,
                                    float *invu, int *info, const int *prev, int *plan2,
                                    int sw_e0, int sw_e1, int u12_r1, hipStream_t s);
/* after the pair's first panel P at k (64 wide) is factored on its own S
 * rows and U12 formed on the second panel's wq S rows [k+64, k+64+wq): for
 * every S column in [k+64, n), L21 = inv(U11)^T Lt on S rows k..k+63 (invu:
 * P's inv(U11), ATLMX_TSLU_W x ATLMX_TSLU_W column-major), then those wq
 * rows -= U12 L21, with the update's operands rounded to bfloat16 (lp = 1)
 * or half (lp = 2) as atlmx_k_supdate_lp does.  One launch in place of the
 * in-place L21 GEMM and the narrow update. */
hipError_t atlmx_k_tslu_l21up(float *S, long lds, int n, int k, int wq, const float *invu,
                              int lp, const int *info, hipStream_t s);
/* the pair P (A columns k..k+63) and Q (the next 64, k + 128 <= n) after
 * both were factored within S rows k..k+127: for the S rows [0, k) and
 * [k+128, n), the row moves of plan (P's followed by Q's, as
 * atlmx_k_tslu_select writes to plan2), then on the rows from k+128
 * U12(P) = U12(P) inv(L_P)^T, U12(Q) -= U12(P) L21(P, Q rows),
 * U12(Q) = U12(Q) inv(L_Q)^T (transposed in S). */
hipError_t atlmx_k_tslu_tail(float *S, long lds, int n, int k, const int *plan,
                             const float *invlp, const float *invlq, const int *info,
                             hipStream_t s);
/* columns dst <- src of S (rows of A) for the moves of plan, entries
 * [ebeg, eend) of each column (S rows, i.e. columns of A) except the gap
 * [gbeg, gend) (none when gend <= gbeg); with lu11 non-NULL, the part of
 * the diagonal block (rows and entries k..k+w-1) in that range <- lu11.
 * Nothing is done when *info is nonzero. */
hipError_t atlmx_k_tslu_swap(float *S, long lds, int ebeg, int eend, int gbeg, int gend,
                             const int *plan, int k, int w, const float *lu11, const int *info,
                             hipStream_t s);
/* Batched factor + solve in one persistent launch (k_batched.hip).
 * kind 0: Cholesky (lower) + solve; kind 1: LU (partial pivoting) + solve.
 * Instance i = order[it] for it = 0.. (taken from *counter, zero on entry):
 * n[i], A[i] (lda[i]), ipiv[i] (kind 1, 0-based), nrhs <= ATLMX_BATCH_MAXRHS
 * right-hand sides B[i] (ldb[i]; B may be NULL with nrhs 0), info[i].
 * maxn bounds every n[i] (LDS sizing); nb: panel width 16 or 32; grid:
 * persistent workgroups. */
#define ATLMX_BATCH_MAXRHS 8
hipError_t atlmx_k_dbatch_factor(int kind, int nb, int count, const int *order, const int *n,
                                 double *const *A, const int *lda, int *const *ipiv, int nrhs,
                                 double *const *B, const int *ldb, int *info, int *counter,
                                 int maxn, int grid, hipStream_t s);
hipError_t atlmx_k_sbatch_factor(int kind, int nb, int count, const int *order, const int *n,
                                 float *const *A, const int *lda, int *const *ipiv, int nrhs,
                                 float *const *B, const int *ldb, int *info, int *counter,
                                 int maxn, int grid, hipStream_t s);
/* kind 2 of atlmx_k_?batch_factor: LU as kind 1, then solve A^T X = B. */
/* Largest order n[i] the factor kernels hold in LDS for kind (0 Cholesky,
 * 1 or 2 LU) and precision (dbl 1: FP64).  Host-only arithmetic. */
int atlmx_k_batch_max_order(int kind, int dbl);
/* Cholesky adjoint of the solve X = A^{-1} B per instance: L (potrf's lower
 * triangle; its strict upper triangle is zeroed), X, Bb = A^{-1} Xb (n x
 * nrhs, ld ldx); writes G (ld ldg) = gradient of <Xb, X> with respect to
 * the lower triangle of A; W: n x n scratch (ld n). */
hipError_t atlmx_k_dbatch_chol_adjoint(int count, const int *order, const int *n,
                                       double *const *L, const int *lda, int nrhs,
                                       double *const *X, double *const *Bb, const int *ldx,
                                       double *const *G, const int *ldg, double *const *W,
                                       int *counter, int grid, hipStream_t s);
hipError_t atlmx_k_sbatch_chol_adjoint(int count, const int *order, const int *n,
                                       float *const *L, const int *lda, int nrhs,
                                       float *const *X, float *const *Bb, const int *ldx,
                                       float *const *G, const int *ldg, float *const *W,
                                       int *counter, int grid, hipStream_t s);
/* C[i] = alpha op(A[i]) op(B[i]) + beta C[i] for count instances of sizes
 * m[i] x n[i] x k[i]; tiles are numbered instance by instance in `order`,
 * tile_start[p] = first tile of instance order[p] (count + 1 entries).
 * variant: 0 vector FMAs (256 threads per tile), 1 MFMA with one
 * wavefront per tile, 2 MFMA with four wavefronts per tile. */
#define ATLMX_BATCH_TILE 64
hipError_t atlmx_k_dbatch_gemm(int count, int ta, int tb, const int *order, const int *m,
                               const int *n,
                               const int *k, double alpha, const double *const *A,
                               const int *lda, const double *const *B, const int *ldb,
                               double beta, double *const *C, const int *ldc,
                               const int *tile_start, int *counter, int variant, int grid,
                               hipStream_t s);
hipError_t atlmx_k_sbatch_gemm(int count, int ta, int tb, const int *order, const int *m,
                               const int *n,
                               const int *k, float alpha, const float *const *A,
                               const int *lda, const float *const *B, const int *ldb,
                               float beta, float *const *C, const int *ldc,
                               const int *tile_start, int *counter, int variant, int grid,
                               hipStream_t s);

/* S = (float) (A diag(colscale))^T (colscale may be NULL), perm = 0..n-1 */
hipError_t atlmx_k_tslu_convert(int n, const double *A, long lda, float *S, long lds, int *perm,
                                const double *colscale, hipStream_t s);
/* scale[j] = the power of two with mx[j] * scale[j] in [0.5, 1) (1 for 0) */
hipError_t atlmx_k_pow2_scale(int n, const double *mx, double *scale, hipStream_t s);
/* X(i, :) = (or +=) scale[i] * D(i, :), D FP32, X FP64 */
hipError_t atlmx_k_scale_s2d(int M, int N, const float *D, long ldd, const double *scale,
                             double *X, long ldx, int accumulate, hipStream_t s);
/* Y(i, :) = (float) X(perm[i], :) */
hipError_t atlmx_k_tslu_gather(int n, int nrhs, const double *X, long ldx, const int *perm,
                               float *Y, long ldy, hipStream_t s);
/* C += alpha A B (FP32, A and B not transposed): the short-K trailing
 * update, C read once into the MFMA accumulators */
hipError_t atlmx_k_supdate(int M, int N, int K, float alpha, const float *A, long lda,
                           const float *B, long ldb, float *C, long ldc, const int *abort_flag,
                           hipStream_t s);
/* the same with A and B rounded to bfloat16 (lp = 1) or IEEE half (lp = 2),
 * round to nearest even, products accumulated in FP32 onto the FP32 C */
hipError_t atlmx_k_supdate_lp(int lp, int M, int N, int K, float alpha, const float *A,
                              long lda, const float *B, long ldb, float *C, long ldc,
                              const int *abort_flag, hipStream_t s);

/* B(M x N) *= alpha */
hipError_t atlmx_k_dscal2d(int M, int N, double alpha, double *B, long ldb,
                           hipStream_t s);
hipError_t atlmx_k_sscal2d(int M, int N, float alpha, float *B, long ldb,
                           hipStream_t s);

/* LU panel: factor the m x w column panel P (ld), w <= ATLMX_PANEL_MAXW
 * (w <= ATLMX_PANEL_MAXW/2 when m > ATLMX_PANEL_TALL), with partial
 * pivoting: in one cooperative launch up to 2*65536 rows, beyond that as
 * three launches per column on the panel in HBM.  ipiv[c] receives the
 * absolute pivot row (row0 + panel row); info (device int) gets row0+c+1
 * for the first exactly-zero pivot if it is still 0.  scratch: device
 * buffer of atlmx_k_panel_scratch_bytes() bytes, initialised with
 * atlmx_k_panel_scratch_init() (which clears every exchange tag); the
 * launches on one scratch take increasing sequence numbers seq = 1, 2, ...
 * below ATLMX_PANEL_SEQ_MAX, and after that the scratch is initialised again
 * and the numbering restarts at 1. */
#define ATLMX_PANEL_MAXW 64
#define ATLMX_PANEL_TALL 65536
#define ATLMX_PANEL_SEQ_MAX (1 << 26)
size_t atlmx_k_panel_scratch_bytes(int m, int w);
hipError_t atlmx_k_panel_scratch_init(void *scratch, hipStream_t s);
hipError_t atlmx_k_dgetf2_panel(int m, int w, double *P, long ld, int row0,
                                int *ipiv, int *info, void *scratch, int seq,
                                hipStream_t s);
hipError_t atlmx_k_sgetf2_panel(int m, int w, float *P, long ld, int row0,
                                int *ipiv, int *info, void *scratch, int seq,
                                hipStream_t s);

/* Apply the row interchanges ipiv[k1..k2) (absolute row numbers below
 * rows_end, forward order when inc > 0, reverse when inc < 0) to the N
 * columns of A.  Up to 256 columns with rows_end-k1 <= 16384 replay the
 * swaps in LDS (one workgroup per column); otherwise one thread per
 * column. */
hipError_t atlmx_k_dlaswp(int N, double *A, long lda, int k1, int k2,
                          const int *ipiv, int inc, int rows_end, hipStream_t s);
hipError_t atlmx_k_slaswp(int N, float *A, long lda, int k1, int k2,
                          const int *ipiv, int inc, int rows_end, hipStream_t s);

/* The same interchanges as one permutation of rows [k1, rows_end)
 * (rows_end-k1 <= ATLMX_LASWP_PERM_MAXROWS): src[i] = the row (relative to
 * k1) whose contents end up in row k1+i; tmp: rows_end-k1 ints of scratch.
 * permute_rows then applies such a map to N columns through tmp
 * (rows_end-k1 by N elements). */
#define ATLMX_LASWP_PERM_MAXROWS 16384
hipError_t atlmx_k_laswp_perm(int k1, int k2, int rows_end, const int *ipiv, int inc,
                              int *src, int *tmp, hipStream_t s);
hipError_t atlmx_k_dpermute_rows(int N, double *A, long lda, int k1, int nr,
                                 const int *src, double *tmp, hipStream_t s);
hipError_t atlmx_k_spermute_rows(int N, float *A, long lda, int k1, int nr,
                                 const int *src, float *tmp, hipStream_t s);

/* Forward interchanges k1..k2 (k2-k1 <= 2*ATLMX_PANEL_MAXW) of rows below
 * rows_end (rows_end-k1 <= ATLMX_LASWP_PLAN_MAXROWS) as a move list:
 * plan[0] = count, plan[2+2i] = destination row, plan[3+2i] = source row;
 * plan holds ATLMX_LASWP_PLAN_INTS ints.  apply moves the rows in columns
 * [c0, c1) minus [s0, s1). */
#define ATLMX_LASWP_PLAN_MAXROWS 32768
#define ATLMX_LASWP_PLAN_INTS (2 + 8 * ATLMX_PANEL_MAXW)
hipError_t atlmx_k_laswp_plan(int k1, int k2, int rows_end, const int *ipiv,
                              int *plan, hipStream_t s);
hipError_t atlmx_k_dlaswp_apply(double *A, long lda, int c0, int c1, int s0,
                                int s1, const int *plan, hipStream_t s);
hipError_t atlmx_k_slaswp_apply(float *A, long lda, int c0, int c1, int s0,
                                int s1, const int *plan, hipStream_t s);

/* Cholesky of one n x n diagonal block (n <= 128) in place (lower or
 * upper).  On a non-positive pivot sets *info = info_base + k + 1 (if still
 * 0) and stops; does nothing when *info is already nonzero. */
#define ATLMX_CHOL_NB 128
hipError_t atlmx_k_dpotf2(int uplo, int n, double *A, long lda, int info_base,
                          int *info, hipStream_t s);
hipError_t atlmx_k_spotf2(int uplo, int n, float *A, long lda, int info_base,
                          int *info, hipStream_t s);
/* The panel's triangular solve after a jb x jb (jb <= 128) diagonal block D
 * is factored (uplo as for potf2): nr vectors of P (Lower: the rows of the
 * nr x jb block below D; Upper: the columns of the jb x nr block right of
 * it) are solved in place against D's factor, using inv = the inverted
 * 64 x 64 diagonal blocks of D from atlmx_k_?tri_inv.  Does nothing when
 * *info is nonzero. */
hipError_t atlmx_k_dchol_panel(int uplo, int nr, int jb, const double *D, long lda,
                               const double *inv, double *P, const int *info, hipStream_t s);
hipError_t atlmx_k_schol_panel(int uplo, int nr, int jb, const float *D, long lda,
                               const float *inv, float *P, const int *info, hipStream_t s);

/* First index of max |x_i| (cblas_i?amax semantics) into *out. */
hipError_t atlmx_k_idamax(int n, const double *x, int incx, int *out,
                          hipStream_t s);
hipError_t atlmx_k_isamax(int n, const float *x, int incx, int *out,
                          hipStream_t s);

/* Triangle copies (k_aux.hip), column-major, for the routines built on the
 * dense GEMM or a dense workspace:
 *   EXPAND_TRI: W (n x n) = A's uplo triangle, zero outside, 1 on the
 *               diagonal when diag is KX_UNIT
 *   EXPAND_SYM: W = the symmetric matrix stored in A's uplo triangle
 *   STORE_TRI : A's uplo triangle = W's (diagonal not written when unit)
 *   IDENTITY  : W (m x n) = I (A unused)
 *   COPY      : A (m x n) = W
 *   TRANSPOSE : W (n x m, ldw >= n) = A^T, A m x n */
#define ATLMX_TC_EXPAND_TRI 0
#define ATLMX_TC_EXPAND_SYM 1
#define ATLMX_TC_STORE_TRI 2
#define ATLMX_TC_IDENTITY 3
#define ATLMX_TC_COPY 4
#define ATLMX_TC_TRANSPOSE 5
hipError_t atlmx_k_dtricopy(int mode, int uplo, int diag, int m, int n, double *A, long lda,
                            double *W, long ldw, hipStream_t s);
hipError_t atlmx_k_stricopy(int mode, int uplo, int diag, int m, int n, float *A, long lda,
                            float *W, long ldw, hipStream_t s);
/* *info (device) = 1 + the first i < n with A(i,i) == 0, else 0 */
hipError_t atlmx_k_ddiag_zero(int n, const double *A, long lda, int *info, hipStream_t s);
hipError_t atlmx_k_sdiag_zero(int n, const float *A, long lda, int *info, hipStream_t s);

/* Householder QR (k_qr.hip).  geqr2_panel: factor the m x w panel P (w <=
 * ATLMX_QR_NB) column by column as LAPACK's geqr2/larfg: R on and above
 * the diagonal, the reflectors' vectors (unit leading entry implied) below,
 * tau[0..min(m,w)); work: atlmx_k_qr_work_doubles() doubles of scratch.
 * qr_vcopy: V (m x w, ld m) = the panel's unit lower trapezoid.  larft:
 * the w x w upper triangular factor T of the block reflector from G =
 * V^T V (ld w) and tau. */
#define ATLMX_QR_NB 32
size_t atlmx_k_qr_work_doubles(void);
hipError_t atlmx_k_dgeqr2_panel(int m, int w, double *P, long ld, double *tau, double *work,
                                hipStream_t s);
hipError_t atlmx_k_sgeqr2_panel(int m, int w, float *P, long ld, float *tau, double *work,
                                hipStream_t s);
hipError_t atlmx_k_dqr_vcopy(int m, int w, const double *P, long ld, double *V, hipStream_t s);
hipError_t atlmx_k_sqr_vcopy(int m, int w, const float *P, long ld, float *V, hipStream_t s);
hipError_t atlmx_k_dlarft(int w, const double *G, const double *tau, double *Tm, hipStream_t s);
hipError_t atlmx_k_slarft(int w, const float *G, const float *tau, float *Tm, hipStream_t s);

/* Fill n ints with v (stream-ordered, no memset node). */
hipError_t atlmx_k_fill_i32(int *p, int n, int v, hipStream_t s);

/* ---- complex (z: double, c: float; interleaved real/imaginary parts;
 * leading dimensions in complex elements unless marked as reals) ---- */
/* out = (ar + i ai) * op(A), op(A) m x n: A(i,p) or, for trans, A(p,i),
 * conjugated when conj.  expand = 1: out is the real 2m x 2n matrix with
 * each element as the 2x2 block [Re -Im; Im Re] (ldo in reals); expand = 0:
 * out is complex m x n (ldo in complex elements).  tri = KX_LOWER/KX_UPPER
 * zeroes op(A) outside that triangle; unit puts 1 on the diagonal. */
hipError_t atlmx_k_zgather(int m, int n, const double *A, long lda, int trans, int conj,
                           double ar, double ai, int expand, double *out, long ldo, int tri,
                           int unit, hipStream_t s);
hipError_t atlmx_k_cgather(int m, int n, const float *A, long lda, int trans, int conj,
                           float ar, float ai, int expand, float *out, long ldo, int tri,
                           int unit, hipStream_t s);
/* C = (br + i bi) C; zero beta writes zeros without reading C */
hipError_t atlmx_k_zscal2d(int m, int n, double br, double bi, double *C, long ldc,
                           hipStream_t s);
hipError_t atlmx_k_cscal2d(int m, int n, float br, float bi, float *C, long ldc, hipStream_t s);
/* Inverses of the 32x32 complex diagonal blocks of the triangle (lower or
 * upper) held expanded in H (2n x 2n reals, ldh in reals), written expanded
 * as 64x64 real blocks (ld 64) at inv + b*64*64; identity past the edge. */
#define ATLMX_CTRI_NB 32
/* largest expanded left operand a complex GEMM builds at once (row chunks) */
#define ATLMX_CPLX_WS_BYTES (1024.0 * 1024.0 * 1024.0)
hipError_t atlmx_k_ztri_inv(int lower, int n, const double *H, long ldh, double *inv,
                            hipStream_t s);
hipError_t atlmx_k_ctri_inv(int lower, int n, const float *H, long ldh, float *inv,
                            hipStream_t s);
/* Unblocked complex LU (ATL_getf2's column loop) of an m x n leaf, n <=
 * ATLMX_CLEAF_MAXW, whose top row is global row row0: ipiv[j] receives
 * row0 + the pivot row (and, when ipiv2 is not NULL, ipiv2[2j + s] = 2 (row0
 * + pivot row) + s), *info = row0 + j + 1 at the first zero pivot (when
 * *info is 0).  One workgroup, the rows in registers where they fit. */
#define ATLMX_CLEAF_MAXW 16
hipError_t atlmx_k_zgetf2(int m, int n, double *A, long lda, int row0, int *ipiv, int *ipiv2,
                          int *info, hipStream_t s);
hipError_t atlmx_k_cgetf2(int m, int n, float *A, long lda, int row0, int *ipiv, int *ipiv2,
                          int *info, hipStream_t s);
/* B := inv(L) B, L unit lower of order n1 <= 64 (its strict lower triangle
 * read), B n1 x nr, in place; one launch. */
hipError_t atlmx_k_ztrsm_llnu(int n1, int nr, const double *L, long lda, double *B, long ldb,
                              hipStream_t s);
hipError_t atlmx_k_ctrsm_llnu(int n1, int nr, const float *L, long lda, float *B, long ldb,
                              hipStream_t s);
/* Cholesky of one n x n (n <= 64) Hermitian diagonal block in place, lower
 * (A = L L^H) or upper (A = U^H U); the real part of the diagonal is the
 * pivot.  A non-positive pivot in column j sets *info = base + j + 1 (when
 * *info is 0) and stops.  One workgroup. */
hipError_t atlmx_k_zpotf2(int upper, int n, double *A, long lda, int base, int *info,
                          hipStream_t s);
hipError_t atlmx_k_cpotf2(int upper, int n, float *A, long lda, int base, int *info,
                          hipStream_t s);
/* beta * C on the lower (lower = 1) or upper triangle of an n x n complex C,
 * beta = 0 storing zeros; herm zeroes the diagonal's imaginary parts (with
 * beta = 1 that is all it does) */
hipError_t atlmx_k_ztri_scale(int n, int lower, double br, double bi, int herm, double *C,
                              long ldc, hipStream_t s);
hipError_t atlmx_k_ctri_scale(int n, int lower, float br, float bi, int herm, float *C,
                              long ldc, hipStream_t s);
/* out = the full n x n Hermitian (herm) or symmetric matrix held in A's
 * lower (lower = 1) or upper triangle */
hipError_t atlmx_k_zsym_expand(int n, int lower, int herm, const double *A, long lda,
                               double *out, long ldo, hipStream_t s);
hipError_t atlmx_k_csym_expand(int n, int lower, int herm, const float *A, long lda, float *out,
                               long ldo, hipStream_t s);
/* C = conj(C) in place, m x n */
hipError_t atlmx_k_zconj(int m, int n, double *C, long ldc, hipStream_t s);
hipError_t atlmx_k_cconj(int m, int n, float *C, long ldc, hipStream_t s);
/* ipiv2[2k + s] = 2 ipiv[k] + s, k in [k1, k2): a complex matrix's row
 * interchanges as those of its interleaved real view */
hipError_t atlmx_k_cpiv2(int k1, int k2, const int *ipiv, int *ipiv2, hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif

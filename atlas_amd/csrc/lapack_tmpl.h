/*
 * lapack_tmpl.h -- precision template of the LAPACK layer: blocked
 * right-looking LU and Cholesky on the GPU, getrs/potrs/gesv/posv, and the
 * clapack C ABI.  Included by lapack.c with T=double/PFX=d and
 * T=float/PFX=s.
 */
#define GETRF_DEV NM(PFX, getrf_dev)
#define GETRS_DEV NM(PFX, getrs_dev)
#define POTRF_DEV NM(PFX, potrf_dev)
#define POTRS_DEV NM(PFX, potrs_dev)
#define PANEL_REC NM(PFX, panel_rec)
#define GETRF_UPDATE NM(PFX, getrf_update)
#define LASWP_ROWS NM(PFX, laswp_rows)

/* Width of the outer LU block (the K of the trailing GEMM) and of the
 * cooperative panel leaf. */
#define LU_NB 128
#define LU_PW ATLMX_PANEL_MAXW
#define CHOL_NB ATLMX_CHOL_NB

/* Forward interchanges k1..k2 applied to columns [c0, c1) minus [s0, s1):
 * as one move list for the whole column range when it fits the plan
 * kernel, else per-column swaps. */
static void LASWP_ROWS(atlmx_ctx *c, T *A, long lda, int M, int k1, int k2, const int *ipiv,
                       int c0, int c1, int s0, int s1, const char *rout)
{
  int pid;
  if (k2 <= k1 || c1 <= c0) return;
  pid = rt_prof_begin(c, PROF_LASWP, 0.0);
  if (k2 - k1 <= 2 * ATLMX_PANEL_MAXW && M - k1 <= ATLMX_LASWP_PLAN_MAXROWS) {
    int *plan = (int *)rt_ws(c, WS_PLAN, ATLMX_LASWP_PLAN_INTS * sizeof(int), rout);
    RT_CHECK(rout, atlmx_k_laswp_plan(k1, k2, M, ipiv, plan, c->stream));
    RT_CHECK(rout, K_LASWP_APPLY(A, lda, c0, c1, s0, s1, plan, c->stream));
  } else {
    const int e0 = (s0 < c1) ? s0 : c1, b1 = (s1 > c0) ? s1 : c0;
    if (e0 > c0)
      RT_CHECK(rout, K_LASWP(e0 - c0, A + (size_t)c0 * lda, lda, k1, k2, ipiv, 1, M,
                             c->stream));
    if (c1 > b1)
      RT_CHECK(rout, K_LASWP(c1 - b1, A + (size_t)b1 * lda, lda, k1, k2, ipiv, 1, M,
                             c->stream));
  }
  rt_prof_end(c, pid);
}

/* Factor the tall panel A[r0:M, r0:r0+w] with partial pivoting.
 * Splits like src/lapack/ATL_getrfC.c:277-300 (left half, laswp + TRSM +
 * GEMM on the right half, right half, back-swap of the left half) down to
 * 32-wide leaves, which run as one cooperative kernel each (k_lu.hip). */
static void PANEL_REC(atlmx_ctx *c, int M, T *A, long lda, int r0, int w, int *ipiv,
                      int *info, void *scratch, const char *rout)
{
  const int leaf = (M - r0 > ATLMX_PANEL_TALL) ? LU_PW / 2 : LU_PW;
  if (w <= leaf) {
    int pid;
    if (++c->panel_set >= ATLMX_PANEL_SEQ_MAX) {     /* restart the exchange tags */
      RT_CHECK(rout, atlmx_k_panel_scratch_init(scratch, c->stream));
      c->panel_set = 1;
    }
    pid = rt_prof_begin(c, PROF_PANEL, (double)(M - r0) * w * 2.0 * sizeof(T));
    RT_CHECK(rout, K_PANEL(M - r0, w, A + r0 + (size_t)r0 * lda, lda, r0, ipiv + r0, info,
                           scratch, c->panel_set, c->stream));
    rt_prof_end(c, pid);
    return;
  } else {
    int wl = (w / 2 / 32) * 32;
    if (wl < 32) wl = 32;
    if (wl > leaf) wl = leaf;
    PANEL_REC(c, M, A, lda, r0, wl, ipiv, info, scratch, rout);
    LASWP_ROWS(c, A, lda, M, r0, r0 + wl, ipiv, r0 + wl, r0 + w, 0, 0, rout);
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_LOWER, KX_NOTRANS, KX_UNIT, wl, w - wl, (T)1,
                      A + r0 + (size_t)r0 * lda, lda, A + r0 + (size_t)(r0 + wl) * lda, lda,
                      NULL, rout);
    if (r0 + wl < M)
      NM(PFX, gemm_dev)(c, KX_NOTRANS, KX_NOTRANS, M - r0 - wl, w - wl, wl, (T)-1,
                        A + (r0 + wl) + (size_t)r0 * lda, lda,
                        A + r0 + (size_t)(r0 + wl) * lda, lda, (T)1,
                        A + (r0 + wl) + (size_t)(r0 + wl) * lda, lda, 0, NULL, rout);
    PANEL_REC(c, M, A, lda, r0 + wl, w - wl, ipiv, info, scratch, rout);
    LASWP_ROWS(c, A, lda, M, r0 + wl, r0 + w, ipiv, r0, r0 + wl, 0, 0, rout);
  }
}

/* Blocked right-looking LU, A = P*L*U in place (column-major M x N), with
 * one panel of lookahead: after panel j's interchanges, the next panel's
 * columns get panel j's update first and that panel is factored on a second
 * stream (its own workspace bank) while the rest of the trailing matrix gets
 * the update on the thread's stream.  The panels still run one after
 * another, so the pivots are those of the plain right-looking order.
 * d_ipiv: device, min(M,N) absolute 0-based pivot rows (clapack
 * convention, src/lapack/ATL_getrfC.c:299,325).  Returns info (1-based
 * first exactly-zero pivot; the factorization still completes,
 * interfaces/lapack/C/src/clapack_dgetrf.c:35-45).  Synchronizes. */
static void GETRF_UPDATE(atlmx_ctx *c, int M, T *A, long lda, int j, int jb, int c0, int nc,
                         const char *rout)
{
  const int jn = j + jb;
  if (nc <= 0) return;
  NM(PFX, trsm_dev)(c, KX_LEFT, KX_LOWER, KX_NOTRANS, KX_UNIT, jb, nc, (T)1,
                    A + j + (size_t)j * lda, lda, A + j + (size_t)c0 * lda, lda, NULL, rout);
  if (jn < M)
    NM(PFX, update_dev)(c, KX_NOTRANS, M - jn, nc, jb, (T)-1, A + jn + (size_t)j * lda, lda,
                        A + j + (size_t)c0 * lda, lda, A + jn + (size_t)c0 * lda, lda, 0, NULL,
                        rout);
}

int GETRF_DEV(atlmx_ctx *c, int M, int N, T *A, long lda, int *d_ipiv, const char *rout)
{
  const int MN = M < N ? M : N;
  int j, info = 0;
  int *d_info;
  void *scratch;
  hipStream_t main = c->stream, aux;
  if (MN <= 0) return 0;
  d_info = (int *)rt_ws(c, WS_INFO, 256, rout);
  scratch = rt_ws(c, WS_PANEL, atlmx_k_panel_scratch_bytes(M, LU_PW), rout);
  if (c->panel_buf != scratch) {
    RT_CHECK(rout, atlmx_k_panel_scratch_init(scratch, c->stream));
    c->panel_buf = scratch;
    c->panel_set = 0;
  }
  aux = rt_aux(c, rout);
  RT_CHECK(rout, atlmx_k_fill_i32(d_info, 1, 0, c->stream));
  PANEL_REC(c, M, A, lda, 0, (MN < LU_NB) ? MN : LU_NB, d_ipiv, d_info, scratch, rout);
  for (j = 0; j < MN; j += LU_NB) {
    const int jb = (MN - j < LU_NB) ? MN - j : LU_NB;
    const int jn = j + jb;
    const int jnb = (jn < MN) ? ((MN - jn < LU_NB) ? MN - jn : LU_NB) : 0;
    LASWP_ROWS(c, A, lda, M, j, jn, d_ipiv, 0, N, j, jn, rout);
    if (jn >= N) break;
    if (jnb > 0) {
      GETRF_UPDATE(c, M, A, lda, j, jb, jn, jnb, rout);
      RT_CHECK(rout, hipEventRecord(c->ev[0], main));
      RT_CHECK(rout, hipStreamWaitEvent(aux, c->ev[0], 0));
      rt_on_aux(c, main, 1);
      PANEL_REC(c, M, A, lda, jn, jnb, d_ipiv, d_info, scratch, rout);
      rt_on_aux(c, main, 0);
      RT_CHECK(rout, hipEventRecord(c->ev[1], aux));
    }
    GETRF_UPDATE(c, M, A, lda, j, jb, jn + jnb, N - jn - jnb, rout);
    if (jnb > 0) RT_CHECK(rout, hipStreamWaitEvent(main, c->ev[1], 0));
  }
  RT_CHECK(rout, hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost,
                                c->stream));
  rt_sync(c, rout);
  return info;
}

/* The row interchanges of getrs on the right-hand sides (ATL_getrs.c's
 * ATL_laswp calls).  Replayed one by one they are a chain of N dependent
 * swaps; here, for up to ATLMX_LASWP_PERM_MAXROWS rows, they become one
 * permutation computed in parallel and a gather of B's rows (through a
 * workspace, at most 1024 columns at a time).  The allocation-free solve
 * (with caller workspaces) keeps the replay; both move the same values. */
static void NM(PFX, laswp_rhs)(atlmx_ctx *c, int NRHS, T *B, long ldb, int N,
                               const int *d_ipiv, int inc, int ws_given, const char *rout)
{
  if (!ws_given && N <= ATLMX_LASWP_PERM_MAXROWS) {
    const int cw = NRHS < 1024 ? NRHS : 1024;
    int *src = (int *)rt_ws(c, WS_LASWP, 2 * (size_t)N * sizeof(int), rout);
    T *tmp = (T *)rt_ws(c, WS_LASWP_B, (size_t)N * cw * sizeof(T), rout);
    int q;
    RT_CHECK(rout, atlmx_k_laswp_perm(0, N, N, d_ipiv, inc, src, src + N, c->stream));
    for (q = 0; q < NRHS; q += cw) {
      const int nq = NRHS - q < cw ? NRHS - q : cw;
      RT_CHECK(rout, K_PERMUTE_ROWS(nq, B + (size_t)q * ldb, ldb, 0, N, src, tmp, c->stream));
    }
  } else {
    RT_CHECK(rout, K_LASWP(NRHS, B, ldb, 0, N, d_ipiv, inc, N, c->stream));
  }
}

/* src/lapack/ATL_getrs.c:8-89 (ColMajor).  NoTrans: P*B, L-unit solve, U
 * solve; Trans: U^T solve, L^T-unit solve, P^T*B.  inv / ctl: NULL, or the
 * two TRSMs' workspaces (see trsm_dev_ws; the solves run one after the
 * other on the stream, so they share them). */
void NM(PFX, getrs_dev_ws)(atlmx_ctx *c, int trans, int N, int NRHS, const T *A, long lda,
                           const int *d_ipiv, T *B, long ldb, T *inv, int *ctl,
                           const char *rout)
{
  if (N <= 0 || NRHS <= 0) return;
  if (trans == KX_NOTRANS) {
    NM(PFX, laswp_rhs)(c, NRHS, B, ldb, N, d_ipiv, 1, inv != NULL, rout);
    NM(PFX, trsm_dev_ws)(c, KX_LEFT, KX_LOWER, KX_NOTRANS, KX_UNIT, N, NRHS, (T)1, A, lda, B,
                         ldb, NULL, inv, ctl, rout);
    NM(PFX, trsm_dev_ws)(c, KX_LEFT, KX_UPPER, KX_NOTRANS, KX_NONUNIT, N, NRHS, (T)1, A, lda,
                         B, ldb, NULL, inv, ctl, rout);
  } else {
    NM(PFX, trsm_dev_ws)(c, KX_LEFT, KX_UPPER, KX_TRANS, KX_NONUNIT, N, NRHS, (T)1, A, lda, B,
                         ldb, NULL, inv, ctl, rout);
    NM(PFX, trsm_dev_ws)(c, KX_LEFT, KX_LOWER, KX_TRANS, KX_UNIT, N, NRHS, (T)1, A, lda, B,
                         ldb, NULL, inv, ctl, rout);
    NM(PFX, laswp_rhs)(c, NRHS, B, ldb, N, d_ipiv, -1, inv != NULL, rout);
  }
}

void GETRS_DEV(atlmx_ctx *c, int trans, int N, int NRHS, const T *A, long lda,
               const int *d_ipiv, T *B, long ldb, const char *rout)
{
  NM(PFX, getrs_dev_ws)(c, trans, N, NRHS, A, lda, d_ipiv, B, ldb, NULL, NULL, rout);
}

/* Blocked right-looking Cholesky (Lower: A = L*L^T, Upper: A = U^T*U).
 * Replaces the recursion of src/lapack/ATL_potrfL.c:204-262 /
 * ATL_potrfU.c:199-257: diagonal block in registers (k_chol.hip), the
 * panel's TRSM in one launch (k_chol.hip's panel kernel on the block's
 * inverted 64 x 64 diagonal blocks), SYRK trailing update on the GEMM kernel
 * restricted to one triangle.  One block of lookahead: step j first
 * updates only the next block column (block row for Upper) from panel j,
 * then factors that block and solves its panel on the second stream while
 * the main stream applies panel j to the rest of the trailing matrix.
 * Every block sees the same updates in the same order as without
 * lookahead, so the result is the same.  A non-positive pivot stops the factorization: the device info flag
 * makes every queued kernel after it return at once.  Synchronizes;
 * returns info. */
static void NM(PFX, potrf_panel)(atlmx_ctx *c, int uplo, int N, T *A, long lda, int j,
                                 int jb, int *d_info, const char *rout)
{
  const int nr = N - j - jb;
  T *Ad = A + j + (size_t)j * lda;
  int pid = rt_prof_begin(c, PROF_POTF2, 0.0);
  RT_CHECK(rout, K_POTF2(uplo, jb, Ad, lda, j, d_info, c->stream));
  rt_prof_end(c, pid);
  if (nr <= 0) return;
  {
    /* the panel's TRSM (Lower: X L^T = A21; Upper: U^T X = A12) in one
     * launch after the diagonal block's 64 x 64 inverses */
    const int nbi = (jb + ATLMX_TRI_NB - 1) / ATLMX_TRI_NB;
    T *inv = (T *)rt_ws(c, WS_TRI_INV, (size_t)nbi * ATLMX_TRI_NB * ATLMX_TRI_NB * sizeof(T),
                        rout);
    T *P = (uplo == KX_LOWER) ? A + (j + jb) + (size_t)j * lda : A + j + (size_t)(j + jb) * lda;
    pid = rt_prof_begin(c, PROF_TRSM_INV, (double)jb * ATLMX_TRI_NB * sizeof(T));
    RT_CHECK(rout, K_TRI_INV(uplo, KX_NONUNIT, jb, Ad, lda, inv, d_info, c->stream));
    rt_prof_end(c, pid);
    pid = rt_prof_begin(c, PROF_MISC, (double)nr * jb * jb);
    RT_CHECK(rout, K_CHOL_PANEL(uplo, nr, jb, Ad, lda, inv, P, d_info, c->stream));
    rt_prof_end(c, pid);
  }
}

/* Trailing update from panel j (columns/rows j..j+jb) of the part of the
 * trailing matrix that starts at diagonal index r0 and is nc wide (Lower:
 * rows r0..N-1 of columns r0..r0+nc-1; Upper: the transposed block row).
 * The GEMM's triangle mask is in the block's own coordinates, which start
 * on the diagonal, so only the referenced triangle is written. */
static void NM(PFX, potrf_update)(atlmx_ctx *c, int uplo, int N, T *A, long lda, int j,
                                  int jb, int r0, int nc, const int *d_info, const char *rout)
{
  if (nc <= 0) return;
  if (uplo == KX_LOWER)
    NM(PFX, update_dev)(c, KX_TRANS, N - r0, nc, jb, (T)-1, A + r0 + (size_t)j * lda, lda,
                        A + r0 + (size_t)j * lda, lda, A + r0 + (size_t)r0 * lda, lda,
                        KX_LOWER, d_info, rout);
  else
    NM(PFX, gemm_dev)(c, KX_TRANS, KX_NOTRANS, nc, N - r0, jb, (T)-1,
                      A + j + (size_t)r0 * lda, lda, A + j + (size_t)r0 * lda, lda, (T)1,
                      A + r0 + (size_t)r0 * lda, lda, KX_UPPER, d_info, rout);
}

int POTRF_DEV(atlmx_ctx *c, int uplo, int N, T *A, long lda, const char *rout)
{
  int j, info = 0;
  int *d_info;
  hipStream_t main = c->stream, aux;
  if (N <= 0) return 0;
  d_info = (int *)rt_ws(c, WS_INFO, 256, rout);
  aux = rt_aux(c, rout);
  RT_CHECK(rout, atlmx_k_fill_i32(d_info, 1, 0, c->stream));
  NM(PFX, potrf_panel)(c, uplo, N, A, lda, 0, (N < CHOL_NB) ? N : CHOL_NB, d_info, rout);
  for (j = 0; j < N; j += CHOL_NB) {
    const int jb = (N - j < CHOL_NB) ? N - j : CHOL_NB;
    const int jn = j + jb;
    const int jnb = (jn < N) ? ((N - jn < CHOL_NB) ? N - jn : CHOL_NB) : 0;
    if (jnb <= 0) break;
    NM(PFX, potrf_update)(c, uplo, N, A, lda, j, jb, jn, jnb, d_info, rout);
    RT_CHECK(rout, hipEventRecord(c->ev[0], main));
    RT_CHECK(rout, hipStreamWaitEvent(aux, c->ev[0], 0));
    rt_on_aux(c, main, 1);
    NM(PFX, potrf_panel)(c, uplo, N, A, lda, jn, jnb, d_info, rout);
    rt_on_aux(c, main, 0);
    RT_CHECK(rout, hipEventRecord(c->ev[1], aux));
    NM(PFX, potrf_update)(c, uplo, N, A, lda, j, jb, jn + jnb, N - jn - jnb, d_info, rout);
    RT_CHECK(rout, hipStreamWaitEvent(main, c->ev[1], 0));
  }
  RT_CHECK(rout, hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost,
                                c->stream));
  rt_sync(c, rout);
  return info;
}

/* src/lapack/ATL_potrs.c:13-86 (ColMajor) */
void POTRS_DEV(atlmx_ctx *c, int uplo, int N, int NRHS, const T *A, long lda, T *B,
               long ldb, const char *rout)
{
  if (N <= 0 || NRHS <= 0) return;
  if (uplo == KX_UPPER) {
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_UPPER, KX_TRANS, KX_NONUNIT, N, NRHS, (T)1, A, lda, B,
                      ldb, NULL, rout);
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_UPPER, KX_NOTRANS, KX_NONUNIT, N, NRHS, (T)1, A, lda,
                      B, ldb, NULL, rout);
  } else {
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_LOWER, KX_NOTRANS, KX_NONUNIT, N, NRHS, (T)1, A, lda,
                      B, ldb, NULL, rout);
    NM(PFX, trsm_dev)(c, KX_LEFT, KX_LOWER, KX_TRANS, KX_NONUNIT, N, NRHS, (T)1, A, lda, B,
                      ldb, NULL, rout);
  }
}

/* ------------------------------------------------------------------------
 * Inverses.  TRTRI (src/lapack/ATL_trtri.c:8-42): info = first exactly-zero
 * diagonal entry when non-unit (A untouched then), else A's triangle is
 * replaced by its inverse -- computed as op(A) X = I on the TRSM (inverted
 * 64x64 diagonal blocks + MFMA GEMM updates) into a dense workspace, whose
 * triangle is stored back (the diagonal left alone when unit).  LAUUM
 * (ATL_lauum.c:6-22): U U^T (upper) or L^T L (lower) into the same
 * triangle, as one triangle-masked GEMM of the expanded triangle with
 * itself.  POTRI = TRTRI + LAUUM (interfaces/lapack/C/src/clapack_dpotri.c).
 * GETRI (ATL_getriC.c): info from U's diagonal, else A^{-1} = the solve of
 * P L U X = I (getrs on the factors) written over the factors.
 * Each synchronizes and returns info.
 * ---------------------------------------------------------------------- */
static int NM(PFX, diag_info)(atlmx_ctx *c, int N, const T *A, long lda, const char *rout)
{
  int info = 0;
  int *d_info = (int *)rt_ws(c, WS_INFO, 256, rout);
  RT_CHECK(rout, K_DIAG_ZERO(N, A, lda, d_info, c->stream));
  RT_CHECK(rout, hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost,
                                c->stream));
  rt_sync(c, rout);
  return info;
}

int NM(PFX, trtri_dev)(atlmx_ctx *c, int uplo, int diag, int N, T *A, long lda,
                       const char *rout)
{
  T *W;
  int info;
  if (N <= 0) return 0;
  if (diag == KX_NONUNIT && (info = NM(PFX, diag_info)(c, N, A, lda, rout)) != 0) return info;
  W = (T *)rt_ws(c, WS_L3A, (size_t)N * N * sizeof(T), rout);
  RT_CHECK(rout, K_TRICOPY(ATLMX_TC_IDENTITY, uplo, diag, N, N, NULL, N, W, N, c->stream));
  NM(PFX, trsm_dev)(c, KX_LEFT, uplo, KX_NOTRANS, diag, N, N, (T)1, A, lda, W, N, NULL, rout);
  RT_CHECK(rout, K_TRICOPY(ATLMX_TC_STORE_TRI, uplo, diag, N, N, A, lda, W, N, c->stream));
  rt_sync(c, rout);
  return 0;
}

void NM(PFX, lauum_dev)(atlmx_ctx *c, int uplo, int N, T *A, long lda, const char *rout)
{
  T *W;
  if (N <= 0) return;
  W = (T *)rt_ws(c, WS_L3B, (size_t)N * N * sizeof(T), rout);
  RT_CHECK(rout, K_TRICOPY(ATLMX_TC_EXPAND_TRI, uplo, KX_NONUNIT, N, N, A, lda, W, N,
                           c->stream));
  if (uplo == KX_LOWER)
    NM(PFX, gemm_dev)(c, KX_TRANS, KX_NOTRANS, N, N, N, (T)1, W, N, W, N, (T)0, A, lda,
                      KX_LOWER, NULL, rout);
  else
    NM(PFX, gemm_dev)(c, KX_NOTRANS, KX_TRANS, N, N, N, (T)1, W, N, W, N, (T)0, A, lda,
                      KX_UPPER, NULL, rout);
  rt_sync(c, rout);
}

int NM(PFX, getri_dev)(atlmx_ctx *c, int N, T *A, long lda, const int *d_ipiv,
                       const char *rout)
{
  T *W;
  int info;
  if (N <= 0) return 0;
  if ((info = NM(PFX, diag_info)(c, N, A, lda, rout)) != 0) return info;
  W = (T *)rt_ws(c, WS_L3A, (size_t)N * N * sizeof(T), rout);
  RT_CHECK(rout, K_TRICOPY(ATLMX_TC_IDENTITY, 0, 0, N, N, NULL, N, W, N, c->stream));
  GETRS_DEV(c, KX_NOTRANS, N, N, A, lda, d_ipiv, W, N, rout);
  RT_CHECK(rout, K_TRICOPY(ATLMX_TC_COPY, 0, 0, N, N, A, lda, W, N, c->stream));
  rt_sync(c, rout);
  return 0;
}

/* ------------------------------------------------------------------------
 * QR (src/lapack/ATL_geqrf.c): blocked Householder QR with 32-wide panels.
 * Each panel is factored column by column (k_qr.hip: larfg and the
 * reflector's application within the panel), then the block reflector
 * H = I - V T V^T (T from larft on V^T V) is applied to the trailing
 * columns as C -= V (T^T (V^T C)) -- three MFMA GEMMs.  Same storage as
 * LAPACK's geqrf: R on and above the diagonal, the reflectors below,
 * tau[0..min(M,N)).
 * ---------------------------------------------------------------------- */
void NM(PFX, geqrf_dev)(atlmx_ctx *c, int M, int N, T *A, long lda, T *d_tau,
                        const char *rout)
{
  const int K = M < N ? M : N, nb = ATLMX_QR_NB;
  double *work;
  T *V, *G, *Tm, *W, *W2;
  int j;
  if (K <= 0) return;
  work = (double *)rt_ws(c, WS_QR, atlmx_k_qr_work_doubles() * sizeof(double), rout);
  V = (T *)rt_ws(c, WS_QRV, ((size_t)M * nb + 2 * (size_t)nb * nb + 2 * (size_t)nb * N) *
                            sizeof(T), rout);
  G = V + (size_t)M * nb;
  Tm = G + (size_t)nb * nb;
  W = Tm + (size_t)nb * nb;
  W2 = W + (size_t)nb * N;
  for (j = 0; j < K; j += nb) {
    const int jb = (K - j < nb) ? K - j : nb, mr = M - j, nc = N - j - jb;
    T *P = A + j + (size_t)j * lda, *Cp = A + j + (size_t)(j + jb) * lda;
    RT_CHECK(rout, K_GEQR2(mr, jb, P, lda, d_tau + j, work, c->stream));
    if (nc <= 0) continue;
    RT_CHECK(rout, K_QR_VCOPY(mr, jb, P, lda, V, c->stream));
    NM(PFX, gemm_dev)(c, KX_TRANS, KX_NOTRANS, jb, jb, mr, (T)1, V, mr, V, mr, (T)0, G, jb, 0,
                      NULL, rout);
    RT_CHECK(rout, K_LARFT(jb, G, d_tau + j, Tm, c->stream));
    NM(PFX, gemm_dev)(c, KX_TRANS, KX_NOTRANS, jb, nc, mr, (T)1, V, mr, Cp, lda, (T)0, W, jb,
                      0, NULL, rout);
    NM(PFX, gemm_dev)(c, KX_TRANS, KX_NOTRANS, jb, nc, jb, (T)1, Tm, jb, W, jb, (T)0, W2, jb, 0,
                      NULL, rout);
    NM(PFX, gemm_dev)(c, KX_NOTRANS, KX_NOTRANS, mr, nc, jb, (T)-1, V, mr, W2, jb, (T)1, Cp,
                      lda, 0, NULL, rout);
  }
}

/* ====================================================================== */
/* public device-pointer API                                                */
/* ====================================================================== */
int NM3(atlmx_, PFX, getrf)(int M, int N, T *A, int lda, int *d_ipiv)
{
  return GETRF_DEV(rt_ctx(), M, N, A, lda, d_ipiv, "atlmx_getrf");
}
void NM3(atlmx_, PFX, getrs)(int trans, int N, int NRHS, const T *A, int lda,
                             const int *d_ipiv, T *B, int ldb)
{
  GETRS_DEV(rt_ctx(), trans == KX_NOTRANS ? KX_NOTRANS : KX_TRANS, N, NRHS, A, lda, d_ipiv,
            B, ldb, "atlmx_getrs");
}
int NM3(atlmx_, PFX, potrf)(int uplo, int N, T *A, int lda)
{
  return POTRF_DEV(rt_ctx(), uplo, N, A, lda, "atlmx_potrf");
}
void NM3(atlmx_, PFX, potrs)(int uplo, int N, int NRHS, const T *A, int lda, T *B, int ldb)
{
  POTRS_DEV(rt_ctx(), uplo, N, NRHS, A, lda, B, ldb, "atlmx_potrs");
}

/* ====================================================================== */
/* ATLAS clapack C ABI                                                      */
/* ====================================================================== */

/* interfaces/lapack/C/src/clapack_dgetrf.c:13-88; RowMajor is the
 * column-major factorization of the transposed view (src/lapack/
 * ATL_getrf.c:12, ATL_getrfR.c). */
int NM3(clapack_, PFX, getrf)(const enum CBLAS_ORDER Order, const int M, const int N,
                              T *A, const int lda, int *ipiv)
{
  static const char *rout = "clapack_" STR(PFX) "getrf";
  const int ierr = atlmx_getrf_check(rout, Order, M, N, lda);
  int m, n, mn, info;
  atlmx_ctx *c;
  long dla;
  T *dA;
  int *dp;
  if (ierr) return ierr;
  m = (Order == CblasColMajor) ? M : N;
  n = (Order == CblasColMajor) ? N : M;
  mn = m < n ? m : n;
  if (mn == 0) return 0;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, m, n, lda, sizeof(T), 1, &dla, rout);
  dp = rt_is_device(ipiv) ? ipiv : (int *)rt_ws(c, WS_IPIV, (size_t)mn * sizeof(int), rout);
  info = GETRF_DEV(c, m, n, dA, dla, dp, rout);
  rt_stage_out(c, A, dA, m, n, lda, dla, sizeof(T), rout);
  if (dp != ipiv)
    RT_CHECK(rout, hipMemcpyAsync(ipiv, dp, (size_t)mn * sizeof(int), hipMemcpyDeviceToHost,
                                  c->stream));
  rt_sync(c, rout);
  return info;
}

static const int *NM(PFX, stage_ipiv)(atlmx_ctx *c, const int *ipiv, int n, const char *rout)
{
  int *dp;
  if (rt_is_device(ipiv)) return ipiv;
  dp = (int *)rt_ws(c, WS_IPIV, (size_t)(n > 0 ? n : 1) * sizeof(int), rout);
  RT_CHECK(rout, hipMemcpyAsync(dp, ipiv, (size_t)n * sizeof(int), hipMemcpyHostToDevice,
                                c->stream));
  return dp;
}

/* interfaces/lapack/C/src/clapack_dgetrs.c; RowMajor holds the factor of
 * A^T, so it is the column-major solve with Trans flipped
 * (src/lapack/ATL_getrs.c:60-88 spells the same with Right-side TRSMs). */
int NM3(clapack_, PFX, getrs)(const enum CBLAS_ORDER Order, const enum CBLAS_TRANSPOSE Trans,
                              const int N, const int NRHS, const T *A, const int lda,
                              const int *ipiv, T *B, const int ldb)
{
  static const char *rout = "clapack_" STR(PFX) "getrs";
  const int ierr = atlmx_getrs_check(rout, Order, Trans, N, NRHS, lda, ldb);
  int t;
  atlmx_ctx *c;
  long dla, dlb;
  T *dA, *dB;
  if (ierr) return ierr;
  if (N == 0 || NRHS == 0) return 0;
  t = (Trans != CblasNoTrans);
  if (Order == CblasRowMajor) t = !t;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, sizeof(T), 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, N, NRHS, ldb, sizeof(T), 1, &dlb, rout);
  GETRS_DEV(c, t ? KX_TRANS : KX_NOTRANS, N, NRHS, dA, dla,
            NM(PFX, stage_ipiv)(c, ipiv, N, rout), dB, dlb, rout);
  rt_stage_out(c, B, dB, N, NRHS, ldb, dlb, sizeof(T), rout);
  rt_sync(c, rout);
  return 0;
}

/* interfaces/lapack/C/src/clapack_dpotrf.c; RowMajor swaps Upper and
 * Lower (src/lapack/ATL_potrf.c:21-26). */
int NM3(clapack_, PFX, potrf)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                              const int N, T *A, const int lda)
{
  static const char *rout = "clapack_" STR(PFX) "potrf";
  const int ierr = atlmx_potrf_check(rout, Order, Uplo, N, lda);
  int low, info;
  atlmx_ctx *c;
  long dla;
  T *dA;
  if (ierr) return ierr;
  if (N == 0) return 0;
  low = (Uplo == CblasLower);
  if (Order == CblasRowMajor) low = !low;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, sizeof(T), 1, &dla, rout);
  info = POTRF_DEV(c, low ? KX_LOWER : KX_UPPER, N, dA, dla, rout);
  rt_stage_out(c, A, dA, N, N, lda, dla, sizeof(T), rout);
  rt_sync(c, rout);
  return info;
}

/* interfaces/lapack/C/src/clapack_dpotrs.c */
int NM3(clapack_, PFX, potrs)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                              const int N, const int NRHS, const T *A, const int lda, T *B,
                              const int ldb)
{
  static const char *rout = "clapack_" STR(PFX) "potrs";
  const int ierr = atlmx_posv_check(rout, Order, Uplo, N, NRHS, lda, ldb);
  int low;
  atlmx_ctx *c;
  long dla, dlb;
  T *dA, *dB;
  if (ierr) return ierr;
  if (N == 0 || NRHS == 0) return 0;
  low = (Uplo == CblasLower);
  if (Order == CblasRowMajor) low = !low;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, sizeof(T), 1, &dla, rout);
  dB = (T *)rt_stage_in(c, WS_STAGE_B, B, N, NRHS, ldb, sizeof(T), 1, &dlb, rout);
  POTRS_DEV(c, low ? KX_LOWER : KX_UPPER, N, NRHS, dA, dla, dB, dlb, rout);
  rt_stage_out(c, B, dB, N, NRHS, ldb, dlb, sizeof(T), rout);
  rt_sync(c, rout);
  return 0;
}

/* interfaces/lapack/C/src/clapack_dgesv.c:29-62 */
int NM3(clapack_, PFX, gesv)(const enum CBLAS_ORDER Order, const int N, const int NRHS,
                             T *A, const int lda, int *ipiv, T *B, const int ldb)
{
  static const char *rout = "clapack_" STR(PFX) "gesv";
  const int ierr = atlmx_gesv_check(rout, Order, N, NRHS, lda, ldb);
  int info;
  atlmx_ctx *c;
  long dla, dlb;
  T *dA, *dB;
  int *dp;
  if (ierr) return ierr;
  if (N == 0) return 0;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, sizeof(T), 1, &dla, rout);
  dp = rt_is_device(ipiv) ? ipiv : (int *)rt_ws(c, WS_IPIV, (size_t)N * sizeof(int), rout);
  info = GETRF_DEV(c, N, N, dA, dla, dp, rout);
  if (!info && NRHS > 0) {
    dB = (T *)rt_stage_in(c, WS_STAGE_B, B, N, NRHS, ldb, sizeof(T), 1, &dlb, rout);
    GETRS_DEV(c, (Order == CblasRowMajor) ? KX_TRANS : KX_NOTRANS, N, NRHS, dA, dla, dp,
              dB, dlb, rout);
    rt_stage_out(c, B, dB, N, NRHS, ldb, dlb, sizeof(T), rout);
  }
  rt_stage_out(c, A, dA, N, N, lda, dla, sizeof(T), rout);
  if (dp != ipiv)
    RT_CHECK(rout, hipMemcpyAsync(ipiv, dp, (size_t)N * sizeof(int), hipMemcpyDeviceToHost,
                                  c->stream));
  rt_sync(c, rout);
  return info;
}

/* interfaces/lapack/C/src/clapack_dposv.c:19-58 */
int NM3(clapack_, PFX, posv)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                             const int N, const int NRHS, T *A, const int lda, T *B,
                             const int ldb)
{
  static const char *rout = "clapack_" STR(PFX) "posv";
  const int ierr = atlmx_posv_check(rout, Order, Uplo, N, NRHS, lda, ldb);
  int info, low;
  atlmx_ctx *c;
  long dla, dlb;
  T *dA, *dB;
  if (ierr) return ierr;
  if (N == 0) return 0;
  low = (Uplo == CblasLower);
  if (Order == CblasRowMajor) low = !low;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, sizeof(T), 1, &dla, rout);
  info = POTRF_DEV(c, low ? KX_LOWER : KX_UPPER, N, dA, dla, rout);
  if (!info && NRHS > 0) {
    dB = (T *)rt_stage_in(c, WS_STAGE_B, B, N, NRHS, ldb, sizeof(T), 1, &dlb, rout);
    POTRS_DEV(c, low ? KX_LOWER : KX_UPPER, N, NRHS, dA, dla, dB, dlb, rout);
    rt_stage_out(c, B, dB, N, NRHS, ldb, dlb, sizeof(T), rout);
  }
  rt_stage_out(c, A, dA, N, N, lda, dla, sizeof(T), rout);
  rt_sync(c, rout);
  return info;
}

/* Order / Uplo / N / lda checks of the inverse routines' C wrappers
 * (interfaces/lapack/C/src/clapack_dpotri.c, clapack_dlauum.c,
 * clapack_dtrtri.c); the position of lda follows the Diag argument when
 * there is one. */
static int NM(PFX, inv_check)(const char *rout, const enum CBLAS_ORDER Order, int has_uplo,
                              const enum CBLAS_UPLO Uplo, int has_diag,
                              const enum CBLAS_DIAG Diag, const int N, const int lda)
{
  int ierr = 0, pos = 2;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, rout, "Order must be %d or %d, but is set to %d\n", CblasRowMajor,
                 CblasColMajor, Order);
  }
  if (has_uplo) {
    if (Uplo != CblasUpper && Uplo != CblasLower) {
      ierr = -pos;
      cblas_xerbla(pos, rout, "Uplo must be %d or %d, but is set to %d\n", CblasUpper,
                   CblasLower, Uplo);
    }
    ++pos;
  }
  if (has_diag) {
    if (Diag != CblasUnit && Diag != CblasNonUnit) {
      ierr = -pos;
      cblas_xerbla(pos, rout, "Diag must be %d or %d, but is set to %d\n", CblasNonUnit,
                   CblasUnit, Diag);
    }
    ++pos;
  }
  if (N < 0) {
    ierr = -pos;
    cblas_xerbla(pos, rout, "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  pos += 2;
  if (lda < N || lda < 1) {
    ierr = -pos;
    cblas_xerbla(pos, rout, "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  return ierr;
}

/* interfaces/lapack/C/src/clapack_dtrtri.c.  Its wrapper only calls
 * ATL_dtrtri when an argument was bad (`if (ierr)`), so it never inverts;
 * here the inverse is computed, which is what ATL_trtri and LAPACK's dtrtri
 * do.  RowMajor is the column-major inverse of the transposed view (Uplo
 * flipped). */
int NM3(clapack_, PFX, trtri)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                              const enum CBLAS_DIAG Diag, const int N, T *A, const int lda)
{
  static const char *rout = "clapack_" STR(PFX) "trtri";
  int ierr = NM(PFX, inv_check)(rout, Order, 1, Uplo, 1, Diag, N, lda), low, info;
  atlmx_ctx *c;
  long dla;
  T *dA;
  if (ierr) return ierr;
  if (N == 0) return 0;
  low = (Uplo == CblasLower);
  if (Order == CblasRowMajor) low = !low;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, sizeof(T), 1, &dla, rout);
  info = NM(PFX, trtri_dev)(c, low ? KX_LOWER : KX_UPPER,
                            Diag == CblasUnit ? KX_UNIT : KX_NONUNIT, N, dA, dla, rout);
  if (!info) rt_stage_out(c, A, dA, N, N, lda, dla, sizeof(T), rout);
  rt_sync(c, rout);
  return info;
}

/* interfaces/lapack/C/src/clapack_dlauum.c; RowMajor: the other triangle
 * of the transposed view holds the same product */
int NM3(clapack_, PFX, lauum)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                              const int N, T *A, const int lda)
{
  static const char *rout = "clapack_" STR(PFX) "lauum";
  int ierr = NM(PFX, inv_check)(rout, Order, 1, Uplo, 0, CblasNonUnit, N, lda), low;
  atlmx_ctx *c;
  long dla;
  T *dA;
  if (ierr) return ierr;
  if (N == 0) return 0;
  low = (Uplo == CblasLower);
  if (Order == CblasRowMajor) low = !low;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, sizeof(T), 1, &dla, rout);
  NM(PFX, lauum_dev)(c, low ? KX_LOWER : KX_UPPER, N, dA, dla, rout);
  rt_stage_out(c, A, dA, N, N, lda, dla, sizeof(T), rout);
  rt_sync(c, rout);
  return 0;
}

/* interfaces/lapack/C/src/clapack_dpotri.c: TRTRI (non-unit) then LAUUM */
int NM3(clapack_, PFX, potri)(const enum CBLAS_ORDER Order, const enum CBLAS_UPLO Uplo,
                              const int N, T *A, const int lda)
{
  static const char *rout = "clapack_" STR(PFX) "potri";
  int ierr = NM(PFX, inv_check)(rout, Order, 1, Uplo, 0, CblasNonUnit, N, lda), low, info;
  atlmx_ctx *c;
  long dla;
  T *dA;
  if (ierr) return ierr;
  if (N == 0) return 0;
  low = (Uplo == CblasLower);
  if (Order == CblasRowMajor) low = !low;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, sizeof(T), 1, &dla, rout);
  info = NM(PFX, trtri_dev)(c, low ? KX_LOWER : KX_UPPER, KX_NONUNIT, N, dA, dla, rout);
  if (!info) {
    NM(PFX, lauum_dev)(c, low ? KX_LOWER : KX_UPPER, N, dA, dla, rout);
    rt_stage_out(c, A, dA, N, N, lda, dla, sizeof(T), rout);
  }
  rt_sync(c, rout);
  return info;
}

/* interfaces/lapack/C/src/clapack_dgetri.c -> src/lapack/ATL_getri.c.
 * RowMajor holds the factors of A^T, whose inverse is the row-major
 * inverse of A, so both orders are the same column-major computation. */
int NM3(clapack_, PFX, getri)(const enum CBLAS_ORDER Order, const int N, T *A, const int lda,
                              const int *ipiv)
{
  static const char *rout = "clapack_" STR(PFX) "getri";
  int ierr = NM(PFX, inv_check)(rout, Order, 0, CblasUpper, 0, CblasNonUnit, N, lda), info;
  atlmx_ctx *c;
  long dla;
  T *dA;
  if (ierr) return ierr;
  if (N == 0) return 0;
  c = rt_ctx();
  dA = (T *)rt_stage_in(c, WS_STAGE_A, A, N, N, lda, sizeof(T), 1, &dla, rout);
  info = NM(PFX, getri_dev)(c, N, dA, dla, NM(PFX, stage_ipiv)(c, ipiv, N, rout), rout);
  if (!info) rt_stage_out(c, A, dA, N, N, lda, dla, sizeof(T), rout);
  rt_sync(c, rout);
  return info;
}

/* interfaces/lapack/C/src/clapack_dgeqrf.c.  RowMajor there is the LQ
 * factorization of the column-major view (ATL_gelqf), which for real types
 * holds the same numbers as the QR of the matrix itself: the storage is
 * transposed into a workspace, factored, and transposed back. */
int NM3(clapack_, PFX, geqrf)(const enum CBLAS_ORDER Order, const int M, const int N, T *A,
                              const int lda, T *TAU)
{
  static const char *rout = "clapack_" STR(PFX) "geqrf";
  int ierr = 0, K;
  atlmx_ctx *c;
  long dla;
  T *dA, *dt;
  if (Order != CblasRowMajor && Order != CblasColMajor) {
    ierr = -1;
    cblas_xerbla(1, rout, "Order must be %d or %d, but is set to %d\n", CblasRowMajor,
                 CblasColMajor, Order);
  }
  if (M < 0) {
    ierr = -2;
    cblas_xerbla(2, rout, "M cannot be less than zero 0,; is set to %d.\n", M);
  }
  if (N < 0) {
    ierr = -3;
    cblas_xerbla(3, rout, "N cannot be less than zero 0,; is set to %d.\n", N);
  }
  if (Order == CblasColMajor) {
    if (lda < M || lda < 1) {
      ierr = -5;
      cblas_xerbla(5, rout, "lda must be >= MAX(M,1): lda=%d M=%d\n", lda, M);
    }
  } else if (lda < N || lda < 1) {
    ierr = -5;
    cblas_xerbla(5, rout, "lda must be >= MAX(N,1): lda=%d N=%d\n", lda, N);
  }
  if (ierr) return ierr;
  K = M < N ? M : N;
  if (K == 0) return 0;
  c = rt_ctx();
  dt = rt_is_device(TAU) ? TAU : (T *)rt_ws(c, WS_TAU, (size_t)K * sizeof(T), rout);
  if (Order == CblasColMajor) {
    dA = (T *)rt_stage_in(c, WS_STAGE_A, A, M, N, lda, sizeof(T), 1, &dla, rout);
    NM(PFX, geqrf_dev)(c, M, N, dA, dla, dt, rout);
    rt_stage_out(c, A, dA, M, N, lda, dla, sizeof(T), rout);
  } else {
    /* the storage is the column-major N x M matrix A^T */
    T *dB = (T *)rt_stage_in(c, WS_STAGE_A, A, N, M, lda, sizeof(T), 1, &dla, rout);
    T *W = (T *)rt_ws(c, WS_STAGE_B, (size_t)M * N * sizeof(T), rout);
    RT_CHECK(rout, K_TRICOPY(ATLMX_TC_TRANSPOSE, 0, 0, N, M, dB, dla, W, M, c->stream));
    NM(PFX, geqrf_dev)(c, M, N, W, M, dt, rout);
    RT_CHECK(rout, K_TRICOPY(ATLMX_TC_TRANSPOSE, 0, 0, M, N, W, M, dB, dla, c->stream));
    rt_stage_out(c, A, dB, N, M, lda, dla, sizeof(T), rout);
  }
  if (dt != TAU)
    RT_CHECK(rout, hipMemcpyAsync(TAU, dt, (size_t)K * sizeof(T), hipMemcpyDeviceToHost,
                                  c->stream));
  rt_sync(c, rout);
  return 0;
}

#undef GETRF_DEV
#undef GETRS_DEV
#undef POTRF_DEV
#undef POTRS_DEV
#undef PANEL_REC
#undef GETRF_UPDATE
#undef LASWP_ROWS

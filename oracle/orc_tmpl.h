/*
 * ORACLE -- test infrastructure only (see oracle.h).
 *
 * Precision template: included by oracle.c once with T=double, P=orc_d and
 * once with T=float, P=orc_s.  The same source is compiled twice the way
 * ATLAS compiles each file with -DDREAL / -DSREAL (makes/Make.ul3:7-23).
 */
#define F_(name) CAT_(P, name)

/* ======================================================================= */
/* Input generators and norms (src/testing/)                               */
/* ======================================================================= */

/* src/testing/ATL_gegen.c:7-20 (gefillgap): lda gap canary -2.56e9 */
void F_(gefillgap)(int M, int N, T *A, int lda)
{
   int i, j;
   if (lda <= M) return;
   for (j = 0; j < N; j++)
      for (i = M; i < lda; i++) A[(size_t)j*lda + i] = (T)ORC_FILLCONST;
}

/* src/testing/ATL_gegen.c:22-44 (gechkgap): count overwrites in the gap */
int F_(gechkgap)(int M, int N, const T *A, int lda)
{
   int i, j, bad = 0;
   if (lda <= M) return 0;
   for (j = 0; j < N; j++)
      for (i = M; i < lda; i++)
         if (A[(size_t)j*lda + i] != (T)ORC_FILLCONST) bad++;
   return bad;
}

/* src/testing/ATL_gegen.c:46-59: seed, fill gap, then column by column
 * A[i] = dumb_rand() (include/atlas_tst.h:18). */
void F_(gegen)(int M, int N, T *A, int lda, int seed)
{
   int i, j;
   orc_srand(seed);
   F_(gefillgap)(M, N, A, lda);
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++) A[(size_t)j*lda + i] = (T)orc_dumb_rand();
}

/* bin/gemmtst.c:85-104 (matgen): like gegen but the gap holds -9e27 */
void F_(matgen)(int M, int N, T *A, int lda, int seed)
{
   int i, j;
   orc_srand(seed);
   for (j = 0; j < N; j++)
   {
      for (i = 0; i < M; i++) A[(size_t)j*lda + i] = (T)orc_dumb_rand();
      for (; i < lda; i++) A[(size_t)j*lda + i] = (T)-9e27;
   }
}

/* src/testing/ATL_epsilon.c:9-21: halving loop until 1+f == 1 */
T F_(epsilon)(void)
{
   volatile T maxval, f1 = (T)0.5;
   T eps;
   const T half = (T)0.5;
   do
   {
      eps = f1;
      f1 *= half;
      maxval = (T)1.0 + f1;
   } while (maxval != (T)1.0);
   return eps;
}

/* src/testing/ATL_genrm1.c:9-26: max column abs-sum; NaN propagates */
T F_(genrm1)(int M, int N, const T *A, int lda)
{
   int i, j;
   T mx = 0, s;
   for (j = 0; j < N; j++)
   {
      s = 0;
      for (i = 0; i < M; i++) s += ORC_ABS(A[(size_t)j*lda + i]);
      if (s != s) return s;
      if (s > mx) mx = s;
   }
   return mx;
}

/* src/testing/ATL_gediffnrm1.c:8-30: 1-norm of (A - B) */
T F_(gediffnrm1)(int M, int N, const T *A, int lda, const T *B, int ldb)
{
   int i, j;
   T mx = 0, s;
   for (j = 0; j < N; j++)
   {
      s = 0;
      for (i = 0; i < M; i++)
         s += ORC_ABS(A[(size_t)j*lda + i] - B[(size_t)j*ldb + i]);
      if (s != s) return s;
      if (s > mx) mx = s;
   }
   return mx;
}

/* src/testing/ATL_infnrm.c:9-35: max |x_i|; a NaN entry is returned as
 * soon as it is met (:30-31). */
T F_(infnrm)(int N, const T *X, int incX)
{
   int i;
   T mx = 0, t;
   for (i = 0; i < N; i++)
   {
      t = ORC_ABS(X[(size_t)i*incX]);
      if (t != t) return t;
      if (t > mx) mx = t;
   }
   return mx;
}

/* ======================================================================= */
/* Level-1 pieces of the LU panel                                          */
/* ======================================================================= */

/* src/blas/reference/level1/ATL_idrefamax.c:112: first index whose |x| is
 * strictly larger than every earlier one; 0 for all-zero input; NaN never
 * wins because (NaN > smax) is false. */
int F_(iamax)(int N, const T *X, int incX)
{
   int i, imax = 0;
   T smax = 0, a;
   for (i = 0; i < N; i++)
   {
      a = ORC_ABS(X[(size_t)i*incX]);
      if (a > smax) { imax = i; smax = a; }
   }
   return imax;
}

/* Pivot search used by the LU restatement: iamax plus, when gap tracking is
 * on, the relative distance between the two largest |x| (for near-tie
 * screening of pivot comparisons). */
static int F_(pivsearch)(int N, const T *X, int gcol)
{
   int i, imax = 0;
   T smax = 0, a;
   for (i = 0; i < N; i++)
   {
      a = ORC_ABS(X[i]);
      if (a > smax) { imax = i; smax = a; }
   }
   if (orc_gaps_ && gcol >= 0 && gcol < orc_ngaps_)
   {
      double second = 0.0, g;
      for (i = 0; i < N; i++)
         if (i != imax && (double)ORC_ABS(X[i]) > second)
            second = ORC_ABS(X[i]);
      g = (smax > 0) ? ((double)smax - second) / (double)smax : 0.0;
      orc_gaps_[gcol] = g;
   }
   return imax;
}

/* src/lapack/ATL_laswp.c:7-110: apply interchanges K1..K2-1 (inci=+1) or
 * K2-1..K1 (inci=-1) to N columns.  ATLAS walks the columns in chunks of
 * 32; the chunking does not change the result, so it is not restated. */
void F_(laswp)(int N, T *A, int lda, int K1, int K2, const int *piv, int inci)
{
   int i, ip, h, k, i1, i2;
   T r;
   if (K2 <= K1 || N <= 0) return;
   if (inci < 0) { i1 = K2 - 1; i2 = K1; }
   else          { i1 = K1;     i2 = K2 - 1; }
   for (k = 0; k < K2 - K1; k++)
   {
      i = (inci > 0) ? i1 + k : i1 - k;
      ip = piv[i];
      if (ip != i)
         for (h = 0; h < N; h++)
         {
            r = A[(size_t)h*lda + i];
            A[(size_t)h*lda + i] = A[(size_t)h*lda + ip];
            A[(size_t)h*lda + ip] = r;
         }
   }
   (void)i2;
}

/* ======================================================================= */
/* GEMM                                                                    */
/* ======================================================================= */

/* src/blas/reference/level3/ATL_drefgemm.c (dispatcher) and
 * ATL_drefgemm{NN,NT,TN,TT}.c loop orders:
 *   NN/NT: per column j, scale C(:,j) by beta (Mdgescal), then for each l
 *          t0 = alpha*B(l,j); C(:,j) += A(:,l)*t0.
 *   TN/TT: per (i,j) dot product t0 = sum_l A(l,i)*B(l,j) (or B(j,l)), then
 *          C(i,j) = beta*C(i,j) (Mdelscal) + alpha*t0.
 * Quick returns as the dispatcher: M==0 || N==0 || ((alpha==0||K==0) &&
 * beta==1); alpha==0 only scales C. */
void F_(gemm_ref)(int TA, int TB, int M, int N, int K, T alpha,
                  const T *A, int lda, const T *B, int ldb, T beta,
                  T *C, int ldc)
{
   int i, j, l;
   T t0;
   const int ta = (TA != ORC_NOTRANS), tb = (TB != ORC_NOTRANS);
   if (M == 0 || N == 0 || ((alpha == 0 || K == 0) && beta == 1)) return;
   if (alpha == 0)
   {
      for (j = 0; j < N; j++)
         for (i = 0; i < M; i++)
         {
            T *c = C + (size_t)j*ldc + i;
            if (beta == 0) *c = 0; else if (beta != 1) *c *= beta;
         }
      return;
   }
   if (!ta)
   {
      for (j = 0; j < N; j++)
      {
         T *c = C + (size_t)j*ldc;
         if (beta == 0) { for (i = 0; i < M; i++) c[i] = 0; }
         else if (beta != 1) { for (i = 0; i < M; i++) c[i] *= beta; }
         for (l = 0; l < K; l++)
         {
            const T *a = A + (size_t)l*lda;
            t0 = alpha * (tb ? B[(size_t)l*ldb + j] : B[(size_t)j*ldb + l]);
            for (i = 0; i < M; i++) c[i] += a[i] * t0;
         }
      }
   }
   else
   {
      for (j = 0; j < N; j++)
         for (i = 0; i < M; i++)
         {
            const T *a = A + (size_t)i*lda;
            T *c = C + (size_t)j*ldc + i;
            t0 = 0;
            if (!tb)
               for (l = 0; l < K; l++) t0 += a[l] * B[(size_t)j*ldb + l];
            else
               for (l = 0; l < K; l++) t0 += a[l] * B[(size_t)l*ldb + j];
            if (beta == 0) *c = 0; else if (beta != 1) *c *= beta;
            *c += alpha * t0;
         }
   }
}

/* Cache-blocked, threaded GEMM used where only the value (not the rounding
 * order) matters: big test residuals and the CPU baseline.  It follows the
 * access-major idea of src/blas/ammm/ATL_iploops.c:349-417 in spirit --
 * pack a KB x NB panel of op(B), stream op(A) panels against it and keep
 * each C block hot -- without restating the generated kernels. */
#define ORC_KB 256
#define ORC_NB 64
#define ORC_MB 96
typedef struct {
   int TA, TB, M, N, K, j0, j1;
   T alpha, beta;
   const T *A, *B;
   T *C;
   int lda, ldb, ldc;
} F_(gemm_job_t);

/* C(mb x nb) += Ap Bp with Ap packed [l][i] (mb per l) and Bp packed
 * [j][l] (kb per j): register tiles of 2 vectors x 4 columns, the vectors
 * GCC generic vectors of 32 bytes (AVX2 where the build targets it). */
typedef T F_(vec_t) __attribute__((vector_size(32), aligned(sizeof(T))));
#define ORC_VL ((int)(32 / sizeof(T)))
static void F_(gemm_micro)(int mb, int nb, int kb, const T *ap, const T *bp, T *C, int ldc)
{
   const int MR = 2 * ORC_VL, NR = 4;
   int i0, j0, i, j, l;
   for (j0 = 0; j0 + NR <= nb; j0 += NR)
   {
      const T *b0 = bp + (size_t)j0*kb, *b1 = b0 + kb, *b2 = b1 + kb, *b3 = b2 + kb;
      for (i0 = 0; i0 + MR <= mb; i0 += MR)
      {
         F_(vec_t) c00 = {0}, c01 = {0}, c02 = {0}, c03 = {0};
         F_(vec_t) c10 = {0}, c11 = {0}, c12 = {0}, c13 = {0};
         for (l = 0; l < kb; l++)
         {
            /* the vector type is aligned to one element: unaligned loads */
            const F_(vec_t) a0 = *(const F_(vec_t) *)(ap + (size_t)l*mb + i0);
            const F_(vec_t) a1 = *(const F_(vec_t) *)(ap + (size_t)l*mb + i0 + ORC_VL);
            c00 += a0 * b0[l]; c10 += a1 * b0[l];
            c01 += a0 * b1[l]; c11 += a1 * b1[l];
            c02 += a0 * b2[l]; c12 += a1 * b2[l];
            c03 += a0 * b3[l]; c13 += a1 * b3[l];
         }
         for (i = 0; i < ORC_VL; i++)
         {
            T *c = C + (size_t)j0*ldc + i0 + i;
            c[0] += c00[i]; c[ORC_VL] += c10[i];
            c[ldc] += c01[i]; c[ldc + ORC_VL] += c11[i];
            c[2 * (size_t)ldc] += c02[i]; c[2 * (size_t)ldc + ORC_VL] += c12[i];
            c[3 * (size_t)ldc] += c03[i]; c[3 * (size_t)ldc + ORC_VL] += c13[i];
         }
      }
      for (i = i0; i < mb; i++)                   /* leftover rows */
         for (j = j0; j < j0 + NR; j++)
         {
            T acc = 0;
            for (l = 0; l < kb; l++) acc += ap[(size_t)l*mb + i] * bp[(size_t)j*kb + l];
            C[(size_t)j*ldc + i] += acc;
         }
   }
   for (j = j0; j < nb; j++)                      /* leftover columns */
      for (i = 0; i < mb; i++)
      {
         T acc = 0;
         for (l = 0; l < kb; l++) acc += ap[(size_t)l*mb + i] * bp[(size_t)j*kb + l];
         C[(size_t)j*ldc + i] += acc;
      }
}
#undef ORC_VL

static void F_(gemm_block)(const F_(gemm_job_t) *g)
{
   const int ncol = g->j1 - g->j0;
   T *bp, *ap;
   int k0, i0, jb, i, j, l;
   if (ncol <= 0) return;
   for (j = g->j0; j < g->j1; j++)
   {
      T *c = g->C + (size_t)j*g->ldc;
      if (g->beta == 0) for (i = 0; i < g->M; i++) c[i] = 0;
      else if (g->beta != 1) for (i = 0; i < g->M; i++) c[i] *= g->beta;
   }
   bp = (T *)malloc(sizeof(T) * ORC_KB * (size_t)ncol);
   ap = (T *)malloc(sizeof(T) * ORC_KB * ORC_MB);
   for (k0 = 0; k0 < g->K; k0 += ORC_KB)
   {
      const int kb = (g->K - k0 < ORC_KB) ? g->K - k0 : ORC_KB;
      /* alpha * op(B) for this thread's columns, [j][l] */
      for (j = 0; j < ncol; j++)
         for (l = 0; l < kb; l++)
         {
            const int gj = g->j0 + j, gl = k0 + l;
            T b = (g->TB == ORC_NOTRANS) ? g->B[(size_t)gj*g->ldb + gl]
                                         : g->B[(size_t)gl*g->ldb + gj];
            bp[(size_t)j*kb + l] = g->alpha * b;
         }
      for (i0 = 0; i0 < g->M; i0 += ORC_MB)
      {
         const int mb = (g->M - i0 < ORC_MB) ? g->M - i0 : ORC_MB;
         /* op(A) block as [l][i], packed once per (k0, i0) */
         if (g->TA == ORC_NOTRANS)
            for (l = 0; l < kb; l++)
               memcpy(ap + (size_t)l*mb, g->A + (size_t)(k0 + l)*g->lda + i0, sizeof(T) * mb);
         else
            for (i = 0; i < mb; i++)
               for (l = 0; l < kb; l++)
                  ap[(size_t)l*mb + i] = g->A[(size_t)(i0 + i)*g->lda + k0 + l];
         for (jb = 0; jb < ncol; jb += ORC_NB)
         {
            const int nb = (ncol - jb < ORC_NB) ? ncol - jb : ORC_NB;
            F_(gemm_micro)(mb, nb, kb, ap, bp + (size_t)jb*kb,
                           g->C + (size_t)(g->j0 + jb)*g->ldc + i0, g->ldc);
         }
      }
   }
   free(bp);
   free(ap);
}

static void *F_(gemm_thread)(void *p)
{
   F_(gemm_block)((const F_(gemm_job_t) *)p);
   return NULL;
}

void F_(gemm_fast)(int TA, int TB, int M, int N, int K, T alpha,
                   const T *A, int lda, const T *B, int ldb, T beta,
                   T *C, int ldc)
{
   int nt = orc_get_threads(), t, per;
   F_(gemm_job_t) jobs[ORC_MAXTHR];
   pthread_t th[ORC_MAXTHR];
   if (M == 0 || N == 0 || ((alpha == 0 || K == 0) && beta == 1)) return;
   if (alpha == 0 || K == 0)
   {
      F_(gemm_ref)(TA, TB, M, N, 0, alpha, A, lda, B, ldb, beta, C, ldc);
      if (K == 0 && alpha != 0 && beta == 1) return;
      return;
   }
   if ((double)M * N * K < 2.0e6) nt = 1;
   per = (N + nt - 1) / nt;
   if (per < 1) per = 1;
   for (t = 0; t < nt; t++)
   {
      jobs[t].TA = TA; jobs[t].TB = TB; jobs[t].M = M; jobs[t].N = N;
      jobs[t].K = K; jobs[t].alpha = alpha; jobs[t].beta = beta;
      jobs[t].A = A; jobs[t].B = B; jobs[t].C = C;
      jobs[t].lda = lda; jobs[t].ldb = ldb; jobs[t].ldc = ldc;
      jobs[t].j0 = t * per;
      jobs[t].j1 = (t + 1) * per < N ? (t + 1) * per : N;
   }
   if (nt == 1) { jobs[0].j0 = 0; jobs[0].j1 = N; F_(gemm_block)(&jobs[0]); return; }
   for (t = 0; t < nt; t++)
      if (jobs[t].j0 < jobs[t].j1)
         pthread_create(&th[t], NULL, F_(gemm_thread), &jobs[t]);
   for (t = 0; t < nt; t++)
      if (jobs[t].j0 < jobs[t].j1) pthread_join(th[t], NULL);
}

/* op = left, lower, no-transpose, unit: the TRSM of the LU recursion,
 * recursive halving with GEMM updates as src/blas/ulevel3/ATL_trsm.c does
 * it, the reference loops at the 64-row leaves. */
static void F_(trsm_llnu)(int M, int N, const T *A, int lda, T *B, int ldb)
{
   int m1;
   if (M <= 64)
   {
      F_(trsm_ref)(ORC_LEFT, ORC_LOWER, ORC_NOTRANS, ORC_UNIT, M, N, (T)1, A, lda, B, ldb);
      return;
   }
   m1 = M / 2;
   F_(trsm_llnu)(m1, N, A, lda, B, ldb);
   F_(gemm_fast)(ORC_NOTRANS, ORC_NOTRANS, M - m1, N, m1, (T)-1, A + m1, lda, B, ldb, (T)1,
                 B + m1, ldb);
   F_(trsm_llnu)(M - m1, N, A + m1 + (size_t)m1*lda, lda, B + m1, ldb);
}

/* interfaces/blas/C/src/cblas_dgemm.c:130-134: ColMajor direct, RowMajor by
 * swapping the operands (A<->B, M<->N, TA<->TB). */
void F_(gemm)(int order, int TA, int TB, int M, int N, int K, T alpha,
              const T *A, int lda, const T *B, int ldb, T beta, T *C, int ldc)
{
   if (order == ORC_COL)
      F_(gemm_fast)(TA, TB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
   else
      F_(gemm_fast)(TB, TA, N, M, K, alpha, B, ldb, A, lda, beta, C, ldc);
}

/* ======================================================================= */
/* TRSM / TRMM / SYRK                                                      */
/* ======================================================================= */
#define AE(i_, j_) A[(size_t)(j_)*lda + (i_)]
#define BE(i_, j_) B[(size_t)(j_)*ldb + (i_)]

/* src/blas/reference/level3/ATL_dreftrsm.c (dispatcher: M==0||N==0 return,
 * alpha==0 zeroes B) and the 16 ATL_dreftrsm{L,R}{L,U}{N,T}{N,U}.c kernels.
 * Each branch below keeps the loop order and the place where alpha is
 * applied of the corresponding file:
 *   LLN*: ATL_dreftrsmLLNN.c  LLT*: ATL_dreftrsmLLTN.c
 *   LUN*: ATL_dreftrsmLUNN.c  LUT*: ATL_dreftrsmLUTN.c
 *   RLN*: ATL_dreftrsmRLNN.c  RLT*: ATL_dreftrsmRLTN.c
 *   RUN*: ATL_dreftrsmRUNN.c  RUT*: ATL_dreftrsmRUTN.c
 * The *U (unit) files are the same loops without the division. */
void F_(trsm_ref)(int side, int uplo, int ta, int diag, int M, int N, T alpha,
                  const T *A, int lda, T *B, int ldb)
{
   int i, j, k;
   T t0;
   const int nu = (diag == ORC_NONUNIT);
   const int tr = (ta != ORC_NOTRANS);
   if (M == 0 || N == 0) return;
   if (alpha == 0)
   {
      for (j = 0; j < N; j++) for (i = 0; i < M; i++) BE(i, j) = 0;
      return;
   }
   if (side == ORC_LEFT)
   {
      if (uplo == ORC_LOWER && !tr)            /* LLN: forward, column axpy */
         for (j = 0; j < N; j++)
         {
            for (i = 0; i < M; i++) BE(i, j) *= alpha;
            for (k = 0; k < M; k++)
            {
               if (nu) BE(k, j) /= AE(k, k);
               for (i = k + 1; i < M; i++) BE(i, j) -= BE(k, j) * AE(i, k);
            }
         }
      else if (uplo == ORC_LOWER && tr)        /* LLT: backward, dot */
         for (j = 0; j < N; j++)
            for (i = M - 1; i >= 0; i--)
            {
               t0 = alpha * BE(i, j);
               for (k = i + 1; k < M; k++) t0 -= AE(k, i) * BE(k, j);
               if (nu) t0 /= AE(i, i);
               BE(i, j) = t0;
            }
      else if (uplo == ORC_UPPER && !tr)       /* LUN: backward, column axpy */
         for (j = 0; j < N; j++)
         {
            for (i = 0; i < M; i++) BE(i, j) *= alpha;
            for (k = M - 1; k >= 0; k--)
            {
               if (nu) BE(k, j) /= AE(k, k);
               for (i = 0; i < k; i++) BE(i, j) -= BE(k, j) * AE(i, k);
            }
         }
      else                                     /* LUT: forward, dot */
         for (j = 0; j < N; j++)
            for (i = 0; i < M; i++)
            {
               t0 = alpha * BE(i, j);
               for (k = 0; k < i; k++) t0 -= AE(k, i) * BE(k, j);
               if (nu) t0 /= AE(i, i);
               BE(i, j) = t0;
            }
   }
   else
   {
      if (uplo == ORC_LOWER && !tr)            /* RLN: columns right to left */
         for (j = N - 1; j >= 0; j--)
         {
            for (i = 0; i < M; i++) BE(i, j) *= alpha;
            for (k = j + 1; k < N; k++)
               for (i = 0; i < M; i++) BE(i, j) -= AE(k, j) * BE(i, k);
            if (nu) for (i = 0; i < M; i++) BE(i, j) /= AE(j, j);
         }
      else if (uplo == ORC_LOWER && tr)        /* RLT: alpha applied last */
         for (k = 0; k < N; k++)
         {
            if (nu) for (i = 0; i < M; i++) BE(i, k) /= AE(k, k);
            for (j = k + 1; j < N; j++)
            {
               t0 = AE(j, k);
               for (i = 0; i < M; i++) BE(i, j) -= t0 * BE(i, k);
            }
            for (i = 0; i < M; i++) BE(i, k) *= alpha;
         }
      else if (uplo == ORC_UPPER && !tr)       /* RUN: columns left to right */
         for (j = 0; j < N; j++)
         {
            for (i = 0; i < M; i++) BE(i, j) *= alpha;
            for (k = 0; k < j; k++)
               for (i = 0; i < M; i++) BE(i, j) -= AE(k, j) * BE(i, k);
            if (nu) for (i = 0; i < M; i++) BE(i, j) /= AE(j, j);
         }
      else                                     /* RUT: alpha applied last */
         for (k = N - 1; k >= 0; k--)
         {
            if (nu) for (i = 0; i < M; i++) BE(i, k) /= AE(k, k);
            for (j = 0; j < k; j++)
            {
               t0 = AE(j, k);
               for (i = 0; i < M; i++) BE(i, j) -= t0 * BE(i, k);
            }
            for (i = 0; i < M; i++) BE(i, k) *= alpha;
         }
   }
}

/* interfaces/blas/C/src/cblas_dtrsm.c:97-105: RowMajor flips Side and Uplo
 * and swaps M/N. */
void F_(trsm)(int order, int side, int uplo, int ta, int diag, int M, int N,
              T alpha, const T *A, int lda, T *B, int ldb)
{
   if (order == ORC_COL)
      F_(trsm_ref)(side, uplo, ta, diag, M, N, alpha, A, lda, B, ldb);
   else
      F_(trsm_ref)(side == ORC_LEFT ? ORC_RIGHT : ORC_LEFT,
                   uplo == ORC_UPPER ? ORC_LOWER : ORC_UPPER,
                   ta, diag, N, M, alpha, A, lda, B, ldb);
}

/* B := alpha*op(A)*B or alpha*B*op(A), A triangular, as a dense product
 * with the triangle expanded (threaded GEMM).  The lutst residual uses it
 * (bin/lutst.c:84-134 calls cblas_trmm), where only the value matters;
 * trmm_ref below keeps the reference loop orders. */
void F_(trmm_fast)(int side, int uplo, int ta, int diag, int M, int N, T alpha,
                   const T *A, int lda, T *B, int ldb)
{
   const int na = (side == ORC_LEFT) ? M : N;
   int i, j;
   T *Tm, *W;
   if (M == 0 || N == 0) return;
   Tm = (T *)calloc((size_t)na * na, sizeof(T));
   W = (T *)malloc((size_t)M * N * sizeof(T));
   for (j = 0; j < na; j++)
      for (i = 0; i < na; i++)
      {
         int in = (uplo == ORC_LOWER) ? (i > j) : (i < j);
         T v = in ? AE(i, j) : 0;
         if (i == j) v = (diag == ORC_UNIT) ? (T)1 : AE(i, i);
         Tm[(size_t)j*na + i] = v;
      }
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++) W[(size_t)j*M + i] = BE(i, j);
   if (side == ORC_LEFT)
      F_(gemm_fast)(ta == ORC_NOTRANS ? ORC_NOTRANS : ORC_TRANS, ORC_NOTRANS,
                    M, N, M, alpha, Tm, na, W, M, 0, B, ldb);
   else
      F_(gemm_fast)(ORC_NOTRANS, ta == ORC_NOTRANS ? ORC_NOTRANS : ORC_TRANS,
                    M, N, N, alpha, W, M, Tm, na, 0, B, ldb);
   free(Tm);
   free(W);
}

/* B := alpha*op(A)*B (Left) or alpha*B*op(A) (Right), A triangular, in the
 * loop order of src/blas/reference/level3/ATL_dreftrmm*.c: alpha == 0
 * zeroes B (ATL_dreftrmm.c:180-183); each branch keeps where alpha and the
 * diagonal enter:
 *   LLN*: ATL_dreftrmmLLNN.c  LUN*: ATL_dreftrmmLUNN.c
 *   LLT*: ATL_dreftrmmLLTN.c  LUT*: ATL_dreftrmmLUTN.c
 *   RLN*: ATL_dreftrmmRLNN.c  RUN*: ATL_dreftrmmRUNN.c
 *   RLT*: ATL_dreftrmmRLTN.c  RUT*: ATL_dreftrmmRUTN.c
 * The *U (unit) files drop the diagonal product. */
void F_(trmm_ref)(int side, int uplo, int ta, int diag, int M, int N, T alpha,
                  const T *A, int lda, T *B, int ldb)
{
   int i, j, k;
   T t0;
   const int nu = (diag == ORC_NONUNIT);
   const int tr = (ta != ORC_NOTRANS);
   if (M == 0 || N == 0) return;
   if (alpha == 0)
   {
      for (j = 0; j < N; j++) for (i = 0; i < M; i++) BE(i, j) = 0;
      return;
   }
   if (side == ORC_LEFT)
   {
      if (uplo == ORC_LOWER && !tr)            /* rows bottom-up, axpy below k */
         for (j = 0; j < N; j++)
            for (k = M - 1; k >= 0; k--)
            {
               t0 = alpha * BE(k, j);
               BE(k, j) = t0;
               if (nu) BE(k, j) *= AE(k, k);
               for (i = k + 1; i < M; i++) BE(i, j) += t0 * AE(i, k);
            }
      else if (uplo == ORC_UPPER && !tr)       /* rows top-down, axpy above k */
         for (j = 0; j < N; j++)
            for (k = 0; k < M; k++)
            {
               t0 = alpha * BE(k, j);
               for (i = 0; i < k; i++) BE(i, j) += t0 * AE(i, k);
               if (nu) t0 *= AE(k, k);
               BE(k, j) = t0;
            }
      else if (uplo == ORC_LOWER)              /* L^T: rows top-down, dot */
         for (j = 0; j < N; j++)
            for (i = 0; i < M; i++)
            {
               t0 = BE(i, j);
               if (nu) t0 *= AE(i, i);
               for (k = i + 1; k < M; k++) t0 += AE(k, i) * BE(k, j);
               BE(i, j) = alpha * t0;
            }
      else                                     /* U^T: rows bottom-up, dot */
         for (j = 0; j < N; j++)
            for (i = M - 1; i >= 0; i--)
            {
               t0 = BE(i, j);
               if (nu) t0 *= AE(i, i);
               for (k = 0; k < i; k++) t0 += AE(k, i) * BE(k, j);
               BE(i, j) = alpha * t0;
            }
   }
   else
   {
      if (uplo == ORC_LOWER && !tr)            /* columns left to right */
         for (j = 0; j < N; j++)
         {
            t0 = alpha;
            if (nu) t0 *= AE(j, j);
            for (i = 0; i < M; i++) BE(i, j) *= t0;
            for (k = j + 1; k < N; k++)
            {
               t0 = alpha * AE(k, j);
               for (i = 0; i < M; i++) BE(i, j) += t0 * BE(i, k);
            }
         }
      else if (uplo == ORC_UPPER && !tr)       /* columns right to left */
         for (j = N - 1; j >= 0; j--)
         {
            t0 = alpha;
            if (nu) t0 *= AE(j, j);
            for (i = 0; i < M; i++) BE(i, j) *= t0;
            for (k = 0; k < j; k++)
            {
               t0 = alpha * AE(k, j);
               for (i = 0; i < M; i++) BE(i, j) += t0 * BE(i, k);
            }
         }
      else if (uplo == ORC_LOWER)              /* B*L^T: k right to left */
         for (k = N - 1; k >= 0; k--)
         {
            for (j = k + 1; j < N; j++)
            {
               t0 = alpha * AE(j, k);
               for (i = 0; i < M; i++) BE(i, j) += t0 * BE(i, k);
            }
            t0 = alpha;
            if (nu) t0 *= AE(k, k);
            for (i = 0; i < M; i++) BE(i, k) = t0 * BE(i, k);
         }
      else                                     /* B*U^T: k left to right */
         for (k = 0; k < N; k++)
         {
            for (j = 0; j < k; j++)
            {
               t0 = alpha * AE(j, k);
               for (i = 0; i < M; i++) BE(i, j) += t0 * BE(i, k);
            }
            t0 = alpha;
            if (nu) t0 *= AE(k, k);
            for (i = 0; i < M; i++) BE(i, k) = t0 * BE(i, k);
         }
   }
}

/* beta scaling of one element as the reference BLAS does it
 * (include/atlas_refmisc.h:194-198): 0 stores zero, 1 leaves it. */
#define ORC_BSCAL(b_, x_) do { if ((b_) == 0) (x_) = 0; else if ((b_) != 1) (x_) *= (b_); } while (0)
#define CE(i_, j_) Cm[(size_t)(j_)*ldc + (i_)]

/* C := alpha*A*B + beta*C (Left) or alpha*B*A + beta*C (Right), A symmetric
 * held in one triangle, in the loop order of
 * src/blas/reference/level3/ATL_drefsymm{LL,LU,RL,RU}.c; quick returns and
 * the alpha == 0 scaling of ATL_drefsymm.c. */
void F_(symm_ref)(int side, int uplo, int M, int N, T alpha, const T *A, int lda,
                  const T *B, int ldb, T beta, T *Cm, int ldc)
{
   int i, j, k;
   T t0, t1;
   if (M == 0 || N == 0 || (alpha == 0 && beta == 1)) return;
   if (alpha == 0)
   {
      for (j = 0; j < N; j++) for (i = 0; i < M; i++) ORC_BSCAL(beta, CE(i, j));
      return;
   }
   if (side == ORC_LEFT)
   {
      for (j = 0; j < N; j++)
      {
         if (uplo == ORC_LOWER)
            for (i = M - 1; i >= 0; i--)
            {
               t0 = alpha * BE(i, j);
               t1 = 0;
               for (k = i + 1; k < M; k++)
               {
                  CE(k, j) += t0 * AE(k, i);
                  t1 += BE(k, j) * AE(k, i);
               }
               ORC_BSCAL(beta, CE(i, j));
               CE(i, j) += t0 * AE(i, i) + alpha * t1;
            }
         else
            for (i = 0; i < M; i++)
            {
               t0 = alpha * BE(i, j);
               t1 = 0;
               for (k = 0; k < i; k++)
               {
                  CE(k, j) += t0 * AE(k, i);
                  t1 += BE(k, j) * AE(k, i);
               }
               ORC_BSCAL(beta, CE(i, j));
               CE(i, j) += t0 * AE(i, i) + alpha * t1;
            }
      }
   }
   else
   {
      for (j = 0; j < N; j++)
      {
         t0 = alpha * AE(j, j);
         for (i = 0; i < M; i++)
         {
            ORC_BSCAL(beta, CE(i, j));
            CE(i, j) += t0 * BE(i, j);
         }
         for (k = 0; k < j; k++)                /* A(k,j) for k < j */
         {
            t0 = alpha * (uplo == ORC_LOWER ? AE(j, k) : AE(k, j));
            for (i = 0; i < M; i++) CE(i, j) += t0 * BE(i, k);
         }
         for (k = j + 1; k < N; k++)            /* A(k,j) for k > j */
         {
            t0 = alpha * (uplo == ORC_LOWER ? AE(k, j) : AE(j, k));
            for (i = 0; i < M; i++) CE(i, j) += t0 * BE(i, k);
         }
      }
   }
}

/* C := alpha*(op(A)*op(B)^T + op(B)*op(A)^T) + beta*C on one triangle, in
 * the loop order of src/blas/reference/level3/ATL_drefsyr2k{LN,LT,UN,UT}.c
 * with the quick returns and alpha == 0 scaling of ATL_drefsyr2k.c. */
void F_(syr2k_ref)(int uplo, int trans, int N, int K, T alpha, const T *A, int lda,
                   const T *B, int ldb, T beta, T *Cm, int ldc)
{
   int i, j, l, i0, i1;
   T t0, t1;
   if (N == 0 || ((alpha == 0 || K == 0) && beta == 1)) return;
   if (alpha == 0)
   {
      for (j = 0; j < N; j++)
      {
         i0 = (uplo == ORC_UPPER) ? 0 : j;
         i1 = (uplo == ORC_UPPER) ? j + 1 : N;
         for (i = i0; i < i1; i++)
         {
            if (beta == 0) CE(i, j) = 0;
            else if (beta != 1) CE(i, j) *= beta;
         }
      }
      return;
   }
   for (j = 0; j < N; j++)
   {
      i0 = (uplo == ORC_UPPER) ? 0 : j;
      i1 = (uplo == ORC_UPPER) ? j + 1 : N;
      if (trans == ORC_NOTRANS)
      {
         for (i = i0; i < i1; i++)
         {
            if (beta == 0) CE(i, j) = 0;
            else if (beta != 1) CE(i, j) *= beta;
         }
         for (l = 0; l < K; l++)
         {
            t0 = alpha * AE(j, l);
            t1 = alpha * BE(j, l);
            for (i = i0; i < i1; i++) CE(i, j) += t1 * AE(i, l) + t0 * BE(i, l);
         }
      }
      else
         for (i = i0; i < i1; i++)
         {
            t0 = 0;
            t1 = 0;
            for (l = 0; l < K; l++)
            {
               t0 += AE(l, i) * BE(l, j);
               t1 += BE(l, i) * AE(l, j);
            }
            ORC_BSCAL(beta, CE(i, j));
            CE(i, j) += alpha * t0 + alpha * t1;
         }
   }
}
#undef CE

/* C := alpha*op(A)*op(A)^T + beta*C on one triangle (the operation of
 * src/blas/ulevel3/ATL_syrk.c:257, used by the Cholesky recursion
 * src/lapack/ATL_potrfL.c:235). */
void F_(syrk_ref)(int uplo, int trans, int N, int K, T alpha, const T *A,
                  int lda, T beta, T *C, int ldc)
{
   int i, j, l;
   for (j = 0; j < N; j++)
   {
      const int i0 = (uplo == ORC_LOWER) ? j : 0;
      const int i1 = (uplo == ORC_LOWER) ? N : j + 1;
      for (i = i0; i < i1; i++)
      {
         T s = 0, *c = C + (size_t)j*ldc + i;
         if (trans == ORC_NOTRANS)
            for (l = 0; l < K; l++) s += AE(i, l) * AE(j, l);
         else
            for (l = 0; l < K; l++) s += AE(l, i) * AE(l, j);
         if (beta == 0) *c = 0; else if (beta != 1) *c *= beta;
         *c += alpha * s;
      }
   }
}
#undef AE
#undef BE

/* ======================================================================= */
/* LU                                                                      */
/* ======================================================================= */

/* src/lapack/ATL_getf2.c:10-59 (LU1) + :61-95 (getf2): unblocked
 * right-looking LU of an M x N panel, one column at a time: iamax, swap the
 * whole panel row, scale by the reciprocal of the pivot when |p| >= SAFMIN
 * (else divide), rank-1 update (ATL_dger with alpha=-1).
 * ATL_getf2.c:55-56 divides Ac[j] instead of Ac[i] in the tiny-pivot branch;
 * this restatement divides each element, as the comment there intends. */
static int F_(getf2_c)(int M, int N, T *A, int lda, int *ipiv, int gcol)
{
   const int MN = M < N ? M : N;
   int i, j, k, ip, iret = 0;
   if (M < 1 || N < 1) return 0;
   for (j = 0; j < MN; j++)
   {
      T *Ac = A + (size_t)j*lda;
      T p;
      ip = j + F_(pivsearch)(M - j, Ac + j, gcol + j);
      ipiv[j] = ip;
      p = Ac[ip];
      if (p != 0)
      {
         if (ip != j)
            for (k = 0; k < N; k++)
            {
               T t = A[(size_t)k*lda + j];
               A[(size_t)k*lda + j] = A[(size_t)k*lda + ip];
               A[(size_t)k*lda + ip] = t;
            }
         if (ORC_ABS(p) >= ORC_SAFMIN)
         {
            const T r = (T)1 / p;
            for (i = j + 1; i < M; i++) Ac[i] *= r;
         }
         else
            for (i = j + 1; i < M; i++) Ac[i] /= p;
      }
      else if (!iret) iret = j + 1;
      /* ATL_dger(M-j-1, N-j-1, -1, x=Ac+j+1, y=row j, A(j+1,j+1)) */
      for (k = j + 1; k < N; k++)
      {
         T *a = A + (size_t)k*lda;
         const T y = a[j];
         for (i = j + 1; i < M; i++) a[i] -= Ac[i] * y;
      }
   }
   return iret;
}

int F_(getf2)(int M, int N, T *A, int lda, int *ipiv)
{
   return F_(getf2_c)(M, N, A, lda, ipiv, 0);
}

/* src/lapack/ATL_getrfC.c:46-136 (LU2): two columns at once; the second
 * column's update and its pivot search ride along the first column's
 * scaling loop. */
static int F_(lu2)(int M, T *A, int lda, int *ipiv, int gcol)
{
   T *A1 = A + lda;
   T t0, t1, scal0, scal1, amax = 0;
   int i, ip, imax = -1, iret = 0;

   ip = F_(pivsearch)(M, A, gcol);
   ipiv[0] = ip;
   t0 = A[ip];
   if (t0 != 0)
   {
      const int recip = ORC_ABS(t0) >= ORC_SAFMIN;
      t1 = A1[ip];
      A[ip] = A[0]; A1[ip] = A1[0];
      A[0] = t0; A1[0] = t1;
      scal0 = recip ? (T)1 / t0 : t0;
      scal1 = -t1;
      for (i = 1; i < M; i++)
      {
         t0 = A[i]; t1 = A1[i];
         if (recip) t0 *= scal0; else t0 /= scal0;
         t1 += t0 * scal1;
         A[i] = t0; A1[i] = t1;
         if (t1 < 0) t1 = -t1;
         if (t1 > amax) { amax = t1; imax = i; }
      }
      if (orc_gaps_ && gcol + 1 < orc_ngaps_ && M > 1)
         (void)F_(pivsearch)(M - 1, A1 + 1, gcol + 1);
      if (amax != 0)
      {
         ipiv[1] = imax;
         t0 = A[imax]; t1 = A1[imax];
         A[imax] = A[1]; A1[imax] = A1[1];
         A[1] = t0; A1[1] = t1;
         if (ORC_ABS(t1) >= ORC_SAFMIN)
         {
            const T r = (T)1 / t1;
            for (i = 2; i < M; i++) A1[i] *= r;
         }
         else
            for (i = 2; i < M; i++) A1[i] /= t1;
      }
      else
      {
         ipiv[1] = (imax != -1) ? imax : 1;
         iret = 2;
      }
   }
   else
   {
      imax = 1 + F_(pivsearch)(M - 1, A1 + 1, gcol + 1);
      amax = A1[imax];
      iret = 1;
      if (amax != 0)
      {
         ipiv[1] = imax;
         t0 = A[imax]; t1 = A1[imax];
         A[imax] = A[1]; A1[imax] = A1[1];
         A[1] = t0; A1[1] = t1;
         if (ORC_ABS(t1) >= ORC_SAFMIN)
         {
            const T r = (T)1 / t1;
            for (i = 2; i < M; i++) A1[i] *= r;
         }
         else
            for (i = 2; i < M; i++) A1[i] /= t1;
      }
      else
         ipiv[1] = (imax != -1) ? imax : 1;
   }
   return iret;
}

/* src/lapack/ATL_getrfC.c:242-361: recursive LU.  Base case getf2 when
 * M*N <= L1C_ELTS (:264-265); otherwise split Nleft = MN/2 rounded down to
 * the kernel LCM (:279-284), factor the left part, laswp + TRSM(L,L,N,U) +
 * GEMM(-1,+1) on the right part, recurse on the bottom-right, offset its
 * pivots and swap the left columns (:285-300).  MN==2 uses LU2 (:308-321),
 * MN==1 scales by the reciprocal (:323-360). ATL_getrfC.c:333-334 divides
 * the first N (not M) elements in the tiny-pivot branch; this restatement
 * divides all M. */
static int F_(getrfC)(int M, int N, T *A, int lda, int *ipiv, int gcol)
{
   const int MN = M < N ? M : N;
   int Nleft, Nright, i, k, ierr = 0;
   if ((size_t)M * N <= (size_t)orc_l1c_elts_)
      return F_(getf2_c)(M, N, A, lda, ipiv, gcol);
   if (MN > 2)
   {
      T *Ac, *An;
      Nleft = MN >> 1;
      if (orc_lcm_ > 1 && Nleft > orc_lcm_) Nleft = (Nleft / orc_lcm_) * orc_lcm_;
      Nright = N - Nleft;
      i = F_(getrfC)(M, Nleft, A, lda, ipiv, gcol);
      if (i && !ierr) ierr = i;
      Ac = A + (size_t)Nleft * lda;
      An = Ac + Nleft;
      F_(laswp)(Nright, Ac, lda, 0, Nleft, ipiv, 1);
      F_(trsm_llnu)(Nleft, Nright, A, lda, Ac, lda);
      F_(gemm_fast)(ORC_NOTRANS, ORC_NOTRANS, M - Nleft, Nright, Nleft, (T)-1,
                    A + Nleft, lda, Ac, lda, (T)1, An, lda);
      i = F_(getrfC)(M - Nleft, Nright, An, lda, ipiv + Nleft, gcol + Nleft);
      if (i && !ierr) ierr = i + Nleft;
      for (i = Nleft; i < MN; i++) ipiv[i] += Nleft;
      F_(laswp)(Nleft, A, lda, Nleft, MN, ipiv, 1);
   }
   else if (MN == 2)
   {
      ierr = F_(lu2)(M, A, lda, ipiv, gcol);
      Nright = N - 2;
      if (Nright > 0)
      {
         T *Ac = A + (size_t)2 * lda;
         F_(laswp)(Nright, Ac, lda, 0, 2, ipiv, 1);
         F_(trsm_ref)(ORC_LEFT, ORC_LOWER, ORC_NOTRANS, ORC_UNIT, 2, Nright,
                      (T)1, A, lda, Ac, lda);
      }
   }
   else if (MN == 1)
   {
      T tmp;
      i = F_(pivsearch)(M, A, gcol);
      ipiv[0] = i;
      tmp = A[i];
      if (tmp != 0)
      {
         if (ORC_ABS(tmp) > ORC_SAFMIN)
         {
            const T r = (T)1 / tmp;
            for (k = 0; k < M; k++) A[k] *= r;
         }
         else
            for (k = 0; k < M; k++) A[k] /= tmp;
         A[i] = A[0];
         A[0] = tmp;
      }
      else ierr = 1;
   }
   return ierr;
}

/* src/lapack/ATL_getrf.c:8-13: ColMajor -> getrfC; RowMajor -> getrfR,
 * which (src/lapack/ATL_getrfR.c) is getrfC applied to the transposed view:
 * the row-major M x N array is the column-major N x M array A^T. */
int F_(getrf)(int order, int M, int N, T *A, int lda, int *ipiv)
{
   if (M <= 0 || N <= 0) return 0;
   if (order == ORC_COL) return F_(getrfC)(M, N, A, lda, ipiv, 0);
   return F_(getrfC)(N, M, A, lda, ipiv, 0);
}

/* src/lapack/ATL_getrs.c:8-89.  ColMajor NoTrans: laswp(+1), L-unit solve,
 * U solve.  ColMajor Trans: U^T solve, L^T-unit solve, laswp(-1).
 * RowMajor: the stored factor is that of A^T (see getrf), so RowMajor
 * NoTrans is the ColMajor Trans sequence on the same memory and vice versa
 * (ATL_getrs.c:67-87 spells the same operations with Right-side TRSMs). */
void F_(getrs)(int order, int trans, int N, int NRHS, const T *A, int lda,
               const int *ipiv, T *B, int ldb)
{
   int t = (trans != ORC_NOTRANS);
   if (N == 0 || NRHS == 0) return;
   if (order == ORC_ROW) t = !t;
   if (!t)
   {
      F_(laswp)(NRHS, B, ldb, 0, N, ipiv, 1);
      F_(trsm_ref)(ORC_LEFT, ORC_LOWER, ORC_NOTRANS, ORC_UNIT, N, NRHS, (T)1,
                   A, lda, B, ldb);
      F_(trsm_ref)(ORC_LEFT, ORC_UPPER, ORC_NOTRANS, ORC_NONUNIT, N, NRHS,
                   (T)1, A, lda, B, ldb);
   }
   else
   {
      F_(trsm_ref)(ORC_LEFT, ORC_UPPER, ORC_TRANS, ORC_NONUNIT, N, NRHS, (T)1,
                   A, lda, B, ldb);
      F_(trsm_ref)(ORC_LEFT, ORC_LOWER, ORC_TRANS, ORC_UNIT, N, NRHS, (T)1,
                   A, lda, B, ldb);
      F_(laswp)(NRHS, B, ldb, 0, N, ipiv, -1);
   }
}

/* ======================================================================= */
/* Cholesky                                                                */
/* ======================================================================= */

/* src/lapack/ATL_potrfL.c:104-202: closed-form leaves for N = 4, 3, 2.
 * They are restated as one loop that keeps the leaves' arithmetic: the
 * diagonal is sqrt'd, column 0 is scaled by the reciprocal, later columns
 * are formed as (a - sum)/l_jj except where the leaf multiplies by the
 * reciprocal (N=4: columns 0 and 1; N=3: column 0). */
static int F_(potrf_leafL)(int N, T *A, int lda)
{
#define L_(i_, j_) A[(size_t)(j_)*lda + (i_)]
   if (N == 1)
   {
      if (L_(0, 0) > 0) { L_(0, 0) = (T)ORC_SQRT(L_(0, 0)); return 0; }
      return 1;
   }
   if (N == 2)
   {
      T L11 = L_(0, 0), L21 = L_(1, 0), L22 = L_(1, 1);
      if (!(L11 > 0)) return 1;
      L_(0, 0) = L11 = (T)ORC_SQRT(L11);
      L_(1, 0) = L21 = L21 / L11;
      L22 -= L21 * L21;
      if (!(L22 > 0)) return 2;
      L_(1, 1) = (T)ORC_SQRT(L22);
      return 0;
   }
   if (N == 3)
   {
      T L11 = L_(0, 0), L21 = L_(1, 0), L31 = L_(2, 0);
      T L22 = L_(1, 1), L32 = L_(2, 1), L33 = L_(2, 2);
      if (!(L11 > 0)) return 1;
      L_(0, 0) = L11 = (T)ORC_SQRT(L11);
      L11 = (T)1 / L11;
      L21 *= L11; L31 *= L11;
      L_(1, 0) = L21; L_(2, 0) = L31;
      L22 -= L21 * L21;
      if (!(L22 > 0)) return 2;
      L22 = (T)ORC_SQRT(L22);
      L32 = (L32 - L31 * L21) / L22;
      L33 -= L31 * L31 + L32 * L32;
      L_(1, 1) = L22; L_(2, 1) = L32;
      if (!(L33 > 0)) return 3;
      L_(2, 2) = (T)ORC_SQRT(L33);
      return 0;
   }
   {  /* N == 4 */
      T L11 = L_(0, 0), L21 = L_(1, 0), L31 = L_(2, 0), L41 = L_(3, 0);
      T L22 = L_(1, 1), L32 = L_(2, 1), L42 = L_(3, 1);
      T L33 = L_(2, 2), L43 = L_(3, 2), L44 = L_(3, 3);
      if (!(L11 > 0)) return 1;
      L_(0, 0) = L11 = (T)ORC_SQRT(L11);
      L11 = (T)1 / L11;
      L21 *= L11; L31 *= L11; L41 *= L11;
      L_(1, 0) = L21; L_(2, 0) = L31; L_(3, 0) = L41;
      L22 -= L21 * L21;
      if (!(L22 > 0)) return 2;
      L_(1, 1) = L22 = (T)ORC_SQRT(L22);
      L22 = (T)1 / L22;
      L32 = (L32 - L31 * L21) * L22;
      L42 = (L42 - L41 * L21) * L22;
      L33 -= L31 * L31 + L32 * L32;
      L_(2, 1) = L32; L_(3, 1) = L42;
      if (!(L33 > 0)) return 3;
      L_(2, 2) = L33 = (T)ORC_SQRT(L33);
      L43 = (L43 - L41 * L31 - L42 * L32) / L33;
      L44 -= L41 * L41 + L42 * L42 + L43 * L43;
      L_(3, 2) = L43;
      if (!(L44 > 0)) return 4;
      L_(3, 3) = (T)ORC_SQRT(L44);
      return 0;
   }
#undef L_
}

/* src/lapack/ATL_potrfL.c:204-262: recursive lower Cholesky.  N > 4: split
 * Nleft = N/2 (rounded to the LCM when Nleft > 2*LCM), factor the left,
 * TRSM(R,L,T,NonUnit) on the bottom-left, SYRK(L,N,-1,+1) on the
 * bottom-right, recurse; info offset by Nleft and the factorization stops
 * at the first non-positive pivot. */
static int F_(potrfL)(int N, T *A, int lda)
{
   int Nleft, Nright, ierr;
   if (N > 4)
   {
      T *Ar, *An;
      Nleft = N >> 1;
      if (orc_lcm_ > 1 && Nleft > (orc_lcm_ << 1)) Nleft = (Nleft / orc_lcm_) * orc_lcm_;
      Nright = N - Nleft;
      ierr = F_(potrfL)(Nleft, A, lda);
      if (ierr) return ierr;
      Ar = A + Nleft;
      An = Ar + (size_t)lda * Nleft;
      F_(trsm_ref)(ORC_RIGHT, ORC_LOWER, ORC_TRANS, ORC_NONUNIT, Nright, Nleft,
                   (T)1, A, lda, Ar, lda);
      F_(syrk_ref)(ORC_LOWER, ORC_NOTRANS, Nright, Nleft, (T)-1, Ar, lda, (T)1,
                   An, lda);
      ierr = F_(potrfL)(Nright, An, lda);
      if (ierr) return ierr + Nleft;
      return 0;
   }
   if (N >= 1) return F_(potrf_leafL)(N, A, lda);
   return 0;
}

/* src/lapack/ATL_potrfU.c:199-257: the upper twin: TRSM(L,U,T,NonUnit) on
 * the top-right, SYRK(U,T,-1,+1) on the bottom-right.  Its closed-form
 * leaves (ATL_potrfU.c:104-197) are the transposes of the lower ones, so the
 * leaf is done by the lower leaf on a transposed copy. */
static int F_(potrfU)(int N, T *A, int lda)
{
   int Nleft, Nright, ierr;
   if (N > 4)
   {
      T *Ac, *An;
      Nleft = N >> 1;
      if (orc_lcm_ > 1 && Nleft > (orc_lcm_ << 1)) Nleft = (Nleft / orc_lcm_) * orc_lcm_;
      Nright = N - Nleft;
      ierr = F_(potrfU)(Nleft, A, lda);
      if (ierr) return ierr;
      Ac = A + (size_t)lda * Nleft;
      An = Ac + Nleft;
      F_(trsm_ref)(ORC_LEFT, ORC_UPPER, ORC_TRANS, ORC_NONUNIT, Nleft, Nright,
                   (T)1, A, lda, Ac, lda);
      F_(syrk_ref)(ORC_UPPER, ORC_TRANS, Nright, Nleft, (T)-1, Ac, lda, (T)1,
                   An, lda);
      ierr = F_(potrfU)(Nright, An, lda);
      if (ierr) return ierr + Nleft;
      return 0;
   }
   if (N >= 1)
   {
      T w[16];
      int i, j, r;
      for (j = 0; j < N; j++)
         for (i = 0; i <= j; i++) w[i * 4 + j] = A[(size_t)j*lda + i];
      r = F_(potrf_leafL)(N, w, 4);
      for (j = 0; j < N; j++)
         for (i = 0; i <= j; i++) A[(size_t)j*lda + i] = w[i * 4 + j];
      return r;
   }
   return 0;
}

/* src/lapack/ATL_potrf.c:7-30: ColMajor Upper->potrfU, Lower->potrfL;
 * RowMajor (real) swaps them. */
int F_(potrf)(int order, int uplo, int N, T *A, int lda)
{
   int low = (uplo == ORC_LOWER);
   if (N <= 0) return 0;
   if (order == ORC_ROW) low = !low;
   return low ? F_(potrfL)(N, A, lda) : F_(potrfU)(N, A, lda);
}

/* src/lapack/ATL_potrs.c:13-86.  ColMajor Upper: U^T then U solves; Lower:
 * L then L^T.  RowMajor stores the other triangle of the same symmetric
 * matrix, so it is the ColMajor sequence with Uplo flipped. */
void F_(potrs)(int order, int uplo, int N, int NRHS, const T *A, int lda,
               T *B, int ldb)
{
   int low = (uplo == ORC_LOWER);
   if (N == 0 || NRHS == 0) return;
   if (order == ORC_ROW) low = !low;
   if (!low)
   {
      F_(trsm_ref)(ORC_LEFT, ORC_UPPER, ORC_TRANS, ORC_NONUNIT, N, NRHS, (T)1,
                   A, lda, B, ldb);
      F_(trsm_ref)(ORC_LEFT, ORC_UPPER, ORC_NOTRANS, ORC_NONUNIT, N, NRHS,
                   (T)1, A, lda, B, ldb);
   }
   else
   {
      F_(trsm_ref)(ORC_LEFT, ORC_LOWER, ORC_NOTRANS, ORC_NONUNIT, N, NRHS,
                   (T)1, A, lda, B, ldb);
      F_(trsm_ref)(ORC_LEFT, ORC_LOWER, ORC_TRANS, ORC_NONUNIT, N, NRHS, (T)1,
                   A, lda, B, ldb);
   }
}

/* interfaces/lapack/C/src/clapack_dgesv.c:61-62: getrf then getrs unless
 * the factor is singular. */
int F_(gesv)(int order, int N, int NRHS, T *A, int lda, int *ipiv, T *B,
             int ldb)
{
   int info = F_(getrf)(order, N, N, A, lda, ipiv);
   if (!info) F_(getrs)(order, ORC_NOTRANS, N, NRHS, A, lda, ipiv, B, ldb);
   return info;
}

/* interfaces/lapack/C/src/clapack_dposv.c:57-58 */
int F_(posv)(int order, int uplo, int N, int NRHS, T *A, int lda, T *B,
             int ldb)
{
   int info = F_(potrf)(order, uplo, N, A, lda);
   if (!info) F_(potrs)(order, uplo, N, NRHS, A, lda, B, ldb);
   return info;
}

/* ======================================================================= */
/* Testers                                                                 */
/* ======================================================================= */

/* bin/lutst.c:84-134 (ATL_LmulUC) + :238-245: form L*U from the packed
 * factor, apply the pivots to a copy of the original A, and return
 * ||P*A - L*U||_1 / (||A||_1 * eps * min(M,N)). */
T F_(lu_resid)(int M, int N, const T *LU, int lda, const int *ipiv,
               const T *A0, int lda0)
{
   const int MN = M < N ? M : N;
   T *C = (T *)calloc((size_t)M * N, sizeof(T));
   T *PA = (T *)malloc((size_t)M * N * sizeof(T));
   T nrmA, eps, r;
   int i, j;
   if (M >= N)
   {
      for (j = 0; j < MN; j++)
      {
         C[(size_t)j*M + j] = 1;
         for (i = j + 1; i < M; i++) C[(size_t)j*M + i] = LU[(size_t)j*lda + i];
      }
      F_(trmm_fast)(ORC_RIGHT, ORC_UPPER, ORC_NOTRANS, ORC_NONUNIT, M, N, (T)1,
                    LU, lda, C, M);
   }
   else
   {
      for (j = 0; j < N; j++)
         for (i = 0; i <= (j < M - 1 ? j : M - 1); i++)
            C[(size_t)j*M + i] = LU[(size_t)j*lda + i];
      F_(trmm_fast)(ORC_LEFT, ORC_LOWER, ORC_NOTRANS, ORC_UNIT, M, N, (T)1,
                    LU, lda, C, M);
   }
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++) PA[(size_t)j*M + i] = A0[(size_t)j*lda0 + i];
   nrmA = F_(genrm1)(M, N, PA, M);
   F_(laswp)(N, PA, M, 0, MN, ipiv, 1);
   r = F_(gediffnrm1)(M, N, PA, M, C, M);
   eps = F_(epsilon)();
   free(C);
   free(PA);
   return r / (nrmA * eps * (T)MN);
}

/* bin/lutst.c:199-253 (lutestC): gegen(seed M*N+lda), getrf, residual, and
 * the count of non-trivial pivots (findnpvt, :136-141). */
T F_(lutest)(int M, int N, int lda, int *npiv)
{
   const int MN = M < N ? M : N;
   T *A = (T *)malloc((size_t)lda * N * sizeof(T));
   T *A0 = (T *)malloc((size_t)lda * N * sizeof(T));
   int *ipiv = (int *)malloc(sizeof(int) * (MN > 0 ? MN : 1));
   int i, n = 0;
   T r;
   F_(gegen)(M, N, A, lda, M * N + lda);
   memcpy(A0, A, (size_t)lda * N * sizeof(T));
   F_(getrf)(ORC_COL, M, N, A, lda, ipiv);
   r = F_(lu_resid)(M, N, A, lda, ipiv, A0, lda);
   for (i = 0; i < MN; i++) if (ipiv[i] != i) n++;
   if (npiv) *npiv = n;
   free(A); free(A0); free(ipiv);
   return r;
}

/* bin/llttst.c:166-191 (lltnrm1, real): 1-norm of the symmetric matrix
 * held in one triangle. */
T F_(lltnrm1)(int uplo, int N, const T *A, int lda)
{
   int i, j;
   T t, nrm = 0;
   if (uplo == ORC_UPPER)
      for (j = 0; j < N; j++)
      {
         t = 0;
         for (i = 0; i < j; i++) t += ORC_ABS(A[(size_t)j*lda + i]);
         for (i = j; i < N; i++) t += ORC_ABS(A[(size_t)i*lda + j]);
         if (t > nrm) nrm = t;
      }
   else
      for (i = 0; i < N; i++)
      {
         t = 0;
         for (j = 0; j < i; j++) t += ORC_ABS(A[(size_t)j*lda + i]);
         for (j = i; j < N; j++) t += ORC_ABS(A[(size_t)i*lda + j]);
         if (t > nrm) nrm = t;
      }
   return nrm;
}

/* bin/llttst.c:225-259 (lltgen): gegen, zero (PADVAL) the unreferenced
 * triangle, then diagonal = |5.3*a_jj| + lltnrm1(A). */
void F_(lltgen)(int uplo, int N, T *A, int lda, int seed)
{
   int i, j;
   T t0, t1;
   F_(gegen)(N, N, A, lda, seed);
   for (j = 0; j < N; j++)
   {
      if (uplo == ORC_UPPER)
         for (i = j + 1; i < N; i++) A[(size_t)j*lda + i] = 0;
      else
         for (i = 0; i < j; i++) A[(size_t)j*lda + i] = 0;
   }
   t1 = F_(lltnrm1)(uplo, N, A, lda);
   for (j = 0; j < N; j++)
   {
      t0 = A[(size_t)j*lda + j] * (T)5.3;
      if (t0 < 0) t0 = -t0;
      A[(size_t)j*lda + j] = t0 + t1;
   }
}

/* bin/llttst.c:23-101 (ATL_LmulLt / ATL_UtmulU), :260-290 (lltdiff) and
 * :329-344: ||A - L*L^T||_1 / (||A||_1 * eps * N) on the referenced
 * triangle. */
T F_(llt_resid)(int uplo, int N, const T *L, int lda, const T *A0, int lda0)
{
   T *C = (T *)calloc((size_t)N * N, sizeof(T));
   T nrmA, r, eps;
   int i, j, k;
   for (j = 0; j < N; j++)
   {
      if (uplo == ORC_LOWER)
         for (i = j; i < N; i++)
         {
            T s = L[(size_t)j*lda + i] * L[(size_t)j*lda + j];
            for (k = 0; k < j; k++)
               s += L[(size_t)k*lda + i] * L[(size_t)k*lda + j];
            C[(size_t)j*N + i] = s - A0[(size_t)j*lda0 + i];
         }
      else
         for (i = 0; i <= j; i++)
         {
            T s = 0;
            for (k = 0; k <= i; k++)
               s += L[(size_t)i*lda + k] * L[(size_t)j*lda + k];
            C[(size_t)j*N + i] = s - A0[(size_t)j*lda0 + i];
         }
   }
   nrmA = F_(lltnrm1)(uplo, N, A0, lda0);
   r = F_(lltnrm1)(uplo, N, C, N);
   eps = F_(epsilon)();
   free(C);
   return r / (nrmA * eps * (T)N);
}

/* bin/llttst.c:292-350 (llttest): lltgen(seed N*1029+lda), potrf, check
 * the unreferenced triangle kept its PADVAL (0), residual. Returns -1 if
 * potrf reported failure or the pad was overwritten. */
T F_(llttest)(int uplo, int N, int lda)
{
   T *A = (T *)malloc((size_t)lda * N * sizeof(T));
   T *A0 = (T *)malloc((size_t)lda * N * sizeof(T));
   T r;
   int i, j, bad = 0;
   F_(lltgen)(uplo, N, A, lda, N * 1029 + lda);
   memcpy(A0, A, (size_t)lda * N * sizeof(T));
   if (F_(potrf)(ORC_COL, uplo, N, A, lda)) bad = 1;
   for (j = 0; j < N; j++)
      for (i = 0; i < N; i++)
      {
         int unref = (uplo == ORC_LOWER) ? (i < j) : (i > j);
         if (unref && A[(size_t)j*lda + i] != 0) bad = 1;
      }
   r = bad ? (T)-1 : F_(llt_resid)(uplo, N, A, lda, A0, lda);
   free(A); free(A0);
   return r;
}

/* bin/slvtst.c:419-441: max_j ||x_j - xc_j||_inf / (N*eps*||A||*||xc_j||_inf) */
T F_(slv_resid)(int N, int NRHS, T nrmA, const T *XC, int ldxc, const T *X,
                int ldx)
{
   const T eps = F_(epsilon)();
   T mx = 0, t, nx, nd;
   int j, i;
   for (j = 0; j < NRHS; j++)
   {
      nx = F_(infnrm)(N, XC + (size_t)j*ldxc, 1);
      nd = 0;
      for (i = 0; i < N; i++)
      {
         T d = ORC_ABS(XC[(size_t)j*ldxc + i] - X[(size_t)j*ldx + i]);
         if (d > nd) nd = d;
      }
      t = nd / ((T)N * eps * nrmA * nx);
      if (t > mx) mx = t;
   }
   return mx;
}

/* bin/gemmtst.c:393-445: elementwise |D - C| <= ferr with
 * ferr = 2*(|alpha|*2K*EPS + |beta|*EPS) + EPS, EPS = 1e-15 (double) or
 * 5e-7 (float) (include/atlas_misc.h:107,122).  NaN in D fails.  Returns
 * the number of failing elements. */
int F_(gemm_ferr_check)(int M, int N, int K, T alpha, T beta, const T *C,
                        int ldc, const T *D, int ldd, double *maxdiff,
                        double *ferr)
{
   const double eps = ORC_GEMMTST_EPS;
   const double fe = 2.0 * (ORC_ABS((double)alpha) * 2.0 * K * eps +
                            ORC_ABS((double)beta) * eps) + eps;
   double mx = 0.0;
   int i, j, bad = 0;
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++)
      {
         const double d = D[(size_t)j*ldd + i], c = C[(size_t)j*ldc + i];
         double f = d - c;
         if (f < 0) f = -f;
         if (f > fe || d != d) { bad++; if (f > mx || d != d) mx = f; }
         else if (f > mx) mx = f;
      }
   if (maxdiff) *maxdiff = mx;
   if (ferr) *ferr = fe;
   return bad;
}

/* doc/TestTime.txt:333-345 (l3blastst TRSM check):
 * ||alpha*B - op(A)*X||_1 / (||A||_1 * ||X||_1 * |alpha| * eps * max(M,N)),
 * with A the triangle actually used (unit diagonal when Diag=Unit). */
T F_(trsm_resid)(int side, int uplo, int ta, int diag, int M, int N, T alpha,
                 const T *A, int lda, const T *X, int ldx, const T *B0,
                 int ldb0)
{
   const int na = (side == ORC_LEFT) ? M : N;
   T *R = (T *)malloc((size_t)M * N * sizeof(T));
   T *Tm = (T *)calloc((size_t)na * na, sizeof(T));
   T nA, nX, nR, eps;
   int i, j;
   for (j = 0; j < na; j++)
      for (i = 0; i < na; i++)
      {
         int in = (uplo == ORC_LOWER) ? (i > j) : (i < j);
         T v = in ? A[(size_t)j*lda + i] : 0;
         if (i == j) v = (diag == ORC_UNIT) ? (T)1 : A[(size_t)j*lda + i];
         Tm[(size_t)j*na + i] = v;
      }
   for (j = 0; j < N; j++)
      for (i = 0; i < M; i++) R[(size_t)j*M + i] = alpha * B0[(size_t)j*ldb0 + i];
   if (side == ORC_LEFT)
      F_(gemm_fast)(ta == ORC_NOTRANS ? ORC_NOTRANS : ORC_TRANS, ORC_NOTRANS,
                    M, N, M, (T)-1, Tm, na, X, ldx, (T)1, R, M);
   else
      F_(gemm_fast)(ORC_NOTRANS, ta == ORC_NOTRANS ? ORC_NOTRANS : ORC_TRANS,
                    M, N, N, (T)-1, X, ldx, Tm, na, (T)1, R, M);
   nA = F_(genrm1)(na, na, Tm, na);
   nX = F_(genrm1)(M, N, X, ldx);
   nR = F_(genrm1)(M, N, R, M);
   eps = F_(epsilon)();
   free(R); free(Tm);
   return nR / (nA * nX * ORC_ABS(alpha) * eps * (T)(M > N ? M : N));
}


/* ======================================================================= */
/* Testers of the inverses (bin/trtritst.c, bin/invtst.c, bin/uumtst.c)    */
/* ======================================================================= */

/* bin/trtritst.c:129-160 (trinrm1): max column abs-sum over the triangle,
 * the diagonal left out when unit (column-major). */
T F_(trinrm1)(int uplo, int diag, int N, const T *A, int lda)
{
   const int skip = (diag == ORC_UNIT);
   int i, j;
   T mx = 0, t;
   for (j = 0; j < N; j++)
   {
      t = 0;
      if (uplo == ORC_UPPER)
         for (i = 0; i < j + 1 - skip; i++) t += ORC_ABS(A[(size_t)j*lda + i]);
      else
         for (i = j + skip; i < N; i++) t += ORC_ABS(A[(size_t)j*lda + i]);
      if (t > mx) mx = t;
   }
   return mx;
}

/* bin/trtritst.c:162-258 (trddom + trigen, column-major): gegen, the
 * unreferenced part (and the diagonal when unit) set to padval, then the
 * off-diagonal triangle scaled by 0.9/N (times the smallest |diagonal|
 * when non-unit) so the matrix is diagonally dominant. */
void F_(trigen)(int uplo, int diag, int N, T *A, int lda, T padval, int seed)
{
   const int skipdiag = (diag == ORC_NONUNIT);
   int i, j;
   T alpha, smallest;
   F_(gegen)(N, N, A, lda, seed);
   for (j = 0; j < N; j++)
   {
      if (uplo == ORC_UPPER)
         for (i = j + skipdiag; i < N; i++) A[(size_t)j*lda + i] = padval;
      else
         for (i = 0; i < j + 1 - skipdiag; i++) A[(size_t)j*lda + i] = padval;
   }
   if (N <= 0) return;
   alpha = (T)0.9 / (T)N;
   if (diag == ORC_NONUNIT)
   {
      smallest = ORC_ABS(A[0]);
      for (i = 1; i < N; i++)
         if (ORC_ABS(A[(size_t)i*lda + i]) < smallest) smallest = ORC_ABS(A[(size_t)i*lda + i]);
      alpha *= smallest;
   }
   for (j = 0; j < N; j++)
   {
      if (uplo == ORC_UPPER)
         for (i = 0; i < j; i++) A[(size_t)j*lda + i] *= alpha;
      else
         for (i = j + 1; i < N; i++) A[(size_t)j*lda + i] *= alpha;
   }
}

/* bin/trtritst.c:300-397 (trtritest's residual): ||inv(A)*A - I||_1 /
 * (trinrm1(A) * eps * N), the product by the reference-order TRMM of the
 * computed inverse (uplo/diag triangle of Ainv) times the original
 * triangle expanded with zeros (and a unit diagonal when unit). */
T F_(trtri_resid)(int uplo, int diag, int N, const T *Ainv, int ldi, const T *A0, int lda0)
{
   T *C = (T *)calloc((size_t)N * N > 0 ? (size_t)N * N : 1, sizeof(T));
   T nrm, r;
   int i, j;
   for (j = 0; j < N; j++)
      for (i = 0; i < N; i++)
      {
         const int in = (uplo == ORC_LOWER) ? (i >= j) : (i <= j);
         C[(size_t)j*N + i] = in ? A0[(size_t)j*lda0 + i] : 0;
         if (i == j && diag == ORC_UNIT) C[(size_t)j*N + i] = 1;
      }
   nrm = F_(trinrm1)(uplo, diag, N, A0, lda0);
   F_(trmm_ref)(ORC_LEFT, uplo, ORC_NOTRANS, diag, N, N, (T)1, Ainv, ldi, C, N);
   for (i = 0; i < N; i++) C[(size_t)i*N + i] -= 1;
   r = F_(genrm1)(N, N, C, N);
   free(C);
   return r / (nrm * F_(epsilon)() * (T)N);
}

/* bin/invtst.c:388-418 (geresid): ||A*AI - I||_1 / (N eps ||A||_1 ||AI||_1) */
T F_(inv_resid)(int N, const T *A, int lda, const T *AI, int ldi)
{
   T *C = (T *)calloc((size_t)N * N > 0 ? (size_t)N * N : 1, sizeof(T));
   T num, den;
   int i;
   F_(gemm_fast)(ORC_NOTRANS, ORC_NOTRANS, N, N, N, (T)1, A, lda, AI, ldi, (T)0, C, N);
   for (i = 0; i < N; i++) C[(size_t)i*N + i] -= 1;
   num = F_(genrm1)(N, N, C, N);
   den = F_(genrm1)(N, N, A, lda) * F_(genrm1)(N, N, AI, ldi) * (T)N * F_(epsilon)();
   free(C);
   return num / den;
}

/* bin/invtst.c:420-463 (poresid): both symmetric matrices held in their
 * uplo triangle, ||A*AI - I||_1 / (N eps synrm(A) synrm(AI)) with synrm
 * the symmetric 1-norm (the same sum as lltnrm1). */
T F_(poinv_resid)(int uplo, int N, const T *A, int lda, const T *AI, int ldi)
{
   T *Sa = (T *)calloc((size_t)N * N > 0 ? (size_t)N * N : 1, sizeof(T));
   T *Si = (T *)calloc((size_t)N * N > 0 ? (size_t)N * N : 1, sizeof(T));
   T *C = (T *)calloc((size_t)N * N > 0 ? (size_t)N * N : 1, sizeof(T));
   T num, den;
   int i, j;
   for (j = 0; j < N; j++)
      for (i = 0; i < N; i++)
      {
         const int in = (uplo == ORC_LOWER) ? (i >= j) : (i <= j);
         Sa[(size_t)j*N + i] = in ? A[(size_t)j*lda + i] : A[(size_t)i*lda + j];
         Si[(size_t)j*N + i] = in ? AI[(size_t)j*ldi + i] : AI[(size_t)i*ldi + j];
      }
   F_(gemm_fast)(ORC_NOTRANS, ORC_NOTRANS, N, N, N, (T)1, Sa, N, Si, N, (T)0, C, N);
   for (i = 0; i < N; i++) C[(size_t)i*N + i] -= 1;
   num = F_(genrm1)(N, N, C, N);
   den = F_(lltnrm1)(uplo, N, A, lda) * F_(lltnrm1)(uplo, N, AI, ldi) * (T)N * F_(epsilon)();
   free(Sa); free(Si); free(C);
   return num / den;
}

/* bin/uumtst.c:300-351 (uumtest's residual): the lauum result R (uplo
 * triangle) against U*U^T (upper) or L^T*L (lower) of the original
 * triangle T0, lltnrm1 of the difference / (lltnrm1(T0) * eps * N). */
T F_(lauum_resid)(int uplo, int N, const T *R, int ldr, const T *T0, int ld0)
{
   T *Tt = (T *)calloc((size_t)N * N > 0 ? (size_t)N * N : 1, sizeof(T));
   T *P = (T *)calloc((size_t)N * N > 0 ? (size_t)N * N : 1, sizeof(T));
   T num, nrm;
   int i, j;
   for (j = 0; j < N; j++)
      for (i = 0; i < N; i++)
      {
         const int in = (uplo == ORC_LOWER) ? (i >= j) : (i <= j);
         Tt[(size_t)j*N + i] = in ? T0[(size_t)j*ld0 + i] : 0;
      }
   if (uplo == ORC_LOWER)
      F_(gemm_fast)(ORC_TRANS, ORC_NOTRANS, N, N, N, (T)1, Tt, N, Tt, N, (T)0, P, N);
   else
      F_(gemm_fast)(ORC_NOTRANS, ORC_TRANS, N, N, N, (T)1, Tt, N, Tt, N, (T)0, P, N);
   for (j = 0; j < N; j++)
      for (i = 0; i < N; i++)
      {
         const int in = (uplo == ORC_LOWER) ? (i >= j) : (i <= j);
         if (in) P[(size_t)j*N + i] = R[(size_t)j*ldr + i] - P[(size_t)j*N + i];
      }
   num = F_(lltnrm1)(uplo, N, P, N);
   nrm = F_(lltnrm1)(uplo, N, T0, ld0);
   free(Tt); free(P);
   return num / (nrm * F_(epsilon)() * (T)N);
}

/* ======================================================================= */
/* QR (src/lapack/ATL_geqr2.c, ATL_larfg.c) and bin/qrtst.c's residual      */
/* ======================================================================= */

/* 2-norm with a running scale (the reference BLAS nrm2 algorithm) */
static T F_(nrm2)(int N, const T *X, int incX)
{
   T scale = 0, ssq = 1, a;
   int i;
   for (i = 0; i < N; i++)
   {
      a = ORC_ABS(X[(size_t)i*incX]);
      if (a != 0)
      {
         if (scale < a) { ssq = 1 + ssq * (scale / a) * (scale / a); scale = a; }
         else ssq += (a / scale) * (a / scale);
      }
   }
   return scale * (T)sqrt((double)ssq);
}

/* src/lapack/ATL_larfg.c (real): H (alpha; x) = (beta; 0) with
 * H = I - tau (1; v)(1; v)^T, beta = -sign(alpha) hypot(alpha, ||x||) */
static void F_(larfg)(int N, T *alpha, T *X, int incX, T *tau)
{
   const T safmin = ORC_SAFMIN / F_(epsilon)();
   T xnorm, beta, betap, rsafmn;
   int j, knt = 0;
   if (N < 1) { *tau = 0; return; }
   xnorm = F_(nrm2)(N - 1, X, incX);
   if (xnorm == 0) { *tau = 0; return; }
   betap = (T)hypot((double)*alpha, (double)xnorm);
   beta = (*alpha >= 0) ? -betap : betap;
   if (betap < safmin)
   {
      rsafmn = 1 / safmin;
      do
      {
         knt++;
         for (j = 0; j < N - 1; j++) X[(size_t)j*incX] *= rsafmn;
         beta *= rsafmn;
         *alpha *= rsafmn;
         betap = ORC_ABS(beta);
      } while (betap < safmin && knt < 20);
      xnorm = F_(nrm2)(N - 1, X, incX);
      beta = (T)hypot((double)*alpha, (double)xnorm);
      if (*alpha > 0) beta = -beta;
   }
   *tau = (beta - *alpha) / beta;
   for (j = 0; j < N - 1; j++) X[(size_t)j*incX] *= 1 / (*alpha - beta);
   for (j = 0; j < knt; j++) beta *= ORC_SAFMIN;
   *alpha = beta;
}

/* src/lapack/ATL_geqr2.c: unblocked Householder QR, A(i:M, i+1:N) -=
 * tau v (v^T A(i:M, i+1:N)) after each column's larfg. */
void F_(geqr2)(int M, int N, T *A, int lda, T *tau)
{
   const int K = M < N ? M : N;
   int i, j, r;
   for (i = 0; i < K; i++)
   {
      T *col = A + (size_t)i*lda + i;
      F_(larfg)(M - i, col, col + 1, 1, tau + i);
      if (tau[i] == 0) continue;
      for (j = i + 1; j < N; j++)
      {
         T *cj = A + (size_t)j*lda + i, w = cj[0];
         for (r = 1; r < M - i; r++) w += col[r] * cj[r];
         w *= tau[i];
         cj[0] -= w;
         for (r = 1; r < M - i; r++) cj[r] -= w * col[r];
      }
   }
}

/* bin/qrtst.c:1105-1307 (the QR case): Q R rebuilt from the factored
 * storage (R on and above the diagonal, reflectors below) by applying
 * H_{k-1} .. H_0 to R, then ||A0 - Q R||_1 / (||A0||_1 eps min(M, N)). */
T F_(qr_resid)(int M, int N, const T *QR, int lda, const T *tau, const T *A0, int lda0)
{
   const int K = M < N ? M : N;
   T *W = (T *)calloc((size_t)M * N > 0 ? (size_t)M * N : 1, sizeof(T));
   T r, nrm;
   int i, j, q;
   for (j = 0; j < N; j++)
      for (i = 0; i <= (j < M - 1 ? j : M - 1); i++) W[(size_t)j*M + i] = QR[(size_t)j*lda + i];
   for (i = K - 1; i >= 0; i--)
   {
      const T *v = QR + (size_t)i*lda + i;
      if (tau[i] == 0) continue;
      for (j = 0; j < N; j++)
      {
         T *wj = W + (size_t)j*M + i, s = wj[0];
         for (q = 1; q < M - i; q++) s += v[q] * wj[q];
         s *= tau[i];
         wj[0] -= s;
         for (q = 1; q < M - i; q++) wj[q] -= s * v[q];
      }
   }
   nrm = F_(genrm1)(M, N, A0, lda0);
   r = F_(gediffnrm1)(M, N, A0, lda0, W, M);
   free(W);
   return r / (nrm * F_(epsilon)() * (T)(K > 0 ? K : 1));
}

#undef F_

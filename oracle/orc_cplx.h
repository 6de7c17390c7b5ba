/*
 * ORACLE -- test infrastructure only (see oracle.h).
 *
 * Complex precision template: included by oracle.c with T=double, P=orc_z
 * and with T=float, P=orc_c.  Matrices are column-major with the real and
 * imaginary parts interleaved (ATLAS's complex storage); leading dimensions
 * count complex elements; alpha and beta point at {real, imaginary}.
 */
#define F_(name) CAT_(P, name)

/* One complex multiply-add, c += a*b, in the operation order of the
 * reference's Mmla (include/atlas_refmisc.h:170-174). */
#define ORC_CMLA(ar, ai, br, bi, cr, ci)      \
   do {                                     \
      (cr) += (ar) * (br) - (ai) * (bi);    \
      (ci) += (ar) * (bi) + (ai) * (br);    \
   } while (0)

/* c = beta*c as the reference's Mzelscal (include/atlas_refmisc.h:233-239):
 * beta = 0 stores zeros, beta = 1 leaves c alone, else Mdscl. */
static void F_(celscal)(const T *beta, T *c)
{
   if (beta[0] == 0 && beta[1] == 0) { c[0] = c[1] = 0; return; }
   if (beta[0] == 1 && beta[1] == 0) return;
   {
      const T r = beta[0] * c[0] - beta[1] * c[1];
      const T i = beta[0] * c[1] + beta[1] * c[0];
      c[0] = r; c[1] = i;
   }
}

/* src/blas/reference/level3/ATL_zrefgemm.c (dispatcher, quick returns,
 * alpha = 0 -> Mzgescal(C, beta)) and the loop orders of
 * ATL_zrefgemm{NN,NT,NC}.c (op(A) = A: per column j, C(:,j) = beta C(:,j),
 * then for l = 0..K-1, t = alpha*op(B)(l,j) by Mmul and C(:,j) += A(:,l) t)
 * and ATL_zrefgemm{T,C}{N,T,C}.c (op(A) = A^T or A^H: per (i,j) the dot
 * product t = sum_l op(A)(i,l) op(B)(l,j) from zero, then C(i,j) =
 * beta C(i,j) by Mzelscal and C(i,j) += alpha t).  Conjugation negates the
 * imaginary part before the product, as the reference's -A[..+1]. */
void F_(gemm_ref)(int TA, int TB, int M, int N, int K, const T *alpha, const T *A, int lda,
                   const T *B, int ldb, const T *beta, T *C, int ldc)
{
   const int ta = (TA != ORC_NOTRANS), ca = (TA == ORC_CONJTRANS);
   const int tb = (TB != ORC_NOTRANS), cb = (TB == ORC_CONJTRANS);
   const int a_zero = (alpha[0] == 0 && alpha[1] == 0);
   int i, j, l;
   if (M == 0 || N == 0 ||
       ((a_zero || K == 0) && beta[0] == 1 && beta[1] == 0)) return;
   if (a_zero)
   {
      for (j = 0; j < N; j++)
         for (i = 0; i < M; i++) F_(celscal)(beta, C + 2*((size_t)j*ldc + i));
      return;
   }
   if (!ta)
   {
      for (j = 0; j < N; j++)
      {
         T *c = C + 2*(size_t)j*ldc;
         for (i = 0; i < M; i++) F_(celscal)(beta, c + 2*i);
         for (l = 0; l < K; l++)
         {
            const T *a = A + 2*(size_t)l*lda;
            const T *b = tb ? B + 2*((size_t)l*ldb + j) : B + 2*((size_t)j*ldb + l);
            const T br = b[0], bi = cb ? -b[1] : b[1];
            const T tr = alpha[0] * br - alpha[1] * bi;
            const T ti = alpha[0] * bi + alpha[1] * br;
            for (i = 0; i < M; i++) ORC_CMLA(a[2*i], a[2*i+1], tr, ti, c[2*i], c[2*i+1]);
         }
      }
   }
   else
   {
      for (j = 0; j < N; j++)
         for (i = 0; i < M; i++)
         {
            const T *a = A + 2*(size_t)i*lda;
            T *c = C + 2*((size_t)j*ldc + i);
            T tr = 0, ti = 0;
            for (l = 0; l < K; l++)
            {
               const T *b = tb ? B + 2*((size_t)l*ldb + j) : B + 2*((size_t)j*ldb + l);
               const T ar = a[2*l], ai = ca ? -a[2*l+1] : a[2*l+1];
               const T br = b[0], bi = cb ? -b[1] : b[1];
               ORC_CMLA(ar, ai, br, bi, tr, ti);
            }
            F_(celscal)(beta, c);
            ORC_CMLA(alpha[0], alpha[1], tr, ti, c[0], c[1]);
         }
   }
}

/* (xr + i xi) / (dr + i di) by Smith's algorithm, the form of the
 * reference's Mddiv (include/atlas_refmisc.h:131-155; D. Knuth, TAOCP
 * vol. 2, p. 195). */
static void F_(cdiv)(T dr, T di, T *xr, T *xi)
{
   T r, den, cr, ci;
   if (ORC_ABS(di) < ORC_ABS(dr))
   {
      r = di / dr; den = dr + di * r;
      cr = (*xr + *xi * r) / den;
      ci = (*xi - *xr * r) / den;
   }
   else
   {
      r = dr / di; den = di + dr * r;
      cr = (*xi + *xr * r) / den;
      ci = (-*xr + *xi * r) / den;
   }
   *xr = cr; *xi = ci;
}

/* The effective left operand of a solve: E(i,k) = op(A)(i,k) for Left, and
 * op(A)(k,i) (so that X op(A) = B is E X^T = B^T) for Right. */
static void F_(trsm_elem)(const T *A, int lda, int tr, int cj, int i, int k, T *er, T *ei)
{
   const T *a = tr ? A + 2*((size_t)i*lda + k) : A + 2*((size_t)k*lda + i);
   *er = a[0];
   *ei = cj ? -a[1] : a[1];
}

/* Complex TRSM, all 24 cases of src/blas/reference/level3/ATL_zreftrsm.c
 * (dispatcher: M = 0 or N = 0 returns, alpha = 0 zeroes B).  Each
 * right-hand side (a column of B for Left, a row for Right) is scaled by
 * alpha and solved against the effective triangle E by substitution in dot
 * form, dividing by the diagonal with Smith's algorithm as Mddiv.  The
 * value, not the reference's operation order, is restated: the tests pin it
 * against ATL_zreftrsm (oracle/_ref) and numpy within rounding. */
void F_(trsm_ref)(int side, int uplo, int TA, int diag, int M, int N, const T *alpha,
                   const T *A, int lda, T *B, int ldb)
{
   const int left = (side == ORC_LEFT);
   const int na = left ? M : N, nr = left ? N : M;
   const int cj = (TA == ORC_CONJTRANS);
   /* E = op(A) (Left) or op(A)^T (Right) in terms of A: transposed? */
   const int tr = left ? (TA != ORC_NOTRANS) : (TA == ORC_NOTRANS);
   const int lower = ((uplo == ORC_LOWER) != tr);
   int r, i, k;
   if (M == 0 || N == 0) return;
   for (r = 0; r < nr; r++)
   {
      /* element q of this right-hand side */
#define XE(q) (left ? B + 2*((size_t)r*ldb + (q)) : B + 2*((size_t)(q)*ldb + r))
      for (i = 0; i < na; i++)
      {
         T *x = XE(i);
         if (alpha[0] == 0 && alpha[1] == 0) { x[0] = x[1] = 0; continue; }
         {
            const T xr = alpha[0] * x[0] - alpha[1] * x[1];
            const T xi = alpha[0] * x[1] + alpha[1] * x[0];
            x[0] = xr; x[1] = xi;
         }
      }
      if (alpha[0] == 0 && alpha[1] == 0) continue;
      for (k = 0; k < na; k++)
      {
         const int ii = lower ? k : na - 1 - k;
         T *x = XE(ii), sr = x[0], si = x[1], er, ei;
         const int j0 = lower ? 0 : ii + 1, j1 = lower ? ii : na;
         int jj;
         for (jj = j0; jj < j1; jj++)
         {
            const T *y = XE(jj);
            F_(trsm_elem)(A, lda, tr, cj, ii, jj, &er, &ei);
            sr -= er * y[0] - ei * y[1];
            si -= er * y[1] + ei * y[0];
         }
         if (diag != ORC_UNIT)
         {
            F_(trsm_elem)(A, lda, tr, cj, ii, ii, &er, &ei);
            F_(cdiv)(er, ei, &sr, &si);
         }
         x[0] = sr; x[1] = si;
      }
#undef XE
   }
}

/* Complex LU, src/lapack/ATL_getf2.c:60-95 with LU1 (:10-58), complex
 * branch.  Per column j: the pivot is the first row of largest |Re| + |Im|
 * at or below j (cblas_izamax = ATL_izrefamax.c: strict >, starting from 0,
 * so NaN never wins and an all-zero column gives row j); ipiv[j] = that
 * row.  A nonzero pivot (|Re| + |Im| != 0) swaps rows across all N columns
 * and the rows below are multiplied by its reciprocal (cplxinvert, here
 * Smith's division of 1), or divided by it when |Re| + |Im| < SAFMIN; a
 * zero pivot makes the return value j + 1 if none was zero before.  The
 * rank-1 update (geru with -1) of the trailing block follows in every case.
 * mingap (may be NULL) receives the smallest relative margin
 * (best - runner-up) / best between the pivot and the next candidate over
 * all columns: tests compare pivots only where it is comfortable. */
int F_(getf2)(int M, int N, T *A, int lda, int *ipiv, T *mingap)
{
   const int MN = M < N ? M : N;
   int i, j, c, info = 0;
   T gap = 1;
   for (j = 0; j < MN; j++)
   {
      T *Aj = A + 2*(size_t)j*lda, best = 0, second = 0, pr, pi, pv;
      int p = j;
      for (i = j; i < M; i++)
      {
         const T v = ORC_ABS(Aj[2*i]) + ORC_ABS(Aj[2*i+1]);
         if (v > best) { second = best; best = v; p = i; }
         else if (v > second) second = v;
      }
      if (best > 0 && M - j > 1 && (best - second) / best < gap) gap = (best - second) / best;
      ipiv[j] = p;
      pr = Aj[2*p]; pi = Aj[2*p+1];
      pv = ORC_ABS(pr) + ORC_ABS(pi);
      if (pv != 0)
      {
         if (p != j)
            for (c = 0; c < N; c++)
            {
               T *a = A + 2*((size_t)c*lda + j), *b = A + 2*((size_t)c*lda + p), t;
               t = a[0]; a[0] = b[0]; b[0] = t;
               t = a[1]; a[1] = b[1]; b[1] = t;
            }
         if (pv >= ORC_SAFMIN)
         {
            T rr = 1, ri = 0;
            F_(cdiv)(pr, pi, &rr, &ri);
            for (i = j + 1; i < M; i++)
            {
               const T xr = Aj[2*i], xi = Aj[2*i+1];
               Aj[2*i] = xr * rr - xi * ri;
               Aj[2*i+1] = xr * ri + xi * rr;
            }
         }
         else
            for (i = j + 1; i < M; i++) F_(cdiv)(pr, pi, &Aj[2*i], &Aj[2*i+1]);
      }
      else if (!info)
         info = j + 1;
      for (c = j + 1; c < N; c++)
      {
         T *ac = A + 2*(size_t)c*lda;
         const T ur = ac[2*j], ui = ac[2*j+1];
         for (i = j + 1; i < M; i++)
         {
            const T lr = Aj[2*i], li = Aj[2*i+1];
            ac[2*i] -= lr * ur - li * ui;
            ac[2*i+1] -= lr * ui + li * ur;
         }
      }
   }
   if (mingap) *mingap = gap;
   return info;
}

/* Complex Cholesky, src/lapack/ATL_potrfL.c / ATL_potrfU.c for TCPLX: A =
 * L L^H (Lower) or U^H U (Upper, handled as the lower factor of A^H, U =
 * L^H), the real part of the diagonal being the pivot.  Column j: d =
 * Re A(j,j) - sum_t |L(j,t)|^2; a d that is not positive (or NaN) returns
 * j + 1 with the columns before j factored; else L(j,j) = sqrt(d) (imaginary
 * part 0) and L(i,j) = (A(i,j) - sum_t L(i,t) conj(L(j,t))) / L(j,j).  The
 * value is restated (dot-product order, not ATLAS's recursion); the tests
 * pin it against numpy's Cholesky. */
int F_(potrf)(int uplo, int N, T *A, int lda)
{
   const int up = (uplo == ORC_UPPER);
   int i, j, t;
   /* element (i,j), i >= j, of the lower view */
#define LV(i, j) (up ? A + 2*((size_t)(i)*lda + (j)) : A + 2*((size_t)(j)*lda + (i)))
#define LVI(p) (up ? -(p)[1] : (p)[1])
   for (j = 0; j < N; j++)
   {
      T d = LV(j, j)[0];
      for (t = 0; t < j; t++)
      {
         const T *l = LV(j, t);
         d -= l[0] * l[0] + LVI(l) * LVI(l);
      }
      if (!(d > 0)) return j + 1;
      d = ORC_SQRT(d);
      LV(j, j)[0] = d;
      LV(j, j)[1] = 0;
      for (i = j + 1; i < N; i++)
      {
         T *a = LV(i, j), sr = a[0], si = LVI(a);
         for (t = 0; t < j; t++)
         {
            const T *x = LV(i, t), *y = LV(j, t);
            const T xr = x[0], xi = LVI(x), yr = y[0], yi = -LVI(y);   /* conj(L(j,t)) */
            sr -= xr * yr - xi * yi;
            si -= xr * yi + xi * yr;
         }
         a[0] = sr / d;
         a[1] = up ? -(si / d) : si / d;
      }
   }
#undef LV
#undef LVI
   return 0;
}

#undef F_

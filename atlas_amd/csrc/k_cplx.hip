/*
 * k_cplx.hip -- the complex (c/z) Level-3 pieces around the real MFMA GEMM.
 *
 * Column-major complex storage (real and imaginary parts interleaved) read
 * as a real matrix of twice the rows is exactly the interleaved real view
 * of the matrix, and a complex product C = A B is the real product
 *     C_re-view = Ahat * B_re-view,
 * where Ahat replaces every element a of A by the 2x2 block
 *     [ Re a  -Im a ]
 *     [ Im a   Re a ].
 * The real product does 2m x 2k x n multiply-adds = 8mnk flops, the complex
 * count, so a complex GEMM is one launch of the real MFMA kernel after an
 * O(mk) expansion of the left operand.  The kernels here build that
 * expansion (with op(A), conjugation, alpha and a triangle mask folded in),
 * plain complex gathers of op(B) and transposes, the complex scaling of C by
 * beta, and the inverses of the diagonal blocks of a complex triangle (for
 * TRSM, stored already expanded so the solve stays a chain of real GEMMs).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "kernels.h"

namespace {

constexpr int CT = 32;   /* tile edge of the gather, in complex elements */

/* ------------------------------------------------------------------------
 * out = alpha * op(A) for an m x n op(A), where op(A)(i,p) is A(i,p)
 * (trans = 0) or A(p,i) (trans = 1), conjugated when conj is set.
 *   expand = 1: out is real, 2m x 2n with leading dimension ldo (reals),
 *               element (i,p) written as its 2x2 block;
 *   expand = 0: out is complex, m x n with leading dimension ldo (complex).
 * tri = KX_LOWER / KX_UPPER keeps op(A)'s elements on and below / above the
 * diagonal and writes zeros elsewhere; unit writes 1 on the diagonal.
 * One 32x32 tile per 256-thread workgroup; the source tile is read along its
 * columns and, for trans, turned through LDS so the writes stay contiguous.
 * ---------------------------------------------------------------------- */
template <typename T>
__global__ __launch_bounds__(256) void cgather_kernel(int m, int n, const T *__restrict__ A,
                                                      long lda, int trans, int conj, T ar,
                                                      T ai, int expand, T *__restrict__ out,
                                                      long ldo, int tri, int unit) {
  __shared__ T sr[CT][CT + 1], si[CT][CT + 1];
  const int tx = threadIdx.x & (CT - 1), ty = threadIdx.x >> 5;
  const int i0 = blockIdx.x * CT, p0 = blockIdx.y * CT;
  /* source rows run along tx in both cases: op row i (trans = 0) or op
   * column p (trans = 1) */
  for (int q = ty; q < CT; q += 8) {
    int sr_, sc_;
    if (!trans) { sr_ = i0 + tx; sc_ = p0 + q; }
    else        { sr_ = p0 + tx; sc_ = i0 + q; }
    const bool in = trans ? (sr_ < n && sc_ < m) : (sr_ < m && sc_ < n);
    T vr = T(0), vi = T(0);
    if (in) {
      const T *a = A + 2 * ((size_t)sc_ * lda + sr_);
      vr = a[0]; vi = a[1];
    }
    if (!trans) { sr[tx][q] = vr; si[tx][q] = vi; }
    else        { sr[q][tx] = vr; si[q][tx] = vi; }
  }
  __syncthreads();
  for (int q = ty; q < CT; q += 8) {
    const int i = i0 + tx, p = p0 + q;
    if (i >= m || p >= n) continue;
    T xr = sr[tx][q], xi = si[tx][q];
    if (conj) xi = -xi;
    T vr = ar * xr - ai * xi, vi = ar * xi + ai * xr;
    if ((tri == KX_LOWER && i < p) || (tri == KX_UPPER && i > p)) { vr = T(0); vi = T(0); }
    if (unit && i == p) { vr = T(1); vi = T(0); }
    if (expand) {
      T *o0 = out + (size_t)(2 * p) * ldo + 2 * i;
      T *o1 = o0 + ldo;
      o0[0] = vr; o0[1] = vi;
      o1[0] = -vi; o1[1] = vr;
    } else {
      T *o = out + 2 * ((size_t)p * ldo + i);
      o[0] = vr; o[1] = vi;
    }
  }
}

/* C = beta * C for an m x n complex C; beta = 0 writes zeros without reading
 * C (the Mzgescal of include/atlas_refmisc.h:341-361). */
template <typename T>
__global__ __launch_bounds__(256) void cscal_kernel(int m, int n, T br, T bi, T *C, long ldc) {
  const long total = (long)m * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long j = e / m, i = e - j * m;
    T *c = C + 2 * (j * ldc + i);
    if (br == T(0) && bi == T(0)) { c[0] = T(0); c[1] = T(0); continue; }
    const T cr = c[0], ci = c[1];
    c[0] = br * cr - bi * ci;
    c[1] = br * ci + bi * cr;
  }
}

/* 1/d by Smith's scaling (R. L. Smith, CACM 5(8), 1962): divide by the
 * larger of |Re d|, |Im d| first so neither square overflows. */
template <typename T>
__device__ inline void crecip(T dr, T di, T &rr, T &ri) {
  if (fabs(di) <= fabs(dr)) {
    const T t = di / dr, den = dr + di * t;
    rr = T(1) / den; ri = -t / den;
  } else {
    const T t = dr / di, den = di + dr * t;
    rr = t / den; ri = T(-1) / den;
  }
}

/* ------------------------------------------------------------------------
 * Inverses of the 32x32 complex diagonal blocks of a triangular matrix
 * held expanded in H (2n x 2n reals, leading dimension ldh, the complex
 * element (i,j) at H[2j*ldh + 2i], H[2j*ldh + 2i + 1]), written expanded:
 * block b as a 64x64 real block at inv + b*64*64 (leading dimension 64).
 * Past the matrix's edge the block is the identity.  One workgroup of 64
 * threads per block; thread c < 32 builds column c of the inverse by
 * substitution, with the block and the columns in LDS.
 * ---------------------------------------------------------------------- */
template <typename T>
__global__ __launch_bounds__(64) void ctri_inv_kernel(int lower, int n, const T *__restrict__ H,
                                                      long ldh, T *__restrict__ inv) {
  __shared__ T Lr[CT][CT + 1], Li[CT][CT + 1];
  __shared__ T Xr[CT][CT + 1], Xi[CT][CT + 1];
  __shared__ T Dr[CT], Di[CT];
  const int tid = threadIdx.x, b0 = blockIdx.x * CT, nb = min(CT, n - b0);
  for (int e = tid; e < CT * CT; e += 64) {
    const int r = e & (CT - 1), c = e >> 5;
    T vr, vi;
    if (r < nb && c < nb) {
      const T *h = H + (size_t)(2 * (b0 + c)) * ldh + 2 * (b0 + r);
      vr = h[0]; vi = h[1];
    } else {
      vr = (r == c) ? T(1) : T(0); vi = T(0);
    }
    Lr[r][c] = vr; Li[r][c] = vi;
  }
  __syncthreads();
  if (tid < CT) crecip(Lr[tid][tid], Li[tid][tid], Dr[tid], Di[tid]);
  __syncthreads();
  if (tid < CT) {
    const int c = tid;
    for (int r = 0; r < CT; ++r) { Xr[r][c] = T(0); Xi[r][c] = T(0); }
    Xr[c][c] = Dr[c]; Xi[c][c] = Di[c];
    if (lower) {
      for (int r = c + 1; r < CT; ++r) {
        T sr = T(0), si = T(0);
        for (int t = c; t < r; ++t) {
          sr += Lr[r][t] * Xr[t][c] - Li[r][t] * Xi[t][c];
          si += Lr[r][t] * Xi[t][c] + Li[r][t] * Xr[t][c];
        }
        Xr[r][c] = -(sr * Dr[r] - si * Di[r]);
        Xi[r][c] = -(sr * Di[r] + si * Dr[r]);
      }
    } else {
      for (int r = c - 1; r >= 0; --r) {
        T sr = T(0), si = T(0);
        for (int t = r + 1; t <= c; ++t) {
          sr += Lr[r][t] * Xr[t][c] - Li[r][t] * Xi[t][c];
          si += Lr[r][t] * Xi[t][c] + Li[r][t] * Xr[t][c];
        }
        Xr[r][c] = -(sr * Dr[r] - si * Di[r]);
        Xi[r][c] = -(sr * Di[r] + si * Dr[r]);
      }
    }
  }
  __syncthreads();
  T *o = inv + (size_t)blockIdx.x * 64 * 64;
  for (int e = tid; e < CT * CT; e += 64) {
    const int r = e & (CT - 1), c = e >> 5;
    const T vr = Xr[r][c], vi = Xi[r][c];
    T *o0 = o + (size_t)(2 * c) * 64 + 2 * r, *o1 = o0 + 64;
    o0[0] = vr; o0[1] = vi;
    o1[0] = -vi; o1[1] = vr;
  }
}

/* x / d by Smith's algorithm (the same scaling as crecip) */
template <typename T>
__device__ inline void cdivide(T dr, T di, T &xr, T &xi) {
  T cr, ci;
  if (fabs(di) < fabs(dr)) {
    const T t = di / dr, den = dr + di * t;
    cr = (xr + xi * t) / den; ci = (xi - xr * t) / den;
  } else {
    const T t = dr / di, den = di + dr * t;
    cr = (xi + xr * t) / den; ci = (-xr + xi * t) / den;
  }
  xr = cr; xi = ci;
}

template <typename T> struct SafMin;
template <> struct SafMin<double> { static constexpr double v = 2.2250738585072014e-308; };
template <> struct SafMin<float> { static constexpr float v = 1.17549435e-38f; };

/* ------------------------------------------------------------------------
 * Unblocked complex LU of an m x n leaf (n <= CLEAF_MAXW columns) whose top
 * row is global row row0, in place, one 1024-thread workgroup: the column
 * loop of src/lapack/ATL_getf2.c:60-95 with LU1 (:10-58) for complex data.
 * Per column j: the pivot is the first row of largest |Re| + |Im| at or
 * below j (cblas_izamax, ATL_izrefamax.c: NaN never wins, all-zero gives
 * row j); ipiv[j] = row0 + pivot row.  A nonzero pivot swaps its row with
 * row j across the leaf and the rows below are multiplied by its
 * reciprocal, or divided by it when |Re| + |Im| < SAFMIN; a zero pivot sets
 * *info = row0 + j + 1 when *info is 0.  Then the rank-1 update of the
 * trailing columns, which runs in both cases.
 * Thread t owns rows j + t, j + t + 1024, ...; pivot row j of the trailing
 * columns is staged in LDS for the update.
 * ---------------------------------------------------------------------- */
constexpr int CLEAF_MAXW = 16;
template <typename T>
__global__ __launch_bounds__(1024) void cgetf2_kernel(int m, int n, T *__restrict__ A, long lda,
                                                      int row0, int *__restrict__ ipiv,
                                                      int *__restrict__ ipiv2,
                                                      int *__restrict__ info) {
  __shared__ T s_v[16];
  __shared__ int s_i[16];
  __shared__ int s_p;
  __shared__ T s_pr, s_pi;
  __shared__ T s_rr[CLEAF_MAXW], s_ri[CLEAF_MAXW];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int mn = min(m, n);
  for (int j = 0; j < mn; ++j) {
    T *Aj = A + 2 * (size_t)j * lda;
    T bv = T(0);
    int bi = j;
    for (int i = j + tid; i < m; i += 1024) {
      const T v = fabs(Aj[2 * i]) + fabs(Aj[2 * i + 1]);
      if (v > bv) { bv = v; bi = i; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const T ov = __shfl_xor(bv, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { s_v[wid] = bv; s_i[wid] = bi; }
    __syncthreads();
    if (tid == 0) {
      T v = s_v[0];
      int p = s_i[0];
      for (int w = 1; w < 16; ++w)
        if (s_v[w] > v || (s_v[w] == v && s_i[w] < p)) { v = s_v[w]; p = s_i[w]; }
      s_p = p;
      s_pr = Aj[2 * p];
      s_pi = Aj[2 * p + 1];
      ipiv[j] = row0 + p;
      if (ipiv2) { ipiv2[2 * j] = 2 * (row0 + p); ipiv2[2 * j + 1] = 2 * (row0 + p) + 1; }
    }
    __syncthreads();
    const int p = s_p;
    const T pr = s_pr, pi = s_pi;
    const T pv = fabs(pr) + fabs(pi);   /* the same in every thread */
    if (pv != T(0)) {
      if (p != j && tid < n) {
        T *a = A + 2 * ((size_t)tid * lda + j), *b = A + 2 * ((size_t)tid * lda + p);
        const T ar = a[0], ai = a[1];
        a[0] = b[0]; a[1] = b[1];
        b[0] = ar; b[1] = ai;
      }
      __syncthreads();
      if (pv >= SafMin<T>::v) {
        T rr, ri;
        crecip(pr, pi, rr, ri);
        for (int i = j + 1 + tid; i < m; i += 1024) {
          const T xr = Aj[2 * i], xi = Aj[2 * i + 1];
          Aj[2 * i] = xr * rr - xi * ri;
          Aj[2 * i + 1] = xr * ri + xi * rr;
        }
      } else {
        for (int i = j + 1 + tid; i < m; i += 1024) {
          T xr = Aj[2 * i], xi = Aj[2 * i + 1];
          cdivide(pr, pi, xr, xi);
          Aj[2 * i] = xr; Aj[2 * i + 1] = xi;
        }
      }
    } else if (tid == 0 && *info == 0) {
      *info = row0 + j + 1;
    }
    /* row j of the trailing columns (after the swap) to LDS */
    __syncthreads();
    if (tid > j && tid < n) {
      s_rr[tid] = A[2 * ((size_t)tid * lda + j)];
      s_ri[tid] = A[2 * ((size_t)tid * lda + j) + 1];
    }
    __syncthreads();
    for (int i = j + 1 + tid; i < m; i += 1024) {
      const T lr = Aj[2 * i], li = Aj[2 * i + 1];
      for (int c = j + 1; c < n; ++c) {
        T *a = A + 2 * ((size_t)c * lda + i);
        a[0] -= lr * s_rr[c] - li * s_ri[c];
        a[1] -= lr * s_ri[c] + li * s_rr[c];
      }
    }
    __syncthreads();
  }
}

/* ------------------------------------------------------------------------
 * The same leaf with its rows in registers: thread t holds rows t,
 * t + 1024, ... (R of them) of the W columns, so a column step costs two
 * LDS exchanges and three barriers instead of dependent global round trips.
 * Rows j and p are exchanged through LDS (the owners of rows j and p write
 * them, then take each other's); the update reads the pivot row from LDS.
 * m <= 1024 R, n <= W.  Also writes ipiv2 (the pivots of the interleaved
 * real view, rows 2p and 2p + 1) when given.
 * ---------------------------------------------------------------------- */
template <typename T, int R, int W>
__global__ __launch_bounds__(1024) void cgetf2_reg_kernel(int m, int n, T *__restrict__ A,
                                                          long lda, int row0,
                                                          int *__restrict__ ipiv,
                                                          int *__restrict__ ipiv2,
                                                          int *__restrict__ info) {
  __shared__ T s_v[16];
  __shared__ int s_i[16];
  __shared__ int s_p;
  __shared__ T s_pr[W], s_pi[W], s_jr[W], s_ji[W];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  T ar[R][W], ai[R][W];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = tid + 1024 * q;
#pragma unroll
    for (int c = 0; c < W; ++c) {
      const bool in = (i < m && c < n);
      ar[q][c] = in ? A[2 * ((size_t)c * lda + i)] : T(0);
      ai[q][c] = in ? A[2 * ((size_t)c * lda + i) + 1] : T(0);
    }
  }
  const int mn = min(m, n);
#pragma unroll
  for (int j = 0; j < W; ++j) {
    if (j >= mn) break;
    T bv = T(0);
    int bi = j;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = tid + 1024 * q;
      if (i >= j && i < m) {
        const T v = fabs(ar[q][j]) + fabs(ai[q][j]);
        if (v > bv) { bv = v; bi = i; }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const T ov = __shfl_xor(bv, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { s_v[wid] = bv; s_i[wid] = bi; }
    __syncthreads();
    if (tid == 0) {
      T v = s_v[0];
      int p = s_i[0];
      for (int w = 1; w < 16; ++w)
        if (s_v[w] > v || (s_v[w] == v && s_i[w] < p)) { v = s_v[w]; p = s_i[w]; }
      s_p = p;
      ipiv[j] = row0 + p;
      if (ipiv2) { ipiv2[2 * j] = 2 * (row0 + p); ipiv2[2 * j + 1] = 2 * (row0 + p) + 1; }
    }
    __syncthreads();
    const int p = s_p;
    /* the owners of rows p and j publish them */
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = tid + 1024 * q;
      if (i == p) {
#pragma unroll
        for (int c = 0; c < W; ++c) { s_pr[c] = ar[q][c]; s_pi[c] = ai[q][c]; }
      }
      if (i == j) {
#pragma unroll
        for (int c = 0; c < W; ++c) { s_jr[c] = ar[q][c]; s_ji[c] = ai[q][c]; }
      }
    }
    __syncthreads();
    const T pr = s_pr[j], pi = s_pi[j];
    const T pv = fabs(pr) + fabs(pi);   /* the same in every thread */
    /* the row the update uses: the pivot row after a swap, else row j */
    const T *ur = (pv != T(0)) ? s_pr : s_jr;
    const T *ui = (pv != T(0)) ? s_pi : s_ji;
    if (pv != T(0)) {
      T rr = T(0), ri = T(0);
      const bool inv = (pv >= SafMin<T>::v);
      if (inv) crecip(pr, pi, rr, ri);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = tid + 1024 * q;
        if (p != j && i == j) {
#pragma unroll
          for (int c = 0; c < W; ++c) { ar[q][c] = s_pr[c]; ai[q][c] = s_pi[c]; }
        } else if (p != j && i == p) {
#pragma unroll
          for (int c = 0; c < W; ++c) { ar[q][c] = s_jr[c]; ai[q][c] = s_ji[c]; }
        }
        if (i > j && i < m) {
          T xr = ar[q][j], xi = ai[q][j];
          if (inv) {
            const T yr = xr * rr - xi * ri;
            xi = xr * ri + xi * rr;
            xr = yr;
          } else {
            cdivide(pr, pi, xr, xi);
          }
          ar[q][j] = xr; ai[q][j] = xi;
        }
      }
    } else if (tid == 0 && *info == 0) {
      *info = row0 + j + 1;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = tid + 1024 * q;
      if (i > j && i < m) {
        const T lr = ar[q][j], li = ai[q][j];
#pragma unroll
        for (int c = j + 1; c < W; ++c) {
          ar[q][c] -= lr * ur[c] - li * ui[c];
          ai[q][c] -= lr * ui[c] + li * ur[c];
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = tid + 1024 * q;
#pragma unroll
    for (int c = 0; c < W; ++c)
      if (i < m && c < n) {
        A[2 * ((size_t)c * lda + i)] = ar[q][c];
        A[2 * ((size_t)c * lda + i) + 1] = ai[q][c];
      }
  }
}

/* ------------------------------------------------------------------------
 * B := inv(L) B for a unit lower triangular L of order n1 <= 64 and B of
 * n1 rows, in place; one workgroup per 64 columns of B, with L and its
 * column block in LDS.  Right-looking: after row t of the solution is final,
 * every row below takes its L(r,t) multiple, one barrier per t.  This is
 * the TRSM(Left, Lower, NoTrans, Unit) of the LU recursion and of the
 * panel's U12, in one launch.
 * ---------------------------------------------------------------------- */
constexpr int STRSM_N = 64;
template <typename T>
__global__ __launch_bounds__(1024) void ctrsm_llnu_kernel(int n1, int nr, const T *__restrict__ L,
                                                          long lda, T *__restrict__ B, long ldb) {
  __shared__ T Lr[STRSM_N][STRSM_N + 1], Li[STRSM_N][STRSM_N + 1];
  __shared__ T Br[STRSM_N][STRSM_N + 1], Bi[STRSM_N][STRSM_N + 1];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * STRSM_N, nc = min(STRSM_N, nr - c0);
  for (int e = tid; e < n1 * n1; e += 1024) {
    const int r = e % n1, c = e / n1;
    if (r > c) {
      const T *l = L + 2 * ((size_t)c * lda + r);
      Lr[r][c] = l[0]; Li[r][c] = l[1];
    }
  }
  for (int e = tid; e < n1 * nc; e += 1024) {
    const int r = e % n1, c = e / n1;
    const T *b = B + 2 * ((size_t)(c0 + c) * ldb + r);
    Br[r][c] = b[0]; Bi[r][c] = b[1];
  }
  __syncthreads();
  for (int t = 0; t + 1 < n1; ++t) {
    const int rows = n1 - 1 - t;
    for (int e = tid; e < rows * nc; e += 1024) {
      const int r = t + 1 + e % rows, c = e / rows;
      const T lr = Lr[r][t], li = Li[r][t], xr = Br[t][c], xi = Bi[t][c];
      Br[r][c] -= lr * xr - li * xi;
      Bi[r][c] -= lr * xi + li * xr;
    }
    __syncthreads();
  }
  for (int e = tid; e < n1 * nc; e += 1024) {
    const int r = e % n1, c = e / n1;
    T *b = B + 2 * ((size_t)(c0 + c) * ldb + r);
    b[0] = Br[r][c]; b[1] = Bi[r][c];
  }
}

/* ------------------------------------------------------------------------
 * Cholesky of one n x n (n <= 64) Hermitian diagonal block in place: A = L
 * L^H from the lower triangle, or A = U^H U from the upper one, which is
 * read into LDS as its conjugate transpose (the lower triangle of A^H) and
 * written back the same way.  Right-looking, one barrier-separated step per
 * column: the pivot is the real part of the diagonal (the imaginary part is
 * not referenced; the factor's diagonal is real), the column below it is
 * divided by the pivot's square root, and the lower triangle of the trailing
 * block takes L(i,j) conj(L(c,j)).  A pivot that is not positive (or NaN)
 * stops the factorization and sets *info = base + j + 1 (when *info is 0);
 * the block is written back as far as it got (src/lapack/ATL_potrfL.c: the
 * info of the first failing column).  One 256-thread workgroup.
 * ---------------------------------------------------------------------- */
constexpr int CPOTF2_N = 64;
template <typename T>
__global__ __launch_bounds__(256) void cpotf2_kernel(int upper, int n, T *__restrict__ A,
                                                     long lda, int base, int *__restrict__ info) {
  __shared__ T Lr[CPOTF2_N][CPOTF2_N + 1], Li[CPOTF2_N][CPOTF2_N + 1];
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += 256) {
    const int r = e % n, c = e / n;
    if (r < c) continue;
    const T *a = upper ? A + 2 * ((size_t)r * lda + c) : A + 2 * ((size_t)c * lda + r);
    Lr[r][c] = a[0];
    Li[r][c] = upper ? -a[1] : a[1];
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const T d = Lr[j][j];
    if (!(d > T(0))) {                          /* uniform: every thread reads d */
      if (tid == 0 && *info == 0) *info = base + j + 1;
      break;
    }
    const T s = sqrt(d), rs = T(1) / s;
    for (int i = j + 1 + tid; i < n; i += 256) { Lr[i][j] *= rs; Li[i][j] *= rs; }
    __syncthreads();
    if (tid == 0) { Lr[j][j] = s; Li[j][j] = T(0); }
    const int rows = n - 1 - j;
    for (int e = tid; e < rows * rows; e += 256) {
      const int i = j + 1 + e % rows, c = j + 1 + e / rows;
      if (i < c) continue;
      /* L(i,c) -= L(i,j) * conj(L(c,j)) */
      const T ar = Lr[i][j], ai = Li[i][j], br = Lr[c][j], bi = -Li[c][j];
      Lr[i][c] -= ar * br - ai * bi;
      Li[i][c] -= ar * bi + ai * br;
    }
    __syncthreads();
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += 256) {
    const int r = e % n, c = e / n;
    if (r < c) continue;
    T *a = upper ? A + 2 * ((size_t)r * lda + c) : A + 2 * ((size_t)c * lda + r);
    a[0] = Lr[r][c];
    a[1] = upper ? -Li[r][c] : Li[r][c];
  }
}

/* C = beta*C on one triangle of an n x n complex C (lower: i >= j), beta = 0
 * storing zeros; herm also stores 0 in the diagonal's imaginary parts (a
 * Hermitian C's diagonal is real: src/blas/reference/level3/ATL_zrefherk*.c
 * set it so at every scaling and update).  With beta = 1 and herm only the
 * diagonal is touched. */
template <typename T>
__global__ __launch_bounds__(256) void ctri_scale_kernel(int n, int lower, T br, T bi, int herm,
                                                         T *C, long ldc) {
  const bool diag_only = (br == T(1) && bi == T(0));
  const long total = diag_only ? (long)n : (long)n * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long j = diag_only ? e : e / n, i = diag_only ? e : e - j * n;
    if (lower ? (i < j) : (i > j)) continue;
    T *c = C + 2 * (j * ldc + i);
    if (!diag_only) {
      if (br == T(0) && bi == T(0)) {
        c[0] = T(0); c[1] = T(0);
      } else if (herm) {
        /* a real beta scales each part; the diagonal's imaginary part is
         * not referenced */
        c[0] = br * c[0];
        c[1] = (i == j) ? T(0) : br * c[1];
      } else {
        const T cr = c[0], ci = c[1];
        c[0] = br * cr - bi * ci;
        c[1] = br * ci + bi * cr;
      }
    }
    if (herm && i == j) c[1] = T(0);
  }
}

/* The full n x n matrix of a Hermitian (herm) or symmetric A held in one
 * triangle: out(i,j) = A(i,j) in the triangle, else A(j,i), conjugated when
 * herm; a Hermitian diagonal's imaginary part is taken as 0 (it is not
 * referenced: ATL_zrefhemm*.c). */
template <typename T>
__global__ __launch_bounds__(256) void csym_expand_kernel(int n, int lower, int herm,
                                                          const T *__restrict__ A, long lda,
                                                          T *__restrict__ out, long ldo) {
  const long total = (long)n * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long j = e / n, i = e - j * n;
    const bool in = lower ? (i >= j) : (i <= j);
    const T *a = in ? A + 2 * (j * lda + i) : A + 2 * (i * lda + j);
    T vr = a[0], vi = a[1];
    if (herm && !in) vi = -vi;
    if (herm && i == j) vi = T(0);
    T *o = out + 2 * (j * ldo + i);
    o[0] = vr; o[1] = vi;
  }
}

/* C = conj(C), m x n complex, in place */
template <typename T>
__global__ __launch_bounds__(256) void cconj_kernel(int m, int n, T *C, long ldc) {
  const long total = (long)m * n;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long j = e / m, i = e - j * m;
    T *c = C + 2 * (j * ldc + i) + 1;
    *c = -*c;
  }
}

/* ipiv2[2k + s] = 2 ipiv[k] + s for k in [k1, k2): the row interchanges of
 * a complex matrix as those of its interleaved real view. */
__global__ __launch_bounds__(256) void cpiv2_kernel(int k1, int k2, const int *__restrict__ ipiv,
                                                    int *__restrict__ ipiv2) {
  const int k = k1 + blockIdx.x * 256 + threadIdx.x;
  if (k >= k2) return;
  const int p = ipiv[k];
  ipiv2[2 * k] = 2 * p;
  ipiv2[2 * k + 1] = 2 * p + 1;
}

template <typename T>
hipError_t launch_gather(int m, int n, const T *A, long lda, int trans, int conj, T ar, T ai,
                         int expand, T *out, long ldo, int tri, int unit, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (!A || !out) return hipErrorInvalidValue;
  if (lda < (trans ? n : m)) return hipErrorInvalidValue;
  if (ldo < (expand ? 2L * m : (long)m)) return hipErrorInvalidValue;
  const unsigned gx = (unsigned)((m + CT - 1) / CT), gy = (unsigned)((n + CT - 1) / CT);
  if (gy > 65535u) return hipErrorInvalidValue;
  hipLaunchKernelGGL((cgather_kernel<T>), dim3(gx, gy), dim3(256), 0, s, m, n, A, lda, trans,
                     conj, ar, ai, expand, out, ldo, tri, unit);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cscal(int m, int n, T br, T bi, T *C, long ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (!C || ldc < m) return hipErrorInvalidValue;
  const long total = (long)m * n;
  const long want = (total + 255) / 256;
  const unsigned g = (unsigned)(want < 8192 ? want : 8192);
  hipLaunchKernelGGL((cscal_kernel<T>), dim3(g), dim3(256), 0, s, m, n, br, bi, C, ldc);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ctri_inv(int lower, int n, const T *H, long ldh, T *inv, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (!H || !inv || ldh < 2L * n) return hipErrorInvalidValue;
  hipLaunchKernelGGL((ctri_inv_kernel<T>), dim3((n + CT - 1) / CT), dim3(64), 0, s, lower, n,
                     H, ldh, inv);
  return hipGetLastError();
}

}  // namespace

extern "C" hipError_t atlmx_k_zgather(int m, int n, const double *A, long lda, int trans,
                                      int conj, double ar, double ai, int expand, double *out,
                                      long ldo, int tri, int unit, hipStream_t s) {
  return launch_gather<double>(m, n, A, lda, trans, conj, ar, ai, expand, out, ldo, tri, unit,
                               s);
}
extern "C" hipError_t atlmx_k_cgather(int m, int n, const float *A, long lda, int trans,
                                      int conj, float ar, float ai, int expand, float *out,
                                      long ldo, int tri, int unit, hipStream_t s) {
  return launch_gather<float>(m, n, A, lda, trans, conj, ar, ai, expand, out, ldo, tri, unit,
                              s);
}
extern "C" hipError_t atlmx_k_zscal2d(int m, int n, double br, double bi, double *C, long ldc,
                                      hipStream_t s) {
  return launch_cscal<double>(m, n, br, bi, C, ldc, s);
}
extern "C" hipError_t atlmx_k_cscal2d(int m, int n, float br, float bi, float *C, long ldc,
                                      hipStream_t s) {
  return launch_cscal<float>(m, n, br, bi, C, ldc, s);
}
extern "C" hipError_t atlmx_k_ztri_inv(int lower, int n, const double *H, long ldh, double *inv,
                                       hipStream_t s) {
  return launch_ctri_inv<double>(lower, n, H, ldh, inv, s);
}
extern "C" hipError_t atlmx_k_ctri_inv(int lower, int n, const float *H, long ldh, float *inv,
                                       hipStream_t s) {
  return launch_ctri_inv<float>(lower, n, H, ldh, inv, s);
}

namespace {
/* the register leaf where its rows fit (8 columns up to 2048 rows, 4 up to
 * 4096), else the global-memory one */
template <typename T>
hipError_t launch_cgetf2(int m, int n, T *A, long lda, int row0, int *ipiv, int *ipiv2,
                         int *info, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (n > CLEAF_MAXW || !A || !ipiv || !info || lda < m) return hipErrorInvalidValue;
  const dim3 g(1), b(1024);
  if (n <= 8 && m <= 1024)
    hipLaunchKernelGGL((cgetf2_reg_kernel<T, 1, 8>), g, b, 0, s, m, n, A, lda, row0, ipiv,
                       ipiv2, info);
  else if (n <= 8 && m <= 2048)
    hipLaunchKernelGGL((cgetf2_reg_kernel<T, 2, 8>), g, b, 0, s, m, n, A, lda, row0, ipiv,
                       ipiv2, info);
  else if (n <= 4 && m <= 4096)
    hipLaunchKernelGGL((cgetf2_reg_kernel<T, 4, 4>), g, b, 0, s, m, n, A, lda, row0, ipiv,
                       ipiv2, info);
  else
    hipLaunchKernelGGL((cgetf2_kernel<T>), g, b, 0, s, m, n, A, lda, row0, ipiv, ipiv2, info);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cpotf2(int upper, int n, T *A, long lda, int base, int *info, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n > CPOTF2_N || !A || !info || lda < n) return hipErrorInvalidValue;
  hipLaunchKernelGGL((cpotf2_kernel<T>), dim3(1), dim3(256), 0, s, upper, n, A, lda, base, info);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ctri_scale(int n, int lower, T br, T bi, int herm, T *C, long ldc,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (!C || ldc < n) return hipErrorInvalidValue;
  const long total = (br == T(1) && bi == T(0)) ? (long)n : (long)n * n;
  const long want = (total + 255) / 256;
  hipLaunchKernelGGL((ctri_scale_kernel<T>), dim3((unsigned)(want < 8192 ? want : 8192)),
                     dim3(256), 0, s, n, lower, br, bi, herm, C, ldc);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_csym_expand(int n, int lower, int herm, const T *A, long lda, T *out, long ldo,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (!A || !out || lda < n || ldo < n) return hipErrorInvalidValue;
  const long want = ((long)n * n + 255) / 256;
  hipLaunchKernelGGL((csym_expand_kernel<T>), dim3((unsigned)(want < 8192 ? want : 8192)),
                     dim3(256), 0, s, n, lower, herm, A, lda, out, ldo);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_cconj(int m, int n, T *C, long ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  if (!C || ldc < m) return hipErrorInvalidValue;
  const long want = ((long)m * n + 255) / 256;
  hipLaunchKernelGGL((cconj_kernel<T>), dim3((unsigned)(want < 8192 ? want : 8192)), dim3(256),
                     0, s, m, n, C, ldc);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ctrsm_llnu(int n1, int nr, const T *L, long lda, T *B, long ldb,
                             hipStream_t s) {
  if (n1 <= 0 || nr <= 0) return hipSuccess;
  if (n1 > STRSM_N || !L || !B || lda < n1 || ldb < n1) return hipErrorInvalidValue;
  const long nb = (nr + STRSM_N - 1) / STRSM_N;
  if (nb > (1L << 30)) return hipErrorInvalidValue;
  hipLaunchKernelGGL((ctrsm_llnu_kernel<T>), dim3((unsigned)nb), dim3(1024), 0, s, n1, nr, L,
                     lda, B, ldb);
  return hipGetLastError();
}
}  // namespace

extern "C" hipError_t atlmx_k_zgetf2(int m, int n, double *A, long lda, int row0, int *ipiv,
                                     int *ipiv2, int *info, hipStream_t s) {
  return launch_cgetf2<double>(m, n, A, lda, row0, ipiv, ipiv2, info, s);
}
extern "C" hipError_t atlmx_k_cgetf2(int m, int n, float *A, long lda, int row0, int *ipiv,
                                     int *ipiv2, int *info, hipStream_t s) {
  return launch_cgetf2<float>(m, n, A, lda, row0, ipiv, ipiv2, info, s);
}
extern "C" hipError_t atlmx_k_ztrsm_llnu(int n1, int nr, const double *L, long lda, double *B,
                                         long ldb, hipStream_t s) {
  return launch_ctrsm_llnu<double>(n1, nr, L, lda, B, ldb, s);
}
extern "C" hipError_t atlmx_k_ctrsm_llnu(int n1, int nr, const float *L, long lda, float *B,
                                         long ldb, hipStream_t s) {
  return launch_ctrsm_llnu<float>(n1, nr, L, lda, B, ldb, s);
}
extern "C" hipError_t atlmx_k_cpiv2(int k1, int k2, const int *ipiv, int *ipiv2,
                                    hipStream_t s) {
  if (k2 <= k1) return hipSuccess;
  if (!ipiv || !ipiv2) return hipErrorInvalidValue;
  hipLaunchKernelGGL(cpiv2_kernel, dim3((k2 - k1 + 255) / 256), dim3(256), 0, s, k1, k2, ipiv,
                     ipiv2);
  return hipGetLastError();
}
extern "C" hipError_t atlmx_k_zpotf2(int upper, int n, double *A, long lda, int base, int *info,
                                     hipStream_t s) {
  return launch_cpotf2<double>(upper, n, A, lda, base, info, s);
}
extern "C" hipError_t atlmx_k_cpotf2(int upper, int n, float *A, long lda, int base, int *info,
                                     hipStream_t s) {
  return launch_cpotf2<float>(upper, n, A, lda, base, info, s);
}
extern "C" hipError_t atlmx_k_zconj(int m, int n, double *C, long ldc, hipStream_t s) {
  return launch_cconj<double>(m, n, C, ldc, s);
}
extern "C" hipError_t atlmx_k_cconj(int m, int n, float *C, long ldc, hipStream_t s) {
  return launch_cconj<float>(m, n, C, ldc, s);
}
extern "C" hipError_t atlmx_k_ztri_scale(int n, int lower, double br, double bi, int herm,
                                         double *C, long ldc, hipStream_t s) {
  return launch_ctri_scale<double>(n, lower, br, bi, herm, C, ldc, s);
}
extern "C" hipError_t atlmx_k_ctri_scale(int n, int lower, float br, float bi, int herm, float *C,
                                         long ldc, hipStream_t s) {
  return launch_ctri_scale<float>(n, lower, br, bi, herm, C, ldc, s);
}
extern "C" hipError_t atlmx_k_zsym_expand(int n, int lower, int herm, const double *A, long lda,
                                          double *out, long ldo, hipStream_t s) {
  return launch_csym_expand<double>(n, lower, herm, A, lda, out, ldo, s);
}
extern "C" hipError_t atlmx_k_csym_expand(int n, int lower, int herm, const float *A, long lda,
                                          float *out, long ldo, hipStream_t s) {
  return launch_csym_expand<float>(n, lower, herm, A, lda, out, ldo, s);
}

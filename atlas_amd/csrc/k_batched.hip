/*
 * k_batched.hip -- thousands of small, uneven-size problems in ONE
 * persistent launch (no ATLAS counterpart: ATLAS has no batched interface;
 * each instance here follows the same algorithms as its single-problem
 * routines, src/lapack/ATL_potrfL.c:104-262, ATL_getrfC.c:46-361,
 * ATL_potrs.c:13-86, ATL_getrs.c:8-89, and GEMM's C = alpha op(A) op(B) +
 * beta C of interfaces/blas/C/src/cblas_dgemm.c:16-134).
 *
 * A persistent grid of workgroups takes work items from one atomic counter
 * in the order the host planned (largest first), so uneven sizes balance
 * themselves.  A factor+solve instance is one workgroup's work:
 *
 *   posv (lower Cholesky): right-looking, NB-column panels: the diagonal
 *     block is factored in LDS, the panel below is solved row by row
 *     against it, and the lower triangle of the trailing matrix is updated
 *     in 64 x 64 tiles (each thread a 4 x 4 piece, the panel's rows staged
 *     in LDS).  The right-hand sides are solved in LDS.
 *   gesv (LU, partial pivoting): the (n-k) x NB panel lives in LDS while it
 *     is factored with partial pivoting (workgroup arg-max per column, ties
 *     to the lower row, NaN never wins, like cblas_idamax), its interchanges
 *     are applied to the other columns, U12 is solved column by column and
 *     the trailing matrix is updated in 64 x 64 tiles.
 *
 * The GEMM work item is one 64 x 64 tile of one instance.
 */
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>
#include "kernels.h"

namespace {

constexpr int NT = 256;
constexpr int TS = 64;                 /* trailing-update / GEMM tile */
constexpr int MAXR = ATLMX_BATCH_MAXRHS;

/* v from lane `src` (a constant) to every lane of the wavefront */
__device__ inline double bcast(double v, int src) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline float bcast(float v, int src) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
template <typename T> __device__ inline T dsqrt(T x);
template <> __device__ inline double dsqrt<double>(double x) { return sqrt(x); }
template <> __device__ inline float dsqrt<float>(float x) { return sqrtf(x); }

/* C(r0.., c0..) -= P_r(rows r0.., K) * P_c(rows c0.., K)^T for one 64 x 64 tile,
 * restricted to r >= c when lower_only.  P_r, P_c: column-major with ld.
 * mr, mc: rows/cols of the tile inside the matrix. */
template <typename T>
__device__ void tile_update_nt(T *C, long ldc, const T *Pr, const T *Pc, long ldp, int K, int mr,
                               int mc, bool lower_only, int r0, int c0, T *s_a, T *s_b) {
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = T(0);
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int kc = min(16, K - k0);
    __syncthreads();
    /* s_a[kk][r] = Pr[r + (k0+kk)*ldp]; s_b[kk][c] likewise */
    for (int e = tid; e < 16 * TS; e += NT) {
      const int kk = e / TS, rr = e % TS;
      s_a[kk * TS + rr] = (kk < kc && rr < mr) ? Pr[(size_t)(k0 + kk) * ldp + rr] : T(0);
      s_b[kk * TS + rr] = (kk < kc && rr < mc) ? Pc[(size_t)(k0 + kk) * ldp + rr] : T(0);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = s_a[kk * TS + ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = s_b[kk * TS + tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = ty * 4 + i, c = tx * 4 + j;
      if (r < mr && c < mc && (!lower_only || r0 + r >= c0 + c))
        C[(size_t)c * ldc + r] -= acc[i][j];
    }
}

/* C(tile) -= L(rows, K) * U(K, cols): L column-major ld ldl, U column-major ld ldu */
template <typename T>
__device__ void tile_update_nn(T *C, long ldc, const T *L, long ldl, const T *U, long ldu, int K,
                               int mr, int mc, T *s_a, T *s_b) {
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = T(0);
  for (int k0 = 0; k0 < K; k0 += 16) {
    const int kc = min(16, K - k0);
    __syncthreads();
    for (int e = tid; e < 16 * TS; e += NT) {
      const int kk = e / TS, rr = e % TS;
      s_a[kk * TS + rr] = (kk < kc && rr < mr) ? L[(size_t)(k0 + kk) * ldl + rr] : T(0);
      /* U(k0+kk, cc): kk fastest for the global read, cc = e / 16 */
      const int cc = e / 16, k2 = e % 16;
      s_b[k2 * TS + cc] = (k2 < kc && cc < mc) ? U[(size_t)cc * ldu + k0 + k2] : T(0);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = s_a[kk * TS + ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = s_b[kk * TS + tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = ty * 4 + i, c = tx * 4 + j;
      if (r < mr && c < mc) C[(size_t)c * ldc + r] -= acc[i][j];
    }
}

/* ---------------------------------------------------------------------- */
/* Trailing updates on the matrix cores, one 64 x 64 tile per wavefront    */
/* ---------------------------------------------------------------------- */
typedef double td4 __attribute__((ext_vector_type(4)));
typedef float tf4 __attribute__((ext_vector_type(4)));
template <typename T> struct TMfma;
template <> struct TMfma<double> {
  typedef td4 acc_t;
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  /* v_mfma_f64_16x16x4f64: lane l, element r holds D[(l>>4) + 4r][l&15] */
  __device__ static int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct TMfma<float> {
  typedef tf4 acc_t;
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  /* v_mfma_f32_16x16x4f32: lane l, element r holds D[4(l>>4) + r][l&15] */
  __device__ static int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

/* C(tile) -= X Y^T for a 64 x 64 tile by one wavefront, K <= KMAX: X(r, k)
 * = X[r + k ldx] (rows r0 of the tile), Y(c, k) = Y[c*ycs + k*yks] (NT:
 * ycs = 1, yks = ld; NN with Y = U^T: ycs = ld, yks = 1).  Operands go from
 * memory straight into the MFMA lanes (lane l: row or column l&15 of a
 * 16-block, k = 4s + (l>>4)), no LDS and no barrier, so the wavefronts of a
 * workgroup work on different tiles at once.  lower_only: only r0 + r >=
 * c0 + c is written. */
template <typename T, int KMAX>
__device__ void wave_tile_update(T *C, long ldc, const T *X, long ldx, const T *Y, long ycs,
                                 long yks, int K, int mr, int mc, bool lower_only, int r0,
                                 int c0) {
  typedef TMfma<T> MF;
  typedef typename MF::acc_t acc_t;
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < KMAX / 4; ++s) {
    const int k = 4 * s + lk;
    if (4 * s >= K) break;                        /* uniform */
    T a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 16 * i + li;
      a[i] = (r < mr && k < K) ? X[(size_t)r + (size_t)k * ldx] : T(0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = 16 * j + li;
      b[j] = (c < mc && k < K) ? Y[(size_t)c * ycs + (size_t)k * yks] : T(0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = MF::mma(a[i], b[j], acc[i][j]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int rr = 16 * i + MF::drow(lane, r), cc = 16 * j + li;
        if (rr < mr && cc < mc && (!lower_only || r0 + rr >= c0 + cc))
          C[(size_t)cc * ldc + rr] -= acc[i][j][r];
      }
}

/* ---------------------------------------------------------------------- */
/* Triangular solve of one instance's right-hand side by one workgroup     */
/* ---------------------------------------------------------------------- */
/* op(A) X = X for the nrhs (<= MAXR) vectors of X in LDS (column r at
 * X + r n), A's triangle in global memory (lower: the stored triangle is
 * the lower one; trans: op(A) = A^T; unit: the diagonal is taken as 1).
 * Left-looking over blocks of SB rows in solve order (top first when
 * lower != trans), RCH right-hand sides at a time: the block's entries less
 * the solved entries' contributions (all four wavefronts, one pass over A
 * for the RCH vectors, then one barrier), then the SB x SB diagonal
 * triangle -- staged in LDS -- by wavefront 0 alone, one row per lane,
 * each solved value broadcast from its lane: no barrier per row.  s_blk
 * holds SB x (SB + 1), s_part 4 RCH SB elements. */
constexpr int SB = 32;
constexpr int RCH = 4;                             /* right-hand sides per pass */
constexpr int RCT = 2;                             /* the transposed GEMV's accumulator set */
template <typename T>
__device__ void wg_trsv(int n, int nrhs, const T *A, long lda, bool lower, bool trans, bool unit,
                        T *X, T *s_blk, T *s_part) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool fwd = lower != trans;
  const int nb = (n + SB - 1) / SB;
  for (int bs = 0; bs < nb; ++bs) {
    const int b = fwd ? bs : nb - 1 - bs;
    const int R0 = b * SB, rb = min(SB, n - R0);
    /* solved range: [0, R0) forward, [R0 + rb, n) backward */
    const int k0 = fwd ? 0 : R0 + rb, k1 = fwd ? R0 : n;
    __syncthreads();
    /* the diagonal block of op(A), [row][col] */
    for (int e = tid; e < SB * SB; e += NT) {
      const int i = e % SB, j = e / SB;
      T v = T(0);
      if (i < rb && j < rb && (fwd ? i >= j : i <= j))
        v = trans ? A[(size_t)(R0 + i) * lda + R0 + j] : A[(size_t)(R0 + j) * lda + R0 + i];
      s_blk[i * (SB + 1) + j] = v;
    }
    for (int rc = 0; rc < nrhs; rc += RCH) {
      if (rc > 0) __syncthreads();               /* wavefront 0 is done with s_part */
      if (!trans) {
        /* op(A)(i, k) = A[k lda + i]: thread (row l, chunk q) sums k = k0 + q, k0 + q + 8, ...;
         * the two chunks of a wavefront are added by a lane exchange */
        constexpr int ST = NT / SB, U = 8;       /* stride, loads in flight */
        const int l = tid & (SB - 1), q = tid / SB;
        T part[RCH];
#pragma unroll
        for (int r = 0; r < RCH; ++r) part[r] = T(0);
        if (l < rb) {
          int k = k0 + q;
          for (; k + (U - 1) * ST < k1; k += U * ST) {
            T a[U];
#pragma unroll
            for (int u = 0; u < U; ++u) a[u] = A[(size_t)(k + u * ST) * lda + R0 + l];
#pragma unroll
            for (int r = 0; r < RCH; ++r)
              if (rc + r < nrhs)
#pragma unroll
                for (int u = 0; u < U; ++u) part[r] += a[u] * X[(size_t)(rc + r) * n + k + u * ST];
          }
          for (; k < k1; k += ST) {
            const T a = A[(size_t)k * lda + R0 + l];
#pragma unroll
            for (int r = 0; r < RCH; ++r)
              if (rc + r < nrhs) part[r] += a * X[(size_t)(rc + r) * n + k];
          }
        }
#pragma unroll
        for (int r = 0; r < RCH; ++r) {
          const T v = part[r] + __shfl_xor(part[r], 32);
          if (lane < 32) s_part[(r * 4 + wave) * SB + l] = v;
        }
      } else {
        /* op(A)(i, k) = A[i lda + k]: wavefront w takes rows w, w + 4, ...,
         * lanes run along k, all loads of a step issued together */
        constexpr int RW = SB / (NT / 64);
        for (int rt = 0; rt < RCH && rc + rt < nrhs; rt += RCT) {
          T part[RW][RCT];
#pragma unroll
          for (int m = 0; m < RW; ++m)
#pragma unroll
            for (int r = 0; r < RCT; ++r) part[m][r] = T(0);
          for (int k = k0 + lane; k < k1; k += 64) {
            T a[RW];
#pragma unroll
            for (int m = 0; m < RW; ++m) {
              const int l = wave + m * (NT / 64);
              a[m] = l < rb ? A[(size_t)(R0 + l) * lda + k] : T(0);
            }
#pragma unroll
            for (int r = 0; r < RCT; ++r)
              if (rc + rt + r < nrhs) {
                const T xk = X[(size_t)(rc + rt + r) * n + k];
#pragma unroll
                for (int m = 0; m < RW; ++m) part[m][r] += a[m] * xk;
              }
          }
#pragma unroll
          for (int m = 0; m < RW; ++m)
#pragma unroll
            for (int r = 0; r < RCT; ++r) {
              T v = part[m][r];
              for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
              if (lane == 0) s_part[((rt + r) * 4) * SB + wave + m * (NT / 64)] = v;
            }
        }
      }
      __syncthreads();
      if (wave == 0) {
        T c[RCH];
#pragma unroll
        for (int r = 0; r < RCH; ++r) {
          c[r] = T(0);
          if (rc + r < nrhs && lane < rb) {
            T sum = s_part[(r * 4) * SB + lane];
            if (!trans)
              for (int w = 1; w < NT / 64; ++w) sum += s_part[(r * 4 + w) * SB + lane];
            c[r] = X[(size_t)(rc + r) * n + R0 + lane] - sum;
          }
        }
        for (int s = 0; s < SB; ++s) {
          const int j = fwd ? s : SB - 1 - s;
          if (j < rb) {
            const T dj = unit ? T(1) : s_blk[j * (SB + 1) + j];
            const T lj = (lane < rb && (fwd ? (lane > j) : (lane < j)))
                             ? s_blk[lane * (SB + 1) + j] : T(0);
#pragma unroll
            for (int r = 0; r < RCH; ++r) {
              if (rc + r < nrhs) {
                if (lane == j) c[r] /= dj;
                const T xj = __shfl(c[r], j);
                c[r] -= lj * xj;
              }
            }
          }
        }
        if (lane < rb)
#pragma unroll
          for (int r = 0; r < RCH; ++r)
            if (rc + r < nrhs) X[(size_t)(rc + r) * n + R0 + lane] = c[r];
      }
    }
  }
  __syncthreads();
}

/* ---------------------------------------------------------------------- */
/* Cholesky factor + solve of one instance by one workgroup                */
/* ---------------------------------------------------------------------- */
template <typename T, int NB>
__device__ void wg_posv(int n, T *A, long lda, int nrhs, T *B, long ldb, int *info, T *s_d,
                        T *s_a, T *s_b, T *s_x, int *s_flag) {
  const int tid = threadIdx.x;
  for (int k = 0; k < n; k += NB) {
    const int kb = min(NB, n - k);
    /* ---- diagonal block: potf2 in LDS (ATL_potrfL.c's leaf, unblocked) ---- */
    __syncthreads();
    for (int e = tid; e < NB * NB; e += NT) {
      const int r = e % NB, c = e / NB;
      s_d[c * (NB + 1) + r] = (r < kb && c < kb && r >= c) ? A[(size_t)(k + c) * lda + k + r] : T(0);
    }
    if (tid == 0) *s_flag = 0;
    __syncthreads();
    /* wavefront 0 alone, one row per lane in registers: column j's values
     * reach the other lanes by readlane, so no barrier per column */
    if (threadIdx.x < 64) {
      const int lane = threadIdx.x;
      T a[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) a[c] = (lane < kb && c <= lane) ? s_d[c * (NB + 1) + lane] : T(0);
      int fail = 0;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        if (j < kb && !fail) {
          const T d = bcast(a[j], j);
          if (!(d > T(0))) {
            fail = j + 1;                         /* uniform: every lane saw the same d */
          } else {
            const T sd = dsqrt(d);
            if (lane == j) a[j] = sd;
            else if (lane > j) a[j] = a[j] / sd;
#pragma unroll
            for (int c = j + 1; c < NB; ++c) {
              const T lcj = bcast(a[j], c);
              if (lane >= c && lane < kb) a[c] -= a[j] * lcj;
            }
          }
        }
      }
      if (lane < kb)
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c <= lane) s_d[c * (NB + 1) + lane] = a[c];
      if (lane == 0) *s_flag = fail;
    }
    __syncthreads();
    if (*s_flag) {
      if (tid == 0) *info = k + *s_flag;
      return;                               /* ATL_potrfL.c:238-240: stop */
    }
    for (int e = tid; e < NB * NB; e += NT) {
      const int r = e % NB, c = e / NB;
      if (r < kb && c < kb && r >= c) A[(size_t)(k + c) * lda + k + r] = s_d[c * (NB + 1) + r];
    }
    /* ---- panel below: row i solves x L_kk^T = a ---- */
    for (int i = k + kb + tid; i < n; i += NT) {
      T x[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) x[c] = (c < kb) ? A[(size_t)(k + c) * lda + i] : T(0);
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c < kb) {
          T v = x[c];
#pragma unroll
          for (int l = 0; l < c; ++l) v -= x[l] * s_d[l * (NB + 1) + c];
          x[c] = v / s_d[c * (NB + 1) + c];
        }
      }
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < kb) A[(size_t)(k + c) * lda + i] = x[c];
    }
    __syncthreads();
    /* ---- trailing lower triangle: A22 -= P P^T in 64 x 64 tiles ---- */
    const int m0 = k + kb, m = n - m0;
    const int nt = (m + TS - 1) / TS;
    const T *P = A + (size_t)k * lda;
    {
      /* the lower triangle's tiles, wavefront w taking every fourth */
      const int wave = threadIdx.x >> 6;
      int t = 0;
      for (int tj = 0; tj < nt; ++tj)
        for (int ti = tj; ti < nt; ++ti, ++t) {
          if ((t & 3) != wave) continue;
          const int r0 = m0 + ti * TS, c0 = m0 + tj * TS;
          wave_tile_update<T, NB>(A + (size_t)c0 * lda + r0, lda, P + r0, lda, P + c0, 1, lda,
                                  kb, min(TS, n - r0), min(TS, n - c0), ti == tj, r0, c0);
        }
    }
  }
  if (tid == 0) *info = 0;
  if (nrhs <= 0) return;
  /* ---- potrs: L y = b, L^T x = y, right-hand sides in LDS; the factor
   * phase's LDS (s_d, s_a, s_b, contiguous) is the solve's staging ---- */
  __syncthreads();
  for (int e = tid; e < n * nrhs; e += NT) s_x[e] = B[(size_t)(e / n) * ldb + e % n];
  wg_trsv(n, nrhs, A, lda, true, false, false, s_x, s_d, s_d + SB * (SB + 1));
  wg_trsv(n, nrhs, A, lda, true, true, false, s_x, s_d, s_d + SB * (SB + 1));
  for (int e = tid; e < n * nrhs; e += NT) B[(size_t)(e / n) * ldb + e % n] = s_x[e];
}

/* ---------------------------------------------------------------------- */
/* LU (partial pivoting) factor + solve of one instance by one workgroup   */
/* ---------------------------------------------------------------------- */
template <typename T, int NB>
__device__ void wg_gesv(int n, T *A, long lda, int *ipiv, int nrhs, T *B, long ldb, int *info,
                        T *s_p, T *s_a, T *s_b, T *s_x, T *s_v, int *s_i, bool trans) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int first0 = 0;                          /* first exactly-zero pivot, 1-based */
  for (int k = 0; k < n; k += NB) {
    const int kb = min(NB, n - k), mp = n - k;
    /* ---- panel rows k..n-1, columns k..k+kb-1 into LDS (s_p[c*mp + r]) ---- */
    __syncthreads();
    for (int e = tid; e < mp * kb; e += NT) {
      const int r = e % mp, c = e / mp;
      s_p[c * mp + r] = A[(size_t)(k + c) * lda + k + r];
    }
    __syncthreads();
    for (int j = 0; j < kb; ++j) {
      /* arg-max |p(r, j)|, r >= j: first maximum, NaN never wins */
      T bv = T(-1);
      int bi = 0x7fffffff;
      for (int r = j + tid; r < mp; r += NT) {
        const T v = fabs(s_p[j * mp + r]);
        if (v > bv) { bv = v; bi = r; }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const T ov = __shfl_xor(bv, off);
        const int oi = __shfl_xor(bi, off);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
      }
      if (lane == 0) { s_v[wave] = bv; s_i[wave] = bi; }
      __syncthreads();
      if (tid == 0) {
        T v = s_v[0];
        int p = s_i[0];
        for (int q = 1; q < NT / 64; ++q)
          if (s_v[q] > v || (s_v[q] == v && s_i[q] < p)) { v = s_v[q]; p = s_i[q]; }
        if (!(v > T(0))) p = j;            /* all zero/NaN: cblas_i?amax answers the first */
        s_i[4] = p;
      }
      __syncthreads();
      const int p = s_i[4];
      if (tid == 0) ipiv[k + j] = k + p;
      if (p != j && tid < kb) {
        const T t = s_p[tid * mp + j];
        s_p[tid * mp + j] = s_p[tid * mp + p];
        s_p[tid * mp + p] = t;
      }
      __syncthreads();
      const T piv = s_p[j * mp + j];
      if (piv == T(0)) {
        if (!first0) first0 = k + j + 1;          /* every thread: uniform control flow */
      } else {
        for (int r = j + 1 + tid; r < mp; r += NT) {
          const T l = s_p[j * mp + r] / piv;
          s_p[j * mp + r] = l;
          for (int c = j + 1; c < kb; ++c) s_p[c * mp + r] -= l * s_p[c * mp + j];
        }
      }
      __syncthreads();
    }
    for (int e = tid; e < mp * kb; e += NT) {
      const int r = e % mp, c = e / mp;
      A[(size_t)(k + c) * lda + k + r] = s_p[c * mp + r];
    }
    /* ---- the panel's interchanges on the other columns ---- */
    for (int c = tid; c < n; c += NT) {
      if (c >= k && c < k + kb) continue;
      T *col = A + (size_t)c * lda;
      for (int j = 0; j < kb; ++j) {
        const int p = ipiv[k + j];
        if (p != k + j) { const T t = col[k + j]; col[k + j] = col[p]; col[p] = t; }
      }
    }
    __syncthreads();
    /* ---- U12 = L11^{-1} A12, one column per thread ---- */
    for (int c = k + kb + tid; c < n; c += NT) {
      T x[NB];
#pragma unroll
      for (int r = 0; r < NB; ++r) x[r] = (r < kb) ? A[(size_t)c * lda + k + r] : T(0);
#pragma unroll
      for (int r = 0; r < NB; ++r) {
        if (r < kb) {
          T v = x[r];
#pragma unroll
          for (int l = 0; l < r; ++l) v -= s_p[l * mp + r] * x[l];
          x[r] = v;
        }
      }
#pragma unroll
      for (int r = 0; r < NB; ++r)
        if (r < kb) A[(size_t)c * lda + k + r] = x[r];
    }
    __syncthreads();
    /* ---- A22 -= L21 U12 ---- */
    const int m0 = k + kb, m = n - m0;
    const int nt = (m + TS - 1) / TS;
    for (int t = wave; t < nt * nt; t += NT / 64) {
      const int ti = t % nt, tj = t / nt;         /* wavefront w takes every fourth tile */
      const int r0 = m0 + ti * TS, c0 = m0 + tj * TS;
      wave_tile_update<T, NB>(A + (size_t)c0 * lda + r0, lda, A + (size_t)k * lda + r0, lda,
                              A + (size_t)c0 * lda + k, lda, 1, kb, min(TS, n - r0),
                              min(TS, n - c0), false, r0, c0);
    }
  }
  if (tid == 0) *info = first0;
  if (nrhs <= 0 || first0) return;
  /* ---- getrs in LDS (ATL_getrs.c:8-89): NoTrans: P b, L y = b (unit),
   * U x = y; Trans (A^T = U^T L^T P): U^T z = b, L^T w = z (unit), x = P^T w.
   * The triangular solves stage through s_a, s_b (contiguous) ---- */
  __syncthreads();
  for (int e = tid; e < n * nrhs; e += NT) s_x[e] = B[(size_t)(e / n) * ldb + e % n];
  __syncthreads();
  if (!trans && tid < nrhs) {
    for (int j = 0; j < n; ++j) {
      const int p = ipiv[j];
      if (p != j) { const T t = s_x[tid * n + j]; s_x[tid * n + j] = s_x[tid * n + p]; s_x[tid * n + p] = t; }
    }
  }
  if (!trans) {
    wg_trsv(n, nrhs, A, lda, true, false, true, s_x, s_a, s_a + SB * (SB + 1));
    wg_trsv(n, nrhs, A, lda, false, false, false, s_x, s_a, s_a + SB * (SB + 1));
  } else {
    wg_trsv(n, nrhs, A, lda, false, true, false, s_x, s_a, s_a + SB * (SB + 1));
    wg_trsv(n, nrhs, A, lda, true, true, true, s_x, s_a, s_a + SB * (SB + 1));
  }
  if (trans && tid < nrhs) {                 /* x = P^T w: the swaps in reverse */
    for (int j = n - 1; j >= 0; --j) {
      const int p = ipiv[j];
      if (p != j) { const T t = s_x[tid * n + j]; s_x[tid * n + j] = s_x[tid * n + p]; s_x[tid * n + p] = t; }
    }
  }
  __syncthreads();
  for (int e = tid; e < n * nrhs; e += NT) B[(size_t)(e / n) * ldb + e % n] = s_x[e];
}

/* ---------------------------------------------------------------------- */
/* Cholesky adjoint of one instance by one workgroup                       */
/* ---------------------------------------------------------------------- */
/* C(tile r0.., c0..) = sum_k Aop(r, k) Bop(k, c) over k in [k0, k1), where
 * Aop(r, k) = A[r*ar + k*ak] and Bop(k, c) = B[k*bk + c*bc] (strides give
 * the transposes); only r >= c is written when lower_only. */
template <typename T>
__device__ void tile_gemm(T *C, long ldc, const T *A, long ar, long ak, const T *B, long bk,
                          long bc, int k0, int k1, int mr, int mc, bool lower_only, int r0,
                          int c0, T *s_a, T *s_b) {
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = T(0);
  for (int kb = k0; kb < k1; kb += 16) {
    const int kc = min(16, k1 - kb);
    __syncthreads();
    for (int e = tid; e < 16 * TS; e += NT) {
      const int kk = e / TS, rr = e % TS;
      s_a[kk * TS + rr] = (kk < kc && rr < mr) ? A[(size_t)rr * ar + (size_t)(kb + kk) * ak] : T(0);
      s_b[kk * TS + rr] = (kk < kc && rr < mc) ? B[(size_t)(kb + kk) * bk + (size_t)rr * bc] : T(0);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = s_a[kk * TS + ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = s_b[kk * TS + tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = ty * 4 + i, c = tx * 4 + j;
      if (r < mr && c < mc && (!lower_only || r0 + r >= c0 + c)) C[(size_t)c * ldc + r] = acc[i][j];
    }
}

/* The adjoint's two triangular solves for n <= NT, in registers: the
 * right-hand sides are independent, so a block of ADJ_CB of them is solved
 * at a time with one unknown per thread.  Step i (from the last) has the
 * owner of unknown i divide by L(i,i) and publish its ADJ_CB values through
 * LDS (two buffers by the parity of i), then every thread t < i subtracts
 * L(i, t) times them: one barrier per step.  L(i, t) is loaded ADJ_PF
 * steps ahead into a ring of registers (the steps run in unrolled groups of
 * ADJ_PF, so every ring index is fixed).
 *   lt = 1: Y = L^{-T} W (W's columns are the right-hand sides; thread t
 *           holds row t of a block of columns);
 *   lt = 0: S = W L^{-1} (W's rows are; thread t holds column t of a block
 *           of rows).
 * Both read L(i, t) = L[t*lda + i] at step i. */
constexpr int ADJ_CB = 16;
constexpr int ADJ_PF = 8;

template <typename T>
__device__ void wg_adj_solve(int n, const T *L, long lda, T *W, int lt,
                             T (*s_y)[ADJ_CB]) {
  const int t = threadIdx.x;
  for (int b0 = 0; b0 < n; b0 += ADJ_CB) {
    const int bb = min(ADJ_CB, n - b0);
    T y[ADJ_CB];
#pragma unroll
    for (int q = 0; q < ADJ_CB; ++q)
      y[q] = (t < n && q < bb) ? (lt ? W[(size_t)(b0 + q) * n + t] : W[(size_t)t * n + b0 + q])
                               : T(0);
    T ring[ADJ_PF];
#pragma unroll
    for (int q = 0; q < ADJ_PF; ++q) {
      const int i = n - 1 - q;
      ring[q] = (i >= 0 && t < i) ? L[(size_t)t * lda + i] : T(0);
    }
    for (int i0 = n - 1; i0 >= 0; i0 -= ADJ_PF) {
#pragma unroll
      for (int q = 0; q < ADJ_PF; ++q) {
        const int i = i0 - q;
        if (i >= 0) {
          const T li = ring[q];
          const int ip = i - ADJ_PF;               /* the step this slot serves next */
          ring[q] = (ip >= 0 && t < ip) ? L[(size_t)t * lda + ip] : T(0);
          T *buf = s_y[i & 1];
          if (t == i) {
            const T rd = T(1) / L[(size_t)i * lda + i];
#pragma unroll
            for (int r = 0; r < ADJ_CB; ++r) {
              y[r] *= rd;
              buf[r] = y[r];
            }
          }
          __syncthreads();
          if (t < i) {
#pragma unroll
            for (int r = 0; r < ADJ_CB; ++r) y[r] -= li * buf[r];
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < ADJ_CB; ++q)
      if (t < n && q < bb) {
        if (lt) W[(size_t)(b0 + q) * n + t] = y[q];
        else W[(size_t)t * n + b0 + q] = y[q];
      }
    __syncthreads();
  }
}

/* Given L (lower triangle of A after potrf; its strict upper triangle is
 * overwritten with zeros here), X = A^{-1} B and Bb = A^{-1} Xb (the solve's
 * adjoint), write to G the gradient of <Xb, X> with respect to the lower
 * triangle of A (the entries potrf reads; strict upper part zero):
 *   Lb = tril(-(Bb X^T + X Bb^T) L)            (potrs adjoint: L's gradient)
 *   P  = Phi(L^T Lb)   (lower triangle, diagonal halved)
 *   S  = L^{-T} P L^{-1}                        (two triangular solves)
 *   G  = tril(S + S^T) - diag(S)  (the symmetric gradient (S + S^T)/2, each
 *        off-diagonal entry counted for both of its positions)
 * (Murray, "Differentiation of the Cholesky decomposition", 2016).  W: an
 * n x n scratch (leading dimension n). */
template <typename T>
__device__ void wg_chol_adjoint(int n, T *L, long lda, int nrhs, const T *Xs, const T *Bb,
                                long ldx, T *G, long ldg, T *W, T *s_a, T *s_b,
                                T (*s_y)[ADJ_CB]) {
  const int tid = threadIdx.x;
  const int nt = (n + TS - 1) / TS;
  /* W = -(Bb X^T + X Bb^T); G = 0; strict upper of L = 0 */
  for (int e = tid; e < n * n; e += NT) {
    const int i = e % n, j = e / n;
    T v = T(0);
    for (int r = 0; r < nrhs; ++r)
      v += Bb[(size_t)r * ldx + i] * Xs[(size_t)r * ldx + j] +
           Xs[(size_t)r * ldx + i] * Bb[(size_t)r * ldx + j];
    W[(size_t)j * n + i] = -v;
    G[(size_t)j * ldg + i] = T(0);
    if (i < j) L[(size_t)j * lda + i] = T(0);
  }
  __syncthreads();
  /* G = tril(W L): G(i, j) = sum_{k >= j} W(i, k) L(k, j) */
  for (int tj = 0; tj < nt; ++tj)
    for (int ti = tj; ti < nt; ++ti) {
      const int r0 = ti * TS, c0 = tj * TS;
      tile_gemm(G + (size_t)c0 * ldg + r0, ldg, W + r0, 1, n, L + (size_t)c0 * lda, 1, lda, c0,
                n, min(TS, n - r0), min(TS, n - c0), ti == tj, r0, c0, s_a, s_b);
    }
  __syncthreads();
  /* W = L^T G on the lower triangle: P(i, j) = sum_{k >= i} L(k, i) G(k, j) */
  for (int tj = 0; tj < nt; ++tj)
    for (int ti = tj; ti < nt; ++ti) {
      const int r0 = ti * TS, c0 = tj * TS;
      tile_gemm(W + (size_t)c0 * n + r0, n, L + (size_t)r0 * lda, lda, 1, G + (size_t)c0 * ldg,
                1, ldg, r0, n, min(TS, n - r0), min(TS, n - c0), ti == tj, r0, c0, s_a, s_b);
    }
  __syncthreads();
  /* Phi: halve the diagonal, zero the strict upper triangle */
  for (int e = tid; e < n * n; e += NT) {
    const int i = e % n, j = e / n;
    if (i < j) W[(size_t)j * n + i] = T(0);
    else if (i == j) W[(size_t)j * n + i] *= T(0.5);
  }
  __syncthreads();
  if (n <= NT) {
    wg_adj_solve<T>(n, L, lda, W, 1, s_y);   /* Y = L^{-T} P */
    wg_adj_solve<T>(n, L, lda, W, 0, s_y);   /* S = Y L^{-1} */
  } else {
  /* Y = L^{-T} P in place: rows from the last; Y(i,:) /= L(i,i), then rows
   * k < i lose L(i, k) Y(i,:) */
  for (int i = n - 1; i >= 0; --i) {
    const T d = L[(size_t)i * lda + i];
    for (int c = tid; c < n; c += NT) W[(size_t)c * n + i] /= d;
    __syncthreads();
    for (int e = tid; e < i * n; e += NT) {
      const int k = e % i, c = e / i;
      W[(size_t)c * n + k] -= L[(size_t)k * lda + i] * W[(size_t)c * n + i];
    }
    __syncthreads();
  }
  /* S = Y L^{-1} in place: columns from the last; S(:,j) /= L(j,j), then
   * columns j' < j lose S(:, j) L(j, j') */
  for (int j = n - 1; j >= 0; --j) {
    const T d = L[(size_t)j * lda + j];
    for (int r = tid; r < n; r += NT) W[(size_t)j * n + r] /= d;
    __syncthreads();
    for (int e = tid; e < n * j; e += NT) {
      const int r = e % n, jp = e / n;
      W[(size_t)jp * n + r] -= W[(size_t)j * n + r] * L[(size_t)jp * lda + j];
    }
    __syncthreads();
  }
  }
  /* G = tril(S + S^T) - diag(S) */
  for (int e = tid; e < n * n; e += NT) {
    const int i = e % n, j = e / n;
    T v = T(0);
    if (i > j) v = W[(size_t)j * n + i] + W[(size_t)i * n + j];
    else if (i == j) v = W[(size_t)j * n + i];
    G[(size_t)j * ldg + i] = v;
  }
  __syncthreads();
}

template <typename T>
struct BatchArgs {
  int count;
  const int *order;              /* instance processing order */
  const int *n;
  T *const *A; const int *lda;
  int *const *ipiv;              /* gesv */
  int nrhs;
  T *const *B; const int *ldb;
  int *info;
  int *counter;                  /* zero on entry; left at count + grid */
};

template <typename T, int NB>
__global__ __launch_bounds__(NT) void posv_batched_kernel(BatchArgs<T> g) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  T *s_d = (T *)smem;                                  /* NB x (NB+1) */
  T *s_a = s_d + NB * (NB + 1);                        /* 16 x 64 */
  T *s_b = s_a + 16 * TS;                              /* 16 x 64 */
  T *s_x = s_b + 16 * TS;                              /* n x nrhs */
  __shared__ int s_item, s_flag;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(g.counter, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= g.count) return;
    const int i = g.order[it];
    wg_posv<T, NB>(g.n[i], g.A[i], g.lda[i], g.nrhs, g.B ? g.B[i] : nullptr,
                   g.ldb ? g.ldb[i] : 1, &g.info[i], s_d, s_a, s_b, s_x, &s_flag);
  }
}

template <typename T, int NB>
__global__ __launch_bounds__(NT) void gesv_batched_kernel(BatchArgs<T> g, int maxn, int trans) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  T *s_p = (T *)smem;                                  /* maxn x NB */
  T *s_a = s_p + (size_t)maxn * NB;                    /* 16 x 64 */
  T *s_b = s_a + 16 * TS;
  T *s_x = s_b + 16 * TS;                              /* maxn x nrhs */
  __shared__ T s_v[4];
  __shared__ int s_i[5];
  __shared__ int s_item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(g.counter, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= g.count) return;
    const int i = g.order[it];
    wg_gesv<T, NB>(g.n[i], g.A[i], g.lda[i], g.ipiv[i], g.nrhs, g.B ? g.B[i] : nullptr,
                   g.ldb ? g.ldb[i] : 1, &g.info[i], s_p, s_a, s_b, s_x, s_v, s_i, trans != 0);
  }
}

/* ---------------------------------------------------------------------- */
/* variable-size batched GEMM: one work item = one 64 x 64 tile           */
/* ---------------------------------------------------------------------- */
template <typename T>
struct GemmBatchArgs {
  int count, ta, tb;
  const int *order;              /* instance order of the tile numbering */
  const int *m, *n, *k;
  T alpha, beta;
  const T *const *A; const int *lda;
  const T *const *B; const int *ldb;
  T *const *C; const int *ldc;
  const int *tile_start;         /* prefix sums of tiles per position in order (count + 1) */
  int *counter;
};

template <typename T>
__global__ __launch_bounds__(NT) void gemm_batched_kernel(GemmBatchArgs<T> g) {
  __shared__ T s_a[16 * TS], s_b[16 * TS];
  __shared__ int s_item;
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int total = g.tile_start[g.count];
  for (;;) {
    __syncthreads();
    if (tid == 0) s_item = atomicAdd(g.counter, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= total) return;
    /* instance: the last i with tile_start[i] <= it */
    int lo = 0, hi = g.count - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (g.tile_start[mid] <= it) lo = mid; else hi = mid - 1;
    }
    const int i = g.order[lo], t = it - g.tile_start[lo];
    const int M = g.m[i], N = g.n[i], K = g.k[i];
    const int tm = (M + TS - 1) / TS;
    const int r0 = (t % tm) * TS, c0 = (t / tm) * TS;
    const T *A = g.A[i], *B = g.B[i];
    const long lda = g.lda[i], ldb = g.ldb[i], ldc = g.ldc[i];
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = T(0);
    for (int k0 = 0; k0 < K; k0 += 16) {
      __syncthreads();
      for (int e = tid; e < 16 * TS; e += NT) {
        /* op(A)(r0+rr, k0+kk), op(B)(k0+kk, c0+cc) */
        int rr, kk;
        if (g.ta == KX_NOTRANS) { kk = e / TS; rr = e % TS; }
        else                    { rr = e / 16; kk = e % 16; }
        const int gr = r0 + rr, gk = k0 + kk;
        s_a[kk * TS + rr] = (gr < M && gk < K)
            ? (g.ta == KX_NOTRANS ? A[(size_t)gk * lda + gr] : A[(size_t)gr * lda + gk]) : T(0);
        int cc, k2;
        if (g.tb == KX_NOTRANS) { cc = e / 16; k2 = e % 16; }
        else                    { k2 = e / TS; cc = e % TS; }
        const int gc = c0 + cc, gk2 = k0 + k2;
        s_b[k2 * TS + cc] = (gc < N && gk2 < K)
            ? (g.tb == KX_NOTRANS ? B[(size_t)gc * ldb + gk2] : B[(size_t)gk2 * ldb + gc]) : T(0);
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < 16; ++kk) {
        T a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) a[q] = s_a[kk * TS + ty * 4 + q];
#pragma unroll
        for (int q = 0; q < 4; ++q) b[q] = s_b[kk * TS + tx * 4 + q];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
          for (int q = 0; q < 4; ++q) acc[p][q] += a[p] * b[q];
      }
    }
    T *C = g.C[i];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        const int r = r0 + ty * 4 + p, c = c0 + tx * 4 + q;
        if (r < M && c < N) {
          T *pc = C + (size_t)c * ldc + r;
          T v = g.alpha * acc[p][q];
          if (g.beta != T(0)) v += g.beta * *pc;
          *pc = v;
        }
      }
  }
}

template <typename T>
struct AdjArgs {
  int count, nrhs;
  const int *order, *n, *lda, *ldx, *ldg;
  T *const *L; T *const *X; T *const *Bb; T *const *G; T *const *W;
  /* L's strict upper triangle is zeroed */
  int *counter;
};

template <typename T>
__global__ __launch_bounds__(NT) void chol_adjoint_kernel(AdjArgs<T> g) {
  __shared__ T s_a[16 * TS], s_b[16 * TS];
  __shared__ T s_y[2][ADJ_CB];
  __shared__ int s_item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(g.counter, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= g.count) return;
    const int i = g.order[it];
    wg_chol_adjoint<T>(g.n[i], g.L[i], g.lda[i], g.nrhs, g.X[i], g.Bb[i], g.ldx[i], g.G[i],
                       g.ldg[i], g.W[i], s_a, s_b, s_y);
  }
}

template <typename T, int NB>
hipError_t launch_posv(const BatchArgs<T> &g, int maxn, int grid, hipStream_t s) {
  const size_t lds = sizeof(T) * ((size_t)NB * (NB + 1) + 2 * 16 * TS + (size_t)maxn * MAXR);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void *)posv_batched_kernel<T, NB>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((posv_batched_kernel<T, NB>), dim3(grid), dim3(NT), lds, s, g);
  return hipGetLastError();
}

template <typename T, int NB>
hipError_t launch_gesv(const BatchArgs<T> &g, int maxn, int grid, int trans, hipStream_t s) {
  const size_t lds = sizeof(T) * ((size_t)maxn * NB + 2 * 16 * TS + (size_t)maxn * MAXR);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void *)gesv_batched_kernel<T, NB>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((gesv_batched_kernel<T, NB>), dim3(grid), dim3(NT), lds, s, g, maxn, trans);
  return hipGetLastError();
}

template <typename T>
hipError_t batch_factor(int kind, int nb, int count, const int *order, const int *n, T *const *A,
                        const int *lda, int *const *ipiv, int nrhs, T *const *B, const int *ldb,
                        int *info, int *counter, int maxn, int grid, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (nrhs < 0 || nrhs > MAXR || maxn <= 0 || grid <= 0) return hipErrorInvalidValue;
  BatchArgs<T> g;
  g.count = count; g.order = order; g.n = n; g.A = A; g.lda = lda; g.ipiv = ipiv;
  g.nrhs = nrhs; g.B = B; g.ldb = ldb; g.info = info; g.counter = counter;
  if (kind == 0) {
    if (nb == 16) return launch_posv<T, 16>(g, maxn, grid, s);
    if (nb == 32) return launch_posv<T, 32>(g, maxn, grid, s);
  } else if (kind == 1 || kind == 2) {      /* 2: solve with A^T */
    if (nb == 16) return launch_gesv<T, 16>(g, maxn, grid, kind == 2, s);
    if (nb == 32) return launch_gesv<T, 32>(g, maxn, grid, kind == 2, s);
  }
  return hipErrorInvalidValue;
}

/* The same work items on the matrix cores: a 64 x 64 tile of C per item,
 * NW wavefronts per tile (1: one wavefront holds the whole tile as 4 x 4
 * MFMA blocks of 16 x 16; 4: a 2 x 2 grid of wavefronts, each 32 x 32 =
 * 2 x 2 blocks), K in slices of 16 staged through LDS as [k][row], the next
 * slice's loads issued before the current slice's MFMAs.  FP64:
 * v_mfma_f64_16x16x4f64 (lane l, element r of the accumulator holds
 * D[(l>>4) + 4r][l&15]); FP32: v_mfma_f32_16x16x4f32 (D[4(l>>4) + r][l&15]).
 * In one MFMA lane l supplies op(A)(row l&15, k slot l>>4) and op(B)(k slot
 * l>>4, column l&15); over the four sub-steps s4 the slots l>>4 = 0..3 take
 * k = 4 (l>>4) + s4, so the 16 k of a slice are each used once.  The
 * choice of NW is the autotuner's. */
typedef double bd4 __attribute__((ext_vector_type(4)));
typedef float bf4 __attribute__((ext_vector_type(4)));
template <typename T> struct BMfma;
template <> struct BMfma<double> {
  typedef bd4 acc_t;
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct BMfma<float> {
  typedef bf4 acc_t;
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename T, int NW>
__global__ __launch_bounds__(64 * NW) void gemm_batched_mfma_kernel(GemmBatchArgs<T> g) {
  typedef BMfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int NTH = 64 * NW;
  constexpr int KS = 16;                       /* K slice */
  constexpr int LDT = TS + 4;                  /* LDS row: [k][row], padded */
  constexpr int PER = KS * TS / NTH;           /* elements of a slice per thread */
  constexpr int BPW = (NW == 1) ? 4 : 2;       /* 16 x 16 blocks per wavefront side */
  __shared__ T s_a[KS * LDT], s_b[KS * LDT];
  __shared__ int s_item;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (NW == 1) ? 0 : (wave & 1) * 32, wn = (NW == 1) ? 0 : (wave >> 1) * 32;
  const int lr = lane & 15, kq = (lane >> 4) * 4;
  const int total = g.tile_start[g.count];
  for (;;) {
    __syncthreads();
    if (tid == 0) s_item = atomicAdd(g.counter, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= total) return;
    int lo = 0, hi = g.count - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (g.tile_start[mid] <= it) lo = mid; else hi = mid - 1;
    }
    const int i = g.order[lo], t = it - g.tile_start[lo];
    const int M = g.m[i], N = g.n[i], K = g.k[i];
    const int tm = (M + TS - 1) / TS;
    const int r0 = (t % tm) * TS, c0 = (t / tm) * TS;
    const T *A = g.A[i], *B = g.B[i];
    const long lda = g.lda[i], ldb = g.ldb[i], ldc = g.ldc[i];
    /* element e of a slice: op(A)(r0 + ra, k0 + ka) and op(B)(k0 + kb, c0 + cb),
     * numbered so that consecutive threads read consecutive addresses */
    T va[PER], vb[PER];
    auto load = [&](int k0) {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int e = tid + NTH * q;
        int ra, ka, kb, cb;
        if (g.ta == KX_NOTRANS) { ka = e / TS; ra = e % TS; } else { ra = e / KS; ka = e % KS; }
        if (g.tb == KX_NOTRANS) { cb = e / KS; kb = e % KS; } else { kb = e / TS; cb = e % TS; }
        const int gr = r0 + ra, gka = k0 + ka, gc = c0 + cb, gkb = k0 + kb;
        va[q] = (gr < M && gka < K)
                    ? (g.ta == KX_NOTRANS ? A[(size_t)gka * lda + gr] : A[(size_t)gr * lda + gka])
                    : T(0);
        vb[q] = (gc < N && gkb < K)
                    ? (g.tb == KX_NOTRANS ? B[(size_t)gc * ldb + gkb] : B[(size_t)gkb * ldb + gc])
                    : T(0);
      }
    };
    auto store = [&]() {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int e = tid + NTH * q;
        int ra, ka, kb, cb;
        if (g.ta == KX_NOTRANS) { ka = e / TS; ra = e % TS; } else { ra = e / KS; ka = e % KS; }
        if (g.tb == KX_NOTRANS) { cb = e / KS; kb = e % KS; } else { kb = e / TS; cb = e % TS; }
        s_a[ka * LDT + ra] = va[q];
        s_b[kb * LDT + cb] = vb[q];
      }
    };
    acc_t acc[BPW][BPW];
#pragma unroll
    for (int a = 0; a < BPW; ++a)
#pragma unroll
      for (int b = 0; b < BPW; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
    if (K > 0) load(0);
    for (int k0 = 0; k0 < K; k0 += KS) {
      __syncthreads();
      store();
      __syncthreads();
      if (k0 + KS < K) load(k0 + KS);
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        T a[BPW], b[BPW];
#pragma unroll
        for (int x = 0; x < BPW; ++x) {
          a[x] = s_a[(kq + s4) * LDT + wm + 16 * x + lr];
          b[x] = s_b[(kq + s4) * LDT + wn + 16 * x + lr];
        }
#pragma unroll
        for (int x = 0; x < BPW; ++x)
#pragma unroll
          for (int y = 0; y < BPW; ++y) acc[x][y] = MF::mma(a[x], b[y], acc[x][y]);
      }
    }
    T *C = g.C[i];
#pragma unroll
    for (int x = 0; x < BPW; ++x)
#pragma unroll
      for (int y = 0; y < BPW; ++y)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int r = r0 + wm + 16 * x + MF::drow(lane, rr), c = c0 + wn + 16 * y + lr;
          if (r < M && c < N) {
            T *pc = C + (size_t)c * ldc + r;
            T v = g.alpha * acc[x][y][rr];
            if (g.beta != T(0)) v += g.beta * *pc;
            *pc = v;
          }
        }
  }
}

/* Largest instance order whose LDS working set fits (panel width 16, the
 * smaller of the two the autotuner tries). */
template <typename T>
int max_order(int kind) {
  const long cap = 160 * 1024 / (long)sizeof(T);
  const long fixed = (kind == 0 ? 16L * 17 : 0L) + 2L * 16 * TS;
  const long per_row = (kind == 0 ? 0L : 16L) + MAXR;
  return (int)((cap - fixed) / per_row);
}

template <typename T>
hipError_t batch_gemm(int count, int ta, int tb, const int *order, const int *m, const int *n,
                      const int *k,
                      T alpha, const T *const *A, const int *lda, const T *const *B,
                      const int *ldb, T beta, T *const *C, const int *ldc,
                      const int *tile_start, int *counter, int variant, int grid,
                      hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if ((ta != KX_NOTRANS && ta != KX_TRANS) || (tb != KX_NOTRANS && tb != KX_TRANS) ||
      grid <= 0 || variant < 0 || variant > 2)
    return hipErrorInvalidValue;
  GemmBatchArgs<T> g;
  g.count = count; g.ta = ta; g.tb = tb; g.order = order; g.m = m; g.n = n; g.k = k;
  g.alpha = alpha;
  g.beta = beta; g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.tile_start = tile_start; g.counter = counter;
  if (variant == 0)
    hipLaunchKernelGGL(gemm_batched_kernel<T>, dim3(grid), dim3(NT), 0, s, g);
  else if (variant == 1)
    hipLaunchKernelGGL((gemm_batched_mfma_kernel<T, 1>), dim3(grid), dim3(64), 0, s, g);
  else
    hipLaunchKernelGGL((gemm_batched_mfma_kernel<T, 4>), dim3(grid), dim3(256), 0, s, g);
  return hipGetLastError();
}

template <typename T>
hipError_t chol_adjoint(int count, const int *order, const int *n, T *const *L, const int *lda,
                        int nrhs, T *const *Xs, T *const *Bb, const int *ldx, T *const *G,
                        const int *ldg, T *const *W, int *counter, int grid, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (nrhs < 0 || grid <= 0) return hipErrorInvalidValue;
  AdjArgs<T> g;
  g.count = count; g.nrhs = nrhs; g.order = order; g.n = n; g.lda = lda; g.ldx = ldx;
  g.ldg = ldg; g.L = L; g.X = Xs; g.Bb = Bb; g.G = G; g.W = W; g.counter = counter;
  hipLaunchKernelGGL(chol_adjoint_kernel<T>, dim3(grid), dim3(NT), 0, s, g);
  return hipGetLastError();
}

}  // namespace

extern "C" hipError_t atlmx_k_dbatch_chol_adjoint(int count, const int *order, const int *n,
                                                  double *const *L, const int *lda, int nrhs,
                                                  double *const *X, double *const *Bb,
                                                  const int *ldx, double *const *G,
                                                  const int *ldg, double *const *W, int *counter,
                                                  int grid, hipStream_t s) {
  return chol_adjoint<double>(count, order, n, L, lda, nrhs, X, Bb, ldx, G, ldg, W, counter,
                              grid, s);
}
extern "C" hipError_t atlmx_k_sbatch_chol_adjoint(int count, const int *order, const int *n,
                                                  float *const *L, const int *lda, int nrhs,
                                                  float *const *X, float *const *Bb,
                                                  const int *ldx, float *const *G,
                                                  const int *ldg, float *const *W, int *counter,
                                                  int grid, hipStream_t s) {
  return chol_adjoint<float>(count, order, n, L, lda, nrhs, X, Bb, ldx, G, ldg, W, counter,
                             grid, s);
}

extern "C" hipError_t atlmx_k_dbatch_factor(int kind, int nb, int count, const int *order,
                                            const int *n, double *const *A, const int *lda,
                                            int *const *ipiv, int nrhs, double *const *B,
                                            const int *ldb, int *info, int *counter, int maxn,
                                            int grid, hipStream_t s) {
  return batch_factor<double>(kind, nb, count, order, n, A, lda, ipiv, nrhs, B, ldb, info,
                              counter, maxn, grid, s);
}
extern "C" hipError_t atlmx_k_sbatch_factor(int kind, int nb, int count, const int *order,
                                            const int *n, float *const *A, const int *lda,
                                            int *const *ipiv, int nrhs, float *const *B,
                                            const int *ldb, int *info, int *counter, int maxn,
                                            int grid, hipStream_t s) {
  return batch_factor<float>(kind, nb, count, order, n, A, lda, ipiv, nrhs, B, ldb, info,
                             counter, maxn, grid, s);
}
extern "C" hipError_t atlmx_k_dbatch_gemm(int count, int ta, int tb, const int *order,
                                          const int *m, const int *n,
                                          const int *k, double alpha, const double *const *A,
                                          const int *lda, const double *const *B,
                                          const int *ldb, double beta, double *const *C,
                                          const int *ldc, const int *tile_start, int *counter,
                                          int variant, int grid, hipStream_t s) {
  return batch_gemm<double>(count, ta, tb, order, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                            tile_start, counter, variant, grid, s);
}
extern "C" hipError_t atlmx_k_sbatch_gemm(int count, int ta, int tb, const int *order,
                                          const int *m, const int *n,
                                          const int *k, float alpha, const float *const *A,
                                          const int *lda, const float *const *B, const int *ldb,
                                          float beta, float *const *C, const int *ldc,
                                          const int *tile_start, int *counter, int variant,
                                          int grid, hipStream_t s) {
  return batch_gemm<float>(count, ta, tb, order, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                           tile_start, counter, variant, grid, s);
}

/* kind 0 Cholesky, 1 LU (2, the transposed solve, as 1) */
extern "C" int atlmx_k_batch_max_order(int kind, int dbl) {
  const int k = kind == 0 ? 0 : 1;
  return dbl ? max_order<double>(k) : max_order<float>(k);
}

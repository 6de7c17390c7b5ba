/*
 * k_chol.hip -- Cholesky of one diagonal block (n <= 128) in registers, and
 * the solve of the panel below (right of) it.
 *
 * Replaces the leaves of ATLAS's recursive Cholesky: the closed-form 4x4,
 * 3x3, 2x2 kernels (src/lapack/ATL_potrfL.c:104-202, ATL_potrfU.c:98-197)
 * and the small-N recursion above them.  The whole block is held in the
 * registers of one 256-thread workgroup and factored right-looking, one
 * column at a time: rank-1 update of the trailing triangle with the
 * unscaled column over the pivot; the square roots and the column scaling
 * are applied once, when the block is written back.  A non-positive (or NaN) pivot stops the
 * factorization and records the 1-based global column in *info, as
 * ATL_potrfL.c:238-240 does; later kernels of the same Cholesky see the
 * nonzero info and do nothing.  Only the referenced triangle is read or
 * written.
 */
#include <hip/hip_runtime.h>
#include "kernels.h"

namespace {

constexpr int CNB = ATLMX_CHOL_NB;
constexpr int NT = 256;

template <typename T>
__device__ inline T dsqrt(T x);
template <>
__device__ inline double dsqrt<double>(double x) { return sqrt(x); }
template <>
__device__ inline float dsqrt<float>(float x) { return sqrtf(x); }

/* 1/x for the rank-1 multipliers: the hardware reciprocal refined by
 * Newton steps (two for double, one for float), within an ulp or so of
 * the rounded quotient and a short dependent chain on every column step. */
__device__ inline double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}
__device__ inline float rcp_nr(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}

/* The block lives in registers: thread (ti, tj) of a 16 x 16 grid holds the
 * 8 x 8 tile of rows ti*8.., columns tj*8.. (tiles above the diagonal stay
 * idle).  Step k: the owners of column k publish it as it stands (below the
 * pivot, zeros above) and the pivot d_k through LDS, and every tile right
 * of column k subtracts a(i,k) * a(j,k) / d_k in registers.  Column k stays
 * unscaled: L(i,k) = a(i,k) / sqrt(d_k) and L(k,k) = sqrt(d_k) are formed
 * for the whole block at write-back, so no square root or division by it
 * sits on the step-to-step chain.  One barrier per step: the column buffers
 * alternate with the parity of k, so a step's writes never meet the
 * previous step's reads, and every thread reads the same d_k, so a failed
 * pivot stops all of them at once.  The eight steps of a tile column are
 * unrolled, so every register index is fixed at compile time. */
template <typename T>
__global__ __launch_bounds__(NT) void potf2_kernel(int lower, int n, T *A, long lda,
                                                   int info_base, int *info) {
  if (*info != 0) return;
  __shared__ T col[2][CNB];
  __shared__ T piv[CNB];
  const int tid = threadIdx.x;
  const int ti = tid >> 4, tj = tid & 15;
  const bool own = ti >= tj;
  T a[8][8];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int i = ti * 8 + r, j = tj * 8 + c;
      a[r][c] = (own && i < n && j < n && i >= j)
                    ? (lower ? A[(size_t)j * lda + i] : A[(size_t)i * lda + j])
                    : T(0);
    }
  int done = n;
  const int nkt = (n + 7) / 8;
  for (int kt = 0; kt < nkt && done == n; ++kt) {
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) {
      const int k = kt * 8 + kc;
      if (k >= n) break;
      T *ck = col[k & 1];
      if (tj == kt && own) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const int i = ti * 8 + r;
          ck[i] = (i > k && i < n) ? a[r][kc] : T(0);
        }
        if (ti == kt) piv[k] = a[kc][kc];
      }
      __syncthreads();
      const T d = piv[k];
      if (!__builtin_amdgcn_readfirstlane((int)(d > T(0)))) {
        if (tid == 0) atomicCAS(info, 0, info_base + k + 1);
        done = k;
        break;
      }
      if (own && tj >= kt) {
        const T rd = rcp_nr(d);
        T x[8], y[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) x[r] = ck[ti * 8 + r];
#pragma unroll
        for (int c = 0; c < 8; ++c) y[c] = ck[tj * 8 + c] * rd;
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
          for (int c = 0; c < 8; ++c) a[r][c] -= x[r] * y[c];
      }
    }
  }
  /* write back the factored columns (the whole block when it succeeded; on
   * failure the columns before the failing one, as the reference leaves
   * them factored), each scaled by its pivot's square root */
  if (own) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int j = tj * 8 + c;
      if (j >= done) continue;
      const T sd = dsqrt(piv[j]), rs = T(1) / sd;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int i = ti * 8 + r;
        if (i < n && i >= j) {
          const T v = (i == j) ? sd : a[r][c] * rs;
          if (lower) A[(size_t)j * lda + i] = v;
          else A[(size_t)i * lda + j] = v;
        }
      }
    }
  }
}

/* The panel's triangular solve after the diagonal block (jb <= 128 rows)
 * is factored: every panel vector v (a row of A21 for Lower, a column of
 * A12 for Upper) becomes y with L y^T = v^T, L the block's lower factor
 * (Lower: L11; Upper: U11^T).  With L split into 64-row blocks, L = [L00 0;
 * L10 L11'] and the inverted diagonal blocks from atlmx_k_?tri_inv:
 *   y0 = v0 inv(L00)^T,   y1 = (v1 - y0 L10^T) inv(L11')^T
 * -- the two-block TRSM ATLAS's Cholesky makes (ATL_potrfL.c:236-246), as
 * three 64-wide products in one launch.  A workgroup takes 64 vectors,
 * each wavefront 16 of them through all three products on the matrix cores
 * (the MFMA's four k lane groups take k = 16 lk + t in step t); a product's
 * result is turned into the next one's operand layout through the
 * wavefront's own slice of LDS, so no workgroup barrier is needed. */
template <typename T> struct CMfma;
template <> struct CMfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  /* v_mfma_f64_16x16x4f64: lane l, element r holds D[(l>>4) + 4r][l&15] */
  __device__ static int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct CMfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  /* v_mfma_f32_16x16x4f32: lane l, element r holds D[4(l>>4) + r][l&15] */
  __device__ static int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename T>
__global__ __launch_bounds__(NT) void chol_panel_solve_kernel(int lower, int nr, int jb,
                                                              const T *__restrict__ D, long lda,
                                                              const T *__restrict__ inv, T *P,
                                                              const int *info) {
  typedef CMfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int B = 64;
  if (*info != 0) return;
  __shared__ T s_w[NT / 64][16][B + 1];           /* per wavefront: 16 vectors x 64 */
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int b0 = min(B, jb), b1 = jb - B;         /* block sizes (b1 <= 0: one block) */
  const int vec = blockIdx.x * 64 + wave * 16 + li;   /* this lane's A-operand vector */
  const bool vok = vec < nr;
  /* v(p, k): Lower P[k lda + p], Upper P[p lda + k] */
  auto vat = [&](int p, int k) -> T * { return lower ? P + (size_t)k * lda + p : P + (size_t)p * lda + k; };
  /* L(i, k) of the factored block: Lower D[k lda + i], Upper D[i lda + k] */
  auto lat = [&](int i, int k) -> T { return lower ? D[(size_t)k * lda + i] : D[(size_t)i * lda + k]; };
  /* inv(L_bb)(i, k): Lower inv_b[k 64 + i], Upper (inv(U_bb)^T) inv_b[i 64 + k] */
  auto iat = [&](int b, int i, int k) -> T {
    const T *ib = inv + (size_t)b * B * B;
    return lower ? ib[(size_t)k * B + i] : ib[(size_t)i * B + k];
  };
  T (*sw)[B + 1] = s_w[wave];

  /* y0 = v0 inv(L00)^T: A operand v0(vec, k), B operand W0(k, i) = inv(L00)(i, k) */
  acc_t acc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = acc_t{0, 0, 0, 0};
  /* every operand of a product is requested before its first MFMA */
  T av[16], wv[16][4];
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int k = 16 * lk + t;
    av[t] = (vok && k < b0) ? *vat(vec, k) : T(0);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 16 * c + li;
      wv[t][c] = (k < b0 && i < b0) ? iat(0, i, k) : T(0);
    }
  }
#pragma unroll
  for (int t = 0; t < 16; ++t)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = MF::mma(av[t], wv[t][c], acc[c]);
  /* y0 to LDS ([vector][i]) and to P */
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int vr = MF::drow(lane, r), i = 16 * c + li;
      const int p = blockIdx.x * 64 + wave * 16 + vr;
      sw[vr][i] = acc[c][r];
      if (p < nr && i < b0) *vat(p, i) = acc[c][r];
    }
  if (b1 <= 0) return;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  /* v1' = v1 - y0 L10^T: A operand y0(vec, k), B operand L10(i, k) = L(64 + i, k) */
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = acc_t{0, 0, 0, 0};
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int k = 16 * lk + t;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 16 * c + li;
      wv[t][c] = (i < b1) ? lat(B + i, k) : T(0);
    }
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const T a = sw[li][16 * lk + t];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = MF::mma(a, wv[t][c], acc[c]);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int vr = MF::drow(lane, r), i = 16 * c + li;
      const int p = blockIdx.x * 64 + wave * 16 + vr;
      const T v1 = (p < nr && i < b1) ? *vat(p, B + i) : T(0);
      sw[vr][i] = v1 - acc[c][r];
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  /* y1 = v1' inv(L11')^T */
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c] = acc_t{0, 0, 0, 0};
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int k = 16 * lk + t;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = 16 * c + li;
      wv[t][c] = (k < b1 && i < b1) ? iat(1, i, k) : T(0);
    }
  }
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const T a = sw[li][16 * lk + t];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] = MF::mma(a, wv[t][c], acc[c]);
  }
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int vr = MF::drow(lane, r), i = 16 * c + li;
      const int p = blockIdx.x * 64 + wave * 16 + vr;
      if (p < nr && i < b1) *vat(p, B + i) = acc[c][r];
    }
}

template <typename T>
hipError_t launch_chol_panel(int uplo, int nr, int jb, const T *D, long lda, const T *inv, T *P,
                             const int *info, hipStream_t s) {
  if (nr <= 0 || jb <= 0) return hipSuccess;
  if (jb > CNB || lda < jb || !D || !inv || !P || !info) return hipErrorInvalidValue;
  hipLaunchKernelGGL(chol_panel_solve_kernel<T>, dim3((unsigned)((nr + 63) / 64)), dim3(NT), 0,
                     s, uplo == KX_LOWER, nr, jb, D, lda, inv, P, info);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_potf2(int uplo, int n, T *A, long lda, int info_base, int *info,
                        hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n > CNB || lda < n || !A || !info) return hipErrorInvalidValue;
  hipLaunchKernelGGL(potf2_kernel<T>, dim3(1), dim3(NT), 0, s, uplo == KX_LOWER, n, A, lda,
                     info_base, info);
  return hipGetLastError();
}

}  // namespace

extern "C" hipError_t atlmx_k_dpotf2(int uplo, int n, double *A, long lda, int info_base,
                                     int *info, hipStream_t s) {
  return launch_potf2<double>(uplo, n, A, lda, info_base, info, s);
}
extern "C" hipError_t atlmx_k_spotf2(int uplo, int n, float *A, long lda, int info_base,
                                     int *info, hipStream_t s) {
  return launch_potf2<float>(uplo, n, A, lda, info_base, info, s);
}
extern "C" hipError_t atlmx_k_dchol_panel(int uplo, int nr, int jb, const double *D, long lda,
                                          const double *inv, double *P, const int *info,
                                          hipStream_t s) {
  return launch_chol_panel<double>(uplo, nr, jb, D, lda, inv, P, info, s);
}
extern "C" hipError_t atlmx_k_schol_panel(int uplo, int nr, int jb, const float *D, long lda,
                                          const float *inv, float *P, const int *info,
                                          hipStream_t s) {
  return launch_chol_panel<float>(uplo, nr, jb, D, lda, inv, P, info, s);
}

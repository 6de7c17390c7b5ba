/*
 * k_trsv.hip -- triangular solve with a few right-hand sides, in ONE
 * launch: the HBM-bound case of TRSM (getrs/potrs with NRHS = 1..4 reads
 * the n^2/2 factor once).  Both orientations keep one output row per lane,
 * so a block's result needs no cross-lane reduction.
 *
 * ATLAS runs this case through the same recursive TRSM as the many-RHS
 * case (src/blas/ulevel3/ATL_trsm.c:13-228).  Here each 64-row block of the
 * solution is owned by one workgroup.  Workgroups take block indices from
 * an atomic counter in solve order, so every block a workgroup waits for
 * was claimed by a workgroup that is already running (no deadlock, no
 * cooperative launch needed).  Block k streams its row (or, transposed, its
 * column) of 64x64 tiles of A, subtracting A_kj * x_j as each x_j is
 * published, then multiplies by the inverted diagonal block (from
 * atlmx_k_?tri_inv) and publishes x_k as tagged words (see below).
 * Published solution values move with device-coherent (sc1) accesses.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "kernels.h"

namespace {

constexpr int TB = ATLMX_TRI_NB;   /* 64 */
constexpr int NT = 256;
constexpr int MAXR = ATLMX_TRSV_MAXRHS;

__device__ inline double ldc(const double *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline float ldc(const float *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void stc(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void stc(float *p, float v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* Tagged publication: x_k's entries are published as 64-bit words
 * (1 << 32 | 32 bits of the value) in pub (zero on entry) -- one word for a
 * float, two (low and high halves) for a double -- so a consumer's lanes
 * poll the values themselves: one coherent round trip per block instead of
 * a flag followed by the data.  All four wavefronts of block k subtract
 * every earlier block except its predecessor (the block solved just before
 * it), each lane polling its own x entries, with no barrier per block; then
 * one wavefront (lane = row, the predecessor's tile and the inverted
 * diagonal block prefetched) waits for the predecessor, subtracts its
 * contribution, multiplies by the inverse and publishes: the chain from one
 * block to the next is one poll and two 64-term dot products per lane. */
template <typename T> struct Tag;
template <> struct Tag<float> {
  static constexpr int W = 1;
  __device__ static void put(unsigned long long *p, float v) {
    __hip_atomic_store(p, (1ull << 32) | (unsigned long long)__float_as_uint(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static float get(const unsigned long long *w) {
    return __uint_as_float((unsigned)w[0]);
  }
};
template <> struct Tag<double> {
  static constexpr int W = 2;
  __device__ static void put(unsigned long long *p, double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    __hip_atomic_store(p, (1ull << 32) | (b & 0xffffffffull), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, (1ull << 32) | (b >> 32), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  __device__ static double get(const unsigned long long *w) {
    return __longlong_as_double((long long)((w[1] << 32) | (w[0] & 0xffffffffull)));
  }
};

__device__ inline unsigned long long take_word(const unsigned long long *p) {
  unsigned long long w;
  while ((w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0ull)
    __builtin_amdgcn_s_sleep(1);
  return w;
}

template <typename T, int NR>
__global__ __launch_bounds__(NT) void trsv_tag_kernel(int n, int nrhs, int lower, int trans,
                                                      const T *__restrict__ A, long lda,
                                                      const T *__restrict__ inv, T *X, long ldx,
                                                      int *ctl, unsigned long long *pub) {
  constexpr int W = Tag<T>::W;
  /* FP32 keeps the predecessor's tile and the inverse in wavefront 0's
   * registers; FP64 would need 256 of them for those alone, so it keeps
   * them in LDS (rows padded to TB + 1) */
  constexpr bool REG = sizeof(T) == 4;
  constexpr int LT = REG ? 1 : TB * (TB + 1);
  __shared__ int s_idx;
  __shared__ T s_part[4][TB][NR];
  __shared__ T s_v[TB][NR];
  __shared__ T s_ap[LT], s_iv[LT];
  const int tid = threadIdx.x;
  const int nblk = (n + TB - 1) / TB;
  const bool fwd = (lower != 0) != (trans != 0);
  if (tid == 0) s_idx = atomicAdd(&ctl[0], 1);
  __syncthreads();
  const int order = s_idx;
  if (order >= nblk) return;
  const int k = fwd ? order : nblk - 1 - order;
  const int r0 = k * TB, rb = min(TB, n - r0);
  const int ln = tid & 63, q = tid >> 6;
  const int pj = order > 0 ? (fwd ? k - 1 : k + 1) : -1;   /* predecessor */

  /* the predecessor's tile row ln and row ln of op(inv_k): wavefront 0's
   * registers (FP32), or LDS filled by the whole workgroup (FP64) */
  T ap[REG ? TB : 1], iv[REG ? TB : 1];
  const int pc0 = pj * TB, pcb = pj >= 0 ? min(TB, n - pc0) : 0;
  const T *ik = inv + (size_t)k * TB * TB;
  if constexpr (REG) {
    if (q == 0) {
#pragma unroll
      for (int m = 0; m < TB; ++m) {
        const bool ok = ln < rb && m < pcb;
        ap[m] = !ok ? T(0)
                    : (trans ? A[(size_t)(r0 + ln) * lda + pc0 + m]
                             : A[(size_t)(pc0 + m) * lda + r0 + ln]);
      }
#pragma unroll
      for (int m = 0; m < TB; ++m)
        iv[m] = trans ? ik[(size_t)ln * TB + m] : ik[(size_t)m * TB + ln];
    }
  } else {
    for (int e = tid; e < TB * TB; e += NT) {
      const int i = e % TB, m = e / TB;           /* row i, column m */
      const bool ok = i < rb && m < pcb;
      s_ap[i * (TB + 1) + m] = !ok ? T(0)
                                   : (trans ? A[(size_t)(r0 + i) * lda + pc0 + m]
                                            : A[(size_t)(pc0 + m) * lda + r0 + i]);
      s_iv[i * (TB + 1) + m] = trans ? ik[(size_t)i * TB + m] : ik[(size_t)m * TB + i];
    }
    /* read only after the barrier below that ends the streaming phase */
  }

  /* every earlier block but the predecessor: thread (row ln, columns q*16..) */
  T acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = T(0);
  /* the tile of block s+1 is loaded while block s's entries are awaited */
  auto load_tile = [&](int s, T (&a)[16]) {
    const int j = fwd ? s : nblk - 1 - s;
    const int c0 = j * TB, cb = min(TB, n - c0);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int cc = q * 16 + t;
      const bool ok = ln < rb && cc < cb;
      a[t] = !ok ? T(0)
                 : (trans ? A[(size_t)(r0 + ln) * lda + c0 + cc]
                          : A[(size_t)(c0 + cc) * lda + r0 + ln]);
    }
  };
  T an[16];
  if (order > 1) load_tile(0, an);
  for (int s = 0; s + 1 < order; ++s) {
    const int j = fwd ? s : nblk - 1 - s;
    const int c0 = j * TB, cb = min(TB, n - c0);
    T a[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) a[t] = an[t];
    if (s + 2 < order) load_tile(s + 1, an);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      if (r >= nrhs) break;
      /* all 16 entries requested at once; only missing words are polled
       * again (unused entries read as the words of zero) */
      const unsigned long long *pp = pub + ((size_t)r * n + c0 + q * 16) * W;
      unsigned long long w[16 * W];
#pragma unroll
      for (int t = 0; t < 16 * W; ++t)
        w[t] = (q * 16 + t / W < cb)
                   ? __hip_atomic_load(pp + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                   : (1ull << 32);
      for (;;) {
        bool done = true;
#pragma unroll
        for (int t = 0; t < 16 * W; ++t) done = done && (w[t] != 0ull);
        if (done) break;
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int t = 0; t < 16 * W; ++t)
          if (w[t] == 0ull)
            w[t] = __hip_atomic_load(pp + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
#pragma unroll
      for (int t = 0; t < 16; ++t) acc[r] += a[t] * Tag<T>::get(&w[t * W]);
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) s_part[q][ln][r] = acc[r];
  __syncthreads();
  if (q != 0) return;

  /* wavefront 0 alone from here: c = b_k - partial sums */
  T c[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r)
    c[r] = (ln < rb && r < nrhs) ? X[(size_t)r * ldx + r0 + ln] -
                                       (s_part[0][ln][r] + s_part[1][ln][r] +
                                        s_part[2][ln][r] + s_part[3][ln][r])
                                 : T(0);
  if (pj >= 0) {
    const int pc0 = pj * TB, pcb = min(TB, n - pc0);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      T v = T(0);
      if (ln < pcb && r < nrhs) {
        unsigned long long w[W];
#pragma unroll
        for (int h = 0; h < W; ++h) w[h] = take_word(&pub[((size_t)r * n + pc0 + ln) * W + h]);
        v = Tag<T>::get(w);
      }
      s_v[ln][r] = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int m = 0; m < TB; ++m) {
      T am;
      if constexpr (REG) am = ap[m]; else am = s_ap[ln * (TB + 1) + m];
#pragma unroll
      for (int r = 0; r < NR; ++r) c[r] -= am * s_v[m][r];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) s_v[ln][r] = c[r];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  T y[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) y[r] = T(0);
#pragma unroll
  for (int m = 0; m < TB; ++m) {
    T im;
    if constexpr (REG) im = iv[m]; else im = s_iv[ln * (TB + 1) + m];
#pragma unroll
    for (int r = 0; r < NR; ++r) y[r] += im * s_v[m][r];
  }
  if (ln < rb) {
#pragma unroll
    for (int r = 0; r < NR; ++r)
      if (r < nrhs) {
        Tag<T>::put(&pub[((size_t)r * n + r0 + ln) * W], y[r]);
        X[(size_t)r * ldx + r0 + ln] = y[r];
      }
  }
}

template <typename T>
hipError_t launch_trsv_tag(int uplo, int trans, int n, int nrhs, const T *A, long lda,
                           const T *inv, T *X, long ldx, int *ctl, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (nrhs > MAXR || lda < n || ldx < n || !ctl || ((uintptr_t)ctl & 7))
    return hipErrorInvalidValue;
  const int nblk = (n + TB - 1) / TB;
  unsigned long long *pub = reinterpret_cast<unsigned long long *>(ctl + 2);
  if (nrhs == 1)
    hipLaunchKernelGGL((trsv_tag_kernel<T, 1>), dim3(nblk), dim3(NT), 0, s, n, nrhs,
                       uplo == KX_LOWER, trans == KX_TRANS, A, lda, inv, X, ldx, ctl, pub);
  else
    hipLaunchKernelGGL((trsv_tag_kernel<T, MAXR>), dim3(nblk), dim3(NT), 0, s, n, nrhs,
                       uplo == KX_LOWER, trans == KX_TRANS, A, lda, inv, X, ldx, ctl, pub);
  return hipGetLastError();
}

/* ---- many right-hand sides ----------------------------------------------
 * op(A) X = B with up to ATLMX_TRSM_TAG_MAXRHS columns in ONE launch, the
 * case getrs/potrs with tens of right-hand sides hit: a chain of 2 * nblk
 * small GEMM launches otherwise.  The work item is (R = 2 consecutive
 * 64-row blocks in solve order, group of RC = 16 columns); items are taken
 * from the atomic counter in solve order, group-minor, so every item an item
 * waits for was claimed by a running workgroup.  An item subtracts
 * op(A)_kj X_j from both its blocks for every earlier block j as X_j's
 * columns of its group are published, each 64x64 by 64x16 product on the
 * matrix cores (wavefront w owns rows 16w..16w+15; the MFMA's four k lane
 * groups take k = 16 lk + t in step t).  Then it solves its first block
 * (B_k minus the sum, times the inverted diagonal block), publishes it,
 * applies it to its second block inside the workgroup and solves that one:
 * two blocks per cross-workgroup hand-over, which is what bounds the chain.
 * Solutions are published as tagged words (as in the few-RHS solve) and
 * stored in place; a published block is staged in LDS once per workgroup
 * and read by all four wavefronts. */
constexpr int RC = 16;
constexpr int RB = 2;     /* 64-row blocks per work item */

template <typename T> struct TMfma;
template <> struct TMfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  /* v_mfma_f64_16x16x4f64: lane l, element r holds D[(l>>4) + 4r][l&15] */
  __device__ static int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct TMfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  /* v_mfma_f32_16x16x4f32: lane l, element r holds D[4(l>>4) + r][l&15] */
  __device__ static int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename T>
__global__ __launch_bounds__(NT) void trsm_tag_kernel(int n, int nrhs, int lower, int trans,
                                                      const T *__restrict__ A, long lda,
                                                      const T *__restrict__ inv, T *X, long ldx,
                                                      int *ctl, unsigned long long *pub) {
  typedef TMfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int W = Tag<T>::W;
  __shared__ int s_idx;
  __shared__ T s_x[2][RB][TB][RC + 1];            /* an earlier item's blocks of X, [row][col] */
  __shared__ T s_c[TB][RC + 1];                   /* B_k minus the partial sums */
  __shared__ T s_iv[RB][TB][TB + 1];              /* op(inv_k)(row, k) stored [k][row] */
  __shared__ T s_b[RB][TB][RC + 1];               /* the item's right-hand sides */
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nblk = (n + TB - 1) / TB;
  const int ngrp = (nrhs + RC - 1) / RC;
  const int nitem = (nblk + RB - 1) / RB;
  const bool fwd = (lower != 0) != (trans != 0);
  if (tid == 0) s_idx = atomicAdd(&ctl[0], 1);
  __syncthreads();
  const int ticket = s_idx;
  if (ticket >= nitem * ngrp) return;
  const int item = ticket / ngrp, g = ticket % ngrp;
  const int o0 = item * RB;                       /* solve-order position of block 0 */
  const int nbh = min(RB, nblk - o0);             /* blocks in this item */
  auto blk = [&](int o) { return fwd ? o : nblk - 1 - o; };
  int kb[RB], rbk[RB];
#pragma unroll
  for (int b = 0; b < RB; ++b) {
    kb[b] = blk(min(o0 + b, nblk - 1));
    rbk[b] = b < nbh ? min(TB, n - kb[b] * TB) : 0;
  }
  const int q0 = g * RC, qb = min(RC, nrhs - q0);
  const int li = lane & 15, lk = lane >> 4;
  const int arow = wave * 16 + li;                /* row this lane gives the A operand */

  /* the item's inverted diagonal blocks and right-hand sides, staged once
   * (off the chain from one item to the next) */
#pragma unroll
  for (int b = 0; b < RB; ++b) {
    const T *ik = inv + (size_t)kb[b] * TB * TB;
    for (int e = tid; e < TB * TB; e += NT) {
      const int r = e % TB, kk = e / TB;          /* op(inv)(r, kk) */
      s_iv[b][kk][r] = trans ? ik[(size_t)r * TB + kk] : ik[(size_t)kk * TB + r];
    }
    for (int e = tid; e < TB * RC; e += NT) {
      const int r = e % TB, cc = e / TB;
      s_b[b][r][cc] = (r < rbk[b] && cc < qb) ? X[(size_t)(q0 + cc) * ldx + kb[b] * TB + r] : T(0);
    }
  }
  __syncthreads();
  /* op(A)(64 kr + arow, 64 j + 16 lk + t), t = 0..15: the lane's elements
   * of tile (kr, j); zero outside the matrix or for a missing block */
  auto load_a = [&](int kr, int rb, int j, T (&a)[16]) {
    const int r0 = kr * TB, c0 = j * TB, cb = min(TB, n - c0);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int kk = 16 * lk + t;
      const bool ok = arow < rb && kk < cb;
      a[t] = !ok ? T(0)
                 : (trans ? A[(size_t)(r0 + arow) * lda + c0 + kk]
                          : A[(size_t)(c0 + kk) * lda + r0 + arow]);
    }
  };
  /* A(k1, k0), for the second block once the first is solved */
  T a10[16];
  load_a(kb[1], rbk[1], kb[0], a10);

  /* one accumulator per (block, earlier block of the pair): four
   * independent MFMA chains (a chain of dependent FP64 MFMAs is bound by
   * their latency, not their issue rate) */
  acc_t acc[RB][RB];
#pragma unroll
  for (int b = 0; b < RB; ++b)
#pragma unroll
    for (int jj = 0; jj < RB; ++jj) acc[b][jj] = acc_t{0, 0, 0, 0};
  const int pr = tid >> 2, pc = (tid & 3) * 4;    /* this thread's 4 entries of a block */
  /* stage blocks [j0, j1) of earlier item i in s_x[par] (the thread's
   * words all requested at once, only missing ones polled again; entries
   * outside a block read as zero), then the products with them */
  auto consume = [&](int i, int par, int j0, int j1, const T (&a)[RB][RB][16]) {
    unsigned long long w[RB][4 * W];
#pragma unroll
    for (int jj = 0; jj < RB; ++jj) {
      const int jbk = blk(i * RB + jj), cb = min(TB, n - jbk * TB);
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int h = 0; h < W; ++h) {
          const bool ok = jj >= j0 && jj < j1 && pr < cb && pc + e < qb;
          w[jj][e * W + h] =
              ok ? __hip_atomic_load(&pub[((size_t)(q0 + pc + e) * n + jbk * TB + pr) * W + h],
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                 : (1ull << 32);
        }
    }
    for (;;) {
      bool done = true;
#pragma unroll
      for (int jj = 0; jj < RB; ++jj)
#pragma unroll
        for (int t = 0; t < 4 * W; ++t) done = done && (w[jj][t] != 0ull);
      if (done) break;
      __builtin_amdgcn_s_sleep(1);
#pragma unroll
      for (int jj = 0; jj < RB; ++jj) {
        const int jbk = blk(i * RB + jj);
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int h = 0; h < W; ++h)
            if (w[jj][e * W + h] == 0ull)
              w[jj][e * W + h] = __hip_atomic_load(
                  &pub[((size_t)(q0 + pc + e) * n + jbk * TB + pr) * W + h], __ATOMIC_RELAXED,
                  __HIP_MEMORY_SCOPE_AGENT);
      }
    }
#pragma unroll
    for (int jj = 0; jj < RB; ++jj) {
      if (jj < j0 || jj >= j1) continue;
      const int cb = min(TB, n - blk(i * RB + jj) * TB);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        s_x[par][jj][pr][pc + e] = (pr < cb && pc + e < qb) ? Tag<T>::get(&w[jj][e * W]) : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < RB; ++jj) {
      if (jj < j0 || jj >= j1) continue;
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const T xv = s_x[par][jj][16 * lk + t][li];
#pragma unroll
        for (int b = 0; b < RB; ++b) acc[b][jj] = MF::mma(a[b][jj][t], xv, acc[b][jj]);
      }
    }
  };
  /* every earlier item; both of its blocks polled at once, except for the
   * item just before, whose first block is used while its second is still
   * being solved.  The four tiles' operands load meanwhile. */
  for (int i = 0; i < item; ++i) {
    const int par = i & 1;
    T a[RB][RB][16];
#pragma unroll
    for (int b = 0; b < RB; ++b)
#pragma unroll
      for (int jj = 0; jj < RB; ++jj) load_a(kb[b], rbk[b], blk(i * RB + jj), a[b][jj]);
    if (i + 1 < item) {
      consume(i, par, 0, RB, a);
    } else {
      consume(i, par, 0, 1, a);
      consume(i, par, 1, RB, a);
    }
  }
  /* block b: c = B_k - sum, X_k = op(inv_k) c, published; block 0's
   * solution then enters block 1's sum (s_x[0][0] is free: the last
   * products above read s_x before the barrier below).  Each 16-step
   * product runs as four chains of four. */
#pragma unroll
  for (int b = 0; b < RB; ++b) {
    if (b >= nbh) break;
    const int r0 = kb[b] * TB, rb = rbk[b];
    acc_t sum = acc[b][0];
#pragma unroll
    for (int jj = 1; jj < RB; ++jj) sum += acc[b][jj];
    if (b > 0) {
      __syncthreads();
      acc_t u[4] = {acc_t{0, 0, 0, 0}, acc_t{0, 0, 0, 0}, acc_t{0, 0, 0, 0}, acc_t{0, 0, 0, 0}};
#pragma unroll
      for (int t = 0; t < 16; ++t) u[t & 3] = MF::mma(a10[t], s_x[0][0][16 * lk + t][li], u[t & 3]);
      sum += (u[0] + u[1]) + (u[2] + u[3]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = wave * 16 + MF::drow(lane, r);
      s_c[row][li] = s_b[b][row][li] - sum[r];
    }
    __syncthreads();
    acc_t yv[4] = {acc_t{0, 0, 0, 0}, acc_t{0, 0, 0, 0}, acc_t{0, 0, 0, 0}, acc_t{0, 0, 0, 0}};
#pragma unroll
    for (int t = 0; t < 16; ++t)
      yv[t & 3] = MF::mma(s_iv[b][16 * lk + t][arow], s_c[16 * lk + t][li], yv[t & 3]);
    const acc_t y = (yv[0] + yv[1]) + (yv[2] + yv[3]);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = wave * 16 + MF::drow(lane, r);
      if (row < rb && li < qb) {
        X[(size_t)(q0 + li) * ldx + r0 + row] = y[r];
        Tag<T>::put(&pub[((size_t)(q0 + li) * n + r0 + row) * W], y[r]);
      }
      if (b + 1 < nbh) s_x[0][0][row][li] = (row < rb && li < qb) ? y[r] : T(0);
    }
  }
}

template <typename T>
hipError_t launch_trsm_tag(int uplo, int trans, int n, int nrhs, const T *A, long lda,
                           const T *inv, T *X, long ldx, int *ctl, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (nrhs > ATLMX_TRSM_TAG_MAXRHS || lda < n || ldx < n || !ctl || ((uintptr_t)ctl & 7))
    return hipErrorInvalidValue;
  const int nitem = (n + RB * TB - 1) / (RB * TB), ngrp = (nrhs + RC - 1) / RC;
  unsigned long long *pub = reinterpret_cast<unsigned long long *>(ctl + 2);
  hipLaunchKernelGGL(trsm_tag_kernel<T>, dim3((unsigned)(nitem * ngrp)), dim3(NT), 0, s, n, nrhs,
                     uplo == KX_LOWER, trans == KX_TRANS, A, lda, inv, X, ldx, ctl, pub);
  return hipGetLastError();
}

}  // namespace

extern "C" hipError_t atlmx_k_dtrsv(int uplo, int trans, int n, int nrhs, const double *A,
                                    long lda, const double *inv, double *X, long ldx, int *ctl,
                                    hipStream_t s) {
  return launch_trsv_tag<double>(uplo, trans, n, nrhs, A, lda, inv, X, ldx, ctl, s);
}
extern "C" hipError_t atlmx_k_strsv(int uplo, int trans, int n, int nrhs, const float *A,
                                    long lda, const float *inv, float *X, long ldx, int *ctl,
                                    hipStream_t s) {
  return launch_trsv_tag<float>(uplo, trans, n, nrhs, A, lda, inv, X, ldx, ctl, s);
}

/* block counter (and padding to 8 bytes), then the tagged words */
extern "C" size_t atlmx_k_trsv_ctl_ints(int n, int dbl) {
  return 2 + (size_t)2 * MAXR * n * (dbl ? 2 : 1);
}

extern "C" hipError_t atlmx_k_dtrsm_tag(int uplo, int trans, int n, int nrhs, const double *A,
                                        long lda, const double *inv, double *X, long ldx,
                                        int *ctl, hipStream_t s) {
  return launch_trsm_tag<double>(uplo, trans, n, nrhs, A, lda, inv, X, ldx, ctl, s);
}
extern "C" hipError_t atlmx_k_strsm_tag(int uplo, int trans, int n, int nrhs, const float *A,
                                        long lda, const float *inv, float *X, long ldx,
                                        int *ctl, hipStream_t s) {
  return launch_trsm_tag<float>(uplo, trans, n, nrhs, A, lda, inv, X, ldx, ctl, s);
}

/* block counter (and padding to 8 bytes), then the tagged words of every
 * right-hand side */
extern "C" size_t atlmx_k_trsm_tag_ctl_ints(int n, int nrhs, int dbl) {
  return 2 + (size_t)2 * nrhs * n * (dbl ? 2 : 1);
}

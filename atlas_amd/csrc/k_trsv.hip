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
 * atlmx_k_?tri_inv) and publishes x_k with a flag.  Published solution
 * values and flags move with device-coherent (sc1) accesses.
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

/* op(A) lower (forward) when lower != trans.  X (n x nrhs, ldx) holds B on
 * entry and the solution on exit.  ctl[0] = block counter, ctl[1 + k] =
 * flag of block k (all zero on entry).  inv: inverted diagonal blocks. */
template <typename T>
__global__ __launch_bounds__(NT) void trsv_kernel(int n, int nrhs, int lower, int trans,
                                                  const T *__restrict__ A, long lda,
                                                  const T *__restrict__ inv, T *X, long ldx,
                                                  int *ctl) {
  __shared__ int s_idx;
  __shared__ T s_part[4][TB][MAXR];
  __shared__ T s_acc[TB][MAXR];
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

  if (!trans) {
    /* thread (row i = ln, column group q): coalesced down the columns */
    T acc[MAXR];
#pragma unroll
    for (int r = 0; r < MAXR; ++r) acc[r] = T(0);
    for (int s = 0; s < order; ++s) {
      const int j = fwd ? s : nblk - 1 - s;
      const int c0 = j * TB, cb = min(TB, n - c0);
      /* the tile of A does not depend on x_j: load it before waiting */
      T a[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int cc = q * 16 + t;
        a[t] = (ln < rb && cc < cb) ? A[(size_t)(c0 + cc) * lda + (r0 + ln)] : T(0);
      }
      if (tid == 0)
        while (__hip_atomic_load(&ctl[1 + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
          __builtin_amdgcn_s_sleep(1);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int cc = q * 16 + t;
        if (cc < cb) {
#pragma unroll
          for (int r = 0; r < MAXR; ++r)
            if (r < nrhs) acc[r] += a[t] * ldc(&X[(size_t)r * ldx + c0 + cc]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < MAXR; ++r) s_part[q][ln][r] = acc[r];
  } else {
    /* thread (row i = ln, column group q): op(A)[r0+i][c0+cc] = A[c0+cc][r0+i]
     * is stored contiguously in cc, so each lane reads its own 16-element
     * (64- or 128-byte) segment and keeps its row's sums -- no cross-lane
     * reduction on the chain from one block to the next */
    T acc[MAXR];
#pragma unroll
    for (int r = 0; r < MAXR; ++r) acc[r] = T(0);
    for (int s = 0; s < order; ++s) {
      const int j = fwd ? s : nblk - 1 - s;
      const int c0 = j * TB, cb = min(TB, n - c0);
      T a[16];
      const T *arow = A + (size_t)(r0 + ln) * lda + c0 + q * 16;
#pragma unroll
      for (int t = 0; t < 16; ++t)
        a[t] = (ln < rb && q * 16 + t < cb) ? arow[t] : T(0);
      if (tid == 0)
        while (__hip_atomic_load(&ctl[1 + j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
          __builtin_amdgcn_s_sleep(1);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int cc = q * 16 + t;
        if (cc < cb) {
#pragma unroll
          for (int r = 0; r < MAXR; ++r)
            if (r < nrhs) acc[r] += a[t] * ldc(&X[(size_t)r * ldx + c0 + cc]);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < MAXR; ++r) s_part[q][ln][r] = acc[r];
  }
  __syncthreads();
  if (q == 0) {
#pragma unroll
    for (int r = 0; r < MAXR; ++r)
      s_acc[ln][r] = (ln < rb && r < nrhs)
                         ? ldc(&X[(size_t)r * ldx + r0 + ln]) -
                               (s_part[0][ln][r] + s_part[1][ln][r] + s_part[2][ln][r] +
                                s_part[3][ln][r])
                         : T(0);
  }
  __syncthreads();
  /* x_k = op(inv_k) * acc: thread (i, q) sums its quarter of row i */
  const T *iv = inv + (size_t)k * TB * TB;
  T y[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; ++r) y[r] = T(0);
  if (ln < rb) {
    for (int t = 0; t < 16; ++t) {
      const int cc = q * 16 + t;
      const T v = trans ? iv[(size_t)ln * TB + cc] : iv[(size_t)cc * TB + ln];
#pragma unroll
      for (int r = 0; r < MAXR; ++r) y[r] += v * s_acc[cc][r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < MAXR; ++r) s_part[q][ln][r] = y[r];
  __syncthreads();
  if (q == 0 && ln < rb) {
#pragma unroll
    for (int r = 0; r < MAXR; ++r)
      if (r < nrhs)
        stc(&X[(size_t)r * ldx + r0 + ln],
            s_part[0][ln][r] + s_part[1][ln][r] + s_part[2][ln][r] + s_part[3][ln][r]);
  }
  /* x_k must reach the coherence point before the flag: wait for the
   * stores' acknowledgements (vmcnt = 0), then the barrier orders the flag */
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (tid == 0) __hip_atomic_store(&ctl[1 + k], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* FP32 variant with tagged publication: x_k's entries are published as
 * 64-bit words (1 << 32 | bits(x)) in pub (zero on entry), so a consumer's
 * lanes poll the values themselves -- one coherent round trip per block
 * instead of a flag followed by the data.  All four wavefronts of block k
 * subtract every earlier block except its predecessor (the block solved
 * just before it), each lane polling its own x entries, with no barrier per
 * block; then one wavefront (lane = row, the predecessor's tile and the
 * inverted diagonal block prefetched in registers) waits for the
 * predecessor, subtracts its contribution, multiplies by the inverse and
 * publishes: the chain from one block to the next is one poll and two
 * 64-term dot products per lane. */
__device__ inline float take(const unsigned long long *p) {
  unsigned long long w;
  while ((w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0ull)
    __builtin_amdgcn_s_sleep(1);
  return __uint_as_float((unsigned)w);
}

__global__ __launch_bounds__(NT) void strsv_tag_kernel(int n, int nrhs, int lower, int trans,
                                                       const float *__restrict__ A, long lda,
                                                       const float *__restrict__ inv, float *X,
                                                       long ldx, int *ctl,
                                                       unsigned long long *pub) {
  __shared__ int s_idx;
  __shared__ float s_part[4][TB][MAXR];
  __shared__ float s_v[TB][MAXR];
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

  /* wavefront 0: the predecessor's tile row ln and row ln of op(inv_k) */
  float ap[TB], iv[TB];
  if (q == 0) {
    const int pc0 = pj * TB, pcb = pj >= 0 ? min(TB, n - pc0) : 0;
#pragma unroll
    for (int m = 0; m < TB; ++m) {
      const bool ok = ln < rb && m < pcb;
      ap[m] = !ok ? 0.0f
                  : (trans ? A[(size_t)(r0 + ln) * lda + pc0 + m]
                           : A[(size_t)(pc0 + m) * lda + r0 + ln]);
    }
    const float *ik = inv + (size_t)k * TB * TB;
#pragma unroll
    for (int m = 0; m < TB; ++m)
      iv[m] = trans ? ik[(size_t)ln * TB + m] : ik[(size_t)m * TB + ln];
  }

  /* every earlier block but the predecessor: thread (row ln, columns q*16..) */
  float acc[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; ++r) acc[r] = 0.0f;
  /* the tile of block s+1 is loaded while block s's entries are awaited */
  auto load_tile = [&](int s, float (&a)[16]) {
    const int j = fwd ? s : nblk - 1 - s;
    const int c0 = j * TB, cb = min(TB, n - c0);
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int cc = q * 16 + t;
      const bool ok = ln < rb && cc < cb;
      a[t] = !ok ? 0.0f
                 : (trans ? A[(size_t)(r0 + ln) * lda + c0 + cc]
                          : A[(size_t)(c0 + cc) * lda + r0 + ln]);
    }
  };
  float an[16];
  if (order > 1) load_tile(0, an);
  for (int s = 0; s + 1 < order; ++s) {
    const int j = fwd ? s : nblk - 1 - s;
    const int c0 = j * TB, cb = min(TB, n - c0);
    float a[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) a[t] = an[t];
    if (s + 2 < order) load_tile(s + 1, an);
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
      if (r >= nrhs) break;
      /* all 16 entries requested at once; only missing ones are polled
       * again (unused entries read as the word of 0.0f) */
      const unsigned long long *pp = pub + (size_t)r * n + c0 + q * 16;
      unsigned long long w[16];
#pragma unroll
      for (int t = 0; t < 16; ++t)
        w[t] = (q * 16 + t < cb)
                   ? __hip_atomic_load(pp + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                   : (1ull << 32);
      for (;;) {
        bool done = true;
#pragma unroll
        for (int t = 0; t < 16; ++t) done = done && (w[t] != 0ull);
        if (done) break;
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int t = 0; t < 16; ++t)
          if (w[t] == 0ull)
            w[t] = __hip_atomic_load(pp + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
#pragma unroll
      for (int t = 0; t < 16; ++t) acc[r] += a[t] * __uint_as_float((unsigned)w[t]);
    }
  }
#pragma unroll
  for (int r = 0; r < MAXR; ++r) s_part[q][ln][r] = acc[r];
  __syncthreads();
  if (q != 0) return;

  /* wavefront 0 alone from here: c = b_k - partial sums */
  float c[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; ++r)
    c[r] = (ln < rb && r < nrhs) ? X[(size_t)r * ldx + r0 + ln] -
                                       (s_part[0][ln][r] + s_part[1][ln][r] +
                                        s_part[2][ln][r] + s_part[3][ln][r])
                                 : 0.0f;
  if (pj >= 0) {
    const int pc0 = pj * TB, pcb = min(TB, n - pc0);
#pragma unroll
    for (int r = 0; r < MAXR; ++r)
      s_v[ln][r] = (ln < pcb && r < nrhs) ? take(&pub[(size_t)r * n + pc0 + ln]) : 0.0f;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int m = 0; m < TB; ++m)
#pragma unroll
      for (int r = 0; r < MAXR; ++r) c[r] -= ap[m] * s_v[m][r];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
#pragma unroll
  for (int r = 0; r < MAXR; ++r) s_v[ln][r] = c[r];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  float y[MAXR];
#pragma unroll
  for (int r = 0; r < MAXR; ++r) y[r] = 0.0f;
#pragma unroll
  for (int m = 0; m < TB; ++m)
#pragma unroll
    for (int r = 0; r < MAXR; ++r) y[r] += iv[m] * s_v[m][r];
  if (ln < rb) {
#pragma unroll
    for (int r = 0; r < MAXR; ++r)
      if (r < nrhs) {
        __hip_atomic_store(&pub[(size_t)r * n + r0 + ln],
                           (1ull << 32) | (unsigned long long)__float_as_uint(y[r]),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        X[(size_t)r * ldx + r0 + ln] = y[r];
      }
  }
}

template <typename T>
hipError_t launch_trsv(int uplo, int trans, int n, int nrhs, const T *A, long lda,
                       const T *inv, T *X, long ldx, int *ctl, hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (nrhs > MAXR || lda < n || ldx < n || !ctl) return hipErrorInvalidValue;
  const int nblk = (n + TB - 1) / TB;
  hipLaunchKernelGGL(trsv_kernel<T>, dim3(nblk), dim3(NT), 0, s, n, nrhs, uplo == KX_LOWER,
                     trans == KX_TRANS, A, lda, inv, X, ldx, ctl);
  return hipGetLastError();
}

}  // namespace

extern "C" hipError_t atlmx_k_dtrsv(int uplo, int trans, int n, int nrhs, const double *A,
                                    long lda, const double *inv, double *X, long ldx, int *ctl,
                                    hipStream_t s) {
  return launch_trsv<double>(uplo, trans, n, nrhs, A, lda, inv, X, ldx, ctl, s);
}
extern "C" hipError_t atlmx_k_strsv(int uplo, int trans, int n, int nrhs, const float *A,
                                    long lda, const float *inv, float *X, long ldx, int *ctl,
                                    hipStream_t s) {
  if (n <= 0 || nrhs <= 0) return hipSuccess;
  if (nrhs > MAXR || lda < n || ldx < n || !ctl || ((uintptr_t)ctl & 7))
    return hipErrorInvalidValue;
  const int nblk = (n + TB - 1) / TB;
  unsigned long long *pub = reinterpret_cast<unsigned long long *>(ctl + 2);
  hipLaunchKernelGGL(strsv_tag_kernel, dim3(nblk), dim3(NT), 0, s, n, nrhs, uplo == KX_LOWER,
                     trans == KX_TRANS, A, lda, inv, X, ldx, ctl, pub);
  return hipGetLastError();
}

extern "C" size_t atlmx_k_trsv_ctl_ints(int n, int dbl) {
  return dbl ? (size_t)(n + TB - 1) / TB + 1 : 2 + (size_t)2 * MAXR * n;
}

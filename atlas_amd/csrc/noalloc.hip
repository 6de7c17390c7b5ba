/*
 * noalloc.hip -- atlmx_dgetrs_noalloc: the FP64 LU solve on device pointers
 * with no allocation of any kind, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It has no counterpart in ATLAS.
 *
 * atlmx_dgetrs needs two scratch buffers: the inverted 64 x 64 diagonal
 * blocks of the triangular factor (nblk * 64 * 64 doubles) and the chained
 * solve's block counter and published solution words (for up to 64
 * right-hand sides; more run 64 per launch, with the same result).  It
 * takes them from the calling thread's grow-only workspaces.  Here they are
 * slices of one file-scope device array, `scratch`, sized for N <=
 * ATLMX_NOALLOC_MAX_N.  Each GPU holds its own copy of the array until the
 * process ends; its address on a device comes from hipGetSymbolAddress, made
 * once per device with that device current, and only that address is ever
 * handed to a kernel or a fill.  The kernels are those of atlmx_dgetrs, in
 * the same order, so the two give bit-identical results.
 *
 * Every call on a GPU shares that GPU's array: the solve's words are cleared by a
 * fill kernel at the start of each solve and the inverses are rewritten in
 * full, so one call never sees the last call's contents, but calls that
 * overlap on one GPU (two threads, or two streams) race on it.
 */
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <stdio.h>
#include <string.h>

extern "C" {
#include "rt.h"
#include "drivers.h"
#include "../../include/atlmx.h"
}

namespace {

constexpr int NB = ATLMX_TRI_NB;                                     /* 64 */
constexpr size_t INV_BYTES =
    (size_t)((ATLMX_NOALLOC_MAX_N + NB - 1) / NB) * NB * NB * sizeof(double);
/* the solve's block counter and tagged words: atlmx_k_trsv_ctl_ints(N, 1)
 * ints (2 + 4 * ATLMX_TRSV_MAXRHS * N) for a few right-hand sides,
 * atlmx_k_trsm_tag_ctl_ints(N, ATLMX_TRSM_TAG_CHUNK, 1) ints
 * (2 + 4 * ATLMX_TRSM_TAG_CHUNK * N) for up to ATLMX_TRSM_TAG_MAXRHS, which
 * are then solved that many columns per launch */
constexpr size_t CTL_RHS =
    ATLMX_TRSM_TAG_CHUNK > ATLMX_TRSV_MAXRHS ? ATLMX_TRSM_TAG_CHUNK : ATLMX_TRSV_MAXRHS;
constexpr size_t CTL_BYTES =
    ((size_t)(2 + 4 * CTL_RHS * (size_t)ATLMX_NOALLOC_MAX_N) * sizeof(int) + 255) &
    ~(size_t)255;
constexpr size_t SCRATCH_BYTES = INV_BYTES + CTL_BYTES;
constexpr int MAX_DEV = 64;

static_assert(INV_BYTES % 256 == 0, "the solve's slice must stay 256-byte aligned");

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

void *scratch_addr[MAX_DEV];           /* device address of `scratch`, per device */
pthread_mutex_t scratch_mu = PTHREAD_MUTEX_INITIALIZER;

/* The array's address on device `dev` (the current device), looked up once. */
void *scratch_on(int dev, const char *rout) {
  void *p = __atomic_load_n(&scratch_addr[dev], __ATOMIC_ACQUIRE);
  if (p) return p;
  pthread_mutex_lock(&scratch_mu);
  p = scratch_addr[dev];
  if (!p) {
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      pthread_mutex_unlock(&scratch_mu);
      rt_fail(rout, "hipGetSymbolAddress(scratch)", e != hipSuccess ? e : hipErrorUnknown);
    }
    __atomic_store_n(&scratch_addr[dev], p, __ATOMIC_RELEASE);
  }
  pthread_mutex_unlock(&scratch_mu);
  return p;
}

}  // namespace

/* Called by atlmx_init: resolve the array on the current device up front. */
extern "C" void atlmx_noalloc_setup(void) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEV) return;
  (void)scratch_on(dev, "atlmx_init");
}

extern "C" size_t atlmx_noalloc_scratch_bytes(void) { return SCRATCH_BYTES; }

extern "C" int atlmx_dgetrs_noalloc(int Trans, int N, int NRHS, const double *A, int lda,
                                    const int *d_ipiv, double *B, int ldb) {
  static const char rout[] = "atlmx_dgetrs_noalloc";
  int bad = 0;
  if (Trans != KX_NOTRANS && Trans != KX_TRANS && Trans != 113) bad = -1;
  else if (N < 0 || N > ATLMX_NOALLOC_MAX_N) bad = -2;
  else if (NRHS < 0) bad = -3;
  else if (lda < (N > 1 ? N : 1)) bad = -5;
  else if (ldb < (N > 1 ? N : 1)) bad = -8;
  if (bad) {
    if (bad == -2)
      fprintf(stderr, "%s: N = %d is outside [0, %d], the order its static scratch "
              "array holds; nothing was done\n", rout, N, ATLMX_NOALLOC_MAX_N);
    else
      fprintf(stderr, "%s: parameter %d was incorrect; nothing was done\n", rout, -bad);
    return bad;
  }
  if (N == 0 || NRHS == 0) return 0;
  int dev = 0;
  RT_CHECK(rout, hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEV) rt_fail(rout, "device ordinal", hipErrorInvalidDevice);
  unsigned char *base = static_cast<unsigned char *>(scratch_on(dev, rout));
  /* a context on the stack: the calling thread's stream if it has one, no
   * workspaces (both scratch buffers are passed in) and no profiler events */
  const atlmx_ctx *tc = rt_ctx_peek();
  atlmx_ctx lc;
  memset(&lc, 0, sizeof lc);
  lc.dev = dev;
  lc.stream = tc ? tc->stream : 0;
  lc.own = lc.stream;
  lc.noprof = 1;
  dgetrs_dev_ws(&lc, Trans == KX_NOTRANS ? KX_NOTRANS : KX_TRANS, N, NRHS, A, lda, d_ipiv, B,
                ldb, reinterpret_cast<double *>(base), reinterpret_cast<int *>(base + INV_BYTES),
                rout);
  return 0;
}

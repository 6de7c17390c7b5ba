/*
 * runtime.c -- host runtime of libatlas_mi355x.
 *
 * ATLAS has no device and no handles: each call mallocs its workspace and
 * frees it (src/blas/ammm/ATL_ammmKMNK.c:53-65,105), and when a malloc
 * fails it recurses on a smaller problem (ATL_ammm.c:293-356).  Here each
 * host thread gets one HIP stream and a set of grow-only device workspaces
 * that persist across calls (hipMalloc is far too slow to do per call), so
 * the library is callable from any host thread with no visible global
 * mutable state.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "rt.h"
#include "../../include/cblas.h"
#include "../../include/atlmx.h"

static __thread atlmx_ctx *tls_ctx = NULL;

/* Every context and profiler buffer ever created, so that process exit can
 * release them while the HIP runtime is still up (an atexit handler
 * registered after the runtime's own initialisation runs before its
 * teardown): streams and events destroyed, workspaces freed, after one
 * device synchronisation. */
#define RT_REG_MAX 1024
static pthread_mutex_t reg_mu = PTHREAD_MUTEX_INITIALIZER;
static atlmx_ctx *reg_ctx[RT_REG_MAX];
static int reg_nctx = 0;
static void *reg_prof[RT_REG_MAX];
static int reg_nprof = 0;
static int reg_hooked = 0;
static void rt_release_prof(void *buf);

static void rt_teardown(void)
{
  int i, b, s;
  pthread_mutex_lock(&reg_mu);
  if (reg_nctx > 0 || reg_nprof > 0) (void)hipDeviceSynchronize();
  for (i = 0; i < reg_nctx; ++i) {
    atlmx_ctx *c = reg_ctx[i];
    for (b = 0; b < 2; ++b)
      for (s = 0; s < WS_NSLOTS; ++s)
        if (c->ws[b][s]) { (void)hipFree(c->ws[b][s]); c->ws[b][s] = NULL; c->ws_bytes[b][s] = 0; }
    if (c->aux) {
      for (s = 0; s < 3; ++s)
        if (c->ev[s]) { (void)hipEventDestroy(c->ev[s]); c->ev[s] = NULL; }
      (void)hipStreamDestroy(c->aux);
      c->aux = NULL;
    }
    c->panel_buf = NULL;
  }
  for (i = 0; i < reg_nprof; ++i) rt_release_prof(reg_prof[i]);
  reg_nctx = reg_nprof = 0;
  pthread_mutex_unlock(&reg_mu);
}

static void rt_register(atlmx_ctx *c, void *prof)
{
  pthread_mutex_lock(&reg_mu);
  if (!reg_hooked) { atexit(rt_teardown); reg_hooked = 1; }
  if (c && reg_nctx < RT_REG_MAX) reg_ctx[reg_nctx++] = c;
  if (prof && reg_nprof < RT_REG_MAX) reg_prof[reg_nprof++] = prof;
  pthread_mutex_unlock(&reg_mu);
}

void rt_fail(const char *rout, const char *what, hipError_t e)
{
  cblas_xerbla(0, rout, "GPU failure in %s: %s (%d)\n", what, hipGetErrorString(e), (int)e);
  abort();   /* a user cblas_xerbla that returns must not let us continue */
}

atlmx_ctx *rt_ctx(void)
{
  atlmx_ctx *c = tls_ctx;
  hipError_t e;
  if (c) return c;
  c = (atlmx_ctx *)calloc(1, sizeof(*c));
  if (!c) { fprintf(stderr, "atlas_mi355x: out of host memory\n"); exit(-1); }
  e = hipGetDevice(&c->dev);
  if (e != hipSuccess) rt_fail("atlas_mi355x", "hipGetDevice", e);
  /* Default: the legacy default stream, so device operands written by the
   * caller on that stream (or any blocking stream) are ready before the
   * library reads them. */
  c->own = 0;
  c->stream = 0;
  tls_ctx = c;
  rt_register(c, NULL);
  return c;
}

int rt_is_device(const void *p)
{
  hipPointerAttribute_t a;
  hipError_t e;
  if (!p) return 0;
  memset(&a, 0, sizeof(a));
  e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) { (void)hipGetLastError(); return 0; }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
         a.type == hipMemoryTypeUnified || a.isManaged;
}

hipStream_t rt_aux(atlmx_ctx *c, const char *rout)
{
  if (!c->aux) {
    RT_CHECK(rout, hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
    RT_CHECK(rout, hipEventCreateWithFlags(&c->ev[0], hipEventDisableTiming));
    RT_CHECK(rout, hipEventCreateWithFlags(&c->ev[1], hipEventDisableTiming));
    RT_CHECK(rout, hipEventCreateWithFlags(&c->ev[2], hipEventDisableTiming));
    RT_CHECK(rout, hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount,
                                         c->dev));
  }
  return c->aux;
}

void rt_sync(atlmx_ctx *c, const char *rout)
{
  RT_CHECK(rout, hipStreamSynchronize(c->stream));
}

void *rt_ws(atlmx_ctx *c, int slot, size_t bytes, const char *rout)
{
  const int b = c->bank;
  if (bytes == 0) bytes = 256;
  if (c->ws_bytes[b][slot] < bytes) {
    size_t nb = bytes + bytes / 8;
    nb = (nb + 4095) & ~(size_t)4095;
    if (c->ws[b][slot]) {
      /* the old buffer may still be in use by queued work */
      RT_CHECK(rout, hipStreamSynchronize(c->stream));
      RT_CHECK(rout, hipFree(c->ws[b][slot]));
      c->ws[b][slot] = NULL;
      c->ws_bytes[b][slot] = 0;
    }
    RT_CHECK(rout, hipMalloc(&c->ws[b][slot], nb));
    c->ws_bytes[b][slot] = nb;
  }
  return c->ws[b][slot];
}

void rt_on_aux(atlmx_ctx *c, hipStream_t main, int on)
{
  c->stream = on ? c->aux : main;
  c->bank = on ? 1 : 0;
}

void *rt_stage_in(atlmx_ctx *c, int slot, const void *h, int rows, int cols, long ld,
                  size_t esz, int copy_in, long *dld, const char *rout)
{
  long pad;
  void *d;
  if (rt_is_device(h)) { *dld = ld; return (void *)h; }
  pad = ((long)(rows > 0 ? rows : 1) + 15) & ~15L;
  d = rt_ws(c, slot, (size_t)pad * (size_t)(cols > 0 ? cols : 1) * esz, rout);
  *dld = pad;
  if (copy_in && rows > 0 && cols > 0)
    RT_CHECK(rout, hipMemcpy2DAsync(d, (size_t)pad * esz, h, (size_t)ld * esz,
                                    (size_t)rows * esz, (size_t)cols, hipMemcpyHostToDevice,
                                    c->stream));
  return d;
}

void rt_stage_out(atlmx_ctx *c, void *h, const void *d, int rows, int cols, long ld,
                  long dld, size_t esz, const char *rout)
{
  if (h == d || rows <= 0 || cols <= 0) return;
  RT_CHECK(rout, hipMemcpy2DAsync(h, (size_t)ld * esz, d, (size_t)dld * esz,
                                  (size_t)rows * esz, (size_t)cols, hipMemcpyDeviceToHost,
                                  c->stream));
}

/* ---------------------------------------------------------------------- */
/* profiler                                                                 */
/* ---------------------------------------------------------------------- */
#define PROF_MAX 65536
typedef struct { hipEvent_t a, b; int cls; double work, bytes; } prof_rec;
static __thread prof_rec *prof_buf = NULL;
static __thread int prof_n = 0;
static __thread int prof_enabled = 0;   /* bit mask of recorded classes */

static void rt_release_prof(void *buf)
{
  prof_rec *r = (prof_rec *)buf;
  int i;
  for (i = 0; i < PROF_MAX; ++i) {
    if (r[i].a) { (void)hipEventDestroy(r[i].a); r[i].a = NULL; }
    if (r[i].b) { (void)hipEventDestroy(r[i].b); r[i].b = NULL; }
  }
}

int rt_prof_on(void) { return prof_enabled; }

int rt_prof_begin(atlmx_ctx *c, int cls, double work)
{
  return rt_prof_begin2(c, cls, work, 0.0);
}

int rt_prof_begin2(atlmx_ctx *c, int cls, double work, double bytes)
{
  prof_rec *r;
  if (!(prof_enabled & (1 << cls)) || prof_n >= PROF_MAX || c->noprof) return -1;
  r = &prof_buf[prof_n];
  if (!r->a) {
    RT_CHECK("atlmx_prof", hipEventCreate(&r->a));
    RT_CHECK("atlmx_prof", hipEventCreate(&r->b));
  }
  r->cls = cls;
  r->work = work;
  r->bytes = bytes;
  RT_CHECK("atlmx_prof", hipEventRecord(r->a, c->stream));
  return prof_n++;
}

void rt_prof_end(atlmx_ctx *c, int id)
{
  if (id < 0) return;
  RT_CHECK("atlmx_prof", hipEventRecord(prof_buf[id].b, c->stream));
}

/* ---- public runtime entry points (declared in include/atlmx.h) ---- */
void atlmx_noalloc_setup(void);   /* noalloc.hip */

int atlmx_init(void)
{
  (void)rt_ctx();
  atlmx_noalloc_setup();
  return 0;
}

atlmx_ctx *rt_ctx_peek(void) { return tls_ctx; }

/* NULL selects the legacy default stream. */
void atlmx_set_stream(void *stream)
{
  atlmx_ctx *c = rt_ctx();
  c->stream = (hipStream_t)stream;
}

void *atlmx_get_stream(void)
{
  return (void *)rt_ctx()->stream;
}

void atlmx_synchronize(void)
{
  atlmx_ctx *c = rt_ctx();
  rt_sync(c, "atlmx_synchronize");
}

/* mask: bit c records launch class c (0 = off; 1 = GEMM only; -1 = all) */
void atlmx_prof_enable(int mask)
{
  if (mask && !prof_buf) {
    prof_buf = (prof_rec *)calloc(PROF_MAX, sizeof(prof_rec));
    if (!prof_buf) return;
    rt_register(NULL, prof_buf);
  }
  prof_enabled = mask;
  prof_n = 0;
}

/* Sum of event-timed durations (ms), launch count and summed algorithmic
 * work (flops or bytes, as each launcher reports) of one launch class. */
int atlmx_prof_query(int cls, double *ms_total, long *launches, double *work_total)
{
  return atlmx_prof_query2(cls, ms_total, launches, work_total, NULL);
}

/* As atlmx_prof_query, plus the summed algorithmic HBM bytes (each operand
 * read once, each output written once) where the launcher states them. */
int atlmx_prof_query2(int cls, double *ms_total, long *launches, double *work_total,
                      double *bytes_total)
{
  atlmx_ctx *c = rt_ctx();
  double ms = 0.0, w = 0.0, by = 0.0;
  long n = 0;
  int i;
  if (prof_n > 0) rt_sync(c, "atlmx_prof_query");
  for (i = 0; i < prof_n; ++i) {
    float t = 0.0f;
    if (prof_buf[i].cls != cls) continue;
    RT_CHECK("atlmx_prof_query", hipEventElapsedTime(&t, prof_buf[i].a, prof_buf[i].b));
    ms += t;
    w += prof_buf[i].work;
    by += prof_buf[i].bytes;
    n++;
  }
  if (ms_total) *ms_total = ms;
  if (launches) *launches = n;
  if (work_total) *work_total = w;
  if (bytes_total) *bytes_total = by;
  return 0;
}

void atlmx_prof_reset(void) { prof_n = 0; }

/*
 * rt.h -- host runtime of libatlas_mi355x (C): per-thread HIP stream and
 * device workspaces, host<->device staging, error handling, and the
 * optional per-launch profiler.
 */
#ifndef ATLMX_RT_H
#define ATLMX_RT_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>
#include "kernels.h"

/* workspace slots; each grows independently and is reused across calls */
enum {
  WS_TRI_INV = 0,   /* TRSM inverted diagonal blocks           */
  WS_PANEL,         /* LU panel exchange buffers + counter     */
  WS_INFO,          /* device info / pivot scratch             */
  WS_STAGE_A,       /* staging of host operands                */
  WS_STAGE_B,
  WS_STAGE_C,
  WS_IPIV,
  WS_PLAN,          /* row-interchange move lists               */
  WS_TRSV,          /* few-RHS triangular solve counter + flags */
  WS_NPINV,         /* no-pivot LU: inverted diagonal-block factors */
  WS_TSLU,          /* tournament pivoting: tree + move list    */
  WS_GMRES,         /* GMRES refinement: Krylov basis + vectors */
  WS_MP0,           /* mixed-precision solver buffers          */
  WS_MP1,
  WS_MP2,
  WS_MP3,
  WS_MP4,
  WS_MPSCALE,       /* mixed precision: column scale factors    */
  WS_L3A,           /* TRMM/SYMM/LAUUM: dense copy of the triangle or
                       symmetric matrix; TRTRI/GETRI: the inverse      */
  WS_L3B,           /* TRMM: copy of B                          */
  WS_QR,            /* QR: panel partial sums, larfg coefficients */
  WS_QRV,           /* QR: V, V^T V, T and the two W blocks     */
  WS_TAU,           /* QR: device copy of TAU                   */
  WS_LASWP,         /* getrs: row-interchange gather map        */
  WS_LASWP_B,       /* getrs: right-hand sides being permuted   */
  WS_CPIV,          /* complex LU: pivots of the interleaved real view */
  WS_CPLX_X,        /* complex HEMM/SYMM/TRMM: the dense matrix */
  WS_CPLX_Y,        /* complex TRMM: copy of B */
  WS_NSLOTS
};

typedef struct atlmx_ctx {
  int dev;
  hipStream_t stream;       /* stream every launch of this thread uses */
  hipStream_t own;          /* default stream of this thread (0: legacy default) */
  /* workspaces, one bank per stream: bank 1 serves work issued on aux, so
   * helpers running on the two streams at once never share a buffer */
  void *ws[2][WS_NSLOTS];
  size_t ws_bytes[2][WS_NSLOTS];
  int bank;
  void *panel_buf;          /* LU panel exchange buffer last initialised */
  int panel_set;            /* sequence number of the last panel launch  */
  hipStream_t aux;          /* second stream for lookahead work (lazily created) */
  hipEvent_t ev[3];         /* ordering events between stream and aux   */
  int ncu;                  /* compute units of the device (set with aux) */
  int noprof;               /* 1: the profiler records nothing for this context */
} atlmx_ctx;

atlmx_ctx *rt_ctx(void);
/* this thread's context if it exists, else NULL (creates nothing) */
atlmx_ctx *rt_ctx_peek(void);

/* Abort through cblas_xerbla's convention (message on stderr, exit(-1)) on
 * any HIP failure: a GPU error never returns garbage. */
void rt_fail(const char *rout, const char *what, hipError_t e);
#define RT_CHECK(rout, call)                                    \
  do {                                                          \
    hipError_t e_ = (call);                                     \
    if (e_ != hipSuccess) rt_fail((rout), #call, e_);           \
  } while (0)

int rt_is_device(const void *p);
void *rt_ws(atlmx_ctx *c, int slot, size_t bytes, const char *rout);
/* the context's second stream and its two events (created on first use) */
hipStream_t rt_aux(atlmx_ctx *c, const char *rout);
/* issue the following work on aux (on = 1, with its workspace bank) or
 * back on the thread's stream (on = 0, returning that stream) */
void rt_on_aux(atlmx_ctx *c, hipStream_t main, int on);
void rt_sync(atlmx_ctx *c, const char *rout);

/* Stage a rows x cols column-major operand.  A device pointer is returned
 * as is (*dld = ld); a host operand is copied (when copy_in) into
 * workspace slot `slot` with a padded leading dimension. */
void *rt_stage_in(atlmx_ctx *c, int slot, const void *h, int rows, int cols, long ld,
                  size_t esz, int copy_in, long *dld, const char *rout);
void rt_stage_out(atlmx_ctx *c, void *h, const void *d, int rows, int cols, long ld,
                  long dld, size_t esz, const char *rout);

/* ---- profiler: hipEvents around selected launches on the ctx stream ---- */
enum { PROF_GEMM = 0, PROF_PANEL, PROF_TRSM_INV, PROF_LASWP, PROF_POTF2, PROF_MISC,
       PROF_TRSV, PROF_UPDATE, PROF_NCLASS };
int rt_prof_on(void);
int rt_prof_begin(atlmx_ctx *c, int cls, double work);
int rt_prof_begin2(atlmx_ctx *c, int cls, double work, double bytes);
void rt_prof_end(atlmx_ctx *c, int id);

#endif

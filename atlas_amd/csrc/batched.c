/*
 * batched.c -- host side of the batched interface (no ATLAS counterpart):
 * plans, the autotuner and hipGraph capture for thousands of small,
 * uneven-size GEMM, Cholesky and LU factor+solve instances that run in one
 * persistent launch (k_batched.hip).
 *
 * A plan is made once for a batch shape: the host orders the instances
 * largest first (so the persistent workgroups that take them from one
 * atomic counter finish together), uploads the size and leading-dimension
 * arrays, picks the launch configuration with the autotuner, and captures
 * the counter reset (a fill kernel, never a memset node) and the persistent
 * kernel into a hipGraph.  Running the plan replays that graph on the
 * calling thread's stream.
 *
 * The autotuner times every candidate configuration -- panel width 16 or
 * 32, one to three persistent workgroups per CU; for GEMM the kernel
 * (vector FMAs, or MFMA with one or four wavefronts per 64 x 64 tile) and
 * the workgroups per CU -- on a replica of up to AT_SAMPLE instances drawn evenly from the
 * ordered batch (well-conditioned synthetic data of the same sizes, so
 * tuning never touches the caller's arrays), and keeps the fastest.
 * Results are cached for the process by (kind, precision, nrhs, size
 * buckets), so later plans of a similar shape skip the timing.
 */
#include <fcntl.h>
#include <math.h>
#include <pthread.h>
#include <sys/stat.h>
#include <unistd.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "rt.h"
#include "../../include/cblas.h"
#include "../../include/atlmx.h"

#define AT_SAMPLE 512
#define AT_CACHE 64

enum { BK_POSV = 0, BK_GESV = 1, BK_GEMM = 2, BK_GESVT = 3, BK_CHOLADJ = 4 };

struct atlmx_batch {
  int kind, dbl, count, nrhs, maxn;
  int nb, grid;                   /* launch configuration */
  int *d_meta;                    /* order | n | lda | ldb  (| m | k | ldc | tile_start) */
  int *d_counter;
  void *const *A, *const *B, *const *C;
  void *const *D, *const *E;      /* Cholesky adjoint: gradient G, scratch W */
  int *const *ipiv;
  int *info;
  int ta, tb;
  double alpha, beta;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int captured;
};

typedef struct { int kind, dbl, nrhs, b_max, b_mean, nb, grid; } at_entry;
static at_entry at_cache[AT_CACHE];
static int at_n = 0;
/* plans may be made from any host thread: the cache is read and extended
 * under this lock (the timing itself runs outside it) */
static pthread_mutex_t at_lock = PTHREAD_MUTEX_INITIALIZER;
static int at_file_read = 0;

/* The tuner's choices also persist across processes in a text file, one
 * JSON object per line keyed by the GPU's architecture name and the shape
 * class: {"arch": "gfx950", "kind": 2, "dbl": 1, "nrhs": 0, "b_max": 9,
 * "b_mean": 8, "nb": 2, "grid": 768}.  The file is $ATLMX_TUNE_CACHE, else
 * $XDG_CACHE_HOME/atlas_mi355x/batch_tune.jsonl, else
 * $HOME/.cache/atlas_mi355x/batch_tune.jsonl; ATLMX_TUNE_CACHE set to the
 * empty string turns persistence off.  Lines for other architectures, or
 * that do not parse, are ignored; a new choice is appended as one short
 * write. */
static int at_path(char *buf, size_t len, int make_dirs)
{
  const char *e = getenv("ATLMX_TUNE_CACHE"), *base;
  char dir[4064];
  if (e) {
    if (!*e) return 0;
    snprintf(buf, len, "%.4000s", e);
    return 1;
  }
  base = getenv("XDG_CACHE_HOME");
  if (base && *base) {
    snprintf(dir, sizeof dir, "%.4000s/atlas_mi355x", base);
  } else {
    base = getenv("HOME");
    if (!base || !*base) return 0;
    snprintf(dir, sizeof dir, "%.4000s/.cache", base);
    if (make_dirs) (void)mkdir(dir, 0755);
    snprintf(dir, sizeof dir, "%.4000s/.cache/atlas_mi355x", base);
  }
  if (make_dirs) (void)mkdir(dir, 0755);
  snprintf(buf, len, "%.4000s/batch_tune.jsonl", dir);
  return 1;
}

static void at_arch(char *buf, size_t len)
{
  int dev = 0;
  hipDeviceProp_t p;
  size_t i;
  snprintf(buf, len, "unknown");
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return;
  snprintf(buf, len, "%.63s", p.gcnArchName);
  for (i = 0; buf[i]; ++i)                 /* "gfx950:sramecc+:xnack-" -> "gfx950" */
    if (buf[i] == ':' || buf[i] == '"') { buf[i] = 0; break; }
}

static int cu_count(void);

/* called with at_lock held */
static void at_file_load(void)
{
  char path[4200], arch[64], line[512], a2[64];
  FILE *f;
  if (at_file_read) return;
  at_file_read = 1;
  if (!at_path(path, sizeof path, 0)) return;
  f = fopen(path, "r");
  if (!f) return;
  at_arch(arch, sizeof arch);
  while (fgets(line, sizeof line, f) && at_n < AT_CACHE) {
    at_entry a;
    if (sscanf(line, " {\"arch\": \"%63[^\"]\", \"kind\": %d, \"dbl\": %d, \"nrhs\": %d, "
                     "\"b_max\": %d, \"b_mean\": %d, \"nb\": %d, \"grid\": %d}",
               a2, &a.kind, &a.dbl, &a.nrhs, &a.b_max, &a.b_mean, &a.nb, &a.grid) != 8)
      continue;
    if (strcmp(a2, arch) != 0) continue;
    /* only configurations the tuner itself could have chosen: the GEMM
     * variant 0..2 or the panel width 16 / 32, at most 16 workgroups per CU */
    if (a.kind == BK_GEMM ? (a.nb < 0 || a.nb > 2) : (a.nb != 16 && a.nb != 32)) continue;
    if (a.grid <= 0 || a.grid > 16 * cu_count()) continue;
    at_cache[at_n++] = a;
  }
  fclose(f);
}

static void at_file_append(const at_entry *a)
{
  char path[4200], arch[64], line[512];
  int fd, len;
  if (!at_path(path, sizeof path, 1)) return;
  at_arch(arch, sizeof arch);
  len = snprintf(line, sizeof line,
                 "{\"arch\": \"%s\", \"kind\": %d, \"dbl\": %d, \"nrhs\": %d, \"b_max\": %d, "
                 "\"b_mean\": %d, \"nb\": %d, \"grid\": %d}\n",
                 arch, a->kind, a->dbl, a->nrhs, a->b_max, a->b_mean, a->nb, a->grid);
  fd = open(path, O_WRONLY | O_CREAT | O_APPEND, 0644);
  if (fd < 0) return;
  if (len > 0 && len < (int)sizeof line) (void)!write(fd, line, (size_t)len);
  close(fd);
}

static int cu_count(void)
{
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess)
      cus = 256;
    else
      cus = p.multiProcessorCount;
  }
  return cus;
}

static int bucket(double v)
{
  int b = 0;
  while (v > 16.0 && b < 30) { v /= 2.0; ++b; }
  return b;
}

/* one launch of the plan's kernel with configuration (nb, grid) */
static hipError_t launch(const atlmx_batch *p, int count, const int *meta, void *const *A,
                         void *const *B, void *const *C, int *const *ipiv, int *info,
                         int nb, int grid, hipStream_t s)
{
  const int *order = meta, *n = meta + count, *lda = meta + 2 * count, *ldb = meta + 3 * count;
  hipError_t e = atlmx_k_fill_i32(p->d_counter, 1, 0, s);
  if (e != hipSuccess) return e;
  if (p->kind == BK_GEMM) {
    const int *m = meta + 4 * count, *k = meta + 5 * count, *ldc = meta + 6 * count;
    const int *ts = meta + 7 * count;
    /* meta layout for GEMM: order | n | lda | ldb | m | k | ldc | tile_start;
     * nb is the kernel variant (0 vector FMAs, 1 / 2 MFMA with one / four
     * wavefronts per tile) */
    if (p->dbl)
      return atlmx_k_dbatch_gemm(count, p->ta, p->tb, order, m, n, k, p->alpha,
                                 (const double *const *)A, lda, (const double *const *)B, ldb,
                                 p->beta, (double *const *)C, ldc, ts, p->d_counter, nb, grid,
                                 s);
    return atlmx_k_sbatch_gemm(count, p->ta, p->tb, order, m, n, k, (float)p->alpha,
                               (const float *const *)A, lda, (const float *const *)B, ldb,
                               (float)p->beta, (float *const *)C, ldc, ts, p->d_counter, nb,
                               grid, s);
  }
  if (p->kind == BK_CHOLADJ) {
    /* meta: order | n | lda | ldx | ldg; A = L, B = X, C = Bb */
    const int *ldg = meta + 4 * count;
    if (p->dbl)
      return atlmx_k_dbatch_chol_adjoint(count, order, n, (double *const *)A, lda, p->nrhs,
                                         (double *const *)B, (double *const *)C, ldb,
                                         (double *const *)p->D, ldg, (double *const *)p->E,
                                         p->d_counter, grid, s);
    return atlmx_k_sbatch_chol_adjoint(count, order, n, (float *const *)A, lda, p->nrhs,
                                       (float *const *)B, (float *const *)C, ldb,
                                       (float *const *)p->D, ldg, (float *const *)p->E,
                                       p->d_counter, grid, s);
  }
  if (p->dbl)
    return atlmx_k_dbatch_factor(p->kind == BK_GESVT ? 2 : p->kind, nb, count, order, n,
                                 (double *const *)A, lda, ipiv,
                                 p->nrhs, (double *const *)B, B ? ldb : NULL, info,
                                 p->d_counter, p->maxn, grid, s);
  return atlmx_k_sbatch_factor(p->kind == BK_GESVT ? 2 : p->kind, nb, count, order, n,
                               (float *const *)A, lda, ipiv,
                               p->nrhs, (float *const *)B, B ? ldb : NULL, info, p->d_counter,
                               p->maxn, grid, s);
}

/* upload ints */
static int *to_dev(const int *h, size_t cnt, const char *rout)
{
  int *d = NULL;
  RT_CHECK(rout, hipMalloc((void **)&d, sizeof(int) * (cnt ? cnt : 1)));
  if (cnt) RT_CHECK(rout, hipMemcpy(d, h, sizeof(int) * cnt, hipMemcpyHostToDevice));
  return d;
}

static unsigned long long lcg(unsigned long long *s)
{
  *s = *s * 6364136223846793005ULL + 1442695040888963407ULL;
  return *s >> 11;
}

/* Time every feasible configuration on a replica of a sample of the batch.
 * Replica instances: the sampled sizes, matrices made diagonally dominant
 * (so both Cholesky and LU run to completion), fresh for every timing. */
static void autotune(atlmx_batch *p, const int *h_order, const int *h_n, const int *h_m,
                     const int *h_k, const char *rout)
{
  const int esz = p->dbl ? 8 : 4;
  const int cus = cu_count();
  const int ns = p->count < AT_SAMPLE ? p->count : AT_SAMPLE;
  int *sn = (int *)malloc(sizeof(int) * ns * 8), *meta;
  size_t *offA = (size_t *)malloc(sizeof(size_t) * ns * 3), total = 0;
  void **hA = (void **)malloc(sizeof(void *) * ns * 3);
  unsigned char *pool = NULL, *pool0 = NULL;
  void *const *dA, *const *dB, *const *dC;
  int **hipiv = NULL, *const *dipiv = NULL, *dinfo = NULL, *ipool = NULL;
  double mean = 0.0, best = 1e30;
  int i, c, nbs[2] = {16, 32}, reps, b_max, b_mean;
  hipEvent_t e0, e1;
  hipStream_t s = 0;
  unsigned long long seed = 0x5DEECE66DULL;
  if (!sn || !offA || !hA) { fprintf(stderr, "%s: out of host memory\n", rout); exit(-1); }
  if (p->kind == BK_CHOLADJ) {
    /* one configuration: two workgroups per CU (184 registers and 16 KB of
     * LDS each), so one's barrier-bound triangular solves overlap the
     * other's */
    p->nb = 0;
    p->grid = 2 * cus;
    free(sn); free(offA); free(hA);
    return;
  }

  for (i = 0; i < p->count; ++i)
    mean += (p->kind == BK_GEMM) ? cbrt((double)h_m[i] * h_n[i] * h_k[i]) : h_n[i];
  mean /= p->count;
  b_max = bucket(p->maxn);
  b_mean = bucket(mean);
  pthread_mutex_lock(&at_lock);
  at_file_load();
  for (i = 0; i < at_n; ++i) {
    const at_entry *a = &at_cache[i];
    if (a->kind == p->kind && a->dbl == p->dbl && a->nrhs == p->nrhs && a->b_max == b_max &&
        a->b_mean == b_mean) {
      p->nb = a->nb;
      p->grid = a->grid;
      pthread_mutex_unlock(&at_lock);
      free(sn); free(offA); free(hA);
      return;
    }
  }
  pthread_mutex_unlock(&at_lock);

  /* sample evenly from the ordered batch; sizes and leading dimensions */
  for (i = 0; i < ns; ++i) {
    const int src = h_order[(int)((long)i * p->count / ns)];
    const int n = h_n[src], m = h_m ? h_m[src] : n, k = h_k ? h_k[src] : n;
    sn[i] = i;                      /* order */
    sn[ns + i] = n;                 /* n */
    if (p->kind == BK_GEMM) {
      sn[2 * ns + i] = p->ta == KX_NOTRANS ? m : k;    /* lda */
      sn[3 * ns + i] = p->tb == KX_NOTRANS ? k : n;    /* ldb */
      sn[4 * ns + i] = m;
      sn[5 * ns + i] = k;
      sn[6 * ns + i] = m;                               /* ldc */
      offA[3 * i] = total; total += (size_t)m * k * esz;
      offA[3 * i + 1] = total; total += (size_t)k * n * esz;
      offA[3 * i + 2] = total; total += (size_t)m * n * esz;
    } else {
      sn[2 * ns + i] = n;
      sn[3 * ns + i] = n;
      offA[3 * i] = total; total += (size_t)n * n * esz;
      offA[3 * i + 1] = total; total += (size_t)n * (p->nrhs > 0 ? p->nrhs : 1) * esz;
      offA[3 * i + 2] = total;
    }
    total = (total + 255) & ~(size_t)255;
  }
  if (p->kind == BK_GEMM) {       /* tile prefix sums */
    int t = 0;
    for (i = 0; i < ns; ++i) {
      sn[7 * ns + i] = t;
      t += ((sn[4 * ns + i] + ATLMX_BATCH_TILE - 1) / ATLMX_BATCH_TILE) *
           ((sn[ns + i] + ATLMX_BATCH_TILE - 1) / ATLMX_BATCH_TILE);
    }
    sn = (int *)realloc(sn, sizeof(int) * (ns * 8 + 1));
    sn[8 * ns] = t;
  }
  meta = to_dev(sn, (size_t)ns * 8 + (p->kind == BK_GEMM ? 1 : 0), rout);

  /* the replica's arrays, filled on the host: diagonally dominant */
  {
    unsigned char *h = (unsigned char *)malloc(total ? total : 1);
    size_t q;
    if (!h) { fprintf(stderr, "%s: out of host memory\n", rout); exit(-1); }
    for (q = 0; q < total / esz; ++q) {
      const double v = (double)(lcg(&seed) % 2000001) * 1e-6 - 1.0;
      if (p->dbl) ((double *)h)[q] = v; else ((float *)h)[q] = (float)v;
    }
    if (p->kind != BK_GEMM)
      for (i = 0; i < ns; ++i) {
        const int n = sn[ns + i];
        int d;
        for (d = 0; d < n; ++d) {
          const size_t o = offA[3 * i] / esz + (size_t)d * n + d;
          if (p->dbl) ((double *)h)[o] = 2.0 * n; else ((float *)h)[o] = 2.0f * n;
        }
      }
    RT_CHECK(rout, hipMalloc((void **)&pool, total ? total : 1));
    RT_CHECK(rout, hipMalloc((void **)&pool0, total ? total : 1));
    RT_CHECK(rout, hipMemcpy(pool0, h, total, hipMemcpyHostToDevice));
    free(h);
  }
  for (i = 0; i < ns; ++i) {
    hA[i] = pool + offA[3 * i];
    hA[ns + i] = pool + offA[3 * i + 1];
    hA[2 * ns + i] = pool + offA[3 * i + 2];
  }
  {
    void **d;
    RT_CHECK(rout, hipMalloc((void **)&d, sizeof(void *) * ns * 3));
    RT_CHECK(rout, hipMemcpy(d, hA, sizeof(void *) * ns * 3, hipMemcpyHostToDevice));
    dA = d; dB = d + ns; dC = d + 2 * ns;
  }
  if (p->kind == BK_GESV || p->kind == BK_GESVT) {
    int **hp = (int **)malloc(sizeof(int *) * ns);
    RT_CHECK(rout, hipMalloc((void **)&ipool, sizeof(int) * (size_t)ns * (p->maxn + 1)));
    for (i = 0; i < ns; ++i) hp[i] = ipool + (size_t)i * (p->maxn + 1);
    RT_CHECK(rout, hipMalloc((void **)&hipiv, sizeof(int *) * ns));
    RT_CHECK(rout, hipMemcpy(hipiv, hp, sizeof(int *) * ns, hipMemcpyHostToDevice));
    dipiv = hipiv;
    free(hp);
  }
  RT_CHECK(rout, hipMalloc((void **)&dinfo, sizeof(int) * ns));
  RT_CHECK(rout, hipEventCreate(&e0));
  RT_CHECK(rout, hipEventCreate(&e1));

  p->nb = p->kind == BK_GEMM ? 2 : 32;
  p->grid = cus;
  /* factorizations: panel width 16 or 32 x one to three workgroups per CU;
   * GEMM: the kernel variant (vector FMAs, MFMA with one or four
   * wavefronts per 64 x 64 tile) x workgroups per CU */
  for (c = 0; c < (p->kind == BK_GEMM ? 3 : 2); ++c)
    for (reps = 1; reps <= (p->kind == BK_GEMM ? 4 : 3); ++reps) {
      const int nb = p->kind == BK_GEMM ? c : nbs[c];
      const int grid = (p->kind == BK_GEMM && c == 1 ? 3 : 1) * reps * cus;
      float ms = 0.0f;
      hipError_t e;
      RT_CHECK(rout, hipMemcpy(pool, pool0, total, hipMemcpyDeviceToDevice));
      e = launch(p, ns, meta, dA, p->nrhs > 0 || p->kind == BK_GEMM ? dB : NULL, dC, dipiv,
                 dinfo, nb, grid, s);
      if (e == hipErrorInvalidValue) { (void)hipGetLastError(); continue; }   /* infeasible */
      RT_CHECK(rout, e);
      RT_CHECK(rout, hipMemcpy(pool, pool0, total, hipMemcpyDeviceToDevice));
      RT_CHECK(rout, hipEventRecord(e0, s));
      RT_CHECK(rout, launch(p, ns, meta, dA, p->nrhs > 0 || p->kind == BK_GEMM ? dB : NULL, dC,
                            dipiv, dinfo, nb, grid, s));
      RT_CHECK(rout, hipEventRecord(e1, s));
      RT_CHECK(rout, hipEventSynchronize(e1));
      RT_CHECK(rout, hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) { best = ms; p->nb = nb; p->grid = grid; }
    }
  RT_CHECK(rout, hipDeviceSynchronize());
  RT_CHECK(rout, hipEventDestroy(e0));
  RT_CHECK(rout, hipEventDestroy(e1));
  RT_CHECK(rout, hipFree(pool));
  RT_CHECK(rout, hipFree(pool0));
  RT_CHECK(rout, hipFree((void *)dA));
  RT_CHECK(rout, hipFree(meta));
  RT_CHECK(rout, hipFree(dinfo));
  if (ipool) RT_CHECK(rout, hipFree(ipool));
  if (hipiv) RT_CHECK(rout, hipFree(hipiv));
  free(sn); free(offA); free(hA);
  pthread_mutex_lock(&at_lock);
  {
    at_entry a = {p->kind, p->dbl, p->nrhs, b_max, b_mean, p->nb, p->grid};
    if (at_n < AT_CACHE) at_cache[at_n++] = a;
    at_file_append(&a);
  }
  pthread_mutex_unlock(&at_lock);
}

static atlmx_batch *plan_create2(int kind, int dbl, int count, int ta, int tb, const int *h_m,
                                 const int *h_n, const int *h_k, double alpha, void *const *A,
                                 const int *h_lda, void *const *B, const int *h_ldb, double beta,
                                 void *const *C, const int *h_ldc, int *const *ipiv, int nrhs,
                                 int *info, void *const *D, void *const *E, const char *rout);

static int cmp_cost_desc(const void *a, const void *b)
{
  const double x = ((const double *)a)[0], y = ((const double *)b)[0];
  return (x < y) - (x > y);
}

static atlmx_batch *plan_create(int kind, int dbl, int count, int ta, int tb, const int *h_m,
                                const int *h_n, const int *h_k, double alpha, void *const *A,
                                const int *h_lda, void *const *B, const int *h_ldb, double beta,
                                void *const *C, const int *h_ldc, int *const *ipiv, int nrhs,
                                int *info, const char *rout)
{
  return plan_create2(kind, dbl, count, ta, tb, h_m, h_n, h_k, alpha, A, h_lda, B, h_ldb, beta,
                      C, h_ldc, ipiv, nrhs, info, NULL, NULL, rout);
}

static atlmx_batch *plan_create2(int kind, int dbl, int count, int ta, int tb, const int *h_m,
                                 const int *h_n, const int *h_k, double alpha, void *const *A,
                                 const int *h_lda, void *const *B, const int *h_ldb, double beta,
                                 void *const *C, const int *h_ldc, int *const *ipiv, int nrhs,
                                 int *info, void *const *D, void *const *E, const char *rout)
{
  atlmx_batch *p;
  double *cost;
  int *h, i, t;
  const size_t nmeta = (size_t)count * (kind == BK_GEMM ? 8 : 5) + 1;
  /* refuse orders the factor kernels cannot hold in LDS before touching
   * the device: a NULL plan and a message, not a failed launch later */
  if (kind == BK_POSV || kind == BK_GESV || kind == BK_GESVT) {
    const int lim = atlmx_k_batch_max_order(kind == BK_POSV ? 0 : 1, dbl);
    for (i = 0; i < count; ++i)
      if (h_n[i] < 0 || h_n[i] > lim) {
        fprintf(stderr, "%s: instance %d has order %d; this plan holds orders 0..%d\n",
                rout, i, h_n[i], lim);
        return NULL;
      }
  }
  if (kind == BK_CHOLADJ)
    for (i = 0; i < count; ++i)
      if (h_n[i] < 0) {
        fprintf(stderr, "%s: instance %d has order %d\n", rout, i, h_n[i]);
        return NULL;
      }
  p = (atlmx_batch *)calloc(1, sizeof(*p));
  if (!p) { fprintf(stderr, "%s: out of host memory\n", rout); exit(-1); }
  p->kind = kind; p->dbl = dbl; p->count = count; p->nrhs = nrhs;
  p->A = A; p->B = B; p->C = C; p->ipiv = ipiv; p->info = info; p->D = D; p->E = E;
  p->ta = ta; p->tb = tb; p->alpha = alpha; p->beta = beta;
  cost = (double *)malloc(sizeof(double) * 2 * count);
  h = (int *)malloc(sizeof(int) * nmeta);
  if (!cost || !h) { fprintf(stderr, "%s: out of host memory\n", rout); exit(-1); }
  p->maxn = 1;
  for (i = 0; i < count; ++i) {
    const double n = h_n[i];
    cost[2 * i] = (kind == BK_GEMM) ? (double)h_m[i] * h_n[i] * h_k[i] : n * n * n;
    cost[2 * i + 1] = i;
    if (h_n[i] > p->maxn) p->maxn = h_n[i];
  }
  qsort(cost, count, 2 * sizeof(double), cmp_cost_desc);
  for (i = 0; i < count; ++i) h[i] = (int)cost[2 * i + 1];
  /* n, lda, ldb are indexed by instance (the kernel reads n[order[it]]) */
  for (i = 0; i < count; ++i) {
    h[count + i] = h_n[i];
    h[2 * count + i] = h_lda[i];
    h[3 * count + i] = h_ldb ? h_ldb[i] : 1;
  }
  if (kind == BK_CHOLADJ)
    for (i = 0; i < count; ++i) h[4 * count + i] = h_ldc[i];     /* ldg */
  if (kind == BK_GEMM) {
    for (i = 0; i < count; ++i) {
      h[4 * count + i] = h_m[i];
      h[5 * count + i] = h_k[i];
      h[6 * count + i] = h_ldc[i];
    }
    for (t = 0, i = 0; i < count; ++i) {        /* tiles, in processing order */
      const int src = h[i];
      h[7 * count + i] = t;
      t += ((h_m[src] + ATLMX_BATCH_TILE - 1) / ATLMX_BATCH_TILE) *
           ((h_n[src] + ATLMX_BATCH_TILE - 1) / ATLMX_BATCH_TILE);
    }
    h[8 * count] = t;
  }
  p->d_meta = to_dev(h, nmeta, rout);
  RT_CHECK(rout, hipMalloc((void **)&p->d_counter, 64));
  {
    int *ho = (int *)malloc(sizeof(int) * count);
    for (i = 0; i < count; ++i) ho[i] = h[i];
    autotune(p, ho, h_n, h_m, h_k, rout);
    free(ho);
  }
  free(cost);
  free(h);
  return p;
}

int atlmx_batch_run(atlmx_batch *p)
{
  static const char *rout = "atlmx_batch_run";
  atlmx_ctx *c = rt_ctx();
  if (!p || p->count <= 0) return 0;
  if (!p->captured) {
    /* capture the counter reset and the persistent kernel once */
    hipStream_t s;
    RT_CHECK(rout, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    RT_CHECK(rout, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    RT_CHECK(rout, launch(p, p->count, p->d_meta, p->A, p->B, p->C, p->ipiv, p->info, p->nb,
                          p->grid, s));
    RT_CHECK(rout, hipStreamEndCapture(s, &p->graph));
    RT_CHECK(rout, hipGraphInstantiate(&p->exec, p->graph, NULL, NULL, 0));
    RT_CHECK(rout, hipStreamDestroy(s));
    p->captured = 1;
  }
  RT_CHECK(rout, hipGraphLaunch(p->exec, c->stream));
  return 0;
}

void atlmx_batch_destroy(atlmx_batch *p)
{
  static const char *rout = "atlmx_batch_destroy";
  if (!p) return;
  RT_CHECK(rout, hipDeviceSynchronize());
  if (p->captured) {
    RT_CHECK(rout, hipGraphExecDestroy(p->exec));
    RT_CHECK(rout, hipGraphDestroy(p->graph));
  }
  RT_CHECK(rout, hipFree(p->d_meta));
  RT_CHECK(rout, hipFree(p->d_counter));
  free(p);
}

void atlmx_batch_config(const atlmx_batch *p, int *nb, int *grid)
{
  if (nb) *nb = p ? p->nb : 0;
  if (grid) *grid = p ? p->grid : 0;
}

/* ---- plans ---- */
atlmx_batch *atlmx_dposv_batch_plan(int count, const int *h_n, double *const *d_A,
                                    const int *h_lda, int nrhs, double *const *d_B,
                                    const int *h_ldb, int *d_info)
{
  static const char *rout = "atlmx_dposv_batch_plan";
  if (count < 0 || nrhs < 0 || nrhs > ATLMX_BATCH_MAXRHS) cblas_xerbla(5, rout, "");
  return plan_create(BK_POSV, 1, count, 0, 0, NULL, h_n, NULL, 0.0, (void *const *)d_A, h_lda,
                     (void *const *)d_B, h_ldb, 0.0, NULL, NULL, NULL, nrhs, d_info, rout);
}

atlmx_batch *atlmx_sposv_batch_plan(int count, const int *h_n, float *const *d_A,
                                    const int *h_lda, int nrhs, float *const *d_B,
                                    const int *h_ldb, int *d_info)
{
  static const char *rout = "atlmx_sposv_batch_plan";
  if (count < 0 || nrhs < 0 || nrhs > ATLMX_BATCH_MAXRHS) cblas_xerbla(5, rout, "");
  return plan_create(BK_POSV, 0, count, 0, 0, NULL, h_n, NULL, 0.0, (void *const *)d_A, h_lda,
                     (void *const *)d_B, h_ldb, 0.0, NULL, NULL, NULL, nrhs, d_info, rout);
}

atlmx_batch *atlmx_dgesv_batch_plan(int count, const int *h_n, double *const *d_A,
                                    const int *h_lda, int *const *d_ipiv, int nrhs,
                                    double *const *d_B, const int *h_ldb, int *d_info)
{
  static const char *rout = "atlmx_dgesv_batch_plan";
  if (count < 0 || nrhs < 0 || nrhs > ATLMX_BATCH_MAXRHS) cblas_xerbla(6, rout, "");
  return plan_create(BK_GESV, 1, count, 0, 0, NULL, h_n, NULL, 0.0, (void *const *)d_A, h_lda,
                     (void *const *)d_B, h_ldb, 0.0, NULL, NULL, d_ipiv, nrhs, d_info, rout);
}

atlmx_batch *atlmx_sgesv_batch_plan(int count, const int *h_n, float *const *d_A,
                                    const int *h_lda, int *const *d_ipiv, int nrhs,
                                    float *const *d_B, const int *h_ldb, int *d_info)
{
  static const char *rout = "atlmx_sgesv_batch_plan";
  if (count < 0 || nrhs < 0 || nrhs > ATLMX_BATCH_MAXRHS) cblas_xerbla(6, rout, "");
  return plan_create(BK_GESV, 0, count, 0, 0, NULL, h_n, NULL, 0.0, (void *const *)d_A, h_lda,
                     (void *const *)d_B, h_ldb, 0.0, NULL, NULL, d_ipiv, nrhs, d_info, rout);
}

atlmx_batch *atlmx_dgemm_batch_plan(int TA, int TB, int count, const int *h_m, const int *h_n,
                                    const int *h_k, double alpha, const double *const *d_A,
                                    const int *h_lda, const double *const *d_B,
                                    const int *h_ldb, double beta, double *const *d_C,
                                    const int *h_ldc)
{
  static const char *rout = "atlmx_dgemm_batch_plan";
  if (TA != CblasNoTrans && TA != CblasTrans) cblas_xerbla(1, rout, "");
  if (TB != CblasNoTrans && TB != CblasTrans) cblas_xerbla(2, rout, "");
  return plan_create(BK_GEMM, 1, count, TA == CblasNoTrans ? KX_NOTRANS : KX_TRANS,
                     TB == CblasNoTrans ? KX_NOTRANS : KX_TRANS, h_m, h_n, h_k, alpha,
                     (void *const *)d_A, h_lda, (void *const *)d_B, h_ldb, beta,
                     (void *const *)d_C, h_ldc, NULL, 0, NULL, rout);
}

atlmx_batch *atlmx_sgemm_batch_plan(int TA, int TB, int count, const int *h_m, const int *h_n,
                                    const int *h_k, float alpha, const float *const *d_A,
                                    const int *h_lda, const float *const *d_B, const int *h_ldb,
                                    float beta, float *const *d_C, const int *h_ldc)
{
  static const char *rout = "atlmx_sgemm_batch_plan";
  if (TA != CblasNoTrans && TA != CblasTrans) cblas_xerbla(1, rout, "");
  if (TB != CblasNoTrans && TB != CblasTrans) cblas_xerbla(2, rout, "");
  return plan_create(BK_GEMM, 0, count, TA == CblasNoTrans ? KX_NOTRANS : KX_TRANS,
                     TB == CblasNoTrans ? KX_NOTRANS : KX_TRANS, h_m, h_n, h_k, alpha,
                     (void *const *)d_A, h_lda, (void *const *)d_B, h_ldb, beta,
                     (void *const *)d_C, h_ldc, NULL, 0, NULL, rout);
}

/* ---- adjoints ---- */
atlmx_batch *atlmx_dgesv_t_batch_plan(int count, const int *h_n, double *const *d_A,
                                      const int *h_lda, int *const *d_ipiv, int nrhs,
                                      double *const *d_B, const int *h_ldb, int *d_info)
{
  static const char *rout = "atlmx_dgesv_t_batch_plan";
  if (count < 0 || nrhs < 0 || nrhs > ATLMX_BATCH_MAXRHS) cblas_xerbla(6, rout, "");
  return plan_create(BK_GESVT, 1, count, 0, 0, NULL, h_n, NULL, 0.0, (void *const *)d_A, h_lda,
                     (void *const *)d_B, h_ldb, 0.0, NULL, NULL, d_ipiv, nrhs, d_info, rout);
}

atlmx_batch *atlmx_sgesv_t_batch_plan(int count, const int *h_n, float *const *d_A,
                                      const int *h_lda, int *const *d_ipiv, int nrhs,
                                      float *const *d_B, const int *h_ldb, int *d_info)
{
  static const char *rout = "atlmx_sgesv_t_batch_plan";
  if (count < 0 || nrhs < 0 || nrhs > ATLMX_BATCH_MAXRHS) cblas_xerbla(6, rout, "");
  return plan_create(BK_GESVT, 0, count, 0, 0, NULL, h_n, NULL, 0.0, (void *const *)d_A, h_lda,
                     (void *const *)d_B, h_ldb, 0.0, NULL, NULL, d_ipiv, nrhs, d_info, rout);
}

atlmx_batch *atlmx_dposv_adjoint_batch_plan(int count, const int *h_n, double *const *d_L,
                                            const int *h_lda, int nrhs, double *const *d_X,
                                            double *const *d_Bb, const int *h_ldx,
                                            double *const *d_G, const int *h_ldg,
                                            double *const *d_W)
{
  static const char *rout = "atlmx_dposv_adjoint_batch_plan";
  if (count < 0 || nrhs < 0) cblas_xerbla(5, rout, "");
  return plan_create2(BK_CHOLADJ, 1, count, 0, 0, NULL, h_n, NULL, 0.0, (void *const *)d_L,
                      h_lda, (void *const *)d_X, h_ldx, 0.0, (void *const *)d_Bb, h_ldg, NULL,
                      nrhs, NULL, (void *const *)d_G, (void *const *)d_W, rout);
}

atlmx_batch *atlmx_sposv_adjoint_batch_plan(int count, const int *h_n, float *const *d_L,
                                            const int *h_lda, int nrhs, float *const *d_X,
                                            float *const *d_Bb, const int *h_ldx,
                                            float *const *d_G, const int *h_ldg,
                                            float *const *d_W)
{
  static const char *rout = "atlmx_sposv_adjoint_batch_plan";
  if (count < 0 || nrhs < 0) cblas_xerbla(5, rout, "");
  return plan_create2(BK_CHOLADJ, 0, count, 0, 0, NULL, h_n, NULL, 0.0, (void *const *)d_L,
                      h_lda, (void *const *)d_X, h_ldx, 0.0, (void *const *)d_Bb, h_ldg, NULL,
                      nrhs, NULL, (void *const *)d_G, (void *const *)d_W, rout);
}

int atlmx_batch_max_order(int kind, int dbl)
{
  return atlmx_k_batch_max_order(kind == 0 ? 0 : 1, dbl ? 1 : 0);
}

/*
 * k_gemm.hip -- gfx950 MFMA GEMM: C = alpha*op(A)*op(B) + beta*C.
 *
 * Replaces ATLAS's access-major engine (src/blas/ammm/ATL_ammm.c:12-356,
 * the loop nests of ATL_iploops.c:19-735 and the generated ammkern_t
 * microkernels, include/atlas_amm.h:46) and its pthread decomposition
 * (src/threads/blas/level3/ATL_tammm_G.c:211-420).  Instead of packing
 * op(A)/op(B) into access-major blocks with a copy pass, each workgroup
 * streams its K-slices of op(A) and op(B) from HBM straight into a
 * double-buffered LDS tile (the "copy" of ATL_iploops.c:397-415 is the
 * global->LDS load), and the C block stays in registers until one fused
 * alpha/beta epilogue (the blk2c of ATL_iploops.c:416-417).  Tile
 * scheduling is on the GPU: one workgroup per 128x128 C tile, numbered so
 * that each of the 8 XCDs works on a compact group of tiles.
 *
 * Tiling (both precisions): 256 threads = 4 wavefronts in a 2x2 grid; each
 * wavefront owns a 64x64 block of C = 4x4 MFMA tiles of 16x16
 * (v_mfma_f64_16x16x4f64 or v_mfma_f32_16x16x4f32).  LDS holds op(A) as
 * [m][k] and op(B) as [n][k] (k contiguous, row padded), so each lane reads
 * its four K values of an MFMA operand with one 16/32-byte LDS load.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "kernels.h"

namespace {

constexpr int BM = ATLMX_GEMM_BM;
constexpr int BN = ATLMX_GEMM_BN;
constexpr int NT = 256;
#ifndef ATLMX_GEMM_GROUP_M
#define ATLMX_GEMM_GROUP_M 8
#endif
constexpr int GROUP_M = ATLMX_GEMM_GROUP_M;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4 acc_t;
  typedef d2 vec_t;                 /* one 16-byte global load */
  static constexpr int VEC = 2;
  static constexpr int BK = 16;     /* 2 x 36 KB LDS: two workgroups per CU */
  static constexpr int PADK = 2;    /* LDS row = 18 doubles = 144 B */
  static constexpr int PADR = 2;    /* [k][row] tiles: 130 doubles per k */
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  /* v_mfma_f64_16x16x4f64: lane l, element r holds D[(l>>4) + 4r][l&15] */
  __device__ static int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
  typedef f4 acc_t;
  typedef f4 vec_t;
  static constexpr int VEC = 4;
  static constexpr int BK = 16;     /* 2 x 20 KB LDS: three workgroups per CU */
  static constexpr int PADK = 4;    /* LDS row = 20 floats = 80 B */
  static constexpr int PADR = 4;    /* [k][row] tiles: 132 floats per k */
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  /* v_mfma_f32_16x16x4f32: lane l, element r holds D[4(l>>4) + r][l&15] */
  __device__ static int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename T>
struct GemmArgs {
  int M, N, K;
  T alpha, beta;
  const T *A; long lda;
  const T *B; long ldb;
  T *C; long ldc;
  int tri;
  int tiles_m, tiles_n;
  int per_xcd;            /* tiles per XCD slot in the remapped numbering */
  int vec_a, vec_b;       /* 16-byte vector loads allowed */
  const int *abort_flag;
};

/* Load one BKxBM (or BKxBN) slice of an operand into registers.
 * KCONT: the operand's K index is contiguous in memory (op(A) with A
 * transposed, or op(B) with B not transposed).  rows = extent in M (or N).
 * Element (r, k) of the slice lives at P[k + r*ld] when KCONT, else at
 * P[r + k*ld]. */
template <typename T, bool KCONT, int ROWS, int BK>
struct SliceLoader {
  typedef Mfma<T> MF;
  static constexpr int VEC = MF::VEC;
  static constexpr int NV = ROWS * BK / VEC / NT;   /* vectors per thread */
  typename MF::vec_t v[NV];

  __device__ void load(const T *P, long ld, int r0, int k0, int rows, int K,
                       bool vec_ok, int tid) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int lin = tid + NT * i;
      int r, k;
      if (KCONT) { r = lin / (BK / VEC); k = (lin % (BK / VEC)) * VEC; }
      else       { k = lin / (ROWS / VEC); r = (lin % (ROWS / VEC)) * VEC; }
      const int gr = r0 + r, gk = k0 + k;
      if (KCONT) {
        if (vec_ok && gr < rows && gk + VEC <= K) {
          v[i] = *reinterpret_cast<const typename MF::vec_t *>(P + (size_t)gr * ld + gk);
        } else {
#pragma unroll
          for (int e = 0; e < VEC; ++e)
            v[i][e] = (gr < rows && gk + e < K) ? P[(size_t)gr * ld + gk + e] : T(0);
        }
      } else {
        if (vec_ok && gk < K && gr + VEC <= rows) {
          v[i] = *reinterpret_cast<const typename MF::vec_t *>(P + (size_t)gk * ld + gr);
        } else {
#pragma unroll
          for (int e = 0; e < VEC; ++e)
            v[i][e] = (gk < K && gr + e < rows) ? P[(size_t)gk * ld + gr + e] : T(0);
        }
      }
    }
  }

  /* S is the LDS tile: [ROWS][BK + PADK] (k contiguous) when KCONT, else
   * [BK][ROWS + PADR] (rows contiguous), so each loaded vector is stored
   * whole */
  static constexpr int LDK = BK + MF::PADK;
  static constexpr int LDR = ROWS + MF::PADR;
  static constexpr int SIZE = KCONT ? ROWS * LDK : BK * LDR;
  __device__ void store(T *S, int tid) const {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int lin = tid + NT * i;
      if (KCONT) {
        const int r = lin / (BK / VEC), k = (lin % (BK / VEC)) * VEC;
        *reinterpret_cast<typename MF::vec_t *>(S + r * LDK + k) = v[i];
      } else {
        const int k = lin / (ROWS / VEC), r = (lin % (ROWS / VEC)) * VEC;
        *reinterpret_cast<typename MF::vec_t *>(S + k * LDR + r) = v[i];
      }
    }
  }
  /* element (r, k) of the tile */
  __device__ static int at(int r, int k) { return KCONT ? r * LDK + k : k * LDR + r; }
};

/* Two workgroups per CU (at most 256 registers per lane, 2 x 72 KB LDS):
 * one workgroup's prologue, barriers and C epilogue overlap the other's
 * MFMA stream, which matters most for the short-K trailing updates of the
 * LU and Cholesky. */
/* KBAND: the K-band variant (tri = KX_KBAND_*), a separate instance so
 * that the other launches keep the plain main loop's code */
template <typename T, int TA, int TB, bool KBAND = false>
__global__ __launch_bounds__(NT, 2) void gemm_kernel(GemmArgs<T> g) {
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int BK = MF::BK;
  constexpr int LDK = BK + MF::PADK;
  constexpr bool A_KCONT = (TA != KX_NOTRANS);   /* op(A)[m][k] = A[k + m*lda] */
  constexpr bool B_KCONT = (TB == KX_NOTRANS);   /* op(B)[k][n] = B[k + n*ldb] */

  if (g.abort_flag && *g.abort_flag) return;

  /* ---- tile scheduling: XCD-compact numbering, then grouped raster ----
   * Workgroups go to the XCDs round-robin, so the compact numbering gives
   * each XCD one contiguous chunk of tiles.  Under a triangle mask that
   * would leave the chunks of one side all live and those of the other all
   * skipped; there the tiles are dealt round-robin instead (tile = bid),
   * which spreads the live ones evenly over the XCDs. */
  const int bid = blockIdx.x;
  const int tile = g.tri ? bid : (bid & 7) * g.per_xcd + (bid >> 3);
  const int ntiles = g.tiles_m * g.tiles_n;
  if (tile >= ntiles) return;
  const int per_group = GROUP_M * g.tiles_n;
  const int group = tile / per_group;
  const int first_m = group * GROUP_M;
  const int gsz = min(g.tiles_m - first_m, GROUP_M);
  const int tm = first_m + (tile % per_group) % gsz;
  const int tn = (tile % per_group) / gsz;
  const int m0 = tm * BM, n0 = tn * BN;

  /* the triangle masks; for the real view of a complex C (KX_CLOWER /
   * KX_CUPPER) real row r is complex row r >> 1 */
  if (g.tri == KX_LOWER && m0 + BM - 1 < n0) return;   /* tile above diagonal */
  if (g.tri == KX_UPPER && n0 + BN - 1 < m0) return;   /* tile below diagonal */
  if (g.tri == KX_CLOWER && ((m0 + BM - 1) >> 1) < n0) return;
  if (g.tri == KX_CUPPER && n0 + BN - 1 < (m0 >> 1)) return;

  typedef SliceLoader<T, A_KCONT, BM, BK> LA;
  typedef SliceLoader<T, B_KCONT, BN, BK> LB;
  __shared__ __attribute__((aligned(16))) T As[2][LA::SIZE];
  __shared__ __attribute__((aligned(16))) T Bs[2][LB::SIZE];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;

  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  LA la;
  LB lb;
  /* the K slices this tile reads: all, or under a K band only those that
   * meet the triangle's nonzeros.  AL / AU: A's real row r and column k
   * are complex row r >> 1 and column k >> 1 (or real ones), so rows
   * [m0, m0 + BM) (m0 even) need k < m0 + BM (lower) or k >= m0 (upper);
   * BL / BU: B's real row k is complex row k >> 1, so columns [n0, n0 + BN)
   * need k >= 2 n0 (lower) or k < 2 (n0 + BN) (upper); RBL / RBU: the same
   * for a real op(B), k >= n0 or k < n0 + BN.  The skipped slices hold
   * zeros. */
  int kb = 0, ke = g.K;
  if constexpr (KBAND) {
    if (g.tri == KX_KBAND_AL) ke = min(ke, m0 + BM);
    if (g.tri == KX_KBAND_AU) kb = m0;
    if (g.tri == KX_KBAND_BL) kb = 2 * n0;
    if (g.tri == KX_KBAND_BU) ke = min(ke, 2 * (n0 + BN));
    if (g.tri == KX_KBAND_RBL) kb = n0;
    if (g.tri == KX_KBAND_RBU) ke = min(ke, n0 + BN);
  }
  const int kt0 = kb / BK, kt1 = (ke + BK - 1) / BK;

  if (kt0 < kt1) {
    la.load(g.A, g.lda, m0, kt0 * BK, g.M, g.K, g.vec_a, tid);
    lb.load(g.B, g.ldb, n0, kt0 * BK, g.N, g.K, g.vec_b, tid);
    la.store(As[0], tid);
    lb.store(Bs[0], tid);
  }
  __syncthreads();

  /* MFMA k-slot (lane>>4) of sub-step s holds K index 4*(lane>>4) + s of a
   * 16-wide K group, so each lane reads its 4 K values with one LDS load.
   * The product is formed transposed, D = B^T * A^T, so a lane's output
   * elements run down a column of C (coalesced stores). */
  const int kq = (lane >> 4) * 4;
  const int lr = lane & 15;
  for (int kt = kt0; kt < kt1; ++kt) {
    const int cur = (kt - kt0) & 1;
    if (kt + 1 < kt1) {
      la.load(g.A, g.lda, m0, (kt + 1) * BK, g.M, g.K, g.vec_a, tid);
      lb.load(g.B, g.ldb, n0, (kt + 1) * BK, g.N, g.K, g.vec_b, tid);
    }
#pragma unroll
    for (int kg = 0; kg < BK / 16; ++kg) {
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        T a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[i] = As[cur][LA::at(wm + i * 16 + lr, kg * 16 + kq + s)];
          b[i] = Bs[cur][LB::at(wn + i * 16 + lr, kg * 16 + kq + s)];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = MF::mma(b[j], a[i], acc[i][j]);
      }
    }
    if (kt + 1 < kt1) {
      la.store(As[cur ^ 1], tid);
      lb.store(Bs[cur ^ 1], tid);
    }
    __syncthreads();
  }

  /* ---- fused epilogue: C = alpha*acc + beta*C ---- */
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        /* acc[i][j] holds D = (C tile)^T: D[n][m] */
        const int row = m0 + wm + i * 16 + lr;
        const int col = n0 + wn + j * 16 + MF::drow(lane, r);
        if (row < g.M && col < g.N) {
          if (g.tri == KX_LOWER && row < col) continue;
          if (g.tri == KX_UPPER && row > col) continue;
          if (g.tri == KX_CLOWER && (row >> 1) < col) continue;
          if (g.tri == KX_CUPPER && (row >> 1) > col) continue;
          T *c = g.C + (size_t)col * g.ldc + row;
          T v = g.alpha * acc[i][j][r];
          if (g.beta != T(0)) v += g.beta * *c;
          *c = v;
        }
      }
}

/* C += alpha A B for the short-K trailing updates (A not transposed, K
 * contiguous in B).  The product is formed untransposed, so a lane's four
 * accumulator elements are four consecutive rows of one column of C: the C
 * tile moves with 16-byte loads and stores.  The C loads are issued right
 * after the first K slice's operand loads and added only after the last
 * MFMA, so they arrive while the MFMAs run; each next K slice (of KB
 * columns) is loaded into registers before the current slice's MFMAs
 * (alpha is applied to A as it is staged).  KB = 32 and no occupancy
 * request leave the compiler room to schedule without spilling (about 180
 * registers, so two workgroups still share a CU): 0.73 ms for a
 * 16128 x 16128 x 128 update, against 1.04 ms for the transposed product
 * with the C tile loaded before the MFMAs. */
template <typename T, int KB>
__global__ __launch_bounds__(NT, 1) void update_kernel(GemmArgs<T> g) {
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  typedef SliceLoader<T, false, BM, KB> LA;
  typedef SliceLoader<T, true, BN, KB> LB;
  if (g.abort_flag && *g.abort_flag) return;
  const int bid = blockIdx.x;
  const int tile = (bid & 7) * g.per_xcd + (bid >> 3);
  const int ntiles = g.tiles_m * g.tiles_n;
  if (tile >= ntiles) return;
  const int per_group = GROUP_M * g.tiles_n;
  const int group = tile / per_group;
  const int first_m = group * GROUP_M;
  const int gsz = min(g.tiles_m - first_m, GROUP_M);
  const int tm = first_m + (tile % per_group) % gsz;
  const int tn = (tile % per_group) / gsz;
  const int m0 = tm * BM, n0 = tn * BN;

  __shared__ __attribute__((aligned(16))) T As[LA::SIZE];
  __shared__ __attribute__((aligned(16))) T Bs[LB::SIZE];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int kq = (lane >> 4) * 4, lr = lane & 15;
  /* accumulator element r of block (i, j): row rb(i) + r, column cb(j) */
  const int rb0 = m0 + wm + 4 * (lane >> 4), cb0 = n0 + wn + lr;
  const bool cvec = (g.ldc % 4 == 0) && (((uintptr_t)g.C % 16) == 0);

  LA la;
  LB lb;
  la.load(g.A, g.lda, m0, 0, g.M, g.K, g.vec_a, tid);
  lb.load(g.B, g.ldb, n0, 0, g.N, g.K, g.vec_b, tid);
  acc_t cv[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = rb0 + i * 16, col = cb0 + j * 16;
      const T *cp = g.C + (size_t)col * g.ldc + row;
      if (col < g.N && row + 3 < g.M && cvec) {
        cv[i][j] = *reinterpret_cast<const acc_t *>(cp);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[i][j][r] = (col < g.N && row + r < g.M) ? cp[r] : T(0);
      }
    }
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  for (int k0 = 0; k0 < g.K; k0 += KB) {
#pragma unroll
    for (int i = 0; i < LA::NV; ++i) la.v[i] *= g.alpha;
    if (k0 > 0) __syncthreads();
    la.store(As, tid);
    lb.store(Bs, tid);
    __syncthreads();
    if (k0 + KB < g.K) {
      la.load(g.A, g.lda, m0, k0 + KB, g.M, g.K, g.vec_a, tid);
      lb.load(g.B, g.ldb, n0, k0 + KB, g.N, g.K, g.vec_b, tid);
    }
#pragma unroll
    for (int kg = 0; kg < KB / 16; ++kg)
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4) {
        T a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          a[i] = As[LA::at(wm + i * 16 + lr, kg * 16 + kq + s4)];
          b[i] = Bs[LB::at(wn + i * 16 + lr, kg * 16 + kq + s4)];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = MF::mma(a[i], b[j], acc[i][j]);
      }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = rb0 + i * 16, col = cb0 + j * 16;
      T *cp = g.C + (size_t)col * g.ldc + row;
      const acc_t v = acc[i][j] + cv[i][j];
      if (col < g.N && row + 3 < g.M && cvec) {
        *reinterpret_cast<acc_t *>(cp) = v;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (col < g.N && row + r < g.M) cp[r] = v[r];
      }
    }
}

/* The FP64 short-K update, C += alpha op(A) op(B) with beta = 1 (the LU
 * and Cholesky trailing updates, K <= 128 or so), op(B) either form.  The
 * FP64 MFMA's accumulator holds rows (l>>4) + 4r of a column, so the
 * product is formed transposed as in gemm_kernel (a lane's elements run down
 * one row of C, and the 16 lanes of a row group read 128 contiguous bytes of
 * a column); the C tile (64 doubles a thread, in registers) is loaded right
 * after the first K slice's operands and added after the last MFMA, so its
 * HBM traffic overlaps the MFMAs.  That takes about 300 registers a lane,
 * so one workgroup runs per CU.  tri: only the referenced triangle of C is
 * read and written (the SYRK form). */
template <int TB>
__global__ __launch_bounds__(NT, 1) void update_d_kernel(GemmArgs<double> g) {
  typedef Mfma<double> MF;
  typedef MF::acc_t acc_t;
  constexpr int BK = MF::BK;
  constexpr bool B_KCONT = (TB == KX_NOTRANS);
  typedef SliceLoader<double, false, BM, BK> LA;
  typedef SliceLoader<double, B_KCONT, BN, BK> LB;
  if (g.abort_flag && *g.abort_flag) return;
  const int bid = blockIdx.x;
  /* masked updates deal their tiles round-robin over the XCDs, as in
   * gemm_kernel */
  const int tile = g.tri ? bid : (bid & 7) * g.per_xcd + (bid >> 3);
  const int ntiles = g.tiles_m * g.tiles_n;
  if (tile >= ntiles) return;
  const int per_group = GROUP_M * g.tiles_n;
  const int group = tile / per_group;
  const int first_m = group * GROUP_M;
  const int gsz = min(g.tiles_m - first_m, GROUP_M);
  const int tm = first_m + (tile % per_group) % gsz;
  const int tn = (tile % per_group) / gsz;
  const int m0 = tm * BM, n0 = tn * BN;
  if (g.tri == KX_LOWER && m0 + BM - 1 < n0) return;
  if (g.tri == KX_UPPER && n0 + BN - 1 < m0) return;

  __shared__ __attribute__((aligned(16))) double As[2][LA::SIZE];
  __shared__ __attribute__((aligned(16))) double Bs[2][LB::SIZE];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int kq = (lane >> 4) * 4, lr = lane & 15;
  const int nk = (g.K + BK - 1) / BK;

  LA la;
  LB lb;
  la.load(g.A, g.lda, m0, 0, g.M, g.K, g.vec_a, tid);
  lb.load(g.B, g.ldb, n0, 0, g.N, g.K, g.vec_b, tid);
  /* the C tile, element r of block (i, j) = C(row, col) as the accumulator
   * holds it: row m0 + wm + 16 i + lr, column n0 + wn + 16 j + drow(r) */
  acc_t cv[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + lr, col = n0 + wn + j * 16 + MF::drow(lane, r);
        const bool in = row < g.M && col < g.N &&
                        !(g.tri == KX_LOWER && row < col) && !(g.tri == KX_UPPER && row > col);
        cv[i][j][r] = in ? g.C[(size_t)col * g.ldc + row] : 0.0;
      }
  la.store(As[0], tid);
  lb.store(Bs[0], tid);
  __syncthreads();
  acc_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) {
      la.load(g.A, g.lda, m0, (kt + 1) * BK, g.M, g.K, g.vec_a, tid);
      lb.load(g.B, g.ldb, n0, (kt + 1) * BK, g.N, g.K, g.vec_b, tid);
    }
#pragma unroll
    for (int s4 = 0; s4 < 4; ++s4) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = As[cur][LA::at(wm + i * 16 + lr, kq + s4)];
        b[i] = Bs[cur][LB::at(wn + i * 16 + lr, kq + s4)];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = MF::mma(b[j], a[i], acc[i][j]);
    }
    if (kt + 1 < nk) {
      la.store(As[cur ^ 1], tid);
      lb.store(Bs[cur ^ 1], tid);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + lr, col = n0 + wn + j * 16 + MF::drow(lane, r);
        const bool in = row < g.M && col < g.N &&
                        !(g.tri == KX_LOWER && row < col) && !(g.tri == KX_UPPER && row > col);
        if (in) g.C[(size_t)col * g.ldc + row] = g.alpha * acc[i][j][r] + cv[i][j][r];
      }
}

/* The same update with its operands rounded to 16 bits: FP32 A and B are
 * converted (round to nearest even) as they are staged into LDS, the
 * products run on v_mfma_f32_16x16x16_{bf16,f16} and accumulate in FP32
 * onto the FP32 C tile.  LP = 1: bfloat16, LP = 2: IEEE half.  LDS holds
 * both operands [row][k] with k contiguous, so a lane's four K values of an
 * MFMA operand are one 8-byte load. */
typedef short s4 __attribute__((ext_vector_type(4)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));

__device__ inline unsigned short to_bf16(float x) {
  const unsigned u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);   /* NaN */
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline unsigned short to_f16(float x) {
  const _Float16 h = (_Float16)x;
  return __builtin_bit_cast(unsigned short, h);
}

template <int LP> struct Low;
template <> struct Low<1> {
  __device__ static unsigned short cvt(float x) { return to_bf16(x); }
  __device__ static f4 mma(uint2 a, uint2 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4, a),
                                                     __builtin_bit_cast(s4, b), c, 0, 0, 0);
  }
};
template <> struct Low<2> {
  __device__ static unsigned short cvt(float x) { return to_f16(x); }
  __device__ static f4 mma(uint2 a, uint2 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(h4, a),
                                                 __builtin_bit_cast(h4, b), c, 0, 0, 0);
  }
};

template <int LP>
__global__ __launch_bounds__(NT, 1) void update_lp_kernel(GemmArgs<float> g) {
  typedef Mfma<float> MF;
  constexpr int KB = 32;
  constexpr int LDH = KB + 8;                     /* 16-bit elements per LDS row */
  typedef SliceLoader<float, false, BM, KB> LA;   /* A: rows contiguous */
  typedef SliceLoader<float, true, BN, KB> LB;    /* B: k contiguous */
  if (g.abort_flag && *g.abort_flag) return;
  const int bid = blockIdx.x;
  const int tile = (bid & 7) * g.per_xcd + (bid >> 3);
  const int ntiles = g.tiles_m * g.tiles_n;
  if (tile >= ntiles) return;
  const int per_group = GROUP_M * g.tiles_n;
  const int group = tile / per_group;
  const int first_m = group * GROUP_M;
  const int gsz = min(g.tiles_m - first_m, GROUP_M);
  const int tm = first_m + (tile % per_group) % gsz;
  const int tn = (tile % per_group) / gsz;
  const int m0 = tm * BM, n0 = tn * BN;

  __shared__ __attribute__((aligned(16))) unsigned short As[BM * LDH];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[BN * LDH];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
  const int kq = (lane >> 4) * 4, lr = lane & 15;
  /* the untransposed product: element r of block (i, j) is row rb(i) + r,
   * column cb(j) of C, so the C tile moves as 16-byte vectors, loaded behind
   * the first slice and added after the last MFMA */
  const int rb0 = m0 + wm + 4 * (lane >> 4), cb0 = n0 + wn + lr;
  const bool cvec = (g.ldc % 4 == 0) && (((uintptr_t)g.C % 16) == 0);

  LA la;
  LB lb;
  la.load(g.A, g.lda, m0, 0, g.M, g.K, g.vec_a, tid);
  lb.load(g.B, g.ldb, n0, 0, g.N, g.K, g.vec_b, tid);
  f4 cv[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = rb0 + i * 16, col = cb0 + j * 16;
      const float *cp = g.C + (size_t)col * g.ldc + row;
      if (col < g.N && row + 3 < g.M && cvec) {
        cv[i][j] = *reinterpret_cast<const f4 *>(cp);
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) cv[i][j][r] = (col < g.N && row + r < g.M) ? cp[r] : 0.0f;
      }
    }
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0, 0, 0, 0};

  for (int k0 = 0; k0 < g.K; k0 += KB) {
    if (k0 > 0) __syncthreads();
#pragma unroll
    for (int i = 0; i < LA::NV; ++i) {            /* 4 rows at one k */
      const int lin = tid + NT * i;
      const int k = lin / (BM / 4), r = (lin % (BM / 4)) * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) As[(r + e) * LDH + k] = Low<LP>::cvt(g.alpha * la.v[i][e]);
    }
#pragma unroll
    for (int i = 0; i < LB::NV; ++i) {            /* 4 k at one row */
      const int lin = tid + NT * i;
      const int r = lin / (KB / 4), k = (lin % (KB / 4)) * 4;
      uint2 w;
      w.x = (unsigned)Low<LP>::cvt(lb.v[i][0]) | ((unsigned)Low<LP>::cvt(lb.v[i][1]) << 16);
      w.y = (unsigned)Low<LP>::cvt(lb.v[i][2]) | ((unsigned)Low<LP>::cvt(lb.v[i][3]) << 16);
      *reinterpret_cast<uint2 *>(&Bs[r * LDH + k]) = w;
    }
    __syncthreads();
    if (k0 + KB < g.K) {
      la.load(g.A, g.lda, m0, k0 + KB, g.M, g.K, g.vec_a, tid);
      lb.load(g.B, g.ldb, n0, k0 + KB, g.N, g.K, g.vec_b, tid);
    }
#pragma unroll
    for (int ks = 0; ks < KB / 16; ++ks) {
      uint2 a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = *reinterpret_cast<const uint2 *>(&As[(wm + i * 16 + lr) * LDH + ks * 16 + kq]);
        b[i] = *reinterpret_cast<const uint2 *>(&Bs[(wn + i * 16 + lr) * LDH + ks * 16 + kq]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = Low<LP>::mma(a[i], b[j], acc[i][j]);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = rb0 + i * 16, col = cb0 + j * 16;
      float *cp = g.C + (size_t)col * g.ldc + row;
      const f4 v = acc[i][j] + cv[i][j];
      if (col < g.N && row + 3 < g.M && cvec) {
        *reinterpret_cast<f4 *>(cp) = v;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (col < g.N && row + r < g.M) cp[r] = v[r];
      }
    }
}

template <typename T>
hipError_t launch_gemm(int ta, int tb, int M, int N, int K, T alpha, const T *A,
                       long lda, const T *B, long ldb, T beta, T *C, long ldc,
                       int tri, const int *abort_flag, hipStream_t s) {
  if (M <= 0 || N <= 0) return hipSuccess;
  if (ta != KX_NOTRANS && ta != KX_TRANS) return hipErrorInvalidValue;
  if (tb != KX_NOTRANS && tb != KX_TRANS) return hipErrorInvalidValue;
  if (K < 0 || ldc < M) return hipErrorInvalidValue;
  if (K > 0 && (!A || !B)) return hipErrorInvalidValue;
  if (K > 0 && lda < (ta == KX_NOTRANS ? M : K)) return hipErrorInvalidValue;
  if (K > 0 && ldb < (tb == KX_NOTRANS ? K : N)) return hipErrorInvalidValue;
  GemmArgs<T> g;
  g.M = M; g.N = N; g.K = K; g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.tri = tri; g.abort_flag = abort_flag;
  g.tiles_m = (M + BM - 1) / BM;
  g.tiles_n = (N + BN - 1) / BN;
  const long ntiles = (long)g.tiles_m * g.tiles_n;
  if (ntiles > (1L << 30)) return hipErrorInvalidValue;
  g.per_xcd = (int)((ntiles + 7) / 8);
  constexpr int VEC = Mfma<T>::VEC;
  g.vec_a = (((uintptr_t)A % 16) == 0) && (lda % VEC == 0);
  g.vec_b = (((uintptr_t)B % 16) == 0) && (ldb % VEC == 0);
  const dim3 grid((unsigned)(8 * g.per_xcd)), block(NT);
  if (tri >= KX_KBAND_AL && tri <= KX_KBAND_RBU) {
    if (ta == KX_NOTRANS && tb == KX_NOTRANS)
      hipLaunchKernelGGL((gemm_kernel<T, KX_NOTRANS, KX_NOTRANS, true>), grid, block, 0, s, g);
    else if (ta == KX_NOTRANS)
      hipLaunchKernelGGL((gemm_kernel<T, KX_NOTRANS, KX_TRANS, true>), grid, block, 0, s, g);
    else if (tb == KX_NOTRANS)
      hipLaunchKernelGGL((gemm_kernel<T, KX_TRANS, KX_NOTRANS, true>), grid, block, 0, s, g);
    else
      hipLaunchKernelGGL((gemm_kernel<T, KX_TRANS, KX_TRANS, true>), grid, block, 0, s, g);
  } else if (ta == KX_NOTRANS && tb == KX_NOTRANS)
    hipLaunchKernelGGL((gemm_kernel<T, KX_NOTRANS, KX_NOTRANS>), grid, block, 0, s, g);
  else if (ta == KX_NOTRANS)
    hipLaunchKernelGGL((gemm_kernel<T, KX_NOTRANS, KX_TRANS>), grid, block, 0, s, g);
  else if (tb == KX_NOTRANS)
    hipLaunchKernelGGL((gemm_kernel<T, KX_TRANS, KX_NOTRANS>), grid, block, 0, s, g);
  else
    hipLaunchKernelGGL((gemm_kernel<T, KX_TRANS, KX_TRANS>), grid, block, 0, s, g);
  return hipGetLastError();
}

}  // namespace

/* C += alpha A B, A and B not transposed, FP32, any K (staged 64 at a time) */
extern "C" hipError_t atlmx_k_supdate(int M, int N, int K, float alpha, const float *A,
                                      long lda, const float *B, long ldb, float *C, long ldc,
                                      const int *abort_flag, hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  if (lda < M || ldb < K || ldc < M || !A || !B || !C) return hipErrorInvalidValue;
  GemmArgs<float> g;
  g.M = M; g.N = N; g.K = K; g.alpha = alpha; g.beta = 1.0f;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.tri = 0; g.abort_flag = abort_flag;
  g.tiles_m = (M + BM - 1) / BM;
  g.tiles_n = (N + BN - 1) / BN;
  const long ntiles = (long)g.tiles_m * g.tiles_n;
  if (ntiles > (1L << 30)) return hipErrorInvalidValue;
  g.per_xcd = (int)((ntiles + 7) / 8);
  g.vec_a = (((uintptr_t)A % 16) == 0) && (lda % 4 == 0);
  g.vec_b = (((uintptr_t)B % 16) == 0) && (ldb % 4 == 0);
  hipLaunchKernelGGL((update_kernel<float, 32>), dim3((unsigned)(8 * g.per_xcd)), dim3(NT), 0, s,
                     g);
  return hipGetLastError();
}

/* C += alpha A op(B) in FP64 with the C tile prefetched (update_d_kernel);
 * tri as atlmx_k_dgemm. */
extern "C" hipError_t atlmx_k_dupdate(int tb, int M, int N, int K, double alpha,
                                      const double *A, long lda, const double *B, long ldb,
                                      double *C, long ldc, int tri, const int *abort_flag,
                                      hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  if ((tb != KX_NOTRANS && tb != KX_TRANS) || lda < M || ldb < (tb == KX_NOTRANS ? K : N) ||
      ldc < M || !A || !B || !C)
    return hipErrorInvalidValue;
  GemmArgs<double> g;
  g.M = M; g.N = N; g.K = K; g.alpha = alpha; g.beta = 1.0;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.tri = tri; g.abort_flag = abort_flag;
  g.tiles_m = (M + BM - 1) / BM;
  g.tiles_n = (N + BN - 1) / BN;
  const long ntiles = (long)g.tiles_m * g.tiles_n;
  if (ntiles > (1L << 30)) return hipErrorInvalidValue;
  g.per_xcd = (int)((ntiles + 7) / 8);
  g.vec_a = (((uintptr_t)A % 16) == 0) && (lda % 2 == 0);
  g.vec_b = (((uintptr_t)B % 16) == 0) && (ldb % 2 == 0);
  const dim3 grid((unsigned)(8 * g.per_xcd)), block(NT);
  if (tb == KX_NOTRANS)
    hipLaunchKernelGGL(update_d_kernel<KX_NOTRANS>, grid, block, 0, s, g);
  else
    hipLaunchKernelGGL(update_d_kernel<KX_TRANS>, grid, block, 0, s, g);
  return hipGetLastError();
}

/* the same with the operands rounded to bfloat16 (lp = 1) or half (lp = 2) */
extern "C" hipError_t atlmx_k_supdate_lp(int lp, int M, int N, int K, float alpha,
                                         const float *A, long lda, const float *B, long ldb,
                                         float *C, long ldc, const int *abort_flag,
                                         hipStream_t s) {
  if (M <= 0 || N <= 0 || K <= 0) return hipSuccess;
  if ((lp != 1 && lp != 2) || lda < M || ldb < K || ldc < M || !A || !B || !C)
    return hipErrorInvalidValue;
  GemmArgs<float> g;
  g.M = M; g.N = N; g.K = K; g.alpha = alpha; g.beta = 1.0f;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.tri = 0; g.abort_flag = abort_flag;
  g.tiles_m = (M + BM - 1) / BM;
  g.tiles_n = (N + BN - 1) / BN;
  const long ntiles = (long)g.tiles_m * g.tiles_n;
  if (ntiles > (1L << 30)) return hipErrorInvalidValue;
  g.per_xcd = (int)((ntiles + 7) / 8);
  g.vec_a = (((uintptr_t)A % 16) == 0) && (lda % 4 == 0);
  g.vec_b = (((uintptr_t)B % 16) == 0) && (ldb % 4 == 0);
  const dim3 grid((unsigned)(8 * g.per_xcd)), block(NT);
  if (lp == 1)
    hipLaunchKernelGGL(update_lp_kernel<1>, grid, block, 0, s, g);
  else
    hipLaunchKernelGGL(update_lp_kernel<2>, grid, block, 0, s, g);
  return hipGetLastError();
}

extern "C" hipError_t atlmx_k_dgemm(int ta, int tb, int M, int N, int K, double alpha,
                                    const double *A, long lda, const double *B, long ldb,
                                    double beta, double *C, long ldc, int tri,
                                    const int *abort_flag, hipStream_t s) {
  return launch_gemm<double>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tri,
                             abort_flag, s);
}

extern "C" hipError_t atlmx_k_sgemm(int ta, int tb, int M, int N, int K, float alpha,
                                    const float *A, long lda, const float *B, long ldb,
                                    float beta, float *C, long ldc, int tri,
                                    const int *abort_flag, hipStream_t s) {
  return launch_gemm<float>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tri,
                            abort_flag, s);
}

/*
 * kernels_mp.h -- launchers of the mixed-precision solver's memory-bound
 * kernels (k_mp.hip).  Device pointers, column-major, 64-bit leading
 * dimensions; each returns a hipError_t.
 */
#ifndef ATLMX_KERNELS_MP_H
#define ATLMX_KERNELS_MP_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MP_MAXRHS 16   /* right-hand sides per residual pass */
#define MP_KSPLIT 16   /* K slices of the residual GEMV      */

hipError_t atlmx_k_d2s(int M, int N, const double *A, long lda, float *S, long lds,
                       hipStream_t s);
hipError_t atlmx_k_s2d(int M, int N, const float *S, long lds, double *A, long lda,
                       hipStream_t s);
/* R = B - A*X (A: M x K, X: K x nrhs, nrhs <= MP_MAXRHS); ws of
 * atlmx_k_resid_ws_bytes(M, nrhs) bytes. */
size_t atlmx_k_resid_ws_bytes(int M, int nrhs);
hipError_t atlmx_k_dresid(int M, int K, int nrhs, const double *A, long lda, const double *X,
                          long ldx, const double *B, long ldb, double *R, long ldr, void *ws,
                          hipStream_t s);
/* out[c] = max_i |X(i,c)|, c < N */
hipError_t atlmx_k_dcolmax(int M, int N, const double *X, long ldx, double *out,
                           hipStream_t s);
/* rowsum[i] = sum_j |A(i,j)|; ws of atlmx_k_resid_ws_bytes(M, 1) bytes */
hipError_t atlmx_k_drowsum(int M, int N, const double *A, long lda, double *rowsum,
                           void *ws, hipStream_t s);
/* out[j] = sum_i V(i,j) w(i) for j < kc (V: M x kc), summed in a fixed
 * order; ws of atlmx_k_dots_ws_bytes(M, kc) bytes. */
size_t atlmx_k_dots_ws_bytes(int M, int kc);
hipError_t atlmx_k_ddots(int M, int kc, const double *V, long ldv, const double *w,
                         double *out, void *ws, hipStream_t s);
/* X += (double)D */
hipError_t atlmx_k_axpy_s2d(int M, int N, const float *D, long ldd, double *X, long ldx,
                            hipStream_t s);

#ifdef __cplusplus
}
#endif
#endif

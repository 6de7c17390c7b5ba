/*
 * wave_argmax.h -- (value, index) arg-max over one 64-lane wavefront with
 * DPP row shifts and row broadcasts (the gfx9-family reduction network),
 * result broadcast to every lane.  Larger value wins; equal values go to
 * the lower index (cblas_i?amax's first-maximum rule).  A lane with nothing
 * to offer passes value -2 and index INT_MAX.
 *
 * Six DPP steps replace the six ds_bpermute rounds of a __shfl_xor
 * butterfly, which cost ~100 cycles of LDS latency each.
 */
#ifndef ATLMX_WAVE_ARGMAX_H
#define ATLMX_WAVE_ARGMAX_H

#include <hip/hip_runtime.h>

namespace atlmx {

template <int CTRL, int RM>
__device__ __forceinline__ int dpp32(int old, int x) {
  return __builtin_amdgcn_update_dpp(old, x, CTRL, RM, 0xf, false);
}

__device__ __forceinline__ void argmax_combine(double &v, int &i, int &g, double v2, int i2,
                                               int g2) {
  if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; g = g2; }
}

template <int CTRL, int RM>
__device__ __forceinline__ void argmax_step(double &v, int &i, int &g) {
  const double ident = -2.0;
  const int lo = dpp32<CTRL, RM>(__double2loint(ident), __double2loint(v));
  const int hi = dpp32<CTRL, RM>(__double2hiint(ident), __double2hiint(v));
  const int i2 = dpp32<CTRL, RM>(0x7fffffff, i);
  const int g2 = dpp32<CTRL, RM>(0, g);
  argmax_combine(v, i, g, __hiloint2double(hi, lo), i2, g2);
}

/* v, i (and a payload g that travels with the winner) reduced over the
 * wavefront; every lane receives the result. */
__device__ __forceinline__ void wave_argmax(double &v, int &i, int &g) {
  argmax_step<0x111, 0xf>(v, i, g);   /* row_shr:1 */
  argmax_step<0x112, 0xf>(v, i, g);   /* row_shr:2 */
  argmax_step<0x114, 0xf>(v, i, g);   /* row_shr:4 */
  argmax_step<0x118, 0xf>(v, i, g);   /* row_shr:8 */
  argmax_step<0x142, 0xa>(v, i, g);   /* row_bcast:15 -> rows 1, 3 */
  argmax_step<0x143, 0xc>(v, i, g);   /* row_bcast:31 -> rows 2, 3 */
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  v = __hiloint2double(hi, lo);
  i = __builtin_amdgcn_readlane(i, 63);
  g = __builtin_amdgcn_readlane(g, 63);
}

}  // namespace atlmx

#endif

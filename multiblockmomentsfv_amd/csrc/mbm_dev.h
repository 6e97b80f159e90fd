/*
 * mbm_dev.h -- device helpers shared by the stage and adjoint kernels:
 * wavefront lane shifts (DPP), wave reductions, and raw buffer access to one
 * block of a layer with scalar component offsets.
 */
#ifndef MBM_DEV_H
#define MBM_DEV_H

#include <hip/hip_runtime.h>

#include "mbm_phys.h"

namespace mbm {

constexpr int NVC = 9;

/* wavefront-wide DPP shifts on the two 32-bit halves of a double:
 * wave_shr:1 (lane i <- lane i-1), wave_shl:1 (lane i <- lane i+1); the lane
 * without a source reads 0 (bound_ctrl), a halo lane the stencil never uses.
 * Without an 'old' operand no copy of the destination is needed first. */
template <int CTRL>
MBM_DEV double dpp_shift(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = (int)b, hi = (int)(b >> 32);
  const int l2 = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xf, 0xf, true);
  const int h2 = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)h2 << 32) | (unsigned int)l2);
}
MBM_DEV double from_left(double v) { return dpp_shift<0x138>(v); }
MBM_DEV double from_right(double v) { return dpp_shift<0x130>(v); }

MBM_DEV double wave_max(double v)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

MBM_DEV void fold_max(double &m, double v) { m = v > m ? v : m; }

MBM_DEV double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

/* Loads the 9 components of one cell (raw, no arithmetic). */
MBM_DEV void load9(const double *__restrict__ p, long long plane, double *U)
{
#pragma unroll
  for (int m = 0; m < NVC; m++) U[m] = p[m * plane];
}

/* Raw buffer access to one block of a layer (gfx9 buffer resource: base,
 * stride 0, byte extent of `elems` doubles, DATA_FORMAT=32 in the last word).
 * The component offset m*plane*8 goes in the scalar offset and the lane's
 * byte offset in one 32-bit VGPR, so nine loads share one address register;
 * accesses past the extent read 0 / are dropped by the hardware. */
MBM_DEV __amdgpu_buffer_rsrc_t block_rsrc(const double *base, long long elems)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(base), (short)0,
                                           (int)(elems * 8), 0x00020000);
}

MBM_DEV void load9b(__amdgpu_buffer_rsrc_t r, int pbytes, unsigned boff, double *U)
{
#pragma unroll
  for (int m = 0; m < NVC; m++)
    U[m] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, boff, m * pbytes, 0));
}

MBM_DEV void store1b(__amdgpu_buffer_rsrc_t r, int soff, unsigned boff, double v)
{
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, boff, soff, 0);
}

} // namespace mbm

#endif

// 64-lane wavefront primitives for gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace drx
{

__device__ __forceinline__ int lane_id () { return __lane_id (); }

__device__ __forceinline__ int32_t wave_reduce_add (int32_t v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor (v, o, 64);
  return v;
}

__device__ __forceinline__ float wave_reduce_add (float v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor (v, o, 64);
  return v;
}

// DPP move of v within the wavefront; rows not in ROW_MASK keep `old`.
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_move (float old, float v)
{
  return __int_as_float (__builtin_amdgcn_update_dpp (__float_as_int (old), __float_as_int (v), CTRL, ROW_MASK, 0xF, false));
}

// Min over the 64 lanes with DPP steps (quad swaps, half-row and row mirrors, row broadcasts): the total
// ends in lane 63 and is read back to every lane.  min/max are exact, so the order does not matter.
__device__ __forceinline__ float wave_reduce_min (float v)
{
  v = fminf (v, dpp_move<0xB1> (v, v));        // quad_perm [1,0,3,2]
  v = fminf (v, dpp_move<0x4E> (v, v));        // quad_perm [2,3,0,1]
  v = fminf (v, dpp_move<0x141> (v, v));       // row_half_mirror
  v = fminf (v, dpp_move<0x140> (v, v));       // row_mirror
  v = fminf (v, dpp_move<0x142, 0xA> (v, v));  // row_bcast15 into rows 1, 3
  v = fminf (v, dpp_move<0x143, 0xC> (v, v));  // row_bcast31 into rows 2, 3
  return __int_as_float (__builtin_amdgcn_readlane (__float_as_int (v), 63));
}

__device__ __forceinline__ float wave_reduce_max (float v)
{
  v = fmaxf (v, dpp_move<0xB1> (v, v));
  v = fmaxf (v, dpp_move<0x4E> (v, v));
  v = fmaxf (v, dpp_move<0x141> (v, v));
  v = fmaxf (v, dpp_move<0x140> (v, v));
  v = fmaxf (v, dpp_move<0x142, 0xA> (v, v));
  v = fmaxf (v, dpp_move<0x143, 0xC> (v, v));
  return __int_as_float (__builtin_amdgcn_readlane (__float_as_int (v), 63));
}

// Order-preserving integer key of a float (any non-NaN value): the sign bit kept, the magnitude bits of a
// negative value flipped, so signed integer order is float order (-0 below +0).  Its own inverse.
__device__ __forceinline__ int32_t float_key (float f)
{
  const int32_t b = __float_as_int (f);
  return b ^ ((b >> 31) & 0x7FFFFFFF);
}
__device__ __forceinline__ float key_float (int32_t k) { return __int_as_float (k ^ ((k >> 31) & 0x7FFFFFFF)); }

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int32_t dpp_move_i (int32_t old, int32_t v)
{
  return __builtin_amdgcn_update_dpp (old, v, CTRL, ROW_MASK, 0xF, false);
}

// DPP permutation of v within each row where every lane has a source lane (quad_perm, mirrors): no `old`
// value is involved, so the move can fold into the instruction that consumes it.
template <int CTRL>
__device__ __forceinline__ int32_t dpp_perm_i (int32_t v)
{
  return __builtin_amdgcn_mov_dpp (v, CTRL, 0xF, 0xF, true);
}

// Min / max of integer keys within each 16-lane row (quad swaps, half-row and row mirrors): every lane of
// a row ends holding its row's result.  Integer min/max need no NaN canonicalisation, so each step can be
// one DPP-modified instruction.
__device__ __forceinline__ int32_t row_reduce_min (int32_t v)
{
  v = min (v, dpp_perm_i<0xB1> (v));
  v = min (v, dpp_perm_i<0x4E> (v));
  v = min (v, dpp_perm_i<0x141> (v));
  v = min (v, dpp_perm_i<0x140> (v));
  return v;
}

__device__ __forceinline__ int32_t row_reduce_max (int32_t v)
{
  v = max (v, dpp_perm_i<0xB1> (v));
  v = max (v, dpp_perm_i<0x4E> (v));
  v = max (v, dpp_perm_i<0x141> (v));
  v = max (v, dpp_perm_i<0x140> (v));
  return v;
}

// Float sum within each 16-lane row by the same four DPP steps: quad sums, then the two quads of each half
// row, then the two halves.  Each step adds values that are equal across the lanes being combined, and
// addition commutes, so every lane of a row ends with the bit-identical sum.
__device__ __forceinline__ float row_reduce_add (float v)
{
  v += __int_as_float (dpp_perm_i<0xB1> (__float_as_int (v)));
  v += __int_as_float (dpp_perm_i<0x4E> (__float_as_int (v)));
  v += __int_as_float (dpp_perm_i<0x141> (__float_as_int (v)));
  v += __int_as_float (dpp_perm_i<0x140> (__float_as_int (v)));
  return v;
}

// Min of integer keys over the 64 lanes, read back to every lane.
__device__ __forceinline__ int32_t wave_reduce_min (int32_t v)
{
  v = row_reduce_min (v);
  v = min (v, dpp_move_i<0x142, 0xA> (v, v));   // row_bcast15 into rows 1, 3
  v = min (v, dpp_move_i<0x143, 0xC> (v, v));   // row_bcast31 into rows 2, 3
  return __builtin_amdgcn_readlane (v, 63);
}

// Max of integer keys over the 64 lanes, read back to every lane.
__device__ __forceinline__ int32_t wave_reduce_max (int32_t v)
{
  v = row_reduce_max (v);
  v = max (v, dpp_move_i<0x142, 0xA> (v, v));   // row_bcast15 into rows 1, 3
  v = max (v, dpp_move_i<0x143, 0xC> (v, v));   // row_bcast31 into rows 2, 3
  return __builtin_amdgcn_readlane (v, 63);
}

// Inclusive prefix sum across the 64 lanes.
__device__ __forceinline__ int32_t wave_scan_incl_add (int32_t v)
{
  const int l = __lane_id ();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1)
  {
    const int32_t t = __shfl_up (v, o, 64);
    if (l >= o) v += t;
  }
  return v;
}

// Bitmask of lanes below this one.
__device__ __forceinline__ uint64_t lanemask_lt () { return (__lane_id () == 0) ? 0ull : (~0ull >> (64 - __lane_id ())); }

} // namespace drx

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

__device__ __forceinline__ float wave_reduce_min (float v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf (v, __shfl_xor (v, o, 64));
  return v;
}

__device__ __forceinline__ float wave_reduce_max (float v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf (v, __shfl_xor (v, o, 64));
  return v;
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

// Learned Newton initializer: a 27 -> 64 -> 64 -> 64 -> 64 -> 3 MLP (four hidden layers of width 64, GELU)
// with BF16 weights and activations and FP32 accumulation on gfx950 MFMA (v_mfma_f32_16x16x32_bf16).
//
// Input (27): the query point and the element's 8 corner control points, each expressed in the frame
// of the corner box (centre 0, half-extent 1 per axis).  Output (3): a correction to the affine guess
// base = (q' + 1) / 2 of the reference coordinates; the initial Newton guess is base + output.
//
// MFMA operand layouts (16x16x32, one wavefront, lane l, j = l % 16, g = l / 16):
//   A: lane holds row j, 8 consecutive k starting at 8g;   B: lane holds column j, the same 8 k;
//   C: lane holds rows 4g .. 4g+3 of column j.
// The forward pass keeps activations in registers between layers: the C tile of a layer is used
// directly as the B operand of the next one, with that layer's weights stored in the matching
// permuted k order (slot s of lane group g holds feature 4g+s for s < 4, 16+4g+s-4 otherwise, within
// each 32-wide k block).  Samples are the N (column) dimension: 64 samples = 4 column blocks.
#pragma once

#include "drx_common.hpp"

#ifndef DRX_MLP_SB
#define DRX_MLP_SB 4        // 16-sample column blocks per pass of the batched inference kernel (4: 15.7 G/s,
                            // 2: 13.4, 1: 9.2 G rows/s with the four-operation GELU)
#endif
#ifndef DRX_MLP_WAVE_SB
#define DRX_MLP_WAVE_SB 2   // the same inside point location (learned locate 1.11 G/s vs 1.08 at 1 and
                            // 1.09 at 4; learned render 19.8 ms vs 22.8 at 1)
#endif

namespace drx
{

typedef __bf16 bf16x8 __attribute__ ((ext_vector_type (8)));
typedef __bf16 bf16x4 __attribute__ ((ext_vector_type (4)));
typedef float floatx4 __attribute__ ((ext_vector_type (4)));

namespace mlp
{
constexpr int IN = 27, IN_PAD = 32, H = 64, OUT = 3, OUT_PAD = 16;
// fp32 master parameters: W_l row-major [out][in_pad], then b_l
constexpr int W1 = 0, B1 = W1 + H * IN_PAD, W2 = B1 + H, B2 = W2 + H * H, W3 = B2 + H, B3 = W3 + H * H,
              W4 = B3 + H, B4 = W4 + H * H, W5 = B4 + H, B5 = W5 + OUT_PAD * H, NPARAM = B5 + OUT_PAD;
// bf16 forward copies, same shapes as the master weights
constexpr int F1 = 0, F2 = F1 + H * IN_PAD, F3 = F2 + H * H, F4 = F3 + H * H, F5 = F4 + H * H,
              NFWD = F5 + OUT_PAD * H;
// bf16 transposed copies of layers 2..5 for the backward pass: [in 64][out_pad], out_pad 64 or 32 (layer 5)
constexpr int T2 = 0, T3 = T2 + H * H, T4 = T3 + H * H, T5 = T4 + H * H, NTRN = T5 + H * 32;
// inference fragments (64 lanes x 8 bf16): layer 1 ob 0..3; layers 2..4 (ob, kb) = 2 ob + kb; layer 5 kb
constexpr int FR1 = 0, FR2 = FR1 + 4, FR3 = FR2 + 8, FR4 = FR3 + 8, FR5 = FR4 + 8, NFRAG = FR5 + 2;
constexpr int NBIAS = 4 * H + OUT_PAD;
constexpr int64_t PACKED_BYTES = (int64_t)NFRAG * 64 * 16 + NBIAS * 4;
// training state blob: params | adam m | adam v (fp32, NPARAM_PAD each) | fwd bf16 | transposed bf16
constexpr int NPARAM_PAD = (NPARAM + 63) & ~63;
constexpr int64_t STATE_BYTES = 3ll * NPARAM_PAD * 4 + (int64_t)NFWD * 2 + (int64_t)NTRN * 2;
constexpr int PSTRIDE = NPARAM_PAD + 64;   // per-block gradient partial row (one block = 64 training rows) (+ loss at NPARAM)

__host__ __device__ constexpr int in_pad (int l) { return l == 1 ? IN_PAD : H; }
__host__ __device__ constexpr int out_pad (int l) { return l == 5 ? OUT_PAD : H; }
__host__ __device__ constexpr int w_off (int l) { return l == 1 ? W1 : l == 2 ? W2 : l == 3 ? W3 : l == 4 ? W4 : W5; }
__host__ __device__ constexpr int b_off (int l) { return l == 1 ? B1 : l == 2 ? B2 : l == 3 ? B3 : l == 4 ? B4 : B5; }
__host__ __device__ constexpr int f_off (int l) { return l == 1 ? F1 : l == 2 ? F2 : l == 3 ? F3 : l == 4 ? F4 : F5; }
__host__ __device__ constexpr int t_off (int l) { return l == 2 ? T2 : l == 3 ? T3 : l == 4 ? T4 : T5; }
__host__ __device__ constexpr int t_pad (int l) { return l == 5 ? 32 : H; }
} // namespace mlp

// GELU x * Phi(x) with the normal CDF Phi replaced by a cubic smoothstep over [-2, 2]:
//   Phi~(x) = 1/2 + u (3/4 - u^2/4),  u = clamp(x / 2, -1, 1)   (|Phi~ - Phi| < 0.029, C1 at +-2)
// Six full-rate VALU operations and no transcendental; training uses the same function.
__device__ __forceinline__ float gelu_cdf (float x, float &u)
{
  u = __builtin_amdgcn_fmed3f (0.5f * x, -1.f, 1.f);
  return fmaf (u, fmaf (-0.25f * u, u, 0.75f), 0.5f);
}
__device__ __forceinline__ float gelu (float x)
{
  float u;
  return x * gelu_cdf (x, u);
}
// d/dx [x Phi~(x)] = Phi~(x) + x Phi~'(x), Phi~'(x) = 3/8 (1 - u^2) inside [-2, 2], 0 outside
__device__ __forceinline__ float gelu_grad (float x)
{
  float u;
  const float h = gelu_cdf (x, u);
  return fabsf (x) < 2.f ? fmaf (x * 0.375f, fmaf (-u, u, 1.f), h) : h;
}

__device__ __forceinline__ floatx4 mfma_bf16 (bf16x8 a, bf16x8 b, floatx4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16 (a, b, c, 0, 0, 0);
}

struct MlpNet
{
  const bf16x8 *frag;   // NFRAG * 64
  const float *bias;    // NBIAS
};

__host__ __device__ inline MlpNet mlp_net (const void *packed)
{
  MlpNet n;
  n.frag = (const bf16x8 *)packed;
  n.bias = (const float *)((const char *)packed + (size_t)mlp::NFRAG * 64 * 16);
  return n;
}

// The same pointer, hidden from the optimiser: weight loads issued after it cannot be hoisted out of the
// loop that calls the MLP (which would keep all 30 fragments, 120 VGPRs, live across the loop).
template <typename T>
__device__ __forceinline__ const T *opaque_ptr (const T *p)
{
  uint64_t v = reinterpret_cast<uint64_t> (p);
  asm volatile ("" : "+s"(v));
  return reinterpret_cast<const T *> (v);
}

// Inference form of GELU: x Phi~(x) / 8 from zq = x / 4.  drx_mlp_pack scales the packed weights so that
// every hidden layer's MFMA yields a quarter of its pre-activation (layer-1 weights and hidden biases
// x 0.25) and the layer after takes the activation divided by 8 (hidden weights x 8 x 0.25 = x 2, output
// weights x 8).  With w = clamp(zq, -1/2, 1/2) = u / 2: 0.75 - u^2/4 = fma(-w, w, 0.75) and
// Phi~ / 2 = fma(w, t, 0.25).  Powers of two scale every product, sum and bf16 rounding exactly, so the
// outputs equal the unscaled network's: four VALU operations per value instead of six.
__device__ __forceinline__ float gelu_from_quarter (float zq)
{
  const float w = __builtin_amdgcn_fmed3f (zq, -0.5f, 0.5f);
  return zq * fmaf (w, fmaf (-w, w, 0.75f), 0.25f);
}

// GELU of two 16-row C tiles (one 32-wide k block of the next layer) packed into its B operand.
__device__ __forceinline__ bf16x8 gelu_pack (const floatx4 &lo, const floatx4 &hi)
{
  bf16x8 v;
#pragma unroll
  for (int r = 0; r < 4; ++r)
  {
    v[r] = (__bf16)gelu_from_quarter (lo[r]);
    v[4 + r] = (__bf16)gelu_from_quarter (hi[r]);
  }
  return v;
}

// Forward pass for SB column blocks of 16 samples.  bin[s] is the layer-1 B operand of block s (lane
// holds input features 8g .. 8g+7 of sample j); y[s] receives the output C tile (rows 0..2 in lanes 0..15).
// Must be called with the whole wavefront active.
template <int SB>
__device__ __forceinline__ void mlp_forward (const MlpNet &net_in, const bf16x8 (&bin)[SB], floatx4 (&y)[SB])
{
  using namespace mlp;
  const int lane = __lane_id (), g = lane >> 4;
  MlpNet net;
  net.frag = opaque_ptr (net_in.frag);
  net.bias = opaque_ptr (net_in.bias);
  bf16x8 b[SB][2];
  {
    floatx4 acc[SB][4];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
    {
      const bf16x8 a = net.frag[(FR1 + ob) * 64 + lane];
      const floatx4 bias = *(const floatx4 *)(net.bias + 16 * ob + 4 * g);
#pragma unroll
      for (int s = 0; s < SB; ++s) acc[s][ob] = mfma_bf16 (a, bin[s], bias);
    }
#pragma unroll
    for (int s = 0; s < SB; ++s)
    {
      b[s][0] = gelu_pack (acc[s][0], acc[s][1]);
      b[s][1] = gelu_pack (acc[s][2], acc[s][3]);
    }
  }
#pragma unroll
  for (int L = 2; L <= 4; ++L)
  {
    const int fr = (L == 2) ? FR2 : (L == 3) ? FR3 : FR4;
    floatx4 acc[SB][4];
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
    {
      const floatx4 bias = *(const floatx4 *)(net.bias + H * (L - 1) + 16 * ob + 4 * g);
      const bf16x8 a0 = net.frag[(fr + 2 * ob) * 64 + lane];
      const bf16x8 a1 = net.frag[(fr + 2 * ob + 1) * 64 + lane];
#pragma unroll
      for (int s = 0; s < SB; ++s) acc[s][ob] = mfma_bf16 (a1, b[s][1], mfma_bf16 (a0, b[s][0], bias));
    }
#pragma unroll
    for (int s = 0; s < SB; ++s)
    {
      b[s][0] = gelu_pack (acc[s][0], acc[s][1]);
      b[s][1] = gelu_pack (acc[s][2], acc[s][3]);
    }
  }
  {
    const floatx4 bias = *(const floatx4 *)(net.bias + 4 * H + 4 * g);
    const bf16x8 a0 = net.frag[(FR5 + 0) * 64 + lane];
    const bf16x8 a1 = net.frag[(FR5 + 1) * 64 + lane];
#pragma unroll
    for (int s = 0; s < SB; ++s) y[s] = mfma_bf16 (a1, b[s][1], mfma_bf16 (a0, b[s][0], bias));
  }
}

// Features of (query, element) from a packed geometry block g of order P (see drx_mesh_pack): the
// query relative to the block origin and the 8 corner control points, in the corner-box frame.
// f[27..31] = 0.  base = (q' + 1) / 2.
template <int P>
__device__ __forceinline__ void mlp_features (const float *__restrict__ g, const float pt[3], float f[mlp::IN_PAD],
                                              float base[3])
{
  constexpr int N1 = P + 1, NP = ((N1 * N1 * N1) + 3) & ~3;
  float c[8][3];
  float lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
#pragma unroll
  for (int k = 0; k < 8; ++k)
  {
    const int qi = ((k & 1) ? P : 0) + N1 * (((k & 2) ? P : 0) + N1 * ((k & 4) ? P : 0));
#pragma unroll
    for (int d = 0; d < 3; ++d)
    {
      c[k][d] = g[4 + d * NP + qi];
      lo[d] = fminf (lo[d], c[k][d]);
      hi[d] = fmaxf (hi[d], c[k][d]);
    }
  }
  float ctr[3], inv[3];
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    ctr[d] = 0.5f * (lo[d] + hi[d]);
    const float half = 0.5f * (hi[d] - lo[d]);
    inv[d] = half > 1e-30f ? 1.f / half : 0.f;
    f[d] = ((pt[d] - g[d]) - ctr[d]) * inv[d];
    base[d] = 0.5f * f[d] + 0.5f;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) f[3 + 3 * k + d] = (c[k][d] - ctr[d]) * inv[d];
#pragma unroll
  for (int k = mlp::IN; k < mlp::IN_PAD; ++k) f[k] = 0.f;
}

__device__ __forceinline__ void mlp_features_mixed (int order, const float *__restrict__ g, const float pt[3],
                                                    float f[mlp::IN_PAD], float base[3])
{
  switch (order)
  {
  case 1: mlp_features<1> (g, pt, f, base); break;
  case 2: mlp_features<2> (g, pt, f, base); break;
  case 3: mlp_features<3> (g, pt, f, base); break;
  default: mlp_features<4> (g, pt, f, base); break;
  }
}

// Orders LDS traffic between the lanes of one wavefront (no s_barrier: the wave runs in lockstep).
__device__ __forceinline__ void wave_lds_sync ()
{
  __builtin_amdgcn_fence (__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier ();
  __builtin_amdgcn_fence (__ATOMIC_ACQUIRE, "wavefront");
}

// Wavefront-wide inference for 64 lanes, each with its own feature row f; lanes without work (has ==
// false) get no output, and 16-lane column blocks without any work are skipped.  lx: 64 x 32 bf16 and
// ly: 64 x 4 floats of this wave's LDS.  Returns the correction in out.
__device__ __forceinline__ void mlp_wave_infer (const MlpNet &net, const float f[mlp::IN_PAD], bool has, float out[3],
                                                __bf16 *lx, float *ly)
{
  const int lane = __lane_id (), j = lane & 15, g = lane >> 4;
  const uint64_t work = __ballot (has);
  bf16x8 *row = (bf16x8 *)(lx + lane * mlp::IN_PAD);
#pragma unroll
  for (int v = 0; v < 4; ++v)
  {
    bf16x8 t;
#pragma unroll
    for (int s = 0; s < 8; ++s) t[s] = (__bf16)f[8 * v + s];
    row[v] = t;
  }
  wave_lds_sync ();
  constexpr int SB = DRX_MLP_WAVE_SB;
#pragma unroll 1
  for (int sb0 = 0; sb0 < 4; sb0 += SB)
  {
    if (((work >> (16 * sb0)) & ((SB == 1) ? 0xFFFFull : ~0ull >> (64 - 16 * SB))) == 0) continue;
    bf16x8 bin[SB];
    floatx4 y[SB];
#pragma unroll
    for (int s = 0; s < SB; ++s) bin[s] = *(const bf16x8 *)(lx + (16 * (sb0 + s) + j) * mlp::IN_PAD + 8 * g);
    mlp_forward<SB> (net, bin, y);
    if (g == 0)
#pragma unroll
      for (int s = 0; s < SB; ++s) *(floatx4 *)(ly + 4 * (16 * (sb0 + s) + j)) = y[s];
  }
  wave_lds_sync ();
  const floatx4 r = *(const floatx4 *)(ly + 4 * lane);
  out[0] = r[0];
  out[1] = r[1];
  out[2] = r[2];
  wave_lds_sync ();   // the next call may overwrite lx / ly
}

} // namespace drx

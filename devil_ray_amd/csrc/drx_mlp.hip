// Learned Newton initializer: feature extraction, hand-written MFMA training (forward, backward, Adam)
// and batched inference.  Network layout and operand conventions are described in drx_mlp.hpp.
#include "drx_common.hpp"
#include "drx_element.hpp"
#include "drx_mlp.hpp"
#include "drx_wave.hpp"

#include <cmath>

namespace drx
{
using namespace mlp;

struct MlpState
{
  float *params, *m, *v;
  __bf16 *fw, *tw;
};

static MlpState state_view (void *blob)
{
  MlpState s;
  char *p = (char *)blob;
  s.params = (float *)p;
  s.m = s.params + NPARAM_PAD;
  s.v = s.m + NPARAM_PAD;
  s.fw = (__bf16 *)(s.v + NPARAM_PAD);
  s.tw = s.fw + NFWD;
  return s;
}

__device__ __forceinline__ uint64_t mix64 (uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// (layer, row, col) of master parameter p; col < 0 for a bias.  Returns the layer (1..5).
__device__ __forceinline__ int param_coords (int p, int &row, int &col)
{
  for (int l = 5; l >= 1; --l)
  {
    if (p >= b_off (l))
    {
      row = p - b_off (l);
      col = -1;
      return l;
    }
    if (p >= w_off (l))
    {
      row = (p - w_off (l)) / in_pad (l);
      col = (p - w_off (l)) % in_pad (l);
      return l;
    }
  }
  row = col = 0;
  return 0;
}

__device__ __forceinline__ bool param_live (int l, int row, int col)
{
  if (col < 0) return l != 5 || row < OUT;
  if (l == 1 && col >= IN) return false;
  if (l == 5 && row >= OUT) return false;
  return true;
}

__device__ __forceinline__ void write_copies (const MlpState &s, int l, int row, int col, float w)
{
  if (col < 0) return;
  s.fw[f_off (l) + row * in_pad (l) + col] = (__bf16)w;
  if (l >= 2) s.tw[t_off (l) + col * t_pad (l) + row] = (__bf16)w;
}

// Uniform init with variance 1/fan_in (last layer scaled by 0.1), zero biases, zero Adam moments.
__global__ void k_mlp_init (MlpState s, uint64_t seed)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < NTRN) s.tw[p] = (__bf16)0.f;
  if (p >= NPARAM_PAD) return;
  s.m[p] = 0.f;
  s.v[p] = 0.f;
  if (p >= NPARAM)
  {
    s.params[p] = 0.f;
    return;
  }
  int row, col;
  const int l = param_coords (p, row, col);
  float w = 0.f;
  if (col >= 0 && param_live (l, row, col))
  {
    const int fan_in = (l == 1) ? IN : H;
    const float a = sqrtf (3.f / (float)fan_in) * (l == 5 ? 0.1f : 1.f);
    const uint64_t h = mix64 (seed * 0x100000001B3ull + (uint64_t)p);
    const float u = (float)(h >> 40) * (1.f / 16777216.f);   // [0, 1)
    w = (2.f * u - 1.f) * a;
  }
  s.params[p] = w;
  if (col >= 0) s.fw[f_off (l) + row * in_pad (l) + col] = (__bf16)w;
}

// Transposed copies need the forward ones written first (separate launch keeps it simple).
__global__ void k_mlp_init_t (MlpState s)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  int row, col;
  const int l = param_coords (p, row, col);
  if (col >= 0 && l >= 2) s.tw[t_off (l) + col * t_pad (l) + row] = (__bf16)s.params[p];
}

// Features + training target for (point, located cell) pairs.  Unlocated rows get zeros.
template <int P>
__global__ void __launch_bounds__ (256) k_mlp_features (const float *__restrict__ mesh, const int64_t *__restrict__ offs,
                                                        const int32_t *__restrict__ orders, int stride,
                                                        const float *__restrict__ pts, const Location *__restrict__ loc,
                                                        int64_t n, bf16x8 *__restrict__ out_x, float *__restrict__ out_t,
                                                        float *__restrict__ out_base)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Location L = loc[i];
  float f[IN_PAD], base[3] = { 0.f, 0.f, 0.f };
  if (L.cell >= 0)
  {
    const float pt[3] = { pts[3 * i], pts[3 * i + 1], pts[3 * i + 2] };
    if constexpr (P == 0)
      mlp_features_mixed (orders[L.cell], mesh + offs[L.cell], pt, f, base);
    else
      mlp_features<P> (mesh + (int64_t)L.cell * stride, pt, f, base);
  }
  else
  {
#pragma unroll
    for (int k = 0; k < IN_PAD; ++k) f[k] = 0.f;
  }
#pragma unroll
  for (int v = 0; v < 4; ++v)
  {
    bf16x8 t;
#pragma unroll
    for (int s = 0; s < 8; ++s) t[s] = (__bf16)f[8 * v + s];
    out_x[4 * i + v] = t;
  }
  if (out_t)
#pragma unroll
    for (int d = 0; d < 3; ++d) out_t[3 * i + d] = (L.cell >= 0) ? L.ref[d] - base[d] : 0.f;
  if (out_base)
#pragma unroll
    for (int d = 0; d < 3; ++d) out_base[3 * i + d] = base[d];
}

// Sum over the 16 lanes that share a lane group (lanes j = 0..15 of group g, one DPP row); every lane of
// the group receives it.
__device__ __forceinline__ float group16_sum (float v) { return row_reduce_add (v); }

// One training step's forward + backward for 64 samples, four wavefronts per block.  Wave w owns sample
// block w (16 samples) in the forward pass, the loss and the deltas, and the weight-gradient tiles of
// input block w (K = all 64 samples, read from the block's feature-major LDS copies), so each gradient
// entry is written by exactly one wave.  Bias gradients and the loss are summed over the four waves in
// wave order (deterministic).  Writes the block's gradient (and squared-error sum at [NPARAM]) to
// partials[blockIdx.x * PSTRIDE ...].
constexpr int kTrainRed = 4 * H + OUT_PAD + 1;   // bias gradients of layers 1-4, layer 5, loss
__global__ void __launch_bounds__ (256) k_mlp_train (const bf16x8 *__restrict__ x, const float *__restrict__ tgt,
                                                     int64_t n, int64_t start, const float *__restrict__ params,
                                                     const __bf16 *__restrict__ fw, const __bf16 *__restrict__ tw,
                                                     float *__restrict__ partials, float grad_scale)
{
  __shared__ __attribute__ ((aligned (16))) __bf16 xs[64 * IN_PAD];   // sample-major inputs
  __shared__ __attribute__ ((aligned (16))) __bf16 xf[IN_PAD * 64];   // feature-major inputs
  __shared__ __attribute__ ((aligned (16))) __bf16 hs[2][64 * H];     // sample-major activations (ping-pong)
  __shared__ __attribute__ ((aligned (16))) __bf16 hf[4][H * 64];     // feature-major activations H1..H4
  __shared__ __attribute__ ((aligned (16))) __bf16 gp[4][4][16 * 64]; // GELU' per wave and lane, [value][lane]
  __shared__ __attribute__ ((aligned (16))) __bf16 dsm[2][64 * H];    // sample-major deltas (ping-pong)
  __shared__ __attribute__ ((aligned (16))) __bf16 dfm[2][H * 64];    // feature-major deltas (ping-pong)
  __shared__ float red[4][kTrainRed];                                 // per-wave bias gradients + loss

  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, j = lane & 15, g = lane >> 4;
  const int sb = wv;   // this wave's sample block
  float *part = partials + (int64_t)blockIdx.x * PSTRIDE;
  const int64_t base_idx = start + (int64_t)blockIdx.x * 64;

  {
    const int smp = threadIdx.x >> 2, v = threadIdx.x & 3;
    const int64_t idx = (base_idx + smp) % n;
    const bf16x8 t = x[4 * idx + v];
    *(bf16x8 *)(xs + smp * IN_PAD + 8 * v) = t;
#pragma unroll
    for (int s = 0; s < 8; ++s) xf[(8 * v + s) * 64 + smp] = t[s];
  }
  __syncthreads ();

  // ---- forward, hidden layers (sample block sb) ----
#pragma unroll
  for (int l = 1; l <= 4; ++l)
  {
    const int ip = in_pad (l);
    const __bf16 *in_s = (l == 1) ? xs : hs[(l - 2) & 1];
    __bf16 *out_s = hs[(l - 1) & 1];
    const __bf16 *W = fw + f_off (l);
#pragma unroll
    for (int ob = 0; ob < 4; ++ob)
    {
      floatx4 acc = *(const floatx4 *)(params + b_off (l) + 16 * ob + 4 * g);
#pragma unroll
      for (int kb = 0; kb < ip / 32; ++kb)
      {
        const bf16x8 a = *(const bf16x8 *)(W + (16 * ob + j) * ip + 32 * kb + 8 * g);
        const bf16x8 b = *(const bf16x8 *)(in_s + (16 * sb + j) * ip + 32 * kb + 8 * g);
        acc = mfma_bf16 (a, b, acc);
      }
      bf16x4 hv;
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        const float z = acc[r];
        hv[r] = (__bf16)gelu (z);
        hf[l - 1][(16 * ob + 4 * g + r) * 64 + 16 * sb + j] = hv[r];
        gp[l - 1][wv][(4 * ob + r) * 64 + lane] = (__bf16)gelu_grad (z);
      }
      *(bf16x4 *)(out_s + (16 * sb + j) * H + 16 * ob + 4 * g) = hv;
    }
    __syncthreads ();
  }

  // ---- output layer + loss gradient ----
  {
    const __bf16 *W = fw + F5;
    const bf16x4 zero4 = { (__bf16)0.f, (__bf16)0.f, (__bf16)0.f, (__bf16)0.f };
    floatx4 acc = *(const floatx4 *)(params + B5 + 4 * g);
#pragma unroll
    for (int kb = 0; kb < 2; ++kb)
    {
      const bf16x8 a = *(const bf16x8 *)(W + j * H + 32 * kb + 8 * g);
      const bf16x8 b = *(const bf16x8 *)(hs[1] + (16 * sb + j) * H + 32 * kb + 8 * g);
      acc = mfma_bf16 (a, b, acc);
    }
    const int smp = 16 * sb + j;
    const int64_t idx = (base_idx + smp) % n;
    float loss = 0.f;
    bf16x4 dv;
#pragma unroll
    for (int r = 0; r < 4; ++r)
    {
      const int f = 4 * g + r;
      float d = 0.f;
      if (f < OUT)
      {
        const float e = acc[r] - tgt[3 * idx + f];
        loss += e * e;
        d = 2.f * e * grad_scale;
      }
      dv[r] = (__bf16)d;
      dfm[0][f * 64 + smp] = dv[r];
      const float t = group16_sum (d);
      if (j == 0) red[wv][4 * H + f] = t;
    }
    *(bf16x4 *)(dsm[0] + smp * H + 4 * g) = dv;
    *(bf16x4 *)(dsm[0] + smp * H + 16 + 4 * g) = zero4;   // k padding of the 16-wide output layer
    loss = wave_reduce_add (loss);
    if (lane == 0) red[wv][4 * H + OUT_PAD] = loss;
  }
  __syncthreads ();

  // ---- backward ----
  int cur = 0;
#pragma unroll
  for (int L = 5; L >= 2; --L)
  {
    // dW_L = delta_L . H_{L-1}^T   (K = 64 samples), input block ib = wv
    {
      const int ib = wv;
      const __bf16 *hT = hf[L - 2];
#pragma unroll
      for (int ob = 0; ob < out_pad (L) / 16; ++ob)
      {
        floatx4 acc = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int kb = 0; kb < 2; ++kb)
        {
          const bf16x8 a = *(const bf16x8 *)(dfm[cur] + (16 * ob + j) * 64 + 32 * kb + 8 * g);
          const bf16x8 b = *(const bf16x8 *)(hT + (16 * ib + j) * 64 + 32 * kb + 8 * g);
          acc = mfma_bf16 (a, b, acc);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) part[w_off (L) + (16 * ob + 4 * g + r) * H + 16 * ib + j] = acc[r];
      }
    }
    // delta_{L-1} = (W_L^T delta_L) * GELU'(z_{L-1}), sample block sb
    const int kp = t_pad (L);
    const __bf16 *WT = tw + t_off (L);
#pragma unroll
    for (int ib = 0; ib < 4; ++ib)
    {
      floatx4 acc = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
      for (int kb = 0; kb < kp / 32; ++kb)
      {
        const bf16x8 a = *(const bf16x8 *)(WT + (16 * ib + j) * kp + 32 * kb + 8 * g);
        const bf16x8 b = *(const bf16x8 *)(dsm[cur] + (16 * sb + j) * H + 32 * kb + 8 * g);
        acc = mfma_bf16 (a, b, acc);
      }
      bf16x4 dv;
#pragma unroll
      for (int r = 0; r < 4; ++r)
      {
        const float d = acc[r] * (float)gp[L - 2][wv][(4 * ib + r) * 64 + lane];
        dv[r] = (__bf16)d;
        dfm[cur ^ 1][(16 * ib + 4 * g + r) * 64 + 16 * sb + j] = dv[r];
        const float t = group16_sum (d);
        if (j == 0) red[wv][(L - 2) * H + 16 * ib + 4 * g + r] = t;
      }
      *(bf16x4 *)(dsm[cur ^ 1] + (16 * sb + j) * H + 16 * ib + 4 * g) = dv;
    }
    __syncthreads ();
    cur ^= 1;
  }
  // dW_1 = delta_1 . X^T, output block ob = wv
  {
    const int ob = wv;
#pragma unroll
    for (int ib = 0; ib < 2; ++ib)
    {
      floatx4 acc = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
      for (int kb = 0; kb < 2; ++kb)
      {
        const bf16x8 a = *(const bf16x8 *)(dfm[cur] + (16 * ob + j) * 64 + 32 * kb + 8 * g);
        const bf16x8 b = *(const bf16x8 *)(xf + (16 * ib + j) * 64 + 32 * kb + 8 * g);
        acc = mfma_bf16 (a, b, acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) part[W1 + (16 * ob + 4 * g + r) * IN_PAD + 16 * ib + j] = acc[r];
    }
  }
  // bias gradients and loss: the four waves' values added in wave order
  for (int t = threadIdx.x; t < kTrainRed; t += 256)
  {
    const float s = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
    part[t < 4 * H ? b_off (t / H + 1) + t % H : t < 4 * H + OUT_PAD ? B5 + (t - 4 * H) : NPARAM] = s;
  }
}

// partials[w * PSTRIDE + p] summed over rows w0..w1-1 in row order; the loads of eight rows are issued
// before they are added, so the rows cost one memory latency per eight instead of one each.
__device__ __forceinline__ float ordered_row_sum (const float *__restrict__ partials, int w0, int w1, int p)
{
  float sum = 0.f;
  int w = w0;
  for (; w + 8 <= w1; w += 8)
  {
    float t[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = partials[(int64_t)(w + k) * PSTRIDE + p];
#pragma unroll
    for (int k = 0; k < 8; ++k) sum += t[k];
  }
  for (; w < w1; ++w) sum += partials[(int64_t)w * PSTRIDE + p];
  return sum;
}

// First level of the gradient sum: group g adds its share of the per-block rows, in order.
constexpr int kGradGroups = 16;
__global__ void __launch_bounds__ (256) k_mlp_grad_groups (const float *__restrict__ partials, int nwaves,
                                                           float *__restrict__ groups)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > NPARAM) return;
  const int g = blockIdx.y;
  const int w0 = (int)((int64_t)nwaves * g / kGradGroups), w1 = (int)((int64_t)nwaves * (g + 1) / kGradGroups);
  groups[(int64_t)g * PSTRIDE + p] = ordered_row_sum (partials, w0, w1, p);
}

// Adam over the gradient (the group sums added in a fixed order: deterministic), then the bf16 copies.
// Thread NPARAM writes the step's mean squared error.
__global__ void __launch_bounds__ (256) k_mlp_adam (MlpState s, const float *__restrict__ partials, int nwaves, float lr,
                                                    float b1, float b2, float eps, float bc1, float bc2,
                                                    float loss_scale, float *__restrict__ loss_out)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > NPARAM) return;
  const float gsum = ordered_row_sum (partials, 0, nwaves, p);
  if (p == NPARAM)
  {
    if (loss_out) *loss_out = gsum * loss_scale;
    return;
  }
  int row, col;
  const int l = param_coords (p, row, col);
  if (!param_live (l, row, col)) return;
  const float m = b1 * s.m[p] + (1.f - b1) * gsum;
  const float v = b2 * s.v[p] + (1.f - b2) * gsum * gsum;
  s.m[p] = m;
  s.v[p] = v;
  const float w = s.params[p] - lr * (m / bc1) / (sqrtf (v / bc2) + eps);
  s.params[p] = w;
  write_copies (s, l, row, col, w);
}

// Inference fragments in MFMA A-operand order (permuted k for layers fed from registers) + fp32 biases,
// scaled for the scaled-activation form of the hidden layers (gelu_from_quarter).
__global__ void k_mlp_pack (const float *__restrict__ params, bf16x8 *__restrict__ frag, float *__restrict__ bias)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < NBIAS)
  {
    float b;
    if (t < 4 * H)
      b = 0.25f * params[b_off (t / H + 1) + t % H];   // hidden layers yield a quarter of their pre-activation
    else
      b = params[B5 + (t - 4 * H)];
    bias[t] = b;
  }
  if (t >= NFRAG * 64) return;
  const int fr = t / 64, lane = t % 64, j = lane & 15, g = lane >> 4;
  int l, ob, kb;
  if (fr < FR2)
  {
    l = 1;
    ob = fr - FR1;
    kb = 0;
  }
  else if (fr < FR5)
  {
    l = 2 + (fr - FR2) / 8;
    ob = ((fr - FR2) % 8) / 2;
    kb = (fr - FR2) % 2;
  }
  else
  {
    l = 5;
    ob = 0;
    kb = fr - FR5;
  }
  const int ip = in_pad (l), row = 16 * ob + j;
  // weight scale of the scaled-activation form (gelu_from_quarter): layer 1 x 0.25, layers 2-4 x 2, output x 8
  const float sc = (l == 1) ? 0.25f : (l == 5) ? 8.f : 2.f;
  bf16x8 v;
#pragma unroll
  for (int s = 0; s < 8; ++s)
  {
    const int k = (l == 1) ? 8 * g + s : 32 * kb + ((s < 4) ? 4 * g + s : 16 + 4 * g + (s - 4));
    v[s] = (__bf16)(sc * params[w_off (l) + row * ip + k]);
  }
  frag[t] = v;
}

// Batched inference: each wavefront takes 64 consecutive rows.  out = base + MLP(x) (base nullable).
template <int SB>
__global__ void __launch_bounds__ (256) k_mlp_infer (MlpNet net, const bf16x8 *__restrict__ x,
                                                     const float *__restrict__ base, int64_t n, float *__restrict__ out)
{
  const int lane = threadIdx.x & 63, j = lane & 15, g = lane >> 4;
  const int64_t s0 = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * 64;
  if (s0 >= n) return;
#pragma unroll
  for (int sb0 = 0; sb0 < 4; sb0 += SB)
  {
    bf16x8 bin[SB];
    floatx4 y[SB];
#pragma unroll
    for (int s = 0; s < SB; ++s)
    {
      const int64_t smp = s0 + 16 * (sb0 + s) + j;
      const bf16x8 z = { (__bf16)0.f, (__bf16)0.f, (__bf16)0.f, (__bf16)0.f,
                         (__bf16)0.f, (__bf16)0.f, (__bf16)0.f, (__bf16)0.f };
      bin[s] = (smp < n) ? x[4 * smp + g] : z;
    }
    mlp_forward<SB> (net, bin, y);
    if (g == 0)
#pragma unroll
      for (int s = 0; s < SB; ++s)
      {
        const int64_t smp = s0 + 16 * (sb0 + s) + j;
        if (smp < n)
#pragma unroll
          for (int r = 0; r < OUT; ++r) out[3 * smp + r] = (base ? base[3 * smp + r] : 0.f) + y[s][r];
      }
  }
}

} // namespace drx

using namespace drx;

extern "C" {

int64_t drx_mlp_param_count (void) { return NPARAM; }
int64_t drx_mlp_state_bytes (void) { return STATE_BYTES; }
int64_t drx_mlp_packed_bytes (void) { return PACKED_BYTES; }

int drx_mlp_init (void *state, uint64_t seed, void *stream)
{
  if (!state) return fail (DRX_ERR_ARG, "drx_mlp_init: null state");
  const MlpState s = state_view (state);
  const int nth = NTRN > NPARAM_PAD ? NTRN : NPARAM_PAD;
  hipLaunchKernelGGL (k_mlp_init, dim3 ((nth + 255) / 256), dim3 (256), 0, as_stream (stream), s, seed);
  DRX_LAUNCH_CHECK ();
  hipLaunchKernelGGL (k_mlp_init_t, dim3 ((NPARAM + 255) / 256), dim3 (256), 0, as_stream (stream), s);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_mlp_features (const drx_mesh *mesh, const float *points, const void *locations, int64_t n, void *out_x,
                      float *out_target, float *out_base, void *stream)
{
  if (!mesh || !mesh->data || mesh->ncomp != 3) return fail (DRX_ERR_ARG, "drx_mlp_features: need a geometry mesh");
  if (n < 0 || (n > 0 && (!points || !locations || !out_x))) return fail (DRX_ERR_ARG, "drx_mlp_features: bad arguments");
  if (n == 0) return DRX_OK;
  const bool mixed = mesh->elem_offset != nullptr;
  if (mixed != (mesh->elem_order != nullptr)) return fail (DRX_ERR_ARG, "drx_mlp_features: mixed mesh needs orders and offsets");
  if (!mixed && (mesh->order < 1 || mesh->order > DRX_MAX_MIXED_ORDER))
    return fail (DRX_ERR_ARG, "drx_mlp_features: order %d unsupported", mesh->order);
  const dim3 grid ((unsigned)((n + 255) / 256)), blk (256);
  const int p = mixed ? 0 : mesh->order;
  auto go = [&] (auto kern) {
    hipLaunchKernelGGL (kern, grid, blk, 0, as_stream (stream), mesh->data, mesh->elem_offset, mesh->elem_order,
                        mesh->stride, points, (const Location *)locations, n, (bf16x8 *)out_x, out_target, out_base);
  };
  switch (p)
  {
  case 0: go (k_mlp_features<0>); break;
  case 1: go (k_mlp_features<1>); break;
  case 2: go (k_mlp_features<2>); break;
  case 3: go (k_mlp_features<3>); break;
  default: go (k_mlp_features<4>); break;
  }
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_mlp_train (void *state, const void *x, const float *target, int64_t n, const drx_mlp_train_opts *o,
                   float *out_loss, void *stream)
{
  if (!state || !x || !target || !o || n <= 0) return fail (DRX_ERR_ARG, "drx_mlp_train: bad arguments");
  if (o->batch <= 0 || o->batch % 64 != 0) return fail (DRX_ERR_ARG, "drx_mlp_train: batch %d is not a positive multiple of 64", o->batch);
  if (o->steps < 0) return fail (DRX_ERR_ARG, "drx_mlp_train: negative step count");
  const int nw = o->batch / 64;
  const hipStream_t st = as_stream (stream);
  float *partials = (float *)arena (SLOT_TREE, (size_t)(nw + kGradGroups) * PSTRIDE * sizeof (float), st);
  if (!partials) return DRX_ERR_HIP;
  float *groups = partials + (size_t)nw * PSTRIDE;
  const MlpState s = state_view (state);
  const double b1 = o->beta1 > 0 ? o->beta1 : 0.9, b2 = o->beta2 > 0 ? o->beta2 : 0.999;
  const float eps = o->eps > 0 ? o->eps : 1e-8f;
  const float lr_final = o->lr_final >= 0 ? o->lr_final : o->lr;
  const int total = o->total_steps > 0 ? o->total_steps : o->step0 + o->steps;
  for (int k = 0; k < o->steps; ++k)
  {
    const int t = o->step0 + k + 1;   // Adam step number (1-based)
    const int64_t start = (o->sample0 + (int64_t)k * o->batch) % n;
    hipLaunchKernelGGL (k_mlp_train, dim3 (nw), dim3 (256), 0, st, (const bf16x8 *)x, target, n, start, s.params, s.fw,
                        s.tw, partials, 1.f / (float)o->batch);
    DRX_LAUNCH_CHECK ();
    // cosine decay from lr to lr_final over total steps
    const double ph = fmin (1.0, (double)(t - 1) / (double)(total > 1 ? total - 1 : 1));
    const float lr = (float)(lr_final + 0.5 * (o->lr - lr_final) * (1.0 + cos (M_PI * ph)));
    hipLaunchKernelGGL (k_mlp_grad_groups, dim3 ((NPARAM + 1 + 255) / 256, kGradGroups), dim3 (256), 0, st, partials, nw,
                        groups);
    DRX_LAUNCH_CHECK ();
    hipLaunchKernelGGL (k_mlp_adam, dim3 ((NPARAM + 1 + 255) / 256), dim3 (256), 0, st, s, groups, kGradGroups, lr, (float)b1,
                        (float)b2, eps, (float)(1.0 - pow (b1, t)), (float)(1.0 - pow (b2, t)),
                        1.f / (float)(o->batch * OUT), out_loss ? out_loss + k : nullptr);
    DRX_LAUNCH_CHECK ();
  }
  return DRX_OK;
}

int drx_mlp_pack (const void *state, void *out_packed, void *stream)
{
  if (!state || !out_packed) return fail (DRX_ERR_ARG, "drx_mlp_pack: null pointer");
  const MlpState s = state_view ((void *)state);
  const MlpNet net = mlp_net (out_packed);
  hipLaunchKernelGGL (k_mlp_pack, dim3 ((NFRAG * 64 + 255) / 256), dim3 (256), 0, as_stream (stream), s.params,
                      (bf16x8 *)net.frag, (float *)net.bias);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

int drx_mlp_infer (const void *packed, const void *x, const float *base, int64_t n, float *out, void *stream)
{
  if (!packed || n < 0 || (n > 0 && (!x || !out))) return fail (DRX_ERR_ARG, "drx_mlp_infer: bad arguments");
  if (n == 0) return DRX_OK;
  const int64_t waves = (n + 63) / 64;
  hipLaunchKernelGGL (k_mlp_infer<DRX_MLP_SB>, dim3 ((unsigned)((waves + 3) / 4)), dim3 (256), 0, as_stream (stream),
                      mlp_net (packed), (const bf16x8 *)x, base, n, out);
  DRX_LAUNCH_CHECK ();
  return DRX_OK;
}

} // extern "C"

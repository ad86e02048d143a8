// Device-side tensor Bernstein hex evaluation, 3x3 LUP solve, dray-contract Newton inverse map,
// FP64 polish and the BVH candidate traversal.  All templated on the element order P so that the
// (P+1)^3 loops unroll; P == 0 selects the runtime-order path for mixed-order meshes.
#pragma once

#include "drx_common.hpp"

namespace drx
{

template <int P> struct Ord
{
  static constexpr int N1 = P + 1;
  static constexpr int NPE = N1 * N1 * N1;
  static constexpr int NP = (NPE + 3) & ~3;   // padded component stride
  static constexpr int GSTRIDE = 4 + 3 * NP;  // floats per geometry block
  static constexpr int FSTRIDE = NP;          // floats per scalar-field block
};

// 1D Bernstein basis of degree P and its derivative at t (T = float or double).
template <int P, typename T>
__device__ __forceinline__ void bernstein (T t, T B[P + 1], T dB[P + 1])
{
  const T s = T (1) - t;
  T b[P + 1];   // degree P-1 basis in b[0..P-1]
  b[0] = T (1);
#pragma unroll
  for (int d = 1; d <= P - 1; ++d)
  {
    b[d] = t * b[d - 1];
#pragma unroll
    for (int i = d - 1; i >= 1; --i) b[i] = s * b[i] + t * b[i - 1];
    b[0] = s * b[0];
  }
  // derivative: P (b_{i-1} - b_i) with out-of-range terms zero
#pragma unroll
  for (int i = 0; i <= P; ++i)
  {
    const T lo = (i >= 1) ? b[i - 1] : T (0);
    const T hi = (i <= P - 1) ? b[i] : T (0);
    dB[i] = T (P) * (lo - hi);
  }
  // raise to degree P
  B[P] = t * b[P - 1];
#pragma unroll
  for (int i = P - 1; i >= 1; --i) B[i] = s * b[i] + t * b[i - 1];
  B[0] = s * b[0];
}

// Select arr[k] for a runtime k with a compare/select chain (keeps arr in registers).
template <int N, typename T>
__device__ __forceinline__ T pick (const T arr[N], int k)
{
  T v = arr[0];
#pragma unroll
  for (int q = 1; q < N; ++q) v = (k == q) ? arr[q] : v;
  return v;
}

// phi(x) - origin and the Jacobian J[c][a] = d phi_c / d x_a of a packed geometry block.
// The w-slabs are walked by a rolled loop so that only one slab of control points is live at a time.
template <int P, typename T>
__device__ __forceinline__ void eval_geom (const float *__restrict__ g, const T x[3], T y[3], T J[3][3])
{
  constexpr int N1 = Ord<P>::N1, NP = Ord<P>::NP;
  T Bu[N1], dBu[N1], Bv[N1], dBv[N1], Bw[N1], dBw[N1];
  bernstein<P, T> (x[0], Bu, dBu);
  bernstein<P, T> (x[1], Bv, dBv);
  bernstein<P, T> (x[2], Bw, dBw);
  T val[3] = { 0, 0, 0 }, du[3] = { 0, 0, 0 }, dv[3] = { 0, 0, 0 }, dw[3] = { 0, 0, 0 };
  const float *X = g + 4;
#pragma unroll 1
  for (int k = 0; k < N1; ++k)
  {
    const T bw = pick<N1, T> (Bw, k), dbw = pick<N1, T> (dBw, k);
    const float *slab = X + N1 * N1 * k;
#pragma unroll
    for (int j = 0; j < N1; ++j)
    {
      T a[3] = { 0, 0, 0 }, b[3] = { 0, 0, 0 };
#pragma unroll
      for (int i = 0; i < N1; ++i)
      {
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
          const T cp = (T)slab[c * NP + N1 * j + i];
          a[c] += Bu[i] * cp;
          b[c] += dBu[i] * cp;
        }
      }
      const T w1 = Bv[j] * bw, w2 = dBv[j] * bw, w3 = Bv[j] * dbw;
#pragma unroll
      for (int c = 0; c < 3; ++c)
      {
        val[c] += w1 * a[c];
        du[c] += w1 * b[c];
        dv[c] += w2 * a[c];
        dw[c] += w3 * a[c];
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c)
  {
    y[c] = val[c];
    J[c][0] = du[c];
    J[c][1] = dv[c];
    J[c][2] = dw[c];
  }
}

template <int P>
__device__ __forceinline__ float eval_scalar (const float *__restrict__ f, const float x[3])
{
  constexpr int N1 = Ord<P>::N1;
  float Bu[N1], dBu[N1], Bv[N1], dBv[N1], Bw[N1], dBw[N1];
  bernstein<P, float> (x[0], Bu, dBu);
  bernstein<P, float> (x[1], Bv, dBv);
  bernstein<P, float> (x[2], Bw, dBw);
  float val = 0.f;
#pragma unroll
  for (int k = 0; k < N1; ++k)
  {
    float sv = 0.f;
#pragma unroll
    for (int j = 0; j < N1; ++j)
    {
      float a = 0.f;
      const int row = N1 * (j + N1 * k);
#pragma unroll
      for (int i = 0; i < N1; ++i) a += Bu[i] * f[row + i];
      sv += Bv[j] * a;
    }
    val += Bw[k] * sv;
  }
  return val;
}

// Crout LUP with partial pivoting, as matrix_mult_inv (matrix.hpp:221-373,461-477).  A is consumed.
template <typename T>
__device__ __forceinline__ void lup_solve3 (T A[3][3], const T b[3], T x[3])
{
  int perm[3] = { 0, 1, 2 };
#pragma unroll
  for (int top = 0; top < 3; ++top)
  {
    int maxr = top;
    T maxv = fabs (A[top][top]);
#pragma unroll
    for (int r = top + 1; r < 3; ++r)
    {
      const T cv = fabs (A[r][top]);
      if (maxv < cv)
      {
        maxv = cv;
        maxr = r;
      }
    }
#pragma unroll
    for (int r = top + 1; r < 3; ++r)
    {
      if (maxr == r)
      {
#pragma unroll
        for (int c = 0; c < 3; ++c)
        {
          const T t = A[r][c];
          A[r][c] = A[top][c];
          A[top][c] = t;
        }
        const int tp = perm[r];
        perm[r] = perm[top];
        perm[top] = tp;
      }
    }
#pragma unroll
    for (int c = top + 1; c < 3; ++c) A[top][c] /= A[top][top];
#pragma unroll
    for (int r = top + 1; r < 3; ++r)
#pragma unroll
      for (int c = top + 1; c < 3; ++c) A[r][c] -= A[r][top] * A[top][c];
  }
  T y[3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
  {
    const int pr = perm[r];
    T v = (pr == 0) ? b[0] : ((pr == 1) ? b[1] : b[2]);
#pragma unroll
    for (int c = 0; c < r; ++c) v -= A[r][c] * y[c];
    y[r] = v / A[r][r];
  }
#pragma unroll
  for (int r = 2; r >= 0; --r)
  {
    T v = y[r];
#pragma unroll
    for (int c = r + 1; c < 3; ++c) v -= A[r][c] * x[c];
    x[r] = v;
  }
}

// inf-norm condition estimate ||J|| * ||J^-1|| via the adjugate.
__device__ __forceinline__ float cond_inf (const float J[3][3])
{
  float C[3][3];
  C[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  C[0][1] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  C[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  C[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  C[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  C[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  C[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  C[2][1] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  C[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const float det = J[0][0] * C[0][0] + J[0][1] * C[1][0] + J[0][2] * C[2][0];
  float nj = 0.f, nc = 0.f;
#pragma unroll
  for (int r = 0; r < 3; ++r)
  {
    nj = fmaxf (nj, fabsf (J[r][0]) + fabsf (J[r][1]) + fabsf (J[r][2]));
    nc = fmaxf (nc, fabsf (C[r][0]) + fabsf (C[r][1]) + fabsf (C[r][2]));
  }
  const float ad = fabsf (det);
  return (ad > 0.f) ? nj * nc / ad : INFINITY;
}

struct NewtonOut
{
  int steps;
  float cond;
};

// eval_inverse_local + IterativeMethod::solve (element.hpp:435-486, newton_solver.hpp:50-73): up to 20
// steps; each clamps x into [0,1], evaluates phi and J, solves J dx = target - phi by LUP, x += dx;
// converged when |x - prev|inf < 1e-4; accepted when converged and x in [-1e-4, 1+1e-4]^3.
// `t` is the target relative to the block's origin.
template <int P>
__device__ __forceinline__ bool newton_f32 (const float *__restrict__ g, const float t[3], float x[3], NewtonOut &no,
                                            bool want_cond)
{
  const float tol = 1e-4f;
  float prev[3] = { x[0], x[1], x[2] };
  int steps = 0;
  bool conv = false;
  float Jl[3][3];
  while (steps < 20 && !conv)
  {
#pragma unroll
    for (int i = 0; i < 3; ++i) x[i] = fminf (1.f, fmaxf (x[i], 0.f));
    float y[3], J[3][3];
    eval_geom<P, float> (g, x, y, J);
    float dy[3], dx[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) dy[i] = t[i] - y[i];
    if (want_cond)
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) Jl[r][c] = J[r][c];
    lup_solve3<float> (J, dy, dx);
#pragma unroll
    for (int i = 0; i < 3; ++i) x[i] = x[i] + dx[i];
    ++steps;
    float res = fmaxf (-(x[0] - prev[0]), x[0] - prev[0]);
#pragma unroll
    for (int i = 1; i < 3; ++i) res = fmaxf (res, fmaxf (-(x[i] - prev[i]), x[i] - prev[i]));
    conv = res < tol;
#pragma unroll
    for (int i = 0; i < 3; ++i) prev[i] = x[i];
  }
  no.steps = steps;
  float mn = 2.f, mx = -1.f;
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    mn = fminf (x[d], mn);
    mx = fmaxf (x[d], mx);
  }
  const bool inside = (mn >= 0.f - tol) && (mx <= 1.f + tol);
  const bool found = conv && inside;
  no.cond = (found && want_cond) ? cond_inf (Jl) : 0.f;
  return found;
}

// FP64 Newton correction (no clamping) from an accepted FP32 solution: two steps in double.
template <int P>
__device__ __forceinline__ void polish_f64 (const float *__restrict__ g, const float pt[3], float x[3])
{
  double xd[3] = { x[0], x[1], x[2] };
  const double td[3] = { (double)pt[0] - (double)g[0], (double)pt[1] - (double)g[1], (double)pt[2] - (double)g[2] };
#pragma unroll 1
  for (int it = 0; it < 2; ++it)
  {
    double y[3], J[3][3], dy[3], dx[3];
    eval_geom<P, double> (g, xd, y, J);
#pragma unroll
    for (int i = 0; i < 3; ++i) dy[i] = td[i] - y[i];
    lup_solve3<double> (J, dy, dx);
#pragma unroll
    for (int i = 0; i < 3; ++i) xd[i] += dx[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) x[i] = (float)xd[i];
}

struct LocateCfg
{
  const float4 *nodes;      // inner nodes, 4 float4 each
  const float4 *leaves;     // packed leaves, 2 float4 each
  const float *mesh;        // packed geometry blocks
  const int64_t *offsets;   // mixed order: per-element block offset (nullptr for uniform)
  const int32_t *orders;    // mixed order: per-element order
  int stride;               // uniform: floats per element block
  bool polish;
  float cond_threshold;
};

struct LocateStats
{
  int newton;
  int candidates;
  int polished;
};

template <int P, bool POLISH>
__device__ __forceinline__ bool try_element_at (const LocateCfg &cfg, const float *g, const float pt[3], float x[3],
                                                LocateStats &st)
{
  const float t[3] = { pt[0] - g[0], pt[1] - g[1], pt[2] - g[2] };
  NewtonOut no;
  const bool found = newton_f32<P> (g, t, x, no, POLISH);
  st.newton += no.steps;
  st.candidates += 1;
  if constexpr (POLISH)
  {
    if (found && !(no.cond <= cfg.cond_threshold))
    {
      polish_f64<P> (g, pt, x);
      st.polished += 1;
    }
  }
  return found;
}

// Runtime-order dispatch for mixed meshes.
template <bool POLISH>
__device__ __forceinline__ bool try_element_mixed (const LocateCfg &cfg, int el, const float pt[3], float x[3],
                                                   LocateStats &st)
{
  const float *g = cfg.mesh + cfg.offsets[el];
  switch (cfg.orders[el])
  {
  case 1: return try_element_at<1, POLISH> (cfg, g, pt, x, st);
  case 2: return try_element_at<2, POLISH> (cfg, g, pt, x, st);
  case 3: return try_element_at<3, POLISH> (cfg, g, pt, x, st);
  default: return try_element_at<4, POLISH> (cfg, g, pt, x, st);
  }
}

constexpr int kStackDepth = 64;
constexpr int kStackLds = 16;   // hot part of each lane's traversal stack lives in LDS
constexpr int kBarrier = -2000000000;

// Per-lane traversal stack: entries [0, kStackLds) in LDS (entry-major, so the 64 lanes of a wave hit
// 64 consecutive words), deeper entries in a private array that is touched only by deep traversals.
struct TravStack
{
  int32_t *lds;       // kStackLds * blockDim.x words, this lane's column at lds[e * bdim + tid]
  int bdim, tid;
  int32_t spill[kStackDepth - kStackLds];
  __device__ __forceinline__ void set (int e, int32_t v)
  {
    if (e < kStackLds)
      lds[e * bdim + tid] = v;
    else
      spill[e - kStackLds] = v;
  }
  __device__ __forceinline__ int32_t get (int e) const
  {
    return (e < kStackLds) ? lds[e * bdim + tid] : spill[e - kStackLds];
  }
};

// DeviceMesh::locate (device_mesh.hpp:247-354): depth-first, left child first, right child pushed when
// the point is in both; the first accepted candidate wins.  Returns the element id or -1; sets *overflow
// if the 64-entry stack would overflow (dray's stack has the same capacity).
template <int P, bool POLISH>
__device__ __forceinline__ int locate_point (const LocateCfg &cfg, TravStack &todo, const float pt[3], float ref[3],
                                             LocateStats &st, bool &overflow)
{
  int cur = 0, sp = 0;
  todo.set (0, kBarrier);
  for (;;)
  {
    // Walk inner nodes until this lane holds a leaf (or the stack is exhausted).  Lanes that reach a
    // leaf early wait here, so the Newton solve below runs with the wave converged.
    while (cur > -1)
    {
      const float4 *nd = cfg.nodes + cur;
      const float4 a = nd[0], b = nd[1], c = nd[2], ch = nd[3];
      bool in_l = true, in_r = true;
      if (pt[0] < a.x) in_l = false;
      if (pt[1] < a.y) in_l = false;
      if (pt[2] < a.z) in_l = false;
      if (pt[0] > a.w) in_l = false;
      if (pt[1] > b.x) in_l = false;
      if (pt[2] > b.y) in_l = false;
      if (pt[0] < b.z) in_r = false;
      if (pt[1] < b.w) in_r = false;
      if (pt[2] < c.x) in_r = false;
      if (pt[0] > c.y) in_r = false;
      if (pt[1] > c.z) in_r = false;
      if (pt[2] > c.w) in_r = false;
      if (!in_l && !in_r)
      {
        cur = todo.get (sp);
        --sp;
      }
      else
      {
        const int lc = __float_as_int (ch.x), rc = __float_as_int (ch.y);
        cur = in_l ? lc : rc;
        if (in_l && in_r)
        {
          if (sp + 1 >= kStackDepth)
          {
            overflow = true;
            return -1;
          }
          todo.set (++sp, rc);
        }
      }
    }
    if (cur == kBarrier) return -1;
    const int leaf = -cur - 1;
    const float4 l0 = cfg.leaves[2 * (int64_t)leaf], l1 = cfg.leaves[2 * (int64_t)leaf + 1];
    const int el = __float_as_int (l0.x);
    float x[3] = { (l0.y + l1.x) * 0.5f, (l0.z + l1.y) * 0.5f, (l0.w + l1.z) * 0.5f };   // subref_center
    bool found;
    if constexpr (P == 0)
      found = try_element_mixed<POLISH> (cfg, el, pt, x, st);
    else
      found = try_element_at<P, POLISH> (cfg, cfg.mesh + (int64_t)el * Ord<P>::GSTRIDE, pt, x, st);
    if (found)
    {
      ref[0] = x[0];
      ref[1] = x[1];
      ref[2] = x[2];
      return el;
    }
    cur = todo.get (sp);
    --sp;
  }
}

} // namespace drx

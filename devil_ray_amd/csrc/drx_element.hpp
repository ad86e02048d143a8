// Device-side tensor Bernstein hex evaluation, 3x3 LUP solve, dray-contract Newton inverse map,
// FP64 polish and the BVH candidate traversal.  All templated on the element order P so that the
// (P+1)^3 loops unroll; P == 0 selects the runtime-order path for mixed-order meshes.
#pragma once

#include "drx_common.hpp"

namespace drx
{

#ifndef DRX_REG_ELEMENT_MAX_ORDER
#define DRX_REG_ELEMENT_MAX_ORDER 0   // elements up to this order are held in registers during Newton
#endif
#ifndef DRX_EVAL_UNROLL_K
#define DRX_EVAL_UNROLL_K 1           // w-slabs of control points in flight per step of eval_geom
#endif
#ifndef DRX_EVAL_ROTATE_MIN_ORDER
#define DRX_EVAL_ROTATE_MIN_ORDER 4   // orders >= this rotate the w-weight arrays per slab instead of selecting
#endif                                // by slab index (measured: faster at order 4, slower at order 3)

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
#pragma unroll DRX_EVAL_UNROLL_K
  for (int k = 0; k < N1; ++k)
  {
    T bw, dbw;
    if constexpr (P >= DRX_EVAL_ROTATE_MIN_ORDER)
    {
      // slab k's w-weights are always in slot 0: the arrays rotate by one each slab, so they stay in
      // registers (the compiler turns a runtime-indexed read of a 5-entry array into scratch memory)
      bw = Bw[0];
      dbw = dBw[0];
#pragma unroll
      for (int i = 0; i < N1 - 1; ++i)
      {
        Bw[i] = Bw[i + 1];
        dBw[i] = dBw[i + 1];
      }
    }
    else
    {
      bw = pick<N1, T> (Bw, k);
      dbw = pick<N1, T> (dBw, k);
    }
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

// Same evaluation from a register copy of the element (geometry relative to the block origin).
template <int P, typename T>
__device__ __forceinline__ void eval_geom_regs (const float (&C)[3][Ord<P>::NPE], const T x[3], T y[3], T J[3][3])
{
  constexpr int N1 = Ord<P>::N1;
  T Bu[N1], dBu[N1], Bv[N1], dBv[N1], Bw[N1], dBw[N1];
  bernstein<P, T> (x[0], Bu, dBu);
  bernstein<P, T> (x[1], Bv, dBv);
  bernstein<P, T> (x[2], Bw, dBw);
  T val[3] = { 0, 0, 0 }, du[3] = { 0, 0, 0 }, dv[3] = { 0, 0, 0 }, dw[3] = { 0, 0, 0 };
#pragma unroll
  for (int k = 0; k < N1; ++k)
  {
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
          const T cp = (T)C[c][N1 * (j + N1 * k) + i];
          a[c] += Bu[i] * cp;
          b[c] += dBu[i] * cp;
        }
      }
      const T w1 = Bv[j] * Bw[k], w2 = dBv[j] * Bw[k], w3 = Bv[j] * dBw[k];
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

// Scalar field value at reference point x; FT = float, or double for an FP64 field (evaluated in double
// at the float32 reference point).
template <int P, typename FT = float>
__device__ __forceinline__ FT eval_scalar (const FT *__restrict__ f, const float x[3])
{
  constexpr int N1 = Ord<P>::N1;
  FT Bu[N1], dBu[N1], Bv[N1], dBv[N1], Bw[N1], dBw[N1];
  bernstein<P, FT> ((FT)x[0], Bu, dBu);
  bernstein<P, FT> ((FT)x[1], Bv, dBv);
  bernstein<P, FT> ((FT)x[2], Bw, dBw);
  FT val = 0;
#pragma unroll
  for (int k = 0; k < N1; ++k)
  {
    FT sv = 0;
#pragma unroll
    for (int j = 0; j < N1; ++j)
    {
      FT a = 0;
      const int row = N1 * (j + N1 * k);
#pragma unroll
      for (int i = 0; i < N1; ++i) a += Bu[i] * f[row + i];
      sv += Bv[j] * a;
    }
    val += Bw[k] * sv;
  }
  return val;
}

// Scalar value and reference-space gradient (for lit shading).
template <int P>
__device__ __forceinline__ float eval_scalar_d (const float *__restrict__ f, const float x[3], float d[3])
{
  constexpr int N1 = Ord<P>::N1;
  float Bu[N1], dBu[N1], Bv[N1], dBv[N1], Bw[N1], dBw[N1];
  bernstein<P, float> (x[0], Bu, dBu);
  bernstein<P, float> (x[1], Bv, dBv);
  bernstein<P, float> (x[2], Bw, dBw);
  float val = 0.f, du = 0.f, dv = 0.f, dw = 0.f;
#pragma unroll
  for (int k = 0; k < N1; ++k)
  {
#pragma unroll
    for (int j = 0; j < N1; ++j)
    {
      float a = 0.f, b = 0.f;
      const int row = N1 * (j + N1 * k);
#pragma unroll
      for (int i = 0; i < N1; ++i)
      {
        a += Bu[i] * f[row + i];
        b += dBu[i] * f[row + i];
      }
      val += Bv[j] * Bw[k] * a;
      du += Bv[j] * Bw[k] * b;
      dv += dBv[j] * Bw[k] * a;
      dw += Bv[j] * dBw[k] * a;
    }
  }
  d[0] = du;
  d[1] = dv;
  d[2] = dw;
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
  bool conv;
};

// eval_inverse_local + IterativeMethod::solve (element.hpp:435-486, newton_solver.hpp:50-73): up to 20
// steps; each clamps x into [0,1], evaluates phi and J, solves J dx = target - phi by LUP, x += dx;
// converged when |x - prev|inf < 1e-4; accepted when converged and x in [-1e-4, 1+1e-4]^3.
// `t` is the target relative to the block's origin.
template <int P, bool REGS>
__device__ __forceinline__ bool newton_f32 (const float *__restrict__ g, const float t[3], float x[3], NewtonOut &no,
                                            bool want_cond, int max_steps = 20)
{
  // REGS: the element's control points are loaded once into registers and reused by every iteration
  float C[REGS ? 3 : 1][REGS ? Ord<P>::NPE : 1];
  if constexpr (REGS)
  {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < Ord<P>::NPE; ++q) C[c][q] = g[4 + c * Ord<P>::NP + q];
  }
  const float tol = 1e-4f;
  float prev[3] = { x[0], x[1], x[2] };
  int steps = 0;
  bool conv = false;
  float Jl[3][3];
  while (steps < max_steps && !conv)
  {
#pragma unroll
    for (int i = 0; i < 3; ++i) x[i] = fminf (1.f, fmaxf (x[i], 0.f));
    float y[3], J[3][3];
    if constexpr (REGS)
      eval_geom_regs<P, float> (C, x, y, J);
    else
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
  no.conv = conv;
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

// Grid cell of a coordinate for the compact traversal nodes (drx_bvh_quantize): monotone in v, so a
// point inside a box lies inside the box's cell range.  The same expression quantizes the boxes and the
// query points.
__host__ __device__ __forceinline__ uint32_t grid_cell (float v, float origin, float cells_per_unit)
{
#pragma clang fp contract(off)
  const float c = floorf ((v - origin) * cells_per_unit);
  return c > 0.f ? (c < 65535.f ? (uint32_t)c : 65535u) : 0u;
}

// polish_f64 as an out-of-line call: the FP64 correction runs only for accepted points whose condition
// estimate exceeds the threshold, and kept out of line its double-precision temporaries do not raise the
// register budget of the walk around it.  Arguments and result by value, so that the caller's point stays
// in registers.
template <int P>
__device__ __attribute__ ((noinline)) float3 polish_f64_call (const float *__restrict__ g, float px, float py, float pz,
                                                              float x0, float x1, float x2)
{
  const float pt[3] = { px, py, pz };
  float x[3] = { x0, x1, x2 };
  polish_f64<P> (g, pt, x);
  return make_float3 (x[0], x[1], x[2]);
}

struct LocateCfg
{
  const float4 *nodes;      // inner nodes, 4 float4 each
  const uint4 *qnodes;      // optional compact inner nodes, 2 uint4 each (then leaf_boxes and the grid are set)
  const float4 *leaf_boxes; // with qnodes: exact leaf boxes, 2 float4 each
  float q_org[3], q_inv[3]; // with qnodes: grid origin and cells per unit length
  const float4 *wide;       // 8-wide nodes, 16 float4 each (nullptr: binary traversal)
  const float4 *leaves;     // packed leaves, 2 float4 each
  const float4 *starts;     // optional Newton start maps, 3 float4 per leaf (drx_bvh_leaf_starts)
  bool centre;              // start at the SubRef centre (used only when starts is null)
  const float *mesh;        // packed geometry blocks
  const int64_t *offsets;   // mixed order: per-element block offset (nullptr for uniform)
  const int32_t *orders;    // mixed order: per-element order
  int stride;               // uniform: floats per element block
  bool polish;
  float cond_threshold;
  int32_t *polished_out;    // optional, per point: 1 where the FP64 correction ran (batched locate only)
  const Location *guess;    // DRX_INIT_GUESS: given starts (per point, or per ray and sample)
  Location *record;         // integrate: the location of every (ray, sample)
  int samples_per_ray;      // integrate: guess / record entries per ray
};

struct LocateStats
{
  int newton;
  int candidates;
  int polished;
  int fallbacks;   // learned-start Newton runs that did not converge and were redone from the plain start
};

template <int P, bool POLISH>
__device__ __forceinline__ bool try_element_at (const LocateCfg &cfg, const float *g, const float pt[3], float x[3],
                                                LocateStats &st)
{
  const float t[3] = { pt[0] - g[0], pt[1] - g[1], pt[2] - g[2] };
  NewtonOut no;
  const bool found = newton_f32<P, (P <= DRX_REG_ELEMENT_MAX_ORDER)> (g, t, x, no, POLISH);
  st.newton += no.steps;
  st.candidates += 1;
  if constexpr (POLISH)
  {
    if (found && !(no.cond <= cfg.cond_threshold))
    {
      const float3 r = polish_f64_call<P> (g, pt[0], pt[1], pt[2], x[0], x[1], x[2]);
      x[0] = r.x;
      x[1] = r.y;
      x[2] = r.z;
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

// Newton's starting point for a candidate leaf (dray hands LocateHack an uninitialised guess,
// device_mesh.hpp:327-334): the element's affine approximation at its centre, i.e. the trilinear map
// through its 8 corner control points differentiated at the reference centre, solved for the point and
// clamped into the candidate's SubRef; the SubRef centre where that approximation is singular.  Nearly
// affine elements start within the Newton tolerance's reach, so most solves converge in two steps instead
// of three.  Plain float operations (no contraction), so the oracle computes the same start bit for bit.
__device__ __forceinline__ void corner_start (const float *g, int p, const float pt[3], const float4 &l0,
                                              const float4 &l1, float x[3])
{
#pragma clang fp contract(off)
  const float lo[3] = { l0.y, l0.z, l0.w }, hi[3] = { l1.x, l1.y, l1.z };
  const int n1 = p + 1, np = (n1 * n1 * n1 + 3) & ~3;
  int idx[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) idx[q] = (q & 1) * p + n1 * (((q >> 1) & 1) * p) + n1 * n1 * ((q >> 2) * p);
  float r[3], eu[3], ev[3], ew[3];
  // one coordinate at a time, so only its 8 corner values are live
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    const float *X = g + 4 + d * np;
    float c[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) c[q] = X[idx[q]];
    const float m = (((c[0] + c[1]) + (c[2] + c[3])) + ((c[4] + c[5]) + (c[6] + c[7]))) * 0.125f;
    r[d] = (pt[d] - g[d]) - m;
    eu[d] = (((c[1] - c[0]) + (c[3] - c[2])) + ((c[5] - c[4]) + (c[7] - c[6]))) * 0.25f;
    ev[d] = (((c[2] - c[0]) + (c[3] - c[1])) + ((c[6] - c[4]) + (c[7] - c[5]))) * 0.25f;
    ew[d] = (((c[4] - c[0]) + (c[5] - c[1])) + ((c[6] - c[2]) + (c[7] - c[3]))) * 0.25f;
    asm volatile ("" ::: "memory");   // finish this coordinate before loading the next one's corners
  }
  // Cramer's rule: x - 1/2 = (det(r, ev, ew), det(eu, r, ew), det(eu, ev, r)) / det(eu, ev, ew)
  const float vw[3] = { ev[1] * ew[2] - ev[2] * ew[1], ev[2] * ew[0] - ev[0] * ew[2], ev[0] * ew[1] - ev[1] * ew[0] };
  const float rw[3] = { r[1] * ew[2] - r[2] * ew[1], r[2] * ew[0] - r[0] * ew[2], r[0] * ew[1] - r[1] * ew[0] };
  const float vr[3] = { ev[1] * r[2] - ev[2] * r[1], ev[2] * r[0] - ev[0] * r[2], ev[0] * r[1] - ev[1] * r[0] };
  const float det = (eu[0] * vw[0] + eu[1] * vw[1]) + eu[2] * vw[2];
  const float inv = 1.0f / det;
  const float dx[3] = { ((r[0] * vw[0] + r[1] * vw[1]) + r[2] * vw[2]) * inv,
                        ((eu[0] * rw[0] + eu[1] * rw[1]) + eu[2] * rw[2]) * inv,
                        ((eu[0] * vr[0] + eu[1] * vr[1]) + eu[2] * vr[2]) * inv };
#pragma unroll
  for (int d = 0; d < 3; ++d)
  {
    const float v = 0.5f + dx[d];
    // NaN / infinity (singular corners) fall back to the SubRef centre
    x[d] = (v == v && fabsf (v) < 1e30f) ? fminf (hi[d], fmaxf (lo[d], v)) : (lo[d] + hi[d]) * 0.5f;
  }
}

// The Newton start for candidate element `el` of leaf `leaf` (leaf record l0, l1): the leaf's stored
// start map when the BVH carries one (read in the same round as the leaf record), else the SubRef centre
// when the caller asked for dray's subref_center start (data_model/subref.hpp:52-56), else corner_start.
// candidate_start_rows takes the start map rows already read (map[], used when cfg.starts is set).
template <int P>
__device__ __forceinline__ void candidate_start_rows (const LocateCfg &cfg, int el, const float4 (&map)[3],
                                                      const float pt[3], const float4 &l0, const float4 &l1, float x[3])
{
  if (cfg.starts)
  {
    // an explicit fma chain, so the start's rounding does not depend on how the compiler contracts
    const float4 r0 = map[0], r1 = map[1], r2 = map[2];
    const float v[3] = { fmaf (r0.x, pt[0], fmaf (r0.y, pt[1], fmaf (r0.z, pt[2], r0.w))),
                         fmaf (r1.x, pt[0], fmaf (r1.y, pt[1], fmaf (r1.z, pt[2], r1.w))),
                         fmaf (r2.x, pt[0], fmaf (r2.y, pt[1], fmaf (r2.z, pt[2], r2.w))) };
    const float lo[3] = { l0.y, l0.z, l0.w }, hi[3] = { l1.x, l1.y, l1.z };
#pragma unroll
    for (int d = 0; d < 3; ++d)
      x[d] = (v[d] == v[d] && fabsf (v[d]) < 1e30f) ? fminf (hi[d], fmaxf (lo[d], v[d])) : (lo[d] + hi[d]) * 0.5f;
  }
  else if (cfg.centre)
  {
    x[0] = (l0.y + l1.x) * 0.5f;
    x[1] = (l0.z + l1.y) * 0.5f;
    x[2] = (l0.w + l1.z) * 0.5f;
  }
  else if constexpr (P == 0)
    corner_start (cfg.mesh + cfg.offsets[el], cfg.orders[el], pt, l0, l1, x);
  else
    corner_start (cfg.mesh + (int64_t)el * Ord<P>::GSTRIDE, P, pt, l0, l1, x);
}

template <int P>
__device__ __forceinline__ void candidate_start (const LocateCfg &cfg, int el, int leaf, const float pt[3],
                                                 const float4 &l0, const float4 &l1, float x[3])
{
  float4 map[3];
  if (cfg.starts)
  {
    const float4 *m = cfg.starts + 3 * (int64_t)leaf;
    map[0] = m[0];
    map[1] = m[1];
    map[2] = m[2];
  }
  candidate_start_rows<P> (cfg, el, map, pt, l0, l1, x);
}

#ifndef DRX_STACK_LDS
#define DRX_STACK_LDS 16   // entries of each lane's traversal stack kept in LDS
#endif
constexpr int kStackDepth = 64;
constexpr int kStackLds = DRX_STACK_LDS;   // hot part of each lane's traversal stack lives in LDS
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

// Element staging in LDS: the lanes of a wavefront (rays of one 8x8 tile, or neighbouring query points)
// mostly test the same few elements.  After traversal each lane holds a candidate; the wavefront takes the
// distinct candidate elements in lane order, up to kStage<P> of them, copies each element's packed block
// into its LDS area in one cooperative round (one float4 per lane) and returns, per lane, the LDS copy of
// its element (or its global block when its element did not fit).  Newton then reads control points from
// LDS instead of re-reading them from the cache hierarchy every iteration; the bits are the same.
#ifndef DRX_STAGE_K
#define DRX_STAGE_K 4   // elements of order <= 3 staged per wavefront
#endif
template <int P> struct Stage
{
  static constexpr int F4 = (Ord<P>::GSTRIDE + 3) / 4;   // float4 per packed element block
  static constexpr int K = P <= 3 ? DRX_STAGE_K : 2;     // elements staged per wavefront
  static constexpr int FLOATS = K * F4 * 4;              // LDS floats per wavefront
};

__device__ __forceinline__ void stage_sync ()
{
  __builtin_amdgcn_fence (__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier ();
  __builtin_amdgcn_fence (__ATOMIC_ACQUIRE, "wavefront");
}

template <int P>
__device__ __forceinline__ const float *stage_element (const LocateCfg &cfg, int el, const float *g, float *stage)
{
  constexpr int F4 = Stage<P>::F4, K = Stage<P>::K;
  constexpr int MIN_ACTIVE = 16;                     // fewer active lanes: read from global memory
  constexpr int R = (F4 + MIN_ACTIVE - 1) / MIN_ACTIVE;   // copy rounds per element at >= MIN_ACTIVE lanes
  // Only the lanes active here take part: lanes whose ray or search has ended are masked off, so the
  // copy is spread over the active lanes by their rank among them.
  const uint64_t active = __ballot (1);
  const int nact = __popcll (active);
  if (nact < MIN_ACTIVE) return g;
  const int lane = __lane_id ();
  const int rank = __popcll (active & ((1ull << lane) - 1));
  uint64_t todo = __ballot (el >= 0);
  int slot = -1;
  int ids[K];
  int n = 0;
#pragma unroll
  for (int s = 0; s < K; ++s)
  {
    ids[s] = -1;
    if (todo)
    {
      const int leader = __ffsll ((unsigned long long)todo) - 1;
      const int e = __builtin_amdgcn_readlane (el, leader);
      const uint64_t mine = __ballot (el == e);
      if (el == e) slot = s;
      todo &= ~mine;
      ids[s] = e;
      n = s + 1;
    }
  }
  if (n == 0) return g;
  stage_sync ();   // earlier readers of this wavefront's stage are done
#pragma unroll
  for (int s = 0; s < K; ++s)
    if (s < n)
    {
      const float4 *src = (const float4 *)(cfg.mesh + (int64_t)ids[s] * Ord<P>::GSTRIDE);
      float4 *dst = (float4 *)(stage + s * F4 * 4);
      float4 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r)
      {
        const int q = rank + r * nact;
        if (q < F4) v[r] = src[q];
      }
#pragma unroll
      for (int r = 0; r < R; ++r)
      {
        const int q = rank + r * nact;
        if (q < F4) dst[q] = v[r];
      }
    }
  stage_sync ();
  return slot >= 0 ? stage + slot * F4 * 4 : g;
}

constexpr int kWalkDone = -1, kWalkOverflow = -2;

// DeviceMesh::locate's traversal (device_mesh.hpp:262-320) from the state (cur, sp) to the next candidate
// leaf: inner nodes are visited depth-first, left child first, the right child pushed when the point is
// in both child boxes.  Over the compact nodes (cfg.qnodes) the point's grid cells are compared with
// each child's cell range, a superset of dray's box test, and each reached leaf's exact box then decides,
// so the leaves returned are dray's candidates in dray's order.  Returns the leaf index with cur = its
// code, its packed record in l0, l1 and its Newton start map in map[] when the BVH has start maps (read
// together with the exact box over the compact nodes, so a candidate costs one memory round trip),
// kWalkDone when the stack is exhausted, kWalkOverflow when the 64-entry stack would overflow.
__device__ __forceinline__ int next_candidate_leaf (const LocateCfg &cfg, TravStack &todo, const float pt[3], int &cur,
                                                    int &sp, float4 &l0, float4 &l1, float4 (&map)[3])
{
  // compact nodes: the point's grid cells, compared with each child box's cell range
  const bool compact = cfg.qnodes != nullptr;
  uint32_t qp[3] = { 0u, 0u, 0u };
  if (compact)
#pragma unroll
    for (int d = 0; d < 3; ++d) qp[d] = grid_cell (pt[d], cfg.q_org[d], cfg.q_inv[d]);
  int leaf;
  for (;;)
  {
    while (cur > -1)
    {
      bool in_l = true, in_r = true;
      int lc, rc;
      if (compact)
      {
        const uint4 A = cfg.qnodes[cur], B = cfg.qnodes[cur + 1];
        if (qp[0] < (A.x & 0xFFFFu)) in_l = false;
        if (qp[1] < (A.x >> 16)) in_l = false;
        if (qp[2] < (A.y & 0xFFFFu)) in_l = false;
        if (qp[0] > (A.y >> 16)) in_l = false;
        if (qp[1] > (A.z & 0xFFFFu)) in_l = false;
        if (qp[2] > (A.z >> 16)) in_l = false;
        if (qp[0] < (A.w & 0xFFFFu)) in_r = false;
        if (qp[1] < (A.w >> 16)) in_r = false;
        if (qp[2] < (B.x & 0xFFFFu)) in_r = false;
        if (qp[0] > (B.x >> 16)) in_r = false;
        if (qp[1] > (B.y & 0xFFFFu)) in_r = false;
        if (qp[2] > (B.y >> 16)) in_r = false;
        lc = (int)B.z;
        rc = (int)B.w;
      }
      else
      {
        const float4 *nd = cfg.nodes + cur;
        const float4 a = nd[0], b = nd[1], c = nd[2], ch = nd[3];
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
        lc = __float_as_int (ch.x);
        rc = __float_as_int (ch.y);
      }
      if (!in_l && !in_r)
      {
        cur = todo.get (sp);
        --sp;
      }
      else
      {
        cur = in_l ? lc : rc;
        if (in_l && in_r)
        {
          if (sp + 1 >= kStackDepth)
          {
            return kWalkOverflow;
          }
          todo.set (++sp, rc);
        }
      }
    }
    if (cur == kBarrier) return kWalkDone;
    leaf = -cur - 1;
    l0 = cfg.leaves[2 * (int64_t)leaf];
    l1 = cfg.leaves[2 * (int64_t)leaf + 1];
    if (cfg.starts)
    {
      const float4 *m = cfg.starts + 3 * (int64_t)leaf;
      map[0] = m[0];
      map[1] = m[1];
      map[2] = m[2];
    }
    if (!compact) break;
    // the cell ranges are conservative: the leaf's exact box (dray's child-box test) decides
    const float4 b0 = cfg.leaf_boxes[2 * (int64_t)leaf], b1 = cfg.leaf_boxes[2 * (int64_t)leaf + 1];
    const bool in = !(pt[0] < b0.x) && !(pt[1] < b0.y) && !(pt[2] < b0.z) && !(pt[0] > b0.w) && !(pt[1] > b1.x) &&
                    !(pt[2] > b1.y);
    if (in) break;
    cur = todo.get (sp);
    --sp;
  }
  return leaf;
}

// DeviceMesh::locate (device_mesh.hpp:247-354): depth-first, left child first, right child pushed when
// the point is in both; the first accepted candidate wins.
//
// Resumable form, one candidate per call: cur / sp carry this lane's depth-first state between calls
// (start a search with cur = 0, sp = 0 and kBarrier in stack entry 0).  The lane walks inner nodes until
// it holds a leaf (lanes that get there early wait, so the Newton solve runs with the wave converged),
// then tests that one candidate.  Returns 1 if it was accepted (ref, cell set), 0 if it was rejected (the
// state then holds the next stack entry: call again with the same point), -1 when the search is
// exhausted or the 64-entry stack (dray's capacity) would overflow (*overflow set).
template <int P, bool POLISH>
__device__ __forceinline__ int locate_point_step (const LocateCfg &cfg, TravStack &todo, const float pt[3],
                                                  float ref[3], LocateStats &st, bool &overflow, int &cur, int &sp,
                                                  int &cell, float *stage = nullptr, const Location *guess = nullptr)
{
  float4 l0, l1, map[3];
  const int leaf = next_candidate_leaf (cfg, todo, pt, cur, sp, l0, l1, map);
  if (leaf < 0)
  {
    if (leaf == kWalkOverflow) overflow = true;
    return -1;
  }
  const int el = __float_as_int (l0.x);
  float x[3];
  candidate_start_rows<P> (cfg, el, map, pt, l0, l1, x);
  if (guess && guess->cell == el)   // a given start for this candidate's element (use_init_guess)
  {
    x[0] = guess->ref[0];
    x[1] = guess->ref[1];
    x[2] = guess->ref[2];
  }
  bool found;
  if constexpr (P == 0)
    found = try_element_mixed<POLISH> (cfg, el, pt, x, st);
  else
  {
    const float *g = cfg.mesh + (int64_t)el * Ord<P>::GSTRIDE;
    if (stage) g = stage_element<P> (cfg, el, g, stage);
    found = try_element_at<P, POLISH> (cfg, g, pt, x, st);
  }
  if (found)
  {
    ref[0] = x[0];
    ref[1] = x[1];
    ref[2] = x[2];
    cell = el;
    return 1;
  }
  cur = todo.get (sp);
  --sp;
  return 0;
}

// The whole search for one point: the element id or -1.
template <int P, bool POLISH>
__device__ __forceinline__ int locate_point (const LocateCfg &cfg, TravStack &todo, const float pt[3], float ref[3],
                                             LocateStats &st, bool &overflow)
{
  int cur = 0, sp = 0, cell = -1;
  todo.set (0, kBarrier);
  for (;;)
  {
    const int r = locate_point_step<P, POLISH> (cfg, todo, pt, ref, st, overflow, cur, sp, cell);
    if (r > 0) return cell;
    if (r < 0) return -1;
  }
}

// Same candidate sequence over the 8-wide collapse: the hit children of a node beyond the first are
// pushed as child codes, highest slot first, so they pop in slot (= binary left-to-right) order.
constexpr int32_t kWideEnd = (int32_t)0x80000000;

template <int P, bool POLISH>
__device__ __forceinline__ int locate_point_wide_step (const LocateCfg &cfg, TravStack &todo, const float pt[3],
                                                       float ref[3], LocateStats &st, bool &overflow, int32_t &cur,
                                                       int &sp, int &cell)
{
  // resumable like locate_point_step: start with cur = 0, sp = 0 (sp counts stack entries)
  {
    while (cur >= 0)
    {
      const float4 *nd = cfg.wide + 16 * (int64_t)cur;
      const float4 lx0 = nd[0], lx1 = nd[1], ly0 = nd[2], ly1 = nd[3], lz0 = nd[4], lz1 = nd[5];
      const float4 hx0 = nd[6], hx1 = nd[7], hy0 = nd[8], hy1 = nd[9], hz0 = nd[10], hz1 = nd[11];
      const float4 c0 = nd[12], c1 = nd[13];
      const float lx[8] = { lx0.x, lx0.y, lx0.z, lx0.w, lx1.x, lx1.y, lx1.z, lx1.w };
      const float ly[8] = { ly0.x, ly0.y, ly0.z, ly0.w, ly1.x, ly1.y, ly1.z, ly1.w };
      const float lz[8] = { lz0.x, lz0.y, lz0.z, lz0.w, lz1.x, lz1.y, lz1.z, lz1.w };
      const float hx[8] = { hx0.x, hx0.y, hx0.z, hx0.w, hx1.x, hx1.y, hx1.z, hx1.w };
      const float hy[8] = { hy0.x, hy0.y, hy0.z, hy0.w, hy1.x, hy1.y, hy1.z, hy1.w };
      const float hz[8] = { hz0.x, hz0.y, hz0.z, hz0.w, hz1.x, hz1.y, hz1.z, hz1.w };
      const int32_t ch[8] = { __float_as_int (c0.x), __float_as_int (c0.y), __float_as_int (c0.z), __float_as_int (c0.w),
                              __float_as_int (c1.x), __float_as_int (c1.y), __float_as_int (c1.z), __float_as_int (c1.w) };
      uint32_t mask = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
      {
        // dray's inclusive point-in-box test (device_mesh.hpp:272-288)
        const bool in = !(pt[0] < lx[k] || pt[1] < ly[k] || pt[2] < lz[k] || pt[0] > hx[k] || pt[1] > hy[k] ||
                          pt[2] > hz[k]);
        mask |= (in ? 1u : 0u) << k;
      }
      if (mask == 0)
      {
        cur = (sp == 0) ? kWideEnd : todo.get (--sp);
        continue;
      }
      const int first = __builtin_ctz (mask);
      int32_t next = ch[0];
#pragma unroll
      for (int k = 7; k >= 0; --k)
      {
        if (k == first) next = ch[k];
        if (k > first && ((mask >> k) & 1u))
        {
          if (sp >= kStackDepth)
          {
            overflow = true;
            return -1;
          }
          todo.set (sp++, ch[k]);
        }
      }
      cur = next;
    }
    if (cur == kWideEnd) return -1;
    const int leaf = -cur - 1;
    const float4 l0 = cfg.leaves[2 * (int64_t)leaf], l1 = cfg.leaves[2 * (int64_t)leaf + 1];
    const int el = __float_as_int (l0.x);
    float x[3];
    candidate_start<P> (cfg, el, leaf, pt, l0, l1, x);
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
      cell = el;
      return 1;
    }
    cur = (sp == 0) ? kWideEnd : todo.get (--sp);
    return 0;
  }
}

template <int P, bool POLISH>
__device__ __forceinline__ int locate_point_wide (const LocateCfg &cfg, TravStack &todo, const float pt[3],
                                                  float ref[3], LocateStats &st, bool &overflow)
{
  int32_t cur = 0;
  int sp = 0, cell = -1;
  for (;;)
  {
    const int r = locate_point_wide_step<P, POLISH> (cfg, todo, pt, ref, st, overflow, cur, sp, cell);
    if (r > 0) return cell;
    if (r < 0) return -1;
  }
}

template <int P, bool POLISH, bool WIDE>
__device__ __forceinline__ int locate_any (const LocateCfg &cfg, TravStack &todo, const float pt[3], float ref[3],
                                           LocateStats &st, bool &overflow)
{
  if constexpr (WIDE)
    return locate_point_wide<P, POLISH> (cfg, todo, pt, ref, st, overflow);
  else
    return locate_point<P, POLISH> (cfg, todo, pt, ref, st, overflow);
}

// ---------------------------------------------------------------------------------------------
// Wave-coherent locate over the 8-wide BVH.  All active lanes of a wavefront call it together.  The
// wave walks the union of its lanes' depth-first paths with wave-uniform node reads; at every wide node
// each lane tests its own point and a per-slot ballot keeps, for each child, exactly the lanes whose
// own traversal would enter it.  Reaching a leaf appends it to the queue (LDS) of every lane in its mask
// that is still searching, so each lane's queue holds its own depth-first candidate sequence.  Queues are
// drained (per-lane Newton, first accept wins) when every searching lane holds a candidate or a queue is
// full; lanes that accepted drop out of all pending masks and the walk resumes.
// ---------------------------------------------------------------------------------------------
constexpr int kWaveStack = 96;   // pending (child, lane mask) entries per wave
// Wave-uniform node reads go through the constant address space so they become scalar loads.
typedef const float __attribute__ ((address_space (4))) cfloat;
constexpr int kCandQ = 4;        // queued candidate leaves per lane

struct WaveScratch
{
  int32_t *st_code;    // kWaveStack entries of this wave
  uint64_t *st_mask;   // kWaveStack entries of this wave
  int32_t *cand;       // kCandQ x 64, this wave's queue: cand[q * 64 + lane]
};

template <int P, bool POLISH>
__device__ __forceinline__ int locate_wave (const LocateCfg &cfg, const WaveScratch &ws, const float pt[3],
                                            float ref[3], LocateStats &st, bool &overflow)
{
  const int lane = __lane_id ();
  const uint64_t me = 1ull << lane;
  uint64_t searching = __ballot (1);
  int result = -1;
  int ncand = 0;
  int sp = 0;
  int32_t cur = 0;
  uint64_t curmask = searching;

  // per-lane drain of the candidate queues; updates `searching`
  auto drain = [&] () {
    for (int q = 0; q < ncand; ++q)
    {
      if (result >= 0) break;
      const int leaf = ws.cand[q * 64 + lane];
      const float4 l0 = cfg.leaves[2 * (int64_t)leaf], l1 = cfg.leaves[2 * (int64_t)leaf + 1];
      const int el = __float_as_int (l0.x);
      float x[3];
      candidate_start<P> (cfg, el, leaf, pt, l0, l1, x);
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
        result = el;
      }
    }
    ncand = 0;
    searching &= __ballot (result < 0);
  };

  auto pop = [&] () {
    for (;;)
    {
      if (sp == 0)
      {
        cur = kWideEnd;
        return;
      }
      --sp;
      cur = __builtin_amdgcn_readfirstlane (ws.st_code[sp]);
      const uint64_t m = ws.st_mask[sp];
      // readfirstlane returns a signed int: widen through uint32_t so bit 31 does not sign-extend
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane ((int)(uint32_t)m);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane ((int)(uint32_t)(m >> 32));
      curmask = ((uint64_t)hi << 32) | (uint64_t)lo;
      curmask &= searching;
      if (curmask) return;
    }
  };

  while (cur != kWideEnd && searching)
  {
    if (cur >= 0)
    {
      cur = __builtin_amdgcn_readfirstlane (cur);
      cfloat *nd = (cfloat *)((const float *)cfg.wide + 64 * (int64_t)cur);
      uint64_t b[8];
      const bool part = (curmask & me) != 0;
#pragma unroll
      for (int k = 0; k < 8; ++k)
      {
        const bool in = part && !(pt[0] < nd[k] || pt[1] < nd[8 + k] || pt[2] < nd[16 + k] || pt[0] > nd[24 + k] ||
                                  pt[1] > nd[32 + k] || pt[2] > nd[40 + k]);
        b[k] = __ballot (in);
      }
      int first = -1;
#pragma unroll
      for (int k = 7; k >= 0; --k)
        if (b[k]) first = k;
      if (first < 0)
      {
        pop ();
        continue;
      }
#pragma unroll
      for (int k = 7; k >= 0; --k)
      {
        if (k > first && b[k])
        {
          if (sp >= kWaveStack)
          {
            overflow = true;
            return -1;
          }
          // every active lane writes the same entry, so each lane later reads back its own store
          ws.st_code[sp] = __float_as_int (nd[48 + k]);
          ws.st_mask[sp] = b[k];
          ++sp;
        }
      }
      cur = __float_as_int (nd[48 + first]);
      curmask = b[first];
    }
    else
    {
      // leaf: queue it for the lanes that reach it
      const uint64_t take = curmask & searching;
      if (take & me) ws.cand[ncand++ * 64 + lane] = -cur - 1;
      const bool full = __ballot (ncand == kCandQ) != 0;
      const bool all_have = (__ballot (ncand > 0) & searching) == searching;
      if (full || all_have) drain ();
      pop ();
    }
  }
  if (searching) drain ();
  return result;
}

} // namespace drx

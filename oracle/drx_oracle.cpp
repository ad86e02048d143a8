// drx oracle -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// A CPU restatement of Devil Ray's (dray's) volume-render / point-location path, written from the
// reference's source as a behavioural spec.  It exists only as the parity checker: tests/, the smoke
// entry point and the CPU-baseline leg of bench.py load it; the shipped library never does.
//
// Build: oracle/build_oracle.sh  (g++ -O2 -fopenmp -ffp-contract=off; x86-64 without FMA, which is
// what dray's own OpenMP build does, so float results are the reference CPU path's bits).
//
// Deviations from dray, deliberate and documented:
//   * The Morton sort is stable: equal codes keep index order (dray: std::sort, unstable,
//     src/dray/linear_bvh_builder.cpp:128-129).
//   * Newton's start (dray passes an uninitialised guess, src/dray/data_model/device_mesh.hpp:327-334
//     with use_init_guess = true) is the element's corner-affine estimate clamped into the candidate's
//     SubRef (corner_start below; the SubRef centre where that is singular), or, after
//     orc_set_newton_start(2), the SubRef centre itself (subref_center, data_model/subref.hpp:52-56).
//   * The FP64 instance is a clean FP64 restatement: the clamp and the inf-norm use fmin/fmax in
//     double (dray's FP64 build still rounds through fminf/fmaxf, src/dray/data_model/element.hpp:449,
//     src/dray/vec.hpp:229-238).
//   * Element order may vary per element (mixed-order meshes); dray has one order per mesh.
//
// Every entry point is extern "C" with plain pointers so ctypes can call it.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int32_t i32;
typedef uint32_t u32;
typedef int64_t i64;

namespace orc
{

// ------------------------------------------------------------------------------------------
// Range / AABB (src/dray/range.hpp:20-208, src/dray/aabb.hpp:23-234).  Float32, AoS order
// [xmin, xmax, ymin, ymax, zmin, zmax].
// ------------------------------------------------------------------------------------------
static inline float inf32 () { return std::numeric_limits<float>::infinity (); }

struct Range
{
  float mn = inf32 (), mx = -inf32 ();
  bool is_empty () const { return mn > mx; }
  void include (float v)            // range.hpp:71-75
  {
    mn = std::fmin (mn, v);
    mx = std::fmax (mx, v);
  }
  void include (const Range &o)     // range.hpp:77-85
  {
    if (!o.is_empty ())
    {
      include (o.mn);
      include (o.mx);
    }
  }
  float center () const { return is_empty () ? std::numeric_limits<float>::quiet_NaN () : 0.5f * (mn + mx); }  // :129-138
  float length () const { return is_empty () ? std::numeric_limits<float>::quiet_NaN () : mx - mn; }           // :149-159
  void scale (float s)              // range.hpp:161-173
  {
    if (is_empty ()) return;
    const float c = center ();
    const float delta = s * 0.5f * length ();
    include (c - delta);
    include (c + delta);
  }
};

struct AABB
{
  Range r[3];
  void include (const AABB &o) { for (int d = 0; d < 3; ++d) r[d].include (o.r[d]); }
  void include_pt (const float p[3]) { for (int d = 0; d < 3; ++d) r[d].include (p[d]); }
  float volume () const             // aabb.hpp:173-180
  {
    float v = 1.f;
    for (int d = 0; d < 3; ++d) v *= r[d].length ();
    return v;
  }
  float surface_area () const       // aabb.hpp:182-195
  {
    float l[3];
    for (int d = 0; d < 3; ++d) l[d] = r[d].length ();
    float a = 2.f * l[0] * l[1];
    a += 2.f * l[1] * l[2];
    a += 2.f * l[2] * l[0];
    return a;
  }
  void scale (float s) { for (int d = 0; d < 3; ++d) r[d].scale (s); }  // aabb.hpp:105-111
  void store (float *o) const { for (int d = 0; d < 3; ++d) { o[2 * d] = r[d].mn; o[2 * d + 1] = r[d].mx; } }
  void load (const float *o) { for (int d = 0; d < 3; ++d) { r[d].mn = o[2 * d]; r[d].mx = o[2 * d + 1]; } }
};

struct SubRef { float lo[3], hi[3]; };   // SubRef<3,Tensor> = Vec<Vec<Float,3>,2> (data_model/subref.hpp:34-35)

static inline int npe_of (int p) { return (p + 1) * (p + 1) * (p + 1); }

// ------------------------------------------------------------------------------------------
// a2: construct_bvh per-element split (data_model/mesh_utils.cpp:514-670), pick_binary_split
// (:482-506), split_subref<Tensor> (data_model/ref_space.hpp:150-171), split_inplace<Tensor>
// (data_model/elem_ops.hpp:1346-1406), positive_get_bounds (data_model/element.hpp:150-160).
// ------------------------------------------------------------------------------------------
static void net_bounds (const float *net, int npe, AABB &box)
{
  box = AABB ();
  for (int i = 0; i < npe; ++i) box.include_pt (net + 3 * i);
}

// de Casteljau split of a tensor net at 0.5 along `axis`; lower=true keeps [lo, mid] (dray's
// "Left" branch, f_lower_t_upper == false), lower=false keeps [mid, hi].
static void split_inplace (float *net, int p, int axis, bool lower)
{
  const int pp[4] = { 1, p + 1, (p + 1) * (p + 1), (p + 1) * (p + 1) * (p + 1) };
  const int stride = pp[axis], chunk_sz = pp[axis + 1], num_chunks = pp[3 - (axis + 1)];
  const float f = 0.5f;
  for (int chunk = 0; chunk < num_chunks; ++chunk)
  {
    float *d = net + 3 * chunk * chunk_sz;
    if (lower)
    {
      for (int ff = 1; ff <= p; ++ff)
        for (int ii = p; ii >= ff; --ii)
          for (int e = 0; e < stride; ++e)
            for (int c = 0; c < 3; ++c)
              d[3 * (ii * stride + e) + c] = d[3 * ((ii - 1) * stride + e) + c] * (1 - f) + d[3 * (ii * stride + e) + c] * f;
    }
    else
    {
      for (int fb = 1; fb <= p; ++fb)
        for (int ii = 0; ii <= p - fb; ++ii)
          for (int e = 0; e < stride; ++e)
            for (int c = 0; c < 3; ++c)
              d[3 * (ii * stride + e) + c] = d[3 * (ii * stride + e) + c] * (1 - f) + d[3 * ((ii + 1) * stride + e) + c] * f;
    }
  }
}

static const int kSplits = 32;            // 2 * (2 << 3), mesh_utils.cpp:528
static const float kBoxScale = (float)1.000001;  // mesh_utils.cpp:519 (double constant passed as float32)

// One element: net (npe Vec3 floats in dray dof order) -> 33 scaled boxes + 33 subrefs.
static void split_element (const float *net0, int p, std::vector<float> &scratch, AABB *boxes, SubRef *refs)
{
  const int npe = npe_of (p);
  scratch.resize ((size_t)(kSplits + 1) * npe * 3);
  float *nets = scratch.data ();
  std::memcpy (nets, net0, sizeof (float) * 3 * npe);
  net_bounds (nets, npe, boxes[0]);
  for (int d = 0; d < 3; ++d) { refs[0].lo[d] = 0.f; refs[0].hi[d] = 1.f; }
  int count = 1;
  for (int i = 0; i < kSplits; ++i)
  {
    int max_id = 0;
    float max_measure = boxes[0].volume ();
    if (max_measure == 0.f) max_measure = boxes[0].surface_area ();
    for (int b = 1; b < count; ++b)
    {
      float m = boxes[b].volume ();
      if (m == 0.f) m = boxes[b].surface_area ();
      if (m > max_measure) { max_id = b; max_measure = m; }
    }
    // pick_binary_split<Tensor>: longest ref axis, ties -> lowest axis
    SubRef &sr = refs[max_id];
    int axis = 0;
    float maxlen = sr.hi[0] - sr.lo[0];
    for (int d = 1; d < 3; ++d)
    {
      const float len = sr.hi[d] - sr.lo[d];
      if (len > maxlen) { axis = d; maxlen = len; }
    }
    // split_subref: lower half stays in place, complement goes to `count`
    const float sp = sr.lo[axis] * (1.0f - 0.5f) + sr.hi[axis] * 0.5f;
    SubRef comp = sr;
    sr.hi[axis] = sp;
    comp.lo[axis] = sp;
    refs[count] = comp;
    float *mother = nets + (size_t)max_id * npe * 3;
    float *dghter = nets + (size_t)count * npe * 3;
    std::memcpy (dghter, mother, sizeof (float) * 3 * npe);
    split_inplace (mother, p, axis, true);
    split_inplace (dghter, p, axis, false);
    net_bounds (mother, npe, boxes[max_id]);
    net_bounds (dghter, npe, boxes[count]);
    ++count;
  }
  for (int i = 0; i <= kSplits; ++i) boxes[i].scale (kBoxScale);
}

// ------------------------------------------------------------------------------------------
// a3-a9: LinearBVHBuilder (linear_bvh_builder.cpp:19-518), morton (morton_codes.hpp:16-51),
// clz (math.hpp:155-187).
// ------------------------------------------------------------------------------------------
static inline u32 expand_bits32 (u32 x)
{
  x = (x | (x << 16)) & 0x030000FF;
  x = (x | (x << 8)) & 0x0300F00F;
  x = (x | (x << 4)) & 0x030C30C3;
  x = (x | (x << 2)) & 0x09249249;
  return x;
}

static inline u32 morton_3d (float x, float y, float z)
{
  x = std::fmin (std::fmax (x * 1024.0f, 0.0f), 1023.0f);
  y = std::fmin (std::fmax (y * 1024.0f, 0.0f), 1023.0f);
  z = std::fmin (std::fmax (z * 1024.0f, 0.0f), 1023.0f);
  return (expand_bits32 ((u32)z) << 2) | (expand_bits32 ((u32)y) << 1) | expand_bits32 ((u32)x);
}

static inline int clz32 (u32 x)
{
  u32 y, n = 32;
  y = x >> 16; if (y != 0) { n -= 16; x = y; }
  y = x >> 8;  if (y != 0) { n -= 8;  x = y; }
  y = x >> 4;  if (y != 0) { n -= 4;  x = y; }
  y = x >> 2;  if (y != 0) { n -= 2;  x = y; }
  y = x >> 1;  if (y != 0) return (int)(n - 2);
  return (int)(n - x);
}

static inline int delta (int a, int b, int inner_size, const u32 *mc)   // :149-166
{
  const bool oor = (b < 0 || b > inner_size);
  const int bb = oor ? 0 : b;
  u32 ex = mc[a] ^ mc[bb];
  const bool tie = (ex == 0);
  ex = tie ? ((u32)a ^ (u32)bb) : ex;
  int count = clz32 (ex);
  if (tie) count += 32;
  return oor ? -1 : count;
}

struct BuildOut
{
  std::vector<u32> mcodes_unsorted, mcodes_sorted;
  std::vector<i32> ids, leafs, left, right, parent;
  std::vector<float> inner_nodes;   // 16 floats per inner node
  AABB bounds;
};

static void bvh_build (const float *aabbs_in, const i32 *prims_in, int n_in, BuildOut &out)
{
  std::vector<AABB> aabbs;
  std::vector<i32> prims;
  const bool zero_case = (n_in == 0);
  if (n_in == 0)
  {
    AABB inv; const float z[3] = { 0, 0, 0 }; inv.include_pt (z);
    aabbs = { inv, inv };
    prims = { 0, 0 };
  }
  else if (n_in == 1)
  {
    AABB a0; a0.load (aabbs_in);
    AABB inv; const float z[3] = { 0, 0, 0 }; inv.include_pt (z);
    aabbs = { a0, inv };
    prims = { prims_in[0], prims_in[0] };
  }
  else
  {
    aabbs.resize (n_in);
    prims.assign (prims_in, prims_in + n_in);
    for (int i = 0; i < n_in; ++i) aabbs[i].load (aabbs_in + 6 * i);
  }
  const int n = (int)aabbs.size ();
  // reduce (:19-58)
  AABB bounds;
  if (!zero_case)
  {
    float mn[3] = { inf32 (), inf32 (), inf32 () }, mx[3] = { -inf32 (), -inf32 (), -inf32 () };
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d)
      {
        mn[d] = std::min (mn[d], aabbs[i].r[d].mn);
        mx[d] = std::max (mx[d], aabbs[i].r[d].mx);
      }
    bounds.include_pt (mn);
    bounds.include_pt (mx);
  }
  out.bounds = bounds;
  // get_mcodes (:60-90)
  float minc[3], inv_ext[3];
  for (int d = 0; d < 3; ++d)
  {
    minc[d] = bounds.r[d].mn;
    const float ext = bounds.r[d].mx - bounds.r[d].mn;
    inv_ext[d] = (ext == .0f) ? 0.f : 1.f / ext;
  }
  std::vector<u32> mc (n);
  for (int i = 0; i < n; ++i)
  {
    const float cx = (aabbs[i].r[0].center () - minc[0]) * inv_ext[0];
    const float cy = (aabbs[i].r[1].center () - minc[1]) * inv_ext[1];
    const float cz = (aabbs[i].r[2].center () - minc[2]) * inv_ext[2];
    mc[i] = morton_3d (cx, cy, cz);
  }
  out.mcodes_unsorted = mc;
  // sort_mcodes (:120-135), stable
  std::vector<i32> ids (n);
  std::iota (ids.begin (), ids.end (), 0);
  std::stable_sort (ids.begin (), ids.end (), [&] (i32 a, i32 b) { return mc[a] < mc[b]; });
  std::vector<u32> smc (n);
  std::vector<AABB> saabb (n);
  std::vector<i32> sprim (n);
  for (int i = 0; i < n; ++i)
  {
    smc[i] = mc[ids[i]];
    saabb[i] = aabbs[ids[i]];
    sprim[i] = prims[ids[i]];
  }
  out.mcodes_sorted = smc;
  out.ids = ids;
  out.leafs = sprim;
  // build_tree (:168-251), Karras 2012
  const int inner = n - 1;
  std::vector<i32> left (inner), right (inner), parent (2 * n - 1);
  for (int i = 0; i < inner; ++i)
  {
    const int d = 0 > (delta (i, i + 1, inner, smc.data ()) - delta (i, i - 1, inner, smc.data ())) ? -1 : 1;
    const int min_delta = delta (i, i - d, inner, smc.data ());
    int lmax = 2;
    while (delta (i, i + lmax * d, inner, smc.data ()) > min_delta) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
      if (delta (i, i + (l + t) * d, inner, smc.data ()) > min_delta) l += t;
    const int j = i + l * d;
    const int delta_node = delta (i, j, inner, smc.data ());
    int s = 0;
    float div = 2.f;
    for (int t = (int)std::ceil ((float)l / div);; div *= 2, t = (int)std::ceil ((float)l / div))
    {
      if (delta (i, i + (s + t) * d, inner, smc.data ()) > delta_node) s += t;
      if (t == 1) break;
    }
    const int split = i + s * d + std::min (d, 0);
    if (std::min (i, j) == split) { parent[split + inner] = i; left[i] = split + inner; }
    else { parent[split] = i; left[i] = split; }
    if (std::max (i, j) == split + 1) { parent[split + inner + 1] = i; right[i] = split + inner + 1; }
    else { parent[split + 1] = i; right[i] = split + 1; }
    if (i == 0) parent[0] = -1;
  }
  // propagate_aabbs (:253-317): min/max merges are order independent, so a post-order pass gives
  // the atomic bottom-up result exactly.
  std::vector<AABB> inner_box (inner);
  {
    std::vector<int> counter (inner, 0);
    for (int leaf = 0; leaf < n; ++leaf)
    {
      int cur = parent[inner + leaf];
      while (cur != -1)
      {
        if (counter[cur]++ == 0) break;
        AABB b;
        const int lc = left[cur], rc = right[cur];
        b.include (lc >= inner ? saabb[lc - inner] : inner_box[lc]);
        b.include (rc >= inner ? saabb[rc - inner] : inner_box[rc]);
        inner_box[cur] = b;
        cur = parent[cur];
      }
    }
  }
  // emit (:319-397)
  out.inner_nodes.assign ((size_t)inner * 16, 0.f);
  for (int node = 0; node < inner; ++node)
  {
    AABB lb, rb;
    int lc = left[node], rc = right[node];
    if (lc >= inner) { lb = saabb[lc - inner]; lc = -(lc - inner + 1); }
    else { lb = inner_box[lc]; lc *= 4; }
    if (rc >= inner) { rb = saabb[rc - inner]; rc = -(rc - inner + 1); }
    else { rb = inner_box[rc]; rc *= 4; }
    float *o = &out.inner_nodes[(size_t)node * 16];
    o[0] = lb.r[0].mn; o[1] = lb.r[1].mn; o[2] = lb.r[2].mn; o[3] = lb.r[0].mx;
    o[4] = lb.r[1].mx; o[5] = lb.r[2].mx; o[6] = rb.r[0].mn; o[7] = rb.r[1].mn;
    o[8] = rb.r[2].mn; o[9] = rb.r[0].mx; o[10] = rb.r[1].mx; o[11] = rb.r[2].mx;
    std::memcpy (&o[12], &lc, 4);
    std::memcpy (&o[13], &rc, 4);
  }
  out.left = left;
  out.right = right;
  out.parent = parent;
}

// ------------------------------------------------------------------------------------------
// a13: Element_impl<3,NC,Tensor,General>::eval_d (data_model/pos_tensor_element.tcc:99-232) with
// binomials C(p-1,k) (integer_utils.hpp:18-70).  dofs: AoS NC values per dof, dray dof order.
// ------------------------------------------------------------------------------------------
template <typename T, int NC> struct DofV { T v[NC]; };

template <typename T, int NC>
static void eval_d (const T *dofs, int p, const T ref[3], T val[NC], T deriv[3][NC])
{
  typedef DofV<T, NC> D;
  auto zero = [] () { D z; for (int c = 0; c < NC; ++c) z.v[c] = T (0); return z; };
  auto axpy = [] (const D &a, T s, const D &b, T t) { D r; for (int c = 0; c < NC; ++c) r.v[c] = a.v[c] * s + b.v[c] * t; return r; };
  auto sub_scale = [] (const D &a, const D &b, T s) { D r; for (int c = 0; c < NC; ++c) r.v[c] = (a.v[c] - b.v[c]) * s; return r; };
  auto load = [&] (int idx) { D r; for (int c = 0; c < NC; ++c) r.v[c] = dofs[NC * idx + c]; return r; };

  const T u = ref[0], v = ref[1], w = ref[2];
  const T ubar = T (1.0 - u), vbar = T (1.0 - v), wbar = T (1.0 - w);
  const int p1 = p, p2 = p, p3 = p;
  int B[32];
  if (p >= 1)
  {
    i64 val_b = 1; int ijk0 = 0, ijk1 = p - 1;
    for (int ii = 0; ii <= p - 1; ++ii)
    {
      B[ii] = (int)val_b;
      i64 t = val_b; t *= ijk1; ijk1--; ijk0++; if (ijk0) t /= ijk0; val_b = t;  // slide_over(0)
    }
  }
  int cidx = 0;
  D val_u, val_v, val_w, deriv_u, duv[2], duvw[3];
  T wpow = 1.0;
  D wL[3], wR[3];
  for (int q = 0; q < 3; ++q) { wL[q] = zero (); wR[q] = zero (); }
  for (int ii = 0; ii <= p3; ++ii)
  {
    T vpow = 1.0;
    D vL[2], vR[2];
    vL[0] = vL[1] = vR[0] = vR[1] = zero ();
    for (int jj = 0; jj <= p2; ++jj)
    {
      T upow = 1.0;
      D uL = zero (), uR = zero ();
      D C = load (cidx++);
      for (int kk = 1; kk <= p1; ++kk)
      {
        const T bu = T (B[kk - 1]) * upow;
        for (int c = 0; c < NC; ++c) uL.v[c] = uL.v[c] * ubar + C.v[c] * bu;
        C = load (cidx++);
        for (int c = 0; c < NC; ++c) uR.v[c] = uR.v[c] * ubar + C.v[c] * bu;
        upow *= u;
      }
      val_u = (p1 > 0) ? axpy (uL, ubar, uR, u) : C;
      deriv_u = sub_scale (uR, uL, T (p1));
      if (jj > 0)
      {
        const T bv = T (B[jj - 1]) * vpow;
        for (int c = 0; c < NC; ++c)
        {
          vR[0].v[c] = vR[0].v[c] * vbar + val_u.v[c] * bv;
          vR[1].v[c] = vR[1].v[c] * vbar + deriv_u.v[c] * bv;
        }
        vpow *= v;
      }
      if (jj < p2)
      {
        const T bv = T (B[jj]) * vpow;
        for (int c = 0; c < NC; ++c)
        {
          vL[0].v[c] = vL[0].v[c] * vbar + val_u.v[c] * bv;
          vL[1].v[c] = vL[1].v[c] * vbar + deriv_u.v[c] * bv;
        }
      }
    }
    val_v = (p2 > 0) ? axpy (vL[0], vbar, vR[0], v) : val_u;
    duv[0] = (p2 > 0) ? axpy (vL[1], vbar, vR[1], v) : deriv_u;
    duv[1] = sub_scale (vR[0], vL[0], T (p2));
    if (ii > 0)
    {
      const T bw = T (B[ii - 1]) * wpow;
      for (int c = 0; c < NC; ++c)
      {
        wR[0].v[c] = wR[0].v[c] * wbar + val_v.v[c] * bw;
        wR[1].v[c] = wR[1].v[c] * wbar + duv[0].v[c] * bw;
        wR[2].v[c] = wR[2].v[c] * wbar + duv[1].v[c] * bw;
      }
      wpow *= w;
    }
    if (ii < p3)
    {
      const T bw = T (B[ii]) * wpow;
      for (int c = 0; c < NC; ++c)
      {
        wL[0].v[c] = wL[0].v[c] * wbar + val_v.v[c] * bw;
        wL[1].v[c] = wL[1].v[c] * wbar + duv[0].v[c] * bw;
        wL[2].v[c] = wL[2].v[c] * wbar + duv[1].v[c] * bw;
      }
    }
  }
  val_w = (p3 > 0) ? axpy (wL[0], wbar, wR[0], w) : val_v;
  duvw[0] = (p3 > 0) ? axpy (wL[1], wbar, wR[1], w) : duv[0];
  duvw[1] = (p3 > 0) ? axpy (wL[2], wbar, wR[2], w) : duv[1];
  duvw[2] = sub_scale (wR[0], wL[0], T (p3));
  for (int c = 0; c < NC; ++c)
  {
    val[c] = val_w.v[c];
    for (int q = 0; q < 3; ++q) deriv[q][c] = duvw[q].v[c];
  }
}

// ------------------------------------------------------------------------------------------
// a14: matrix_mult_inv = Crout LUP with partial pivoting + solve (matrix.hpp:221-373,461-477).
// A(row,col); epsilon<float>=1e-4f, epsilon<double>=(double)1e-8f (math.hpp:25-26,58-66).
// ------------------------------------------------------------------------------------------
template <typename T> static inline T lup_eps ();
template <> inline float lup_eps<float> () { return 1e-4f; }
template <> inline double lup_eps<double> () { return (double)1e-8f; }

template <typename T>
static void lup_solve (const T Ain[3][3], const T b[3], T x[3], bool &valid)
{
  T A[3][3];
  std::memcpy (A, Ain, sizeof (A));
  int perm[3] = { 0, 1, 2 };
  valid = true;
  for (int top = 0; top < 3; ++top)
  {
    int maxr = top;
    T maxv = std::fabs (A[maxr][top]);
    for (int r = top + 1; r < 3; ++r)
    {
      const T cv = std::fabs (A[r][top]);
      if (maxv < cv) { maxv = cv; maxr = r; }
    }
    if (maxv < lup_eps<T> ()) valid = false;
    if (maxr != top)
    {
      for (int c = 0; c < 3; ++c) std::swap (A[maxr][c], A[top][c]);
      std::swap (perm[maxr], perm[top]);
    }
    for (int c = top + 1; c < 3; ++c) A[top][c] /= A[top][top];
    for (int r = top + 1; r < 3; ++r)
      for (int c = top + 1; c < 3; ++c) A[r][c] -= A[r][top] * A[top][c];
  }
  T y[3];
  for (int r = 0; r < 3; ++r)
  {
    y[r] = b[perm[r]];
    for (int c = 0; c < r; ++c) y[r] -= A[r][c] * y[c];
    y[r] /= A[r][r];
  }
  for (int r = 2; r >= 0; --r)
  {
    x[r] = y[r];
    for (int c = r + 1; c < 3; ++c) x[r] -= A[r][c] * x[c];
  }
}

// ------------------------------------------------------------------------------------------
// a12/a15: eval_inverse_local + IterativeMethod::solve (element.hpp:435-486,
// newton_solver.hpp:50-73), QuadRefSpace::is_inside (pos_tensor_element.tcc:49-61).
// ------------------------------------------------------------------------------------------
template <typename T> static inline T clamp01 (T v) { return std::fmin (T (1), std::fmax (v, T (0))); }

template <typename T>
static bool newton_inverse (const T *dofs, int p, const T target[3], T x[3], int *steps_out)
{
  const T tol = T (1e-4f);
  const int max_steps = 20;
  int steps = 0;
  bool converged = false;
  T prev[3] = { x[0], x[1], x[2] };
  while (steps < max_steps && !converged)
  {
    for (int i = 0; i < 3; ++i) x[i] = clamp01 (x[i]);
    T y[3], J[3][3];
    eval_d<T, 3> (dofs, p, x, y, J);
    T dy[3], A[3][3], dx[3];
    for (int i = 0; i < 3; ++i) dy[i] = target[i] - y[i];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) A[r][c] = J[c][r];   // set_col(c, d/dref_c)
    bool valid;
    lup_solve<T> (A, dy, dx, valid);
    for (int i = 0; i < 3; ++i) x[i] = x[i] + dx[i];
    ++steps;
    T res = std::fmax (-(x[0] - prev[0]), x[0] - prev[0]);
    for (int i = 1; i < 3; ++i) res = std::fmax (res, std::fmax (-(x[i] - prev[i]), x[i] - prev[i]));
    converged = res < tol;
    for (int i = 0; i < 3; ++i) prev[i] = x[i];
  }
  if (steps_out) *steps_out = steps;
  T mn = 2, mx = -1;
  for (int d = 0; d < 3; ++d) { mn = std::min (x[d], mn); mx = std::max (x[d], mx); }
  const bool inside = (mn >= T (0.f) - tol) && (mx <= T (1.f) + tol);
  return converged && inside;
}

// ------------------------------------------------------------------------------------------
// Mesh / BVH views for the query routines.
// ------------------------------------------------------------------------------------------
template <typename T> struct MeshView
{
  const i32 *ctrl_idx;   // dray GridFunction: element e's dofs at ctrl_idx[off(e) + local]
  const T *values;       // NC values per control point
  const i32 *orders;     // nullable: per-element order
  const i64 *offsets;    // nullable: per-element start into ctrl_idx
  int order;
  int n_el;
  int order_of (int e) const { return orders ? orders[e] : order; }
  i64 off_of (int e) const { return offsets ? offsets[e] : (i64)e * npe_of (order); }
  template <int NC> void gather (int e, T *out) const
  {
    const int p = order_of (e), n = npe_of (p);
    const i64 o = off_of (e);
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < NC; ++c) out[NC * i + c] = values[(i64)NC * ctrl_idx[o + i] + c];
  }
};

struct BVHView
{
  const float *inner;    // 16 floats per node
  const i32 *leaf_nodes;
  const i32 *aabb_ids;
  const float *subrefs;  // 6 floats per original aabb: lo xyz, hi xyz
};

struct LocStats { i64 newton_steps = 0; i64 candidates = 0; };

// a10/a11: DeviceMesh::locate (device_mesh.hpp:247-354)
// Newton's start for a candidate (dray passes an uninitialised guess, device_mesh.hpp:327-334; the GPU
// path uses the same rule): the element's affine approximation at its centre -- the trilinear map through
// its 8 corner control points, differentiated at the reference centre -- solved for the point by Cramer's
// rule and clamped into the candidate's SubRef; the SubRef centre where that is singular.
template <typename T>
static void corner_start (const T *dofs, int p, const T pt[3], const float *sr, T x[3])
{
  const int n1 = p + 1;
  T c[8][3];
  for (int q = 0; q < 8; ++q)
  {
    const int idx = (q & 1) * p + n1 * (((q >> 1) & 1) * p) + n1 * n1 * ((q >> 2) * p);
    for (int d = 0; d < 3; ++d) c[q][d] = dofs[3 * idx + d];
  }
  T r[3], eu[3], ev[3], ew[3];
  for (int d = 0; d < 3; ++d)
  {
    const T m = (((c[0][d] + c[1][d]) + (c[2][d] + c[3][d])) + ((c[4][d] + c[5][d]) + (c[6][d] + c[7][d]))) * T (0.125);
    r[d] = pt[d] - m;
    eu[d] = (((c[1][d] - c[0][d]) + (c[3][d] - c[2][d])) + ((c[5][d] - c[4][d]) + (c[7][d] - c[6][d]))) * T (0.25);
    ev[d] = (((c[2][d] - c[0][d]) + (c[3][d] - c[1][d])) + ((c[6][d] - c[4][d]) + (c[7][d] - c[5][d]))) * T (0.25);
    ew[d] = (((c[4][d] - c[0][d]) + (c[5][d] - c[1][d])) + ((c[6][d] - c[2][d]) + (c[7][d] - c[3][d]))) * T (0.25);
  }
  const T vw[3] = { ev[1] * ew[2] - ev[2] * ew[1], ev[2] * ew[0] - ev[0] * ew[2], ev[0] * ew[1] - ev[1] * ew[0] };
  const T rw[3] = { r[1] * ew[2] - r[2] * ew[1], r[2] * ew[0] - r[0] * ew[2], r[0] * ew[1] - r[1] * ew[0] };
  const T vr[3] = { ev[1] * r[2] - ev[2] * r[1], ev[2] * r[0] - ev[0] * r[2], ev[0] * r[1] - ev[1] * r[0] };
  const T det = (eu[0] * vw[0] + eu[1] * vw[1]) + eu[2] * vw[2];
  const T inv = T (1) / det;
  const T dx[3] = { ((r[0] * vw[0] + r[1] * vw[1]) + r[2] * vw[2]) * inv, ((eu[0] * rw[0] + eu[1] * rw[1]) + eu[2] * rw[2]) * inv,
                    ((eu[0] * vr[0] + eu[1] * vr[1]) + eu[2] * vr[2]) * inv };
  for (int d = 0; d < 3; ++d)
  {
    const T v = T (0.5) + dx[d];
    const T lo = T (sr[d]), hi = T (sr[3 + d]);
    x[d] = (v == v && std::fabs (v) < T (1e30)) ? std::min (hi, std::max (lo, v)) : T ((sr[d] + sr[3 + d]) * 0.5f);
  }
}

// Newton start rule of locate_point: 0 = corner_start, 2 = the candidate's SubRef centre, subref_center
// (data_model/subref.hpp:52-56), which is dray's start when use_init_guess is false.  Set by
// orc_set_newton_start; read-only while a locate runs.
static int g_start_mode = 0;

template <typename T>
static i32 locate_point (const MeshView<T> &mesh, const BVHView &bvh, const T pt[3], T ref_out[3],
                         std::vector<T> &dofs, LocStats *st)
{
  i32 todo[64];
  i32 cur = 0, sp = 0;
  const i32 barrier = -2000000000;
  todo[sp] = barrier;
  ref_out[0] = ref_out[1] = ref_out[2] = T (-1);
  while (cur != barrier)
  {
    if (cur > -1)
    {
      const float *nd = bvh.inner + (size_t)cur * 4;   // cur is the float4 offset
      bool in_l = true, in_r = true;
      if (pt[0] < nd[0]) in_l = false;
      if (pt[1] < nd[1]) in_l = false;
      if (pt[2] < nd[2]) in_l = false;
      if (pt[0] > nd[3]) in_l = false;
      if (pt[1] > nd[4]) in_l = false;
      if (pt[2] > nd[5]) in_l = false;
      if (pt[0] < nd[6]) in_r = false;
      if (pt[1] < nd[7]) in_r = false;
      if (pt[2] < nd[8]) in_r = false;
      if (pt[0] > nd[9]) in_r = false;
      if (pt[1] > nd[10]) in_r = false;
      if (pt[2] > nd[11]) in_r = false;
      if (!in_l && !in_r)
      {
        cur = todo[sp];
        --sp;
      }
      else
      {
        i32 lc, rc;
        std::memcpy (&lc, &nd[12], 4);
        std::memcpy (&rc, &nd[13], 4);
        cur = in_l ? lc : rc;
        if (in_l && in_r) todo[++sp] = rc;
      }
    }
    else
    {
      const i32 leaf = -cur - 1;
      const i32 el = bvh.leaf_nodes[leaf];
      const float *sr = bvh.subrefs + (size_t)6 * bvh.aabb_ids[leaf];
      const int p = mesh.order_of (el);
      dofs.resize ((size_t)3 * npe_of (p));
      mesh.template gather<3> (el, dofs.data ());
      T x[3];
      if (g_start_mode == 2)
        for (int d = 0; d < 3; ++d) x[d] = T ((sr[d] + sr[3 + d]) * 0.5f);
      else
        corner_start<T> (dofs.data (), p, pt, sr, x);
      int steps = 0;
      const bool found = newton_inverse<T> (dofs.data (), p, pt, x, &steps);
      if (st) { st->newton_steps += steps; st->candidates += 1; }
      if (found)
      {
        for (int d = 0; d < 3; ++d) ref_out[d] = x[d];
        return el;
      }
      cur = todo[sp];
      --sp;
    }
  }
  return -1;
}

template <typename T>
static T field_eval (const MeshView<T> &field, int el, const T ref[3], std::vector<T> &dofs)
{
  const int p = field.order_of (el);
  dofs.resize (npe_of (p));
  field.template gather<1> (el, dofs.data ());
  T val[1], der[3][1];
  eval_d<T, 1> (dofs.data (), p, ref, val, der);
  return val[0];
}

// ------------------------------------------------------------------------------------------
// Rays (ray.hpp:20-28), ray setup (ray.cpp:102-241), colour map (device_color_map.hpp:61-77),
// blend (rendering/colors.hpp:15-23), integrate (rendering/volume.cpp:69-228).
// ------------------------------------------------------------------------------------------
struct Ray { float dir[3]; float near_; float orig[3]; float far_; i32 pixel_id; };
static_assert (sizeof (Ray) == 36, "Ray is 36 bytes");
struct Partial { i32 pixel_id; float depth; float rgba[4]; };
static_assert (sizeof (Partial) == 24, "VolumePartial is 24 bytes");

static inline float rcp_safe (float f) { return 1.0f / ((std::fabs (f) < 1e-8f) ? (std::signbit (f) ? -1e-8f : 1e-8f) : f); }

static bool ray_slab (const Ray &ray, const AABB &mb, float &near_o, float &far_o)
{
  const float idx = rcp_safe (ray.dir[0]), idy = rcp_safe (ray.dir[1]), idz = rcp_safe (ray.dir[2]);
  const float ox = ray.orig[0] * idx, oy = ray.orig[1] * idy, oz = ray.orig[2] * idz;
  const float xmin = mb.r[0].mn * idx - ox, ymin = mb.r[1].mn * idy - oy, zmin = mb.r[2].mn * idz - oz;
  const float xmax = mb.r[0].mx * idx - ox, ymax = mb.r[1].mx * idy - oy, zmax = mb.r[2].mx * idz - oz;
  const float min_int = ray.near_;
  const float mind = std::max (std::max (std::max (std::min (ymin, ymax), std::min (xmin, xmax)), std::min (zmin, zmax)), min_int);
  float maxd = std::min (std::min (std::max (ymin, ymax), std::max (xmin, xmax)), std::max (zmin, zmax));
  maxd = std::min (maxd, ray.far_);
  near_o = mind;
  far_o = maxd;
  return maxd > mind;
}

template <typename T> struct ColorMapT
{
  const float *lut;   // rgba per entry
  int size;
  T mn, mx, inv_range;
  void color (T scalar, float out[4]) const
  {
    const T s = std::min (mx, std::max (mn, scalar));   // clamp (math.hpp:235-238)
    const float normalized = static_cast<float> ((s - mn) * inv_range);
    int idx = static_cast<i32> (normalized * float (size - 1));
    idx = std::min (std::max (idx, 0), size - 1);
    for (int c = 0; c < 4; ++c) out[c] = lut[4 * idx + c];
  }
};

static inline void blend (float f[4], const float b[4])
{
  const float a = b[3] * (1.f - f[3]);
  f[0] = f[0] + b[0] * a;
  f[1] = f[1] + b[1] * a;
  f[2] = f[2] + b[2] * a;
  f[3] = a + f[3];
}

struct RayStats { i64 located = 0; i64 newton = 0; i64 candidates = 0; };

// Vec::normalize (vec.hpp:211-221): magnitude = sqrtf(sum of squares), then component-wise division.
static void vnormalize (float v[3])
{
  float s = 0.f;
  for (int i = 0; i < 3; ++i) s += v[i] * v[i];
  const float m = sqrtf (s);
  for (int i = 0; i < 3; ++i) v[i] /= m;
}

// PointLight (rendering/point_light.hpp:15-33); dray's defaults: amb 0.4, diff 0.75, spec 0.8, spec_pow 15.
struct Light { float pos[3], amb[3], diff[3], spec[3], spec_pow; };
static_assert (sizeof (Light) == 52, "PointLight is 13 floats");

static inline float dot3 (const float a[3], const float b[3])
{
  float s = a[0] * b[0];
  s += a[1] * b[1];
  s += a[2] * b[2];
  return s;
}
static inline float clampf (float v, float lo, float hi) { return std::min (hi, std::max (lo, v)); }   // math.hpp:235-238

// VolumeShader::scalar_gradient + shaded_color (rendering/volume_shader.hpp:39-158): physical field
// gradient = ref-space gradient x J^-1 (matrix_inverse, matrix.hpp:431-457: LUP, one solve per unit
// column), then per light ambient + diffuse + Blinn-Phong specular on the colour-map sample, with the
// gradient flipped toward the light / half vector (the flips persist across lights) and the colour
// channels clamped to [0, 1] after each light.  Samples with alpha <= 0.01 contribute nothing.
template <typename T>
static void shaded_color (const MeshView<T> &mesh, const MeshView<T> &field, int el, const T ref[3], const Ray &ray,
                          const ColorMapT<T> &cmap, const Light *lights, int n_lights, std::vector<T> &scratch,
                          float out[4])
{
  const int p = mesh.order_of (el);
  const int npe = npe_of (p);
  scratch.resize ((size_t)4 * npe);
  T *gd = scratch.data (), *fd = gd + 3 * npe;
  mesh.template gather<3> (el, gd);
  field.template gather<1> (el, fd);
  T wp[3], jac[3][3], sv[1], fdr[3][1];
  eval_d<T, 3> (gd, p, ref, wp, jac);
  eval_d<T, 1> (fd, p, ref, sv, fdr);
  T J[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) J[r][c] = jac[c][r];   // set_col(rdim, d phi / d ref_rdim)
  T inv[3][3];
  for (int c = 0; c < 3; ++c)
  {
    T e[3] = { T (0), T (0), T (0) };
    e[c] = T (1);
    T x[3];
    bool valid;
    lup_solve<T> (J, e, x, valid);
    for (int r = 0; r < 3; ++r) inv[r][c] = x[r];
  }
  T grad[3];
  for (int c = 0; c < 3; ++c)
  {
    T sum = fdr[0][0] * inv[0][c];
    sum = sum + fdr[1][0] * inv[1][c];
    sum = sum + fdr[2][0] * inv[2][c];
    grad[c] = sum;
  }
  float col[4];
  cmap.color (sv[0], col);
  float acc[4] = { 0.f, 0.f, 0.f, 0.f };
  if (col[3] > 0.01)
  {
    T m2 = grad[0] * grad[0];
    m2 += grad[1] * grad[1];
    m2 += grad[2] * grad[2];
    const T mag = std::sqrt (m2);
    float fg[3];
    for (int i = 0; i < 3; ++i) fg[i] = float (grad[i] / mag);
    const float view[3] = { -ray.dir[0], -ray.dir[1], -ray.dir[2] };
    for (int l = 0; l < n_lights; ++l)
    {
      const Light &L = lights[l];
      float ld[3];
      for (int i = 0; i < 3; ++i) ld[i] = L.pos[i] - float (wp[i]);
      vnormalize (ld);
      if (dot3 (ld, fg) < 0.f)
        for (int i = 0; i < 3; ++i) fg[i] = -fg[i];
      const float diffuse = clampf (dot3 (ld, fg), 0.f, 1.f);
      float sc[4] = { L.amb[0] * col[0], L.amb[1] * col[1], L.amb[2] * col[2], col[3] };
      for (int c = 0; c < 3; ++c) sc[c] += diffuse * L.diff[c] * col[c];
      float hv[3];
      for (int i = 0; i < 3; ++i) hv[i] = 0.5f * (view[i] + ld[i]);
      vnormalize (hv);
      if (dot3 (fg, hv) < 0.f)
        for (int i = 0; i < 3; ++i) fg[i] = -fg[i];
      const float doth = clampf (dot3 (fg, hv), 0.f, 1.f);
      const float intensity = std::pow (doth, L.spec_pow);
      for (int c = 0; c < 3; ++c) sc[c] += intensity * L.spec[c] * col[c];
      for (int c = 0; c < 4; ++c) acc[c] += sc[c];
      for (int c = 0; c < 3; ++c) acc[c] = clampf (acc[c], 0.f, 1.f);
    }
  }
  for (int c = 0; c < 4; ++c) out[c] = acc[c];
}

// T: geometry, reference coordinates and distances; F: the scalar field and the colour map (F = T as in
// dray, where both are Float; F = double over T = float for an FP64 field on FP32 geometry).
// mesh64 (nullable): locate each sample point (computed in T) with the FP64 walk and Newton on the same
// geometry in double, the reference point then rounded to T: which samples are located, and where, as
// exact arithmetic decides at the sample points of a T-precision march.
template <typename T, typename F = T>
static void integrate_ray (const MeshView<T> &mesh, const MeshView<F> &field, const BVHView &bvh,
                           const Ray &ray, T sample_dist, const ColorMapT<F> &cmap, const Light *lights,
                           int n_lights, Partial *out5, std::vector<T> &scratch, RayStats &rs,
                           const MeshView<double> *mesh64 = nullptr)
{
  std::vector<F> fscratch;
  std::vector<double> scratch64;
  auto locate = [&] (const T pt[3], T ref[3], LocStats *ls) -> i32 {
    if (!mesh64) return locate_point<T> (mesh, bvh, pt, ref, scratch, ls);
    const double p64[3] = { double (pt[0]), double (pt[1]), double (pt[2]) };
    double r64[3] = { 0, 0, 0 };
    const i32 c = locate_point<double> (*mesh64, bvh, p64, r64, scratch64, ls);
    for (int d = 0; d < 3; ++d) ref[d] = T (r64[d]);
    return c;
  };
  for (int s = 0; s < 5; ++s) { out5[s].pixel_id = -1; out5[s].depth = inf32 (); for (int c = 0; c < 4; ++c) out5[s].rgba[c] = 0.f; }
  T distance = T (ray.near_) + sample_dist;
  int segment = 0;
  Partial partial;
  partial.pixel_id = ray.pixel_id;
  for (int c = 0; c < 4; ++c) partial.rgba[c] = 0.f;
  LocStats ls;
  for (int s = 0; s < 5; ++s)
  {
    bool found = false;
    i32 cell = -1;
    T ref[3];
    while (distance < T (ray.far_) && !found)
    {
      T pt[3];
      for (int d = 0; d < 3; ++d) pt[d] = T (ray.orig[d]) + distance * T (ray.dir[d]);
      cell = locate (pt, ref, &ls);
      if (cell != -1) { found = true; rs.located++; }
      else distance += sample_dist;
    }
    if (distance >= T (ray.far_)) break;
    partial.depth = float (distance);
    for (int c = 0; c < 4; ++c) partial.rgba[c] = 0.f;
    do
    {
      float sc[4];
      if constexpr (std::is_same<T, F>::value)
      {
        if (lights)   // use_lighting: VolumeShader::shaded_color, else VolumeShader::color (volume.cpp:162-171)
          shaded_color<T> (mesh, field, cell, ref, ray, cmap, lights, n_lights, scratch, sc);
        else
          cmap.color (field_eval<T> (field, cell, ref, scratch), sc);
      }
      else
      {
        const F rf[3] = { F (ref[0]), F (ref[1]), F (ref[2]) };
        cmap.color (field_eval<F> (field, cell, rf, fscratch), sc);
      }
      blend (partial.rgba, sc);
      distance += sample_dist;
      T pt[3];
      for (int d = 0; d < 3; ++d) pt[d] = T (ray.orig[d]) + distance * T (ray.dir[d]);
      cell = locate (pt, ref, &ls);
      found = cell != -1;
      if (found) rs.located++;
    } while (distance < T (ray.far_) && found && partial.rgba[3] < 0.95f);
    out5[segment] = partial;
    ++segment;
    if (distance >= T (ray.far_) || partial.rgba[3] > 0.95f) break;
  }
  rs.newton += ls.newton_steps;
  rs.candidates += ls.candidates;
}

// Camera::gen_perspective per-pixel part (rendering/camera.cpp:457-481) and the host set-up
// (:402-440, set_fov :132-163).
struct CamSetup { float nlook[3], dx[3], dy[3], pos[3]; int w, h; };

static CamSetup camera_setup (int w, int h, float fov_y, const float pos[3], const float look_at[3], const float up_in[3], float zoom)
{
  const double pi = 3.14159265358979323846264338327950288;
  // set_fov
  const float fov_y_rad = (fov_y * static_cast<float> (pi)) / 180.0f;
  const float vertical = std::tan (0.5f * fov_y_rad);
  const float aspect = float (w) / float (h);
  const float horizontal = aspect * vertical;
  const float fov_x_rad = 2.0f * std::atan (horizontal);
  const float fov_x = 180.0f * (fov_x_rad / static_cast<float> (pi));
  float up[3] = { up_in[0], up_in[1], up_in[2] };
  vnormalize (up);   // set_up normalises
  float look[3];
  for (int i = 0; i < 3; ++i) look[i] = look_at[i] - pos[i];
  const float thx = tanf ((fov_x * float (pi) / 180.f) * .5f);
  const float thy = tanf ((fov_y * float (pi) / 180.f) * .5f);
  float ru[3] = { look[1] * up[2] - look[2] * up[1], look[2] * up[0] - look[0] * up[2], look[0] * up[1] - look[1] * up[0] };
  float ruf[3] = { ru[0], ru[1], ru[2] };
  vnormalize (ru);
  float rv[3] = { ruf[1] * look[2] - ruf[2] * look[1], ruf[2] * look[0] - ruf[0] * look[2], ruf[0] * look[1] - ruf[1] * look[0] };
  vnormalize (rv);
  CamSetup cs;
  for (int i = 0; i < 3; ++i)
  {
    cs.dx[i] = ru[i] * (2 * thx / (float)w);
    cs.dy[i] = rv[i] * (2 * thy / (float)h);
    if (zoom > 0) { cs.dx[i] = cs.dx[i] / zoom; cs.dy[i] = cs.dy[i] / zoom; }
    cs.nlook[i] = look[i];
    cs.pos[i] = pos[i];
  }
  vnormalize (cs.nlook);
  cs.w = w;
  cs.h = h;
  return cs;
}

static Ray camera_ray (const CamSetup &cs, int idx)
{
  Ray r;
  const int i = idx % cs.w, j = idx / cs.w;
  for (int d = 0; d < 3; ++d) r.orig[d] = cs.pos[d];
  r.near_ = 0.f;
  r.far_ = inf32 ();
  r.pixel_id = j * cs.w + i;
  const float a = (2.f * float (i) - float (cs.w)) / 2.0f;
  const float b = (2.f * float (j) - float (cs.w)) / 2.0f;   // dray uses w for both axes (camera.cpp:469-470)
  for (int d = 0; d < 3; ++d)
  {
    const float t1 = cs.dx[d] * a;
    const float t2 = cs.dy[d] * b;
    r.dir[d] = (cs.nlook[d] + t1) + t2;
    if (r.dir[d] == 0.f) r.dir[d] += 0.0000001f;
  }
  vnormalize (r.dir);
  return r;
}

static int nthreads_or_default (int nt)
{
#ifdef _OPENMP
  return nt > 0 ? nt : omp_get_max_threads ();
#else
  (void)nt;
  return 1;
#endif
}

} // namespace orc

using namespace orc;

extern "C" {

int orc_num_threads_max ()
{
#ifdef _OPENMP
  return omp_get_max_threads ();
#else
  return 1;
#endif
}

// ---- elements --------------------------------------------------------------------------
int orc_split_elements (const i32 *ctrl_idx, const float *values, const i32 *orders, const i64 *offsets,
                        int order, int n_el, float *out_aabbs, i32 *out_prims, float *out_subrefs, int nthreads)
{
  MeshView<float> m { ctrl_idx, values, orders, offsets, order, n_el };
  const int nt = nthreads_or_default (nthreads);
#pragma omp parallel num_threads(nt)
  {
    std::vector<float> net, scratch;
    AABB boxes[kSplits + 1];
    SubRef refs[kSplits + 1];
#pragma omp for schedule(static)
    for (int e = 0; e < n_el; ++e)
    {
      const int p = m.order_of (e);
      net.resize ((size_t)3 * npe_of (p));
      m.gather<3> (e, net.data ());
      split_element (net.data (), p, scratch, boxes, refs);
      for (int i = 0; i <= kSplits; ++i)
      {
        const i64 k = (i64)e * (kSplits + 1) + i;
        boxes[i].store (out_aabbs + 6 * k);
        out_prims[k] = e;
        for (int d = 0; d < 3; ++d) { out_subrefs[6 * k + d] = refs[i].lo[d]; out_subrefs[6 * k + 3 + d] = refs[i].hi[d]; }
      }
    }
  }
  return 0;
}

// ---- LBVH ----------------------------------------------------------------------------------
// Outputs sized for m = max(n, 2) leaves: inner (m-1)*16 floats, leaf_nodes m, aabb_ids m,
// mcodes_sorted m, mcodes_unsorted m (optional), left/right (m-1), parent (2m-1) (optional).
int orc_bvh_build (const float *aabbs, const i32 *prims, int n, float *out_inner, i32 *out_leaf_nodes,
                   i32 *out_aabb_ids, float *out_bounds, u32 *out_mcodes_sorted, u32 *out_mcodes_unsorted,
                   i32 *out_left, i32 *out_right, i32 *out_parent)
{
  BuildOut b;
  bvh_build (aabbs, prims, n, b);
  const int m = (int)b.leafs.size ();
  std::memcpy (out_inner, b.inner_nodes.data (), sizeof (float) * b.inner_nodes.size ());
  std::memcpy (out_leaf_nodes, b.leafs.data (), sizeof (i32) * m);
  std::memcpy (out_aabb_ids, b.ids.data (), sizeof (i32) * m);
  b.bounds.store (out_bounds);
  if (out_mcodes_sorted) std::memcpy (out_mcodes_sorted, b.mcodes_sorted.data (), sizeof (u32) * m);
  if (out_mcodes_unsorted) std::memcpy (out_mcodes_unsorted, b.mcodes_unsorted.data (), sizeof (u32) * m);
  if (out_left) std::memcpy (out_left, b.left.data (), sizeof (i32) * (m - 1));
  if (out_right) std::memcpy (out_right, b.right.data (), sizeof (i32) * (m - 1));
  if (out_parent) std::memcpy (out_parent, b.parent.data (), sizeof (i32) * (2 * m - 1));
  return m;
}

u32 orc_morton_3d (float x, float y, float z) { return morton_3d (x, y, z); }
int orc_clz (u32 x) { return clz32 (x); }

// ---- element evaluation / LUP ------------------------------------------------------------
void orc_eval_d_f32 (const float *dofs, int p, int ncomp, const float ref[3], float *val, float *deriv)
{
  if (ncomp == 3) { float v[3], d[3][3]; eval_d<float, 3> (dofs, p, ref, v, d); std::memcpy (val, v, sizeof v); std::memcpy (deriv, d, sizeof d); }
  else { float v[1], d[3][1]; eval_d<float, 1> (dofs, p, ref, v, d); val[0] = v[0]; for (int q = 0; q < 3; ++q) deriv[q] = d[q][0]; }
}
void orc_eval_d_f64 (const double *dofs, int p, int ncomp, const double ref[3], double *val, double *deriv)
{
  if (ncomp == 3) { double v[3], d[3][3]; eval_d<double, 3> (dofs, p, ref, v, d); std::memcpy (val, v, sizeof v); std::memcpy (deriv, d, sizeof d); }
  else { double v[1], d[3][1]; eval_d<double, 1> (dofs, p, ref, v, d); val[0] = v[0]; for (int q = 0; q < 3; ++q) deriv[q] = d[q][0]; }
}
int orc_lup_solve_f32 (const float *A9, const float *b, float *x)
{
  float A[3][3]; std::memcpy (A, A9, sizeof A); bool valid; lup_solve<float> (A, b, x, valid); return valid ? 1 : 0;
}
int orc_lup_solve_f64 (const double *A9, const double *b, double *x)
{
  double A[3][3]; std::memcpy (A, A9, sizeof A); bool valid; lup_solve<double> (A, b, x, valid); return valid ? 1 : 0;
}
int orc_newton_f32 (const float *dofs, int p, const float target[3], float x[3], int *steps)
{ return newton_inverse<float> (dofs, p, target, x, steps) ? 1 : 0; }
int orc_newton_f64 (const double *dofs, int p, const double target[3], double x[3], int *steps)
{ return newton_inverse<double> (dofs, p, target, x, steps) ? 1 : 0; }

// ---- locate -------------------------------------------------------------------------------
#define ORC_LOCATE_BODY(T)                                                                      \
  MeshView<T> m { ctrl_idx, values, orders, offsets, order, n_el };                            \
  BVHView bv { inner, leaf_nodes, aabb_ids, subrefs };                                         \
  const int nt = nthreads_or_default (nthreads);                                               \
  i64 tot_steps = 0, tot_cand = 0;                                                             \
  _Pragma("omp parallel num_threads(nt) reduction(+:tot_steps,tot_cand)")                      \
  {                                                                                            \
    std::vector<T> scratch;                                                                    \
    _Pragma("omp for schedule(static)")                                                        \
    for (i64 i = 0; i < npts; ++i)                                                             \
    {                                                                                          \
      LocStats ls;                                                                             \
      T ref[3];                                                                                \
      out_cells[i] = locate_point<T> (m, bv, pts + 3 * i, ref, scratch, &ls);                  \
      for (int d = 0; d < 3; ++d) out_refs[3 * i + d] = ref[d];                                \
      if (out_steps) out_steps[i] = (i32)ls.newton_steps;                                      \
      tot_steps += ls.newton_steps;                                                            \
      tot_cand += ls.candidates;                                                               \
    }                                                                                          \
  }                                                                                            \
  if (stats2) { stats2[0] = tot_steps; stats2[1] = tot_cand; }                                 \
  return 0;

int orc_locate_f32 (const i32 *ctrl_idx, const float *values, const i32 *orders, const i64 *offsets, int order, int n_el,
                    const float *inner, const i32 *leaf_nodes, const i32 *aabb_ids, const float *subrefs,
                    const float *pts, i64 npts, i32 *out_cells, float *out_refs, i32 *out_steps, i64 *stats2, int nthreads)
{ ORC_LOCATE_BODY (float) }

int orc_locate_f64 (const i32 *ctrl_idx, const double *values, const i32 *orders, const i64 *offsets, int order, int n_el,
                    const float *inner, const i32 *leaf_nodes, const i32 *aabb_ids, const float *subrefs,
                    const double *pts, i64 npts, i32 *out_cells, double *out_refs, i32 *out_steps, i64 *stats2, int nthreads)
{ ORC_LOCATE_BODY (double) }

int orc_set_newton_start (int mode)
{
  if (mode != 0 && mode != 2) return 1;
  g_start_mode = mode;
  return 0;
}

// Brute force: try every element (no BVH) in index order, Newton from the element centre, first
// accept wins.  Used on tiny meshes to show the BVH traversal loses no hits.
int orc_locate_brute_f64 (const i32 *ctrl_idx, const double *values, const i32 *orders, const i64 *offsets, int order,
                          int n_el, const double *pts, i64 npts, i32 *out_cells, double *out_refs)
{
  MeshView<double> m { ctrl_idx, values, orders, offsets, order, n_el };
  std::vector<double> dofs;
  for (i64 i = 0; i < npts; ++i)
  {
    out_cells[i] = -1;
    for (int e = 0; e < n_el; ++e)
    {
      const int p = m.order_of (e);
      dofs.resize ((size_t)3 * npe_of (p));
      m.gather<3> (e, dofs.data ());
      double x[3] = { 0.5, 0.5, 0.5 };
      if (newton_inverse<double> (dofs.data (), p, pts + 3 * i, x, nullptr))
      {
        out_cells[i] = e;
        for (int d = 0; d < 3; ++d) out_refs[3 * i + d] = x[d];
        break;
      }
    }
  }
  return 0;
}

// FP64 refinement to the exact inverse image: unclamped Newton in double on the given element,
// iterated until |dx|inf < 1e-14 (at most 60 steps).  The "FP64 oracle" reference coordinates.
int orc_refine_f64 (const i32 *ctrl_idx, const double *values, const i32 *orders, const i64 *offsets, int order, int n_el,
                    const double *pts, const i32 *cells, const double *ref_in, i64 npts, double *ref_out, int nthreads)
{
  MeshView<double> m { ctrl_idx, values, orders, offsets, order, n_el };
  const int nt = nthreads_or_default (nthreads);
#pragma omp parallel num_threads(nt)
  {
    std::vector<double> dofs;
#pragma omp for schedule(static)
    for (i64 i = 0; i < npts; ++i)
    {
      double x[3] = { ref_in[3 * i], ref_in[3 * i + 1], ref_in[3 * i + 2] };
      if (cells[i] >= 0)
      {
        const int p = m.order_of (cells[i]);
        dofs.resize ((size_t)3 * npe_of (p));
        m.gather<3> (cells[i], dofs.data ());
        for (int it = 0; it < 60; ++it)
        {
          double y[3], J[3][3], A[3][3], dy[3], dx[3];
          eval_d<double, 3> (dofs.data (), p, x, y, J);
          for (int k = 0; k < 3; ++k) dy[k] = pts[3 * i + k] - y[k];
          for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) A[r][c] = J[c][r];
          bool valid;
          lup_solve<double> (A, dy, dx, valid);
          double mx = 0;
          for (int k = 0; k < 3; ++k) { x[k] += dx[k]; mx = std::max (mx, std::fabs (dx[k])); }
          if (!(mx >= 1e-14)) break;
        }
      }
      for (int k = 0; k < 3; ++k) ref_out[3 * i + k] = x[k];
    }
  }
  return 0;
}

int orc_field_eval_f32 (const i32 *ctrl_idx, const float *values, const i32 *orders, const i64 *offsets, int order, int n_el,
                        const i32 *cells, const float *refs, i64 n, float *out)
{
  MeshView<float> f { ctrl_idx, values, orders, offsets, order, n_el };
  std::vector<float> s;
  for (i64 i = 0; i < n; ++i)
    if (cells[i] != -1) out[i] = field_eval<float> (f, cells[i], refs + 3 * i, s);
  return 0;
}

// ---- rays / camera / integrate / composite ------------------------------------------------
void orc_camera_perspective (int w, int h, float fov_y, const float *pos, const float *look_at, const float *up, float zoom,
                             void *out_rays /* Ray[w*h] */)
{
  CamSetup cs = camera_setup (w, h, fov_y, pos, look_at, up, zoom);
  Ray *r = (Ray *)out_rays;
  for (int i = 0; i < w * h; ++i) r[i] = camera_ray (cs, i);
}

// remove_missed_rays (ray.cpp:102-115): returns the number of active rays written to out_rays.
int orc_remove_missed_rays (const void *rays_in, int n, const float *bounds6, void *out_rays)
{
  AABB b; b.load (bounds6);
  AABB mb = b;
  mb.scale (1.001f);
  const Ray *r = (const Ray *)rays_in;
  Ray *o = (Ray *)out_rays;
  int k = 0;
  for (int i = 0; i < n; ++i)
  {
    float nr, fr;
    if (ray_slab (r[i], mb, nr, fr))
    {
      Ray q = r[i];
      // calc_ray_start on the compacted ray reruns the slab with the ray's own near/far
      float n2, f2;
      ray_slab (q, mb, n2, f2);
      q.near_ = n2;
      q.far_ = f2;
      o[k++] = q;
    }
  }
  return k;
}

// Volume sample spacing (volume.cpp:89-91): |bounds diagonal| / samples in float32.
float orc_sample_dist (const float *bounds6, int samples)
{
  float d[3];
  for (int i = 0; i < 3; ++i) d[i] = bounds6[2 * i + 1] - bounds6[2 * i];
  float s = 0.f;
  for (int i = 0; i < 3; ++i) s += d[i] * d[i];
  return sqrtf (s) / float (samples);
}

#define ORC_INTEGRATE_BODY(T, F)                                                                       \
  MeshView<T> m { ctrl_idx, values, orders, offsets, order, n_el };                                   \
  MeshView<double> m64v { ctrl_idx, values64, orders, offsets, order, n_el };                         \
  const MeshView<double> *m64 = values64 ? &m64v : nullptr;                                           \
  MeshView<F> f { f_ctrl_idx, f_values, orders, f_offsets, order, n_el };                             \
  BVHView bv { inner, leaf_nodes, aabb_ids, subrefs };                                                \
  ColorMapT<F> cm;                                                                                    \
  cm.lut = lut; cm.size = lut_size; cm.mn = F (fmin_); cm.mx = F (fmax_);                             \
  { const F len = cm.mx - cm.mn;                                                                     \
    cm.inv_range = F (1) / ((std::fabs (len) < F (1e-8)) ? (std::signbit (len) ? F (-1e-8) : F (1e-8)) : len); } \
  const Ray *rays = (const Ray *)rays_v;                                                              \
  Partial *parts = (Partial *)out_partials;                                                           \
  const int nt = nthreads_or_default (nthreads);                                                      \
  i64 tl = 0, tn = 0, tc = 0;                                                                         \
  _Pragma("omp parallel num_threads(nt) reduction(+:tl,tn,tc)")                                       \
  {                                                                                                   \
    std::vector<T> scratch;                                                                           \
    _Pragma("omp for schedule(static)")                                                               \
    for (i64 i = 0; i < n_rays; ++i)                                                                  \
    {                                                                                                 \
      RayStats rs;                                                                                    \
      integrate_ray<T, F> (m, f, bv, rays[i], T (sample_dist), cm, (const Light *)lights, n_lights,    \
                        parts + 5 * i, scratch, rs, m64);                                             \
      if (ray_located) ray_located[i] = (i32)rs.located;                                              \
      tl += rs.located; tn += rs.newton; tc += rs.candidates;                                         \
    }                                                                                                 \
  }                                                                                                   \
  if (stats3) { stats3[0] = tl; stats3[1] = tn; stats3[2] = tc; }                                     \
  return 0;

// Rays must already be active (remove_missed_rays applied).  Writes 5 partial slots per ray
// (unused slots carry depth = +inf), like integrate_partials before compact_partials.
int orc_integrate_f32 (const i32 *ctrl_idx, const float *values, const i32 *f_ctrl_idx, const float *f_values,
                       const i32 *orders, const i64 *offsets, const i64 *f_offsets, int order, int n_el,
                       const float *inner, const i32 *leaf_nodes, const i32 *aabb_ids, const float *subrefs,
                       const void *rays_v, i64 n_rays, float sample_dist, const float *lut, int lut_size,
                       double fmin_, double fmax_, void *out_partials, i32 *ray_located, i64 *stats3, int nthreads,
                       const float *lights, int n_lights)
{
  const double *values64 = nullptr;
  ORC_INTEGRATE_BODY (float, float)
}

// The FP32 march and colouring, each sample located with the FP64 walk and Newton on values64 (the same
// geometry in double).
int orc_integrate_f32_locate64 (const i32 *ctrl_idx, const float *values, const i32 *f_ctrl_idx, const float *f_values,
                                const i32 *orders, const i64 *offsets, const i64 *f_offsets, int order, int n_el,
                                const float *inner, const i32 *leaf_nodes, const i32 *aabb_ids, const float *subrefs,
                                const void *rays_v, i64 n_rays, float sample_dist, const float *lut, int lut_size,
                                double fmin_, double fmax_, void *out_partials, i32 *ray_located, i64 *stats3,
                                int nthreads, const float *lights, int n_lights, const double *values64)
{
  ORC_INTEGRATE_BODY (float, float)
}

int orc_integrate_f64 (const i32 *ctrl_idx, const double *values, const i32 *f_ctrl_idx, const double *f_values,
                       const i32 *orders, const i64 *offsets, const i64 *f_offsets, int order, int n_el,
                       const float *inner, const i32 *leaf_nodes, const i32 *aabb_ids, const float *subrefs,
                       const void *rays_v, i64 n_rays, float sample_dist, const float *lut, int lut_size,
                       double fmin_, double fmax_, void *out_partials, i32 *ray_located, i64 *stats3, int nthreads,
                       const float *lights, int n_lights)
{
  const double *values64 = nullptr;
  ORC_INTEGRATE_BODY (double, double)
}

// FP32 geometry with an FP64 scalar field (and an FP64 colour-map range): unlit only (lights ignored).
int orc_integrate_f32_field64 (const i32 *ctrl_idx, const float *values, const i32 *f_ctrl_idx, const double *f_values,
                               const i32 *orders, const i64 *offsets, const i64 *f_offsets, int order, int n_el,
                               const float *inner, const i32 *leaf_nodes, const i32 *aabb_ids, const float *subrefs,
                               const void *rays_v, i64 n_rays, float sample_dist, const float *lut, int lut_size,
                               double fmin_, double fmax_, void *out_partials, i32 *ray_located, i64 *stats3,
                               int nthreads, const float *lights, int n_lights)
{
  lights = nullptr;
  (void)n_lights;
  const double *values64 = nullptr;
  ORC_INTEGRATE_BODY (float, double)
}

// One shaded_color sample (for testing the lit path on its own): element `cell` at reference point ref,
// ray direction dir; out = RGBA.
#define ORC_SHADE_BODY(T)                                                                              \
  MeshView<T> m { ctrl_idx, values, orders, offsets, order, n_el };                                   \
  MeshView<T> f { f_ctrl_idx, f_values, orders, f_offsets, order, n_el };                             \
  ColorMapT<T> cm;                                                                                    \
  cm.lut = lut; cm.size = lut_size; cm.mn = T (fmin_); cm.mx = T (fmax_);                             \
  { const T len = cm.mx - cm.mn;                                                                     \
    cm.inv_range = T (1) / ((std::fabs (len) < T (1e-8)) ? (std::signbit (len) ? T (-1e-8) : T (1e-8)) : len); } \
  Ray ray {};                                                                                         \
  for (int d = 0; d < 3; ++d) ray.dir[d] = dir[d];                                                    \
  std::vector<T> scratch;                                                                             \
  const T r[3] = { T (ref[0]), T (ref[1]), T (ref[2]) };                                              \
  shaded_color<T> (m, f, cell, r, ray, cm, (const Light *)lights, n_lights, scratch, out4);

void orc_shade_f32 (const i32 *ctrl_idx, const float *values, const i32 *f_ctrl_idx, const float *f_values,
                    const i32 *orders, const i64 *offsets, const i64 *f_offsets, int order, int n_el, int cell,
                    const double *ref, const float *dir, const float *lut, int lut_size, double fmin_, double fmax_,
                    const float *lights, int n_lights, float *out4)
{ ORC_SHADE_BODY (float) }

void orc_shade_f64 (const i32 *ctrl_idx, const double *values, const i32 *f_ctrl_idx, const double *f_values,
                    const i32 *orders, const i64 *offsets, const i64 *f_offsets, int order, int n_el, int cell,
                    const double *ref, const float *dir, const float *lut, int lut_size, double fmin_, double fmax_,
                    const float *lights, int n_lights, float *out4)
{ ORC_SHADE_BODY (double) }

// Renderer's default light (rendering/renderer.cpp:141-159): 0.1 x |look - pos| above the camera along
// the camera's orthonormalised up; out = 13 floats (pos, amb, diff, spec, spec_pow) with dray's defaults.
void orc_default_light (const float *pos, const float *look_at, const float *up_in, float *out13)
{
  float up[3] = { up_in[0], up_in[1], up_in[2] };
  vnormalize (up);
  float look[3];
  for (int i = 0; i < 3; ++i) look[i] = look_at[i] - pos[i];
  float m2 = 0.f;
  for (int i = 0; i < 3; ++i) m2 += look[i] * look[i];
  const float mag = sqrtf (m2);
  float right[3] = { look[1] * up[2] - look[2] * up[1], look[2] * up[0] - look[0] * up[2], look[0] * up[1] - look[1] * up[0] };
  vnormalize (right);
  float mu[3] = { right[1] * look[2] - right[2] * look[1], right[2] * look[0] - right[0] * look[2],
                  right[0] * look[1] - right[1] * look[0] };
  vnormalize (mu);
  Light L;
  for (int i = 0; i < 3; ++i)
  {
    L.pos[i] = pos[i] + .1f * mag * mu[i];
    L.amb[i] = 0.4f;
    L.diff[i] = 0.75f;
    L.spec[i] = 0.8f;
  }
  L.spec_pow = 15.f;
  std::memcpy (out13, &L, sizeof L);
}

// compact_partials (volume.cpp:42-67): keep depth != +inf, order preserved.  Returns count.
int orc_compact_partials (const void *in, i64 n, void *out)
{
  const Partial *p = (const Partial *)in;
  Partial *o = (Partial *)out;
  int k = 0;
  for (i64 i = 0; i < n; ++i)
    if (p[i].depth != inf32 ()) o[k++] = p[i];
  return k;
}

// Volume::save compositing (volume.cpp:369-416) + Framebuffer::composite_background
// (framebuffer.cpp:146-162; background white).  out_rgba: 4 floats per pixel.
void orc_composite (const void *parts_v, int n, int w, int h, float *out_rgba);

// Renderer::render's cross-domain compositing (rendering/renderer.cpp:307-330 via apcomp's
// PartialCompositor, which is not part of the reference sources): partials sorted by (pixel id, depth),
// ties kept in input order, then composited per pixel front to back as Volume::save does.
void orc_composite_domains (const void *parts_v, int n, int w, int h, float *out_rgba)
{
  const Partial *p = (const Partial *)parts_v;
  std::vector<Partial> v (p, p + n);
  std::stable_sort (v.begin (), v.end (), [] (const Partial &a, const Partial &b) {
    return a.pixel_id != b.pixel_id ? a.pixel_id < b.pixel_id : a.depth < b.depth;
  });
  orc_composite (v.data (), n, w, h, out_rgba);
}

void orc_composite (const void *parts_v, int n, int w, int h, float *out_rgba)
{
  const Partial *p = (const Partial *)parts_v;
  for (int i = 0; i < 4 * w * h; ++i) out_rgba[i] = 0.f;
  if (n > 0)
  {
    float color[4];
    std::memcpy (color, p[0].rgba, sizeof color);
    int cur = -1;
    for (int i = 0; i < n; ++i)
    {
      if (cur != p[i].pixel_id)
      {
        if (cur != -1) std::memcpy (out_rgba + 4 * cur, color, sizeof color);
        std::memcpy (color, p[i].rgba, sizeof color);
        cur = p[i].pixel_id;
      }
      else
      {
        const float a = (1.f - color[3]);   // pre_mult_alpha_blend_host (colors.hpp:45-53)
        color[0] = color[0] + p[i].rgba[0] * a;
        color[1] = color[1] + p[i].rgba[1] * a;
        color[2] = color[2] + p[i].rgba[2] * a;
        color[3] = color[3] + p[i].rgba[3] * a;
      }
      if (i == n - 1) std::memcpy (out_rgba + 4 * cur, color, sizeof color);
    }
  }
  const float bg[4] = { 1.f, 1.f, 1.f, 1.f };
  for (int i = 0; i < w * h; ++i)
  {
    float *c = out_rgba + 4 * i;
    if (c[3] < 1.f)
    {
      const float a = (1.f - c[3]);   // blend_pre_alpha (colors.hpp:25-33)
      c[0] = c[0] + bg[0] * a;
      c[1] = c[1] + bg[1] * a;
      c[2] = c[2] + bg[2] * a;
      c[3] = c[3] + bg[3] * a;
    }
  }
}

// Renderer::render's volume compositing (rendering/renderer.cpp:307-330; apcomp's PartialCompositor is not
// part of the reference sources: partials sorted by (pixel id, depth), ties in input order, each pixel's
// run blended front to back with pre_mult_alpha_blend_host) followed by detail::partials_to_framebuffer
// (renderer.cpp:97-138): blend = 0 replaces the pixel; blend = 1 puts the volume over the framebuffer
// colour (colour = volume + (1 - volume alpha) * framebuffer).  depths (nullable) take the run's first
// partial's depth.  Pixels without partials are left as they are.
void orc_partials_to_framebuffer (const void *parts_v, int n, int w, int h, float *colors, float *depths, int blend)
{
  const Partial *p = (const Partial *)parts_v;
  std::vector<Partial> v (p, p + n);
  std::stable_sort (v.begin (), v.end (), [] (const Partial &a, const Partial &b) {
    return a.pixel_id != b.pixel_id ? a.pixel_id < b.pixel_id : a.depth < b.depth;
  });
  int i = 0;
  while (i < n)
  {
    const int pid = v[i].pixel_id;
    float c[4];
    std::memcpy (c, v[i].rgba, sizeof c);
    const float d = v[i].depth;
    int k = i + 1;
    for (; k < n && v[k].pixel_id == pid; ++k)
    {
      const float a = 1.f - c[3];
      for (int q = 0; q < 4; ++q) c[q] = c[q] + v[k].rgba[q] * a;
    }
    if (pid >= 0 && pid < w * h)
    {
      float *fb = colors + 4 * (i64)pid;
      if (blend)
      {
        const float o = 1.f - c[3];
        for (int q = 0; q < 4; ++q) c[q] = c[q] + o * fb[q];
      }
      std::memcpy (fb, c, sizeof c);
      if (depths) depths[pid] = d;
    }
    i = k;
  }
}

// Framebuffer::composite_background (framebuffer.cpp:146-162) with blend_pre_alpha (colors.hpp:25-33).
void orc_framebuffer_composite_background (float *colors, int npix, const float *bg)
{
  for (int i = 0; i < npix; ++i)
  {
    float *c = colors + 4 * (i64)i;
    if (c[3] < 1.f)
    {
      const float a = 1.f - c[3];
      for (int q = 0; q < 4; ++q) c[q] = c[q] + bg[q] * a;
    }
  }
}

// Camera::reset_to_bounds / elevate / azimuth (rendering/camera.cpp:485-552) with rotate()
// (transform_3d.hpp:126-157) and transform_vector (:53-60), float operation order as in dray.
// cam = 7 floats in/out: pos[3], look_at[3], fov_y; up[3] read only.  mode 0 reset_to_bounds(bounds6,
// pads[3]), 1 elevate(arg), 2 azimuth(arg).
static void orbit_rotate (float deg, float axis[3], float v[3])
{
  const float rad = 0.01745329251994329547437168059786927f * deg;
  vnormalize (axis);
  const float sn = sinf (rad), cs = cosf (rad);
  const float m[3][3] = {
    { axis[0] * axis[0] * (1 - cs) + cs, axis[0] * axis[1] * (1 - cs) - axis[2] * sn, axis[0] * axis[2] * (1 - cs) + axis[1] * sn },
    { axis[1] * axis[0] * (1 - cs) + axis[2] * sn, axis[1] * axis[1] * (1 - cs) + cs, axis[1] * axis[2] * (1 - cs) - axis[0] * sn },
    { axis[2] * axis[0] * (1 - cs) - axis[1] * sn, axis[2] * axis[1] * (1 - cs) + axis[0] * sn, axis[2] * axis[2] * (1 - cs) + cs } };
  float r[3];
  for (int i = 0; i < 3; ++i)
  {
    float d = 0.f;
    for (int j = 0; j < 3; ++j) d += m[i][j] * v[j];
    d += 0.f * 0.f;
    r[i] = d;
  }
  std::memcpy (v, r, sizeof r);
}

void orc_camera_orbit (float *cam, const float *up_in, int mode, float arg, const float *bounds6, const double *pads)
{
  float *pos = cam, *look_at = cam + 3;
  if (mode == 0)
  {
    float db[6] = { 3.402823466e+38f, -3.402823466e+38f, 3.402823466e+38f, -3.402823466e+38f, 3.402823466e+38f,
                    -3.402823466e+38f };
    for (int d = 0; d < 3; ++d)
    {
      const double pad = pads[d] * ((double)bounds6[2 * d + 1] - (double)bounds6[2 * d]);
      const float hi = (float)((double)bounds6[2 * d + 1] + pad), lo = (float)((double)bounds6[2 * d] - pad);
      for (float x : { hi, lo })
      {
        db[2 * d] = db[2 * d] < x ? db[2 * d] : x;
        db[2 * d + 1] = db[2 * d + 1] > x ? db[2 * d + 1] : x;
      }
    }
    float pd[3] = { pos[0] - look_at[0], pos[1] - look_at[1], pos[2] - look_at[2] };
    vnormalize (pd);
    float c[3], e[3];
    for (int d = 0; d < 3; ++d)
    {
      c[d] = 0.5f * (db[2 * d] + db[2 * d + 1]);
      e[d] = db[2 * d + 1] - db[2 * d];
    }
    float m2 = 0.f;
    for (int d = 0; d < 3; ++d) m2 += e[d] * e[d];
    const float diag = sqrtf (m2);
    for (int d = 0; d < 3; ++d)
    {
      look_at[d] = c[d];
      pos[d] = c[d] + pd[d] * diag * 1.0f;
    }
    cam[6] = 60.f;
    return;
  }
  float look[3] = { look_at[0] - pos[0], look_at[1] - pos[1], look_at[2] - pos[2] };
  float m2 = 0.f;
  for (int d = 0; d < 3; ++d) m2 += look[d] * look[d];
  const float dist = sqrtf (m2);
  float right[3] = { look[1] * up_in[2] - look[2] * up_in[1], look[2] * up_in[0] - look[0] * up_in[2],
                     look[0] * up_in[1] - look[1] * up_in[0] };
  float up[3] = { right[1] * look[2] - right[2] * look[1], right[2] * look[0] - right[0] * look[2],
                  right[0] * look[1] - right[1] * look[0] };
  vnormalize (look);
  vnormalize (right);
  vnormalize (up);
  if (mode == 1) orbit_rotate (-arg, right, look);
  else orbit_rotate (arg, up, look);
  for (int d = 0; d < 3; ++d) pos[d] = look_at[d] - look[d] * dist;
}

// ColorTable::sample (color_table.cpp:152-166) of a table given as control points, after
// correct_opacity(factor) (:304-324; factor <= 0 skips the correction).
// rgb_pts: n_rgb x (pos, r, g, b, a); alpha_pts: n_alpha x (pos, alpha).
void orc_color_table_sample (const float *rgb_pts, int n_rgb, const float *alpha_pts, int n_alpha, int smooth,
                             float factor, int num_samples, float alpha_scale, float *out_rgba)
{
  std::vector<float> ap (alpha_pts, alpha_pts + 2 * n_alpha);
  if (factor > 0.f)
    for (int i = 0; i < n_alpha; ++i) ap[2 * i + 1] = 1.f - std::pow ((1.f - ap[2 * i + 1]), factor);
  for (int i = 0; i < num_samples; ++i)
  {
    const float s = static_cast<float> (i) / static_cast<float> (num_samples - 1);
    float c[4];
    // map_rgb (:168-225)
    if (n_rgb == 0) { c[0] = c[1] = c[2] = 0.5f; c[3] = 1.f; }
    else if (n_rgb == 1 || s <= rgb_pts[0]) { for (int k = 0; k < 4; ++k) c[k] = rgb_pts[1 + k]; }
    else if (s >= rgb_pts[5 * (n_rgb - 1)]) { for (int k = 0; k < 4; ++k) c[k] = rgb_pts[5 * (n_rgb - 1) + 1 + k]; }
    else
    {
      int sec;
      for (sec = 1; sec < n_rgb - 1; ++sec) if (s < rgb_pts[5 * sec]) break;
      const int fst = sec - 1;
      const float seg = rgb_pts[5 * sec] - rgb_pts[5 * fst];
      const float a = (seg == 0.f) ? .5f : (s - rgb_pts[5 * fst]) / seg;
      for (int k = 0; k < 4; ++k)
      {
        const float c0 = rgb_pts[5 * fst + 1 + k], c1 = rgb_pts[5 * sec + 1 + k];
        c[k] = smooth ? (c0 * (1.0f - a) + c1 * a) : (a < .5 ? c0 : c1);
      }
    }
    // map_alpha (:227-283)
    float al;
    if (n_alpha == 0) al = 1.f;
    else if (n_alpha == 1 || s <= ap[0]) al = ap[1];
    else if (s >= ap[2 * (n_alpha - 1)]) al = ap[2 * (n_alpha - 1) + 1];
    else
    {
      int sec;
      for (sec = 1; sec < n_alpha - 1; ++sec) if (s < ap[2 * sec]) break;
      const int fst = sec - 1;
      const float seg = ap[2 * sec] - ap[2 * fst];
      const float a = (seg == 0.f) ? .5f : (s - ap[2 * fst]) / seg;
      const float a0 = ap[2 * fst + 1], a1 = ap[2 * sec + 1];
      al = smooth ? (a0 * (1.f - a) + a1 * a) : (a < .5 ? a0 : a1);
    }
    c[3] = al;
    c[3] *= alpha_scale;
    for (int k = 0; k < 4; ++k) out_rgba[4 * i + k] = c[k];
  }
}

} // extern "C"

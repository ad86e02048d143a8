/*
 * drx.h -- C ABI of the MI355X (gfx950) volume-render / point-location library (libdrx.so).
 *
 * Plain pointers and sizes only.  Every pointer argument is DEVICE memory unless its comment says
 * "host".  Every function returns DRX_OK (0) or a non-zero status; drx_last_error() returns the
 * message of the last failure on the calling host thread.  Functions enqueue on `stream` (a
 * hipStream_t passed as void*, NULL = the null stream) and, unless noted, return without waiting;
 * functions that hand back a host-visible count synchronise the stream first.  The library never frees
 * caller memory; temporaries live in an arena owned by the library with one set of buffers per (device,
 * stream): calls on different streams never share them, and calls on one stream are ordered by it, so
 * any number of streams may be used at once (two host threads must not issue calls on the same stream
 * at the same time).  drx_release frees the current device's arena.
 *
 * Reference interfaces replaced (LLNL/devil_ray, paths relative to its src/dray/):
 *   drx_bvh_split_elements  construct_bvh per-element split          data_model/mesh_utils.cpp:514-670
 *   drx_bvh_build           LinearBVHBuilder::construct               linear_bvh_builder.hpp:20-21,
 *                                                                     linear_bvh_builder.cpp:406-518
 *   drx_locate              Mesh::locate / UnstructuredMesh::locate   data_model/mesh.hpp:34,
 *                           (DeviceMesh::locate per point)            data_model/unstructured_mesh.cpp:123-148,
 *                                                                     data_model/device_mesh.hpp:247-354
 *   drx_field_eval          Field::eval                               data_model/field.hpp:53,
 *                                                                     data_model/unstructured_field.cpp:213-246
 *   drx_camera_rays         Camera::create_rays / gen_perspective     rendering/camera.cpp:208-243,402-483
 *   drx_remove_missed_rays  remove_missed_rays                        ray.cpp:102-241
 *   drx_volume_integrate    Volume::integrate -> integrate_partials   rendering/volume.hpp:40,
 *                           (+ VolumeShader::color / shaded_color)    rendering/volume.cpp:69-228,323-351,
 *                                                                     rendering/volume_shader.hpp:39-167
 *   drx_default_light       detail::default_light                     rendering/renderer.cpp:141-159
 *   drx_composite_domains   Renderer::render volume partial compositing  rendering/renderer.cpp:307-330
 *   drx_composite           Volume::save compositing + Framebuffer::  rendering/volume.cpp:369-416,
 *                           composite_background                      rendering/framebuffer.cpp:146-162
 *   drx_partials_to_framebuffer  apcomp PartialCompositor::composite  rendering/renderer.cpp:307-330,
 *                           + detail::partials_to_framebuffer         rendering/renderer.cpp:97-138
 *   drx_framebuffer_clear   Framebuffer::clear                        rendering/framebuffer.cpp:126-144
 *   drx_framebuffer_composite_background  Framebuffer::composite_background  rendering/framebuffer.cpp:146-162
 *   drx_save_png            Framebuffer::save -> PNGEncoder           rendering/framebuffer.cpp:44-51,
 *                                                                     utils/png_encoder.cpp:56-116
 *   drx_color_table_sample  ColorTable::correct_opacity + sample      color_table.cpp:152-324
 *   drx_color_table_preset  ColorTable (const std::string &name)      color_table.cpp:378-1330
 *   drx_camera_reset_to_bounds, drx_camera_elevate, drx_camera_azimuth
 *                           Camera::reset_to_bounds / elevate / azimuth  rendering/camera.cpp:485-552
 *   drx_mesh_pack           (no counterpart: GridFunction -> element-local SoA conversion done once at
 *                           upload; replaces DeviceMesh::get_elem's indirect reads, device_mesh.hpp:79-96)
 *   drx_mlp_*               (no counterpart: learned Newton initial guess; dray starts Newton from
 *                           LocateHack::eval_inverse's guess, device_mesh.hpp:131-165,327-334)
 */
#ifndef DRX_H
#define DRX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum drx_status
{
  DRX_OK = 0,
  DRX_ERR_ARG = 1,         /* bad argument (null pointer, size, unsupported order) */
  DRX_ERR_HIP = 2,         /* a HIP runtime call failed */
  DRX_ERR_UNSUPPORTED = 3, /* valid request this build does not implement */
  DRX_ERR_OVERFLOW = 4,    /* a fixed-capacity structure overflowed (traversal stack) */
  DRX_ERR_STATE = 5        /* library not initialised / wrong device */
};

/* Highest element order the GPU kernels accept for a mesh of one order (dray allows up to 20,
 * constants.hpp:14).  Orders above DRX_MAX_MIXED_ORDER run the binary walk with the plain Newton start
 * only; mixed-order meshes and the learned start take orders up to DRX_MAX_MIXED_ORDER. */
#define DRX_MAX_ORDER 8
#define DRX_MAX_MIXED_ORDER 4
/* BVH leaves per element: 32 binary de Casteljau splits -> 33 sub-boxes (mesh_utils.cpp:528). */
#define DRX_SPLITS_PER_ELEMENT 33

/* Element-local packed mesh (output of drx_mesh_pack).  Geometry (ncomp == 3) blocks are
 *   [ox, oy, oz, 0, X[npe_pad], Y[npe_pad], Z[npe_pad]]
 * with X = x - ox etc. (control points relative to the element's box centre, computed in double and
 * rounded once), npe = (p+1)^3 in dray dof order i + (p+1)(j + (p+1)k), npe_pad = npe rounded up to 4.
 * Scalar-field (ncomp == 1) blocks are [F[npe_pad]].  Uniform-order meshes place element e at
 * e * stride floats; mixed-order meshes give per-element orders and float offsets. */
typedef struct drx_mesh
{
  const float *data;
  const int64_t *elem_offset; /* device, NULL for uniform order */
  const int32_t *elem_order;  /* device, NULL for uniform order */
  int32_t n_el;
  int32_t order;              /* uniform order; for mixed meshes the maximum order */
  int32_t ncomp;              /* 3 = geometry, 1 = scalar field */
  int32_t stride;             /* floats per element block (uniform order); doubles for an FP64 field */
  const double *data64;       /* scalar fields only, nullable: FP64 blocks [F[npe_pad]] of doubles
                                 (drx_mesh_pack_f64), used instead of `data` (which may then be NULL) */
} drx_mesh;

/* BVH in dray's layout (bvh.hpp:15-39): inner node k = 16 floats
 *   [L.min xyz, L.max x | L.max yz, R.min xy | R.min z, R.max xyz | lchild, rchild (int32 bits), 0, 0]
 * child >= 0: float4 offset (4 * inner index) of an inner node; child < 0: leaf -(child) - 1. */
typedef struct drx_bvh
{
  const float *inner_nodes;   /* (n_leaves - 1) * 16 floats */
  const int32_t *leaf_nodes;  /* element id per leaf, sorted (m_leaf_nodes) */
  const int32_t *aabb_ids;    /* original aabb index per leaf (m_aabb_ids) */
  const float *subrefs;       /* 6 floats (lo xyz, hi xyz) per ORIGINAL aabb (mesh m_ref_aabbs) */
  const float *leaf_pack;     /* optional, 8 floats per leaf: {el_id bits, lo xyz, hi xyz, 0}; see drx_bvh_pack_leaves */
  const float *wide_nodes;    /* optional 8-wide collapse (drx_bvh_widen); used for traversal when set */
  int32_t n_leaves;
  float bounds[6];            /* host copy of m_bounds: xmin, xmax, ymin, ymax, zmin, zmax */
  const float *leaf_starts;   /* optional, 12 floats per leaf: Newton start map (drx_bvh_leaf_starts) */
  const uint32_t *qnodes;     /* optional, 8 words per inner node: compact traversal nodes (drx_bvh_quantize);
                                 binary walks use them, with leaf_boxes and qgrid, when set */
  const float *leaf_boxes;    /* with qnodes: 8 floats per leaf, its exact box {lo xyz, hi x, hi yz, 0, 0} */
  float qgrid[6];             /* with qnodes: grid origin xyz, grid cells per unit length xyz */
} drx_bvh;

/* Location (location.hpp:16-21): {int32 cell (-1 = not found), float ref[3]} = 16 bytes. */
/* Ray (ray.hpp:20-28): {float dir[3], near, orig[3], far; int32 pixel_id} = 36 bytes.   */
/* VolumePartial (volume_partial.hpp:16-21): {int32 pixel_id; float depth; float rgba[4]} = 24 bytes. */

enum drx_init_mode
{
  DRX_INIT_AFFINE = 0,        /* default: Newton starts at the element's corner-affine estimate clamped into
                                 the candidate's SubRef (the leaf's start map when bvh->leaf_starts is set,
                                 drx_bvh_leaf_starts); the SubRef centre where that estimate is singular */
  DRX_INIT_MLP = 1,           /* learned initial guess (opts->mlp); plain Newton from the DRX_INIT_AFFINE
                                 start decides acceptance when the guess has not converged */
  DRX_INIT_SUBREF_CENTRE = 2, /* Newton starts at the candidate's SubRef centre, subref_center
                                 (data_model/subref.hpp:52-56): dray's start when use_init_guess is false */
  DRX_INIT_GUESS = 3          /* given initial guesses (LocateHack::eval_inverse with use_init_guess true,
                                 device_mesh.hpp:131-165): the candidate whose element is the guess's cell
                                 starts at the guess's reference point, the others at the DRX_INIT_AFFINE
                                 start; opts->guess: one Location per point (drx_locate) or per (ray,
                                 sample) (drx_volume_integrate, binary walk) */
};

typedef struct drx_locate_opts
{
  int32_t init_mode;          /* drx_init_mode */
  int32_t fp64_polish;        /* 1: FP64 Newton correction for accepted points with cond(J) > cond_threshold */
  float cond_threshold;       /* e.g. 1e4 */
  int32_t *newton_steps;      /* device, nullable: Newton steps summed over candidates, per point */
  int32_t *candidates;        /* device, nullable: candidate leaves tried, per point */
  int32_t traversal;          /* drx_traversal: how a wavefront walks the BVH (same results either way) */
  int32_t coherent_order;     /* drx_locate: 1 = process queries in Morton order (results stay in input order) */
  const void *mlp;            /* DRX_INIT_MLP: packed initializer (drx_mlp_pack), device */
  int32_t mlp_newton_steps;   /* DRX_INIT_MLP: Newton budget from the learned guess before the plain
                                 fallback (<= 0: 4) */
  int32_t launch;             /* drx_launch: how drx_locate's streamed walks (binary, 8-wide per lane,
                                 learned start) are launched, and whether drx_volume_integrate takes the
                                 sample-parallel kernel; the results are the same either way */
  int32_t chunk;              /* points per wavefront chunk for those walks (0: tuned, else 1024) */
  int32_t *polished;          /* device, nullable: drx_locate writes 1 where the point's accepted solution took
                                 the FP64 correction (condition estimate > cond_threshold), else 0 */
  int32_t mlp_fallback;       /* DRX_INIT_MLP: the plain start that takes over (and decides acceptance) when
                                 the learned guess has not converged: 0 the DRX_INIT_AFFINE start, 1 the
                                 SubRef centre (dray's subref_center) */
  const void *guess;          /* DRX_INIT_GUESS: device Locations (see drx_init_mode) */
  void *record;               /* drx_volume_integrate, nullable: device Locations, the location found for
                                 every (ray, sample) (cell -1 where none), entry ray * samples_per_ray +
                                 sample in the order the march locates them; with opts->guess it reads the
                                 same entries (the binary walk then runs for every frame size) */
  int32_t samples_per_ray;    /* entries per ray of guess / record in drx_volume_integrate (samples beyond
                                 it are neither guessed nor recorded) */
} drx_locate_opts;

enum drx_launch
{
  DRX_LAUNCH_AUTO = 0,        /* the form drx_locate_autotune picked for this device, order and walk, else chunked */
  DRX_LAUNCH_CHUNKED = 1,     /* one wavefront per chunk of points */
  DRX_LAUNCH_PERSISTENT = 2,  /* about one wavefront per resident slot; wavefronts take chunks from a queue */
  DRX_LAUNCH_SAMPLES = 3      /* drx_volume_integrate: several wavefronts per 64 rays locate and shade
                                 consecutive samples in parallel and one of them runs each ray's march over
                                 them (same partials and statistics; binary walk, plain start, orders up to
                                 DRX_MAX_MIXED_ORDER).  Under DRX_LAUNCH_AUTO and DRX_TRAVERSE_AUTO a render
                                 takes it for frames of fewer than 300 active rays per SIMD; DRX_LAUNCH_CHUNKED
                                 or DRX_LAUNCH_PERSISTENT keep one lane per ray */
};

/* Training schedule of the learned initializer (drx_mlp_train): Adam, cosine decay lr -> lr_final over
 * total_steps.  Step k reads rows [sample0 + k*batch, +batch) of the data set, wrapping around. */
typedef struct drx_mlp_train_opts
{
  int32_t batch;              /* rows per step, a multiple of 64 */
  int32_t steps;              /* steps this call runs */
  int32_t step0;              /* Adam steps already taken (bias correction, schedule position) */
  int32_t total_steps;        /* length of the schedule (<= 0: step0 + steps) */
  int64_t sample0;            /* first row of this call */
  float lr, lr_final;         /* lr_final < 0: constant lr */
  float beta1, beta2, eps;    /* <= 0: 0.9, 0.999, 1e-8 */
} drx_mlp_train_opts;

enum drx_traversal
{
  DRX_TRAVERSE_AUTO = 0,      /* library default: DRX_TRAVERSE_BINARY (fastest measured on MI355X), except
                                 drx_volume_integrate over a BVH of at most 65,536 leaves with its 8-wide
                                 collapse and at least 48 active rays per SIMD: DRX_TRAVERSE_WAVE */
  DRX_TRAVERSE_LANE = 1,      /* every lane walks the 8-wide collapse (bvh->wide_nodes) depth-first */
  DRX_TRAVERSE_WAVE = 2,      /* the wavefront walks the union of its lanes' paths over the 8-wide
                                 collapse with uniform node reads */
  DRX_TRAVERSE_BINARY = 3     /* every lane walks dray's binary tree depth-first */
};

typedef struct drx_integrate_stats
{
  int64_t located_samples;    /* successful locate() calls inside the march */
  int64_t newton_steps;       /* Newton iterations over all candidates */
  int64_t candidates;         /* candidate leaves tried */
  int64_t polished;           /* samples that took the FP64 correction */
  float kernel_ms;            /* duration of the integrate kernel alone (HIP events on `stream`) */
  float total_ms;             /* the whole call: integrate + partial compaction (HIP events) */
  int32_t walk;               /* the walk that ran: 0 binary, 1 8-wide per lane, 2 wave-coherent, 3 learned start,
                                 4 binary with given starts / record, 5 binary, sample-parallel launch */
} drx_integrate_stats;

/* PointLight (rendering/point_light.hpp:15-33); dray's defaults: amb 0.4, diff 0.75, spec 0.8, spec_pow 15. */
typedef struct drx_point_light
{
  float pos[3];
  float amb[3];
  float diff[3];
  float spec[3];
  float spec_pow;
} drx_point_light;

#define DRX_MAX_LIGHTS 8

/* Volume shading (rendering/volume.cpp:162-171): use_lighting 0 = VolumeShader::color (colour map only),
 * 1 = VolumeShader::shaded_color with lights[0..n_lights) (host array, n_lights <= DRX_MAX_LIGHTS). */
typedef struct drx_shading
{
  int32_t use_lighting;
  int32_t n_lights;
  const drx_point_light *lights;
} drx_shading;

typedef struct drx_camera
{
  int32_t width, height;
  float fov_y;                /* degrees */
  float pos[3];
  float look_at[3];
  float up[3];
  float zoom;
} drx_camera;

/* ---- library state ------------------------------------------------------------------------- */
int drx_init (int device);                 /* select device, check it is gfx950 */
const char *drx_last_error (void);
const char *drx_version (void);
int drx_release (void);                    /* free the library's arena on the current device */

/* ---- mesh packing -------------------------------------------------------------------------- */
/* Host helper: floats per element block for order p and ncomp. */
int32_t drx_mesh_stride (int32_t order, int32_t ncomp);
/* Host helper for mixed-order meshes: fills h_offsets[n_el] (host) and returns total floats; -1 (with the
 * message in drx_last_error) when an element's order is outside 1..DRX_MAX_MIXED_ORDER. */
int64_t drx_mesh_pack_layout (const int32_t *h_orders, int32_t n_el, int32_t ncomp, int64_t *h_offsets);
/* dray GridFunction (ctrl_idx + ncomp values per control point) -> packed blocks in `out`.
 * orders/offsets (device) are NULL for uniform order; offsets index ctrl_idx per element and
 * out_offsets (device) index `out` per element. */
int drx_mesh_pack (const int32_t *ctrl_idx, const float *values, int32_t ncomp, const int32_t *orders,
                   const int64_t *offsets, const int64_t *out_offsets, int32_t order, int32_t n_el,
                   float *out, void *stream);

/* An FP64 scalar field (dray built with Float = double, types.hpp:29-33): dray GridFunction (ctrl_idx + one
 * double per control point) -> packed double blocks [F[npe_pad]] in `out`; orders/offsets/out_offsets as in
 * drx_mesh_pack (offsets counted in doubles; drx_mesh_pack_layout with ncomp 1 gives them). */
int drx_mesh_pack_f64 (const int32_t *ctrl_idx, const double *values, const int32_t *orders, const int64_t *offsets,
                       const int64_t *out_offsets, int32_t order, int32_t n_el, double *out, void *stream);

/* ---- LBVH ---------------------------------------------------------------------------------- */
/* 33 sub-boxes per element, bit-exact with dray's float32 split: out_aabbs 6*33*n_el floats,
 * out_prims 33*n_el, out_subrefs 6*33*n_el.  Reads dray's GridFunction (ctrl_idx + Vec3 values). */
int drx_bvh_split_elements (const int32_t *ctrl_idx, const float *values, const int32_t *orders,
                            const int64_t *offsets, int32_t order, int32_t n_el, float *out_aabbs,
                            int32_t *out_prims, float *out_subrefs, void *stream);
/* LinearBVHBuilder::construct(aabbs, prim_ids).  m = max(n, 2) leaves (dray pads 0/1 boxes):
 * out_inner (m-1)*16 floats, out_leaf_nodes m, out_aabb_ids m; out_bounds (host) 6 floats;
 * out_mcodes (nullable) m sorted Morton codes.  Synchronises `stream`. */
int drx_bvh_build (const float *aabbs, const int32_t *prims, int32_t n, float *out_inner,
                   int32_t *out_leaf_nodes, int32_t *out_aabb_ids, float *out_bounds, uint32_t *out_mcodes,
                   void *stream);
/* Fused leaf records for traversal: 8 floats per leaf {el_id bits, subref lo xyz, hi xyz, 0}. */
int drx_bvh_pack_leaves (const drx_bvh *bvh, float *out_leaf_pack, void *stream);
/* Newton start data per leaf (no dray counterpart; dray starts Newton from an uninitialised guess,
 * device_mesh.hpp:327-334): 12 floats per leaf, rows {M_d0, M_d1, M_d2, b_d} for d = 0..2 of the affine
 * map x0 = M p + b that inverts the leaf element's corner-affine approximation (the trilinear map
 * through its 8 corner control points, differentiated at the reference centre); b is NaN where that
 * approximation is singular.  A candidate's Newton starts at x0 clamped into its SubRef (SubRef centre
 * if NaN).  Optional: without it the same start is computed from the element's corners at each
 * candidate, one more dependent memory access. */
int drx_bvh_leaf_starts (const drx_mesh *mesh, const drx_bvh *bvh, float *out_leaf_starts, void *stream);
/* Compact traversal nodes (no dray counterpart; dray walks the 64-byte nodes above, bvh.hpp:25-38,
 * device_mesh.hpp:262-320).  Per inner node 8 uint32: {L lo x | L lo y << 16, L lo z | L hi x << 16,
 * L hi y | L hi z << 16, R lo x | R lo y << 16, R lo z | R hi x << 16, R hi y | R hi z << 16, lchild, rchild},
 * where each box corner is its cell on a 16-bit grid over the bounds, cell(v) = clamp(floor((v - origin)
 * * cells_per_unit), 0, 65535) in float32; lchild/rchild >= 0: 2 * inner index (uint4 offset), < 0: leaf
 * -(child) - 1.  Since cell() is monotone, a point in a box lies in its cell range, so a walk that tests
 * cells visits every subtree dray visits; it tests each leaf's exact float32 box (out_leaf_boxes, 8 floats
 * per leaf) before Newton, so the candidates and their order are dray's.  out_grid (host, 6 floats):
 * origin xyz, cells per unit xyz; copy it into drx_bvh.qgrid.  Needs inner_nodes, n_leaves, bounds. */
int drx_bvh_quantize (const drx_bvh *bvh, uint32_t *out_qnodes, float *out_leaf_boxes, float *out_grid, void *stream);
/* 8-wide traversal copy of the binary BVH: children in binary left-to-right order, exact float boxes,
 * so traversal visits the same candidate leaves in the same order as dray's binary depth-first walk.
 * Node = 64 floats: lo_x[8] lo_y[8] lo_z[8] hi_x[8] hi_y[8] hi_z[8] child[8] (int32 bits) pad[8].
 * out_wide NULL: only count.  *out_count (host) = number of wide nodes.  Synchronises. */
int drx_bvh_widen (const drx_bvh *bvh, float *out_wide, int64_t capacity_nodes, int64_t *out_count, void *stream);

/* ---- queries ------------------------------------------------------------------------------- */
int drx_locate (const drx_mesh *mesh, const drx_bvh *bvh, const float *points, int64_t n,
                void *out_locations, const drx_locate_opts *opts, void *stream);
/* Autotuner for drx_locate's streamed walks (no dray counterpart; dray launches one thread per point,
 * unstructured_mesh.cpp:123-148): times every chunk size in {256, 512, 1024, 2048, 4096} with each launch
 * form (one warm-up and two timed calls each, HIP events on `stream`) on the caller's points and options,
 * and records the fastest for the current device, the mesh's order, opts->fp64_polish and the walk (binary,
 * 8-wide per lane, or the learned start's), where DRX_LAUNCH_AUTO and chunk 0 then use it.  out_locations
 * (n Locations) is overwritten.  out_launch, out_chunk: the choice; out_ms (nullable, 10 floats): ms per
 * call of each candidate, chunk-major, chunked before persistent.  Not for the wave walk (DRX_ERR_ARG). */
/* Allocation-free drx_locate.  No counterpart in dray: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as drx_locate; it calls no hipMalloc/hipFree and keeps its scratch (the
 * traversal-overflow flag, the persistent launch's chunk counter and, with opts->coherent_order, the Morton
 * sort buffers) in a static device array of 4.13 MiB that each GPU the library runs on holds until
 * the process ends.  opts->coherent_order is limited to n <= DRX_LOCATE_NOALLOC_MAX_ORDERED points
 * (DRX_ERR_ARG beyond it); without it any n is accepted.  All calls on one GPU share that array: calls
 * that overlap on one GPU (two streams, two host threads) race and must be serialised by the caller. */
#define DRX_LOCATE_NOALLOC_MAX_ORDERED 262144
int drx_locate_noalloc (const drx_mesh *mesh, const drx_bvh *bvh, const float *points, int64_t n,
                        void *out_locations, const drx_locate_opts *opts, void *stream);
int drx_locate_autotune (const drx_mesh *mesh, const drx_bvh *bvh, const float *points, int64_t n,
                         void *out_locations, const drx_locate_opts *opts, int32_t *out_launch, int32_t *out_chunk,
                         float *out_ms, void *stream);
/* Field::eval: out_values[i] = field at locations[i] (untouched where cell == -1); an FP64 field
 * (field->data64) is evaluated in double and rounded to float32 on output. */
int drx_field_eval (const drx_mesh *field, const void *locations, int64_t n, float *out_values, void *stream);

/* ---- rendering ----------------------------------------------------------------------------- */
int drx_camera_rays (const drx_camera *cam, void *out_rays, void *stream);
/* Compacts rays hitting bounds (scaled x1.001) and sets near/far.  out_count is host; synchronises. */
int drx_remove_missed_rays (const void *rays, int32_t n, const float *bounds6_host, void *out_rays,
                            int32_t *out_count, void *stream);
/* Tile-major processing order for rays that came from a width x height camera (pixel_id = j*w+i):
 * fills out_order[n] with a permutation of 0..n-1 that walks 8x8 pixel tiles, tiles ordered by their
 * longest [near, far] ray segment, longest first (ties in tile order).  Results of the integration do
 * not depend on the order; it sets which rays share a wavefront and which run first. */
int drx_ray_tile_order (const void *rays, int32_t n, int32_t width, int32_t height, int32_t *out_order,
                        void *stream);
/* Volume::integrate on active rays: samples = |bounds diagonal|/sample_dist steps, field colour from
 * the (already opacity-corrected) LUT of lut_size RGBA entries over [field_min, field_max] (the colour map's
 * Float range: rounded to float32 for an FP32 field, used in double for an FP64 field, field->data64, which
 * is evaluated in double as dray's Float = double build does; unlit only), lit per
 * `shading` (NULL = unlit), front-to-back blend, <= 5 partials per ray.  out_partials must hold
 * 5*n_rays partials; on return it holds the compacted partials (depth != inf, ray order) and *out_count
 * (host) their number.  ray_order may be NULL.  stats (host, nullable).  Synchronises. */
int drx_volume_integrate (const drx_mesh *mesh, const drx_mesh *field, const drx_bvh *bvh,
                          const void *active_rays, int32_t n_rays, const float *lut, int32_t lut_size,
                          double field_min, double field_max, float sample_dist, const int32_t *ray_order,
                          const drx_locate_opts *opts, const drx_shading *shading, void *out_partials,
                          int32_t *out_count, drx_integrate_stats *stats, void *stream);
/* Renderer::render's cross-domain volume compositing (rendering/renderer.cpp:307-330, done there by
 * apcomp's PartialCompositor): partials of any number of domains in any order are sorted by (pixel id,
 * depth), blended front to back per pixel and put over a white background.  Pixel ids must be in
 * [0, width*height). */
int drx_composite_domains (const void *partials, int32_t n, int32_t width, int32_t height, float *out_rgba,
                           void *stream);
/* Renderer's default light for a camera (rendering/renderer.cpp:141-159), host. */
int drx_default_light (const drx_camera *cam, drx_point_light *out_light);
/* Partials (grouped per pixel, as integrate emits them) -> RGBA image over a white background. */
int drx_composite (const void *partials, int32_t n, int32_t width, int32_t height, float *out_rgba,
                   void *stream);

/* ---- framebuffer (rendering/framebuffer.hpp) ------------------------------------------------ */
/* A framebuffer is `colors` (RGBA, 4 floats per pixel) and `depths` (1 float per pixel) in device
 * memory, pixel id = x + y * width, row 0 at the bottom of the image. */
/* Framebuffer::clear: colors = color, depths = +infinity (depths nullable). */
int drx_framebuffer_clear (float *colors, float *depths, int32_t n_pixels, const float color[4], void *stream);
/* Renderer::render's volume compositing: partials of any number of domains, in any order, are sorted
 * by (pixel id, depth) (ties keep input order), blended front to back per pixel and written as
 * detail::partials_to_framebuffer does: blend = 0 replaces the pixel's colour, blend = 1 puts the
 * volume over the colour the framebuffer holds (colour = volume + (1 - volume alpha) * framebuffer);
 * depths (nullable) take each pixel's nearest partial depth.  Pixels without partials are untouched. */
int drx_partials_to_framebuffer (const void *partials, int32_t n, int32_t width, int32_t height, float *colors,
                                 float *depths, int32_t blend, void *stream);
/* Framebuffer::composite_background: pixels with alpha < 1 are put over bg (blend_pre_alpha). */
int drx_framebuffer_composite_background (float *colors, int32_t n_pixels, const float bg[4], void *stream);
/* Framebuffer::save: host RGBA floats -> 8-bit RGBA PNG file at `path` (rows flipped so row 0 is the
 * bottom of the picture; each channel (uint8)(v * 255), clamped to [0, 255]). */
int drx_save_png (const float *rgba, int32_t width, int32_t height, const char *path);

/* ---- learned Newton initializer ----------------------------------------------------------- */
/* MLP 27 -> 64 -> 64 -> 64 -> 64 -> 3, GELU x*Phi(x) with Phi a cubic smoothstep over [-2, 2], BF16
 * weights/activations, FP32 accumulate.
 * Features (bf16, 32 per row, 27 used): the query and the element's 8 corner control points in the frame
 * of the corner box; the network predicts a correction to base = (q' + 1) / 2.                     */
int64_t drx_mlp_param_count (void);   /* fp32 master parameters */
int64_t drx_mlp_state_bytes (void);   /* training state: master weights, Adam moments, bf16 copies */
int64_t drx_mlp_packed_bytes (void);  /* inference blob (drx_mlp_pack) */
int drx_mlp_init (void *state, uint64_t seed, void *stream);
/* Features of (points[i], locations[i].cell) rows: out_x 32 bf16 per row; out_target (nullable) 3 floats
 * per row = located ref - base (training target); out_base (nullable) 3 floats.  Unlocated rows: zeros. */
int drx_mlp_features (const drx_mesh *mesh, const float *points, const void *locations, int64_t n, void *out_x,
                      float *out_target, float *out_base, void *stream);
/* opts->steps Adam steps on mean squared error; out_loss (device, nullable) gets each step's MSE. */
int drx_mlp_train (void *state, const void *x, const float *target, int64_t n, const drx_mlp_train_opts *opts,
                   float *out_loss, void *stream);
int drx_mlp_pack (const void *state, void *out_packed, void *stream);
/* out[3 i + d] = base[3 i + d] (or 0 when base is NULL) + MLP(x_i)[d]. */
int drx_mlp_infer (const void *packed, const void *x, const float *base, int64_t n, float *out, void *stream);

/* ---- host-side colour table ---------------------------------------------------------------- */
/* ColorTable with rgb points (pos, r, g, b, a) and alpha points (pos, alpha), host arrays; opacity
 * corrected with factor (<= 0: none) then sampled to num_samples RGBA entries (host out). */
int drx_color_table_sample (const float *rgb_pts, int32_t n_rgb, const float *alpha_pts, int32_t n_alpha,
                            int32_t smooth, float factor, int32_t num_samples, float alpha_scale,
                            float *out_rgba_host);
/* The rgb control points (pos, r, g, b, 1) of a named table, host.  "" and "default" name cool2warm (the
 * ColorMap default); named tables are smooth except "sharp"; an unknown name prints a notice and takes
 * YlOrRd.  rgb_pts null or capacity < *n_pts: only *n_pts and *smooth are written. */
int drx_color_table_preset (const char *name, float *rgb_pts, int32_t capacity, int32_t *n_pts, int32_t *smooth);

/* ---- camera orbits (host, float32 arithmetic in dray's operation order) ------------------------ */
/* Look at the centre of bounds (xmin, xmax, ymin, ymax, zmin, zmax) padded by pads[d] x extent from one
 * padded diagonal away along the current viewing direction; fov_y becomes 60 degrees. */
int drx_camera_reset_to_bounds (drx_camera *cam, const float *bounds6, const double *pads);
/* Orbit the camera position about look_at by degrees about its right (elevate) or up (azimuth) axis. */
int drx_camera_elevate (drx_camera *cam, float degrees);
int drx_camera_azimuth (drx_camera *cam, float degrees);

#ifdef __cplusplus
}
#endif
#endif /* DRX_H */

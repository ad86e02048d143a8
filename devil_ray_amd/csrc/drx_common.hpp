// Shared host/device definitions of libdrx (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/drx.h"

namespace drx
{

// ---- error state (per host thread) ----------------------------------------------------------
void set_error (const char *fmt, ...);
int fail (int status, const char *fmt, ...);

#define DRX_HIP(call)                                                                           \
  do                                                                                            \
  {                                                                                             \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return ::drx::fail (DRX_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString (e_),  \
                          __FILE__, __LINE__);                                                  \
  } while (0)

#define DRX_LAUNCH_CHECK()                                                                      \
  do                                                                                            \
  {                                                                                             \
    hipError_t e_ = hipGetLastError ();                                                         \
    if (e_ != hipSuccess)                                                                       \
      return ::drx::fail (DRX_ERR_HIP, "kernel launch failed: %s (%s:%d)", hipGetErrorString (e_), \
                          __FILE__, __LINE__);                                                  \
  } while (0)

// ---- per-(device, stream) arena of growable scratch slots -------------------------------------------
// Slots are independent buffers; a call may hold several at once.  Each stream has its own set, so calls
// on different streams never share scratch; growing a slot is stream-ordered.  Contents are not preserved
// across growth.
enum ArenaSlot
{
  SLOT_SORT_KEYS_A = 0,
  SLOT_SORT_KEYS_B,
  SLOT_SORT_VALS_A,
  SLOT_SORT_VALS_B,
  SLOT_SORT_HIST,
  SLOT_TREE,
  SLOT_TREE2,
  SLOT_SCAN,
  SLOT_PARTIALS,
  SLOT_FLAGS,
  SLOT_STATS,
  SLOT_PYRAMID,
  SLOT_COUNT
};
void *arena (int slot, size_t bytes, hipStream_t stream);   // nullptr + error set on failure
int arena_release ();

inline hipStream_t as_stream (void *s) { return reinterpret_cast<hipStream_t> (s); }

// ---- POD records shared with the reference layout ---------------------------------------------
struct Ray
{
  float dir[3];
  float near_;
  float orig[3];
  float far_;
  int32_t pixel_id;
};
static_assert (sizeof (Ray) == 36, "Ray layout");

struct Partial
{
  int32_t pixel_id;
  float depth;
  float rgba[4];
};
static_assert (sizeof (Partial) == 24, "VolumePartial layout");

struct Location
{
  int32_t cell;
  float ref[3];
};
static_assert (sizeof (Location) == 16, "Location layout");

inline int npe_of (int p) { return (p + 1) * (p + 1) * (p + 1); }
inline int npe_pad_of (int p) { return (npe_of (p) + 3) & ~3; }

// Exclusive scan of int32 flags (device), n elements; writes scan[n] and *total (host).  The block sums
// live in `sums` (scan_sums_len(n) int32) when given, else in the arena's scan slot.
int exclusive_scan_i32 (const int32_t *in, int32_t *out, int32_t n, int32_t *total_host, hipStream_t st,
                        int32_t *sums = nullptr);
int32_t scan_sums_len (int32_t n);

// Stable LSD radix sort of (key, value) pairs over the low `bits` bits (8-bit digits); ping-pongs between
// A and B; *keys_out / *vals_out point at whichever holds the result.  hist holds radix_hist_len(n) int32;
// scan_sums (nullable: arena) scan_sums_len(256 * radix_tiles(n)) int32.
int radix_sort_pairs (uint32_t *kA, int32_t *vA, uint32_t *kB, int32_t *vB, int32_t *hist, int n, int bits,
                      hipStream_t st, uint32_t **keys_out, int32_t **vals_out, int32_t *scan_sums = nullptr);
int radix_tiles (int n);
inline int64_t radix_hist_len (int n) { return 2 * 256 * (int64_t)radix_tiles (n) + 8; }

// Caller-provided buffers for morton_order_points (n keys/values each, hist and scan sums as above).
struct SortScratch
{
  uint32_t *kA, *kB;
  int32_t *vA, *vB, *hist, *scan_sums;
};
// Permutation of n query points in Morton order over bounds6 (host); lives in `ws` when given, else in
// the arena's sort slots.
int morton_order_points (const float *pts, int n, const float *bounds6, hipStream_t st, const int32_t **order_out,
                         const SortScratch *ws = nullptr);

// Resolves the current device's address of drx_locate_noalloc's static scratch array (once per device).
int locate_noalloc_setup ();

} // namespace drx

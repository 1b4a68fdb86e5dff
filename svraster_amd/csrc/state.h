// Layout of the three opaque per-call state buffers (uint8 tensors on the Python side).
// The reference keeps the same three roles (cuda/src/raster_state.h:41-78) with its own private
// layout; this layout is private to this library and every sub-array is 256-byte aligned.
//
//   geometry state  (V views x P voxels): view table, counters, n_dup, inclusive scan,
//                   packed pixel bboxes, octant bitsets, scan scratch
//   binning state   (K keys): sorted keys/values, the ping-pong pair, radix-sort scratch, and
//                   the contributing entries with their pixel masks (forward -> backward)
//   image state     (pixels, tiles): tile ranges, last-contributor bound, per-pixel contributor
//                   count, work-queue counter, view table copy
#pragma once
#include <stddef.h>
#include <stdint.h>
#include "common.h"

namespace svr {

size_t scan_partial_count(long long n);
size_t sort_tile_count(long long n);
size_t sort_scratch_count(long long n);

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Carver {
  size_t off = 0;
  template <typename T>
  size_t take(long long count) {
    const size_t o = off;
    off = align256(off + sizeof(T) * (size_t)(count > 0 ? count : 0));
    return o;
  }
};

// Geometry state.
struct GeomLayout {
  size_t views, misc, n_dup, scan, bbox, quad, partial, tbox, total;
  GeomLayout(long long n_views, long long P) {
    Carver c;
    const long long VP = n_views * P;
    views = c.take<ViewDesc>(n_views);
    misc = c.take<uint32_t>(64);
    n_dup = c.take<uint32_t>(VP);
    scan = c.take<uint32_t>(VP);
    bbox = c.take<uint2>(VP);
    quad = c.take<uint8_t>(VP);
    partial = c.take<uint32_t>((long long)scan_partial_count(VP) + 1);
    tbox = c.take<uint2>(VP);
    total = c.off;
  }
};

struct GeomPtrs {
  ViewDesc* views;
  uint32_t* misc;      // [0] = low zero rank bits shared by all visible voxels; [8] = peak key
                       // count of the batched preprocesses since the last svr_batch_key_peak reset
  uint32_t* n_dup;
  uint32_t* scan;
  uint2* bbox;
  uint8_t* quad;
  uint32_t* partial;
  // Per (view, voxel): the pixels whose centre lies inside the projected corners' rectangle (with
  // a 1/32-pixel margin), packed like bbox and inside it.  Only such pixels can hit the voxel, so
  // the compositors take their candidate pixels from it.
  uint2* tbox;
  GeomPtrs(void* base, long long n_views, long long P) {
    GeomLayout L(n_views, P);
    char* b = (char*)base;
    views = (ViewDesc*)(b + L.views);
    misc = (uint32_t*)(b + L.misc);
    n_dup = (uint32_t*)(b + L.n_dup);
    scan = (uint32_t*)(b + L.scan);
    bbox = (uint2*)(b + L.bbox);
    quad = (uint8_t*)(b + L.quad);
    partial = (uint32_t*)(b + L.partial);
    tbox = (uint2*)(b + L.tbox);
  }
};

// Binning state.
struct BinLayout {
  size_t keys, vals, keys_alt, vals_alt, hist, partial, total;
  explicit BinLayout(long long K) {
    Carver c;
    keys = c.take<uint64_t>(K);
    vals = c.take<uint32_t>(K);
    keys_alt = c.take<uint64_t>(K);
    vals_alt = c.take<uint32_t>(K);
    const long long hn = 512LL * (long long)sort_tile_count(K);   // up to 9-bit digits
    hist = c.take<uint32_t>(hn);
    partial = c.take<uint32_t>((long long)sort_scratch_count(K));
    total = c.off;
  }
};

struct BinPtrs {
  uint64_t* keys;
  uint32_t* vals;
  uint64_t* keys_alt;
  uint32_t* vals_alt;
  uint32_t* hist;
  uint32_t* partial;
  BinPtrs(void* base, long long K) {
    BinLayout L(K);
    char* b = (char*)base;
    keys = (uint64_t*)(b + L.keys);
    vals = (uint32_t*)(b + L.vals);
    keys_alt = (uint64_t*)(b + L.keys_alt);
    vals_alt = (uint32_t*)(b + L.vals_alt);
    hist = (uint32_t*)(b + L.hist);
    partial = (uint32_t*)(b + L.partial);
  }
};

// Contributing entries, written by the forward compositor for the backward (K keys): each tile's
// entries that composited into at least one of its pixels, packed in list order from the start
// of the tile's list range, and per entry the pixels of each of the tile's four 8x8 blocks that
// composited it (bit = 8 * row + column).
struct ContribLayout {
  size_t cvals, pmask, total;
  explicit ContribLayout(long long K) {
    Carver c;
    cvals = c.take<uint32_t>(K);
    pmask = c.take<uint64_t>(4 * K);
    total = c.off;
  }
};

struct ContribPtrs {
  uint32_t* cvals;     // value (octant << 29 | voxel) of each contributing entry
  uint64_t* pmask;     // [entry][block]: pixels of the block that composited the entry
  ContribPtrs(void* base, long long K) {
    ContribLayout L(K);
    char* b = (char*)base;
    cvals = (uint32_t*)(b + L.cvals);
    pmask = (uint64_t*)(b + L.pmask);
  }
};

// The per-view binning state: the sort buffers, then the contributing entries.
inline size_t raster_bin_bytes(long long K) { return BinLayout(K).total + ContribLayout(K).total; }

// Image state.
struct ImageLayout {
  size_t views, misc, ranges, tile_last, n_contrib, cranges, ds_part, tile_view, total;
  ImageLayout(long long n_views, long long n_pix, long long n_tiles) {
    Carver c;
    views = c.take<ViewDesc>(n_views);
    misc = c.take<uint32_t>(64);
    ranges = c.take<uint2>(n_tiles);
    tile_last = c.take<uint32_t>(n_tiles);
    n_contrib = c.take<uint32_t>(n_pix);
    cranges = c.take<uint2>(n_tiles);
    ds_part = c.take<float2>(n_tiles);
    tile_view = c.take<uint32_t>(n_tiles);
    total = c.off;
  }
};

struct ImagePtrs {
  ViewDesc* views;
  uint32_t* misc;      // [0] = forward work counter, [1] = backward work counter
  uint2* ranges;
  uint32_t* tile_last;
  uint32_t* n_contrib;
  uint2* cranges;      // per tile: its contributing entries [x, y) in the contributing-entry arrays
  float2* ds_part;     // per backward workgroup: its dL/ds partial sum as an unevaluated (hi, lo) pair
  uint32_t* tile_view; // per tile: its view (filled at multi-view set-up)
  ImagePtrs(void* base, long long n_views, long long n_pix, long long n_tiles) {
    ImageLayout L(n_views, n_pix, n_tiles);
    char* b = (char*)base;
    views = (ViewDesc*)(b + L.views);
    misc = (uint32_t*)(b + L.misc);
    ranges = (uint2*)(b + L.ranges);
    tile_last = (uint32_t*)(b + L.tile_last);
    n_contrib = (uint32_t*)(b + L.n_contrib);
    cranges = (uint2*)(b + L.cranges);
    ds_part = (float2*)(b + L.ds_part);
    tile_view = (uint32_t*)(b + L.tile_view);
  }
};

// Batched rank-order binning state (V views, P voxels): per (voxel, octant) pair its key count
// over all views, the counts in rank order and their inclusive scan, and its view mask.
struct PairLayout {
  size_t pair_cnt, cnt_ranked, scan_ranked, partial, vmask, total;
  PairLayout(long long P, long long V) {
    Carver c;
    const long long n = 8 * P;
    pair_cnt = c.take<uint32_t>(n);
    cnt_ranked = c.take<uint32_t>(n);
    scan_ranked = c.take<uint32_t>(n);
    partial = c.take<uint32_t>((long long)scan_partial_count(n) + 1);
    vmask = c.take<uint32_t>(n * ((V + 31) / 32));
    total = c.off;
  }
};

struct PairPtrs {
  uint32_t* pair_cnt;
  uint32_t* cnt_ranked;
  uint32_t* scan_ranked;
  uint32_t* partial;
  uint32_t* vmask;      // per pair: the views (32 per word) whose octant set holds its octant
  PairPtrs(void* base, long long P, long long V) {
    PairLayout L(P, V);
    char* b = (char*)base;
    pair_cnt = (uint32_t*)(b + L.pair_cnt);
    cnt_ranked = (uint32_t*)(b + L.cnt_ranked);
    scan_ranked = (uint32_t*)(b + L.scan_ranked);
    partial = (uint32_t*)(b + L.partial);
    vmask = (uint32_t*)(b + L.vmask);
  }
};

// Launch helpers implemented in integer_stage.hip.
hipError_t launch_init_view(ViewDesc* out, const ViewDesc& base, const float* w2c, const float* c2w,
                            const float* bg, hipStream_t st);
hipError_t launch_preprocess(const ViewDesc* views, int n_views, int P, const float* centers,
                             const float* lengths, const uint8_t* is_leaf, const int64_t* octpaths,
                             int32_t* out_n_dup, uint32_t* n_dup, uint2* bboxes, uint2* tboxes, uint8_t* quads,
                             uint32_t* tz_min, hipStream_t st);
hipError_t launch_scan(const uint32_t* in, const uint32_t* n_dev, long long n_cap, uint32_t* partial,
                       uint32_t* out, bool inclusive, hipStream_t st);
hipError_t launch_duplicate(const ViewDesc* views, int n_views, int P, const int64_t* octpaths,
                            const uint32_t* n_dup, const uint32_t* scan, const uint2* bboxes,
                            const uint8_t* quads, int shift, int qb, int rank_bits, uint64_t* keys,
                            uint32_t* vals, long long cap, hipStream_t st);
hipError_t launch_radix_sort(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                             const uint32_t* n_dev, long long n_cap, int nbits, uint32_t* hist,
                             uint32_t* scan_partial, bool* result_in_alt, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, long long n, uint32_t v, hipStream_t st);
hipError_t launch_grid_pts_link(long long N, const int64_t* octpath, const int8_t* octlevel, uint64_t* keys,
                                uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t* hist,
                                uint32_t* sort_partial, uint32_t* scan_partial, uint32_t* flags, uint32_t* rank,
                                int64_t* vox_key, int64_t* grid_key, uint32_t* n_unique, hipStream_t st);
hipError_t launch_tile_ranges(const uint64_t* keys, const uint32_t* n_dev, long long n_cap, int rank_bits,
                              uint2* ranges, hipStream_t st);
int tile_sort_passes(long long n_tiles);
hipError_t launch_tile_sort(uint32_t* tiles, uint32_t* vals, uint32_t* tiles_alt, uint32_t* vals_alt,
                            const uint32_t* n_dev, long long n_cap, long long n_tiles, uint32_t* hist,
                            uint32_t* scan_partial, hipStream_t st);
hipError_t launch_pair_rank(long long n_pairs, const int64_t* octpaths, uint64_t* keys, uint32_t* vals,
                            hipStream_t st);
// pair_slots (may be null): rank-order slot of every (voxel, octant) pair; then the key counts go
// straight to pair_cnt[pair_slots[8 v + q]] (rank order), else to pair_cnt[8 v + q].
hipError_t launch_preprocess_views(const ViewDesc* views, int V, int P, const float* centers, const float* lengths,
                                   const uint8_t* is_leaf, uint32_t* n_dup, uint2* bboxes, uint2* tboxes,
                                   uint8_t* quads, const uint32_t* pair_slots, uint32_t* pair_cnt,
                                   uint32_t* pair_vmask, hipStream_t st);
// slots[order[i]] = i: the inverse of the pair rank order.
hipError_t launch_pair_slots(long long n_pairs, const uint32_t* order, uint32_t* slots, hipStream_t st);
// *peak = max(*peak, *total): the largest key count asked of a fixed-capacity binning state.
hipError_t launch_key_peak(const uint32_t* total, uint32_t* peak, hipStream_t st);
hipError_t launch_pair_gather(long long n_pairs, const uint32_t* order, const uint32_t* pair_cnt,
                              uint32_t* cnt_ranked, hipStream_t st);
hipError_t launch_pair_emit(const ViewDesc* views, int V, int P, long long n_pairs, const uint32_t* order,
                            const uint32_t* cnt_ranked, const uint32_t* scan_ranked, const uint2* bboxes,
                            const uint32_t* vmask, uint32_t* tiles, uint32_t* vals, long long cap, hipStream_t st);
hipError_t launch_tile_ranges_u32(const uint32_t* tiles, const uint32_t* n_dev, long long n_cap, uint2* ranges,
                                  hipStream_t st);

// Render launch helpers implemented in render_fwd.hip / render_bwd.hip.
struct FwdArgs {
  const ViewDesc* views; int n_views; int total_tiles;
  const uint2* ranges; const uint32_t* vals;
  const uint2* bboxes; long long P;
  const uint2* tboxes;             // candidate pixel rectangles (GeomPtrs::tbox), indexed like bboxes
  const float* centers; const float* lengths; const float* geos; const float* rgbs;
  const VoxelRecord* vrec;         // per-voxel records (multi-view path); null: centers / lengths / geos
  long long rgb_view_stride;       // floats between the colour arrays of consecutive views (0: shared)
  const float* s_val;
  uint32_t* tile_last; uint32_t* n_contrib;
  float* color; float* depth; float* normal; float* T; float* max_w;
  // Optional fused MSE colour gradient: dL_dcolor = (color - mse_target) * 2 / (3 W H) per view,
  // both laid out like color (null: not written).
  const float* mse_target; float* dL_dcolor;
  uint32_t* work_counter;          // null: one workgroup per tile
  const uint32_t* tile_view = nullptr;   // per tile: its view (work-queue launches)
  uint32_t* cvals; uint64_t* pmask; uint2* cranges;   // contributing entries (written)
  int need_depth, need_normal;
  int variant;                     // 0: pair-compacted (default), 1: one entry at a time
  int halves = 2;                  // entries staged per round: 64 * halves (pair-compacted kernel)
  int blocks_per_cu = 8;           // persistent grid size with a work counter
};
hipError_t launch_render_forward(const FwdArgs& a, hipStream_t st);

struct BwdArgs {
  const ViewDesc* views; int n_views; int total_tiles;
  const uint2* ranges; const uint32_t* tile_last; const uint32_t* vals;
  const uint2* bboxes; long long P;
  const float* centers; const float* lengths; const float* geos; const float* rgbs;
  const VoxelRecord* vrec;         // per-voxel records (multi-view path); null: centers / lengths / geos
  long long rgb_view_stride;
  const float* s_val;
  const float* T; const uint32_t* n_contrib;
  const float* dL_dcolor; const float* dL_ddepth; const float* dL_dnormal; const float* dL_dT;
  const float* gt_color; float lambda_R_concen;
  float* dL_dvox;                  // [P,12]
  float* dL_drgb_views;            // [V,P,3] per-view colour gradients (null: into dL_dvox)
  float* dL_ds;                    // [1], accumulated into
  float2* ds_part;                 // [launch grid]: per-workgroup dL/ds partials (hi, lo)
  // Corner-SDF gradients straight into the grid points (the gather's transpose fused into the
  // flush; null: into dL_dvox slots 0-7).
  const int64_t* vox_key; float* dL_dgrid;
  uint32_t* work_counter;
  const uint32_t* tile_view = nullptr;   // per tile: its view (work-queue launches)
  // Contributing entries per tile and their pixel masks, as the forward wrote them.
  const uint32_t* cvals; const uint64_t* pmask; const uint2* cranges;
  int need_depth, need_normal;
  int variant;                     // 0: pair-compacted (default), 1: one entry at a time
  int halves = 2;                  // entries staged per round: 64 * halves (pair-compacted kernel)
  int blocks_per_cu = 8;           // persistent grid size with a work counter
};
hipError_t launch_render_backward(const BwdArgs& a, hipStream_t st);

}  // namespace svr

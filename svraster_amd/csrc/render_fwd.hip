// Forward SDF compositor for gfx950.
//
// Reference: FORWARD::renderCUDA<..., n_samp = 3, SDF> (cuda/src/forward.cu:42-564).
//
// One 256-thread workgroup per 256-pixel tile (16x16, 32x8 or 8x32), four wavefronts of 64
// pixels.  The tile's sorted entries are staged into LDS with the per-entry constants the pixel
// work needs (box corners relative to the camera and 1/size, corner SDFs with their z differences,
// colour, packed bbox, normalised linear normal).  Two kernels:
//   render_fwd_kernel        one entry at a time, wave-uniform, wave w on the 8x8 quadrant w
//                            (lane l: pixel (l&7, l>>3) of it); 256 entries per round;
//   render_fwd_pairs_kernel  the default, pair-compacted (below), wave w on an interleaved 8x8
//                            lattice of the tile.
// Optionally the workgroups pull tiles from a global (view, tile) queue so one launch can serve
// many views.
#include "common.h"
#include "state.h"

namespace svr {

__device__ inline int find_view(const ViewDesc* views, int n_views, int g) {
  int lo = 0, hi = n_views - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (views[mid].tile_base <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

struct StagedEntry {
  float4 ch;    // centre xyz, half size
  float4 ml;    // box minimum xyz, 1/size
  float4 g0, g1;
  float4 rgb;   // rgb, value bits (octant << 29 | voxel)
  float4 nrm;   // normalised linear normal, 1/|lin_n|
};

__device__ inline void stage_entry(const FwdArgs& a, uint32_t e, int view, const float* __restrict__ rgbs,
                                   bool need_normal, float4* s_ch, float4* s_ml, float4* s_g0, float4* s_g1,
                                   float4* s_rgb, float4* s_nrm, uint2* s_bb, int slot, const Ray& r,
                                   bool cull = false, float s = 0.f, float cam_scale = 0.f) {
  const uint32_t val = a.vals[e];
  const uint32_t vox = val & kVoxMask;
  uint2 bb = a.tboxes[(long long)view * a.P + vox];
  VoxelGeom v;
  load_voxel_geom(v, a.vrec, a.centers, a.lengths, a.geos, vox);
  const float cx = v.cx, cy = v.cy, cz = v.cz, l = v.l;
  const float4 g0 = v.g0, g1 = v.g1;
  stage_segment(cx, cy, cz, l, v.l_inv, g0, g1, r, s_ch[slot], s_ml[slot], s_g0[slot], s_g1[slot]);
  if (cull) {
    // An entry that cannot reach the minimum alpha gets an empty bbox (no pixel masks).
    const float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
    const float ps = cam_scale + fmaxf(fabsf(cx), fmaxf(fabsf(cy), fabsf(cz)));
    if (!voxel_may_contribute(g, s, ps, l)) bb = make_uint2(0xFFFFFFFFu, 0u);
  }
  s_bb[slot] = bb;
  s_rgb[slot] = make_float4(rgbs[3 * vox + 0], rgbs[3 * vox + 1], rgbs[3 * vox + 2], __uint_as_float(val));
  if (need_normal) {
    const float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
    float nx, ny, nz, rn;
    lin_normal(g, nx, ny, nz, rn);
    s_nrm[slot] = make_float4(nx * rn, ny * rn, nz * rn, rn);
  }
}

// One staged entry held in registers between its global loads and its LDS store, so the loads
// of the next batch are in flight while the current batch is composited.
struct EntryRegs {
  uint32_t val;
  uint2 bb;
  VoxelGeom v;
  float r, g, b;
};

__device__ inline void fetch_entry(EntryRegs& e, uint32_t val, int view, const FwdArgs& a,
                                   const float* __restrict__ rgbs) {
  const uint32_t vox = val & kVoxMask;
  e.val = val;
  e.bb = a.tboxes[(long long)view * a.P + vox];
  load_voxel_geom(e.v, a.vrec, a.centers, a.lengths, a.geos, vox);
  e.r = rgbs[3 * vox + 0]; e.g = rgbs[3 * vox + 1]; e.b = rgbs[3 * vox + 2];
}

// LDS copy of a fetched entry.  With kCull, an entry that cannot reach the minimum alpha gets an
// empty bbox (no pixel masks); the multi-view path has left such voxels out of binning already.
template <bool kCull>
__device__ inline void store_entry(const EntryRegs& e, bool need_normal, float s, float cam_scale, const Ray& r,
                                   float4* s_ch, float4* s_ml, float4* s_g0, float4* s_g1, float4* s_rgb, float4* s_nrm,
                                   uint2* s_bb, int slot) {
  const VoxelGeom& v = e.v;
  stage_segment(v.cx, v.cy, v.cz, v.l, v.l_inv, v.g0, v.g1, r, s_ch[slot], s_ml[slot], s_g0[slot], s_g1[slot]);
  const float g[8] = {v.g0.x, v.g0.y, v.g0.z, v.g0.w, v.g1.x, v.g1.y, v.g1.z, v.g1.w};
  const float ps = cam_scale + fmaxf(fabsf(v.cx), fmaxf(fabsf(v.cy), fabsf(v.cz)));
  s_bb[slot] = (!kCull || voxel_may_contribute(g, s, ps, v.l)) ? e.bb : make_uint2(0xFFFFFFFFu, 0u);
  s_rgb[slot] = make_float4(e.r, e.g, e.b, __uint_as_float(e.val));
  if (need_normal) {
    float nx, ny, nz, rn;
    lin_normal(g, nx, ny, nz, rn);
    s_nrm[slot] = make_float4(nx * rn, ny * rn, nz * rn, rn);
  }
}

// Median depth refined with 16 sub-steps inside the median voxel (forward.cu:450-521).
__device__ inline float median_depth(const FwdArgs& a, const Ray& r, int v, float dmed_T, float s) {
  VoxelGeom e;
  load_voxel_geom(e, a.vrec, a.centers, a.lengths, a.geos, (uint32_t)v);
  const float cx = e.cx, cy = e.cy, cz = e.cz, l = e.l;
  const float gg[8] = {e.g0.x, e.g0.y, e.g0.z, e.g0.w, e.g1.x, e.g1.y, e.g1.z, e.g1.w};
  float ta, tb;
  ray_box(r, cx, cy, cz, 0.5f * l, ta, tb);
  const float l_inv = e.l_inv;
  const float step = (tb - ta) * (1.f / 16.f);
  float Dmed = ta - 0.5f * step;
  float qx = (r.ox + ta * r.dx - (cx - 0.5f * l)) * l_inv;
  float qy = (r.oy + ta * r.dy - (cy - 0.5f * l)) * l_inv;
  float qz = (r.oz + ta * r.dz - (cz - 0.5f * l)) * l_inv;
  const float sx = step * r.dx * l_inv, sy = step * r.dy * l_inv, sz = step * r.dz * l_inv;
  const float ks = sdf_cdf_scale(s);
  float phi_prev = sdf_cdf(trilerp(gg, qx, qy, qz), ks);
  float Tm = dmed_T;
  for (int k = 0; k < 16 && Tm > 0.5f; ++k) {
    Dmed += step;
    qx += sx; qy += sy; qz += sz;
    const float phi = sdf_cdf(trilerp(gg, qx, qy, qz), ks);
    Tm *= 1.f - (phi_prev - phi) / (phi_prev + kPhiEps);
    phi_prev = phi;
  }
  return Dmed;
}

// The MSE target of pixel (px, py) (fused colour gradient), loaded at the start of the tile so
// that the loads are in flight while the tile is composited.
struct PixelTarget {
  float t0 = 0.f, t1 = 0.f, t2 = 0.f;
};

__device__ inline PixelTarget load_target(const FwdArgs& a, const ViewDesc& V, bool inside, int px, int py) {
  PixelTarget g;
  if (a.dL_dcolor && inside) {
    const long long HW = (long long)V.W * V.H;
    const long long o = V.pix_base * 3 + (long long)py * V.W + px;
    g.t0 = a.mse_target[o];
    g.t1 = a.mse_target[o + HW];
    g.t2 = a.mse_target[o + 2 * HW];
  }
  return g;
}

template <bool kDepth, bool kNormal>
__device__ inline void write_pixel(const FwdArgs& a, const ViewDesc& V, const Ray& r, bool inside, int px, int py,
                                   float T, float C0, float C1, float C2, float D, float Dmed, float N0,
                                   float N1, float N2, uint32_t last, const PixelTarget& tg) {
  if (!inside) return;
  const long long HW = (long long)V.W * V.H;
  const long long pid = (long long)py * V.W + px;
  const long long pb = V.pix_base;
  const float c0 = C0 + T * V.bg[0], c1 = C1 + T * V.bg[1], c2 = C2 + T * V.bg[2];
  a.color[pb * 3 + 0 * HW + pid] = c0;
  a.color[pb * 3 + 1 * HW + pid] = c1;
  a.color[pb * 3 + 2 * HW + pid] = c2;
  if (a.dL_dcolor) {
    // gradient of the view's mean squared error over its 3 W H colour values
    const float sc = 2.f / (3.f * (float)HW);
    const long long o = pb * 3 + pid;
    a.dL_dcolor[o] = (c0 - tg.t0) * sc;
    a.dL_dcolor[o + HW] = (c1 - tg.t1) * sc;
    a.dL_dcolor[o + 2 * HW] = (c2 - tg.t2) * sc;
  }
  a.T[pb + pid] = T;
  a.n_contrib[pb + pid] = last;
  if (kDepth) {
    a.depth[pb * 4 + 0 * HW + pid] = D * r.norm_inv;
    a.depth[pb * 4 + 2 * HW + pid] = Dmed * r.norm_inv;
  }
  if (kNormal) {
    a.normal[pb * 3 + 0 * HW + pid] = N0;
    a.normal[pb * 3 + 1 * HW + pid] = N1;
    a.normal[pb * 3 + 2 * HW + pid] = N2;
  }
}

template <bool kDepth, bool kNormal, bool kMaxW>
__global__ __launch_bounds__(256) void render_fwd_kernel(FwdArgs a) {
  __shared__ float4 s_ch[256], s_ml[256], s_g0[256], s_g1[256], s_rgb[256];
  __shared__ float4 s_nrm[kNormal ? 256 : 1];
  __shared__ uint2 s_bb[256];
  __shared__ uint32_t s_red[4];
  __shared__ int s_tile;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const float s = *a.s_val;
  const uint64_t lt_mask = (1ull << lane) - 1ull;

  for (int iter = 0;; ++iter) {
    int g;
    if (a.work_counter) {
      if (t == 0) s_tile = (int)atomicAdd(a.work_counter, 1u);
      __syncthreads();
      g = __builtin_amdgcn_readfirstlane(s_tile);   // uniform: tile and view data in scalar registers
      __syncthreads();
      if (g >= a.total_tiles) return;
    } else {
      if (iter > 0) return;
      g = blockIdx.x;
    }
    const int vi = find_view(a.views, a.n_views, g);
    const ViewDesc& V = a.views[vi];
    const int W = V.W, H = V.H;
    const int lt = g - V.tile_base;
    const int tx = lt % V.tiles_x, ty = lt / V.tiles_x;
    int px, py;
    tile_quadrant(V, tx, ty, w, px, py);
    px += lane & 7;
    py += lane >> 3;
    const bool inside = px < W && py < H;
    const Ray r = pixel_ray(V, px, py);
    const uint2 rg = a.ranges[g];
    const int n = (int)(rg.y - rg.x);
    const PixelTarget tg = load_target(a, V, inside, px, py);

    float T = 1.f, C0 = 0.f, C1 = 0.f, C2 = 0.f, D = 0.f, N0 = 0.f, N1 = 0.f, N2 = 0.f;
    uint32_t last = 0;
    int dmed_vox = -1;
    float dmed_T = 0.f;
    bool done = !inside;

    // Every staged entry is handed to the backward with its pixel masks (empty where no pixel of
    // the block composited it): the tile's contributing range is [start, tile_last).
    for (int base = 0; base < n; base += 256) {
      if (__syncthreads_count(done) == 256) break;
      if (base + t < n) {
        stage_entry(a, rg.x + base + t, vi, a.rgbs + (long long)vi * a.rgb_view_stride, kNormal, s_ch, s_ml,
                    s_g0, s_g1, s_rgb, s_nrm, s_bb, t, r);
        a.cvals[rg.x + base + t] = a.vals[rg.x + base + t];
      }
      __syncthreads();
      const int cnt = min(256, n - base);
      uint64_t cm0 = 0ull, cm1 = 0ull, cm2 = 0ull, cm3 = 0ull;   // entries this pixel composited
      for (int j = 0; j < cnt; ++j) {
        if (__all(done)) break;
        if (done) continue;
        const uint2 bb = s_bb[j];
        const float4 rgb = s_rgb[j];
        const uint32_t val = __float_as_uint(rgb.w);
        if (!pixel_in_bbox(px, py, bb.x, bb.y) || (val >> 29) != r.octant) continue;
        const float4 box = s_ch[j];
        float ta, tb;
        ray_box_st(r, box, s_ml[j], ta, tb);
        if (ta > tb) continue;
        Segment seg;
        eval_segment_st(r, ta, tb, box, s_g0[j], s_g1[j], s, seg);
        if (seg.alpha < kMinAlpha) continue;
        const float wgt = T * seg.alpha;
        C0 += wgt * rgb.x; C1 += wgt * rgb.y; C2 += wgt * rgb.z;
        if (kDepth) {
          float sub[3];
          sub_alphas(seg, sub);
          D += T * depth_value(sub, ta, tb);
          if (T > 0.5f) { dmed_vox = (int)(val & kVoxMask); dmed_T = T; }
        }
        if (kNormal) {
          const float4 nn = s_nrm[j];
          N0 += wgt * nn.x; N1 += wgt * nn.y; N2 += wgt * nn.z;
        }
        T *= (1.f - seg.alpha);
        done = T < kEarlyStopT;
        last = (uint32_t)(base + j + 1);
        if (kMaxW) atomicMax(reinterpret_cast<uint32_t*>(a.max_w) + (val & kVoxMask), __float_as_uint(wgt));
        const uint64_t bit = 1ull << (j & 63);
        const int q = j >> 6;   // wave-uniform
        if (q == 0) cm0 |= bit;
        else if (q == 1) cm1 |= bit;
        else if (q == 2) cm2 |= bit;
        else cm3 |= bit;
      }
      const long long o = (long long)rg.x + base;
      const uint64_t e0 = transpose64(cm0, lane), e1 = transpose64(cm1, lane);
      const uint64_t e2 = transpose64(cm2, lane), e3 = transpose64(cm3, lane);
      if (lane < cnt) a.pmask[4 * (o + lane) + w] = e0;
      if (64 + lane < cnt) a.pmask[4 * (o + 64 + lane) + w] = e1;
      if (128 + lane < cnt) a.pmask[4 * (o + 128 + lane) + w] = e2;
      if (192 + lane < cnt) a.pmask[4 * (o + 192 + lane) + w] = e3;
    }

    float Dmed = 0.f;
    if (kDepth && inside && dmed_vox >= 0) Dmed = median_depth(a, r, dmed_vox, dmed_T, s);
    write_pixel<kDepth, kNormal>(a, V, r, inside, px, py, T, C0, C1, C2, D, Dmed, N0, N1, N2, last, tg);
    // tile_last = max over the tile's pixels of (range start + contributors) (forward.cu:561).
    const uint32_t tl = wave_max_u32(rg.x + (inside ? last : 0u));
    if (lane == 0) s_red[w] = tl;
    __syncthreads();
    if (t == 0) {
      const uint32_t tlast = max(max(s_red[0], s_red[1]), max(s_red[2], s_red[3]));
      a.tile_last[g] = tlast;
      a.cranges[g] = make_uint2(rg.x, tlast);
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// Pair compositor (default).  Most voxels cover a handful of a tile's pixels and most (pixel,
// voxel) candidates have alpha below the minimum, so the cost must follow the pairs that exist,
// not the entries times the pixels.  Wave w takes the interleaved lattice of the tile's pixels
// (wave_lattice: strides 2x2 on 16x16 tiles), so every entry's footprint falls about evenly on the
// four waves and their work per round is balanced.  Per staged batch of 128 entries each wave:
//   1. builds the 64-bit mask of every entry's pixels in its lattice (candidate rectangle --
//      the pixel centres inside the projected corners' rectangle, within the reference's bbox --,
//      octant, pixels still unsaturated; empty for an entry whose corner SDFs cannot reach the
//      minimum alpha),
//      lane l holding entries l and l + 64, and prefix-sums the counts: entry j owns the run
//      [pref_j, pref_j + |m_j|) of an entry-major pair list;
//   2. walks that list 64 pairs per instruction (the entry of a position from a prefix maximum
//      over run heads, the pixel as the k-th set bit of the entry's mask): ray/box test, then the
//      SDF segment for the hits, alpha (and depth value) stored per position;
//   3. records the contributing pairs (alpha >= kMinAlpha) in a 128-bit entry set per pixel;
//   4. composites each pixel over its set in entry order.
// Per-pixel arithmetic and order are those of the one-entry-at-a-time kernel, so the outputs are
// identical bit for bit.  Each round's contributing entries (and, per contributing entry and 8x8
// quadrant, the pixels that composited it) are written for the backward.
// ------------------------------------------------------------------------------------------
constexpr int kPairWin = 256;     // pair positions held per wave at once

// Pixels of the wave's 8x8 lattice (bit = 8 * row + column) inside a packed bbox.
__device__ inline uint64_t bbox_block_mask(uint32_t b0, uint32_t b1, const WaveLattice& L) {
  const int mx = (1 << L.kx) - 1, my = (1 << L.ky) - 1;
  const int x0 = max(((int)(b0 >> 16) - L.ox + mx) >> L.kx, 0), x1 = min(((int)(b1 >> 16) - L.ox) >> L.kx, 7);
  const int y0 = max(((int)(b0 & 0xFFFFu) - L.oy + my) >> L.ky, 0), y1 = min(((int)(b1 & 0xFFFFu) - L.oy) >> L.ky, 7);
  if (x0 > x1 || y0 > y1) return 0ull;
  const uint32_t cols = (0xFFu >> (7 - x1)) & (0xFFu << x0);     // columns x0..x1 of a row
  const uint32_t rep = __builtin_amdgcn_perm(0u, cols, 0u);       // that byte in all four bytes
  const uint64_t rows = (~0ull << (8 * y0)) & (~0ull >> (8 * (7 - y1)));   // bytes of rows y0..y1
  return ((uint64_t)rep << 32 | rep) & rows;
}

__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}


// p + popcount(m): two v_bcnt with accumulator.  Written as such because the compiler otherwise
// distributes a following scale-by-4 of the sum over its three terms (three more shifts).
__device__ inline uint32_t pref_plus_rank(uint32_t p, uint64_t m) {
  uint32_t r, k;
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"((uint32_t)m), "v"(p));
  asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(k) : "v"((uint32_t)(m >> 32)), "v"(r));
  return k;
}

// Mask of entry slot j for the wave's pixels: bbox, octant of the pixel rays, live pixels.
__device__ inline uint64_t entry_block_mask(const uint2* s_bb, const float4* s_rgb, int j, const WaveLattice& L,
                                            const uint64_t* s_om, uint64_t live) {
  const uint2 bb = s_bb[j];
  const uint32_t q = __float_as_uint(s_rgb[j].w) >> 29;
  return bbox_block_mask(bb.x, bb.y, L) & live & s_om[q];
}

template <bool kDepth, bool kNormal, bool kMaxW, bool kQueue, int kHalves>
__global__ __launch_bounds__(256, kDepth ? 3 : 4) void render_fwd_pairs_kernel(FwdArgs a) {
  constexpr int kStage = 64 * kHalves;   // entries staged per round (kStage or half of it)
  __shared__ float4 s_ch[kStage], s_ml[kStage], s_g0[kStage], s_g1[kStage], s_rgb[kStage];
  __shared__ float4 s_nrm[kNormal ? kStage : 1];
  __shared__ uint2 s_bb[kStage];
  __shared__ float4 s_rd[4][64];              // direction xyz
  __shared__ float4 s_ri[4][64];              // 1/direction xyz
  __shared__ uint64_t s_om[4][8];             // pixels of the block whose ray has octant q
  __shared__ uint64_t s_m[4][kStage];     // per entry: its pixels in the block
  __shared__ uint16_t s_pref[4][kStage];  // per entry: first position of its run
  __shared__ unsigned long long s_cs[4][2][64];   // per pixel: contributing entries (lo, hi halves)
  __shared__ uint8_t s_own[4][kPairWin];      // 1 + entry whose run starts at the position, else 0
  __shared__ float s_pa[4][kPairWin];         // alpha per position (read for contributing pairs only)
  __shared__ float s_pd[4][kDepth ? kPairWin : 1];
  __shared__ unsigned long long s_em[kStage][4];      // per staged entry and 8x8 quadrant: pixels that composited it
  __shared__ uint64_t s_uu[kHalves];          // per staging wave: its entries that some pixel composited
  __shared__ uint32_t s_red[4];
  __shared__ int s_tile[2];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const float s = *a.s_val;
  const uint64_t lt_mask = (1ull << lane) - 1ull;

  for (int iter = 0;; ++iter) {
    int g;
    if (kQueue) {
      // Two slots by iteration parity: a slot is rewritten two tiles later, after every wave has
      // passed the barriers of the tile in between, so one barrier per tile suffices.
      if (t == 0) s_tile[iter & 1] = (int)atomicAdd(a.work_counter, 1u);
      __syncthreads();
      g = __builtin_amdgcn_readfirstlane(s_tile[iter & 1]);   // uniform: tile and view data in scalar registers
      if (g >= a.total_tiles) return;
    } else {
      if (iter > 0) return;
      g = blockIdx.x;
    }
    const int vi = kQueue ? (int)a.tile_view[g] : find_view(a.views, a.n_views, g);
    const ViewDesc& V = a.views[vi];
    const int W = V.W, H = V.H;
    const int lt = g - V.tile_base;
    const int tx = lt % V.tiles_x, ty = lt / V.tiles_x;
    const WaveLattice L = wave_lattice(V, tx, ty, w);
    const int px = lattice_px(L, lane), py = lattice_py(L, lane);
    const bool inside = px < W && py < H;
    const Ray r = pixel_ray(V, px, py);
    const PixelTarget tg = load_target(a, V, inside, px, py);
    const float cam_scale = fmaxf(fabsf(r.ox), fmaxf(fabsf(r.oy), fabsf(r.oz)));
    s_rd[w][lane] = make_float4(r.dx, r.dy, r.dz, 0.f);
    s_ri[w][lane] = make_float4(r.ix, r.iy, r.iz, 0.f);
    {
      uint64_t om = 0ull;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint64_t b = __ballot(inside && r.octant == (uint32_t)q);
        om = (lane == q) ? b : om;
      }
      if (lane < 8) s_om[w][lane] = om;
    }
    const uint2 rg = a.ranges[g];
    const int n = (int)(rg.y - rg.x);

    float T = 1.f, C0 = 0.f, C1 = 0.f, C2 = 0.f, D = 0.f, N0 = 0.f, N1 = 0.f, N2 = 0.f;
    uint32_t last = 0;
    int dmed_vox = -1;
    float dmed_T = 0.f;
    bool done = !inside;

    // Entry data one batch ahead in registers, and its value one more batch ahead.
    const float* rgbs_v = a.rgbs + (long long)vi * a.rgb_view_stride;
    EntryRegs er;
    uint32_t nval = 0;
    if (t < kStage) {
      if (t < n) fetch_entry(er, a.vals[rg.x + t], vi, a, rgbs_v);
      if (kStage + t < n) nval = a.vals[rg.x + kStage + t];
    }

    // A round's contributing entries are known once every wave has finished it, so they are
    // written at the top of the next round: packed after the tile's earlier ones (cc so far), the
    // staging lane writing its entry's value and its four quadrant pixel masks.  The masks are
    // built during the composite: every pixel ORs its bit into s_em[entry][its quadrant] for each
    // entry it composites.  Between the two barriers at the top of a round the staging lanes read
    // and clear their rows (no wave composites then), and each staging wave leaves the ballot of
    // its composited entries in s_uu.
    uint64_t qbit;   // this lane's pixel as a bit of its 8x8 quadrant's mask
    int qw;          // that quadrant
    {
      const int xt = px - tx * V.tile_w, yt = py - ty * V.tile_h;
      qw = (xt >> 3) + (V.tile_w >> 3) * (yt >> 3);
      qbit = 1ull << ((xt & 7) + 8 * (yt & 7));
    }
    unsigned long long* const s_em_q = &s_em[0][qw];
    uint32_t cc = 0;
    uint32_t sval = 0;   // value of the entry this lane staged in the last round
    if (t < kStage) { s_em[t][0] = 0ull; s_em[t][1] = 0ull; s_em[t][2] = 0ull; s_em[t][3] = 0ull; }
    for (int base = 0;; base += kStage) {
      const int ndone = __syncthreads_count(done || base >= n);
      const bool fin = ndone == 256;
      uint64_t e0 = 0ull, e1 = 0ull, e2 = 0ull, e3 = 0ull;
      if (t < kStage) {
        e0 = s_em[t][0]; e1 = s_em[t][1]; e2 = s_em[t][2]; e3 = s_em[t][3];
        s_em[t][0] = 0ull; s_em[t][1] = 0ull; s_em[t][2] = 0ull; s_em[t][3] = 0ull;
      }
      const bool comp = ((e0 | e1) | (e2 | e3)) != 0ull;
      const uint64_t ub = __ballot(comp);
      if (w < kHalves && lane == 0) s_uu[w] = ub;
      const uint32_t pval = sval;
      if (!fin && t < kStage && base + t < n) {
        store_entry<!kQueue>(er, kNormal, s, cam_scale, r, s_ch, s_ml, s_g0, s_g1, s_rgb, s_nrm, s_bb, t);
        sval = er.val;
      }
      __syncthreads();
      {
        const uint64_t u0 = s_uu[0], u1 = kHalves > 1 ? s_uu[kHalves - 1] : 0ull;
        const uint32_t n0 = (uint32_t)__popcll(u0);
        if (comp) {   // waves 0 and 1 staged entries 0-63 and 64-127
          const long long idx = (long long)rg.x + cc + (w == 0 ? 0u : n0) + (uint32_t)__popcll(ub & lt_mask);
          a.cvals[idx] = pval;
          ulonglong2* pm = reinterpret_cast<ulonglong2*>(a.pmask + 4 * idx);
          pm[0] = make_ulonglong2(e0, e1);
          pm[1] = make_ulonglong2(e2, e3);
        }
        cc += n0 + (uint32_t)__popcll(u1);
      }
      if (fin) break;
      if (t < kStage) {
        if (base + kStage + t < n) fetch_entry(er, nval, vi, a, rgbs_v);
        if (base + 2 * kStage + t < n) nval = a.vals[rg.x + base + 2 * kStage + t];
      }
      const int cnt = min(kStage, n - base);
      const uint64_t live = __ballot(!done);

      // 1. entry masks (empty once every pixel is saturated) and the entry-major run offsets.
      const uint64_t m_lo = lane < cnt ? entry_block_mask(s_bb, s_rgb, lane, L, s_om[w], live) : 0ull;
      const uint64_t m_hi = (kHalves > 1 && lane + 64 < cnt)
                                ? entry_block_mask(s_bb, s_rgb, lane + 64, L, s_om[w], live) : 0ull;
      const uint32_t c_lo = (uint32_t)__popcll(m_lo), c_hi = (uint32_t)__popcll(m_hi);
      if (__ballot((c_lo | c_hi) != 0u) != 0ull) {
        const uint32_t i_lo = wave_incl_scan_u32(c_lo);
        const uint32_t i_hi = kHalves > 1 ? wave_incl_scan_u32(c_hi) : 0u;
        const uint32_t t_lo = (uint32_t)__builtin_amdgcn_readlane((int)i_lo, 63);
        const uint32_t total = t_lo + (uint32_t)__builtin_amdgcn_readlane((int)i_hi, 63);
        const uint32_t p_lo = i_lo - c_lo, p_hi = t_lo + i_hi - c_hi;
        s_m[w][lane] = m_lo;
        s_pref[w][lane] = (uint16_t)p_lo;
        if (kHalves > 1) {
          s_m[w][lane + 64] = m_hi;
          s_pref[w][lane + 64] = (uint16_t)p_hi;
        }
        uint32_t carry = 0;   // 1 + entry owning the last position handled so far

        for (uint32_t w0 = 0; w0 < total; w0 += kPairWin) {
          const uint32_t w1 = min(total, w0 + (uint32_t)kPairWin);
          const uint32_t np = w1 - w0;
          for (uint32_t k = lane; k < np; k += 64) s_own[w][k] = 0;
          s_cs[w][0][lane] = 0ull;
          s_cs[w][1][lane] = 0ull;
          wave_lds_sync();
          if (c_lo && p_lo >= w0 && p_lo < w1) s_own[w][p_lo - w0] = (uint8_t)(lane + 1);
          if (c_hi && p_hi >= w0 && p_hi < w1) s_own[w][p_hi - w0] = (uint8_t)(lane + 65);
          wave_lds_sync();

          // 2. + 3. per position: ray/box test, alpha (and depth value) of the segment; contributing
          // pairs into the pixel's set.  Misses are evaluated too (their alpha is not used): about
          // half of the positions miss, and listing the hits for a second pass cost more than the
          // segments it saved.
          for (uint32_t c0 = 0; c0 < np; c0 += 64) {
            const uint32_t k = c0 + lane;
            const bool valid = k < np;
            uint32_t h = valid ? (uint32_t)s_own[w][k] : 0u;
            h = max(wave_incl_max_u32(h), carry);
            carry = (uint32_t)__builtin_amdgcn_readlane((int)h, 63);
            if (valid) {
              const int j = (int)h - 1;
              const int bit = select_set_bit(s_m[w][j], (int)(w0 + k - (uint32_t)s_pref[w][j]));
              const Ray pr = pair_ray(r, s_rd[w][bit], s_ri[w][bit]);
              const float4 box = s_ch[j];
              float ta, tb;
              ray_box_st(pr, box, s_ml[j], ta, tb);
              const bool hit = !(ta > tb);
              Segment seg;
              eval_segment_st(pr, ta, tb, box, s_g0[j], s_g1[j], s, seg);
              s_pa[w][k] = seg.alpha;
              if (kDepth) {
                float sub[3];
                sub_alphas(seg, sub);
                s_pd[w][k] = depth_value(sub, ta, tb);
              }
              if (hit && seg.alpha >= kMinAlpha) atomicOr(&s_cs[w][j >> 6][bit], 1ull << (j & 63));
            }
          }
          wave_lds_sync();
          // 4. composite this pixel's contributing pairs of the window, in entry order.
          // The 128-bit set as the word being consumed and the one after it (entries jb + bit).
          // The word being consumed is kept non-zero while entries remain: it takes the next word
          // as soon as it empties, so the loop tests one word.
          uint64_t cur = s_cs[w][0][lane], nxt = s_cs[w][1][lane];
          int jb = 0;
          if (cur == 0ull) { cur = nxt; nxt = 0ull; jb = 64; }
          const float* const pa = &s_pa[w][0] - w0;   // indexed by absolute position
          const float* const pd = &s_pd[w][0] - w0;
          while (cur != 0ull && !done) {
            const int j = jb + __builtin_ctzll(cur);
            cur &= cur - 1ull;
            if (cur == 0ull) { cur = nxt; nxt = 0ull; jb = 64; }
            const uint32_t k = pref_plus_rank(s_pref[w][j], s_m[w][j] & lt_mask);
            const float alpha = pa[k];
            const float4 rgb = s_rgb[j];
            const float wgt = T * alpha;
            C0 += wgt * rgb.x; C1 += wgt * rgb.y; C2 += wgt * rgb.z;
            if (kDepth) {
              D += T * pd[k];
              if (T > 0.5f) { dmed_vox = (int)(__float_as_uint(rgb.w) & kVoxMask); dmed_T = T; }
            }
            if (kNormal) {
              const float4 nn = s_nrm[j];
              N0 += wgt * nn.x; N1 += wgt * nn.y; N2 += wgt * nn.z;
            }
            T *= (1.f - alpha);
            done = T < kEarlyStopT;
            last = (uint32_t)(base + j + 1);
            atomicOr(s_em_q + 4 * j, qbit);
            if (kMaxW)
              atomicMax(reinterpret_cast<uint32_t*>(a.max_w) + (__float_as_uint(rgb.w) & kVoxMask), __float_as_uint(wgt));
          }
          wave_lds_sync();
        }
      }
    }

    float Dmed = 0.f;
    if (kDepth && inside && dmed_vox >= 0) Dmed = median_depth(a, r, dmed_vox, dmed_T, s);
    write_pixel<kDepth, kNormal>(a, V, r, inside, px, py, T, C0, C1, C2, D, Dmed, N0, N1, N2, last, tg);
    const uint32_t tl = wave_max_u32(rg.x + (inside ? last : 0u));
    if (lane == 0) s_red[w] = tl;
    __syncthreads();
    if (t == 0) {
      a.tile_last[g] = max(max(s_red[0], s_red[1]), max(s_red[2], s_red[3]));
      a.cranges[g] = make_uint2(rg.x, rg.x + cc);
    }
    __syncthreads();
  }
}

template <bool D, bool N>
static hipError_t launch_fwd_mw(const FwdArgs& a, hipStream_t st, unsigned grid) {
  if (a.variant == 1) {
    if (a.max_w) render_fwd_kernel<D, N, true><<<grid, 256, 0, st>>>(a);
    else render_fwd_kernel<D, N, false><<<grid, 256, 0, st>>>(a);
  } else {
    // The work-queue (multi-view) launch is its own instantiation, so profiles tell it apart; it
    // also comes with 64-entry rounds (halves = 1) for the tuner.
    if (a.work_counter) {
      if (a.max_w) render_fwd_pairs_kernel<D, N, true, true, 2><<<grid, 256, 0, st>>>(a);
      else if (a.halves == 1) render_fwd_pairs_kernel<D, N, false, true, 1><<<grid, 256, 0, st>>>(a);
      else render_fwd_pairs_kernel<D, N, false, true, 2><<<grid, 256, 0, st>>>(a);
    } else {
      if (a.max_w) render_fwd_pairs_kernel<D, N, true, false, 2><<<grid, 256, 0, st>>>(a);
      else render_fwd_pairs_kernel<D, N, false, false, 2><<<grid, 256, 0, st>>>(a);
    }
  }
  return hipGetLastError();
}

hipError_t launch_render_forward(const FwdArgs& a, hipStream_t st) {
  if (a.total_tiles <= 0) return hipSuccess;
  unsigned grid = (unsigned)a.total_tiles;
  if (a.work_counter) grid = (unsigned)min(a.total_tiles, 256 * max(1, a.blocks_per_cu));
  if (a.need_depth && a.need_normal) return launch_fwd_mw<true, true>(a, st, grid);
  if (a.need_depth) return launch_fwd_mw<true, false>(a, st, grid);
  if (a.need_normal) return launch_fwd_mw<false, true>(a, st, grid);
  return launch_fwd_mw<false, false>(a, st, grid);
}

}  // namespace svr

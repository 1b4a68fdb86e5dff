// Reverse-order backward of the SDF compositor for gfx950.
//
// Reference: BACKWARD::renderCUDA<..., n_samp = 3, SDF> (cuda/src/backward.cu:38-851) and its
// host side (backward.cu:986-1109).
//
// Same tile / wave / LDS staging as the forward, walking each tile's list back to front from
// the saved last-contributor bound.  Transmittance is recovered per pixel by T <- T / (1 - a)
// from the saved final T; the alpha is recomputed by the exact same device function as the
// forward, so the recovery divides by the same (1 - a) the forward multiplied by.
//
// Per-voxel gradients are reduced across the wave before touching global memory: for each
// staged entry the 12 gradient values [8 corner SDFs | rgb | |dL/da * a|] of all lanes whose
// pixel used the entry are summed with DPP row reductions (or read from the single contributing
// lane), then lanes 0..11 issue one vector atomic add of 12 consecutive floats.  dL/ds, a
// single scalar summed over every (pixel, entry) pair with terms of both signs, is accumulated
// per lane with compensated FP32 sums, reduced per tile into a (hi, lo) partial and summed over
// the tiles in a fixed order by a one-workgroup kernel: deterministic, no float atomics.
#include "common.h"
#include "state.h"

namespace svr {

__device__ inline int find_view_b(const ViewDesc* views, int n_views, int g) {
  int lo = 0, hi = n_views - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (views[mid].tile_base <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// One staged entry held in registers between its global loads and its LDS store, so the loads
// of the next batch are in flight while the current batch is processed.
struct EntryRegsB {
  uint32_t val;
  VoxelGeom v;
  float r, g, b;
};

__device__ inline void fetch_entry_b(EntryRegsB& e, const BwdArgs& a, uint32_t val, const float* __restrict__ rgbs) {
  const uint32_t vox = val & kVoxMask;
  e.val = val;
  load_voxel_geom(e.v, a.vrec, a.centers, a.lengths, a.geos, vox);
  e.r = rgbs[3 * vox]; e.g = rgbs[3 * vox + 1]; e.b = rgbs[3 * vox + 2];
}

__device__ inline void store_entry_b(const EntryRegsB& e, bool need_normal, const Ray& r, float4* s_ch, float4* s_ml,
                                     float4* s_g0, float4* s_g1, float4* s_rgb, float4* s_nrm, int slot) {
  const VoxelGeom& v = e.v;
  stage_segment(v.cx, v.cy, v.cz, v.l, v.l_inv, v.g0, v.g1, r, s_ch[slot], s_ml[slot], s_g0[slot], s_g1[slot]);
  s_rgb[slot] = make_float4(e.r, e.g, e.b, __uint_as_float(e.val));
  if (need_normal) {
    const float g[8] = {v.g0.x, v.g0.y, v.g0.z, v.g0.w, v.g1.x, v.g1.y, v.g1.z, v.g1.w};
    float nx, ny, nz, rn;
    lin_normal(g, nx, ny, nz, rn);
    s_nrm[slot] = make_float4(nx * rn, ny * rn, nz * rn, rn);
  }
}

__device__ inline void wave_lds_sync_b() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// dL/ds of one tile: the workgroup's compensated partials (per lane, then per wave, then the
// four waves in order) written to part[g] as a (hi, lo) pair, and the lane accumulators reset.
// Every tile slot of the launch is written once, by the workgroup that rendered the tile, so
// the sum over tiles (ds_finalize_kernel) does not depend on which workgroup took which tile.
__device__ inline void flush_tile_ds(float& hi, float& lo, float2* s_red, float2* part, int g, int t, int lane,
                                     int w) {
  const float2 ws = wave_comp_sum(hi, lo);
  if (lane == 0) s_red[w] = ws;
  __syncthreads();
  if (t == 0) {
    float h = 0.f, l = 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      comp_add(h, l, s_red[k].x);
      l += s_red[k].y;
    }
    part[g] = make_float2(h, l);
  }
  hi = 0.f;
  lo = 0.f;
}

// Sum of the per-tile (hi, lo) partials in a fixed order, added to *dL_ds.  One workgroup.
__global__ __launch_bounds__(256) void ds_finalize_kernel(const float2* __restrict__ part, int n,
                                                          float* __restrict__ dL_ds) {
  __shared__ float2 s_w[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  float hi = 0.f, lo = 0.f;
  for (int i = t; i < n; i += 256) {
    const float2 p = part[i];
    comp_add(hi, lo, p.x);
    lo += p.y;
  }
  const float2 ws = wave_comp_sum(hi, lo);
  if (lane == 0) s_w[w] = ws;
  __syncthreads();
  if (t == 0) {
    float h = 0.f, l = 0.f;
    comp_add(h, l, *dL_ds);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      comp_add(h, l, s_w[k].x);
      l += s_w[k].y;
    }
    *dL_ds = h + l;
  }
}

template <bool kDepth, bool kNormal, bool kConcen>
__global__ __launch_bounds__(256) void render_bwd_kernel(BwdArgs a) {
  __shared__ float4 s_ch[256], s_ml[256], s_g0[256], s_g1[256], s_rgb[256];
  __shared__ float4 s_nrm[kNormal ? 256 : 1];
  __shared__ uint2 s_bb[256];
  __shared__ float2 s_red[4];
  __shared__ int s_tile;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const float s = *a.s_val;
  float ds_hi = 0.f, ds_lo = 0.f;

  for (int iter = 0;; ++iter) {
    int g;
    if (a.work_counter) {
      if (t == 0) s_tile = (int)atomicAdd(a.work_counter, 1u);
      __syncthreads();
      g = __builtin_amdgcn_readfirstlane(s_tile);   // uniform: tile and view data in scalar registers
      __syncthreads();
      if (g >= a.total_tiles) break;
    } else {
      if (iter > 0) break;
      g = blockIdx.x;
    }
    const int vi = find_view_b(a.views, a.n_views, g);
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
    const uint32_t start = a.ranges[g].x;
    const uint32_t end = a.tile_last[g];
    const int n = (int)(end - start);

    const long long HW = (long long)W * H;
    const long long pid = (long long)py * W + px;
    const long long pb = V.pix_base;
    float T = 0.f, dC0 = 0.f, dC1 = 0.f, dC2 = 0.f, last_dT = 0.f, dD = 0.f;
    float dN0 = 0.f, dN1 = 0.f, dN2 = 0.f, gt0 = 0.f, gt1 = 0.f, gt2 = 0.f;
    uint32_t last = 0;
    if (inside) {
      T = a.T[pb + pid];
      last = a.n_contrib[pb + pid];
      dC0 = a.dL_dcolor[pb * 3 + 0 * HW + pid];
      dC1 = a.dL_dcolor[pb * 3 + 1 * HW + pid];
      dC2 = a.dL_dcolor[pb * 3 + 2 * HW + pid];
      last_dT = (a.dL_dT ? a.dL_dT[pb + pid] : 0.f) + (V.bg[0] * dC0 + V.bg[1] * dC1 + V.bg[2] * dC2);
      if (kDepth) dD = a.dL_ddepth[pb * 4 + pid] * r.norm_inv;
      if (kNormal) {
        dN0 = a.dL_dnormal[pb * 3 + 0 * HW + pid];
        dN1 = a.dL_dnormal[pb * 3 + 1 * HW + pid];
        dN2 = a.dL_dnormal[pb * 3 + 2 * HW + pid];
      }
      if (kConcen) {
        gt0 = a.gt_color[pb * 3 + 0 * HW + pid];
        gt1 = a.gt_color[pb * 3 + 1 * HW + pid];
        gt2 = a.gt_color[pb * 3 + 2 * HW + pid];
      }
    }
    const float wR = kConcen ? a.lambda_R_concen / (float)(W * H) : 0.f;

    for (int base = 0; base < n; base += 256) {
      __syncthreads();
      if (base + t < n) {
        const uint32_t e = end - 1 - (uint32_t)(base + t);
        const uint32_t val = a.vals[e];
        const uint32_t vox = val & kVoxMask;
        s_bb[t] = a.bboxes[(long long)vi * a.P + vox];
        VoxelGeom vg;
        load_voxel_geom(vg, a.vrec, a.centers, a.lengths, a.geos, vox);
        const float4 g0 = vg.g0, g1 = vg.g1;
        stage_segment(vg.cx, vg.cy, vg.cz, vg.l, vg.l_inv, g0, g1, r, s_ch[t], s_ml[t], s_g0[t], s_g1[t]);
        const float* rgbv = a.rgbs + (long long)vi * a.rgb_view_stride;
        s_rgb[t] = make_float4(rgbv[3 * vox], rgbv[3 * vox + 1], rgbv[3 * vox + 2], __uint_as_float(val));
        if (kNormal) {
          const float gg[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
          float nx, ny, nz, rn;
          lin_normal(gg, nx, ny, nz, rn);
          s_nrm[t] = make_float4(nx * rn, ny * rn, nz * rn, rn);
        }
      }
      __syncthreads();
      const int cnt = min(256, n - base);
      for (int j = 0; j < cnt; ++j) {
        const uint32_t rel = end - 1 - (uint32_t)(base + j) - start;
        bool act = inside && rel < last;
        float gv[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) gv[k] = 0.f;
        const float4 rgb = s_rgb[j];
        const uint32_t val = __float_as_uint(rgb.w);
        if (act) {
          const uint2 bb = s_bb[j];
          act = pixel_in_bbox(px, py, bb.x, bb.y) && (val >> 29) == r.octant;
        }
        float ta = 0.f, tb = 0.f;
        if (act) {
          ray_box_st(r, s_ch[j], s_ml[j], ta, tb);
          act = !(ta > tb);
        }
        if (act) {
          Segment seg;
          eval_segment_st(r, ta, tb, s_ch[j], s_g0[j], s_g1[j], s, seg);
          const float alpha = seg.alpha;
          act = !(alpha < kMinAlpha);
          if (act) {
            T = T / (1.f - alpha);
            const float ptw = alpha * T;
            float dptw = dC0 * rgb.x + dC1 * rgb.y + dC2 * rgb.z;
            float4 nn = make_float4(0.f, 0.f, 0.f, 0.f);
            if (kNormal) {
              nn = s_nrm[j];
              dptw += dN0 * nn.x + dN1 * nn.y + dN2 * nn.z;
            }
            const float dLda = T * (dptw - last_dT);
            last_dT += alpha * (dptw - last_dT);

            // Main alpha: a = (phi0 - phi3) / (phi0 + eps).
            const float p0 = seg.phi[0], p3 = seg.phi[3];
            const float rden = 1.f / (p0 + kPhiEps);
            const float q0 = p0 * (1.f - p0), q3 = p3 * (1.f - p3);
            float c[4];
            c[0] = dLda * s * q0 * (1.f - alpha) * rden;
            c[1] = 0.f;
            c[2] = 0.f;
            c[3] = -dLda * s * q3 * rden;
            comp_add(ds_hi, ds_lo, dLda * (q0 * seg.sdf[0] * (1.f - alpha) - q3 * seg.sdf[3]) * rden);

            if (kDepth) {
              // Depth through the three sub-interval alphas (backward.cu:338-388, 682-768).
              float sub[3];
              sub_alphas(seg, sub);
              const float step = kThird * (tb - ta);
              const float t0 = ta + 0.5f * step, t1 = ta + 1.5f * step, t2 = ta + 2.5f * step;
              const float a0 = sub[0], a1 = sub[1], a2 = sub[2];
              const float k = dD * T;
              const float gk[3] = {k * (t0 + a1 * a2 * t2 - a1 * t1 - a2 * t2),
                                   k * (t1 + a0 * a2 * t2 - a0 * t1 - a2 * t2),
                                   k * (t2 + a0 * a1 * t2 - a0 * t2 - a1 * t2)};
#pragma unroll
              for (int m = 0; m < 3; ++m) {
                const float pk = seg.phi[m], pk1 = seg.phi[m + 1];
                const float dk = 1.f / (pk + kPhiEps);
                const float raw = (pk - pk1) * dk;
                if (raw > 0.f) {
                  c[m] += gk[m] * pk1 * dk * dk * s * pk * (1.f - pk);
                  c[m + 1] += -gk[m] * dk * s * pk1 * (1.f - pk1);
                }
              }
              last_dT += dD * (a0 * t0 + (1.f - a0) * a1 * t1 + (1.f - a0) * (1.f - a1) * a2 * t2);
            }

            // Scatter the per-sample SDF sensitivities to the 8 corners.
#pragma unroll
            for (int m = 0; m < 4; ++m) {
              const float fm = (float)m;
              float wts[8];
              trilinear_weights(seg.qx + seg.sx * fm, seg.qy + seg.sy * fm, seg.qz + seg.sz * fm, wts);
#pragma unroll
              for (int i = 0; i < 8; ++i) gv[i] += c[m] * wts[i];
            }

            if (kNormal) {
              // N += w * lin_n / |lin_n|: normalise Jacobian, then the +-1 corner pattern.
              const float sx = ptw * dN0, sy = ptw * dN1, sz = ptw * dN2;
              const float pr = sx * nn.x + sy * nn.y + sz * nn.z;
              const float lx = nn.w * (sx - pr * nn.x);
              const float ly = nn.w * (sy - pr * nn.y);
              const float lz = nn.w * (sz - pr * nn.z);
#pragma unroll
              for (int i = 0; i < 8; ++i)
                gv[i] += ((i & 4) ? lx : -lx) + ((i & 2) ? ly : -ly) + ((i & 1) ? lz : -lz);
            }

            gv[8] = ptw * dC0;
            gv[9] = ptw * dC1;
            gv[10] = ptw * dC2;
            if (kConcen) {
              gv[8] += wR * ptw * 2.f * (rgb.x - gt0);
              gv[9] += wR * ptw * 2.f * (rgb.y - gt1);
              gv[10] += wR * ptw * 2.f * (rgb.z - gt2);
            }
            gv[11] = fabsf(dLda * alpha);
          }
        }
        const uint64_t m = __ballot(act);
        if (m == 0) continue;
        // Lane k of `mine` receives the wave total of gradient slot k, then lanes 0..11 add 12
        // consecutive floats with one vector atomic.
        float mine = 0.f;
        if (__popcll(m) == 1) {
          const int src = __ffsll((long long)m) - 1;
#pragma unroll
          for (int k = 0; k < 12; ++k) {
            const float sk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(gv[k]), src));
            mine = (lane == k) ? sk : mine;
          }
        } else {
#pragma unroll
          for (int k = 0; k < 12; ++k) {
            const float sk = wave_sum_uniform(gv[k]);
            mine = (lane == k) ? sk : mine;
          }
        }
        if (lane < 12) {
          const uint32_t vox = val & kVoxMask;
          float* dst = a.dL_dvox + (long long)vox * 12 + lane;
          if (a.dL_drgb_views && lane >= 8 && lane < 11)
            dst = a.dL_drgb_views + ((long long)vi * a.P + vox) * 3 + (lane - 8);
          if (a.dL_dgrid && lane < 8) dst = a.dL_dgrid + a.vox_key[(long long)vox * 8 + lane];
          atomicAdd(dst, mine);
        }
      }
    }
    flush_tile_ds(ds_hi, ds_lo, s_red, a.ds_part, g, t, lane, w);
  }
}

// ------------------------------------------------------------------------------------------
// Pair backward (default).  The forward hands over, per tile, its contributing entries (those
// that composited into at least one pixel) and, per entry and 8x8 block, the pixels that
// composited it.  Those (pixel, entry) pairs are exactly the ones the reverse walk visits: alpha
// at least the minimum, list position before the pixel's last contributor.  Per staged batch
// (entries in back-to-front order), each wave:
//   1. reads its block's pixel masks of the batch's entries and prefix-sums their counts: entry j owns
//      the run [pref_j, pref_j + |m_j|) of an entry-major pair list, cut into windows of at most
//      128 positions and 64 entries; each pixel's entry set in a window is ORed together in LDS
//      while step 2 visits the pairs;
//   2. evaluates alpha (and the depth value) of every pair, 64 per instruction;
//   3. walks each pixel's set in back-to-front order: T <- T / (1 - a), dL/dalpha and the running
//      dL/dT -- the only sequential part, a few FMAs per pair;
//   4. computes the SDF-segment derivatives of every pair, 64 per instruction, summed over each
//      entry's run (segmented DPP sum) into per-entry LDS accumulators shared by the four waves;
//      after the batch, one global atomic per (entry, gradient slot).
// The per-voxel global atomics fall from 12 per (pixel, entry) in the reference to 12 per
// (tile, entry).
// ------------------------------------------------------------------------------------------
constexpr int kBwdWin = 512;   // pair positions per window (the four waves together)

template <bool kDepth, bool kNormal, bool kConcen, bool kQueue, int kHalves>
__global__ __launch_bounds__(256, kDepth || kNormal || kConcen ? 3 : (kHalves == 1 ? 5 : 4)) void render_bwd_pairs_kernel(BwdArgs a) {
  constexpr int kStage = 64 * kHalves;   // entries staged per round (two 64-entry halves or one)
  __shared__ float4 s_ch[kStage], s_ml[kStage], s_g0[kStage], s_g1[kStage], s_rgb[kStage];
  __shared__ float4 s_nrm[kNormal ? kStage : 1];
  __shared__ float s_acc[kStage][12];
  __shared__ uint32_t s_used[kStage];
  __shared__ float4 s_rd[4][64];           // ray direction xyz, dL/ddepth
  __shared__ float4 s_ri[4][64];           // 1/direction xyz, dL/dcolor x
  __shared__ float2 s_dc[4][64];           // dL/dcolor y, z
  __shared__ unsigned long long s_cs[4][64];   // per pixel: its entries (relative to the window base) in the window
  __shared__ float4 s_dn[4][kNormal ? 64 : 1];
  __shared__ float4 s_gt[4][kConcen ? 64 : 1];
  __shared__ uint64_t s_m[4][kStage];      // per block and entry: the block's pixels it contributed to
  __shared__ uint16_t s_pref[kStage + 1];  // per entry: first position of its run; [kStage] = total
  __shared__ uint32_t s_qo[kStage];        // per entry: offsets of blocks 1, 2, 3 in its run (8 bits each)
  __shared__ uint16_t s_hd[kBwdWin];       // entry | block << 7 | pixel << 9 per window position
  __shared__ float s_pa[kBwdWin];          // alpha per position; after the walk, T before the entry
  __shared__ float s_pd[kDepth ? kBwdWin : 1];
  __shared__ float s_pl[kBwdWin];          // dL/dalpha per position
  __shared__ uint64_t s_rsb[kBwdWin / 64];  // run starts in the window, one bit per position
  __shared__ float2 s_red[4];
  __shared__ int s_tile[2];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const float s = *a.s_val;
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  float ds_hi = 0.f, ds_lo = 0.f;
  if (t < kBwdWin / 64) s_rsb[t] = 0ull;   // wave 0 alone writes it from here on (marks, clears)

  for (int iter = 0;; ++iter) {
    int g;
    if (kQueue) {
      // Two slots by iteration parity: a slot is rewritten two tiles later, after every wave has
      // passed the barriers of the tile in between, so one barrier per tile suffices.
      if (t == 0) s_tile[iter & 1] = (int)atomicAdd(a.work_counter, 1u);
      __syncthreads();
      g = __builtin_amdgcn_readfirstlane(s_tile[iter & 1]);   // uniform: tile and view data in scalar registers
      if (g >= a.total_tiles) break;
    } else {
      if (iter > 0) break;
      g = blockIdx.x;
    }
    const int vi = kQueue ? (int)a.tile_view[g] : find_view_b(a.views, a.n_views, g);
    const ViewDesc& V = a.views[vi];
    const int W = V.W, H = V.H;
    const int lt = g - V.tile_base;
    const int tx = lt % V.tiles_x, ty = lt / V.tiles_x;
    int qx, qy;
    tile_quadrant(V, tx, ty, w, qx, qy);
    const int px = qx + (lane & 7), py = qy + (lane >> 3);
    const bool inside = px < W && py < H;
    const Ray r = pixel_ray(V, px, py);
    // The tile's contributing entries: cvals / pmask [cr.x, cr.y), walked back to front.
    const uint2 cr = a.cranges[g];
    const int n = (int)(cr.y - cr.x);
    const long long top = (long long)cr.y - 1;   // entry of batch slot 0 in the first batch

    const long long HW = (long long)W * H;
    const long long pid = (long long)py * W + px;
    const long long pb = V.pix_base;
    float T = 0.f, dC0 = 0.f, dC1 = 0.f, dC2 = 0.f, last_dT = 0.f, dD = 0.f;
    float dN0 = 0.f, dN1 = 0.f, dN2 = 0.f, gt0 = 0.f, gt1 = 0.f, gt2 = 0.f;
    if (inside) {
      T = a.T[pb + pid];
      dC0 = a.dL_dcolor[pb * 3 + 0 * HW + pid];
      dC1 = a.dL_dcolor[pb * 3 + 1 * HW + pid];
      dC2 = a.dL_dcolor[pb * 3 + 2 * HW + pid];
      last_dT = (a.dL_dT ? a.dL_dT[pb + pid] : 0.f) + (V.bg[0] * dC0 + V.bg[1] * dC1 + V.bg[2] * dC2);
      if (kDepth) dD = a.dL_ddepth[pb * 4 + pid] * r.norm_inv;
      if (kNormal) {
        dN0 = a.dL_dnormal[pb * 3 + 0 * HW + pid];
        dN1 = a.dL_dnormal[pb * 3 + 1 * HW + pid];
        dN2 = a.dL_dnormal[pb * 3 + 2 * HW + pid];
      }
      if (kConcen) {
        gt0 = a.gt_color[pb * 3 + 0 * HW + pid];
        gt1 = a.gt_color[pb * 3 + 1 * HW + pid];
        gt2 = a.gt_color[pb * 3 + 2 * HW + pid];
      }
    }
    const float wR = kConcen ? a.lambda_R_concen / (float)(W * H) : 0.f;
    s_rd[w][lane] = make_float4(r.dx, r.dy, r.dz, dD);
    s_ri[w][lane] = make_float4(r.ix, r.iy, r.iz, dC0);
    s_dc[w][lane] = make_float2(dC1, dC2);
    if (kNormal) s_dn[w][lane] = make_float4(dN0, dN1, dN2, 0.f);
    if (kConcen) s_gt[w][lane] = make_float4(gt0, gt1, gt2, 0.f);

    // Each wave's block masks of the next batch are loaded a batch ahead (lane: slots lane and
    // lane + 64).
    const float* rgbs_v = a.rgbs + (long long)vi * a.rgb_view_stride;
    const uint64_t rgb_row0 = (uint64_t)vi * (uint64_t)a.P;   // this view's first row of dL_drgb_views
    // Staging (threads t < kStage, one entry each).  With 128-entry batches (four workgroups per
    // CU) the entry data is loaded a batch ahead into registers and its value two batches ahead;
    // with 64-entry batches the registers go to occupancy (five workgroups per CU, no spills): the
    // value is loaded a batch ahead and the entry data when the batch is staged.
    constexpr bool kPrefetch = kHalves > 1;
    EntryRegsB er;
    uint32_t nval = 0;
    if (t < kStage) {
      if (kPrefetch) {
        if (t < n) fetch_entry_b(er, a, a.cvals[top - t], rgbs_v);
        if (kStage + t < n) nval = a.cvals[top - kStage - t];
      } else if (t < n) {
        nval = a.cvals[top - t];
      }
    }
    uint64_t nm_lo = lane < n ? a.pmask[4 * (top - lane) + w] : 0ull;
    uint64_t nm_hi = (kHalves > 1 && 64 + lane < n) ? a.pmask[4 * (top - 64 - lane) + w] : 0ull;

    for (int base = 0; base < n; base += kStage) {
      __syncthreads();
      const int cnt = min(kStage, n - base);
      if (t < cnt) {
        if (!kPrefetch) fetch_entry_b(er, a, nval, rgbs_v);
        store_entry_b(er, kNormal, r, s_ch, s_ml, s_g0, s_g1, s_rgb, s_nrm, t);
        s_used[t] = 0u;
      }
      for (int k = t; k < kStage * 12; k += 256) (&s_acc[0][0])[k] = 0.f;
      // This wave's block masks of the batch (loaded during the previous batch).
      s_m[w][lane] = nm_lo;
      if (kHalves > 1) s_m[w][lane + 64] = nm_hi;
      __syncthreads();
      if (t < kStage) {
        if (kPrefetch) {
          if (base + kStage + t < n) fetch_entry_b(er, a, nval, rgbs_v);
          if (base + 2 * kStage + t < n) nval = a.cvals[top - (base + 2 * kStage + t)];
        } else if (base + kStage + t < n) {
          nval = a.cvals[top - (base + kStage + t)];
        }
      }
      {
        const int nb = base + kStage;
        nm_lo = nb + lane < n ? a.pmask[4 * (top - nb - lane) + w] : 0ull;
        nm_hi = (kHalves > 1 && nb + 64 + lane < n) ? a.pmask[4 * (top - nb - 64 - lane) + w] : 0ull;
      }
      // 1. The batch's pair list, shared by the four waves: entry e owns the positions
      // [pref_e, pref_e + c_e), its block-0 pairs first, then blocks 1, 2, 3, pixels ascending.
      if (w == 0) {
        uint32_t c_lo, qo_lo, c_hi = 0u, qo_hi = 0u;
        {
          const uint32_t a0 = (uint32_t)__popcll(s_m[0][lane]), a1 = (uint32_t)__popcll(s_m[1][lane]);
          const uint32_t a2 = (uint32_t)__popcll(s_m[2][lane]), a3 = (uint32_t)__popcll(s_m[3][lane]);
          qo_lo = a0 | ((a0 + a1) << 8) | ((a0 + a1 + a2) << 16);
          c_lo = a0 + a1 + a2 + a3;
        }
        if (kHalves > 1) {
          const uint32_t a0 = (uint32_t)__popcll(s_m[0][lane + 64]), a1 = (uint32_t)__popcll(s_m[1][lane + 64]);
          const uint32_t a2 = (uint32_t)__popcll(s_m[2][lane + 64]), a3 = (uint32_t)__popcll(s_m[3][lane + 64]);
          qo_hi = a0 | ((a0 + a1) << 8) | ((a0 + a1 + a2) << 16);
          c_hi = a0 + a1 + a2 + a3;
        }
        const uint32_t i_lo = wave_incl_scan_u32(c_lo);
        const uint32_t i_hi = kHalves > 1 ? wave_incl_scan_u32(c_hi) : 0u;
        const uint32_t t_lo = (uint32_t)__builtin_amdgcn_readlane((int)i_lo, 63);
        s_pref[lane] = (uint16_t)(i_lo - c_lo);
        s_qo[lane] = qo_lo;
        if (kHalves > 1) {
          s_pref[lane + 64] = (uint16_t)(t_lo + i_hi - c_hi);
          s_qo[lane + 64] = qo_hi;
        }
        if (lane == 63) s_pref[kStage] = (uint16_t)(t_lo + i_hi);
      }
      __syncthreads();
      const uint32_t total = s_pref[kStage];
      // A window holds at most kBwdWin positions of at most 64 consecutive entries [ja, ja + 64),
      // so each pixel's entries in it form one 64-bit set (ORed together while the alphas are
      // evaluated).  With 128 entries per batch, ja is 0 or 64.  The window bounds come from
      // shared data, so every wave takes the same windows and meets the same barriers.
      int ja = 0;
      const int qsh = 8 * (w - 1);   // this block's offset field in s_qo (w > 0)
      for (uint32_t w0 = 0, w1 = 0; w0 < total; w0 = w1) {
        if (kHalves > 1 && ja == 0 && (uint32_t)s_pref[64] <= w0) ja = 64;
        const uint32_t lim = (kHalves > 1 && ja == 0) ? (uint32_t)s_pref[64] : total;
        w1 = min(min(total, lim), w0 + (uint32_t)kBwdWin);
        const uint32_t np = w1 - w0;
        s_cs[w][lane] = 0ull;
        // Run starts of the window's entries as a bitmap over its positions (wave 0; the bitmap
        // was cleared after the previous window's alpha pass).  Every entry has at least one
        // pair, so the starts are distinct.
        const uint32_t pl = (uint32_t)s_pref[ja + lane];   // run start of entry ja + lane
        if (w == 0 && pl > w0 && pl < w1) atomicOr(&s_rsb[(pl - w0) >> 6], 1ull << ((pl - w0) & 63u));
        __syncthreads();

        // 2. alpha (and depth value) per position, 64-position chunks dealt to the waves in turn:
        // the owning entry (the entries starting at or before the chunk's first position, plus
        // the run starts between it and the position), its block and pixel.
        for (uint32_t c0 = 64u * (uint32_t)w; c0 < np; c0 += 256u) {
          const uint32_t k = c0 + lane;
          const int j0 = ja - 1 + (int)__popcll(__ballot(pl <= w0 + c0));
          const uint64_t starts = s_rsb[c0 >> 6] & ~1ull;   // run starts after the chunk's first position
          if (k < np) {
            const uint32_t p = w0 + k;
            const int j = j0 + (int)__popcll(starts & (~0ull >> (63 - lane)));
            const uint32_t rr = p - (uint32_t)s_pref[j];
            const uint32_t qo = s_qo[j];
            const uint32_t o1 = qo & 255u, o2 = (qo >> 8) & 255u, o3 = qo >> 16;
            const int q = (int)(rr >= o1) + (int)(rr >= o2) + (int)(rr >= o3);
            const uint32_t off = q == 0 ? 0u : (q == 1 ? o1 : (q == 2 ? o2 : o3));
            const int bit = select_set_bit(s_m[q][j], (int)(rr - off));
            const Ray pr = pair_ray(r, s_rd[q][bit], s_ri[q][bit]);
            const float4 box = s_ch[j];
            float ta, tb;
            ray_box_st(pr, box, s_ml[j], ta, tb);
            Segment seg;
            eval_segment_st(pr, ta, tb, box, s_g0[j], s_g1[j], s, seg);
            s_pa[k] = seg.alpha;
            if (kDepth) {
              float sub[3];
              sub_alphas(seg, sub);
              s_pd[k] = depth_value(sub, ta, tb);
            }
            s_hd[k] = (uint16_t)((uint32_t)j | ((uint32_t)q << 7) | ((uint32_t)bit << 9));
            atomicOr(&s_cs[q][bit], 1ull << (j - ja));
          }
        }
        __syncthreads();

        // 3. sequential reverse walk per pixel (wave w: block w) over its entries in the window
        // (ascending slot = back to front along the list), software-pipelined: the next entry's
        // position is looked up while this one is walked.
        if (w == 0 && lane < kBwdWin / 64) s_rsb[lane] = 0ull;   // no wave reads it until the next window
        uint64_t cs = s_cs[w][lane];
        int jc = -1;
        uint32_t kc = 0;
        if (cs != 0ull) {
          jc = ja + __builtin_ctzll(cs);
          kc = (uint32_t)s_pref[jc] + (w == 0 ? 0u : ((s_qo[jc] >> qsh) & 255u)) +
               (uint32_t)__popcll(s_m[w][jc] & lt_mask);
        }
        while (jc >= 0) {
          const int j = jc;
          const uint32_t kk = kc - w0;
          cs &= cs - 1ull;
          const float alpha = s_pa[kk];
          const float4 rgb = s_rgb[j];
          jc = -1;
          if (cs != 0ull) {
            jc = ja + __builtin_ctzll(cs);
            kc = (uint32_t)s_pref[jc] + (w == 0 ? 0u : ((s_qo[jc] >> qsh) & 255u)) +
                 (uint32_t)__popcll(s_m[w][jc] & lt_mask);
          }
          T = div_rn(T, 1.f - alpha);   // 1 - alpha > 2^-21, T normal: the IEEE quotient
          float dptw = dC0 * rgb.x + dC1 * rgb.y + dC2 * rgb.z;
          if (kNormal) {
            const float4 nn = s_nrm[j];
            dptw += dN0 * nn.x + dN1 * nn.y + dN2 * nn.z;
          }
          const float dLda = T * (dptw - last_dT);
          last_dT += alpha * (dptw - last_dT);
          if (kDepth) last_dT += dD * s_pd[kk];
          s_pl[kk] = dLda;
          s_pa[kk] = T;
        }
        __syncthreads();

        // 4. per-pair derivatives, chunks dealt to the waves in turn; each entry's run (all its
        // blocks) is summed across lanes (segmented DPP sum) and its last lane in the chunk adds
        // the 12 sums to the workgroup's LDS accumulators.
        for (uint32_t c0 = 64u * (uint32_t)w; c0 < np; c0 += 256u) {
          const uint32_t k = c0 + lane;
          const bool valid = k < np;
          const uint32_t d = valid ? (uint32_t)s_hd[k] : 0u;
          const uint32_t jj = valid ? (d & 127u) : 0xFFFFFFFEu;
          uint32_t jp = (uint32_t)__shfl_up((int)jj, 1, 64);
          if (lane == 0) jp = 0xFFFFFFFFu;
          const uint32_t jn = (uint32_t)__shfl_down((int)jj, 1, 64);
          const bool head = jj != jp;
          const int seg_start = (int)wave_incl_max_u32(head ? (uint32_t)lane : 0u);
          const bool tail = valid && (lane == 63 || k + 1 == np || jn != jj);
          float gv[12];
#pragma unroll
          for (int q = 0; q < 12; ++q) gv[q] = 0.f;
          const int j = (int)(d & 127u);
          if (valid) {
            const int qb = (int)((d >> 7) & 3u), bit = (int)(d >> 9);
            const float Tp = s_pa[k];
            const float dLda = s_pl[k];
            const float4 rdd = s_rd[qb][bit];
            const float4 rid = s_ri[qb][bit];
            const Ray pr = pair_ray(r, rdd, rid);
            const float4 box = s_ch[j];
            float ta, tb;
            ray_box_st(pr, box, s_ml[j], ta, tb);
            Segment seg;
            eval_segment_st(pr, ta, tb, box, s_g0[j], s_g1[j], s, seg);
            const float alpha = seg.alpha;
            const float ptw = alpha * Tp;
            const float p0v = seg.phi[0], p3v = seg.phi[3];
            const float rden = rcp_rn(p0v + kPhiEps);   // p0 + eps in [1e-6, 2]: the IEEE quotient
            const float q0 = p0v * (1.f - p0v), q3 = p3v * (1.f - p3v);
            float c[4];
            c[0] = dLda * s * q0 * (1.f - alpha) * rden;
            c[1] = 0.f;
            c[2] = 0.f;
            c[3] = -dLda * s * q3 * rden;
            comp_add(ds_hi, ds_lo, dLda * (q0 * seg.sdf[0] * (1.f - alpha) - q3 * seg.sdf[3]) * rden);
            if (kDepth) {
              float sub[3];
              sub_alphas(seg, sub);
              const float step = kThird * (tb - ta);
              const float t0 = ta + 0.5f * step, t1 = ta + 1.5f * step, t2 = ta + 2.5f * step;
              const float a0 = sub[0], a1 = sub[1], a2 = sub[2];
              const float kd = rdd.w * Tp;
              const float gk[3] = {kd * (t0 + a1 * a2 * t2 - a1 * t1 - a2 * t2),
                                   kd * (t1 + a0 * a2 * t2 - a0 * t1 - a2 * t2),
                                   kd * (t2 + a0 * a1 * t2 - a0 * t2 - a1 * t2)};
#pragma unroll
              for (int mm = 0; mm < 3; ++mm) {
                const float pk = seg.phi[mm], pk1 = seg.phi[mm + 1];
                const float dk = rcp_rn(pk + kPhiEps);
                if ((pk - pk1) * dk > 0.f) {
                  c[mm] += gk[mm] * pk1 * dk * dk * s * pk * (1.f - pk);
                  c[mm + 1] += -gk[mm] * dk * s * pk1 * (1.f - pk1);
                }
              }
            }
#pragma unroll
            for (int mm = 0; mm < 4; ++mm) {
              if (!kDepth && (mm == 1 || mm == 2)) continue;   // c[1] = c[2] = 0 without depth
              const float fm = (float)mm;
              float wts[8];
              if (mm == 0) trilinear_weights(seg.qx, seg.qy, seg.qz, wts);   // the entry point itself
              else trilinear_weights(seg.qx + seg.sx * fm, seg.qy + seg.sy * fm, seg.qz + seg.sz * fm, wts);
#pragma unroll
              for (int i = 0; i < 8; ++i) gv[i] += c[mm] * wts[i];
            }
            const float2 dcyz = s_dc[qb][bit];
            const float4 dcv = make_float4(rid.w, dcyz.x, dcyz.y, 0.f);
            if (kNormal) {
              const float4 nn = s_nrm[j];
              const float4 dnv = s_dn[qb][bit];
              const float sx = ptw * dnv.x, sy = ptw * dnv.y, sz = ptw * dnv.z;
              const float pr2 = sx * nn.x + sy * nn.y + sz * nn.z;
              const float lx = nn.w * (sx - pr2 * nn.x);
              const float ly = nn.w * (sy - pr2 * nn.y);
              const float lz = nn.w * (sz - pr2 * nn.z);
#pragma unroll
              for (int i = 0; i < 8; ++i)
                gv[i] += ((i & 4) ? lx : -lx) + ((i & 2) ? ly : -ly) + ((i & 1) ? lz : -lz);
            }
            gv[8] = ptw * dcv.x;
            gv[9] = ptw * dcv.y;
            gv[10] = ptw * dcv.z;
            if (kConcen) {
              const float4 gtv = s_gt[qb][bit];
              const float4 rgb = s_rgb[j];
              gv[8] += wR * ptw * 2.f * (rgb.x - gtv.x);
              gv[9] += wR * ptw * 2.f * (rgb.y - gtv.y);
              gv[10] += wR * ptw * 2.f * (rgb.z - gtv.z);
            }
            gv[11] = fabsf(dLda * alpha);
          }
          wave_segmented_sum<12>(gv, lane, seg_start);
          if (tail) {
#pragma unroll
            for (int q = 0; q < 12; ++q) atomicAdd(&s_acc[j][q], gv[q]);
            s_used[j] = 1u;
          }
        }
      }
      __syncthreads();
      // Flush the batch's per-entry sums: one global atomic per (entry, slot) that received any.
      // Thread t takes the sums t, t + 256, ... of the cnt x 12; as 256 = 21 * 12 + 4, each step
      // advances 21 entries and 4 slots (no division per step).
      {
        int j = t / 12, slot = t - 12 * j;
        for (int k = t; k < cnt * 12; k += 256) {
          if (s_used[j]) {
            const float v = s_acc[j][slot];
            const uint64_t vox = __float_as_uint(s_rgb[j].w) & kVoxMask;
            float* dst = a.dL_dvox + ((vox << 3) + (vox << 2)) + slot;
            if (a.dL_drgb_views && slot >= 8 && slot < 11) {
              const uint64_t row = rgb_row0 + vox;
              dst = a.dL_drgb_views + ((row << 1) + row) + (slot - 8);
            }
            if (a.dL_dgrid && slot < 8) dst = a.dL_dgrid + a.vox_key[(vox << 3) + slot];
            atomicAdd(dst, v);
          }
          j += 21;
          slot += 4;
          if (slot >= 12) { slot -= 12; ++j; }
        }
      }
    }
    flush_tile_ds(ds_hi, ds_lo, s_red, a.ds_part, g, t, lane, w);
  }
}

template <bool D, bool N>
static hipError_t launch_bwd_c(const BwdArgs& a, hipStream_t st, unsigned grid) {
  const bool concen = a.lambda_R_concen > 0.f && a.gt_color;
  if (a.variant == 1) {
    if (concen) render_bwd_kernel<D, N, true><<<grid, 256, 0, st>>>(a);
    else render_bwd_kernel<D, N, false><<<grid, 256, 0, st>>>(a);
  } else {
    // The work-queue (multi-view) launch is its own instantiation, so profiles tell it apart.
    if (a.work_counter) {
      if (concen) render_bwd_pairs_kernel<D, N, true, true, 2><<<grid, 256, 0, st>>>(a);
      else if (a.halves == 1) render_bwd_pairs_kernel<D, N, false, true, 1><<<grid, 256, 0, st>>>(a);
      else render_bwd_pairs_kernel<D, N, false, true, 2><<<grid, 256, 0, st>>>(a);
    } else {
      if (concen) render_bwd_pairs_kernel<D, N, true, false, 2><<<grid, 256, 0, st>>>(a);
      else if (a.halves == 1) render_bwd_pairs_kernel<D, N, false, false, 1><<<grid, 256, 0, st>>>(a);
      else render_bwd_pairs_kernel<D, N, false, false, 2><<<grid, 256, 0, st>>>(a);
    }
  }
  return hipGetLastError();
}

hipError_t launch_render_backward(const BwdArgs& a, hipStream_t st) {
  if (a.total_tiles <= 0) return hipSuccess;

  unsigned grid = (unsigned)a.total_tiles;
  if (a.work_counter) grid = (unsigned)min(a.total_tiles, 256 * max(1, a.blocks_per_cu));
  hipError_t e;
  if (a.need_depth && a.need_normal) e = launch_bwd_c<true, true>(a, st, grid);
  else if (a.need_depth) e = launch_bwd_c<true, false>(a, st, grid);
  else if (a.need_normal) e = launch_bwd_c<false, true>(a, st, grid);
  else e = launch_bwd_c<false, false>(a, st, grid);
  if (e != hipSuccess) return e;
  ds_finalize_kernel<<<1, 256, 0, st>>>(a.ds_part, a.total_tiles, a.dL_ds);
  return hipGetLastError();
}

}  // namespace svr

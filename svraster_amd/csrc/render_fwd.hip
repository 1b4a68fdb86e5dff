// Forward SDF compositor for gfx950.
//
// Reference: FORWARD::renderCUDA<..., n_samp = 3, SDF> (cuda/src/forward.cu:42-564).
//
// One 256-thread workgroup per 16x16 tile; wave w covers the 8x8 quadrant (w&1, w>>1) and
// lane l the pixel (l&7, l>>3) inside it.  The tile's sorted entries are staged 256 at a time
// into LDS with the per-entry constants the pixel loop needs (box centre / half size, box
// minimum and 1/size, 8 corner SDFs, colour, packed bbox, normalised linear normal), so the
// per-pixel loop reads only LDS broadcasts.  The loop over staged entries is wave-uniform: a
// lane whose pixel misses the entry is simply masked, and a wave whose 64 pixels are all
// saturated leaves the batch early.  Optionally the workgroups pull tiles from a global
// (view, tile) queue so one launch can serve many views.
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

__device__ inline void stage_entry(const uint32_t* __restrict__ vals, uint32_t e, int view, long long P,
                                   const uint2* __restrict__ bboxes, const float* __restrict__ centers,
                                   const float* __restrict__ lengths, const float* __restrict__ geos,
                                   const float* __restrict__ rgbs, bool need_normal,
                                   float4* s_ch, float4* s_ml, float4* s_g0, float4* s_g1,
                                   float4* s_rgb, float4* s_nrm, uint2* s_bb, int slot) {
  const uint32_t val = vals[e];
  const uint32_t vox = val & kVoxMask;
  s_bb[slot] = bboxes[(long long)view * P + vox];
  const float cx = centers[3 * vox + 0], cy = centers[3 * vox + 1], cz = centers[3 * vox + 2];
  const float l = lengths[vox];
  s_ch[slot] = make_float4(cx, cy, cz, 0.5f * l);
  s_ml[slot] = make_float4(cx - 0.5f * l, cy - 0.5f * l, cz - 0.5f * l, 1.f / l);
  const float4 g0 = reinterpret_cast<const float4*>(geos)[2 * vox];
  const float4 g1 = reinterpret_cast<const float4*>(geos)[2 * vox + 1];
  s_g0[slot] = g0;
  s_g1[slot] = g1;
  s_rgb[slot] = make_float4(rgbs[3 * vox + 0], rgbs[3 * vox + 1], rgbs[3 * vox + 2], __uint_as_float(val));
  if (need_normal) {
    const float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
    float nx, ny, nz, rn;
    lin_normal(g, nx, ny, nz, rn);
    s_nrm[slot] = make_float4(nx * rn, ny * rn, nz * rn, rn);
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

  for (int iter = 0;; ++iter) {
    int g;
    if (a.work_counter) {
      if (t == 0) s_tile = (int)atomicAdd(a.work_counter, 1u);
      __syncthreads();
      g = s_tile;
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
    const int px = tx * kTile + (w & 1) * 8 + (lane & 7);
    const int py = ty * kTile + (w >> 1) * 8 + (lane >> 3);
    const bool inside = px < W && py < H;
    const Ray r = pixel_ray(V, px, py);
    const uint2 rg = a.ranges[g];
    const int n = (int)(rg.y - rg.x);

    float T = 1.f, C0 = 0.f, C1 = 0.f, C2 = 0.f, D = 0.f, N0 = 0.f, N1 = 0.f, N2 = 0.f;
    uint32_t last = 0;
    int dmed_vox = -1;
    float dmed_T = 0.f;
    bool done = !inside;

    for (int base = 0; base < n; base += 256) {
      if (__syncthreads_count(done) == 256) break;
      if (base + t < n)
        stage_entry(a.vals, rg.x + base + t, vi, a.P, a.bboxes, a.centers, a.lengths, a.geos,
                    a.rgbs + (long long)vi * a.rgb_view_stride, kNormal, s_ch, s_ml, s_g0, s_g1, s_rgb,
                    s_nrm, s_bb, t);
      __syncthreads();
      const int cnt = min(256, n - base);
      for (int j = 0; j < cnt; ++j) {
        if (__all(done)) break;
        if (done) continue;
        const uint2 bb = s_bb[j];
        const float4 rgb = s_rgb[j];
        const uint32_t val = __float_as_uint(rgb.w);
        if (!pixel_in_bbox(px, py, bb.x, bb.y) || (val >> 29) != r.octant) continue;
        const float4 ch = s_ch[j];
        float ta, tb;
        ray_box(r, ch.x, ch.y, ch.z, ch.w, ta, tb);
        if (ta > tb) continue;
        const float4 ml = s_ml[j];
        const float4 g0 = s_g0[j], g1 = s_g1[j];
        const float gg[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
        Segment seg;
        eval_segment(r, ta, tb, ml.x, ml.y, ml.z, ml.w, gg, s, seg);
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
      }
    }

    float Dmed = 0.f;
    if (kDepth && inside && dmed_vox >= 0) {
      // Median depth refined with 16 sub-steps inside the median voxel (forward.cu:450-521).
      const int v = dmed_vox;
      const float cx = a.centers[3 * v], cy = a.centers[3 * v + 1], cz = a.centers[3 * v + 2];
      const float l = a.lengths[v];
      float gg[8];
      for (int k = 0; k < 8; ++k) gg[k] = a.geos[8 * v + k];
      float ta, tb;
      ray_box(r, cx, cy, cz, 0.5f * l, ta, tb);
      const float l_inv = 1.f / l;
      const float step = (tb - ta) * (1.f / 16.f);
      Dmed = ta - 0.5f * step;
      float qx = (r.ox + ta * r.dx - (cx - 0.5f * l)) * l_inv;
      float qy = (r.oy + ta * r.dy - (cy - 0.5f * l)) * l_inv;
      float qz = (r.oz + ta * r.dz - (cz - 0.5f * l)) * l_inv;
      const float sx = step * r.dx * l_inv, sy = step * r.dy * l_inv, sz = step * r.dz * l_inv;
      float phi_prev = fast_sigmoid(s * trilerp(gg, qx, qy, qz));
      float Tm = dmed_T;
      for (int k = 0; k < 16 && Tm > 0.5f; ++k) {
        Dmed += step;
        qx += sx; qy += sy; qz += sz;
        const float phi = fast_sigmoid(s * trilerp(gg, qx, qy, qz));
        Tm *= 1.f - (phi_prev - phi) / (phi_prev + kPhiEps);
        phi_prev = phi;
      }
    }

    if (inside) {
      const long long HW = (long long)W * H;
      const long long pid = (long long)py * W + px;
      const long long pb = V.pix_base;
      a.color[pb * 3 + 0 * HW + pid] = C0 + T * V.bg[0];
      a.color[pb * 3 + 1 * HW + pid] = C1 + T * V.bg[1];
      a.color[pb * 3 + 2 * HW + pid] = C2 + T * V.bg[2];
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
    // tile_last = max over the tile's pixels of (range start + contributors) (forward.cu:561).
    const uint32_t tl = wave_max_u32(rg.x + (inside ? last : 0u));
    if (lane == 0) s_red[w] = tl;
    __syncthreads();
    if (t == 0) a.tile_last[g] = max(max(s_red[0], s_red[1]), max(s_red[2], s_red[3]));
    __syncthreads();
  }
}

template <bool D, bool N>
static hipError_t launch_fwd_mw(const FwdArgs& a, hipStream_t st, unsigned grid) {
  if (a.max_w) render_fwd_kernel<D, N, true><<<grid, 256, 0, st>>>(a);
  else render_fwd_kernel<D, N, false><<<grid, 256, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t launch_render_forward(const FwdArgs& a, hipStream_t st) {
  if (a.total_tiles <= 0) return hipSuccess;
  unsigned grid = (unsigned)a.total_tiles;
  if (a.work_counter) grid = (unsigned)min(a.total_tiles, 256 * 8);
  if (a.need_depth && a.need_normal) return launch_fwd_mw<true, true>(a, st, grid);
  if (a.need_depth) return launch_fwd_mw<true, false>(a, st, grid);
  if (a.need_normal) return launch_fwd_mw<false, true>(a, st, grid);
  return launch_fwd_mw<false, false>(a, st, grid);
}

}  // namespace svr

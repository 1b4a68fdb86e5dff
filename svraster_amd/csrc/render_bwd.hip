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
// lane), then lanes 0..11 issue one vector atomic add of 12 consecutive floats.  dL/ds is kept
// per lane and reduced once per workgroup.
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

template <bool kDepth, bool kNormal, bool kConcen>
__global__ __launch_bounds__(256) void render_bwd_kernel(BwdArgs a) {
  __shared__ float4 s_ch[256], s_ml[256], s_g0[256], s_g1[256], s_rgb[256];
  __shared__ float4 s_nrm[kNormal ? 256 : 1];
  __shared__ uint2 s_bb[256];
  __shared__ float s_red[4];
  __shared__ int s_tile;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const float s = *a.s_val;
  float ds_acc = 0.f;

  for (int iter = 0;; ++iter) {
    int g;
    if (a.work_counter) {
      if (t == 0) s_tile = (int)atomicAdd(a.work_counter, 1u);
      __syncthreads();
      g = s_tile;
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
    const int px = tx * kTile + (w & 1) * 8 + (lane & 7);
    const int py = ty * kTile + (w >> 1) * 8 + (lane >> 3);
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
      last_dT = a.dL_dT[pb + pid] + (V.bg[0] * dC0 + V.bg[1] * dC1 + V.bg[2] * dC2);
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
        const float cx = a.centers[3 * vox], cy = a.centers[3 * vox + 1], cz = a.centers[3 * vox + 2];
        const float l = a.lengths[vox];
        s_ch[t] = make_float4(cx, cy, cz, 0.5f * l);
        s_ml[t] = make_float4(cx - 0.5f * l, cy - 0.5f * l, cz - 0.5f * l, 1.f / l);
        const float4 g0 = reinterpret_cast<const float4*>(a.geos)[2 * vox];
        const float4 g1 = reinterpret_cast<const float4*>(a.geos)[2 * vox + 1];
        s_g0[t] = g0;
        s_g1[t] = g1;
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
          const float4 ch = s_ch[j];
          ray_box(r, ch.x, ch.y, ch.z, ch.w, ta, tb);
          act = !(ta > tb);
        }
        if (act) {
          const float4 ml = s_ml[j];
          const float4 g0 = s_g0[j], g1 = s_g1[j];
          const float gg[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
          Segment seg;
          eval_segment(r, ta, tb, ml.x, ml.y, ml.z, ml.w, gg, s, seg);
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
            ds_acc += dLda * (q0 * seg.sdf[0] * (1.f - alpha) - q3 * seg.sdf[3]) * rden;

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
          atomicAdd(dst, mine);
        }
      }
    }
  }
  // One dL/ds atomic per workgroup.
  const float wsum = wave_sum_uniform(ds_acc);
  if (lane == 0) s_red[w] = wsum;
  __syncthreads();
  if (t == 0) atomicAdd(a.dL_ds, (s_red[0] + s_red[1]) + (s_red[2] + s_red[3]));
}

template <bool D, bool N>
static hipError_t launch_bwd_c(const BwdArgs& a, hipStream_t st, unsigned grid) {
  if (a.lambda_R_concen > 0.f && a.gt_color) render_bwd_kernel<D, N, true><<<grid, 256, 0, st>>>(a);
  else render_bwd_kernel<D, N, false><<<grid, 256, 0, st>>>(a);
  return hipGetLastError();
}

hipError_t launch_render_backward(const BwdArgs& a, hipStream_t st) {
  if (a.total_tiles <= 0) return hipSuccess;
  unsigned grid = (unsigned)a.total_tiles;
  if (a.work_counter) grid = (unsigned)min(a.total_tiles, 256 * 8);
  if (a.need_depth && a.need_normal) return launch_bwd_c<true, true>(a, st, grid);
  if (a.need_depth) return launch_bwd_c<true, false>(a, st, grid);
  if (a.need_normal) return launch_bwd_c<false, true>(a, st, grid);
  return launch_bwd_c<false, false>(a, st, grid);
}

}  // namespace svr

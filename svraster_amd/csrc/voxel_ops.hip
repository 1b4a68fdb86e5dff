// Per-voxel operators around the compositor on gfx950: spherical-harmonics colour and its
// transpose, corner-SDF gather and scatter-add, the Adam step, and octree index conversions.
//
// References:
//   SH            cuda/src/sh_compute.cu:33-299 (degree <= 3, direction = voxel centre minus
//                 camera position per voxel, +0.5 offset, clamp >= 0 in forward only)
//   corner gather cuda/src/geo_params_gather.cu:22-72, 134-173
//   Adam          cuda/src/adam_step.cu:21-158
//   octree utils  cuda/src/utils.cu:129-189
#include "common.h"
#include "ops.h"

namespace svr {

// Real SH basis constants, degrees 0..3 (textbook values).
__constant__ float kC0 = 0.28209479177387814f;
__constant__ float kC1 = 0.4886025119029199f;
__constant__ float kC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
                             -1.0925484305920792f, 0.5462742152960396f};
__constant__ float kC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
                             0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f,
                             -0.5900435899266435f};

// Basis values b[0..15] for direction (x, y, z); returns the number used at degree `deg`.
__device__ inline int sh_basis(int deg, float x, float y, float z, float b[16]) {
  b[0] = kC0;
  if (deg < 1) return 1;
  b[1] = -kC1 * y; b[2] = kC1 * z; b[3] = -kC1 * x;
  if (deg < 2) return 4;
  const float xx = x * x, yy = y * y, zz = z * z, xy = x * y, yz = y * z, xz = x * z;
  b[4] = kC2[0] * xy;
  b[5] = kC2[1] * yz;
  b[6] = kC2[2] * (2.f * zz - xx - yy);
  b[7] = kC2[3] * xz;
  b[8] = kC2[4] * (xx - yy);
  if (deg < 3) return 9;
  b[9] = kC3[0] * y * (3.f * xx - yy);
  b[10] = kC3[1] * xy * z;
  b[11] = kC3[2] * y * (4.f * zz - xx - yy);
  b[12] = kC3[3] * z * (2.f * zz - 3.f * xx - 3.f * yy);
  b[13] = kC3[4] * x * (4.f * zz - xx - yy);
  b[14] = kC3[5] * z * (xx - yy);
  b[15] = kC3[6] * x * (xx - 3.f * yy);
  return 16;
}

__device__ inline void sh_dir(const float* __restrict__ centers, const float* __restrict__ cam, int v,
                              float& x, float& y, float& z) {
  x = centers[3 * v] - cam[0];
  y = centers[3 * v + 1] - cam[1];
  z = centers[3 * v + 2] - cam[2];
  const float rn = rsqrtf(x * x + y * y + z * z);
  x *= rn; y *= rn; z *= rn;
}

// Per-voxel SH colour of one view (voxels t of the list, or of idx).  Each wavefront moves the
// coefficient rows of its 64 voxels through LDS, 15 floats (5 coefficients) of every row at a
// time: consecutive lanes read consecutive floats of a row, so a load instruction touches a few
// cache lines rather than one per lane.  kDeg is the SH degree; a row holds
// 3 * ((kDeg + 1)^2 - 1) used floats of its 3 * (M - 1).
template <int kDeg>
struct ShRows {
  static constexpr int kNb = (kDeg + 1) * (kDeg + 1);
  static constexpr int kL = 3 * (kNb - 1);     // floats used per row
  static constexpr int kC = 15;                // floats per chunk (odd: conflict-free LDS rows)
  static constexpr int kChunks = (kL + kC - 1) / kC;
  static constexpr int kRowStride = (kL % 2 == 1 || kL == 0) ? (kL > 0 ? kL : 1) : kL + 1;   // whole rows, odd
};

template <int kDeg>
__global__ __launch_bounds__(256) void sh_forward_kernel(int n, const int64_t* __restrict__ idx, int M,
                                                         const float* __restrict__ centers,
                                                         const float* __restrict__ cam,
                                                         const float* __restrict__ sh0,
                                                         const float* __restrict__ shs,
                                                         float* __restrict__ rgbs) {
  using R = ShRows<kDeg>;
  __shared__ float s_row[4][64 * R::kC];
  __shared__ int s_vox[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t0 = blockIdx.x * 256 + w * 64;
  if (t0 >= n) return;                       // whole wavefront past the end
  const int cnt = min(64, n - t0);
  const bool valid = lane < cnt;
  const int v = valid ? (idx ? (int)idx[t0 + lane] : t0 + lane) : 0;
  float r0 = 0.f, r1 = 0.f, r2 = 0.f;
  float b[16];
  if (valid) {
    float x, y, z;
    sh_dir(centers, cam, v, x, y, z);
    sh_basis(kDeg, x, y, z, b);
    r0 = b[0] * sh0[3 * v]; r1 = b[0] * sh0[3 * v + 1]; r2 = b[0] * sh0[3 * v + 2];
  }
  if (R::kL > 0) {
    s_vox[w][lane] = v;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const long long stride = 3LL * (M - 1);
#pragma unroll
    for (int c = 0; c < R::kChunks; ++c) {
      const int lc = min(R::kC, R::kL - R::kC * c);   // floats of this chunk (constant once unrolled)
#pragma unroll
      for (int i = 0; i < R::kC; ++i) {
        if (i >= lc) break;
        const int e = lane + 64 * i;
        const int row = e / lc, f = e - row * lc;
        if (row < cnt) s_row[w][row * lc + f] = shs[(long long)s_vox[w][row] * stride + R::kC * c + f];
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int kk = 0; kk < R::kC / 3; ++kk) {
        if (3 * kk >= lc) break;
        const float* cf = &s_row[w][lane * lc + 3 * kk];
        const float bk = b[1 + (R::kC / 3) * c + kk];
        r0 += bk * cf[0];
        r1 += bk * cf[1];
        r2 += bk * cf[2];
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (!valid) return;
  r0 += 0.5f; r1 += 0.5f; r2 += 0.5f;
  rgbs[3 * v] = fmaxf(r0, 0.f);
  rgbs[3 * v + 1] = fmaxf(r1, 0.f);
  rgbs[3 * v + 2] = fmaxf(r2, 0.f);
}

// Transpose of the linear part; the forward clamp is not masked (sh_compute.cu:118-121).  Rows
// of dL/dshs are assembled in LDS and written with consecutive lanes on consecutive floats.
template <int kDeg>
__global__ __launch_bounds__(256) void sh_backward_kernel(int n, const int64_t* __restrict__ idx, int M,
                                                          const float* __restrict__ centers,
                                                          const float* __restrict__ cam,
                                                          const float* __restrict__ dL_drgb,
                                                          float* __restrict__ dsh0,
                                                          float* __restrict__ dshs) {
  constexpr int kL = ShRows<kDeg>::kL, kS = ShRows<kDeg>::kRowStride;
  __shared__ float s_row[4][64 * kS];
  __shared__ int s_vox[4][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t0 = blockIdx.x * 256 + w * 64;
  if (t0 >= n) return;
  const int cnt = min(64, n - t0);
  const bool valid = lane < cnt;
  const int v = valid ? (idx ? (int)idx[t0 + lane] : t0 + lane) : 0;
  float b[16];
  float g0 = 0.f, g1 = 0.f, g2 = 0.f;
  if (valid) {
    float x, y, z;
    sh_dir(centers, cam, v, x, y, z);
    sh_basis(kDeg, x, y, z, b);
    g0 = dL_drgb[3 * v]; g1 = dL_drgb[3 * v + 1]; g2 = dL_drgb[3 * v + 2];
    dsh0[3 * v] = b[0] * g0; dsh0[3 * v + 1] = b[0] * g1; dsh0[3 * v + 2] = b[0] * g2;
  }
  const int lw = min(kL, 3 * (M - 1));   // floats of the row that exist
  if (kL > 0 && lw > 0) {
    s_vox[w][lane] = v;
    if (valid) {
#pragma unroll
      for (int k = 1; k < ShRows<kDeg>::kNb; ++k) {
        float* c = &s_row[w][lane * kS + 3 * (k - 1)];
        c[0] = b[k] * g0; c[1] = b[k] * g1; c[2] = b[k] * g2;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const long long stride = 3LL * (M - 1);
#pragma unroll
    for (int i = 0; i < kL; ++i) {
      const int e = lane + 64 * i;
      const int row = e / kL, f = e - row * kL;
      if (row < cnt && f < lw) dshs[(long long)s_vox[w][row] * stride + f] = s_row[w][row * kS + f];
    }
  }
}

__global__ __launch_bounds__(256) void gather_geo_kernel(int n, const int64_t* __restrict__ vox_key,
                                                         const int64_t* __restrict__ care,
                                                         const float* __restrict__ grid_pts,
                                                         float* __restrict__ geos) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long v = care ? care[t] : t;
  const int64_t* k = vox_key + v * 8;
  float4 a, b;
  a.x = grid_pts[k[0]]; a.y = grid_pts[k[1]]; a.z = grid_pts[k[2]]; a.w = grid_pts[k[3]];
  b.x = grid_pts[k[4]]; b.y = grid_pts[k[5]]; b.z = grid_pts[k[6]]; b.w = grid_pts[k[7]];
  reinterpret_cast<float4*>(geos)[2 * v] = a;
  reinterpret_cast<float4*>(geos)[2 * v + 1] = b;
}

__global__ __launch_bounds__(256) void gather_geo_bw_kernel(int n, const int64_t* __restrict__ vox_key,
                                                            const int64_t* __restrict__ care,
                                                            const float* __restrict__ dgeos, int stride,
                                                            float* __restrict__ dgrid) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long v = care ? care[t] : t;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const float g = dgeos[v * stride + i];
    if (g != 0.f) atomicAdd(dgrid + vox_key[v * 8 + i], g);
  }
}

// Batched views: colour of every (view, visible voxel) pair, direction from that view's camera.
// One voxel per thread, looping over the views, so each voxel's coefficients are read once.
// The SH gradient rows of a block of 256 voxels are contiguous ([P, M-1, 3], at most 45 floats
// for M <= 16); the backward passes them through LDS kShVox voxels at a time, whole rows, so each
// pass reads and writes one contiguous span of the row arrays (gradient, or parameter and Adam
// moments) exactly once, coalesced, with 11.5 KB of LDS (the odd row stride keeps the per-thread
// LDS writes free of bank conflicts).
constexpr int kShRowMax = 45;
constexpr int kShVox = 64;

// Bit k set when view v0 + k sees voxel i (non-empty octant set, i.e. n_dup > 0); the loads are
// independent, so they are all in flight at once instead of one dependent load per view.
__device__ inline uint32_t visible_views(const uint8_t* __restrict__ quads, int P, int i, int v0, int nv) {
  uint32_t vis = 0u;
#pragma unroll 8
  for (int k = 0; k < nv; ++k) vis |= (quads[(long long)(v0 + k) * P + i] != 0 ? 1u : 0u) << k;
  return vis;
}

template <int kDeg>
__global__ __launch_bounds__(256) void sh_views_forward_kernel(int P, int V, const uint8_t* __restrict__ quads,
                                                               const ViewDesc* __restrict__ views, int M,
                                                               const float* __restrict__ centers,
                                                               const float* __restrict__ sh0,
                                                               const float* __restrict__ shs,
                                                               float* __restrict__ rgbs_views) {
  constexpr int NB = (kDeg + 1) * (kDeg + 1);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  float c[NB][3];
  bool loaded = false;
  for (int v0 = 0; v0 < V; v0 += 32) {
    const int nvw = min(32, V - v0);
    const uint32_t vis = visible_views(quads, P, i, v0, nvw);
    if (vis && !loaded) {          // coefficients only for voxels some view sees
      c[0][0] = sh0[3 * i]; c[0][1] = sh0[3 * i + 1]; c[0][2] = sh0[3 * i + 2];
      const float* cs = shs + (long long)i * (M - 1) * 3;
#pragma unroll
      for (int k = 1; k < NB; ++k) {
        c[k][0] = cs[3 * (k - 1)]; c[k][1] = cs[3 * (k - 1) + 1]; c[k][2] = cs[3 * (k - 1) + 2];
      }
      loaded = true;
    }
    // all lanes on the same view at once: coalesced writes into the [V, P, 3] colours
    for (int k = 0; k < nvw; ++k) {
      const bool on = (vis >> k) & 1u;
      if (!__any(on)) continue;
      if (!on) continue;
      const int v = v0 + k;
      const long long o = (long long)v * P + i;
      const float* c2w = views[v].c2w;
      const float cam[3] = {c2w[3], c2w[7], c2w[11]};
      float x, y, z;
      sh_dir(centers, cam, i, x, y, z);
      float b[16];
      sh_basis(kDeg, x, y, z, b);
      float r0 = 0.f, r1 = 0.f, r2 = 0.f;
#pragma unroll
      for (int kk = 0; kk < NB; ++kk) {
        r0 += b[kk] * c[kk][0]; r1 += b[kk] * c[kk][1]; r2 += b[kk] * c[kk][2];
      }
      rgbs_views[3 * o] = fmaxf(r0 + 0.5f, 0.f);
      rgbs_views[3 * o + 1] = fmaxf(r1 + 0.5f, 0.f);
      rgbs_views[3 * o + 2] = fmaxf(r2 + 0.5f, 0.f);
    }
  }
}

// Transpose summed over the views, one voxel per thread: no atomics, and every output row is
// written (zeros for voxels no view saw), so the outputs need no separate clearing.  The
// per-view colour gradients it consumes are reset to zero, so the backward compositor's
// accumulation buffer is clear again for the next step without a separate fill.
// Adam (adam_step.cu:21-102).  lerp(v0, v1, t) = fma(t, v1, fma(-t, v0, v0)).
__device__ inline float adam_lerp(float v0, float v1, float t) { return fmaf(t, v1, fmaf(-t, v0, v0)); }

// One element: m = lerp(g, m, b1), v = lerp(g^2, v, b2), p -= step_size * m / denom.
template <bool kBiased>
__device__ inline void adam_update(float& p, float& m, float& v, float g, float step_size, float b1, float b2,
                                   float rsqrt_bc2, float eps) {
#pragma clang fp contract(off)
  const float mv = adam_lerp(g, m, b1);
  const float vv = adam_lerp(g * g, v, b2);
  const float denom = kBiased ? sqrtf(vv) + eps : fmaf(sqrtf(vv), rsqrt_bc2, eps);
  p -= step_size * (mv / denom);
  m = mv;
  v = vv;
}

// Unbiased step whose bias corrections come from a table indexed by a step counter held in
// device memory: entry t-1 = {lr / (1 - b1^t), 1 / sqrt(1 - b2^t)}, rounded to float from the
// same double expressions as svr_unbiased_adam_step; steps past the table use its last entry.
__device__ inline float2 adam_table_entry(const AdamTable& T) {
  const int t = min(max(*T.step, 1), T.n);
  return T.tab[t - 1];
}

// kAdam: instead of writing dL/dsh0 and dL/dshs, apply the unbiased Adam step of `A` to sh0/shs
// right here, so the SH gradients never reach HBM.  Dense (A.sparse == 0, the reference
// trainer's default): every element of every voxel steps, a zero gradient included.  Sparse:
// elements whose gradient is exactly 0 are left alone.  Same gradient values and Adam
// arithmetic as the unfused pair of calls.
template <int kDeg, bool kAdam>
__global__ __launch_bounds__(256) void sh_views_backward_kernel(int P, int V, const uint8_t* __restrict__ quads,
                                                                const ViewDesc* __restrict__ views, int M,
                                                                const float* __restrict__ centers,
                                                                float* __restrict__ drgb_views,
                                                                float* __restrict__ dsh0,
                                                                float* __restrict__ dshs, ShAdam A) {
  constexpr int NB = (kDeg + 1) * (kDeg + 1);
  __shared__ float s_rows[kShVox * kShRowMax];
  const int t = threadIdx.x;
  const long long i0 = (long long)blockIdx.x * 256;
  const int nv = (int)min(256LL, (long long)P - i0);
  const int row = (M - 1) * 3;
  float2 h0 = make_float2(0.f, 0.f), hs = h0;
  if (kAdam) {
    h0 = adam_table_entry(A.t0);
    hs = adam_table_entry(A.ts);
  }
  float acc[NB][3];
#pragma unroll
  for (int k = 0; k < NB; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.f;
  if (t < nv) {
    const int i = (int)i0 + t;
    for (int v0 = 0; v0 < V; v0 += 32) {
      const int nvw = min(32, V - v0);
      const uint32_t vis = visible_views(quads, P, i, v0, nvw);
      // all lanes on the same view at once, so the per-view gradient rows are read and cleared
      // with coalesced accesses ([V, P, 3] layout)
      for (int k = 0; k < nvw; ++k) {
        const bool on = (vis >> k) & 1u;
        if (!__any(on)) continue;
        if (!on) continue;
        const int v = v0 + k;
        const long long o = (long long)v * P + i;
        const float g0 = drgb_views[3 * o], g1 = drgb_views[3 * o + 1], g2 = drgb_views[3 * o + 2];
        if (g0 == 0.f && g1 == 0.f && g2 == 0.f) continue;     // no contribution in this view
        drgb_views[3 * o] = 0.f;
        drgb_views[3 * o + 1] = 0.f;
        drgb_views[3 * o + 2] = 0.f;
        const float* c2w = views[v].c2w;
        const float cam[3] = {c2w[3], c2w[7], c2w[11]};
        float x, y, z;
        sh_dir(centers, cam, i, x, y, z);
        float b[16];
        sh_basis(kDeg, x, y, z, b);
#pragma unroll
        for (int kk = 0; kk < NB; ++kk) {
          acc[kk][0] += b[kk] * g0; acc[kk][1] += b[kk] * g1; acc[kk][2] += b[kk] * g2;
        }
      }
    }
    if (kAdam) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float g = acc[0][c];
        if (!A.sparse || g != 0.0f) {
          const long long e = 3LL * i + c;
          float pv = A.sh0[e], mv = A.m0[e], vv = A.v0[e];
          adam_update<false>(pv, mv, vv, g, h0.x, A.t0.b1, A.t0.b2, h0.y, A.t0.eps);
          A.sh0[e] = pv; A.m0[e] = mv; A.v0[e] = vv;
        }
      }
    } else {
      dsh0[3 * i] = acc[0][0]; dsh0[3 * i + 1] = acc[0][1]; dsh0[3 * i + 2] = acc[0][2];
    }
  }
  // Rows through LDS, kShVox voxels at a time (element q of a row is coefficient 1 + q / 3,
  // channel q % 3; elements past the active degree are zero).
  for (int p0 = 0; p0 < nv; p0 += kShVox) {
    const int pn = min(kShVox, nv - p0);
    if (t >= p0 && t < p0 + pn) {
      float* c = s_rows + (t - p0) * row;
#pragma unroll
      for (int q = 0; q < kShRowMax; ++q)
        if (q < row) c[q] = (1 + q / 3 < NB) ? acc[1 + q / 3 < NB ? 1 + q / 3 : 0][q % 3] : 0.f;
    }
    __syncthreads();
    const long long e0 = (i0 + p0) * row;
    const int n = pn * row;
    if (kAdam) {
      // The group's elements are contiguous in sh / m / v: whole float4s (the group starts
      // 16-byte aligned when the arrays do, as (i0 + p0) * row is a multiple of 4), with the
      // loads of three float4 steps issued before their updates; a tail of n % 4 elements, and
      // unaligned arrays, take the scalar loop.
      const bool vec = ((reinterpret_cast<uintptr_t>(A.shs + e0) | reinterpret_cast<uintptr_t>(A.ms + e0) |
                         reinterpret_cast<uintptr_t>(A.vs + e0)) & 15u) == 0;
      const int n4 = vec ? (n >> 2) : 0;
      float4* P4 = reinterpret_cast<float4*>(A.shs + e0);
      float4* M4 = reinterpret_cast<float4*>(A.ms + e0);
      float4* V4 = reinterpret_cast<float4*>(A.vs + e0);
      const float4* G4 = reinterpret_cast<const float4*>(s_rows);
      for (int k0 = t; k0 < n4; k0 += 3 * 256) {
        float4 pv[3], mv[3], vv[3];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
          const int k = k0 + u * 256;
          if (k < n4) { pv[u] = P4[k]; mv[u] = M4[k]; vv[u] = V4[k]; }
        }
#pragma unroll
        for (int u = 0; u < 3; ++u) {
          const int k = k0 + u * 256;
          if (k >= n4) continue;
          const float4 g = G4[k];
          float* pp = &pv[u].x;
          float* mm = &mv[u].x;
          float* qq = &vv[u].x;
          const float gg[4] = {g.x, g.y, g.z, g.w};
#pragma unroll
          for (int c = 0; c < 4; ++c)
            if (!A.sparse || gg[c] != 0.0f)
              adam_update<false>(pp[c], mm[c], qq[c], gg[c], hs.x, A.ts.b1, A.ts.b2, hs.y, A.ts.eps);
          P4[k] = pv[u]; M4[k] = mv[u]; V4[k] = vv[u];
        }
      }
      for (int k = 4 * n4 + t; k < n; k += 256) {
        const float g = s_rows[k];
        if (A.sparse && g == 0.0f) continue;
        const long long e = e0 + k;
        float pv = A.shs[e], mv = A.ms[e], vv = A.vs[e];
        adam_update<false>(pv, mv, vv, g, hs.x, A.ts.b1, A.ts.b2, hs.y, A.ts.eps);
        A.shs[e] = pv; A.ms[e] = mv; A.vs[e] = vv;
      }
    } else {
      for (int k = t; k < n; k += 256) dshs[e0 + k] = s_rows[k];
    }
    __syncthreads();
  }
}

template <bool kSparse, bool kBiased>
__global__ __launch_bounds__(256) void adam_kernel(long long n, float* __restrict__ p,
                                                   const float* __restrict__ g, float* __restrict__ m,
                                                   float* __restrict__ v, float step_size, float b1,
                                                   float b2, float rsqrt_bc2, float eps) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gv = g[i];
  if (kSparse && gv == 0.0f) return;
  float pv = p[i], mv = m[i], vv = v[i];
  adam_update<kBiased>(pv, mv, vv, gv, step_size, b1, b2, rsqrt_bc2, eps);
  p[i] = pv;
  m[i] = mv;
  v[i] = vv;
}

template <bool kSparse>
__global__ __launch_bounds__(256) void adam_table_kernel(long long n, float* __restrict__ p,
                                                         const float* __restrict__ g, float* __restrict__ m,
                                                         float* __restrict__ v, AdamTable T) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gv = g[i];
  if (kSparse && gv == 0.0f) return;
  const float2 h = adam_table_entry(T);
  float pv = p[i], mv = m[i], vv = v[i];
  adam_update<false>(pv, mv, vv, gv, h.x, T.b1, T.b2, h.y, T.eps);
  p[i] = pv;
  m[i] = mv;
  v[i] = vv;
}

__global__ void ijk_to_octpath_kernel(int n, const int64_t* __restrict__ ijk, const int8_t* __restrict__ lv,
                                      int64_t* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int i = (int)ijk[3 * t], j = (int)ijk[3 * t + 1], k = (int)ijk[3 * t + 2];
  int64_t path = 0;
  for (int l = 0; l < kMaxLevels; ++l) {
    path |= (int64_t)(((i & 1) << 2) | ((j & 1) << 1) | (k & 1)) << (3 * l);
    i >>= 1; j >>= 1; k >>= 1;
  }
  out[t] = path << (3 * (kMaxLevels - lv[t]));
}

__global__ void octpath_to_ijk_kernel(int n, const int64_t* __restrict__ path_in, const int8_t* __restrict__ lv,
                                      int64_t* __restrict__ ijk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int64_t path = path_in[t] >> (3 * (kMaxLevels - lv[t]));
  int i = 0, j = 0, k = 0;
  for (int l = 0; l < kMaxLevels; ++l) {
    const int b = (int)(path & 7);
    i |= ((b >> 2) & 1) << l;
    j |= ((b >> 1) & 1) << l;
    k |= (b & 1) << l;
    path >>= 3;
  }
  ijk[3 * t] = i; ijk[3 * t + 1] = j; ijk[3 * t + 2] = k;
}

// Per-voxel corner values into the grid points, averaged over the voxels sharing each point
// (agg_voxel_into_grid_pts, adaptive.py:336-360: index_reduce "mean", include_self False).
// Sums and counts by atomics, then one division per point; points no voxel names are 0.
__global__ __launch_bounds__(256) void grid_pts_accumulate_kernel(long long n, int C, const int64_t* __restrict__ vox_key,
                                                                  const float* __restrict__ vox_val,
                                                                  float* __restrict__ sum,
                                                                  uint32_t* __restrict__ count) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long corner = t / C;
  const int c = (int)(t - corner * C);
  const long long key = vox_key[corner];
  atomicAdd(sum + key * C + c, vox_val[t]);
  if (c == 0) atomicAdd(count + key, 1u);
}

__global__ __launch_bounds__(256) void grid_pts_mean_kernel(long long n, int C, const uint32_t* __restrict__ count,
                                                            float* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t k = count[t / C];
  out[t] = k ? out[t] / (float)k : 0.f;
}

__global__ void fill_f32_kernel(long long n, float* __restrict__ p, float v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

__global__ void fill_u32v_kernel(long long n, uint32_t* __restrict__ p, uint32_t v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// ------------------------------------------------------------------------------------------
static inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_grid_pts_mean(long long N, int C, const int64_t* vox_key, const float* vox_val, long long M,
                                uint32_t* count, float* out, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  fill_f32_kernel<<<blocks_for(M * C), 256, 0, st>>>(M * C, out, 0.f);
  fill_u32v_kernel<<<blocks_for(M), 256, 0, st>>>(M, count, 0u);
  if (N > 0)
    grid_pts_accumulate_kernel<<<blocks_for(8 * N * C), 256, 0, st>>>(8 * N * C, C, vox_key, vox_val, out, count);
  grid_pts_mean_kernel<<<blocks_for(M * C), 256, 0, st>>>(M * C, C, count, out);
  return hipGetLastError();
}

hipError_t launch_sh_forward(int n, const int64_t* idx, int deg, int M, const float* centers,
                             const float* cam, const float* sh0, const float* shs, float* rgbs,
                             hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned g = blocks_for(n);
  if (deg == 0) sh_forward_kernel<0><<<g, 256, 0, st>>>(n, idx, M, centers, cam, sh0, shs, rgbs);
  else if (deg == 1) sh_forward_kernel<1><<<g, 256, 0, st>>>(n, idx, M, centers, cam, sh0, shs, rgbs);
  else if (deg == 2) sh_forward_kernel<2><<<g, 256, 0, st>>>(n, idx, M, centers, cam, sh0, shs, rgbs);
  else sh_forward_kernel<3><<<g, 256, 0, st>>>(n, idx, M, centers, cam, sh0, shs, rgbs);
  return hipGetLastError();
}

hipError_t launch_sh_backward(int n, const int64_t* idx, int deg, int M, const float* centers,
                              const float* cam, const float* dL_drgb, float* dsh0, float* dshs,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned g = blocks_for(n);
  if (deg == 0) sh_backward_kernel<0><<<g, 256, 0, st>>>(n, idx, M, centers, cam, dL_drgb, dsh0, dshs);
  else if (deg == 1) sh_backward_kernel<1><<<g, 256, 0, st>>>(n, idx, M, centers, cam, dL_drgb, dsh0, dshs);
  else if (deg == 2) sh_backward_kernel<2><<<g, 256, 0, st>>>(n, idx, M, centers, cam, dL_drgb, dsh0, dshs);
  else sh_backward_kernel<3><<<g, 256, 0, st>>>(n, idx, M, centers, cam, dL_drgb, dsh0, dshs);
  return hipGetLastError();
}

hipError_t launch_sh_views_forward(int P, int V, const uint8_t* quads, const ViewDesc* views, int deg, int M,
                                   const float* centers, const float* sh0, const float* shs, float* rgbs_views,
                                   hipStream_t st) {
  if (P <= 0 || V <= 0) return hipSuccess;
  const unsigned nb = blocks_for(P);
  switch (deg) {
    case 0: sh_views_forward_kernel<0><<<nb, 256, 0, st>>>(P, V, quads, views, M, centers, sh0, shs, rgbs_views); break;
    case 1: sh_views_forward_kernel<1><<<nb, 256, 0, st>>>(P, V, quads, views, M, centers, sh0, shs, rgbs_views); break;
    case 2: sh_views_forward_kernel<2><<<nb, 256, 0, st>>>(P, V, quads, views, M, centers, sh0, shs, rgbs_views); break;
    default: sh_views_forward_kernel<3><<<nb, 256, 0, st>>>(P, V, quads, views, M, centers, sh0, shs, rgbs_views); break;
  }
  return hipGetLastError();
}

hipError_t launch_sh_views_backward(int P, int V, const uint8_t* quads, const ViewDesc* views, int deg, int M,
                                    const float* centers, float* drgb_views, float* dsh0, float* dshs,
                                    const ShAdam* adam, hipStream_t st) {
  if (P <= 0) return hipSuccess;
  const unsigned nb = blocks_for(P);
  const ShAdam A = adam ? *adam : ShAdam{};
#define SVR_SHB(D)                                                                                           \
  if (adam) sh_views_backward_kernel<D, true><<<nb, 256, 0, st>>>(P, V, quads, views, M, centers, drgb_views,  \
                                                                  dsh0, dshs, A);                           \
  else sh_views_backward_kernel<D, false><<<nb, 256, 0, st>>>(P, V, quads, views, M, centers, drgb_views, dsh0, \
                                                              dshs, A);
  switch (deg) {
    case 0: SVR_SHB(0) break;
    case 1: SVR_SHB(1) break;
    case 2: SVR_SHB(2) break;
    default: SVR_SHB(3) break;
  }
#undef SVR_SHB
  return hipGetLastError();
}

// Leaf voxels that can still produce a pair alpha >= kMinAlpha under the current corner SDFs
// and s (voxel_may_contribute); the rest can be left out of binning without changing any pixel.
__global__ __launch_bounds__(256) void voxel_activity_kernel(int P, const uint8_t* __restrict__ is_leaf,
                                                             const float* __restrict__ geos,
                                                             const float* __restrict__ s_val,
                                                             const float* __restrict__ centers,
                                                             const float* __restrict__ lengths, float cam_scale,
                                                             uint8_t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  uint8_t keep = 0;
  if (is_leaf[i]) {
    const float4 g0 = reinterpret_cast<const float4*>(geos)[2 * i];
    const float4 g1 = reinterpret_cast<const float4*>(geos)[2 * i + 1];
    const float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
    const float c = fmaxf(fabsf(centers[3 * i]), fmaxf(fabsf(centers[3 * i + 1]), fabsf(centers[3 * i + 2])));
    keep = voxel_may_contribute(g, *s_val, cam_scale + c, lengths[i]) ? 1 : 0;
  }
  out[i] = keep;
}

// Voxel records of the multi-view compositors (VoxelRecord): corner SDFs gathered from the grid
// points through vox_key (the reference's corner gather, geo_params_gather.cu:134-152), centre,
// size and 1/size, one 64-byte record per voxel; and, when bin_leaf is given, the voxels that
// enter binning: leaves whose corner SDFs can reach the minimum alpha (with cull; every leaf
// without).
__global__ __launch_bounds__(256) void voxel_records_kernel(int P, const int64_t* __restrict__ vox_key,
                                                            const float* __restrict__ grid_pts,
                                                            const float* __restrict__ centers,
                                                            const float* __restrict__ lengths,
                                                            const uint8_t* __restrict__ is_leaf,
                                                            const float* __restrict__ s_val, float cam_scale, int cull,
                                                            VoxelRecord* __restrict__ rec,
                                                            uint8_t* __restrict__ bin_leaf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const int64_t* k = vox_key + 8LL * i;
  const int64_t k0 = k[0], k1 = k[1], k2 = k[2], k3 = k[3], k4 = k[4], k5 = k[5], k6 = k[6], k7 = k[7];
  const float4 g0 = make_float4(grid_pts[k0], grid_pts[k1], grid_pts[k2], grid_pts[k3]);
  const float4 g1 = make_float4(grid_pts[k4], grid_pts[k5], grid_pts[k6], grid_pts[k7]);
  const float cx = centers[3 * i], cy = centers[3 * i + 1], cz = centers[3 * i + 2];
  const float l = lengths[i];
  VoxelRecord* R = rec + i;
  R->cl = make_float4(cx, cy, cz, l);
  R->g0 = g0;
  R->g1 = g1;
  R->aux = make_float4(1.f / l, 0.f, 0.f, 0.f);
  if (bin_leaf) {
    uint8_t keep = 0;
    if (is_leaf[i]) {
      const float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
      const float c = fmaxf(fabsf(cx), fmaxf(fabsf(cy), fabsf(cz)));
      keep = (!cull || voxel_may_contribute(g, *s_val, cam_scale + c, l)) ? 1 : 0;
    }
    bin_leaf[i] = keep;
  }
}

hipError_t launch_voxel_records(int P, const int64_t* vox_key, const float* grid_pts, const float* centers,
                                const float* lengths, const uint8_t* is_leaf, const float* s_val, float cam_scale,
                                int cull, VoxelRecord* rec, uint8_t* bin_leaf, hipStream_t st) {
  if (P <= 0) return hipSuccess;
  voxel_records_kernel<<<blocks_for(P), 256, 0, st>>>(P, vox_key, grid_pts, centers, lengths, is_leaf, s_val,
                                                       cam_scale, cull, rec, bin_leaf);
  return hipGetLastError();
}

hipError_t launch_voxel_activity(int P, const uint8_t* is_leaf, const float* geos, const float* s_val,
                                 const float* centers, const float* lengths, float cam_scale, uint8_t* out,
                                 hipStream_t st) {
  if (P <= 0) return hipSuccess;
  voxel_activity_kernel<<<blocks_for(P), 256, 0, st>>>(P, is_leaf, geos, s_val, centers, lengths, cam_scale, out);
  return hipGetLastError();
}

hipError_t launch_gather_geo(int n, const int64_t* vox_key, const int64_t* care, const float* grid_pts,
                             float* geos, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  gather_geo_kernel<<<blocks_for(n), 256, 0, st>>>(n, vox_key, care, grid_pts, geos);
  return hipGetLastError();
}

hipError_t launch_gather_geo_bw(int n, const int64_t* vox_key, const int64_t* care, const float* dgeos,
                                float* dgrid, hipStream_t st, int stride) {
  if (n <= 0) return hipSuccess;
  gather_geo_bw_kernel<<<blocks_for(n), 256, 0, st>>>(n, vox_key, care, dgeos, stride, dgrid);
  return hipGetLastError();
}

hipError_t launch_adam(bool sparse, bool biased, long long n, float* p, const float* g, float* m, float* v,
                       float step_size, float b1, float b2, float rsqrt_bc2, float eps, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned nb = blocks_for(n);
  if (sparse && biased) adam_kernel<true, true><<<nb, 256, 0, st>>>(n, p, g, m, v, step_size, b1, b2, rsqrt_bc2, eps);
  else if (sparse) adam_kernel<true, false><<<nb, 256, 0, st>>>(n, p, g, m, v, step_size, b1, b2, rsqrt_bc2, eps);
  else if (biased) adam_kernel<false, true><<<nb, 256, 0, st>>>(n, p, g, m, v, step_size, b1, b2, rsqrt_bc2, eps);
  else adam_kernel<false, false><<<nb, 256, 0, st>>>(n, p, g, m, v, step_size, b1, b2, rsqrt_bc2, eps);
  return hipGetLastError();
}

hipError_t launch_adam_table(bool sparse, long long n, float* p, const float* g, float* m, float* v,
                             const AdamTable& T, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned nb = blocks_for(n);
  if (sparse) adam_table_kernel<true><<<nb, 256, 0, st>>>(n, p, g, m, v, T);
  else adam_table_kernel<false><<<nb, 256, 0, st>>>(n, p, g, m, v, T);
  return hipGetLastError();
}

hipError_t launch_ijk_to_octpath(int n, const int64_t* ijk, const int8_t* lv, int64_t* out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  ijk_to_octpath_kernel<<<blocks_for(n), 256, 0, st>>>(n, ijk, lv, out);
  return hipGetLastError();
}

hipError_t launch_octpath_to_ijk(int n, const int64_t* path, const int8_t* lv, int64_t* ijk, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  octpath_to_ijk_kernel<<<blocks_for(n), 256, 0, st>>>(n, path, lv, ijk);
  return hipGetLastError();
}

}  // namespace svr

// Shared device helpers for the gfx950 sparse-voxel SDF rasterizer.
//
// Conventions restated from the reference (cuda/src/config.h:14-39, cuda/src/auxiliary.h):
//   * 16x16 pixel tiles; a tile is one 256-thread workgroup = four wave64s, each wave an
//     8x8 pixel quadrant of the tile.
//   * octpath: 48 rank bits (16 levels x 3 bits), x = bit2, y = bit1, z = bit0.
//   * sorted value = octant << 29 | voxel id.
//   * corner index of the 8 corner SDFs = x<<2 | y<<1 | z.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace svr {

constexpr int kTile = 16;
constexpr int kTilePixels = kTile * kTile;
constexpr int kMaxLevels = 16;
constexpr int kRankBits = 3 * kMaxLevels;
constexpr float kMinAlpha = 1e-5f;
constexpr float kEarlyStopT = 1e-4f;
constexpr float kPhiEps = 1e-6f;
constexpr float kThird = 0.3333333f;       // depth sub-step factor used by the reference
constexpr uint32_t kVoxMask = (1u << 29) - 1u;

constexpr int kVoxGeoTriinterp = 2, kVoxGeoTriinterp1 = 1, kVoxGeoTriinterp3 = 3;
constexpr int kDensityExpLinear11 = 0, kDensitySDF = 1;
constexpr int kCamPersp = 0, kCamOrtho = 10;

// Range in which rcp_rn is the correctly rounded reciprocal: |x| in [2^-100, 2^100].
constexpr float kRcpRnLo = 0x1p-100f, kRcpRnHi = 0x1p100f;

// 1/x rounded to nearest for |x| in [kRcpRnLo, kRcpRnHi]: the hardware reciprocal (within 1 ulp)
// and one Newton step, whose residual 1 - x y is exact in an FMA.  The result equals the IEEE
// quotient 1.0f / x for every float32 of that range on gfx950; svr_selfcheck_rcp_rn checks all
// 2^23 significands of any exponent range on the device, against the IEEE division.
__device__ inline float rcp_rn(float x) {
  const float y = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, y, 1.f), y, y);
}

// x / y rounded to nearest for y in rcp_rn's range and x, x / y normal numbers (or x = 0): the
// product q0 = x * (1/y) with the correctly rounded reciprocal is within an ulp of the quotient,
// its remainder x - q0 y is exact in an FMA, and one correction q0 + (x - q0 y) * (1/y) rounds to
// the correctly rounded quotient (Markstein's theorem for FMA division).  Six instructions instead
// of the nine-to-ten of the IEEE division sequence; svr_selfcheck_div_rn compares it with 1.0f * x / y
// on random operands on the device.
__device__ inline float div_rn(float x, float y) {
#pragma clang fp contract(off)
  const float yi = rcp_rn(y);
  const float q0 = x * yi;
  return __builtin_fmaf(__builtin_fmaf(-q0, y, x), yi, q0);
}

// One camera view as the kernels see it.  Built on the device from the caller's matrices so
// that no host synchronisation is needed, and stored in the per-view state buffers.
struct alignas(16) ViewDesc {
  int W, H, tiles_x, tiles_y;
  float tanfovx, tanfovy, cx, cy;
  float focal_x, focal_y, near_, pad0;
  float w2c[12];
  float c2w[12];
  float bg[3];
  int tile_base;          // first global tile index of this view (batched launches)
  long long pix_base;     // first pixel of this view in concatenated per-pixel buffers
  int tile_w, tile_h;     // tile shape: 16x16 (the reference's), or 32x8 / 8x32 (batched path)
};

// The four 8x8 pixel quadrants of a 256-pixel tile, one per wavefront: quadrant origin of
// wavefront w of tile (tx, ty).  16x16 tiles are 2x2 quadrants, 32x8 tiles 4x1, 8x32 tiles 1x4.
__host__ __device__ inline void tile_quadrant(const ViewDesc& v, int tx, int ty, int w, int& qx, int& qy) {
  const int across = v.tile_w >> 3;
  qx = tx * v.tile_w + (w % across) * 8;
  qy = ty * v.tile_h + (w / across) * 8;
}

// The pixels of wavefront w of tile (tx, ty): an 8x8 lattice with strides (2^kx, 2^ky),
// 2^kx * 2^ky = 4, interleaved with the lattices of the tile's other three wavefronts, so the four
// see alike shares of every entry's footprint (the work per staged entry is balanced across the
// workgroup).  16x16 tiles: strides 2x2, 32x8: 4x1, 8x32: 1x4.  Lane l holds lattice point
// (l & 7, l >> 3), the pixel (ox + (l & 7) * 2^kx, oy + (l >> 3) * 2^ky).
struct WaveLattice {
  int ox, oy, kx, ky;
};

__host__ __device__ inline WaveLattice wave_lattice(const ViewDesc& v, int tx, int ty, int w) {
  const int kx = v.tile_w == 32 ? 2 : (v.tile_w == 16 ? 1 : 0);
  WaveLattice L;
  L.kx = kx;
  L.ky = 2 - kx;
  L.ox = tx * v.tile_w + (w & ((1 << kx) - 1));
  L.oy = ty * v.tile_h + (w >> kx);
  return L;
}

__host__ __device__ inline int lattice_px(const WaveLattice& L, int lane) { return L.ox + ((lane & 7) << L.kx); }
__host__ __device__ inline int lattice_py(const WaveLattice& L, int lane) { return L.oy + ((lane >> 3) << L.ky); }

__host__ __device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Octant of a direction / of a corner relative to the camera (auxiliary.h:76-84).
__device__ inline uint32_t ray_octant(float x, float y, float z) {
  return (uint32_t(x < 0.f) << 2) | (uint32_t(y < 0.f) << 1) | uint32_t(z < 0.f);
}

// Pixel-centre ray in camera space (auxiliary.h:87-95).
__device__ inline void camera_ray(const ViewDesc& v, int px, int py, float& x, float& y) {
#pragma clang fp contract(off)
  x = ((float)px + 0.5f - v.cx) * 2.f * v.tanfovx / (float)v.W;
  y = ((float)py + 0.5f - v.cy) * 2.f * v.tanfovy / (float)v.H;
}

struct Ray {
  float ox, oy, oz;
  float dx, dy, dz;
  float ix, iy, iz;     // 1/d
  float norm_inv;       // 1/|cam_rd|
  uint32_t octant;
};

// World ray of pixel (px, py): rotate the camera ray by c2w and divide by its camera-space
// length (forward.cu:124-133).
__device__ inline Ray pixel_ray(const ViewDesc& v, int px, int py) {
#pragma clang fp contract(off)
  float cx_, cy_;
  camera_ray(v, px, py, cx_, cy_);
  const float n = sqrtf(cx_ * cx_ + cy_ * cy_ + 1.f);
  Ray r;
  r.norm_inv = n <= kRcpRnHi ? rcp_rn(n) : 1.f / n;   // n >= 1: the IEEE quotient 1 / n
  const float* m = v.c2w;
  r.dx = (m[0] * cx_ + m[1] * cy_ + m[2]) * r.norm_inv;
  r.dy = (m[4] * cx_ + m[5] * cy_ + m[6]) * r.norm_inv;
  r.dz = (m[8] * cx_ + m[9] * cy_ + m[10]) * r.norm_inv;
  r.ox = m[3]; r.oy = m[7]; r.oz = m[11];
  // |d| <= 1: the IEEE quotient from rcp_rn unless a component is below its range (or zero)
  r.ix = fabsf(r.dx) >= kRcpRnLo ? rcp_rn(r.dx) : 1.f / r.dx;
  r.iy = fabsf(r.dy) >= kRcpRnLo ? rcp_rn(r.dy) : 1.f / r.dy;
  r.iz = fabsf(r.dz) >= kRcpRnLo ? rcp_rn(r.dz) : 1.f / r.dz;
  r.octant = ray_octant(r.dx, r.dy, r.dz);
  return r;
}

// Ray of another pixel of the same view: shared origin, direction and its inverse from LDS
// (both computed by pixel_ray).
__device__ inline Ray pair_ray(const Ray& r, const float4 d, const float4 inv) {
  Ray pr = r;
  pr.dx = d.x; pr.dy = d.y; pr.dz = d.z;
  pr.ix = inv.x; pr.iy = inv.y; pr.iz = inv.z;
  return pr;
}

// Slab test of a ray against a cube of half size h whose centre relative to the ray origin is
// (dx, dy, dz) (auxiliary.h:251-264, where dir = vox_c - ro).  The compositors stage the relative
// centre once per entry, since every ray of a view starts at the camera centre.
__device__ inline void ray_box_rel(const Ray& r, float dx, float dy, float dz, float h, float& a, float& b) {
  const float x0 = (dx - h) * r.ix, x1 = (dx + h) * r.ix;
  const float y0 = (dy - h) * r.iy, y1 = (dy + h) * r.iy;
  const float z0 = (dz - h) * r.iz, z1 = (dz + h) * r.iz;
  a = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fminf(z0, z1));
  b = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
}

// Slab test against a cube of centre c and half size h.
__device__ inline void ray_box(const Ray& r, float cx, float cy, float cz, float h, float& a, float& b) {
  ray_box_rel(r, cx - r.ox, cy - r.oy, cz - r.oz, h, a, b);
}

__device__ inline bool pixel_in_bbox(int px, int py, uint32_t b0, uint32_t b1) {
  return px >= int(b0 >> 16) && py >= int(b0 & 0xFFFFu) &&
         px <= int(b1 >> 16) && py <= int(b1 & 0xFFFFu);
}

// Trilinear interpolation of 8 corner values at local coordinate q in [0,1]^3, as nested
// lerps (7 lerps).  Forward and backward both call this so alpha is bit-identical in the two
// passes, which keeps the T / (1 - alpha) recovery exact.
__device__ inline float trilerp(const float g[8], float x, float y, float z) {
#pragma clang fp contract(off)
  const float c00 = fmaf(z, g[1] - g[0], g[0]);
  const float c01 = fmaf(z, g[3] - g[2], g[2]);
  const float c10 = fmaf(z, g[5] - g[4], g[4]);
  const float c11 = fmaf(z, g[7] - g[6], g[6]);
  const float c0 = fmaf(y, c01 - c00, c00);
  const float c1 = fmaf(y, c11 - c10, c10);
  return fmaf(x, c1 - c0, c0);
}

__device__ inline void trilinear_weights(float x, float y, float z, float w[8]) {
  const float x0 = 1.f - x, y0 = 1.f - y, z0 = 1.f - z;
  const float a00 = x0 * y0, a01 = x0 * y, a10 = x * y0, a11 = x * y;
  w[0] = a00 * z0; w[1] = a00 * z; w[2] = a01 * z0; w[3] = a01 * z;
  w[4] = a10 * z0; w[5] = a10 * z; w[6] = a11 * z0; w[7] = a11 * z;
}

// The logistic CDF of s * sdf, 1 / (1 + 2^(sdf * k)) with k = sdf_cdf_scale(s) = -s log2(e): one
// multiply (k is loop-invariant, hoisted out of the callers' pair loops), the hardware exp2 and
// reciprocal.  Every alpha of the compositors (forward, backward, median depth) comes from it.
__device__ inline float sdf_cdf_scale(float s) { return s * -1.44269504088896341f; }
__device__ inline float sdf_cdf(float sdf, float k) {
#pragma clang fp contract(off)
  return __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(sdf * k));
}

// Voxel record of the multi-view path, 64 bytes per voxel (one aligned segment per staged
// entry): centre xyz and size, the eight corner SDFs (gathered from the grid points), 1/size and
// three unused floats.
struct VoxelRecord {
  float4 cl;    // centre xyz, size
  float4 g0, g1;
  float4 aux;   // 1/size (IEEE quotient), 0, 0, 0
};

// Geometry of voxel v for the compositors: from the 64-byte record when there is one, else from
// the separate centre / size / corner-SDF arrays (the one-view drop-in path).
struct VoxelGeom {
  float cx, cy, cz, l, l_inv;
  float4 g0, g1;
};

__device__ inline void load_voxel_geom(VoxelGeom& e, const VoxelRecord* __restrict__ rec,
                                       const float* __restrict__ centers, const float* __restrict__ lengths,
                                       const float* __restrict__ geos, uint32_t v) {
  if (rec) {
    const VoxelRecord& R = rec[v];
    const float4 cl = R.cl;
    e.cx = cl.x; e.cy = cl.y; e.cz = cl.z; e.l = cl.w;
    e.g0 = R.g0; e.g1 = R.g1;
    e.l_inv = R.aux.x;
  } else {
    e.cx = centers[3 * v]; e.cy = centers[3 * v + 1]; e.cz = centers[3 * v + 2];
    e.l = lengths[v];
    e.g0 = reinterpret_cast<const float4*>(geos)[2 * v];
    e.g1 = reinterpret_cast<const float4*>(geos)[2 * v + 1];
    e.l_inv = 1.f / e.l;
  }
}


// Per-entry constants of the SDF segment as the compositors stage them, once per entry and view.
// Every ray of a view starts at the camera centre o, so the box corners are kept relative to it,
// and the corner SDFs with the differences the z lerps take:
//   box = (lo.x, lo.y, lo.z, 1/l), hi = (hi.x, hi.y, hi.z, 0), lo/hi = (c - o) -/+ l/2;
//   gd0 = (g0, g1 - g0, g2, g3 - g2), gd1 = (g4, g5 - g4, g6, g7 - g6).
// ray_box_st and trilerp_d on these give the same values as ray_box and trilerp on the raw data.
// l_inv is 1.f / l (IEEE quotient), from the caller or the voxel record.
__device__ inline void stage_segment(float cx, float cy, float cz, float l, float l_inv, float4 g0, float4 g1,
                                     const Ray& r, float4& box, float4& hi, float4& gd0, float4& gd1) {
#pragma clang fp contract(off)
  const float h = 0.5f * l;
  const float dx = cx - r.ox, dy = cy - r.oy, dz = cz - r.oz;
  box = make_float4(dx - h, dy - h, dz - h, l_inv);
  hi = make_float4(dx + h, dy + h, dz + h, 0.f);
  gd0 = make_float4(g0.x, g0.y - g0.x, g0.z, g0.w - g0.z);
  gd1 = make_float4(g1.x, g1.y - g1.x, g1.z, g1.w - g1.z);
}

__device__ inline void ray_box_st(const Ray& r, const float4 box, const float4 hi, float& a, float& b) {
  const float x0 = box.x * r.ix, x1 = hi.x * r.ix;
  const float y0 = box.y * r.iy, y1 = hi.y * r.iy;
  const float z0 = box.z * r.iz, z1 = hi.z * r.iz;
  a = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fminf(z0, z1));
  b = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fmaxf(z0, z1));
}

__device__ inline float trilerp_d(const float4 gd0, const float4 gd1, float x, float y, float z) {
#pragma clang fp contract(off)
  const float c00 = fmaf(z, gd0.y, gd0.x);
  const float c01 = fmaf(z, gd0.w, gd0.z);
  const float c10 = fmaf(z, gd1.y, gd1.x);
  const float c11 = fmaf(z, gd1.w, gd1.z);
  const float c0 = fmaf(y, c01 - c00, c00);
  const float c1 = fmaf(y, c11 - c10, c10);
  return fmaf(x, c1 - c0, c0);
}

// The SDF segment of one (ray, voxel) pair: SDF at the entry point and three equal steps to
// the exit point, their logistic CDFs, and the NeuS alpha (forward.cu:255-306).
struct Segment {
  float qx, qy, qz;        // entry point in voxel-local [0,1]^3
  float sx, sy, sz;        // local step between samples
  float sdf[4];
  float phi[4];
  float alpha;
};

__device__ inline void eval_segment(const Ray& r, float a, float b, float minx, float miny,
                                    float minz, float l_inv, const float g[8], float s,
                                    Segment& seg) {
#pragma clang fp contract(off)
  seg.qx = (fmaf(a, r.dx, r.ox) - minx) * l_inv;
  seg.qy = (fmaf(a, r.dy, r.oy) - miny) * l_inv;
  seg.qz = (fmaf(a, r.dz, r.oz) - minz) * l_inv;
  const float st = ((b - a) * (1.f / 3.f)) * l_inv;
  seg.sx = st * r.dx; seg.sy = st * r.dy; seg.sz = st * r.dz;
  const float ks = sdf_cdf_scale(s);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float fk = (float)k;
    seg.sdf[k] = k == 0 ? trilerp(g, seg.qx, seg.qy, seg.qz)
                        : trilerp(g, fmaf(seg.sx, fk, seg.qx), fmaf(seg.sy, fk, seg.qy), fmaf(seg.sz, fk, seg.qz));
    seg.phi[k] = sdf_cdf(seg.sdf[k], ks);
  }
  const float raw = (seg.phi[0] - seg.phi[3]) * __builtin_amdgcn_rcpf(seg.phi[0] + kPhiEps);
  seg.alpha = fmaxf(raw, 0.f);
}

// eval_segment on staged constants: the local coordinate of the entry point is
// (a d - lo) / l, the ray origin having been folded into lo.
__device__ inline void eval_segment_st(const Ray& r, float a, float b, const float4 box, const float4 gd0,
                                       const float4 gd1, float s, Segment& seg) {
#pragma clang fp contract(off)
  seg.qx = fmaf(a, r.dx, -box.x) * box.w;
  seg.qy = fmaf(a, r.dy, -box.y) * box.w;
  seg.qz = fmaf(a, r.dz, -box.z) * box.w;
  const float st = ((b - a) * (1.f / 3.f)) * box.w;
  seg.sx = st * r.dx; seg.sy = st * r.dy; seg.sz = st * r.dz;
  const float ks = sdf_cdf_scale(s);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float fk = (float)k;
    seg.sdf[k] = k == 0 ? trilerp_d(gd0, gd1, seg.qx, seg.qy, seg.qz)
                        : trilerp_d(gd0, gd1, fmaf(seg.sx, fk, seg.qx), fmaf(seg.sy, fk, seg.qy),
                                    fmaf(seg.sz, fk, seg.qz));
    seg.phi[k] = sdf_cdf(seg.sdf[k], ks);
  }
  const float raw = (seg.phi[0] - seg.phi[3]) * __builtin_amdgcn_rcpf(seg.phi[0] + kPhiEps);
  seg.alpha = fmaxf(raw, 0.f);
}

// Value of lane (lane ^ kMask) without going through LDS: DPP for masks 1..8 (quad_perm
// [1,0,3,2] and [2,3,0,1]; half-row mirror then quad reverse for 4; row rotate by 8), and the
// gfx950 row-swap instructions for 16 and 32.
template <int kMask>
__device__ inline uint32_t xor_lane(uint32_t x, int lane) {
  if constexpr (kMask == 1) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, true);
  } else if constexpr (kMask == 2) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, true);
  } else if constexpr (kMask == 4) {
    const int y = __builtin_amdgcn_mov_dpp((int)x, 0x141, 0xF, 0xF, true);
    return (uint32_t)__builtin_amdgcn_mov_dpp(y, 0x1B, 0xF, 0xF, true);
  } else if constexpr (kMask == 8) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x128, 0xF, 0xF, true);
  } else if constexpr (kMask == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return (lane & 16) ? r[0] : r[1];
  } else {
    static_assert(kMask == 32, "xor mask must be a power of two below 64");
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return (lane & 32) ? r[0] : r[1];
  }
}

template <int kSh>
__device__ inline uint64_t transpose_stage(uint64_t x, int lane) {
  const uint32_t plo = xor_lane<kSh>((uint32_t)x, lane);
  const uint32_t phi = xor_lane<kSh>((uint32_t)(x >> 32), lane);
  const uint64_t p = ((uint64_t)phi << 32) | plo;
  constexpr uint64_t m = ~0ull / ((1ull << kSh) + 1ull);   // bit positions whose bit kSh is clear
  return (lane & kSh) ? ((x & ~m) | ((p & ~m) >> kSh)) : ((x & m) | ((p & m) << kSh));
}

// Transposes a 64x64 bit matrix held one row per lane (lane i, bit b = A[i][b]): lane b returns
// column b (bit i = A[i][b]).  Six butterfly stages swap the off-diagonal blocks of size 32..1.
__device__ inline uint64_t transpose64(uint64_t x, int lane) {
  x = transpose_stage<32>(x, lane);
  x = transpose_stage<16>(x, lane);
  x = transpose_stage<8>(x, lane);
  x = transpose_stage<4>(x, lane);
  x = transpose_stage<2>(x, lane);
  return transpose_stage<1>(x, lane);
}

// Inclusive wave64 prefix sum with DPP: row shifts by 1, 2, 4, 8 (zero fill), then the last lane
// of row 0 / 2 broadcast into rows 1 / 3 and the last lane of row 1 into rows 2 and 3.
__device__ inline uint32_t wave_incl_scan_u32(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);
  return x;
}

// Inclusive wave64 prefix maximum (same DPP pattern as wave_incl_scan_u32; 0 is the identity).
__device__ inline uint32_t wave_incl_max_u32(uint32_t x) {
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false));
  x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false));
  return x;
}

// Segmented inclusive sum over the wave of N floats per lane (N a multiple of 4); lane l's
// segment starts at lane seg_start <= l.  After the call the last lane of each segment holds the
// segment's sums.  DPP row shifts by 1, 2, 4, 8 inside 16-lane rows, then the last lane of the
// previous row (rows 1 and 3) and lane 31 (rows 2 and 3) for segments that started in an earlier
// row.  Each step adds the shifted partial sum times 1 or 0 (whether the source lane is in the
// lane's segment): x = fma(dpp(x), c, x) is one v_fmac_f32 with the DPP shift on its first source.
// The compiler does not fold a DPP move into fmac, so the steps are written as v_fmac_f32_dpp
// directly, four values per statement (a value's DPP read follows its previous write by at least
// three other instructions, and each statement starts with two wait states for the write before
// it).  Rows outside row_mask keep their value, as do lanes whose source is out of the row
// (bound_ctrl: source reads 0).  Sums of finite values equal those of a select-and-add form
// (adding +0 leaves a non-zero sum unchanged).
#define SVR_FMAC_DPP4(CTRL, V, Q, C)                                                       \
  asm volatile("s_nop 1\n\t"                                                               \
               "v_fmac_f32_dpp %0, %0, %4 " CTRL "\n\t"                                     \
               "v_fmac_f32_dpp %1, %1, %4 " CTRL "\n\t"                                     \
               "v_fmac_f32_dpp %2, %2, %4 " CTRL "\n\t"                                     \
               "v_fmac_f32_dpp %3, %3, %4 " CTRL                                            \
               : "+v"(V[Q]), "+v"(V[Q + 1]), "+v"(V[Q + 2]), "+v"(V[Q + 3])                  \
               : "v"(C))

template <int N>
__device__ inline void wave_segmented_sum(float (&v)[N], int lane, int seg_start) {
  static_assert(N % 4 == 0, "values are summed in groups of four");
  const int lr = lane & 15;
  const float c1 = (lr >= 1 && lane - 1 >= seg_start) ? 1.f : 0.f;
  const float c2 = (lr >= 2 && lane - 2 >= seg_start) ? 1.f : 0.f;
  const float c4 = (lr >= 4 && lane - 4 >= seg_start) ? 1.f : 0.f;
  const float c8 = (lr >= 8 && lane - 8 >= seg_start) ? 1.f : 0.f;
  const float cr = (((lane >> 4) & 1) && seg_start < (lane & ~15)) ? 1.f : 0.f;
  const float cw = (lane >= 32 && seg_start < 32) ? 1.f : 0.f;
#pragma unroll
  for (int q = 0; q < N; q += 4) SVR_FMAC_DPP4("row_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1", v, q, c1);
#pragma unroll
  for (int q = 0; q < N; q += 4) SVR_FMAC_DPP4("row_shr:2 row_mask:0xf bank_mask:0xf bound_ctrl:1", v, q, c2);
#pragma unroll
  for (int q = 0; q < N; q += 4) SVR_FMAC_DPP4("row_shr:4 row_mask:0xf bank_mask:0xf bound_ctrl:1", v, q, c4);
#pragma unroll
  for (int q = 0; q < N; q += 4) SVR_FMAC_DPP4("row_shr:8 row_mask:0xf bank_mask:0xf bound_ctrl:1", v, q, c8);
#pragma unroll
  for (int q = 0; q < N; q += 4) SVR_FMAC_DPP4("row_bcast:15 row_mask:0xa bank_mask:0xf", v, q, cr);
#pragma unroll
  for (int q = 0; q < N; q += 4) SVR_FMAC_DPP4("row_bcast:31 row_mask:0xc bank_mask:0xf", v, q, cw);
}
#undef SVR_FMAC_DPP4

// Compensated FP32 summation.  two_sum is Knuth's error-free transformation: s = fl(a + b) and
// e = (a + b) - s exactly.  A running (hi, lo) pair adds each term to hi and its rounding error to
// lo (Neumaier), so hi + lo carries the sum of many terms of mixed sign to about one rounding of
// the result instead of one rounding per term.
__device__ inline void two_sum(float a, float b, float& s, float& e) {
#pragma clang fp contract(off)
  s = a + b;
  const float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__device__ inline void comp_add(float& hi, float& lo, float x) {
#pragma clang fp contract(off)
  float s, e;
  two_sum(hi, x, s, e);
  hi = s;
  lo += e;
}

// Sum of the (hi, lo) pairs of the 64 lanes, returned in every lane (xor butterfly; every lane
// forms the same sums, so the result is the same in all lanes).
__device__ inline float2 wave_comp_sum(float hi, float lo) {
#pragma clang fp contract(off)
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const float ohi = __shfl_xor(hi, off, 64);
    const float olo = __shfl_xor(lo, off, 64);
    float s, e;
    two_sum(hi, ohi, s, e);
    hi = s;
    lo = (lo + olo) + e;
  }
  return make_float2(hi, lo);
}

// Position of the k-th (0-based) set bit of m; k < popcount(m).  After picking the 32-bit half,
// the position p is the largest with popcount(x & (2^p - 1)) <= k, found by a binary search on
// the bits of p that counts the bits below each candidate with one bit-field extract (no
// running remainder, no shifts of x).
__device__ inline int select_set_bit(uint64_t m, int k) {
  const uint32_t lo = (uint32_t)m;
  const int c32 = __popc(lo);
  const bool up = k >= c32;
  const uint32_t x = up ? (uint32_t)(m >> 32) : lo;
  const int kk = up ? k - c32 : k;
  int p = (__popc(x & 0xFFFFu) <= kk) ? 16 : 0;
  p = (__popc(__builtin_amdgcn_ubfe(x, 0u, (uint32_t)p + 8u)) <= kk) ? p + 8 : p;
  p = (__popc(__builtin_amdgcn_ubfe(x, 0u, (uint32_t)p + 4u)) <= kk) ? p + 4 : p;
  p = (__popc(__builtin_amdgcn_ubfe(x, 0u, (uint32_t)p + 2u)) <= kk) ? p + 2 : p;
  p = (__popc(__builtin_amdgcn_ubfe(x, 0u, (uint32_t)p + 1u)) <= kk) ? p + 1 : p;
  return up ? p + 32 : p;
}

// Conservative per-voxel test: false only if no ray segment through the voxel can reach
// alpha >= kMinAlpha, so skipping the voxel leaves every pixel unchanged (the compositors skip
// such pairs anyway, forward.cu:303-305, backward.cu:428).
//
// The main alpha is (phi(v0) - phi(v3)) / (phi(v0) + 1e-6) with phi = sigmoid(s * v) and v0, v3
// trilinear SDF values at the segment ends.  For v in [lo, hi] the fraction is at most
// B = (sigmoid(a) - sigmoid(b)) / (sigmoid(a) + 1e-6), a = s * hi, b = s * lo.  [lo, hi] is the
// corner range widened for (i) trilinear extrapolation when rounding puts a sample slightly
// outside the voxel (local coordinate error eta <= 2e-6 (|pos| + l) / l, giving at most 8 eta
// times the range), plus a fixed quarter of the range, and (ii) 1e-6 of the largest |corner| for
// the rounding of the lerps.  The float evaluation of phi (s * v product, exp, rcp) has relative
// error below 2.7e-6, which can add at most 5.4e-6 to the computed alpha, so B < 3e-6 guarantees
// a computed alpha below 8.5e-6 < kMinAlpha.  B is evaluated without cancellation as
//   sigmoid(a) - sigmoid(b) = (1 - e^-(a - b)) sigmoid(a) sigmoid(-b),  a - b = s (hi - lo),
// with expm1 for the first factor, so its relative error stays near 1e-6.  NaN corners or s keep
// the voxel.
__device__ inline float sigmoid_acc(float x) { return 1.f / (1.f + expf(-x)); }

__device__ inline bool voxel_may_contribute(const float g[8], float s, float pos_scale, float l) {
  float gmin = g[0], gmax = g[0], sum = g[0];
#pragma unroll
  for (int k = 1; k < 8; ++k) {
    gmin = fminf(gmin, g[k]);
    gmax = fmaxf(gmax, g[k]);
    sum += g[k];
  }
  if (__builtin_isnan(sum)) return true;
  const float R = gmax - gmin;
  const float eta = 2e-6f * (pos_scale + l) / l;
  const float sl = R * (0.25f + 8.f * eta) + 1e-6f * fmaxf(fabsf(gmin), fabsf(gmax));
  const float a = s * (gmax + sl), b = s * (gmin - sl);
  const float d = s * (R + 2.f * sl);
  const float sa = sigmoid_acc(a);
  const float B = (-expm1f(-d)) * sa * sigmoid_acc(-b) / (sa + 1e-6f);
  return !(B < 3e-6f);
}

// Depth of one segment from its three sub-interval alphas (forward.cu:349-357).
__device__ inline void sub_alphas(const Segment& seg, float sub[3]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float raw = (seg.phi[k] - seg.phi[k + 1]) * __builtin_amdgcn_rcpf(seg.phi[k] + kPhiEps);
    sub[k] = fmaxf(raw, 0.f);
  }
}

__device__ inline float depth_value(const float sub[3], float a, float b) {
  const float step = kThird * (b - a);
  const float t0 = a + 0.5f * step, t1 = a + 1.5f * step, t2 = a + 2.5f * step;
  return sub[0] * t0 + (1.f - sub[0]) * sub[1] * t1 + (1.f - sub[0]) * (1.f - sub[1]) * sub[2] * t2;
}

// Unnormalised "linear" normal from corner sums and its reciprocal norm (forward.cu:422-433).
__device__ inline void lin_normal(const float g[8], float& nx, float& ny, float& nz, float& rn) {
  nx = (g[4] + g[5] + g[6] + g[7]) - (g[0] + g[1] + g[2] + g[3]);
  ny = (g[2] + g[3] + g[6] + g[7]) - (g[0] + g[1] + g[4] + g[5]);
  nz = (g[1] + g[3] + g[5] + g[7]) - (g[0] + g[2] + g[4] + g[6]);
  rn = rsqrtf(nx * nx + ny * ny + nz * nz + 1e-15f);
}

// Tile rectangle of a packed pixel bbox, clamped to the view (auxiliary.h:97-111, there with
// 16x16 tiles).  Tile sizes are powers of two (8, 16 or 32), so the divisions are shifts.
__device__ inline void bbox_tiles(uint32_t b0, uint32_t b1, const ViewDesc& v, int& tx0, int& ty0, int& tx1,
                                  int& ty1) {
  const int sx = __builtin_ctz((unsigned)v.tile_w), sy = __builtin_ctz((unsigned)v.tile_h);
  tx0 = max(0, min(v.tiles_x - 1, int(b0 >> 16) >> sx));
  ty0 = max(0, min(v.tiles_y - 1, int(b0 & 0xFFFFu) >> sy));
  tx1 = max(0, min(v.tiles_x - 1, int(b1 >> 16) >> sx));
  ty1 = max(0, min(v.tiles_y - 1, int(b1 & 0xFFFFu) >> sy));
}

__device__ inline uint64_t rep3(uint32_t q) {
  uint64_t v = 0;
#pragma unroll
  for (int i = 0; i < kMaxLevels; ++i) v = (v << 3) | q;
  return v;
}

// Wave64 helpers.
//
// Sum over the 64 lanes (all lanes must be active), returned in every lane: DPP quad swaps and
// row mirrors give each 16-lane row its sum, then the four row sums are read back as scalars.
__device__ inline float wave_sum_uniform(float x) {
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x4E, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x141, 0xF, 0xF, false));
  x += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x140, 0xF, 0xF, false));
  const int v = __float_as_int(x);
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(v, 0));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(v, 16));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(v, 32));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(v, 48));
  return (r0 + r1) + (r2 + r3);
}

// Minimum over the currently active lanes of the wave (inactive lanes contribute nothing).
__device__ inline uint32_t wave_min_active_u32(uint32_t v) {
  const uint64_t act = __ballot(1);
  uint32_t m = v;
  for (uint64_t b = act; b; b &= b - 1) {
    const int l = __ffsll((long long)b) - 1;
    m = min(m, (uint32_t)__builtin_amdgcn_readlane((int)v, l));
  }
  return m;
}

__device__ inline uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, 64));
  return v;
}

}  // namespace svr

// Camera utilities of the reference's utils module: per-pixel world ray directions, depth map to
// world points, the view-cone test of points and the eight per-octant rank codes of a voxel
// (cuda/src/utils.cu:23-127 kernels, 254-327 hosts).  All are elementwise and HBM-bound: one
// thread per pixel / point / voxel, planar outputs written with unit-stride stores.
#include "common.h"
#include "ops.h"

namespace svr {

namespace {

// The 3x4 part of a row-major 4x4 matrix in device memory, read once per workgroup into LDS.
struct Mat34 { float m[12]; };

__device__ inline void load_mat34(const float* __restrict__ src, float* dst) {
  if (threadIdx.x < 12) dst[threadIdx.x] = src[threadIdx.x];
  __syncthreads();
}

// Camera-space direction of pixel p of a W x H image (auxiliary.h:87-95, utils.cu:60-64) and
// its rotation by c2w (auxiliary.h:235-243), products summed left to right without contraction.
__device__ inline void pixel_world_dir(int p, int W, int H, float cx, float cy, float tfx, float tfy,
                                       const float* c2w, float& dx, float& dy, float& dz) {
#pragma clang fp contract(off)
  const float yid = (float)(p / W);
  const float xid = (float)(p % W);
  const float x = (xid + 0.5f - cx) * 2.f * tfx / (float)W;
  const float y = (yid + 0.5f - cy) * 2.f * tfy / (float)H;
  dx = c2w[0] * x + c2w[1] * y + c2w[2] * 1.f;
  dy = c2w[4] * x + c2w[5] * y + c2w[6] * 1.f;
  dz = c2w[8] * x + c2w[9] * y + c2w[10] * 1.f;
}

__global__ __launch_bounds__(256) void compute_rd_kernel(int W, int H, float cx, float cy, float tfx, float tfy,
                                                         const float* __restrict__ c2w_in, float* __restrict__ rd) {
  __shared__ float c2w[12];
  load_mat34(c2w_in, c2w);
  const long long n = (long long)W * H;
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  float dx, dy, dz;
  pixel_world_dir((int)p, W, H, cx, cy, tfx, tfy, c2w, dx, dy, dz);
  rd[p] = dx;
  rd[n + p] = dy;
  rd[2 * n + p] = dz;
}

__global__ __launch_bounds__(256) void depth2pts_kernel(int W, int H, float cx, float cy, float tfx, float tfy,
                                                        const float* __restrict__ c2w_in,
                                                        const float* __restrict__ depth, float* __restrict__ pts) {
#pragma clang fp contract(off)
  __shared__ float c2w[12];
  load_mat34(c2w_in, c2w);
  const long long n = (long long)W * H;
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  float dx, dy, dz;
  pixel_world_dir((int)p, W, H, cx, cy, tfx, tfy, c2w, dx, dy, dz);
  const float d = depth[p];
  pts[p] = c2w[3] + d * dx;
  pts[n + p] = c2w[7] + d * dy;
  pts[2 * n + p] = c2w[11] + d * dz;
}

// utils.cu:23-43: camera-space point in front of the near plane and inside the frustum cone.
__global__ __launch_bounds__(256) void is_in_cone_kernel(long long N, float tfx, float tfy, float near_plane,
                                                         const float* __restrict__ w2c_in,
                                                         const float* __restrict__ pts, uint8_t* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ float w2c[12];
  load_mat34(w2c_in, w2c);
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
  const float cx = w2c[0] * x + w2c[1] * y + w2c[2] * z + w2c[3];
  const float cy = w2c[4] * x + w2c[5] * y + w2c[6] * z + w2c[7];
  const float cz = w2c[8] * x + w2c[9] * y + w2c[10] * z + w2c[11];
  out[i] = (cz > near_plane) && (fabsf(cx) <= tfx * cz) && (fabsf(cy) <= tfy * cz);
}

// utils.cu:107-127: rank of a voxel for each of the eight ray octants, [8, P] (octant-major).
__global__ __launch_bounds__(256) void voxel_order_rank_kernel(int P, const int64_t* __restrict__ paths,
                                                               int64_t* __restrict__ out) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= P) return;
  const uint64_t path = (uint64_t)paths[i];
#pragma unroll
  for (uint32_t q = 0; q < 8; ++q) out[(long long)q * P + i] = (int64_t)(path ^ rep3(q));
}

// Device check of rcp_rn against the IEEE division 1.0f / x: every significand of the
// exponents [e_lo, e_lo + gridDim.y), both signs; mismatches counted into out[0].
__global__ __launch_bounds__(256) void selfcheck_rcp_rn_kernel(int e_lo, unsigned long long* __restrict__ out) {
  const uint32_t m = blockIdx.x * 256u + threadIdx.x;
  const uint32_t bits = ((uint32_t)(e_lo + (int)blockIdx.y + 127) << 23) | m;
  uint32_t bad = 0;
#pragma unroll
  for (int sgn = 0; sgn < 2; ++sgn) {
    const float x = __uint_as_float(bits | ((uint32_t)sgn << 31));
    bad += (uint32_t)(__float_as_uint(rcp_rn(x)) != __float_as_uint(1.0f / x));
  }
  if (bad != 0u) atomicAdd(out, (unsigned long long)bad);
}

// Operands of div_rn's self-check: y with exponent in [-30, 0] and x with exponent in [-60, 0],
// random significands (a hash of the sample index and the seed), every 4096th x zero.
__device__ inline uint32_t mix32(uint32_t h) {
  h ^= h >> 16; h *= 0x7FEB352Du; h ^= h >> 15; h *= 0x846CA68Bu; h ^= h >> 16;
  return h;
}

__global__ __launch_bounds__(256) void selfcheck_div_rn_kernel(uint32_t seed, unsigned long long* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const uint32_t h0 = mix32(i ^ mix32(seed)), h1 = mix32(h0 + 0x9E3779B9u), h2 = mix32(h1 + 0x9E3779B9u);
  const float y = __uint_as_float(((127u - (h0 >> 27) % 31u) << 23) | (h1 & 0x7FFFFFu));
  const float x = (i & 4095u) == 0u ? 0.f : __uint_as_float(((127u - (h2 >> 26) % 61u) << 23) | (h2 & 0x7FFFFFu));
  const uint32_t bad = (uint32_t)(__float_as_uint(div_rn(x, y)) != __float_as_uint(x / y));
  if (bad != 0u) atomicAdd(out, (unsigned long long)bad);
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_selfcheck_rcp_rn(int e_lo, int e_hi, unsigned long long* out, hipStream_t st) {
  selfcheck_rcp_rn_kernel<<<dim3((1u << 23) / 256u, (unsigned)(e_hi - e_lo + 1)), 256, 0, st>>>(e_lo, out);
  return hipGetLastError();
}

hipError_t launch_selfcheck_div_rn(int n_blocks, uint32_t seed, unsigned long long* out, hipStream_t st) {
  selfcheck_div_rn_kernel<<<(unsigned)n_blocks, 256, 0, st>>>(seed, out);
  return hipGetLastError();
}

hipError_t launch_compute_rd(int W, int H, float cx, float cy, float tfx, float tfy, const float* c2w, float* rd,
                             hipStream_t st) {
  const long long n = (long long)W * H;
  if (n == 0) return hipSuccess;
  compute_rd_kernel<<<blocks_for(n), 256, 0, st>>>(W, H, cx, cy, tfx, tfy, c2w, rd);
  return hipGetLastError();
}

hipError_t launch_depth2pts(int W, int H, float cx, float cy, float tfx, float tfy, const float* c2w,
                            const float* depth, float* pts, hipStream_t st) {
  const long long n = (long long)W * H;
  if (n == 0) return hipSuccess;
  depth2pts_kernel<<<blocks_for(n), 256, 0, st>>>(W, H, cx, cy, tfx, tfy, c2w, depth, pts);
  return hipGetLastError();
}

hipError_t launch_is_in_cone(long long N, float tfx, float tfy, float near_plane, const float* w2c,
                             const float* pts, uint8_t* out, hipStream_t st) {
  if (N == 0) return hipSuccess;
  is_in_cone_kernel<<<blocks_for(N), 256, 0, st>>>(N, tfx, tfy, near_plane, w2c, pts, out);
  return hipGetLastError();
}

hipError_t launch_voxel_order_rank(int P, const int64_t* paths, int64_t* out, hipStream_t st) {
  if (P == 0) return hipSuccess;
  voxel_order_rank_kernel<<<blocks_for(P), 256, 0, st>>>(P, paths, out);
  return hipGetLastError();
}

}  // namespace svr

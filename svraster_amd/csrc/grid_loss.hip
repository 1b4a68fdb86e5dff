// Grid smoothness backward on gfx950: eikonal and Laplacian losses sampled on the dense
// valid-gradient grid table, plus the table build itself.
//
// References:
//   grid eikonal   cuda/src/ge_compute.cu:14-117 (host 120-161)
//   grid Laplacian cuda/src/ls_compute.cu:14-109 (host 112-155)
//   table helpers  cuda/src/auxiliary.h:360-455 (neighbour validity, key search, sampling,
//                  corner scatter)
//   table build    cuda/src/utils.cu:190-252, 362-408
//
// Each active grid key samples the SDF at the cell (Laplacian) and its six face neighbours,
// all offset by one random vector per key, by locating the covering voxel with a binary search
// over the sorted key table and interpolating its 8 corners; the loss gradient is scattered to
// those corners with atomics.  The reference seeds the offset from the clock; here it comes
// from a seeded PCG hash of (seed, key index) or from an explicit per-key array, so results are
// reproducible.
#include "common.h"
#include "ops.h"

namespace svr {

__device__ inline uint32_t pcg_hash(uint32_t x) {
  const uint32_t state = x * 747796405u + 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

__device__ inline float hash_unit(uint32_t x) { return (float)(pcg_hash(x) >> 8) * (1.0f / 16777216.0f); }

__device__ inline void random_offset(const float* randv, uint32_t seed, int tid, float& rx, float& ry, float& rz) {
  if (randv) {
    rx = randv[3 * tid]; ry = randv[3 * tid + 1]; rz = randv[3 * tid + 2];
  } else {
    const uint32_t b = pcg_hash(seed) ^ (uint32_t)tid * 3u;
    rx = hash_unit(b); ry = hash_unit(b + 0x9E3779B9u); rz = hash_unit(b + 0x3C6EF372u);
  }
}

__device__ inline bool neighbours_valid(int x, int y, int z, int res, const bool* __restrict__ mask) {
  if (x <= 0 || x >= res - 1 || y <= 0 || y >= res - 1 || z <= 0 || z >= res - 1) return false;
  const long long rr = (long long)res * res;
  const long long k = x + (long long)y * res + z * rr;
  return mask[k + 1] && mask[k - 1] && mask[k + res] && mask[k - res] && mask[k + rr] && mask[k - rr];
}

// Same midpoint rule as the reference's search, so duplicate keys resolve to the same entry.
__device__ inline int key_search(const int32_t* __restrict__ keys, int size, int key) {
  int l = 0, r = size - 1;
  while (l <= r) {
    const int m = (l + r) / 2;
    const int km = keys[m];
    if (km == key) return m;
    if (km < key) l = m + 1; else r = m - 1;
  }
  return -1;
}



// SDF at grid point (gx+rx, gy+ry, gz+rz); returns voxel id (-1 when not covered) and weights.
__device__ inline float grid_sample(const GridTable& G, int gx, int gy, int gz, float rx, float ry, float rz,
                                    float w[8], int& vox) {
  const int key = gx + gy * G.res + gz * G.res * G.res;
  const int idx = G.dense ? G.dense[key] : key_search(G.keys, G.n_keys, key);
  vox = -1;
  if (idx < 0) return 0.f;
  const int v = G.key2vox[idx];
  if (v < 0 || v >= G.n_vox) return 0.f;
  vox = v;
  const float size = G.vox_size[v];
  const float qx = ((float)gx + rx - G.vox_coord[3 * v]) / size;
  const float qy = ((float)gy + ry - G.vox_coord[3 * v + 1]) / size;
  const float qz = ((float)gz + rz - G.vox_coord[3 * v + 2]) / size;
  trilinear_weights(qx, qy, qz, w);
  float sdf = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    w[i] = fminf(fmaxf(w[i], 0.f), 1.f);
    sdf += G.grid_pts[G.vox_key[(long long)v * 8 + i]] * w[i];
  }
  return sdf;
}

__device__ inline void grid_scatter(const GridTable& G, int vox, const float w[8], float g, float* __restrict__ out) {
  if (vox < 0) return;
#pragma unroll
  for (int i = 0; i < 8; ++i) atomicAdd(out + G.vox_key[(long long)vox * 8 + i], g * w[i]);
}

__global__ __launch_bounds__(256) void grid_eikonal_kernel(GridTable G, const bool* __restrict__ mask,
                                                           const int32_t* __restrict__ active, int A,
                                                           float weight, float inv, uint32_t seed,
                                                           const int32_t* __restrict__ step,
                                                           const float* __restrict__ randv,
                                                           float* __restrict__ grad) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= A) return;
  const int gk = G.keys[active[tid]];
  const int x = gk % G.res, y = (gk / G.res) % G.res, z = gk / (G.res * G.res);
  if (!mask[gk] || !neighbours_valid(x, y, z, G.res, mask)) return;
  float rx, ry, rz;
  random_offset(randv, step ? seed + 0x9E3779B9u * (uint32_t)*step : seed, tid, rx, ry, rz);
  const int dx[6] = {1, -1, 0, 0, 0, 0}, dy[6] = {0, 0, 1, -1, 0, 0}, dz[6] = {0, 0, 0, 0, 1, -1};
  float w[6][8], sd[6];
  int vox[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) sd[i] = grid_sample(G, x + dx[i], y + dy[i], z + dz[i], rx, ry, rz, w[i], vox[i]);
  const float gx = (sd[0] - sd[1]) * 0.5f, gy = (sd[2] - sd[3]) * 0.5f, gz = (sd[4] - sd[5]) * 0.5f;
  const float gn = sqrtf(gx * gx + gy * gy + gz * gz + 1e-6f);
  const float gw = gn * 0.5f * inv;
  const float dLdg = 2.f * (gw - 1.f) * weight;
  const float lx = dLdg * gx / gn * 0.5f * inv;
  const float ly = dLdg * gy / gn * 0.5f * inv;
  const float lz = dLdg * gz / gn * 0.5f * inv;
  grid_scatter(G, vox[0], w[0], lx, grad);
  grid_scatter(G, vox[1], w[1], -lx, grad);
  grid_scatter(G, vox[2], w[2], ly, grad);
  grid_scatter(G, vox[3], w[3], -ly, grad);
  grid_scatter(G, vox[4], w[4], lz, grad);
  grid_scatter(G, vox[5], w[5], -lz, grad);
}

// One grid key of the Laplacian loss (ls_compute.cu:14-109): samples at the cell and its six
// face neighbours, all shifted by the key's random offset; returns false when the key or a
// neighbour is outside the valid mask.  lx, ly, lz are the per-axis loss coefficients: sample
// +-x receives lx, +-y ly, +-z lz and the centre -2 (lx + ly + lz), each times its trilinear
// weights.
__device__ inline bool laplacian_key(const GridTable& G, const bool* __restrict__ mask, int gk, int tid,
                                     uint32_t seed, const float* __restrict__ randv, float weight, float inv,
                                     float (&w)[7][8], int (&vox)[7], float& lx, float& ly, float& lz) {
  const int x = gk % G.res, y = (gk / G.res) % G.res, z = gk / (G.res * G.res);
  if (!mask[gk] || !neighbours_valid(x, y, z, G.res, mask)) return false;
  float rx, ry, rz;
  random_offset(randv, seed, tid, rx, ry, rz);
  const int dx[7] = {0, 1, -1, 0, 0, 0, 0}, dy[7] = {0, 0, 0, 1, -1, 0, 0}, dz[7] = {0, 0, 0, 0, 0, 1, -1};
  float sd[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) sd[i] = grid_sample(G, x + dx[i], y + dy[i], z + dz[i], rx, ry, rz, w[i], vox[i]);
  const float r0 = sd[1] + sd[2] - 2.f * sd[0];
  const float r1 = sd[3] + sd[4] - 2.f * sd[0];
  const float r2 = sd[5] + sd[6] - 2.f * sd[0];
  const float lr = sqrtf(r0 * r0 + r1 * r1 + r2 * r2 + 1e-10f);
  const float sz = weight * 0.5f * inv * inv;
  lx = r0 / lr * sz;
  ly = r1 / lr * sz;
  lz = r2 / lr * sz;
  return true;
}

__device__ inline uint32_t step_seed(uint32_t seed, const int32_t* __restrict__ step) {
  return step ? seed + 0x9E3779B9u * (uint32_t)*step : seed;
}

__global__ __launch_bounds__(256) void grid_laplacian_kernel(GridTable G, const bool* __restrict__ mask,
                                                             const int32_t* __restrict__ active, int A,
                                                             float weight, float inv, uint32_t seed,
                                                             const int32_t* __restrict__ step,
                                                             const float* __restrict__ randv,
                                                             float* __restrict__ grad) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= A) return;
  float w[7][8], lx, ly, lz;
  int vox[7];
  if (!laplacian_key(G, mask, G.keys[active[tid]], tid, step_seed(seed, step), randv, weight, inv, w, vox, lx, ly,
                     lz))
    return;
  const float lc = -2.f * (lx + ly + lz);
  grid_scatter(G, vox[0], w[0], lc, grad);
  grid_scatter(G, vox[1], w[1], lx, grad);
  grid_scatter(G, vox[2], w[2], lx, grad);
  grid_scatter(G, vox[3], w[3], ly, grad);
  grid_scatter(G, vox[4], w[4], ly, grad);
  grid_scatter(G, vox[5], w[5], lz, grad);
  grid_scatter(G, vox[6], w[6], lz, grad);
}

// Two-pass ("pull") form of the same loss -- and of the eikonal loss, alone or with it -- with
// every table entry active once, taken in table order (voxel by voxel; entry e of a voxel with cell box lo..lo+s is cell
// lo + (i % s, i / s % s, i / s^2), i = e - first entry of the voxel, x fastest).  It needs a
// table whose voxels' entries are exactly their cell boxes in that order with power-of-two
// sizes, whose cells are unique, and whose mask is its own (true exactly on its cells).  The
// random offset of the key at entry e is drawn with index e.
//
// The table's voxels are numbered compactly (t = 0..T-1, tvox[t] = voxel id); per step, a record
// per table voxel holds its box origin and size and its 8 corner SDFs (12 floats), so a sample
// costs one cell -> voxel lookup and three 16-byte loads.  Both passes compute every lookup
// address first and issue the loads together (no load waits on a branch), since the time is
// in load latency, not arithmetic.
//
// Pass 1, one lane per entry: the key's coefficients (zero when it fails the mask test).
// Samples inside the lane's own voxel box use its own record, and such neighbours are in the
// mask by construction; samples across a face read the mask and their voxel's record.
// Pass 2, one lane per entry: the samples that land in cell c come from key c (centre sample)
// and from its six face neighbours (their sample toward c), all inside c's voxel, so the lane
// sums the seven weighted contributions into that voxel's 8 corners; lanes of one voxel are
// summed across the wavefront and one lane per voxel run adds the 8 sums to the gradient.  Same
// terms as the per-key scatter (the same products of coefficient and weight), summed in another
// order.
struct VoxelBox {
  int lx, ly, lz, s;
  float cx, cy, cz, inv;
};

// Power-of-two sizes: 1 / size is exact and t / size == t * (1 / size).
__device__ inline VoxelBox record_box(float4 r) {
  VoxelBox b;
  b.cx = r.x; b.cy = r.y; b.cz = r.z;
  b.lx = (int)r.x; b.ly = (int)r.y; b.lz = (int)r.z; b.s = (int)r.w;
  b.inv = __builtin_amdgcn_rcpf(r.w);
  return b;
}

// One unsigned compare per axis, combined without short-circuit branches.
__device__ inline bool in_box(const VoxelBox& b, int x, int y, int z) {
  const unsigned s = (unsigned)b.s;
  return ((unsigned)(x - b.lx) < s) & ((unsigned)(y - b.ly) < s) & ((unsigned)(z - b.lz) < s);
}

// Local coordinates of cell (x, y, z) + r in the voxel box (the reference divides by the size).
__device__ inline void box_local(const VoxelBox& b, int x, int y, int z, float rx, float ry, float rz, float& qx,
                                 float& qy, float& qz) {
  qx = ((float)x + rx - b.cx) * b.inv;
  qy = ((float)y + ry - b.cy) * b.inv;
  qz = ((float)z + rz - b.cz) * b.inv;
}

// SDF of the voxel (box b, corner SDFs g0, g1) at cell (x, y, z) + r: clamped trilinear weights
// and the weighted corner sum, as grid_sample computes them.
__device__ inline float record_sample(const VoxelBox& b, float4 g0, float4 g1, int x, int y, int z, float rx,
                                      float ry, float rz) {
  float qx, qy, qz, w[8];
  box_local(b, x, y, z, rx, ry, rz, qx, qy, qz);
  trilinear_weights(qx, qy, qz, w);
  const float g[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
  float sdf = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) sdf += g[i] * fminf(fmaxf(w[i], 0.f), 1.f);
  return sdf;
}

__global__ __launch_bounds__(256) void grid_voxel_records_kernel(int T, const int32_t* __restrict__ tvox,
                                                                 const float* __restrict__ coord,
                                                                 const float* __restrict__ size,
                                                                 const int64_t* __restrict__ vox_key,
                                                                 const float* __restrict__ grid_pts,
                                                                 float4* __restrict__ rec) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const long long v = tvox[t];
  float g[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) g[i] = grid_pts[vox_key[v * 8 + i]];
  rec[3LL * t] = make_float4(coord[3 * v], coord[3 * v + 1], coord[3 * v + 2], size[v]);
  rec[3LL * t + 1] = make_float4(g[0], g[1], g[2], g[3]);
  rec[3LL * t + 2] = make_float4(g[4], g[5], g[6], g[7]);
}

struct TableOrder {
  const int32_t* keys;       // [G] cell of each entry
  const int32_t* tv;         // [G] compact table voxel of each entry
  const int32_t* cell_entry; // [res^3] entry of each cell, -1 if none
  const int32_t* cell_tv;    // [res^3] compact table voxel of each cell, -1 if none
  const int32_t* tvox;       // [T] voxel id of each compact table voxel
  const float4* rec;         // [T][3] box + corner SDF records
  int G, T, res, lg;         // lg = log2(res): the table's resolution is a power of two
};

// Pass 1 for the enabled terms.  Both terms sample the six face neighbours of the key at the
// key's random offset (the Laplacian also the key's own cell), so one set of samples serves both.
// Laplacian only: three coefficients per entry (lx, ly, lz), coef[k * G + e].  With the eikonal:
// seven, the coefficient of each sample of the entry -- centre -2 (lx + ly + lz), then +x, -x,
// +y, -y, +z, -z: lx +- ex, ly +- ey, lz +- ez, where e = 2 (|g| / (2 h) - 1) w_ge g / |g| / (2 h)
// and g the central differences of the six samples (ge_compute.cu:76-114).  Term weights come from
// wts[0] (eikonal) and wts[1] (Laplacian) in device memory when wts is given, so a captured
// graph follows a weight schedule without being captured again.
template <bool kLap, bool kEik>
__global__ __launch_bounds__(256) void grid_smooth_coef_kernel(TableOrder O, const bool* __restrict__ mask,
                                                               const float* __restrict__ wts, float w_lap,
                                                               float inv, uint32_t seed,
                                                               const int32_t* __restrict__ step,
                                                               float* __restrict__ coef) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = O.G, res = O.res;
  if (e >= n) return;
  const int c = O.keys[e];
  const int t = O.tv[e];
  const int x = c & (res - 1), y = (c >> O.lg) & (res - 1), z = c >> (2 * O.lg);
  const unsigned inner = (unsigned)(res - 2);   // x, y, z in 1 .. res-2
  bool valid = ((unsigned)t < (unsigned)O.T) & ((unsigned)(x - 1) < inner) & ((unsigned)(y - 1) < inner) &
               ((unsigned)(z - 1) < inner);
  const int ts = valid ? t : 0;
  const float4 r0 = O.rec[3LL * ts], g0 = O.rec[3LL * ts + 1], g1 = O.rec[3LL * ts + 2];
  const VoxelBox b = record_box(r0);
  const int dx[7] = {0, 1, -1, 0, 0, 0, 0}, dy[7] = {0, 0, 0, 1, -1, 0, 0}, dz[7] = {0, 0, 0, 0, 0, 1, -1};
  // neighbour cells (the lane's own cell when the key is invalid, so every address is in range)
  int cell[7], tq[7];
  bool ob[7];
  cell[0] = c; ob[0] = false;
#pragma unroll
  for (int i = 1; i < 7; ++i) {
    const int qx = x + dx[i], qy = y + dy[i], qz = z + dz[i];
    ob[i] = valid & !in_box(b, qx, qy, qz);
    cell[i] = valid ? qx + (qy << O.lg) + (qz << (2 * O.lg)) : c;
  }
  bool m[7];
#pragma unroll
  for (int i = 1; i < 7; ++i) {
    m[i] = mask[cell[i]];
    tq[i] = O.cell_tv[cell[i]];
  }
#pragma unroll
  for (int i = 1; i < 7; ++i) {
    valid = valid & (!ob[i] | m[i]);
    tq[i] = ob[i] ? tq[i] : ts;
  }
  float4 qr[7], qg0[7], qg1[7];
#pragma unroll
  for (int i = 1; i < 7; ++i) {
    const int tt = (unsigned)tq[i] < (unsigned)O.T ? tq[i] : ts;
    qr[i] = O.rec[3LL * tt];
    qg0[i] = O.rec[3LL * tt + 1];
    qg1[i] = O.rec[3LL * tt + 2];
  }
  const float wl = wts ? wts[1] : w_lap;
  const float we = kEik ? (wts ? wts[0] : 0.f) : 0.f;
  float lx = 0.f, ly = 0.f, lz = 0.f, ex = 0.f, ey = 0.f, ez = 0.f;
  if (valid) {
    float rx, ry, rz;
    random_offset(nullptr, step_seed(seed, step), e, rx, ry, rz);
    float sd[7];
    sd[0] = kLap ? record_sample(b, g0, g1, x, y, z, rx, ry, rz) : 0.f;
#pragma unroll
    for (int i = 1; i < 7; ++i) {
      // an in-box neighbour has tq[i] == ts, so its record is the key's own
      const int qx = x + dx[i], qy = y + dy[i], qz = z + dz[i];
      const float v = record_sample(record_box(qr[i]), qg0[i], qg1[i], qx, qy, qz, rx, ry, rz);
      sd[i] = (unsigned)tq[i] < (unsigned)O.T ? v : 0.f;
    }
    if (kLap) {
      const float r0s = sd[1] + sd[2] - 2.f * sd[0];
      const float r1s = sd[3] + sd[4] - 2.f * sd[0];
      const float r2s = sd[5] + sd[6] - 2.f * sd[0];
      const float lr = sqrtf(r0s * r0s + r1s * r1s + r2s * r2s + 1e-10f);
      const float sz = wl * 0.5f * inv * inv;
      lx = r0s / lr * sz;
      ly = r1s / lr * sz;
      lz = r2s / lr * sz;
    }
    if (kEik) {
      const float gx = (sd[1] - sd[2]) * 0.5f, gy = (sd[3] - sd[4]) * 0.5f, gz = (sd[5] - sd[6]) * 0.5f;
      const float gn = sqrtf(gx * gx + gy * gy + gz * gz + 1e-6f);
      const float gw = gn * 0.5f * inv;
      const float dLdg = 2.f * (gw - 1.f) * we;
      ex = dLdg * gx / gn * 0.5f * inv;
      ey = dLdg * gy / gn * 0.5f * inv;
      ez = dLdg * gz / gn * 0.5f * inv;
    }
  }
  if (!kEik) {
    coef[e] = lx;
    coef[(long long)n + e] = ly;
    coef[2LL * n + e] = lz;
  } else {
    coef[e] = -2.f * (lx + ly + lz);
    coef[(long long)n + e] = lx + ex;
    coef[2LL * n + e] = lx - ex;
    coef[3LL * n + e] = ly + ey;
    coef[4LL * n + e] = ly - ey;
    coef[5LL * n + e] = lz + ez;
    coef[6LL * n + e] = lz - ez;
  }
}

template <bool kEik>
__global__ __launch_bounds__(256) void grid_laplacian_pull_kernel(TableOrder O, const int64_t* __restrict__ vox_key,
                                                                  uint32_t seed, const int32_t* __restrict__ step,
                                                                  const float* __restrict__ coef,
                                                                  float* __restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int n = O.G, res = O.res;
  const long long n3 = (long long)n;
  const bool live = e < n;
  const int es = live ? e : 0;
  const int c = O.keys[es];
  const int t0 = O.tv[es];
  const int t = (live & ((unsigned)t0 < (unsigned)O.T)) ? t0 : -1;
  const float4 r0 = O.rec[3LL * (t >= 0 ? t : 0)];
  const VoxelBox b = record_box(r0);
  const int x = c & (res - 1), y = (c >> O.lg) & (res - 1), z = c >> (2 * O.lg);
  // the cell, then the keys whose +x, -x, +y, -y, +z, -z sample lands in it
  const int kx[7] = {0, -1, 1, 0, 0, 0, 0}, ky[7] = {0, 0, 0, -1, 1, 0, 0}, kz[7] = {0, 0, 0, 0, 0, -1, 1};
  int ek[7];
  bool use[7];
  int cell[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    const int ax = x + kx[i], ay = y + ky[i], az = z + kz[i];
    const bool ingrid = ((ax | ay | az) & -res) == 0;   // res is a power of two
    const bool inb = in_box(b, ax, ay, az);
    use[i] = (t >= 0) & ingrid;
    ek[i] = inb ? es + (kz[i] * b.s + ky[i]) * b.s + kx[i] : -1;
    cell[i] = (!inb & ingrid) ? ax + (ay << O.lg) + (az << (2 * O.lg)) : c;
  }
  int looked[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) looked[i] = O.cell_entry[cell[i]];
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    if (ek[i] < 0) ek[i] = looked[i];
    use[i] = use[i] & (ek[i] >= 0);
    if (!use[i]) ek[i] = es;
  }
  float l[7];
  if (kEik) {
    // per-sample coefficients: the key at ek[i] contributes its sample toward this cell, which
    // is its sample i (centre, +x seen from c - x, -x seen from c + x, ...)
#pragma unroll
    for (int i = 0; i < 7; ++i) l[i] = coef[i * n3 + ek[i]];
  } else {
    const float cx0 = coef[ek[0]], cy0 = coef[n3 + ek[0]], cz0 = coef[2 * n3 + ek[0]];
    l[0] = -2.f * (cx0 + cy0 + cz0);
    l[1] = coef[ek[1]];
    l[2] = coef[ek[2]];
    l[3] = coef[n3 + ek[3]];
    l[4] = coef[n3 + ek[4]];
    l[5] = coef[2 * n3 + ek[5]];
    l[6] = coef[2 * n3 + ek[6]];
  }
  float acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = 0.f;
  const uint32_t sd = step_seed(seed, step);
#pragma unroll
  for (int i = 0; i < 7; ++i) {
    if (!use[i] || l[i] == 0.f) continue;
    float rx, ry, rz;
    random_offset(nullptr, sd, ek[i], rx, ry, rz);
    // The sample lies in this cell, inside the voxel's own box, so its local coordinates are in
    // [0, 1] and the trilinear weights need no clamping; l is folded into the xy products.
    float qx, qy, qz;
    box_local(b, x, y, z, rx, ry, rz, qx, qy, qz);
    const float lx0 = l[i] * (1.f - qx), lx1 = l[i] * qx;
    const float a00 = lx0 * (1.f - qy), a01 = lx0 * qy, a10 = lx1 * (1.f - qy), a11 = lx1 * qy;
    const float z0 = 1.f - qz;
    acc[0] = fmaf(a00, z0, acc[0]); acc[1] = fmaf(a00, qz, acc[1]);
    acc[2] = fmaf(a01, z0, acc[2]); acc[3] = fmaf(a01, qz, acc[3]);
    acc[4] = fmaf(a10, z0, acc[4]); acc[5] = fmaf(a10, qz, acc[5]);
    acc[6] = fmaf(a11, z0, acc[6]); acc[7] = fmaf(a11, qz, acc[7]);
  }
  const int v = t >= 0 ? t : -1;
  const int prev = __shfl_up(v, 1);
  const int next = __shfl_down(v, 1);
  const bool head = lane == 0 || prev != v;
  const int seg_start = (int)wave_incl_max_u32(head ? (uint32_t)lane : 0u);
  wave_segmented_sum<8>(acc, lane, seg_start);
  const bool tail = lane == 63 || next != v;
  if (tail && v >= 0) {
    const long long vid = O.tvox[v];
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (acc[q] != 0.f) atomicAdd(grad + vox_key[vid * 8 + q], acc[q]);
  }
}

// Valid-gradient table: voxels (leaves with >= 0.9 cells, non-leaves with 0.9..1.5 cells) whose
// centre lies inside the region are rasterised into the res^3 grid.  Two passes (count, then
// write at the scanned offset) make the key order deterministic: voxel order, then x, y, z.
__device__ inline bool table_cell_range(int i, const float* __restrict__ center, const float* __restrict__ size,
                                        const float* __restrict__ scene_center, float extent, int res,
                                        const bool* __restrict__ is_leaf, int lo[3], int hi[3]) {
#pragma clang fp contract(off)
  const float cells = (size[i] / extent) * (float)res;
  const bool leaf = is_leaf[i];
  if (!leaf && (cells > 1.5f || cells < 0.9f)) return false;
  if (leaf && cells < 0.9f) return false;
  for (int d = 0; d < 3; ++d) {
    const float mn = scene_center[d] - 0.5f * extent, mx = scene_center[d] + 0.5f * extent;
    const float c = center[3 * i + d];
    if (c < mn || c > mx) return false;
  }
  for (int d = 0; d < 3; ++d) {
    const float mn = scene_center[d] - 0.5f * extent, mx = scene_center[d] + 0.5f * extent;
    const float c = center[3 * i + d];
    const float a = c - 0.5f * size[i], b = c + 0.5f * size[i];
    lo[d] = (int)roundf((a - mn) / (mx - mn) * (float)res);
    hi[d] = (int)roundf((b - mn) / (mx - mn) * (float)res);
  }
  return true;
}

__global__ __launch_bounds__(256) void table_count_kernel(int N, const float* __restrict__ center,
                                                          const float* __restrict__ size,
                                                          const float* __restrict__ scene_center, float extent,
                                                          int res, const bool* __restrict__ is_leaf,
                                                          uint32_t* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int lo[3], hi[3];
  uint32_t c = 0;
  if (table_cell_range(i, center, size, scene_center, extent, res, is_leaf, lo, hi)) {
    const int nx = max(hi[0] - lo[0], 0), ny = max(hi[1] - lo[1], 0), nz = max(hi[2] - lo[2], 0);
    c = (uint32_t)(nx * ny * nz);
  }
  count[i] = c;
}

__global__ __launch_bounds__(256) void table_write_kernel(int N, const float* __restrict__ center,
                                                          const float* __restrict__ size,
                                                          const float* __restrict__ scene_center, float extent,
                                                          int res, const bool* __restrict__ is_leaf,
                                                          const uint32_t* __restrict__ scan, bool* __restrict__ mask,
                                                          int32_t* __restrict__ keys, int32_t* __restrict__ key2vox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int lo[3], hi[3];
  if (!table_cell_range(i, center, size, scene_center, extent, res, is_leaf, lo, hi)) return;
  uint32_t off = i ? scan[i - 1] : 0u;
  for (int x = lo[0]; x < hi[0]; ++x)
    for (int y = lo[1]; y < hi[1]; ++y)
      for (int z = lo[2]; z < hi[2]; ++z) {
        const int k = x + res * (y + res * z);
        mask[k] = true;
        keys[off] = k;
        key2vox[off] = i;
        ++off;
      }
}

// ------------------------------------------------------------------------------------------
hipError_t launch_grid_loss(bool laplacian, const GridTable& G, const bool* mask, const int32_t* active, int A,
                            float weight, float inv, uint32_t seed, const int32_t* step, const float* randv,
                            float* grad, hipStream_t st) {
  if (A <= 0) return hipSuccess;
  const unsigned nb = (unsigned)((A + 255) / 256);
  if (laplacian) grid_laplacian_kernel<<<nb, 256, 0, st>>>(G, mask, active, A, weight, inv, seed, step, randv, grad);
  else grid_eikonal_kernel<<<nb, 256, 0, st>>>(G, mask, active, A, weight, inv, seed, step, randv, grad);
  return hipGetLastError();
}

hipError_t launch_grid_laplacian_pull(const GridOrderArgs& A, hipStream_t st) {
  if (A.G <= 0) return hipSuccess;
  if (A.res <= 0 || (A.res & (A.res - 1)) != 0) return hipErrorInvalidValue;
  if (A.T > 0)
    grid_voxel_records_kernel<<<(unsigned)((A.T + 255) / 256), 256, 0, st>>>(A.T, A.tvox, A.coord, A.size, A.vox_key,
                                                                             A.grid_pts, reinterpret_cast<float4*>(A.rec));
  TableOrder O;
  O.keys = A.keys; O.tv = A.tv; O.cell_entry = A.cell_entry; O.cell_tv = A.cell_tv; O.tvox = A.tvox;
  O.rec = reinterpret_cast<const float4*>(A.rec); O.G = A.G; O.T = A.T; O.res = A.res;
  O.lg = __builtin_ctz((unsigned)A.res);
  const unsigned nb = (unsigned)((A.G + 255) / 256);
  const bool lap = (A.terms & 1) != 0, eik = (A.terms & 2) != 0;
  if (lap && eik) {
    grid_smooth_coef_kernel<true, true><<<nb, 256, 0, st>>>(O, A.mask, A.weights, A.weight, A.inv, A.seed, A.step,
                                                            A.coef);
  } else if (eik) {
    grid_smooth_coef_kernel<false, true><<<nb, 256, 0, st>>>(O, A.mask, A.weights, A.weight, A.inv, A.seed, A.step,
                                                             A.coef);
  } else {
    grid_smooth_coef_kernel<true, false><<<nb, 256, 0, st>>>(O, A.mask, A.weights, A.weight, A.inv, A.seed, A.step,
                                                             A.coef);
  }
  if (eik) grid_laplacian_pull_kernel<true><<<nb, 256, 0, st>>>(O, A.vox_key, A.seed, A.step, A.coef, A.grad);
  else grid_laplacian_pull_kernel<false><<<nb, 256, 0, st>>>(O, A.vox_key, A.seed, A.step, A.coef, A.grad);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Per-voxel regularisers of the corner SDFs.
//
// Total variation (tv_compute.cu:23-159): per voxel and channel, corner i receives
// w * sum over its three edges of (+1 if its value exceeds the neighbour's, else -1), w = weight
// (times 0.01 * vox_size_inv[voxel] unless no_tv_s).  With tv_sparse a voxel takes part only
// when one of its corners already has a non-zero gradient; the reference reads the gradient
// while other threads add to it, so here the test is made for every voxel first, on the
// gradient as it was before the launch (one pass writes a flag per voxel and channel).
__global__ __launch_bounds__(256) void tv_sparse_flags_kernel(long long NC, int C, const int64_t* __restrict__ vox_key,
                                                              const float* __restrict__ grad,
                                                              uint8_t* __restrict__ flags) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NC) return;
  const long long n = idx / C;
  const int c = (int)(idx - n * C);
  bool any = false;
#pragma unroll
  for (int i = 0; i < 8; ++i) any |= grad[vox_key[n * 8 + i] * C + c] != 0.f;
  flags[idx] = any ? 1 : 0;
}

template <bool kNoTvS>
__global__ __launch_bounds__(256) void total_variation_kernel(long long NC, int C, const float* __restrict__ grid_pts,
                                                              const int64_t* __restrict__ vox_key, float weight,
                                                              const float* __restrict__ vox_size_inv,
                                                              const uint8_t* __restrict__ flags,
                                                              float* __restrict__ grad) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= NC) return;
  if (flags && !flags[idx]) return;
  const long long n = idx / C;
  const int c = (int)(idx - n * C);
  long long key[8];
  float v[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    key[i] = vox_key[n * 8 + i];
    v[i] = grid_pts[key[i] * C + c];
  }
  float w = weight;
  if (!kNoTvS) w *= 0.01f * vox_size_inv[n];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int g = (v[i] > v[i ^ 1] ? 1 : -1) + (v[i] > v[i ^ 2] ? 1 : -1) + (v[i] > v[i ^ 4] ? 1 : -1);
    atomicAdd(grad + key[i] * C + c, w * (float)g);
  }
}

// Voxel gradient (vg_compute.cu:23-75): per active voxel, the SDF gradient at the centre from
// the four edge differences per axis, g = (1 / 4h) sum_i s_i S_i; corner i receives
// 4 (|g|^2 - 2) weight (g . s_i) / 4 -- the reference's own constants (2, and 1/4 in place of
// 1/(4h) in the chain rule), kept as they are.
__global__ __launch_bounds__(256) void voxel_gradient_kernel(int A, const int32_t* __restrict__ active,
                                                             const float* __restrict__ grid_pts,
                                                             const int64_t* __restrict__ vox_key,
                                                             const float* __restrict__ vox_size_inv, float weight,
                                                             float* __restrict__ grad) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= A) return;
  const long long v = active[tid];
  const float inv4h = 0.25f * vox_size_inv[v];
  long long key[8];
  float S[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    key[i] = vox_key[v * 8 + i];
    S[i] = grid_pts[key[i]];
  }
  // corner i = (bit 2: x, bit 1: y, bit 0: z); s = 2 * bit - 1
  float gx = 0.f, gy = 0.f, gz = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    gx += ((i & 4) ? S[i] : -S[i]);
    gy += ((i & 2) ? S[i] : -S[i]);
    gz += ((i & 1) ? S[i] : -S[i]);
  }
  gx *= inv4h; gy *= inv4h; gz *= inv4h;
  const float g2 = gx * gx + gy * gy + gz * gz;
  const float scale = 4.f * (g2 - 2.f) * weight;
  const float dx = scale * gx, dy = scale * gy, dz = scale * gz;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const float d = ((i & 4) ? dx : -dx) + ((i & 2) ? dy : -dy) + ((i & 1) ? dz : -dz);
    atomicAdd(grad + key[i], d * 0.25f);
  }
}

// Points loss (pl_compute.cu:14-66): each point (in grid cells) inside the grid and on a valid
// cell takes the voxel the table assigns to its cell, interpolates the SDF there with
// unclamped trilinear weights (auxiliary.h:319-334) and adds 2 weight sdf w_i to corner i
// (auxiliary.h:443-455), i.e. the gradient of weight * sdf^2.
__global__ __launch_bounds__(256) void points_loss_kernel(int Mp, const float* __restrict__ pts, GridTable G,
                                                          const bool* __restrict__ mask, float weight,
                                                          float* __restrict__ grad) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= Mp) return;
  const float px = pts[3 * tid], py = pts[3 * tid + 1], pz = pts[3 * tid + 2];
  const int ix = (int)px, iy = (int)py, iz = (int)pz;
  const int res = G.res;
  if (ix < 0 || ix >= res || iy < 0 || iy >= res || iz < 0 || iz >= res) return;
  const int cell = ix + iy * res + iz * res * res;
  if (!mask[cell]) return;
  const int idx = G.dense ? G.dense[cell] : key_search(G.keys, G.n_keys, cell);
  if (idx < 0) return;
  const int v = G.key2vox[idx];
  if (v < 0 || v >= G.n_vox) return;
  const float size = G.vox_size[v];
  const float qx = (px - G.vox_coord[3 * v]) / size;
  const float qy = (py - G.vox_coord[3 * v + 1]) / size;
  const float qz = (pz - G.vox_coord[3 * v + 2]) / size;
  float w[8];
  trilinear_weights(qx, qy, qz, w);
  float sdf = 0.f;
#pragma unroll
  for (int i = 0; i < 8; ++i) sdf += w[i] * G.grid_pts[G.vox_key[(long long)v * 8 + i]];
  const float d = 2.f * sdf * weight;
#pragma unroll
  for (int i = 0; i < 8; ++i) atomicAdd(grad + G.vox_key[(long long)v * 8 + i], d * w[i]);
}

hipError_t launch_total_variation(long long N, int C, const float* grid_pts, const int64_t* vox_key, float weight,
                                  const float* vox_size_inv, bool no_tv_s, uint8_t* sparse_flags, float* grad,
                                  hipStream_t st) {
  const long long NC = N * C;
  if (NC <= 0) return hipSuccess;
  const unsigned nb = (unsigned)((NC + 255) / 256);
  if (sparse_flags) tv_sparse_flags_kernel<<<nb, 256, 0, st>>>(NC, C, vox_key, grad, sparse_flags);
  if (no_tv_s) total_variation_kernel<true><<<nb, 256, 0, st>>>(NC, C, grid_pts, vox_key, weight, vox_size_inv,
                                                                 sparse_flags, grad);
  else total_variation_kernel<false><<<nb, 256, 0, st>>>(NC, C, grid_pts, vox_key, weight, vox_size_inv,
                                                         sparse_flags, grad);
  return hipGetLastError();
}

hipError_t launch_voxel_gradient(int A, const int32_t* active, const float* grid_pts, const int64_t* vox_key,
                                 const float* vox_size_inv, float weight, float* grad, hipStream_t st) {
  if (A <= 0) return hipSuccess;
  voxel_gradient_kernel<<<(unsigned)((A + 255) / 256), 256, 0, st>>>(A, active, grid_pts, vox_key, vox_size_inv,
                                                                     weight, grad);
  return hipGetLastError();
}

hipError_t launch_points_loss(int Mp, const float* pts, const GridTable& G, const bool* mask, float weight,
                              float* grad, hipStream_t st) {
  if (Mp <= 0) return hipSuccess;
  points_loss_kernel<<<(unsigned)((Mp + 255) / 256), 256, 0, st>>>(Mp, pts, G, mask, weight, grad);
  return hipGetLastError();
}

// Cell -> table index for every cell of the res^3 grid: the entry key_search returns (so
// duplicate keys resolve exactly as the search does), -1 where the cell has no key.  Turns the
// six or seven binary searches per active key into one load each.
__global__ __launch_bounds__(256) void grid_dense_index_kernel(long long n_cells, const int32_t* __restrict__ keys,
                                                               int G, int32_t* __restrict__ out) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  out[c] = key_search(keys, G, (int)c);
}

hipError_t launch_grid_dense_index(long long n_cells, const int32_t* keys, int G, int32_t* out, hipStream_t st) {
  if (n_cells <= 0) return hipSuccess;
  grid_dense_index_kernel<<<(unsigned)((n_cells + 255) / 256), 256, 0, st>>>(n_cells, keys, G, out);
  return hipGetLastError();
}

hipError_t launch_table_count(int N, const float* center, const float* size, const float* scene_center,
                              float extent, int res, const bool* is_leaf, uint32_t* count, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  table_count_kernel<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(N, center, size, scene_center, extent, res,
                                                                   is_leaf, count);
  return hipGetLastError();
}

hipError_t launch_table_write(int N, const float* center, const float* size, const float* scene_center,
                              float extent, int res, const bool* is_leaf, const uint32_t* scan, bool* mask,
                              int32_t* keys, int32_t* key2vox, hipStream_t st) {
  if (N <= 0) return hipSuccess;
  table_write_kernel<<<(unsigned)((N + 255) / 256), 256, 0, st>>>(N, center, size, scene_center, extent, res,
                                                                   is_leaf, scan, mask, keys, key2vox);
  return hipGetLastError();
}

}  // namespace svr

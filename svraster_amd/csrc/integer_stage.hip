// Integer stage of the rasterizer on gfx950: voxel preprocess, prefix scan, (tile, rank) key
// duplication, in-house stable LSD radix sort and tile-range extraction.
//
// Reference semantics:
//   preprocess        cuda/src/preprocess.cu:35-154 (+ focal length at 195-198)
//   scan + count      cuda/src/forward.cu:881-897
//   duplicateWithKeys cuda/src/forward.cu:749-802, key codec cuda/src/auxiliary.h:40-74
//   radix sort        cuda/src/forward.cu:917-926 (CUB SortPairs over the key bits)
//   tile ranges       cuda/src/forward.cu:807-829, 928
//
// Every kernel here takes a view table and a (view, voxel) index space so that one launch can
// serve one view (the drop-in path) or many views (the batched path).  Counts that only the
// device knows (the number of duplicated keys) are read from device memory, so the batched
// path can run without host synchronisation.
#include "common.h"
#include "state.h"

namespace svr {

// ------------------------------------------------------------------------------------------
// View table construction (no host synchronisation: matrices are device pointers).
// ------------------------------------------------------------------------------------------
__global__ void init_view_kernel(ViewDesc* out, ViewDesc base, const float* w2c, const float* c2w,
                                 const float* bg) {
  if (threadIdx.x != 0) return;
  ViewDesc v = base;
  for (int i = 0; i < 12; ++i) { v.w2c[i] = w2c[i]; v.c2w[i] = c2w[i]; }
  if (bg) { v.bg[0] = bg[0]; v.bg[1] = bg[1]; v.bg[2] = bg[2]; }
  *out = v;
}

// ------------------------------------------------------------------------------------------
// Preprocess.  The floating-point steps are evaluated without contraction so that the packed
// bounding boxes are bit-identical to an IEEE FP32 restatement with the same operation order.
// ------------------------------------------------------------------------------------------
// bbox, octant bits and key count of voxel (c, half size r) in view v (preprocess.cu:60-153);
// shared by the per-view and the voxel-major batched kernels so both give identical bits.
// tb: the pixels whose centre lies in the projected corners' rectangle widened by kTboxMargin,
// clipped to bb (packed like bb; empty when x0 > x1 or y0 > y1).  A ray through a pixel centre
// outside it cannot meet the cube when every corner is in front of the near plane (the cube's
// projection is the convex hull of its corners' projections); the margin covers FP32 rounding of
// the projection and of the compositors' ray / box test (a few 1e-3 pixel at most).  With a
// corner behind the near plane the projection is unbounded and tb = bb.
// kTight (the multi-view path): bin over tb's tiles instead of bb's, and leave the (view, voxel)
// out (no keys, no octants) when tb holds no pixel; bb is then tb.  The per-view path keeps the
// reference's keys (kTight = false).
constexpr float kTboxMargin = 0.03125f;

template <bool kTight>
__device__ inline void preprocess_one(const ViewDesc& v, bool leaf, float vcx, float vcy, float vcz, float vlen,
                                      uint32_t& nd, uint2& bb, uint2& tb, uint32_t& qbits, uint32_t& ntiles) {
#pragma clang fp contract(off)
  nd = 0;
  ntiles = 0;
  bb = make_uint2(0u, 0u);
  tb = bb;
  qbits = 0;
  if (leaf) {
    const float r = 0.5f * vlen;
    const float rox = v.c2w[3], roy = v.c2w[7], roz = v.c2w[11];
    const float rx = vcx - rox, ry = vcy - roy, rz = vcz - roz;
    const float d2 = rx * rx + ry * ry + rz * rz;
    if (!(d2 < v.near_ * v.near_)) {
      float mnx = 1e9f, mny = 1e9f, mxx = -1e9f, mxy = -1e9f;
      const float* m = v.w2c;
      // Corner c = (x, y, z) bits (c & 4, c & 2, c & 1): coordinate centre -/+ r per axis
      // (centre + r * (-1) is centre - r exactly), in camera space with the operation order
      // ((m0 x + m1 y) + m2 z) + m3.
      const float wx[2] = {vcx + r * -1.f, vcx + r * 1.f};
      const float wy[2] = {vcy + r * -1.f, vcy + r * 1.f};
      const float wz[2] = {vcz + r * -1.f, vcz + r * 1.f};
      float ccx[8], ccy[8], ccz[8];
      bool front = true;
      // !(z < near) && z >= 2^-100 is z >= max(near, 2^-100) (both false for a NaN z)
      const float zlo = fmaxf(v.near_, kRcpRnLo);
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float x = wx[(c >> 2) & 1], y = wy[(c >> 1) & 1], z = wz[c & 1];
        ccx[c] = m[0] * x + m[1] * y + m[2] * z + m[3];
        ccy[c] = m[4] * x + m[5] * y + m[6] * z + m[7];
        ccz[c] = m[8] * x + m[9] * y + m[10] * z + m[11];
        front = front && ccz[c] >= zlo && ccz[c] <= kRcpRnHi;
      }
      if (front) {
        // Every corner in front of the near plane with depth in rcp_rn's range (the common
        // case): the reciprocal without the division sequence (same bits), straight-line min /
        // max (no NaN can arise) and the octant set as the product of the per-axis sides, the
        // same values the per-corner loop below gives.
        float ux[8], uy[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const float inv_z = rcp_rn(ccz[c]);
          ux[c] = ccx[c] * inv_z;
          uy[c] = ccy[c] * inv_z;
        }
        mnx = fminf(mnx, fminf(fminf(fminf(ux[0], ux[1]), fminf(ux[2], ux[3])), fminf(fminf(ux[4], ux[5]), fminf(ux[6], ux[7]))));
        mny = fminf(mny, fminf(fminf(fminf(uy[0], uy[1]), fminf(uy[2], uy[3])), fminf(fminf(uy[4], uy[5]), fminf(uy[6], uy[7]))));
        mxx = fmaxf(mxx, fmaxf(fmaxf(fmaxf(ux[0], ux[1]), fmaxf(ux[2], ux[3])), fmaxf(fmaxf(ux[4], ux[5]), fmaxf(ux[6], ux[7]))));
        mxy = fmaxf(mxy, fmaxf(fmaxf(fmaxf(uy[0], uy[1]), fmaxf(uy[2], uy[3])), fmaxf(fmaxf(uy[4], uy[5]), fmaxf(uy[6], uy[7]))));
        // Bit b of xs: some corner has x-octant bit b (x < camera x); likewise ys, zs.
        const uint32_t xs = (1u << uint32_t(wx[0] < rox)) | (1u << uint32_t(wx[1] < rox));
        const uint32_t ys = (1u << uint32_t(wy[0] < roy)) | (1u << uint32_t(wy[1] < roy));
        const uint32_t zs = (1u << uint32_t(wz[0] < roz)) | (1u << uint32_t(wz[1] < roz));
        const uint32_t yz = ((ys & 1u) ? zs : 0u) | ((ys & 2u) ? (zs << 2) : 0u);
        qbits = ((xs & 1u) ? yz : 0u) | ((xs & 2u) ? (yz << 4) : 0u);
      } else {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          if (ccz[c] < v.near_) continue;
          const float inv_z = 1.0f / ccz[c];
          const float ux = ccx[c] * inv_z, uy = ccy[c] * inv_z;
          mnx = fminf(mnx, ux); mny = fminf(mny, uy);
          mxx = fmaxf(mxx, ux); mxy = fmaxf(mxy, uy);
          qbits |= 1u << ((uint32_t(wx[(c >> 2) & 1] < rox) << 2) | (uint32_t(wy[(c >> 1) & 1] < roy) << 1) |
                          uint32_t(wz[c & 1] < roz));
        }
      }
      const float cxh = v.cx - 0.5f, cyh = v.cy - 0.5f;
      const float bx0 = fmaxf(v.focal_x * mnx + cxh, 0.0f);
      const float by0 = fmaxf(v.focal_y * mny + cyh, 0.0f);
      const float bx1 = fminf(v.focal_x * mxx + cxh, (float)v.W);
      const float by1 = fminf(v.focal_y * mxy + cyh, (float)v.H);
      if (!(bx0 > bx1 || by0 > by1)) {
        if (kTight && front) {
          // The candidate rectangle alone: for 0 <= x < 2^18 (x -/+ 1/32 exact), ceil(x - 1/32)
          // >= rint(x) and floor(x + 1/32) <= rint(x) (ties to even included), so it lies within
          // the reference bbox without clamping to it.
          const uint32_t tx0 = (uint32_t)ceilf(bx0 - kTboxMargin), ty0 = (uint32_t)ceilf(by0 - kTboxMargin);
          const uint32_t tx1 = (uint32_t)floorf(bx1 + kTboxMargin), ty1 = (uint32_t)floorf(by1 + kTboxMargin);
          tb = make_uint2((tx0 << 16) | ty0, (tx1 << 16) | ty1);
          bb = tb;
          if (tx0 > tx1 || ty0 > ty1) qbits = 0;
        } else {
          const uint32_t rx0 = (uint32_t)__builtin_rintf(bx0), ry0 = (uint32_t)__builtin_rintf(by0);
          const uint32_t rx1 = (uint32_t)__builtin_rintf(bx1), ry1 = (uint32_t)__builtin_rintf(by1);
          bb.x = (rx0 << 16) | ry0;
          bb.y = (rx1 << 16) | ry1;
          tb = bb;
          if (front) {
            const uint32_t tx0 = max(rx0, (uint32_t)ceilf(bx0 - kTboxMargin));
            const uint32_t ty0 = max(ry0, (uint32_t)ceilf(by0 - kTboxMargin));
            const uint32_t tx1 = min(rx1, (uint32_t)floorf(bx1 + kTboxMargin));
            const uint32_t ty1 = min(ry1, (uint32_t)floorf(by1 + kTboxMargin));
            tb = make_uint2((tx0 << 16) | ty0, (tx1 << 16) | ty1);
          }
        }
        int tx0, ty0, tx1, ty1;
        bbox_tiles(bb.x, bb.y, v, tx0, ty0, tx1, ty1);
        ntiles = (uint32_t)((1 + ty1 - ty0) * (1 + tx1 - tx0));
        nd = ntiles * (uint32_t)__popc(qbits);
      }
    }
  }
  if (nd == 0) { bb = make_uint2(0u, 0u); tb = bb; qbits = 0; ntiles = 0; }
}

__global__ __launch_bounds__(256) void preprocess_kernel(
    const ViewDesc* __restrict__ views, int P, const float* __restrict__ centers,
    const float* __restrict__ lengths, const uint8_t* __restrict__ is_leaf,
    const int64_t* __restrict__ octpaths, int32_t* __restrict__ out_n_dup,
    uint32_t* __restrict__ n_dup, uint2* __restrict__ bboxes, uint2* __restrict__ tboxes,
    uint8_t* __restrict__ quads, uint32_t* __restrict__ tz_min) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int view = blockIdx.y;
  if (i >= P) return;
  const ViewDesc& v = views[view];
  const long long o = (long long)view * P + i;
  uint32_t nd, ntiles;
  uint2 bb, tb;
  uint32_t qbits;
  preprocess_one<false>(v, is_leaf[i] != 0, centers[3 * i + 0], centers[3 * i + 1], centers[3 * i + 2], lengths[i], nd,
                        bb, tb, qbits, ntiles);
  if (out_n_dup) out_n_dup[o] = (int32_t)nd;
  n_dup[o] = nd;
  bboxes[o] = bb;
  tboxes[o] = tb;
  quads[o] = (uint8_t)qbits;
  if (tz_min) {
    // Number of low rank bits (a multiple of 3) that are zero in this voxel's octpath; reduced
    // across the wave, and the wave's atomic is skipped when it cannot lower the current minimum
    // (all waves hit one address, so issuing every atomic would serialise them).
    uint32_t tz = (uint32_t)kRankBits;
    if (nd) {
      const uint64_t p = (uint64_t)octpaths[i] & ((1ull << kRankBits) - 1ull);
      tz = p ? (uint32_t)__builtin_ctzll(p) : (uint32_t)kRankBits;
      tz = min(tz - tz % 3u, (uint32_t)kRankBits);
    }
    tz = wave_min_active_u32(tz);
    if (__lane_id() == __ffsll((long long)__ballot(1)) - 1 && tz < (uint32_t)kRankBits &&
        tz < __hip_atomic_load(tz_min, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMin(tz_min, tz);
  }
}

// ------------------------------------------------------------------------------------------
// Reduce-then-scan over u32 (tile = 256 threads x 16 items).
// ------------------------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t& total, uint32_t* lds_w) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)x, o, 64);
    if (lane >= o) x += t;
  }
  if (lane == 63) lds_w[w] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
  for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
    const uint32_t c = lds_w[k];
    if (k < w) wbase += c;
    tot += c;
  }
  __syncthreads();
  total = tot;
  return wbase + x - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(
    const uint32_t* __restrict__ in, const uint32_t* __restrict__ n_dev, long long n_cap,
    uint32_t* __restrict__ partial) {
  __shared__ uint32_t lw[4];
  const long long n = n_dev ? min((long long)*n_dev, n_cap) : n_cap;
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const long long i = base + k;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_exclusive_scan(s, tot, lw);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

// Single workgroup: exclusive scan of the per-tile partial sums in place.
__global__ __launch_bounds__(1024) void scan_partials_kernel(uint32_t* __restrict__ partial, int nb) {
  __shared__ uint32_t lw[16];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    const int i = base + threadIdx.x;
    const uint32_t v = i < nb ? partial[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(v, tot, lw);
    const uint32_t c = carry;
    if (i < nb) partial[i] = c + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry = c + tot;
    __syncthreads();
  }
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(
    const uint32_t* __restrict__ in, const uint32_t* __restrict__ n_dev, long long n_cap,
    const uint32_t* __restrict__ partial, uint32_t* __restrict__ out, int inclusive) {
  __shared__ uint32_t lw[4];
  const long long n = n_dev ? min((long long)*n_dev, n_cap) : n_cap;
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const long long i = base + k;
    v[k] = i < n ? in[i] : 0u;
    s += v[k];
  }
  uint32_t tot;
  uint32_t run = block_exclusive_scan(s, tot, lw) + partial[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const long long i = base + k;
    const uint32_t nx = run + v[k];
    if (i < n) out[i] = inclusive ? nx : run;
    run = nx;
  }
}

// ------------------------------------------------------------------------------------------
// Duplication: one (compressed key, value) pair per (octant, tile) of each visible voxel, in
// the reference's order (octants ascending, then tile rows, then tile columns).
//
// Compressed key = global tile << rank_bits | (rank >> shift) << qb | (qb ? octant : 0), where
// rank = octpath ^ repeat3(octant), `shift` is a multiple of 3 below which every visible
// voxel's octpath is zero, and qb = 3 when shift >= 3.  Below `shift` the rank bits are the
// octant pattern, which orders like the octant itself, so sorting the compressed key gives the
// same order as sorting the reference's (tile << 48 | rank) key.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void duplicate_kernel(
    const ViewDesc* __restrict__ views, int P, const int64_t* __restrict__ octpaths,
    const uint32_t* __restrict__ n_dup, const uint32_t* __restrict__ scan,
    const uint2* __restrict__ bboxes, const uint8_t* __restrict__ quads, int shift, int qb,
    int rank_bits, uint64_t* __restrict__ keys, uint32_t* __restrict__ vals, long long cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int view = blockIdx.y;
  if (i >= P) return;
  const long long o = (long long)view * P + i;
  const uint32_t nd = n_dup[o];
  if (nd == 0) return;
  const ViewDesc& v = views[view];
  uint32_t off = scan[o] - nd;
  const uint2 bb = bboxes[o];
  int tx0, ty0, tx1, ty1;
  bbox_tiles(bb.x, bb.y, v, tx0, ty0, tx1, ty1);
  const uint64_t path = (uint64_t)octpaths[i];
  const uint32_t qbits = quads[o];
  for (uint32_t q = 0; q < 8; ++q) {
    if (!(qbits & (1u << q))) continue;
    const uint64_t rank = ((path ^ rep3(q)) & ((1ull << kRankBits) - 1ull)) >> shift;
    const uint64_t rk = (rank << qb) | (qb ? (uint64_t)q : 0ull);
    const uint32_t val = (q << 29) | (uint32_t)i;
    for (int y = ty0; y <= ty1; ++y)
      for (int x = tx0; x <= tx1; ++x) {
        const uint64_t tile = (uint64_t)(v.tile_base + y * v.tiles_x + x);
        if ((long long)off < cap) {          // capacity guard for the preallocated batched path
          keys[off] = (tile << rank_bits) | rk;
          vals[off] = val;
        }
        ++off;
      }
  }
}

// ------------------------------------------------------------------------------------------
// Stable LSD radix sort over kBits-bit digits (8 for the 64-bit rank keys, 9 for 32-bit tile
// ids).  Per pass: per-tile digit histograms written tile-major (each workgroup one contiguous
// row), a column scan of that table (per digit over tiles, then over digits), then a stable
// scatter in which each wave64 ranks its 64 items with kBits ballots (items with equal digits get
// consecutive slots in lane order) and reads its row of output offsets.
// ------------------------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kMaxDigitBits = 9;

template <typename K, int kBits>
__global__ __launch_bounds__(kSortThreads) void radix_histogram_kernel(
    const K* __restrict__ keys, const uint32_t* __restrict__ n_dev, long long n_cap,
    int shift, uint32_t* __restrict__ hist, int nb) {
  constexpr int kD = 1 << kBits;
  __shared__ uint32_t h[kD];
  for (int d = threadIdx.x; d < kD; d += kSortThreads) h[d] = 0;
  __syncthreads();
  const long long n = n_dev ? min((long long)*n_dev, n_cap) : n_cap;
  const long long base = (long long)blockIdx.x * kSortTile;
  // Runs of equal digits in consecutive lanes (frequent in the pass over the high digit) are
  // counted by their first lane alone, so they do not serialise on one LDS address.
  const int lane = threadIdx.x & 63;
#pragma unroll 4
  for (int r = 0; r < kSortItems; ++r) {
    const long long i = base + r * kSortThreads + threadIdx.x;
    const bool valid = i < n;
    const uint32_t d = valid ? (uint32_t)(keys[i] >> shift) & (uint32_t)(kD - 1) : (uint32_t)kD;
    const uint32_t prev = (uint32_t)__shfl_up((int)d, 1);
    const bool head = valid && (lane == 0 || prev != d);
    // run end: the next head or invalid lane above this one (valid lanes form a prefix)
    const uint64_t stops = __ballot(!valid || lane == 0 || prev != d) & ~((2ull << lane) - 1ull);
    const int next = stops ? (int)__builtin_ctzll(stops) : 64;
    if (head) atomicAdd(&h[d], (uint32_t)(next - lane));
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kD; d += kSortThreads) hist[(long long)blockIdx.x * kD + d] = h[d];
}

// Column scan of the tile-major histogram table hist[tile][digit] (nb rows of kD), in place:
// entry (b, d) becomes the output position of tile b's first key with digit d -- the keys with
// smaller digits plus those with digit d in earlier tiles.  Rows are grouped in chunks of
// kColChunk: per chunk and digit a sum, the sums transposed digit-major and scanned, then each
// chunk walks its rows from its prefix.  Every pass over the table reads whole rows.
constexpr int kColChunk = 128;

int col_chunks(int nb) { return (nb + kColChunk - 1) / kColChunk; }

template <int kD>
__global__ __launch_bounds__(kD) void hist_chunk_sum_kernel(const uint32_t* __restrict__ hist, int nb,
                                                            uint32_t* __restrict__ csum_t, int nc) {
  const int d = threadIdx.x, c = blockIdx.x;
  const int b0 = c * kColChunk, b1 = min(nb, b0 + kColChunk);
  uint32_t s = 0;
#pragma unroll 8
  for (int b = b0; b < b1; ++b) s += hist[(long long)b * kD + d];
  csum_t[(long long)d * nc + c] = s;   // digit-major, for the scan
}

template <int kD>
__global__ __launch_bounds__(kD) void hist_chunk_apply_kernel(uint32_t* __restrict__ hist, int nb,
                                                              const uint32_t* __restrict__ cpre_t, int nc) {
  const int d = threadIdx.x, c = blockIdx.x;
  const int b0 = c * kColChunk, b1 = min(nb, b0 + kColChunk);
  uint32_t run = cpre_t[(long long)d * nc + c];
#pragma unroll 8
  for (int b = b0; b < b1; ++b) {
    const uint32_t v = hist[(long long)b * kD + d];
    hist[(long long)b * kD + d] = run;
    run += v;
  }
}

// Stable scatter of one tile of kSortTile keys.  Wavefront w ranks the contiguous quarter
// [w * kSortTile / 4, (w + 1) * kSortTile / 4) of the tile, 64 keys at a time, against its own
// per-digit running counts (kBits ballots per round, no workgroup barrier); the counts then give
// every key its slot in the tile sorted by digit, the keys are reordered through LDS, and each
// thread writes consecutive slots, so a wavefront's stores cover runs of consecutive addresses.
template <typename K, int kBits>
__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(
    const K* __restrict__ keys_in, const uint32_t* __restrict__ vals_in,
    const uint32_t* __restrict__ n_dev, long long n_cap, int shift,
    const uint32_t* __restrict__ offsets, int nb, K* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out) {
  constexpr int kD = 1 << kBits;
  constexpr int kPerThread = kD / kSortThreads;          // digits per thread in the digit scan
  constexpr int kQuarter = kSortTile / 4;
  static_assert(kD % kSortThreads == 0 && kSortItems * 64 == kQuarter, "tile shape");
  __shared__ uint32_t wrun[4][kD];   // per wavefront: keys of digit d so far, then its first slot
  __shared__ uint32_t adj[kD];       // output position of slot 0 if it held digit d
  __shared__ K s_key[kSortTile];
  __shared__ uint32_t s_val[kSortTile];
  __shared__ uint32_t s_w[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long long n = n_dev ? min((long long)*n_dev, n_cap) : n_cap;
  const long long base = (long long)blockIdx.x * kSortTile;
  if (base >= n) return;                                  // uniform across the block
  const int n_tile = (int)min((long long)kSortTile, n - base);
  for (int d = t; d < kD; d += kSortThreads) {
#pragma unroll
    for (int k = 0; k < 4; ++k) wrun[k][d] = 0;
  }
  __syncthreads();
  // 1. per wavefront, rank of each key among the earlier keys of its digit in the quarter
  const uint64_t lt_mask = (1ull << lane) - 1ull;
  K key[kSortItems];
  uint32_t val[kSortItems], pos[kSortItems];
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const int li = w * kQuarter + r * 64 + lane;         // index within the tile
    const bool valid = li < n_tile;
    key[r] = valid ? keys_in[base + li] : (K)0;
    val[r] = valid ? vals_in[base + li] : 0u;
  }
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    const bool valid = w * kQuarter + r * 64 + lane < n_tile;
    const uint32_t d = valid ? (uint32_t)(key[r] >> shift) & (uint32_t)(kD - 1) : (uint32_t)kD;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kBits; ++b) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
    uint32_t before = 0;
    if (valid) before = wrun[w][d];
    // LDS accesses of one wavefront complete in order: every lane has read the count above
    // before the leader's update lands.
    if (valid && rank == 0) wrun[w][d] = before + (uint32_t)__popcll(peers);
    pos[r] = valid ? before + rank : 0u;
  }
  __syncthreads();
  // 2. first slot of each digit in the sorted tile, and of each wavefront's keys within it
  uint32_t tot[kPerThread], sum = 0;
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const int d = t * kPerThread + j;
    tot[j] = wrun[0][d] + wrun[1][d] + wrun[2][d] + wrun[3][d];
    sum += tot[j];
  }
  uint32_t all;
  uint32_t start = block_exclusive_scan(sum, all, s_w);
#pragma unroll
  for (int j = 0; j < kPerThread; ++j) {
    const int d = t * kPerThread + j;
    adj[d] = offsets[(long long)blockIdx.x * kD + d] - start;
    uint32_t s0 = start;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t c = wrun[k][d];
      wrun[k][d] = s0;
      s0 += c;
    }
    start += tot[j];
  }
  __syncthreads();
  // 3. keys into their slots, then out in slot order
#pragma unroll
  for (int r = 0; r < kSortItems; ++r) {
    if (w * kQuarter + r * 64 + lane < n_tile) {
      const uint32_t d = (uint32_t)(key[r] >> shift) & (uint32_t)(kD - 1);
      const uint32_t slot = wrun[w][d] + pos[r];
      s_key[slot] = key[r];
      s_val[slot] = val[r];
    }
  }
  __syncthreads();
  for (int sl = t; sl < n_tile; sl += kSortThreads) {
    const K k = s_key[sl];
    const uint32_t dst = adj[(uint32_t)(k >> shift) & (uint32_t)(kD - 1)] + (uint32_t)sl;
    keys_out[dst] = k;
    vals_out[dst] = s_val[sl];
  }
}

// ------------------------------------------------------------------------------------------
// Tile ranges of the sorted list (ranges must be zero-filled first).
// ------------------------------------------------------------------------------------------
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, long long n, uint32_t v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void tile_ranges_kernel(
    const uint64_t* __restrict__ keys, const uint32_t* __restrict__ n_dev, long long n_cap,
    int rank_bits, uint2* __restrict__ ranges) {
  const long long n = n_dev ? min((long long)*n_dev, n_cap) : n_cap;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t tile = (uint32_t)(keys[i] >> rank_bits);
  if (i == 0 || (uint32_t)(keys[i - 1] >> rank_bits) != tile) ranges[tile].x = (uint32_t)i;
  if (i == n - 1 || (uint32_t)(keys[i + 1] >> rank_bits) != tile) ranges[tile].y = (uint32_t)(i + 1);
}

// Tile ranges of a sorted list of 32-bit tile ids (ranges zero-filled first).
// Four sorted entries per thread (one 16-byte load; the tiles buffer is 256-byte aligned), compared
// with their neighbours; ~0u stands for "no entry" (tile ids are far below it).
__global__ __launch_bounds__(256) void tile_ranges_u32_kernel(
    const uint32_t* __restrict__ tiles, const uint32_t* __restrict__ n_dev, long long n_cap,
    uint2* __restrict__ ranges) {
  const long long n = n_dev ? min((long long)*n_dev, n_cap) : n_cap;
  const long long i0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i0 >= n) return;
  uint32_t t[6];
  t[0] = i0 > 0 ? tiles[i0 - 1] : ~0u;
  if (i0 + 4 <= n) {
    const uint4 q = *reinterpret_cast<const uint4*>(tiles + i0);
    t[1] = q.x; t[2] = q.y; t[3] = q.z; t[4] = q.w;
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) t[k + 1] = i0 + k < n ? tiles[i0 + k] : ~0u;
  }
  t[5] = i0 + 4 < n ? tiles[i0 + 4] : ~0u;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (i0 + k >= n) break;
    if (t[k] != t[k + 1]) ranges[t[k + 1]].x = (uint32_t)(i0 + k);
    if (t[k + 2] != t[k + 1]) ranges[t[k + 1]].y = (uint32_t)(i0 + k + 1);
  }
}

// ------------------------------------------------------------------------------------------
// Batched binning in rank order.  Within a tile the reference orders entries by the 48-bit
// rank octpath ^ repeat3(octant) (auxiliary.h:40-59), a key that does not depend on the view.
// Every (voxel, octant) pair is therefore sorted once by rank (pair_rank_kernel + the radix
// sort; the order is fixed while the voxel set is), and each step writes its keys pair by pair
// in that order and sorts them stably by global tile id only, which leaves each tile's entries
// in the reference's order.
// ------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pair_rank_kernel(long long n_pairs, const int64_t* __restrict__ octpaths,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pairs) return;
  const uint32_t q = (uint32_t)(i & 7);
  keys[i] = ((uint64_t)octpaths[i >> 3] ^ rep3(q)) & ((1ull << kRankBits) - 1ull);
  vals[i] = (uint32_t)i;                       // voxel << 3 | octant
}

// Voxel-major preprocess of all views: the voxel is read once, its bbox / octants / n_dup are
// written per view, and per octant q the number of keys it emits over all views (tiles of the
// views whose octant set holds q) goes to pair_cnt[8 * voxel + q].  Each (view, voxel) is binned
// over the tiles of its candidate-pixel rectangle (preprocess_one<true>): every tile list is the
// reference's list without the entries that cannot meet any pixel of the tile, in the same order.
__global__ __launch_bounds__(256) void preprocess_views_kernel(
    const ViewDesc* __restrict__ views, int V, int P, const float* __restrict__ centers,
    const float* __restrict__ lengths, const uint8_t* __restrict__ is_leaf, uint32_t* __restrict__ n_dup,
    uint2* __restrict__ bboxes, uint2* __restrict__ tboxes, uint8_t* __restrict__ quads,
    const uint32_t* __restrict__ pair_slots, uint32_t* __restrict__ pair_cnt, uint32_t* __restrict__ pair_vmask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const bool leaf = is_leaf[i] != 0;
  const float cx = centers[3 * i + 0], cy = centers[3 * i + 1], cz = centers[3 * i + 2];
  const float len = lengths[i];
  const int nw = (V + 31) >> 5;
  uint32_t c8[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
  uint32_t m8[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};   // views (of the current 32) that see octant q
  for (int v = 0; v < V; ++v) {
    uint32_t nd, qbits, nt;   // nt: tiles of the bbox, the keys per octant
    uint2 bb, tb;
    preprocess_one<true>(views[v], leaf, cx, cy, cz, len, nd, bb, tb, qbits, nt);
    const long long o = (long long)v * P + i;
    if (n_dup) n_dup[o] = nd;
    quads[o] = (uint8_t)qbits;
    if (nd) {
      bboxes[o] = bb;        // read only where n_dup > 0
      tboxes[o] = tb;
      const uint32_t vbit = 1u << (v & 31);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const uint32_t sq = 0u - ((qbits >> q) & 1u);   // all ones when the view sees octant q
        c8[q] += nt & sq;
        m8[q] |= vbit & sq;
      }
    }
    if ((v & 31) == 31 || v == V - 1) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        pair_vmask[(8LL * i + q) * nw + (v >> 5)] = m8[q];
        m8[q] = 0u;
      }
    }
  }
  if (pair_slots) {
    // Straight into rank order: eight scattered stores instead of a gather pass over all pairs.
    const uint4* sl = reinterpret_cast<const uint4*>(pair_slots + 8LL * i);
    const uint4 s0 = sl[0], s1 = sl[1];
    pair_cnt[s0.x] = c8[0]; pair_cnt[s0.y] = c8[1]; pair_cnt[s0.z] = c8[2]; pair_cnt[s0.w] = c8[3];
    pair_cnt[s1.x] = c8[4]; pair_cnt[s1.y] = c8[5]; pair_cnt[s1.z] = c8[6]; pair_cnt[s1.w] = c8[7];
  } else {
    uint4* out = reinterpret_cast<uint4*>(pair_cnt + 8LL * i);
    out[0] = make_uint4(c8[0], c8[1], c8[2], c8[3]);
    out[1] = make_uint4(c8[4], c8[5], c8[6], c8[7]);
  }
}

__global__ __launch_bounds__(256) void pair_slots_kernel(long long n_pairs, const uint32_t* __restrict__ order,
                                                         uint32_t* __restrict__ slots) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_pairs) slots[order[i]] = (uint32_t)i;
}

// Key counts in rank order.
__global__ __launch_bounds__(256) void pair_gather_kernel(long long n_pairs, const uint32_t* __restrict__ order,
                                                          const uint32_t* __restrict__ pair_cnt,
                                                          uint32_t* __restrict__ cnt_ranked) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_pairs) cnt_ranked[i] = pair_cnt[order[i]];
}

// Keys of each (voxel, octant) pair at its rank-order offset: global tile id and value
// octant << 29 | voxel, view by view and tile by tile.  A workgroup takes 256 consecutive pairs and
// lists their (pair, view) tile rectangles in LDS: each pair's thread enters its rectangles' views,
// then each thread completes kEmitRectsPerThread consecutive rectangles (view box, tile count) and a
// workgroup scan over the tile counts gives every rectangle its output offset.  Every thread then
// writes whole output positions, finding the rectangle that holds its position by a binary search
// over the offsets.  So the loads run in parallel across the workgroup, the work per thread does not
// depend on how the keys are spread over the pairs (a coarse voxel can own thousands of keys while
// its neighbours in rank order own none), and the stores are coalesced; the small LDS footprint
// keeps 8 wavefronts per SIMD resident to overlap the workgroups' chains of dependent loads.
constexpr int kEmitRectsPerThread = 4;
constexpr int kEmitRects = 256 * kEmitRectsPerThread;

__global__ __launch_bounds__(256) void pair_emit_kernel(
    const ViewDesc* __restrict__ views, int V, int P, long long n_pairs, const uint32_t* __restrict__ order,
    const uint32_t* __restrict__ cnt_ranked, const uint32_t* __restrict__ scan_ranked,
    const uint2* __restrict__ bboxes, const uint32_t* __restrict__ vmask, uint32_t* __restrict__ tiles,
    uint32_t* __restrict__ vals, long long cap) {
  __shared__ uint32_t s_off[kEmitRects];    // output offset of the rectangle, from the block's first key
  __shared__ uint32_t s_tile0[kEmitRects];  // global tile id of its top-left tile
  __shared__ uint32_t s_shape[kEmitRects];  // view, then width | row stride << 16 (tiles)
  __shared__ uint32_t s_val[kEmitRects];    // octant << 29 | voxel
  __shared__ uint32_t s_w[4];
  const long long i0 = (long long)blockIdx.x * 256;
  const long long i = i0 + threadIdx.x;
  const long long i_last = min(n_pairs, i0 + 256) - 1;
  // Output range of the block: [first, end), the same for every thread.
  const long long first = (long long)scan_ranked[i0] - cnt_ranked[i0];
  const long long end = min((long long)scan_ranked[i_last], cap);
  if (first >= end) return;
  const int nw = (V + 31) >> 5;
  uint32_t c = 0, pv = 0, nrect = 0;
  if (i < n_pairs) {
    c = cnt_ranked[i];
    pv = order[i];
  }
  const uint32_t* pm = vmask + (8LL * (pv >> 3) + (pv & 7u)) * nw;   // the pair's view mask words
  if (c)
    for (int wi = 0; wi < nw; ++wi) nrect += __popc(pm[wi]);
  uint32_t total;
  const uint32_t rbase = block_exclusive_scan(nrect, total, s_w);
  uint32_t base = 0;   // output offset of the batch's first rectangle
  for (uint32_t b0 = 0; b0 < total; b0 += kEmitRects) {
    const uint32_t nb = min(total - b0, (uint32_t)kEmitRects);
    // 1. views of the batch's rectangles, in pair order and view order within a pair
    if (nrect && rbase < b0 + nb && rbase + nrect > b0) {
      uint32_t k = rbase;
      for (int wi = 0; wi < nw && k < b0 + nb; ++wi)
        for (uint32_t m = pm[wi]; m && k < b0 + nb; m &= m - 1u, ++k)
          if (k >= b0) {
            s_shape[k - b0] = (uint32_t)((wi << 5) + __builtin_ctz(m));
            s_val[k - b0] = pv;
          }
    }
    __syncthreads();
    // 2. boxes and tile counts of kEmitRectsPerThread consecutive rectangles per thread
    uint32_t n[kEmitRectsPerThread], cum = 0;
#pragma unroll
    for (int j = 0; j < kEmitRectsPerThread; ++j) {
      const uint32_t r = threadIdx.x * kEmitRectsPerThread + j;
      n[j] = 0;
      if (r < nb) {
        const int v = (int)s_shape[r];
        const uint32_t pr = s_val[r];
        const ViewDesc& vd = views[v];
        const uint2 bb = bboxes[(long long)v * P + (pr >> 3)];
        int tx0, ty0, tx1, ty1;
        bbox_tiles(bb.x, bb.y, vd, tx0, ty0, tx1, ty1);
        const uint32_t tw = (uint32_t)(tx1 - tx0 + 1);
        n[j] = tw * (uint32_t)(ty1 - ty0 + 1);
        s_tile0[r] = (uint32_t)(vd.tile_base + ty0 * vd.tiles_x + tx0);
        s_shape[r] = tw | ((uint32_t)vd.tiles_x << 16);
        s_val[r] = ((pr & 7u) << 29) | (pr >> 3);
      }
      cum += n[j];
    }
    uint32_t batch_keys;
    uint32_t o = base + block_exclusive_scan(cum, batch_keys, s_w);
#pragma unroll
    for (int j = 0; j < kEmitRectsPerThread; ++j) {
      const uint32_t r = threadIdx.x * kEmitRectsPerThread + j;
      if (r < nb) s_off[r] = o;
      o += n[j];
    }
    __syncthreads();
    // 3. the batch's output positions, clipped to the block's range
    const uint32_t p_end = (uint32_t)min((long long)(base + batch_keys), end - first);
    for (uint32_t p = base + threadIdx.x; p < p_end; p += 256) {
      uint32_t r = 0;   // last rectangle whose offset is <= p
#pragma unroll
      for (uint32_t step = kEmitRects / 2; step > 0; step >>= 1) {
        const uint32_t t = r + step;
        if (t < nb && s_off[t] <= p) r = t;
      }
      const uint32_t shape = s_shape[r], tw = shape & 0xFFFFu;
      const uint32_t rel = p - s_off[r], ty = rel / tw;
      tiles[first + p] = s_tile0[r] + ty * (shape >> 16) + (rel - ty * tw);
      vals[first + p] = s_val[r];
    }
    base += batch_keys;
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// Host-side launch helpers (used by the C-ABI entry points).
// ------------------------------------------------------------------------------------------
hipError_t launch_init_view(ViewDesc* out, const ViewDesc& base, const float* w2c,
                            const float* c2w, const float* bg, hipStream_t st) {
  init_view_kernel<<<1, 64, 0, st>>>(out, base, w2c, c2w, bg);
  return hipGetLastError();
}

hipError_t launch_preprocess(const ViewDesc* views, int n_views, int P, const float* centers,
                             const float* lengths, const uint8_t* is_leaf, const int64_t* octpaths,
                             int32_t* out_n_dup, uint32_t* n_dup, uint2* bboxes, uint2* tboxes, uint8_t* quads,
                             uint32_t* tz_min, hipStream_t st) {
  if (P <= 0 || n_views <= 0) return hipSuccess;
  dim3 grid(ceil_div(P, 256), n_views);
  preprocess_kernel<<<grid, 256, 0, st>>>(views, P, centers, lengths, is_leaf, octpaths, out_n_dup,
                                           n_dup, bboxes, tboxes, quads, tz_min);
  return hipGetLastError();
}

size_t scan_partial_count(long long n) { return (size_t)((n + kScanTile - 1) / kScanTile); }

__global__ void key_peak_kernel(const uint32_t* __restrict__ total, uint32_t* __restrict__ peak) {
  if (threadIdx.x == 0) atomicMax(peak, *total);
}

hipError_t launch_key_peak(const uint32_t* total, uint32_t* peak, hipStream_t st) {
  key_peak_kernel<<<1, 64, 0, st>>>(total, peak);
  return hipGetLastError();
}

hipError_t launch_scan(const uint32_t* in, const uint32_t* n_dev, long long n_cap, uint32_t* partial,
                       uint32_t* out, bool inclusive, hipStream_t st) {
  if (n_cap <= 0) return hipSuccess;
  const int nb = (int)scan_partial_count(n_cap);
  scan_reduce_kernel<<<nb, kScanThreads, 0, st>>>(in, n_dev, n_cap, partial);
  scan_partials_kernel<<<1, 1024, 0, st>>>(partial, nb);
  scan_apply_kernel<<<nb, kScanThreads, 0, st>>>(in, n_dev, n_cap, partial, out, inclusive ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_duplicate(const ViewDesc* views, int n_views, int P, const int64_t* octpaths,
                            const uint32_t* n_dup, const uint32_t* scan, const uint2* bboxes,
                            const uint8_t* quads, int shift, int qb, int rank_bits, uint64_t* keys,
                            uint32_t* vals, long long cap, hipStream_t st) {
  if (P <= 0 || n_views <= 0) return hipSuccess;
  dim3 grid(ceil_div(P, 256), n_views);
  duplicate_kernel<<<grid, 256, 0, st>>>(views, P, octpaths, n_dup, scan, bboxes, quads, shift, qb,
                                          rank_bits, keys, vals, cap);
  return hipGetLastError();
}

size_t sort_tile_count(long long n) { return (size_t)((n + kSortTile - 1) / kSortTile); }

// Scratch of the histogram column scan for n keys (up to 9-bit digits): the digit-major chunk
// sums, scanned in place, and that scan's partials.
size_t sort_scratch_count(long long n) {
  const long long m = (1LL << kMaxDigitBits) * (long long)col_chunks((int)sort_tile_count(n));
  return (size_t)(m + (long long)scan_partial_count(m) + 1);
}

template <int kD>
static hipError_t hist_column_scan(uint32_t* hist, int nb, uint32_t* scratch, hipStream_t st) {
  const int nc = col_chunks(nb);
  const long long m = (long long)kD * nc;
  hist_chunk_sum_kernel<kD><<<nc, kD, 0, st>>>(hist, nb, scratch, nc);
  hipError_t e = launch_scan(scratch, nullptr, m, scratch + m, scratch, false, st);
  if (e != hipSuccess) return e;
  hist_chunk_apply_kernel<kD><<<nc, kD, 0, st>>>(hist, nb, scratch, nc);
  return hipGetLastError();
}

// Sorts (keys, vals) by bits [0, nbits).  Buffers ping-pong; returns in `*result_in_alt`
// whether the sorted data ended in the alt buffers.
hipError_t launch_radix_sort(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                             const uint32_t* n_dev, long long n_cap, int nbits, uint32_t* hist,
                             uint32_t* scan_partial, bool* result_in_alt, hipStream_t st) {
  *result_in_alt = false;
  if (n_cap <= 0) return hipSuccess;
  const int nb = (int)sort_tile_count(n_cap);
  uint64_t* ki = keys; uint32_t* vi = vals;
  uint64_t* ko = keys_alt; uint32_t* vo = vals_alt;
  for (int shift = 0; shift < nbits; shift += 8) {
    radix_histogram_kernel<uint64_t, 8><<<nb, kSortThreads, 0, st>>>(ki, n_dev, n_cap, shift, hist, nb);
    hipError_t e = hist_column_scan<256>(hist, nb, scan_partial, st);
    if (e != hipSuccess) return e;
    radix_scatter_kernel<uint64_t, 8><<<nb, kSortThreads, 0, st>>>(ki, vi, n_dev, n_cap, shift, hist, nb, ko, vo);
    std::swap(ki, ko);
    std::swap(vi, vo);
    *result_in_alt = !*result_in_alt;
  }
  return hipGetLastError();
}

int tile_sort_passes(long long n_tiles) {
  int bits = 1;
  while ((1LL << bits) < n_tiles) ++bits;
  return (bits + kMaxDigitBits - 1) / kMaxDigitBits;
}

// Stable sort of (32-bit tile id, value) pairs by tile id, 9-bit digits.  The data starts in
// the buffers that make the last pass end in (tiles, vals) (see tile_sort_passes).
hipError_t launch_tile_sort(uint32_t* tiles, uint32_t* vals, uint32_t* tiles_alt, uint32_t* vals_alt,
                            const uint32_t* n_dev, long long n_cap, long long n_tiles, uint32_t* hist,
                            uint32_t* scan_partial, hipStream_t st) {
  if (n_cap <= 0) return hipSuccess;
  const int passes = tile_sort_passes(n_tiles);
  const int nb = (int)sort_tile_count(n_cap);
  uint32_t* ki = (passes % 2) ? tiles_alt : tiles;
  uint32_t* vi = (passes % 2) ? vals_alt : vals;
  uint32_t* ko = (passes % 2) ? tiles : tiles_alt;
  uint32_t* vo = (passes % 2) ? vals : vals_alt;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * kMaxDigitBits;
    radix_histogram_kernel<uint32_t, kMaxDigitBits><<<nb, kSortThreads, 0, st>>>(ki, n_dev, n_cap, shift, hist, nb);
    hipError_t e = hist_column_scan<1 << kMaxDigitBits>(hist, nb, scan_partial, st);
    if (e != hipSuccess) return e;
    radix_scatter_kernel<uint32_t, kMaxDigitBits><<<nb, kSortThreads, 0, st>>>(ki, vi, n_dev, n_cap, shift, hist, nb,
                                                                               ko, vo);
    std::swap(ki, ko);
    std::swap(vi, vo);
  }
  return hipGetLastError();
}

hipError_t launch_pair_rank(long long n_pairs, const int64_t* octpaths, uint64_t* keys, uint32_t* vals,
                            hipStream_t st) {
  if (n_pairs <= 0) return hipSuccess;
  pair_rank_kernel<<<(unsigned)((n_pairs + 255) / 256), 256, 0, st>>>(n_pairs, octpaths, keys, vals);
  return hipGetLastError();
}

hipError_t launch_preprocess_views(const ViewDesc* views, int V, int P, const float* centers, const float* lengths,
                                   const uint8_t* is_leaf, uint32_t* n_dup, uint2* bboxes, uint2* tboxes,
                                   uint8_t* quads, const uint32_t* pair_slots, uint32_t* pair_cnt,
                                   uint32_t* pair_vmask, hipStream_t st) {
  if (P <= 0 || V <= 0) return hipSuccess;
  preprocess_views_kernel<<<ceil_div(P, 256), 256, 0, st>>>(views, V, P, centers, lengths, is_leaf, n_dup, bboxes,
                                                            tboxes, quads, pair_slots, pair_cnt, pair_vmask);
  return hipGetLastError();
}

hipError_t launch_pair_slots(long long n_pairs, const uint32_t* order, uint32_t* slots, hipStream_t st) {
  if (n_pairs <= 0) return hipSuccess;
  pair_slots_kernel<<<(unsigned)((n_pairs + 255) / 256), 256, 0, st>>>(n_pairs, order, slots);
  return hipGetLastError();
}

hipError_t launch_pair_gather(long long n_pairs, const uint32_t* order, const uint32_t* pair_cnt,
                              uint32_t* cnt_ranked, hipStream_t st) {
  if (n_pairs <= 0) return hipSuccess;
  pair_gather_kernel<<<(unsigned)((n_pairs + 255) / 256), 256, 0, st>>>(n_pairs, order, pair_cnt, cnt_ranked);
  return hipGetLastError();
}

hipError_t launch_pair_emit(const ViewDesc* views, int V, int P, long long n_pairs, const uint32_t* order,
                            const uint32_t* cnt_ranked, const uint32_t* scan_ranked, const uint2* bboxes,
                            const uint32_t* vmask, uint32_t* tiles, uint32_t* vals, long long cap, hipStream_t st) {
  if (n_pairs <= 0) return hipSuccess;
  pair_emit_kernel<<<(unsigned)((n_pairs + 255) / 256), 256, 0, st>>>(views, V, P, n_pairs, order, cnt_ranked,
                                                                      scan_ranked, bboxes, vmask, tiles, vals, cap);
  return hipGetLastError();
}

hipError_t launch_tile_ranges_u32(const uint32_t* tiles, const uint32_t* n_dev, long long n_cap, uint2* ranges,
                                  hipStream_t st) {
  if (n_cap <= 0) return hipSuccess;
  tile_ranges_u32_kernel<<<(unsigned)((n_cap + 1023) / 1024), 256, 0, st>>>(tiles, n_dev, n_cap, ranges);
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, long long n, uint32_t v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  fill_u32_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, n, v);
  return hipGetLastError();
}

hipError_t launch_tile_ranges(const uint64_t* keys, const uint32_t* n_dev, long long n_cap, int rank_bits,
                              uint2* ranges, hipStream_t st) {
  if (n_cap <= 0) return hipSuccess;
  tile_ranges_kernel<<<(unsigned)((n_cap + 255) / 256), 256, 0, st>>>(keys, n_dev, n_cap, rank_bits, ranges);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Voxel <-> grid-point link (octree_utils.py:129-144, 203-222): the 8 corners of every voxel as
// integer coordinates at the finest level, deduplicated.  The reference takes
// unique(dim=0, return_inverse=True) over the [8N, 3] rows, i.e. grid points in lexicographic
// (x, y, z) order; the same order is the order of the packed code x << 34 | y << 17 | z
// (coordinates <= 2^16 take 17 bits), so a stable radix sort of the codes, a flag where the
// code changes and a scan give the grid points in that order and every corner's index.
__global__ __launch_bounds__(256) void corner_codes_kernel(long long n, const int64_t* __restrict__ octpath,
                                                           const int8_t* __restrict__ octlevel,
                                                           uint64_t* __restrict__ codes,
                                                           uint32_t* __restrict__ idx) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long v = t >> 3;
  const int c = (int)(t & 7);
  const int lv = octlevel[v];
  uint64_t path = (uint64_t)octpath[v] >> (3 * (kMaxLevels - lv));
  uint64_t i = 0, j = 0, k = 0;
  for (int l = 0; l < kMaxLevels; ++l) {
    const uint64_t b = path & 7u;
    i |= ((b >> 2) & 1u) << l;
    j |= ((b >> 1) & 1u) << l;
    k |= (b & 1u) << l;
    path >>= 3;
  }
  const int sh = kMaxLevels - lv;
  const uint64_t x = (i + ((c >> 2) & 1)) << sh, y = (j + ((c >> 1) & 1)) << sh, z = (k + (c & 1)) << sh;
  codes[t] = (x << 34) | (y << 17) | z;
  idx[t] = (uint32_t)t;
}

__global__ __launch_bounds__(256) void code_change_kernel(long long n, const uint64_t* __restrict__ codes,
                                                          uint32_t* __restrict__ flags) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  flags[p] = (p == 0 || codes[p] != codes[p - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void link_write_kernel(long long n, const uint64_t* __restrict__ codes,
                                                         const uint32_t* __restrict__ idx,
                                                         const uint32_t* __restrict__ flags,
                                                         const uint32_t* __restrict__ rank_incl,
                                                         int64_t* __restrict__ vox_key,
                                                         int64_t* __restrict__ grid_key) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const long long r = (long long)rank_incl[p] - 1;
  vox_key[idx[p]] = r;
  if (flags[p]) {
    const uint64_t c = codes[p];
    grid_key[3 * r] = (int64_t)(c >> 34);
    grid_key[3 * r + 1] = (int64_t)((c >> 17) & 0x1FFFFu);
    grid_key[3 * r + 2] = (int64_t)(c & 0x1FFFFu);
  }
}

hipError_t launch_grid_pts_link(long long N, const int64_t* octpath, const int8_t* octlevel, uint64_t* keys,
                                uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t* hist,
                                uint32_t* sort_partial, uint32_t* scan_partial, uint32_t* flags, uint32_t* rank,
                                int64_t* vox_key, int64_t* grid_key, uint32_t* n_unique, hipStream_t st) {
  const long long n = 8 * N;
  if (n <= 0) return hipSuccess;
  const unsigned nb = (unsigned)((n + 255) / 256);
  corner_codes_kernel<<<nb, 256, 0, st>>>(n, octpath, octlevel, keys, vals);
  bool in_alt = false;
  hipError_t e = launch_radix_sort(keys, vals, keys_alt, vals_alt, nullptr, n, 51, hist, sort_partial, &in_alt, st);
  if (e != hipSuccess) return e;
  const uint64_t* ks = in_alt ? keys_alt : keys;
  const uint32_t* vs = in_alt ? vals_alt : vals;
  code_change_kernel<<<nb, 256, 0, st>>>(n, ks, flags);
  e = launch_scan(flags, nullptr, n, scan_partial, rank, true, st);
  if (e != hipSuccess) return e;
  link_write_kernel<<<nb, 256, 0, st>>>(n, ks, vs, flags, rank, vox_key, grid_key);
  e = hipMemcpyAsync(n_unique, rank + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace svr

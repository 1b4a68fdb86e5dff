// C-ABI entry points of libsvraster_hip.so (declared in include/svraster_hip.h).
//
// Host orchestration of one rasterizer call, restating the steps of the reference's host
// functions (cuda/src/preprocess.cu:158-228, forward.cu:832-982 and 986-1096,
// backward.cu:986-1109, sh_compute.cu:244-299, geo_params_gather.cu:134-173,
// adam_step.cu:106-158, ge_compute.cu:120-161, ls_compute.cu:112-155, utils.cu:329-408) on
// top of this library's kernels.  All launches go to the caller's stream.
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/svraster_hip.h"
#include "common.h"
#include "ops.h"
#include "state.h"

using namespace svr;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, const char* a = "", long long b = 0) {
  char buf[512];
  snprintf(buf, sizeof(buf), fmt, a, b);
  g_err = buf;
  return code;
}

int fail_hip(hipError_t e, const char* where) {
  char buf[512];
  snprintf(buf, sizeof(buf), "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), where);
  g_err = buf;
  return 2;
}

#define SVR_TRY(expr)                                      \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return fail_hip(_e, #expr);      \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Optional HIP-event timing of the compositor launches (enabled by svr_timing_enable).
constexpr int kTimedKinds = 2;  // 0 = forward compositor, 1 = backward compositor
struct TimedLaunch { hipEvent_t a, b; };
std::mutex g_timing_mu;
bool g_timing_on = false;
std::vector<TimedLaunch> g_timing[kTimedKinds];

struct LaunchTimer {
  int kind; hipStream_t st; TimedLaunch ev{nullptr, nullptr}; bool on = false;
  LaunchTimer(int k, hipStream_t s) : kind(k), st(s) {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    on = g_timing_on;
    if (on && hipEventCreate(&ev.a) == hipSuccess && hipEventCreate(&ev.b) == hipSuccess) {
      (void)hipEventRecord(ev.a, st);
    } else {
      on = false;
    }
  }
  ~LaunchTimer() {
    if (!on) return;
    (void)hipEventRecord(ev.b, st);
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_timing[kind].push_back(ev);
  }
};

// Runtime options (svr_set_option): compositor variant, 0 = pair-compacted, 1 = one entry at a time;
// batch_depth = entries staged per round by the multi-view pair-compacted compositors (64 or 128);
// persistent_blocks_per_cu = workgroups per CU of the multi-view persistent launches.
int g_compositor_variant = 0;
int g_batch_depth = 128;
// Entries staged per round by the reverse compositor (64 or 128): with 64 its LDS and register
// budget fit five workgroups per CU instead of four.
int g_bwd_batch_depth = 64;
int g_blocks_per_cu = 8;

int tile_bits_for(long long n_tiles) {
  int b = 0;
  while ((1LL << b) < n_tiles) ++b;
  return b;
}

ViewDesc make_view(int W, int H, float tanfovx, float tanfovy, float cx, float cy, float near_plane,
                   int tile_w, int tile_h) {
  ViewDesc v;
  memset(&v, 0, sizeof(v));
  v.W = W; v.H = H;
  v.tile_w = tile_w; v.tile_h = tile_h;
  v.tiles_x = ceil_div(W, tile_w);
  v.tiles_y = ceil_div(H, tile_h);
  v.tanfovx = tanfovx; v.tanfovy = tanfovy; v.cx = cx; v.cy = cy;
  // Focal lengths in FP32 with the reference's operation order (preprocess.cu:197-198).
  v.focal_x = (0.5f * (float)W) / tanfovx;
  v.focal_y = (0.5f * (float)H) / tanfovy;
  v.near_ = near_plane;
  v.tile_base = 0;
  v.pix_base = 0;
  return v;
}

int check_modes(int vox_geo_mode, int density_mode, int cam_mode) {
  if (density_mode != kDensitySDF)
    return fail(1, "density_mode %s%lld is not supported (only SDF = 1)", "", density_mode);
  if (vox_geo_mode != kVoxGeoTriinterp3)
    return fail(1, "vox_geo_mode %s%lld is not supported (only triinterp3 = 3)", "", vox_geo_mode);
  if (cam_mode != kCamPersp)
    return fail(1, "cam_mode %s%lld is not supported (only perspective = 0)", "", cam_mode);
  return 0;
}

int check_image(int W, int H, int P) {
  if (W <= 0 || H <= 0 || W > 65535 || H > 65535)
    return fail(1, "image size must be within 1..65535 (got W=%s%lld)", "", W);
  if ((long long)ceil_div(W, kTile) * ceil_div(H, kTile) > 65536)
    return fail(1, "at most 65536 tiles per view%s (got %lld)", "", (long long)ceil_div(W, kTile) * ceil_div(H, kTile));
  if (P < 0 || P >= (1 << 29)) return fail(1, "voxel count must be below 2^29%s (got %lld)", "", P);
  return 0;
}

struct KeyLayout {
  int shift, qb, rank_bits, nbits;
};

KeyLayout key_layout(int rank_shift, long long n_tiles) {
  KeyLayout k;
  k.shift = rank_shift;
  k.qb = rank_shift >= 3 ? 3 : 0;
  k.rank_bits = kRankBits - rank_shift + k.qb;
  k.nbits = k.rank_bits + tile_bits_for(n_tiles);
  return k;
}

// Fixed device scratch of the allocation-free entry points: per-grid-point voxel counts of
// svr_agg_voxel_into_grid_pts_noalloc, at most kScratchPoints grid points (16 MiB).  Each GPU
// the library runs on holds its own copy until the process ends.  Host code must not pass the
// array's name to a kernel (in host code it is only a stand-in for the symbol): scratch_ptr()
// looks up its device address once per device with hipGetSymbolAddress and keeps it.
constexpr long long kScratchPoints = 1LL << 22;
static __device__ __align__(256) uint32_t scratch[kScratchPoints];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (nullptr and a message on failure).
uint32_t* scratch_ptr() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    fail(2, "no current HIP device%s (device %lld)", "", dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fail_hip(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
      return nullptr;
    }
    g_scratch_ptr[dev] = p;
  }
  return reinterpret_cast<uint32_t*>(g_scratch_ptr[dev]);
}

}  // namespace

extern "C" {

const char* svr_last_error(void) { return g_err.c_str(); }

void svr_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(g_timing_mu);
  g_timing_on = on != 0;
}

int svr_timing_read(int kind, int reset, double* total_ms, long long* launches) {
  g_err.clear();
  *total_ms = 0.0;
  *launches = 0;
  if (kind < 0 || kind >= kTimedKinds) return fail(1, "timing kind out of range%s (%lld)", "", kind);
  std::vector<TimedLaunch> evs;
  {
    std::lock_guard<std::mutex> lk(g_timing_mu);
    evs = g_timing[kind];
    if (reset) g_timing[kind].clear();
  }
  for (auto& e : evs) {
    SVR_TRY(hipEventSynchronize(e.b));
    float ms = 0.f;
    SVR_TRY(hipEventElapsedTime(&ms, e.a, e.b));
    *total_ms += ms;
    *launches += 1;
    if (reset) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
  }
  return 0;
}
int svr_abi_version(void) { return 5; }

int svr_set_option(const char* name, int value) {
  g_err.clear();
  if (name && strcmp(name, "compositor_variant") == 0) {
    if (value < 0 || value > 1) return fail(1, "compositor_variant must be 0 or 1%s (got %lld)", "", value);
    g_compositor_variant = value;
    return 0;
  }
  if (name && strcmp(name, "batch_depth") == 0) {
    if (value != 64 && value != 128) return fail(1, "batch_depth must be 64 or 128%s (got %lld)", "", value);
    g_batch_depth = value;
    return 0;
  }
  if (name && strcmp(name, "bwd_batch_depth") == 0) {
    if (value != 64 && value != 128) return fail(1, "bwd_batch_depth must be 64 or 128%s (got %lld)", "", value);
    g_bwd_batch_depth = value;
    return 0;
  }
  if (name && strcmp(name, "persistent_blocks_per_cu") == 0) {
    if (value < 1 || value > 32) return fail(1, "persistent_blocks_per_cu must be 1..32%s (got %lld)", "", value);
    g_blocks_per_cu = value;
    return 0;
  }
  return fail(1, "unknown option %s%lld", name ? name : "(null)", 0);
}

int svr_get_option(const char* name) {
  if (name && strcmp(name, "compositor_variant") == 0) return g_compositor_variant;
  if (name && strcmp(name, "batch_depth") == 0) return g_batch_depth;
  if (name && strcmp(name, "bwd_batch_depth") == 0) return g_bwd_batch_depth;
  if (name && strcmp(name, "persistent_blocks_per_cu") == 0) return g_blocks_per_cu;
  return -1;
}

size_t svr_geom_state_bytes(int n_views, int P) { return GeomLayout(n_views, P).total; }
size_t svr_binning_state_bytes(long long K) { return raster_bin_bytes(K); }
size_t svr_image_state_bytes(int n_views, long long n_pixels, long long n_tiles) {
  return ImageLayout(n_views, n_pixels, n_tiles).total;
}

int svr_rasterize_preprocess(int W, int H, float tanfovx, float tanfovy, float cx, float cy,
                             const float* w2c, const float* c2w, int cam_mode, float near_plane,
                             const int64_t* octree_paths, const float* vox_centers,
                             const float* vox_lengths, const uint8_t* is_leaf, int P,
                             int32_t* out_n_dup, void* geom_buffer, void* stream) {
  g_err.clear();
  if (cam_mode != kCamPersp) return fail(1, "cam_mode %s%lld is not supported (only perspective = 0)", "", cam_mode);
  if (int e = check_image(W, H, P)) return e;
  if (P == 0) return 0;
  hipStream_t st = as_stream(stream);
  GeomPtrs gp(geom_buffer, 1, P);
  const ViewDesc v = make_view(W, H, tanfovx, tanfovy, cx, cy, near_plane, kTile, kTile);
  SVR_TRY(launch_init_view(gp.views, v, w2c, c2w, nullptr, st));
  SVR_TRY(launch_fill_u32(gp.misc, 1, (uint32_t)kRankBits, st));
  SVR_TRY(launch_preprocess(gp.views, 1, P, vox_centers, vox_lengths, is_leaf, octree_paths, out_n_dup,
                            gp.n_dup, gp.bbox, gp.tbox, gp.quad, gp.misc, st));
  return 0;
}

int svr_rasterize_count(void* geom_buffer, int P, long long* K_out, int* rank_shift_out, void* stream) {
  g_err.clear();
  *K_out = 0;
  *rank_shift_out = 0;
  if (P <= 0) return 0;
  hipStream_t st = as_stream(stream);
  GeomPtrs gp(geom_buffer, 1, P);
  SVR_TRY(launch_scan(gp.n_dup, nullptr, P, gp.partial, gp.scan, true, st));
  uint32_t host[2] = {0, 0};
  SVR_TRY(hipMemcpyAsync(&host[0], gp.scan + (P - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SVR_TRY(hipMemcpyAsync(&host[1], gp.misc, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SVR_TRY(hipStreamSynchronize(st));
  *K_out = (long long)host[0];
  *rank_shift_out = (int)(host[1] > (uint32_t)kRankBits ? kRankBits : host[1]);
  return 0;
}

int svr_rasterize_forward(int vox_geo_mode, int density_mode, int W, int H, float tanfovx,
                          float tanfovy, float cx, float cy, const float* w2c, const float* c2w,
                          const float* background, int cam_mode, int need_depth,
                          int need_distortion, int need_normal, const int64_t* octree_paths,
                          const float* vox_centers, const float* vox_lengths, const float* geos,
                          const float* rgbs, int P, void* geom_buffer, long long K, int rank_shift,
                          void* binning_buffer, void* image_buffer, float* out_color,
                          float* out_depth, float* out_normal, float* out_T, float* max_w,
                          const float* s_val, void* stream) {
  g_err.clear();
  if (int e = check_modes(vox_geo_mode, density_mode, cam_mode)) return e;
  if (int e = check_image(W, H, P)) return e;
  if (need_distortion) return fail(1, "distortion output is not supported%s%lld", "", 0);
  if (rank_shift < 0 || rank_shift > kRankBits || rank_shift % 3) return fail(1, "bad rank_shift %s%lld", "", rank_shift);
  if (K < 0 || K >= (1LL << 31)) return fail(1, "key count out of range%s (%lld)", "", K);
  if (P == 0) return 0;
  hipStream_t st = as_stream(stream);
  const long long n_tiles = (long long)ceil_div(W, kTile) * ceil_div(H, kTile);
  GeomPtrs gp(geom_buffer, 1, P);
  BinPtrs bp(binning_buffer, K);
  ImagePtrs ip(image_buffer, 1, (long long)W * H, n_tiles);
  const ViewDesc v = make_view(W, H, tanfovx, tanfovy, cx, cy, 0.f, kTile, kTile);
  SVR_TRY(launch_init_view(ip.views, v, w2c, c2w, background, st));
  const uint32_t* n_dev = gp.scan + (P - 1);

  const KeyLayout kl = key_layout(rank_shift, n_tiles);
  const int passes = (kl.nbits + 7) / 8;
  // Duplicate into the buffer pair that makes the last sort pass land in the primary pair.
  const bool start_alt = (passes % 2) == 1;
  uint64_t* k0 = start_alt ? bp.keys_alt : bp.keys;
  uint32_t* v0 = start_alt ? bp.vals_alt : bp.vals;
  uint64_t* k1 = start_alt ? bp.keys : bp.keys_alt;
  uint32_t* v1 = start_alt ? bp.vals : bp.vals_alt;
  if (K > 0) {
    SVR_TRY(launch_duplicate(ip.views, 1, P, octree_paths, gp.n_dup, gp.scan, gp.bbox, gp.quad, kl.shift, kl.qb,
                             kl.rank_bits, k0, v0, K, st));
    bool in_alt = false;
    SVR_TRY(launch_radix_sort(k0, v0, k1, v1, n_dev, K, kl.nbits, bp.hist, bp.partial, &in_alt, st));
  }
  SVR_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(ip.ranges), 2 * n_tiles, 0u, st));
  if (K > 0) SVR_TRY(launch_tile_ranges(bp.keys, n_dev, K, kl.rank_bits, ip.ranges, st));

  FwdArgs a;
  memset(&a, 0, sizeof(a));
  a.views = ip.views; a.n_views = 1; a.total_tiles = (int)n_tiles;
  a.ranges = ip.ranges; a.vals = bp.vals;
  a.bboxes = gp.bbox; a.tboxes = gp.tbox; a.P = P;
  a.centers = vox_centers; a.lengths = vox_lengths; a.geos = geos; a.rgbs = rgbs;
  a.s_val = s_val;
  a.tile_last = ip.tile_last; a.n_contrib = ip.n_contrib;
  a.color = out_color; a.depth = out_depth; a.normal = out_normal; a.T = out_T; a.max_w = max_w;
  a.work_counter = nullptr;
  const ContribPtrs cp((char*)binning_buffer + BinLayout(K).total, K);
  a.cvals = cp.cvals; a.pmask = cp.pmask; a.cranges = ip.cranges;
  a.need_depth = need_depth; a.need_normal = need_normal;
  a.variant = g_compositor_variant;
  a.halves = g_batch_depth / 64;
  a.blocks_per_cu = g_blocks_per_cu;
  {
    LaunchTimer timer(0, st);
    SVR_TRY(launch_render_forward(a, st));
  }
  return 0;
}

int svr_rasterize_backward(long long K, int vox_geo_mode, int density_mode, int W, int H,
                           float tanfovx, float tanfovy, float cx, float cy, const float* w2c,
                           const float* c2w, const float* background, const float* vox_centers,
                           const float* vox_lengths, const float* geos, const float* rgbs, int P,
                           void* geom_buffer, void* binning_buffer, void* image_buffer,
                           const float* out_T, const float* dL_dcolor, const float* dL_ddepth,
                           const float* dL_dnormal, const float* dL_dT, float lambda_R_concen,
                           const float* gt_color, float lambda_ascending, float lambda_dist,
                           int need_depth, int need_normal, const float* s_val, float* dL_dvox,
                           float* dL_ds, void* stream) {
  g_err.clear();
  if (int e = check_modes(vox_geo_mode, density_mode, kCamPersp)) return e;
  if (int e = check_image(W, H, P)) return e;
  if (lambda_ascending != 0.f) return fail(1, "lambda_ascending > 0 is not supported%s%lld", "", 0);
  if (lambda_dist != 0.f) return fail(1, "lambda_dist > 0 is not supported%s%lld", "", 0);
  if (need_depth && !dL_ddepth) return fail(1, "need_depth requires dL_ddepth%s%lld", "", 0);
  if (need_normal && !dL_dnormal) return fail(1, "need_normal requires dL_dnormal%s%lld", "", 0);
  if (lambda_R_concen > 0.f && !gt_color) return fail(1, "lambda_R_concen > 0 requires gt_color%s%lld", "", 0);
  if (P == 0 || K <= 0) return 0;
  hipStream_t st = as_stream(stream);
  const long long n_tiles = (long long)ceil_div(W, kTile) * ceil_div(H, kTile);
  GeomPtrs gp(geom_buffer, 1, P);
  BinPtrs bp(binning_buffer, K);
  ImagePtrs ip(image_buffer, 1, (long long)W * H, n_tiles);
  const ViewDesc v = make_view(W, H, tanfovx, tanfovy, cx, cy, 0.f, kTile, kTile);
  SVR_TRY(launch_init_view(ip.views, v, w2c, c2w, background, st));
  BwdArgs a;
  memset(&a, 0, sizeof(a));
  a.views = ip.views; a.n_views = 1; a.total_tiles = (int)n_tiles;
  a.ranges = ip.ranges; a.tile_last = ip.tile_last; a.vals = bp.vals;
  a.bboxes = gp.bbox; a.P = P;
  a.centers = vox_centers; a.lengths = vox_lengths; a.geos = geos; a.rgbs = rgbs;
  a.s_val = s_val;
  a.T = out_T; a.n_contrib = ip.n_contrib;
  a.dL_dcolor = dL_dcolor; a.dL_ddepth = dL_ddepth; a.dL_dnormal = dL_dnormal; a.dL_dT = dL_dT;
  a.gt_color = gt_color; a.lambda_R_concen = lambda_R_concen;
  a.dL_dvox = dL_dvox; a.dL_ds = dL_ds; a.ds_part = ip.ds_part;
  a.work_counter = nullptr;
  a.need_depth = need_depth; a.need_normal = need_normal;
  a.variant = g_compositor_variant;
  a.halves = g_bwd_batch_depth / 64;
  a.blocks_per_cu = g_blocks_per_cu;
  const ContribPtrs cp((char*)binning_buffer + BinLayout(K).total, K);
  a.cvals = cp.cvals; a.pmask = cp.pmask; a.cranges = ip.cranges;
  {
    LaunchTimer timer(1, st);
    SVR_TRY(launch_render_backward(a, st));
  }
  return 0;
}

int svr_sh_compute(int degree, const int64_t* idx, int n_idx, const float* vox_centers, int P,
                   const float* cam_pos, const float* sh0, const float* shs, int M, float* rgbs,
                   void* stream) {
  g_err.clear();
  if (degree < 0 || degree > 3) return fail(1, "SH degree must be 0..3%s (got %lld)", "", degree);
  if (M < (degree + 1) * (degree + 1)) return fail(1, "shs has too few coefficients%s (M=%lld)", "", M);
  const int n = idx ? n_idx : P;
  SVR_TRY(launch_sh_forward(n, idx, degree, M, vox_centers, cam_pos, sh0, shs, rgbs, as_stream(stream)));
  return 0;
}

int svr_sh_compute_bw(int degree, int M, const int64_t* idx, int n_idx, const float* vox_centers,
                      int P, const float* cam_pos, const float* dL_drgbs, float* dL_dsh0,
                      float* dL_dshs, void* stream) {
  g_err.clear();
  if (degree < 0 || degree > 3) return fail(1, "SH degree must be 0..3%s (got %lld)", "", degree);
  const int n = idx ? n_idx : P;
  SVR_TRY(launch_sh_backward(n, idx, degree, M, vox_centers, cam_pos, dL_drgbs, dL_dsh0, dL_dshs,
                             as_stream(stream)));
  return 0;
}

int svr_gather_triinterp_geo_params(const int64_t* vox_key, int P, const int64_t* care_idx,
                                    int n_care, const float* grid_pts, float* geos, void* stream) {
  g_err.clear();
  (void)P;
  SVR_TRY(launch_gather_geo(n_care, vox_key, care_idx, grid_pts, geos, as_stream(stream)));
  return 0;
}

int svr_gather_triinterp_geo_params_bw(const int64_t* vox_key, int P, const int64_t* care_idx,
                                       int n_care, const float* dL_dgeos, float* dL_dgrid_pts,
                                       void* stream) {
  g_err.clear();
  (void)P;
  SVR_TRY(launch_gather_geo_bw(n_care, vox_key, care_idx, dL_dgeos, dL_dgrid_pts, as_stream(stream)));
  return 0;
}

int svr_batch_gather_bw(int P, const int64_t* vox_key, const float* dL_dvox, float* dL_dgrid_pts, void* stream) {
  g_err.clear();
  if (P < 0) return fail(1, "bad voxel count%s %lld", "", P);
  if (P > 0 && (!vox_key || !dL_dvox || !dL_dgrid_pts))
    return fail(1, "vox_key, dL_dvox and dL_dgrid_pts are required%s%lld", "", 0);
  SVR_TRY(launch_gather_geo_bw(P, vox_key, nullptr, dL_dvox, dL_dgrid_pts, as_stream(stream), 12));
  return 0;
}

int svr_unbiased_adam_step(int sparse, float* param, const float* grad, float* exp_avg,
                           float* exp_avg_sq, long long numel, double step, double lr,
                           double beta1, double beta2, float eps, void* stream) {
  g_err.clear();
  // Bias corrections in double on the host (adam_step.cu:117-122).
  const double bc1 = 1.0 - pow(beta1, step);
  const double bc2 = 1.0 - pow(beta2, step);
  const double step_size = lr / bc1;
  const double rsqrt_bc2 = 1.0 / sqrt(bc2);
  SVR_TRY(launch_adam(sparse != 0, false, numel, param, grad, exp_avg, exp_avg_sq, (float)step_size,
                      (float)beta1, (float)beta2, (float)rsqrt_bc2, eps, as_stream(stream)));
  return 0;
}

int svr_adam_bias_table(double lr, double beta1, double beta2, int n, float* out) {
  g_err.clear();
  if (n <= 0 || !out) return fail(1, "table length%s %lld must be positive", "", n);
  for (int t = 1; t <= n; ++t) {
    // Same double expressions as svr_unbiased_adam_step (adam_step.cu:117-122).
    const double bc1 = 1.0 - pow(beta1, (double)t);
    const double bc2 = 1.0 - pow(beta2, (double)t);
    out[2 * (t - 1)] = (float)(lr / bc1);
    out[2 * (t - 1) + 1] = (float)(1.0 / sqrt(bc2));
  }
  return 0;
}

static AdamTable adam_table(const int32_t* step, const float* table, int n, float b1, float b2, float eps) {
  AdamTable T;
  T.step = step; T.tab = reinterpret_cast<const float2*>(table); T.n = n; T.b1 = b1; T.b2 = b2; T.eps = eps;
  return T;
}

int svr_sparse_adam_step_dev(int sparse, float* param, const float* grad, float* exp_avg, float* exp_avg_sq,
                             long long numel, const int32_t* step, const float* table, int n_table, float beta1,
                             float beta2, float eps, void* stream) {
  g_err.clear();
  if (!step || !table || n_table <= 0) return fail(1, "step counter and bias table%s are required", "", 0);
  SVR_TRY(launch_adam_table(sparse != 0, numel, param, grad, exp_avg, exp_avg_sq,
                            adam_table(step, table, n_table, beta1, beta2, eps), as_stream(stream)));
  return 0;
}

int svr_biased_adam_step(int sparse, float* param, const float* grad, float* exp_avg,
                         float* exp_avg_sq, long long numel, float lr, float beta1, float beta2,
                         float eps, void* stream) {
  g_err.clear();
  SVR_TRY(launch_adam(sparse != 0, true, numel, param, grad, exp_avg, exp_avg_sq, lr, beta1, beta2, 1.f, eps,
                      as_stream(stream)));
  return 0;
}

static int grid_loss_common(bool lap, const float* grid_pts, long long n_pts, const int64_t* vox_key, int N,
                            const float* coord, const float* size, int res, const uint8_t* mask,
                            const int32_t* keys, const int32_t* key2vox, int G, const int32_t* active, int A,
                            float weight, float inv, unsigned seed, const float* randv, float* grad,
                            void* stream, const int32_t* dense = nullptr, const int32_t* step = nullptr) {
  g_err.clear();
  if (res <= 2) return fail(1, "grid_res must exceed 2%s (got %lld)", "", res);
  if ((long long)res * res * res >= (1LL << 31)) return fail(1, "grid_res%s %lld too large (res^3 must fit int32)", "", res);
  GridTable T;
  T.grid_pts = grid_pts; T.vox_key = vox_key; T.keys = keys; T.key2vox = key2vox;
  T.vox_coord = coord; T.vox_size = size; T.n_keys = G; T.res = res; T.n_vox = N; T.n_pts = n_pts;
  T.dense = dense;
  SVR_TRY(launch_grid_loss(lap, T, reinterpret_cast<const bool*>(mask), active, A, weight, inv, seed, step, randv,
                           grad, as_stream(stream)));
  return 0;
}

int svr_grid_eikonal_bw(const float* grid_pts, long long n_pts, const int64_t* vox_key, int N,
                        const float* grid_voxel_coord, const float* grid_voxel_size, int grid_res,
                        const uint8_t* grid_mask, const int32_t* grid_keys,
                        const int32_t* grid2voxel, int G, const int32_t* active_list, int A,
                        float weight, float vox_size_inv, unsigned int seed, const float* randv,
                        float* grid_pts_grad, void* stream) {
  return grid_loss_common(false, grid_pts, n_pts, vox_key, N, grid_voxel_coord, grid_voxel_size, grid_res,
                          grid_mask, grid_keys, grid2voxel, G, active_list, A, weight, vox_size_inv, seed, randv,
                          grid_pts_grad, stream);
}

int svr_laplacian_smoothness_bw(const float* grid_pts, long long n_pts, const int64_t* vox_key,
                                int N, const float* grid_voxel_coord,
                                const float* grid_voxel_size, int grid_res,
                                const uint8_t* grid_mask, const int32_t* grid_keys,
                                const int32_t* grid2voxel, int G, const int32_t* active_list,
                                int A, float weight, float vox_size_inv, unsigned int seed,
                                const float* randv, float* grid_pts_grad, void* stream) {
  return grid_loss_common(true, grid_pts, n_pts, vox_key, N, grid_voxel_coord, grid_voxel_size, grid_res,
                          grid_mask, grid_keys, grid2voxel, G, active_list, A, weight, vox_size_inv, seed, randv,
                          grid_pts_grad, stream);
}

int svr_total_variation_bw(const float* grid_pts, long long M, int C, const int64_t* vox_key, long long N,
                           float weight, const float* vox_size_inv, int no_tv_s, int tv_sparse,
                           uint8_t* sparse_ws, float* grid_pts_grad, void* stream) {
  g_err.clear();
  if (C <= 0) return fail(1, "channel count%s %lld must be positive", "", C);
  if (N <= 0) return 0;
  if (tv_sparse && !sparse_ws) return fail(1, "tv_sparse needs a workspace of N*C bytes%s%lld", "", 0);
  if (!no_tv_s && !vox_size_inv) return fail(1, "vox_size_inv is required unless no_tv_s%s%lld", "", 0);
  (void)M;
  SVR_TRY(launch_total_variation(N, C, grid_pts, vox_key, weight, vox_size_inv, no_tv_s != 0,
                                 tv_sparse ? sparse_ws : nullptr, grid_pts_grad, as_stream(stream)));
  return 0;
}

int svr_voxel_gradient_bw(const float* grid_pts, long long M, const int64_t* vox_key, long long N,
                          const float* vox_size_inv, const int32_t* active_list, int A, float weight,
                          float* grid_pts_grad, void* stream) {
  g_err.clear();
  (void)M; (void)N;
  SVR_TRY(launch_voxel_gradient(A, active_list, grid_pts, vox_key, vox_size_inv, weight, grid_pts_grad,
                                as_stream(stream)));
  return 0;
}

int svr_points_loss_bw(const float* points_in_grid, int Mp, const float* grid_pts, long long M,
                       const int64_t* vox_key, int N, const float* grid_voxel_coord, const float* grid_voxel_size,
                       int grid_res, const uint8_t* grid_mask, const int32_t* grid_keys, const int32_t* grid2voxel,
                       int G, float weight, float* grid_pts_grad, void* stream) {
  g_err.clear();
  if (grid_res <= 0 || (long long)grid_res * grid_res * grid_res >= (1LL << 31))
    return fail(1, "grid_res%s %lld out of range", "", grid_res);
  GridTable T;
  T.grid_pts = grid_pts; T.vox_key = vox_key; T.keys = grid_keys; T.key2vox = grid2voxel;
  T.vox_coord = grid_voxel_coord; T.vox_size = grid_voxel_size; T.n_keys = G; T.res = grid_res; T.n_vox = N;
  T.n_pts = M;
  SVR_TRY(launch_points_loss(Mp, points_in_grid, T, reinterpret_cast<const bool*>(grid_mask), weight,
                             grid_pts_grad, as_stream(stream)));
  return 0;
}

int svr_grid_dense_index(const int32_t* grid_keys, int G, int grid_res, int32_t* out, void* stream) {
  g_err.clear();
  if (grid_res <= 2 || (long long)grid_res * grid_res * grid_res >= (1LL << 31))
    return fail(1, "grid_res%s %lld out of range", "", grid_res);
  SVR_TRY(launch_grid_dense_index((long long)grid_res * grid_res * grid_res, grid_keys, G, out, as_stream(stream)));
  return 0;
}

int svr_grid_smoothness_bw_step(int laplacian, const float* grid_pts, long long n_pts, const int64_t* vox_key, int N,
                                const float* grid_voxel_coord, const float* grid_voxel_size, int grid_res,
                                const uint8_t* grid_mask, const int32_t* grid_keys, const int32_t* grid2voxel, int G,
                                const int32_t* dense_index, const int32_t* active_list, int A, float weight,
                                float vox_size_inv, unsigned int seed, const int32_t* step, float* grid_pts_grad,
                                void* stream) {
  g_err.clear();
  if (!step || !active_list) return fail(1, "step counter and active list%s are required", "", 0);
  return grid_loss_common(laplacian != 0, grid_pts, n_pts, vox_key, N, grid_voxel_coord, grid_voxel_size, grid_res,
                          grid_mask, grid_keys, grid2voxel, G, active_list, A, weight, vox_size_inv, seed, nullptr,
                          grid_pts_grad, stream, dense_index, step);
}

static int grid_table_order(const float* grid_pts, const int64_t* vox_key, int N, const float* grid_voxel_coord,
                            const float* grid_voxel_size, int grid_res, const uint8_t* grid_mask,
                            const int32_t* table_keys, const int32_t* table_tvox, int G, const int32_t* cell_entry,
                            const int32_t* cell_tvox, const int32_t* tvox, int T, int terms, float weight,
                            const float* weights, float vox_size_inv, unsigned int seed, const int32_t* step,
                            float* rec_ws, float* coef_ws, float* grid_pts_grad, void* stream) {
  if (!step || !table_keys || !table_tvox || !cell_entry || !cell_tvox || !tvox || !rec_ws || !coef_ws)
    return fail(1, "missing table-order argument%s (all pointers are required)", "", 0);
  if (grid_res <= 2 || (long long)grid_res * grid_res * grid_res >= (1LL << 31))
    return fail(1, "grid_res%s %lld out of range", "", grid_res);
  if (T < 0 || T > N) return fail(1, "table voxel count%s %lld out of range", "", T);
  if (terms < 1 || terms > 3) return fail(1, "terms must be 1 (Laplacian), 2 (eikonal) or 3 (both)%s, got %lld", "",
                                          terms);
  if ((terms & 2) && !weights) return fail(1, "the eikonal term reads its weight from `weights`%s", "", 0);
  GridOrderArgs A;
  A.grid_pts = grid_pts; A.vox_key = vox_key; A.coord = grid_voxel_coord; A.size = grid_voxel_size;
  A.mask = reinterpret_cast<const bool*>(grid_mask);
  A.keys = table_keys; A.tv = table_tvox; A.cell_entry = cell_entry; A.cell_tv = cell_tvox; A.tvox = tvox;
  A.rec = rec_ws; A.coef = coef_ws; A.grad = grid_pts_grad;
  A.G = G; A.T = T; A.res = grid_res; A.weight = weight; A.inv = vox_size_inv; A.seed = seed; A.step = step;
  A.terms = terms; A.weights = weights;
  SVR_TRY(launch_grid_laplacian_pull(A, as_stream(stream)));
  return 0;
}

int svr_grid_laplacian_bw_table_order(const float* grid_pts, long long n_pts, const int64_t* vox_key, int N,
                                      const float* grid_voxel_coord, const float* grid_voxel_size, int grid_res,
                                      const uint8_t* grid_mask, const int32_t* table_keys,
                                      const int32_t* table_tvox, int G, const int32_t* cell_entry,
                                      const int32_t* cell_tvox, const int32_t* tvox, int T, float weight,
                                      float vox_size_inv, unsigned int seed, const int32_t* step, float* rec_ws,
                                      float* coef_ws, float* grid_pts_grad, void* stream) {
  g_err.clear();
  (void)n_pts;
  return grid_table_order(grid_pts, vox_key, N, grid_voxel_coord, grid_voxel_size, grid_res, grid_mask, table_keys,
                          table_tvox, G, cell_entry, cell_tvox, tvox, T, 1, weight, nullptr, vox_size_inv, seed,
                          step, rec_ws, coef_ws, grid_pts_grad, stream);
}

int svr_grid_smoothness_bw_table_order(const float* grid_pts, long long n_pts, const int64_t* vox_key, int N,
                                       const float* grid_voxel_coord, const float* grid_voxel_size, int grid_res,
                                       const uint8_t* grid_mask, const int32_t* table_keys,
                                       const int32_t* table_tvox, int G, const int32_t* cell_entry,
                                       const int32_t* cell_tvox, const int32_t* tvox, int T, int terms,
                                       const float* weights, float vox_size_inv, unsigned int seed,
                                       const int32_t* step, float* rec_ws, float* coef_ws, float* grid_pts_grad,
                                       void* stream) {
  g_err.clear();
  (void)n_pts;
  if (!weights) return fail(1, "weights (device float[2]: eikonal, Laplacian) are required%s", "", 0);
  return grid_table_order(grid_pts, vox_key, N, grid_voxel_coord, grid_voxel_size, grid_res, grid_mask, table_keys,
                          table_tvox, G, cell_entry, cell_tvox, tvox, T, terms, 0.f, weights, vox_size_inv, seed,
                          step, rec_ws, coef_ws, grid_pts_grad, stream);
}

// Workspace of svr_build_grid_pts_link for N voxels: a binning layout for the 8N corner codes,
// then flags and ranks (u32 each) and a scan partial large enough for 8N.
struct LinkLayout {
  size_t bin, flags, rank, partial, nuniq, total;
  explicit LinkLayout(long long N) {
    const long long n = 8 * N;
    Carver c;
    bin = c.take<uint8_t>((long long)BinLayout(n).total);
    flags = c.take<uint32_t>(n);
    rank = c.take<uint32_t>(n);
    partial = c.take<uint32_t>((long long)scan_partial_count(n) + 1);
    nuniq = c.take<uint32_t>(1);
    total = c.off;
  }
};

size_t svr_grid_pts_link_workspace_bytes(long long N) { return N > 0 ? LinkLayout(N).total : 0; }

int svr_build_grid_pts_link(const int64_t* octpath, const int8_t* octlevel, long long N, void* ws,
                            int64_t* vox_key, int64_t* grid_pts_key, long long* n_grid_pts, void* stream) {
  g_err.clear();
  if (N < 0 || 8 * N >= (1LL << 32)) return fail(1, "voxel count%s %lld out of range", "", N);
  if (n_grid_pts) *n_grid_pts = 0;
  if (N == 0) return 0;
  hipStream_t st = as_stream(stream);
  const LinkLayout L(N);
  uint8_t* base = static_cast<uint8_t*>(ws);
  BinPtrs b(base + L.bin, 8 * N);
  uint32_t* flags = reinterpret_cast<uint32_t*>(base + L.flags);
  uint32_t* rank = reinterpret_cast<uint32_t*>(base + L.rank);
  uint32_t* partial = reinterpret_cast<uint32_t*>(base + L.partial);
  uint32_t* nuniq = reinterpret_cast<uint32_t*>(base + L.nuniq);
  SVR_TRY(launch_grid_pts_link(N, octpath, octlevel, b.keys, b.vals, b.keys_alt, b.vals_alt, b.hist, b.partial,
                               partial, flags, rank, vox_key, grid_pts_key, nuniq, st));
  uint32_t m = 0;
  SVR_TRY(hipMemcpyAsync(&m, nuniq, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SVR_TRY(hipStreamSynchronize(st));
  if (n_grid_pts) *n_grid_pts = (long long)m;
  return 0;
}

int svr_agg_voxel_into_grid_pts(const int64_t* vox_key, long long N, const float* vox_val, int C, long long M,
                                uint32_t* count_ws, float* out, void* stream) {
  g_err.clear();
  if (C <= 0 || M < 0 || N < 0) return fail(1, "bad sizes%s %lld", "", C);
  SVR_TRY(launch_grid_pts_mean(N, C, vox_key, vox_val, M, count_ws, out, as_stream(stream)));
  return 0;
}

int svr_agg_voxel_into_grid_pts_noalloc(const int64_t* vox_key, long long N, const float* vox_val, int C,
                                        long long M, float* out, void* stream) {
  g_err.clear();
  if (C <= 0 || M < 0 || N < 0) return fail(1, "bad sizes%s %lld", "", C);
  if (M > kScratchPoints)
    return fail(1, "svr_agg_voxel_into_grid_pts_noalloc takes at most 4194304 grid points%s (got %lld); "
                   "use svr_agg_voxel_into_grid_pts with a workspace", "", M);
  uint32_t* count = scratch_ptr();
  if (!count) return 2;
  // The counts are cleared by a fill kernel inside launch_grid_pts_mean: the scratch keeps the
  // previous call's contents, and no memset node is recorded if the caller captures the call.
  SVR_TRY(launch_grid_pts_mean(N, C, vox_key, vox_val, M, count, out, as_stream(stream)));
  return 0;
}

int svr_ijk_2_octpath(const int64_t* ijk, const int8_t* octlevel, int n, int64_t* octpath, void* stream) {
  g_err.clear();
  SVR_TRY(launch_ijk_to_octpath(n, ijk, octlevel, octpath, as_stream(stream)));
  return 0;
}

int svr_octpath_2_ijk(const int64_t* octpath, const int8_t* octlevel, int n, int64_t* ijk, void* stream) {
  g_err.clear();
  SVR_TRY(launch_octpath_to_ijk(n, octpath, octlevel, ijk, as_stream(stream)));
  return 0;
}

int svr_compute_rd(int W, int H, float cx, float cy, float tanfovx, float tanfovy, const float* c2w,
                   float* out_rd, void* stream) {
  g_err.clear();
  if (W < 0 || H < 0) return fail(1, "image size must be non-negative%s (got %lld)", "", (long long)W * H);
  if ((long long)W * H > 0 && (!c2w || !out_rd)) return fail(1, "c2w and out_rd are required%s", "", 0);
  SVR_TRY(launch_compute_rd(W, H, cx, cy, tanfovx, tanfovy, c2w, out_rd, as_stream(stream)));
  return 0;
}

int svr_depth2pts(int W, int H, float cx, float cy, float tanfovx, float tanfovy, const float* c2w,
                  const float* depth, float* out_pts, void* stream) {
  g_err.clear();
  if (W < 0 || H < 0) return fail(1, "image size must be non-negative%s (got %lld)", "", (long long)W * H);
  if ((long long)W * H > 0 && (!c2w || !depth || !out_pts))
    return fail(1, "c2w, depth and out_pts are required%s", "", 0);
  SVR_TRY(launch_depth2pts(W, H, cx, cy, tanfovx, tanfovy, c2w, depth, out_pts, as_stream(stream)));
  return 0;
}

int svr_is_in_cone(float tanfovx, float tanfovy, float near_plane, const float* w2c, const float* pts,
                   long long N, uint8_t* out_is_in, void* stream) {
  g_err.clear();
  if (N < 0) return fail(1, "point count must be non-negative%s (got %lld)", "", N);
  if (N > 0 && (!w2c || !pts || !out_is_in)) return fail(1, "w2c, pts and out_is_in are required%s", "", 0);
  SVR_TRY(launch_is_in_cone(N, tanfovx, tanfovy, near_plane, w2c, pts, out_is_in, as_stream(stream)));
  return 0;
}

int svr_voxel_order_rank(const int64_t* octree_paths, int P, int64_t* out_ranks, void* stream) {
  g_err.clear();
  if (P < 0) return fail(1, "voxel count must be non-negative%s (got %lld)", "", P);
  if (P > 0 && (!octree_paths || !out_ranks)) return fail(1, "octree_paths and out_ranks are required%s", "", 0);
  SVR_TRY(launch_voxel_order_rank(P, octree_paths, out_ranks, as_stream(stream)));
  return 0;
}

int svr_selfcheck_rcp_rn(int e_lo, int e_hi, unsigned long long* mismatches, void* stream) {
  g_err.clear();
  if (e_lo < -126 || e_hi > 127 || e_lo > e_hi)
    return fail(1, "exponent range must lie in [-126, 127]%s (got %lld)", "", (long long)e_lo);
  if (!mismatches) return fail(1, "mismatches is required%s", "", 0);
  SVR_TRY(launch_selfcheck_rcp_rn(e_lo, e_hi, mismatches, as_stream(stream)));
  return 0;
}

int svr_selfcheck_div_rn(int n_blocks, unsigned seed, unsigned long long* mismatches, void* stream) {
  g_err.clear();
  if (n_blocks < 1 || n_blocks > (1 << 22)) return fail(1, "n_blocks must be 1..2^22%s (got %lld)", "", (long long)n_blocks);
  if (!mismatches) return fail(1, "mismatches is required%s", "", 0);
  SVR_TRY(launch_selfcheck_div_rn(n_blocks, (uint32_t)seed, mismatches, as_stream(stream)));
  return 0;
}

// ------------------------------------------------------------------------------------------
// Batched views: one workspace holding the geometry state of V views x P voxels, a binning
// state of fixed capacity and the image state of all views' tiles and pixels.  Every launch
// covers all views; key counts stay on the device, so a whole step can be captured as a graph.
// ------------------------------------------------------------------------------------------
}  // extern "C"

namespace {
struct BatchPtrs {
  GeomPtrs g;
  BinPtrs b;
  ContribPtrs c;
  ImagePtrs im;
  PairPtrs pr;
  BatchPtrs(void* ws, int V, int P, long long n_pix, long long n_tiles, long long cap)
      : g(ws, V, P),
        b((char*)ws + GeomLayout(V, P).total, cap),
        c((char*)ws + GeomLayout(V, P).total + BinLayout(cap).total, cap),
        im((char*)ws + GeomLayout(V, P).total + raster_bin_bytes(cap), V, n_pix, n_tiles),
        pr((char*)ws + GeomLayout(V, P).total + raster_bin_bytes(cap) + ImageLayout(V, n_pix, n_tiles).total, P, V) {}
};

int check_batch(int V, int P, long long n_pix, long long n_tiles, long long cap) {
  if (V <= 0 || V > 4096) return fail(1, "view count must be 1..4096%s (got %lld)", "", V);
  if (P <= 0 || P >= (1 << 29)) return fail(1, "voxel count must be 1..2^29%s (got %lld)", "", P);
  if ((long long)V * P >= (1LL << 32)) return fail(1, "views x voxels must stay below 2^32%s (%lld)", "", (long long)V * P);
  if (n_tiles <= 0 || n_tiles >= (1LL << 31) || n_pix <= 0) return fail(1, "bad tile/pixel totals%s%lld", "", n_tiles);
  if (cap <= 0 || cap >= (1LL << 31)) return fail(1, "key capacity must be 1..2^31%s (got %lld)", "", cap);
  return 0;
}
}  // namespace

extern "C" {

size_t svr_batch_workspace_bytes(int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap) {
  return GeomLayout(n_views, P).total + raster_bin_bytes(K_cap) + ImageLayout(n_views, n_pixels, n_tiles).total +
         PairLayout(P, n_views).total;
}

int svr_batch_setup(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                    const int* W, const int* H, const float* tanfovx, const float* tanfovy,
                    const float* cx, const float* cy, const float* near_plane, const float* w2c,
                    const float* c2w, const float* background, int tile_w, int tile_h, void* stream) {
  g_err.clear();
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  if (!((tile_w == 16 && tile_h == 16) || (tile_w == 32 && tile_h == 8) || (tile_w == 8 && tile_h == 32)))
    return fail(1, "tile shape must be 16x16, 32x8 or 8x32%s (got width %lld)", "", tile_w);
  hipStream_t st = as_stream(stream);
  BatchPtrs bp(ws, n_views, P, n_pixels, n_tiles, K_cap);
  long long tile_base = 0, pix_base = 0;
  for (int v = 0; v < n_views; ++v) {
    if (int e = check_image(W[v], H[v], P)) return e;
    ViewDesc d = make_view(W[v], H[v], tanfovx[v], tanfovy[v], cx[v], cy[v], near_plane[v], tile_w, tile_h);
    d.tile_base = (int)tile_base;
    d.pix_base = pix_base;
    SVR_TRY(launch_init_view(bp.g.views + v, d, w2c + 16 * v, c2w + 16 * v, background, st));
    if (tile_base + (long long)d.tiles_x * d.tiles_y > n_tiles)
      return fail(1, "tile/pixel totals do not match the views%s (tiles over %lld)", "", n_tiles);
    SVR_TRY(launch_fill_u32(bp.im.tile_view + tile_base, (long long)d.tiles_x * d.tiles_y, (uint32_t)v, st));
    tile_base += (long long)d.tiles_x * d.tiles_y;
    pix_base += (long long)W[v] * H[v];
  }
  if (tile_base != n_tiles || pix_base != n_pixels)
    return fail(1, "tile/pixel totals do not match the views%s (tiles %lld)", "", tile_base);
  SVR_TRY(hipMemsetAsync(bp.g.misc + 8, 0, sizeof(uint32_t), st));
  return 0;
}

size_t svr_batch_pair_order_bytes(int P) {
  return P > 0 ? BinLayout(8LL * P).total : 0;
}

int svr_batch_pair_order(int P, const int64_t* octree_paths, void* tmp, uint32_t* order_out, void* stream) {
  g_err.clear();
  if (P <= 0 || P >= (1 << 29)) return fail(1, "voxel count must be 1..2^29%s (got %lld)", "", P);
  hipStream_t st = as_stream(stream);
  const long long n = 8LL * P;
  BinPtrs t(tmp, n);
  SVR_TRY(launch_pair_rank(n, octree_paths, t.keys, t.vals, st));
  bool in_alt = false;
  SVR_TRY(launch_radix_sort(t.keys, t.vals, t.keys_alt, t.vals_alt, nullptr, n, kRankBits, t.hist, t.partial,
                            &in_alt, st));
  SVR_TRY(hipMemcpyAsync(order_out, in_alt ? t.vals_alt : t.vals, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice,
                         st));
  return 0;
}

int svr_batch_pair_slots(int P, const uint32_t* pair_order, uint32_t* slots_out, void* stream) {
  g_err.clear();
  if (P <= 0 || P >= (1 << 29)) return fail(1, "voxel count must be 1..2^29%s (got %lld)", "", P);
  if (!pair_order || !slots_out) return fail(1, "pair_order and slots_out are required%s%lld", "", 0);
  SVR_TRY(launch_pair_slots(8LL * P, pair_order, slots_out, as_stream(stream)));
  return 0;
}

int svr_batch_preprocess(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                         const float* vox_centers, const float* vox_lengths, const uint8_t* is_leaf,
                         const uint32_t* pair_order, const uint32_t* pair_slots, void* stream) {
  g_err.clear();
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  if (!pair_order && !pair_slots) return fail(1, "pair_order or pair_slots is required%s%lld", "", 0);
  hipStream_t st = as_stream(stream);
  BatchPtrs bp(ws, n_views, P, n_pixels, n_tiles, K_cap);
  const long long n = 8LL * P;
  if (pair_slots) {
    SVR_TRY(launch_preprocess_views(bp.g.views, n_views, P, vox_centers, vox_lengths, is_leaf, nullptr, bp.g.bbox,
                                    bp.g.tbox, bp.g.quad, pair_slots, bp.pr.cnt_ranked, bp.pr.vmask, st));
  } else {
    SVR_TRY(launch_preprocess_views(bp.g.views, n_views, P, vox_centers, vox_lengths, is_leaf, nullptr, bp.g.bbox,
                                    bp.g.tbox, bp.g.quad, nullptr, bp.pr.pair_cnt, bp.pr.vmask, st));
    SVR_TRY(launch_pair_gather(n, pair_order, bp.pr.pair_cnt, bp.pr.cnt_ranked, st));
  }
  SVR_TRY(launch_scan(bp.pr.cnt_ranked, nullptr, n, bp.pr.partial, bp.pr.scan_ranked, true, st));
  SVR_TRY(launch_key_peak(bp.pr.scan_ranked + (n - 1), bp.g.misc + 8, st));
  return 0;
}

int svr_voxel_activity(int P, const uint8_t* is_leaf, const float* geos, const float* s_val,
                       const float* vox_centers, const float* vox_lengths, float cam_scale, uint8_t* out,
                       void* stream) {
  g_err.clear();
  if (P < 0) return fail(1, "bad voxel count%s %lld", "", P);
  if (!(cam_scale >= 0.f)) return fail(1, "cam_scale must be >= 0%s %lld", "", 0);
  SVR_TRY(launch_voxel_activity(P, is_leaf, geos, s_val, vox_centers, vox_lengths, cam_scale, out,
                                as_stream(stream)));
  return 0;
}

int svr_batch_voxel_records(int P, const int64_t* vox_key, const float* grid_pts, const float* vox_centers,
                            const float* vox_lengths, const uint8_t* is_leaf, const float* s_val, float cam_scale,
                            int cull, float* out_records, uint8_t* out_bin_leaf, void* stream) {
  g_err.clear();
  if (P < 0) return fail(1, "bad voxel count%s %lld", "", P);
  if (P > 0 && (!vox_key || !grid_pts || !vox_centers || !vox_lengths || !out_records))
    return fail(1, "vox_key, grid_pts, centres, sizes and out_records are required%s%lld", "", 0);
  if (reinterpret_cast<uintptr_t>(out_records) % 16 != 0) return fail(1, "out_records must be 16-byte aligned%s%lld", "", 0);
  if (out_bin_leaf && (!is_leaf || !s_val)) return fail(1, "out_bin_leaf needs is_leaf and s_val%s%lld", "", 0);
  if (!(cam_scale >= 0.f)) return fail(1, "cam_scale must be >= 0%s %lld", "", 0);
  SVR_TRY(launch_voxel_records(P, vox_key, grid_pts, vox_centers, vox_lengths, is_leaf, s_val, cam_scale, cull,
                               reinterpret_cast<VoxelRecord*>(out_records), out_bin_leaf, as_stream(stream)));
  return 0;
}

int svr_batch_sh_forward(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                         const float* vox_centers, const float* sh0, const float* shs, int degree, int M,
                         float* rgbs_views, void* stream) {
  g_err.clear();
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  if (degree < 0 || degree > 3 || M < (degree + 1) * (degree + 1)) return fail(1, "bad SH degree%s %lld", "", degree);
  if (M > 16) return fail(1, "batched SH supports at most 16 coefficients per channel%s (got %lld)", "", M);
  BatchPtrs bp(ws, n_views, P, n_pixels, n_tiles, K_cap);
  SVR_TRY(launch_sh_views_forward(P, n_views, bp.g.quad, bp.g.views, degree, M, vox_centers, sh0, shs, rgbs_views,
                                  as_stream(stream)));
  return 0;
}

int svr_batch_bin(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                  const uint32_t* pair_order, void* stream) {
  g_err.clear();
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  hipStream_t st = as_stream(stream);
  BatchPtrs bp(ws, n_views, P, n_pixels, n_tiles, K_cap);
  const long long n = 8LL * P;
  const uint32_t* n_dev = bp.pr.scan_ranked + (n - 1);
  // Keys are written where the last tile-sort pass ends in (keys, vals).
  uint32_t* tiles = reinterpret_cast<uint32_t*>(bp.b.keys);
  uint32_t* tiles_alt = reinterpret_cast<uint32_t*>(bp.b.keys_alt);
  const bool odd = tile_sort_passes(n_tiles) % 2 == 1;
  SVR_TRY(launch_pair_emit(bp.g.views, n_views, P, n, pair_order, bp.pr.cnt_ranked, bp.pr.scan_ranked, bp.g.bbox,
                           bp.pr.vmask, odd ? tiles_alt : tiles, odd ? bp.b.vals_alt : bp.b.vals, K_cap, st));
  SVR_TRY(launch_tile_sort(tiles, bp.b.vals, tiles_alt, bp.b.vals_alt, n_dev, K_cap, n_tiles, bp.b.hist,
                           bp.b.partial, st));
  SVR_TRY(launch_fill_u32(reinterpret_cast<uint32_t*>(bp.im.ranges), 2 * n_tiles, 0u, st));
  SVR_TRY(launch_tile_ranges_u32(tiles, n_dev, K_cap, bp.im.ranges, st));
  return 0;
}

int svr_batch_composite(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                        const float* vox_records, const float* rgbs_views, const float* s_val, int need_depth,
                        int need_normal, float* out_color, float* out_depth, float* out_normal, float* out_T,
                        const float* mse_target, float* out_dL_dcolor, void* stream) {
  g_err.clear();
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  if ((mse_target == nullptr) != (out_dL_dcolor == nullptr))
    return fail(1, "mse_target and out_dL_dcolor go together%s%lld", "", 0);
  hipStream_t st = as_stream(stream);
  BatchPtrs bp(ws, n_views, P, n_pixels, n_tiles, K_cap);
  SVR_TRY(launch_fill_u32(bp.im.misc, 2, 0u, st));
  FwdArgs a;
  memset(&a, 0, sizeof(a));
  a.views = bp.g.views; a.n_views = n_views; a.total_tiles = (int)n_tiles;
  a.ranges = bp.im.ranges; a.vals = bp.b.vals;
  a.bboxes = bp.g.bbox; a.tboxes = bp.g.tbox; a.P = P;
  a.vrec = reinterpret_cast<const VoxelRecord*>(vox_records); a.rgbs = rgbs_views;
  a.rgb_view_stride = 3LL * P;
  a.s_val = s_val;
  a.tile_last = bp.im.tile_last; a.n_contrib = bp.im.n_contrib;
  a.color = out_color; a.depth = out_depth; a.normal = out_normal; a.T = out_T; a.max_w = nullptr;
  a.mse_target = mse_target; a.dL_dcolor = out_dL_dcolor;
  a.work_counter = bp.im.misc;
  a.tile_view = bp.im.tile_view;
  a.cvals = bp.c.cvals; a.pmask = bp.c.pmask; a.cranges = bp.im.cranges;
  a.need_depth = need_depth; a.need_normal = need_normal;
  a.variant = g_compositor_variant;
  a.halves = g_batch_depth / 64;
  a.blocks_per_cu = g_blocks_per_cu;
  {
    LaunchTimer timer(0, st);
    SVR_TRY(launch_render_forward(a, st));
  }
  return 0;
}

int svr_batch_rasterize(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                        const float* vox_records, const float* rgbs_views, const float* s_val, const uint32_t* pair_order, int need_depth,
                        int need_normal, float* out_color, float* out_depth, float* out_normal, float* out_T,
                        const float* mse_target, float* out_dL_dcolor, void* stream) {
  if (int e = svr_batch_bin(ws, n_views, P, n_pixels, n_tiles, K_cap, pair_order, stream)) return e;
  return svr_batch_composite(ws, n_views, P, n_pixels, n_tiles, K_cap, vox_records, rgbs_views, s_val, need_depth,
                             need_normal, out_color, out_depth, out_normal, out_T, mse_target, out_dL_dcolor, stream);
}

int svr_batch_backward(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                       const float* vox_records, const float* rgbs_views, const float* s_val, const float* out_T, const float* dL_dcolor,
                       const float* dL_ddepth, const float* dL_dnormal, const float* dL_dT, int need_depth,
                       int need_normal, float* dL_dvox, float* dL_drgb_views, float* dL_ds,
                       const int64_t* vox_key, float* dL_dgrid_pts, void* stream) {
  g_err.clear();
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  if ((vox_key == nullptr) != (dL_dgrid_pts == nullptr))
    return fail(1, "vox_key and dL_dgrid_pts go together%s%lld", "", 0);
  if (need_depth && !dL_ddepth) return fail(1, "need_depth requires dL_ddepth%s%lld", "", 0);
  if (need_normal && !dL_dnormal) return fail(1, "need_normal requires dL_dnormal%s%lld", "", 0);
  hipStream_t st = as_stream(stream);
  BatchPtrs bp(ws, n_views, P, n_pixels, n_tiles, K_cap);
  SVR_TRY(launch_fill_u32(bp.im.misc + 1, 1, 0u, st));
  BwdArgs a;
  memset(&a, 0, sizeof(a));
  a.views = bp.g.views; a.n_views = n_views; a.total_tiles = (int)n_tiles;
  a.ranges = bp.im.ranges; a.tile_last = bp.im.tile_last; a.vals = bp.b.vals;
  a.bboxes = bp.g.bbox; a.P = P;
  a.vrec = reinterpret_cast<const VoxelRecord*>(vox_records); a.rgbs = rgbs_views;
  a.rgb_view_stride = 3LL * P;
  a.s_val = s_val;
  a.T = out_T; a.n_contrib = bp.im.n_contrib;
  a.dL_dcolor = dL_dcolor; a.dL_ddepth = dL_ddepth; a.dL_dnormal = dL_dnormal; a.dL_dT = dL_dT;
  a.gt_color = nullptr; a.lambda_R_concen = 0.f;
  a.dL_dvox = dL_dvox; a.dL_drgb_views = dL_drgb_views; a.dL_ds = dL_ds; a.ds_part = bp.im.ds_part;
  a.vox_key = vox_key; a.dL_dgrid = dL_dgrid_pts;
  a.work_counter = bp.im.misc + 1;
  a.tile_view = bp.im.tile_view;
  a.need_depth = need_depth; a.need_normal = need_normal;
  a.variant = g_compositor_variant;
  a.halves = g_bwd_batch_depth / 64;
  a.blocks_per_cu = g_blocks_per_cu;
  a.cvals = bp.c.cvals; a.pmask = bp.c.pmask; a.cranges = bp.im.cranges;
  {
    LaunchTimer timer(1, st);
    SVR_TRY(launch_render_backward(a, st));
  }
  return 0;
}

int svr_batch_sh_backward(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                          const float* vox_centers, float* dL_drgb_views, int degree, int M,
                          float* dL_dsh0, float* dL_dshs, void* stream) {
  g_err.clear();
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  if (degree < 0 || degree > 3 || M < (degree + 1) * (degree + 1)) return fail(1, "bad SH degree%s %lld", "", degree);
  if (M > 16) return fail(1, "batched SH supports at most 16 coefficients per channel%s (got %lld)", "", M);
  BatchPtrs bp(ws, n_views, P, n_pixels, n_tiles, K_cap);
  SVR_TRY(launch_sh_views_backward(P, n_views, bp.g.quad, bp.g.views, degree, M, vox_centers, dL_drgb_views,
                                   dL_dsh0, dL_dshs, nullptr, as_stream(stream)));
  return 0;
}

int svr_batch_sh_backward_adam(void* ws, int n_views, int P, long long n_pixels, long long n_tiles,
                               long long K_cap, const float* vox_centers, float* dL_drgb_views, int degree, int M,
                               float* sh0, float* sh0_m, float* sh0_v, const float* table_sh0, float* shs,
                               float* shs_m, float* shs_v, const float* table_shs, int n_table,
                               const int32_t* step, float beta1, float beta2, float eps, int sparse,
                               void* stream) {
  g_err.clear();
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  if (degree < 0 || degree > 3 || M < (degree + 1) * (degree + 1)) return fail(1, "bad SH degree%s %lld", "", degree);
  if (M > 16) return fail(1, "batched SH supports at most 16 coefficients per channel%s (got %lld)", "", M);
  if (!step || !table_sh0 || !table_shs || n_table <= 0) return fail(1, "step counter and bias tables%s are required", "", 0);
  BatchPtrs bp(ws, n_views, P, n_pixels, n_tiles, K_cap);
  ShAdam A;
  A.sh0 = sh0; A.m0 = sh0_m; A.v0 = sh0_v;
  A.shs = shs; A.ms = shs_m; A.vs = shs_v;
  A.t0 = adam_table(step, table_sh0, n_table, beta1, beta2, eps);
  A.ts = adam_table(step, table_shs, n_table, beta1, beta2, eps);
  A.sparse = sparse ? 1 : 0;
  SVR_TRY(launch_sh_views_backward(P, n_views, bp.g.quad, bp.g.views, degree, M, vox_centers, dL_drgb_views,
                                   nullptr, nullptr, &A, as_stream(stream)));
  return 0;
}

int svr_batch_state_offsets(int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                            long long* out) {
  g_err.clear();
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  const size_t g = GeomLayout(n_views, P).total;
  const BinLayout bl(K_cap);
  const size_t im = g + raster_bin_bytes(K_cap);
  const ImageLayout il(n_views, n_pixels, n_tiles);
  out[0] = (long long)(im + il.ranges);
  out[1] = (long long)(g + bl.vals);
  out[2] = (long long)(im + il.n_contrib);
  out[3] = (long long)(im + il.tile_last);
  out[4] = (long long)(im + il.cranges);
  return 0;
}

int svr_batch_key_peak(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                       int reset, long long* peak_out, void* stream) {
  g_err.clear();
  *peak_out = 0;
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  hipStream_t st = as_stream(stream);
  BatchPtrs bp(ws, n_views, P, n_pixels, n_tiles, K_cap);
  uint32_t k = 0;
  SVR_TRY(hipMemcpyAsync(&k, bp.g.misc + 8, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SVR_TRY(hipStreamSynchronize(st));
  *peak_out = k;
  if (reset) {
    const uint32_t zero = 0;
    SVR_TRY(hipMemcpyAsync(bp.g.misc + 8, &zero, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    SVR_TRY(hipStreamSynchronize(st));
  }
  return 0;
}

int svr_batch_key_count(void* ws, int n_views, int P, long long n_pixels, long long n_tiles, long long K_cap,
                        long long* K_out, void* stream) {
  g_err.clear();
  *K_out = 0;
  if (int e = check_batch(n_views, P, n_pixels, n_tiles, K_cap)) return e;
  hipStream_t st = as_stream(stream);
  BatchPtrs bp(ws, n_views, P, n_pixels, n_tiles, K_cap);
  uint32_t k = 0;
  SVR_TRY(hipMemcpyAsync(&k, bp.pr.scan_ranked + (8LL * P - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SVR_TRY(hipStreamSynchronize(st));
  *K_out = k;
  return 0;
}

size_t svr_valid_gradient_table_workspace_bytes(int N) {
  Carver c;
  c.take<uint32_t>(N);                                    // per-voxel cell counts
  c.take<uint32_t>(N);                                    // inclusive scan
  c.take<uint32_t>((long long)scan_partial_count(N) + 1); // scan scratch
  return c.off;
}

static void table_ptrs(void* ws, int N, uint32_t** count, uint32_t** scan, uint32_t** partial) {
  Carver c;
  const size_t o0 = c.take<uint32_t>(N);
  const size_t o1 = c.take<uint32_t>(N);
  const size_t o2 = c.take<uint32_t>((long long)scan_partial_count(N) + 1);
  *count = (uint32_t*)((char*)ws + o0);
  *scan = (uint32_t*)((char*)ws + o1);
  *partial = (uint32_t*)((char*)ws + o2);
}

int svr_valid_gradient_table_count(const float* vox_center, const float* vox_size,
                                   const float* scene_center, float inside_extent,
                                   int grid_res_pow2, const uint8_t* is_leaf, int N,
                                   void* workspace, long long* count_out, void* stream) {
  g_err.clear();
  *count_out = 0;
  if (grid_res_pow2 < 0 || grid_res_pow2 > 10) return fail(1, "grid_res_pow2 must be 0..10%s (got %lld)", "", grid_res_pow2);
  if (N <= 0) return 0;
  hipStream_t st = as_stream(stream);
  uint32_t *cnt, *scan, *part;
  table_ptrs(workspace, N, &cnt, &scan, &part);
  const int res = 1 << grid_res_pow2;
  SVR_TRY(launch_table_count(N, vox_center, vox_size, scene_center, inside_extent, res,
                             reinterpret_cast<const bool*>(is_leaf), cnt, st));
  SVR_TRY(launch_scan(cnt, nullptr, N, part, scan, true, st));
  uint32_t total = 0;
  SVR_TRY(hipMemcpyAsync(&total, scan + (N - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  SVR_TRY(hipStreamSynchronize(st));
  *count_out = total;
  return 0;
}

int svr_valid_gradient_table_fill(const float* vox_center, const float* vox_size,
                                  const float* scene_center, float inside_extent,
                                  int grid_res_pow2, const uint8_t* is_leaf, int N,
                                  void* workspace, uint8_t* grid_mask, int32_t* grid_keys,
                                  int32_t* grid2voxel, void* stream) {
  g_err.clear();
  if (N <= 0) return 0;
  uint32_t *cnt, *scan, *part;
  table_ptrs(workspace, N, &cnt, &scan, &part);
  SVR_TRY(launch_table_write(N, vox_center, vox_size, scene_center, inside_extent, 1 << grid_res_pow2,
                             reinterpret_cast<const bool*>(is_leaf), scan, reinterpret_cast<bool*>(grid_mask),
                             grid_keys, grid2voxel, as_stream(stream)));
  return 0;
}

}  // extern "C"

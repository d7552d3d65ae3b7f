// C ABI of libdroplet_hip.so: device context, errors, buffers, renderer and
// scene handles, the standalone stable sort and the numerics self-check.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <unordered_set>
#include <vector>

#include "dl_common.h"
#include "dl_internal.h"
#include "batch.h"
#include "render.h"

struct dl_renderer {
  dl::Renderer r;
};
struct dl_scene {
  dl::Scene s;
};
struct dl_batch {
  dl::Batch b;
};

namespace dl {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

void fail(int code, const std::string &msg) { throw Error{code, msg}; }

void check_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  const int code = (e == hipErrorOutOfMemory) ? DL_ERR_OOM : DL_ERR_DEVICE;
  fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

DeviceBuffer::~DeviceBuffer() { release(); }

void DeviceBuffer::release() {
  if (ptr_) (void)hipFree(ptr_);
  ptr_ = nullptr;
  cap_ = 0;
}

void DeviceBuffer::reserve(size_t bytes) {
  if (bytes <= cap_) return;
  release();
  // grow by 1/8 headroom so frames of slightly different size reuse memory
  size_t want = bytes + bytes / 8;
  want = (want + 255) & ~(size_t)255;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    want = (bytes + 255) & ~(size_t)255;
    e = hipMalloc(&p, want);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    fail(DL_ERR_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  ptr_ = p;
  cap_ = want;
}

void Context::make_current() const { DL_HIP(hipSetDevice(device)); }

// ---- scene registry ---------------------------------------------------------
namespace {
std::mutex g_scene_mu;
std::unordered_set<uint64_t> g_live_scenes;
std::atomic<uint64_t> g_scene_serial{0};
}  // namespace

Scene::Scene() : serial(++g_scene_serial) {
  std::lock_guard<std::mutex> lk(g_scene_mu);
  g_live_scenes.insert(serial);
}

Scene::~Scene() {
  std::lock_guard<std::mutex> lk(g_scene_mu);
  g_live_scenes.erase(serial);
}

bool scene_alive(uint64_t serial) {
  std::lock_guard<std::mutex> lk(g_scene_mu);
  return g_live_scenes.count(serial) != 0;
}

static void upload_array(Context &c, DeviceBuffer &b, const float *src, size_t count, const char *name) {
  if (count == 0) {
    b.reserve(256);
    return;
  }
  if (!src) fail(DL_ERR_INVALID, std::string(name) + " is null");
  b.reserve(count * sizeof(float));
  c.staging.to_device(b.get(), src, count * sizeof(float), c.stream);
}

void Scene::upload(const dl_splats &sp) {
  if (sp.sh_bands > 3) fail(DL_ERR_INVALID, "shBands must be 0..3");
  const size_t n = sp.count;
  const uint32_t M = sp.sh_bands == 0 ? 0 : sp.sh_bands == 1 ? 3 : sp.sh_bands == 2 ? 8 : 15;
  ctx->make_current();
  upload_array(*ctx, pos, sp.positions, 3 * n, "positions");
  upload_array(*ctx, rot, sp.rotations, 4 * n, "rotations");
  upload_array(*ctx, ls, sp.log_scales, 3 * n, "logScales");
  upload_array(*ctx, op, sp.opacity_logits, n, "opacityLogits");
  upload_array(*ctx, dc, sp.colors_dc, 3 * n, "colorsDC");
  upload_array(*ctx, sh, sp.colors_sh, 3 * (size_t)M * n, "colorsSH");
  count = sp.count;
  bands = sp.sh_bands;
  sh_coeffs = M;
  build_color_records();
}

// Splat i's record: DC, then the 3 * M SH values the renderer reads at
// colorsSH[3 M i ..] (for a PLY scene with more coefficients per channel
// than M(bands) that is the reference's own stride-3M read), then zeros.
__global__ __launch_bounds__(256) void k_color_records(const float *__restrict__ dc, const float *__restrict__ sh,
                                                       uint32_t n, uint32_t m3, uint32_t stride,
                                                       float *__restrict__ col) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;  // one float of the output
  if (e >= (size_t)n * stride) return;
  const size_t i = e / stride;
  const uint32_t j = (uint32_t)(e - i * stride);
  float v = 0.0f;
  if (j < 3) v = dc[3 * i + j];
  else if (j < 3 + m3) v = sh[i * m3 + (j - 3)];
  col[e] = v;
}

void Scene::build_color_records() {
  ctx->make_current();
  hipStream_t s = ctx->stream;
  const uint32_t stride = color_record_floats(bands);
  const uint32_t M = bands == 0 ? 0 : bands == 1 ? 3 : bands == 2 ? 8 : 15;
  const size_t total = (size_t)count * stride;
  col.reserve(std::max<size_t>(total * sizeof(float), 256));
  if (total) {
    k_color_records<<<(uint32_t)((total + 255) / 256), 256, 0, s>>>(dc.as<const float>(), sh.as<const float>(),
                                                                   count, 3 * M, stride, col.as<float>());
    DL_HIP(hipGetLastError());
  }
  DL_HIP(hipStreamSynchronize(s));
}

SceneView Scene::view() const {
  return SceneView{pos.as<const float>(), rot.as<const float>(), ls.as<const float>(), op.as<const float>(),
                   dc.as<const float>(), sh.as<const float>(), col.as<const float>(), count, bands};
}

// ---------------------------------------------------------------------------
// numerics self-check kernel
// ---------------------------------------------------------------------------
// A kernel of known VALU work, to calibrate the SQ instruction counters
// against: each lane runs `iters` iterations of 16 f32 FMAs (8 independent
// chains of 2), so a wave issues 16 * iters v_fma_f32 plus a few setup
// instructions, and 2 * 16 * iters * 64 FP32 operations; the loop counter
// is uniform, so the loop control is scalar.
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t iters, float *__restrict__ out) {
  float a[8];
#pragma unroll
  for (int k = 0; k < 8; k++) a[k] = (float)(threadIdx.x + 7 * k) * 1e-3f;
  const float m = 0.999f, c = 1e-4f;
#pragma unroll 1
  for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
      a[k] = __builtin_fmaf(a[k], m, c);
      a[k] = __builtin_fmaf(a[k], m, c);
    }
  }
  out[blockIdx.x * 256 + threadIdx.x] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

__global__ void k_debug_math(int op, const float *a, const float *b, float *out, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float r = 0.0f;
  switch (op) {
    case 0: r = dl_expf(a[i]); break;
    case 1: r = __builtin_sqrtf(a[i]); break;
    case 2: r = a[i] / b[i]; break;
    case 3: r = bits_f32(sortable_key((float)((double)a[i] * (double)b[i]))); break;
    case 4: r = dl_sigmoidf(a[i]); break;
    case 5: r = dl_expf_nonpos(a[i]); break;
    default: break;
  }
  out[i] = r;
}

}  // namespace dl

using namespace dl;

// ---- autotuner --------------------------------------------------------------
namespace {

// The autotuner searches in two stages. Stage 1, around `base`: radix digit
// widths 8 / 6 / 4 and pass schedules (full lists; first pass n/32 .. n/4
// with growth 4, 8 or 16; the default rule), n = the largest scene. Stage 2,
// around the stage-1 winner: tile sides 16 / 8 / 32 px and rasterizer
// workgroups of one tile or 2x2 tiles (32-px tiles take one 4-wave
// workgroup each). The batch runs stage 1 only, without 4-bit digits.
uint32_t digits_of(const dl_renderer_opts &o) { return o.sort_digit_bits ? o.sort_digit_bits : 8; }
uint32_t tile_of(const dl_renderer_opts &o) { return o.tile_size ? o.tile_size : 16; }
uint32_t group_of(const dl_renderer_opts &o) { return tile_of(o) == 32 || o.raster_group != 2 ? 1 : 2; }

bool same_options(const dl_renderer_opts &a, const dl_renderer_opts &b) {
  return digits_of(a) == digits_of(b) && a.first_pass == b.first_pass && a.pass_growth == b.pass_growth &&
         (a.full_lists != 0) == (b.full_lists != 0) && tile_of(a) == tile_of(b) && group_of(a) == group_of(b);
}

// Stage-1 candidates other than base itself, which the tuner times first.
std::vector<dl_renderer_opts> schedule_candidates(const dl_renderer_opts &base, uint64_t n) {
  std::vector<dl_renderer_opts> out;
  const uint32_t digits[] = {8, 6, 4};
  struct Sched {
    uint32_t div, growth, full;
  };
  const Sched scheds[] = {{0, 0, 0},  {0, 0, 1},  {32, 8, 0}, {16, 4, 0}, {16, 8, 0},
                          {12, 8, 0}, {8, 8, 0},  {8, 16, 0}, {4, 8, 0}};
  for (uint32_t d : digits)
    for (const Sched &sc : scheds) {
      dl_renderer_opts o = base;
      o.sort_digit_bits = d;
      o.full_lists = sc.full;
      o.first_pass = sc.div ? (uint32_t)std::max<uint64_t>(65536, n / sc.div) : 0;
      o.pass_growth = sc.growth;
      if (!same_options(o, base)) out.push_back(o);
    }
  return out;
}

// Stage-2 candidates around base (base's shape excluded).
std::vector<dl_renderer_opts> shape_candidates(const dl_renderer_opts &base) {
  std::vector<dl_renderer_opts> out;
  const uint32_t shapes[][2] = {{16, 1}, {16, 2}, {8, 1}, {8, 2}, {32, 1}};
  for (const auto &sh : shapes) {
    dl_renderer_opts o = base;
    o.tile_size = sh[0];
    o.raster_group = sh[1];
    if (!same_options(o, base)) out.push_back(o);
  }
  return out;
}

// Frames run before the first timing, so that the default schedule is not
// timed while the GPU's clocks are still rising.
constexpr int kTuneSettle = 8;
constexpr int kTuneReps = 5;

template <class F>
float wall_ms_per_call(F &&f, int reps) {
  f();  // warm: plans, allocations, graph capture
  const auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < reps; k++) f();
  const auto t1 = std::chrono::steady_clock::now();
  return (float)(std::chrono::duration<double, std::milli>(t1 - t0).count() / reps);
}

}  // namespace

extern "C" {

const char *dl_last_error(void) { return g_last_error.c_str(); }

int dl_context_create(const dl_device_opts *opts, dl_context **out) {
  return guard([&] {
    if (!out) fail(DL_ERR_INVALID, "out is null");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
      (void)hipGetLastError();
      fail(DL_ERR_DEVICE, "No HIP device available");
    }
    const int dev = opts ? opts->device_index : 0;
    if (dev < 0 || dev >= count) fail(DL_ERR_INVALID, "device_index out of range");
    std::unique_ptr<dl_context> c(new dl_context());
    c->c.device = dev;
    DL_HIP(hipSetDevice(dev));
    hipDeviceProp_t prop;
    DL_HIP(hipGetDeviceProperties(&prop, dev));
    c->c.num_cus = prop.multiProcessorCount;
    c->c.arch = prop.gcnArchName;
    DL_HIP(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
    // the device address of the no-alloc sort's scratch (once per device)
    noalloc_setup(dev);
    *out = c.release();
  });
}

void dl_context_destroy(dl_context *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  ctx->c.sort_keys.release();
  ctx->c.sort_vals.release();
  ctx->c.sort_tmp_keys.release();
  ctx->c.sort_tmp_vals.release();
  ctx->c.sort_ws.release();
  if (ctx->c.stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

int dl_context_device_name(dl_context *ctx, char *buf, size_t cap) {
  return guard([&] {
    if (!ctx || !buf || cap == 0) fail(DL_ERR_INVALID, "null argument");
    std::strncpy(buf, ctx->c.arch.c_str(), cap - 1);
    buf[cap - 1] = 0;
  });
}

int dl_renderer_create(dl_context *ctx, const dl_renderer_opts *opts, dl_renderer **out) {
  return guard([&] {
    if (!ctx || !out) fail(DL_ERR_INVALID, "null argument");
    std::unique_ptr<dl_renderer> r(new dl_renderer());
    r->r.ctx = &ctx->c;
    if (opts) r->r.opts = *opts;
    r->r.max_cov = (opts && opts->max_coverage_per_splat) ? opts->max_coverage_per_splat : kMaxCoverageDefault;
    r->r.own_scene.ctx = &ctx->c;
    *out = r.release();
  });
}

void dl_renderer_destroy(dl_renderer *r) {
  if (!r) return;
  (void)hipSetDevice(r->r.ctx->device);
  (void)hipStreamSynchronize(r->r.ctx->stream);
  delete r;
}

static dl_background bg_or_default(const dl_background *bg) {
  return bg ? *bg : dl_background{0.0f, 0.0f, 0.0f, 1.0f};
}

int dl_render(dl_renderer *r, const dl_splats *splats, const dl_camera *cam, const dl_background *bg,
              uint8_t *out_rgba) {
  return guard([&] {
    if (!r || !splats || !cam || !out_rgba) fail(DL_ERR_INVALID, "null argument");
    r->r.own_scene.upload(*splats);
    r->r.render(r->r.own_scene.view(), *cam, bg_or_default(bg));
    r->r.read_image(out_rgba);
  });
}

int dl_scene_create(dl_context *ctx, const dl_splats *splats, dl_scene **out) {
  return guard([&] {
    if (!ctx || !splats || !out) fail(DL_ERR_INVALID, "null argument");
    std::unique_ptr<dl_scene> sc(new dl_scene());
    sc->s.ctx = &ctx->c;
    sc->s.upload(*splats);
    *out = sc.release();
  });
}

void dl_scene_destroy(dl_scene *scene) {
  if (!scene) return;
  (void)hipSetDevice(scene->s.ctx->device);
  (void)hipStreamSynchronize(scene->s.ctx->stream);
  delete scene;
}

int dl_scene_create_from_ply(dl_context *ctx, const uint8_t *bytes, uint64_t len, int flip_y, dl_scene **out) {
  return guard([&] {
    if (!ctx || !out) fail(DL_ERR_INVALID, "null argument");
    *out = nullptr;
    std::unique_ptr<dl_scene> sc(new dl_scene());
    sc->s.ctx = &ctx->c;
    sc->s.load_ply(bytes, (size_t)len, flip_y != 0);
    *out = sc.release();
  });
}

int dl_scene_info(const dl_scene *scene, uint32_t *count, uint32_t *sh_bands, uint32_t *sh_coeffs_per_channel) {
  return guard([&] {
    if (!scene) fail(DL_ERR_INVALID, "null scene");
    if (count) *count = scene->s.count;
    if (sh_bands) *sh_bands = scene->s.bands;
    if (sh_coeffs_per_channel) *sh_coeffs_per_channel = scene->s.sh_coeffs;
  });
}

int dl_scene_read(dl_scene *scene, float *positions, float *rotations, float *log_scales, float *opacity_logits,
                  float *colors_dc, float *colors_sh) {
  return guard([&] {
    if (!scene) fail(DL_ERR_INVALID, "null scene");
    Scene &s = scene->s;
    s.ctx->make_current();
    hipStream_t st = s.ctx->stream;
    const size_t n = s.count;
    auto get = [&](float *dst, const DeviceBuffer &b, size_t floats) {
      if (dst && floats) s.ctx->staging.to_host(dst, b.get(), floats * 4, st);
    };
    get(positions, s.pos, 3 * n);
    get(rotations, s.rot, 4 * n);
    get(log_scales, s.ls, 3 * n);
    get(opacity_logits, s.op, n);
    get(colors_dc, s.dc, 3 * n);
    get(colors_sh, s.sh, 3 * (size_t)s.sh_coeffs * n);
    DL_HIP(hipStreamSynchronize(st));
  });
}

int dl_scene_bounds(dl_scene *scene, double bounds_min[3], double bounds_max[3]) {
  return guard([&] {
    if (!scene || !bounds_min || !bounds_max) fail(DL_ERR_INVALID, "null argument");
    scene->s.bounds(bounds_min, bounds_max);
  });
}

int dl_render_scene(dl_renderer *r, const dl_scene *scene, const dl_camera *cam, const dl_background *bg,
                    uint8_t *out_rgba) {
  return guard([&] {
    if (!r || !scene || !cam) fail(DL_ERR_INVALID, "null argument");
    if (scene->s.ctx != r->r.ctx) fail(DL_ERR_INVALID, "scene and renderer belong to different contexts");
    r->r.render(scene->s.view(), *cam, bg_or_default(bg));
    if (out_rgba) r->r.read_image(out_rgba);
  });
}

int dl_render_forward(dl_renderer *r, const dl_scene *scene, const dl_camera *cam, const dl_background *bg,
                      float *out_rgb) {
  return guard([&] {
    if (!r || !scene || !cam) fail(DL_ERR_INVALID, "null argument");
    if (scene->s.ctx != r->r.ctx) fail(DL_ERR_INVALID, "scene and renderer belong to different contexts");
    r->r.forward_train(scene->s.view(), scene->s.serial, *cam, bg_or_default(bg), out_rgb);
  });
}

int dl_render_backward(dl_renderer *r, const float *grad_rgb, dl_grads *out) {
  return guard([&] {
    if (!r || !grad_rgb || !out) fail(DL_ERR_INVALID, "null argument");
    r->r.backward(grad_rgb, *out);
  });
}

int dl_renderer_read_image(dl_renderer *r, uint8_t *out_rgba) {
  return guard([&] {
    if (!r || !out_rgba) fail(DL_ERR_INVALID, "null argument");
    r->r.read_image(out_rgba);
  });
}

int dl_renderer_set_timing(dl_renderer *r, int enabled) {
  return guard([&] {
    if (!r) fail(DL_ERR_INVALID, "null renderer");
    r->r.timing = enabled != 0;
  });
}

int dl_renderer_set_counting(dl_renderer *r, int enabled) {
  return guard([&] {
    if (!r) fail(DL_ERR_INVALID, "null renderer");
    if (enabled < 0 || enabled > 2) fail(DL_ERR_INVALID, "counting must be 0, 1 or 2");
    r->r.counting = enabled == 1;
    r->r.tallying = enabled == 2;
  });
}

int dl_renderer_set_pair_budget(dl_renderer *r, uint64_t pairs) {
  return guard([&] {
    if (!r) fail(DL_ERR_INVALID, "null renderer");
    if (pairs > 0xFFFFFFF0ull) fail(DL_ERR_INVALID, "the pair budget must be below 2^32 - 16");
    r->r.pair_budget = pairs;
    r->r.spec_cap = 0;
  });
}

int dl_renderer_row_iters(dl_renderer *r, uint64_t *out, uint32_t cap, uint32_t *rows) {
  return guard([&] {
    if (!r || !rows) fail(DL_ERR_INVALID, "null argument");
    Renderer &R = r->r;
    *rows = R.row_iters_rows;
    if (!out || R.row_iters_rows == 0) return;
    if (cap < R.row_iters_rows) fail(DL_ERR_CAPACITY, "row_iters buffer too small");
    R.ctx->make_current();
    DL_HIP(hipMemcpyAsync(out, R.row_iters.get(), (size_t)R.row_iters_rows * 8, hipMemcpyDeviceToHost, R.ctx->stream));
    DL_HIP(hipStreamSynchronize(R.ctx->stream));
  });
}

int dl_renderer_stats(dl_renderer *r, dl_render_stats *out) {
  return guard([&] {
    if (!r || !out) fail(DL_ERR_INVALID, "null argument");
    *out = r->r.stats;
  });
}

int dl_renderer_set_taps(dl_renderer *r, int enabled) {
  return guard([&] {
    if (!r) fail(DL_ERR_INVALID, "null renderer");
    r->r.keep_taps = enabled != 0;
  });
}

int dl_renderer_tap(dl_renderer *r, int which, void *dst, size_t bytes, size_t *written) {
  return guard([&] {
    if (!r) fail(DL_ERR_INVALID, "null renderer");
    if (which < 0 || which >= (int)r->r.taps.size()) fail(DL_ERR_INVALID, "unknown tap");
    const auto &t = r->r.taps[which];
    if (written) *written = t.size();
    if (!dst) return;
    if (bytes < t.size()) fail(DL_ERR_CAPACITY, "tap buffer too small");
    if (!t.empty()) std::memcpy(dst, t.data(), t.size());
  });
}


int dl_renderer_autotune(dl_renderer *r, const dl_scene *scene, const dl_camera *cam, dl_tune_result *out) {
  return guard([&] {
    if (!r || !scene || !cam || !out) fail(DL_ERR_INVALID, "null argument");
    if (scene->s.ctx != r->r.ctx) fail(DL_ERR_INVALID, "scene and renderer belong to different contexts");
    Renderer &R = r->r;
    const dl_renderer_opts before = R.opts;
    // restores the caller's options and switches if a candidate throws
    struct Restore {
      Renderer &R;
      dl_renderer_opts opts;
      bool keep_taps, timing, counting, done = false;
      ~Restore() {
        if (done) return;
        R.opts = opts;
        R.keep_taps = keep_taps;
        R.timing = timing;
        R.counting = counting;
      }
    } restore{R, before, R.keep_taps, R.timing, R.counting};
    const bool keep_taps = R.keep_taps, timing = R.timing, counting = R.counting;
    R.keep_taps = R.timing = R.counting = false;
    const dl_background bg{0.0f, 0.0f, 0.0f, 1.0f};
    auto frame = [&] { R.render(scene->s.view(), *cam, bg); };
    for (int k = 0; k < kTuneSettle; k++) frame();
    dl_tune_result res{};
    res.default_ms = wall_ms_per_call(frame, kTuneReps);
    res.chosen = before;
    res.chosen_ms = res.default_ms;
    for (const auto &o : schedule_candidates(before, scene->s.count)) {
      R.opts = o;
      const float ms = wall_ms_per_call(frame, kTuneReps);
      res.candidates++;
      if (ms < res.chosen_ms) {
        res.chosen_ms = ms;
        res.chosen = o;
      }
    }
    for (const auto &o : shape_candidates(res.chosen)) {
      R.opts = o;
      const float ms = wall_ms_per_call(frame, kTuneReps);
      res.candidates++;
      // a shape that fell back to 16-px tiles (maxCoverage truncates a splat
      // there) renders at 16 px whatever it asks: not a candidate
      if (R.stats.tile_size != tile_of(o)) continue;
      if (ms < res.chosen_ms) {
        res.chosen_ms = ms;
        res.chosen = o;
      }
    }
    R.opts = res.chosen;
    R.keep_taps = keep_taps;
    R.timing = timing;
    R.counting = counting;
    restore.done = true;
    *out = res;
  });
}

int dl_renderer_tune_candidates(const dl_renderer_opts *base, uint64_t num_splats, dl_renderer_opts *out,
                                uint32_t cap, uint32_t *count) {
  return guard([&] {
    if (!base || !count) fail(DL_ERR_INVALID, "null argument");
    std::vector<dl_renderer_opts> all = schedule_candidates(*base, num_splats);
    for (const auto &o : shape_candidates(*base)) all.push_back(o);
    *count = (uint32_t)all.size();
    if (out)
      for (size_t k = 0; k < all.size() && k < cap; k++) out[k] = all[k];
  });
}

int dl_batch_autotune(dl_batch *b, dl_scene *const *scenes, uint32_t num_scenes, const dl_view *views,
                      uint32_t num_views, dl_tune_result *out) {
  return guard([&] {
    if (!b || !scenes || !views || !out) fail(DL_ERR_INVALID, "null argument");
    std::vector<SceneView> sv(num_scenes);
    uint64_t nmax = 0;
    for (uint32_t k = 0; k < num_scenes; k++) {
      if (!scenes[k]) fail(DL_ERR_INVALID, "null scene");
      if (scenes[k]->s.ctx != b->b.ctx) fail(DL_ERR_INVALID, "scene and batch belong to different contexts");
      sv[k] = scenes[k]->s.view();
      nmax = std::max<uint64_t>(nmax, sv[k].count);
    }
    Batch &B = b->b;
    const dl_renderer_opts before = B.opts;
    struct Restore {
      Batch &B;
      dl_renderer_opts opts;
      bool done = false;
      ~Restore() {
        if (done) return;
        B.opts = opts;
        B.layout.clear();
      }
    } restore{B, before};
    auto call = [&] { B.render(sv, views, num_views); };
    for (int k = 0; k < 2; k++) call();
    dl_tune_result res{};
    res.default_ms = wall_ms_per_call(call, 2);
    res.chosen = before;
    res.chosen_ms = res.default_ms;
    for (const auto &o : schedule_candidates(before, nmax)) {
      if (o.sort_digit_bits == 4) continue;  // 8 passes over every view's keys: never faster here
      B.opts = o;
      B.layout.clear();  // a new schedule is a new plan (and graph)
      const float ms = wall_ms_per_call(call, 2);
      res.candidates++;
      if (ms < res.chosen_ms) {
        res.chosen_ms = ms;
        res.chosen = o;
      }
    }
    B.opts = res.chosen;
    B.layout.clear();
    restore.done = true;
    *out = res;
  });
}

int dl_batch_create(dl_context *ctx, const dl_renderer_opts *opts, dl_batch **out) {
  return guard([&] {
    if (!ctx || !out) fail(DL_ERR_INVALID, "null argument");
    std::unique_ptr<dl_batch> b(new dl_batch());
    b->b.ctx = &ctx->c;
    if (opts) b->b.opts = *opts;
    b->b.max_cov = (opts && opts->max_coverage_per_splat) ? opts->max_coverage_per_splat : kMaxCoverageDefault;
    if (opts && opts->chunk_cap) b->b.item_cap = opts->chunk_cap;
    *out = b.release();
  });
}

void dl_batch_destroy(dl_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->b.ctx->device);
  (void)hipStreamSynchronize(b->b.ctx->stream);
  delete b;
}

int dl_batch_set_graph(dl_batch *b, int enabled) {
  return guard([&] {
    if (!b) fail(DL_ERR_INVALID, "null batch");
    b->b.use_graph = enabled != 0;
  });
}

int dl_batch_render(dl_batch *b, dl_scene *const *scenes, uint32_t num_scenes, const dl_view *views,
                    uint32_t num_views) {
  return guard([&] {
    if (!b || !scenes || !views) fail(DL_ERR_INVALID, "null argument");
    std::vector<SceneView> sv(num_scenes);
    for (uint32_t k = 0; k < num_scenes; k++) {
      if (!scenes[k]) fail(DL_ERR_INVALID, "null scene");
      if (scenes[k]->s.ctx != b->b.ctx) fail(DL_ERR_INVALID, "scene and batch belong to different contexts");
      sv[k] = scenes[k]->s.view();
    }
    b->b.render(sv, views, num_views);
  });
}

int dl_batch_set_pair_budget(dl_batch *b, uint64_t pairs) {
  return guard([&] {
    if (!b) fail(DL_ERR_INVALID, "null batch");
    if (pairs > 0xFFFFFFF0ull) fail(DL_ERR_INVALID, "the pair budget must be below 2^32 - 16");
    b->b.pair_budget = pairs;
    b->b.layout.clear();
    b->b.break_layout.clear();
  });
}

int dl_batch_read_image(dl_batch *b, uint32_t view, uint8_t *out_rgba) {
  return guard([&] {
    if (!b || !out_rgba) fail(DL_ERR_INVALID, "null argument");
    b->b.read_image(view, out_rgba);
  });
}

int dl_batch_get_stats(dl_batch *b, dl_batch_stats *out) {
  return guard([&] {
    if (!b || !out) fail(DL_ERR_INVALID, "null argument");
    *out = b->b.stats;
  });
}

int dl_sort_pairs_u32(dl_context *ctx, uint32_t *keys, uint32_t *vals, uint32_t n, uint32_t key_bits) {
  return guard([&] {
    if (!ctx) fail(DL_ERR_INVALID, "null context");
    if (n == 0) return;
    if (!keys || !vals) fail(DL_ERR_INVALID, "null keys/vals");
    if (key_bits == 0 || key_bits > 32) fail(DL_ERR_INVALID, "key_bits must be 1..32");
    Context &c = ctx->c;
    c.make_current();
    hipStream_t s = c.stream;
    const size_t bytes = (size_t)n * 4;
    c.sort_keys.reserve(bytes);
    c.sort_vals.reserve(bytes);
    c.sort_tmp_keys.reserve(bytes);
    c.sort_tmp_vals.reserve(bytes);
    c.sort_ws.reserve(radix_workspace_bytes(n, (int)key_bits));
    DL_HIP(hipMemcpyAsync(c.sort_keys.get(), keys, bytes, hipMemcpyHostToDevice, s));
    DL_HIP(hipMemcpyAsync(c.sort_vals.get(), vals, bytes, hipMemcpyHostToDevice, s));
    const bool in_tmp = radix_sort_pairs(s, c.sort_keys.as<uint32_t>(), c.sort_vals.as<uint32_t>(),
                                         c.sort_tmp_keys.as<uint32_t>(), c.sort_tmp_vals.as<uint32_t>(), n,
                                         (int)key_bits, c.sort_ws.get());
    DL_HIP(hipMemcpyAsync(keys, in_tmp ? c.sort_tmp_keys.get() : c.sort_keys.get(), bytes, hipMemcpyDeviceToHost, s));
    DL_HIP(hipMemcpyAsync(vals, in_tmp ? c.sort_tmp_vals.get() : c.sort_vals.get(), bytes, hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
  });
}

int dl_sort_pairs_u32_noalloc(dl_context *ctx, uint32_t *d_keys, uint32_t *d_vals, uint32_t n, uint32_t key_bits) {
  return guard([&] {
    if (!ctx) fail(DL_ERR_INVALID, "null context");
    if (n == 0) return;
    if (!d_keys || !d_vals) fail(DL_ERR_INVALID, "null keys/vals");
    if (key_bits == 0 || key_bits > 32) fail(DL_ERR_INVALID, "key_bits must be 1..32");
    noalloc_sort_pairs(ctx->c, d_keys, d_vals, n, (int)key_bits);
  });
}

uint64_t dl_sort_pairs_u32_noalloc_max(void) { return noalloc_max_pairs(); }

int dl_scan_u32(dl_context *ctx, const uint32_t *in, uint32_t *out, uint64_t n, uint32_t repeat, uint32_t *total) {
  return guard([&] {
    if (!ctx || !total) fail(DL_ERR_INVALID, "null argument");
    if (n && (!in || !out)) fail(DL_ERR_INVALID, "null in/out");
    if (repeat == 0) fail(DL_ERR_INVALID, "repeat must be >= 1");
    Context &c = ctx->c;
    c.make_current();
    hipStream_t s = c.stream;
    const size_t bytes = (size_t)n * 4;
    DeviceBuffer d_in, d_out, d_total, ws;
    d_in.reserve(std::max<size_t>(bytes, 4));
    d_out.reserve(std::max<size_t>(bytes, 4));
    d_total.reserve(4);
    ws.reserve(scan_workspace_bytes(n));
    if (n) DL_HIP(hipMemcpyAsync(d_in.get(), in, bytes, hipMemcpyHostToDevice, s));
    for (uint32_t r = 0; r < repeat; r++)
      exclusive_scan_u32(s, d_in.as<uint32_t>(), d_out.as<uint32_t>(), n, d_total.as<uint32_t>(), ws.get());
    if (n) DL_HIP(hipMemcpyAsync(out, d_out.get(), bytes, hipMemcpyDeviceToHost, s));
    DL_HIP(hipMemcpyAsync(total, d_total.get(), 4, hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
  });
}

int dl_debug_valu_probe(dl_context *ctx, uint32_t blocks, uint32_t iters, double *checksum) {
  return guard([&] {
    if (!ctx || blocks == 0) fail(DL_ERR_INVALID, "null context or no blocks");
    Context &c = ctx->c;
    c.make_current();
    hipStream_t s = c.stream;
    DeviceBuffer out;
    out.reserve((size_t)blocks * 256 * 4);
    k_valu_probe<<<blocks, 256, 0, s>>>(iters, out.as<float>());
    DL_HIP(hipGetLastError());
    std::vector<float> h((size_t)blocks * 256);
    DL_HIP(hipMemcpyAsync(h.data(), out.get(), h.size() * 4, hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
    double sum = 0.0;
    for (float v : h) sum += v;
    if (checksum) *checksum = sum;
  });
}

int dl_context_synchronize(dl_context *ctx) {
  return guard([&] {
    if (!ctx) fail(DL_ERR_INVALID, "null context");
    ctx->c.make_current();
    DL_HIP(hipDeviceSynchronize());
  });
}

int dl_debug_math(dl_context *ctx, int op, const float *a, const float *b, float *out, uint32_t n) {
  return guard([&] {
    if (!ctx || !a || !out) fail(DL_ERR_INVALID, "null argument");
    if ((op == 2 || op == 3) && !b) fail(DL_ERR_INVALID, "op needs b");
    if (op < 0 || op > 5) fail(DL_ERR_INVALID, "unknown op");
    if (n == 0) return;
    Context &c = ctx->c;
    c.make_current();
    hipStream_t s = c.stream;
    DeviceBuffer da, db, dout;
    da.reserve((size_t)n * 4);
    dout.reserve((size_t)n * 4);
    DL_HIP(hipMemcpyAsync(da.get(), a, (size_t)n * 4, hipMemcpyHostToDevice, s));
    if (b) {
      db.reserve((size_t)n * 4);
      DL_HIP(hipMemcpyAsync(db.get(), b, (size_t)n * 4, hipMemcpyHostToDevice, s));
    }
    k_debug_math<<<div_up(n, 256), 256, 0, s>>>(op, da.as<float>(), db.as<float>(), dout.as<float>(), n);
    DL_HIP(hipGetLastError());
    DL_HIP(hipMemcpyAsync(out, dout.get(), (size_t)n * 4, hipMemcpyDeviceToHost, s));
    DL_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"

// context.hip -- context lifetime, errors, handles, transfer functions and
// the blue-noise atlas.  Replaces gprtContextCreate / gprtDeviceTextureCreate
// and friends (hostCode.cpp:280-365).
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

namespace aarbf {

int set_error(aarbf_context *ctx, int status, const std::string &msg) {
  if (ctx) ctx->last_error = msg;
  return status;
}

int hip_error(aarbf_context *ctx, hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return set_error(ctx, e == hipErrorOutOfMemory ? AARBF_ERR_OUT_OF_MEMORY : AARBF_ERR_HIP, m);
}

void *scratch(aarbf_context *ctx, size_t bytes) {
  if (bytes <= ctx->scratch_bytes) return ctx->scratch;
  if (ctx->scratch) {
    hipStreamSynchronize(ctx->stream);
    hipFree(ctx->scratch);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  size_t want = bytes + bytes / 4 + 4096;
  if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
    ctx->scratch = nullptr;
    return nullptr;
  }
  ctx->scratch_bytes = want;
  return ctx->scratch;
}

TF *get_tf(aarbf_context *ctx, uint64_t h) {
  return ctx->tfs.count(h) ? reinterpret_cast<TF *>(h) : nullptr;
}
STBN *get_stbn(aarbf_context *ctx, uint64_t h) {
  return ctx->stbns.count(h) ? reinterpret_cast<STBN *>(h) : nullptr;
}
BVH *get_bvh(aarbf_context *ctx, uint64_t h) {
  return ctx->bvhs.count(h) ? reinterpret_cast<BVH *>(h) : nullptr;
}

BVHView make_view(const BVH &b) {
  BVHView v{};
  v.nodes = b.nodes;
  v.groups = b.groups;
  v.log2K = b.log2K;
  v.top = b.num_levels - 1;
  for (uint32_t l = 0; l < b.num_levels; ++l) {
    v.counts[l] = b.counts[l];
    v.level_ofs[l] = (uint32_t)b.level_ofs[l];
    v.group_ofs[l] = (uint32_t)b.group_ofs[l];
  }
  return v;
}

} // namespace aarbf

using namespace aarbf;

extern "C" {

const char *aarbf_status_string(int status) {
  switch (status) {
  case AARBF_OK: return "ok";
  case AARBF_ERR_INVALID_ARGUMENT: return "invalid argument";
  case AARBF_ERR_OUT_OF_MEMORY: return "out of memory";
  case AARBF_ERR_HIP: return "HIP runtime error";
  case AARBF_ERR_IO: return "I/O error";
  case AARBF_ERR_FORMAT: return "malformed particle file";
  case AARBF_ERR_UNSUPPORTED: return "unsupported";
  case AARBF_ERR_NO_DEVICE: return "no HIP device";
  default: return "unknown status";
  }
}

int aarbf_context_create(int device, aarbf_context **out) {
  if (!out) return AARBF_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return AARBF_ERR_NO_DEVICE;
  if (device < 0 || device >= count) return AARBF_ERR_INVALID_ARGUMENT;
  auto *ctx = new (std::nothrow) aarbf_context();
  if (!ctx) return AARBF_ERR_OUT_OF_MEMORY;
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&ctx->d_stats, sizeof(uint64_t) * 32) != hipSuccess ||
      hipMalloc(&ctx->d_counter, sizeof(uint32_t) * 4) != hipSuccess ||
      hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipMemset(ctx->d_stats, 0, sizeof(uint64_t) * 32) != hipSuccess ||
      aarbf::noalloc_init(device) != AARBF_OK) {
    if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
    if (ctx->d_stats) hipFree(ctx->d_stats);
    if (ctx->d_counter) hipFree(ctx->d_counter);
    delete ctx;
    return AARBF_ERR_HIP;
  }
  ctx->stream = ctx->own_stream;
  *out = ctx;
  return AARBF_OK;
}

int aarbf_context_destroy(aarbf_context *ctx) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  for (uint64_t h : ctx->tfs) {
    TF *t = reinterpret_cast<TF *>(h);
    hipFree(t->table);
    delete t;
  }
  for (uint64_t h : ctx->stbns) {
    STBN *s = reinterpret_cast<STBN *>(h);
    hipFree(s->rgba);
    delete s;
  }
  for (uint64_t h : ctx->bvhs) {
    BVH *b = reinterpret_cast<BVH *>(h);
    hipFree(b->nodes);
    if (b->groups) hipFree(b->groups);
    delete b;
  }
  if (ctx->scratch) hipFree(ctx->scratch);
  if (ctx->baked) hipFree(ctx->baked);
  if (ctx->lat_buf) hipFree(ctx->lat_buf);
  for (int k = 0; k < 2; ++k) {
    if (ctx->stage[k]) hipHostFree(ctx->stage[k]);
    if (ctx->stage_ev[k]) hipEventDestroy(ctx->stage_ev[k]);
  }
  for (auto &g : ctx->graphs) hipGraphExecDestroy(g.second);
  if (ctx->d_stats) hipFree(ctx->d_stats);
  if (ctx->d_counter) hipFree(ctx->d_counter);
  for (hipEvent_t &e : ctx->tev)
    if (e) hipEventDestroy(e);
  if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return AARBF_OK;
}

int aarbf_set_stream(aarbf_context *ctx, void *s) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  ctx->stream = s ? static_cast<hipStream_t>(s) : ctx->own_stream;
  return AARBF_OK;
}

void *aarbf_get_stream(aarbf_context *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int aarbf_sync(aarbf_context *ctx) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return AARBF_OK;
}

const char *aarbf_last_error(const aarbf_context *ctx) {
  return ctx ? ctx->last_error.c_str() : "null context";
}

int aarbf_set_render_mode(aarbf_context *ctx, int mode) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, mode >= 0 && mode <= 2, "aarbf_set_render_mode: mode must be 0, 1 or 2");
  ctx->render_mode = mode;
  return AARBF_OK;
}

int aarbf_set_tuning(aarbf_context *ctx, int param, int value) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  switch (param) {
  case AARBF_TUNE_STEP_WAVES:
    AARBF_REQUIRE(ctx, value == 0 || (value >= 4 && value <= 6), "aarbf_set_tuning: step waves must be 0, 4, 5 or 6");
    ctx->tune_step_waves = value;
    return AARBF_OK;
  case AARBF_TUNE_STEP_CAP:
    AARBF_REQUIRE(ctx, value >= 16 && value <= 65536, "aarbf_set_tuning: step cap must be in [16, 65536]");
    ctx->tune_step_cap = value;
    return AARBF_OK;
  case AARBF_TUNE_GRAPH_ROUNDS:
    AARBF_REQUIRE(ctx, value >= 0 && value <= 64 && value % 2 == 0,
                  "aarbf_set_tuning: graph rounds must be 0 or an even number up to 64");
    ctx->tune_graph_rounds = value;
    return AARBF_OK;
  case AARBF_TUNE_WALK_BUDGET:
    AARBF_REQUIRE(ctx, value >= 1 && value <= 65536, "aarbf_set_tuning: walk budget must be in [1, 65536]");
    ctx->tune_walk_budget = value;
    return AARBF_OK;
  case AARBF_TUNE_DIFF_WAVES:
    AARBF_REQUIRE(ctx, value == 0 || value == 1, "aarbf_set_tuning: diff waves must be 0 or 1");
    ctx->tune_diff_waves = value;
    return AARBF_OK;
  case AARBF_TUNE_STEP_MIN_LIVE:
    AARBF_REQUIRE(ctx, value >= 0 && value <= 64, "aarbf_set_tuning: step min-live lanes must be in [0, 64]");
    ctx->tune_step_min_live = value;
    return AARBF_OK;
  default:
    return set_error(ctx, AARBF_ERR_INVALID_ARGUMENT, "aarbf_set_tuning: unknown parameter");
  }
}

int aarbf_set_stats(aarbf_context *ctx, int enable) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  ctx->stats = enable != 0;
  return AARBF_OK;
}

int aarbf_render_stats_read(aarbf_context *ctx, aarbf_render_stats *out, int reset) {
  if (!ctx || !out) return AARBF_ERR_INVALID_ARGUMENT;
  uint64_t v[8];
  AARBF_HIP(ctx, hipMemcpyAsync(v, ctx->d_stats, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
  AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  out->samples = v[0];
  out->steps = v[1];
  out->queries = v[2];
  out->nodes = v[3];
  out->leaves = v[4];
  out->particles = v[5];
  out->covering = v[6];
  out->exhausted = v[7];
  if (reset) AARBF_HIP(ctx, hipMemsetAsync(ctx->d_stats, 0, sizeof(v), ctx->stream));
  return AARBF_OK;
}

int aarbf_set_timing(aarbf_context *ctx, int enable) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  if (enable)
    for (hipEvent_t &e : ctx->tev)
      if (!e) AARBF_HIP(ctx, hipEventCreate(&e));
  ctx->timing = enable != 0;
  return AARBF_OK;
}

int aarbf_batch_latency(aarbf_context *ctx, double *ms, uint32_t n) {
  if (!ctx || !ms) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, ctx->lat_n > 0 && n == ctx->lat_n, "aarbf_batch_latency: n differs from the last batch");
  std::vector<unsigned long long> h(2 * (size_t)n + 1);
  AARBF_HIP(ctx, hipMemcpyAsync(h.data(), ctx->lat_buf, sizeof(unsigned long long) * h.size(),
                                hipMemcpyDeviceToHost, ctx->stream));
  AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  int khz = 0;
  AARBF_HIP(ctx, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
  AARBF_REQUIRE(ctx, khz > 0, "aarbf_batch_latency: no wall clock rate");
  const unsigned long long t0 = h[2 * (size_t)n];
  for (uint32_t i = 0; i < n; ++i) {
    const unsigned long long t = h[2 * (size_t)i + 1];
    ms[i] = t >= t0 && t ? (double)(t - t0) / (double)khz : -1.0;
  }
  return AARBF_OK;
}

int aarbf_read_timing(aarbf_context *ctx, double *ms, uint64_t *launches, int reset) {
  if (!ctx || !ms || !launches) return AARBF_ERR_INVALID_ARGUMENT;
  for (int k = 0; k < 4; ++k) {
    ms[k] = ctx->tms[k];
    launches[k] = ctx->tlaunch[k];
    if (reset) {
      ctx->tms[k] = 0.0;
      ctx->tlaunch[k] = 0;
    }
  }
  return AARBF_OK;
}

int aarbf_render_histogram(aarbf_context *ctx, uint64_t *out, int reset) {
  if (!ctx || !out) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_HIP(ctx, hipMemcpyAsync(out, ctx->d_stats + 8, sizeof(uint64_t) * 24, hipMemcpyDeviceToHost, ctx->stream));
  AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (reset) AARBF_HIP(ctx, hipMemsetAsync(ctx->d_stats + 8, 0, sizeof(uint64_t) * 24, ctx->stream));
  return AARBF_OK;
}

// ---- transfer functions (hostCode.cpp:355-365,524-572) --------------------

int aarbf_tf_create(aarbf_context *ctx, const float *rgba, uint32_t n, uint64_t *handle) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, rgba && handle && n >= 1 && n <= 65536, "aarbf_tf_create: need 1..65536 texels");
  TF *t = new (std::nothrow) TF();
  if (!t) return AARBF_ERR_OUT_OF_MEMORY;
  t->n = n;
  hipError_t e = hipMalloc(&t->table, sizeof(float4) * n);
  if (e == hipSuccess) e = hipMemcpy(t->table, rgba, sizeof(float4) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (t->table) hipFree(t->table);
    delete t;
    return hip_error(ctx, e, "aarbf_tf_create");
  }
  *handle = reinterpret_cast<uint64_t>(t);
  ctx->tfs.insert(*handle);
  return AARBF_OK;
}

// sRGB electro-optical transfer function (IEC 61966-2-1), applied to the RGB
// of R8G8B8A8_SRGB texels on sampling.
static float srgb_to_linear(uint8_t v) {
  double c = v / 255.0;
  double l = c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
  return (float)l;
}

int aarbf_tf_create_rgba8(aarbf_context *ctx, const uint8_t *rgba8, uint32_t n, int srgb, uint64_t *handle) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, rgba8 && handle && n >= 1 && n <= 65536, "aarbf_tf_create_rgba8: need 1..65536 texels");
  std::vector<float> f(4 * (size_t)n);
  for (uint32_t i = 0; i < n; ++i)
    for (int c = 0; c < 4; ++c) {
      uint8_t v = rgba8[4 * i + c];
      f[4 * i + c] = (srgb && c < 3) ? srgb_to_linear(v) : (float)v / 255.0f;
    }
  return aarbf_tf_create(ctx, f.data(), n, handle);
}

int aarbf_tf_create_from_marks(aarbf_context *ctx, const aarbf_tf_mark *marks, uint32_t n, int kind,
                               uint64_t *handle) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  uint8_t texels[4 * AARBF_TF_TEXELS];
  AARBF_REQUIRE(ctx, aarbf_tf_texels_from_marks(marks, n, kind, texels) == AARBF_OK,
                "aarbf_tf_create_from_marks: need 1 or more marks with positions in [0, 1] and a valid kind");
  return aarbf_tf_create_rgba8(ctx, texels, AARBF_TF_TEXELS, 1, handle);
}

int aarbf_tf_destroy(aarbf_context *ctx, uint64_t handle) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  TF *t = get_tf(ctx, handle);
  AARBF_REQUIRE(ctx, t, "aarbf_tf_destroy: unknown handle");
  hipStreamSynchronize(ctx->stream);
  hipFree(t->table);
  delete t;
  ctx->tfs.erase(handle);
  return AARBF_OK;
}

// ---- blue-noise atlas (hostCode.cpp:336-352) ------------------------------

int aarbf_stbn_create(aarbf_context *ctx, const uint8_t *rgba8, int width, int height, uint64_t *handle) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, rgba8 && handle && width > 0 && height > 0, "aarbf_stbn_create: bad atlas");
  STBN *s = new (std::nothrow) STBN();
  if (!s) return AARBF_ERR_OUT_OF_MEMORY;
  s->width = width;
  s->height = height;
  size_t bytes = (size_t)width * height * 4;
  hipError_t e = hipMalloc(&s->rgba, bytes);
  if (e == hipSuccess) e = hipMemcpy(s->rgba, rgba8, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (s->rgba) hipFree(s->rgba);
    delete s;
    return hip_error(ctx, e, "aarbf_stbn_create");
  }
  *handle = reinterpret_cast<uint64_t>(s);
  ctx->stbns.insert(*handle);
  return AARBF_OK;
}

int aarbf_stbn_destroy(aarbf_context *ctx, uint64_t handle) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  STBN *s = get_stbn(ctx, handle);
  AARBF_REQUIRE(ctx, s, "aarbf_stbn_destroy: unknown handle");
  hipStreamSynchronize(ctx->stream);
  hipFree(s->rgba);
  delete s;
  ctx->stbns.erase(handle);
  return AARBF_OK;
}

// Layout probe used by the Python bindings to check their struct mirrors.
int aarbf_handle_counts(const aarbf_context *ctx, uint32_t counts[3]) {
  if (!ctx || !counts) return AARBF_ERR_INVALID_ARGUMENT;
  counts[0] = (uint32_t)ctx->tfs.size();
  counts[1] = (uint32_t)ctx->stbns.size();
  counts[2] = (uint32_t)ctx->bvhs.size();
  return AARBF_OK;
}

uint32_t aarbf_abi_layout(uint32_t *out, uint32_t n) {
  uint32_t v[] = {
      (uint32_t)sizeof(aarbf_push_constants),
      (uint32_t)offsetof(aarbf_push_constants, light),
      (uint32_t)offsetof(aarbf_push_constants, camera),
      (uint32_t)offsetof(aarbf_push_constants, rbfRadius),
      (uint32_t)offsetof(aarbf_push_constants, world),
      (uint32_t)offsetof(aarbf_push_constants, numParticles),
      (uint32_t)offsetof(aarbf_push_constants, colormap),
      (uint32_t)offsetof(aarbf_push_constants, radii),
      (uint32_t)sizeof(aarbf_bounds_constants),
      (uint32_t)offsetof(aarbf_bounds_constants, particlesPerLeaf),
      (uint32_t)offsetof(aarbf_bounds_constants, radiusmap),
      (uint32_t)sizeof(aarbf_taa_constants),
      (uint32_t)offsetof(aarbf_taa_constants, imageTexture),
      (uint32_t)sizeof(aarbf_raygen_data),
      (uint32_t)offsetof(aarbf_raygen_data, fbSize),
      (uint32_t)offsetof(aarbf_raygen_data, globalAABBMax),
      (uint32_t)sizeof(aarbf_render_stats),
      (uint32_t)sizeof(aarbf_request),
      (uint32_t)offsetof(aarbf_request, pc),
      (uint32_t)offsetof(aarbf_request, spp),
      (uint32_t)sizeof(aarbf_tune_result),
      (uint32_t)offsetof(aarbf_tune_result, walk_budget),
      (uint32_t)offsetof(aarbf_tune_result, step_min_live),
      (uint32_t)sizeof(aarbf_camera_state),
      (uint32_t)offsetof(aarbf_camera_state, firstFrame),
      (uint32_t)sizeof(aarbf_camera_input),
      (uint32_t)offsetof(aarbf_camera_input, left_shift),
  };
  uint32_t m = (uint32_t)(sizeof(v) / sizeof(v[0]));
  for (uint32_t i = 0; i < n && i < m; ++i) out[i] = v[i];
  return m;
}

} // extern "C"

// internal.h -- shared host/device definitions of the aarbf library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "aarbf.h"

namespace aarbf {

// ---- transfer function / blue noise / BVH objects behind the handles ----

struct TF {
  float4 *table = nullptr; // device RGBA texels
  uint32_t n = 0;
};

struct STBN {
  uint8_t *rgba = nullptr; // device RGBA8 texels
  int width = 0, height = 0;
};

constexpr int kMaxLevels = 32;

// Hilbert-packed BVH: level 0 holds one (widened) box per leaf, node j of
// level l+1 bounds nodes [j*K, j*K+K) of level l.  Every node is two float4
// (lo.xyz, hi.xyz) in one allocation; level_ofs is in float4 units.
//
// For K <= 8 the internal nodes are also stored quantised: node j of level
// l >= 1 owns one QGroup holding its own box corner and per-axis step and
// its K children's boxes as 8-bit offsets, rounded outwards (80 bytes per
// group instead of K * 32 bytes of float boxes).
struct QGroup {
  float4 a;       // lo.x, lo.y, lo.z, step.x
  float4 b;       // step.y, step.z, 0, 0
  uint32_t q[12]; // bytes: lo.x[8], lo.y[8], lo.z[8], hi.x[8], hi.y[8], hi.z[8]; lo.x > hi.x marks an empty child
};

struct BVH {
  uint32_t K = 8, log2K = 3;
  uint32_t num_levels = 0; // including level 0
  uint32_t counts[kMaxLevels] = {};
  uint64_t level_ofs[kMaxLevels] = {};
  float4 *nodes = nullptr;
  uint64_t total_nodes = 0;
  QGroup *groups = nullptr;            // K <= 8 only
  uint64_t group_ofs[kMaxLevels] = {}; // first group of level l (l >= 1)
  uint64_t total_groups = 0;
};

// Device view of a BVH passed by value to kernels.
struct BVHView {
  const float4 *nodes;
  const QGroup *groups;
  uint32_t log2K;
  uint32_t top; // root level
  uint32_t counts[kMaxLevels];
  uint32_t level_ofs[kMaxLevels]; // in float4 units (< 2^32 for <= 2^30 nodes)
  uint32_t group_ofs[kMaxLevels];
};

} // namespace aarbf

struct aarbf_context {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  std::string last_error;
  bool stats = false;
  uint64_t *d_stats = nullptr; // 8 counters
  uint32_t *d_counter = nullptr; // work counter of the persistent render kernel
  int render_mode = 0;           // 0 wavefront gather/step, 1 per-pixel point queries, 2 majorant tracking
  int diff_estimator = 0;        // differentiable renderer: 0 expected marching, 1 ratio tracking
  int cus = 0;                   // compute units of the device
  int tune_step_waves = 0;       // occupancy target of the step kernel (0 = default, 4 waves per SIMD)
  int tune_step_cap = 8192;      // tracker steps per ray per wavefront round (base, see step_budget)
  int tune_graph_rounds = 0;     // wavefront rounds per captured hipGraph (0 = host-driven rounds)
  int tune_walk_budget = 256;    // nodes a one-lane gather walk opens before handing its ray to a wave
  int tune_step_min_live = 16;   // step-kernel waves leave their loop below this many live lanes (0 = never)
  int tune_diff_waves = 1;       // differentiable expected marching: 0 a lane per sample, 1 a wave per sample
  std::vector<std::pair<std::string, hipGraphExec_t>> graphs; // captured rounds, keyed by their arguments
  // kernel timing (HIP events around each kernel class of the render rounds)
  bool timing = false;
  hipEvent_t tev[4] = {nullptr, nullptr, nullptr, nullptr};
  double tms[4] = {0, 0, 0, 0};   // gather, step, dense step, other
  uint64_t tlaunch[4] = {0, 0, 0, 0};
  // grow-only device scratch
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  // baked particle payload
  float4 *baked = nullptr;
  size_t baked_count = 0;
  int payload = 0;          // 0 FP32 (x, y, z, r) + (alpha^3, rgb^2.2); 1 FP16 leaf-local
  int baked_fmt = 0;        // format and size of the last bake
  uint32_t baked_n = 0;
  // per-view latency record of the last aarbf_render_batch: 2 words per view
  // (samples finished, wall clock at the last) and the clock at the start
  unsigned long long *lat_buf = nullptr;
  // pinned staging buffers of the streaming importer (aarbf_import_step_device)
  void *stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  uint32_t lat_cap = 0, lat_n = 0;
  std::set<uint64_t> tfs, stbns, bvhs;
};

namespace aarbf {

int set_error(aarbf_context *ctx, int status, const std::string &msg);
int hip_error(aarbf_context *ctx, hipError_t e, const char *what);
// Ensures ctx->scratch holds at least `bytes`; returns nullptr on failure.
void *scratch(aarbf_context *ctx, size_t bytes);
TF *get_tf(aarbf_context *ctx, uint64_t h);
STBN *get_stbn(aarbf_context *ctx, uint64_t h);
BVH *get_bvh(aarbf_context *ctx, uint64_t h);
BVHView make_view(const BVH &b);
// Takes the current device's address of the allocation-free render entry
// point's scratch array (render.hip); called once per device.
int noalloc_init(int device);

} // namespace aarbf

#define AARBF_HIP(ctx, call)                                                   \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) return aarbf::hip_error((ctx), e_, #call);           \
  } while (0)

#define AARBF_CHECK_LAUNCH(ctx, what)                                          \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) return aarbf::hip_error((ctx), e_, what);            \
  } while (0)

#define AARBF_REQUIRE(ctx, cond, msg)                                          \
  do {                                                                         \
    if (!(cond)) return aarbf::set_error((ctx), AARBF_ERR_INVALID_ARGUMENT, (msg)); \
  } while (0)

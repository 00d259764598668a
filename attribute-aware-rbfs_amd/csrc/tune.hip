// tune.hip -- the schedule autotuner: times a representative batch of
// views under candidate schedule parameters and keeps the fastest.  No
// counterpart in the reference (its GPRT/Vulkan ray-tracing pipeline has no
// software schedule to tune).
//
// The parameters are the step budget per ray per round (how deep each
// round drains a ray before it re-queues), the step kernel's occupancy
// target (waves per SIMD, which sets its register budget, i.e. the
// workgroup shape on a CU) and the number of rounds captured per hipGraph
// (0 = host-driven rounds), the node budget of a one-lane gather walk
// before it hands its ray to a one-wave gather (how many walks the wave
// gather takes depends on the scene), and the live-lane count below which a
// step-kernel wave leaves its loop so its rays are repacked in the next
// round.  None changes an image, only its cost.  The BVH
// leaf size is a property of the scene and is tuned one level up, by
// rebuilding the scene per leaf size (aarbf.tune in the Python package).
//
// Search: coordinate descent from the current setting -- the step budget
// over {2048, 4096, 8192, 16384}, then the occupancy over {4, 5, 6}, then the
// graph rounds over {0, 8}, then the walk budget over {256, 1024, 65536},
// then the step kernel's live-lane threshold over {0, 16} -- each candidate
// timed on the stream with HIP
// events (best of `reps` runs after one warm-up run).  The views render
// into the tuner's own buffers; the caller's images are not touched.
#include <vector>

#include "internal.h"

using namespace aarbf;

extern "C" {

int aarbf_autotune(aarbf_context *ctx, const aarbf_request *reqs, uint32_t n, uint32_t reps,
                   aarbf_tune_result *out) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, reqs && n >= 1 && out, "aarbf_autotune: no requests");
  AARBF_REQUIRE(ctx, !ctx->timing && !ctx->stats, "aarbf_autotune: disable timing and statistics first");
  hipSetDevice(ctx->device);
  reps = reps ? reps : 1;
  // private image + accum buffers
  size_t px = 0;
  for (uint32_t i = 0; i < n; ++i) {
    AARBF_REQUIRE(ctx, reqs[i].width > 0 && reqs[i].height > 0, "aarbf_autotune: empty request");
    px += (size_t)reqs[i].width * (size_t)reqs[i].height;
  }
  float *buf = nullptr;
  AARBF_HIP(ctx, hipMalloc(&buf, px * 2 * 4 * sizeof(float)));
  std::vector<aarbf_request> R(reqs, reqs + n);
  size_t off = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const size_t p = (size_t)R[i].width * (size_t)R[i].height;
    R[i].rg.imageBuffer = (uint64_t)(buf + 4 * off);
    R[i].rg.accumBuffer = (uint64_t)(buf + 4 * (px + off));
    off += p;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = AARBF_OK;
  const int keep_cap = ctx->tune_step_cap, keep_waves = ctx->tune_step_waves, keep_graph = ctx->tune_graph_rounds,
            keep_budget = ctx->tune_walk_budget, keep_live = ctx->tune_step_min_live;
  auto run = [&](int cap, int waves, int graph, int budget, int live, double &ms) -> int {
    ctx->tune_step_cap = cap;
    ctx->tune_step_waves = waves;
    ctx->tune_graph_rounds = graph;
    ctx->tune_walk_budget = budget;
    ctx->tune_step_min_live = live;
    ms = 1e300;
    for (uint32_t r = 0; r <= reps; ++r) {
      AARBF_HIP(ctx, hipMemsetAsync(buf, 0, px * 2 * 4 * sizeof(float), ctx->stream));
      AARBF_HIP(ctx, hipEventRecord(e0, ctx->stream));
      int st = aarbf_render_batch(ctx, R.data(), n);
      if (st != AARBF_OK) return st;
      AARBF_HIP(ctx, hipEventRecord(e1, ctx->stream));
      AARBF_HIP(ctx, hipEventSynchronize(e1));
      float t = 0.0f;
      AARBF_HIP(ctx, hipEventElapsedTime(&t, e0, e1));
      if (r > 0 && t < ms) ms = t; // run 0 warms up (and captures graphs)
    }
    return AARBF_OK;
  };
  do {
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      rc = set_error(ctx, AARBF_ERR_HIP, "aarbf_autotune: events");
      break;
    }
    int best_cap = keep_cap, best_waves = keep_waves, best_graph = keep_graph, best_budget = keep_budget,
        best_live = keep_live;
    double best = 0.0;
    if ((rc = run(best_cap, best_waves, best_graph, best_budget, best_live, best)) != AARBF_OK) break;
    out->baseline_ms = best;
    uint32_t tried = 1;
    const int caps[] = {2048, 4096, 8192, 16384}, waves[] = {4, 5, 6}, graphs[] = {0, 8},
              budgets[] = {256, 1024, 65536}, lives[] = {0, 16};
    for (int c : caps) {
      if (c == best_cap) continue;
      double ms;
      if ((rc = run(c, best_waves, best_graph, best_budget, best_live, ms)) != AARBF_OK) break;
      ++tried;
      if (ms < best) best = ms, best_cap = c;
    }
    if (rc != AARBF_OK) break;
    for (int w : waves) {
      if (w == (best_waves ? best_waves : 4)) continue;
      double ms;
      if ((rc = run(best_cap, w, best_graph, best_budget, best_live, ms)) != AARBF_OK) break;
      ++tried;
      if (ms < best) best = ms, best_waves = w;
    }
    if (rc != AARBF_OK) break;
    for (int g : graphs) {
      if (g == best_graph) continue;
      double ms;
      if ((rc = run(best_cap, best_waves, g, best_budget, best_live, ms)) != AARBF_OK) break;
      ++tried;
      if (ms < best) best = ms, best_graph = g;
    }
    if (rc != AARBF_OK) break;
    for (int b : budgets) {
      if (b == best_budget) continue;
      double ms;
      if ((rc = run(best_cap, best_waves, best_graph, b, best_live, ms)) != AARBF_OK) break;
      ++tried;
      if (ms < best) best = ms, best_budget = b;
    }
    if (rc != AARBF_OK) break;
    for (int l : lives) {
      if (l == best_live) continue;
      double ms;
      if ((rc = run(best_cap, best_waves, best_graph, best_budget, l, ms)) != AARBF_OK) break;
      ++tried;
      if (ms < best) best = ms, best_live = l;
    }
    if (rc != AARBF_OK) break;
    out->step_cap = best_cap;
    out->step_waves = best_waves;
    out->graph_rounds = best_graph;
    out->walk_budget = best_budget;
    out->step_min_live = best_live;
    out->best_ms = best;
    out->candidates = tried;
  } while (false);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipStreamSynchronize(ctx->stream);
  hipFree(buf);
  if (rc == AARBF_OK) {
    ctx->tune_step_cap = out->step_cap;
    ctx->tune_step_waves = out->step_waves;
    ctx->tune_graph_rounds = out->graph_rounds;
    ctx->tune_walk_budget = out->walk_budget;
    ctx->tune_step_min_live = out->step_min_live;
  } else {
    ctx->tune_step_cap = keep_cap;
    ctx->tune_step_waves = keep_waves;
    ctx->tune_graph_rounds = keep_graph;
    ctx->tune_walk_budget = keep_budget;
    ctx->tune_step_min_live = keep_live;
  }
  return rc;
}

} // extern "C"

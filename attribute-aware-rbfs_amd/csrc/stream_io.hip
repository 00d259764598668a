// stream_io.hip -- the importer's device load path: one timestep of the
// importParticles.h:8-32 format read straight into device memory.
//
// The reference reads every timestep into host vectors and copies the
// current one through a mapped buffer (hostCode.cpp:138,759-763).  Here
// the binary file is read in 64 MiB chunks into two pinned staging buffers
// (allocated once per context); the copy of one chunk to the device runs
// on the context's stream while the next chunk is read, so a large
// timestep streams at the slower of disk and PCIe bandwidth.
#include <cstdio>
#include <fstream>
#include <string>

#include "internal.h"

using namespace aarbf;

namespace {
constexpr size_t kStageBytes = size_t(64) << 20;

int ensure_stage(aarbf_context *ctx) {
  for (int k = 0; k < 2; ++k) {
    if (!ctx->stage[k]) AARBF_HIP(ctx, hipHostMalloc(&ctx->stage[k], kStageBytes, hipHostMallocDefault));
    if (!ctx->stage_ev[k]) AARBF_HIP(ctx, hipEventCreateWithFlags(&ctx->stage_ev[k], hipEventDisableTiming));
  }
  return AARBF_OK;
}
} // namespace

extern "C" int aarbf_import_step_device(aarbf_context *ctx, const char *path, uint32_t step, uint64_t particles,
                                        uint64_t capacity, uint64_t *count) {
  if (!ctx) return AARBF_ERR_INVALID_ARGUMENT;
  AARBF_REQUIRE(ctx, path && count && (particles || capacity == 0), "aarbf_import_step_device: null argument");
  *count = 0;
  std::ifstream hrf(path);
  if (!hrf) return set_error(ctx, AARBF_ERR_IO, std::string("aarbf_import_step_device: cannot open ") + path);
  std::string name;
  uint32_t k = 0;
  bool found = false;
  while (std::getline(hrf, name))
    if (k++ == step) {
      found = true;
      break;
    }
  AARBF_REQUIRE(ctx, found, "aarbf_import_step_device: the header lists fewer timesteps");
  const std::string file = std::string(path) + "_" + name;
  FILE *f = std::fopen(file.c_str(), "rb");
  if (!f) return set_error(ctx, AARBF_ERR_IO, "aarbf_import_step_device: cannot open " + file);
  uint64_t n = 0;
  if (std::fread(&n, 8, 1, f) != 1) {
    std::fclose(f);
    return set_error(ctx, AARBF_ERR_FORMAT, "aarbf_import_step_device: no particle count in " + file);
  }
  *count = n;
  if (n > capacity) {
    std::fclose(f);
    return set_error(ctx, AARBF_ERR_INVALID_ARGUMENT,
                     "aarbf_import_step_device: the timestep has more particles than the buffer holds");
  }
  int st = ensure_stage(ctx);
  if (st) {
    std::fclose(f);
    return st;
  }
  hipSetDevice(ctx->device);
  const uint64_t bytes = 16 * n;
  uint64_t done = 0;
  int buf = 0;
  while (done < bytes) {
    const size_t len = (size_t)std::min<uint64_t>(kStageBytes, bytes - done);
    // the staging buffer's previous copy must have finished before it is refilled
    hipError_t e = hipEventSynchronize(ctx->stage_ev[buf]);
    if (e == hipSuccess && std::fread(ctx->stage[buf], 1, len, f) != len) {
      std::fclose(f);
      hipStreamSynchronize(ctx->stream);
      return set_error(ctx, AARBF_ERR_FORMAT, "aarbf_import_step_device: " + file + " is shorter than its count");
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync((char *)particles + done, ctx->stage[buf], len, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->stage_ev[buf], ctx->stream);
    if (e != hipSuccess) {
      std::fclose(f);
      hipStreamSynchronize(ctx->stream);
      return hip_error(ctx, e, "aarbf_import_step_device");
    }
    done += len;
    buf ^= 1;
  }
  std::fclose(f);
  // the staging buffers are reused by the next call: the copies finish here
  AARBF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return AARBF_OK;
}

// dropin_snippet.cpp -- the C++ binding a reference maintainer adds to
// hostCode.cpp, run end to end against include/aarbf.h and
// libaarbf.so: the reference's particle file -> device pre-process
// (hostCode.cpp:183-250) -> RGBA8 sRGB transfer functions (:355-365) ->
// GenRBFBounds + BVH with the reference's radius model (pc.radii = 0,
// :770-789) -> per frame ParticleRBFRayGen + CompositeGui + history copy,
// with accumID / frameID incremented as the render loop does (:813-830).
//
// usage: dropin_snippet <particle header> <texels: 3 x 64 RGBA8> <out prefix>
//                       W H frames ppl rbfRadius frameID from.xyz at.xyz up.xyz cosFovy
// writes <out>_image.f32, <out>_accum.f32, <out>_taa.f32 (float4 per pixel)
// and <out>_frame.u32 (BGRA8 words) after the last frame.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "aarbf.h"

#define CHECK(call)                                                                           \
  do {                                                                                        \
    int st_ = (call);                                                                         \
    if (st_ != AARBF_OK) {                                                                    \
      std::fprintf(stderr, "%s failed: %s (%s)\n", #call, aarbf_status_string(st_),          \
                   ctx ? aarbf_last_error(ctx) : "");                                         \
      return 1;                                                                               \
    }                                                                                         \
  } while (0)

#define HIPCHECK(call)                                                                        \
  do {                                                                                        \
    if ((call) != hipSuccess) {                                                               \
      std::fprintf(stderr, "%s failed\n", #call);                                             \
      return 1;                                                                               \
    }                                                                                         \
  } while (0)

static bool write_file(const char *prefix, const char *suffix, const void *data, size_t bytes) {
  char path[4096];
  std::snprintf(path, sizeof(path), "%s%s", prefix, suffix);
  FILE *f = std::fopen(path, "wb");
  if (!f) return false;
  bool ok = std::fwrite(data, 1, bytes, f) == bytes;
  return std::fclose(f) == 0 && ok;
}

int main(int argc, char **argv) {
  aarbf_context *ctx = nullptr;
  if (argc != 20) {
    std::fprintf(stderr, "usage: %s header texels out W H frames ppl rbfRadius frameID from(3) at(3) up(3) cosFovy\n",
                 argv[0]);
    return 2;
  }
  const char *header = argv[1], *texel_path = argv[2], *out = argv[3];
  const int W = std::atoi(argv[4]), H = std::atoi(argv[5]), frames = std::atoi(argv[6]);
  const uint16_t ppl = (uint16_t)std::atoi(argv[7]);
  const float rbfRadius = (float)std::atof(argv[8]);
  const int frameID = std::atoi(argv[9]);
  float from[3], at[3], up[3];
  for (int k = 0; k < 3; ++k) {
    from[k] = (float)std::atof(argv[10 + k]);
    at[k] = (float)std::atof(argv[13 + k]);
    up[k] = (float)std::atof(argv[16 + k]);
  }
  const float cosFovy = (float)std::atof(argv[19]);

  // importParticles (importParticles.h:8-32): timestep 0
  aarbf_timesteps *ts = nullptr;
  CHECK(aarbf_import_particles(header, &ts));
  const float *xyzw = nullptr;
  uint64_t n = 0;
  const char *name = nullptr;
  CHECK(aarbf_timesteps_get(ts, 0, &xyzw, &n, &name));
  std::vector<unsigned char> texels(3 * 64 * 4);
  FILE *tf = std::fopen(texel_path, "rb");
  if (!tf || std::fread(texels.data(), 1, texels.size(), tf) != texels.size()) return 3;
  std::fclose(tf);

  CHECK(aarbf_context_create(0, &ctx));
  const uint32_t numAABBs = (uint32_t)((n + ppl - 1) / ppl);
  float *d_particles = nullptr, *d_aabbs = nullptr;
  uint64_t *d_codes = nullptr;
  float *d_image = nullptr, *d_accum = nullptr, *d_taa = nullptr, *d_taaPrev = nullptr;
  uint32_t *d_frame = nullptr;
  const size_t px = (size_t)W * H;
  HIPCHECK(hipMalloc(&d_particles, 16 * n));
  HIPCHECK(hipMalloc(&d_codes, 8 * n));
  HIPCHECK(hipMalloc(&d_aabbs, 24 * (size_t)numAABBs));
  HIPCHECK(hipMalloc(&d_image, 16 * px));
  HIPCHECK(hipMalloc(&d_accum, 16 * px));
  HIPCHECK(hipMalloc(&d_taa, 16 * px));
  HIPCHECK(hipMalloc(&d_taaPrev, 16 * px));
  HIPCHECK(hipMalloc(&d_frame, 4 * px));
  HIPCHECK(hipMemset(d_accum, 0, 16 * px));
  HIPCHECK(hipMemset(d_taa, 0, 16 * px));
  HIPCHECK(hipMemset(d_taaPrev, 0, 16 * px));
  HIPCHECK(hipMemcpy(d_particles, xyzw, 16 * n, hipMemcpyHostToDevice)); // replaces Map / memcpy / Unmap

  // pre-process on the device (hostCode.cpp:183-250)
  float aabb[6] = {1e38f, 1e38f, 1e38f, -1e38f, -1e38f, -1e38f};
  CHECK(aarbf_spacetime_aabb(ctx, (uint64_t)d_particles, n, aabb));
  CHECK(aarbf_hilbert_codes(ctx, (uint64_t)d_particles, n, aabb, (uint64_t)d_codes));
  CHECK(aarbf_sort_particles(ctx, (uint64_t)d_particles, (uint64_t)d_codes, n, 0));
  float range[2] = {1e20f, -1e20f};
  CHECK(aarbf_attribute_range(ctx, (uint64_t)d_particles, n, range));
  CHECK(aarbf_normalize_attributes(ctx, (uint64_t)d_particles, n, range));

  // the reference's three 64-texel R8G8B8A8_SRGB transfer functions
  uint64_t cm = 0, rm = 0, dm = 0;
  CHECK(aarbf_tf_create_rgba8(ctx, texels.data(), 64, 1, &cm));
  CHECK(aarbf_tf_create_rgba8(ctx, texels.data() + 256, 64, 1, &rm));
  CHECK(aarbf_tf_create_rgba8(ctx, texels.data() + 512, 64, 1, &dm));

  aarbf_bounds_constants bc = {(uint64_t)d_aabbs, numAABBs, (uint64_t)d_particles, (uint32_t)n, ppl, rbfRadius, rm, 0};
  CHECK(aarbf_gen_bounds(ctx, &bc));
  uint64_t world = 0;
  CHECK(aarbf_bvh_build(ctx, (uint64_t)d_aabbs, numAABBs, 0, &world));

  // rtConstants as hostCode.cpp:395-409,481-486 fill them
  aarbf_push_constants pc = {};
  pc.accumID = 1;
  pc.frameID = (uint16_t)frameID;
  pc.visualizeAttributes = 1;
  pc.particlesPerLeaf = ppl;
  pc.light.azimuth = 0.0f;
  pc.light.elevation = 0.0f;
  pc.light.ambient = 0.5f;
  CHECK(aarbf_camera_basis(from, at, up, cosFovy, W, H, &pc));
  pc.rbfRadius = rbfRadius;
  pc.unit = rbfRadius * 0.1f;
  pc.world = world;
  pc.particles = (uint64_t)d_particles;
  pc.numParticles = (uint32_t)n;
  pc.colormap = cm;
  pc.radiusmap = rm;
  pc.densitymap = dm;
  pc.radii = 0; // the reference's radius model: rbfRadius * radiusmap(w)
  aarbf_raygen_data rg = {0, (uint64_t)d_image, (uint64_t)d_accum, {W, H}, {aabb[0], aabb[1], aabb[2]},
                          {aabb[3], aabb[4], aabb[5]}};
  aarbf_taa_constants tc = {0, 0, {W, H}, (uint64_t)d_image, 0, (uint64_t)d_frame, (uint64_t)d_taa,
                            (uint64_t)d_taaPrev};
  for (int f = 0; f < frames; ++f) {
    CHECK(aarbf_render_rbf(ctx, &rg, &pc, W, H));
    CHECK(aarbf_taa_resolve(ctx, &tc));
    pc.accumID++;
    pc.frameID++;
  }
  CHECK(aarbf_sync(ctx));

  std::vector<float> buf(4 * px);
  std::vector<uint32_t> words(px);
  HIPCHECK(hipMemcpy(buf.data(), d_image, 16 * px, hipMemcpyDeviceToHost));
  if (!write_file(out, "_image.f32", buf.data(), 16 * px)) return 4;
  HIPCHECK(hipMemcpy(buf.data(), d_accum, 16 * px, hipMemcpyDeviceToHost));
  if (!write_file(out, "_accum.f32", buf.data(), 16 * px)) return 4;
  HIPCHECK(hipMemcpy(buf.data(), d_taa, 16 * px, hipMemcpyDeviceToHost));
  if (!write_file(out, "_taa.f32", buf.data(), 16 * px)) return 4;
  HIPCHECK(hipMemcpy(words.data(), d_frame, 4 * px, hipMemcpyDeviceToHost));
  if (!write_file(out, "_frame.u32", words.data(), 4 * px)) return 4;

  CHECK(aarbf_bvh_destroy(ctx, world));
  CHECK(aarbf_tf_destroy(ctx, cm));
  CHECK(aarbf_tf_destroy(ctx, rm));
  CHECK(aarbf_tf_destroy(ctx, dm));
  for (void *p : {(void *)d_particles, (void *)d_codes, (void *)d_aabbs, (void *)d_image, (void *)d_accum,
                  (void *)d_taa, (void *)d_taaPrev, (void *)d_frame})
    HIPCHECK(hipFree(p));
  CHECK(aarbf_context_destroy(ctx));
  aarbf_timesteps_destroy(ts);
  std::printf("dropin_snippet: %llu particles, %dx%d, %d frames\n", (unsigned long long)n, W, H, frames);
  return 0;
}

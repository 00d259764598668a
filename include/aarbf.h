/*
 * aarbf.h -- C ABI of the MI355X attribute-aware RBF renderer.
 *
 * The structs mirror the reference's shared host/device structs field by
 * field (sharedCode.h:33-117), with GPRT buffer/texture/accel handles
 * replaced by 64-bit values: device pointers for buffers, and handles
 * returned by this library for transfer-function tables, the blue-noise
 * atlas and the BVH.  Fields that have no counterpart in the reference are
 * appended after the reference fields and marked "extension".
 *
 * Each entry point replaces one GPRT call sequence of hostCode.cpp; the
 * comment above it cites the reference call it stands in for.
 *
 * Conventions:
 *   - every function returns an aarbf_status (0 = OK); the message of the
 *     last failure on a context is available from aarbf_last_error();
 *   - no C++ exception crosses this boundary;
 *   - device work is enqueued on the context's HIP stream (the library's
 *     own stream, or one supplied with aarbf_set_stream) and is
 *     asynchronous unless the function returns host data;
 *   - a context is not thread-safe: one host thread drives it, as the
 *     reference's single render loop does (hostCode.cpp:489-832).
 */
#ifndef AARBF_H
#define AARBF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AARBF_MAX_DEPTH 10000 /* sharedCode.h:27 */
#define AARBF_STBN_TILE 256   /* deviceCodeRBF.hlsl:245-248 */

typedef enum aarbf_status {
  AARBF_OK = 0,
  AARBF_ERR_INVALID_ARGUMENT = 1,
  AARBF_ERR_OUT_OF_MEMORY = 2,
  AARBF_ERR_HIP = 3,
  AARBF_ERR_IO = 4,
  AARBF_ERR_FORMAT = 5,
  AARBF_ERR_UNSUPPORTED = 6,
  AARBF_ERR_NO_DEVICE = 7
} aarbf_status;

typedef struct aarbf_context aarbf_context;
typedef struct aarbf_timesteps aarbf_timesteps;

typedef struct aarbf_float3 {
  float x, y, z;
} aarbf_float3;

typedef struct aarbf_int2 {
  int32_t x, y;
} aarbf_int2;

/* sharedCode.h:33-73 (PushConstants). */
typedef struct aarbf_push_constants {
  uint16_t accumID;
  uint16_t frameID;
  uint16_t enableBlueNoise; /* 1: stochastic marching driven by the STBN atlas */
  uint16_t showNoise;
  uint16_t visualizeAttributes;
  uint16_t particlesPerLeaf;
  struct {
    float azimuth;
    float elevation;
    float ambient;
  } light;
  struct {
    aarbf_float3 pos;
    aarbf_float3 dir_00;
    aarbf_float3 dir_du;
    aarbf_float3 dir_dv;
  } camera;
  float rbfRadius;
  float unit;
  uint64_t world;        /* BVH handle from aarbf_bvh_build */
  uint64_t particles;    /* device float4[numParticles]: x, y, z, attribute */
  uint32_t numParticles;
  uint64_t colormap;     /* transfer-function handles from aarbf_tf_create* */
  uint64_t radiusmap;
  uint64_t densitymap;
  /* extension: device float[numParticles] of per-particle radii; 0 = use
   * rbfRadius for every particle (the reference's behaviour).  The radius of
   * particle i is radii[i] * radiusmap(w_i).  rbfRadius still sets the
   * 2*rbfRadius margin of the ray clip box (deviceCodeRBF.hlsl:218-219), so
   * callers set it to at least max(radii[i] * radiusmap) / 2. */
  uint64_t radii;
} aarbf_push_constants;

/* sharedCode.h:76-86 (BoundsConstants). */
typedef struct aarbf_bounds_constants {
  uint64_t aabbs; /* device float[6 * numAABBs]: min.xyz, max.xyz per leaf */
  uint32_t numAABBs;
  uint64_t particles;
  uint32_t numParticles;
  uint16_t particlesPerLeaf;
  float rbfRadius;
  uint64_t radiusmap;
  uint64_t radii; /* extension, as in aarbf_push_constants */
} aarbf_bounds_constants;

/* sharedCode.h:88-105 (TAAConstants). */
typedef struct aarbf_taa_constants {
  uint16_t disableTAA;
  uint16_t showNoise;
  aarbf_int2 fbSize;
  uint64_t imageTexture;  /* device float4[W*H]: this frame's render */
  uint64_t guiTexture;    /* device uint32 RGBA8[W*H] overlay, or 0 for none */
  uint64_t frameBuffer;   /* device uint32 BGRA8[W*H], or 0 to skip packing */
  uint64_t taaBuffer;     /* device float4[W*H] */
  uint64_t taaPrevBuffer; /* device float4[W*H] */
} aarbf_taa_constants;

/* sharedCode.h:108-117 (RayGenData). */
typedef struct aarbf_raygen_data {
  uint64_t stbnTexture; /* STBN handle from aarbf_stbn_create, or 0 */
  uint64_t imageBuffer; /* device float4[W*H] */
  uint64_t accumBuffer; /* device float4[W*H] */
  aarbf_int2 fbSize;
  aarbf_float3 globalAABBMin;
  aarbf_float3 globalAABBMax;
} aarbf_raygen_data;

/* Counters gathered by the render kernels when statistics are enabled. */
typedef struct aarbf_render_stats {
  uint64_t samples;       /* ParticleRBFRayGen invocations */
  uint64_t steps;         /* tracker iterations (primary + shadow) */
  uint64_t queries;       /* BVH walks: per-step point queries and per-segment gathers */
  uint64_t nodes;         /* BVH node boxes tested (levels >= 1) */
  uint64_t leaves;        /* leaf boxes tested */
  uint64_t particles;     /* particle spheres tested */
  uint64_t covering;      /* particle spheres that covered the query point */
  uint64_t exhausted;     /* samples whose tracker hit AARBF_MAX_DEPTH */
} aarbf_render_stats;

/* ---- context ------------------------------------------------------------ */

/* Replaces gprtContextCreate(&GPU) (hostCode.cpp:280-281). */
int aarbf_context_create(int device, aarbf_context **out);
int aarbf_context_destroy(aarbf_context *ctx);
/* Use an existing HIP stream (hipStream_t) for all later work; NULL
 * restores the context's own stream. */
int aarbf_set_stream(aarbf_context *ctx, void *hip_stream);
void *aarbf_get_stream(aarbf_context *ctx);
int aarbf_sync(aarbf_context *ctx);
const char *aarbf_last_error(const aarbf_context *ctx);
const char *aarbf_status_string(int status);
/* Render schedule: 0 (default) wavefront rounds of per-segment candidate
 * gathering and tracker steps, 1 one lane per pixel with a BVH point query
 * per step; both produce the reference's images bit for bit.  2: delta
 * tracking with local majorants (empty space skipped, faint regions
 * stepped at their own density bound): the same expected image, different
 * random-number use, so per-sample results differ; stochastic marching is
 * rendered as in mode 0.  No counterpart in the reference. */
int aarbf_set_render_mode(aarbf_context *ctx, int mode);
/* Schedule parameters of the default renderer (none changes the image):
 * AARBF_TUNE_STEP_WAVES: occupancy target of the step kernel in waves per
 * SIMD (0 = default 4, or 4..6); AARBF_TUNE_STEP_CAP: tracker steps
 * per ray per wavefront round (16..65536, default 8192) for a call of up
 * to 8x the lanes the GPU keeps resident (4 waves per SIMD: 2,097,152
 * pixels on 256 CUs); a call with more pixels scales it by
 * (pixels / that)^(2/3), to the nearest power of two, at most 65536;
 * AARBF_TUNE_GRAPH_ROUNDS: 0 (default) = the host reads the queue lengths
 * back after every round and sizes the next round's grids; an even R in
 * 2..64 = R rounds with fixed grids are captured once into a hipGraph
 * (cached per set of arguments; kernels read their counts on the device)
 * and replayed with one read-back per replay; AARBF_TUNE_WALK_BUDGET: tree
 * nodes a one-lane gather walk may open before it hands its ray to a
 * one-wave-per-ray gather in the same round (1..65536, default 256);
 * AARBF_TUNE_DIFF_WAVES: the differentiable renderer's expected marching
 * (estimator 0) with one lane per (pixel, sample) (0) or one wave per
 * (pixel, sample), its lanes on consecutive march points (1, default;
 * ratio tracking always runs one lane per sample);
 * AARBF_TUNE_STEP_MIN_LIVE: a step-kernel wave leaves its loop once fewer
 * than this many of its 64 lanes are still stepping (their rays go on in
 * the next round, packed into full waves with the others); 0 = never,
 * else 1..64 (default 16), applied only while the step queue holds at
 * least 65,536 rays.  No counterpart in the reference. */
#define AARBF_TUNE_STEP_WAVES 1
#define AARBF_TUNE_STEP_CAP 2
#define AARBF_TUNE_GRAPH_ROUNDS 3
#define AARBF_TUNE_WALK_BUDGET 4
#define AARBF_TUNE_DIFF_WAVES 5
#define AARBF_TUNE_STEP_MIN_LIVE 6
int aarbf_set_tuning(aarbf_context *ctx, int param, int value);
/* Enable (1) or disable (0) render statistics; read them back (and reset
 * them) with aarbf_render_stats_read. */
int aarbf_set_stats(aarbf_context *ctx, int enable);
int aarbf_render_stats_read(aarbf_context *ctx, aarbf_render_stats *out, int reset);
/* Walk-cost histogram of the segment gathers since the last reset (with
 * statistics enabled): out[b] (b < 18) counts gathers that tested between
 * 2^b and 2^(b+1)-1 BVH nodes; out[18..20] = step-kernel wave iterations,
 * lanes active in them, iterations in which some lane tested candidates;
 * out[21..23] the same for the dense-segment kernel.  No counterpart in
 * the reference. */
int aarbf_render_histogram(aarbf_context *ctx, uint64_t *out, int reset);
/* Kernel timing of the default renderer: with timing enabled, HIP events on
 * the context's stream bracket each kernel of the render rounds; ms[k] and
 * launches[k] accumulate the time and launch count of the segment gathers
 * (k = 0), the step kernel (1), the dense-segment step kernel (2) and the
 * set-up / fold kernels (3).  No counterpart in the reference. */
int aarbf_set_timing(aarbf_context *ctx, int enable);
int aarbf_read_timing(aarbf_context *ctx, double *ms, uint64_t *launches, int reset);
/* Live handles on the context: counts[0] transfer functions, counts[1]
 * blue-noise atlases, counts[2] BVHs (for leak checks).  No counterpart in
 * the reference. */
int aarbf_handle_counts(const aarbf_context *ctx, uint32_t counts[3]);
/* Writes sizes and offsets of the structs above into out[0..n) and returns
 * how many there are; bindings in other languages use it to check their
 * struct mirrors.  No counterpart in the reference. */
uint32_t aarbf_abi_layout(uint32_t *out, uint32_t n);

/* ---- particle files (importParticles.h:8-32) ---------------------------- */

/* Reads the header file at `path` (one timestep name per line) and, for each
 * name, the binary file path + "_" + name: a little-endian uint64 count
 * followed by count float4 (x, y, z, attribute).  Unlike the reference, a
 * missing or truncated timestep file is an error. */
int aarbf_import_particles(const char *path, aarbf_timesteps **out);
uint32_t aarbf_timesteps_count(const aarbf_timesteps *ts);
int aarbf_timesteps_get(const aarbf_timesteps *ts, uint32_t step, const float **xyzw,
                        uint64_t *count, const char **name);
void aarbf_timesteps_destroy(aarbf_timesteps *ts);
/* The device load path of the same format: timestep `step` of the header
 * file `path` read straight into device float4 particles[capacity], in
 * chunks through pinned staging buffers whose copies to the device overlap
 * the next read (the reference reads into host vectors, importParticles.h
 * :17-29, then copies through a mapped buffer, hostCode.cpp:759-763).
 * *count receives the timestep's particle count, also when it exceeds
 * capacity (then AARBF_ERR_INVALID_ARGUMENT and nothing is copied).
 * Synchronises the context's stream. */
int aarbf_import_step_device(aarbf_context *ctx, const char *path, uint32_t step, uint64_t particles,
                             uint64_t capacity, uint64_t *count);
/* Writes the same format (header plus one binary file per step). */
int aarbf_export_particles(const char *path, uint32_t num_steps, const char *const *names,
                           const float *const *xyzw, const uint64_t *counts);

/* ---- pre-process (hostCode.cpp:183-250), on the device ------------------ */

/* Folds the xyz bounds of n device float4 particles into aabb[6] (min.xyz,
 * max.xyz, host memory, updated in place; start from {1e38,...,-1e38,...}
 * as hostCode.cpp:183 does).  Synchronises. */
int aarbf_spacetime_aabb(aarbf_context *ctx, uint64_t particles, uint64_t n, float aabb[6]);
/* 48-bit Hilbert codes of the particle positions quantised to 16 bits per
 * axis inside aabb (hostCode.cpp:196-209 + hilbert.h:181-225), written to
 * device uint64[n]. */
int aarbf_hilbert_codes(aarbf_context *ctx, uint64_t particles, uint64_t n, const float aabb[6],
                        uint64_t codes);
/* Sorts particles (and codes) in place by (code, x, y, z, w)
 * (hostCode.cpp:212-221).  perm, if non-zero, receives device uint32[n]:
 * the original index of each sorted particle. */
int aarbf_sort_particles(aarbf_context *ctx, uint64_t particles, uint64_t codes, uint64_t n,
                         uint64_t perm);
/* Folds min/max of the attribute (w) into range[2] (host, in place; start
 * from {1e20, -1e20} as hostCode.cpp:227-228).  Synchronises. */
int aarbf_attribute_range(aarbf_context *ctx, uint64_t particles, uint64_t n, float range[2]);
/* w <- (w - min) / (max - min) when max > min (hostCode.cpp:243-250). */
int aarbf_normalize_attributes(aarbf_context *ctx, uint64_t particles, uint64_t n,
                               const float range[2]);

/* ---- transfer functions and blue noise ---------------------------------- */

/* An RGBA float table of n texels, sampled like a linearly filtered,
 * clamp-to-edge 1D texture (the reference's 64-texel TF textures,
 * hostCode.cpp:355-365).  rgba is host memory, 4*n floats. */
int aarbf_tf_create(aarbf_context *ctx, const float *rgba, uint32_t n, uint64_t *handle);
/* The same from RGBA8 texels; srgb != 0 decodes RGB with the sRGB transfer
 * function as an R8G8B8A8_SRGB texture does (alpha stays linear). */
int aarbf_tf_create_rgba8(aarbf_context *ctx, const uint8_t *rgba8, uint32_t n, int srgb,
                          uint64_t *handle);
int aarbf_tf_destroy(aarbf_context *ctx, uint64_t handle);
/* The reference's transfer-function editing (hostCode.cpp:416-436,
 * 524-572): a gradient of marks sampled at the 64 positions i / 63 and
 * written as RGBA8 bytes.  The gradient is linear in RGBA between marks
 * (sorted by position, each in [0, 1]) and takes the first / last mark's
 * value outside them.  AARBF_TF_COLORMAP writes make_8bit(pow(c, 1 / 2.2))
 * for RGB and make_8bit(alpha); AARBF_TF_SCALAR (radius and density maps)
 * writes make_8bit(red) and zeros; make_8bit(f) = clamp(int(f * 256), 0,
 * 255).  out: 64 * 4 bytes.  Host only. */
typedef struct aarbf_tf_mark {
  float position;
  float rgba[4];
} aarbf_tf_mark;
#define AARBF_TF_TEXELS 64
#define AARBF_TF_COLORMAP 0
#define AARBF_TF_SCALAR 1
int aarbf_tf_texels_from_marks(const aarbf_tf_mark *marks, uint32_t n, int kind, uint8_t *out);
/* aarbf_tf_texels_from_marks followed by aarbf_tf_create_rgba8(srgb = 1):
 * the table the reference's shaders sample after an edit of its gradient
 * widget. */
int aarbf_tf_create_from_marks(aarbf_context *ctx, const aarbf_tf_mark *marks, uint32_t n, int kind,
                               uint64_t *handle);
/* The blue-noise atlas: RGBA8 texels, width x height (the reference's
 * stbn.png is an 8x16 grid of 256x256 tiles, hostCode.cpp:336-352). */
int aarbf_stbn_create(aarbf_context *ctx, const uint8_t *rgba8, int width, int height,
                      uint64_t *handle);
int aarbf_stbn_destroy(aarbf_context *ctx, uint64_t handle);

/* ---- bounds and BVH ----------------------------------------------------- */

/* GenRBFBounds (deviceCodeBounds.hlsl:29-82), replacing
 * gprtComputeLaunch1D(context, GenRBFBounds, ...) (hostCode.cpp:770-779). */
int aarbf_gen_bounds(aarbf_context *ctx, const aarbf_bounds_constants *bc);

/* Builds a BVH over numAABBs leaf boxes (device float[6*numAABBs], the
 * layout written by aarbf_gen_bounds; NaN boxes are inactive).  Replaces
 * gprtAccelBuild(BLAS) + gprtAccelBuild(TLAS) + gprtAccelGetHandle
 * (hostCode.cpp:468-469,781-789).  `branching` is the node width (2..16; 0
 * picks the default).  The handle goes into aarbf_push_constants.world. */
int aarbf_bvh_build(aarbf_context *ctx, uint64_t aabbs, uint32_t numAABBs, uint32_t branching,
                    uint64_t *handle);
/* Recomputes node boxes for new leaf boxes with the same topology
 * (gprtAccelUpdate, hostCode.cpp:790-794). */
int aarbf_bvh_refit(aarbf_context *ctx, uint64_t handle, uint64_t aabbs);
int aarbf_bvh_destroy(aarbf_context *ctx, uint64_t handle);
/* Topology queries for tests: number of levels (leaves are level 0) and
 * the node count of each; node boxes of one level copied to host float[8*n]
 * (lo.xyz, pad, hi.xyz, pad). */
int aarbf_bvh_levels(aarbf_context *ctx, uint64_t handle, uint32_t *num_levels,
                     uint32_t *counts, uint32_t max_levels);
int aarbf_bvh_read_level(aarbf_context *ctx, uint64_t handle, uint32_t level, float *boxes);
/* Quantised copy of level `level` (>= 1, branching <= 8): 80 bytes per node
 * (20 uint32): lo.xyz and step.x as floats, step.y, step.z, 2 unused, then
 * 48 bytes of child bounds -- lo.x[8], lo.y[8], lo.z[8], hi.x[8], hi.y[8],
 * hi.z[8] -- dequantised as lo + q * step; lo.x > hi.x marks no child. */
int aarbf_bvh_read_groups(aarbf_context *ctx, uint64_t handle, uint32_t level, uint32_t *groups);

/* ---- rendering ---------------------------------------------------------- */

/* One launch of ParticleRBFRayGen (deviceCodeRBF.hlsl:206-336) over a W x H
 * image: replaces gprtRayGenSetParameters + gprtRayGenLaunch2D
 * (hostCode.cpp:370-378,813).  Delta tracking, or stochastic marching when
 * pc->enableBlueNoise is set (then rg->stbnTexture is required). */
int aarbf_render_rbf(aarbf_context *ctx, const aarbf_raygen_data *rg,
                     const aarbf_push_constants *pc, int width, int height);
/* spp consecutive launches fused into one: launch s uses frameID + s and
 * accumID + s, exactly as spp calls of aarbf_render_rbf with the counters
 * incremented between them (hostCode.cpp:826-827). */
int aarbf_render_rbf_spp(aarbf_context *ctx, const aarbf_raygen_data *rg,
                         const aarbf_push_constants *pc, int width, int height, uint32_t spp);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc: aarbf_render_rbf_spp on the one-lane-per-pixel
 * schedule (render mode 1; the same image bit for bit), allocating nothing.
 * Its only scratch buffer, the particle payload baked from the particles
 * and transfer functions (36 bytes per particle), lives in a static array
 * of each device (AARBF_NOALLOC_MAX_PARTICLES particles, 36 MiB, held until
 * the process ends), set up when a context is created on the device; calls
 * with more particles are rejected.  All calls on one device share that
 * array: calls that overlap in time on one device (several contexts,
 * threads or streams) race and must be serialised by the caller.  The
 * payload format of aarbf_set_payload applies; aarbf_payload_read does not
 * see this call's payload. */
#define AARBF_NOALLOC_MAX_PARTICLES 1048576
int aarbf_render_rbf_spp_noalloc(aarbf_context *ctx, const aarbf_raygen_data *rg,
                                 const aarbf_push_constants *pc, int width, int height, uint32_t spp);
/* Particle payload the renderer bakes from the particles and TFs: 0 (default)
 * FP32, 32 bytes per particle; 1 FP16 leaf-local, 16 bytes per particle --
 * positions and radius as 16-bit fixed point in the frame of their chunk of
 * 32 particles (origin and per-axis step computed in double, stored as float;
 * the radius shrunk by the position's rounding shift, so each decoded sphere
 * stays inside the original one, and coded over the chunk's radius range),
 * alpha as 16-bit unorm, rgb^2.2 as binary16.  FP16 renders a slightly
 * different scene; accumulation stays FP32.  No counterpart in the
 * reference. */
int aarbf_set_payload(aarbf_context *ctx, int format);
/* The decoded payload of the last render (x, y, z, r, alpha^3, r^2.2, g^2.2,
 * b^2.2 per particle) into device float out[8 * n]. */
int aarbf_payload_read(aarbf_context *ctx, uint64_t out, uint32_t n);
/* ---- batched rendering (no counterpart in the reference) -------------------
 * One view of a batch: what one aarbf_render_rbf_spp call would take. */
typedef struct aarbf_request {
  aarbf_raygen_data rg;    /* image / accum buffers, STBN atlas, global AABB */
  aarbf_push_constants pc; /* camera, frameID, accumID, enableBlueNoise, light, ... */
  int width, height;
  uint32_t spp;
} aarbf_request;
/* Renders n views of one scene in one wavefront schedule: the views' pixels
 * share every gather / step round, so small views do not leave the GPU idle.
 * Each view's image equals that of aarbf_render_rbf_spp(ctx, &rg, &pc, width,
 * height, spp) bit for bit.  Views may differ in size, spp, camera, frame /
 * accum counters, light, showNoise and enableBlueNoise (stochastic marching
 * or delta tracking, mixed in one batch); world, particles, numParticles,
 * particlesPerLeaf, radii, transfer functions, rbfRadius, unit,
 * visualizeAttributes and the global AABB must match request 0.  Render
 * mode 1 does not apply (the wavefront schedule is used); mode 2 applies
 * when no view uses marching.  Records each view's latency. */
int aarbf_render_batch(aarbf_context *ctx, const aarbf_request *reqs, uint32_t n);
/* Per view of the last aarbf_render_batch (n = its view count): ms from the
 * start of the batch's GPU work to the end of the view's last sample (device
 * wall clock), or -1 if it did not finish.  Synchronises the stream. */
int aarbf_batch_latency(aarbf_context *ctx, double *ms, uint32_t n);
/* Schedule autotuner: renders the views (into its own buffers; the
 * callers' images are untouched) under candidate schedule parameters --
 * AARBF_TUNE_STEP_CAP over {2048, 4096, 8192, 16384}, AARBF_TUNE_STEP_WAVES
 * over {4, 5, 6}, AARBF_TUNE_GRAPH_ROUNDS over {0, 8},
 * AARBF_TUNE_WALK_BUDGET over {256, 1024, 65536},
 * AARBF_TUNE_STEP_MIN_LIVE over {0, 16}, by coordinate descent
 * from the current setting, each timed as the best of `reps` runs after a
 * warm-up -- and keeps the fastest in the context.  The BVH leaf size is a
 * scene property, tuned by rebuilding the scene per leaf size around this
 * call.  No counterpart in the reference. */
typedef struct aarbf_tune_result {
  int32_t step_cap, step_waves, graph_rounds;
  uint32_t candidates;  /* settings timed */
  double baseline_ms;   /* the setting at entry */
  double best_ms;       /* the setting kept */
  int32_t walk_budget;
  int32_t step_min_live; /* AARBF_TUNE_STEP_MIN_LIVE kept */
} aarbf_tune_result;
int aarbf_autotune(aarbf_context *ctx, const aarbf_request *reqs, uint32_t n, uint32_t reps,
                   aarbf_tune_result *out);
/* ParticleSplatRayGen (deviceCodeSplat.hlsl:58-154).  Rays are clipped to
 * the global AABB grown by one rbfRadius, as in the reference
 * (deviceCodeSplat.hlsl:66-67); with per-particle radii set rbfRadius to at
 * least the largest radii[i] * radiusmap, or samples of spheres reaching
 * past that box are clipped. */
int aarbf_render_splat(aarbf_context *ctx, const aarbf_raygen_data *rg,
                       const aarbf_push_constants *pc, int width, int height);
/* ---- differentiable rendering (no counterpart in the reference) -------------
 * The expectation of stochastic marching (deviceCodeRBF.hlsl:67-139, primary
 * ray, ambient 1) over its blue-noise threshold: each sample marches
 * t_k = tenter + unit * (r1 + k), r1 its first LCG draw, and the pixel is the
 * mean over spp samples of sum_k (F(L_k) - F(L_{k-1})) c_k with L_k the
 * accumulated densitymap values, F(L) = min(1, (1 - e^-L) / 0.99) and c_k the
 * attribute-weighted colour.  Needs per-particle radii (pc->radii) and
 * visualizeAttributes.  image: device float4 [height][width] (rgb, 1).
 * path: 0, or a device float4 [height][width][spp] buffer that receives one
 * vector per sample for aarbf_diff_backward (it then replays each sample
 * once instead of twice). */
int aarbf_diff_forward(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc,
                       int width, int height, uint32_t spp, uint64_t image, uint64_t path);
/* Adds dLoss/dparameters to the gradient buffers (device float, zeroed by
 * the caller): g_pos [3N] (x, y, z per particle), g_radii [N], g_attr [N],
 * g_colormap [4 * colormap texels], g_densitymap / g_radiusmap [texels]
 * (the red channel they are read from).  grad_image: device float4
 * [height][width], rgb used; path: the forward's path buffer (same scene
 * and view) or 0.  Stores nothing per step. */
int aarbf_diff_backward(aarbf_context *ctx, const aarbf_raygen_data *rg, const aarbf_push_constants *pc,
                        int width, int height, uint32_t spp, uint64_t path, uint64_t grad_image,
                        uint64_t g_pos, uint64_t g_radii, uint64_t g_attr, uint64_t g_colormap,
                        uint64_t g_densitymap, uint64_t g_radiusmap);
/* Estimator of the differentiable renderer (default 0): 0 the expectation
 * of stochastic marching over its threshold, as described above; 1 ratio
 * tracking, the expectation of delta tracking (deviceCodeRBF.hlsl:141-203,
 * primary ray, ambient 1) over its accept draws: tentative points at the
 * tracker's free flights t_i = t_{i-1} - ln(1 - xi_i) unit from t = 0 (one
 * LCG draw each after the sample's three set-up draws), pixel = mean over
 * samples of sum_i T_{i-1} D_i c_i, T_i = T_{i-1} (1 - D_i), each sample
 * ending after the first point with T_i < 2^-20.  path then holds each
 * sample's colour.  No counterpart in the reference. */
int aarbf_set_diff_estimator(aarbf_context *ctx, int estimator);
/* CompositeGui's TAA resolve and BGRA8 packing (deviceCodeCommon.hlsl:67-115)
 * plus the taa -> taaPrev history copy (hostCode.cpp:819-830).  The GUI
 * layer (guiTexture, R8G8B8A8_SRGB texels, hostCode.cpp:328-330) is decoded
 * with the sRGB transfer function and composited with over() on top of the
 * pixel; frameBuffer receives B | G << 8 | R << 16 | A << 24, each channel
 * clamped to [0, 1] and rounded as (uint)(v * 255 + 0.5). */
int aarbf_taa_resolve(aarbf_context *ctx, const aarbf_taa_constants *tc);

/* ---- host helpers of the reference's render loop (no device work) ------- */

/* getLightDirection (sharedCode.h:127-132) evaluated on the host in
 * binary32, as the renderer uses it: out = (sin(a) cos(b), sin(a) sin(b),
 * cos(a)), a = elevation * 3.14 - 0.5 * 3.14, b = azimuth * 2 * 3.14. */
int aarbf_light_direction(float azimuth, float elevation, float out[3]);
/* The camera basis the reference writes into rtConstants.camera
 * (hostCode.cpp:669-682, 712-719): pos = look_from,
 * d00 = normalize(look_at - look_from), du = cos_fovy * (width / height) *
 * normalize(cross(d00, look_up)), dv = cos_fovy * normalize(cross(du, d00)),
 * then d00 -= du / 2 and d00 -= dv / 2.  cos_fovy is the reference's
 * image-plane scale (2 tan(fovy / 2) for a vertical field of view fovy).
 * Writes pc->camera only. */
int aarbf_camera_basis(const float look_from[3], const float look_at[3], const float look_up[3],
                       float cos_fovy, int width, int height, aarbf_push_constants *pc);
/* The reference's interactive camera (hostCode.cpp:253-271 state,
 * 593-731 per frame). */
typedef struct aarbf_camera_state {
  aarbf_float3 lookFrom, lookAt, lookUp;
  float cosFovy;
  int32_t firstFrame; /* set to 1 before the first frame */
} aarbf_camera_state;
typedef struct aarbf_camera_input {
  float dx, dy;                 /* cursor movement since the last frame, pixels */
  int32_t left, right, middle;  /* mouse buttons held and not captured by a GUI */
  int32_t key_x, key_y, key_z;  /* up-axis keys */
  int32_t left_shift;           /* negates the up axis a key selects */
} aarbf_camera_input;
/* One frame of the controller: an X / Y / Z key sets the up axis (negated
 * with left shift); the left button (or an up-axis key, or the first frame)
 * orbits look_from about look_at by -dx * 2 pi / width around the up axis
 * and -dy * pi / height around cross(up, normalize(look_at - position));
 * the right button zooms (look_from - look_at scaled by 0.95 for dy > 0,
 * 1.05 for dy < 0; only pc->camera.pos changes); the middle button pans
 * look_from and look_at by (right * dx - up * dy) * 0.001 * diagonal.  The
 * camera basis in pc is recomputed as the reference recomputes it, and
 * pc->accumID is reset to 1 when the camera moved; *changed (may be NULL)
 * tells whether it did. */
int aarbf_camera_update(aarbf_camera_state *st, const aarbf_camera_input *in, int width, int height,
                        float diagonal, aarbf_push_constants *pc, int *changed);

#ifdef __cplusplus
}
#endif

#endif /* AARBF_H */

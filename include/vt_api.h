/*
 * vt_api.h - C ABI of the MI355X voxel path tracer.
 *
 * This is the boundary that replaces the GLSL program "ABI" of voxelToy:
 * the uniforms and textures bound by Renderer (src/shaders/shader.h:43-81,
 * src/renderer/glResources.h:9-24, src/renderer/renderer.cpp:215-312) and
 * the draw calls that launch pathTracer.fs, accumulation.fs and
 * voxelize.{vs,gs} (src/renderer/renderer.cpp:584-611,
 * src/voxelize/gpuVoxelizer.cpp:46-72).  Plain C types only; every
 * function returns VT_OK or a negative status and leaves a message for
 * vt_last_error().
 *
 * Conventions:
 *   - a context owns all device memory and one HIP stream on one device;
 *     host arrays passed in are copied (glTexImage* semantics);
 *   - calls on one context are not reentrant; use one context per host
 *     thread;
 *   - images are in GL row order: row 0 is the bottom row;
 *   - matrices are row-major float[16] (m[r*4+c]), as Imath stores them
 *     and as the reference uploads them with transpose = GL_TRUE.
 */
#ifndef VT_API_H_
#define VT_API_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; vt_last_error() holds the message of the calling thread's
 * last failure.  No C++ exception leaves the library: a host allocation that
 * cannot be made returns VT_ERR_LIMIT ("out of memory"), any other internal
 * exception VT_ERR_STATE. */
#define VT_OK 0
#define VT_ERR_ARG -1
#define VT_ERR_HIP -2
#define VT_ERR_STATE -3
#define VT_ERR_LIMIT -4

typedef struct vt_context vt_context;

/* Camera uniforms (pathTracer.fs:14-23, generateRay.h).  lens_model:
 * 0 pinhole, 1 thin lens, 2 orthographic (CameraParameters::CameraLensModel,
 * src/camera/cameraParameters.h:23-28).  focal_distance is
 * FocalDistanceData.focalDistance (focalDistanceDevice.h:5-8). */
typedef struct vt_camera {
    float inv_modelview[16];
    float proj[16];
    float inv_proj[16];
    float viewport[4];      /* x, y, w, h */
    float win_z;            /* gl_FragCoord.z of the full-screen quad */
    float lens_radius;
    float focal_distance;
    int32_t lens_model;
} vt_camera;

/* Environment uniforms (pathTracer.fs:25-33).  With use_image != 0 the RGB
 * float image (row 0 = top of the equirect map, v = 0) is sampled
 * bilinearly with GL_CLAMP borders, and cdf_u ((w'+1) x h'), cdf_v (h'+1)
 * and integral come from vt_env_build_cdf or the caller. */
typedef struct vt_environment {
    int32_t use_image;
    float top[3];
    float bottom[3];
    int32_t image_w, image_h;
    const float* image_rgb;
    int32_t cdf_u_w, cdf_u_h;
    const float* cdf_u;
    int32_t cdf_v_n;
    const float* cdf_v;
    float integral;
    float rotation_radians;
} vt_environment;

/* Integrator settings (RenderSettings src/renderer/renderSettings.h:6-29
 * and the pathTracer.fs uniforms :37-44, SelectVoxelData). */
typedef struct vt_settings {
    int32_t max_bounces;
    float exposure;
    float gamma;
    int32_t selected_voxel[3];   /* out of range = no selection */
    float wireframe_opacity;
    float wireframe_thickness;
} vt_settings;

/* Work counters of the last vt_render (filled when stats are enabled). */
typedef struct vt_stats {
    uint64_t rays;          /* camera samples + shadow + continuation traversals */
    uint64_t visits;        /* voxel fetches made by the DDA */
    uint64_t interactions;  /* surface shading events */
    uint64_t samples;
    uint64_t env_samples;   /* importance-sampled image-environment draws */
    uint64_t env_evals;     /* image-environment lookups by direction */
    double kernel_ms;       /* HIP-event time of the last render launch */
} vt_stats;

/* ---- context ------------------------------------------------------------ */
const char* vt_last_error(void);
int vt_create(int device, vt_context** out);
void vt_destroy(vt_context* ctx);
int vt_synchronize(vt_context* ctx);
/* Number of compute units of the context's device. */
int vt_device_cus(vt_context* ctx);

/* ---- volume -------------------------------------------------------------
 * Replaces Renderer::createVoxelDataTexture (src/renderer/renderer.cpp:833-940):
 * the GL_R32I material-offset grid (x fastest, -1 = empty), the GL_R32F
 * material records and the GL_R32I emissive-voxel list.  Bounds follow
 * the reference (cubic voxels, longest axis 1000 units, centred) unless
 * bmin/bmax are given (non-NULL).  At most 255 distinct occupied offsets;
 * offsets must be -1 or lie in [0, mat_size). */
int vt_set_volume(vt_context* ctx, int32_t nx, int32_t ny, int32_t nz, const int32_t* material_offsets,
                  const float* mat_data, int32_t mat_size, const int32_t* emissive, int32_t n_emissive,
                  const float* bmin, const float* bmax);
/* Same volume from a uint8 ID grid: id_offset[256] gives the material
 * offset of each ID (-1 = empty). */
int vt_set_volume_u8(vt_context* ctx, int32_t nx, int32_t ny, int32_t nz, const uint8_t* ids,
                     const int32_t* id_offset, const float* mat_data, int32_t mat_size,
                     const int32_t* emissive, int32_t n_emissive, const float* bmin, const float* bmax);
/* Volume bounds and voxel size actually in use (6 + 3 floats). */
int vt_get_bounds(vt_context* ctx, float* bmin, float* bmax, float* voxel_size);
/* Dense readback of the grid as material offsets (-1 empty), x fastest. */
int vt_read_volume(vt_context* ctx, int32_t* material_offsets);
/* Material record edits: Renderer::updateMaterialColor/Value
 * (src/renderer/renderer.cpp:1205-1235). */
int vt_update_material(vt_context* ctx, int32_t offset, const float* values, int32_t count);
int vt_read_materials(vt_context* ctx, float* mat_data, int32_t capacity, int32_t* size);

/* ---- render state ------------------------------------------------------- */
int vt_set_camera(vt_context* ctx, const vt_camera* cam);
int vt_set_environment(vt_context* ctx, const vt_environment* env);
int vt_set_settings(vt_context* ctx, const vt_settings* st);
/* Allocate the W x H accumulation target (sampleTexture/averageTexture,
 * src/renderer/renderer.cpp:763-795) and clear it.  fp64 != 0 also keeps
 * double-precision running sums of every sample. */
int vt_set_frame(vt_context* ctx, int32_t width, int32_t height, int32_t fp64);
int vt_reset_accum(vt_context* ctx);
int vt_set_stats(vt_context* ctx, int32_t enabled);
/* Kernel organisation of vt_render (results are identical):
 *   0 = persistent megakernel (one lane renders a pixel's passes start to end)
 *   1 = wavefront (generate / traverse / shade kernels over ray queues; at
 *       most max_paths paths in flight, 0 = default 16M).
 * Default 1. */
int vt_set_render_mode(vt_context* ctx, int32_t mode, uint64_t max_paths);
/* Wavefront traversal tuning (results are identical for every value):
 * dda_budget   DDA steps a lane takes between refill checks (1..4096),
 * refill_min   idle lanes of a 64-lane wave that trigger a refill (1..64),
 * refill_chunk queue entries a wave reserves per atomic (64..65536, or -1
 *              for exactly the idle lanes);
 * 0 selects the default of each.  No counterpart in the reference (its
 * traversal is one fragment-shader loop). */
int vt_set_traversal_tuning(vt_context* ctx, int32_t dda_budget, int32_t refill_min, int32_t refill_chunk);

/* ---- rendering ----------------------------------------------------------
 * Render `count` progressive passes: pass k uses accumulation index
 * n = n0 + k (accumulation.fs: avg = (s + avg*n)/(n+1)) and RNG frame
 * min(n, rng_frame_max) (src/renderer/renderer.cpp:594).  Equivalent to
 * `count` calls of Renderer::render(). */
int vt_render(vt_context* ctx, uint64_t seed, uint32_t n0, uint32_t count, uint32_t rng_frame_max);
/* RGBA32F running mean, W*H*4 floats, row 0 = bottom. */
int vt_read_accum(vt_context* ctx, float* rgba);
/* FP64 running sums (W*H*3 doubles) when vt_set_frame(fp64=1). */
int vt_read_sum64(vt_context* ctx, double* rgb);
int vt_get_stats(vt_context* ctx, vt_stats* out);
/* Resume of a progressive render: write back a running mean (and FP64 sums)
 * read earlier with vt_read_accum / vt_read_sum64 (same layouts, the frame
 * of the same size), so that vt_render(ctx, seed, n, ...) continues the
 * interrupted render exactly.  No counterpart in the reference (its
 * accumulation lives in GL textures). */
int vt_write_accum(vt_context* ctx, const float* rgba);
int vt_write_sum64(vt_context* ctx, const double* rgb);
/* Wavefront renders of at least 2 passes and 4M paths run their first and
 * second half of the passes on two streams of the device (1 = one stream).
 * Images are identical (the halves accumulate in pass order); the second
 * half's paths take their own path buffers, so the in-flight capacity
 * (vt_set_render_mode max_paths) is split between the halves.  Default 1;
 * no counterpart in the reference. */
int vt_set_overlap(vt_context* ctx, int32_t streams);
/* Per-launch timing of the traversal kernel: when enabled, each wavefront
 * vt_render brackets every traversal launch (up to 256 per call) with HIP
 * events on the stream it runs on; vt_get_launch_timing returns the summed
 * durations (ms), the number of timed launches of the last vt_render and,
 * when union_ms is non-null, the time during which at least one of them
 * ran (less than the sum when two streams overlap).  Measurement only (no
 * counterpart in the reference). */
int vt_set_launch_timing(vt_context* ctx, int32_t enabled);
int vt_get_launch_timing(vt_context* ctx, double* traverse_ms, uint32_t* launches, double* union_ms);

/* ---- batches -------------------------------------------------------------
 * Many scenes (contexts on one device, each with its own volume, camera,
 * environment, settings and frame) rendered together: scenes of at least
 * 2^22 paths (width x height x count; vt_batch_set_tuning) get their own
 * wavefront launches on their own context streams, concurrently; the others
 * share one wavefront launch whose queues hold the paths of all of them.
 * With a graph, every launch of the batch (the solo scenes' branches and the
 * merged launch) is recorded once as one HIP graph and replayed.  Each context's accumulation afterwards equals what
 * vt_render(ctx, seeds[i], n0[i], counts[i], rng_frame_max) gives.  The
 * traversal tuning of ctxs[0] applies to the shared launch, each scene's
 * own tuning to its own launches, and the path capacity (vt_set_render_mode
 * max_paths) of each launch's context, unless vt_batch_set_tuning or
 * vt_autotune_batch set them for the batch.  No counterpart in the reference
 * (one scene per GL context); the contexts must outlive the batch. */
typedef struct vt_batch vt_batch;
int vt_batch_create(vt_context* const* ctxs, int32_t n, vt_batch** out);
/* flags: bit 0 = launch through a HIP graph (recorded on first use and
 * re-recorded when the batch's shape changes), bit 1 = collect stats. */
int vt_batch_render(vt_batch* b, const uint64_t* seeds, const uint32_t* n0, const uint32_t* counts,
                    uint32_t rng_frame_max, int32_t flags);
/* Counters of the last batch render (summed over scenes) and its time. */
int vt_batch_get_stats(vt_batch* b, vt_stats* out);
void vt_batch_destroy(vt_batch* b);
/* Tuning of the batch's launches (results are identical for every value):
 * solo_paths   paths from which a scene gets its own launches (0 = 2^22);
 * max_paths    paths in flight per launch, merged and solo (0 = each
 *              launch's context's vt_set_render_mode capacity);
 * dda_budget, refill_min, refill_chunk  the merged launch's traversal
 *              tuning as in vt_set_traversal_tuning (0 = the first
 *              context's).  No counterpart in the reference. */
typedef struct vt_batch_tuning {
    uint64_t solo_paths, max_paths;
    int32_t dda_budget, refill_min, refill_chunk;
    double ms;              /* vt_autotune_batch: time of the chosen candidate */
    int32_t candidates;     /* vt_autotune_batch: number timed */
} vt_batch_tuning;
int vt_batch_set_tuning(vt_batch* b, uint64_t solo_paths, uint64_t max_paths, int32_t dda_budget, int32_t refill_min,
                        int32_t refill_chunk);
/* The tuning in effect, defaults resolved. */
int vt_batch_get_tuning(vt_batch* b, vt_batch_tuning* out);
/* Times graph replays of vt_batch_render(b, seeds, n0, counts,
 * rng_frame_max) over the solo threshold, the path capacity and the merged
 * launch's traversal tuning (coordinate descent, best of two replays per
 * candidate), keeps the fastest and reports it.  Every candidate renders
 * identical images; every scene's accumulation is restored afterwards. */
int vt_autotune_batch(vt_batch* b, const uint64_t* seeds, const uint32_t* n0, const uint32_t* counts,
                      uint32_t rng_frame_max, vt_batch_tuning* out);

/* ---- autotuning ------------------------------------------------------------
 * Times `passes`-pass renders of the context's current scene under each
 * kernel organisation and traversal tuning (both render modes, step budget,
 * refill threshold, queue chunk), applies the fastest and reports it.  Every
 * candidate renders the identical image; the accumulation is restored
 * afterwards.  No counterpart in the reference. */
typedef struct vt_tuning {
    int32_t render_mode;
    int32_t dda_budget, refill_min, refill_chunk;
    uint64_t max_paths;
    double ms;              /* time of the chosen candidate */
    int32_t candidates;     /* number timed */
} vt_tuning;
int vt_autotune(vt_context* ctx, uint32_t passes, vt_tuning* out);

/* ---- learned radiance cache ---------------------------------------------------
 * A 288-64-64-64-64-64-3 ReLU MLP over a 12-band sin/cos encoding of a
 * surface query (12 raw floats: position normalised to the volume bounds,
 * (direction+1)/2, (normal+1)/2, albedo) predicting outgoing RGB radiance.
 * Inference runs FP16 matrix-core tiles with FP32 accumulation; training
 * keeps FP32 master weights: forward as the inference, backward and weight
 * gradients on FP32 matrix-core tiles (relative L2 loss, Adam).  Parameter
 * vector layout: W1[64][288], b1,
 * W2..W5[64][64], b2..b5, W6[3][64], b6.  No counterpart in the reference. */
typedef struct vt_cache vt_cache;
int vt_cache_create(vt_context* ctx, uint64_t seed, vt_cache** out);
void vt_cache_destroy(vt_cache* cache);
int32_t vt_cache_num_params(void);
int vt_cache_get_params(vt_cache* cache, float* params);
int vt_cache_set_params(vt_cache* cache, const float* params);
/* n queries (raw, n x 12) -> radiance (n x 3), host arrays. */
int vt_cache_infer(vt_cache* cache, int64_t n, const float* raw, float* out);
/* Gradient of the loss over n samples (targets n x 3) without updating. */
int vt_cache_gradient(vt_cache* cache, int64_t n, const float* raw, const float* target, float* grad, float* loss);
/* One Adam step (lr; beta1 0.9, beta2 0.99, eps 1e-8) on n samples. */
int vt_cache_train(vt_cache* cache, int64_t n, const float* raw, const float* target, float lr, float* loss);

/* Render `count` passes (pass n0+k, one sample per pixel each) with the
 * cache: paths end into the cache at their vertex of bounce cache_bounce
 * (-1: never), except about train_paths training paths per pass, which
 * trace train_extra more bounces (-1: to the end, as the plain integrator)
 * before ending into the cache themselves, and give one sample per shaded
 * vertex before that end (target = radiance gathered after the vertex,
 * including the cache's prediction at the end, / throughput at the vertex);
 * after each pass the cache takes Adam steps (learning rate lr) over those
 * samples in minibatches of at most `batch`.  Statistics of the last call
 * in *out (may be NULL). */
typedef struct vt_cache_stats {
    uint64_t queries;       /* paths that ended into the cache */
    uint64_t samples;       /* training samples used */
    uint32_t steps;         /* Adam steps taken */
    float last_loss;        /* loss of the last step */
    double ms;              /* wall time of the call */
} vt_cache_stats;
int vt_render_cached(vt_context* ctx, vt_cache* cache, uint64_t seed, uint32_t n0, uint32_t count,
                     int32_t cache_bounce, uint32_t train_paths, int32_t train_extra, float lr, uint32_t batch,
                     vt_cache_stats* out);

/* ---- probes -------------------------------------------------------------
 * First hit of every pixel's primary ray of frame `frame`: rec[8] per
 * pixel = kind (0 AABB miss, 1 traversal miss, 2 voxel, 3 ground), voxel
 * x, y, z, DDA steps (-1 if the march started outside the grid), face
 * normal nx, ny, nz. */
int vt_trace_first_hit(vt_context* ctx, uint64_t seed, uint32_t frame, int32_t* rec);
/* Injected rays (origin[3], dir[3] each), same record. */
int vt_trace_rays(vt_context* ctx, int32_t n, const float* origins, const float* dirs, int32_t* rec);

/* Allocation-free variant of vt_trace_rays.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (picking on every
 * mouse move in an interactive loop, serving loops and harnesses that forbid
 * hipMalloc, code around stream capture).  Same arguments, records and errors
 * as vt_trace_rays for n <= VT_NOALLOC_MAX_RAYS (VT_ERR_LIMIT beyond, nothing
 * launched).  Rays and records are staged in a fixed device array of
 * VT_NOALLOC_MAX_RAYS x 56 bytes (896 KiB), one per device, that the library
 * holds until the process ends; calls that would overlap on one device take
 * turns on it (each returns after its records are on the host). */
#define VT_NOALLOC_MAX_RAYS 16384
int vt_trace_rays_noalloc(vt_context* ctx, int32_t n, const float* origins, const float* dirs, int32_t* rec);

/* ---- voxelizer ----------------------------------------------------------
 * GPUVoxelizer::voxelizeMesh (src/voxelize/gpuVoxelizer.cpp:46-72 with
 * voxelize.vs/gs).  Vertices are float xyz, triangles uint32 index
 * triples; model_transform is the row-major mesh transform
 * (computeMeshTransform, src/renderer/import.cpp:46-64).  mode 0 = THIN
 * (reference default), 1 = FAT.  The grid must already exist
 * (vt_set_volume); covered voxels are set to material offset
 * `material_offset` (must be a known offset).  If occupancy != NULL the
 * dense byte occupancy (1 = covered) of this call alone is also written
 * there.  vt_get_stats then reports the voxelizer kernel's time. */
int vt_voxelize(vt_context* ctx, const float* vertices, int64_t n_vertices, const uint32_t* indices,
                int64_t n_triangles, const float* model_transform, int32_t mode, int32_t material_offset,
                uint8_t* occupancy);

/* ---- host helpers (same arithmetic as the reference host code) ---------
 * updateCamera matrices (src/renderer/renderer.cpp:343-446). */
int vt_camera_matrices(const float* eye, const float* target, float fov_y, float near_d, float far_d,
                       int32_t lens_model, int32_t width, int32_t height, float* modelview,
                       float* inv_modelview, float* proj, float* inv_proj);
/* calculateCDF (src/renderer/image.cpp:104-319) over an RGB float image:
 * writes the function-image size, CDF U ((w'+1) x h'), CDF V (h'+1) and the
 * integral.  cdf_u/cdf_v may be NULL to query sizes. */
int vt_env_build_cdf(const float* rgb, int32_t w, int32_t h, int32_t* cdf_u_w, int32_t* cdf_u_h,
                     float* cdf_u, int32_t* cdf_v_n, float* cdf_v, float* integral);
/* computeMeshTransform (src/renderer/import.cpp:46-64): row-major 4x4 that
 * fits the bounds into the unit cube with a one-voxel margin. */
int vt_mesh_transform(const float* bmin, const float* bmax, const int32_t* res, float* M);
/* pruneInteriorEmissiveVoxels (src/renderer/import.cpp:227-250); returns
 * the new count. */
int32_t vt_prune_emissive(const int32_t* material_offsets, int32_t nx, int32_t ny, int32_t nz,
                          int32_t* emissive, int32_t n_emissive);

/* ---- diagnostics (not part of the renderer's interface) -------------------
 * Device evaluation of the elementary functions used by the kernels, for
 * checking that device and host arithmetic agree bit for bit.  fn: 0 sqrt,
 * 1 a/b, 2 sin, 3 cos, 4 acos, 5 atan2(a,b), 6 log2, 7 exp2, 8 pow(a,b). */
int vt_device_math(vt_context* ctx, int32_t fn, int32_t n, const float* a, const float* b, float* out);

/* Diagnostic: streaming copy of `bytes` bytes between two device buffers,
 * `iters` times, with a plain vectorized copy kernel; returns the mean kernel
 * time.  A known-size reference for hardware byte counters and the HBM
 * roofline (bench.py --calibrate). */
int vt_copy_probe(vt_context* ctx, uint64_t bytes, int32_t iters, double* kernel_ms);

#ifdef __cplusplus
}
#endif
#endif

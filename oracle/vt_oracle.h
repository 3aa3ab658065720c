/*
 * vt_oracle.h - CPU restatement of voxelToy's rendering hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * kernels: it is loaded by the tests, by the smoke check and by the CPU
 * baseline leg of the benchmark, and nothing in the product links it.
 *
 * Every function follows one reference function op for op in IEEE single
 * precision with contraction off (built with -ffp-contract=off); the
 * reference file:line is cited at each definition in vt_oracle.c.  Scene
 * data is taken in the reference's own formats: an int32 material-offset
 * grid (x fastest, -1 = empty, src/renderer/renderer.cpp:863-872), the float
 * material-record array (src/renderer/loaders/voxLoader.cpp:241-253), the
 * int32 emissive voxel list, row-major camera matrices as updateCamera
 * builds them (src/renderer/renderer.cpp:343-446) and the CDF tables of
 * src/renderer/image.cpp:104-319.
 */
#ifndef VT_ORACLE_H_
#define VT_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene {
    int32_t res[3];
    float bmin[3], bmax[3], vsize[3];
    const int32_t* offsets;      /* res[0]*res[1]*res[2], x fastest; -1 empty */
    const float* mat;            /* material records */
    int32_t mat_size;
    const int32_t* emissive;     /* linear voxel indices */
    int32_t n_emissive;
} orc_scene;

typedef struct orc_camera {
    float inv_mv[16];    /* row-major, m[r*4+c] */
    float proj[16];
    float inv_proj[16];
    float viewport[4];
    float win_z;         /* gl_FragCoord.z of the full-screen quad */
    float lens_radius;
    float focal_distance;
    int32_t lens_model;  /* 0 pinhole, 1 thin lens, 2 orthographic */
} orc_camera;

typedef struct orc_env {
    int32_t use_image;
    float top[3], bottom[3];
    int32_t img_w, img_h;
    const float* img;            /* RGB float, row 0 = v = 0 */
    int32_t cdfu_w, cdfu_h;      /* (w'+1) x h' */
    const float* cdfu;
    int32_t cdfv_n;              /* h'+1 */
    const float* cdfv;
    float integral;
    float rotation;              /* radians */
} orc_env;

typedef struct orc_settings {
    int32_t max_bounces;
    float exposure;
    float gamma;
    int32_t selected[3];         /* SelectVoxelData.index; out of range = none */
    float wire_opacity;
    float wire_thickness;
} orc_settings;

typedef struct orc_counters {
    uint64_t rays;          /* camera samples + shadow + continuation traversals */
    uint64_t visits;        /* voxel fetches made by the DDA */
    uint64_t interactions;  /* bounce-loop iterations (surface shading events) */
    uint64_t samples;
    uint64_t env_samples;   /* importance-sampled image-environment draws */
    uint64_t env_evals;     /* image-environment lookups by direction */
} orc_counters;

/* --- elementary pieces, exposed for unit tests --------------------------- */
float orc_aabb(const float o[3], const float d[3], const float bmin[3], const float bmax[3]);
int orc_max_steps(const int32_t res[3]);
/* Returns 1 on hit.  vp receives the final voxel position, *steps the number
 * of DDA advances, *wrote 0 when the start point was outside the grid. */
int orc_raymarch(const orc_scene* s, const float o[3], const float d[3], int max_steps,
                 int32_t vp[3], int32_t* steps, int32_t* wrote);
/* traverse(): voxel hit or ground hit (final y < 0). kind: 0 miss, 2 voxel, 3 ground */
int orc_traverse(const orc_scene* s, const float o[3], const float d[3],
                 int32_t vp[3], int32_t* steps, int32_t* kind);
/* voxelSpaceToWorldSpace: out[12] = position, tangent, normal, binormal */
void orc_basis(const orc_scene* s, const int32_t vp[3], const float o[3], const float d[3], float out[12]);
void orc_generate_ray(const orc_camera* cam, float fx, float fy, const float u[4], float o[3], float d[3]);
void orc_pcg32_raw(uint64_t initstate, uint64_t initseq, int n, uint32_t* out);
void orc_pcg32_stream(uint64_t seed, uint32_t pixel, uint32_t frame, int n, float* out);
void orc_env_sample_uv(const orc_env* e, float su, float sv, float uv[2], int32_t idx[2]);
void orc_env_texture(const orc_env* e, float u, float v, float rgb[3]);
void orc_uv_from_dir(const float d[3], float rotation, float uv[2]);
float orc_powf(float x, float y);
void orc_sincosf(float x, float* s, float* c);
float orc_atan2f(float y, float x);
float orc_acosf(float x);
float orc_log2f(float x);
float orc_exp2f(float x);

/* --- probes ------------------------------------------------------------- */
/* Per-pixel first hit of the primary ray for frame `frame`.  rec[8] per
 * pixel: kind (0 AABB miss, 1 traversal miss, 2 voxel, 3 ground), vx, vy, vz,
 * steps (-1 when the march started outside the grid), nx, ny, nz (face
 * normal of voxelSpaceToWorldSpace as integers).  Pixels bottom-up. */
void orc_first_hit_image(const orc_scene* s, const orc_camera* cam, int32_t W, int32_t H,
                         uint64_t seed, uint32_t frame, int32_t* rec);
/* Injected rays: same record per ray (kind 1 = miss). */
void orc_trace_rays(const orc_scene* s, int32_t n, const float* o, const float* d, int32_t* rec);

/* --- integrator ---------------------------------------------------------- */
/* One sample of pathTracer.fs main for pixel (px, py) at `frame`; rgb
 * receives the tonemapped sample (alpha is 1). */
void orc_sample(const orc_scene* s, const orc_env* e, const orc_camera* cam, const orc_settings* st,
                int32_t W, int32_t H, int32_t px, int32_t py, uint64_t seed, uint32_t frame,
                uint32_t rng_frame, float rgb[3], orc_counters* cnt);
/* Progressive render of frames n0 .. n0+count-1 over rows [y0, y1).  accum
 * is W*H*4 floats (running mean, accumulation.fs), sum64 optional W*H*3
 * doubles (running sums of the same samples).  rng frame = min(n,
 * rng_frame_max).  nthreads > 1 uses OpenMP over pixels. */
void orc_render(const orc_scene* s, const orc_env* e, const orc_camera* cam, const orc_settings* st,
                int32_t W, int32_t H, int32_t y0, int32_t y1, uint64_t seed, uint32_t n0, uint32_t count,
                uint32_t rng_frame_max, float* accum, double* sum64, orc_counters* cnt, int32_t nthreads);

/* --- voxelizer (voxelize.vs + voxelize.gs) ------------------------------- */
/* mode 0 THIN, 1 FAT.  M is the row-major model transform (as uploaded with
 * transpose).  occ is res[0]*res[1]*res[2] bytes, x fastest; covered voxels
 * get `value`. */
void orc_voxelize(const float* verts, const uint32_t* idx, int64_t ntri, const float M[16],
                  const int32_t res[3], int32_t mode, uint8_t value, uint8_t* occ);
void orc_mesh_transform(const float bmin[3], const float bmax[3], const int32_t res[3], float M[16]);

/* --- host-side preparation ---------------------------------------------- */
/* calculateCDF core over an already filtered function image f (w x h). */
void orc_calculate_cdf(const float* f, int32_t w, int32_t h, float* cdfu, float* cdfv, float* integral);
/* generateImageFunction restatement: resize (if > max_size) + luminance +
 * 3x3 gaussian.  out must hold w'*h'; returns w', h'. */
void orc_image_function(const float* rgb, int32_t w, int32_t h, int32_t max_size,
                        float* out, int32_t* ow, int32_t* oh);
/* pruneInteriorEmissiveVoxels; returns the new count, list edited in place */
int32_t orc_prune_emissive(const int32_t* offsets, const int32_t res[3], int32_t* list, int32_t n);
/* updateCamera: params = eye[3], target[3], fovY, near, far, dist(for ortho);
 * lens_model; W,H.  Writes mv, inv_mv, proj, inv_proj (row-major). */
void orc_camera_matrices(const float eye[3], const float target[3], float fov_y, float near_d,
                         float far_d, int32_t lens_model, int32_t W, int32_t H,
                         float mv[16], float inv_mv[16], float proj[16], float inv_proj[16]);

#ifdef __cplusplus
}
#endif
#endif

/*
 * vt_renderer.h - C ABI over the renderer object (the C++ class
 * voxeltoy::Renderer, which mirrors the reference's Renderer,
 * src/renderer/renderer.h:27-122).  Every function returns VT_OK (0) or a
 * negative VT_ERR_* code unless stated otherwise; the message of the last
 * failure is vtr_get_status().  A renderer created with device -1 has no
 * GPU context: loading, camera and settings work, render() fails.
 */
#ifndef VT_RENDERER_H
#define VT_RENDERER_H

#include "vt_api.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vtr_renderer vtr_renderer;

/* RenderSettings (src/renderer/renderSettings.h:6-29) plus the tonemap
 * constants of pathTracer.fs:294. */
typedef struct vtr_settings {
    int32_t max_bounces;
    int32_t max_samples;
    int32_t width, height;             /* image resolution */
    int32_t viewport[4];
    float wireframe_opacity, wireframe_thickness;
    const char* background_image;      /* .pfm path, NULL or "" = gradient */
    float background_top[3], background_bottom[3];
    int32_t background_rotation_degrees;
    float exposure, gamma;
} vtr_settings;

/* CameraParameters state (src/camera/cameraParameters.h:18-129). */
typedef struct vtr_camera_state {
    float eye[3], target[3];
    float fov_y, near_d, far_d;
    float focal_length, focal_distance, lens_radius;
    float film[2];
    int32_t lens_model;                /* 0 pinhole, 1 thin lens, 2 orthographic */
} vtr_camera_state;

/* Picking actions (src/renderer/actions.h:7-13). */
enum { VTR_PA_SELECT_FOCAL_POINT = 0, VTR_PA_SELECT_ACTIVE_VOXEL = 1, VTR_PA_ADD_VOXEL = 2, VTR_PA_REMOVE_VOXEL = 3 };

int vtr_create(int device, vtr_renderer** out);
void vtr_destroy(vtr_renderer* r);
const char* vtr_get_status(vtr_renderer* r);
/* The device context the renderer drives (NULL without a device). */
vt_context* vtr_context(vtr_renderer* r);

int vtr_initialize(vtr_renderer* r, const char* shader_path);
int vtr_resize_frame(vtr_renderer* r, int32_t fb_w, int32_t fb_h, int32_t vp_x, int32_t vp_y, int32_t vp_w,
                     int32_t vp_h);
/* One progressive pass: returns 0 (RR_SAMPLES_PENDING), 1
 * (RR_FINISHED_RENDERING) or a negative error. */
int vtr_render(vtr_renderer* r);
int vtr_reset_render(vtr_renderer* r);
int vtr_set_seed(vtr_renderer* r, uint64_t seed);
int32_t vtr_number_samples(vtr_renderer* r);

int vtr_load_vox_file(vtr_renderer* r, const char* path);
int vtr_load_mesh(vtr_renderer* r, const char* path);
/* Dense material-offset grid (x fastest, -1 empty), material data and
 * emissive list (copied). */
int vtr_load_volume(vtr_renderer* r, int32_t nx, int32_t ny, int32_t nz, const int32_t* offsets,
                    const float* mat_data, int32_t mat_size, const int32_t* emissive, int32_t n_emissive);
/* Host copy of the current volume: res[3], bounds; offsets may be NULL. */
int vtr_get_volume(vtr_renderer* r, int32_t* res, float* bmin, float* bmax, int32_t* offsets);
int vtr_get_emissive(vtr_renderer* r, int32_t* out, int32_t capacity, int32_t* count);
int vtr_save_image(vtr_renderer* r, const char* path);
/* CPUVoxelizer::voxelizeMesh (src/voxelize/cpuVoxelizer.h:5-13) on host
 * threads, with the mesh placed like vt_voxelize places it (row-major model
 * transform, resolution res[3]); occupancy (res[0] x res[1] x res[2] bytes,
 * x fastest) gets 1 at every covered voxel.  mode 0 THIN, 1 FAT. */
int vtr_cpu_voxelize(const float* vertices, int64_t n_vertices, const uint32_t* indices, int64_t n_triangles,
                     const float* model_transform, const int32_t* res, int32_t mode, uint8_t* occupancy);
/* Checkpoint / resume of the progressive render (frame size, sample count,
 * seed, running mean); load into a renderer with the saved one's scene,
 * camera and settings.  No counterpart in the reference. */
int vtr_save_checkpoint(vtr_renderer* r, const char* path);
int vtr_load_checkpoint(vtr_renderer* r, const char* path);

int vtr_get_render_settings(vtr_renderer* r, vtr_settings* out);
/* Copy the settings in and apply them (Renderer::updateRenderSettings). */
int vtr_update_render_settings(vtr_renderer* r, const vtr_settings* s);

int vtr_camera_get(vtr_renderer* r, vtr_camera_state* out);
int vtr_camera_set_eye_target(vtr_renderer* r, const float* eye, const float* target);
int vtr_camera_look_at(vtr_renderer* r, const float* target);
int vtr_camera_set_distance_from_target(vtr_renderer* r, float distance);
int vtr_camera_orbit(vtr_renderer* r, float theta, float phi);
int vtr_camera_focus_on_bounds(vtr_renderer* r);
int vtr_camera_set_lens_model(vtr_renderer* r, int32_t model);
int vtr_camera_set_focal_length(vtr_renderer* r, float length);
int vtr_camera_set_film_size(vtr_renderer* r, float w, float h);
int vtr_camera_set_lens_radius(vtr_renderer* r, float radius);
int vtr_camera_set_fstop(vtr_renderer* r, float fstop);
int vtr_camera_set_fov_y(vtr_renderer* r, float fov);
/* Camera::setCameraController (src/camera/camera.h:22-28, camera.cpp:25-35):
 * 0 = CCM_ORBIT, 1 = CCM_FLY; selects how vtr_on_mouse_move and
 * vtr_on_key_press move the camera. */
int vtr_camera_set_controller(vtr_renderer* r, int32_t mode);
/* The vt_camera block the renderer uploads for the current state. */
int vtr_camera_block(vtr_renderer* r, vt_camera* out);

int vtr_on_mouse_move(vtr_renderer* r, int32_t dx, int32_t dy, int32_t buttons);
int vtr_on_key_press(vtr_renderer* r, int32_t key);

/* Queue a picking action for the next render (x, y normalised, y down). */
int vtr_request_action(vtr_renderer* r, float x, float y, float dx, float dy, int32_t action, int32_t restart);
/* Results of the picking services. */
float vtr_focal_distance(vtr_renderer* r);
int vtr_selected_voxel(vtr_renderer* r, int32_t* xyz);

/* MeshLoader::loadFromOBJ (src/mesh/meshLoader.cpp:11-75) without a
 * renderer: counts (and, when the arrays are large enough, copies) the
 * vertices (xyz floats) and triangle indices, plus the bounds. */
int vtr_load_obj(const char* path, float* vertices, int64_t vertex_capacity, int64_t* n_vertices,
                 uint32_t* indices, int64_t index_capacity, int64_t* n_triangles, float* bmin, float* bmax);

int vtr_get_material_data(vtr_renderer* r, float* out, int32_t capacity, int32_t* size);
int vtr_update_material_color(vtr_renderer* r, uint32_t offset, const float* rgb);
int vtr_update_material_value(vtr_renderer* r, uint32_t offset, float value);

#ifdef __cplusplus
}
#endif
#endif

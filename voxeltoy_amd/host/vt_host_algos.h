/*
 * vt_host_algos.h - host-side preparation steps of the render path, written
 * with the same float arithmetic as the reference host code:
 *   - camera matrices of Renderer::updateCamera (src/renderer/renderer.cpp:343-446)
 *   - the environment CDF builder (src/renderer/image.cpp:104-425)
 *   - interior emissive-voxel pruning (src/renderer/import.cpp:180-250)
 *   - the mesh-to-unit-cube transform (src/renderer/import.cpp:46-64)
 * Built with -ffp-contract=off.
 */
#pragma once

#include <cstdint>
#include <vector>

namespace vt {
namespace host {

/* Row-major 4x4 inverse computed in double and rounded to float. */
void inverse4(const float* m, float* out);

/* updateCamera: eye/target define the getBasis frame
 * (src/camera/cameraParameters.cpp:43-50). */
void camera_matrices(const float eye[3], const float target[3], float fov_y, float near_d, float far_d,
                     int lens_model, int width, int height, float mv[16], float inv_mv[16], float proj[16],
                     float inv_proj[16]);

/* calculateCDF core over an already filtered function image. */
void calculate_cdf(const float* f, int w, int h, float* cdf_u, float* cdf_v, float* integral);

/* generateImageFunction: luminance, resize to max_size (Lanczos-3 widened
 * by the minification ratio), normalised 3x3 Gaussian with black borders. */
void image_function(const float* rgb, int w, int h, int max_size, std::vector<float>& out, int& ow, int& oh);

/* pruneInteriorEmissiveVoxels; returns the new count (list edited in place). */
int32_t prune_emissive(const int32_t* offsets, const int32_t res[3], int32_t* list, int32_t n);

/* computeMeshTransform (row-major). */
void mesh_transform(const float bmin[3], const float bmax[3], const int32_t res[3], float M[16]);

}  // namespace host
}  // namespace vt

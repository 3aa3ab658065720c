/*
 * renderer.cpp - voxeltoy::Renderer over the vt_* C ABI, and the vtr_* C ABI
 * over voxeltoy::Renderer.  Method-by-method correspondence with the
 * reference (src/renderer/renderer.cpp, import.cpp, actions.cpp, services/)
 * is noted at each function.
 */
#include "renderer.h"

#include "vt_host_algos.h"
#include "vt_renderer.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>

namespace voxeltoy {

/* ---- GPUVoxelizer (gpuVoxelizer.cpp:46-72) ------------------------------------ */
bool GPUVoxelizer::voxelizeMesh(const Mesh* mesh, const M44f& M, const V3i& resolution, int32_t materialOffset,
                                int32_t mode)
{
    (void)resolution;   /* the grid of the context defines the resolution */
    if (!m_ctx || !mesh) return false;
    const int r = vt_voxelize(m_ctx, mesh->vertices().data(), (int64_t)mesh->numVertices(), mesh->indices().data(),
                              (int64_t)mesh->numTriangles(), M.data(), mode, materialOffset, nullptr);
    if (r != VT_OK) {
        if (m_logger) (*m_logger)(std::string("voxelize failed: ") + vt_last_error());
        return false;
    }
    return true;
}

/* computeMeshTransform (import.cpp:46-64). */
M44f computeMeshTransform(const Box3f& bounds, const V3i& res)
{
    const float bmin[3] = { bounds.min.x, bounds.min.y, bounds.min.z };
    const float bmax[3] = { bounds.max.x, bounds.max.y, bounds.max.z };
    const int32_t r[3] = { res.x, res.y, res.z };
    M44f m;
    vt::host::mesh_transform(bmin, bmax, r, m.data());
    return m;
}

/* ---- Renderer ------------------------------------------------------------------ */

/* Renderer::Renderer (renderer.cpp:41-65): camera looks at the origin from
 * distance 100 at f/16; 512x512 image, 1 bounce, 2048^2 samples. */
Renderer::Renderer() : Renderer(0) {}

Renderer::Renderer(int device) : m_device(device)
{
    m_camera.controller().lookAt(V3f(0.0f, 0.0f, 0.0f));
    m_camera.controller().setDistanceFromTarget(100.0f);
    m_camera.setFStop(16.0f);
}

Renderer::~Renderer()
{
    if (m_ctx) vt_destroy(m_ctx);
}

void Renderer::info(const std::string& msg)
{
    m_status = msg;
    if (m_logger) (*m_logger)(msg);
}

void Renderer::fail(const std::string& msg)
{
    m_failed = true;
    info(msg);
}

/* Renderer::initialize (renderer.cpp:76-147).  There are no shaders to
 * compile; the path is kept for interface compatibility only. */
void Renderer::initialize(const std::string& shaderPath)
{
    (void)shaderPath;
    if (m_device >= 0 && !m_ctx) {
        if (vt_create(m_device, &m_ctx) != VT_OK) {
            m_ctx = nullptr;
            fail(std::string("renderer initialisation failed: ") + vt_last_error());
            return;
        }
    }
    m_initialized = true;
    createVoxelDataTexture(V3i(16));
    ensureFrame();
    updateCamera();
    updateRenderSettings();
    info("Renderer initialized.");
}

void Renderer::reloadShaders(const std::string& shaderPath) { (void)shaderPath; }

/* Allocate the accumulation target when the image size changed. */
bool Renderer::ensureFrame()
{
    const int W = m_renderSettings.m_imageResolution.x, H = m_renderSettings.m_imageResolution.y;
    if (!m_ctx) return false;
    if (W == m_frameW && H == m_frameH) return true;
    if (vt_set_frame(m_ctx, W, H, 0) != VT_OK) {
        fail(std::string("frame allocation failed: ") + vt_last_error());
        return false;
    }
    m_frameW = W;
    m_frameH = H;
    return true;
}

/* renderer.cpp:571-579: the film keeps 36 mm on its longer side and takes
 * the image aspect. */
void Renderer::fitFilmToImage()
{
    const float aspect = (float)m_renderSettings.m_imageResolution.x / m_renderSettings.m_imageResolution.y;
    if (aspect >= 1.0f)
        m_camera.setFilmSize(CameraParameters::FILM_SIZE_35MM, CameraParameters::FILM_SIZE_35MM / aspect);
    else
        m_camera.setFilmSize(CameraParameters::FILM_SIZE_35MM * aspect, CameraParameters::FILM_SIZE_35MM);
}

/* Renderer::resizeFrame (renderer.cpp:448-554). */
void Renderer::resizeFrame(int fbW, int fbH, int vx, int vy, int vw, int vh)
{
    m_numberSamples = 0;
    m_renderSettings.m_imageResolution = V2i(std::max(1, fbW), std::max(1, fbH));
    m_renderSettings.m_viewport[0] = vx;
    m_renderSettings.m_viewport[1] = vy;
    m_renderSettings.m_viewport[2] = vw;
    m_renderSettings.m_viewport[3] = vh;
    fitFilmToImage();
    ensureFrame();
    updateCamera();
}

/* The uniforms updateCamera uploads (renderer.cpp:343-446) as one block:
 * matrices, the integrator viewport (0, 0, W, H), the depth of the
 * full-screen quad at z = near under identity matrices, lens parameters and
 * the focal distance the focal-point service maintains. */
vt_camera Renderer::cameraBlock() const
{
    vt_camera c;
    std::memset(&c, 0, sizeof(c));
    const CameraParameters& p = m_camera.parameters();
    const float eye[3] = { p.eye().x, p.eye().y, p.eye().z };
    const float tgt[3] = { p.target().x, p.target().y, p.target().z };
    float mv[16];
    vt::host::camera_matrices(eye, tgt, p.fovY(), p.nearDistance(), p.farDistance(), (int)p.lensModel(),
                              m_renderSettings.m_imageResolution.x, m_renderSettings.m_imageResolution.y, mv,
                              c.inv_modelview, c.proj, c.inv_proj);
    c.viewport[0] = 0.0f;
    c.viewport[1] = 0.0f;
    c.viewport[2] = (float)m_renderSettings.m_imageResolution.x;
    c.viewport[3] = (float)m_renderSettings.m_imageResolution.y;
    c.win_z = 0.5f * p.nearDistance() + 0.5f;
    c.lens_radius = p.lensRadius();
    c.focal_distance = m_focalDistance;
    c.lens_model = (int32_t)p.lensModel();
    return c;
}

void Renderer::updateCamera()
{
    if (!m_initialized || !m_ctx) return;
    const vt_camera c = cameraBlock();
    if (vt_set_camera(m_ctx, &c) != VT_OK) fail(std::string("camera update failed: ") + vt_last_error());
}

/* Renderer::updateRenderSettings + loadBackgroundImage (renderer.cpp:947-1106):
 * a new background image is loaded and its sampling tables rebuilt once;
 * the accumulation restarts. */
void Renderer::updateRenderSettings()
{
    if (!m_initialized) return;
    const RenderSettings& s = m_renderSettings;
    if (m_ctx) {
        vt_settings st;
        std::memset(&st, 0, sizeof(st));
        st.max_bounces = s.m_pathtracerMaxNumBounces;
        st.exposure = s.m_exposure;
        st.gamma = s.m_gamma;
        st.selected_voxel[0] = m_selectedVoxel.x;
        st.selected_voxel[1] = m_selectedVoxel.y;
        st.selected_voxel[2] = m_selectedVoxel.z;
        st.wireframe_opacity = s.m_wireframeOpacity;
        st.wireframe_thickness = s.m_wireframeThickness;
        if (vt_set_settings(m_ctx, &st) != VT_OK) fail(std::string("settings rejected: ") + vt_last_error());

        vt_environment env;
        std::memset(&env, 0, sizeof(env));
        env.top[0] = s.m_backgroundColor[0].x; env.top[1] = s.m_backgroundColor[0].y; env.top[2] = s.m_backgroundColor[0].z;
        env.bottom[0] = s.m_backgroundColor[1].x; env.bottom[1] = s.m_backgroundColor[1].y;
        env.bottom[2] = s.m_backgroundColor[1].z;
        env.rotation_radians = (float)(s.m_backgroundRotationDegrees * M_PI / 180.0f);
        std::vector<float> img, cu, cv;
        if (!s.m_backgroundImage.empty()) {
            int w = 0, h = 0;
            std::string err;
            if (!readImage(s.m_backgroundImage, img, w, h, &err)) {
                fail("background image: " + err);
                return;
            }
            int32_t cuw = 0, cuh = 0, cvn = 0;
            float integral = 0.0f;
            if (vt_env_build_cdf(img.data(), w, h, &cuw, &cuh, nullptr, &cvn, nullptr, &integral) != VT_OK) {
                fail(std::string("background CDF: ") + vt_last_error());
                return;
            }
            cu.resize((size_t)cuw * cuh);
            cv.resize((size_t)cvn);
            vt_env_build_cdf(img.data(), w, h, &cuw, &cuh, cu.data(), &cvn, cv.data(), &integral);
            env.use_image = 1;
            env.image_w = w; env.image_h = h; env.image_rgb = img.data();
            env.cdf_u_w = cuw; env.cdf_u_h = cuh; env.cdf_u = cu.data();
            env.cdf_v_n = cvn; env.cdf_v = cv.data();
            env.integral = integral;
            m_currentBackgroundImage = s.m_backgroundImage;
        }
        if (vt_set_environment(m_ctx, &env) != VT_OK) fail(std::string("environment rejected: ") + vt_last_error());
    }
    resetRender();
}

/* Renderer::resetRender (renderer.cpp:941-945). */
void Renderer::resetRender()
{
    updateCamera();
    m_numberSamples = 0;
}

/* Renderer::render (renderer.cpp:556-645): one pass with sample count
 * n = m_numberSamples; the RNG frame is min(n, maxSamples - 1). */
Renderer::RenderResult Renderer::render()
{
    if (!m_initialized) return RR_FINISHED_RENDERING;
    if (!m_ctx) {
        fail("render: no device context");
        return RR_FINISHED_RENDERING;
    }
    processPendingActions();
    fitFilmToImage();
    if (!ensureFrame()) return RR_FINISHED_RENDERING;
    updateCamera();
    const int max_s = m_renderSettings.m_pathtracerMaxSamples;
    const uint32_t rng_max = max_s > 0 ? (uint32_t)(max_s - 1) : 0u;
    if (vt_render(m_ctx, m_seed, (uint32_t)m_numberSamples, 1, rng_max) != VT_OK) {
        fail(std::string("render failed: ") + vt_last_error());
        return RR_FINISHED_RENDERING;
    }
    vt_stats st;
    if (vt_get_stats(m_ctx, &st) == VT_OK && st.kernel_ms > 0.0) {
        std::ostringstream ss;
        ss << "fps " << 1000.0 / st.kernel_ms;
        m_status = ss.str();
    }
    if (m_numberSamples++ < m_renderSettings.m_pathtracerMaxSamples) return RR_SAMPLES_PENDING;
    return RR_FINISHED_RENDERING;
}

/* Renderer::createVoxelDataTexture (renderer.cpp:833-940): cubic voxels,
 * longest axis 1000 units, centred. */
bool Renderer::createVoxelDataTexture(const V3i& res, const int32_t* offsets, const float* mat, size_t n_mat,
                                      const int32_t* emissive, size_t n_emissive)
{
    m_volumeResolution = res;
    const float voxelSize = 1000.0f / (float)std::max(res.x, std::max(res.y, res.z));
    const V3f bs(voxelSize * res.x, voxelSize * res.y, voxelSize * res.z);
    m_volumeBounds = Box3f(-bs * 0.5f, bs * 0.5f);
    const size_t nvox = (size_t)res.x * res.y * res.z;
    if (offsets) m_offsets.assign(offsets, offsets + nvox);
    else m_offsets.assign(nvox, -1);
    m_materials.assign(mat, mat + n_mat);
    m_emissive.assign(emissive, emissive + n_emissive);
    return uploadVolume();
}

bool Renderer::uploadVolume()
{
    if (!m_ctx) return true;
    const float bmin[3] = { m_volumeBounds.min.x, m_volumeBounds.min.y, m_volumeBounds.min.z };
    const float bmax[3] = { m_volumeBounds.max.x, m_volumeBounds.max.y, m_volumeBounds.max.z };
    if (vt_set_volume(m_ctx, m_volumeResolution.x, m_volumeResolution.y, m_volumeResolution.z, m_offsets.data(),
                      m_materials.data(), (int32_t)m_materials.size(), m_emissive.data(), (int32_t)m_emissive.size(),
                      bmin, bmax) != VT_OK) {
        fail(std::string("volume upload failed: ") + vt_last_error());
        return false;
    }
    return true;
}

bool Renderer::loadVolume(const V3i& res, const std::vector<int32_t>& offsets, const std::vector<float>& mat,
                          const std::vector<int32_t>& emissive)
{
    if ((size_t)res.x * res.y * res.z != offsets.size() || res.x <= 0 || res.y <= 0 || res.z <= 0) {
        fail("loadVolume: grid size does not match the resolution");
        return false;
    }
    const bool ok = createVoxelDataTexture(res, offsets.data(), mat.data(), mat.size(), emissive.data(),
                                           emissive.size());
    resetRender();
    return ok;
}

/* pruneInteriorEmissiveVoxels (import.cpp:180-250). */
void Renderer::pruneInteriorEmissiveVoxels(const std::vector<int32_t>& voxelMaterials, V3i& res,
                                           std::vector<int32_t>& emissive)
{
    const int32_t r[3] = { res.x, res.y, res.z };
    const int32_t n = vt::host::prune_emissive(voxelMaterials.data(), r, emissive.data(), (int32_t)emissive.size());
    emissive.resize((size_t)n);
}

/* Renderer::loadVoxFile (import.cpp:11-44). */
void Renderer::loadVoxFile(const std::string& file)
{
    std::vector<int32_t> offsets, emissive;
    std::vector<float> mat;
    V3i res;
    MagicaVoxelLoader loader;
    if (!loader.load(file, offsets, mat, emissive, res)) {
        fail("loadVoxFile: " + loader.error());
        return;
    }
    pruneInteriorEmissiveVoxels(offsets, res, emissive);
    createVoxelDataTexture(res, offsets.data(), mat.data(), mat.size(), emissive.data(), emissive.size());
    m_camera.controller().setDistanceFromTarget(m_volumeBounds.size().length() * 0.5f);
    resetRender();
}

/* Renderer::loadMesh (import.cpp:66-130): 64^3 grid, one grey Lambert
 * record at offset 0, THIN voxelization of the mesh fitted to the unit cube. */
void Renderer::loadMesh(const std::string& file)
{
    std::string err;
    Mesh* mesh = MeshLoader::loadFromOBJ(file.c_str(), &err);
    if (!mesh) {
        fail("loadMesh: " + err);
        return;
    }
    const float mat[7] = { (float)Material::MT_LAMBERT, 0.0f, 0.0f, 0.0f, 0.75f, 0.75f, 0.75f };
    createVoxelDataTexture(V3i(64), nullptr, mat, 7, nullptr, 0);
    const M44f M = computeMeshTransform(mesh->bounds(), m_volumeResolution);
    if (m_ctx) {
        GPUVoxelizer vox(m_ctx, m_logger);
        if (vox.voxelizeMesh(mesh, M, m_volumeResolution, 0)) {
            std::vector<int32_t> grid(m_offsets.size());
            if (vt_read_volume(m_ctx, grid.data()) == VT_OK) m_offsets.swap(grid);
        }
    }
    delete mesh;
    resetRender();
}

/* Renderer::saveImage (renderer.cpp:1108-1140). */
void Renderer::saveImage(const std::string& file)
{
    if (!m_initialized || !m_ctx) return;
    const int W = m_frameW, H = m_frameH;
    std::vector<float> px((size_t)W * H * 4);
    if (vt_read_accum(m_ctx, px.data()) != VT_OK) {
        fail(std::string("saveImage: ") + vt_last_error());
        return;
    }
    std::string err;
    if (!writeImage(file, px.data(), W, H, &err)) fail("saveImage: " + err);
}

/* Checkpoint file: 8-byte tag, then int32 width, height, sample count and a
 * zero word, uint64 seed, then width x height x 4 floats (row 0 = bottom). */
namespace {
const char kCheckpointTag[8] = { 'V', 'T', 'C', 'K', 'P', 'T', '0', '1' };
struct CheckpointHeader {
    char tag[8];
    int32_t width, height, samples, zero;
    uint64_t seed;
};
}  // namespace

bool Renderer::saveCheckpoint(const std::string& file)
{
    if (!m_initialized || !m_ctx || m_frameW <= 0) {
        fail("saveCheckpoint: nothing rendered");
        return false;
    }
    std::vector<float> px((size_t)m_frameW * m_frameH * 4);
    if (vt_read_accum(m_ctx, px.data()) != VT_OK) {
        fail(std::string("saveCheckpoint: ") + vt_last_error());
        return false;
    }
    CheckpointHeader h;
    std::memcpy(h.tag, kCheckpointTag, sizeof(h.tag));
    h.width = m_frameW;
    h.height = m_frameH;
    h.samples = m_numberSamples;
    h.zero = 0;
    h.seed = m_seed;
    std::ofstream f(file, std::ios::binary);
    f.write((const char*)&h, sizeof(h));
    f.write((const char*)px.data(), (std::streamsize)(px.size() * sizeof(float)));
    if (!f) {
        fail("saveCheckpoint: cannot write " + file);
        return false;
    }
    return true;
}

bool Renderer::loadCheckpoint(const std::string& file)
{
    if (!m_initialized || !m_ctx) {
        fail("loadCheckpoint: renderer not initialized");
        return false;
    }
    std::ifstream f(file, std::ios::binary);
    CheckpointHeader h;
    if (!f.read((char*)&h, sizeof(h)) || std::memcmp(h.tag, kCheckpointTag, sizeof(h.tag)) != 0) {
        fail("loadCheckpoint: " + file + " is not a checkpoint");
        return false;
    }
    if (h.width != m_renderSettings.m_imageResolution.x || h.height != m_renderSettings.m_imageResolution.y ||
        h.samples < 0) {
        fail("loadCheckpoint: checkpoint frame does not match the renderer's image resolution");
        return false;
    }
    std::vector<float> px((size_t)h.width * h.height * 4);
    if (!f.read((char*)px.data(), (std::streamsize)(px.size() * sizeof(float)))) {
        fail("loadCheckpoint: " + file + " is truncated");
        return false;
    }
    fitFilmToImage();
    if (!ensureFrame()) return false;
    updateCamera();
    if (vt_write_accum(m_ctx, px.data()) != VT_OK) {
        fail(std::string("loadCheckpoint: ") + vt_last_error());
        return false;
    }
    m_numberSamples = h.samples;
    m_seed = h.seed;
    return true;
}

/* Renderer::onMouseMove (renderer.cpp:658-669): pixel motion normalised by
 * the image resolution, handed to the active camera controller. */
bool Renderer::onMouseMove(int dx, int dy, int buttons)
{
    const float ndx = (float)dx / m_renderSettings.m_imageResolution.x;
    const float ndy = (float)dy / m_renderSettings.m_imageResolution.y;
    if (m_camera.controller().onMouseMove(ndx, ndy, buttons)) {
        updateCamera();
        m_numberSamples = 0;
        return true;
    }
    return false;
}

/* Renderer::onKeyPress (renderer.cpp:670-692): the active controller first
 * (the fly controller's W/A/S/D), then 'F' focuses on the volume bounds.
 * The space key's switch to the edit-mode preview integrator has no
 * counterpart here. */
bool Renderer::onKeyPress(int key)
{
    if (m_camera.controller().onKeyPress(key)) {
        updateCamera();
        m_numberSamples = 0;
        return true;
    }
    if (key == 'F' || key == 'f') {
        m_camera.controller().focusOnBounds(m_volumeBounds);
        m_numberSamples = 0;
        updateCamera();
        return true;
    }
    return false;
}

/* Renderer::requestAction (actions.cpp:5-19). */
void Renderer::requestAction(float x, float y, float dx, float dy, Action::PICKING_ACTION action, bool restart)
{
    Action a;
    a.m_point = V2f(x, y);
    a.m_velocity = V2f(dx, dy);
    a.m_type = action;
    a.m_invalidatesRender = restart;
    m_scheduledActions.push_back(a);
}

/* rayAABBIntersection (aabb.h:1-32) on the host. */
static float ray_aabb(const float o[3], const float d[3], const float bmin[3], const float bmax[3])
{
    float tmin = -INFINITY, tmax = INFINITY;
    for (int k = 0; k < 3; ++k) {
        const float inv = 1.0f / d[k];
        const float t0 = (bmin[k] - o[k]) * inv, t1 = (bmax[k] - o[k]) * inv;
        tmin = std::max(tmin, std::min(t0, t1));
        tmax = std::min(tmax, std::max(t0, t1));
    }
    if (tmax < 0.0f || tmin > tmax) return -1.0f;
    return std::max(0.0f, tmin);
}

/* Picking ray through window position (px, py): the pinhole branch of
 * generateRay (the services compile with PINHOLE), without the pixel jitter
 * (the reference draws it from an arbitrary noise texel). */
bool Renderer::pickRay(float px, float py, float origin[3], float dir[3]) const
{
    const vt_camera c = cameraBlock();
    const float nx = ((2.0f * px) - (2.0f * c.viewport[0])) / c.viewport[2] - 1.0f;
    const float ny = ((2.0f * py) - (2.0f * c.viewport[1])) / c.viewport[3] - 1.0f;
    const float nz = (2.0f * c.win_z) - 1.0f;
    const float P32 = c.proj[14], P22 = c.proj[10], P23 = c.proj[11];
    const float w = P32 / (nz - (P22 / P23));
    float e[4];
    const float v[4] = { nx * w, ny * w, nz * w, w };
    for (int r = 0; r < 4; ++r)
        e[r] = ((c.inv_proj[r * 4] * v[0] + c.inv_proj[r * 4 + 1] * v[1]) + c.inv_proj[r * 4 + 2] * v[2]) +
               c.inv_proj[r * 4 + 3] * v[3];
    const float el = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    if (!(el > 0.0f)) return false;
    const float ed[3] = { e[0] / el, e[1] / el, e[2] / el };
    float wd[3];
    for (int r = 0; r < 3; ++r) {
        origin[r] = c.inv_modelview[r * 4 + 3];
        wd[r] = (c.inv_modelview[r * 4] * ed[0] + c.inv_modelview[r * 4 + 1] * ed[1]) + c.inv_modelview[r * 4 + 2] * ed[2];
    }
    const float wl = std::sqrt(wd[0] * wd[0] + wd[1] * wd[1] + wd[2] * wd[2]);
    if (!(wl > 0.0f)) return false;
    for (int r = 0; r < 3; ++r) dir[r] = wd[r] / wl;
    return true;
}

/* processPendingActions (actions.cpp:21-57) with the services:
 * focal point (focalDistance.vs:45-81), active voxel (selectVoxel.vs:36-71),
 * add voxel (addVoxel.vs:15-41), remove voxel (removeVoxel.vs:8-12). */
void Renderer::processPendingActions()
{
    for (const Action& a : m_scheduledActions) {
        if (a.m_invalidatesRender) m_numberSamples = 0;
        const float px = a.m_point.x * m_renderSettings.m_imageResolution.x;
        const float py = (1.0f - a.m_point.y) * m_renderSettings.m_imageResolution.y;
        const float vx = a.m_velocity.x * m_renderSettings.m_imageResolution.x;
        const float vy = a.m_velocity.y * m_renderSettings.m_imageResolution.y;
        float o[3], d[3];
        int32_t rec[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
        bool hit = false;
        float entry[3] = { 0, 0, 0 };
        const float bmin[3] = { m_volumeBounds.min.x, m_volumeBounds.min.y, m_volumeBounds.min.z };
        const float bmax[3] = { m_volumeBounds.max.x, m_volumeBounds.max.y, m_volumeBounds.max.z };
        if (m_ctx && (a.m_type == Action::PA_SELECT_FOCAL_POINT || a.m_type == Action::PA_SELECT_ACTIVE_VOXEL) &&
            pickRay(px, py, o, d)) {
            const float t = ray_aabb(o, d, bmin, bmax);
            if (t >= 0.0f) {
                for (int k = 0; k < 3; ++k) entry[k] = o[k] + t * d[k];
                if (vt_trace_rays(m_ctx, 1, entry, d, rec) == VT_OK) hit = rec[0] == 2 || rec[0] == 3;
            }
        }
        switch (a.m_type) {
            case Action::PA_SELECT_FOCAL_POINT: {
                m_focalDistance = 99999999.0f;
                if (hit) {
                    const V3f vs = m_volumeBounds.size();
                    const float sz[3] = { vs.x / m_volumeResolution.x, vs.y / m_volumeResolution.y,
                                          vs.z / m_volumeResolution.z };
                    float vmin[3], vmax[3];
                    for (int k = 0; k < 3; ++k) {
                        vmin[k] = (float)rec[1 + k] * sz[k] + bmin[k];
                        vmax[k] = vmin[k] + sz[k];
                    }
                    m_focalDistance = ray_aabb(o, d, vmin, vmax);
                }
                break;
            }
            case Action::PA_SELECT_ACTIVE_VOXEL:
                m_selectedVoxel = V3i(0);
                if (hit) {
                    m_selectedVoxel = V3i(rec[1], rec[2], rec[3]);
                    m_selectedNormal = V3f((float)rec[5], (float)rec[6], (float)rec[7]);
                }
                if (m_ctx) {
                    vt_settings st;
                    std::memset(&st, 0, sizeof(st));
                    const RenderSettings& s = m_renderSettings;
                    st.max_bounces = s.m_pathtracerMaxNumBounces;
                    st.exposure = s.m_exposure;
                    st.gamma = s.m_gamma;
                    st.selected_voxel[0] = m_selectedVoxel.x;
                    st.selected_voxel[1] = m_selectedVoxel.y;
                    st.selected_voxel[2] = m_selectedVoxel.z;
                    st.wireframe_opacity = s.m_wireframeOpacity;
                    st.wireframe_thickness = s.m_wireframeThickness;
                    vt_set_settings(m_ctx, &st);
                }
                break;
            case Action::PA_ADD_VOXEL:
            case Action::PA_REMOVE_VOXEL: {
                V3i c = m_selectedVoxel;
                int32_t off = -1;
                if (a.m_type == Action::PA_ADD_VOXEL) {
                    /* the new voxel goes next to the selected one, along the
                     * screen-space motion's dominant camera axis or else the
                     * selected face normal; it copies the selected voxel's
                     * material (record 0 when the selection is empty) */
                    V3f n = m_selectedNormal;
                    if (vx != 0.0f || vy != 0.0f) {
                        const vt_camera cb = cameraBlock();
                        const V3f right(cb.inv_modelview[0], cb.inv_modelview[4], cb.inv_modelview[8]);
                        const V3f up(cb.inv_modelview[1], cb.inv_modelview[5], cb.inv_modelview[9]);
                        auto axis = [](const V3f& v) {
                            const V3f av(std::fabs(v.x), std::fabs(v.y), std::fabs(v.z));
                            auto sgn = [](float f) { return (float)((f > 0) - (f < 0)); };
                            if (av.x >= av.y && av.x >= av.z) return V3f(sgn(v.x), 0, 0);
                            if (av.y >= av.x && av.y >= av.z) return V3f(0, sgn(v.y), 0);
                            return V3f(0, 0, sgn(v.z));
                        };
                        const bool horiz = std::fabs(vx) >= std::fabs(vy);
                        n = horiz ? axis(right) * (float)((vx > 0) - (vx < 0)) : axis(up) * (float)((vy > 0) - (vy < 0));
                    }
                    const V3i src = c;
                    c = V3i(src.x + (int)n.x, src.y + (int)n.y, src.z + (int)n.z);
                    const bool src_in = src.x >= 0 && src.y >= 0 && src.z >= 0 && src.x < m_volumeResolution.x &&
                                        src.y < m_volumeResolution.y && src.z < m_volumeResolution.z;
                    off = 0;
                    if (src_in) {
                        const int32_t so = m_offsets[(size_t)src.x + (size_t)m_volumeResolution.x *
                                                                        ((size_t)src.y + (size_t)m_volumeResolution.y * src.z)];
                        if (so >= 0) off = so;
                    }
                    if (m_materials.empty()) break;
                }
                if (c.x < 0 || c.y < 0 || c.z < 0 || c.x >= m_volumeResolution.x || c.y >= m_volumeResolution.y ||
                    c.z >= m_volumeResolution.z)
                    break;
                const int32_t lin = c.x + m_volumeResolution.x * (c.y + m_volumeResolution.y * c.z);
                m_offsets[(size_t)lin] = off;
                m_emissive.erase(std::remove(m_emissive.begin(), m_emissive.end(), lin), m_emissive.end());
                if (off >= 0 && (size_t)off + 3 < m_materials.size() &&
                    (m_materials[off + 1] + m_materials[off + 2] + m_materials[off + 3]) / 3 > 0.0f)
                    m_emissive.push_back(lin);
                uploadVolume();
                break;
            }
        }
    }
    m_scheduledActions.clear();
}

/* Renderer::getMaterials (renderer.cpp:1142-1203). */
std::vector<Material::SerializedData> Renderer::getMaterials() const
{
    std::string err;
    std::vector<Material::SerializedData> r = Material::serializeAll(m_materials, &err);
    if (!err.empty() && m_logger) (*m_logger)(err);
    return r;
}

/* updateMaterialColor / updateMaterialValue (renderer.cpp:1205-1235): edit
 * the record in place; the accumulation is not restarted. */
void Renderer::updateMaterialColor(unsigned int off, const float color[3])
{
    if (!m_initialized || (size_t)off + 3 > m_materials.size()) return;
    std::memcpy(&m_materials[off], color, 3 * sizeof(float));
    if (m_ctx && vt_update_material(m_ctx, (int32_t)off, color, 3) != VT_OK)
        fail(std::string("updateMaterialColor: ") + vt_last_error());
}

void Renderer::updateMaterialValue(unsigned int off, float value)
{
    if (!m_initialized || (size_t)off >= m_materials.size()) return;
    m_materials[off] = value;
    if (m_ctx && vt_update_material(m_ctx, (int32_t)off, &value, 1) != VT_OK)
        fail(std::string("updateMaterialValue: ") + vt_last_error());
}

}  // namespace voxeltoy

/* ---- C ABI ------------------------------------------------------------------------ */
using voxeltoy::Renderer;

struct vtr_renderer {
    Renderer r;
    std::string bg;     /* storage behind vtr_settings.background_image */
    explicit vtr_renderer(int device) : r(device) {}
};

/* Run one renderer call and map its failure flag to a status code. */
template <typename F>
static int guarded(vtr_renderer* r, int err, F&& f)
{
    if (!r) return VT_ERR_ARG;
    r->r.clearFailure();
    f();
    return r->r.failed() ? err : VT_OK;
}

/* No C++ exception leaves the C ABI: an entry point's body runs in no_throw,
 * and a failed host allocation (containers sized from files or the caller's
 * input) or any other exception becomes an error code. */
template <typename F>
static int no_throw(F&& f) noexcept
{
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return VT_ERR_LIMIT;
    } catch (...) {
        return VT_ERR_STATE;
    }
}

extern "C" {

int vtr_create(int device, vtr_renderer** out)
{
    return no_throw([&]() -> int {
        if (!out) return VT_ERR_ARG;
        *out = new vtr_renderer(device);
        return VT_OK;
    });
}

void vtr_destroy(vtr_renderer* r) { delete r; }

const char* vtr_get_status(vtr_renderer* r) { return r ? r->r.getStatus().c_str() : "null renderer"; }
vt_context* vtr_context(vtr_renderer* r) { return r ? r->r.context() : nullptr; }

int vtr_initialize(vtr_renderer* r, const char* shader_path)
{
    return no_throw([&]() -> int {
        return guarded(r, VT_ERR_STATE, [&] { r->r.initialize(shader_path ? shader_path : ""); });
    });
}

int vtr_resize_frame(vtr_renderer* r, int32_t w, int32_t h, int32_t vx, int32_t vy, int32_t vw, int32_t vh)
{
    return no_throw([&]() -> int {
        if (!r || w <= 0 || h <= 0 || w > 16384 || h > 16384) return VT_ERR_ARG;
        r->r.resizeFrame(w, h, vx, vy, vw, vh);
        return VT_OK;
    });
}

int vtr_render(vtr_renderer* r)
{
    return no_throw([&]() -> int {
        if (!r) return VT_ERR_ARG;
        if (!r->r.context()) return VT_ERR_STATE;
        Renderer::RenderResult rr = Renderer::RR_FINISHED_RENDERING;
        const int e = guarded(r, VT_ERR_STATE, [&] { rr = r->r.render(); });
        if (e != VT_OK) return e;
        return rr == Renderer::RR_SAMPLES_PENDING ? 0 : 1;
    });
}

int vtr_reset_render(vtr_renderer* r)
{
    return no_throw([&]() -> int {
        if (!r) return VT_ERR_ARG;
        r->r.resetRender();
        return VT_OK;
    });
}

int vtr_set_seed(vtr_renderer* r, uint64_t seed)
{
    return no_throw([&]() -> int {
        if (!r) return VT_ERR_ARG;
        r->r.setSeed(seed);
        return VT_OK;
    });
}

int32_t vtr_number_samples(vtr_renderer* r) { return r ? r->r.numberSamples() : VT_ERR_ARG; }

int vtr_load_vox_file(vtr_renderer* r, const char* path)
{
    return no_throw([&]() -> int {
        if (!path) return VT_ERR_ARG;
        return guarded(r, VT_ERR_ARG, [&] { r->r.loadVoxFile(path); });
    });
}

int vtr_load_mesh(vtr_renderer* r, const char* path)
{
    return no_throw([&]() -> int {
        if (!path) return VT_ERR_ARG;
        return guarded(r, VT_ERR_ARG, [&] { r->r.loadMesh(path); });
    });
}

int vtr_load_volume(vtr_renderer* r, int32_t nx, int32_t ny, int32_t nz, const int32_t* offsets,
                    const float* mat, int32_t mat_size, const int32_t* emissive, int32_t n_emissive)
{
    return no_throw([&]() -> int {
        if (!r || nx <= 0 || ny <= 0 || nz <= 0 || !offsets || mat_size < 0 || n_emissive < 0 ||
            (mat_size > 0 && !mat) || (n_emissive > 0 && !emissive))
            return VT_ERR_ARG;
        const size_t n = (size_t)nx * ny * nz;
        std::vector<int32_t> o(offsets, offsets + n), e(emissive, emissive + n_emissive);
        std::vector<float> m(mat, mat + mat_size);
        return guarded(r, VT_ERR_ARG, [&] { r->r.loadVolume(voxeltoy::V3i(nx, ny, nz), o, m, e); });
    });
}

int vtr_get_volume(vtr_renderer* r, int32_t* res, float* bmin, float* bmax, int32_t* offsets)
{
    return no_throw([&]() -> int {
        if (!r) return VT_ERR_ARG;
        const voxeltoy::V3i v = r->r.volumeResolution();
        const voxeltoy::Box3f b = r->r.volumeBounds();
        if (res) { res[0] = v.x; res[1] = v.y; res[2] = v.z; }
        if (bmin) { bmin[0] = b.min.x; bmin[1] = b.min.y; bmin[2] = b.min.z; }
        if (bmax) { bmax[0] = b.max.x; bmax[1] = b.max.y; bmax[2] = b.max.z; }
        if (offsets) std::memcpy(offsets, r->r.hostOffsets().data(), r->r.hostOffsets().size() * sizeof(int32_t));
        return VT_OK;
    });
}

int vtr_get_emissive(vtr_renderer* r, int32_t* out, int32_t cap, int32_t* count)
{
    return no_throw([&]() -> int {
        if (!r || !count) return VT_ERR_ARG;
        const std::vector<int32_t>& e = r->r.hostEmissive();
        *count = (int32_t)e.size();
        if (out) std::memcpy(out, e.data(), std::min<size_t>(e.size(), (size_t)std::max(0, cap)) * sizeof(int32_t));
        return VT_OK;
    });
}

int vtr_save_image(vtr_renderer* r, const char* path)
{
    return no_throw([&]() -> int {
        if (!path) return VT_ERR_ARG;
        if (r && !r->r.context()) return VT_ERR_STATE;
        return guarded(r, VT_ERR_ARG, [&] { r->r.saveImage(path); });
    });
}

int vtr_save_checkpoint(vtr_renderer* r, const char* path)
{
    return no_throw([&]() -> int {
        if (!r || !path) return VT_ERR_ARG;
        return r->r.saveCheckpoint(path) ? VT_OK : VT_ERR_STATE;
    });
}

int vtr_load_checkpoint(vtr_renderer* r, const char* path)
{
    return no_throw([&]() -> int {
        if (!r || !path) return VT_ERR_ARG;
        return r->r.loadCheckpoint(path) ? VT_OK : VT_ERR_STATE;
    });
}

int vtr_get_render_settings(vtr_renderer* r, vtr_settings* o)
{
    return no_throw([&]() -> int {
        if (!r || !o) return VT_ERR_ARG;
        const voxeltoy::RenderSettings& s = r->r.renderSettings();
        o->max_bounces = s.m_pathtracerMaxNumBounces;
        o->max_samples = s.m_pathtracerMaxSamples;
        o->width = s.m_imageResolution.x;
        o->height = s.m_imageResolution.y;
        for (int k = 0; k < 4; ++k) o->viewport[k] = s.m_viewport[k];
        o->wireframe_opacity = s.m_wireframeOpacity;
        o->wireframe_thickness = s.m_wireframeThickness;
        r->bg = s.m_backgroundImage;
        o->background_image = r->bg.c_str();
        for (int k = 0; k < 3; ++k) {
            o->background_top[k] = s.m_backgroundColor[0][k];
            o->background_bottom[k] = s.m_backgroundColor[1][k];
        }
        o->background_rotation_degrees = s.m_backgroundRotationDegrees;
        o->exposure = s.m_exposure;
        o->gamma = s.m_gamma;
        return VT_OK;
    });
}

int vtr_update_render_settings(vtr_renderer* r, const vtr_settings* i)
{
    return no_throw([&]() -> int {
        if (!r || !i) return VT_ERR_ARG;
        if (i->max_bounces < 0 || i->width <= 0 || i->height <= 0 || !(i->gamma != 0.0f)) return VT_ERR_ARG;
        voxeltoy::RenderSettings& s = r->r.renderSettings();
        const bool resized = i->width != s.m_imageResolution.x || i->height != s.m_imageResolution.y;
        s.m_pathtracerMaxNumBounces = i->max_bounces;
        s.m_pathtracerMaxSamples = i->max_samples;
        s.m_wireframeOpacity = i->wireframe_opacity;
        s.m_wireframeThickness = i->wireframe_thickness;
        s.m_backgroundImage = i->background_image ? i->background_image : "";
        for (int k = 0; k < 3; ++k) {
            s.m_backgroundColor[0][k] = i->background_top[k];
            s.m_backgroundColor[1][k] = i->background_bottom[k];
        }
        s.m_backgroundRotationDegrees = i->background_rotation_degrees;
        s.m_exposure = i->exposure;
        s.m_gamma = i->gamma;
        return guarded(r, VT_ERR_ARG, [&] {
            if (resized)
                r->r.resizeFrame(i->width, i->height, i->viewport[0], i->viewport[1], i->viewport[2], i->viewport[3]);
            for (int k = 0; k < 4; ++k) s.m_viewport[k] = i->viewport[k];
            r->r.updateRenderSettings();
        });
    });
}

int vtr_camera_get(vtr_renderer* r, vtr_camera_state* o)
{
    return no_throw([&]() -> int {
        if (!r || !o) return VT_ERR_ARG;
        const voxeltoy::CameraParameters& p = r->r.camera().parameters();
        for (int k = 0; k < 3; ++k) { o->eye[k] = p.eye()[k]; o->target[k] = p.target()[k]; }
        o->fov_y = p.fovY();
        o->near_d = p.nearDistance();
        o->far_d = p.farDistance();
        o->focal_length = p.focalLength();
        o->focal_distance = p.focalDistance();
        o->lens_radius = p.lensRadius();
        o->film[0] = p.filmSize().x;
        o->film[1] = p.filmSize().y;
        o->lens_model = (int32_t)p.lensModel();
        return VT_OK;
    });
}

#define VTR_CAM(body)                \
    if (!r) return VT_ERR_ARG;       \
    body;                            \
    r->r.resetRender();              \
    return VT_OK;

int vtr_camera_set_eye_target(vtr_renderer* r, const float* e, const float* t)
{
    return no_throw([&]() -> int {
        if (!e || !t) return VT_ERR_ARG;
        VTR_CAM(r->r.camera().controller().setEyeTarget(voxeltoy::V3f(e[0], e[1], e[2]), voxeltoy::V3f(t[0], t[1], t[2])))
    });
}
int vtr_camera_look_at(vtr_renderer* r, const float* t)
{
    return no_throw([&]() -> int {
        if (!t) return VT_ERR_ARG;
        VTR_CAM(r->r.camera().controller().lookAt(voxeltoy::V3f(t[0], t[1], t[2])))
    });
}
int vtr_camera_set_distance_from_target(vtr_renderer* r, float d)
{
    return no_throw([&]() -> int {
        VTR_CAM(r->r.camera().controller().setDistanceFromTarget(d))
    });
}
int vtr_camera_orbit(vtr_renderer* r, float theta, float phi) { VTR_CAM(r->r.camera().controller().orbit(theta, phi)) }
int vtr_camera_focus_on_bounds(vtr_renderer* r) { VTR_CAM(r->r.camera().controller().focusOnBounds(r->r.volumeBounds())) }
int vtr_camera_set_lens_model(vtr_renderer* r, int32_t m)
{
    return no_throw([&]() -> int {
        if (m < 0 || m > 2) return VT_ERR_ARG;
        VTR_CAM(r->r.camera().setLensModel((voxeltoy::CameraParameters::CameraLensModel)m))
    });
}
int vtr_camera_set_focal_length(vtr_renderer* r, float l) { VTR_CAM(r->r.camera().setFocalLength(l)) }
int vtr_camera_set_film_size(vtr_renderer* r, float w, float h) { VTR_CAM(r->r.camera().setFilmSize(w, h)) }
int vtr_camera_set_lens_radius(vtr_renderer* r, float rad) { VTR_CAM(r->r.camera().setLensRadius(rad)) }
int vtr_camera_set_fstop(vtr_renderer* r, float f) { VTR_CAM(r->r.camera().setFStop(f)) }
int vtr_camera_set_fov_y(vtr_renderer* r, float f) { VTR_CAM(r->r.camera().setFovY(f)) }

int vtr_camera_set_controller(vtr_renderer* r, int32_t mode)
{
    return no_throw([&]() -> int {
        if (!r || (mode != 0 && mode != 1)) return VT_ERR_ARG;
        r->r.camera().setCameraController(mode == 1 ? voxeltoy::Camera::CCM_FLY : voxeltoy::Camera::CCM_ORBIT);
        return VT_OK;
    });
}

int vtr_camera_block(vtr_renderer* r, vt_camera* out)
{
    return no_throw([&]() -> int {
        if (!r || !out) return VT_ERR_ARG;
        *out = r->r.cameraBlock();
        return VT_OK;
    });
}

int vtr_on_mouse_move(vtr_renderer* r, int32_t dx, int32_t dy, int32_t buttons)
{
    return no_throw([&]() -> int {
        if (!r) return VT_ERR_ARG;
        return r->r.onMouseMove(dx, dy, buttons) ? 1 : 0;
    });
}

int vtr_on_key_press(vtr_renderer* r, int32_t key)
{
    return no_throw([&]() -> int {
        if (!r) return VT_ERR_ARG;
        return r->r.onKeyPress(key) ? 1 : 0;
    });
}

int vtr_request_action(vtr_renderer* r, float x, float y, float dx, float dy, int32_t action, int32_t restart)
{
    return no_throw([&]() -> int {
        if (!r || action < 0 || action > 3) return VT_ERR_ARG;
        r->r.requestAction(x, y, dx, dy, (voxeltoy::Action::PICKING_ACTION)action, restart != 0);
        return VT_OK;
    });
}

float vtr_focal_distance(vtr_renderer* r) { return r ? r->r.focalDistance() : -1.0f; }

int vtr_selected_voxel(vtr_renderer* r, int32_t* xyz)
{
    return no_throw([&]() -> int {
        if (!r || !xyz) return VT_ERR_ARG;
        const voxeltoy::V3i v = r->r.selectedVoxel();
        xyz[0] = v.x; xyz[1] = v.y; xyz[2] = v.z;
        return VT_OK;
    });
}

int vtr_load_obj(const char* path, float* vertices, int64_t vcap, int64_t* nv, uint32_t* indices, int64_t icap,
                 int64_t* nt, float* bmin, float* bmax)
{
    return no_throw([&]() -> int {
        if (!path || !nv || !nt) return VT_ERR_ARG;
        std::vector<float> v;
        std::vector<uint32_t> idx;
        std::string err;
        if (!voxeltoy::MeshLoader::loadFromOBJ(path, v, idx, &err)) return VT_ERR_ARG;
        *nv = (int64_t)(v.size() / 3);
        *nt = (int64_t)(idx.size() / 3);
        if (vertices && vcap >= *nv) std::memcpy(vertices, v.data(), v.size() * sizeof(float));
        if (indices && icap >= *nt) std::memcpy(indices, idx.data(), idx.size() * sizeof(uint32_t));
        const voxeltoy::Box3f b = voxeltoy::computeBounds(v.data(), v.size() / 3);
        if (bmin) { bmin[0] = b.min.x; bmin[1] = b.min.y; bmin[2] = b.min.z; }
        if (bmax) { bmax[0] = b.max.x; bmax[1] = b.max.y; bmax[2] = b.max.z; }
        return VT_OK;
    });
}

int vtr_get_material_data(vtr_renderer* r, float* out, int32_t cap, int32_t* size)
{
    return no_throw([&]() -> int {
        if (!r || !size) return VT_ERR_ARG;
        const std::vector<float>& m = r->r.hostMaterials();
        *size = (int32_t)m.size();
        if (out) std::memcpy(out, m.data(), std::min<size_t>(m.size(), (size_t)std::max(0, cap)) * sizeof(float));
        return VT_OK;
    });
}

int vtr_update_material_color(vtr_renderer* r, uint32_t off, const float* rgb)
{
    return no_throw([&]() -> int {
        if (!r || !rgb) return VT_ERR_ARG;
        if ((size_t)off + 3 > r->r.hostMaterials().size()) return VT_ERR_ARG;
        return guarded(r, VT_ERR_STATE, [&] { r->r.updateMaterialColor(off, rgb); });
    });
}

int vtr_update_material_value(vtr_renderer* r, uint32_t off, float v)
{
    return no_throw([&]() -> int {
        if (!r) return VT_ERR_ARG;
        if ((size_t)off >= r->r.hostMaterials().size()) return VT_ERR_ARG;
        return guarded(r, VT_ERR_STATE, [&] { r->r.updateMaterialValue(off, v); });
    });
}

}  // extern "C"

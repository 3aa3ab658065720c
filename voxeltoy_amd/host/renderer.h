/*
 * renderer.h - the renderer object a host application drives, mirroring the
 * public interface of the reference's Renderer (src/renderer/renderer.h:27-122),
 * RenderSettings (src/renderer/renderSettings.h:6-29), Action
 * (src/renderer/actions.h:5-24), Logger (src/log/logger.h:7-11) and
 * GPUVoxelizer (src/voxelize/gpuVoxelizer.h:9-25), on top of the vt_* C ABI.
 *
 * Semantics kept from the reference:
 *   - render() runs one progressive pass (1 spp for every pixel) and folds it
 *     into the running mean; it returns RR_SAMPLES_PENDING while fewer than
 *     m_pathtracerMaxSamples passes were made (renderer.cpp:556-645);
 *   - the film aspect follows the image resolution at every render() and
 *     resizeFrame() (renderer.cpp:571-579);
 *   - volumes get cubic voxels with the longest axis 1000 units, centred
 *     (renderer.cpp:845-850); .vox loading prunes interior emissive voxels
 *     and sets the camera distance to half the bounds diagonal
 *     (import.cpp:11-44); meshes are voxelized into a 64^3 grid
 *     (import.cpp:66-84);
 *   - the picking actions run at the start of the next render()
 *     (actions.cpp:21-57): focal point (thin-lens focal distance = distance
 *     to the voxel under the cursor), active voxel (painted red by the
 *     integrator), add voxel, remove voxel;
 *   - errors go to the Logger and getStatus(); no exceptions.
 * Differences (the reference's behaviour there is undefined or tied to
 * OpenGL): shaders do not exist, so initialize() ignores its path; images
 * are written as .pfm or .ppm; meshes are voxelized with material offset 0
 * of a one-record material table (the reference's GL path writes to stale
 * image bindings); add/remove voxel edit the material-offset grid directly.
 */
#pragma once

#include "camera.h"
#include "scene_io.h"
#include "vt_api.h"

#include <cstdint>
#include <string>
#include <vector>

namespace voxeltoy {

class Logger {
public:
    virtual ~Logger() = default;
    virtual void operator()(const std::string& msg) = 0;
};

struct RenderSettings {
    int m_pathtracerMaxNumBounces = 1;
    int m_pathtracerMaxSamples = 2048 * 2048;
    V2i m_imageResolution = V2i(512, 512);
    int m_viewport[4] = { 0, 0, 512, 512 };
    float m_wireframeOpacity = 0.0f;
    float m_wireframeThickness = 0.01f;
    /* path of an environment image (.pfm, .hdr/.pic, .ppm/.pgm); empty = gradient */
    std::string m_backgroundImage;
    V3f m_backgroundColor[2] = { V3f(153.0f / 255 * 2, 187.0f / 255 * 2, 201.0f / 255 * 2),
                                 V3f(77.0f / 255, 64.0f / 255, 50.0f / 255) };   /* top, bottom */
    int m_backgroundRotationDegrees = 0;
    /* tonemap of each sample: pow(L * exposure, 1 / gamma) (pathTracer.fs:294) */
    float m_exposure = 1.0f;
    float m_gamma = 2.2f;
};

struct Action {
    enum PICKING_ACTION { PA_SELECT_FOCAL_POINT, PA_SELECT_ACTIVE_VOXEL, PA_ADD_VOXEL, PA_REMOVE_VOXEL };
    PICKING_ACTION m_type = PA_SELECT_FOCAL_POINT;
    V2f m_point;       /* normalised screen position, y down (as UIs report it) */
    V2f m_velocity;    /* normalised screen-space motion */
    bool m_invalidatesRender = true;
};

/* Voxelize a mesh into the context's current grid: covered voxels get
 * material offset `materialOffset` (the reference's texture-unit handle). */
class GPUVoxelizer {
public:
    explicit GPUVoxelizer(vt_context* ctx, Logger* logger = nullptr) : m_ctx(ctx), m_logger(logger) {}
    bool voxelizeMesh(const Mesh* mesh, const M44f& meshTransform, const V3i& resolution, int32_t materialOffset,
                      int32_t mode = 0);

private:
    vt_context* m_ctx;
    Logger* m_logger;
};

/* CPUVoxelizer (src/voxelize/cpuVoxelizer.h:5-13): the same voxelization on
 * host threads (0 = one per hardware thread).  The first form takes
 * vertices already in voxel space (model transform and resolution applied,
 * as the reference's caller does, import.cpp:98-106); the second places a
 * mesh as GPUVoxelizer's vertex stage does.  Covered voxels of voxelStorage
 * (x fastest) are set to 1; mode 0 THIN (the reference's setting), 1 FAT.
 * Every triangle is voxelized: the reference's task split misses
 * triangles when the mesh has fewer triangles than threads
 * (cpuVoxelizer.cpp:289-299). */
class CPUVoxelizer {
public:
    static void voxelizeMesh(const V3f* vertices, const unsigned int* indices, unsigned int numTriangles,
                             const V3i& voxelDimensions, unsigned char* voxelStorage, int mode = 0, int threads = 0);
    static void voxelizeMesh(const Mesh* mesh, const M44f& meshTransform, const V3i& resolution,
                             unsigned char* voxelStorage, int mode = 0, int threads = 0);
};

/* computeMeshTransform (import.cpp:46-64): fit the bounds into the unit cube
 * with a one-voxel margin along the major axis. */
M44f computeMeshTransform(const Box3f& bounds, const V3i& voxelResolution);

class Renderer {
public:
    Renderer();
    explicit Renderer(int device);
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    /* Create the device context and default resources (a 16^3 empty grid). */
    void initialize(const std::string& shaderPath);
    void resizeFrame(int frameBufferWidth, int frameBufferHeight, int viewportX, int viewportY, int viewportW,
                     int viewportH);

    enum RenderResult { RR_SAMPLES_PENDING, RR_FINISHED_RENDERING };
    RenderResult render();
    void reloadShaders(const std::string& shaderPath);

    void loadMesh(const std::string& file);
    void loadVoxFile(const std::string& file);
    void pruneInteriorEmissiveVoxels(const std::vector<int32_t>& voxelMaterials, V3i& volumeResolution,
                                     std::vector<int32_t>& emissiveVoxelIndices);
    void saveImage(const std::string& file);
    void resetRender();
    /* Checkpoint of the progressive render (no counterpart in the reference):
     * frame size, sample count, RNG seed and the RGBA32F running mean.
     * loadCheckpoint restores them into a renderer holding the saved one's
     * scene, camera and settings (which the checkpoint does not carry), so
     * that rendering on gives the image of the uninterrupted render.
     * Failures go to the logger and return false. */
    bool saveCheckpoint(const std::string& file);
    bool loadCheckpoint(const std::string& file);

    /* Mouse and key input through the camera's active controller
     * (Camera::setCameraController: orbit or fly), pixel motion normalised
     * by the image resolution; 'F' focuses on the volume bounds. */
    bool onMouseMove(int dx, int dy, int buttons);
    bool onKeyPress(int key);

    Camera& camera() { return m_camera; }
    RenderSettings& renderSettings() { return m_renderSettings; }
    void updateRenderSettings();
    const std::string& getStatus() const { return m_status; }
    void setLogger(Logger* logger) { m_logger = logger; }

    void requestAction(float x, float y, float dx, float dy, Action::PICKING_ACTION action,
                       bool restartAccumulation);

    std::vector<Material::SerializedData> getMaterials() const;
    void updateMaterialColor(unsigned int dataOffset, const float color[3]);
    void updateMaterialValue(unsigned int dataOffset, float value);

    /* ---- additions for programmatic hosts and tests ---- */
    /* Replace the volume with a dense offset grid (x fastest, -1 empty). */
    bool loadVolume(const V3i& resolution, const std::vector<int32_t>& voxelMaterials,
                    const std::vector<float>& materialData, const std::vector<int32_t>& emissiveVoxelIndices);
    vt_context* context() { return m_ctx; }
    bool initialized() const { return m_initialized; }
    /* set by every failure reported to the logger since clearFailure() */
    bool failed() const { return m_failed; }
    void clearFailure() { m_failed = false; }
    /* host copies of the current volume */
    const std::vector<int32_t>& hostOffsets() const { return m_offsets; }
    const std::vector<float>& hostMaterials() const { return m_materials; }
    const std::vector<int32_t>& hostEmissive() const { return m_emissive; }
    void setSeed(uint64_t seed) { m_seed = seed; }
    int numberSamples() const { return m_numberSamples; }
    /* thin-lens focal distance in use (99999999 until a focal point is picked) */
    float focalDistance() const { return m_focalDistance; }
    void setFocalDistance(float d) { m_focalDistance = d; }
    V3i selectedVoxel() const { return m_selectedVoxel; }
    Box3f volumeBounds() const { return m_volumeBounds; }
    V3i volumeResolution() const { return m_volumeResolution; }
    /* The camera block updateCamera() uploads (matrices, viewport, lens). */
    vt_camera cameraBlock() const;

private:
    void updateCamera();
    void fitFilmToImage();
    void info(const std::string& msg);
    void fail(const std::string& msg);
    bool uploadVolume();
    bool createVoxelDataTexture(const V3i& resolution, const int32_t* voxelMaterials = nullptr,
                                const float* materialData = nullptr, size_t materialDataSize = 0,
                                const int32_t* emissiveVoxelIndices = nullptr, size_t numEmissiveVoxels = 0);
    void processPendingActions();
    bool pickRay(float px, float py, float origin[3], float dir[3]) const;
    bool ensureFrame();

    bool m_initialized = false;
    int m_device = 0;
    vt_context* m_ctx = nullptr;
    Camera m_camera;
    RenderSettings m_renderSettings;
    Box3f m_volumeBounds;
    V3i m_volumeResolution = V3i(16);
    int m_numberSamples = 0;
    uint64_t m_seed = 0;
    std::string m_status;
    Logger* m_logger = nullptr;
    std::vector<Action> m_scheduledActions;
    float m_focalDistance = 99999999.0f;
    V3i m_selectedVoxel = V3i(0);
    V3f m_selectedNormal = V3f(1.0f, 0.0f, 0.0f);
    int m_frameW = 0, m_frameH = 0;
    std::string m_currentBackgroundImage;
    bool m_failed = false;
    std::vector<int32_t> m_offsets;
    std::vector<float> m_materials;
    std::vector<int32_t> m_emissive;
};

}  // namespace voxeltoy

/*
 * vt_api.hip - host side of the C ABI declared in include/vt_api.h.
 *
 * A context owns one HIP stream (default priority) on one device and all
 * device buffers: the Morton-bricked grid and its occupancy bits, 256
 * decoded material records, the emissive list, the environment image and
 * CDF tables, and the RGBA32F accumulation target (plus optional FP64
 * sums).  Host arrays are copied on upload.
 */
#include "../../include/vt_api.h"
#include "vt_kernels.h"
#include "vt_cache.h"
#include "../host/vt_host_algos.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

/* No C++ exception leaves the C ABI: an entry point's body runs in no_throw,
 * and a failed host allocation (containers sized from the caller's input)
 * or any other exception becomes an error code.  The messages are short
 * enough to be stored without allocating. */
template <typename F>
int no_throw(F&& f) noexcept
{
    try {
        return f();
    } catch (const std::bad_alloc&) {
        g_err.assign("out of memory");
        return VT_ERR_LIMIT;
    } catch (...) {
        g_err.assign("internal error");
        return VT_ERR_STATE;
    }
}

/* Wavefront defaults (vt_set_render_mode, vt_set_traversal_tuning). */
#define VT_DEFAULT_MAX_PATHS (16ull << 20)
#define VT_TIMED_LAUNCHES 256     /* traversal launches timed per render call */
#define VT_OVERLAP_MIN_PATHS (1ull << 22)   /* paths of a render worth two streams */
#define VT_DEFAULT_DDA_BUDGET 32
#define VT_DEFAULT_REFILL_MIN 8
#define VT_DEFAULT_REFILL_CHUNK 256
/* (x, y) columns above which the voxelizer gives a triangle a whole wave */
#define VT_VOX_BIG_AREA 32
/* paths from which a scene of a batch gets its own launches */
#ifndef VT_BATCH_SOLO_PATHS
#define VT_BATCH_SOLO_PATHS (1ull << 22)
#endif

/* On failure the runtime's last-error state is cleared too, so a failed
 * call is not reported again by the next launch check. */
#define VT_HIP(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            (void)hipGetLastError();                                                      \
            return fail(VT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));  \
        }                                                                                 \
    } while (0)

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

/* A call's temporary device buffer, freed when the call returns, the early
 * error returns of VT_HIP included. */
template <typename T>
struct DevTmp {
    T* p = nullptr;
    DevTmp() = default;
    DevTmp(const DevTmp&) = delete;
    DevTmp& operator=(const DevTmp&) = delete;
    ~DevTmp() { dfree(p); }
};

}  // namespace

/* ---- staging memory of vt_trace_rays_noalloc -------------------------------------
 * A fixed device array, one copy per device, held until the process ends:
 * origins (VT_NOALLOC_MAX_RAYS x 12 B), directions (x 12 B) and hit records
 * (x 32 B).  Its host-side name is a stand-in, not a device address, so host
 * code uses only the address that hipGetSymbolAddress returns for each
 * device (looked up once per device by vt_create, with that device current). */
#define VT_STR_(x) #x
#define VT_STR(x) VT_STR_(x)
namespace {
constexpr size_t VT_NOALLOC_O = 0;
constexpr size_t VT_NOALLOC_D = (size_t)VT_NOALLOC_MAX_RAYS * 12u;
constexpr size_t VT_NOALLOC_R = (size_t)VT_NOALLOC_MAX_RAYS * 24u;
constexpr size_t VT_NOALLOC_BYTES = (size_t)VT_NOALLOC_MAX_RAYS * 56u;
constexpr int VT_MAX_DEVICES = 64;
static_assert(VT_NOALLOC_R % 256 == 0, "record staging must stay aligned");

std::mutex g_noalloc_lock;                       /* guards the table and serializes calls */
void* g_noalloc_ptr[VT_MAX_DEVICES] = {};

/* Device address of the current device's copy of `scratch`. */
int noalloc_lookup(int device);
}  // namespace

static __device__ __align__(256) unsigned char scratch[VT_NOALLOC_BYTES];

namespace {
int noalloc_lookup(int device)
{
    std::lock_guard<std::mutex> lk(g_noalloc_lock);
    if (g_noalloc_ptr[device]) return VT_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return fail(VT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_noalloc_ptr[device] = ptr;
    return VT_OK;
}
}  // namespace

struct vt_context {
    /* batches and caches built on this context; destroying the context
     * releases their device memory and leaves them unusable (their own
     * destroy call then only frees the handle) */
    std::vector<vt_batch*> batches;
    std::vector<vt_cache*> caches;
    int device = 0;
    int cus = 0;
    size_t lds_limit = 65536;       /* dynamic LDS per workgroup */
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timed = false;
    /* second stream: a wavefront render of many passes runs its two halves
     * on two streams, so each half's launch tails overlap the other's bulk */
    int overlap = 1;
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_half = nullptr, ev_join = nullptr;
    vt::WfPaths wf2{};
    uint64_t wf2_alloc = 0;
    /* per-launch traversal timing (wavefront mode) */
    std::vector<hipEvent_t> lt_ev;
    int lt_used = 0;
    bool lt_on = false;

    /* volume */
    bool has_volume = false;
    vt::BrickDims dims{};
    float bmin[3] = { 0, 0, 0 }, bmax[3] = { 0, 0, 0 }, vs[3] = { 0, 0, 0 };
    uint8_t* d_bricks = nullptr;
    uint64_t* d_occ64 = nullptr;
    uint32_t* d_bits = nullptr;
    uint32_t bit_words = 0;
    uint64_t* d_occp = nullptr;     /* padded levels of the wave march */
    uint64_t* d_occr = nullptr;     /* padded masks with reversed bits, for the packed march */
    uint32_t* d_bitsp = nullptr;
    uint32_t bit_words_p = 0;
    int32_t* d_emissive = nullptr;
    int32_t n_emissive = 0;
    int4* d_lights = nullptr;       /* emissive entries decoded: x, y, z, slot */
    std::vector<float> mat;
    int32_t slot_off[256];
    int32_t n_slots = 0;            /* slots 1..n_slots used */
    vt::MatRec* d_recs = nullptr;

    /* environment */
    vt_environment env{};
    float* d_img = nullptr;
    float4* d_quad = nullptr;       /* bilinear quads of the image */
    float* d_cdfu = nullptr;
    float* d_cdfv = nullptr;

    vt_camera cam{};
    vt_settings st{};

    /* frame */
    int32_t W = 0, H = 0;
    float4* d_accum = nullptr;
    double* d_sum64 = nullptr;

    int render_mode = 1;
    uint64_t max_paths = VT_DEFAULT_MAX_PATHS;
    int32_t dda_budget = VT_DEFAULT_DDA_BUDGET;
    int32_t refill_min = VT_DEFAULT_REFILL_MIN;
    int32_t refill_chunk = VT_DEFAULT_REFILL_CHUNK;
    vt::WfPaths wf{};
    uint64_t wf_alloc = 0;           /* slots allocated in wf */

    uint32_t* d_tile = nullptr;
    unsigned long long* d_stats = nullptr;
    bool stats_enabled = false;
    bool stats_valid = false;

    vt_context() { for (int i = 0; i < 256; ++i) slot_off[i] = -1; }
};

namespace {

float mat_at(const std::vector<float>& m, int64_t i) { return (i < 0 || i >= (int64_t)m.size()) ? 0.0f : m[(size_t)i]; }

/* Decode the record a material offset points at (materials.h:7-56): type
 * at off, emission at off+1..3, albedo/reflectance at off+4..6, roughness
 * at off+7; reads past the array return 0 as texelFetch does. */
vt::MatRec decode(const std::vector<float>& m, int32_t off)
{
    vt::MatRec r;
    const float t = mat_at(m, off);
    r.type = (t >= 0.0f && t < 3.0f) ? (int32_t)t : -1;
    for (int k = 0; k < 3; ++k) {
        r.e[k] = r.type < 0 ? 0.0f : mat_at(m, (int64_t)off + 1 + k);
        r.r[k] = mat_at(m, (int64_t)off + 4 + k);
    }
    r.rough = mat_at(m, (int64_t)off + 7);
    return r;
}

int upload_records(vt_context* c)
{
    vt::MatRec recs[256];
    recs[0] = decode(c->mat, 0);
    for (int s = 1; s < 256; ++s) recs[s] = decode(c->mat, c->slot_off[s] < 0 ? 0 : c->slot_off[s]);
    if (!c->d_recs) VT_HIP(hipMalloc(&c->d_recs, sizeof(recs)));
    VT_HIP(hipMemcpyAsync(c->d_recs, recs, sizeof(recs), hipMemcpyHostToDevice, c->stream));
    VT_HIP(hipStreamSynchronize(c->stream));
    return VT_OK;
}

/* Rebuild both occupancy levels from the slot bytes. */
vt::TraceArgs trace_args(const vt_context* c);

/* After any change of the bricks or the emissive list: block occupancy and
 * the emissive light table. */
int refresh_occupancy(vt_context* c)
{
    VT_HIP(vt::launch_block_masks(c->dims, c->d_bricks, c->d_occ64, c->d_bits, c->d_occp, c->d_bitsp, c->d_occr,
                                  c->stream));
    dfree(c->d_lights);
    if (c->n_emissive > 0 && c->d_emissive) {
        VT_HIP(hipMalloc(&c->d_lights, (size_t)c->n_emissive * 16u));
        const vt::TraceArgs a = trace_args(c);
        VT_HIP(vt::launch_light_table(a.g, c->d_lights, c->stream));
    }
    return VT_OK;
}

vt::BrickDims make_dims(int32_t nx, int32_t ny, int32_t nz)
{
    vt::BrickDims d;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.bx = (nx + 7) / 8; d.by = (ny + 7) / 8; d.bz = (nz + 7) / 8;
    int32_t m = std::max(d.bx, std::max(d.by, d.bz));
    int32_t p = 1;
    while (p < m) p <<= 1;
    d.bdim = p;
    d.n_bricks = (uint64_t)p * p * p;
    d.kx = (nx + 3) / 4; d.ky = (ny + 3) / 4; d.kz = (nz + 3) / 4;
    d.n_blocks = (uint64_t)d.kx * d.ky * d.kz;
    return d;
}

/* Reference bounds (renderer.cpp:845-850): cubic voxels, longest axis 1000. */
void reference_bounds(int32_t nx, int32_t ny, int32_t nz, float* bmin, float* bmax)
{
    const float sizeMultiplier = 1000.0f;
    const float voxelSize = sizeMultiplier / (float)std::max(nx, std::max(ny, nz));
    const float bs[3] = { voxelSize * (float)nx, voxelSize * (float)ny, voxelSize * (float)nz };
    for (int k = 0; k < 3; ++k) {
        bmin[k] = -bs[k] * 0.5f;
        bmax[k] = bs[k] * 0.5f;
    }
}

int common_volume_setup(vt_context* c, int32_t nx, int32_t ny, int32_t nz, const float* mat_data, int32_t mat_size,
                        const float* bmin, const float* bmax)
{
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > 4096 || ny > 4096 || nz > 4096)
        return fail(VT_ERR_ARG, "volume resolution must be in [1, 4096] per axis");
    if (mat_size < 0 || (mat_size > 0 && !mat_data)) return fail(VT_ERR_ARG, "bad material data");
    const vt::BrickDims d = make_dims(nx, ny, nz);
    c->has_volume = false;
    dfree(c->d_bricks);
    dfree(c->d_occ64);
    dfree(c->d_bits);
    dfree(c->d_occp);
    dfree(c->d_bitsp);
    dfree(c->d_occr);
    dfree(c->d_emissive);
    c->dims = d;
    if (bmin && bmax) {
        for (int k = 0; k < 3; ++k) { c->bmin[k] = bmin[k]; c->bmax[k] = bmax[k]; }
    } else {
        reference_bounds(nx, ny, nz, c->bmin, c->bmax);
    }
    const int32_t res[3] = { nx, ny, nz };
    for (int k = 0; k < 3; ++k) c->vs[k] = (c->bmax[k] - c->bmin[k]) / (float)res[k];
    c->mat.assign(mat_data, mat_data + mat_size);
    for (int i = 0; i < 256; ++i) c->slot_off[i] = -1;
    c->n_slots = 0;
    VT_HIP(hipMalloc(&c->d_bricks, d.n_bricks * 512u));
    VT_HIP(hipMalloc(&c->d_occ64, d.n_blocks * 8u));
    c->bit_words = (uint32_t)((d.n_blocks + 31) / 32);
    VT_HIP(hipMalloc(&c->d_bits, (size_t)c->bit_words * 4u));
    const uint64_t np = (uint64_t)(d.kx + 2) * (uint64_t)(d.ky + 2) * (uint64_t)(d.kz + 2);
    VT_HIP(hipMalloc(&c->d_occp, np * 8u));
    c->bit_words_p = (uint32_t)((np + 31) / 32);
    VT_HIP(hipMalloc(&c->d_bitsp, (size_t)c->bit_words_p * 4u));
    if (nx <= VT_PACKED_MAX_EXTENT && ny <= VT_PACKED_MAX_EXTENT && nz <= VT_PACKED_MAX_EXTENT)
        VT_HIP(hipMalloc(&c->d_occr, np * 8u));
    VT_HIP(vt::launch_fill_u8(c->d_bricks, d.n_bricks * 512u, 0, c->stream));
    return VT_OK;
}

int finish_volume(vt_context* c, const int32_t* emissive, int32_t n_emissive, const std::vector<uint8_t>& occupied_of_linear,
                  bool have_occ)
{
    (void)occupied_of_linear;
    (void)have_occ;
    c->n_emissive = n_emissive;
    if (n_emissive > 0) {
        VT_HIP(hipMalloc(&c->d_emissive, (size_t)n_emissive * 4u));
        VT_HIP(hipMemcpyAsync(c->d_emissive, emissive, (size_t)n_emissive * 4u, hipMemcpyHostToDevice, c->stream));
    }
    int r = refresh_occupancy(c);
    if (r != VT_OK) return r;
    r = upload_records(c);
    if (r != VT_OK) return r;
    VT_HIP(hipStreamSynchronize(c->stream));
    c->has_volume = true;
    return VT_OK;
}

int check_emissive(const int32_t* emissive, int32_t n, int64_t nvox, const std::vector<uint8_t>& occ)
{
    if (n < 0 || (n > 0 && !emissive)) return fail(VT_ERR_ARG, "bad emissive list");
    for (int32_t i = 0; i < n; ++i) {
        if (emissive[i] < 0 || (int64_t)emissive[i] >= nvox)
            return fail(VT_ERR_ARG, "emissive voxel index out of range at " + std::to_string(i));
        if (!occ[(size_t)emissive[i]])
            return fail(VT_ERR_ARG, "emissive voxel " + std::to_string(emissive[i]) + " is empty");
    }
    return VT_OK;
}

vt::GridView grid_view(const vt_context* c)
{
    vt::GridView g;
    g.nx = c->dims.nx; g.ny = c->dims.ny; g.nz = c->dims.nz;
    const int32_t res[3] = { g.nx, g.ny, g.nz };
    {
        /* int(2 * ceil(length(vec3(voxelResolution))))  dda.h:98 */
        const float rx = (float)res[0], ry = (float)res[1], rz = (float)res[2];
        g.max_steps = (int32_t)(2.0f * std::ceil(std::sqrt((rx * rx + ry * ry) + rz * rz)));
    }
    g.rx = (float)res[0]; g.ry = (float)res[1]; g.rz = (float)res[2];
    for (int k = 0; k < 3; ++k) {
        g.bmin[k] = c->bmin[k];
        g.bmax[k] = c->bmax[k];
        g.vs[k] = c->vs[k];
        g.ext[k] = 1.0f / (c->bmax[k] - c->bmin[k]);
    }
    g.bricks = c->d_bricks;
    g.bx4 = (uint32_t)c->dims.kx;
    g.by4 = (uint32_t)c->dims.ky;
    g.occ64 = c->d_occ64;
    g.bits = c->d_bits;
    g.bit_words = c->bit_words;
    g.bxp = (uint32_t)c->dims.kx + 2u;
    g.byp = (uint32_t)c->dims.ky + 2u;
    g.occp = c->d_occp;
    g.bitsp = c->d_bitsp;
    g.bit_words_p = c->bit_words_p;
    g.occr = c->d_occr;
    g.wide = ((uint64_t)(c->dims.kx + 2) * (uint64_t)(c->dims.ky + 2) * (uint64_t)(c->dims.kz + 2)) > (1ull << 29);
    g.n_emissive = c->n_emissive;
    g.emissive = c->d_emissive;
    g.lights = c->d_lights;
    return g;
}

vt::TraceArgs trace_args(const vt_context* c)
{
    vt::TraceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.g = grid_view(c);
    a.e.use_image = c->env.use_image;
    for (int k = 0; k < 3; ++k) { a.e.top[k] = c->env.top[k]; a.e.bottom[k] = c->env.bottom[k]; }
    a.e.w = c->env.image_w; a.e.h = c->env.image_h; a.e.img = c->d_img; a.e.quad = c->d_quad;
    a.e.cu_w = c->env.cdf_u_w; a.e.cu_h = c->env.cdf_u_h; a.e.cdfu = c->d_cdfu;
    a.e.cv_n = c->env.cdf_v_n; a.e.cdfv = c->d_cdfv;
    a.e.integral = c->env.integral; a.e.rotation = c->env.rotation_radians;
    std::memcpy(a.c.inv_mv, c->cam.inv_modelview, sizeof(a.c.inv_mv));
    std::memcpy(a.c.inv_proj, c->cam.inv_proj, sizeof(a.c.inv_proj));
    a.c.P32 = c->cam.proj[2 * 4 + 3];
    a.c.P22 = c->cam.proj[2 * 4 + 2];
    a.c.P23 = c->cam.proj[3 * 4 + 2];
    std::memcpy(a.c.viewport, c->cam.viewport, sizeof(a.c.viewport));
    a.c.win_z = c->cam.win_z;
    a.c.lens_radius = c->cam.lens_radius;
    a.c.focal_distance = c->cam.focal_distance;
    a.c.lens_model = c->cam.lens_model;
    a.s.max_bounces = c->st.max_bounces;
    a.s.exposure = c->st.exposure;
    a.s.inv_gamma = 1.0f / c->st.gamma;
    for (int k = 0; k < 3; ++k) a.s.sel[k] = c->st.selected_voxel[k];
    a.s.wire_opacity = c->st.wireframe_opacity;
    a.s.wire_thickness = c->st.wireframe_thickness;
    a.recs = c->d_recs;
    a.W = c->W; a.H = c->H;
    a.accum = c->d_accum;
    a.sum64 = c->d_sum64;
    a.tile_counter = c->d_tile;
    a.tiles_x = (c->W + 7) / 8;
    a.n_tiles = a.tiles_x * ((c->H + 7) / 8);
    a.cdfv_lds_bytes = c->env.use_image ? ((c->env.cdf_v_n * 4 + 15) / 16) * 16 : 0;
    const size_t base = 256 * sizeof(vt::MatRec) + (size_t)a.cdfv_lds_bytes;
    a.lds_bits = (base + (size_t)c->bit_words * 4u <= c->lds_limit) ? 1 : 0;
    a.lds_bits_p = ((size_t)c->bit_words_p * 4u <= c->lds_limit) ? 1 : 0;
    a.dda_budget = c->dda_budget;
    a.refill_min = c->refill_min;
    a.refill_chunk = c->refill_chunk;
    return a;
}

void free_paths(vt::WfPaths& P, uint64_t& alloc)
{
    for (vt::PathBuf& B : P.s) {
        dfree(B.o); dfree(B.nd); dfree(B.thr); dfree(B.rad); dfree(B.dir); dfree(B.fac);
        dfree(B.meta); dfree(B.rng);
    }
    dfree(P.hit); dfree(P.smp);
    dfree(P.qa0); dfree(P.qb0); dfree(P.qt0); dfree(P.qa1); dfree(P.qb1); dfree(P.qt1); dfree(P.cnt);
    P.cap = 0;
    alloc = 0;
}

void free_wavefront(vt_context* c) { free_paths(c->wf, c->wf_alloc); }

/* Path storage for `slots` paths (reallocated only when it grows). */
int alloc_paths(vt::WfPaths& P, uint64_t& alloc, uint64_t slots)
{
    if (alloc >= slots) { P.cap = slots; return VT_OK; }
    free_paths(P, alloc);
    const size_t n = (size_t)slots;
    for (vt::PathBuf& B : P.s) {
        VT_HIP(hipMalloc(&B.o, n * 16)); VT_HIP(hipMalloc(&B.nd, n * 16)); VT_HIP(hipMalloc(&B.thr, n * 16));
        VT_HIP(hipMalloc(&B.rad, n * 16)); VT_HIP(hipMalloc(&B.dir, n * 16)); VT_HIP(hipMalloc(&B.fac, n * 16));
        VT_HIP(hipMalloc(&B.meta, n * 16)); VT_HIP(hipMalloc(&B.rng, n * 8));
    }
    /* queues and hits: two entries (shadow and continuation ray) per path */
    VT_HIP(hipMalloc(&P.hit, 2 * n * 8));
    VT_HIP(hipMalloc(&P.smp, n * 16));
    VT_HIP(hipMalloc(&P.qa0, 2 * n * 16)); VT_HIP(hipMalloc(&P.qb0, 2 * n * 8)); VT_HIP(hipMalloc(&P.qt0, 2 * n * 4));
    VT_HIP(hipMalloc(&P.qa1, 2 * n * 16)); VT_HIP(hipMalloc(&P.qb1, 2 * n * 8)); VT_HIP(hipMalloc(&P.qt1, 2 * n * 4));
    VT_HIP(hipMalloc(&P.cnt, 16));
    P.cap = slots;
    alloc = slots;
    return VT_OK;
}

int ensure_wavefront(vt_context* c, uint64_t slots) { return alloc_paths(c->wf, c->wf_alloc, slots); }

}  // namespace

extern "C" {

const char* vt_last_error(void) { return g_err.c_str(); }

int vt_create(int device, vt_context** out)
{
    return no_throw([&]() -> int {
        if (!out) return fail(VT_ERR_ARG, "null output pointer");
        int n = 0;
        VT_HIP(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) return fail(VT_ERR_ARG, "no such device");
        if (device >= VT_MAX_DEVICES) return fail(VT_ERR_LIMIT, "device index beyond the library's table");
        VT_HIP(hipSetDevice(device));
        if (int r = noalloc_lookup(device)) return r;
        vt_context* c = new vt_context();
        c->device = device;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete c; return fail(VT_ERR_HIP, "device query failed"); }
        c->cus = prop.multiProcessorCount;
        c->lds_limit = prop.sharedMemPerBlock;
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
            hipMalloc(&c->d_tile, 4) != hipSuccess || hipMalloc(&c->d_stats, 6 * sizeof(unsigned long long)) != hipSuccess) {
            vt_destroy(c);
            return fail(VT_ERR_HIP, "context resource creation failed");
        }
        /* defaults: pathTracer.fs uniform defaults and Renderer ctor (renderer.cpp:41-65) */
        c->env.use_image = 0;
        const float top[3] = { 153.0f / 255 * 2, 187.0f / 255 * 2, 201.0f / 255 * 2 };
        const float bot[3] = { 77.0f / 255, 64.0f / 255, 50.0f / 255 };
        for (int k = 0; k < 3; ++k) { c->env.top[k] = top[k]; c->env.bottom[k] = bot[k]; }
        c->st.max_bounces = 1;
        c->st.exposure = 1.0f;
        c->st.gamma = 2.2f;
        c->st.selected_voxel[0] = c->st.selected_voxel[1] = c->st.selected_voxel[2] = INT32_MIN;
        c->st.wireframe_opacity = 0.0f;
        c->st.wireframe_thickness = 0.01f;
        *out = c;
        return VT_OK;
    });
}

static void batch_release(vt_batch* b);
static void cache_release(vt_cache* k);

void vt_destroy(vt_context* c)
{
    if (!c) return;
    while (!c->caches.empty()) cache_release(c->caches.back());
    while (!c->batches.empty()) batch_release(c->batches.back());
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dfree(c->d_bricks); dfree(c->d_occ64); dfree(c->d_bits); dfree(c->d_occp); dfree(c->d_bitsp); dfree(c->d_occr); dfree(c->d_emissive); dfree(c->d_lights); dfree(c->d_recs);
    dfree(c->d_img); dfree(c->d_quad); dfree(c->d_cdfu); dfree(c->d_cdfv);
    dfree(c->d_accum); dfree(c->d_sum64); dfree(c->d_tile); dfree(c->d_stats);
    free_wavefront(c);
    free_paths(c->wf2, c->wf2_alloc);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_half) (void)hipEventDestroy(c->ev_half);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream2) (void)hipStreamDestroy(c->stream2);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipEvent_t e : c->lt_ev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int vt_synchronize(vt_context* c)
{
    return no_throw([&]() -> int {
        if (!c) return fail(VT_ERR_ARG, "null context");
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_device_cus(vt_context* c) { return c ? c->cus : 0; }

int vt_set_volume(vt_context* c, int32_t nx, int32_t ny, int32_t nz, const int32_t* offsets, const float* mat_data,
                  int32_t mat_size, const int32_t* emissive, int32_t n_emissive, const float* bmin, const float* bmax)
{
    return no_throw([&]() -> int {
        if (!c || !offsets) return fail(VT_ERR_ARG, "null argument");
        VT_HIP(hipSetDevice(c->device));
        int r = common_volume_setup(c, nx, ny, nz, mat_data, mat_size, bmin, bmax);
        if (r != VT_OK) return r;
        const int64_t nvox = (int64_t)nx * ny * nz;
        /* distinct occupied offsets -> slots, in increasing offset order */
        std::vector<uint8_t> used((size_t)std::max(mat_size, 1), 0);
        std::vector<uint8_t> occ((size_t)nvox, 0);
        for (int64_t i = 0; i < nvox; ++i) {
            const int32_t o = offsets[i];
            if (o < 0) continue;
            if (o >= mat_size) return fail(VT_ERR_ARG, "material offset " + std::to_string(o) + " outside material data");
            used[(size_t)o] = 1;
            occ[(size_t)i] = 1;
        }
        std::vector<uint8_t> lut((size_t)std::max(mat_size, 1), 0);
        int32_t ns = 0;
        for (int32_t o = 0; o < mat_size; ++o) {
            if (!used[(size_t)o]) continue;
            if (ns == 255) return fail(VT_ERR_LIMIT, "more than 255 distinct material offsets in the grid");
            ++ns;
            lut[(size_t)o] = (uint8_t)ns;
            c->slot_off[ns] = o;
        }
        c->n_slots = ns;
        r = check_emissive(emissive, n_emissive, nvox, occ);
        if (r != VT_OK) return r;
        DevTmp<int32_t> d_dense;
        DevTmp<uint8_t> d_lut;
        DevTmp<uint32_t> d_bad;
        VT_HIP(hipMalloc(&d_dense.p, (size_t)nvox * 4u));
        VT_HIP(hipMalloc(&d_lut.p, lut.size()));
        VT_HIP(hipMalloc(&d_bad.p, 4));
        VT_HIP(hipMemcpyAsync(d_dense.p, offsets, (size_t)nvox * 4u, hipMemcpyHostToDevice, c->stream));
        VT_HIP(hipMemcpyAsync(d_lut.p, lut.data(), lut.size(), hipMemcpyHostToDevice, c->stream));
        VT_HIP(vt::launch_fill_u32(d_bad.p, 1, 0u, c->stream));
        VT_HIP(vt::launch_brick_from_offsets(c->dims, d_dense.p, d_lut.p, mat_size, c->d_bricks, d_bad.p, c->stream));
        uint32_t bad = 0;
        VT_HIP(hipMemcpyAsync(&bad, d_bad.p, 4, hipMemcpyDeviceToHost, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        if (bad) return fail(VT_ERR_ARG, "grid conversion found an unmapped material offset");
        return finish_volume(c, emissive, n_emissive, occ, true);
    });
}

int vt_set_volume_u8(vt_context* c, int32_t nx, int32_t ny, int32_t nz, const uint8_t* ids, const int32_t* id_offset,
                     const float* mat_data, int32_t mat_size, const int32_t* emissive, int32_t n_emissive,
                     const float* bmin, const float* bmax)
{
    return no_throw([&]() -> int {
        if (!c || !ids || !id_offset) return fail(VT_ERR_ARG, "null argument");
        VT_HIP(hipSetDevice(c->device));
        int r = common_volume_setup(c, nx, ny, nz, mat_data, mat_size, bmin, bmax);
        if (r != VT_OK) return r;
        uint8_t remap[256];
        int32_t ns = 0;
        for (int id = 0; id < 256; ++id) {
            remap[id] = 0;
            if (id_offset[id] < 0) continue;
            if (id_offset[id] >= mat_size) return fail(VT_ERR_ARG, "id " + std::to_string(id) + " offset outside material data");
            if (ns == 255) return fail(VT_ERR_LIMIT, "more than 255 occupied ids");
            ++ns;
            remap[id] = (uint8_t)ns;
            c->slot_off[ns] = id_offset[id];
        }
        c->n_slots = ns;
        const int64_t nvox = (int64_t)nx * ny * nz;
        std::vector<uint8_t> occ;
        if (n_emissive > 0) {
            occ.assign((size_t)nvox, 0);
            for (int64_t i = 0; i < nvox; ++i) occ[(size_t)i] = remap[ids[i]] != 0;
        }
        r = check_emissive(emissive, n_emissive, nvox, occ);
        if (r != VT_OK) return r;
        DevTmp<uint8_t> d_dense;
        DevTmp<uint8_t> d_remap;
        VT_HIP(hipMalloc(&d_dense.p, (size_t)nvox));
        VT_HIP(hipMalloc(&d_remap.p, 256));
        VT_HIP(hipMemcpyAsync(d_dense.p, ids, (size_t)nvox, hipMemcpyHostToDevice, c->stream));
        VT_HIP(hipMemcpyAsync(d_remap.p, remap, 256, hipMemcpyHostToDevice, c->stream));
        VT_HIP(vt::launch_brick_from_ids(c->dims, d_dense.p, d_remap.p, c->d_bricks, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return finish_volume(c, emissive, n_emissive, occ, false);
    });
}

int vt_get_bounds(vt_context* c, float* bmin, float* bmax, float* vsz)
{
    return no_throw([&]() -> int {
        if (!c || !c->has_volume) return fail(VT_ERR_STATE, "no volume");
        for (int k = 0; k < 3; ++k) {
            if (bmin) bmin[k] = c->bmin[k];
            if (bmax) bmax[k] = c->bmax[k];
            if (vsz) vsz[k] = c->vs[k];
        }
        return VT_OK;
    });
}

int vt_read_volume(vt_context* c, int32_t* offsets)
{
    return no_throw([&]() -> int {
        if (!c || !offsets) return fail(VT_ERR_ARG, "null argument");
        if (!c->has_volume) return fail(VT_ERR_STATE, "no volume");
        VT_HIP(hipSetDevice(c->device));
        const int64_t nvox = (int64_t)c->dims.nx * c->dims.ny * c->dims.nz;
        int32_t table[256];
        table[0] = -1;
        for (int s = 1; s < 256; ++s) table[s] = c->slot_off[s];
        DevTmp<int32_t> d_dense;
        DevTmp<int32_t> d_table;
        VT_HIP(hipMalloc(&d_dense.p, (size_t)nvox * 4u));
        VT_HIP(hipMalloc(&d_table.p, sizeof(table)));
        VT_HIP(hipMemcpyAsync(d_table.p, table, sizeof(table), hipMemcpyHostToDevice, c->stream));
        VT_HIP(vt::launch_dense_from_bricks(c->dims, c->d_bricks, d_table.p, d_dense.p, c->stream));
        VT_HIP(hipMemcpyAsync(offsets, d_dense.p, (size_t)nvox * 4u, hipMemcpyDeviceToHost, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_update_material(vt_context* c, int32_t offset, const float* values, int32_t count)
{
    return no_throw([&]() -> int {
        if (!c || !values || count < 0) return fail(VT_ERR_ARG, "bad argument");
        if (offset < 0 || (int64_t)offset + count > (int64_t)c->mat.size())
            return fail(VT_ERR_ARG, "material update outside the material data");
        for (int32_t i = 0; i < count; ++i) c->mat[(size_t)offset + i] = values[i];
        VT_HIP(hipSetDevice(c->device));
        return upload_records(c);
    });
}

int vt_read_materials(vt_context* c, float* mat_data, int32_t capacity, int32_t* size)
{
    return no_throw([&]() -> int {
        if (!c) return fail(VT_ERR_ARG, "null context");
        if (size) *size = (int32_t)c->mat.size();
        if (mat_data) std::memcpy(mat_data, c->mat.data(), sizeof(float) * std::min<size_t>(c->mat.size(), (size_t)std::max(capacity, 0)));
        return VT_OK;
    });
}

int vt_set_camera(vt_context* c, const vt_camera* cam)
{
    return no_throw([&]() -> int {
        if (!c || !cam) return fail(VT_ERR_ARG, "null argument");
        if (cam->lens_model < 0 || cam->lens_model > 2) return fail(VT_ERR_ARG, "lens_model must be 0, 1 or 2");
        c->cam = *cam;
        return VT_OK;
    });
}

int vt_set_environment(vt_context* c, const vt_environment* env)
{
    return no_throw([&]() -> int {
        if (!c || !env) return fail(VT_ERR_ARG, "null argument");
        vt_environment e = *env;
        if (e.use_image) {
            if (e.image_w <= 0 || e.image_h <= 0 || !e.image_rgb || e.cdf_u_w < 2 || e.cdf_u_h < 1 || !e.cdf_u ||
                e.cdf_v_n < 2 || !e.cdf_v)
                return fail(VT_ERR_ARG, "image environment needs an image and its CDF tables");
            if (e.cdf_u_h != e.cdf_v_n - 1) return fail(VT_ERR_ARG, "CDF U height must equal CDF V size - 1");
            if (e.cdf_v_n > 8193) return fail(VT_ERR_LIMIT, "CDF V larger than 8193 entries");
            if ((int64_t)e.image_w * e.image_h > (1ll << 28) || (int64_t)e.cdf_u_w * e.cdf_u_h > (1ll << 28))
                return fail(VT_ERR_LIMIT, "environment image or CDF U larger than 2^28 entries");
        }
        VT_HIP(hipSetDevice(c->device));
        /* The new tables are built in buffers of their own and replace the
         * context's only once every step has succeeded: a rejected or failed
         * call leaves the previous environment in use, whole. */
        float* img = nullptr;
        float4* quad = nullptr;
        float* cdfu = nullptr;
        float* cdfv = nullptr;
        if (e.use_image) {
            const size_t ni = (size_t)e.image_w * e.image_h * 3, nu = (size_t)e.cdf_u_w * e.cdf_u_h;
            hipError_t he = hipMalloc(&img, ni * 4);
            if (he == hipSuccess) he = hipMalloc(&cdfu, nu * 4);
            if (he == hipSuccess) he = hipMalloc(&cdfv, (size_t)e.cdf_v_n * 4);
            if (he == hipSuccess) he = hipMalloc(&quad, (size_t)(e.image_w + 1) * (size_t)(e.image_h + 1) * 64);
            if (he == hipSuccess) he = hipMemcpyAsync(img, e.image_rgb, ni * 4, hipMemcpyHostToDevice, c->stream);
            if (he == hipSuccess) he = hipMemcpyAsync(cdfu, e.cdf_u, nu * 4, hipMemcpyHostToDevice, c->stream);
            if (he == hipSuccess)
                he = hipMemcpyAsync(cdfv, e.cdf_v, (size_t)e.cdf_v_n * 4, hipMemcpyHostToDevice, c->stream);
            if (he == hipSuccess) he = vt::launch_env_quads(img, e.image_w, e.image_h, quad, c->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
            if (he != hipSuccess) {
                (void)hipGetLastError();
                (void)hipStreamSynchronize(c->stream);
                dfree(img); dfree(cdfu); dfree(cdfv); dfree(quad);
                return fail(VT_ERR_HIP, std::string("environment upload failed: ") + hipGetErrorString(he));
            }
        }
        /* earlier renders may still read the old tables */
        hipError_t hs = hipStreamSynchronize(c->stream);
        if (hs == hipSuccess && c->stream2) hs = hipStreamSynchronize(c->stream2);
        if (hs != hipSuccess) {
            (void)hipGetLastError();
            dfree(img); dfree(cdfu); dfree(cdfv); dfree(quad);
            return fail(VT_ERR_HIP, std::string("environment upload: waiting for earlier renders failed: ") +
                                        hipGetErrorString(hs));
        }
        dfree(c->d_img); dfree(c->d_quad); dfree(c->d_cdfu); dfree(c->d_cdfv);
        c->d_img = img; c->d_quad = quad; c->d_cdfu = cdfu; c->d_cdfv = cdfv;
        e.image_rgb = nullptr; e.cdf_u = nullptr; e.cdf_v = nullptr;
        c->env = e;
        return VT_OK;
    });
}

int vt_set_settings(vt_context* c, const vt_settings* st)
{
    return no_throw([&]() -> int {
        if (!c || !st) return fail(VT_ERR_ARG, "null argument");
        if (st->max_bounces < 0) return fail(VT_ERR_ARG, "max_bounces must be >= 0");
        c->st = *st;
        return VT_OK;
    });
}

int vt_set_frame(vt_context* c, int32_t w, int32_t h, int32_t fp64)
{
    return no_throw([&]() -> int {
        if (!c) return fail(VT_ERR_ARG, "null context");
        if (w <= 0 || h <= 0 || (int64_t)w * h > (1ll << 28)) return fail(VT_ERR_ARG, "bad frame size");
        VT_HIP(hipSetDevice(c->device));
        VT_HIP(hipStreamSynchronize(c->stream));
        dfree(c->d_accum); dfree(c->d_sum64);
        c->W = w; c->H = h;
        VT_HIP(hipMalloc(&c->d_accum, (size_t)w * h * sizeof(float4)));
        if (fp64) VT_HIP(hipMalloc(&c->d_sum64, (size_t)w * h * 3 * sizeof(double)));
        return vt_reset_accum(c);
    });
}

int vt_reset_accum(vt_context* c)
{
    return no_throw([&]() -> int {
        if (!c || !c->d_accum) return fail(VT_ERR_STATE, "no frame");
        VT_HIP(hipSetDevice(c->device));
        VT_HIP(vt::launch_fill_u32((uint32_t*)c->d_accum, (uint64_t)c->W * c->H * 4, 0u, c->stream));
        if (c->d_sum64) VT_HIP(vt::launch_fill_u32((uint32_t*)c->d_sum64, (uint64_t)c->W * c->H * 6, 0u, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_set_stats(vt_context* c, int32_t enabled)
{
    return no_throw([&]() -> int {
        if (!c) return fail(VT_ERR_ARG, "null context");
        c->stats_enabled = enabled != 0;
        return VT_OK;
    });
}

int vt_set_overlap(vt_context* c, int32_t streams)
{
    return no_throw([&]() -> int {
        if (!c) return fail(VT_ERR_ARG, "null context");
        if (streams != 1 && streams != 2) return fail(VT_ERR_ARG, "streams must be 1 or 2");
        c->overlap = streams;
        if (streams == 1) {
            VT_HIP(hipSetDevice(c->device));
            VT_HIP(hipStreamSynchronize(c->stream));
            free_paths(c->wf2, c->wf2_alloc);
        }
        return VT_OK;
    });
}

int vt_set_launch_timing(vt_context* c, int32_t enabled)
{
    return no_throw([&]() -> int {
        if (!c) return fail(VT_ERR_ARG, "null context");
        VT_HIP(hipSetDevice(c->device));
        if (enabled && c->lt_ev.empty()) {
            c->lt_ev.assign(2 * VT_TIMED_LAUNCHES, nullptr);
            for (hipEvent_t& e : c->lt_ev) VT_HIP(hipEventCreate(&e));
        }
        c->lt_on = enabled != 0;
        c->lt_used = 0;
        return VT_OK;
    });
}

int vt_get_launch_timing(vt_context* c, double* traverse_ms, uint32_t* launches, double* union_ms)
{
    return no_throw([&]() -> int {
        if (!c || !traverse_ms || !launches) return fail(VT_ERR_ARG, "null argument");
        VT_HIP(hipSetDevice(c->device));
        VT_HIP(hipStreamSynchronize(c->stream));
        double sum = 0.0;
        std::vector<std::pair<float, float>> iv;
        for (int k = 0; k < c->lt_used; ++k) {
            float ms = 0.0f, t0 = 0.0f, t1 = 0.0f;
            VT_HIP(hipEventElapsedTime(&ms, c->lt_ev[2 * k], c->lt_ev[2 * k + 1]));
            sum += ms;
            if (union_ms && c->timed) {
                VT_HIP(hipEventElapsedTime(&t0, c->ev0, c->lt_ev[2 * k]));
                VT_HIP(hipEventElapsedTime(&t1, c->ev0, c->lt_ev[2 * k + 1]));
                iv.emplace_back(t0, t1);
            }
        }
        *traverse_ms = sum;
        *launches = (uint32_t)c->lt_used;
        if (union_ms) {
            /* time during which at least one traversal launch ran (launches of
             * the two streams overlap) */
            std::sort(iv.begin(), iv.end());
            double u = 0.0, lo = 0.0, hi = -1.0;
            for (const auto& p : iv) {
                if (p.first > hi) {
                    if (hi > lo) u += hi - lo;
                    lo = p.first;
                    hi = p.second;
                } else if (p.second > hi) {
                    hi = p.second;
                }
            }
            if (hi > lo) u += hi - lo;
            *union_ms = iv.empty() ? sum : u;
        }
        return VT_OK;
    });
}

int vt_set_render_mode(vt_context* c, int32_t mode, uint64_t max_paths)
{
    return no_throw([&]() -> int {
        if (!c) return fail(VT_ERR_ARG, "null context");
        if (mode != 0 && mode != 1) return fail(VT_ERR_ARG, "mode must be 0 (megakernel) or 1 (wavefront)");
        if (max_paths > (1ull << 30)) return fail(VT_ERR_LIMIT, "max_paths above 2^30");
        c->render_mode = mode;
        c->max_paths = max_paths ? max_paths : VT_DEFAULT_MAX_PATHS;
        return VT_OK;
    });
}

int vt_set_traversal_tuning(vt_context* c, int32_t dda_budget, int32_t refill_min, int32_t refill_chunk)
{
    return no_throw([&]() -> int {
        if (!c) return fail(VT_ERR_ARG, "null context");
        if (dda_budget < 0 || dda_budget > 4096) return fail(VT_ERR_ARG, "dda_budget must be in [0, 4096]");
        if (refill_min < 0 || refill_min > 64) return fail(VT_ERR_ARG, "refill_min must be in [0, 64]");
        if (refill_chunk != -1 && refill_chunk != 0 && (refill_chunk < 64 || refill_chunk > 65536))
            return fail(VT_ERR_ARG, "refill_chunk must be -1, 0 or in [64, 65536]");
        c->dda_budget = dda_budget ? dda_budget : VT_DEFAULT_DDA_BUDGET;
        c->refill_min = refill_min ? refill_min : VT_DEFAULT_REFILL_MIN;
        c->refill_chunk = refill_chunk == 0 ? VT_DEFAULT_REFILL_CHUNK : (refill_chunk < 0 ? 0 : refill_chunk);
        return VT_OK;
    });
}

int vt_render(vt_context* c, uint64_t seed, uint32_t n0, uint32_t count, uint32_t rng_frame_max)
{
    return no_throw([&]() -> int {
        if (!c) return fail(VT_ERR_ARG, "null context");
        if (!c->has_volume) return fail(VT_ERR_STATE, "no volume");
        if (!c->d_accum) return fail(VT_ERR_STATE, "no frame");
        if (count == 0) return VT_OK;
        VT_HIP(hipSetDevice(c->device));
        vt::TraceArgs a = trace_args(c);
        a.seed = seed; a.n0 = n0; a.count = count; a.rng_frame_max = rng_frame_max;
        a.stats = c->stats_enabled ? c->d_stats : nullptr;
        VT_HIP(vt::launch_reset_counter(c->d_tile, a.stats, c->stream));
        if (c->render_mode == 1) {
            const uint64_t total = (uint64_t)c->W * c->H * count;
            c->lt_used = 0;
            const vt::LaunchTimer lt = { c->lt_ev.data(), (int)c->lt_ev.size() / 2, &c->lt_used };
            const vt::LaunchTimer* ltp = c->lt_on ? &lt : nullptr;
            if (c->overlap > 1 && count >= 2 && total >= VT_OVERLAP_MIN_PATHS) {
                /* passes [n0, n0 + c0) on the context stream, the rest on the
                 * second stream; the second half accumulates after the first */
                const uint32_t c0 = count / 2, c1 = count - c0;
                const uint64_t WH = (uint64_t)c->W * c->H, half_cap = std::max<uint64_t>(c->max_paths / 2, WH);
                if (!c->stream2) {
                    VT_HIP(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
                    VT_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
                    VT_HIP(hipEventCreateWithFlags(&c->ev_half, hipEventDisableTiming));
                    VT_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
                }
                int r = ensure_wavefront(c, std::min<uint64_t>(WH * c0, half_cap));
                if (r == VT_OK) r = alloc_paths(c->wf2, c->wf2_alloc, std::min<uint64_t>(WH * c1, half_cap));
                if (r != VT_OK) return r;
                vt::TraceArgs a1 = a;
                a.count = c0;
                a1.n0 = n0 + c0;
                a1.count = c1;
                VT_HIP(hipEventRecord(c->ev0, c->stream));
                VT_HIP(hipEventRecord(c->ev_fork, c->stream));
                VT_HIP(hipStreamWaitEvent(c->stream2, c->ev_fork, 0));
                VT_HIP(vt::launch_render_wavefront(a, c->wf, c->cus, c->stream, c->stats_enabled, ltp));
                VT_HIP(hipEventRecord(c->ev_half, c->stream));
                VT_HIP(vt::launch_render_wavefront(a1, c->wf2, c->cus, c->stream2, c->stats_enabled, ltp, c->ev_half));
                VT_HIP(hipEventRecord(c->ev_join, c->stream2));
                VT_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
                VT_HIP(hipEventRecord(c->ev1, c->stream));
            } else {
                int r = ensure_wavefront(c, std::min<uint64_t>(total, c->max_paths));
                if (r != VT_OK) return r;
                VT_HIP(hipEventRecord(c->ev0, c->stream));
                VT_HIP(vt::launch_render_wavefront(a, c->wf, c->cus, c->stream, c->stats_enabled, ltp));
                VT_HIP(hipEventRecord(c->ev1, c->stream));
            }
        } else {
            const int per_cu = vt::render_blocks_per_cu(a, c->stats_enabled);
            const int grid = std::max(1, c->cus * per_cu);
            VT_HIP(hipEventRecord(c->ev0, c->stream));
            VT_HIP(vt::launch_render(a, grid, c->stream, c->stats_enabled));
            VT_HIP(hipEventRecord(c->ev1, c->stream));
        }
        c->timed = true;
        c->stats_valid = c->stats_enabled;
        return VT_OK;
    });
}

/* ---- batches ---------------------------------------------------------------- */
struct vt_batch {
    std::vector<vt_context*> ctxs;
    vt_context* lead = nullptr;
    std::vector<vt::TraceArgs> h_tab;
    std::vector<uint64_t> h_base;         /* path prefix (n + 1), then pixel prefix (n + 1) */
    vt::TraceArgs* d_tab = nullptr;
    uint64_t* d_base = nullptr;
    unsigned long long* d_stats = nullptr;
    vt::WfPaths wf{};
    uint64_t wf_alloc = 0;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> graph_key;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t fork = nullptr;            /* start of the solo scenes' branches inside the graph */
    std::vector<hipEvent_t> join;         /* per scene: end of its own launches on its context stream */
    bool timed = false, stats_valid = false;
    /* tuning (vt_batch_set_tuning, vt_autotune_batch); 0 = the default */
    uint64_t solo_paths = VT_BATCH_SOLO_PATHS;
    uint64_t max_paths = 0;               /* 0: each launch's own context's capacity */
    int32_t dda_budget = 0, refill_min = 0, refill_chunk = 0;   /* merged launch; 0: the lead context's */
};

int vt_batch_create(vt_context* const* ctxs, int32_t n, vt_batch** out)
{
    return no_throw([&]() -> int {
        if (!ctxs || !out || n <= 0 || n > 65536) return fail(VT_ERR_ARG, "need 1..65536 contexts");
        for (int32_t i = 0; i < n; ++i) {
            if (!ctxs[i]) return fail(VT_ERR_ARG, "null context in batch");
            if (ctxs[i]->device != ctxs[0]->device) return fail(VT_ERR_ARG, "batch contexts must share one device");
        }
        {
            /* one scene per context: a context listed twice would accumulate
             * two scenes' samples into one frame in the same launch */
            std::vector<const vt_context*> sorted(ctxs, ctxs + n);
            std::sort(sorted.begin(), sorted.end());
            if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
                return fail(VT_ERR_ARG, "a context appears more than once in the batch");
        }
        VT_HIP(hipSetDevice(ctxs[0]->device));
        vt_batch* b = new vt_batch();
        b->ctxs.assign(ctxs, ctxs + n);
        b->lead = ctxs[0];
        b->h_tab.resize((size_t)n);
        b->h_base.resize(2 * ((size_t)n + 1));
        for (vt_context* c : b->ctxs)
            if (std::find(c->batches.begin(), c->batches.end(), b) == c->batches.end()) c->batches.push_back(b);
        if (hipMalloc(&b->d_tab, sizeof(vt::TraceArgs) * (size_t)n) != hipSuccess ||
            hipMalloc(&b->d_base, sizeof(uint64_t) * b->h_base.size()) != hipSuccess ||
            hipMalloc(&b->d_stats, 6 * sizeof(unsigned long long)) != hipSuccess ||
            hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            vt_batch_destroy(b);
            return fail(VT_ERR_HIP, "batch allocation failed");
        }
        *out = b;
        return VT_OK;
    });
}

/* Free the batch's device resources and detach it from its contexts. */
static void batch_release(vt_batch* b)
{
    if (!b->lead) return;
    (void)hipSetDevice(b->lead->device);
    if (b->lead->stream) (void)hipStreamSynchronize(b->lead->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    dfree(b->d_tab); dfree(b->d_base); dfree(b->d_stats);
    free_paths(b->wf, b->wf_alloc);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->fork) (void)hipEventDestroy(b->fork);
    b->ev0 = b->ev1 = b->fork = nullptr;
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    b->join.clear();
    (void)hipGetLastError();
    for (vt_context* c : b->ctxs) c->batches.erase(std::remove(c->batches.begin(), c->batches.end(), b), c->batches.end());
    b->ctxs.clear();
    b->lead = nullptr;
}

int vt_batch_set_tuning(vt_batch* b, uint64_t solo_paths, uint64_t max_paths, int32_t dda_budget, int32_t refill_min,
                        int32_t refill_chunk)
{
    return no_throw([&]() -> int {
        if (!b) return fail(VT_ERR_ARG, "null batch");
        if (!b->lead) return fail(VT_ERR_STATE, "a context of the batch was destroyed");
        if (max_paths != 0 && (max_paths < 4096 || max_paths > (1ull << 28)))
            return fail(VT_ERR_ARG, "max_paths must be 0 or in [4096, 2^28]");
        if (dda_budget < 0 || dda_budget > 4096) return fail(VT_ERR_ARG, "dda_budget must be in [0, 4096]");
        if (refill_min < 0 || refill_min > 64) return fail(VT_ERR_ARG, "refill_min must be in [0, 64]");
        if (refill_chunk != -1 && refill_chunk != 0 && (refill_chunk < 64 || refill_chunk > 65536))
            return fail(VT_ERR_ARG, "refill_chunk must be -1, 0 or in [64, 65536]");
        b->solo_paths = solo_paths ? solo_paths : VT_BATCH_SOLO_PATHS;
        b->max_paths = max_paths;
        b->dda_budget = dda_budget;
        b->refill_min = refill_min;
        b->refill_chunk = refill_chunk;
        return VT_OK;
    });
}

int vt_batch_get_tuning(vt_batch* b, vt_batch_tuning* out)
{
    return no_throw([&]() -> int {
        if (!b || !out) return fail(VT_ERR_ARG, "null argument");
        if (!b->lead) return fail(VT_ERR_STATE, "a context of the batch was destroyed");
        std::memset(out, 0, sizeof(*out));
        out->solo_paths = b->solo_paths;
        out->max_paths = b->max_paths ? b->max_paths : b->lead->max_paths;
        out->dda_budget = b->dda_budget ? b->dda_budget : b->lead->dda_budget;
        out->refill_min = b->refill_min ? b->refill_min : b->lead->refill_min;
        out->refill_chunk = b->refill_chunk ? (b->refill_chunk < 0 ? -1 : b->refill_chunk) : b->lead->refill_chunk;
        return VT_OK;
    });
}

int vt_autotune_batch(vt_batch* b, const uint64_t* seeds, const uint32_t* n0, const uint32_t* counts,
                      uint32_t rng_frame_max, vt_batch_tuning* out)
{
    return no_throw([&]() -> int {
        if (!b || !seeds || !counts || !out) return fail(VT_ERR_ARG, "null argument");
        if (!b->lead) return fail(VT_ERR_STATE, "a context of the batch was destroyed");
        VT_HIP(hipSetDevice(b->lead->device));
        /* the scenes' accumulations, restored at the end */
        struct Saved { vt_context* c; void* acc; void* sum; size_t acc_b, sum_b; };
        std::vector<Saved> saved;
        int rc = VT_OK;
        auto release = [&]() {
            for (Saved& v : saved) { dfree(v.acc); dfree(v.sum); }
        };
        for (vt_context* c : b->ctxs) {
            if (!c->has_volume || !c->d_accum) { release(); return fail(VT_ERR_STATE, "batch scene without a volume or frame"); }
            Saved v = { c, nullptr, nullptr, (size_t)c->W * c->H * sizeof(float4),
                        c->d_sum64 ? (size_t)c->W * c->H * 3 * sizeof(double) : 0 };
            hipError_t he = hipStreamSynchronize(c->stream);
            if (he == hipSuccess) he = hipMalloc(&v.acc, v.acc_b);
            if (he == hipSuccess && v.sum_b) he = hipMalloc(&v.sum, v.sum_b);
            if (he == hipSuccess) he = hipMemcpyAsync(v.acc, c->d_accum, v.acc_b, hipMemcpyDeviceToDevice, c->stream);
            if (he == hipSuccess && v.sum_b) he = hipMemcpyAsync(v.sum, c->d_sum64, v.sum_b, hipMemcpyDeviceToDevice, c->stream);
            saved.push_back(v);
            if (he != hipSuccess) {
                (void)hipGetLastError();
                for (vt_context* d : b->ctxs) (void)hipStreamSynchronize(d->stream);
                release();
                return fail(VT_ERR_HIP, std::string("batch autotune: saving the accumulations failed: ") + hipGetErrorString(he));
            }
        }
        const uint64_t t_solo = b->solo_paths, t_max = b->max_paths;
        const int32_t t_bud = b->dda_budget, t_ref = b->refill_min, t_chk = b->refill_chunk;
        vt_batch_tuning best;
        std::memset(&best, 0, sizeof(best));
        best.solo_paths = t_solo; best.max_paths = t_max;
        best.dda_budget = t_bud; best.refill_min = t_ref; best.refill_chunk = t_chk;
        best.ms = 1e30;
        int ncand = 0;
        /* one graph recording (untimed), then the best of two replays */
        auto time_it = [&](const vt_batch_tuning& t, double& ms) -> int {
            b->solo_paths = t.solo_paths; b->max_paths = t.max_paths;
            b->dda_budget = t.dda_budget; b->refill_min = t.refill_min; b->refill_chunk = t.refill_chunk;
            int r = vt_batch_render(b, seeds, n0, counts, rng_frame_max, 1);
            ms = 1e30;
            for (int rep = 0; rep < 2 && r == VT_OK; ++rep) {
                r = vt_batch_render(b, seeds, n0, counts, rng_frame_max, 1);
                float e = 0.0f;
                if (r == VT_OK && hipEventElapsedTime(&e, b->ev0, b->ev1) == hipSuccess) ms = std::min(ms, (double)e);
            }
            ++ncand;
            return r;
        };
        /* coordinate descent: solo threshold, path capacity, then the merged
         * launch's step budget, refill threshold and queue chunk */
        const uint64_t solo_c[] = { 1ull << 20, 1ull << 21, 1ull << 22, 1ull << 23, 1ull << 24, ~0ull };
        const uint64_t max_c[] = { 4ull << 20, 8ull << 20, 16ull << 20 };
        const int32_t bud_c[] = { 16, 32, 64 }, ref_c[] = { 8, 16 }, chk_c[] = { 128, 256, 512 };
        for (int stage = 0; stage < 5 && rc == VT_OK; ++stage) {
            const int nc = stage == 0 ? 6 : stage == 1 ? 3 : stage == 2 ? 3 : stage == 3 ? 2 : 3;
            vt_batch_tuning stage_best = best;
            for (int k = 0; k < nc && rc == VT_OK; ++k) {
                vt_batch_tuning t = best;
                if (stage == 0) t.solo_paths = solo_c[k];
                if (stage == 1) t.max_paths = max_c[k];
                if (stage == 2) t.dda_budget = bud_c[k];
                if (stage == 3) t.refill_min = ref_c[k];
                if (stage == 4) t.refill_chunk = chk_c[k];
                double ms = 0.0;
                rc = time_it(t, ms);
                if (rc == VT_OK && ms < stage_best.ms) { stage_best = t; stage_best.ms = ms; }
            }
            best = stage_best;
        }
        if (rc == VT_OK) {
            b->solo_paths = best.solo_paths; b->max_paths = best.max_paths;
            b->dda_budget = best.dda_budget; b->refill_min = best.refill_min; b->refill_chunk = best.refill_chunk;
        } else {
            b->solo_paths = t_solo; b->max_paths = t_max;
            b->dda_budget = t_bud; b->refill_min = t_ref; b->refill_chunk = t_chk;
        }
        const std::string err = rc == VT_OK ? std::string() : g_err;
        /* restore every scene's accumulation */
        int rr = VT_OK;
        for (Saved& v : saved) {
            hipError_t he = hipStreamSynchronize(b->lead->stream);
            if (he == hipSuccess) he = hipMemcpyAsync(v.c->d_accum, v.acc, v.acc_b, hipMemcpyDeviceToDevice, v.c->stream);
            if (he == hipSuccess && v.sum_b)
                he = hipMemcpyAsync(v.c->d_sum64, v.sum, v.sum_b, hipMemcpyDeviceToDevice, v.c->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(v.c->stream);
            if (he != hipSuccess && rr == VT_OK) {
                (void)hipGetLastError();
                rr = fail(VT_ERR_HIP, std::string("batch autotune: restoring an accumulation failed: ") + hipGetErrorString(he));
            }
        }
        release();
        if (rc != VT_OK) return fail(rc, err);
        if (rr != VT_OK) return rr;
        const int r = vt_batch_get_tuning(b, out);
        out->ms = best.ms;
        out->candidates = ncand;
        return r;
    });
}

void vt_batch_destroy(vt_batch* b)
{
    if (!b) return;
    batch_release(b);
    delete b;
}

int vt_batch_render(vt_batch* b, const uint64_t* seeds, const uint32_t* n0, const uint32_t* counts,
                    uint32_t rng_frame_max, int32_t flags)
{
    return no_throw([&]() -> int {
        if (!b || !seeds || !counts) return fail(VT_ERR_ARG, "null argument");
        if (!b->lead) return fail(VT_ERR_STATE, "a context of the batch was destroyed");
        const size_t n = b->ctxs.size();
        vt_context* L = b->lead;
        VT_HIP(hipSetDevice(L->device));
        const bool stats = (flags & 2) != 0, graph = (flags & 1) != 0;
        /* Scenes with at least b->solo_paths paths fill the GPU by themselves:
         * they get the single-scene launches (the grid as scalars) on their own
         * context streams, concurrently with each other and with one batched
         * launch of all other scenes on the lead stream, where merging saves
         * their many small launches.  With a graph, all of it is one graph: the
         * solo scenes' launches branch off the lead stream and join it again. */
        std::vector<std::pair<size_t, vt::TraceArgs>> solo;
        std::vector<size_t> merged;
        for (size_t i = 0; i < n; ++i) {
            vt_context* c = b->ctxs[i];
            if (!c->has_volume) return fail(VT_ERR_STATE, "batch scene without a volume");
            if (!c->d_accum) return fail(VT_ERR_STATE, "batch scene without a frame");
            vt::TraceArgs a = trace_args(c);
            a.seed = seeds[i];
            a.n0 = n0 ? n0[i] : 0u;
            a.count = counts[i];
            a.rng_frame_max = rng_frame_max;
            a.stats = stats ? b->d_stats : nullptr;
            const uint64_t paths = (uint64_t)c->W * (uint64_t)c->H * counts[i];
            if (paths == 0) continue;
            if (paths >= b->solo_paths) {
                solo.emplace_back(i, a);
            } else {
                b->h_tab[merged.size()] = a;
                merged.push_back(i);
            }
        }
        const size_t m = merged.size();
        uint64_t* pb = b->h_base.data();
        uint64_t* xb = pb + m + 1;
        pb[0] = xb[0] = 0;
        int max_bounces = 0;
        for (size_t k = 0; k < m; ++k) {
            vt_context* c = b->ctxs[merged[k]];
            const uint64_t WH = (uint64_t)c->W * (uint64_t)c->H;
            pb[k + 1] = pb[k] + WH * counts[merged[k]];
            xb[k + 1] = xb[k] + WH;
            max_bounces = std::max(max_bounces, c->st.max_bounces);
        }
        const uint64_t total = pb[m], pixels = xb[m];
        if (total == 0 && solo.empty()) return VT_OK;
        /* uploads and clears still queued on the scenes' own streams */
        for (vt_context* c : b->ctxs) VT_HIP(hipStreamSynchronize(c->stream));
        if (total > (1ull << 40)) return fail(VT_ERR_LIMIT, "batch too large");
        /* the previous launch may still read the table */
        VT_HIP(hipStreamSynchronize(L->stream));
        if (m) {
            VT_HIP(hipMemcpy(b->d_tab, b->h_tab.data(), sizeof(vt::TraceArgs) * m, hipMemcpyHostToDevice));
            VT_HIP(hipMemcpy(b->d_base, b->h_base.data(), sizeof(uint64_t) * 2 * (m + 1), hipMemcpyHostToDevice));
        }
        if (m) {
            const int r = alloc_paths(b->wf, b->wf_alloc, std::min<uint64_t>(total, b->max_paths ? b->max_paths : L->max_paths));
            if (r != VT_OK) return r;
        }
        if (b->join.size() != n) {
            for (hipEvent_t e : b->join)
                if (e) (void)hipEventDestroy(e);
            b->join.assign(n, nullptr);
            for (hipEvent_t& e : b->join) VT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
        for (const auto& sa : solo) {
            vt_context* c = b->ctxs[sa.first];
            const uint64_t paths = (uint64_t)sa.second.W * (uint64_t)sa.second.H * sa.second.count;
            const int r = ensure_wavefront(c, std::min<uint64_t>(paths, b->max_paths ? b->max_paths : c->max_paths));
            if (r != VT_OK) return r;
        }
        vt::BatchTable B;
        B.tab = b->d_tab;
        B.path_base = b->d_base;
        B.pix_base = b->d_base + m + 1;
        B.n = (uint32_t)m;
        B.stats = stats ? b->d_stats : nullptr;
        B.packed = 1;
        for (size_t k = 0; k < m; ++k)
            if (!b->h_tab[k].g.occr) B.packed = 0;
        vt::TraceArgs a0 = trace_args(L);
        if (b->dda_budget) a0.dda_budget = b->dda_budget;
        if (b->refill_min) a0.refill_min = b->refill_min;
        if (b->refill_chunk) a0.refill_chunk = b->refill_chunk < 0 ? 0 : b->refill_chunk;
        if (stats) VT_HIP(vt::launch_reset_counter(L->d_tile, b->d_stats, L->stream));
        /* every launch of the batch: the solo scenes on their own streams,
         * forked from and joined back to the lead stream, and the merged launch */
        auto issue = [&](hipEvent_t fork) -> hipError_t {
            hipError_t e = hipSuccess;
            if (!solo.empty()) e = hipEventRecord(fork, L->stream);
            for (const auto& sa : solo) {
                vt_context* c = b->ctxs[sa.first];
                if (e == hipSuccess && c->stream != L->stream) e = hipStreamWaitEvent(c->stream, fork, 0);
                if (e == hipSuccess) e = vt::launch_render_wavefront(sa.second, c->wf, c->cus, c->stream, stats);
                if (e == hipSuccess) e = hipEventRecord(b->join[sa.first], c->stream);
            }
            if (e == hipSuccess && m)
                e = vt::launch_render_batch(a0, B, b->wf, total, pixels, max_bounces, L->cus, L->stream, stats);
            for (const auto& sa : solo)
                if (e == hipSuccess) e = hipStreamWaitEvent(L->stream, b->join[sa.first], 0);
            return e;
        };
        if (graph) {
            /* the graph holds every launch argument: re-recorded when any of
             * them changes (shape, tuning, path buffers, solo scenes' state) */
            std::vector<uint64_t> key = { (uint64_t)m, total, pixels, b->wf.cap, (uint64_t)max_bounces,
                                          (uint64_t)stats, (uint64_t)a0.dda_budget, (uint64_t)a0.refill_min,
                                          (uint64_t)a0.refill_chunk, (uint64_t)(uintptr_t)b->wf.s[0].o,
                                          (uint64_t)B.packed };
            key.insert(key.end(), merged.begin(), merged.end());
            for (const auto& sa : solo) {
                const vt_context* c = b->ctxs[sa.first];
                key.push_back(~(uint64_t)sa.first);
                key.push_back((uint64_t)(uintptr_t)c->wf.s[0].o);
                key.push_back(c->wf.cap);
                uint64_t w[(sizeof(vt::TraceArgs) + 7) / 8] = {};
                std::memcpy(w, &sa.second, sizeof(vt::TraceArgs));
                key.insert(key.end(), w, w + (sizeof(vt::TraceArgs) + 7) / 8);
            }
            if (!b->exec || key != b->graph_key) {
                if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
                b->graph_key.clear();
                hipGraph_t g = nullptr;
                VT_HIP(hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal));
                const hipError_t le = issue(b->fork);
                const hipError_t ce = hipStreamEndCapture(L->stream, &g);
                if (le != hipSuccess || ce != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    (void)hipGetLastError();
                    return fail(VT_ERR_HIP, std::string("batch graph capture failed: ") +
                                                hipGetErrorString(le != hipSuccess ? le : ce));
                }
                const hipError_t ie = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (ie != hipSuccess) {
                    b->exec = nullptr;
                    return fail(VT_ERR_HIP, std::string("batch graph instantiation failed: ") + hipGetErrorString(ie));
                }
                b->graph_key = key;
            }
            VT_HIP(hipEventRecord(b->ev0, L->stream));
            VT_HIP(hipGraphLaunch(b->exec, L->stream));
        } else {
            VT_HIP(hipEventRecord(b->ev0, L->stream));
            VT_HIP(issue(b->fork));
        }
        VT_HIP(hipEventRecord(b->ev1, L->stream));
        /* each context's later reads and renders run on its own stream */
        VT_HIP(hipStreamSynchronize(L->stream));
        b->timed = true;
        b->stats_valid = stats;
        return VT_OK;
    });
}

int vt_batch_get_stats(vt_batch* b, vt_stats* out)
{
    return no_throw([&]() -> int {
        if (!b || !out) return fail(VT_ERR_ARG, "null argument");
        if (!b->lead) return fail(VT_ERR_STATE, "a context of the batch was destroyed");
        VT_HIP(hipSetDevice(b->lead->device));
        VT_HIP(hipStreamSynchronize(b->lead->stream));
        std::memset(out, 0, sizeof(*out));
        if (b->stats_valid) {
            unsigned long long s[6];
            VT_HIP(hipMemcpy(s, b->d_stats, sizeof(s), hipMemcpyDeviceToHost));
            out->rays = s[0]; out->visits = s[1]; out->interactions = s[2];
            out->samples = s[3]; out->env_samples = s[4]; out->env_evals = s[5];
        }
        if (b->timed) {
            float ms = 0.0f;
            VT_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            out->kernel_ms = ms;
        }
        return VT_OK;
    });
}

/* ---- learned radiance cache ------------------------------------------------------- */
#define VT_CACHE_LOSS_SLOTS 1024

struct vt_cache {
    vt_context* ctx = nullptr;
    float* d_params = nullptr;
    float* d_grad = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    _Float16* d_pack = nullptr;
    float* d_bias = nullptr;
    float* d_tpack = nullptr;
    float* d_part = nullptr;           /* [MLP_GRAD_SPLITS][MLP_N_PARAMS] */
    float* d_loss = nullptr;           /* [VT_CACHE_LOSS_SLOTS] loss sum per batch */
    uint32_t* d_steps = nullptr;       /* Adam step count */
    uint32_t* d_nsamp = nullptr;       /* sample count of host-supplied batches */
    unsigned long long* d_tally = nullptr;   /* queries, samples over a cached render */
    /* training scratch for batches of up to 16 * tiles samples */
    uint32_t tiles = 0;
    _Float16 *d_X = nullptr, *d_H = nullptr;
    float *d_DZ = nullptr, *d_DY = nullptr;
    /* host-supplied samples */
    int64_t cap = 0;
    float *d_raw = nullptr, *d_tgt = nullptr;
    /* inference staging */
    int64_t icap = 0;
    float *d_iraw = nullptr, *d_iout = nullptr;
    /* cache rendering buffers */
    uint64_t qcap = 0;
    float *d_qraw = nullptr, *d_qout = nullptr;
    uint4* d_qmeta = nullptr;
    float4 *d_qrad = nullptr, *d_qthr = nullptr;
    uint32_t* d_counts = nullptr;      /* [0] queries, [1] samples */
    uint64_t vcap = 0;                 /* train slots x vertices */
    float* d_vrec = nullptr;
    uint32_t* d_vrec_n = nullptr;
    uint64_t vslots = 0;
    uint64_t scap = 0;
    float *d_sraw = nullptr, *d_stgt = nullptr;
};

namespace {

int cache_repack(vt_cache* c)
{
    VT_HIP(vt::launch_mlp_pack(c->d_params, c->d_pack, c->d_bias, c->d_tpack, c->ctx->stream));
    return VT_OK;
}

/* training scratch for batches of up to `batch` samples */
int cache_train_scratch(vt_cache* c, uint32_t batch)
{
    const uint32_t tiles = (batch + 15u) / 16u;
    if (tiles <= c->tiles) return VT_OK;
    dfree(c->d_X); dfree(c->d_H); dfree(c->d_DZ); dfree(c->d_DY);
    c->tiles = 0;
    const size_t t16 = (size_t)tiles * 16u;
    VT_HIP(hipMalloc(&c->d_X, t16 * vt::MLP_IN * 2));
    VT_HIP(hipMalloc(&c->d_H, t16 * 5 * vt::MLP_W * 2));
    VT_HIP(hipMalloc(&c->d_DZ, t16 * 5 * vt::MLP_W * 4));
    VT_HIP(hipMalloc(&c->d_DY, t16 * vt::MLP_OUT * 4));
    c->tiles = tiles;
    return VT_OK;
}

/* host samples -> device arrays, count -> d_nsamp */
int cache_upload_samples(vt_cache* c, int64_t n, const float* raw, const float* target)
{
    hipStream_t s = c->ctx->stream;
    if (n > c->cap) {
        dfree(c->d_raw); dfree(c->d_tgt);
        c->cap = 0;
        VT_HIP(hipMalloc(&c->d_raw, (size_t)n * vt::MLP_RAW * 4));
        VT_HIP(hipMalloc(&c->d_tgt, (size_t)n * vt::MLP_OUT * 4));
        c->cap = n;
    }
    VT_HIP(hipMemcpyAsync(c->d_raw, raw, (size_t)n * vt::MLP_RAW * 4, hipMemcpyHostToDevice, s));
    VT_HIP(hipMemcpyAsync(c->d_tgt, target, (size_t)n * vt::MLP_OUT * 4, hipMemcpyHostToDevice, s));
    VT_HIP(vt::launch_fill_u32(c->d_nsamp, 1, (uint32_t)n, s));
    return VT_OK;
}

/* One training batch: forward/backward, gradient GEMM, then the summed
 * gradient into `grad` (grad != 0, Adam state untouched) or an Adam step
 * with learning rate lr.  The batch's loss sum is added to d_loss[slot]. */
int cache_train_batch(vt_cache* c, const vt::TrainBatch& b, float* grad, float lr, int slot)
{
    hipStream_t s = c->ctx->stream;
    int r = cache_train_scratch(c, b.batch);
    if (r != VT_OK) return r;
    vt::TrainScratch S;
    S.X = c->d_X; S.H = c->d_H; S.DZ = c->d_DZ; S.DY = c->d_DY; S.part = c->d_part;
    S.tiles = c->tiles;
    S.loss = c->d_loss + slot;
    S.steps = grad ? nullptr : c->d_steps;
    VT_HIP(vt::launch_mlp_train_fb(c->d_pack, c->d_bias, c->d_tpack, b, S, c->ctx->cus, s));
    VT_HIP(vt::launch_mlp_train_grad(b, S, s));
    VT_HIP(vt::launch_mlp_train_update(b, S, grad, c->d_params, c->d_m, c->d_v, c->d_pack, c->d_bias, c->d_tpack, lr,
                                       0.9f, 0.99f, 1e-8f, s));
    return VT_OK;
}

/* gradient or Adam step over n host samples in one batch; mean loss out */
int cache_host_step(vt_cache* c, int64_t n, const float* raw, const float* target, float* grad_host, float lr,
                    float* loss)
{
    if (n > (int64_t)0x7fffffff) return fail(VT_ERR_LIMIT, "too many samples in one batch");
    hipStream_t s = c->ctx->stream;
    int r = cache_upload_samples(c, n, raw, target);
    if (r != VT_OK) return r;
    VT_HIP(vt::launch_fill_u32((uint32_t*)c->d_loss, 1, 0u, s));
    const vt::TrainBatch b = { c->d_raw, c->d_tgt, c->d_nsamp, (uint32_t)n, 0u, (uint32_t)n };
    r = cache_train_batch(c, b, grad_host ? c->d_grad : nullptr, lr, 0);
    if (r != VT_OK) return r;
    if (grad_host)
        VT_HIP(hipMemcpyAsync(grad_host, c->d_grad, (size_t)vt::MLP_N_PARAMS * 4, hipMemcpyDeviceToHost, s));
    float l = 0.0f;
    VT_HIP(hipMemcpyAsync(&l, c->d_loss, 4, hipMemcpyDeviceToHost, s));
    VT_HIP(hipStreamSynchronize(s));
    if (loss) *loss = l / (float)(vt::MLP_OUT * n);
    return VT_OK;
}

}  // namespace

int32_t vt_cache_num_params(void) { return vt::MLP_N_PARAMS; }

int vt_cache_create(vt_context* ctx, uint64_t seed, vt_cache** out)
{
    return no_throw([&]() -> int {
        if (!ctx || !out) return fail(VT_ERR_ARG, "null argument");
        VT_HIP(hipSetDevice(ctx->device));
        vt_cache* c = new vt_cache();
        c->ctx = ctx;
        ctx->caches.push_back(c);
        const size_t pb = (size_t)vt::MLP_N_PARAMS * 4;
        /* the pack layouts must be exact inverses (checked once, on the host) */
        for (int e = 0; e < vt::MLP_PACK; ++e) {
            const int p = vt::pack_source(e);
            if (p < 0) continue;
            int layer = 5;
            while (layer > 0 && p < vt::mlp_w_off(layer)) --layer;
            const int in = layer == 0 ? vt::MLP_IN : vt::MLP_W, local = p - vt::mlp_w_off(layer);
            if (vt::pack_dest(layer, local / in, local % in) != e) {
                vt_cache_destroy(c);
                return fail(VT_ERR_STATE, "internal error: FP16 weight pack mapping is not invertible");
            }
        }
        for (int e = 0; e < vt::MLP_TPACK; ++e) {
            const int p = vt::tpack_source(e);
            if (p < 0) continue;
            int layer = 5;
            while (layer > 0 && p < vt::mlp_w_off(layer)) --layer;
            const int local = p - vt::mlp_w_off(layer);
            if (layer == 0 || vt::tpack_dest(layer, local / vt::MLP_W, local % vt::MLP_W) != e) {
                vt_cache_destroy(c);
                return fail(VT_ERR_STATE, "internal error: FP32 weight pack mapping is not invertible");
            }
        }
        if (hipMalloc(&c->d_params, pb) != hipSuccess || hipMalloc(&c->d_grad, pb) != hipSuccess ||
            hipMalloc(&c->d_m, pb) != hipSuccess || hipMalloc(&c->d_v, pb) != hipSuccess ||
            hipMalloc(&c->d_pack, (size_t)vt::MLP_PACK * 2) != hipSuccess ||
            hipMalloc(&c->d_bias, (size_t)vt::MLP_BIAS * 4) != hipSuccess ||
            hipMalloc(&c->d_tpack, (size_t)vt::MLP_TPACK * 4) != hipSuccess ||
            hipMalloc(&c->d_part, (size_t)vt::MLP_GRAD_SPLITS * vt::MLP_N_PARAMS * 4) != hipSuccess ||
            hipMalloc(&c->d_loss, VT_CACHE_LOSS_SLOTS * 4) != hipSuccess || hipMalloc(&c->d_steps, 4) != hipSuccess ||
            hipMalloc(&c->d_nsamp, 4) != hipSuccess || hipMalloc(&c->d_tally, 16) != hipSuccess) {
            vt_cache_destroy(c);
            return fail(VT_ERR_HIP, "cache allocation failed");
        }
        /* He-uniform weights (bound sqrt(6 / fan_in)), output layer scaled by
         * 0.1, zero biases; PCG32 stream `seed` */
        std::vector<float> p((size_t)vt::MLP_N_PARAMS, 0.0f);
        /* PCG32 (XSH-RR), initstate = seed, stream 0x5bd1e995 */
        uint64_t st = 0u;
        const uint64_t inc = (0x5bd1e995ull << 1u) | 1u;
        auto next = [&]() {
            const uint64_t old = st;
            st = old * 6364136223846793005ull + inc;
            const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u), rot = (uint32_t)(old >> 59u);
            return (xs >> rot) | (xs << ((32u - rot) & 31u));
        };
        next();
        st += seed;
        next();
        for (int L = 0; L < 6; ++L) {
            const int in = L == 0 ? vt::MLP_IN : vt::MLP_W, outn = L == 5 ? vt::MLP_OUT : vt::MLP_W;
            const float bound = std::sqrt(6.0f / (float)in) * (L == 5 ? 0.1f : 1.0f);
            for (int e = 0; e < outn * in; ++e) {
                const float u = (float)(next() >> 8) * (1.0f / 16777216.0f);
                p[(size_t)vt::mlp_w_off(L) + e] = (2.0f * u - 1.0f) * bound;
            }
        }
        *out = c;
        const int r = vt_cache_set_params(c, p.data());
        if (r != VT_OK) { vt_cache_destroy(c); *out = nullptr; }
        return r;
    });
}

/* Free the cache's device memory and detach it from its context. */
static void cache_release(vt_cache* c)
{
    if (!c->ctx) return;
    (void)hipSetDevice(c->ctx->device);
    (void)hipStreamSynchronize(c->ctx->stream);
    dfree(c->d_params); dfree(c->d_grad); dfree(c->d_m); dfree(c->d_v); dfree(c->d_pack); dfree(c->d_bias);
    dfree(c->d_tpack); dfree(c->d_part); dfree(c->d_loss); dfree(c->d_steps); dfree(c->d_nsamp); dfree(c->d_tally);
    dfree(c->d_X); dfree(c->d_H); dfree(c->d_DZ); dfree(c->d_DY); dfree(c->d_raw);
    dfree(c->d_tgt); dfree(c->d_iraw); dfree(c->d_iout);
    dfree(c->d_qraw); dfree(c->d_qout); dfree(c->d_qmeta); dfree(c->d_qrad); dfree(c->d_qthr); dfree(c->d_counts);
    dfree(c->d_vrec); dfree(c->d_vrec_n); dfree(c->d_sraw); dfree(c->d_stgt);
    (void)hipGetLastError();
    auto& v = c->ctx->caches;
    v.erase(std::remove(v.begin(), v.end(), c), v.end());
    c->ctx = nullptr;
    c->tiles = 0; c->cap = 0; c->icap = 0; c->qcap = 0; c->vcap = 0; c->vslots = 0; c->scap = 0;
}

void vt_cache_destroy(vt_cache* c)
{
    if (!c) return;
    cache_release(c);
    delete c;
}

int vt_cache_get_params(vt_cache* c, float* p)
{
    return no_throw([&]() -> int {
        if (!c || !p) return fail(VT_ERR_ARG, "null argument");
        if (!c->ctx) return fail(VT_ERR_STATE, "the cache's context was destroyed");
        VT_HIP(hipSetDevice(c->ctx->device));
        VT_HIP(hipMemcpyAsync(p, c->d_params, (size_t)vt::MLP_N_PARAMS * 4, hipMemcpyDeviceToHost, c->ctx->stream));
        VT_HIP(hipStreamSynchronize(c->ctx->stream));
        return VT_OK;
    });
}

int vt_cache_set_params(vt_cache* c, const float* p)
{
    return no_throw([&]() -> int {
        if (!c || !p) return fail(VT_ERR_ARG, "null argument");
        if (!c->ctx) return fail(VT_ERR_STATE, "the cache's context was destroyed");
        VT_HIP(hipSetDevice(c->ctx->device));
        hipStream_t s = c->ctx->stream;
        VT_HIP(hipMemcpyAsync(c->d_params, p, (size_t)vt::MLP_N_PARAMS * 4, hipMemcpyHostToDevice, s));
        VT_HIP(vt::launch_fill_u32((uint32_t*)c->d_grad, vt::MLP_N_PARAMS, 0u, s));
        VT_HIP(vt::launch_fill_u32((uint32_t*)c->d_m, vt::MLP_N_PARAMS, 0u, s));
        VT_HIP(vt::launch_fill_u32((uint32_t*)c->d_v, vt::MLP_N_PARAMS, 0u, s));
        VT_HIP(vt::launch_fill_u32(c->d_steps, 1, 0u, s));
        int r = cache_repack(c);
        if (r != VT_OK) return r;
        VT_HIP(hipStreamSynchronize(s));
        return VT_OK;
    });
}

int vt_cache_infer(vt_cache* c, int64_t n, const float* raw, float* out)
{
    return no_throw([&]() -> int {
        if (!c || n < 0 || (n > 0 && (!raw || !out))) return fail(VT_ERR_ARG, "bad argument");
        if (!c->ctx) return fail(VT_ERR_STATE, "the cache's context was destroyed");
        if (n == 0) return VT_OK;
        VT_HIP(hipSetDevice(c->ctx->device));
        hipStream_t s = c->ctx->stream;
        if (n > c->icap) {
            dfree(c->d_iraw); dfree(c->d_iout);
            c->icap = 0;
            VT_HIP(hipMalloc(&c->d_iraw, (size_t)n * vt::MLP_RAW * 4));
            VT_HIP(hipMalloc(&c->d_iout, (size_t)n * vt::MLP_OUT * 4));
            c->icap = n;
        }
        VT_HIP(hipMemcpyAsync(c->d_iraw, raw, (size_t)n * vt::MLP_RAW * 4, hipMemcpyHostToDevice, s));
        VT_HIP(vt::launch_mlp_infer(c->d_pack, c->d_bias, c->d_iraw, c->d_iout, n, c->ctx->cus, s));
        VT_HIP(hipMemcpyAsync(out, c->d_iout, (size_t)n * vt::MLP_OUT * 4, hipMemcpyDeviceToHost, s));
        VT_HIP(hipStreamSynchronize(s));
        return VT_OK;
    });
}

int vt_cache_gradient(vt_cache* c, int64_t n, const float* raw, const float* target, float* grad, float* loss)
{
    return no_throw([&]() -> int {
        if (!c || n <= 0 || !raw || !target || !grad) return fail(VT_ERR_ARG, "bad argument");
        if (!c->ctx) return fail(VT_ERR_STATE, "the cache's context was destroyed");
        VT_HIP(hipSetDevice(c->ctx->device));
        return cache_host_step(c, n, raw, target, grad, 0.0f, loss);
    });
}

int vt_cache_train(vt_cache* c, int64_t n, const float* raw, const float* target, float lr, float* loss)
{
    return no_throw([&]() -> int {
        if (!c || n <= 0 || !raw || !target) return fail(VT_ERR_ARG, "bad argument");
        if (!c->ctx) return fail(VT_ERR_STATE, "the cache's context was destroyed");
        if (!(lr > 0.0f)) return fail(VT_ERR_ARG, "learning rate must be positive");
        VT_HIP(hipSetDevice(c->ctx->device));
        return cache_host_step(c, n, raw, target, nullptr, lr, loss);
    });
}

int vt_render_cached(vt_context* c, vt_cache* k, uint64_t seed, uint32_t n0, uint32_t count, int32_t cache_bounce,
                     uint32_t train_paths, int32_t train_extra, float lr, uint32_t batch, vt_cache_stats* out)
{
    return no_throw([&]() -> int {
        if (!c || !k || k->ctx != c) return fail(VT_ERR_ARG, "cache must belong to the context (and be alive)");
        if (!c->has_volume || !c->d_accum) return fail(VT_ERR_STATE, "no volume or no frame");
        if (cache_bounce < -1 || cache_bounce > 64) return fail(VT_ERR_ARG, "cache_bounce must be in [-1, 64]");
        if (train_extra < -1 || train_extra > 64) return fail(VT_ERR_ARG, "train_extra must be in [-1, 64]");
        if (train_paths > 0 && !(lr > 0.0f)) return fail(VT_ERR_ARG, "learning rate must be positive");
        if (batch == 0) batch = 16384;
        if (batch > (1u << 20)) return fail(VT_ERR_LIMIT, "training batch above 2^20 samples");
        VT_HIP(hipSetDevice(c->device));
        const auto t0 = std::chrono::steady_clock::now();
        hipStream_t s = c->stream;
        const uint64_t WH = (uint64_t)c->W * (uint64_t)c->H;
        if (WH > 0xffffffffull) return fail(VT_ERR_LIMIT, "frame too large");
        int r = ensure_wavefront(c, std::max<uint64_t>(WH, std::min<uint64_t>(WH, c->max_paths)));
        if (r != VT_OK) return r;
        /* buffers */
        if (k->qcap < WH) {
            dfree(k->d_qraw); dfree(k->d_qout); dfree(k->d_qmeta); dfree(k->d_qrad); dfree(k->d_qthr);
            k->qcap = 0;
            VT_HIP(hipMalloc(&k->d_qraw, WH * 12 * 4));
            VT_HIP(hipMalloc(&k->d_qout, WH * 3 * 4));
            VT_HIP(hipMalloc(&k->d_qmeta, WH * 16));
            VT_HIP(hipMalloc(&k->d_qrad, WH * 16));
            VT_HIP(hipMalloc(&k->d_qthr, WH * 16));
            k->qcap = WH;
        }
        if (!k->d_counts) VT_HIP(hipMalloc(&k->d_counts, 16));
        const uint32_t maxv = (uint32_t)std::max(1, c->st.max_bounces);
        const uint32_t stride = train_paths ? (uint32_t)std::max<uint64_t>(1, WH / train_paths) : (uint32_t)WH + 1u;
        const uint64_t slots = train_paths ? (WH + stride - 1) / stride : 0;
        if (k->vcap < slots * maxv || k->vslots < slots) {
            dfree(k->d_vrec); dfree(k->d_vrec_n);
            k->vcap = k->vslots = 0;
            VT_HIP(hipMalloc(&k->d_vrec, std::max<uint64_t>(1, slots * maxv) * 18 * 4));
            VT_HIP(hipMalloc(&k->d_vrec_n, std::max<uint64_t>(1, slots) * 4));
            k->vcap = slots * maxv;
            k->vslots = slots;
        }
        const uint64_t scap = std::max<uint64_t>(1, slots * maxv);
        if (k->scap < scap) {
            dfree(k->d_sraw); dfree(k->d_stgt);
            k->scap = 0;
            VT_HIP(hipMalloc(&k->d_sraw, scap * 12 * 4));
            VT_HIP(hipMalloc(&k->d_stgt, scap * 3 * 4));
            k->scap = scap;
        }
        if (scap > 0xffffffffull) return fail(VT_ERR_LIMIT, "too many training samples per pass");
        const uint32_t nb = train_paths ? (uint32_t)((scap + batch - 1) / batch) : 0u;
        if (nb > VT_CACHE_LOSS_SLOTS) return fail(VT_ERR_LIMIT, "too many training batches per pass");
        vt_cache_stats st;
        std::memset(&st, 0, sizeof(st));
        uint32_t steps0 = 0;
        VT_HIP(hipMemcpyAsync(&steps0, k->d_steps, 4, hipMemcpyDeviceToHost, s));
        VT_HIP(hipStreamSynchronize(s));
        VT_HIP(vt::launch_fill_u32((uint32_t*)k->d_tally, 4, 0u, s));
        for (uint32_t pass = 0; pass < count; ++pass) {
            const uint32_t n = n0 + pass;
            vt::TraceArgs a = trace_args(c);
            a.seed = seed;
            a.n0 = n;
            a.count = 1;
            a.rng_frame_max = 0xffffffffu;
            a.stats = nullptr;
            vt::CacheRender C;
            std::memset(&C, 0, sizeof(C));
            C.cache_bounce = cache_bounce;
            C.train_bounce = (cache_bounce < 0 || train_extra < 0) ? -1 : cache_bounce + train_extra;
            C.train_stride = stride;
            /* the training paths move across the image from pass to pass */
            C.train_offset = (uint32_t)(((uint64_t)n * 2654435761u) % stride);
            C.train_slots = (uint32_t)slots;
            C.max_vertices = maxv;
            C.vrec = k->d_vrec;
            C.vrec_n = k->d_vrec_n;
            C.q_raw = k->d_qraw;
            C.q_meta = k->d_qmeta;
            C.q_rad = k->d_qrad;
            C.q_thr = k->d_qthr;
            C.q_out = k->d_qout;
            C.q_count = k->d_counts;
            C.q_done = k->d_counts + 2;
            C.s_raw = k->d_sraw;
            C.s_tgt = k->d_stgt;
            C.s_count = k->d_counts + 1;
            C.s_cap = (uint32_t)k->scap;
            VT_HIP(vt::launch_fill_u32(k->d_counts, 4, 0u, s));
            VT_HIP(vt::launch_render_cached(a, c->wf, C, k->d_pack, k->d_bias, c->cus, s));
            VT_HIP(vt::launch_cache_tally(k->d_counts, (uint32_t)k->scap, k->d_tally, s));
            /* training batches over the samples the pass stored: as many as hold
             * samples (reading the count back costs one round trip, launching
             * every possible batch a few microseconds per empty one) */
            uint32_t nb_used = 0;
            if (nb) {
                VT_HIP(vt::launch_fill_u32((uint32_t*)k->d_loss, nb, 0u, s));
                uint32_t stored = 0;
                VT_HIP(hipMemcpyAsync(&stored, k->d_counts + 1, 4, hipMemcpyDeviceToHost, s));
                VT_HIP(hipStreamSynchronize(s));
                const uint64_t used = std::min<uint64_t>(stored, k->scap);
                nb_used = (uint32_t)((used + batch - 1) / batch);
            }
            for (uint32_t bi = 0; bi < nb_used; ++bi) {
                const vt::TrainBatch tb = { k->d_sraw, k->d_stgt, k->d_counts + 1, (uint32_t)k->scap, bi * batch, batch };
                r = cache_train_batch(k, tb, nullptr, lr, (int)bi);
                if (r != VT_OK) return r;
            }
        }
        /* statistics: totals, Adam steps taken, loss of the last pass's last batch */
        unsigned long long tally[2] = { 0, 0 };
        uint32_t counts[2] = { 0, 0 }, steps1 = 0;
        std::vector<float> losses(std::max<uint32_t>(nb, 1u), 0.0f);
        VT_HIP(hipMemcpyAsync(tally, k->d_tally, 16, hipMemcpyDeviceToHost, s));
        VT_HIP(hipMemcpyAsync(counts, k->d_counts, 8, hipMemcpyDeviceToHost, s));
        VT_HIP(hipMemcpyAsync(&steps1, k->d_steps, 4, hipMemcpyDeviceToHost, s));
        if (nb) VT_HIP(hipMemcpyAsync(losses.data(), k->d_loss, (size_t)nb * 4, hipMemcpyDeviceToHost, s));
        VT_HIP(hipStreamSynchronize(s));
        st.queries = tally[0];
        st.samples = train_paths ? tally[1] : 0;
        st.steps = steps1 - steps0;
        const uint64_t last_n = std::min<uint64_t>(counts[1], k->scap);
        for (uint32_t bi = 0; bi < nb; ++bi) {
            const uint64_t lo = (uint64_t)bi * batch;
            if (lo >= last_n) break;
            const uint64_t nbat = std::min<uint64_t>(batch, last_n - lo);
            st.last_loss = losses[bi] / (float)(vt::MLP_OUT * nbat);
        }
        st.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        c->timed = false;
        if (out) *out = st;
        return VT_OK;
    });
}

/* ---- autotuning --------------------------------------------------------------- */
int vt_autotune(vt_context* c, uint32_t passes, vt_tuning* out)
{
    return no_throw([&]() -> int {
        if (!c || !out) return fail(VT_ERR_ARG, "null argument");
        if (!c->has_volume || !c->d_accum) return fail(VT_ERR_STATE, "autotune needs a volume and a frame");
        if (passes == 0) passes = 1;
        VT_HIP(hipSetDevice(c->device));
        const size_t acc_bytes = (size_t)c->W * c->H * sizeof(float4);
        const size_t sum_bytes = c->d_sum64 ? (size_t)c->W * c->H * 3 * sizeof(double) : 0;
        DevTmp<void> save_acc;
        DevTmp<void> save_sum;
        {
            hipError_t he = hipMalloc(&save_acc.p, acc_bytes);
            if (he == hipSuccess && sum_bytes) he = hipMalloc(&save_sum.p, sum_bytes);
            if (he == hipSuccess)
                he = hipMemcpyAsync(save_acc.p, c->d_accum, acc_bytes, hipMemcpyDeviceToDevice, c->stream);
            if (he == hipSuccess && sum_bytes)
                he = hipMemcpyAsync(save_sum.p, c->d_sum64, sum_bytes, hipMemcpyDeviceToDevice, c->stream);
            if (he != hipSuccess) {
                (void)hipGetLastError();
                (void)hipStreamSynchronize(c->stream);
                return fail(VT_ERR_HIP, std::string("autotune: saving the accumulation failed: ") + hipGetErrorString(he));
            }
        }
        const bool was_stats = c->stats_enabled;
        c->stats_enabled = false;
        const int mode0 = c->render_mode, b0 = c->dda_budget, r0 = c->refill_min, k0 = c->refill_chunk;
        struct Cand { int mode, budget, refill, chunk; };
        std::vector<Cand> cands;
        cands.push_back({ 0, b0, r0, k0 });
        for (int budget : { 16, 32, 64 })
            for (int refill : { 8, 16, 24 })
                for (int chunk : { 128, 256, 512 }) cands.push_back({ 1, budget, refill, chunk });
        double best = 1e30;
        Cand pick = { mode0, b0, r0, k0 };
        int rc = VT_OK;
        for (size_t i = 0; i <= cands.size() && rc == VT_OK; ++i) {
            const Cand& k = cands[i == 0 ? 1 : i - 1];   /* the first run warms up */
            c->render_mode = k.mode;
            c->dda_budget = k.budget;
            c->refill_min = k.refill;
            c->refill_chunk = k.chunk;
            rc = vt_render(c, 0, 0, passes, 0xffffffffu);
            if (rc != VT_OK) break;
            float ms = 0.0f;
            if (hipStreamSynchronize(c->stream) != hipSuccess || hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) {
                rc = fail(VT_ERR_HIP, "autotune timing failed");
                break;
            }
            if (i > 0 && ms < best) {
                best = ms;
                pick = k;
            }
        }
        /* the accumulation and the context's settings are restored before any
         * error is reported */
        hipError_t he = hipMemcpyAsync(c->d_accum, save_acc.p, acc_bytes, hipMemcpyDeviceToDevice, c->stream);
        if (he == hipSuccess && sum_bytes)
            he = hipMemcpyAsync(c->d_sum64, save_sum.p, sum_bytes, hipMemcpyDeviceToDevice, c->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
        c->stats_enabled = was_stats;
        c->timed = false;
        if (he != hipSuccess) {
            (void)hipGetLastError();
            c->render_mode = mode0; c->dda_budget = b0; c->refill_min = r0; c->refill_chunk = k0;
            return fail(VT_ERR_HIP, std::string("autotune: restoring the accumulation failed: ") + hipGetErrorString(he));
        }
        if (rc != VT_OK) {
            c->render_mode = mode0; c->dda_budget = b0; c->refill_min = r0; c->refill_chunk = k0;
            return rc;
        }
        c->render_mode = pick.mode;
        c->dda_budget = pick.budget;
        c->refill_min = pick.refill;
        c->refill_chunk = pick.chunk;
        out->render_mode = pick.mode;
        out->dda_budget = pick.budget;
        out->refill_min = pick.refill;
        out->refill_chunk = pick.chunk;
        out->max_paths = c->max_paths;
        out->ms = best;
        out->candidates = (int32_t)cands.size();
        return VT_OK;
    });
}

int vt_read_accum(vt_context* c, float* rgba)
{
    return no_throw([&]() -> int {
        if (!c || !rgba) return fail(VT_ERR_ARG, "null argument");
        if (!c->d_accum) return fail(VT_ERR_STATE, "no frame");
        VT_HIP(hipSetDevice(c->device));
        VT_HIP(hipMemcpyAsync(rgba, c->d_accum, (size_t)c->W * c->H * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_read_sum64(vt_context* c, double* rgb)
{
    return no_throw([&]() -> int {
        if (!c || !rgb) return fail(VT_ERR_ARG, "null argument");
        if (!c->d_sum64) return fail(VT_ERR_STATE, "frame has no FP64 sums");
        VT_HIP(hipSetDevice(c->device));
        VT_HIP(hipMemcpyAsync(rgb, c->d_sum64, (size_t)c->W * c->H * 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_write_accum(vt_context* c, const float* rgba)
{
    return no_throw([&]() -> int {
        if (!c || !rgba) return fail(VT_ERR_ARG, "null argument");
        if (!c->d_accum) return fail(VT_ERR_STATE, "no frame");
        VT_HIP(hipSetDevice(c->device));
        VT_HIP(hipMemcpyAsync(c->d_accum, rgba, (size_t)c->W * c->H * sizeof(float4), hipMemcpyHostToDevice, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_write_sum64(vt_context* c, const double* rgb)
{
    return no_throw([&]() -> int {
        if (!c || !rgb) return fail(VT_ERR_ARG, "null argument");
        if (!c->d_sum64) return fail(VT_ERR_STATE, "frame has no FP64 sums");
        VT_HIP(hipSetDevice(c->device));
        VT_HIP(hipMemcpyAsync(c->d_sum64, rgb, (size_t)c->W * c->H * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_get_stats(vt_context* c, vt_stats* out)
{
    return no_throw([&]() -> int {
        if (!c || !out) return fail(VT_ERR_ARG, "null argument");
        VT_HIP(hipSetDevice(c->device));
        VT_HIP(hipStreamSynchronize(c->stream));
        std::memset(out, 0, sizeof(*out));
        if (c->stats_valid) {
            unsigned long long s[6];
            VT_HIP(hipMemcpy(s, c->d_stats, sizeof(s), hipMemcpyDeviceToHost));
            out->rays = s[0]; out->visits = s[1]; out->interactions = s[2];
            out->samples = s[3]; out->env_samples = s[4]; out->env_evals = s[5];
        }
        if (c->timed) {
            float ms = 0.0f;
            VT_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            out->kernel_ms = ms;
        }
        return VT_OK;
    });
}

int vt_trace_first_hit(vt_context* c, uint64_t seed, uint32_t frame, int32_t* rec)
{
    return no_throw([&]() -> int {
        if (!c || !rec) return fail(VT_ERR_ARG, "null argument");
        if (!c->has_volume || c->W <= 0) return fail(VT_ERR_STATE, "volume and frame required");
        VT_HIP(hipSetDevice(c->device));
        vt::TraceArgs a = trace_args(c);
        a.seed = seed; a.n0 = frame;
        const size_t bytes = (size_t)c->W * c->H * 8 * 4;
        DevTmp<int32_t> d;
        VT_HIP(hipMalloc(&d.p, bytes));
        VT_HIP(vt::launch_first_hit(a, d.p, c->stream));
        VT_HIP(hipMemcpyAsync(rec, d.p, bytes, hipMemcpyDeviceToHost, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_trace_rays(vt_context* c, int32_t n, const float* o, const float* dir, int32_t* rec)
{
    return no_throw([&]() -> int {
        if (!c || n < 0 || (n > 0 && (!o || !dir || !rec))) return fail(VT_ERR_ARG, "bad argument");
        if (!c->has_volume) return fail(VT_ERR_STATE, "no volume");
        if (n == 0) return VT_OK;
        VT_HIP(hipSetDevice(c->device));
        vt::TraceArgs a = trace_args(c);
        DevTmp<float> d_o;
        DevTmp<float> d_d;
        DevTmp<int32_t> d_r;
        VT_HIP(hipMalloc(&d_o.p, (size_t)n * 12));
        VT_HIP(hipMalloc(&d_d.p, (size_t)n * 12));
        VT_HIP(hipMalloc(&d_r.p, (size_t)n * 32));
        VT_HIP(hipMemcpyAsync(d_o.p, o, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        VT_HIP(hipMemcpyAsync(d_d.p, dir, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        VT_HIP(vt::launch_trace_rays(a, n, d_o.p, d_d.p, d_r.p, c->stream));
        VT_HIP(hipMemcpyAsync(rec, d_r.p, (size_t)n * 32, hipMemcpyDeviceToHost, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_trace_rays_noalloc(vt_context* c, int32_t n, const float* o, const float* dir, int32_t* rec)
{
    return no_throw([&]() -> int {
        if (!c || n < 0 || (n > 0 && (!o || !dir || !rec))) return fail(VT_ERR_ARG, "bad argument");
        if (n > VT_NOALLOC_MAX_RAYS)
            return fail(VT_ERR_LIMIT, "vt_trace_rays_noalloc: more rays than its staging limit of "
                                      VT_STR(VT_NOALLOC_MAX_RAYS) " (use vt_trace_rays)");
        if (!c->has_volume) return fail(VT_ERR_STATE, "no volume");
        if (n == 0) return VT_OK;
        VT_HIP(hipSetDevice(c->device));
        /* one call at a time uses the device's staging array; a call returns only
         * after its records are back on the host */
        std::lock_guard<std::mutex> lk(g_noalloc_lock);
        unsigned char* base = (unsigned char*)g_noalloc_ptr[c->device];
        if (!base) return fail(VT_ERR_STATE, "staging array of the device was not looked up");
        float* d_o = (float*)(base + VT_NOALLOC_O);
        float* d_d = (float*)(base + VT_NOALLOC_D);
        int32_t* d_r = (int32_t*)(base + VT_NOALLOC_R);
        vt::TraceArgs a = trace_args(c);
        VT_HIP(hipMemcpyAsync(d_o, o, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        VT_HIP(hipMemcpyAsync(d_d, dir, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        VT_HIP(vt::launch_trace_rays(a, n, d_o, d_d, d_r, c->stream));
        VT_HIP(hipMemcpyAsync(rec, d_r, (size_t)n * 32, hipMemcpyDeviceToHost, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_voxelize(vt_context* c, const float* verts, int64_t nv, const uint32_t* idx, int64_t ntri, const float* M,
                int32_t mode, int32_t material_offset, uint8_t* occupancy)
{
    return no_throw([&]() -> int {
        if (!c || nv < 0 || ntri < 0 || !M || (nv > 0 && !verts) || (ntri > 0 && !idx))
            return fail(VT_ERR_ARG, "bad argument");
        if (mode != 0 && mode != 1) return fail(VT_ERR_ARG, "mode must be 0 (THIN) or 1 (FAT)");
        if (!c->has_volume) return fail(VT_ERR_STATE, "no volume: call vt_set_volume first");
        for (int64_t i = 0; i < 3 * ntri; ++i)
            if ((int64_t)idx[i] >= nv) return fail(VT_ERR_ARG, "triangle index out of range");
        if (material_offset < 0 || material_offset >= (int32_t)c->mat.size())
            return fail(VT_ERR_ARG, "material offset outside material data");
        int slot = 0;
        for (int s = 1; s <= c->n_slots; ++s)
            if (c->slot_off[s] == material_offset) slot = s;
        if (!slot) {
            if (c->n_slots == 255) return fail(VT_ERR_LIMIT, "no free material slot");
            slot = ++c->n_slots;
            c->slot_off[slot] = material_offset;
            int r = upload_records(c);
            if (r != VT_OK) return r;
        }
        VT_HIP(hipSetDevice(c->device));
        vt::VoxArgs a;
        std::memset(&a, 0, sizeof(a));
        a.d = c->dims;
        a.ntri = ntri;
        std::memcpy(a.M, M, sizeof(a.M));
        a.mode = mode;
        a.slot = (uint8_t)slot;
        a.bricks = c->d_bricks;
        DevTmp<float> d_v;
        DevTmp<uint32_t> d_i;
        DevTmp<uint8_t> d_occ;
        const int64_t nvox = (int64_t)c->dims.nx * c->dims.ny * c->dims.nz;
        if (ntri > 0) {
            VT_HIP(hipMalloc(&d_v.p, (size_t)nv * 12));
            VT_HIP(hipMalloc(&d_i.p, (size_t)ntri * 12));
            VT_HIP(hipMemcpyAsync(d_v.p, verts, (size_t)nv * 12, hipMemcpyHostToDevice, c->stream));
            VT_HIP(hipMemcpyAsync(d_i.p, idx, (size_t)ntri * 12, hipMemcpyHostToDevice, c->stream));
        }
        if (occupancy) {
            VT_HIP(hipMalloc(&d_occ.p, (size_t)nvox));
            VT_HIP(vt::launch_fill_u8(d_occ.p, (uint64_t)nvox, 0, c->stream));
        }
        a.verts = d_v.p; a.idx = d_i.p; a.occ = d_occ.p;
        /* the list of large triangles (a wave each) and its counter */
        DevTmp<uint32_t> d_big;
        if (ntri > 0) {
            VT_HIP(hipMalloc(&d_big.p, ((size_t)ntri + 1) * 4));
            VT_HIP(vt::launch_fill_u32(d_big.p + ntri, 1, 0u, c->stream));
            a.big = d_big.p;
            a.big_count = d_big.p + ntri;
            a.big_cap = (uint32_t)std::min<int64_t>(ntri, 0x7fffffff);   /* beyond it, triangles keep their thread */
            a.big_area = VT_VOX_BIG_AREA;
        }
        VT_HIP(hipEventRecord(c->ev0, c->stream));
        VT_HIP(vt::launch_voxelize(a, c->cus, c->stream));
        VT_HIP(hipEventRecord(c->ev1, c->stream));
        c->timed = true;           /* vt_get_stats: kernel_ms of the voxelizer launch */
        c->stats_valid = false;
        {
            int r = refresh_occupancy(c);
            if (r != VT_OK) return r;
        }
        if (occupancy) VT_HIP(hipMemcpyAsync(occupancy, d_occ.p, (size_t)nvox, hipMemcpyDeviceToHost, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_camera_matrices(const float* eye, const float* target, float fov_y, float near_d, float far_d, int32_t lens_model,
                       int32_t width, int32_t height, float* mv, float* inv_mv, float* proj, float* inv_proj)
{
    return no_throw([&]() -> int {
        if (!eye || !target || !mv || !inv_mv || !proj || !inv_proj || width <= 0 || height <= 0)
            return fail(VT_ERR_ARG, "bad argument");
        vt::host::camera_matrices(eye, target, fov_y, near_d, far_d, lens_model, width, height, mv, inv_mv, proj, inv_proj);
        return VT_OK;
    });
}

int vt_env_build_cdf(const float* rgb, int32_t w, int32_t h, int32_t* cu_w, int32_t* cu_h, float* cdf_u, int32_t* cv_n,
                     float* cdf_v, float* integral)
{
    return no_throw([&]() -> int {
        if (!rgb || w <= 0 || h <= 0 || !cu_w || !cu_h || !cv_n) return fail(VT_ERR_ARG, "bad argument");
        std::vector<float> f;
        int fw = 0, fh = 0;
        vt::host::image_function(rgb, w, h, 512, f, fw, fh);
        *cu_w = fw + 1;
        *cu_h = fh;
        *cv_n = fh + 1;
        if (!cdf_u || !cdf_v) return VT_OK;
        float integ = 0.0f;
        vt::host::calculate_cdf(f.data(), fw, fh, cdf_u, cdf_v, &integ);
        if (integral) *integral = integ;
        return VT_OK;
    });
}

int vt_mesh_transform(const float* bmin, const float* bmax, const int32_t* res, float* M)
{
    return no_throw([&]() -> int {
        if (!bmin || !bmax || !res || !M) return fail(VT_ERR_ARG, "null argument");
        if (res[0] <= 0 || res[1] <= 0 || res[2] <= 0) return fail(VT_ERR_ARG, "resolution must be positive");
        vt::host::mesh_transform(bmin, bmax, res, M);
        return VT_OK;
    });
}

int32_t vt_prune_emissive(const int32_t* offsets, int32_t nx, int32_t ny, int32_t nz, int32_t* emissive, int32_t n)
{
    return no_throw([&]() -> int {
        if (!offsets || (n > 0 && !emissive)) return fail(VT_ERR_ARG, "bad argument");
        const int32_t res[3] = { nx, ny, nz };
        return vt::host::prune_emissive(offsets, res, emissive, n);
    });
}

int vt_device_math(vt_context* c, int32_t fn, int32_t n, const float* a, const float* b, float* out)
{
    return no_throw([&]() -> int {
        if (!c || n < 0 || (n > 0 && (!a || !b || !out))) return fail(VT_ERR_ARG, "bad argument");
        if (n == 0) return VT_OK;
        VT_HIP(hipSetDevice(c->device));
        DevTmp<float> da, db, dout;
        VT_HIP(hipMalloc(&da.p, (size_t)n * 4));
        VT_HIP(hipMalloc(&db.p, (size_t)n * 4));
        VT_HIP(hipMalloc(&dout.p, (size_t)n * 4));
        VT_HIP(hipMemcpyAsync(da.p, a, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        VT_HIP(hipMemcpyAsync(db.p, b, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        VT_HIP(vt::launch_device_math(fn, n, da.p, db.p, dout.p, c->stream));
        VT_HIP(hipMemcpyAsync(out, dout.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
        VT_HIP(hipStreamSynchronize(c->stream));
        return VT_OK;
    });
}

int vt_copy_probe(vt_context* c, uint64_t bytes, int32_t iters, double* kernel_ms)
{
    return no_throw([&]() -> int {
        if (!c || bytes < 16 || bytes % 16 || iters < 1 || !kernel_ms) return fail(VT_ERR_ARG, "bad argument");
        VT_HIP(hipSetDevice(c->device));
        DevTmp<uint8_t> a, b;
        VT_HIP(hipMalloc(&a.p, bytes));
        VT_HIP(hipMalloc(&b.p, bytes));
        VT_HIP(vt::launch_fill_u8(a.p, bytes, 1, c->stream));
        VT_HIP(vt::launch_copy(a.p, b.p, bytes, c->stream));
        VT_HIP(hipEventRecord(c->ev0, c->stream));
        for (int32_t i = 0; i < iters; ++i) VT_HIP(vt::launch_copy(a.p, b.p, bytes, c->stream));
        VT_HIP(hipEventRecord(c->ev1, c->stream));
        VT_HIP(hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        VT_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->timed = false;
        *kernel_ms = ms / iters;
        return VT_OK;
    });
}

}  // extern "C"

/*
 * cpu_voxelizer.cpp - CPUVoxelizer (src/voxelize/cpuVoxelizer.h:5-13,
 * cpuVoxelizer.cpp:42-300): voxelize.gs's triangle voxelization on host
 * threads, for hosts without a device.  The arithmetic is the device
 * kernel's (voxelize_kernel, csrc/vt_volume.hip), built without FMA
 * contraction, so both mark the same voxels.
 */
#include "renderer.h"
#include "vt_renderer.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>
#include <new>

namespace voxeltoy {
namespace {

struct V2 { float x, y; };
inline float dot2(V2 a, float bx, float by) { return a.x * bx + a.y * by; }

/* int(float) as llvmpipe's x86 code computes it (cvttss2si): non-finite or
 * out-of-range values give INT_MIN; sums wrap as 32-bit integers. */
inline int32_t cvt_i32(float f)
{
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)f;
}
inline int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

/* One triangle, vertices in voxel space (cpuVoxelizer.cpp:238-276 with
 * voxelizeTriPostSwizzle :116-236): dominant-axis swizzle, the three planes'
 * edge functions, the z-range from the plane, each covered voxel set to 1
 * (unswizzled). */
void voxelize_triangle(const float* v0p, const float* v1p, const float* v2p, const int32_t res[3], int mode,
                       unsigned char* occ)
{
    float v[3][3] = { { v0p[0], v0p[1], v0p[2] }, { v1p[0], v1p[1], v1p[2] }, { v2p[0], v2p[1], v2p[2] } };
    for (auto& p : v)
        for (float c : p)
            if (!(std::fabs(c) < INFINITY)) return;     /* non-finite vertex: skipped */
    const float rf[3] = { (float)res[0], (float)res[1], (float)res[2] };
    const float e01[3] = { v[1][0] - v[0][0], v[1][1] - v[0][1], v[1][2] - v[0][2] };
    const float e12[3] = { v[2][0] - v[1][0], v[2][1] - v[1][1], v[2][2] - v[1][2] };
    float n[3] = { e01[1] * e12[2] - e01[2] * e12[1], e01[2] * e12[0] - e01[0] * e12[2],
                   e01[0] * e12[1] - e01[1] * e12[0] };
    const float anx = std::fabs(n[0]), any = std::fabs(n[1]), anz = std::fabs(n[2]);
    int lut;
    auto rot = [](float* a, int i0, int i1, int i2) {
        const float t[3] = { a[i0], a[i1], a[i2] };
        a[0] = t[0]; a[1] = t[1]; a[2] = t[2];
    };
    if (anx >= any && anx >= anz) {
        for (auto& p : v) rot(p, 1, 2, 0);
        rot(n, 1, 2, 0);
        lut = 0;
    } else if (any >= anx && any >= anz) {
        for (auto& p : v) rot(p, 2, 0, 1);
        rot(n, 2, 0, 1);
        lut = 1;
    } else {
        lut = 2;
    }
    int32_t mn[3], mx[3];
    for (int k = 0; k < 3; ++k) {
        const float amn = std::min(std::min(v[0][k], v[1][k]), v[2][k]);
        const float amx = std::max(std::max(v[0][k], v[1][k]), v[2][k]);
        mn[k] = (int32_t)std::min(std::max(std::floor(amn), 0.0f), rf[k]);
        mx[k] = (int32_t)std::min(std::max(std::ceil(amx), 0.0f), rf[k]);
    }
    V2 nxy[3], nyz[3], nzx[3];
    float dxy[3], dyz[3], dzx[3];
    for (int k = 0; k < 3; ++k) {
        const float* a = v[k];
        const float* b = v[(k + 1) % 3];
        const float ek[3] = { b[0] - a[0], b[1] - a[1], b[2] - a[2] };
        nxy[k] = n[2] >= 0.0f ? V2{ -ek[1], ek[0] } : V2{ ek[1], -ek[0] };
        nyz[k] = n[0] >= 0.0f ? V2{ -ek[2], ek[1] } : V2{ ek[2], -ek[1] };
        nzx[k] = n[1] >= 0.0f ? V2{ -ek[0], ek[2] } : V2{ ek[0], -ek[2] };
        if (mode == 0) {
            dxy[k] = dot2(nxy[k], 0.5f - a[0], 0.5f - a[1]) + 0.5f * std::max(std::fabs(nxy[k].x), std::fabs(nxy[k].y));
            dyz[k] = dot2(nyz[k], 0.5f - a[1], 0.5f - a[2]) + 0.5f * std::max(std::fabs(nyz[k].x), std::fabs(nyz[k].y));
            dzx[k] = dot2(nzx[k], 0.5f - a[2], 0.5f - a[0]) + 0.5f * std::max(std::fabs(nzx[k].x), std::fabs(nzx[k].y));
        } else {
            dxy[k] = (-dot2(nxy[k], a[0], a[1]) + std::max(0.0f, nxy[k].x)) + std::max(0.0f, nxy[k].y);
            dyz[k] = (-dot2(nyz[k], a[1], a[2]) + std::max(0.0f, nyz[k].x)) + std::max(0.0f, nyz[k].y);
            dzx[k] = (-dot2(nzx[k], a[2], a[0]) + std::max(0.0f, nzx[k].x)) + std::max(0.0f, nzx[k].y);
        }
    }
    const float np[3] = { n[2] < 0.0f ? -n[0] : n[0], n[2] < 0.0f ? -n[1] : n[1], n[2] < 0.0f ? -n[2] : n[2] };
    const float dTri = (np[0] * v[0][0] + np[1] * v[0][1]) + np[2] * v[0][2];
    const float dMin = mode == 0 ? dTri - (np[0] * 0.5f + np[1] * 0.5f)
                                 : (dTri - std::max(np[0], 0.0f)) - std::max(np[1], 0.0f);
    const float dMax = mode == 0 ? dMin : (dTri - std::min(np[0], 0.0f)) - std::min(np[1], 0.0f);
    const float nzInv = 1.0f / np[2];
    for (int32_t x = mn[0]; x < mx[0]; ++x) {
        const float fx = (float)x;
        for (int32_t y = mn[1]; y < mx[1]; ++y) {
            const float fy = (float)y;
            if (!(dxy[0] + dot2(nxy[0], fx, fy) >= 0.0f && dxy[1] + dot2(nxy[1], fx, fy) >= 0.0f &&
                  dxy[2] + dot2(nxy[2], fx, fy) >= 0.0f))
                continue;
            const float dnp = np[0] * fx + np[1] * fy;
            const float zMinInt = (-dnp + dMin) * nzInv;
            const float zMaxInt = mode == 0 ? zMinInt : (-dnp + dMax) * nzInv;
            const float zMinFloor = std::floor(zMinInt), zMaxCeil = std::ceil(zMaxInt);
            int32_t zMin = wrap_add(cvt_i32(zMinFloor), -(int32_t)(zMinFloor == zMinInt));
            int32_t zMax = wrap_add(cvt_i32(zMaxCeil), (int32_t)(zMaxCeil == zMaxInt));
            zMin = std::max(zMin, mn[2]);
            zMax = std::min(zMax, mx[2]);
            for (int32_t z = zMin; z < zMax; ++z) {
                const float fz = (float)z;
                bool ok = true;
                for (int k = 0; k < 3; ++k)
                    ok = ok && dyz[k] + dot2(nyz[k], fy, fz) >= 0.0f && dzx[k] + dot2(nzx[k], fz, fx) >= 0.0f;
                if (!ok) continue;
                int32_t c0, c1, c2;
                if (lut == 0) { c0 = z; c1 = x; c2 = y; }
                else if (lut == 1) { c0 = y; c1 = z; c2 = x; }
                else { c0 = x; c1 = y; c2 = z; }
                if ((uint32_t)c0 >= (uint32_t)res[0] || (uint32_t)c1 >= (uint32_t)res[1] || (uint32_t)c2 >= (uint32_t)res[2])
                    continue;
                /* several threads may mark one voxel: relaxed atomic stores of the same value */
                __atomic_store_n(&occ[(size_t)c0 + (size_t)c1 * res[0] + (size_t)c2 * res[0] * res[1]],
                                 (unsigned char)1, __ATOMIC_RELAXED);
            }
        }
    }
}

}  // namespace

void CPUVoxelizer::voxelizeMesh(const V3f* vertices, const unsigned int* indices, unsigned int numTriangles,
                                const V3i& voxelDimensions, unsigned char* voxelStorage, int mode, int threads)
{
    if (!vertices || !indices || !voxelStorage || numTriangles == 0) return;
    const int32_t res[3] = { voxelDimensions.x, voxelDimensions.y, voxelDimensions.z };
    unsigned nt = threads > 0 ? (unsigned)threads : std::max(1u, std::thread::hardware_concurrency());
    nt = std::min(nt, numTriangles);
    auto work = [&](unsigned t) {
        const unsigned from = (unsigned)((uint64_t)numTriangles * t / nt);
        const unsigned to = (unsigned)((uint64_t)numTriangles * (t + 1) / nt);
        for (unsigned i = from; i < to; ++i)
            voxelize_triangle(&vertices[indices[3 * i]].x, &vertices[indices[3 * i + 1]].x,
                              &vertices[indices[3 * i + 2]].x, res, mode, voxelStorage);
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
}

void CPUVoxelizer::voxelizeMesh(const Mesh* mesh, const M44f& M, const V3i& resolution, unsigned char* voxelStorage,
                                int mode, int threads)
{
    if (!mesh) return;
    /* voxelize.vs:20-28: (M * (v, 1)).xyz * resolution, M row-major */
    const float* m = M.data();
    const std::vector<float>& src = mesh->vertices();
    const float rf[3] = { (float)resolution.x, (float)resolution.y, (float)resolution.z };
    std::vector<V3f> vs(mesh->numVertices());
    for (size_t i = 0; i < vs.size(); ++i) {
        const float x = src[3 * i], y = src[3 * i + 1], z = src[3 * i + 2];
        float o[3];
        for (int r = 0; r < 3; ++r) o[r] = ((m[r * 4 + 0] * x + m[r * 4 + 1] * y) + m[r * 4 + 2] * z) + m[r * 4 + 3] * 1.0f;
        vs[i] = V3f(o[0] * rf[0], o[1] * rf[1], o[2] * rf[2]);
    }
    voxelizeMesh(vs.data(), mesh->indices().data(), (unsigned int)mesh->numTriangles(), resolution, voxelStorage, mode,
                 threads);
}

}  // namespace voxeltoy

int vtr_cpu_voxelize(const float* vertices, int64_t n_vertices, const uint32_t* indices, int64_t n_triangles,
                     const float* model_transform, const int32_t* res, int32_t mode, uint8_t* occupancy)
{
    /* no C++ exception leaves the C ABI: a failed allocation or thread
     * start becomes an error code */
    try {
        using namespace voxeltoy;
        if (!model_transform || !res || !occupancy || n_vertices < 0 || n_triangles < 0 || n_triangles > 0xffffffffLL ||
            (n_vertices > 0 && !vertices) || (n_triangles > 0 && !indices) || (mode != 0 && mode != 1) || res[0] <= 0 ||
            res[1] <= 0 || res[2] <= 0)
            return VT_ERR_ARG;
        for (int64_t i = 0; i < 3 * n_triangles; ++i)
            if ((int64_t)indices[i] >= n_vertices) return VT_ERR_ARG;
        Mesh mesh(std::vector<float>(vertices, vertices + 3 * n_vertices),
                  std::vector<uint32_t>(indices, indices + 3 * n_triangles));
        M44f M;
        std::memcpy(M.data(), model_transform, sizeof(float) * 16);
        CPUVoxelizer::voxelizeMesh(&mesh, M, V3i(res[0], res[1], res[2]), occupancy, mode);
        return VT_OK;
    } catch (const std::bad_alloc&) {
        return VT_ERR_LIMIT;
    } catch (...) {
        return VT_ERR_STATE;
    }
}

/*
 * vt_volume.hip - grid layout conversion and the triangle voxelizer.
 *
 * The voxelizer kernel restates voxelize.vs (:20-28) and voxelize.gs
 * (swizzleTri :83-137, voxelizeTriPostSwizzle :146-260, main :262-279) with
 * one thread per triangle; covered voxels are written straight into the
 * Morton-bricked grid (and optionally into a dense byte image).  Duplicate
 * writes of the same value by neighbouring triangles are the same benign
 * race the geometry shader has (voxelize.gs:142).
 */
#include "vt_kernels.h"

namespace vt {

__device__ __forceinline__ uint64_t brick_byte(int32_t x, int32_t y, int32_t z)
{
    return (uint64_t)brick_morton((uint32_t)x >> 3, (uint32_t)y >> 3, (uint32_t)z >> 3) * 512u +
           local_morton((uint32_t)x, (uint32_t)y, (uint32_t)z);
}

__global__ __launch_bounds__(256) void brick_from_offsets_kernel(BrickDims d, const int32_t* __restrict__ dense,
                                                                 const uint8_t* __restrict__ lut, int32_t lut_n,
                                                                 uint8_t* __restrict__ bricks, uint32_t* bad)
{
    const int64_t n = (int64_t)d.nx * d.ny * d.nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = (int32_t)(i % d.nx);
        const int64_t r = i / d.nx;
        const int32_t y = (int32_t)(r % d.ny), z = (int32_t)(r / d.ny);
        const int32_t off = dense[i];
        uint8_t slot = 0;
        if (off >= 0) {
            if (off < lut_n && lut[off] != 0) slot = lut[off];
            else atomicOr(bad, 1u);
        }
        bricks[brick_byte(x, y, z)] = slot;
    }
}

__global__ __launch_bounds__(256) void brick_from_ids_kernel(BrickDims d, const uint8_t* __restrict__ dense,
                                                             const uint8_t* __restrict__ remap,
                                                             uint8_t* __restrict__ bricks)
{
    const int64_t n = (int64_t)d.nx * d.ny * d.nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = (int32_t)(i % d.nx);
        const int64_t r = i / d.nx;
        const int32_t y = (int32_t)(r % d.ny), z = (int32_t)(r / d.ny);
        bricks[brick_byte(x, y, z)] = remap[dense[i]];
    }
}

/* occ64[b] for block b = bx + kx*(by + ky*bz): bit (x + 4y + 16z) set when
 * local voxel (x, y, z) of the block is occupied.  The block's 64 slot bytes
 * are one Morton-ordered run inside its brick. */
__global__ __launch_bounds__(256) void block_masks_kernel(BrickDims d, const uint8_t* __restrict__ bricks,
                                                          uint64_t* __restrict__ occ64)
{
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < d.n_blocks;
         w += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t bx = (uint32_t)(w % (uint64_t)d.kx);
        const uint64_t r = w / (uint64_t)d.kx;
        const uint32_t by = (uint32_t)(r % (uint64_t)d.ky), bz = (uint32_t)(r / (uint64_t)d.ky);
        const uint32_t sb = (bx & 1u) | ((by & 1u) << 1) | ((bz & 1u) << 2);
        const uint4* p = (const uint4*)(bricks + (uint64_t)brick_morton(bx >> 1, by >> 1, bz >> 1) * 512u + sb * 64u);
        uint64_t m = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 v = p[q];
            const uint32_t wd[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
            for (int k = 0; k < 4; ++k)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const uint32_t mo = (uint32_t)(q * 16 + k * 4 + b);   /* Morton index in the block */
                    const uint32_t bit = (mo & 1u) | ((mo >> 2) & 2u) |            /* x bits 0, 3 */
                                         (((mo >> 1) & 1u) << 2) | (((mo >> 4) & 1u) << 3) |   /* y */
                                         (((mo >> 2) & 1u) << 4) | (((mo >> 5) & 1u) << 5);    /* z */
                    if ((wd[k] >> (8 * b)) & 0xffu) m |= 1ull << bit;
                }
        }
        occ64[w] = m;
    }
}

/* bits[i] bit k = (occ64[32i + k] != 0). */
__global__ __launch_bounds__(256) void block_bits_kernel(uint64_t n_blocks, const uint64_t* __restrict__ occ64,
                                                         uint32_t* __restrict__ bits)
{
    const uint64_t n_words = (n_blocks + 31) / 32;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words;
         i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t v = 0;
        for (uint32_t k = 0; k < 32u && i * 32u + k < n_blocks; ++k)
            if (occ64[i * 32u + k]) v |= 1u << k;
        bits[i] = v;
    }
}

/* occp over the padded block grid (kx+2) x (ky+2) x (kz+2): border blocks
 * all set; interior block (bx, by, bz) = grid block (bx-1, by-1, bz-1)
 * with the bits of its voxels outside the grid set (grids whose extents
 * are not multiples of 4). */
__global__ __launch_bounds__(256) void block_masks_padded_kernel(BrickDims d, const uint64_t* __restrict__ occ64,
                                                                 uint64_t* __restrict__ occp)
{
    const uint32_t px = (uint32_t)d.kx + 2u, py = (uint32_t)d.ky + 2u, pz = (uint32_t)d.kz + 2u;
    const uint64_t n = (uint64_t)px * py * pz;
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < n; w += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t bx = (uint32_t)(w % px);
        const uint64_t r = w / px;
        const uint32_t by = (uint32_t)(r % py), bz = (uint32_t)(r / py);
        if (bx == 0 || by == 0 || bz == 0 || bx == px - 1u || by == py - 1u || bz == pz - 1u) {
            occp[w] = ~0ull;
            continue;
        }
        const uint32_t gx = bx - 1u, gy = by - 1u, gz = bz - 1u;
        uint64_t m = occ64[gx + (uint64_t)d.kx * (gy + (uint64_t)d.ky * gz)];
        for (uint32_t l = 0; l < 64u; ++l) {
            const int32_t x = (int32_t)(4u * gx + (l & 3u)), y = (int32_t)(4u * gy + ((l >> 2) & 3u)),
                          z = (int32_t)(4u * gz + (l >> 4));
            if (x >= d.nx || y >= d.ny || z >= d.nz) m |= 1ull << l;
        }
        occp[w] = m;
    }
}

/* occr: occp with each mask's bits reversed (voxel l of the block at bit
 * 63 - l): the packed march tests a voxel by shifting its bit into the sign. */
__global__ __launch_bounds__(256) void block_masks_reversed_kernel(const uint64_t* __restrict__ occp,
                                                                   uint64_t* __restrict__ occr, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        occr[i] = __builtin_bitreverse64(occp[i]);
}

__global__ __launch_bounds__(256) void dense_from_bricks_kernel(BrickDims d, const uint8_t* __restrict__ bricks,
                                                                const int32_t* __restrict__ slot_off,
                                                                int32_t* __restrict__ dense)
{
    const int64_t n = (int64_t)d.nx * d.ny * d.nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t x = (int32_t)(i % d.nx);
        const int64_t r = i / d.nx;
        const int32_t y = (int32_t)(r % d.ny), z = (int32_t)(r / d.ny);
        dense[i] = slot_off[bricks[brick_byte(x, y, z)]];
    }
}

__global__ void fill_u8_kernel(uint8_t* p, uint64_t n, uint8_t v)
{
    const uint64_t n16 = n / 16;
    uint4 vv;
    const uint32_t w = (uint32_t)v * 0x01010101u;
    vv.x = vv.y = vv.z = vv.w = w;
    uint4* p16 = (uint4*)p;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        p16[i] = vv;
    for (uint64_t i = n16 * 16 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void fill_u32_kernel(uint32_t* p, uint64_t n, uint32_t v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

static unsigned grid_for(uint64_t n)
{
    uint64_t b = (n + 255) / 256;
    if (b > 65536u * 8u) b = 65536u * 8u;
    if (b < 1) b = 1;
    return (unsigned)b;
}

hipError_t launch_brick_from_offsets(const BrickDims& d, const int32_t* dense, const uint8_t* lut, int32_t lut_n,
                                     uint8_t* bricks, uint32_t* bad, hipStream_t s)
{
    brick_from_offsets_kernel<<<grid_for((uint64_t)d.nx * d.ny * d.nz), 256, 0, s>>>(d, dense, lut, lut_n, bricks, bad);
    return hipGetLastError();
}

hipError_t launch_brick_from_ids(const BrickDims& d, const uint8_t* dense, const uint8_t* remap, uint8_t* bricks,
                                 hipStream_t s)
{
    brick_from_ids_kernel<<<grid_for((uint64_t)d.nx * d.ny * d.nz), 256, 0, s>>>(d, dense, remap, bricks);
    return hipGetLastError();
}

hipError_t launch_block_masks(const BrickDims& d, const uint8_t* bricks, uint64_t* occ64, uint32_t* bits,
                              uint64_t* occp, uint32_t* bitsp, uint64_t* occr, hipStream_t s)
{
    block_masks_kernel<<<grid_for(d.n_blocks), 256, 0, s>>>(d, bricks, occ64);
    block_bits_kernel<<<grid_for((d.n_blocks + 31) / 32), 256, 0, s>>>(d.n_blocks, occ64, bits);
    const uint64_t np = (uint64_t)(d.kx + 2) * (uint64_t)(d.ky + 2) * (uint64_t)(d.kz + 2);
    block_masks_padded_kernel<<<grid_for(np), 256, 0, s>>>(d, occ64, occp);
    block_bits_kernel<<<grid_for((np + 31) / 32), 256, 0, s>>>(np, occp, bitsp);
    if (occr) block_masks_reversed_kernel<<<grid_for(np), 256, 0, s>>>(occp, occr, np);
    return hipGetLastError();
}

hipError_t launch_dense_from_bricks(const BrickDims& d, const uint8_t* bricks, const int32_t* slot_off, int32_t* dense,
                                    hipStream_t s)
{
    dense_from_bricks_kernel<<<grid_for((uint64_t)d.nx * d.ny * d.nz), 256, 0, s>>>(d, bricks, slot_off, dense);
    return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t* p, uint64_t n, uint8_t v, hipStream_t s)
{
    fill_u8_kernel<<<grid_for(n / 16 + 1), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s)
{
    const uint64_t n16 = bytes / 16;
    copy_kernel<<<grid_for(n16), 256, 0, s>>>((const uint4*)src, (uint4*)dst, n16);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s)
{
    fill_u32_kernel<<<grid_for(n), 256, 0, s>>>(p, n, v);
    return hipGetLastError();
}

/* ---- voxelizer ---------------------------------------------------------------------------- */

/* int(float) with x86 cvttss2si semantics (non-finite or out of range ->
 * INT_MIN), the value the reference's llvmpipe path produces for the
 * z-range of degenerate triangles; sums wrap as 32-bit integers. */
__device__ __forceinline__ int32_t cvt_i32(float f)
{
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int32_t)f;
}
__device__ __forceinline__ int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

struct V2 { float x, y; };
__device__ __forceinline__ float dot2(V2 a, float bx, float by) { return a.x * bx + a.y * by; }
__device__ __forceinline__ bool finite3(f3 v)
{
    return vt_absf(v.x) < vt_inff() && vt_absf(v.y) < vt_inff() && vt_absf(v.z) < vt_inff();
}

/* Per-triangle state of voxelize.gs after the set-up (:55-208): the
 * swizzled vertices' edge functions in the three planes, the z-range plane
 * and the (x, y) bounds of the column loop. */
struct VoxTri {
    int32_t lut;                 /* swizzle: 0 (dominant x), 1 (dominant y), 2 (dominant z) */
    int32_t mn[3], mx[3];
    V2 nxy[3], nyz[3], nzx[3];
    float dxy[3], dyz[3], dzx[3];
    f3 np;
    float dMin, dMax, nzInv;
};

/* voxelize.vs:20-28 and the set-up of voxelize.gs; false for a triangle
 * with a non-finite vertex (skipped). */
__device__ __forceinline__ bool vox_setup(const VoxArgs& a, int64_t tri, VoxTri& T)
{
    const float rf[3] = { (float)a.d.nx, (float)a.d.ny, (float)a.d.nz };
    f3 v[3];
    bool finite = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float* p = a.verts + 3 * (size_t)a.idx[3 * tri + k];
        float ls[4];
        mat4_mul(a.M, p[0], p[1], p[2], 1.0f, ls);
        v[k] = mk(ls[0] * rf[0], ls[1] * rf[1], ls[2] * rf[2]);
        finite = finite && finite3(v[k]);
    }
    if (!finite) return false;
    f3 n = cross(v[1] - v[0], v[2] - v[1]);
    const float anx = vt_absf(n.x), any = vt_absf(n.y), anz = vt_absf(n.z);
    if (anx >= any && anx >= anz) {
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = mk(v[k].y, v[k].z, v[k].x);
        n = mk(n.y, n.z, n.x);
        T.lut = 0;
    } else if (any >= anx && any >= anz) {
#pragma unroll
        for (int k = 0; k < 3; ++k) v[k] = mk(v[k].z, v[k].x, v[k].y);
        n = mk(n.z, n.x, n.y);
        T.lut = 1;
    } else {
        T.lut = 2;
    }
    const float amn[3] = { vt_minf(vt_minf(v[0].x, v[1].x), v[2].x), vt_minf(vt_minf(v[0].y, v[1].y), v[2].y),
                           vt_minf(vt_minf(v[0].z, v[1].z), v[2].z) };
    const float amx[3] = { vt_maxf(vt_maxf(v[0].x, v[1].x), v[2].x), vt_maxf(vt_maxf(v[0].y, v[1].y), v[2].y),
                           vt_maxf(vt_maxf(v[0].z, v[1].z), v[2].z) };
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        T.mn[k] = (int32_t)vt_minf(vt_maxf(floorf(amn[k]), 0.0f), rf[k]);
        T.mx[k] = (int32_t)vt_minf(vt_maxf(ceilf(amx[k]), 0.0f), rf[k]);
    }
    const f3 ee[3] = { v[1] - v[0], v[2] - v[1], v[0] - v[2] };
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const f3 ek = ee[k];
        T.nxy[k] = (n.z >= 0.0f) ? V2{ -ek.y, ek.x } : V2{ ek.y, -ek.x };
        T.nyz[k] = (n.x >= 0.0f) ? V2{ -ek.z, ek.y } : V2{ ek.z, -ek.y };
        T.nzx[k] = (n.y >= 0.0f) ? V2{ -ek.x, ek.z } : V2{ ek.x, -ek.z };
        const f3 vk = v[k];
        if (a.mode == 0) {
            T.dxy[k] = dot2(T.nxy[k], 0.5f - vk.x, 0.5f - vk.y) +
                       0.5f * vt_maxf(vt_absf(T.nxy[k].x), vt_absf(T.nxy[k].y));
            T.dyz[k] = dot2(T.nyz[k], 0.5f - vk.y, 0.5f - vk.z) +
                       0.5f * vt_maxf(vt_absf(T.nyz[k].x), vt_absf(T.nyz[k].y));
            T.dzx[k] = dot2(T.nzx[k], 0.5f - vk.z, 0.5f - vk.x) +
                       0.5f * vt_maxf(vt_absf(T.nzx[k].x), vt_absf(T.nzx[k].y));
        } else {
            T.dxy[k] = (-dot2(T.nxy[k], vk.x, vk.y) + vt_maxf(0.0f, T.nxy[k].x)) + vt_maxf(0.0f, T.nxy[k].y);
            T.dyz[k] = (-dot2(T.nyz[k], vk.y, vk.z) + vt_maxf(0.0f, T.nyz[k].x)) + vt_maxf(0.0f, T.nyz[k].y);
            T.dzx[k] = (-dot2(T.nzx[k], vk.z, vk.x) + vt_maxf(0.0f, T.nzx[k].x)) + vt_maxf(0.0f, T.nzx[k].y);
        }
    }
    T.np = (n.z < 0.0f) ? -n : n;
    const float dTri = dot(T.np, v[0]);
    T.dMin = (a.mode == 0) ? dTri - (T.np.x * 0.5f + T.np.y * 0.5f)
                           : (dTri - vt_maxf(T.np.x, 0.0f)) - vt_maxf(T.np.y, 0.0f);
    T.dMax = (a.mode == 0) ? T.dMin : (dTri - vt_minf(T.np.x, 0.0f)) - vt_minf(T.np.y, 0.0f);
    T.nzInv = 1.0f / T.np.z;
    return true;
}

/* One (x, y) column of the triangle's loop (voxelize.gs:181-231): the 2D
 * edge test in the dominant plane, the z-range from the plane, the tests in
 * the other two planes, and the write of every covered voxel (unswizzled). */
__device__ __forceinline__ void vox_column(const VoxArgs& a, const VoxTri& T, int32_t x, int32_t y)
{
    const int32_t res[3] = { a.d.nx, a.d.ny, a.d.nz };
    const float fx = (float)x, fy = (float)y;
    const float d0 = T.dxy[0] + dot2(T.nxy[0], fx, fy);
    const float d1 = T.dxy[1] + dot2(T.nxy[1], fx, fy);
    const float d2 = T.dxy[2] + dot2(T.nxy[2], fx, fy);
    if (!(d0 >= 0.0f && d1 >= 0.0f && d2 >= 0.0f)) return;
    const float dnp = T.np.x * fx + T.np.y * fy;
    const float zMinInt = (-dnp + T.dMin) * T.nzInv;
    const float zMaxInt = (a.mode == 0) ? zMinInt : (-dnp + T.dMax) * T.nzInv;
    const float zMinFloor = floorf(zMinInt), zMaxCeil = ceilf(zMaxInt);
    int32_t zMin = wrap_add(cvt_i32(zMinFloor), -(int32_t)(zMinFloor == zMinInt));
    int32_t zMax = wrap_add(cvt_i32(zMaxCeil), (int32_t)(zMaxCeil == zMaxInt));
    if (zMin < T.mn[2]) zMin = T.mn[2];
    if (zMax > T.mx[2]) zMax = T.mx[2];
    for (int32_t z = zMin; z < zMax; ++z) {
        const float fz = (float)z;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            ok = ok && (T.dyz[k] + dot2(T.nyz[k], fy, fz) >= 0.0f);
            ok = ok && (T.dzx[k] + dot2(T.nzx[k], fz, fx) >= 0.0f);
        }
        if (!ok) continue;
        int32_t c0, c1, c2;
        if (T.lut == 0) { c0 = z; c1 = x; c2 = y; }
        else if (T.lut == 1) { c0 = y; c1 = z; c2 = x; }
        else { c0 = x; c1 = y; c2 = z; }
        if ((uint32_t)c0 >= (uint32_t)res[0] || (uint32_t)c1 >= (uint32_t)res[1] || (uint32_t)c2 >= (uint32_t)res[2])
            continue;
        if (a.bricks) a.bricks[brick_byte(c0, c1, c2)] = a.slot;
        if (a.occ) a.occ[(size_t)c0 + (size_t)c1 * res[0] + (size_t)c2 * (size_t)res[0] * res[1]] = 1u;
    }
}

/* One thread per triangle.  Workgroup i works on triangle chunk
 * (i % 8) * per + i / 8: workgroups go to the eight XCDs in turn, so each
 * XCD takes one contiguous run of the triangle list (neighbouring triangles,
 * whose voxels share 64-byte lines, are then written from one L2 instead of
 * being merged from eight).  A triangle whose column loop exceeds
 * a.big_area columns is listed for voxelize_big_kernel instead. */
#define VOX_XCDS 8u
__global__ __launch_bounds__(256) void voxelize_kernel(VoxArgs a)
{
    const uint32_t per = (gridDim.x + VOX_XCDS - 1u) / VOX_XCDS;
    const uint64_t chunk = (uint64_t)(blockIdx.x % VOX_XCDS) * per + blockIdx.x / VOX_XCDS;
    const int64_t tri = (int64_t)chunk * blockDim.x + threadIdx.x;
    if (tri >= a.ntri) return;
    VoxTri T;
    if (!vox_setup(a, tri, T)) return;
    const int64_t w = T.mx[0] > T.mn[0] ? T.mx[0] - T.mn[0] : 0, h = T.mx[1] > T.mn[1] ? T.mx[1] - T.mn[1] : 0;
    if (a.big && w * h > a.big_area && tri < (int64_t)0xffffffff) {
        const uint32_t slot = atomicAdd(a.big_count, 1u);
        if (slot < a.big_cap) {
            a.big[slot] = (uint32_t)tri;
            return;
        }
    }
    for (int32_t x = T.mn[0]; x < T.mx[0]; ++x)
        for (int32_t y = T.mn[1]; y < T.mx[1]; ++y) vox_column(a, T, x, y);
}

/* The listed large triangles, one wave each: the lanes take the triangle's
 * (x, y) columns in turn (consecutive lanes, consecutive y). */
__global__ __launch_bounds__(256) void voxelize_big_kernel(VoxArgs a)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t n = *a.big_count < a.big_cap ? *a.big_count : a.big_cap;
    for (uint32_t i = wave; i < n; i += waves) {
        VoxTri T;
        if (!vox_setup(a, (int64_t)a.big[i], T)) continue;
        const int64_t w = T.mx[0] - T.mn[0], h = T.mx[1] - T.mn[1];
        for (int64_t k = lane; k < w * h; k += 64) vox_column(a, T, T.mn[0] + (int32_t)(k / h), T.mn[1] + (int32_t)(k % h));
    }
}

/* Bilinear quad table of an RGB image (EnvView::quad): one 64-byte entry per
 * (i0 + 1, j0 + 1) in [0, w] x [0, h], texels outside the image zero. */
__global__ __launch_bounds__(256) void env_quads_kernel(const float* __restrict__ img, int32_t w, int32_t h,
                                                        float4* __restrict__ quads)
{
    const uint64_t n = (uint64_t)(w + 1) * (uint64_t)(h + 1);
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (uint64_t)gridDim.x * blockDim.x) {
        const int32_t i0 = (int32_t)(e % (uint64_t)(w + 1)) - 1, j0 = (int32_t)(e / (uint64_t)(w + 1)) - 1;
        float t[12];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int32_t i = i0 + (k & 1), j = j0 + (k >> 1);
            const bool ok = i >= 0 && i < w && j >= 0 && j < h;
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) t[3 * k + ch] = ok ? img[((size_t)j * w + i) * 3 + ch] : 0.0f;
        }
        float4* q = quads + e * 4u;
        q[0] = make_float4(t[0], t[1], t[2], t[3]);
        q[1] = make_float4(t[4], t[5], t[6], t[7]);
        q[2] = make_float4(t[8], t[9], t[10], t[11]);
        q[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

hipError_t launch_env_quads(const float* img, int32_t w, int32_t h, float4* quads, hipStream_t s)
{
    const uint64_t n = (uint64_t)(w + 1) * (uint64_t)(h + 1);
    const uint64_t g = (n + 255) / 256;
    env_quads_kernel<<<(unsigned)(g < 65536 ? g : 65536), 256, 0, s>>>(img, w, h, quads);
    return hipGetLastError();
}

hipError_t launch_voxelize(const VoxArgs& a, int cus, hipStream_t s)
{
    if (a.ntri <= 0) return hipSuccess;
    /* a multiple of the XCD count, so the chunk mapping covers every chunk */
    const uint64_t g = (uint64_t)(a.ntri + 255) / 256;
    const uint64_t g8 = (g + VOX_XCDS - 1) / VOX_XCDS * VOX_XCDS;
    voxelize_kernel<<<(unsigned)g8, 256, 0, s>>>(a);
    if (a.big) {
        const uint64_t waves = (uint64_t)a.big_cap < (uint64_t)cus * 32u ? (uint64_t)a.big_cap : (uint64_t)cus * 32u;
        voxelize_big_kernel<<<(unsigned)((waves + 3) / 4), 256, 0, s>>>(a);
    }
    return hipGetLastError();
}

}  // namespace vt

/*
 * vt_device.h - device-side data views and per-ray functions of the voxel
 * path tracer (gfx950).
 *
 * Grid layout in HBM: the dense voxel grid is split into 8x8x8 bricks of one
 * byte per voxel (the material slot, 0 = empty).  Bricks are stored in Morton
 * order of their brick coordinates (over the power-of-two cube that covers
 * the grid) and voxels inside a brick in Morton order of their local
 * coordinates, so a 64-byte line holds a 4x4x4 block.  The slot bytes are
 * read only at a hit.
 *
 * The traversal reads two occupancy levels over the 4x4x4 blocks, both in
 * linear block order (bx + BX*(by + BY*bz), BX = ceil(nx/4), ...): one bit
 * per block (set when any voxel of the block is occupied), staged in LDS when
 * it fits, and one 64-bit voxel mask per block (bit x + 4y + 16z for local
 * voxel (x, y, z)).  Entering a block costs one bit test; only non-empty
 * blocks load their 8-byte mask.
 *
 * The traversal replays the reference 3D-DDA (src/shaders/shared/dda.h:7-100)
 * step by step: `dis` is a running FP32 sum, so empty space skips the loads
 * but not the compare/add sequence, and hit voxel, face and step count are
 * identical to the reference arithmetic.  All float expressions keep the
 * reference operand order and are compiled with -ffp-contract=off.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define VT_FN __device__ __forceinline__
#include "vtmath.h"

namespace vt {

/* ---- small float3 with explicit, unfused arithmetic ----------------------- */
struct f3 { float x, y, z; };
__device__ __forceinline__ f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ f3 normalize(f3 v) { const float inv = 1.0f / sqrtf(dot(v, v)); return v * inv; }
__device__ __forceinline__ float length(f3 v) { return sqrtf(dot(v, v)); }

/* mat4 (row-major) x vec4, column products summed left to right. */
__device__ __forceinline__ void mat4_mul(const float* m, float x, float y, float z, float w, float* o)
{
#pragma unroll
    for (int r = 0; r < 4; ++r)
        o[r] = ((m[r * 4 + 0] * x + m[r * 4 + 1] * y) + m[r * 4 + 2] * z) + m[r * 4 + 3] * w;
}

/* ---- views ------------------------------------------------------------------ */
struct MatRec {                 /* one decoded material record (32 bytes) */
    int32_t type;               /* 0 matte, 1 metal, 2 plastic, -1 none */
    float e[3];                 /* emission */
    float r[3];                 /* albedo / reflectance */
    float rough;                /* Blinn exponent (the reference uses roughness unmapped) */
};

struct GridView {
    int32_t nx, ny, nz;
    int32_t max_steps;
    float rx, ry, rz;           /* resolution as float */
    float bmin[3], bmax[3], vs[3], ext[3];
    const uint8_t* bricks;      /* Morton bricks of 512 bytes */
    uint32_t bx4, by4;          /* 4x4x4 block grid extents in x and y */
    const uint64_t* occ64;      /* per-block voxel masks, linear block order */
    const uint32_t* bits;       /* per-block non-empty bits (global copy) */
    uint32_t bit_words;
    /* the same two levels over the block grid padded by one block on every
     * side (grid block (bx, by, bz) is padded block (bx+1, by+1, bz+1)), in
     * which every voxel outside the grid reads as occupied: the wave march
     * stops at the first voxel outside the grid without a bounds test */
    uint32_t bxp, byp;          /* padded block extents in x and y */
    const uint64_t* occp;
    const uint32_t* bitsp;
    uint32_t bit_words_p;
    /* occp with each mask's bits reversed (voxel l of the block at bit
     * 63 - l), for the packed-coordinate march; null when an extent of the
     * grid exceeds VT_PACKED_MAX_EXTENT */
    const uint64_t* occr;
    int32_t wide;               /* padded masks span 4 GiB or more (byte offsets need 64 bits) */
    int32_t n_emissive;
    const int32_t* emissive;    /* linear voxel indices */
    const int4* lights;         /* per emissive entry: voxel x, y, z and its material slot */
};

struct EnvView {
    int32_t use_image;
    float top[3], bottom[3];
    int32_t w, h;
    const float* img;           /* RGB */
    /* bilinear quads: entry (j0 + 1) * (w + 1) + (i0 + 1) holds texels
     * (i0, j0), (i0 + 1, j0), (i0, j0 + 1), (i0 + 1, j0 + 1) as 12 floats
     * (zero outside the image, the GL_CLAMP border) in one 64-byte line */
    const float4* quad;
    int32_t cu_w, cu_h;
    const float* cdfu;
    int32_t cv_n;
    const float* cdfv;
    float integral, rotation;
};

struct CamView {
    float inv_mv[16];
    float inv_proj[16];
    float P32, P22, P23;        /* proj[2][3], proj[2][2], proj[3][2] (row, col) */
    float viewport[4];
    float win_z;
    float lens_radius;
    float focal_distance;
    int32_t lens_model;
};

struct Settings {
    int32_t max_bounces;
    float exposure;
    float inv_gamma;            /* 1.0f / gamma */
    int32_t sel[3];
    float wire_opacity, wire_thickness;
};

/* ---- Morton indices -------------------------------------------------------------- */
__device__ __forceinline__ uint32_t spread3_10(uint32_t v)
{
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__device__ __forceinline__ uint32_t brick_morton(uint32_t bx, uint32_t by, uint32_t bz)
{
    return spread3_10(bx) | (spread3_10(by) << 1) | (spread3_10(bz) << 2);
}
__device__ __forceinline__ uint32_t spread3_3(uint32_t v) { return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4); }
__device__ __forceinline__ uint32_t local_morton(uint32_t x, uint32_t y, uint32_t z)
{
    return spread3_3(x & 7u) | (spread3_3(y & 7u) << 1) | (spread3_3(z & 7u) << 2);
}
/* Linear index of the 4x4x4 block holding in-grid voxel (x, y, z); fits
 * 24-bit multiplies since block extents are at most 1024. */
/* a * b + c on 24-bit operands as one full-rate instruction (the compiler
 * otherwise may pick a 64-bit multiply-add) */
__device__ __forceinline__ uint32_t mad24(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

__device__ __forceinline__ uint32_t block_index(const GridView& g, uint32_t x, uint32_t y, uint32_t z)
{
    return mad24(mad24(z >> 2, g.by4, y >> 2), g.bx4, x >> 2);
}
/* Bit of voxel (x, y, z) in its 4x4x4 block mask. */
__device__ __forceinline__ uint32_t block_bit(uint32_t x, uint32_t y, uint32_t z)
{
    return (x & 3u) | ((y & 3u) << 2) | ((z & 3u) << 4);
}

/* ---- rayAABBIntersection (aabb.h:1-32) -------------------------------------------- */
__device__ __forceinline__ float aabb(f3 o, f3 d, f3 bmin, f3 bmax)
{
    const f3 inv = mk(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    const f3 t0 = (bmin - o) * inv;
    const f3 t1 = (bmax - o) * inv;
    const float tminf = vt_maxf(vt_minf(t0.x, t1.x), vt_maxf(vt_minf(t0.y, t1.y), vt_minf(t0.z, t1.z)));
    const float tmaxf = vt_minf(vt_maxf(t0.x, t1.x), vt_minf(vt_maxf(t0.y, t1.y), vt_maxf(t0.z, t1.z)));
    if (tmaxf < 0.0f) return -1.0f;
    if (tminf > tmaxf) return -1.0f;
    return vt_maxf(0.0f, tminf);
}

/* ---- 3D-DDA (dda.h:7-61) -----------------------------------------------------------
 * Returns 1 on hit, 0 when the march ended without a hit (position written),
 * -1 when the start point lies outside the grid (position not written).
 * `slot` receives the hit voxel's material slot. */
struct Hit { int32_t x, y, z, steps; uint32_t slot; };

/* Smallest of three floats in one instruction.  The DDA's distances are
 * finite and non-negative (never NaN), where this equals fminf(fminf(a, b), c);
 * the instruction form avoids the NaN-quieting the library min inserts. */
__device__ __forceinline__ float min3f(float a, float b, float c)
{
    float r;
    asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

/* Key of the 4x4x4 block of an in-grid voxel (10 bits per axis). */
__device__ __forceinline__ uint32_t block_key(int32_t x, int32_t y, int32_t z)
{
    return ((uint32_t)x >> 2) | (((uint32_t)y >> 2) << 10) | (((uint32_t)z >> 2) << 20);
}

/* The grid-independent part of raymarch's set-up (dda.h:7-35): origin
 * moved by sign(d)*0.001, direction clamped by mix(d, 1e-5, step(|d|, 1e-5)),
 * its reciprocal and sign.  Computed where a ray is made, so a traversal
 * that takes a new ray skips three divisions. */
struct PreRay {
    f3 o, inv;
    int32_t sx, sy, sz;
};

__device__ __forceinline__ PreRay prepare_ray(f3 o, f3 d)
{
    PreRay r;
    r.o = o + mk(vt_signf(d.x), vt_signf(d.y), vt_signf(d.z)) * 0.001f;
    /* mix(d, 1e-5, step(|d|, 1e-5)) as an exact select */
    const float dx = vt_stepf(vt_absf(d.x), 1e-5f) != 0.0f ? 1e-5f : d.x;
    const float dy = vt_stepf(vt_absf(d.y), 1e-5f) != 0.0f ? 1e-5f : d.y;
    const float dz = vt_stepf(vt_absf(d.z), 1e-5f) != 0.0f ? 1e-5f : d.z;
    r.inv = mk(1.0f / dx, 1.0f / dy, 1.0f / dz);
    r.sx = dx > 0.0f ? 1 : (dx < 0.0f ? -1 : 0);
    r.sy = dy > 0.0f ? 1 : (dy < 0.0f ? -1 : 0);
    r.sz = dz > 0.0f ? 1 : (dz < 0.0f ? -1 : 0);
    return r;
}

/* Queue encoding of a prepared ray in 24 bytes: (origin, 1/d.x) and
 * (1/d.y, 1/d.z).  The step signs are not stored: the clamped direction
 * component is never zero, so its sign is the sign bit of its reciprocal
 * (also for a reciprocal of +-0 from an infinite component), and a NaN
 * component (sign 0 in prepare_ray) has a NaN reciprocal. */
__device__ __forceinline__ int32_t step_sign(float inv)
{
    return inv != inv ? 0 : ((__float_as_uint(inv) >> 31) ? -1 : 1);
}
__device__ __forceinline__ float4 queue_a(const PreRay& r) { return make_float4(r.o.x, r.o.y, r.o.z, r.inv.x); }
__device__ __forceinline__ float2 queue_b(const PreRay& r) { return make_float2(r.inv.y, r.inv.z); }
__device__ __forceinline__ PreRay unpack_ray(float4 qa, float2 qb)
{
    PreRay r;
    r.o = mk(qa.x, qa.y, qa.z);
    r.inv = mk(qa.w, qb.x, qb.y);
    r.sx = step_sign(r.inv.x);
    r.sy = step_sign(r.inv.y);
    r.sz = step_sign(r.inv.z);
    return r;
}

/* The march as resumable state, for traversal kernels whose lanes take a
 * new ray as soon as theirs finishes.  init() returns false when the start
 * point lies outside the grid (steps = -1, nothing else written); advance()
 * runs at most `budget` steps and returns 0 while the march goes on, 1 on a
 * hit, 2 when it left the grid or ran out of steps. */
struct DdaState {
    int32_t x, y, z, steps;
    float disx, disy, disz, ax, ay, az;
    int32_t stx, sty, stz;
    uint32_t key;               /* linear index of the block of `mask` */
    uint64_t mask;

    __device__ __forceinline__ bool init(const GridView& g, f3 o, f3 d) { return init(g, prepare_ray(o, d)); }

    __device__ __forceinline__ bool init(const GridView& g, const PreRay& r)
    {
        const float vox = ((r.o.x - g.bmin[0]) * g.ext[0]) * g.rx;
        const float voy = ((r.o.y - g.bmin[1]) * g.ext[1]) * g.ry;
        const float voz = ((r.o.z - g.bmin[2]) * g.ext[2]) * g.rz;
        const float fx = floorf(vox), fy = floorf(voy), fz = floorf(voz);
        steps = -1;
        if (fx < 0.0f || fy < 0.0f || fz < 0.0f || fx >= g.rx || fy >= g.ry || fz >= g.rz) return false;
        disx = (((fx - vox) + 0.5f) + (float)r.sx * 0.5f) * r.inv.x;
        disy = (((fy - voy) + 0.5f) + (float)r.sy * 0.5f) * r.inv.y;
        disz = (((fz - voz) + 0.5f) + (float)r.sz * 0.5f) * r.inv.z;
        /* (mask*sign)*inc with mask in {0,1} is |inc| when the axis steps */
        ax = vt_absf(r.inv.x); ay = vt_absf(r.inv.y); az = vt_absf(r.inv.z);
        stx = r.sx; sty = r.sy; stz = r.sz;
        x = (int32_t)fx; y = (int32_t)fy; z = (int32_t)fz;
        steps = 0;
        key = 0xffffffffu;
        mask = 0ull;
        return true;
    }

    /* One loop iteration per step in the order of dda.h:37-56: step limit,
     * grid bounds, occupancy, then the compare/add step.  The axes that step
     * are those whose dis equals the minimum of the three, i.e.
     * lessThanEqual(dis, min(dis.yzx, dis.zxy)); dis is a running FP32 sum
     * replayed step by step.  Lanes that stop keep their state and step no
     * further (predicated updates), so the loop has one wave-uniform exit and
     * no divergent branches except the block reload. */
    template <bool COUNT>
    __device__ __forceinline__ int advance(const GridView& g, const uint32_t* __restrict__ bits, int budget,
                                           uint32_t& visits)
    {
        const int32_t limit = budget < g.max_steps - steps ? steps + budget : g.max_steps;
        bool done = false;
        /* block-grid strides held in vector registers across the loop */
        uint32_t bx4 = g.bx4, by4 = g.by4;
        asm volatile("" : "+v"(bx4), "+v"(by4));
        for (;;) {
            const bool inside = (uint32_t)x < (uint32_t)g.nx && (uint32_t)y < (uint32_t)g.ny &&
                                (uint32_t)z < (uint32_t)g.nz;
            const bool more = steps < limit;
            /* the block's linear index doubles as its key (meaningful for
             * in-grid voxels only, the only ones that reload) */
            const uint32_t bk = mad24(mad24((uint32_t)z >> 2, by4, (uint32_t)y >> 2), bx4, (uint32_t)x >> 2);
            if (inside && bk != key) {
                key = bk;
                mask = ((bits[bk >> 5] >> (bk & 31u)) & 1u) ? g.occ64[bk] : 0ull;
            }
            /* the voxel test only when some lane's block has an occupied voxel
             * (a wave-uniform branch: rays in empty space skip it) */
            bool hit = false;
            if (__builtin_amdgcn_ballot_w64(!done && inside && mask != 0ull) != 0ull) {
                const uint32_t bit = ((uint32_t)x & 3u) | (((uint32_t)y & 3u) << 2) | (((uint32_t)z & 3u) << 4);
                hit = inside && ((mask >> bit) & 1ull);
            }
            if (COUNT && !done && more && inside) visits++;
            done = done || !more || !inside || hit;
            /* stopped lanes step no axis, so the step can run before the
             * exit test without changing their state */
            const float m = min3f(disx, disy, disz);
            const bool mx = !done && disx == m, my = !done && disy == m, mz = !done && disz == m;
            /* select form of the masked add: adding +0 instead of skipping can
             * only change the sign of a zero dis, which nothing observes */
            disx = disx + (mx ? ax : 0.0f);
            disy = disy + (my ? ay : 0.0f);
            disz = disz + (mz ? az : 0.0f);
            x += mx ? stx : 0;
            y += my ? sty : 0;
            z += mz ? stz : 0;
            steps += done ? 0 : 1;
            if (__builtin_amdgcn_ballot_w64(!done) == 0ull) break;
        }
        /* every lane holds the state it stopped in */
        if (steps >= limit) return steps >= g.max_steps ? 2 : 0;
        const bool inside = (uint32_t)x < (uint32_t)g.nx && (uint32_t)y < (uint32_t)g.ny &&
                            (uint32_t)z < (uint32_t)g.nz;
        return inside ? 1 : 2;
    }

    /* The march of advance() (same steps, same order of the tests, same
     * result) for a whole wave, written for the vector issue rate: the
     * marching lanes are the execution mask (a lane that stops leaves it),
     * the axis updates of a step run under the lane mask of that axis
     * instead of as selects, the step count follows a scalar iteration
     * counter, and which lanes sit in a non-empty block is a scalar lane
     * mask maintained at block changes.  The wave runs at most `budget`
     * iterations and no lane passes its max_steps.  It reads the padded
     * block levels of the grid (bxp, byp, occp; `bits` = bitsp or its LDS
     * copy), where the voxels outside the grid are set, so leaving the grid
     * needs no test of its own.  LDSB: `bits` points into LDS; UNI: every
     * lane marches the same grid and its padded masks span less than 4 GiB
     * (else grid, bits and masks are per lane, as in batch launches). */
    template <bool COUNT, bool LDSB, bool UNI>
    __device__ __forceinline__ int advance_wave(const GridView& g, const uint32_t* __restrict__ bits, int budget,
                                                uint32_t& visits);

    /* advance_wave (same steps, same order of the tests, same result) with
     * the voxel coordinates packed into one register for grids whose
     * extents are at most VT_PACKED_MAX_EXTENT (g.occr set).  UNI: every
     * lane marches the same grid (else g is the lane's own, as in batch
     * launches); LDSB: `bits` (the padded block bits) is in LDS and is
     * tested before a block's mask is loaded. */
    template <bool COUNT, bool UNI, bool LDSB>
    __device__ __forceinline__ int advance_wave_packed(const GridView& g, const uint32_t* __restrict__ bits,
                                                       int budget, uint32_t& visits);

    /* traverse() result of a finished march: kind and hit record. */
    __device__ __forceinline__ int finish(const GridView& g, int status, Hit& h) const
    {
        h.x = x; h.y = y; h.z = z; h.steps = steps;
        h.slot = status == 1 ? g.bricks[(size_t)brick_morton((uint32_t)x >> 3, (uint32_t)y >> 3, (uint32_t)z >> 3) *
                                            512u +
                                        local_morton((uint32_t)x, (uint32_t)y, (uint32_t)z)]
                             : 0u;
        if (status == 1) return 2;
        return y < 0 ? 3 : 0;
    }
};

/* ---- wave march: the loop of DdaState::advance_wave ---------------------------------
 * Coordinates inside the loop are offset by +4 (one block), so the voxel at
 * x = -1 has a non-negative block index in the padded block grid.  Per
 * iteration, for the marching lanes (exec):
 *   key  = linear index of the voxel's padded 4x4x4 block; lanes whose key
 *          changed reload: block bit from `bits`, then the block's voxel
 *          mask; ne = lanes whose block is non-empty (kept in bit 31 of key
 *          between calls)
 *   hit  = ne & voxel bit set (tested only when some lane is in a non-empty
 *          block); voxels outside the grid are set in the padded masks, so
 *          a hit is also how a ray leaves the grid
 *   lanes that hit stop: steps += iteration, they leave exec
 *   step: m = min3(dis); each axis whose dis equals m adds its increment
 *   to dis and its sign to the coordinate, under that axis's lane mask.
 * At the end lanes still marching add the iteration count to steps, the
 * entry mask is restored and t1 is 1 on the lanes still marching. */
#define VT_WAVE_DDA_HEAD                                                                  \
    "s_mov_b64 %[live], exec\n\t"                                                         \
    "v_cmp_gt_i32_e64 %[ne], 0, %[key]\n\t"                                               \
    "v_bfe_u32 %[key], %[key], 0, 31\n\t"                                                 \
    "s_mov_b32 %[it], 0\n"                                                                \
    ".Lvt_top%=:\n\t"                                                                     \
    "s_cmp_ge_u32 %[it], %[n]\n\t"                                                        \
    "s_cbranch_scc1 .Lvt_end%=\n\t"                                                       \
    "v_lshrrev_b32 %[t0], 2, %[z]\n\t"                                                    \
    "v_lshrrev_b32 %[t1], 2, %[y]\n\t"                                                    \
    "v_mad_u32_u24 %[t0], %[t0], %[by4], %[t1]\n\t"                                       \
    "v_lshrrev_b32 %[t1], 2, %[x]\n\t"                                                    \
    "v_mad_u32_u24 %[t0], %[t0], %[bx4], %[t1]\n\t"                                       \
    "v_cmp_ne_u32_e64 %[s0], %[t0], %[key]\n\t"                                           \
    "s_and_saveexec_b64 %[s1], %[s0]\n\t"                                                 \
    "s_cbranch_scc0 .Lvt_norl%=\n\t"                                                      \
    "v_mov_b32 %[key], %[t0]\n\t"                                                         \
    "v_lshrrev_b32 %[t1], 5, %[t0]\n\t"
/* between HEAD and MID: t1 = 32-bit word of the block bits holding bit t0 */
#define VT_WAVE_DDA_BITS_LDS                                                              \
    "v_lshl_add_u32 %[t1], %[t1], 2, %[bits]\n\t"                                         \
    "ds_read_b32 %[t1], %[t1]\n\t"                                                        \
    "s_waitcnt lgkmcnt(0)\n\t"
#define VT_WAVE_DDA_BITS_GLOBAL                                                           \
    "v_lshlrev_b32 %[t1], 2, %[t1]\n\t"                                                   \
    "global_load_dword %[t1], %[t1], %[bits]\n\t"                                         \
    "s_waitcnt vmcnt(0)\n\t"
#define VT_WAVE_DDA_BITS_LANE                                                             \
    "v_mad_u64_u32 %[t64], vcc, %[t1], 4, %[bits]\n\t"                                    \
    "global_load_dword %[t1], %[t64], off\n\t"                                            \
    "s_waitcnt vmcnt(0)\n\t"
#define VT_WAVE_DDA_MID                                                                   \
    "v_bfe_u32 %[t1], %[t1], %[t0], 1\n\t"                                                \
    "v_cmp_ne_u32_e64 %[s0], 0, %[t1]\n\t"                                                \
    "s_andn2_b64 %[ne], %[ne], exec\n\t"                                                  \
    "s_or_b64 %[ne], %[ne], %[s0]\n\t"                                                    \
    "s_and_b64 exec, exec, %[s0]\n\t"                                                     \
    "s_cbranch_execz .Lvt_norl%=\n\t"
/* between MID and TAIL: mask = voxel mask of block t0 */
#define VT_WAVE_DDA_MASK_GLOBAL                                                           \
    "v_lshlrev_b32 %[t1], 3, %[t0]\n\t"                                                   \
    "global_load_dwordx2 %[mask], %[t1], %[occ]\n\t"                                      \
    "s_waitcnt vmcnt(0)\n"
#define VT_WAVE_DDA_MASK_LANE                                                             \
    "v_mad_u64_u32 %[t64], vcc, %[t0], 8, %[occ]\n\t"                                     \
    "global_load_dwordx2 %[mask], %[t64], off\n\t"                                        \
    "s_waitcnt vmcnt(0)\n"
#define VT_WAVE_DDA_TAIL                                                                  \
    ".Lvt_norl%=:\n\t"                                                                    \
    "s_mov_b64 exec, %[s1]\n\t"                                                           \
    "s_and_b64 %[s0], %[ne], exec\n\t"                                                    \
    "s_cbranch_scc0 .Lvt_go%=\n\t"                                                        \
    "s_and_saveexec_b64 %[s1], %[s0]\n\t"                                                 \
    "v_lshlrev_b32 %[t1], 2, %[y]\n\t"                                                    \
    "v_bfi_b32 %[t0], 3, %[x], %[t1]\n\t"                                                 \
    "v_lshlrev_b32 %[t1], 4, %[z]\n\t"                                                    \
    "v_bfi_b32 %[t0], 15, %[t0], %[t1]\n\t"                                               \
    "v_not_b32 %[t0], %[t0]\n\t"                                                          \
    "v_lshlrev_b64 %[t64], %[t0], %[mask]\n\t"                                            \
    "v_cmp_gt_i64_e64 %[s0], 0, %[t64]\n\t"                                               \
    "s_mov_b64 exec, %[s1]\n\t"                                                           \
    "s_cmp_eq_u64 %[s0], 0\n\t"                                                           \
    "s_cbranch_scc1 .Lvt_go%=\n\t"                                                        \
    "s_andn2_b64 %[s1], exec, %[s0]\n\t"                                                  \
    "s_mov_b64 exec, %[s0]\n\t"                                                           \
    "v_add_u32 %[steps], %[it], %[steps]\n\t"                                             \
    "s_mov_b64 exec, %[s1]\n\t"                                                           \
    "s_cbranch_execz .Lvt_out%=\n"                                                        \
    ".Lvt_go%=:\n\t"                                                                      \
    "v_min3_f32 %[t0], %[dx], %[dy], %[dz]\n\t"                                           \
    "s_mov_b64 %[s1], exec\n\t"                                                           \
    "v_cmp_eq_f32_e64 %[s0], %[dx], %[t0]\n\t"                                            \
    "v_cmp_eq_f32_e64 %[s2], %[dy], %[t0]\n\t"                                            \
    "v_cmp_eq_f32_e64 vcc, %[dz], %[t0]\n\t"                                              \
    "s_mov_b64 exec, %[s0]\n\t"                                                           \
    "v_add_f32 %[dx], %[dx], %[ax]\n\t"                                                   \
    "v_add_u32 %[x], %[x], %[sx]\n\t"                                                     \
    "s_mov_b64 exec, %[s2]\n\t"                                                           \
    "v_add_f32 %[dy], %[dy], %[ay]\n\t"                                                   \
    "v_add_u32 %[y], %[y], %[sy]\n\t"                                                     \
    "s_mov_b64 exec, vcc\n\t"                                                             \
    "v_add_f32 %[dz], %[dz], %[az]\n\t"                                                   \
    "v_add_u32 %[z], %[z], %[sz]\n\t"                                                     \
    "s_mov_b64 exec, %[s1]\n\t"                                                           \
    "s_add_u32 %[it], %[it], 1\n\t"                                                       \
    "s_branch .Lvt_top%=\n"                                                               \
    ".Lvt_end%=:\n\t"                                                                     \
    "v_add_u32 %[steps], %[it], %[steps]\n"                                               \
    ".Lvt_out%=:\n\t"                                                                     \
    "s_mov_b64 %[s0], exec\n\t"                                                           \
    "s_and_b64 exec, %[live], %[ne]\n\t"                                                  \
    "v_or_b32 %[key], 0x80000000, %[key]\n\t"                                             \
    "s_mov_b64 exec, %[live]\n\t"                                                         \
    "v_cndmask_b32_e64 %[t1], 0, 1, %[s0]\n\t"

#define VT_WAVE_DDA_OUTS                                                                  \
    [x] "+v"(xo), [y] "+v"(yo), [z] "+v"(zo), [dx] "+v"(disx), [dy] "+v"(disy), [dz] "+v"(disz),         \
        [steps] "+v"(steps), [key] "+v"(key), [mask] "+v"(mask), [t0] "=&v"(t0), [t1] "=&v"(t1),         \
        [t64] "=&v"(t64), [live] "=&s"(s_live), [ne] "=&s"(s_ne), [s0] "=&s"(s_0), [s1] "=&s"(s_1),      \
        [s2] "=&s"(s_2), [it] "=&s"(it)
#define VT_WAVE_DDA_INS                                                                   \
    [ax] "v"(ax), [ay] "v"(ay), [az] "v"(az), [sx] "v"(stx), [sy] "v"(sty), [sz] "v"(stz), [n] "s"(n)

template <bool COUNT, bool LDSB, bool UNI>
__device__ __forceinline__ int DdaState::advance_wave(const GridView& g, const uint32_t* __restrict__ bits,
                                                      int budget, uint32_t& visits)
{
    static_assert(UNI || !LDSB, "per-lane grids keep their block bits in global memory");
    /* iterations: the budget, cut to the fewest steps any lane has left */
    const int room = g.max_steps - steps;
    int n = budget;
    for (uint64_t near = __builtin_amdgcn_ballot_w64(room < budget); near; near &= near - 1ull) {
        const int r = __builtin_amdgcn_readlane(room, (int)__builtin_ctzll(near));
        n = r < n ? r : n;
    }
    n = n > 0 ? n : 0;
    const int32_t steps0 = steps;
    int32_t xo = x + 4, yo = y + 4, zo = z + 4;
    uint32_t t0, t1;
    uint64_t t64, s_live, s_ne, s_0, s_1, s_2;
    uint32_t it;
    if constexpr (LDSB) {
        const uint32_t lds = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const uint32_t*)bits;
        asm volatile(VT_WAVE_DDA_HEAD VT_WAVE_DDA_BITS_LDS VT_WAVE_DDA_MID VT_WAVE_DDA_MASK_GLOBAL VT_WAVE_DDA_TAIL
                     : VT_WAVE_DDA_OUTS
                     : VT_WAVE_DDA_INS, [bx4] "s"(g.bxp), [by4] "s"(g.byp), [bits] "s"(lds), [occ] "s"(g.occp)
                     : "vcc", "scc", "memory");
    } else if constexpr (UNI) {
        asm volatile(VT_WAVE_DDA_HEAD VT_WAVE_DDA_BITS_GLOBAL VT_WAVE_DDA_MID VT_WAVE_DDA_MASK_GLOBAL
                         VT_WAVE_DDA_TAIL
                     : VT_WAVE_DDA_OUTS
                     : VT_WAVE_DDA_INS, [bx4] "s"(g.bxp), [by4] "s"(g.byp), [bits] "s"(bits), [occ] "s"(g.occp)
                     : "vcc", "scc", "memory");
    } else {
        asm volatile(VT_WAVE_DDA_HEAD VT_WAVE_DDA_BITS_LANE VT_WAVE_DDA_MID VT_WAVE_DDA_MASK_LANE VT_WAVE_DDA_TAIL
                     : VT_WAVE_DDA_OUTS
                     : VT_WAVE_DDA_INS, [bx4] "v"(g.bxp), [by4] "v"(g.byp), [bits] "v"(bits), [occ] "v"(g.occp)
                     : "vcc", "scc", "memory");
    }
    x = xo - 4; y = yo - 4; z = zo - 4;
    const bool marching = t1 != 0u;
    const bool inside = (uint32_t)x < (uint32_t)g.nx && (uint32_t)y < (uint32_t)g.ny && (uint32_t)z < (uint32_t)g.nz;
    /* a visit per iteration spent inside the grid under the limit: every
     * step, plus the voxel a hit stopped in */
    if (COUNT) visits += (uint32_t)(steps - steps0) + (!marching && inside ? 1u : 0u);
    if (marching) return steps >= g.max_steps ? 2 : 0;
    return inside ? 1 : 2;
}

/* ---- packed wave march ---------------------------------------------------------------
 * The loop of advance_wave over P = (x+4) | (y+4) << 10 | (z+4) << 20, the
 * padded coordinates of the voxel in 10-bit fields.  An axis step adds that
 * axis's sign in its field (+-1, +-1 << 10, +-1 << 20); no field leaves
 * [3, extent + 4] because the march stops at the first voxel outside the
 * grid.  Per iteration, for the marching lanes:
 *   key  = P & 0x3fcff3fc (each field without its two low bits): the
 *          voxel's padded block (bx, by, bz) at bits 2, 12 and 22, so a
 *          block change is one AND and one compare; lanes whose key changed
 *          form the block's linear index bx + bxp (by + byp bz) from it and
 *          reload: (LDS: the block's bit first, then) the block's mask from
 *          g.occr; ne = lanes whose block is non-empty
 *   hit  = bit (x&3) + 4 (y&3) + 16 (z&3) of the mask: the byte sum of
 *          P & 0x00300c03 (x&3 in byte 0, (y&3) << 2 in byte 1, (z&3) << 4
 *          in byte 2) is that index, and the masks are stored bit-reversed,
 *          so shifting left by it puts the voxel's bit in the sign
 *   step, step count and exit as in advance_wave. */
#define VT_PACKED_MAX_EXTENT 1016
#define VT_WAVE_DDA_PACKED_HEAD                                                           \
    "s_mov_b64 %[live], exec\n\t"                                                         \
    "v_cmp_gt_i32_e64 %[ne], 0, %[key]\n\t"                                               \
    "v_bfe_u32 %[key], %[key], 0, 31\n\t"                                                 \
    "s_mov_b32 %[it], 0\n"                                                                \
    ".Lvp_top%=:\n\t"                                                                     \
    "s_cmp_ge_u32 %[it], %[n]\n\t"                                                        \
    "s_cbranch_scc1 .Lvp_end%=\n\t"                                                       \
    "v_and_b32 %[t0], 0x3fcff3fc, %[p]\n\t"                                               \
    "v_cmp_ne_u32_e64 %[s0], %[t0], %[key]\n\t"                                           \
    "s_and_saveexec_b64 %[s1], %[s0]\n\t"                                                 \
    "s_cbranch_scc0 .Lvp_norl%=\n\t"                                                      \
    "v_mov_b32 %[key], %[t0]\n\t"                                                         \
    "v_lshrrev_b32 %[t1], 22, %[t0]\n\t"                                                  \
    "v_bfe_u32 %[t2], %[t0], 12, 8\n\t"                                                   \
    "v_mad_u32_u24 %[t1], %[t1], %[byp], %[t2]\n\t"                                       \
    "v_bfe_u32 %[t2], %[t0], 2, 8\n\t"                                                    \
    "v_mad_u32_u24 %[t1], %[t1], %[bxp], %[t2]\n\t"
/* between HEAD and TAIL: t1 = linear index of the block; mask = its mask
 * after its block bit in LDS (LDS), or directly from the lane's own array
 * (LANE, batch launches); ne updated for the lanes that reloaded */
#define VT_WAVE_DDA_PACKED_LANE                                                           \
    "v_mad_u64_u32 %[t64], vcc, %[t1], 8, %[occ]\n\t"                                     \
    "global_load_dwordx2 %[mask], %[t64], off\n\t"                                        \
    "s_waitcnt vmcnt(0)\n\t"                                                              \
    "v_cmp_ne_u64_e64 %[s0], 0, %[mask]\n\t"                                              \
    "s_andn2_b64 %[ne], %[ne], exec\n\t"                                                  \
    "s_or_b64 %[ne], %[ne], %[s0]\n"
#define VT_WAVE_DDA_PACKED_LDS                                                            \
    "v_lshrrev_b32 %[t2], 5, %[t1]\n\t"                                                   \
    "v_lshl_add_u32 %[t2], %[t2], 2, %[bits]\n\t"                                         \
    "ds_read_b32 %[t2], %[t2]\n\t"                                                        \
    "s_waitcnt lgkmcnt(0)\n\t"                                                            \
    "v_bfe_u32 %[t2], %[t2], %[t1], 1\n\t"                                                \
    "v_cmp_ne_u32_e64 %[s0], 0, %[t2]\n\t"                                                \
    "s_andn2_b64 %[ne], %[ne], exec\n\t"                                                  \
    "s_or_b64 %[ne], %[ne], %[s0]\n\t"                                                    \
    "s_and_b64 exec, exec, %[s0]\n\t"                                                     \
    "s_cbranch_execz .Lvp_norl%=\n\t"                                                     \
    "v_lshlrev_b32 %[t1], 3, %[t1]\n\t"                                                   \
    "global_load_dwordx2 %[mask], %[t1], %[occ]\n\t"                                      \
    "s_waitcnt vmcnt(0)\n"
#define VT_WAVE_DDA_PACKED_TAIL                                                           \
    ".Lvp_norl%=:\n\t"                                                                    \
    "s_mov_b64 exec, %[s1]\n\t"                                                           \
    "s_and_b64 %[s0], %[ne], exec\n\t"                                                    \
    "s_cbranch_scc0 .Lvp_go%=\n\t"                                                        \
    "s_and_saveexec_b64 %[s1], %[s0]\n\t"                                                 \
    "v_and_b32 %[t0], 0x300c03, %[p]\n\t"                                                 \
    "v_sad_u8 %[t0], %[t0], 0, 0\n\t"                                                     \
    "v_lshlrev_b64 %[t64], %[t0], %[mask]\n\t"                                            \
    "v_cmp_gt_i64_e64 %[s0], 0, %[t64]\n\t"                                               \
    "s_mov_b64 exec, %[s1]\n\t"                                                           \
    "s_cmp_eq_u64 %[s0], 0\n\t"                                                           \
    "s_cbranch_scc1 .Lvp_go%=\n\t"                                                        \
    "s_andn2_b64 %[s1], exec, %[s0]\n\t"                                                  \
    "s_mov_b64 exec, %[s0]\n\t"                                                           \
    "v_add_u32 %[steps], %[it], %[steps]\n\t"                                             \
    "s_mov_b64 exec, %[s1]\n\t"                                                           \
    "s_cbranch_execz .Lvp_out%=\n"                                                        \
    ".Lvp_go%=:\n\t"                                                                      \
    "v_min3_f32 %[t0], %[dx], %[dy], %[dz]\n\t"                                           \
    "s_mov_b64 %[s1], exec\n\t"                                                           \
    "v_cmp_eq_f32_e64 %[s0], %[dx], %[t0]\n\t"                                            \
    "v_cmp_eq_f32_e64 %[s2], %[dy], %[t0]\n\t"                                            \
    "v_cmp_eq_f32_e64 vcc, %[dz], %[t0]\n\t"                                              \
    "s_mov_b64 exec, %[s0]\n\t"                                                           \
    "v_add_f32 %[dx], %[dx], %[ax]\n\t"                                                   \
    "v_add_u32 %[p], %[p], %[ix]\n\t"                                                     \
    "s_mov_b64 exec, %[s2]\n\t"                                                           \
    "v_add_f32 %[dy], %[dy], %[ay]\n\t"                                                   \
    "v_add_u32 %[p], %[p], %[iy]\n\t"                                                     \
    "s_mov_b64 exec, vcc\n\t"                                                             \
    "v_add_f32 %[dz], %[dz], %[az]\n\t"                                                   \
    "v_add_u32 %[p], %[p], %[iz]\n\t"                                                     \
    "s_mov_b64 exec, %[s1]\n\t"                                                           \
    "s_add_u32 %[it], %[it], 1\n\t"                                                       \
    "s_branch .Lvp_top%=\n"                                                               \
    ".Lvp_end%=:\n\t"                                                                     \
    "v_add_u32 %[steps], %[it], %[steps]\n"                                               \
    ".Lvp_out%=:\n\t"                                                                     \
    "s_mov_b64 %[s0], exec\n\t"                                                           \
    "s_and_b64 exec, %[live], %[ne]\n\t"                                                  \
    "v_or_b32 %[key], 0x80000000, %[key]\n\t"                                             \
    "s_mov_b64 exec, %[live]\n\t"                                                         \
    "v_cndmask_b32_e64 %[t1], 0, 1, %[s0]\n\t"

template <bool COUNT, bool UNI, bool LDSB>
__device__ __forceinline__ int DdaState::advance_wave_packed(const GridView& g, const uint32_t* __restrict__ bits,
                                                             int budget, uint32_t& visits)
{
    static_assert(UNI == LDSB, "one grid: block bits in LDS; per-lane grids: masks read directly");
    const int room = g.max_steps - steps;
    int n = budget;
    for (uint64_t near = __builtin_amdgcn_ballot_w64(room < budget); near; near &= near - 1ull) {
        const int r = __builtin_amdgcn_readlane(room, (int)__builtin_ctzll(near));
        n = r < n ? r : n;
    }
    n = n > 0 ? n : 0;
    const int32_t steps0 = steps;
    uint32_t p = (uint32_t)(x + 4) | ((uint32_t)(y + 4) << 10) | ((uint32_t)(z + 4) << 20);
    const int32_t ix = stx, iy = (int32_t)((uint32_t)sty << 10), iz = (int32_t)((uint32_t)stz << 20);
    uint32_t t0, t1, t2;
    uint64_t t64, s_live, s_ne, s_0, s_1, s_2;
    uint32_t it;
#define VT_WAVE_DDA_PACKED_OUTS                                                           \
    [p] "+v"(p), [dx] "+v"(disx), [dy] "+v"(disy), [dz] "+v"(disz), [steps] "+v"(steps), [key] "+v"(key), \
        [mask] "+v"(mask), [t0] "=&v"(t0), [t1] "=&v"(t1), [t2] "=&v"(t2), [t64] "=&v"(t64),              \
        [live] "=&s"(s_live), [ne] "=&s"(s_ne), [s0] "=&s"(s_0), [s1] "=&s"(s_1), [s2] "=&s"(s_2),          \
        [it] "=&s"(it)
#define VT_WAVE_DDA_PACKED_INS                                                            \
    [ax] "v"(ax), [ay] "v"(ay), [az] "v"(az), [ix] "v"(ix), [iy] "v"(iy), [iz] "v"(iz), [n] "s"(n)
    if constexpr (LDSB) {
        const uint32_t lds = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) const uint32_t*)bits;
        asm volatile(VT_WAVE_DDA_PACKED_HEAD VT_WAVE_DDA_PACKED_LDS VT_WAVE_DDA_PACKED_TAIL
                     : VT_WAVE_DDA_PACKED_OUTS
                     : VT_WAVE_DDA_PACKED_INS, [occ] "s"(g.occr), [bxp] "s"(g.bxp), [byp] "s"(g.byp), [bits] "s"(lds)
                     : "vcc", "scc", "memory");
    } else {
        asm volatile(VT_WAVE_DDA_PACKED_HEAD VT_WAVE_DDA_PACKED_LANE VT_WAVE_DDA_PACKED_TAIL
                     : VT_WAVE_DDA_PACKED_OUTS
                     : VT_WAVE_DDA_PACKED_INS, [occ] "v"(g.occr), [bxp] "v"(g.bxp), [byp] "v"(g.byp)
                     : "vcc", "scc", "memory");
    }
#undef VT_WAVE_DDA_PACKED_OUTS
#undef VT_WAVE_DDA_PACKED_INS
    x = (int32_t)(p & 1023u) - 4;
    y = (int32_t)((p >> 10) & 1023u) - 4;
    z = (int32_t)(p >> 20) - 4;
    const bool marching = t1 != 0u;
    const bool inside = (uint32_t)x < (uint32_t)g.nx && (uint32_t)y < (uint32_t)g.ny && (uint32_t)z < (uint32_t)g.nz;
    if (COUNT) visits += (uint32_t)(steps - steps0) + (!marching && inside ? 1u : 0u);
    if (marching) return steps >= g.max_steps ? 2 : 0;
    return inside ? 1 : 2;
}

/* ---- 3D-DDA (dda.h:7-61) -----------------------------------------------------------
 * Returns 1 on hit, 0 when the march ended without a hit (position written),
 * -1 when the start point lies outside the grid (position not written).
 * `slot` receives the hit voxel's material slot. */
template <bool COUNT>
__device__ __forceinline__ int dda(const GridView& g, const uint32_t* __restrict__ bits, f3 o, f3 d, Hit& h,
                                   uint32_t& visits)
{
    DdaState D;
    if (!D.init(g, o, d)) {
        h.steps = -1;
        return -1;
    }
    const int status = D.advance<COUNT>(g, bits, g.max_steps, visits);
    D.finish(g, status, h);
    return status == 1 ? 1 : 0;
}

/* traverse (dda.h:63-79): 2 voxel hit, 3 ground hit (final y < 0), 0 miss. */
enum { KIND_MISS = 0, KIND_VOXEL = 2, KIND_GROUND = 3 };

template <bool COUNT>
__device__ __forceinline__ int traverse(const GridView& g, const uint32_t* bits, f3 o, f3 d, Hit& h, uint32_t& visits)
{
    const int r = dda<COUNT>(g, bits, o, d, h, visits);
    if (r > 0) return KIND_VOXEL;
    if (r < 0) return KIND_MISS;
    return h.y < 0 ? KIND_GROUND : KIND_MISS;
}

/* Slot of a voxel position: texelFetch semantics, outside the grid reads
 * offset 0, which is record index 0 of the decoded table. */
__device__ __forceinline__ uint32_t fetch_slot(const GridView& g, const uint32_t* bits, int32_t x, int32_t y, int32_t z)
{
    if ((uint32_t)x >= (uint32_t)g.nx || (uint32_t)y >= (uint32_t)g.ny || (uint32_t)z >= (uint32_t)g.nz) return 0u;
    const uint32_t bm = brick_morton((uint32_t)x >> 3, (uint32_t)y >> 3, (uint32_t)z >> 3);
    return g.bricks[(size_t)bm * 512u + local_morton((uint32_t)x, (uint32_t)y, (uint32_t)z)];
}

/* ---- voxelSpaceToWorldSpace (coordinates.h:59-82) ---------------------------------- */
struct Basis { f3 p, t, n, b; };

__device__ __forceinline__ Basis voxel_basis(const GridView& g, float vx, float vy, float vz, f3 o, f3 d)
{
    Basis B;
    const f3 vs = mk(g.vs[0], g.vs[1], g.vs[2]);
    const f3 vmin = mk(vx, vy, vz) * vs + mk(g.bmin[0], g.bmin[1], g.bmin[2]);
    const f3 vmax = vmin + vs;
    const float t = aabb(o, d, vmin, vmax);
    B.p = o + d * t;
    const f3 c = vmin + vs * 0.5f;
    const f3 c2h = B.p - c;
    const float ax = vt_absf(c2h.x), ay = vt_absf(c2h.y), az = vt_absf(c2h.z);
    const f3 m = mk(vt_stepf(ay, ax) * vt_stepf(az, ax), vt_stepf(ax, ay) * vt_stepf(az, ay),
                    vt_stepf(ax, az) * vt_stepf(ay, az));
    B.n = m * mk(vt_signf(c2h.x), vt_signf(c2h.y), vt_signf(c2h.z));
    const float k = 0.57735026919f;
    const f3 t0 = cross(B.n, mk(k, k, k));
    B.b = normalize(cross(t0, B.n));
    B.t = normalize(cross(B.b, B.n));
    return B;
}

__device__ __forceinline__ f3 local_to_world(f3 v, const Basis& B) { return (B.t * v.x + B.n * v.y) + B.b * v.z; }
__device__ __forceinline__ f3 world_to_local(f3 v, const Basis& B) { return mk(dot(v, B.t), dot(v, B.n), dot(v, B.b)); }

/* ---- camera rays (generateRay.h, coordinates.h:1-27) ------------------------------- */
__device__ __forceinline__ void screen_to_eye(const CamView& c, float wx, float wy, float wz, float* eye)
{
    const float nx = ((2.0f * wx) - (2.0f * c.viewport[0])) / c.viewport[2] - 1.0f;
    const float ny = ((2.0f * wy) - (2.0f * c.viewport[1])) / c.viewport[3] - 1.0f;
    const float nz = ((2.0f * wz) - 0.0f - 1.0f) / (1.0f - 0.0f);
    const float w = c.P32 / (nz - (c.P22 / c.P23));
    mat4_mul(c.inv_proj, nx * w, ny * w, nz * w, w, eye);
}

__device__ __forceinline__ void gen_ray(const CamView& c, float fx, float fy, const float* u, f3& o, f3& d)
{
    float tmp[4];
    if (c.lens_model == 0) {
        float es[4];
        screen_to_eye(c, fx + (u[0] - 0.5f), fy + (u[1] - 0.5f), c.win_z + 0.0f, es);
        mat4_mul(c.inv_mv, 0.0f, 0.0f, 0.0f, 1.0f, tmp);
        o = mk(tmp[0], tmp[1], tmp[2]);
        const f3 ed = normalize(mk(es[0], es[1], es[2]));
        mat4_mul(c.inv_mv, ed.x, ed.y, ed.z, 0.0f, tmp);
        const float l2 = ((tmp[0] * tmp[0] + tmp[1] * tmp[1]) + tmp[2] * tmp[2]) + tmp[3] * tmp[3];
        const float inv = 1.0f / sqrtf(l2);
        d = mk(tmp[0] * inv, tmp[1] * inv, tmp[2] * inv);
    } else if (c.lens_model == 1) {
        const float r = sqrtf(u[0]);
        float sn, cs;
        vt_sincosf((2.0f * VT_PI_F) * u[1], &sn, &cs);
        const float lx = (r * cs) * c.lens_radius, ly = (r * sn) * c.lens_radius;
        float eo[4];
        screen_to_eye(c, fx, fy, c.win_z, eo);
        const f3 esO = mk(eo[0], eo[1], eo[2]);
        const f3 esD = normalize(esO);
        const float t = c.focal_distance / -(esD.z);
        const f3 esF = esO + esD * t;
        mat4_mul(c.inv_mv, esF.x, esF.y, esF.z, 1.0f, tmp);
        const f3 wsF = mk(tmp[0], tmp[1], tmp[2]);
        mat4_mul(c.inv_mv, lx, ly, 0.0f, 1.0f, tmp);
        o = mk(tmp[0], tmp[1], tmp[2]);
        d = normalize(wsF - o);
    } else {
        float eye[4];
        const float nx = (fx / c.viewport[2]) * 2.0f - 1.0f;
        const float ny = (fy / c.viewport[3]) * 2.0f - 1.0f;
        const float nz = ((2.0f * c.win_z) - 0.0f - 1.0f) / (1.0f - 0.0f);
        mat4_mul(c.inv_proj, nx, ny, nz, 1.0f, eye);
        mat4_mul(c.inv_mv, eye[0], eye[1], eye[2], eye[3], tmp);
        o = mk(tmp[0], tmp[1], tmp[2]);
        mat4_mul(c.inv_mv, 0.0f, 0.0f, -1.0f, 0.0f, tmp);
        d = mk(tmp[0], tmp[1], tmp[2]);
    }
}

/* ---- environment (lights.h, coordinates.h:97-116, envMapSample.h) ------------------- */
__device__ __forceinline__ f3 tex_bilinear(const EnvView& e, float u, float v)
{
    if (vt_isnan(u) || vt_isnan(v)) return mk(vt_nanf(), vt_nanf(), vt_nanf());
    u = vt_minf(vt_maxf(u, 0.0f), 1.0f);
    v = vt_minf(vt_maxf(v, 0.0f), 1.0f);
    const float s = u * (float)e.w - 0.5f;
    const float t = v * (float)e.h - 0.5f;
    const float fs = floorf(s), ft = floorf(t);
    const int32_t i0 = (int32_t)fs, j0 = (int32_t)ft;
    const float a = s - fs, b = t - ft;
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    /* i0, j0 lie in [-1, w - 1] x [-1, h - 1]: one quad holds the four texels */
    const float4* q = e.quad + ((size_t)(j0 + 1) * (size_t)(e.w + 1) + (size_t)(i0 + 1)) * 4u;
    const float4 q0 = q[0], q1 = q[1], q2 = q[2];
    const float t00[3] = { q0.x, q0.y, q0.z }, t10[3] = { q0.w, q1.x, q1.y };
    const float t01[3] = { q1.z, q1.w, q2.x }, t11[3] = { q2.y, q2.z, q2.w };
    float out[3];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) out[ch] = ((w00 * t00[ch] + w10 * t10[ch]) + w01 * t01[ch]) + w11 * t11[ch];
    return mk(out[0], out[1], out[2]);
}

__device__ __forceinline__ void uv_from_dir(f3 d, float rot, float& u, float& v)
{
    const float theta = vt_acosf(d.y);
    const float phi = (vt_atan2f(d.z, d.x) + VT_PI_F) + rot;
    const float a = phi / VT_TWO_PI_F;
    u = a - floorf(a);
    v = theta / VT_PI_F;
}

__device__ __forceinline__ f3 dir_from_uv(float u, float v, float rot)
{
    const float phi = ((u * VT_TWO_PI_F) - VT_PI_F) - rot;
    const float theta = v * VT_PI_F;
    float st, ct, sp, cp;
    vt_sincosf(theta, &st, &ct);
    vt_sincosf(phi, &sp, &cp);
    return mk(st * cp, ct, st * sp);
}

__device__ __forceinline__ float luminance(f3 c) { return (c.x * 0.2126f + c.y * 0.7152f) + c.z * 0.0722f; }

__device__ __forceinline__ f3 background(const EnvView& e, f3 v)
{
    if (e.use_image) {
        float u, w;
        uv_from_dir(v, e.rotation, u, w);
        return tex_bilinear(e, u, w);
    }
    const float bias = vt_maxf(0.0f, v.y);
    return mk(e.bottom[0], e.bottom[1], e.bottom[2]) * (1.0f - bias) + mk(e.top[0], e.top[1], e.top[2]) * bias;
}

__device__ __forceinline__ int bsearch_cdf(const float* cdf, int hi, float x)
{
    int lo = 0;
    if (hi < 1) return 0;
    while (lo != hi - 1) {
        const int mid = (lo + hi) / 2;
        if (x < cdf[mid]) hi = mid; else lo = mid;
    }
    return lo;
}

__device__ __forceinline__ float cdf_at(const float* cdf, int n, int i) { return (i < 0 || i >= n) ? 0.0f : cdf[i]; }

/* sampleEnvironmentTexture (envMapSample.h:23-65); cdfv may point to LDS. */
__device__ __forceinline__ void env_sample_uv(const EnvView& e, const float* cdfv, float su, float sv, float& u, float& v)
{
    const int sizeU = e.cu_w, sizeV = e.cv_n;
    const int row = bsearch_cdf(cdfv, sizeV - 1, sv);
    const float* rowp = e.cdfu + (size_t)row * (size_t)sizeU;
    const int col = bsearch_cdf(rowp, sizeU - 1, su);
    const float loU = cdf_at(rowp, sizeU, col), upU = cdf_at(rowp, sizeU, col + 1);
    const float du = (su - loU) / (upU - loU);
    const float loV = cdf_at(cdfv, sizeV, row), upV = cdf_at(cdfv, sizeV, row + 1);
    const float dv = (sv - loV) / (upV - loV);
    u = ((float)col + du) / (float)(sizeU - 1);
    v = ((float)row + dv) / (float)(sizeV - 1);
}

/* ---- BSDFs (lambertian.h, microfacet.h) -------------------------------------------- */
__device__ __forceinline__ float power_heuristic(float f, float g) { return (f * f) / ((f * f) + (g * g)); }

__device__ __forceinline__ float G_term(f3 wo, f3 wi, f3 wh)
{
    const float WoDotWh = vt_absf(dot(wo, wh));
    return vt_minf(1.0f, vt_minf(((2.0f * wh.y) * wo.y) / WoDotWh, ((2.0f * wh.y) * wi.y) / WoDotWh));
}

__device__ __forceinline__ void D_term(f3 refl, float e, f3 wo, f3 wh, f3& f, float& pdf)
{
    const float pc = vt_powf(wh.y, e);
    const float fv = ((e + 2.0f) * VT_INV_TWOPI_F) * pc;
    const float woDotWh = dot(wo, wh);
    pdf = (woDotWh <= 0.0f) ? 0.0f : ((e + 1.0f) * pc) / ((VT_TWO_PI_F * 4.0f) * woDotWh);
    f = refl * fv;
}

__device__ __forceinline__ float F_term(float c)
{
    const float ior = 7.3f;
    const float sq = (1.0f - ior) / (1.0f + ior);
    const float r0 = sq * sq;
    return r0 + (1.0f - r0) * vt_powf(1.0f - c, 5.0f);
}

__device__ __forceinline__ void eval_microfacet(f3 refl, float e, f3 wo, f3 wi, f3& f, float& pdf)
{
    const float cWi = wi.y, cWo = wo.y;
    const f3 wh = normalize(wi + wo);
    const float cH = dot(wi, wh);
    D_term(refl, e, wo, wh, f, pdf);
    const float k = (G_term(wo, wi, wh) * F_term(cH)) / ((4.0f * cWo) * cWi);
    f = f * k;
}

__device__ __forceinline__ f3 rec_r(const MatRec& m) { return mk(m.r[0], m.r[1], m.r[2]); }

__device__ __forceinline__ void eval_material(const MatRec& m, f3 wo, f3 wi, f3& f, float& pdf)
{
    if (m.type == 0) {
        pdf = wi.y / VT_PI_F;
        f = rec_r(m) / VT_PI_F;
    } else if (m.type == 1) {
        eval_microfacet(rec_r(m), m.rough, wo, wi, f, pdf);
    } else if (m.type == 2) {
        eval_microfacet(mk(1.0f, 1.0f, 1.0f), m.rough, wo, wi, f, pdf);
    } else {
        f = mk(0.0f, 0.0f, 0.0f);
        pdf = 0.0f;
    }
}

/* sampleMaterialBSDF (materials.h:30-43).  Draws one rand() except for an
 * unknown material type. */
__device__ __forceinline__ f3 sample_material(const MatRec& m, f3 wo, vt_pcg32& rng, f3& f, float& pdf)
{
    if (m.type < 0) { f = mk(0.0f, 0.0f, 0.0f); pdf = 0.0f; return mk(0.0f, 0.0f, 0.0f); }
    float u[4];
    vt_rand4(&rng, u);
    if (m.type == 0) {
        const float r = sqrtf(u[0]);
        float sn, cs;
        vt_sincosf((2.0f * VT_PI_F) * u[1], &sn, &cs);
        const float px = r * cs, py = r * sn;
        const float y = sqrtf(vt_maxf(0.0f, (1.0f - px * px) - py * py));
        f = rec_r(m) / VT_PI_F;
        pdf = y / VT_PI_F;
        return mk(px, y, py);
    }
    const f3 refl = m.type == 1 ? rec_r(m) : mk(1.0f, 1.0f, 1.0f);
    const float e = m.rough;
    const float cosT = vt_powf(u[0], 1.0f / (e + 1.0f));
    const float sinT = sqrtf(vt_maxf(0.0f, 1.0f - cosT * cosT));
    const float phi = (u[1] * 2.0f) * VT_PI_F;
    float sp, cp;
    vt_sincosf(phi, &sp, &cp);
    const f3 wh = mk(sinT * cp, cosT, sinT * sp);
    const f3 wi = (-wo) + wh * (2.0f * dot(wo, wh));
    D_term(refl, e, wo, wh, f, pdf);
    const float cWi = wi.y, cWo = wo.y;
    const f3 wh2 = normalize(wi + wo);
    const float cH = dot(wi, wh2);
    const f3 t1 = refl * f;
    const f3 t2 = t1 * G_term(wo, wi, wh2);
    const f3 t3 = t2 * F_term(cH);
    f = t3 / ((4.0f * cWo) * cWi);
    return wi;
}

}  // namespace vt

/*
 * vt_kernels.h - kernel argument blocks and launchers shared by the HIP
 * translation units of the voxel path tracer.
 */
#pragma once

#include "vt_device.h"

namespace vt {

struct TraceArgs {
    GridView g;
    EnvView e;
    CamView c;
    Settings s;
    const MatRec* recs;          /* 256 records: [0] = offset 0 (ground), [slot] */
    int32_t W, H;
    uint64_t seed;
    uint32_t n0, count, rng_frame_max;
    float4* accum;
    double* sum64;
    unsigned long long* stats;   /* 6 counters */
    uint32_t* tile_counter;
    int32_t tiles_x, n_tiles;
    int32_t lds_bits;            /* stage occupancy bits in LDS */
    int32_t lds_bits_p;          /* wavefront traversal: stage the padded block bits in LDS */
    int32_t cdfv_lds_bytes;
    int32_t dda_budget;          /* wavefront traversal: DDA steps between refill checks */
    int32_t refill_min;          /* wavefront traversal: idle lanes that trigger a refill */
    int32_t refill_chunk;        /* wavefront traversal: queue entries a wave reserves per atomic */
};

struct Counters {
    uint32_t rays, visits, interactions, samples, env_samples, env_evals;
};

enum { PH_PRIMARY = 0, PH_SHADOW = 1, PH_CONT = 2, PH_COMB = 3 /* shade round: shadow result, then CONT */ };

/* Registers of one path between traversals. */
struct PathState {
    f3 ray_o, ray_d;             /* the ray to trace next */
    f3 basis_o;                  /* origin voxelSpaceToWorldSpace re-intersects from */
    f3 throughput, radiance;
    f3 direct;                   /* light-sample contribution pending its shadow ray */
    f3 factor;                   /* f * |cos| / pdf of the sampled continuation */
    f3 next_d;                   /* continuation direction */
    float bsdf_pdf;
    int32_t ex, ey, ez;          /* emissive voxel the shadow ray must reach */
    int32_t light_is_voxel;
    int32_t phase, bounces;
    vt_pcg32 rng;
};

/* Wavefront path storage: one slot per path of a chunk of the launch's
 * (pass, pixel) space, structure-of-arrays in HBM. */
/* Path state by queue position: the fields of the path whose ray sits at
 * position j of a ray queue are at index j of that queue's buffer, so shade
 * kernels read and write them contiguously (a surviving path moves to its
 * position in the next queue). */
struct PathBuf {
    float4* o;        /* ray_o (or basis origin after generate).xyz, bsdf_pdf */
    float4* nd;       /* continuation direction (camera direction after generate), pixel (bits) */
    float4* thr;      /* throughput times the continuation factor, sample slot of the chunk (bits) */
    float4* rad;      /* radiance if the light sample is occluded, emissive target x | y << 16 (bits) */
    float4* dir;      /* radiance if it is visible, target z | light_is_voxel << 15 | bounces << 16 (bits) */
    float4* fac;      /* batch and cache launches only: w = scene or training slot + 1 (bits) */
    uint4* meta;      /* after generate only: pixel, sample slot, bounces (0), scene or training slot + 1 */
    uint64_t* rng;    /* PCG32 state (increment = 2*pixel+1) */
};

/* Traversal result in 8 bytes: voxel coordinates + 1 (they lie in [-1, n]
 * for n <= 4096) in 16-bit fields, kind in the top half of the second word. */
__device__ __forceinline__ uint2 pack_hit(int32_t x, int32_t y, int32_t z, int32_t kind)
{
    return make_uint2((uint32_t)(x + 1) | ((uint32_t)(y + 1) << 16), (uint32_t)(z + 1) | ((uint32_t)kind << 16));
}
__device__ __forceinline__ int32_t unpack_hit(uint2 r, Hit& h)
{
    h.x = (int32_t)(r.x & 0xffffu) - 1;
    h.y = (int32_t)(r.x >> 16) - 1;
    h.z = (int32_t)(r.y & 0xffffu) - 1;
    return (int32_t)(r.y >> 16);
}

struct WfPaths {
    PathBuf s[2];     /* state of the paths in queue 0 / queue 1 */
    uint2* hit;       /* traversal result per queue entry (pack_hit): x, y, z (each + 1), kind */
    float4* smp;      /* finished tonemapped sample of each path of the chunk */
    float4* qa0;      /* ray queue 0 (2 x cap entries: a path's shadow and continuation rays at 2k, 2k + 1):
                         origin.xyz moved by sign(d)*0.001, 1/clamped d.x */
    float2* qb0;      /*              1/clamped d.y, 1/clamped d.z */
    uint32_t* qt0;    /*              scene of the ray (batch launches only) */
    float4* qa1;      /* ray queue 1 */
    float2* qb1;
    uint32_t* qt1;
    uint32_t* cnt;    /* queue lengths: cnt[0] for q0, cnt[1] for q1 */
    uint64_t cap;     /* slots */
};

/* A batch of scenes rendered by one wavefront launch: per-scene arguments
 * and the prefix sums of their path counts (W*H*count) and pixel counts
 * (W*H), n + 1 entries each. */
struct BatchTable {
    const TraceArgs* tab;
    const uint64_t* path_base;
    const uint64_t* pix_base;
    uint32_t n;
    unsigned long long* stats;   /* 6 counters or null */
    int32_t packed;              /* every scene has packed-march masks (GridView::occr) */
};

/* Radiance-cache rendering of one pass: paths that are not training paths
 * end into the cache at their vertex of bounce cache_bounce (query appended
 * to q_*); training paths (path index % train_stride == train_offset, slot =
 * index / train_stride) run the plain integrator and record each shaded
 * vertex (query, throughput, radiance so far: 18 floats), and when they end
 * append one training sample per vertex: target = (R_end - R_v) / T_v. */
struct CacheRender {
    int32_t cache_bounce;   /* bounce at which paths end into the cache (-1 never) */
    int32_t train_bounce;   /* the same for training paths */
    uint32_t train_stride, train_offset, train_slots, max_vertices;
    float* vrec;            /* [train_slots][max_vertices][18] */
    uint32_t* vrec_n;       /* [train_slots] */
    float* q_raw;           /* [paths][12] */
    uint4* q_meta;          /* [paths] metadata of the path that ended into the cache */
    float4* q_rad;          /* [paths] its radiance and throughput at that point */
    float4* q_thr;
    float* q_out;           /* [paths][3] */
    uint32_t* q_count;
    uint32_t* q_done;       /* queries already answered */
    float* s_raw;           /* [s_cap][12] */
    float* s_tgt;           /* [s_cap][3] */
    uint32_t* s_count;
    uint32_t s_cap;
};

size_t render_lds_bytes(const TraceArgs& a);
/* One pass (a.count must be 1) of cache rendering into a.accum. */
hipError_t launch_render_cached(const TraceArgs& a, const WfPaths& P, const CacheRender& C, const void* pack,
                                const float* bias, int cus, hipStream_t s);
/* Wavefront render of every scene of B (paths of all scenes share the
 * queues; rounds follow the largest max_bounces, paths of scenes with fewer
 * bounces finish earlier).  a0 supplies the traversal tuning. */
hipError_t launch_render_batch(const TraceArgs& a0, const BatchTable& B, const WfPaths& P, uint64_t total_paths,
                               uint64_t total_pixels, int max_bounces, int cus, hipStream_t s, bool stats);
/* Events recorded on the launch stream around traversal launches: pair k
 * brackets the k-th timed launch (at most cap pairs; *used counts them). */
struct LaunchTimer {
    hipEvent_t* ev;
    int cap;
    int* used;
};

/* Wavefront render of the launch in a: path chunks of at most P.cap slots,
 * 1 + max_bounces traverse/shade rounds per chunk (each round after the
 * first traces a shadow and a continuation ray per path).  With acc_wait, every
 * accumulation launch first waits for that event (passes accumulate in
 * order when two launches render consecutive passes on two streams). */
hipError_t launch_render_wavefront(const TraceArgs& a, const WfPaths& P, int cus, hipStream_t s, bool stats,
                                   const LaunchTimer* timer = nullptr, hipEvent_t acc_wait = nullptr);
hipError_t launch_render(const TraceArgs& a, int grid, hipStream_t s, bool stats);
int render_blocks_per_cu(const TraceArgs& a, bool stats);
hipError_t launch_first_hit(const TraceArgs& a, int32_t* out, hipStream_t s);
hipError_t launch_trace_rays(const TraceArgs& a, int32_t n, const float* o, const float* d, int32_t* out,
                             hipStream_t s);
hipError_t launch_reset_counter(uint32_t* c, unsigned long long* stats, hipStream_t s);
hipError_t launch_device_math(int32_t fn, int32_t n, const float* a, const float* b, float* out, hipStream_t s);

/* ---- volume layout --------------------------------------------------------------------- */
struct BrickDims {
    int32_t nx, ny, nz;          /* voxels */
    int32_t bx, by, bz;          /* bricks per axis */
    int32_t bdim;                /* power-of-two cube edge in bricks (Morton range) */
    uint64_t n_bricks;           /* bdim^3 */
    int32_t kx, ky, kz;          /* 4x4x4 blocks per axis (ceil(n/4)) */
    uint64_t n_blocks;           /* kx*ky*kz */
};

/* Dense int32 offsets -> bricked slots via lut[offset] (offset in
 * [0, lut_n)), -1 (or any negative) -> slot 0.  Out-of-range offsets set
 * *bad. */
hipError_t launch_brick_from_offsets(const BrickDims& d, const int32_t* dense, const uint8_t* lut, int32_t lut_n,
                                     uint8_t* bricks, uint32_t* bad, hipStream_t s);
/* Dense uint8 ids -> bricked slots via remap[256]. */
hipError_t launch_brick_from_ids(const BrickDims& d, const uint8_t* dense, const uint8_t* remap, uint8_t* bricks,
                                 hipStream_t s);
/* Per 4x4x4 block voxel masks in linear block order (occ64 holds n_blocks
 * words) and the per-block non-empty bits ((n_blocks+31)/32 words); then
 * the same over the block grid padded by one block on every side
 * ((kx+2)(ky+2)(kz+2) words in occp), where voxels outside the grid are
 * set. */
hipError_t launch_block_masks(const BrickDims& d, const uint8_t* bricks, uint64_t* occ64, uint32_t* bits,
                              uint64_t* occp, uint32_t* bitsp, uint64_t* occr, hipStream_t s);
/* Bricked slots -> dense int32 offsets via slot_off[256]. */
hipError_t launch_dense_from_bricks(const BrickDims& d, const uint8_t* bricks, const int32_t* slot_off, int32_t* dense,
                                    hipStream_t s);
/* Fill a byte buffer (graph-safe replacement for hipMemsetAsync). */
hipError_t launch_fill_u8(uint8_t* p, uint64_t n, uint8_t v, hipStream_t s);
/* Emissive list -> (x, y, z, slot) per entry, from the current grid. */
hipError_t launch_light_table(const GridView& g, int4* out, hipStream_t s);
hipError_t launch_env_quads(const float* img, int32_t w, int32_t h, float4* quads, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);
/* Streaming 16-byte-vector copy (bytes must be a multiple of 16). */
hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t s);

/* ---- voxelizer ------------------------------------------------------------------------- */
struct VoxArgs {
    BrickDims d;
    const float* verts;
    const uint32_t* idx;
    int64_t ntri;
    float M[16];
    int32_t mode;                /* 0 THIN, 1 FAT */
    uint8_t slot;
    uint8_t* bricks;             /* may be null */
    uint8_t* occ;                /* dense occupancy, may be null */
    /* triangles with more than big_area (x, y) columns go to a list of
     * big_cap entries (counter big_count, zero at launch) that one wave per
     * triangle then voxelizes; big = null keeps every triangle on one thread */
    uint32_t* big;
    uint32_t* big_count;
    uint32_t big_cap;
    int32_t big_area;
};
hipError_t launch_voxelize(const VoxArgs& a, int cus, hipStream_t s);

}  // namespace vt

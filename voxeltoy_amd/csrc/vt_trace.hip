/*
 * vt_trace.hip - path-tracing kernels for gfx950.
 *
 * render_kernel is a persistent kernel: every wavefront pulls 8x8 pixel
 * tiles from a global tile counter until the frame is exhausted, and each
 * lane renders all requested progressive passes of its pixel, keeping the
 * running mean in registers (accumulation.fs:10-18 applied pass by pass, so
 * the result equals `count` separate passes).  The per-sample integrator is
 * pathTracer.fs main (:172-296) with directLighting (:64-170).
 *
 * LDS holds the 256 decoded material records, the brick occupancy bits
 * (when they fit) and the environment's marginal CDF.
 */
#include "vt_kernels.h"

namespace vt {

struct Counters {
    uint32_t rays, visits, interactions, samples, env_samples, env_evals;
};

template <bool STATS>
__device__ __forceinline__ f3 eval_env(const EnvView& e, f3 wi, float& pdf, Counters& cnt)
{
    const f3 L = background(e, wi);
    if (e.use_image) {
        if (STATS) cnt.env_evals++;
        pdf = luminance(L) / e.integral;
    } else {
        pdf = 1.0f / (4.0f * VT_PI_F);
    }
    return L;
}

/* directLighting (pathTracer.fs:64-170) */
template <bool STATS>
__device__ __forceinline__ f3 direct_lighting(const TraceArgs& a, const MatRec* recs, const uint32_t* bits,
                                              const float* cdfv, const MatRec& m, const Basis& B, f3 wo,
                                              vt_pcg32& rng, Counters& cnt)
{
    const GridView& g = a.g;
    float u[4];
    vt_rand4(&rng, u);
    const int numLights = g.n_emissive + 1;
    const int li = (int)(u[0] * (float)numLights);
    const bool emissive = li < numLights - 1;
    f3 L, wi;
    float pdf;
    int32_t ex = 0, ey = 0, ez = 0;
    if (emissive) {
        int32_t idx = g.emissive[li];
        const int32_t dz = g.nx * g.ny, dy = g.nx;
        ez = idx / dz; idx -= ez * dz;
        ey = idx / dy; idx -= ey * dy;
        ex = idx;
        const MatRec& me = recs[fetch_slot(g, bits, ex, ey, ez)];
        L = mk(me.e[0], me.e[1], me.e[2]) * 10.0f;
        const f3 res = mk(g.rx, g.ry, g.rz);
        const f3 vsE = mk((float)ex, (float)ey, (float)ez);
        const f3 bmin = mk(g.bmin[0], g.bmin[1], g.bmin[2]), bmax = mk(g.bmax[0], g.bmax[1], g.bmax[2]);
        f3 wsE = mk(vsE.x / res.x, vsE.y / res.y, vsE.z / res.z) * (bmax - bmin) + bmin;
        wsE = wsE + mk(u[1], u[2], u[3]) * mk(g.vs[0], g.vs[1], g.vs[2]);
        const f3 toLight = wsE - B.p;
        const float r = length(toLight);
        wi = toLight / r;
        const Basis LB = voxel_basis(g, vsE.x, vsE.y, vsE.z, B.p, wi);
        const float area = (6.0f * g.vs[0]) * g.vs[1];
        const float jac = (r * r) / vt_absf(dot(-wi, LB.n));
        pdf = jac / area;
    } else if (a.e.use_image) {
        float uu, vv;
        env_sample_uv(a.e, cdfv, u[1], u[2], uu, vv);
        wi = dir_from_uv(uu, vv, a.e.rotation);
        L = tex_bilinear(a.e, uu, vv);
        pdf = luminance(L) / a.e.integral;
        if (STATS) cnt.env_samples++;
    } else {
        const float y = u[1];
        const float r = sqrtf(vt_maxf(0.0f, 1.0f - y * y));
        const float phi = (u[2] * 2.0f) * VT_PI_F;
        float sn, cs;
        vt_sincosf(phi, &sn, &cs);
        pdf = 1.0f / (2.0f * VT_PI_F);
        wi = local_to_world(mk(r * cs, y, r * sn), B);
        L = background(a.e, wi);
    }
    pdf = pdf / (float)numLights;
    Hit h;
    if (STATS) cnt.rays++;
    const int kind = traverse<STATS>(g, bits, B.p, wi, h, cnt.visits);
    if (emissive) {
        if (kind == KIND_MISS || h.x != ex || h.y != ey || h.z != ez) return mk(0.0f, 0.0f, 0.0f);
    } else {
        if (kind != KIND_MISS) return mk(0.0f, 0.0f, 0.0f);
    }
    const f3 lsWo = world_to_local(wo, B);
    const f3 lsWi = world_to_local(wi, B);
    f3 f;
    float bpdf;
    eval_material(m, lsWo, lsWi, f, bpdf);
    const float mis = power_heuristic(pdf, bpdf);
    const float cosv = vt_absf(dot(wi, B.n));
    return ((f * L) * cosv) * mis / pdf;
}

/* pathTracer.fs main (:172-296) for one pixel and one pass. */
template <bool STATS>
__device__ __forceinline__ f3 trace_sample(const TraceArgs& a, const MatRec* recs, const uint32_t* bits,
                                           const float* cdfv, int32_t px, int32_t py, uint32_t rng_frame,
                                           Counters& cnt)
{
    const GridView& g = a.g;
    const Settings& st = a.s;
    vt_pcg32 rng;
    vt_rng_init(&rng, a.seed, (uint32_t)(py * a.W + px), rng_frame);
    float u[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    if (a.c.lens_model == 0 || a.c.lens_model == 1) vt_rand4(&rng, u);
    f3 o, d;
    gen_ray(a.c, (float)px + 0.5f, (float)py + 0.5f, u, o, d);
    if (STATS) { cnt.samples++; cnt.rays++; }
    const float t = aabb(o, d, mk(g.bmin[0], g.bmin[1], g.bmin[2]), mk(g.bmax[0], g.bmax[1], g.bmax[2]));
    f3 radiance = mk(0.0f, 0.0f, 0.0f);
    bool escaped = false;
    Hit h;
    if (t < 0.0f) {
        escaped = true;
    } else {
        const f3 entry = o + d * t;
        if (traverse<STATS>(g, bits, entry, d, h, cnt.visits) == KIND_MISS) escaped = true;
    }
    if (escaped) {
        radiance = background(a.e, d);
        if (STATS && a.e.use_image) cnt.env_evals++;
    } else {
        f3 throughput = mk(1.0f, 1.0f, 1.0f);
        int bounces = 0;
        while (bounces < st.max_bounces) {
            if (STATS) cnt.interactions++;
            const Basis B = voxel_basis(g, (float)h.x, (float)h.y, (float)h.z, o, d);
            const MatRec& m = recs[h.slot];   /* ground hits carry slot 0 = record of offset 0 */
            if (h.x == st.sel[0] && h.y == st.sel[1] && h.z == st.sel[2]) {
                radiance = radiance + mk(1.0f, 0.0f, 0.0f);
                break;
            }
            const f3 wo = -d;
            const f3 lsWo = world_to_local(wo, B);
            if (bounces == 0) radiance = radiance + throughput * mk(m.e[0], m.e[1], m.e[2]);
            radiance = radiance + throughput * direct_lighting<STATS>(a, recs, bits, cdfv, m, B, wo, rng, cnt);
            f3 f;
            float pdf;
            const f3 lsWi = sample_material(m, lsWo, rng, f, pdf);
            if (st.wire_opacity > 0.0f) {
                const f3 q = B.p - mk(g.bmin[0], g.bmin[1], g.bmin[2]);
                const f3 ext = mk(g.bmax[0] - g.bmin[0], g.bmax[1] - g.bmin[1], g.bmax[2] - g.bmin[2]);
                const f3 vc = mk(q.x / ext.x, q.y / ext.y, q.z / ext.z) * mk(g.rx, g.ry, g.rz);
                const f3 uvw = mk((float)h.x, (float)h.y, (float)h.z) - vc;
                const float ux = vt_absf(dot(mk(B.n.y, B.n.z, B.n.x), uvw));
                const float uy = vt_absf(dot(mk(B.n.z, B.n.x, B.n.y), uvw));
                const float th = st.wire_thickness;
                float w = ((vt_stepf(th, ux) * vt_stepf(ux, 1.0f - th)) * vt_stepf(th, uy)) * vt_stepf(uy, 1.0f - th);
                w = (1.0f - st.wire_opacity) + st.wire_opacity * w;
                f = f * w;
            }
            const f3 wi = local_to_world(lsWi, B);
            const float cosv = vt_absf(dot(wi, B.n));
            throughput = throughput * ((f * cosv) / pdf);
            o = B.p;
            d = wi;
            if (STATS) cnt.rays++;
            if (traverse<STATS>(g, bits, o, d, h, cnt.visits) == KIND_MISS) {
                float lpdf;
                const f3 L = eval_env<STATS>(a.e, d, lpdf, cnt);
                const float mis = power_heuristic(pdf, lpdf);
                radiance = radiance + (throughput * L) * mis;
                break;
            }
            bounces++;
        }
    }
    const float ig = st.inv_gamma;
    return mk(vt_powf(radiance.x * st.exposure, ig), vt_powf(radiance.y * st.exposure, ig),
              vt_powf(radiance.z * st.exposure, ig));
}

/* Stage records, CDF marginal and (optionally) occupancy bits in LDS. */
__device__ __forceinline__ void stage_lds(const TraceArgs& a, uint8_t* smem, MatRec*& recs, float*& cdfv,
                                          const uint32_t*& bits)
{
    uint32_t* w = (uint32_t*)smem;
    const uint32_t* src = (const uint32_t*)a.recs;
    for (int i = threadIdx.x; i < 256 * 8; i += blockDim.x) w[i] = src[i];
    recs = (MatRec*)smem;
    float* cv = (float*)(smem + 256 * sizeof(MatRec));
    if (a.e.use_image) {
        for (int i = threadIdx.x; i < a.e.cv_n; i += blockDim.x) cv[i] = a.e.cdfv[i];
    }
    cdfv = cv;
    if (a.lds_bits) {
        uint32_t* sb = (uint32_t*)(smem + 256 * sizeof(MatRec) + a.cdfv_lds_bytes);
        for (uint32_t i = threadIdx.x; i < a.g.bit_words; i += blockDim.x) sb[i] = a.g.bits[i];
        bits = sb;
    } else {
        bits = a.g.bits;
    }
    __syncthreads();
}

template <bool STATS>
__global__ __launch_bounds__(256) void render_kernel(TraceArgs a)
{
    extern __shared__ __align__(16) uint8_t smem[];
    MatRec* recs;
    float* cdfv;
    const uint32_t* bits;
    stage_lds(a, smem, recs, cdfv, bits);
    const int lane = threadIdx.x & 63;
    Counters cnt = { 0, 0, 0, 0, 0, 0 };
    uint64_t tot[6] = { 0, 0, 0, 0, 0, 0 };
    for (;;) {
        int tile = 0;
        if (lane == 0) tile = (int)atomicAdd(a.tile_counter, 1u);
        tile = __shfl(tile, 0, 64);
        if (tile >= a.n_tiles) break;
        const int tx = tile % a.tiles_x, ty = tile / a.tiles_x;
        const int px = tx * 8 + (lane & 7), py = ty * 8 + (lane >> 3);
        if (px < a.W && py < a.H) {
            const size_t p = (size_t)py * (size_t)a.W + (size_t)px;
            float4 acc = a.accum[p];
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            if (a.sum64) { s0 = a.sum64[3 * p]; s1 = a.sum64[3 * p + 1]; s2 = a.sum64[3 * p + 2]; }
            for (uint32_t k = 0; k < a.count; ++k) {
                const uint32_t n = a.n0 + k;
                const uint32_t rf = n < a.rng_frame_max ? n : a.rng_frame_max;
                const f3 c = trace_sample<STATS>(a, recs, bits, cdfv, px, py, rf, cnt);
                const float fn = (float)n, fn1 = (float)(n + 1u);
                acc.x = (c.x + acc.x * fn) / fn1;
                acc.y = (c.y + acc.y * fn) / fn1;
                acc.z = (c.z + acc.z * fn) / fn1;
                acc.w = (1.0f + acc.w * fn) / fn1;
                if (a.sum64) { s0 += (double)c.x; s1 += (double)c.y; s2 += (double)c.z; }
            }
            a.accum[p] = acc;
            if (a.sum64) { a.sum64[3 * p] = s0; a.sum64[3 * p + 1] = s1; a.sum64[3 * p + 2] = s2; }
        }
        if (STATS) {
            tot[0] += cnt.rays; tot[1] += cnt.visits; tot[2] += cnt.interactions;
            tot[3] += cnt.samples; tot[4] += cnt.env_samples; tot[5] += cnt.env_evals;
            cnt = Counters{ 0, 0, 0, 0, 0, 0 };
        }
    }
    if (STATS) {
        __shared__ unsigned long long red[6];
        if (threadIdx.x < 6) red[threadIdx.x] = 0ull;
        __syncthreads();
        for (int i = 0; i < 6; ++i) atomicAdd(&red[i], (unsigned long long)tot[i]);
        __syncthreads();
        if (threadIdx.x < 6) atomicAdd(&a.stats[threadIdx.x], red[threadIdx.x]);
    }
}

/* First hit of the primary ray of every pixel (the parity probe). */
__device__ __forceinline__ void hit_record(const GridView& g, const uint32_t* bits, f3 o, f3 d, f3 basis_o,
                                           int32_t* rec)
{
    Hit h;
    h.x = h.y = h.z = 0;
    uint32_t dummy = 0;
    const int kind = traverse<false>(g, bits, o, d, h, dummy);
    rec[1] = h.x; rec[2] = h.y; rec[3] = h.z; rec[4] = h.steps;
    if (h.steps < 0) { rec[1] = rec[2] = rec[3] = 0; }
    rec[5] = rec[6] = rec[7] = 0;
    if (kind == KIND_MISS) { rec[0] = 1; return; }
    rec[0] = kind;
    const Basis B = voxel_basis(g, (float)h.x, (float)h.y, (float)h.z, basis_o, d);
    rec[5] = (int32_t)B.n.x; rec[6] = (int32_t)B.n.y; rec[7] = (int32_t)B.n.z;
}

__global__ __launch_bounds__(256) void first_hit_kernel(TraceArgs a, int32_t* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)a.W * a.H) return;
    const int32_t px = (int32_t)(i % a.W), py = (int32_t)(i / a.W);
    int32_t* rec = out + 8 * i;
    vt_pcg32 rng;
    vt_rng_init(&rng, a.seed, (uint32_t)(py * a.W + px), a.n0);
    float u[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    if (a.c.lens_model == 0 || a.c.lens_model == 1) vt_rand4(&rng, u);
    f3 o, d;
    gen_ray(a.c, (float)px + 0.5f, (float)py + 0.5f, u, o, d);
    const GridView& g = a.g;
    const float t = aabb(o, d, mk(g.bmin[0], g.bmin[1], g.bmin[2]), mk(g.bmax[0], g.bmax[1], g.bmax[2]));
    if (t < 0.0f) {
        for (int k = 0; k < 8; ++k) rec[k] = 0;
        return;
    }
    hit_record(g, g.bits, o + d * t, d, o, rec);
}

__global__ __launch_bounds__(256) void trace_rays_kernel(TraceArgs a, int32_t n, const float* orig, const float* dir,
                                                         int32_t* out)
{
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const f3 o = mk(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]);
    const f3 d = mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
    hit_record(a.g, a.g.bits, o, d, o, out + 8 * (size_t)i);
}

__global__ void reset_counter_kernel(uint32_t* c, unsigned long long* stats)
{
    if (threadIdx.x == 0) *c = 0u;
    if (stats && threadIdx.x < 6) stats[threadIdx.x] = 0ull;
}

__global__ void device_math_kernel(int32_t fn, int32_t n, const float* a, const float* b, float* out)
{
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const float x = a[i], y = b[i];
    float r = 0.0f, s, c;
    switch (fn) {
        case 0: r = sqrtf(x); break;
        case 1: r = x / y; break;
        case 2: vt_sincosf(x, &s, &c); r = s; break;
        case 3: vt_sincosf(x, &s, &c); r = c; break;
        case 4: r = vt_acosf(x); break;
        case 5: r = vt_atan2f(x, y); break;
        case 6: r = vt_log2f(x); break;
        case 7: r = vt_exp2f(x); break;
        case 8: r = vt_powf(x, y); break;
        default: r = 0.0f; break;
    }
    out[i] = r;
}

/* ---- launchers ------------------------------------------------------------------ */
size_t render_lds_bytes(const TraceArgs& a)
{
    return 256 * sizeof(MatRec) + (size_t)a.cdfv_lds_bytes + (a.lds_bits ? (size_t)a.g.bit_words * 4u : 0u);
}

hipError_t launch_render(const TraceArgs& a, int grid, hipStream_t s, bool stats)
{
    const size_t lds = render_lds_bytes(a);
    if (stats) render_kernel<true><<<grid, 256, lds, s>>>(a);
    else render_kernel<false><<<grid, 256, lds, s>>>(a);
    return hipGetLastError();
}

int render_blocks_per_cu(const TraceArgs& a, bool stats)
{
    int blocks = 0;
    const size_t lds = render_lds_bytes(a);
    hipError_t e;
    if (stats) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, render_kernel<true>, 256, lds);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, render_kernel<false>, 256, lds);
    if (e != hipSuccess || blocks < 1) blocks = 1;
    return blocks;
}

hipError_t launch_first_hit(const TraceArgs& a, int32_t* out, hipStream_t s)
{
    const int64_t n = (int64_t)a.W * a.H;
    first_hit_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, out);
    return hipGetLastError();
}

hipError_t launch_trace_rays(const TraceArgs& a, int32_t n, const float* o, const float* d, int32_t* out,
                             hipStream_t s)
{
    trace_rays_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(a, n, o, d, out);
    return hipGetLastError();
}

hipError_t launch_reset_counter(uint32_t* c, unsigned long long* stats, hipStream_t s)
{
    reset_counter_kernel<<<1, 64, 0, s>>>(c, stats);
    return hipGetLastError();
}

hipError_t launch_device_math(int32_t fn, int32_t n, const float* a, const float* b, float* out, hipStream_t s)
{
    device_math_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(fn, n, a, b, out);
    return hipGetLastError();
}

}  // namespace vt

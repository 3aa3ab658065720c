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

#include <algorithm>
#include "vt_cache.h"

namespace vt {


/* pathTracer.fs main (:172-296) with directLighting (:64-170), written as a
 * per-path state machine with a single traversal site.  Each path vertex is
 * shaded once: the light sample (first rand(), :77) and the BSDF sample
 * (second rand(), :255) are both drawn before any ray is traced, the shadow
 * ray's contribution is kept aside and added after its traversal (:133-164
 * only needs visibility), then the continuation ray is traced (:282).  Draw
 * order, operand order and accumulation order equal the reference's,
 * including throughput * vec3(0) for occluded light samples. */

/* Camera ray of one pass of one pixel; returns false when the ray misses the
 * volume bounds (S.radiance then holds the background). */
template <bool STATS>
__device__ __forceinline__ bool start_path(const TraceArgs& a, int32_t px, int32_t py, uint32_t rng_frame, PathState& S,
                                           Counters& cnt)
{
    const GridView& g = a.g;
    vt_rng_init(&S.rng, a.seed, (uint32_t)(py * a.W + px), rng_frame);
    float u[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    if (a.c.lens_model == 0 || a.c.lens_model == 1) vt_rand4(&S.rng, u);
    gen_ray(a.c, (float)px + 0.5f, (float)py + 0.5f, u, S.basis_o, S.ray_d);
    if (STATS) { cnt.samples++; cnt.rays++; }
    const float t = aabb(S.basis_o, S.ray_d, mk(g.bmin[0], g.bmin[1], g.bmin[2]), mk(g.bmax[0], g.bmax[1], g.bmax[2]));
    if (t < 0.0f) {
        S.radiance = background(a.e, S.ray_d);
        if (STATS && a.e.use_image) cnt.env_evals++;
        return false;
    }
    S.ray_o = S.basis_o + S.ray_d * t;
    S.throughput = mk(1.0f, 1.0f, 1.0f);
    S.radiance = mk(0.0f, 0.0f, 0.0f);
    S.direct = S.factor = S.next_d = mk(0.0f, 0.0f, 0.0f);
    S.bsdf_pdf = 0.0f;
    S.ex = S.ey = S.ez = 0;
    S.light_is_voxel = 0;
    S.phase = PH_PRIMARY;
    S.bounces = 0;
    return true;
}

/* Consume the traversal result of the ray in S; returns 1 when S holds a
 * new ray to trace, 0 when the path is finished.  With VERT, a shaded
 * vertex also yields the radiance-cache query of the vertex in vraw (12
 * floats: position in volume units [0,1]^3, (wo+1)/2, (normal+1)/2, albedo)
 * with *vrec set, taken after the bounce-0 emission and before any random
 * draw; when the vertex is at bounce `cache_bounce` the step returns 2
 * there instead (the path ends into the cache). */
template <bool STATS, bool VERT = false>
__device__ __forceinline__ int shade_step(const TraceArgs& a, const MatRec* recs, const uint32_t* bits,
                                          const float* cdfv, PathState& S, const Hit& h, int kind, Counters& cnt,
                                          int cache_bounce = -1, float* vraw = nullptr, bool* vrec = nullptr)
{
    const GridView& g = a.g;
    const Settings& st = a.s;
    if (S.phase == PH_SHADOW) {
        const bool visible = S.light_is_voxel ? (kind != KIND_MISS && h.x == S.ex && h.y == S.ey && h.z == S.ez)
                                              : (kind == KIND_MISS);
        S.radiance = S.radiance + S.throughput * (visible ? S.direct : mk(0.0f, 0.0f, 0.0f));
        S.throughput = S.throughput * S.factor;
        S.basis_o = S.ray_o;
        S.ray_d = S.next_d;
        S.phase = PH_CONT;
        if (STATS) cnt.rays++;
        return true;
    }
    if (kind == KIND_MISS) {
        if (S.phase == PH_PRIMARY) {
            S.radiance = background(a.e, S.ray_d);
            if (STATS && a.e.use_image) cnt.env_evals++;
        } else {
            float lpdf;
            const f3 L = background(a.e, S.ray_d);
            if (a.e.use_image) {
                if (STATS) cnt.env_evals++;
                lpdf = luminance(L) / a.e.integral;
            } else {
                lpdf = 1.0f / (4.0f * VT_PI_F);
            }
            S.radiance = S.radiance + (S.throughput * L) * power_heuristic(S.bsdf_pdf, lpdf);
        }
        return false;
    }
    if (S.phase == PH_CONT) S.bounces++;
    if (S.bounces >= st.max_bounces) return false;
    /* ---- shade the vertex (pathTracer.fs:214-280) ---- */
    if (STATS) cnt.interactions++;
    const Basis B = voxel_basis(g, (float)h.x, (float)h.y, (float)h.z, S.basis_o, S.ray_d);
    const MatRec& m = recs[h.slot];   /* ground hits carry slot 0 = record of offset 0 */
    if (h.x == st.sel[0] && h.y == st.sel[1] && h.z == st.sel[2]) {
        S.radiance = S.radiance + mk(1.0f, 0.0f, 0.0f);
        return false;
    }
    const f3 wo = -S.ray_d;
    const f3 lsWo = world_to_local(wo, B);
    if (S.bounces == 0) S.radiance = S.radiance + S.throughput * mk(m.e[0], m.e[1], m.e[2]);
    if (VERT) {
        vraw[0] = (B.p.x - g.bmin[0]) * g.ext[0];
        vraw[1] = (B.p.y - g.bmin[1]) * g.ext[1];
        vraw[2] = (B.p.z - g.bmin[2]) * g.ext[2];
        vraw[3] = (wo.x + 1.0f) * 0.5f;
        vraw[4] = (wo.y + 1.0f) * 0.5f;
        vraw[5] = (wo.z + 1.0f) * 0.5f;
        vraw[6] = (B.n.x + 1.0f) * 0.5f;
        vraw[7] = (B.n.y + 1.0f) * 0.5f;
        vraw[8] = (B.n.z + 1.0f) * 0.5f;
        vraw[9] = m.r[0];
        vraw[10] = m.r[1];
        vraw[11] = m.r[2];
        *vrec = true;
        if (S.bounces == cache_bounce) return 2;
    }
    /* light sample (directLighting :77-124) */
    float u[4];
    vt_rand4(&S.rng, u);
    const int numLights = g.n_emissive + 1;
    const int li = (int)(u[0] * (float)numLights);
    S.light_is_voxel = li < numLights - 1;
    f3 L, wl;
    float lpdf;
    if (S.light_is_voxel) {
        /* the emissive list entry decoded once per grid (light_table_kernel) */
        const int4 lt = g.lights[li];
        S.ex = lt.x; S.ey = lt.y; S.ez = lt.z;
        const MatRec& me = recs[lt.w];
        L = mk(me.e[0], me.e[1], me.e[2]) * 10.0f;
        const f3 vsE = mk((float)S.ex, (float)S.ey, (float)S.ez);
        const f3 bmin = mk(g.bmin[0], g.bmin[1], g.bmin[2]), bmax = mk(g.bmax[0], g.bmax[1], g.bmax[2]);
        f3 wsE = mk(vsE.x / g.rx, vsE.y / g.ry, vsE.z / g.rz) * (bmax - bmin) + bmin;
        wsE = wsE + mk(u[1], u[2], u[3]) * mk(g.vs[0], g.vs[1], g.vs[2]);
        const f3 toLight = wsE - B.p;
        const float r = length(toLight);
        wl = toLight / r;
        const Basis LB = voxel_basis(g, vsE.x, vsE.y, vsE.z, B.p, wl);
        const float area = (6.0f * g.vs[0]) * g.vs[1];
        const float jac = (r * r) / vt_absf(dot(-wl, LB.n));
        lpdf = jac / area;
    } else if (a.e.use_image) {
        float uu, vv;
        env_sample_uv(a.e, cdfv, u[1], u[2], uu, vv);
        wl = dir_from_uv(uu, vv, a.e.rotation);
        L = tex_bilinear(a.e, uu, vv);
        lpdf = luminance(L) / a.e.integral;
        if (STATS) cnt.env_samples++;
    } else {
        const float y = u[1];
        const float r = sqrtf(vt_maxf(0.0f, 1.0f - y * y));
        const float phi = (u[2] * 2.0f) * VT_PI_F;
        float sn, cs;
        vt_sincosf(phi, &sn, &cs);
        lpdf = 1.0f / (2.0f * VT_PI_F);
        wl = local_to_world(mk(r * cs, y, r * sn), B);
        L = background(a.e, wl);
    }
    lpdf = lpdf / (float)numLights;
    {
        f3 f;
        float bpdf;
        eval_material(m, lsWo, world_to_local(wl, B), f, bpdf);
        const float mis = power_heuristic(lpdf, bpdf);
        S.direct = ((f * L) * vt_absf(dot(wl, B.n))) * mis / lpdf;
    }
    /* BSDF sample (:255-276) */
    f3 f;
    const f3 lsWi = sample_material(m, lsWo, S.rng, f, S.bsdf_pdf);
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
    S.next_d = local_to_world(lsWi, B);
    S.factor = (f * vt_absf(dot(S.next_d, B.n))) / S.bsdf_pdf;
    S.ray_o = B.p;
    S.ray_d = wl;
    S.phase = PH_SHADOW;
    if (STATS) cnt.rays++;
    return true;
}

__device__ __forceinline__ f3 tonemap(const Settings& st, f3 r)
{
    const float ig = st.inv_gamma;
    return mk(vt_powf(r.x * st.exposure, ig), vt_powf(r.y * st.exposure, ig), vt_powf(r.z * st.exposure, ig));
}

/* One pass of one pixel, start to finish (megakernel form). */
template <bool STATS>
__device__ __forceinline__ f3 trace_sample(const TraceArgs& a, const MatRec* recs, const uint32_t* bits,
                                           const float* cdfv, int32_t px, int32_t py, uint32_t rng_frame,
                                           Counters& cnt)
{
    PathState S;
    if (start_path<STATS>(a, px, py, rng_frame, S, cnt)) {
        for (;;) {
            Hit h;
            const int kind = traverse<STATS>(a.g, bits, S.ray_o, S.ray_d, h, cnt.visits);
            if (!shade_step<STATS>(a, recs, bits, cdfv, S, h, kind, cnt)) break;
        }
    }
    return tonemap(a.s, S.radiance);
}

/* Stage records, CDF marginal and (when LDSB) the brick occupancy bits in
 * LDS.  LDSB is a template parameter so the traversal's bit reads compile to
 * LDS instructions rather than generic (flat) loads. */
template <bool LDSB>
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
    if (LDSB) {
        uint32_t* sb = (uint32_t*)(smem + 256 * sizeof(MatRec) + a.cdfv_lds_bytes);
        for (uint32_t i = threadIdx.x; i < a.g.bit_words; i += blockDim.x) sb[i] = a.g.bits[i];
        bits = sb;
    } else {
        bits = a.g.bits;
    }
    __syncthreads();
}

template <bool STATS, bool LDSB>
__global__ __launch_bounds__(256) void render_kernel(TraceArgs a)
{
    extern __shared__ __align__(16) uint8_t smem[];
    MatRec* recs;
    float* cdfv;
    const uint32_t* bits;
    stage_lds<LDSB>(a, smem, recs, cdfv, bits);
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
            for (uint32_t k = 0; k < a.count; ++k) {
                const uint32_t n = a.n0 + k;
                const uint32_t rf = n < a.rng_frame_max ? n : a.rng_frame_max;
                const f3 c = trace_sample<STATS>(a, recs, bits, cdfv, px, py, rf, cnt);
                /* accumulation.fs:17, one pass at a time (the running mean
                 * stays in memory between passes, as the ping-pong textures do) */
                const float fn = (float)n, fn1 = (float)(n + 1u);
                float4 acc = a.accum[p];
                acc.x = (c.x + acc.x * fn) / fn1;
                acc.y = (c.y + acc.y * fn) / fn1;
                acc.z = (c.z + acc.z * fn) / fn1;
                acc.w = (1.0f + acc.w * fn) / fn1;
                a.accum[p] = acc;
                if (a.sum64) {
                    a.sum64[3 * p] += (double)c.x;
                    a.sum64[3 * p + 1] += (double)c.y;
                    a.sum64[3 * p + 2] += (double)c.z;
                }
            }
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

/* ---- wavefront form ---------------------------------------------------------------
 * generate -> (traverse -> shade) x (1 + max_bounces) -> accumulate, per
 * chunk of paths; after the camera rays, each round traces every path's
 * shadow ray and continuation ray together.  The traversal kernel only runs the DDA, so it needs few
 * registers and keeps many waves resident; shading consumes the hits and
 * appends surviving paths to the next queue with one atomic per wave. */

/* Path state in HBM, by queue position (PathBuf).  Every path of a
 * traversal round is in the same phase (round 0 PRIMARY, odd rounds SHADOW,
 * even rounds CONT); a shade kernel reads the fields its phase needs at the
 * path's position j in the current buffer and writes the fields the next
 * phase needs at its position k in the next queue's buffer:
 *   generate -> o (= basis origin), nd (= camera direction), meta, rng
 *   PRIMARY  <- o, nd, meta, rng           -> o, nd, thr, rad, dir, fac, tgt, meta, rng
 *   SHADOW   <- o, nd, thr, rad, dir, fac, tgt, meta, rng -> o, nd, thr, rad, meta, rng
 *   CONT     <- o, nd, thr, rad, meta, rng -> o, nd, thr, rad, dir, fac, tgt, meta, rng
 * After SHADOW, basis_o = ray_o and ray_d = next_d, so CONT reads them from
 * o and nd. */
/* A camera path's state: origin (w: pixel) and direction (w: sample
 * slot); with TAG (batch and cache launches) also the metadata record, for
 * the scene or training tag.  Bounces are 0. */
template <bool TAG>
__device__ __forceinline__ void store_generated(const PathBuf& N, uint32_t k, const PathState& S, uint32_t pix,
                                                uint32_t slot, uint32_t tag)
{
    N.o[k] = make_float4(S.basis_o.x, S.basis_o.y, S.basis_o.z, __uint_as_float(pix));
    N.nd[k] = make_float4(S.ray_d.x, S.ray_d.y, S.ray_d.z, __uint_as_float(slot));
    if (TAG) N.meta[k] = make_uint4(pix, slot, (uint32_t)S.bounces, tag);
    N.rng[k] = S.rng.state;
}

template <int PHASE, bool TAG = false>
__device__ __forceinline__ void load_phase(const PathBuf& B, uint32_t j, PathState& S, uint4& me)
{
    S.phase = PHASE;
    S.direct = S.factor = S.next_d = mk(0.0f, 0.0f, 0.0f);
    S.ex = S.ey = S.ez = S.light_is_voxel = 0;
    S.bsdf_pdf = 0.0f;
    S.rng.state = B.rng[j];
    const float4 o = B.o[j], nd = B.nd[j];
    if (PHASE == PH_SHADOW) {
        /* the state store_cont writes: the throughput already times the
         * continuation factor, the radiance without and with the light
         * sample (S.radiance, S.direct); pixel, slot, light target and
         * bounce count in the w words, the tag in fac.w (TAG only) */
        const float4 th = B.thr[j], rd = B.rad[j], di = B.dir[j];
        const uint32_t exy = __float_as_uint(rd.w), ezb = __float_as_uint(di.w);
        me = make_uint4(__float_as_uint(nd.w), __float_as_uint(th.w), ezb >> 16,
                        TAG ? __float_as_uint(B.fac[j].w) : 0u);
        S.bounces = (int32_t)me.z;
        S.rng.inc = ((uint64_t)me.x << 1u) | 1u;
        S.ray_o = mk(o.x, o.y, o.z);
        S.bsdf_pdf = o.w;
        S.next_d = mk(nd.x, nd.y, nd.z);
        S.basis_o = S.ray_o;
        S.ray_d = S.next_d;
        S.throughput = mk(th.x, th.y, th.z);
        S.radiance = mk(rd.x, rd.y, rd.z);
        S.direct = mk(di.x, di.y, di.z);
        S.ex = (int32_t)(exy & 0xffffu); S.ey = (int32_t)(exy >> 16);
        S.ez = (int32_t)(ezb & 0x7fffu); S.light_is_voxel = (int32_t)((ezb >> 15) & 1u);
        return;
    }
    if (TAG) me = B.meta[j];
    else me = make_uint4(__float_as_uint(o.w), __float_as_uint(nd.w), 0u, 0u);
    S.bounces = (int32_t)me.z;
    S.rng.inc = ((uint64_t)me.x << 1u) | 1u;
    if (PHASE == PH_PRIMARY) {
        S.basis_o = mk(o.x, o.y, o.z);
        S.ray_d = mk(nd.x, nd.y, nd.z);
        S.ray_o = S.basis_o;
        S.throughput = mk(1.0f, 1.0f, 1.0f);
        S.radiance = mk(0.0f, 0.0f, 0.0f);
    } else {
        const float4 th = B.thr[j], rd = B.rad[j];
        S.basis_o = mk(o.x, o.y, o.z);
        S.ray_o = S.basis_o;
        S.bsdf_pdf = o.w;
        S.ray_d = mk(nd.x, nd.y, nd.z);
        S.throughput = mk(th.x, th.y, th.z);
        S.radiance = mk(rd.x, rd.y, rd.z);
    }
}

/* Store after a shade step that continues the path (next: the result of
 * its shadow ray, then its continuation) at position k of the next queue's
 * buffer.  What the shadow result does to the path (pathTracer.fs:133-164
 * as shade_step applies it) is formed here with the same operations:
 * radiance + throughput * 0 and radiance + throughput * direct (the
 * radiance for an occluded and a visible light sample) and throughput *
 * factor, so the round after the shadow ray only picks one.  Metadata in
 * the w components: nd.w pixel, thr.w sample slot, rad.w ex | ey << 16,
 * dir.w ez | light_is_voxel << 15 | bounces << 16 (voxel coordinates below
 * 2^15), fac.w the scene or training tag (TAG only). */
template <bool TAG>
__device__ __forceinline__ void store_cont(const PathBuf& N, uint32_t k, const PathState& S, uint4 me)
{
    const uint32_t exy = ((uint32_t)S.ex & 0xffffu) | ((uint32_t)S.ey << 16);
    const uint32_t ezb = ((uint32_t)S.ez & 0x7fffu) | ((uint32_t)(S.light_is_voxel != 0) << 15) |
                         ((uint32_t)S.bounces << 16);
    const f3 r_occluded = S.radiance + S.throughput * mk(0.0f, 0.0f, 0.0f);
    const f3 r_visible = S.radiance + S.throughput * S.direct;
    const f3 t_next = S.throughput * S.factor;
    N.o[k] = make_float4(S.ray_o.x, S.ray_o.y, S.ray_o.z, S.bsdf_pdf);
    N.nd[k] = make_float4(S.next_d.x, S.next_d.y, S.next_d.z, __uint_as_float(me.x));
    N.thr[k] = make_float4(t_next.x, t_next.y, t_next.z, __uint_as_float(me.y));
    N.rad[k] = make_float4(r_occluded.x, r_occluded.y, r_occluded.z, __uint_as_float(exy));
    N.dir[k] = make_float4(r_visible.x, r_visible.y, r_visible.z, __uint_as_float(ezb));
    if (TAG) N.fac[k] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(me.w));
    N.rng[k] = S.rng.state;
}

template <bool STATS>
__device__ __forceinline__ void flush_counters(unsigned long long* stats, const Counters& cnt)
{
    if (!STATS) return;
    __shared__ unsigned long long red[6];
    if (threadIdx.x < 6) red[threadIdx.x] = 0ull;
    __syncthreads();
    atomicAdd(&red[0], (unsigned long long)cnt.rays);
    atomicAdd(&red[1], (unsigned long long)cnt.visits);
    atomicAdd(&red[2], (unsigned long long)cnt.interactions);
    atomicAdd(&red[3], (unsigned long long)cnt.samples);
    atomicAdd(&red[4], (unsigned long long)cnt.env_samples);
    atomicAdd(&red[5], (unsigned long long)cnt.env_evals);
    __syncthreads();
    if (threadIdx.x < 6 && red[threadIdx.x]) atomicAdd(&stats[threadIdx.x], red[threadIdx.x]);
}

/* Append the ray (o, d) for the threads with `keep` to a ray queue (length
 * counter *n) and return each such thread's position; the queue stores the
 * rays themselves, already prepared for the march (prepare_ray, 24 bytes in
 * qa and qb) and, when qt is given (batch launches), the scene index `tag`,
 * so the traversal reads them in queue order.  One atomic per workgroup of NT
 * threads: every queue counter is a single device address that all XCDs
 * contend for.  Every thread of the workgroup must call it. */
template <int NT>
__device__ __forceinline__ uint32_t block_append(bool keep, uint32_t tag, f3 o, f3 d, float4* qa, float2* qb,
                                                 uint32_t* qt, uint32_t* n)
{
    __shared__ uint32_t wsum[NT / 64];
    __shared__ uint32_t bbase;
    const unsigned long long mask = __ballot(keep);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) wsum[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < NT / 64; ++w) { const uint32_t c = wsum[w]; wsum[w] = t; t += c; }
        bbase = t ? atomicAdd(n, t) : 0u;
    }
    __syncthreads();
    const uint32_t j = bbase + wsum[wave] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (keep) {
        const PreRay r = prepare_ray(o, d);
        qa[j] = queue_a(r);
        qb[j] = queue_b(r);
        if (qt) qt[j] = tag;
    }
    __syncthreads();
    return j;
}

/* The same with two rays per thread with `keep`, sharing the origin o (a
 * shadow ray along d0 and a continuation ray along d1): they take queue
 * entries 2k and 2k + 1, k being the returned position. */
template <int NT>
__device__ __forceinline__ uint32_t block_append2(bool keep, uint32_t tag, f3 o, f3 d0, f3 d1, float4* qa,
                                                  float2* qb, uint32_t* qt, uint32_t* n)
{
    __shared__ uint32_t wsum2[NT / 64];
    __shared__ uint32_t bbase2;
    const unsigned long long mask = __ballot(keep);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) wsum2[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < NT / 64; ++w) { const uint32_t c = wsum2[w]; wsum2[w] = t; t += c; }
        bbase2 = t ? atomicAdd(n, 2u * t) / 2u : 0u;
    }
    __syncthreads();
    const uint32_t k = bbase2 + wsum2[wave] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (keep) {
        const PreRay r0 = prepare_ray(o, d0), r1 = prepare_ray(o, d1);
        qa[2 * k] = queue_a(r0);
        qa[2 * k + 1] = queue_a(r1);
        const float2 b0 = queue_b(r0), b1 = queue_b(r1);
        ((float4*)qb)[k] = make_float4(b0.x, b0.y, b1.x, b1.y);
        if (qt) qt[2 * k] = qt[2 * k + 1] = tag;
    }
    __syncthreads();
    return k;
}

/* Workgroup-aggregated reservation of one entry of a counter for every
 * thread with `want`; returns the entry (undefined without `want`).  Every
 * thread of the workgroup must call it. */
template <int NT>
__device__ __forceinline__ uint32_t block_reserve(bool want, uint32_t* counter)
{
    __shared__ uint32_t rsum[NT / 64];
    __shared__ uint32_t rbase;
    const unsigned long long mask = __ballot(want);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (lane == 0) rsum[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < NT / 64; ++w) { const uint32_t c = rsum[w]; rsum[w] = t; t += c; }
        rbase = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    const uint32_t idx = rbase + rsum[wave] + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    __syncthreads();
    return idx;
}

#define WF_GENERATE_THREADS 1024
#define WF_SHADE_THREADS 256

/* Scene of global path index gi in a batch: base[s] <= gi < base[s + 1]. */
__device__ __forceinline__ uint32_t scene_of(const uint64_t* __restrict__ base, uint32_t n, uint64_t gi)
{
    uint32_t lo = 0, hi = n;       /* invariant: base[lo] <= gi < base[hi] */
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= gi) lo = mid;
        else hi = mid;
    }
    return lo;
}

/* Camera paths of the chunk [g0, g0 + n) of the launch's (scene,) pass and
 * pixel space: a path index enumerates passes outermost and pixels row by
 * row within a pass (within a scene, scenes back to back in a batch). */
template <bool STATS, bool BATCH, bool CACHE = false>
__global__ __launch_bounds__(WF_GENERATE_THREADS) void wf_generate_kernel(TraceArgs a0, BatchTable B, WfPaths P,
                                                                         uint64_t g0, uint32_t n,
                                                                         CacheRender C = CacheRender())
{
    Counters cnt = { 0, 0, 0, 0, 0, 0 };
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        bool keep = false;
        uint32_t sid = 0, gpix = 0, gtag = 0;
        f3 ro = mk(0.0f, 0.0f, 0.0f), rd = ro;
        PathState G;
        if (i < n) {
            uint64_t gi = g0 + i;
            if (BATCH) {
                sid = scene_of(B.path_base, B.n, gi);
                gi -= B.path_base[sid];
            }
            const TraceArgs& a = BATCH ? B.tab[sid] : a0;
            const uint64_t WH = (uint64_t)a.W * (uint64_t)a.H;
            const uint32_t k = (uint32_t)(gi / WH), pix = (uint32_t)(gi % WH);
            const int32_t px = (int32_t)(pix % (uint32_t)a.W), py = (int32_t)(pix / (uint32_t)a.W);
            const uint32_t nn = a.n0 + k;
            const uint32_t rf = nn < a.rng_frame_max ? nn : a.rng_frame_max;
            PathState S;
            uint32_t tag = sid;
            if (CACHE && gi % C.train_stride == C.train_offset && gi / C.train_stride < C.train_slots) {
                tag = (uint32_t)(gi / C.train_stride) + 1u;   /* training slot + 1 */
                C.vrec_n[tag - 1u] = 0u;
            }
            if (start_path<STATS>(a, px, py, rf, S, cnt)) {
                keep = true;
                ro = S.ray_o;
                rd = S.ray_d;
            } else {
                const f3 c = tonemap(a.s, S.radiance);
                P.smp[i] = make_float4(c.x, c.y, c.z, 1.0f);
            }
            G = S;
            gpix = pix;
            gtag = tag;
        }
        const uint32_t q =
            block_append<WF_GENERATE_THREADS>(keep, sid, ro, rd, P.qa0, P.qb0, BATCH ? P.qt0 : nullptr, &P.cnt[0]);
        if (keep) store_generated<BATCH || CACHE>(P.s[0], q, G, gpix, i, gtag);
    }
    flush_counters<STATS>(BATCH ? B.stats : a0.stats, cnt);
}

/* Persistent traversal: each lane marches its ray a.dda_budget steps at a
 * time; whenever at least a.refill_min lanes of the wave are idle they take
 * the next entries of the wave's private chunk of the queue, from the queue's end.
 * A wave reserves a.refill_chunk entries per atomic on the queue head
 * (0 = exactly the idle lanes), so few waves contend for the counter, and
 * lanes stay busy until the queue drains instead of idling behind the
 * longest ray of their wave.  Workgroups of 1024 threads share one LDS copy
 * of the block occupancy bits. */
#define WF_TRAVERSE_THREADS 1024

/* SGPRs are allocated in steps of 16.  A build of this kernel whose
 * allocation rose from 80 to 96 SGPRs, with its hot loop unchanged
 * instruction for instruction, took 1.49 ms per launch instead of 1.10 ms,
 * although the compiler reported 8 waves per SIMD for both (two 1024-thread
 * workgroups per CU need all 8): the cap keeps the allocation at 80. */
#define WF_TRAVERSE_SGPRS 80
template <bool STATS, bool LDSB, bool BATCH>
__global__ __launch_bounds__(WF_TRAVERSE_THREADS) __attribute__((amdgpu_num_sgpr(WF_TRAVERSE_SGPRS)))
void wf_traverse_kernel(TraceArgs a, BatchTable B, WfPaths P, int sel)
{
    extern __shared__ __align__(16) uint8_t smem[];
    /* the march reads the padded block bits only */
    const uint32_t* bits = a.g.bitsp;
    if (LDSB) {
        uint32_t* sb = (uint32_t*)smem;
        for (uint32_t i = threadIdx.x; i < a.g.bit_words_p; i += blockDim.x) sb[i] = a.g.bitsp[i];
        bits = sb;
        __syncthreads();
    }
    const float4* qa = sel ? P.qa1 : P.qa0;
    const float2* qb = sel ? P.qb1 : P.qb0;
    const uint32_t* qt = sel ? P.qt1 : P.qt0;
    const uint32_t n = P.cnt[sel];
    uint32_t* head = &P.cnt[2 + sel];
    if (blockIdx.x == 0 && threadIdx.x == 0) P.cnt[sel ^ 1] = 0u;
    Counters cnt = { 0, 0, 0, 0, 0, 0 };
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t refill_min = (uint32_t)a.refill_min;
    /* A short queue is spread over the waves: no wave takes more than its
     * fair share (at least one ray per lane), so the launch is not held by
     * a few waves tracing several rays per lane one after another. */
    uint32_t chunk = (uint32_t)a.refill_chunk;
    if (chunk) {
        const uint32_t waves = gridDim.x * (blockDim.x >> 6);
        const uint32_t per = (n + waves - 1u) / waves;
        uint32_t fair = ((per + 63u) >> 6) << 6;      /* >= 64: a wave never takes fewer than its idle lanes */
        fair = fair < 64u ? 64u : fair;
        if (fair < chunk) chunk = fair;
    }
    const int budget = a.dda_budget;
    DdaState D;
    GridView gl;                   /* the lane's scene grid (batch only) */
    const GridView& G = BATCH ? gl : a.g;
    uint32_t jq = 0;               /* queue position of the lane's ray (hits are stored there) */
    bool has = false;
    /* wave-uniform: the wave's reserved queue range [c_pos, c_end) */
    uint32_t c_pos = 0, c_end = 0;
    bool head_done = false;       /* the queue head has passed n */
    for (;;) {
        const unsigned long long idle = __ballot(!has);
        const uint32_t n_idle = (uint32_t)__popcll(idle);
        const bool exhausted = head_done && c_pos >= c_end;
        if (!exhausted && (n_idle >= refill_min || n_idle == 64u)) {
            const uint32_t avail = c_end - c_pos;
            const bool grab = avail < n_idle && !head_done;
            const uint32_t take = chunk ? chunk : n_idle - avail;
            uint32_t nb = 0;
            if (grab) {
                if (lane == 0) nb = atomicAdd(head, take);
                nb = __shfl(nb, 0, 64);
                if (nb + take >= n) head_done = true;
            }
            if (!has) {
                const uint32_t p = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                uint32_t j = 0xffffffffu;
                if (p < avail) j = c_pos + p;
                else if (grab && nb + (p - avail) < n) j = nb + (p - avail);
                if (j != 0xffffffffu) {
                    /* the queue is taken from its end: paths are queued in
                     * pixel order, and the last rows (towards the horizon)
                     * hold the longest rays, which then start first instead
                     * of prolonging the end of the launch */
                    j = n - 1u - j;
                    float4 o = qa[j];
                    float2 d = qb[j];
                    /* both parts of the entry in flight together: one
                     * memory latency per refill instead of two */
                    asm volatile("" : "+v"(o.x), "+v"(o.y), "+v"(o.z), "+v"(o.w), "+v"(d.x), "+v"(d.y));
                    jq = j;
                    if (BATCH) gl = B.tab[qt[j]].g;
                    if (D.init(G, unpack_ray(o, d))) {
                        has = true;
                    } else {
                        P.hit[j] = pack_hit(0, 0, 0, KIND_MISS);
                    }
                }
            }
            if (grab) {
                c_end = nb + take < n ? nb + take : n;
                c_pos = nb + (n_idle - avail);
                if (c_pos > c_end) c_pos = c_end;
            } else {
                c_pos += n_idle < avail ? n_idle : avail;
            }
        }
        if (!__ballot(has)) {
            if (head_done && c_pos >= c_end) break;
            continue;
        }
        if (has) {
            /* per-lane grids (batch) and grids whose padded masks span 4 GiB
             * address the masks with 64-bit per-lane arithmetic */
            int status;
            if (BATCH && B.packed)
                status = D.advance_wave_packed<STATS, false, false>(G, nullptr, budget, cnt.visits);
            else if (BATCH || G.wide)
                status = D.advance_wave<STATS, false, false>(G, G.bitsp, budget, cnt.visits);
            else if (LDSB && G.occr)
                status = D.advance_wave_packed<STATS, true, true>(G, bits, budget, cnt.visits);
            else
                status = D.advance_wave<STATS, LDSB, true>(G, bits, budget, cnt.visits);
            if (status) {
                /* kind and voxel only: the shade kernel reads the material slot */
                const int kind = status == 1 ? KIND_VOXEL : (D.y < 0 ? KIND_GROUND : KIND_MISS);
                P.hit[jq] = pack_hit(D.x, D.y, D.z, kind);
                has = false;
            }
        }
    }
    flush_counters<STATS>(BATCH ? B.stats : a.stats, cnt);
}

/* Largest training target kept: above it a sample is an outlier of the
 * path estimate (throughput ratios of near-zero-pdf bounces), whose
 * relative-loss gradient would dominate whole batches. */
#define CACHE_TARGET_MAX 1.0e4f

/* Target of a recorded vertex e (18 floats: query, throughput, radiance):
 * (radiance at the end - radiance at the vertex) / throughput, per channel;
 * false for a zero throughput channel or a non-finite or too large target. */
__device__ __forceinline__ bool vertex_target(const float* e, f3 r_end, float* t)
{
    if (!(e[12] > 0.0f && e[13] > 0.0f && e[14] > 0.0f)) return false;
    t[0] = (r_end.x - e[15]) / e[12];
    t[1] = (r_end.y - e[16]) / e[13];
    t[2] = (r_end.z - e[17]) / e[14];
    for (int c = 0; c < 3; ++c) {
        if (!(t[c] <= CACHE_TARGET_MAX)) return false;   /* also NaN */
        t[c] = t[c] > 0.0f ? t[c] : 0.0f;                /* rounding of R_end - R_v */
    }
    return true;
}

/* Training samples of a finished training path: one per recorded vertex
 * with a valid target. */
__device__ __forceinline__ void emit_training(const CacheRender& C, uint32_t slot, f3 r_end)
{
    const uint32_t nv = C.vrec_n[slot];
    const float* rec = C.vrec + (size_t)slot * C.max_vertices * 18u;
    uint32_t ok = 0;
    float t[3];
    for (uint32_t v = 0; v < nv; ++v) ok += vertex_target(rec + v * 18u, r_end, t) ? 1u : 0u;
    if (!ok) return;
    uint32_t base = atomicAdd(C.s_count, ok);
    for (uint32_t v = 0; v < nv; ++v) {
        const float* e = rec + v * 18u;
        if (!vertex_target(e, r_end, t)) continue;
        if (base < C.s_cap) {
            for (int k = 0; k < 12; ++k) C.s_raw[(size_t)base * 12u + k] = e[k];
            for (int c = 0; c < 3; ++c) C.s_tgt[(size_t)base * 3u + c] = t[c];
        }
        base++;
    }
}

/* Shading of one round.  PHASE PH_PRIMARY: the camera rays' hits (one
 * queue entry per path).  PH_COMB: every path's shadow ray (entry 2j) and
 * continuation ray (entry 2j + 1), traced in the same traversal launch:
 * the shadow result is applied first, then the continuation hit is shaded,
 * exactly the order of the reference's loop.  A path that goes on appends
 * its next shadow and continuation rays, which leave from the same point. */
template <bool STATS, int PHASE, bool BATCH, bool CACHE = false>
__global__ __launch_bounds__(WF_SHADE_THREADS) __attribute__((amdgpu_waves_per_eu(5))) void wf_shade_kernel(
    TraceArgs a0, BatchTable B, WfPaths P, int sel, CacheRender C = CacheRender())
{
    extern __shared__ __align__(16) uint8_t smem[];
    MatRec* recs0 = nullptr;
    float* cdfv0 = nullptr;
    const uint32_t* bits0 = nullptr;
    if (!BATCH) stage_lds<false>(a0, smem, recs0, cdfv0, bits0);
    float4* qao = sel ? P.qa0 : P.qa1;
    float2* qbo = sel ? P.qb0 : P.qb1;
    uint32_t* qto = BATCH ? (sel ? P.qt0 : P.qt1) : nullptr;
    const PathBuf& Bc = P.s[sel];
    const PathBuf& Bn = P.s[sel ^ 1];
    const uint32_t n = PHASE == PH_PRIMARY ? P.cnt[sel] : P.cnt[sel] / 2u;
    if (blockIdx.x == 0 && threadIdx.x == 0) P.cnt[2 + (sel ^ 1)] = 0u;
    Counters cnt = { 0, 0, 0, 0, 0, 0 };
    /* Shadow and continuation rounds: each batch of a workgroup's paths is
     * sorted so that the paths whose continuation ray left the grid (the sky
     * branch) come after those that hit a voxel or the ground (the shading
     * branch).  A wave then runs one of the two branches with most of its
     * lanes instead of both with some of them; in the first bounce about
     * half the continuation rays leave the grid.  Every path is shaded from
     * its own state and hit, so the results do not change, only the order
     * of the appended rays. */
    constexpr bool SORT = !CACHE && PHASE != PH_PRIMARY;
    constexpr uint32_t NW = WF_SHADE_THREADS / 64;
    __shared__ uint32_t s_cls[NW][3];
    __shared__ uint32_t s_src[SORT ? WF_SHADE_THREADS : 1];
    __shared__ uint4 s_hit[SORT ? WF_SHADE_THREADS : 1];
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        uint32_t j = base + threadIdx.x;
        uint4 hraw = make_uint4(0u, 0u, 0u, 0u);
        if (SORT) {
            uint32_t cls = 2u;                              /* no path */
            if (j < n) {
                hraw = ((const uint4*)P.hit)[j];
                Hit hh;
                cls = unpack_hit(make_uint2(hraw.z, hraw.w), hh) == KIND_MISS ? 1u : 0u;
            }
            /* counting sort over the workgroup: class, then wave, then lane */
            const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
            uint32_t rank = 0u;
            for (uint32_t k = 0; k < 3u; ++k) {
                const unsigned long long m = __ballot(cls == k);
                if (lane == 0) s_cls[wv][k] = (uint32_t)__popcll(m);
                if (cls == k) rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            }
            __syncthreads();
            uint32_t pos = rank;
            for (uint32_t k = 0; k < 3u; ++k)
                for (uint32_t w = 0; w < NW; ++w)
                    if (k < cls || (k == cls && w < wv)) pos += s_cls[w][k];
            s_src[pos] = threadIdx.x;
            s_hit[pos] = hraw;
            __syncthreads();
            j = base + s_src[threadIdx.x];
            hraw = s_hit[threadIdx.x];
        }
        bool keep = false, query = false;
        uint32_t sid = 0;
        f3 ro = mk(0.0f, 0.0f, 0.0f), rd = ro, rn = ro;
        float vraw[12];
        PathState S;
        uint4 me = make_uint4(0u, 0u, 0u, 0u);
        if (j < n) {
            load_phase<PHASE == PH_PRIMARY ? PH_PRIMARY : PH_SHADOW, BATCH || CACHE>(Bc, j, S, me);
            if (BATCH) sid = me.w;
            const TraceArgs& a = BATCH ? B.tab[sid] : a0;
            const MatRec* recs = BATCH ? a.recs : recs0;
            const float* cdfv = BATCH ? a.e.cdfv : cdfv0;
            const uint32_t* bits = BATCH ? a.g.bits : bits0;
            const uint32_t slot1 = CACHE ? me.w : 0u;   /* training slot + 1 */
            Hit h;
            h.steps = 0;
            /* the shadow and continuation results of a path are adjacent: one 16-byte load */
            uint2 hv;
            if (PHASE != PH_PRIMARY) {
                const uint4 hp = SORT ? hraw : ((const uint4*)P.hit)[j];
                hv = make_uint2(hp.z, hp.w);
                /* the shadow ray's result (never ends the path): it picks the
                 * radiance with or without the light sample (store_cont) */
                const int ks = unpack_hit(make_uint2(hp.x, hp.y), h);
                const bool visible = S.light_is_voxel ? (ks != KIND_MISS && h.x == S.ex && h.y == S.ey && h.z == S.ez)
                                                      : (ks == KIND_MISS);
                if (visible) S.radiance = S.direct;
                S.phase = PH_CONT;
                if (STATS) cnt.rays++;
            } else {
                hv = P.hit[j];
            }
            const int kind = unpack_hit(hv, h);
            /* the hit voxel's material slot is read here, where occupancy
             * hides the load, not by the traversal (ground: slot 0) */
            h.slot = kind == KIND_VOXEL ? fetch_slot(a.g, bits, h.x, h.y, h.z) : 0u;
            bool vrec = false;
            const int st = shade_step<STATS, CACHE>(a, recs, bits, cdfv, S, h, kind, cnt,
                                                    slot1 ? C.train_bounce : C.cache_bounce, vraw, &vrec);
            /* a training path records its vertices, except the one where it
             * ends into the cache (whose target would be the prediction) */
            if (CACHE && vrec && slot1 && st != 2) {
                const uint32_t slot = slot1 - 1u, nv = C.vrec_n[slot];
                if (nv < C.max_vertices) {
                    float* e = C.vrec + ((size_t)slot * C.max_vertices + nv) * 18u;
                    for (int k = 0; k < 12; ++k) e[k] = vraw[k];
                    e[12] = S.throughput.x; e[13] = S.throughput.y; e[14] = S.throughput.z;
                    e[15] = S.radiance.x; e[16] = S.radiance.y; e[17] = S.radiance.z;
                    C.vrec_n[slot] = nv + 1u;
                }
            }
            keep = st == 1;
            query = CACHE && st == 2;
            if (keep) {
                ro = S.ray_o;
                rd = S.ray_d;       /* shadow ray */
                rn = S.next_d;      /* continuation ray */
            } else if (!query) {
                const f3 c = tonemap(a.s, S.radiance);
                P.smp[me.y] = make_float4(c.x, c.y, c.z, 1.0f);
                if (CACHE && slot1) emit_training(C, slot1 - 1u, S.radiance);
            }
        }
        if (CACHE) {
            const uint32_t qi = block_reserve<WF_SHADE_THREADS>(query, C.q_count);
            if (query) {
                for (int k = 0; k < 12; ++k) C.q_raw[(size_t)qi * 12u + k] = vraw[k];
                C.q_meta[qi] = me;
                C.q_rad[qi] = make_float4(S.radiance.x, S.radiance.y, S.radiance.z, 0.0f);
                C.q_thr[qi] = make_float4(S.throughput.x, S.throughput.y, S.throughput.z, 0.0f);
            }
        }
        const uint32_t k = block_append2<WF_SHADE_THREADS>(keep, sid, ro, rd, rn, qao, qbo, qto, &P.cnt[sel ^ 1]);
        if (keep) store_cont<BATCH || CACHE>(Bn, k, S, me);
    }
    flush_counters<STATS>(BATCH ? B.stats : a0.stats, cnt);
}

/* accumulation.fs:17 over the chunk's samples, pass by pass per pixel (of
 * every scene of a batch: pixels of the scenes back to back). */
template <bool BATCH>
__global__ __launch_bounds__(256) void wf_accumulate_kernel(TraceArgs a0, BatchTable B, WfPaths P, uint64_t g0,
                                                            uint64_t g1, uint64_t n_pixels)
{
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_pixels;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t sid = 0;
        uint64_t p = q, pbase = 0;
        if (BATCH) {
            sid = scene_of(B.pix_base, B.n, q);
            p = q - B.pix_base[sid];
            pbase = B.path_base[sid];
        }
        const TraceArgs& a = BATCH ? B.tab[sid] : a0;
        const uint64_t WH = (uint64_t)a.W * (uint64_t)a.H;
        /* passes of this scene whose paths fall in [g0, g1) */
        const uint64_t lo = g0 > pbase ? g0 - pbase : 0, hi = g1 - pbase;
        if (g1 <= pbase || lo >= WH * a.count) continue;
        const uint64_t last = (hi - 1) / WH;
        const uint32_t k_lo = (uint32_t)(lo / WH), k_hi = last < a.count - 1 ? (uint32_t)last : a.count - 1;
        float4 acc = a.accum[p];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0;
        if (a.sum64) { s0 = a.sum64[3 * p]; s1 = a.sum64[3 * p + 1]; s2 = a.sum64[3 * p + 2]; }
        bool any = false;
        for (uint32_t k = k_lo; k <= k_hi; ++k) {
            const uint64_t gi = pbase + (uint64_t)k * WH + p;
            if (gi < g0 || gi >= g1) continue;
            any = true;
            const float4 c = P.smp[gi - g0];
            const uint32_t n = a.n0 + k;
            const float fn = (float)n, fn1 = (float)(n + 1u);
            acc.x = (c.x + acc.x * fn) / fn1;
            acc.y = (c.y + acc.y * fn) / fn1;
            acc.z = (c.z + acc.z * fn) / fn1;
            acc.w = (1.0f + acc.w * fn) / fn1;
            if (a.sum64) { s0 += (double)c.x; s1 += (double)c.y; s2 += (double)c.z; }
        }
        if (any) {
            a.accum[p] = acc;
            if (a.sum64) { a.sum64[3 * p] = s0; a.sum64[3 * p + 1] = s1; a.sum64[3 * p + 2] = s2; }
        }
    }
}

__global__ void wf_reset_kernel(uint32_t* cnt)
{
    if (threadIdx.x < 4) cnt[threadIdx.x] = 0u;
}

/* Resident workgroups per CU of a kernel (at least 1). */
template <typename K>
static int blocks_per_cu(K kernel, int threads, size_t lds)
{
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, threads, lds) != hipSuccess || b < 1) b = 1;
    return b;
}

/* generate -> (traverse -> shade) x rounds -> accumulate per chunk of at
 * most P.cap paths.  `a` describes the single scene (or, in a batch, carries
 * the traversal tuning and the LDS sizes of scene 0). */
template <bool STATS, bool LDSB, bool BATCH>
static hipError_t wf_run(const TraceArgs& a, const BatchTable& B, const WfPaths& P, uint64_t total,
                         uint64_t n_pixels, int rounds, int cus, hipStream_t s, const LaunchTimer* timer = nullptr,
                         hipEvent_t acc_wait = nullptr)
{
    const size_t lds_t = (!BATCH && LDSB) ? (size_t)a.g.bit_words_p * 4u : 0u;
    const size_t lds_s = BATCH ? 0 : 256 * sizeof(MatRec) + (size_t)a.cdfv_lds_bytes;
    const int bt = blocks_per_cu(wf_traverse_kernel<STATS, LDSB, BATCH>, WF_TRAVERSE_THREADS, lds_t);
    const int bs0 = blocks_per_cu(wf_shade_kernel<STATS, PH_PRIMARY, BATCH>, WF_SHADE_THREADS, lds_s);
    const int bs1 = blocks_per_cu(wf_shade_kernel<STATS, PH_COMB, BATCH>, WF_SHADE_THREADS, lds_s);
    for (uint64_t g0 = 0; g0 < total; g0 += P.cap) {
        const uint64_t g1 = g0 + P.cap < total ? g0 + P.cap : total;
        const uint32_t n = (uint32_t)(g1 - g0);
        wf_reset_kernel<<<1, 64, 0, s>>>(P.cnt);
        const uint32_t gn = (n + WF_GENERATE_THREADS - 1) / WF_GENERATE_THREADS;
        const unsigned gg = (unsigned)(gn < (uint32_t)cus * 2u ? gn : (uint32_t)cus * 2u);
        wf_generate_kernel<STATS, BATCH><<<gg, WF_GENERATE_THREADS, 0, s>>>(a, B, P, g0, n);
        int sel = 0;
        for (int r = 0; r < rounds; ++r) {
            const bool timed = timer && *timer->used < timer->cap;
            if (timed) {
                const hipError_t e = hipEventRecord(timer->ev[2 * *timer->used], s);
                if (e != hipSuccess) return e;
            }
            wf_traverse_kernel<STATS, LDSB, BATCH><<<cus * bt, WF_TRAVERSE_THREADS, lds_t, s>>>(a, B, P, sel);
            if (timed) {
                const hipError_t e = hipEventRecord(timer->ev[2 * *timer->used + 1], s);
                if (e != hipSuccess) return e;
                ++*timer->used;
            }
            if (r == 0)
                wf_shade_kernel<STATS, PH_PRIMARY, BATCH><<<cus * bs0, WF_SHADE_THREADS, lds_s, s>>>(a, B, P, sel);
            else
                wf_shade_kernel<STATS, PH_COMB, BATCH><<<cus * bs1, WF_SHADE_THREADS, lds_s, s>>>(a, B, P, sel);
            sel ^= 1;
        }
        const uint64_t gp = (n_pixels + 255) / 256;
        const unsigned ga = (unsigned)(gp < (uint64_t)cus * 8u ? gp : (uint64_t)cus * 8u);
        if (acc_wait) {
            const hipError_t e = hipStreamWaitEvent(s, acc_wait, 0);
            if (e != hipSuccess) return e;
        }
        wf_accumulate_kernel<BATCH><<<ga, 256, 0, s>>>(a, B, P, g0, g1, n_pixels);
    }
    return hipGetLastError();
}

hipError_t launch_render_wavefront(const TraceArgs& a, const WfPaths& P, int cus, hipStream_t s, bool stats,
                                   const LaunchTimer* timer, hipEvent_t acc_wait)
{
    const BatchTable B = { nullptr, nullptr, nullptr, 0u, nullptr, 0 };
    const uint64_t WH = (uint64_t)a.W * (uint64_t)a.H;
    const uint64_t total = WH * a.count;
    const int rounds = 1 + (a.s.max_bounces > 0 ? a.s.max_bounces : 0);
    const bool lds_bits = a.lds_bits_p;
    if (stats)
        return lds_bits ? wf_run<true, true, false>(a, B, P, total, WH, rounds, cus, s, timer, acc_wait)
                        : wf_run<true, false, false>(a, B, P, total, WH, rounds, cus, s, timer, acc_wait);
    return lds_bits ? wf_run<false, true, false>(a, B, P, total, WH, rounds, cus, s, timer, acc_wait)
                    : wf_run<false, false, false>(a, B, P, total, WH, rounds, cus, s, timer, acc_wait);
}

hipError_t launch_render_batch(const TraceArgs& a0, const BatchTable& B, const WfPaths& P, uint64_t total_paths,
                               uint64_t total_pixels, int max_bounces, int cus, hipStream_t s, bool stats)
{
    const int rounds = 1 + (max_bounces > 0 ? max_bounces : 0);
    if (stats) return wf_run<true, false, true>(a0, B, P, total_paths, total_pixels, rounds, cus, s);
    return wf_run<false, false, true>(a0, B, P, total_paths, total_pixels, rounds, cus, s);
}

/* Paths that ended into the cache since the last apply (queries
 * [*q_done, *q_count)): radiance += throughput * max(0, cache); a training
 * path that ended so emits its samples with that radiance as the end. */
__global__ __launch_bounds__(256) void cache_apply_kernel(TraceArgs a, WfPaths P, CacheRender C)
{
    const uint32_t lo = *C.q_done, n = *C.q_count;
    for (uint32_t i = lo + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint4 me = C.q_meta[i];
        const float4 r = C.q_rad[i], t = C.q_thr[i];
        const float y0 = C.q_out[(size_t)i * 3u], y1 = C.q_out[(size_t)i * 3u + 1], y2 = C.q_out[(size_t)i * 3u + 2];
        /* a channel the cache predicts as 0 (or less) adds nothing */
        const f3 rad = mk(y0 > 0.0f ? r.x + t.x * y0 : r.x, y1 > 0.0f ? r.y + t.y * y1 : r.y,
                          y2 > 0.0f ? r.z + t.z * y2 : r.z);
        const f3 c = tonemap(a.s, rad);
        P.smp[me.y] = make_float4(c.x, c.y, c.z, 1.0f);
        if (me.w) emit_training(C, me.w - 1u, rad);
    }
}

__global__ void cache_mark_kernel(CacheRender C)
{
    if (threadIdx.x == 0) *C.q_done = *C.q_count;
}

template <bool LDSB>
static hipError_t wf_run_cached(const TraceArgs& a, const WfPaths& P, const CacheRender& C, const void* pack,
                                const float* bias, int cus, hipStream_t s)
{
    const BatchTable B = { nullptr, nullptr, nullptr, 0u, nullptr, 0 };
    const size_t lds_t = LDSB ? (size_t)a.g.bit_words_p * 4u : 0u;
    const size_t lds_s = 256 * sizeof(MatRec) + (size_t)a.cdfv_lds_bytes;
    const int bt = blocks_per_cu(wf_traverse_kernel<false, LDSB, false>, WF_TRAVERSE_THREADS, lds_t);
    const int bs0 = blocks_per_cu(wf_shade_kernel<false, PH_PRIMARY, false, true>, WF_SHADE_THREADS, lds_s);
    const int bs1 = blocks_per_cu(wf_shade_kernel<false, PH_COMB, false, true>, WF_SHADE_THREADS, lds_s);
    const uint64_t WH = (uint64_t)a.W * (uint64_t)a.H;
    if (WH > P.cap) return hipErrorInvalidValue;
    const uint32_t n = (uint32_t)WH;
    int rounds = 1 + (a.s.max_bounces > 0 ? a.s.max_bounces : 0);
    const int qround = C.cache_bounce >= 0 ? C.cache_bounce : -1;
    /* without training paths nothing goes past the cache bounce */
    const int tround = C.train_slots == 0 ? qround : (C.train_bounce >= 0 ? C.train_bounce : -1);
    /* round b shades the vertices of bounce b, so once the plain and the
     * training paths have ended into the cache no round is left */
    if (qround >= 0 && tround >= 0) rounds = std::min(rounds, 1 + std::max(qround, tround));
    wf_reset_kernel<<<1, 64, 0, s>>>(P.cnt);
    const uint32_t gn = (n + WF_GENERATE_THREADS - 1) / WF_GENERATE_THREADS;
    const unsigned gg = (unsigned)(gn < (uint32_t)cus * 2u ? gn : (uint32_t)cus * 2u);
    wf_generate_kernel<false, false, true><<<gg, WF_GENERATE_THREADS, 0, s>>>(a, B, P, 0, n, C);
    int sel = 0;
    for (int r = 0; r < rounds; ++r) {
        wf_traverse_kernel<false, LDSB, false><<<cus * bt, WF_TRAVERSE_THREADS, lds_t, s>>>(a, B, P, sel);
        if (r == 0)
            wf_shade_kernel<false, PH_PRIMARY, false, true><<<cus * bs0, WF_SHADE_THREADS, lds_s, s>>>(a, B, P, sel, C);
        else
            wf_shade_kernel<false, PH_COMB, false, true><<<cus * bs1, WF_SHADE_THREADS, lds_s, s>>>(a, B, P, sel, C);
        if (r == qround || r == tround) {
            hipError_t e = launch_mlp_infer_dev((const _Float16*)pack, bias, C.q_raw, C.q_out, C.q_done, C.q_count, n,
                                                cus, s);
            if (e != hipSuccess) return e;
            cache_apply_kernel<<<(unsigned)cus * 4u, 256, 0, s>>>(a, P, C);
            cache_mark_kernel<<<1, 64, 0, s>>>(C);
        }
        sel ^= 1;
    }
    const uint64_t gp = (WH + 255) / 256;
    const unsigned ga = (unsigned)(gp < (uint64_t)cus * 8u ? gp : (uint64_t)cus * 8u);
    wf_accumulate_kernel<false><<<ga, 256, 0, s>>>(a, B, P, 0, WH, WH);
    return hipGetLastError();
}

hipError_t launch_render_cached(const TraceArgs& a, const WfPaths& P, const CacheRender& C, const void* pack,
                                const float* bias, int cus, hipStream_t s)
{
    return a.lds_bits_p ? wf_run_cached<true>(a, P, C, pack, bias, cus, s)
                        : wf_run_cached<false>(a, P, C, pack, bias, cus, s);
}

/* Emissive list entry i -> (x, y, z, material slot) of its voxel: the
 * linear index decoded as coordinates.h:118-128 builds it, the slot read
 * with texelFetch semantics (offset 0 outside the grid). */
__global__ __launch_bounds__(256) void light_table_kernel(GridView g, int4* __restrict__ out)
{
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < g.n_emissive; i += gridDim.x * blockDim.x) {
        int32_t idx = g.emissive[i];
        const int32_t dz = g.nx * g.ny, dy = g.nx;
        const int32_t z = idx / dz;
        idx -= z * dz;
        const int32_t y = idx / dy;
        idx -= y * dy;
        out[i] = make_int4(idx, y, z, (int32_t)fetch_slot(g, g.bits, idx, y, z));
    }
}

hipError_t launch_light_table(const GridView& g, int4* out, hipStream_t s)
{
    if (g.n_emissive <= 0) return hipSuccess;
    const int32_t blocks = (g.n_emissive + 255) / 256;
    light_table_kernel<<<blocks < 4096 ? blocks : 4096, 256, 0, s>>>(g, out);
    return hipGetLastError();
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

static void (*render_ptr(bool stats, bool ldsb))(TraceArgs)
{
    if (stats) return ldsb ? render_kernel<true, true> : render_kernel<true, false>;
    return ldsb ? render_kernel<false, true> : render_kernel<false, false>;
}

hipError_t launch_render(const TraceArgs& a, int grid, hipStream_t s, bool stats)
{
    const size_t lds = render_lds_bytes(a);
    render_ptr(stats, a.lds_bits != 0)<<<grid, 256, lds, s>>>(a);
    return hipGetLastError();
}

int render_blocks_per_cu(const TraceArgs& a, bool stats)
{
    int blocks = 0;
    const size_t lds = render_lds_bytes(a);
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, render_ptr(stats, a.lds_bits != 0),
                                                                       256, lds);
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

/*
 * vt_cache.hip - kernels of the learned radiance cache (see vt_cache.h).
 *
 * Inference runs on the matrix cores: v_mfma_f32_16x16x32_f16, FP16
 * operands, FP32 accumulation.  Each wave evaluates 16 queries at a time
 * with the queries as the MFMA N dimension (columns) and the output features
 * as M (rows): a layer is Y^T = W * X^T, so a 16x16 output tile, held by
 * lane l as rows 4*(l/16)+i of column l%16, is already laid out as the B
 * operand of the next layer; no transposes go through LDS.  Weights are
 * staged once per workgroup in LDS, pre-swizzled so that every MFMA A
 * operand is one 16-byte LDS read per lane, with the column order each
 * K-step's B operand implies.
 *
 * Training (mixed precision, FP32 master weights):
 *  - forward exactly as inference (FP16 operands), keeping the FP16 layer
 *    inputs; backward on v_mfma_f32_16x16x4_f32 with FP32 deltas (the
 *    relative loss's gradients span too wide a range for FP16), transposed
 *    weights pre-swizzled in the same lane order as the D tiles, so the
 *    backward chain also stays inside the lanes;
 *  - the weight gradients dW = DZ * A^T as an FP32-MFMA GEMM over the
 *    samples (K), split in MLP_GRAD_SPLITS K ranges whose partial sums are
 *    written (not atomically added) and summed in a fixed order by the
 *    update kernel, so a step is deterministic;
 *  - Adam on the FP32 master weights, which also rewrites the FP16 pack, the
 *    bias table and the FP32 transposed pack entry of every parameter.
 */
#include "vt_cache.h"

namespace vt {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

/* ---- weight packs -------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* __restrict__ P, _Float16* __restrict__ pack,
                                                       float* __restrict__ bias, float* __restrict__ tpack)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < MLP_PACK) {
        const int src = pack_source(e);
        pack[e] = src < 0 ? (_Float16)0.0f : (_Float16)P[src];
    }
    if (e < MLP_TPACK) {
        const int src = tpack_source(e);
        tpack[e] = src < 0 ? 0.0f : P[src];
    }
    if (e < MLP_BIAS) {
        const int layer = e / MLP_W, o = e % MLP_W;
        float b = 0.0f;
        if (layer < 5) b = P[mlp_b_off(layer) + o];
        else if (o < MLP_OUT) b = P[mlp_b_off(5) + o];
        bias[e] = b;
    }
}

/* ---- MFMA inference ---------------------------------------------------------------- */
#define INFER_THREADS 512

/* bias + ReLU of the four 16-row tiles of a 64-wide layer, converted to the
 * two K-step B operands of the next layer */
__device__ __forceinline__ void relu_to_b(const f4* acc, const float* b, int g, h8* B)
{
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float z = acc[m][i] + b[16 * m + 4 * g + i];
            z = z > 0.0f ? z : 0.0f;
            B[m >> 1][(m & 1) * 4 + i] = (_Float16)z;
        }
}

__global__ __launch_bounds__(INFER_THREADS) void mlp_infer_kernel(const h8* __restrict__ pack,
                                                                  const float* __restrict__ bias,
                                                                  const float* __restrict__ raw,
                                                                  float* __restrict__ out, int64_t n,
                                                                  const uint32_t* __restrict__ lo_dev,
                                                                  const uint32_t* __restrict__ n_dev)
{
    if (n_dev) n = *n_dev;
    const int64_t lo = lo_dev ? (int64_t)*lo_dev : 0;
    extern __shared__ __align__(16) uint8_t smem[];
    h8* wp = (h8*)smem;
    float* bs = (float*)(smem + MLP_PACK * 2);
    for (int i = threadIdx.x; i < MLP_PACK / 8; i += blockDim.x) wp[i] = pack[i];
    for (int i = threadIdx.x; i < MLP_BIAS; i += blockDim.x) bs[i] = bias[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const h8* w1 = wp;
    const h8* wh = wp + MLP_PACK_L1 / 8;
    const h8* wo = wp + (MLP_PACK_L1 + 4 * MLP_PACK_HID) / 8;
    for (int64_t tile = lo / 16 + wave; tile * 16 < n; tile += waves) {
        const int64_t q = tile * 16 + r;
        const bool valid = q >= lo && q < n;
        h8 B1[9];
        {
            float enc[72];
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) encode_scalar(valid ? raw[q * MLP_RAW + 3 * g + jj] : 0.0f, enc + 24 * jj);
#pragma unroll
            for (int s = 0; s < 9; ++s)
#pragma unroll
                for (int j = 0; j < 8; ++j) B1[s][j] = (_Float16)enc[8 * s + j];
        }
        f4 acc[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            acc[m] = (f4){ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
            for (int s = 0; s < 9; ++s)
                acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w1[(m * 9 + s) * 64 + lane], B1[s], acc[m], 0, 0, 0);
        }
        h8 B[2];
        relu_to_b(acc, bs, g, B);
#pragma unroll
        for (int L = 0; L < 4; ++L) {
            const h8* w = wh + L * (MLP_PACK_HID / 8);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                acc[m] = (f4){ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w[(m * 2 + s) * 64 + lane], B[s], acc[m], 0, 0, 0);
            }
            relu_to_b(acc, bs + MLP_W * (L + 1), g, B);
        }
        f4 o = (f4){ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int s = 0; s < 2; ++s) o = __builtin_amdgcn_mfma_f32_16x16x32_f16(wo[s * 64 + lane], B[s], o, 0, 0, 0);
        if (valid && g == 0) {
            out[q * 3 + 0] = o[0] + bs[5 * MLP_W + 0];
            out[q * 3 + 1] = o[1] + bs[5 * MLP_W + 1];
            out[q * 3 + 2] = o[2] + bs[5 * MLP_W + 2];
        }
    }
}

/* ---- MFMA training -------------------------------------------------------------------- */
#define FB_THREADS 512

__device__ __forceinline__ uint32_t batch_count(const TrainBatch& b)
{
    uint32_t n = b.n_dev ? *b.n_dev : b.cap;
    n = n < b.cap ? n : b.cap;
    n = n > b.off ? n - b.off : 0u;
    return n < b.batch ? n : b.batch;
}

/* Forward (FP16 MFMA) + relative-L2 loss + backward (FP32 MFMA) of one
 * batch, 16 samples per wave and tile.  Loss (Mueller et al. 2021, "Real-time
 * neural radiance caching", eq. 3): (y - t)^2 / (sg(y)^2 + 0.01) per channel,
 * averaged over channels and samples.  Samples past n (up to the tile) get
 * dL/dy = 0, so they add nothing to the gradient GEMM. */
__global__ __launch_bounds__(FB_THREADS) void mlp_train_fb_kernel(const h8* __restrict__ pack,
                                                                  const float* __restrict__ bias,
                                                                  const f4* __restrict__ tpack, TrainBatch b,
                                                                  TrainScratch S)
{
    const uint32_t n = batch_count(b);
    if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0u && S.steps) *S.steps += 1u;
    if (n == 0u) return;
    extern __shared__ __align__(16) uint8_t smem[];
    h8* wp = (h8*)smem;
    float* bs = (float*)(smem + MLP_PACK * 2);
    f4* wt = (f4*)(smem + MLP_PACK * 2 + MLP_BIAS * 4);
    for (int i = threadIdx.x; i < MLP_PACK / 8; i += blockDim.x) wp[i] = pack[i];
    for (int i = threadIdx.x; i < MLP_BIAS; i += blockDim.x) bs[i] = bias[i];
    for (int i = threadIdx.x; i < MLP_TPACK / 4; i += blockDim.x) wt[i] = tpack[i];
    __syncthreads();
    const float* wto = (const float*)(wt + 4 * MLP_TPACK_HID / 4);
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const h8* w1 = wp;
    const h8* wh = wp + MLP_PACK_L1 / 8;
    const h8* wo = wp + (MLP_PACK_L1 + 4 * MLP_PACK_HID) / 8;
    const float inv = 1.0f / (float)(MLP_OUT * n);
    float lsum = 0.0f;
    const uint32_t tiles = (n + 15u) / 16u;
    for (uint32_t tile = wave; tile < tiles; tile += waves) {
        const uint32_t ql = tile * 16u + (uint32_t)r;
        const bool valid = ql < n;
        const size_t qg = (size_t)b.off + ql;
        _Float16* xt = S.X + (size_t)tile * (MLP_IN * 16) + 72 * 16 * g + r;
        _Float16* ht = S.H + (size_t)tile * (5 * MLP_W * 16) + 4 * 16 * g + r;
        float* zt = S.DZ + (size_t)tile * (5 * MLP_W * 16) + 4 * 16 * g + r;
        float* yt = S.DY + (size_t)tile * (MLP_OUT * 16) + r;
        h8 B1[9];
#pragma unroll
        for (int jj = 0; jj < 3; ++jj) {
            float e[24];
            encode_scalar(valid ? b.raw[qg * MLP_RAW + 3 * g + jj] : 0.0f, e);
#pragma unroll
            for (int k = 0; k < 24; ++k) {
                const int f = 24 * jj + k;
                B1[f >> 3][f & 7] = (_Float16)e[k];
                xt[16 * f] = B1[f >> 3][f & 7];
            }
        }
        f4 acc[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            acc[m] = (f4){ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
            for (int s = 0; s < 9; ++s)
                acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w1[(m * 9 + s) * 64 + lane], B1[s], acc[m], 0, 0, 0);
        }
        h8 Bh[5][2];
        relu_to_b(acc, bs, g, Bh[0]);
#pragma unroll
        for (int L = 0; L < 4; ++L) {
            const h8* w = wh + L * (MLP_PACK_HID / 8);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                acc[m] = (f4){ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
                for (int s = 0; s < 2; ++s)
                    acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w[(m * 2 + s) * 64 + lane], Bh[L][s], acc[m], 0, 0, 0);
            }
            relu_to_b(acc, bs + MLP_W * (L + 1), g, Bh[L + 1]);
            __builtin_amdgcn_sched_barrier(0);
        }
        /* layer inputs A_1..A_5: feature 16m + 4g + i of sample ql */
#pragma unroll
        for (int L = 0; L < 5; ++L)
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    ht[16 * (L * MLP_W + 16 * m + i)] = Bh[L][m >> 1][(m & 1) * 4 + i];
        f4 o = (f4){ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int s = 0; s < 2; ++s) o = __builtin_amdgcn_mfma_f32_16x16x32_f16(wo[s * 64 + lane], Bh[4][s], o, 0, 0, 0);
        /* loss and dL/dy in lanes 0..15 (rows 0..2 of the output tile) */
        float dy[MLP_OUT] = { 0.0f, 0.0f, 0.0f };
        if (g == 0) {
#pragma unroll
            for (int c = 0; c < MLP_OUT; ++c) {
                if (valid) {
                    const float y = o[c] + bs[5 * MLP_W + c];
                    const float t = b.target[qg * MLP_OUT + c];
                    const float w = 1.0f / (y * y + 0.01f);
                    const float d = y - t;
                    lsum += d * d * w;
                    dy[c] = (2.0f * d * w) * inv;
                }
                yt[16 * c] = dy[c];
            }
        }
        /* B operand of W_6^T: lane group c carries dL/dy_c of sample r */
        const float e0 = __shfl(dy[0], r, 64), e1 = __shfl(dy[1], r, 64), e2 = __shfl(dy[2], r, 64);
        const float bdy = g == 0 ? e0 : (g == 1 ? e1 : (g == 2 ? e2 : 0.0f));
        f4 dh[4];
#pragma unroll
        for (int m = 0; m < 4; ++m)
            dh[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(wto[m * 64 + lane], bdy, (f4){ 0.0f, 0.0f, 0.0f, 0.0f }, 0,
                                                         0, 0);
#pragma unroll
        for (int L = 4; L >= 0; --L) {
            f4 dz[4];
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    dz[m][i] = (float)Bh[L][m >> 1][(m & 1) * 4 + i] > 0.0f ? dh[m][i] : 0.0f;
                    zt[16 * (L * MLP_W + 16 * m + i)] = dz[m][i];
                }
            if (L > 0) {
                const f4* w = wt + (L - 1) * (MLP_TPACK_HID / 4);
#pragma unroll
                for (int m = 0; m < 4; ++m) dh[m] = (f4){ 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
                for (int tq = 0; tq < 4; ++tq)
#pragma unroll
                    for (int m = 0; m < 4; ++m) {
                        const f4 a = w[(m * 4 + tq) * 64 + lane];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            dh[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], dz[tq][u], dh[m], 0, 0, 0);
                    }
            }
            /* keep the next layer's LDS reads from being hoisted (registers) */
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
    if (lane == 0 && lsum != 0.0f) atomicAdd(S.loss, lsum);
}

/* Column blocks of the weight-gradient GEMM: (layer, first 16-column tile,
 * tiles); column `in` of a layer is its bias (input 1). */
__device__ __forceinline__ void grad_item(int item, int& layer, int& nt0, int& nts)
{
    if (item < 5) { layer = 0; nt0 = 4 * item; nts = item < 4 ? 4 : 3; return; }
    layer = 1 + (item - 5) / 2;
    const int k = (item - 5) % 2;
    nt0 = 4 * k;
    nts = k == 0 ? 4 : 1;
}

/* part[split][param] = sum over the split's samples of dZ[o][q] * A[i][q]
 * (A[in][q] = 1 for the bias), FP32 MFMA with 16 samples per K step. */
__global__ __launch_bounds__(64) void mlp_train_grad_kernel(TrainBatch b, TrainScratch S)
{
    const uint32_t n = batch_count(b);
    if (n == 0u) return;
    int layer, nt0, nts;
    grad_item(blockIdx.x, layer, nt0, nts);
    const int split = blockIdx.y, splits = gridDim.y;
    const int in = layer == 0 ? MLP_IN : MLP_W, out = layer == 5 ? MLP_OUT : MLP_W;
    /* per-tile strides and the layer's first row inside a tile */
    const size_t dstride = layer == 5 ? MLP_OUT * 16 : 5 * MLP_W * 16;
    const size_t astride = layer == 0 ? MLP_IN * 16 : 5 * MLP_W * 16;
    const float* D = layer == 5 ? S.DY : S.DZ + (size_t)layer * MLP_W * 16;
    const _Float16* A = layer == 0 ? S.X : S.H + (size_t)(layer - 1) * MLP_W * 16;
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const uint32_t tiles = (n + 15u) / 16u;
    const uint32_t per = (tiles + splits - 1) / splits;
    const uint32_t t0 = split * per, t1 = t0 + per < tiles ? t0 + per : tiles;
    f4 acc[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[m][c] = (f4){ 0.0f, 0.0f, 0.0f, 0.0f };
    const int mts = out > 16 ? 4 : 1;
    for (uint32_t t = t0; t < t1; ++t) {
        const float* Dt = D + (size_t)t * dstride + 4 * g;
        const _Float16* At = A + (size_t)t * astride + 4 * g;
        f4 dv[4], av[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int row = 16 * m + r;
            dv[m] = (m < mts && row < out) ? *(const f4*)(Dt + 16 * row) : (f4){ 0.0f, 0.0f, 0.0f, 0.0f };
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int col = 16 * (nt0 + c) + r;
            if (c < nts && col < in) {
                typedef _Float16 h4 __attribute__((ext_vector_type(4)));
                const h4 h = *(const h4*)(At + 16 * col);
                av[c] = (f4){ (float)h[0], (float)h[1], (float)h[2], (float)h[3] };
            } else {
                const float one = (c < nts && col == in) ? 1.0f : 0.0f;
                av[c] = (f4){ one, one, one, one };
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    if (m < mts && c < nts)
                        acc[m][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(dv[m][u], av[c][u], acc[m][c], 0, 0, 0);
    }
    float* part = S.part + (size_t)split * MLP_N_PARAMS;
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            if (m >= mts || c >= nts) continue;
            const int i = 16 * (nt0 + c) + r;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int o = 16 * m + 4 * g + v;
                if (o >= out || i > in) continue;
                part[i < in ? mlp_w_off(layer) + o * in + i : mlp_b_off(layer) + o] = acc[m][c][v];
            }
        }
}

/* Sum of the split partials (fixed order); then either the gradient out or
 * Adam (Kingma & Ba 2015, bias-corrected, step count on the device) and the
 * packs' entries of the parameter. */
__global__ __launch_bounds__(256) void mlp_train_update_kernel(TrainBatch b, TrainScratch S, float* __restrict__ grad,
                                                               float* __restrict__ P, float* __restrict__ M,
                                                               float* __restrict__ V, _Float16* __restrict__ pack,
                                                               float* __restrict__ bias, float* __restrict__ tpack,
                                                               float lr, float b1, float b2, float eps)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_N_PARAMS || batch_count(b) == 0u) return;
    float gsum = 0.0f;
    for (int s = 0; s < MLP_GRAD_SPLITS; ++s) gsum += S.part[(size_t)s * MLP_N_PARAMS + p];
    if (grad) {
        grad[p] = gsum;
        return;
    }
    const float t = (float)*S.steps;
    const float m = b1 * M[p] + (1.0f - b1) * gsum;
    const float v = b2 * V[p] + (1.0f - b2) * gsum * gsum;
    M[p] = m;
    V[p] = v;
    const float c1 = 1.0f - powf(b1, t), c2 = 1.0f - powf(b2, t);
    const float w = P[p] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[p] = w;
    /* locate the parameter: layer, weight (o, i) or bias o */
    int layer = 5;
    while (layer > 0 && p < mlp_w_off(layer)) --layer;
    const int in = layer == 0 ? MLP_IN : MLP_W;
    const int local = p - mlp_w_off(layer);
    if (p >= mlp_b_off(layer)) {
        const int o = p - mlp_b_off(layer);
        bias[layer * MLP_W + o] = w;
        return;
    }
    const int o = local / in, i = local % in;
    pack[pack_dest(layer, o, i)] = (_Float16)w;
    if (layer > 0) tpack[tpack_dest(layer, o, i)] = w;
}

/* ---- launchers -------------------------------------------------------------------------- */
hipError_t launch_mlp_pack(const float* params, _Float16* pack, float* bias, float* tpack, hipStream_t s)
{
    mlp_pack_kernel<<<(MLP_PACK + 255) / 256, 256, 0, s>>>(params, pack, bias, tpack);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const _Float16* pack, const float* bias, const float* raw, float* out, int64_t n, int cus,
                            hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const size_t lds = (size_t)MLP_PACK * 2 + (size_t)MLP_BIAS * 4;
    static bool attr = false;
    if (!attr) {
        const hipError_t e = hipFuncSetAttribute((const void*)mlp_infer_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    const int64_t tiles = (n + 15) / 16;
    const int64_t want = (tiles + INFER_THREADS / 64 - 1) / (INFER_THREADS / 64);
    const int64_t cap = (int64_t)cus * 2;
    mlp_infer_kernel<<<(unsigned)(want < cap ? want : cap), INFER_THREADS, lds, s>>>((const h8*)pack, bias, raw, out,
                                                                                    n, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_mlp_infer_dev(const _Float16* pack, const float* bias, const float* raw, float* out,
                                const uint32_t* lo_dev, const uint32_t* n_dev, int64_t n_max, int cus, hipStream_t s)
{
    if (n_max <= 0) return hipSuccess;
    const size_t lds = (size_t)MLP_PACK * 2 + (size_t)MLP_BIAS * 4;
    const hipError_t e = hipFuncSetAttribute((const void*)mlp_infer_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)lds);
    if (e != hipSuccess) return e;
    const int64_t tiles = (n_max + 15) / 16;
    const int64_t want = (tiles + INFER_THREADS / 64 - 1) / (INFER_THREADS / 64);
    const int64_t cap = (int64_t)cus * 2;
    mlp_infer_kernel<<<(unsigned)(want < cap ? want : cap), INFER_THREADS, lds, s>>>((const h8*)pack, bias, raw, out,
                                                                                    n_max, lo_dev, n_dev);
    return hipGetLastError();
}

/* Running totals of a cached render's queries and (stored) training samples. */
__global__ void cache_tally_kernel(const uint32_t* __restrict__ counts, uint32_t scap,
                                   unsigned long long* __restrict__ tally)
{
    if (threadIdx.x == 0) {
        tally[0] += counts[0];
        tally[1] += counts[1] < scap ? counts[1] : scap;
    }
}

hipError_t launch_cache_tally(const uint32_t* counts, uint32_t scap, unsigned long long* tally, hipStream_t s)
{
    cache_tally_kernel<<<1, 64, 0, s>>>(counts, scap, tally);
    return hipGetLastError();
}

hipError_t launch_mlp_train_fb(const _Float16* pack, const float* bias, const float* tpack, const TrainBatch& b,
                               const TrainScratch& S, int cus, hipStream_t s)
{
    if (b.batch == 0u || 16ull * S.tiles < b.batch) return hipErrorInvalidValue;
    const size_t lds = (size_t)MLP_PACK * 2 + (size_t)MLP_BIAS * 4 + (size_t)MLP_TPACK * 4;
    static bool attr = false;
    if (!attr) {
        const hipError_t e = hipFuncSetAttribute((const void*)mlp_train_fb_kernel,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        attr = true;
    }
    const uint32_t tiles = (b.batch + 15u) / 16u;
    const uint32_t want = (tiles + FB_THREADS / 64 - 1) / (FB_THREADS / 64);
    const uint32_t grid = want < (uint32_t)cus ? want : (uint32_t)cus;
    mlp_train_fb_kernel<<<grid, FB_THREADS, lds, s>>>((const h8*)pack, bias, (const f4*)tpack, b, S);
    return hipGetLastError();
}

hipError_t launch_mlp_train_grad(const TrainBatch& b, const TrainScratch& S, hipStream_t s)
{
    if (b.batch == 0u || 16ull * S.tiles < b.batch) return hipErrorInvalidValue;
    mlp_train_grad_kernel<<<dim3(MLP_GRAD_ITEMS, MLP_GRAD_SPLITS), 64, 0, s>>>(b, S);
    return hipGetLastError();
}

hipError_t launch_mlp_train_update(const TrainBatch& b, const TrainScratch& S, float* grad, float* params, float* m,
                                   float* v, _Float16* pack, float* bias, float* tpack, float lr, float b1, float b2,
                                   float eps, hipStream_t s)
{
    mlp_train_update_kernel<<<(MLP_N_PARAMS + 255) / 256, 256, 0, s>>>(b, S, grad, params, m, v, pack, bias, tpack, lr,
                                                                       b1, b2, eps);
    return hipGetLastError();
}

}  // namespace vt

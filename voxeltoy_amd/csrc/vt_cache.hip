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
 * Training is FP32 on the vector ALUs: a per-sample forward/backward kernel
 * that writes feature-major activations and deltas, a tiled GEMM kernel that
 * reduces dW = DZ * X^T over the samples, and Adam.
 */
#include "vt_cache.h"

namespace vt {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

/* ---- weight pack ------------------------------------------------------------------- */
/* Pack entry e -> parameter index (or -1 for padding).  Layout per layer:
 * [m (16-row output tile)][s (32-wide K step)][lane][j (8 halves)].  Lane l
 * supplies row o = 16m + l%16; its K slots j are input columns
 *   layer 1: 72*(l/16) + 8s + j (lane group g encodes raw scalars 3g..3g+2),
 *   others : 32s + 4g + j (j < 4), 32s + 16 + 4g + (j - 4) (j >= 4),
 * matching the B operand built from the previous layer's output tiles. */
__host__ __device__ inline int pack_source(int e)
{
    int layer, local;
    if (e < MLP_PACK_L1) { layer = 0; local = e; }
    else if (e < MLP_PACK_L1 + 4 * MLP_PACK_HID) { layer = 1 + (e - MLP_PACK_L1) / MLP_PACK_HID; local = (e - MLP_PACK_L1) % MLP_PACK_HID; }
    else { layer = 5; local = e - MLP_PACK_L1 - 4 * MLP_PACK_HID; }
    const int steps = layer == 0 ? 9 : 2;
    const int j = local & 7, lane = (local >> 3) & 63, ms = local >> 9;
    const int s = ms % steps, m = ms / steps;
    const int g = lane >> 4, o = 16 * m + (lane & 15);
    const int in = layer == 0 ? MLP_IN : MLP_W, out = layer == 5 ? MLP_OUT : MLP_W;
    if (o >= out) return -1;
    const int i = layer == 0 ? 72 * g + 8 * s + j : (j < 4 ? 32 * s + 4 * g + j : 32 * s + 16 + 4 * g + (j - 4));
    return mlp_w_off(layer) + o * in + i;
}

__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* __restrict__ P, _Float16* __restrict__ pack,
                                                       float* __restrict__ bias)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < MLP_PACK) {
        const int src = pack_source(e);
        pack[e] = src < 0 ? (_Float16)0.0f : (_Float16)P[src];
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
                                                                  const uint32_t* __restrict__ n_dev)
{
    if (n_dev) n = *n_dev;
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
    for (int64_t tile = wave; tile * 16 < n; tile += waves) {
        const int64_t q = tile * 16 + r;
        const bool valid = q < n;
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

/* ---- FP32 training ------------------------------------------------------------------ */
/* One thread per sample: forward, relative-L2 loss (Mueller et al. 2021,
 * "Real-time neural radiance caching", eq. 3: (y - t)^2 / (sg(y)^2 + 0.01)
 * per channel, averaged), backward.  All arrays feature-major (f * n + q). */
__global__ __launch_bounds__(256) void mlp_fwd_bwd_kernel(const float* __restrict__ P, const float* __restrict__ raw,
                                                          const float* __restrict__ target, int64_t n,
                                                          float* __restrict__ X, float* __restrict__ H,
                                                          float* __restrict__ DZ, float* __restrict__ DY,
                                                          float* __restrict__ loss_sum)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float lsum = 0.0f;
    if (q < n) {
        float h[MLP_W], acc[MLP_W];
        /* layer 1 over the encoding, feature by feature */
        const float* W = P + mlp_w_off(0);
        const float* b = P + mlp_b_off(0);
#pragma unroll
        for (int o = 0; o < MLP_W; ++o) acc[o] = b[o];
        for (int j = 0; j < MLP_RAW; ++j) {
            float e[24];
            encode_scalar(raw[q * MLP_RAW + j], e);
#pragma unroll
            for (int t = 0; t < 24; ++t) {
                const int f = 24 * j + t;
                X[(int64_t)f * n + q] = e[t];
#pragma unroll
                for (int o = 0; o < MLP_W; ++o) acc[o] += W[o * MLP_IN + f] * e[t];
            }
        }
#pragma unroll
        for (int o = 0; o < MLP_W; ++o) {
            h[o] = acc[o] > 0.0f ? acc[o] : 0.0f;
            H[(int64_t)o * n + q] = h[o];
        }
        for (int L = 1; L < 5; ++L) {
            W = P + mlp_w_off(L);
            b = P + mlp_b_off(L);
#pragma unroll
            for (int o = 0; o < MLP_W; ++o) acc[o] = b[o];
            for (int i = 0; i < MLP_W; ++i) {
                const float hi = h[i];
#pragma unroll
                for (int o = 0; o < MLP_W; ++o) acc[o] += W[o * MLP_W + i] * hi;
            }
#pragma unroll
            for (int o = 0; o < MLP_W; ++o) {
                h[o] = acc[o] > 0.0f ? acc[o] : 0.0f;
                H[((int64_t)L * MLP_W + o) * n + q] = h[o];
            }
        }
        W = P + mlp_w_off(5);
        b = P + mlp_b_off(5);
        float dy[MLP_OUT];
#pragma unroll
        for (int c = 0; c < MLP_OUT; ++c) {
            float y = b[c];
            for (int i = 0; i < MLP_W; ++i) y += W[c * MLP_W + i] * h[i];
            const float t = target[q * MLP_OUT + c];
            const float w = 1.0f / (y * y + 0.01f);
            const float d = y - t;
            lsum += d * d * w;
            dy[c] = (2.0f * d * w) / (float)(MLP_OUT * n);
            DY[(int64_t)c * n + q] = dy[c];
        }
        /* backward: dh = W6^T dy, then layer by layer */
        float dh[MLP_W];
#pragma unroll
        for (int i = 0; i < MLP_W; ++i) dh[i] = W[i] * dy[0] + W[MLP_W + i] * dy[1] + W[2 * MLP_W + i] * dy[2];
        for (int L = 4; L >= 0; --L) {
            float dz[MLP_W];
#pragma unroll
            for (int o = 0; o < MLP_W; ++o) {
                const float hv = H[((int64_t)L * MLP_W + o) * n + q];
                dz[o] = hv > 0.0f ? dh[o] : 0.0f;
                DZ[((int64_t)L * MLP_W + o) * n + q] = dz[o];
            }
            if (L > 0) {
                W = P + mlp_w_off(L);
#pragma unroll
                for (int i = 0; i < MLP_W; ++i) dh[i] = 0.0f;
                for (int o = 0; o < MLP_W; ++o) {
                    const float d = dz[o];
#pragma unroll
                    for (int i = 0; i < MLP_W; ++i) dh[i] += W[o * MLP_W + i] * d;
                }
            }
        }
    }
    /* wave-level loss reduction, one atomic per wave */
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
    if ((threadIdx.x & 63) == 0 && lsum != 0.0f) atomicAdd(loss_sum, lsum);
}

/* dW[o][i] += sum_q A[o][q] * B[i][q] and db[o] += sum_q A[o][q] over the
 * workgroup's K chunk; 64x64 output tiles, 4x4 per thread, K staged in LDS
 * 32 samples at a time. */
#define GRAD_K 32
__global__ __launch_bounds__(256) void mlp_grad_kernel(const float* __restrict__ A, const float* __restrict__ B,
                                                       int OUT, int IN, int64_t n, int64_t kchunk,
                                                       float* __restrict__ dW, float* __restrict__ db)
{
    __shared__ float As[GRAD_K][65];
    __shared__ float Bs[GRAD_K][65];
    const int o0 = blockIdx.y * 64, i0 = blockIdx.x * 64;
    const int64_t k_begin = (int64_t)blockIdx.z * kchunk;
    const int64_t k_end = k_begin + kchunk < n ? k_begin + kchunk : n;
    const int to = threadIdx.x >> 4, ti = threadIdx.x & 15;
    float c[4][4];
    float dbs[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int e = 0; e < 4; ++e) c[a][e] = 0.0f;
    for (int64_t k0 = k_begin; k0 < k_end; k0 += GRAD_K) {
        for (int e = threadIdx.x; e < GRAD_K * 64; e += 256) {
            const int kk = e % GRAD_K, rr = e / GRAD_K;
            const int64_t k = k0 + kk;
            As[kk][rr] = (k < k_end && o0 + rr < OUT) ? A[(int64_t)(o0 + rr) * n + k] : 0.0f;
            Bs[kk][rr] = (k < k_end && i0 + rr < IN) ? B[(int64_t)(i0 + rr) * n + k] : 0.0f;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < GRAD_K; ++kk) {
            float a[4], bb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = As[kk][4 * to + u];
                bb[u] = Bs[kk][4 * ti + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                dbs[u] += a[u];
#pragma unroll
                for (int v = 0; v < 4; ++v) c[u][v] += a[u] * bb[v];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int o = o0 + 4 * to + u;
        if (o >= OUT) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int i = i0 + 4 * ti + v;
            if (i < IN && c[u][v] != 0.0f) atomicAdd(&dW[(int64_t)o * IN + i], c[u][v]);
        }
        if (blockIdx.x == 0 && ti == 0 && dbs[u] != 0.0f) atomicAdd(&db[o], dbs[u]);
    }
}

/* Adam (Kingma & Ba 2015) with bias correction; clears the gradient. */
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ P, float* __restrict__ G, float* __restrict__ M,
                                                   float* __restrict__ V, int n, float lr, float b1, float b2,
                                                   float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float m = b1 * M[i] + (1.0f - b1) * g;
    const float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
    G[i] = 0.0f;
}

/* ---- launchers -------------------------------------------------------------------------- */
hipError_t launch_mlp_pack(const float* params, _Float16* pack, float* bias, hipStream_t s)
{
    mlp_pack_kernel<<<(MLP_PACK + 255) / 256, 256, 0, s>>>(params, pack, bias);
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
                                                                                    n, nullptr);
    return hipGetLastError();
}

hipError_t launch_mlp_infer_dev(const _Float16* pack, const float* bias, const float* raw, float* out,
                                const uint32_t* n_dev, int64_t n_max, int cus, hipStream_t s)
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
                                                                                    n_max, n_dev);
    return hipGetLastError();
}

hipError_t launch_mlp_fwd_bwd(const float* params, const float* raw, const float* target, int64_t n, float* X,
                              float* H, float* DZ, float* DY, float* loss_sum, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    mlp_fwd_bwd_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(params, raw, target, n, X, H, DZ, DY, loss_sum);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float* X, const float* H, const float* DZ, const float* DY, int64_t n, float* grad,
                           hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t kchunk = 2048;
    const unsigned kz = (unsigned)((n + kchunk - 1) / kchunk);
    for (int L = 0; L < 6; ++L) {
        const int in = L == 0 ? MLP_IN : MLP_W, out = L == 5 ? MLP_OUT : MLP_W;
        const float* A = L == 5 ? DY : DZ + (int64_t)L * MLP_W * n;
        const float* B = L == 0 ? X : H + (int64_t)(L - 1) * MLP_W * n;
        dim3 grid((unsigned)((in + 63) / 64), (unsigned)((out + 63) / 64), kz);
        mlp_grad_kernel<<<grid, 256, 0, s>>>(A, B, out, in, n, kchunk, grad + mlp_w_off(L), grad + mlp_b_off(L));
    }
    return hipGetLastError();
}

hipError_t launch_adam(float* params, float* grad, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       int step, hipStream_t s)
{
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    adam_kernel<<<(n + 255) / 256, 256, 0, s>>>(params, grad, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace vt

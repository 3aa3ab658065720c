/*
 * vt_cache.h - the learned radiance cache: a 288-64-64-64-64-64-3 ReLU MLP
 * over a 12-band sin/cos encoding of (position, direction, normal, albedo).
 *
 * Parameters (FP32 master copy, row-major W[out][in] then b[out], layer by
 * layer): W1 64x288, b1, W2..W5 64x64, b2..b5, W6 3x64, b6.
 *
 * Raw query (12 floats): position normalised to the volume bounds [0,1]^3,
 * outgoing direction (d+1)/2, normal (n+1)/2, albedo.  Encoding of scalar
 * v (feature index 24*j + 2*k + {0: sin, 1: cos}, k = 0..11):
 *     (s, c) = sincos(pi * v);  per band: emit s, c; then s, c = 2sc, 1 - 2s^2
 * (the double-angle recurrence, evaluated in FP32 without contraction).
 */
#pragma once

#include "vt_device.h"

namespace vt {

enum {
    MLP_IN = 288, MLP_W = 64, MLP_OUT = 3, MLP_RAW = 12, MLP_BANDS = 12,
    MLP_HIDDEN = 5,                                   /* layers with 64 outputs */
    MLP_N_PARAMS = MLP_W * MLP_IN + MLP_W + 4 * (MLP_W * MLP_W + MLP_W) + MLP_OUT * MLP_W + MLP_OUT,
    /* FP16 weight pack for the MFMA inference (A operands, pre-swizzled) */
    MLP_PACK_L1 = 4 * 9 * 64 * 8, MLP_PACK_HID = 4 * 2 * 64 * 8, MLP_PACK_OUT = 1 * 2 * 64 * 8,
    MLP_PACK = MLP_PACK_L1 + 4 * MLP_PACK_HID + MLP_PACK_OUT,
    MLP_BIAS = 5 * MLP_W + 16,
    /* FP32 transposed weights for the backward (A operands of
     * v_mfma_f32_16x16x4_f32): 4 hidden layers of 4096, output layer 256 */
    MLP_TPACK_HID = 4 * 4 * 64 * 4, MLP_TPACK_OUT = 4 * 64,
    MLP_TPACK = 4 * MLP_TPACK_HID + MLP_TPACK_OUT,
    /* training: column blocks of the weight-gradient GEMM, K splits */
    MLP_GRAD_ITEMS = 15, MLP_GRAD_SPLITS = 64,
};

/* Offsets of each layer's weights and biases in the parameter vector. */
__host__ __device__ inline int mlp_w_off(int layer)   /* layer 0..5 */
{
    return layer == 0 ? 0 : MLP_W * MLP_IN + MLP_W + (layer - 1) * (MLP_W * MLP_W + MLP_W);
}
__host__ __device__ inline int mlp_b_off(int layer)
{
    const int in = layer == 0 ? MLP_IN : MLP_W, out = layer == 5 ? MLP_OUT : MLP_W;
    return mlp_w_off(layer) + out * in;
}

/* FP16 pack entry e -> parameter index (or -1 for padding).  Layout per
 * layer: [m (16-row output tile)][s (32-wide K step)][lane][j (8 halves)].
 * Lane l supplies row o = 16m + l%16; its K slots j are input columns
 *   layer 1: 72*(l/16) + 8s + j (lane group g encodes raw scalars 3g..3g+2),
 *   others : 32s + 16*(j/4) + 4*(l/16) + j%4,
 * the order in which a 16x16 D tile (lane l: rows 4*(l/16)+i of column
 * l%16) becomes the next layer's B operand without leaving the lane. */
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
    const int i = layer == 0 ? 72 * g + 8 * s + j : 32 * s + 16 * (j >> 2) + 4 * g + (j & 3);
    return mlp_w_off(layer) + o * in + i;
}

/* Inverse of pack_source for weight (layer, o, i). */
__host__ __device__ inline int pack_dest(int layer, int o, int i)
{
    const int m = o >> 4;
    if (layer == 0) {
        const int g = i / 72, rem = i % 72, s = rem >> 3, j = rem & 7;
        return ((m * 9 + s) * 64 + 16 * g + (o & 15)) * 8 + j;
    }
    const int base = MLP_PACK_L1 + (layer - 1) * MLP_PACK_HID;
    const int s = i >> 5, w = i & 31, g = (w & 15) >> 2, j = 4 * (w >> 4) + (w & 3);
    return base + ((m * 2 + s) * 64 + 16 * g + (o & 15)) * 8 + j;
}

/* FP32 transposed pack entry e -> parameter index (or -1).  Hidden layer L
 * (1..4), entries [(L-1)*4096, L*4096): [mt][tq][lane][u]; at K-step
 * t = 4tq + u lane l supplies A[row 16mt + l%16][k = l/16] = W_L[o][i] with
 * o = 16tq + 4*(l/16) + u, i = 16mt + l%16 (the order of the D tiles).
 * Output layer, 256 entries [mt][lane]: W_6[l/16][16mt + l%16] for l/16 < 3. */
__host__ __device__ inline int tpack_source(int e)
{
    if (e < 4 * MLP_TPACK_HID) {
        const int layer = 1 + e / MLP_TPACK_HID, local = e % MLP_TPACK_HID;
        const int u = local & 3, lane = (local >> 2) & 63, tq = (local >> 8) & 3, mt = local >> 10;
        const int o = 16 * tq + 4 * (lane >> 4) + u, i = 16 * mt + (lane & 15);
        return mlp_w_off(layer) + o * MLP_W + i;
    }
    const int local = e - 4 * MLP_TPACK_HID, lane = local & 63, mt = local >> 6, c = lane >> 4;
    if (c >= MLP_OUT) return -1;
    return mlp_w_off(5) + c * MLP_W + 16 * mt + (lane & 15);
}

/* Inverse of tpack_source for weight (layer >= 1, o, i). */
__host__ __device__ inline int tpack_dest(int layer, int o, int i)
{
    if (layer == 5) return 4 * MLP_TPACK_HID + (i >> 4) * 64 + 16 * o + (i & 15);
    const int mt = i >> 4, tq = o >> 4, g = (o & 15) >> 2, u = o & 3;
    return (layer - 1) * MLP_TPACK_HID + ((mt * 4 + tq) * 64 + 16 * g + (i & 15)) * 4 + u;
}

/* Encoding of 3 raw scalars (72 features, in feature order) - used by one
 * lane group of the MFMA inference and by the training kernel. */
__device__ __forceinline__ void encode_scalar(float v, float* out24)
{
    float s, c;
    vt_sincosf(VT_PI_F * v, &s, &c);
#pragma unroll
    for (int k = 0; k < MLP_BANDS; ++k) {
        out24[2 * k] = s;
        out24[2 * k + 1] = c;
        const float s2 = (2.0f * s) * c;
        const float c2 = 1.0f - (2.0f * s) * s;
        s = s2;
        c = c2;
    }
}

/* A training batch: samples [off, off + n) of raw [.][12] / target [.][3],
 * n = min(min(*n_dev, cap) - off, batch) (0 when off is past the end); the
 * count is read on the device, so batches can follow a render that
 * produced the samples without a host round trip. */
struct TrainBatch {
    const float* raw;
    const float* target;
    const uint32_t* n_dev;
    uint32_t cap, off, batch;
};

/* Training scratch for batches of at most 16 * tiles samples, stored per
 * 16-sample tile, feature-major inside a tile ([tile][feature][16]): the
 * forward/backward kernel writes with constant offsets from one address per
 * lane, the gradient GEMM reads 4 consecutive samples per lane. */
struct TrainScratch {
    _Float16* X;        /* [tiles][288][16] encoded inputs (FP16, as the MFMA sees them) */
    _Float16* H;        /* [tiles][5*64][16] hidden activations after ReLU (FP16) */
    float* DZ;          /* [tiles][5*64][16] dL/d(pre-activation) of the hidden layers */
    float* DY;          /* [tiles][3][16] dL/dy */
    float* part;        /* [MLP_GRAD_SPLITS][MLP_N_PARAMS] partial gradients */
    uint32_t tiles;
    float* loss;        /* += sum of the batch's per-channel losses */
    uint32_t* steps;    /* += 1 per non-empty batch (Adam's step count) */
};

/* Launchers (all on stream s). */
/* params -> FP16 inference pack, bias table, FP32 transposed pack */
hipError_t launch_mlp_pack(const float* params, _Float16* pack, float* bias, float* tpack, hipStream_t s);
hipError_t launch_mlp_infer(const _Float16* pack, const float* bias, const float* raw, float* out, int64_t n,
                            int cus, hipStream_t s);
/* The same for queries [*lo_dev, *n_dev) (read on the device; *n_dev at most n_max). */
hipError_t launch_mlp_infer_dev(const _Float16* pack, const float* bias, const float* raw, float* out,
                                const uint32_t* lo_dev, const uint32_t* n_dev, int64_t n_max, int cus, hipStream_t s);
/* Training step of one batch: forward (FP16 MFMA, as inference) and
 * backward (FP32 MFMA) into the scratch, the weight-gradient GEMM into
 * K-split partials, then either the summed gradient into grad (grad != 0)
 * or the Adam update of params / m / v that also rewrites the packs. */
/* tally[0] += counts[0] (queries), tally[1] += min(counts[1], scap) (samples) */
hipError_t launch_cache_tally(const uint32_t* counts, uint32_t scap, unsigned long long* tally, hipStream_t s);
hipError_t launch_mlp_train_fb(const _Float16* pack, const float* bias, const float* tpack, const TrainBatch& b,
                               const TrainScratch& S, int cus, hipStream_t s);
hipError_t launch_mlp_train_grad(const TrainBatch& b, const TrainScratch& S, hipStream_t s);
hipError_t launch_mlp_train_update(const TrainBatch& b, const TrainScratch& S, float* grad, float* params, float* m,
                                   float* v, _Float16* pack, float* bias, float* tpack, float lr, float b1, float b2,
                                   float eps, hipStream_t s);

}  // namespace vt

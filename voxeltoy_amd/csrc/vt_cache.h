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

struct CacheArgs;   /* defined in vt_cache.hip */

/* Launchers (all on stream s). */
hipError_t launch_mlp_pack(const float* params, _Float16* pack, float* bias, hipStream_t s);
hipError_t launch_mlp_infer(const _Float16* pack, const float* bias, const float* raw, float* out, int64_t n,
                            int cus, hipStream_t s);
/* The same with the query count read on the device (*n_dev, at most n_max). */
hipError_t launch_mlp_infer_dev(const _Float16* pack, const float* bias, const float* raw, float* out,
                                const uint32_t* n_dev, int64_t n_max, int cus, hipStream_t s);
/* Forward + backward of the relative-L2 loss for n samples: writes the
 * encoded inputs X (288 x n), activations H (5 x 64 x n), deltas DZ
 * (5 x 64 x n), DY (3 x n) (all feature-major) and adds the loss sum. */
hipError_t launch_mlp_fwd_bwd(const float* params, const float* raw, const float* target, int64_t n, float* X,
                              float* H, float* DZ, float* DY, float* loss_sum, hipStream_t s);
/* grad (zeroed by the caller) += dL/dparams from the buffers above. */
hipError_t launch_mlp_grad(const float* X, const float* H, const float* DZ, const float* DY, int64_t n, float* grad,
                           hipStream_t s);
hipError_t launch_adam(float* params, float* grad, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       int step, hipStream_t s);

}  // namespace vt

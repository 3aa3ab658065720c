/*
 * vt_host_algos.cpp - host-side preparation steps (see vt_host_algos.h).
 */
#include "vt_host_algos.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace vt {
namespace host {

namespace {

struct V3 {
    float x, y, z;
};

/* Imath Vec3::normalized(): divide each component by the length. */
V3 imath_normalized(V3 v)
{
    const float l = std::sqrt((v.x * v.x + v.y * v.y) + v.z * v.z);
    if (l == 0.0f) return V3{ 0.0f, 0.0f, 0.0f };
    return V3{ v.x / l, v.y / l, v.z / l };
}

V3 imath_cross(V3 a, V3 b) { return V3{ a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x }; }
float imath_dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

double lanczos3(double x)
{
    x = std::fabs(x);
    if (x >= 3.0) return 0.0;
    if (x < 1e-9) return 1.0;
    const double px = M_PI * x;
    return 3.0 * std::sin(px) * std::sin(px / 3.0) / (px * px);
}

/* Resample every row of a w x h single-channel image to dw columns. */
void resample_rows(const std::vector<float>& src, int sw, int h, std::vector<float>& dst, int dw)
{
    dst.assign((size_t)dw * h, 0.0f);
    const double ratio = (double)sw / (double)dw;
    const double scale = std::max(ratio, 1.0);
    const double rad = 3.0 * scale;
    for (int x = 0; x < dw; ++x) {
        const double center = ((double)x + 0.5) * ratio;
        const int i0 = (int)std::floor(center - rad), i1 = (int)std::ceil(center + rad);
        double wsum = 0.0;
        for (int i = i0; i <= i1; ++i) wsum += lanczos3(((double)i + 0.5 - center) / scale);
        for (int y = 0; y < h; ++y) {
            double acc = 0.0;
            for (int i = i0; i <= i1; ++i) {
                const int ci = std::min(std::max(i, 0), sw - 1);
                acc += lanczos3(((double)i + 0.5 - center) / scale) * (double)src[(size_t)y * sw + ci];
            }
            dst[(size_t)y * dw + x] = (float)(acc / wsum);
        }
    }
}

void transpose(const std::vector<float>& src, int w, int h, std::vector<float>& dst)
{
    dst.assign((size_t)w * h, 0.0f);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) dst[(size_t)x * h + y] = src[(size_t)y * w + x];
}

}  // namespace

void inverse4(const float* m, float* out)
{
    double a[16], inv[16];
    for (int i = 0; i < 16; ++i) a[i] = (double)m[i];
    inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] + a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
    inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] - a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
    inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] + a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
    inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] - a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
    inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] - a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
    inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] + a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
    inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] - a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
    inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] + a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
    inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] + a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
    inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] - a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
    inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] + a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
    inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] - a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
    inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] - a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
    inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] + a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
    inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] - a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
    inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] + a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
    const double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
    for (int i = 0; i < 16; ++i) out[i] = (det == 0.0) ? 0.0f : (float)(inv[i] / det);
}

void camera_matrices(const float eye[3], const float target[3], float fov_y, float near_d, float far_d,
                     int lens_model, int width, int height, float mv[16], float inv_mv[16], float proj[16],
                     float inv_proj[16])
{
    const V3 e{ eye[0], eye[1], eye[2] };
    const V3 fwd = imath_normalized(V3{ target[0] - eye[0], target[1] - eye[1], target[2] - eye[2] });
    const V3 right = imath_normalized(imath_cross(V3{ 0.0f, 1.0f, 0.0f }, fwd));
    const V3 up = imath_cross(fwd, right);
    const float m[16] = { right.x, right.y, right.z, -imath_dot(e, right),
                          up.x, up.y, up.z, -imath_dot(e, up),
                          -fwd.x, -fwd.y, -fwd.z, imath_dot(e, fwd),
                          0.0f, 0.0f, 0.0f, 1.0f };
    std::memcpy(mv, m, sizeof(m));
    float p[16] = { 0.0f };
    const float a = (float)width / (float)height;
    if (lens_model == 2) {
        const V3 d{ target[0] - eye[0], target[1] - eye[1], target[2] - eye[2] };
        const float dist = std::sqrt(imath_dot(d, d));
        const float left = -std::tan(fov_y / 2.0f) * dist;
        const float rgt = -left;
        const float bottom = -std::tan(fov_y / 2.0f) * dist / a;
        const float top = -bottom;
        const float n = -near_d, f = -far_d;
        p[0] = 2.0f / (rgt - left);
        p[3] = -(rgt + left) / (rgt - left);
        p[5] = 2.0f / (top - bottom);
        p[7] = -(top + bottom) / (top - bottom);
        p[10] = -2.0f / (f - n);
        p[11] = -(f + n) / (f - n);
        p[15] = 1.0f;
    } else {
        const float n = near_d, f = far_d;
        const float ef = 1.0f / std::tan(fov_y / 2.0f);
        p[0] = ef / a;
        p[5] = ef;
        p[10] = (f + n) / (n - f);
        p[11] = ((2.0f * f) * n) / (n - f);
        p[14] = -1.0f;
    }
    std::memcpy(proj, p, sizeof(p));
    inverse4(mv, inv_mv);
    inverse4(p, inv_proj);
}

void calculate_cdf(const float* fimg, int w, int h, float* cdfu, float* cdfv, float* integral)
{
    const unsigned W = (unsigned)w, H = (unsigned)h, cw = W + 1;
    std::vector<float> fu((size_t)W * H), fv(H + 1);
    const float iW = (float)W, iH = (float)H, iA = iW * iH;
    float sum = 0.0f;
    for (unsigned y = 0; y < H; ++y) {
        const float sinTheta = (float)std::sin(M_PI * (double)((float)y + 0.5f) / (double)iH);
        for (unsigned x = 0; x < W; ++x) {
            const float v = fimg[(size_t)y * W + x];
            const float value = (0.0f < v) ? v : 0.0f;
            fu[(size_t)y * W + x] = value * sinTheta;
            sum += value * sinTheta;
        }
    }
    float integ = sum / iA;
    integ = (float)((double)integ * (2.0 * M_PI * M_PI));
    *integral = integ;
    const float stepsW = (float)W;
    for (unsigned y = 0; y < H; ++y) {
        float* row = cdfu + (size_t)y * cw;
        row[0] = 0.0f;
        for (unsigned x = 1; x <= W; ++x) row[x] = row[x - 1] + fu[(size_t)y * W + x - 1] / stepsW;
        const float rowInt = row[W];
        fv[y] = rowInt;
        if (rowInt > 0.0f) {
            for (unsigned x = 1; x <= W; ++x) row[x] /= rowInt;
        } else {
            for (unsigned x = 1; x <= W; ++x) row[x] = (float)x / stepsW;
        }
    }
    cdfv[0] = 0.0f;
    const float stepsH = (float)H;
    for (unsigned y = 1; y <= H; ++y) cdfv[y] = cdfv[y - 1] + fv[y - 1] / stepsH;
    const float imgInt = cdfv[H];
    if (imgInt > 0.0f) {
        for (unsigned y = 1; y <= H; ++y) cdfv[y] /= imgInt;
    } else {
        for (unsigned y = 1; y <= H; ++y) cdfv[y] = (float)y / iH;
    }
}

void image_function(const float* rgb, int w, int h, int max_size, std::vector<float>& out, int& ow, int& oh)
{
    /* the RGB image is resized first, channel by channel, and the luminance
     * taken from the resized channels (image.cpp:309-335 orders the steps
     * so) */
    int nw = w, nh = h;
    std::vector<float> chan[3];
    for (int c = 0; c < 3; ++c) {
        chan[c].resize((size_t)w * h);
        for (size_t i = 0; i < chan[c].size(); ++i) chan[c][i] = rgb[3 * i + c];
    }
    if (std::max(w, h) > max_size) {
        const int mxd = std::max(w, h);
        nw = (int)((float)w / (float)mxd * (float)max_size);
        nh = (int)((float)h / (float)mxd * (float)max_size);
        for (int c = 0; c < 3; ++c) {
            std::vector<float> t1, t2, t3;
            resample_rows(chan[c], w, h, t1, nw);
            transpose(t1, nw, h, t2);
            resample_rows(t2, h, nw, t3, nh);
            transpose(t3, nh, nw, chan[c]);
        }
    }
    std::vector<float> lum((size_t)nw * nh);
    for (size_t i = 0; i < lum.size(); ++i) lum[i] = (chan[0][i] * 0.2126f + chan[1][i] * 0.7152f) + chan[2][i] * 0.0722f;
    const float g1 = std::exp(-2.0f * ((2.0f / 3.0f) * (2.0f / 3.0f)));
    const float k1[3] = { g1, 1.0f, g1 };
    float ksum = 0.0f;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) ksum += k1[i] * k1[j];
    out.assign((size_t)nw * nh, 0.0f);
    for (int y = 0; y < nh; ++y) {
        for (int x = 0; x < nw; ++x) {
            float acc = 0.0f;
            for (int j = -1; j <= 1; ++j) {
                for (int i = -1; i <= 1; ++i) {
                    const int xx = x + i, yy = y + j;
                    const float v = (xx < 0 || yy < 0 || xx >= nw || yy >= nh) ? 0.0f : lum[(size_t)yy * nw + xx];
                    acc += ((k1[i + 1] * k1[j + 1]) / ksum) * v;
                }
            }
            out[(size_t)y * nw + x] = acc;
        }
    }
    ow = nw;
    oh = nh;
}

int32_t prune_emissive(const int32_t* offsets, const int32_t res[3], int32_t* list, int32_t n)
{
    static const int nb[6][3] = { { 1, 0, 0 }, { -1, 0, 0 }, { 0, 1, 0 }, { 0, -1, 0 }, { 0, 0, 1 }, { 0, 0, -1 } };
    const int64_t sx = res[0], sxy = (int64_t)res[0] * res[1];
    for (int32_t i = 0; i < n; ++i) {
        const int32_t v = list[i];
        const int32_t z = v / (res[0] * res[1]);
        const int32_t rem = v - z * res[0] * res[1];
        const int32_t y = rem / res[0];
        const int32_t x = rem - y * res[0];
        bool visible = false;
        for (int k = 0; k < 6 && !visible; ++k) {
            const int32_t a = x + nb[k][0], b = y + nb[k][1], c = z + nb[k][2];
            if (a < 0 || a >= res[0] || b < 0 || b >= res[1] || c < 0 || c >= res[2]) {
                visible = true;
            } else if (offsets[a + b * sx + c * sxy] < 0) {
                visible = true;
            }
        }
        if (!visible) {
            list[i] = list[n - 1];
            --n;
            --i;
        }
    }
    return n;
}

void mesh_transform(const float bmin[3], const float bmax[3], const int32_t res[3], float M[16])
{
    const float size[3] = { bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2] };
    const float margin[3] = { 1.0f / (float)res[0], 1.0f / (float)res[1], 1.0f / (float)res[2] };
    int major = 0;
    if (size[1] > size[major]) major = 1;
    if (size[2] > size[major]) major = 2;
    const float s = (float)((1.0 - 2.0 * (double)margin[major]) / (double)size[major]);
    std::memset(M, 0, 16 * sizeof(float));
    M[0] = M[5] = M[10] = s;
    M[15] = 1.0f;
    for (int k = 0; k < 3; ++k) M[4 * k + 3] = (-bmin[k] + margin[k] / s) * s;
}

}  // namespace host
}  // namespace vt

// Reads environment images in the formats voxeltoy::readImage takes besides
// PFM, from files this program writes with known contents:
//   - Radiance RGBE, run-length encoded scanlines (one run and literals per
//     channel) and flat scanlines, "-Y h +X w" (top row first) and "+Y"
//     (bottom row first): (r, g, b) * 2^(e - 136), 0 when e = 0;
//   - binary PPM 8-bit and PGM 16-bit: value / maxval;
// and checks unsupported or truncated files are refused.  Prints OK.
#include "scene_io.h"

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

using namespace voxeltoy;

static int fails = 0;
#define CHECK(c)                                                            \
    do {                                                                    \
        if (!(c)) {                                                         \
            std::fprintf(stderr, "%s:%d: %s\n", __FILE__, __LINE__, #c);    \
            ++fails;                                                        \
        }                                                                   \
    } while (0)

static void put(const std::string& path, const std::vector<unsigned char>& bytes)
{
    FILE* f = std::fopen(path.c_str(), "wb");
    std::fwrite(bytes.data(), 1, bytes.size(), f);
    std::fclose(f);
}

static std::vector<unsigned char> text(const std::string& s) { return std::vector<unsigned char>(s.begin(), s.end()); }

static float rgbe(int v, int e) { return e ? (float)v * std::ldexp(1.0f, e - 136) : 0.0f; }

int main(int argc, char** argv)
{
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int W = 10, H = 2;
    // pixel (x, y) of the image, top row y = 0: RGBE (r, g, b, e)
    auto px = [](int x, int y, int c) {
        const int r = 16 * x + 3, g = 200 - 7 * x, b = y ? 128 : 5 + x, e = y ? 130 : (x == 4 ? 0 : 128 + x % 3);
        const int v[4] = { r, g, b, e };
        return v[c];
    };
    // run-length encoded, top row first
    {
        auto bytes = text("#?RADIANCE\n# written by image_io.cpp\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n-Y 2 +X 10\n");
        for (int y = 0; y < H; ++y) {
            bytes.insert(bytes.end(), { 2, 2, 0, (unsigned char)W });
            for (int c = 0; c < 4; ++c) {
                const bool constant = (c == 3 && y == 1) || (c == 2 && y == 1);
                if (constant) {             // one run of W equal bytes
                    bytes.push_back((unsigned char)(128 + W));
                    bytes.push_back((unsigned char)px(0, y, c));
                } else {                    // W literal bytes
                    bytes.push_back((unsigned char)W);
                    for (int x = 0; x < W; ++x) bytes.push_back((unsigned char)px(x, y, c));
                }
            }
        }
        put(dir + "/rle.hdr", bytes);
        std::vector<float> img;
        int w = 0, h = 0;
        std::string err;
        CHECK(readImage(dir + "/rle.hdr", img, w, h, &err));
        CHECK(w == W && h == H && img.size() == (size_t)W * H * 3);
        for (int y = 0; y < H && w == W; ++y)
            for (int x = 0; x < W; ++x)
                for (int c = 0; c < 3; ++c) CHECK(img[((size_t)y * W + x) * 3 + c] == rgbe(px(x, y, c), px(x, y, 3)));
    }
    // flat scanlines, bottom row first (+Y): the file's first row is the image's last
    {
        auto bytes = text("#?RGBE\n\n+Y 2 +X 10\n");
        for (int fy = 0; fy < H; ++fy)
            for (int x = 0; x < W; ++x)
                for (int c = 0; c < 4; ++c) bytes.push_back((unsigned char)px(x, H - 1 - fy, c));
        put(dir + "/flat.hdr", bytes);
        std::vector<float> img;
        int w = 0, h = 0;
        CHECK(readImage(dir + "/flat.hdr", img, w, h));
        for (int y = 0; y < H && w == W; ++y)
            for (int x = 0; x < W; ++x)
                for (int c = 0; c < 3; ++c) CHECK(img[((size_t)y * W + x) * 3 + c] == rgbe(px(x, y, c), px(x, y, 3)));
    }
    // 8-bit PPM and 16-bit PGM
    {
        auto bytes = text("P6\n# comment\n3 2\n255\n");
        for (int i = 0; i < 18; ++i) bytes.push_back((unsigned char)(i * 13));
        put(dir + "/rgb.ppm", bytes);
        std::vector<float> img;
        int w = 0, h = 0;
        CHECK(readImage(dir + "/rgb.ppm", img, w, h));
        CHECK(w == 3 && h == 2);
        for (int i = 0; i < 18 && img.size() == 18; ++i) CHECK(img[i] == (float)(i * 13) / 255.0f);
        auto g = text("P5 2 1 65535\n");
        g.insert(g.end(), { 0x12, 0x34, 0xff, 0xff });
        put(dir + "/grey.pgm", g);
        CHECK(readImage(dir + "/grey.pgm", img, w, h));
        CHECK(w == 2 && h == 1 && img.size() == 6);
        if (img.size() == 6) {
            for (int c = 0; c < 3; ++c) CHECK(img[c] == (float)0x1234 / 65535.0f && img[3 + c] == 1.0f);
        }
    }
    // refused: other extensions, truncated data, unknown pixel formats
    {
        std::vector<float> img;
        int w = 0, h = 0;
        std::string err;
        CHECK(!readImage(dir + "/x.exr", img, w, h, &err) && err.find("unsupported") != std::string::npos);
        put(dir + "/short.hdr", text("#?RADIANCE\n\n-Y 4 +X 10\n\x02\x02"));
        CHECK(!readImage(dir + "/short.hdr", img, w, h));
        put(dir + "/xyz.hdr", text("#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n\x01\x01\x01\x80"));
        CHECK(!readImage(dir + "/xyz.hdr", img, w, h));
        put(dir + "/short.ppm", text("P6 4 4 255\n\x01\x02"));
        CHECK(!readImage(dir + "/short.ppm", img, w, h));
        // a header asking for 65535 x 65535 pixels of a few-byte file: refused
        // before any pixel buffer is sized from it
        put(dir + "/huge.hdr", text("#?RADIANCE\n\n-Y 65535 +X 65535\n\x02\x02\xff\xff"));
        err.clear();
        CHECK(!readImage(dir + "/huge.hdr", img, w, h, &err) && err.find("truncated") != std::string::npos);
    }
    if (fails) return 1;
    std::printf("OK\n");
    return 0;
}

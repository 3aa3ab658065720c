/*
 * scene_io.cpp - material records, .vox and OBJ loading, image output.
 * See scene_io.h for the file formats and the reference files each part
 * follows.
 */
#include "scene_io.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace voxeltoy {

/* ---- material records (material.cpp:1-99) --------------------------------------- */
namespace Material {

static SerializedData serialize_float(const char* name, size_t offset, float value, float lo, float hi)
{
    SerializedData r;
    r.m_propertyName = name;
    r.m_dataOffset = offset;
    r.m_propertyType = SerializedData::PROPERTY_TYPE_FLOAT;
    r.m_dataRangeFrom = lo;
    r.m_dataRangeTo = hi;
    r.m_value = value;
    return r;
}

static SerializedData serialize_color(const char* name, size_t offset, const float rgb[3])
{
    SerializedData r;
    r.m_propertyName = name;
    r.m_dataOffset = offset;
    r.m_propertyType = SerializedData::PROPERTY_TYPE_COLOR;
    r.m_childProperties.push_back(serialize_float("red", offset, rgb[0], 0, 1));
    r.m_childProperties.push_back(serialize_float("green", offset + 1, rgb[1], 0, 1));
    r.m_childProperties.push_back(serialize_float("blue", offset + 2, rgb[2], 0, 1));
    return r;
}

static SerializedData group(const char* name, size_t offset)
{
    SerializedData d;
    d.m_propertyName = name;
    d.m_dataOffset = offset;
    return d;
}

SerializedData serializeLambert(const LambertMaterialData& m, size_t offset)
{
    SerializedData d = group("Lambert Material", offset);
    d.m_childProperties.push_back(serialize_color("emission", offset + 1, m.emission));
    d.m_childProperties.push_back(serialize_color("albedo", offset + 4, m.albedo));
    return d;
}

SerializedData serializeMetal(const MetalMaterialData& m, size_t offset)
{
    SerializedData d = group("Metal Material", offset);
    d.m_childProperties.push_back(serialize_color("emission", offset + 1, m.emission));
    d.m_childProperties.push_back(serialize_color("reflectance", offset + 4, m.reflectance));
    d.m_childProperties.push_back(serialize_float("roughness", offset + 7, m.roughness, 0, 1000));
    return d;
}

SerializedData serializePlastic(const PlasticMaterialData& m, size_t offset)
{
    SerializedData d = group("Plastic Material", offset);
    d.m_childProperties.push_back(serialize_color("emission", offset + 1, m.emission));
    d.m_childProperties.push_back(serialize_color("diffuseAlbedo", offset + 4, m.diffuseAlbedo));
    d.m_childProperties.push_back(serialize_float("roughness", offset + 7, m.roughness, 0, 1000));
    return d;
}

size_t recordSize(float type)
{
    switch ((int)type) {
        case MT_LAMBERT: return 1 + sizeof(LambertMaterialData) / sizeof(float);
        case MT_METAL: return 1 + sizeof(MetalMaterialData) / sizeof(float);
        case MT_PLASTIC: return 1 + sizeof(PlasticMaterialData) / sizeof(float);
        default: return 0;
    }
}

/* Renderer::getMaterials (renderer.cpp:1142-1203): records are read back to
 * back; an unknown type ends the walk. */
std::vector<SerializedData> serializeAll(const std::vector<float>& data, std::string* error)
{
    std::vector<SerializedData> out;
    size_t off = 0;
    while (off < data.size()) {
        const float t = data[off];
        const size_t n = recordSize(t);
        const bool known = (t == (float)MT_LAMBERT || t == (float)MT_METAL || t == (float)MT_PLASTIC);
        if (!known || off + n > data.size()) {
            if (error) *error = "Unrecognized material type. Aborting.";
            break;
        }
        const float* p = &data[off + 1];
        if (t == (float)MT_LAMBERT) {
            LambertMaterialData m;
            std::memcpy(&m, p, sizeof(m));
            out.push_back(serializeLambert(m, off));
        } else if (t == (float)MT_METAL) {
            MetalMaterialData m;
            std::memcpy(&m, p, sizeof(m));
            out.push_back(serializeMetal(m, off));
        } else {
            PlasticMaterialData m;
            std::memcpy(&m, p, sizeof(m));
            out.push_back(serializePlastic(m, off));
        }
        off += n;
    }
    return out;
}

}  // namespace Material

/* ---- loader base (voxLoader.cpp:8-91) ------------------------------------------ */
void VoxLoader::storeMaterialData(Material::MaterialType type, const float* data, size_t n,
                                  std::vector<float>& storage)
{
    const size_t off = storage.size();
    storage.resize(off + 1 + n);
    storage[off] = (float)type;
    std::memcpy(&storage[off + 1], data, n * sizeof(float));
}

void VoxLoader::generateMaterialLambert(V3f e, V3f a, std::vector<float>& out)
{
    Material::LambertMaterialData d = { { e.x, e.y, e.z }, { a.x, a.y, a.z } };
    storeMaterialData(Material::MT_LAMBERT, reinterpret_cast<const float*>(&d), sizeof(d) / sizeof(float), out);
}

void VoxLoader::generateMaterialMetal(V3f e, V3f r, float rough, std::vector<float>& out)
{
    Material::MetalMaterialData d = { { e.x, e.y, e.z }, { r.x, r.y, r.z }, rough };
    storeMaterialData(Material::MT_METAL, reinterpret_cast<const float*>(&d), sizeof(d) / sizeof(float), out);
}

void VoxLoader::generateMaterialPlastic(V3f e, V3f a, float rough, std::vector<float>& out)
{
    Material::PlasticMaterialData d = { { e.x, e.y, e.z }, { a.x, a.y, a.z }, rough };
    storeMaterialData(Material::MT_PLASTIC, reinterpret_cast<const float*>(&d), sizeof(d) / sizeof(float), out);
}

float VoxLoader::materialEmissiveness(const float* record)
{
    const float t = record[0];
    if (t != (float)Material::MT_LAMBERT && t != (float)Material::MT_METAL && t != (float)Material::MT_PLASTIC)
        return 0.0f;
    return (record[1] + record[2] + record[3]) / 3;
}

/* ---- MagicaVoxel (magicaVoxel.cpp:11-321) ---------------------------------------- */
/* MagicaVoxel's default palette: index 0 unused, 1..215 the 6x6x6 cube over
 * channel levels ff, cc, 99, 66, 33, 00 (blue fastest, red slowest), then
 * ten-step ramps (ee .. 11) of red, green, blue and grey. */
void MagicaVoxelLoader::defaultPalette(uint8_t rgba[256 * 4])
{
    static const uint8_t L[6] = { 0xff, 0xcc, 0x99, 0x66, 0x33, 0x00 };
    static const uint8_t R[10] = { 0xee, 0xdd, 0xbb, 0xaa, 0x88, 0x77, 0x55, 0x44, 0x22, 0x11 };
    std::memset(rgba, 0, 4);
    for (int i = 1; i < 256; ++i) {
        uint8_t r, g, b;
        if (i <= 215) {
            const int k = i - 1;
            r = L[k / 36]; g = L[(k / 6) % 6]; b = L[k % 6];
        } else {
            const int j = i - 216, v = R[j % 10];
            const int ramp = j / 10;
            r = (ramp == 0 || ramp == 3) ? v : 0;
            g = (ramp == 1 || ramp == 3) ? v : 0;
            b = (ramp == 2 || ramp == 3) ? v : 0;
        }
        rgba[4 * i + 0] = r; rgba[4 * i + 1] = g; rgba[4 * i + 2] = b; rgba[4 * i + 3] = 0xff;
    }
}

namespace {

struct VoxReader {
    const std::vector<uint8_t>& buf;
    size_t pos = 0;
    bool ok = true;
    explicit VoxReader(const std::vector<uint8_t>& b) : buf(b) {}
    int32_t i32()
    {
        if (pos + 4 > buf.size()) { ok = false; pos = buf.size(); return 0; }
        int32_t v;
        std::memcpy(&v, &buf[pos], 4);
        pos += 4;
        return v;
    }
};

int32_t fourcc(const char* s) { return (int32_t)((uint32_t)(uint8_t)s[0] | ((uint32_t)(uint8_t)s[1] << 8) |
                                                 ((uint32_t)(uint8_t)s[2] << 16) | ((uint32_t)(uint8_t)s[3] << 24)); }

bool read_file(const std::string& path, std::vector<uint8_t>& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

}  // namespace

bool MagicaVoxelLoader::load(const std::string& filePath, std::vector<int32_t>& voxelMaterials,
                             std::vector<float>& materialData, std::vector<int32_t>& emissiveVoxelIndices,
                             V3i& voxelResolution)
{
    m_error.clear();
    std::vector<uint8_t> buf;
    if (!read_file(filePath, buf)) { m_error = "failed to open file"; return false; }
    VoxReader rd(buf);
    if (rd.i32() != fourcc("VOX ")) { m_error = "magic number does not match"; return false; }
    if (rd.i32() != 150) { m_error = "version does not match"; return false; }
    const int32_t main_id = rd.i32();
    const int32_t main_content = rd.i32();
    const int32_t main_children = rd.i32();
    if (!rd.ok || main_id != fourcc("MAIN")) { m_error = "main chunk is not found"; return false; }
    const size_t main_end = rd.pos + (size_t)std::max(0, main_content) + (size_t)std::max(0, main_children);
    rd.pos += (size_t)std::max(0, main_content);

    int32_t sx = 0, sy = 0, sz = 0;
    std::vector<uint8_t> voxels;   /* x, y, z, colour index per voxel */
    uint8_t palette[256 * 4];
    bool custom = false;
    std::memset(palette, 0, sizeof(palette));
    while (rd.ok && rd.pos < main_end && rd.pos + 12 <= buf.size()) {
        const int32_t id = rd.i32(), content = rd.i32(), children = rd.i32();
        if (content < 0 || children < 0) { m_error = "corrupt chunk"; return false; }
        const size_t end = rd.pos + (size_t)content + (size_t)children;
        if (id == fourcc("SIZE")) {
            sx = rd.i32(); sy = rd.i32(); sz = rd.i32();
        } else if (id == fourcc("XYZI")) {
            const int32_t n = rd.i32();
            if (n < 0) { m_error = "negative number of voxels"; return false; }
            if (rd.pos + 4 * (size_t)n > buf.size()) { m_error = "truncated voxel chunk"; return false; }
            voxels.assign(buf.begin() + rd.pos, buf.begin() + rd.pos + 4 * (size_t)n);
        } else if (id == fourcc("RGBA")) {
            /* file colour k is palette entry k + 1; the 256th is reserved */
            if (rd.pos + 255 * 4 > buf.size()) { m_error = "truncated palette chunk"; return false; }
            custom = true;
            std::memcpy(palette + 4, &buf[rd.pos], 255 * 4);
        }
        rd.pos = end;
    }
    if (!rd.ok) { m_error = "truncated file"; return false; }
    if (sx <= 0 || sy <= 0 || sz <= 0) { m_error = "missing or invalid SIZE chunk"; return false; }
    /* what the volume takes: at most 4096 voxels per axis, and linear voxel
     * indices (the emissive list) that fit in 31 bits */
    if (sx > 4096 || sy > 4096 || sz > 4096 || (uint64_t)sx * (uint64_t)sy * (uint64_t)sz >= (1ull << 31)) {
        m_error = "SIZE chunk beyond the volume's limits";
        return false;
    }
    if (!custom) defaultPalette(palette);

    /* axis conversion: grid (x, y, z) = file (x, z, y) */
    voxelResolution = V3i(sx, sz, sy);
    const size_t nvox = (size_t)sx * (size_t)sy * (size_t)sz;
    voxelMaterials.assign(nvox, -1);
    materialData.clear();
    emissiveVoxelIndices.clear();
    int32_t offset_of[256];
    for (int i = 0; i < 256; ++i) offset_of[i] = -1;
    for (size_t i = 0; i + 3 < voxels.size(); i += 4) {
        const int x = voxels[i], y = voxels[i + 1], z = voxels[i + 2], c = voxels[i + 3];
        if (c == 254) continue;   /* treated as empty, magicaVoxel.cpp:295 */
        if (x >= sx || y >= sy || z >= sz) { m_error = "voxel outside the SIZE box"; return false; }
        const size_t v = (size_t)x + (size_t)z * (size_t)voxelResolution.x +
                         (size_t)y * (size_t)voxelResolution.x * (size_t)voxelResolution.y;
        if (offset_of[c] < 0) {
            offset_of[c] = (int32_t)materialData.size();
            const V3f albedo((float)palette[4 * c + 0] / 255, (float)palette[4 * c + 1] / 255,
                             (float)palette[4 * c + 2] / 255);
            generateMaterialLambert(V3f(0.0f), albedo, materialData);
        }
        voxelMaterials[v] = offset_of[c];
    }
    for (size_t v = 0; v < nvox; ++v) {
        const int32_t off = voxelMaterials[v];
        if (off >= 0 && materialEmissiveness(&materialData[(size_t)off]) > 0.0f)
            emissiveVoxelIndices.push_back((int32_t)v);
    }
    return true;
}

/* ---- meshes (mesh.cpp:61-75, meshLoader.cpp:11-75) ------------------------------ */
Box3f computeBounds(const float* v, size_t n)
{
    Box3f b;
    for (size_t i = 0; i < n; ++i) b.extendBy(V3f(v[3 * i], v[3 * i + 1], v[3 * i + 2]));
    return b;
}

Mesh::Mesh(std::vector<float> vertices, std::vector<uint32_t> indices)
    : m_vertices(std::move(vertices)), m_indices(std::move(indices))
{
    m_bounds = computeBounds(m_vertices.data(), numVertices());
}

namespace {

/* Position index of one face corner token ("7", "7/2", "7//3", "-1/..."),
 * 0-based, or -1 if malformed / out of range. */
long corner_index(const std::string& tok, long n_positions)
{
    const size_t slash = tok.find('/');
    const std::string s = tok.substr(0, slash);
    if (s.empty()) return -1;
    char* end = nullptr;
    const long i = std::strtol(s.c_str(), &end, 10);
    if (*end != '\0' || i == 0) return -1;
    const long k = i > 0 ? i - 1 : n_positions + i;
    return (k >= 0 && k < n_positions) ? k : -1;
}

}  // namespace

bool MeshLoader::loadFromOBJ(const char* filePath, std::vector<float>& vertices, std::vector<uint32_t>& indices,
                             std::string* error)
{
    vertices.clear();
    indices.clear();
    std::ifstream f(filePath);
    if (!f) {
        if (error) *error = std::string("cannot open ") + filePath;
        return false;
    }
    std::vector<float> pos;
    std::unordered_map<long, uint32_t> remap;   /* position index -> vertex of the current group */
    std::string line;
    long line_no = 0;
    bool group_has_faces = false;
    while (std::getline(f, line)) {
        ++line_no;
        std::istringstream ls(line);
        std::string tag;
        if (!(ls >> tag) || tag[0] == '#') continue;
        if (tag == "v") {
            float x = 0, y = 0, z = 0;
            if (!(ls >> x >> y >> z)) {
                if (error) *error = "bad vertex at line " + std::to_string(line_no);
                return false;
            }
            pos.push_back(x); pos.push_back(y); pos.push_back(z);
        } else if (tag == "o" || tag == "g") {
            if (group_has_faces) remap.clear();   /* a new shape re-indexes its positions */
            group_has_faces = false;
        } else if (tag == "f") {
            std::vector<uint32_t> poly;
            std::string tok;
            const long np = (long)(pos.size() / 3);
            while (ls >> tok) {
                const long k = corner_index(tok, np);
                if (k < 0) {
                    if (error) *error = "bad face index at line " + std::to_string(line_no);
                    return false;
                }
                auto it = remap.find(k);
                uint32_t v;
                if (it == remap.end()) {
                    v = (uint32_t)(vertices.size() / 3);
                    vertices.push_back(pos[3 * k]); vertices.push_back(pos[3 * k + 1]); vertices.push_back(pos[3 * k + 2]);
                    remap.emplace(k, v);
                } else {
                    v = it->second;
                }
                poly.push_back(v);
            }
            for (size_t i = 1; i + 1 < poly.size(); ++i) {
                indices.push_back(poly[0]); indices.push_back(poly[i]); indices.push_back(poly[i + 1]);
            }
            group_has_faces = true;
        }
    }
    return true;
}

Mesh* MeshLoader::loadFromOBJ(const char* filePath, std::string* error)
{
    std::vector<float> v;
    std::vector<uint32_t> idx;
    if (!loadFromOBJ(filePath, v, idx, error)) return nullptr;
    return new Mesh(std::move(v), std::move(idx));
}

/* ---- image output (renderer.cpp:1108-1140) -------------------------------------- */
static bool ends_with(const std::string& s, const char* suf)
{
    const size_t n = std::strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
    return true;
}

bool writeImage(const std::string& file, const float* rgba, int w, int h, std::string* error)
{
    if (w <= 0 || h <= 0 || !rgba) {
        if (error) *error = "empty image";
        return false;
    }
    const bool pfm = ends_with(file, ".pfm"), ppm = ends_with(file, ".ppm");
    if (!pfm && !ppm) {
        if (error) *error = "unsupported image extension (use .pfm or .ppm): " + file;
        return false;
    }
    FILE* fp = std::fopen(file.c_str(), "wb");
    if (!fp) {
        if (error) *error = "cannot write " + file;
        return false;
    }
    bool ok = true;
    if (pfm) {
        /* PFM stores the bottom row first: buffer row order */
        std::fprintf(fp, "PF\n%d %d\n-1.0\n", w, h);
        std::vector<float> row((size_t)w * 3);
        for (int y = 0; y < h && ok; ++y) {
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c) row[(size_t)x * 3 + c] = rgba[((size_t)y * w + x) * 4 + c];
            ok = std::fwrite(row.data(), sizeof(float), row.size(), fp) == row.size();
        }
    } else {
        std::fprintf(fp, "P6\n%d %d\n255\n", w, h);
        std::vector<uint8_t> row((size_t)w * 3);
        for (int y = h - 1; y >= 0 && ok; --y) {   /* top row first: the reference's flip */
            for (int x = 0; x < w; ++x)
                for (int c = 0; c < 3; ++c) {
                    float v = rgba[((size_t)y * w + x) * 4 + c];
                    v = std::isnan(v) ? 0.0f : std::min(1.0f, std::max(0.0f, v));
                    row[(size_t)x * 3 + c] = (uint8_t)std::lround(v * 255.0f);
                }
            ok = std::fwrite(row.data(), 1, row.size(), fp) == row.size();
        }
    }
    if (std::fclose(fp) != 0) ok = false;
    if (!ok && error) *error = "write failed: " + file;
    return ok;
}

bool readImagePFM(const std::string& file, std::vector<float>& rgb, int& w, int& h, std::string* error)
{
    std::vector<uint8_t> buf;
    if (!read_file(file, buf)) {
        if (error) *error = "cannot open " + file;
        return false;
    }
    /* header: magic, width height, scale (negative = little-endian), each
     * token followed by whitespace; the pixel data follows one whitespace */
    size_t pos = 0;
    auto token = [&](std::string& t) {
        while (pos < buf.size() && std::isspace(buf[pos])) ++pos;
        t.clear();
        while (pos < buf.size() && !std::isspace(buf[pos])) t.push_back((char)buf[pos++]);
        return !t.empty();
    };
    std::string magic, sw, sh, ss;
    if (!token(magic) || !token(sw) || !token(sh) || !token(ss) || (magic != "PF" && magic != "Pf")) {
        if (error) *error = "not a PFM file: " + file;
        return false;
    }
    ++pos;   /* the single whitespace before the data */
    w = std::atoi(sw.c_str());
    h = std::atoi(sh.c_str());
    const double scale = std::atof(ss.c_str());
    const int ch = magic == "PF" ? 3 : 1;
    if (w <= 0 || h <= 0 || scale == 0.0 || buf.size() < pos + (size_t)w * h * ch * 4) {
        if (error) *error = "truncated or invalid PFM: " + file;
        return false;
    }
    const bool little = scale < 0.0;
    uint32_t probe = 1;
    const bool host_little = *reinterpret_cast<uint8_t*>(&probe) == 1;
    rgb.assign((size_t)w * h * 3, 0.0f);
    for (int y = 0; y < h; ++y)            /* stored bottom row first */
        for (int x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) {
                const size_t src = pos + (((size_t)y * w + x) * ch + (ch == 3 ? c : 0)) * 4;
                uint8_t b[4] = { buf[src], buf[src + 1], buf[src + 2], buf[src + 3] };
                if (little != host_little) { std::swap(b[0], b[3]); std::swap(b[1], b[2]); }
                float v;
                std::memcpy(&v, b, 4);
                rgb[(((size_t)(h - 1 - y)) * w + x) * 3 + c] = v;
            }
    return true;
}

/* Radiance RGBE (.hdr / .pic): "#?RADIANCE" or "#?RGBE" header lines up to
 * an empty line, the resolution line "-Y h +X w" (rows top down) or "+Y h +X w"
 * (bottom up), then the scanlines, each either run-length encoded per
 * channel (starting 2, 2, w >> 8, w & 255; runs of count - 128 copies of one
 * byte, or count literal bytes) or 4 flat bytes per pixel.  A pixel is
 * (r, g, b) * 2^(e - 136), 0 for e = 0, the conversion of the format's
 * reference reader (rgbe2float), which OpenImageIO's reader uses too. */
static bool read_hdr(const std::vector<uint8_t>& buf, std::vector<float>& rgb, int& w, int& h, std::string* error)
{
    size_t pos = 0;
    auto line = [&](std::string& t) {
        t.clear();
        while (pos < buf.size() && buf[pos] != '\n') t.push_back((char)buf[pos++]);
        if (pos >= buf.size()) return false;
        ++pos;
        return true;
    };
    std::string t;
    if (!line(t) || (t.rfind("#?RADIANCE", 0) != 0 && t.rfind("#?RGBE", 0) != 0)) {
        if (error) *error = "not a Radiance HDR file";
        return false;
    }
    for (;;) {
        if (!line(t)) { if (error) *error = "truncated HDR header"; return false; }
        if (t.empty()) break;
        if (t.rfind("FORMAT=", 0) == 0 && t != "FORMAT=32-bit_rle_rgbe") {
            if (error) *error = "unsupported HDR pixel format " + t;
            return false;
        }
    }
    if (!line(t)) { if (error) *error = "missing HDR resolution line"; return false; }
    char ys[3] = { 0 }, xs[3] = { 0 };
    if (std::sscanf(t.c_str(), "%2s %d %2s %d", ys, &h, xs, &w) != 4 || std::strcmp(xs, "+X") != 0 ||
        (std::strcmp(ys, "-Y") != 0 && std::strcmp(ys, "+Y") != 0) || w <= 0 || h <= 0 || w > 65535 || h > 65535) {
        if (error) *error = "unsupported HDR orientation or size: " + t;
        return false;
    }
    const bool top_down = std::strcmp(ys, "-Y") == 0;
    /* the shortest scanline a file can hold (run-length: a 4-byte header and
     * at least one 2-byte run per 127 pixels of each channel; flat: 4 bytes
     * per pixel), so a header cannot ask for more pixels than the file has */
    const size_t min_row = (w >= 8 && w < 32768) ? 4u + 8u * (size_t)((w + 126) / 127) : (size_t)w * 4u;
    if (buf.size() - pos < (size_t)h * min_row) {
        if (error) *error = "truncated HDR pixels";
        return false;
    }
    rgb.assign((size_t)w * h * 3, 0.0f);
    std::vector<uint8_t> row((size_t)w * 4);
    for (int y = 0; y < h; ++y) {
        const bool rle = w >= 8 && w < 32768 && pos + 4 <= buf.size() && buf[pos] == 2 && buf[pos + 1] == 2 &&
                         (((int)buf[pos + 2] << 8) | buf[pos + 3]) == w && !(buf[pos + 2] & 0x80);
        if (rle) {
            pos += 4;
            for (int c = 0; c < 4; ++c)
                for (int x = 0; x < w;) {
                    if (pos >= buf.size()) { if (error) *error = "truncated HDR scanline"; return false; }
                    int n = buf[pos++];
                    if (n > 128) {
                        n -= 128;
                        if (x + n > w || pos >= buf.size()) { if (error) *error = "bad HDR run"; return false; }
                        const uint8_t v = buf[pos++];
                        for (int k = 0; k < n; ++k) row[(size_t)(x++) * 4 + c] = v;
                    } else {
                        if (n == 0 || x + n > w || pos + n > buf.size()) { if (error) *error = "bad HDR run"; return false; }
                        for (int k = 0; k < n; ++k) row[(size_t)(x++) * 4 + c] = buf[pos++];
                    }
                }
        } else {
            if (pos + (size_t)w * 4 > buf.size()) { if (error) *error = "truncated HDR pixels"; return false; }
            std::memcpy(row.data(), &buf[pos], (size_t)w * 4);
            pos += (size_t)w * 4;
        }
        const int dst_row = top_down ? y : h - 1 - y;
        for (int x = 0; x < w; ++x) {
            const uint8_t* q = &row[(size_t)x * 4];
            const float f = q[3] ? std::ldexp(1.0f, (int)q[3] - 136) : 0.0f;
            float* o = &rgb[((size_t)dst_row * w + x) * 3];
            o[0] = q[3] ? (float)q[0] * f : 0.0f;
            o[1] = q[3] ? (float)q[1] * f : 0.0f;
            o[2] = q[3] ? (float)q[2] * f : 0.0f;
        }
    }
    return true;
}

/* Binary PPM / PGM (P6 / P5, maxval up to 65535): values / maxval, the
 * normalised float conversion of OpenImageIO's integer formats. */
static bool read_pnm(const std::vector<uint8_t>& buf, std::vector<float>& rgb, int& w, int& h, std::string* error)
{
    size_t pos = 0;
    auto token = [&](std::string& t) {
        for (;;) {
            while (pos < buf.size() && std::isspace(buf[pos])) ++pos;
            if (pos < buf.size() && buf[pos] == '#') {
                while (pos < buf.size() && buf[pos] != '\n') ++pos;
                continue;
            }
            break;
        }
        t.clear();
        while (pos < buf.size() && !std::isspace(buf[pos])) t.push_back((char)buf[pos++]);
        return !t.empty();
    };
    std::string magic, sw, sh, sm;
    if (!token(magic) || (magic != "P6" && magic != "P5") || !token(sw) || !token(sh) || !token(sm)) {
        if (error) *error = "not a binary PPM/PGM file";
        return false;
    }
    ++pos;
    w = std::atoi(sw.c_str());
    h = std::atoi(sh.c_str());
    const int maxval = std::atoi(sm.c_str());
    const int ch = magic == "P6" ? 3 : 1, bytes = maxval > 255 ? 2 : 1;
    if (w <= 0 || h <= 0 || maxval <= 0 || maxval > 65535 || buf.size() < pos + (size_t)w * h * ch * bytes) {
        if (error) *error = "truncated or invalid PPM/PGM";
        return false;
    }
    rgb.assign((size_t)w * h * 3, 0.0f);
    for (size_t i = 0; i < (size_t)w * h; ++i)
        for (int c = 0; c < 3; ++c) {
            const size_t src = pos + (i * ch + (ch == 3 ? c : 0)) * bytes;
            const int v = bytes == 2 ? ((int)buf[src] << 8) | buf[src + 1] : buf[src];
            rgb[i * 3 + c] = (float)v / (float)maxval;
        }
    return true;
}

bool readImage(const std::string& file, std::vector<float>& rgb, int& w, int& h, std::string* error)
{
    if (ends_with(file, ".pfm")) return readImagePFM(file, rgb, w, h, error);
    const bool hdr = ends_with(file, ".hdr") || ends_with(file, ".pic");
    const bool pnm = ends_with(file, ".ppm") || ends_with(file, ".pgm");
    if (!hdr && !pnm) {
        if (error) *error = "unsupported image extension (use .pfm, .hdr, .pic, .ppm or .pgm): " + file;
        return false;
    }
    std::vector<uint8_t> buf;
    if (!read_file(file, buf)) {
        if (error) *error = "cannot open " + file;
        return false;
    }
    std::string err;
    const bool ok = hdr ? read_hdr(buf, rgb, w, h, &err) : read_pnm(buf, rgb, w, h, &err);
    if (!ok && error) *error = err + ": " + file;
    return ok;
}

}  // namespace voxeltoy

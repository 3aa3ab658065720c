/*
 * scene_io.h - scene inputs and outputs of the renderer:
 *   - material record layout and serialisation for UIs
 *     (src/renderer/material/material.h:1-56, material.cpp:1-99),
 *   - the .vox loader interface and MagicaVoxel loader
 *     (src/renderer/loaders/voxLoader.h:12-49, magicaVoxel.cpp:11-321),
 *   - triangle meshes and the OBJ loader
 *     (src/mesh/mesh.h:7-33, src/mesh/meshLoader.cpp:11-75),
 *   - image output of the accumulated frame (Renderer::saveImage,
 *     src/renderer/renderer.cpp:1108-1140) as PFM (float RGB) or PPM.
 * Errors are reported as false / empty results plus a message, never by
 * exceptions.
 */
#pragma once

#include "vt_math_types.h"

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace voxeltoy {

namespace Material {

enum MaterialType { MT_LAMBERT = 0, MT_METAL = 1, MT_PLASTIC = 2 };

/* Record layout in the material data array: [type][emission rgb][albedo or
 * reflectance rgb][roughness (metal, plastic)]. */
struct LambertMaterialData {
    float emission[3];
    float albedo[3];
};
struct MetalMaterialData {
    float emission[3];
    float reflectance[3];
    float roughness;
};
struct PlasticMaterialData {
    float emission[3];
    float diffuseAlbedo[3];
    float roughness;
};

/* One editable property (or a group of them) and where it lives in the
 * material data array. */
struct SerializedData {
    std::string m_propertyName;
    float m_dataRangeFrom = 0.0f, m_dataRangeTo = 0.0f;
    float m_value = 0.0f;
    size_t m_dataOffset = 0;
    enum { PROPERTY_TYPE_FLOAT, PROPERTY_TYPE_COLOR } m_propertyType = PROPERTY_TYPE_FLOAT;
    std::vector<SerializedData> m_childProperties;
};

SerializedData serializeLambert(const LambertMaterialData&, size_t offset);
SerializedData serializeMetal(const MetalMaterialData&, size_t offset);
SerializedData serializePlastic(const PlasticMaterialData&, size_t offset);

/* Walk a material data array record by record (stops at an unknown type,
 * reporting it through *error when given). */
std::vector<SerializedData> serializeAll(const std::vector<float>& data, std::string* error = nullptr);

/* Floats a record of this type occupies, type field included (0 if unknown). */
size_t recordSize(float type);

}  // namespace Material

/* Voxel file loader: a dense grid of material offsets (-1 = empty, x
 * fastest, then y, then z), the material data array the offsets point into,
 * and the linear indices of voxels whose material emits light. */
class VoxLoader {
public:
    virtual ~VoxLoader() = default;
    virtual bool load(const std::string& filePath, std::vector<int32_t>& voxelMaterials,
                      std::vector<float>& materialData, std::vector<int32_t>& emissiveVoxelIndices,
                      V3i& voxelResolution) = 0;
    const std::string& error() const { return m_error; }

protected:
    void generateMaterialLambert(V3f emission, V3f albedo, std::vector<float>& materialData);
    void generateMaterialMetal(V3f emission, V3f reflectance, float roughness, std::vector<float>& materialData);
    void generateMaterialPlastic(V3f emission, V3f diffuseAlbedo, float roughness, std::vector<float>& materialData);
    /* mean emission of the record starting at `record` (its type field) */
    static float materialEmissiveness(const float* record);
    std::string m_error;

private:
    void storeMaterialData(Material::MaterialType type, const float* data, size_t n, std::vector<float>& storage);
};

/* MagicaVoxel .vox (version 150: SIZE, XYZI, RGBA chunks).  File axes are
 * converted as the reference does (file z is up: grid y = file z, grid z =
 * file y); colour index 254 is treated as empty; every palette colour used
 * becomes one Lambert record with the palette colour as albedo.  The
 * emissive list holds the voxels whose record has non-zero mean emission
 * (none for .vox input, whose records never emit). */
class MagicaVoxelLoader : public VoxLoader {
public:
    bool load(const std::string& filePath, std::vector<int32_t>& voxelMaterials, std::vector<float>& materialData,
              std::vector<int32_t>& emissiveVoxelIndices, V3i& voxelResolution) override;
    /* The default palette (RGBA bytes) used when the file has no RGBA chunk. */
    static void defaultPalette(uint8_t rgba[256 * 4]);
};

/* Indexed triangle mesh (positions only). */
class Mesh {
public:
    Mesh() = default;
    Mesh(std::vector<float> vertices, std::vector<uint32_t> indices);
    const std::vector<float>& vertices() const { return m_vertices; }
    const std::vector<uint32_t>& indices() const { return m_indices; }
    size_t numVertices() const { return m_vertices.size() / 3; }
    size_t numTriangles() const { return m_indices.size() / 3; }
    Box3f bounds() const { return m_bounds; }

private:
    std::vector<float> m_vertices;
    std::vector<uint32_t> m_indices;
    Box3f m_bounds;
};

Box3f computeBounds(const float* vertices, size_t numVertices);

class MeshLoader {
public:
    /* Wavefront OBJ: 'v' positions and 'f' faces (v, v/vt, v//vn, v/vt/vn,
     * negative indices relative to the end); polygons are split into fans
     * (0, i, i+1); each 'o'/'g' group keeps only the positions its faces use,
     * in order of first use, as tinyobjloader 0.9 does.  Returns false (and a
     * message) on unreadable files or out-of-range indices. */
    static bool loadFromOBJ(const char* filePath, std::vector<float>& vertices, std::vector<uint32_t>& indices,
                            std::string* error = nullptr);
    /* nullptr when the file cannot be read */
    static Mesh* loadFromOBJ(const char* filePath, std::string* error = nullptr);
};

/* Write a W x H RGBA float image whose row 0 is the bottom row (the
 * accumulation buffer's order) top row first, as the reference's flipped
 * write does: ".pfm" -> little-endian float RGB PFM (which stores bottom
 * row first, so rows are kept in buffer order), ".ppm" -> 8-bit RGB with
 * values clamped to [0, 1].  Returns false on I/O errors or other
 * extensions. */
bool writeImage(const std::string& file, const float* rgba, int width, int height, std::string* error = nullptr);

/* Read a float PFM ("PF" colour or "Pf" grey, either endianness) into RGB
 * floats with row 0 = the top row (the environment image convention).
 * Returns false (and a message) on malformed files. */
bool readImagePFM(const std::string& file, std::vector<float>& rgb, int& width, int& height,
                  std::string* error = nullptr);

/* loadImage (src/renderer/image.cpp:28-59) for the formats this library
 * reads without OpenImageIO: float PFM, Radiance RGBE (.hdr, .pic; flat or
 * run-length encoded) and binary PPM/PGM (8 or 16 bits, values / maxval).
 * RGB floats, row 0 = the top row. */
bool readImage(const std::string& file, std::vector<float>& rgb, int& width, int& height,
               std::string* error = nullptr);

}  // namespace voxeltoy

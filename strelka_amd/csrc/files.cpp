// files.cpp - on-disk formats (host only): skip-link BVH files, PFM and
// OpenEXR images.
//
// The reference keeps its BVH in memory only and writes frames as PNG through
// HioImage (src/hdRunner/main.cpp:339-356); converged references and RMSE
// tooling need lossless float images, and rebuilding a large tree for every
// run is avoidable, so these two plain formats are this library's own.
//
// BVH file: 16-byte header {"STKBVH01", uint32 node count, uint32 0}, then the
// nodes as StkBvhNode records (32 B, little-endian), the reference's skip-link
// layout (include/render/bvh.h:17-24) that stk_set_bvh accepts.
// PFM: "PF\n<w> <h>\n-1\n" then RGB float32 little-endian rows, bottom row
// first (the Portable Float Map convention).
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/stk/stk.h"

namespace {

const char kBvhMagic[8] = {'S', 'T', 'K', 'B', 'V', 'H', '0', '1'};

struct File {
    FILE* f = nullptr;
    File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
    ~File()
    {
        if (f) std::fclose(f);
    }
};

} // namespace

extern "C" {

STK_API stk_status stk_bvh_save_nodes(const StkBvhNode* nodes, uint32_t numNodes, const char* path)
{
    if (!path || (numNodes && !nodes)) return STK_E_INVALID;
    File file(path, "wb");
    if (!file.f) return STK_E_INVALID;
    uint32_t hdr[2] = {numNodes, 0u};
    if (std::fwrite(kBvhMagic, 1, 8, file.f) != 8 || std::fwrite(hdr, 4, 2, file.f) != 2) return STK_E_INVALID;
    if (numNodes && std::fwrite(nodes, sizeof(StkBvhNode), numNodes, file.f) != numNodes) return STK_E_INVALID;
    return std::fflush(file.f) == 0 ? STK_OK : STK_E_INVALID;
}

STK_API stk_status stk_bvh_load_nodes(const char* path, StkBvhNode* nodes, uint32_t capacity, uint32_t* numNodes)
{
    if (!path || !numNodes) return STK_E_INVALID;
    File file(path, "rb");
    if (!file.f) return STK_E_INVALID;
    char magic[8];
    uint32_t hdr[2];
    if (std::fread(magic, 1, 8, file.f) != 8 || std::memcmp(magic, kBvhMagic, 8) != 0 ||
        std::fread(hdr, 4, 2, file.f) != 2 || hdr[1] != 0u)
        return STK_E_INVALID;
    *numNodes = hdr[0];
    if (!nodes) return STK_OK; // size query
    if (capacity < hdr[0]) return STK_E_INVALID;
    if (hdr[0] && std::fread(nodes, sizeof(StkBvhNode), hdr[0], file.f) != hdr[0]) return STK_E_INVALID;
    char extra;
    return std::fread(&extra, 1, 1, file.f) == 0 ? STK_OK : STK_E_INVALID; // no trailing bytes
}

STK_API stk_status stk_image_write_pfm(const char* path, const float* rgba, uint32_t width, uint32_t height)
{
    if (!path || !rgba || width == 0 || height == 0) return STK_E_INVALID;
    File file(path, "wb");
    if (!file.f) return STK_E_INVALID;
    const std::string hdr = "PF\n" + std::to_string(width) + " " + std::to_string(height) + "\n-1\n";
    if (std::fwrite(hdr.data(), 1, hdr.size(), file.f) != hdr.size()) return STK_E_INVALID;
    std::vector<float> row((size_t)width * 3);
    for (uint32_t y = 0; y < height; ++y) {
        const float* src = rgba + (size_t)(height - 1 - y) * width * 4;
        for (uint32_t x = 0; x < width; ++x)
            for (int c = 0; c < 3; ++c) row[(size_t)x * 3 + c] = src[(size_t)x * 4 + c];
        if (std::fwrite(row.data(), sizeof(float), row.size(), file.f) != row.size()) return STK_E_INVALID;
    }
    return std::fflush(file.f) == 0 ? STK_OK : STK_E_INVALID;
}

// OpenEXR 2.0 single-part scanline file, uncompressed, channels B, G, R as
// 32-bit floats (the header lists channels in name order and every scanline
// stores them in that order): magic 20000630, version 2, the required header
// attributes, a table of one 64-bit file offset per scanline, then per
// scanline {int32 y, int32 byte count, B row, G row, R row}.  Top row first
// (increasing y), as stk_render writes its image.
namespace {

struct ExrHeader {
    std::vector<unsigned char> b;
    void bytes(const void* p, size_t n)
    {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        b.insert(b.end(), c, c + n);
    }
    void str(const char* s) { bytes(s, std::strlen(s) + 1); }
    void i32(int32_t v) { bytes(&v, 4); }
    void f32(float v) { bytes(&v, 4); }
    void attr(const char* name, const char* type, int32_t size)
    {
        str(name);
        str(type);
        i32(size);
    }
};

} // namespace

STK_API stk_status stk_image_write_exr(const char* path, const float* rgba, uint32_t width, uint32_t height)
{
    if (!path || !rgba || width == 0 || height == 0 || width > (1u << 24) || height > (1u << 24))
        return STK_E_INVALID;
    const char* names[3] = {"B", "G", "R"};
    const int comp[3] = {2, 1, 0}; // RGBA component of each listed channel
    ExrHeader h;
    const uint32_t magic = 20000630u, version = 2u;
    h.bytes(&magic, 4);
    h.bytes(&version, 4);
    h.attr("channels", "chlist", 3 * (2 + 16) + 1);
    for (const char* n : names) {
        h.str(n);
        h.i32(2);                          // pixel type FLOAT
        const unsigned char lin[4] = {0, 0, 0, 0}; // pLinear + 3 reserved bytes
        h.bytes(lin, 4);
        h.i32(1); // x sampling
        h.i32(1); // y sampling
    }
    h.bytes("", 1);
    h.attr("compression", "compression", 1);
    h.bytes("", 1); // NO_COMPRESSION
    for (const char* win : {"dataWindow", "displayWindow"}) {
        h.attr(win, "box2i", 16);
        h.i32(0);
        h.i32(0);
        h.i32((int32_t)width - 1);
        h.i32((int32_t)height - 1);
    }
    h.attr("lineOrder", "lineOrder", 1);
    h.bytes("", 1); // INCREASING_Y
    h.attr("pixelAspectRatio", "float", 4);
    h.f32(1.0f);
    h.attr("screenWindowCenter", "v2f", 8);
    h.f32(0.0f);
    h.f32(0.0f);
    h.attr("screenWindowWidth", "float", 4);
    h.f32(1.0f);
    h.bytes("", 1); // end of header
    File file(path, "wb");
    if (!file.f) return STK_E_INVALID;
    if (std::fwrite(h.b.data(), 1, h.b.size(), file.f) != h.b.size()) return STK_E_INVALID;
    const uint64_t rowBytes = (uint64_t)width * 3 * sizeof(float);
    std::vector<uint64_t> offsets(height);
    const uint64_t first = h.b.size() + (uint64_t)height * 8;
    for (uint32_t y = 0; y < height; ++y) offsets[y] = first + (uint64_t)y * (8 + rowBytes);
    if (std::fwrite(offsets.data(), 8, height, file.f) != height) return STK_E_INVALID;
    std::vector<float> row((size_t)width * 3);
    for (uint32_t y = 0; y < height; ++y) {
        const float* src = rgba + (size_t)y * width * 4;
        for (int c = 0; c < 3; ++c)
            for (uint32_t x = 0; x < width; ++x) row[(size_t)c * width + x] = src[(size_t)x * 4 + comp[c]];
        const int32_t lead[2] = {(int32_t)y, (int32_t)rowBytes};
        if (std::fwrite(lead, 4, 2, file.f) != 2 || std::fwrite(row.data(), sizeof(float), row.size(), file.f) != row.size())
            return STK_E_INVALID;
    }
    return std::fflush(file.f) == 0 ? STK_OK : STK_E_INVALID;
}

// Reads a little-endian RGB PFM ("PF", negative scale) into RGBA rows, top row
// first (alpha 1).  rgba may be null to query the size.
STK_API stk_status stk_image_read_pfm(const char* path, float* rgba, size_t capacityFloats, uint32_t* width,
                                      uint32_t* height)
{
    if (!path || !width || !height) return STK_E_INVALID;
    File file(path, "rb");
    if (!file.f) return STK_E_INVALID;
    char tag[3] = {0, 0, 0};
    unsigned w = 0, h = 0;
    double scale = 0.0;
    if (std::fscanf(file.f, "%2s %u %u %lf", tag, &w, &h, &scale) != 4 || std::strcmp(tag, "PF") != 0 || w == 0 ||
        h == 0 || !(scale < 0.0))
        return STK_E_INVALID;
    if (std::fgetc(file.f) != '\n') return STK_E_INVALID; // the single separator before the data
    *width = w;
    *height = h;
    if (!rgba) return STK_OK;
    if (capacityFloats < (size_t)w * h * 4) return STK_E_INVALID;
    std::vector<float> row((size_t)w * 3);
    for (uint32_t y = 0; y < h; ++y) {
        if (std::fread(row.data(), sizeof(float), row.size(), file.f) != row.size()) return STK_E_INVALID;
        float* dst = rgba + (size_t)(h - 1 - y) * w * 4;
        for (uint32_t x = 0; x < w; ++x) {
            for (int c = 0; c < 3; ++c) dst[(size_t)x * 4 + c] = row[(size_t)x * 3 + c];
            dst[(size_t)x * 4 + 3] = 1.0f;
        }
    }
    return STK_OK;
}

} // extern "C"

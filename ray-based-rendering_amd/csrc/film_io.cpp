// Film output: the reference's 8-bit conversion (gamma or Reinhard + gamma), PNG writing, and
// HDR films as PFM (Portable Float Map), which the reference does not have.  Host code; films
// are the C ABI's float RGB arrays, row-major from the top row.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rbr.h"

namespace {

// Film.cpp's per-channel encode: pow(v, 1/gamma), clamped with common.h's
// max(lo, min(hi, v)) (a NaN becomes 1), scaled, rounded by +0.5 and truncated.
unsigned char encode(float value, float invGamma) {
    const float p = std::pow(value, invGamma);
    const float hiClamped = p < 1.0f ? p : 1.0f;                 // std::min(1, p)
    const float c = 0.0f < hiClamped ? hiClamped : 0.0f;        // std::max(0, .)
    return static_cast<unsigned char>(c * 255.0f + 0.5f);
}

// CRC-32 (IEEE 802.3, reflected polynomial 0xEDB88320) as PNG chunks use it.
uint32_t crc32(const unsigned char* p, size_t n, uint32_t crc = 0xFFFFFFFFu) {
    static const std::array<uint32_t, 256> table = [] {
        std::array<uint32_t, 256> t{};
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[i] = c;
        }
        return t;
    }();
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
    return crc;
}

void put32(std::vector<unsigned char>& v, uint32_t x) {
    v.push_back(static_cast<unsigned char>(x >> 24));
    v.push_back(static_cast<unsigned char>(x >> 16));
    v.push_back(static_cast<unsigned char>(x >> 8));
    v.push_back(static_cast<unsigned char>(x));
}

void chunk(std::vector<unsigned char>& out, const char* type, const std::vector<unsigned char>& data) {
    put32(out, static_cast<uint32_t>(data.size()));
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put32(out, crc32(out.data() + start, out.size() - start) ^ 0xFFFFFFFFu);
}

bool writeFile(const char* path, const void* data, size_t n) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, 1, n, f) == n;
    return std::fclose(f) == 0 && ok;
}

}  // namespace

extern "C" {

int rbr_film_to_rgb8(const float* rgb, int width, int height, float gamma, int tone_map, float exposure,
                     unsigned char* out) {
    if (!rgb || !out || width < 1 || height < 1 || !(gamma > 0.0f)) return RBR_ERR_INVALID;
    const float invGamma = 1.0f / gamma;
    const size_t n = static_cast<size_t>(width) * height * 3;
    for (size_t i = 0; i < n; ++i) {
        float c = rgb[i];
        if (tone_map) {
            c = c * exposure;
            c = c / (c + 1.0f);   // Reinhard
        }
        out[i] = encode(c, invGamma);
    }
    return RBR_OK;
}

// PNG with stored (uncompressed) deflate blocks: filter type 0 per row, zlib header 0x78 0x01,
// Adler-32 trailer.
int rbr_film_write_png(const char* path, const unsigned char* rgb8, int width, int height) {
    if (!path || !rgb8 || width < 1 || height < 1) return RBR_ERR_INVALID;
    const size_t row = static_cast<size_t>(width) * 3 + 1;
    std::vector<unsigned char> raw(row * height);
    for (int y = 0; y < height; ++y) {
        raw[y * row] = 0;
        std::memcpy(&raw[y * row + 1], rgb8 + static_cast<size_t>(y) * width * 3, row - 1);
    }
    std::vector<unsigned char> z = {0x78, 0x01};
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back(static_cast<unsigned char>(len));
        z.push_back(static_cast<unsigned char>(len >> 8));
        z.push_back(static_cast<unsigned char>(~len));
        z.push_back(static_cast<unsigned char>(~len >> 8));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;
    for (unsigned char c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put32(z, (b << 16) | a);
    std::vector<unsigned char> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<unsigned char> ihdr;
    put32(ihdr, static_cast<uint32_t>(width));
    put32(ihdr, static_cast<uint32_t>(height));
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});   // 8 bits, RGB, deflate, filter 0, no interlace
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    return writeFile(path, png.data(), png.size()) ? RBR_OK : RBR_ERR_INVALID;
}

// PFM: "PF\n<w> <h>\n-1.0\n" then little-endian FP32 RGB rows from the bottom row up.
int rbr_film_write_pfm(const char* path, const float* rgb, int width, int height) {
    if (!path || !rgb || width < 1 || height < 1) return RBR_ERR_INVALID;
    const std::string head = "PF\n" + std::to_string(width) + " " + std::to_string(height) + "\n-1.0\n";
    const size_t rowf = static_cast<size_t>(width) * 3;
    std::vector<unsigned char> buf(head.begin(), head.end());
    buf.resize(head.size() + rowf * height * 4);
    for (int y = 0; y < height; ++y)
        std::memcpy(&buf[head.size() + static_cast<size_t>(height - 1 - y) * rowf * 4], rgb + y * rowf, rowf * 4);
    return writeFile(path, buf.data(), buf.size()) ? RBR_OK : RBR_ERR_INVALID;
}

// Reads a little-endian RGB PFM ("PF", negative scale).  With rgb == nullptr only the size is
// returned; otherwise rgb must hold capacity >= width * height * 3 floats.
int rbr_film_read_pfm(const char* path, int* width, int* height, float* rgb, size_t capacity) {
    if (!path || !width || !height) return RBR_ERR_INVALID;
    FILE* f = std::fopen(path, "rb");
    if (!f) return RBR_ERR_INVALID;
    char magic[3] = {0, 0, 0};
    int w = 0, h = 0;
    float scale = 0.0f;
    int rc = RBR_ERR_INVALID;
    if (std::fscanf(f, "%2s %d %d %f", magic, &w, &h, &scale) == 4 && std::strcmp(magic, "PF") == 0 && w > 0 &&
        h > 0 && scale < 0.0f && std::fgetc(f) != EOF) {
        *width = w;
        *height = h;
        rc = RBR_OK;
        if (rgb) {
            const size_t rowf = static_cast<size_t>(w) * 3;
            if (capacity < rowf * h) rc = RBR_ERR_INVALID;
            for (int y = h - 1; y >= 0 && rc == RBR_OK; --y)
                if (std::fread(rgb + y * rowf, 4, rowf, f) != rowf) rc = RBR_ERR_INVALID;
        }
    }
    std::fclose(f);
    return rc;
}

}  // extern "C"

// PNG dump of the tone-mapped output image, the file-side counterpart of the
// reference's copy of its B8G8R8A8_UNORM output image to the swapchain
// (ApplicationReSTIRGI.cpp:247-267; format from XVKSwapChain.cpp:112).
//
// 8-bit RGB, no interlace, filter type 0 on every row, and a zlib stream of
// stored (uncompressed) deflate blocks: no compression library is needed and
// the bytes are a plain function of the pixels.  The alpha channel is dropped
// (the passes write 255 everywhere).  CRC-32 (ISO 3309 polynomial 0xEDB88320,
// as PNG specifies) and Adler-32 are computed here.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "xr_host.h"

namespace {

uint32_t crc_table_entry(uint32_t n) {
    uint32_t c = n;
    for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
    return c;
}

struct Crc {
    uint32_t table[256];
    Crc() {
        for (uint32_t n = 0; n < 256; ++n) table[n] = crc_table_entry(n);
    }
    uint32_t update(uint32_t crc, const uint8_t* p, size_t n) const {
        uint32_t c = crc ^ 0xFFFFFFFFu;
        for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
        return c ^ 0xFFFFFFFFu;
    }
};

void be32(std::vector<uint8_t>& out, uint32_t v) {
    out.push_back((uint8_t)(v >> 24));
    out.push_back((uint8_t)(v >> 16));
    out.push_back((uint8_t)(v >> 8));
    out.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t>& out, const Crc& crc, const char type[4], const std::vector<uint8_t>& data) {
    be32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    be32(out, crc.update(0, out.data() + start, out.size() - start));
}

}  // namespace

extern "C" {

int xrh_png_encode(const uint32_t* bgra8, uint32_t width, uint32_t height, uint8_t* dst, size_t dst_bytes,
                   size_t* written) {
    if (!bgra8 || !written || width == 0 || height == 0 || width > (1u << 24) || height > (1u << 24))
        return XR_ERR_INVALID_ARGUMENT;
    // raw scanlines: filter byte 0 + RGB
    const size_t row = 1 + 3 * (size_t)width;
    std::vector<uint8_t> raw;
    std::vector<uint8_t> png;
    try {
        raw.resize(row * height);
        for (uint32_t y = 0; y < height; ++y) {
            uint8_t* r = raw.data() + row * y;
            r[0] = 0;
            for (uint32_t x = 0; x < width; ++x) {
                const uint32_t p = bgra8[(size_t)y * width + x];  // bytes B, G, R, A
                r[1 + 3 * x] = (uint8_t)(p >> 16);
                r[2 + 3 * x] = (uint8_t)(p >> 8);
                r[3 + 3 * x] = (uint8_t)p;
            }
        }
        // zlib stream: header 0x78 0x01, stored blocks of at most 65535 bytes,
        // Adler-32 of the raw data
        std::vector<uint8_t> z;
        z.reserve(raw.size() + raw.size() / 65535 * 5 + 16);
        z.push_back(0x78);
        z.push_back(0x01);
        size_t off = 0;
        do {
            const size_t n = raw.size() - off < 65535 ? raw.size() - off : 65535;
            const bool last = off + n == raw.size();
            z.push_back(last ? 1 : 0);
            z.push_back((uint8_t)n);
            z.push_back((uint8_t)(n >> 8));
            z.push_back((uint8_t)~n);
            z.push_back((uint8_t)(~n >> 8));
            z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
            off += n;
        } while (off < raw.size());
        uint32_t a = 1, b = 0;
        for (uint8_t v : raw) {
            a = (a + v) % 65521u;
            b = (b + a) % 65521u;
        }
        be32(z, (b << 16) | a);

        static const Crc crc;
        static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
        png.insert(png.end(), sig, sig + 8);
        std::vector<uint8_t> ihdr;
        be32(ihdr, width);
        be32(ihdr, height);
        ihdr.push_back(8);  // bit depth
        ihdr.push_back(2);  // colour type: RGB
        ihdr.push_back(0);  // deflate
        ihdr.push_back(0);  // filter method 0
        ihdr.push_back(0);  // no interlace
        chunk(png, crc, "IHDR", ihdr);
        chunk(png, crc, "IDAT", z);
        chunk(png, crc, "IEND", std::vector<uint8_t>());
    } catch (...) {
        return XR_ERR_OUT_OF_MEMORY;
    }
    *written = png.size();
    if (!dst) return XR_OK;  // size query
    if (dst_bytes < png.size()) return XR_ERR_INVALID_ARGUMENT;
    memcpy(dst, png.data(), png.size());
    return XR_OK;
}

int xrh_write_png(const char* path, const uint32_t* bgra8, uint32_t width, uint32_t height) {
    if (!path) return XR_ERR_INVALID_ARGUMENT;
    size_t n = 0;
    int rc = xrh_png_encode(bgra8, width, height, nullptr, 0, &n);
    if (rc) return rc;
    std::vector<uint8_t> buf;
    try {
        buf.resize(n);
    } catch (...) {
        return XR_ERR_OUT_OF_MEMORY;
    }
    rc = xrh_png_encode(bgra8, width, height, buf.data(), buf.size(), &n);
    if (rc) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) return XR_ERR_INVALID_ARGUMENT;
    const bool ok = fwrite(buf.data(), 1, n, f) == n;
    return (fclose(f) == 0 && ok) ? XR_OK : XR_ERR_INVALID_ARGUMENT;
}

}  // extern "C"

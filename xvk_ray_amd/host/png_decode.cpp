// PNG decoding for glTF base-colour textures: the reference decodes them with
// stb_image through tinygltf, asking for 4 components when an image has 3
// (tinygltf LoadImageData), and uploads them as R8G8B8A8_UNORM
// (src/Asset/TextureImage.cpp:25, :51).  This decoder writes RGBA8 for every
// PNG colour type: grey replicated into R, G, B; alpha 255 unless the image
// has alpha or a tRNS chunk; 16-bit samples keep their high byte and samples
// below 8 bits are scaled by 255 / (2^depth - 1), as stb_image does.
//
// Inflate (RFC 1951) with canonical Huffman decoding one bit at a time, in
// the shape of Mark Adler's puff.c (zlib's contrib/puff, a reference inflater
// written for clarity: counts and symbols per code length, decoded bit by
// bit), inside the zlib wrapper (RFC 1950, Adler-32 checked); PNG filters 0-4 with
// the Paeth predictor (PNG 1.2, section 6); chunk CRCs checked.  Interlaced
// (Adam7) images are refused.
#include <stdint.h>
#include <string.h>

#include <string>
#include <vector>

#include "xr_host.h"

namespace {

struct DecodeError {
    const char* msg;
};
[[noreturn]] void fail(const char* m) { throw DecodeError{m}; }

// ------------------------------------------------------------- inflate ----
class Inflater {
public:
    Inflater(const uint8_t* p, size_t n) : p_(p), n_(n) {}

    // expect: the size the image needs; a stream that inflates past it is
    // refused rather than followed.
    std::vector<uint8_t> run(size_t expect) {
        limit_ = expect;
        out_.reserve(expect);
        int last;
        do {
            last = (int)bits(1);
            const uint32_t type = bits(2);
            if (type == 0) stored();
            else if (type == 1) fixed();
            else if (type == 2) dynamic();
            else fail("deflate: reserved block type");
        } while (!last);
        return std::move(out_);
    }
    size_t consumed() const { return pos_; }

private:
    struct Huffman {
        uint16_t count[16];
        uint16_t symbol[320];
    };

    const uint8_t* p_;
    size_t n_;
    size_t pos_ = 0;
    uint32_t buf_ = 0;
    int cnt_ = 0;
    std::vector<uint8_t> out_;
    size_t limit_ = 0;

    void room(size_t more) {
        if (out_.size() + more > limit_) fail("deflate: more data than the image holds");
    }

    uint32_t bits(int need) {
        uint32_t v = buf_;
        while (cnt_ < need) {
            if (pos_ >= n_) fail("deflate: out of input");
            v |= (uint32_t)p_[pos_++] << cnt_;
            cnt_ += 8;
        }
        buf_ = v >> need;
        cnt_ -= need;
        return v & ((1u << need) - 1u);
    }

    // Canonical code from code lengths; returns the number of unused codes
    // (negative: over-subscribed).
    static int build(Huffman* h, const uint8_t* length, int n) {
        memset(h->count, 0, sizeof(h->count));
        for (int s = 0; s < n; ++s) h->count[length[s]]++;
        if (h->count[0] == n) return 0;
        int left = 1;
        for (int len = 1; len < 16; ++len) {
            left <<= 1;
            left -= h->count[len];
            if (left < 0) return left;
        }
        uint16_t offs[16];
        offs[1] = 0;
        for (int len = 1; len < 15; ++len) offs[len + 1] = (uint16_t)(offs[len] + h->count[len]);
        for (int s = 0; s < n; ++s)
            if (length[s] != 0) h->symbol[offs[length[s]]++] = (uint16_t)s;
        return left;
    }

    int decode(const Huffman& h) {
        int code = 0, first = 0, index = 0;
        for (int len = 1; len < 16; ++len) {
            code |= (int)bits(1);
            const int count = h.count[len];
            if (code - count < first) return h.symbol[index + (code - first)];
            index += count;
            first += count;
            first <<= 1;
            code <<= 1;
        }
        fail("deflate: bad code");
    }

    void stored() {
        buf_ = 0;
        cnt_ = 0;  // discard to the byte boundary
        if (pos_ + 4 > n_) fail("deflate: short stored header");
        const uint32_t len = p_[pos_] | (p_[pos_ + 1] << 8);
        const uint32_t nlen = p_[pos_ + 2] | (p_[pos_ + 3] << 8);
        pos_ += 4;
        if (len != (~nlen & 0xFFFFu)) fail("deflate: stored length mismatch");
        if (pos_ + len > n_) fail("deflate: short stored block");
        room(len);
        out_.insert(out_.end(), p_ + pos_, p_ + pos_ + len);
        pos_ += len;
    }

    void codes(const Huffman& lit, const Huffman& dist) {
        static const uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                              31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
        static const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                              2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
        static const uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,
                                               33,  49,  65,  97,  129, 193,  257,  385,  513,  769,
                                               1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
        static const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2,  3,  3,  4,  4,  5,  5,  6,
                                               6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
        for (;;) {
            int sym = decode(lit);
            if (sym < 256) {
                room(1);
                out_.push_back((uint8_t)sym);
                continue;
            }
            if (sym == 256) return;
            sym -= 257;
            if (sym >= 29) fail("deflate: bad length symbol");
            const uint32_t len = kLenBase[sym] + bits(kLenExtra[sym]);
            const int ds = decode(dist);
            if (ds >= 30) fail("deflate: bad distance symbol");
            const size_t d = kDistBase[ds] + bits(kDistExtra[ds]);
            if (d > out_.size()) fail("deflate: distance too far back");
            room(len);
            const size_t from = out_.size() - d;
            for (uint32_t k = 0; k < len; ++k) out_.push_back(out_[from + k]);
        }
    }

    struct FixedCodes {
        Huffman lit, dist;
        FixedCodes() {
            uint8_t l[288];
            int s = 0;
            for (; s < 144; ++s) l[s] = 8;
            for (; s < 256; ++s) l[s] = 9;
            for (; s < 280; ++s) l[s] = 7;
            for (; s < 288; ++s) l[s] = 8;
            build(&lit, l, 288);
            uint8_t d[30];
            for (int k = 0; k < 30; ++k) d[k] = 5;
            build(&dist, d, 30);
        }
    };

    void fixed() {
        static const FixedCodes c;  // thread-safe one-time initialisation
        codes(c.lit, c.dist);
    }

    void dynamic() {
        static const uint8_t kOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
        const int nlen = (int)bits(5) + 257, ndist = (int)bits(5) + 1, ncode = (int)bits(4) + 4;
        if (nlen > 286 || ndist > 30) fail("deflate: bad dynamic counts");
        uint8_t lengths[320];
        memset(lengths, 0, sizeof(lengths));
        for (int k = 0; k < ncode; ++k) lengths[kOrder[k]] = (uint8_t)bits(3);
        Huffman lencode;
        if (build(&lencode, lengths, 19) != 0) fail("deflate: incomplete code-length code");
        int idx = 0;
        while (idx < nlen + ndist) {
            int sym = decode(lencode);
            if (sym < 16) {
                lengths[idx++] = (uint8_t)sym;
                continue;
            }
            uint8_t val = 0;
            int rep;
            if (sym == 16) {
                if (idx == 0) fail("deflate: repeat with no previous length");
                val = lengths[idx - 1];
                rep = 3 + (int)bits(2);
            } else if (sym == 17) {
                rep = 3 + (int)bits(3);
            } else {
                rep = 11 + (int)bits(7);
            }
            if (idx + rep > nlen + ndist) fail("deflate: too many lengths");
            while (rep--) lengths[idx++] = val;
        }
        if (lengths[256] == 0) fail("deflate: no end-of-block code");
        Huffman lit, dist;
        const int l = build(&lit, lengths, nlen);
        if (l < 0 || (l > 0 && nlen - lit.count[0] != 1)) fail("deflate: bad literal/length code");
        const int d = build(&dist, lengths + nlen, ndist);
        if (d < 0 || (d > 0 && ndist - dist.count[0] != 1)) fail("deflate: bad distance code");
        codes(lit, dist);
    }
};

struct CrcTable {
    uint32_t t[256];
    CrcTable() {
        for (uint32_t k = 0; k < 256; ++k) {
            uint32_t c = k;
            for (int j = 0; j < 8; ++j) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[k] = c;
        }
    }
};

uint32_t crc32_of(const uint8_t* p, size_t n) {
    static const CrcTable table;  // thread-safe one-time initialisation
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = table.t[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return c ^ 0xFFFFFFFFu;
}

uint32_t be32(const uint8_t* p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

int paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = p > a ? p - a : a - p;
    const int pb = p > b ? p - b : b - p;
    const int pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

struct Image {
    uint32_t w = 0, h = 0;
    std::vector<uint8_t> rgba;
};

Image decode_png(const uint8_t* data, size_t n) {
    static const uint8_t kSig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (n < 8 || memcmp(data, kSig, 8) != 0) fail("not a PNG file");
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    bool ended = false;
    while (!ended) {
        if (pos + 12 > n) fail("PNG: truncated chunk");
        const uint32_t len = be32(data + pos);
        if (len > n - pos - 12) fail("PNG: chunk past the end of the file");
        const uint8_t* type = data + pos + 4;
        const uint8_t* body = data + pos + 8;
        if (crc32_of(type, len + 4) != be32(body + len)) fail("PNG: chunk CRC mismatch");
        if (!memcmp(type, "IHDR", 4)) {
            if (len != 13) fail("PNG: bad IHDR");
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            if (body[10] != 0 || body[11] != 0) fail("PNG: unknown compression or filter method");
            interlace = body[12];
        } else if (!memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!memcmp(type, "IEND", 4)) {
            ended = true;
        } else if (!(type[0] & 0x20)) {
            fail("PNG: unknown critical chunk");
        }
        pos += 12 + (size_t)len;
    }
    if (w == 0 || h == 0 || w > 16384 || h > 16384) fail("PNG: image size out of range (1..16384)");
    if (interlace != 0) fail("PNG: interlaced images are not read");
    int channels;
    switch (ctype) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: fail("PNG: bad colour type");
    }
    const bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!depth_ok) fail("PNG: bad bit depth for the colour type");
    if (ctype == 3 && (plte.empty() || plte.size() % 3)) fail("PNG: palette image without a palette");
    // zlib wrapper
    if (idat.size() < 6) fail("PNG: no image data");
    const uint8_t cmf = idat[0], flg = idat[1];
    if ((cmf & 15) != 8 || ((cmf << 8) | flg) % 31 != 0 || (flg & 0x20)) fail("PNG: bad zlib header");
    const size_t stride = ((size_t)w * channels * depth + 7) / 8;
    Inflater inf(idat.data() + 2, idat.size() - 2);
    std::vector<uint8_t> raw = inf.run((stride + 1) * h);
    if (raw.size() < (stride + 1) * h) fail("PNG: image data too short");
    const size_t after = 2 + inf.consumed();
    if (after + 4 <= idat.size()) {
        uint32_t a = 1, b = 0;
        for (uint8_t v : raw) {
            a = (a + v) % 65521u;
            b = (b + a) % 65521u;
        }
        if (be32(idat.data() + after) != ((b << 16) | a)) fail("PNG: Adler-32 mismatch");
    }
    // unfilter in place
    const size_t bpp = ((size_t)channels * depth + 7) / 8;
    std::vector<uint8_t> prev(stride, 0), cur(stride);
    std::vector<uint8_t> pix((size_t)stride * h);
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* line = raw.data() + (stride + 1) * y;
        const int filter = line[0];
        for (size_t i = 0; i < stride; ++i) {
            const int x = line[1 + i];
            const int a = i >= bpp ? cur[i - bpp] : 0;
            const int b = prev[i];
            const int c = i >= bpp ? prev[i - bpp] : 0;
            int v;
            switch (filter) {
                case 0: v = x; break;
                case 1: v = x + a; break;
                case 2: v = x + b; break;
                case 3: v = x + ((a + b) >> 1); break;
                case 4: v = x + paeth(a, b, c); break;
                default: fail("PNG: bad filter type");
            }
            cur[i] = (uint8_t)v;
        }
        memcpy(pix.data() + stride * y, cur.data(), stride);
        std::swap(prev, cur);
    }
    // samples -> RGBA8
    Image img;
    img.w = w;
    img.h = h;
    img.rgba.resize((size_t)w * h * 4);
    const int maxv = (1 << depth) - 1;
    auto sample = [&](const uint8_t* row, size_t k) -> uint32_t {  // k-th sample of the row, native depth
        if (depth == 8) return row[k];
        if (depth == 16) return ((uint32_t)row[2 * k] << 8) | row[2 * k + 1];
        const size_t bit = k * depth;
        return (row[bit / 8] >> (8 - depth - (bit % 8))) & (uint32_t)maxv;
    };
    auto to8 = [&](uint32_t v) -> uint8_t {
        if (depth == 16) return (uint8_t)(v >> 8);
        if (depth == 8) return (uint8_t)v;
        return (uint8_t)(v * (255 / maxv));
    };
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t* row = pix.data() + stride * y;
        for (uint32_t x = 0; x < w; ++x) {
            uint8_t* o = img.rgba.data() + ((size_t)y * w + x) * 4;
            if (ctype == 3) {
                const uint32_t i = sample(row, x);
                if (3 * i + 2 >= plte.size()) fail("PNG: palette index out of range");
                o[0] = plte[3 * i];
                o[1] = plte[3 * i + 1];
                o[2] = plte[3 * i + 2];
                o[3] = i < trns.size() ? trns[i] : 255;
                continue;
            }
            uint32_t s[4];
            for (int c = 0; c < channels; ++c) s[c] = sample(row, (size_t)x * channels + c);
            if (ctype == 0 || ctype == 4) {
                o[0] = o[1] = o[2] = to8(s[0]);
                o[3] = ctype == 4 ? to8(s[1]) : 255;
                if (ctype == 0 && trns.size() >= 2 && s[0] == (((uint32_t)trns[0] << 8) | trns[1])) o[3] = 0;
            } else {
                o[0] = to8(s[0]);
                o[1] = to8(s[1]);
                o[2] = to8(s[2]);
                o[3] = ctype == 6 ? to8(s[3]) : 255;
                if (ctype == 2 && trns.size() >= 6 && s[0] == (((uint32_t)trns[0] << 8) | trns[1]) &&
                    s[1] == (((uint32_t)trns[2] << 8) | trns[3]) && s[2] == (((uint32_t)trns[4] << 8) | trns[5]))
                    o[3] = 0;
            }
        }
    }
    return img;
}

}  // namespace

extern "C" {

int xrh_png_decode(const uint8_t* data, size_t bytes, uint32_t* width, uint32_t* height, uint8_t* rgba8,
                   size_t rgba_bytes) {
    if (!data || !width || !height) return XR_ERR_INVALID_ARGUMENT;
    try {
        Image img = decode_png(data, bytes);
        *width = img.w;
        *height = img.h;
        if (!rgba8) return XR_OK;  // size query
        if (rgba_bytes < img.rgba.size()) return XR_ERR_INVALID_ARGUMENT;
        memcpy(rgba8, img.rgba.data(), img.rgba.size());
        return XR_OK;
    } catch (const DecodeError&) {
        return XR_ERR_INVALID_ARGUMENT;
    } catch (const std::bad_alloc&) {
        return XR_ERR_OUT_OF_MEMORY;
    }
}

}  // extern "C"

namespace xrh_internal {

// For the glTF loader: decodes to RGBA8 or returns false with a message.
bool png_to_rgba(const uint8_t* data, size_t bytes, uint32_t* w, uint32_t* h, std::vector<uint8_t>* rgba,
                 std::string* err) {
    try {
        Image img = decode_png(data, bytes);
        *w = img.w;
        *h = img.h;
        *rgba = std::move(img.rgba);
        return true;
    } catch (const DecodeError& e) {
        *err = e.msg;
        return false;
    }
}

}  // namespace xrh_internal

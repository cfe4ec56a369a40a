// glTF 2.0 ingestion into the AScene layout, as the reference loads its
// scenes: AglTFModel::loadglTFFile / loadMaterials / loadNode
// (src/Asset/Model.cpp:69-260, tinygltf's LoadASCIIFromFile underneath) and
// the concatenation of models in AScene::AScene (src/Asset/Scene.cpp:16-50).
//
// What is reproduced, quirks included:
//   * scene 0's root nodes, each loaded depth first with its children BEFORE
//     its own mesh (Model.cpp:120-126 precede :128);
//   * node matrix = T * mat4(R) * S built with glm's operation order, or the
//     node's `matrix` when present (which then wins, Model.cpp:103-118);
//     world matrix = root * ... * parent * node, multiplied in from the left
//     one parent at a time (:134-139);
//   * positions: world * vec4(p, 1), xyz kept (no divide);
//   * normals: transpose(inverse(world)) * vec4(normalize(n), 1), xyz kept, not
//     renormalised; a primitive without NORMAL gets normalize(0) = NaN
//     (:159-160);
//   * TEXCOORD_0 or (0, 0); materialIndex = the primitive's material (-1 when
//     absent), plus the model's material offset in the scene (Scene.cpp:44-47);
//   * attributes read as tightly packed float32 from accessor + view offsets
//     (byteStride is not consulted, :146-158); indices u8 / u16 / u32 plus the
//     primitive's first vertex within the model (:169-196);
//   * materials: baseColorFactor (4th component 1 when only 3 are given),
//     else (0, 0, 0, 0) as the value-initialised AMaterial; the base colour
//     texture's index, else -1; model DiffuseLight when the material is named
//     "light", else Lambertian (:69-94);
//   * one instance per file: (indexOffset, vertexOffset) into the
//     concatenated buffers (Scene.cpp:36-42).
// The matrix arithmetic follows glm 0.9.9.8 (3rd/glm): mat4 * vec4 as
// (m0 x + m1 y) + (m2 z + m3 w), mat4 * mat4 column by column left to right,
// translate / scale / mat3_cast / inverse as glm writes them, in float, with
// no contraction.
//
// Where the reference would read out of bounds or crash (no indices, an index
// type it does not handle, an accessor past its buffer, a non-float
// attribute), this loader returns an error instead.
//
// The matrix routines (inverse, mat * vec order) restate OpenGL Mathematics
// (GLM) 0.9.9.8 so the vertex arrays match the reference's bit for bit.  GLM
// is used under its MIT licence:
//   Copyright (c) 2005 - G-Truc Creation
//   Permission is hereby granted, free of charge, to any person obtaining a
//   copy of this software and associated documentation files (the
//   "Software"), to deal in the Software without restriction, including
//   without limitation the rights to use, copy, modify, merge, publish,
//   distribute, sublicense, and/or sell copies of the Software, and to permit
//   persons to whom the Software is furnished to do so, subject to the
//   following conditions: The above copyright notice and this permission
//   notice shall be included in all copies or substantial portions of the
//   Software.  THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND,
//   EXPRESS OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
//   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT.  IN
//   NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM,
//   DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR
//   OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE
//   USE OR OTHER DEALINGS IN THE SOFTWARE.
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "scene_internal.h"

namespace {

// ---------------------------------------------------------------- JSON ----
struct Json {
    enum Kind { kNull, kBool, kNumber, kString, kArray, kObject } kind = kNull;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> items;                          // array elements
    std::vector<std::pair<std::string, Json>> members;  // object members, file order

    const Json* get(const char* key) const {
        if (kind != kObject) return nullptr;
        for (const auto& m : members)
            if (m.first == key) return &m.second;
        return nullptr;
    }
};

struct ParseError {
    std::string msg;
};

class Parser {
public:
    Parser(const char* p, const char* end) : p_(p), end_(end) {}
    Json document() {
        Json v = value(0);
        ws();
        if (p_ != end_) fail("trailing characters after the JSON document");
        return v;
    }

private:
    const char* p_;
    const char* end_;

    [[noreturn]] void fail(const char* what) { throw ParseError{std::string("JSON: ") + what}; }
    void ws() {
        while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) ++p_;
    }
    bool lit(const char* s) {
        const size_t n = strlen(s);
        if ((size_t)(end_ - p_) >= n && memcmp(p_, s, n) == 0) {
            p_ += n;
            return true;
        }
        return false;
    }
    Json value(int depth) {
        if (depth > 256) fail("nesting deeper than 256");
        ws();
        if (p_ >= end_) fail("unexpected end of input");
        Json v;
        const char c = *p_;
        if (c == '{') {
            v.kind = Json::kObject;
            ++p_;
            ws();
            if (p_ < end_ && *p_ == '}') {
                ++p_;
                return v;
            }
            for (;;) {
                ws();
                if (p_ >= end_ || *p_ != '"') fail("object key expected");
                std::string key = string();
                ws();
                if (p_ >= end_ || *p_ != ':') fail("':' expected");
                ++p_;
                Json m = value(depth + 1);
                v.members.emplace_back(std::move(key), std::move(m));
                ws();
                if (p_ < end_ && *p_ == ',') { ++p_; continue; }
                if (p_ < end_ && *p_ == '}') { ++p_; return v; }
                fail("',' or '}' expected");
            }
        }
        if (c == '[') {
            v.kind = Json::kArray;
            ++p_;
            ws();
            if (p_ < end_ && *p_ == ']') {
                ++p_;
                return v;
            }
            for (;;) {
                v.items.push_back(value(depth + 1));
                ws();
                if (p_ < end_ && *p_ == ',') { ++p_; continue; }
                if (p_ < end_ && *p_ == ']') { ++p_; return v; }
                fail("',' or ']' expected");
            }
        }
        if (c == '"') {
            v.kind = Json::kString;
            v.str = string();
            return v;
        }
        if (lit("true")) { v.kind = Json::kBool; v.b = true; return v; }
        if (lit("false")) { v.kind = Json::kBool; v.b = false; return v; }
        if (lit("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) {
            // strtod on a bounded copy: correctly rounded to the nearest double
            const char* q = p_;
            if (q < end_ && *q == '-') ++q;
            while (q < end_ && ((*q >= '0' && *q <= '9') || *q == '.' || *q == 'e' || *q == 'E' || *q == '+' ||
                                *q == '-'))
                ++q;
            const std::string tok(p_, q);
            char* stop = nullptr;
            errno = 0;
            v.num = strtod(tok.c_str(), &stop);
            if (stop != tok.c_str() + tok.size()) fail("malformed number");
            v.kind = Json::kNumber;
            p_ = q;
            return v;
        }
        fail("unexpected character");
    }
    static void utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (end_ - p_ < 4) fail("short \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            const char h = *p_++;
            v <<= 4;
            if (h >= '0' && h <= '9') v |= (uint32_t)(h - '0');
            else if (h >= 'a' && h <= 'f') v |= (uint32_t)(h - 'a' + 10);
            else if (h >= 'A' && h <= 'F') v |= (uint32_t)(h - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string string() {
        ++p_;  // opening quote
        std::string out;
        while (p_ < end_ && *p_ != '"') {
            char c = *p_++;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p_ >= end_) break;
            c = *p_++;
            switch (c) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00 && end_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
                        p_ += 2;
                        const uint32_t lo = hex4();
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        if (p_ >= end_) fail("unterminated string");
        ++p_;  // closing quote
        return out;
    }
};

// ----------------------------------------------------- glm-order math ----
// Column-major 4x4: m[c][r].
struct M4 {
    float m[4][4];
};
struct V4 {
    float v[4];
};

M4 identity() {
    M4 r;
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) r.m[c][k] = c == k ? 1.0f : 0.0f;
    return r;
}

// glm operator*(mat4, vec4): (m[0]*v0 + m[1]*v1) + (m[2]*v2 + m[3]*v3).
V4 mul(const M4& a, const V4& x) {
    V4 r;
    for (int k = 0; k < 4; ++k) {
        const float add0 = a.m[0][k] * x.v[0] + a.m[1][k] * x.v[1];
        const float add1 = a.m[2][k] * x.v[2] + a.m[3][k] * x.v[3];
        r.v[k] = add0 + add1;
    }
    return r;
}

// glm operator*(mat4, mat4): column i = ((A0 b0 + A1 b1) + A2 b2) + A3 b3.
M4 mul(const M4& a, const M4& b) {
    M4 r;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k)
            r.m[i][k] = a.m[0][k] * b.m[i][0] + a.m[1][k] * b.m[i][1] + a.m[2][k] * b.m[i][2] + a.m[3][k] * b.m[i][3];
    return r;
}

// glm::translate(m, v): column 3 = ((m0 v0 + m1 v1) + m2 v2) + m3.
M4 translate(const M4& a, const float t[3]) {
    M4 r = a;
    for (int k = 0; k < 4; ++k) r.m[3][k] = a.m[0][k] * t[0] + a.m[1][k] * t[1] + a.m[2][k] * t[2] + a.m[3][k];
    return r;
}

// glm::scale(m, v): columns 0..2 times v, column 3 kept.
M4 scale(const M4& a, const float s[3]) {
    M4 r = a;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k < 4; ++k) r.m[c][k] = a.m[c][k] * s[c];
    return r;
}

// glm::mat4_cast(quat) (gtc/quaternion.inl mat3_cast), q = (x, y, z, w).
M4 rotation(const float q[4]) {
    const float x = q[0], y = q[1], z = q[2], w = q[3];
    const float qxx = x * x, qyy = y * y, qzz = z * z;
    const float qxz = x * z, qxy = x * y, qyz = y * z;
    const float qwx = w * x, qwy = w * y, qwz = w * z;
    M4 r = identity();
    r.m[0][0] = 1.0f - 2.0f * (qyy + qzz);
    r.m[0][1] = 2.0f * (qxy + qwz);
    r.m[0][2] = 2.0f * (qxz - qwy);
    r.m[1][0] = 2.0f * (qxy - qwz);
    r.m[1][1] = 1.0f - 2.0f * (qxx + qzz);
    r.m[1][2] = 2.0f * (qyz + qwx);
    r.m[2][0] = 2.0f * (qxz + qwy);
    r.m[2][1] = 2.0f * (qyz - qwx);
    r.m[2][2] = 1.0f - 2.0f * (qxx + qyy);
    return r;
}

// glm::inverse for mat4 (detail/func_matrix.inl compute_inverse<4, 4>):
// 2x2 cofactors, four cofactor columns with alternating signs, the
// determinant from the first row, then a multiply by its reciprocal.
M4 inverse(const M4& a) {
    const float(*m)[4] = a.m;
    const float c00 = m[2][2] * m[3][3] - m[3][2] * m[2][3];
    const float c02 = m[1][2] * m[3][3] - m[3][2] * m[1][3];
    const float c03 = m[1][2] * m[2][3] - m[2][2] * m[1][3];
    const float c04 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
    const float c06 = m[1][1] * m[3][3] - m[3][1] * m[1][3];
    const float c07 = m[1][1] * m[2][3] - m[2][1] * m[1][3];
    const float c08 = m[2][1] * m[3][2] - m[3][1] * m[2][2];
    const float c10 = m[1][1] * m[3][2] - m[3][1] * m[1][2];
    const float c11 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
    const float c12 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
    const float c14 = m[1][0] * m[3][3] - m[3][0] * m[1][3];
    const float c15 = m[1][0] * m[2][3] - m[2][0] * m[1][3];
    const float c16 = m[2][0] * m[3][2] - m[3][0] * m[2][2];
    const float c18 = m[1][0] * m[3][2] - m[3][0] * m[1][2];
    const float c19 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
    const float c20 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
    const float c22 = m[1][0] * m[3][1] - m[3][0] * m[1][1];
    const float c23 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
    const float f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const float v0[4] = {m[1][0], m[0][0], m[0][0], m[0][0]};
    const float v1[4] = {m[1][1], m[0][1], m[0][1], m[0][1]};
    const float v2[4] = {m[1][2], m[0][2], m[0][2], m[0][2]};
    const float v3[4] = {m[1][3], m[0][3], m[0][3], m[0][3]};
    M4 inv;
    for (int k = 0; k < 4; ++k) {
        const float sa = (k % 2 == 0) ? 1.0f : -1.0f;  // SignA (+,-,+,-); SignB is its negation
        // (a*b - c*d) + e*f, each Inv vector as glm writes it
        const float i0 = v1[k] * f0[k] - v2[k] * f1[k] + v3[k] * f2[k];
        const float i1 = v0[k] * f0[k] - v2[k] * f3[k] + v3[k] * f4[k];
        const float i2 = v0[k] * f1[k] - v1[k] * f3[k] + v3[k] * f5[k];
        const float i3 = v0[k] * f2[k] - v1[k] * f4[k] + v2[k] * f5[k];
        inv.m[0][k] = i0 * sa;
        inv.m[1][k] = i1 * -sa;
        inv.m[2][k] = i2 * sa;
        inv.m[3][k] = i3 * -sa;
    }
    // Row0 = (inv[0][0], inv[1][0], inv[2][0], inv[3][0]); Dot0 = m[0] * Row0
    const float d0 = m[0][0] * inv.m[0][0], d1 = m[0][1] * inv.m[1][0];
    const float d2 = m[0][2] * inv.m[2][0], d3 = m[0][3] * inv.m[3][0];
    const float det = (d0 + d1) + (d2 + d3);
    const float one_over = 1.0f / det;
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) inv.m[c][k] = inv.m[c][k] * one_over;
    return inv;
}

M4 transpose(const M4& a) {
    M4 r;
    for (int c = 0; c < 4; ++c)
        for (int k = 0; k < 4; ++k) r.m[c][k] = a.m[k][c];
    return r;
}

// glm::normalize(vec3) = v * (1 / sqrt((x*x + y*y) + z*z)).
void normalize3(const float in[3], float out[3]) {
    const float d = (in[0] * in[0] + in[1] * in[1]) + in[2] * in[2];
    const float s = 1.0f / sqrtf(d);
    for (int k = 0; k < 3; ++k) out[k] = in[k] * s;
}

// -------------------------------------------------------------- loader ----
struct LoadError {
    std::string msg;
};
[[noreturn]] void fail(const std::string& m) { throw LoadError{m}; }

std::string read_file(const std::string& path) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) fail("cannot open " + path);
    std::string data;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) data.append(buf, n);
    const bool bad = ferror(f) != 0;
    fclose(f);
    if (bad) fail("cannot read " + path);
    return data;
}

std::string percent_decode(const std::string& s) {
    std::string out;
    for (size_t i = 0; i < s.size(); ++i) {
        if (s[i] == '%' && i + 2 < s.size() && isxdigit((unsigned char)s[i + 1]) &&
            isxdigit((unsigned char)s[i + 2])) {
            out += (char)strtol(s.substr(i + 1, 2).c_str(), nullptr, 16);
            i += 2;
        } else {
            out += s[i];
        }
    }
    return out;
}

std::string base64_decode(const char* p, const char* end) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+') return 62;
        if (c == '/') return 63;
        return -1;
    };
    std::string out;
    uint32_t acc = 0;
    int bits = 0;
    for (; p < end; ++p) {
        if (*p == '=') break;
        const int v = val(*p);
        if (v < 0) fail("bad base64 data URI");
        acc = (acc << 6) | (uint32_t)v;
        bits += 6;
        if (bits >= 8) {
            bits -= 8;
            out += (char)((acc >> bits) & 0xFF);
        }
    }
    return out;
}

const Json& need(const Json* v, const std::string& what) {
    if (!v) fail("missing " + what);
    return *v;
}
double num(const Json* v, const std::string& what) {
    const Json& j = need(v, what);
    if (j.kind != Json::kNumber) fail(what + " is not a number");
    return j.num;
}
double num_or(const Json* v, double dflt, const std::string& what) { return v ? num(v, what) : dflt; }
size_t index_of(const Json* v, size_t bound, const std::string& what) {
    const double d = num(v, what);
    if (!(d >= 0.0) || d != floor(d) || d >= (double)bound) fail(what + " out of range");
    return (size_t)d;
}
// accessor.count as an element count whose byte size (count * elem_bytes)
// cannot overflow: a non-negative integer below 2^50 / elem_bytes (any larger
// count runs past every buffer, which data_of then reports).
size_t count_of(const Json& acc, size_t elem_bytes) {
    const double d = num(acc.get("count"), "accessor.count");
    if (!(d >= 0.0) || d != floor(d) || d > 1125899906842624.0 / (double)elem_bytes)
        fail("accessor.count out of range");
    return (size_t)d;
}
// A signed 32-bit integer field (material or texture index; -1 allowed).
int32_t int32_of(const Json* v, const std::string& what) {
    const double d = num(v, what);
    if (!(d >= -2147483648.0 && d <= 2147483647.0) || d != floor(d)) fail(what + " out of range");
    return (int32_t)d;
}
const Json& array_of(const Json& root, const char* key) {
    static const Json empty = [] {
        Json j;
        j.kind = Json::kArray;
        return j;
    }();
    const Json* a = root.get(key);
    if (!a) return empty;
    if (a->kind != Json::kArray) fail(std::string(key) + " is not an array");
    return *a;
}

struct Model {
    std::vector<xr_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<xr_material> materials;
    std::vector<xrh_texture_data> textures;
};

class GltfLoader {
public:
    GltfLoader(const std::string& path, bool untextured) : path_(path), untextured_(untextured) {}

    Model load() {
        const std::string text = read_file(path_);
        try {
            root_ = Parser(text.data(), text.data() + text.size()).document();
        } catch (const ParseError& e) {
            fail(path_ + ": " + e.msg);
        }
        if (root_.kind != Json::kObject) fail(path_ + ": top level is not an object");
        load_buffers();
        load_materials();
        if (!untextured_) load_textures();
        const Json& scenes = array_of(root_, "scenes");
        if (scenes.items.empty()) fail(path_ + ": no scenes");
        const Json& nodes = array_of(root_, "nodes");
        const Json& roots = array_of(scenes.items[0], "nodes");
        for (const Json& r : roots.items) {
            const size_t i = index_of(&r, nodes.items.size(), "scenes[0].nodes[]");
            load_node(nodes, i, nullptr, 0);
        }
        return std::move(model_);
    }

private:
    struct NodeRef {
        M4 matrix;
        const NodeRef* parent;
    };

    std::string path_;
    bool untextured_;
    Json root_;
    std::vector<std::string> buffers_;
    Model model_;

    std::string dir() const {
        const size_t s = path_.find_last_of('/');
        return s == std::string::npos ? std::string() : path_.substr(0, s + 1);
    }

    void load_buffers() {
        for (const Json& b : array_of(root_, "buffers").items) {
            const double len = num(b.get("byteLength"), "buffers[].byteLength");
            const Json* uri = b.get("uri");
            if (!uri || uri->kind != Json::kString) fail(path_ + ": buffer without a uri (GLB is not read)");
            std::string data;
            const std::string& u = uri->str;
            if (u.compare(0, 5, "data:") == 0) {
                const size_t comma = u.find(',');
                if (comma == std::string::npos || u.find(";base64") == std::string::npos || u.find(";base64") > comma)
                    fail(path_ + ": data URI is not base64");
                data = base64_decode(u.data() + comma + 1, u.data() + u.size());
            } else {
                data = read_file(dir() + percent_decode(u));
            }
            if ((double)data.size() < len) fail(path_ + ": buffer shorter than its byteLength");
            buffers_.push_back(std::move(data));
        }
    }

    void load_materials() {
        for (const Json& m : array_of(root_, "materials").items) {
            xr_material mat;
            memset(&mat, 0, sizeof(mat));
            mat.base_color_texture_index = -1;
            const Json* pbr = m.get("pbrMetallicRoughness");
            if (pbr && pbr->kind == Json::kObject) {
                const Json* f = pbr->get("baseColorFactor");
                if (f) {
                    if (f->kind != Json::kArray || f->items.size() < 3) fail(path_ + ": bad baseColorFactor");
                    for (int k = 0; k < 3; ++k) mat.base_color[k] = (float)num(&f->items[k], "baseColorFactor[]");
                    mat.base_color[3] = f->items.size() > 3 ? (float)num(&f->items[3], "baseColorFactor[]") : 1.0f;
                }
                const Json* t = pbr->get("baseColorTexture");
                if (t) {
                    const Json* idx = t->get("index");
                    mat.base_color_texture_index = idx ? int32_of(idx, "baseColorTexture.index") : -1;
                }
            }
            if (untextured_) mat.base_color_texture_index = -1;
            const Json* name = m.get("name");
            mat.material_model = (name && name->kind == Json::kString && name->str == "light")
                                     ? XR_MATERIAL_DIFFUSE_LIGHT
                                     : XR_MATERIAL_LAMBERTIAN;
            model_.materials.push_back(mat);
        }
    }

    // AglTFModel::loadImages / loadTextures (Model.cpp:19-67): texture i is
    // the image its `source` names, decoded to RGBA8.
    void load_textures() {
        const Json& images = array_of(root_, "images");
        std::vector<xrh_texture_data> decoded(images.items.size());
        std::vector<bool> done(images.items.size(), false);
        for (const Json& t : array_of(root_, "textures").items) {
            const size_t src = index_of(t.get("source"), images.items.size(), "textures[].source");
            if (!done[src]) {
                decoded[src] = decode_image(images.items[src], src);
                done[src] = true;
            }
            model_.textures.push_back(decoded[src]);
        }
    }

    xrh_texture_data decode_image(const Json& im, size_t index) {
        const std::string what = path_ + ": image " + std::to_string(index);
        std::string bytes;
        const Json* uri = im.get("uri");
        if (uri && uri->kind == Json::kString) {
            const std::string& u = uri->str;
            if (u.compare(0, 5, "data:") == 0) {
                const size_t comma = u.find(',');
                if (comma == std::string::npos || u.find(";base64") > comma) fail(what + ": data URI is not base64");
                bytes = base64_decode(u.data() + comma + 1, u.data() + u.size());
            } else {
                bytes = read_file(dir() + percent_decode(u));
            }
        } else if (const Json* bv = im.get("bufferView")) {
            const Json& views = array_of(root_, "bufferViews");
            const Json& view = views.items[index_of(bv, views.items.size(), "image.bufferView")];
            const size_t bi = index_of(view.get("buffer"), buffers_.size(), "bufferView.buffer");
            const double off = num_or(view.get("byteOffset"), 0.0, "bufferView.byteOffset");
            const double len = num(view.get("byteLength"), "bufferView.byteLength");
            if (off < 0.0 || len < 0.0 || off + len > (double)buffers_[bi].size()) fail(what + ": bufferView past its buffer");
            bytes = buffers_[bi].substr((size_t)off, (size_t)len);
        } else {
            fail(what + ": neither uri nor bufferView");
        }
        xrh_texture_data t;
        std::string err;
        if (!xrh_internal::png_to_rgba(reinterpret_cast<const uint8_t*>(bytes.data()), bytes.size(), &t.width,
                                       &t.height, &t.rgba8, &err))
            fail(what + ": " + err + " (PNG images only; XRH_GLTF_UNTEXTURED loads the scene without them)");
        return t;
    }

    // Float32 attribute data of an accessor: count * comps floats, tightly
    // packed from (view offset + accessor offset), as Model.cpp reads it.
    const float* attribute(size_t acc_index, int comps, size_t* count) {
        const Json& acc = array_of(root_, "accessors").items[acc_index];
        if (num(acc.get("componentType"), "accessor.componentType") != 5126.0)
            fail(path_ + ": vertex attribute is not float32");
        const size_t n = count_of(acc, comps * sizeof(float));
        const uint8_t* base = data_of(acc, n * comps * sizeof(float));
        *count = n;
        return reinterpret_cast<const float*>(base);
    }

    const uint8_t* data_of(const Json& acc, size_t bytes) {
        const Json& views = array_of(root_, "bufferViews");
        const size_t vi = index_of(acc.get("bufferView"), views.items.size(), "accessor.bufferView");
        const Json& view = views.items[vi];
        const size_t bi = index_of(view.get("buffer"), buffers_.size(), "bufferView.buffer");
        const double off = num_or(acc.get("byteOffset"), 0.0, "accessor.byteOffset") +
                           num_or(view.get("byteOffset"), 0.0, "bufferView.byteOffset");
        const std::string& buf = buffers_[bi];
        if (off < 0.0 || off + (double)bytes > (double)buf.size()) fail(path_ + ": accessor runs past its buffer");
        return reinterpret_cast<const uint8_t*>(buf.data()) + (size_t)off;
    }

    void load_node(const Json& nodes, size_t index, const NodeRef* parent, int depth) {
        if (depth > 64) fail(path_ + ": node hierarchy deeper than 64 (or cyclic)");
        const Json& in = nodes.items[index];
        NodeRef node;
        node.parent = parent;
        node.matrix = identity();
        auto vec = [&](const char* key, size_t n, float* out) -> bool {
            const Json* a = in.get(key);
            if (!a || a->kind != Json::kArray || a->items.size() != n) return false;
            for (size_t k = 0; k < n; ++k) out[k] = (float)num(&a->items[k], std::string("node.") + key);
            return true;
        };
        float t[3], r[4], s[3], mat[16];
        if (vec("translation", 3, t)) node.matrix = translate(node.matrix, t);
        if (vec("rotation", 4, r)) node.matrix = mul(node.matrix, rotation(r));
        if (vec("scale", 3, s)) node.matrix = scale(node.matrix, s);
        if (vec("matrix", 16, mat)) {
            for (int c = 0; c < 4; ++c)
                for (int k = 0; k < 4; ++k) node.matrix.m[c][k] = mat[4 * c + k];
        }
        for (const Json& c : array_of(in, "children").items)
            load_node(nodes, index_of(&c, nodes.items.size(), "node.children[]"), &node, depth + 1);
        const Json* mesh = in.get("mesh");
        if (!mesh) return;
        const Json& meshes = array_of(root_, "meshes");
        const Json& m = meshes.items[index_of(mesh, meshes.items.size(), "node.mesh")];
        M4 world = node.matrix;
        for (const NodeRef* p = node.parent; p; p = p->parent) world = mul(p->matrix, world);
        const M4 nrm = transpose(inverse(world));
        const size_t n_acc = array_of(root_, "accessors").items.size();
        for (const Json& prim : array_of(m, "primitives").items) {
            const uint32_t vertex_start = (uint32_t)model_.vertices.size();
            const Json* attrs = prim.get("attributes");
            if (!attrs || attrs->kind != Json::kObject) fail(path_ + ": primitive without attributes");
            const Json* pa = attrs->get("POSITION");
            if (!pa) fail(path_ + ": primitive without POSITION");
            size_t count = 0, nn = 0, nt = 0;
            const float* pos = attribute(index_of(pa, n_acc, "POSITION"), 3, &count);
            const float* nrmv = nullptr;
            const float* uv = nullptr;
            if (const Json* na = attrs->get("NORMAL")) {
                nrmv = attribute(index_of(na, n_acc, "NORMAL"), 3, &nn);
                if (nn < count) fail(path_ + ": NORMAL shorter than POSITION");
            }
            if (const Json* ta = attrs->get("TEXCOORD_0")) {
                uv = attribute(index_of(ta, n_acc, "TEXCOORD_0"), 2, &nt);
                if (nt < count) fail(path_ + ": TEXCOORD_0 shorter than POSITION");
            }
            const Json* pm = prim.get("material");
            const int32_t material = pm ? int32_of(pm, "primitive.material") : -1;
            for (size_t v = 0; v < count; ++v) {
                xr_vertex out;
                float p[3], n3[3] = {0.0f, 0.0f, 0.0f};
                memcpy(p, pos + 3 * v, sizeof(p));
                if (nrmv) memcpy(n3, nrmv + 3 * v, sizeof(n3));
                const V4 wp = mul(world, V4{{p[0], p[1], p[2], 1.0f}});
                float nn3[3];
                normalize3(n3, nn3);
                const V4 wn = mul(nrm, V4{{nn3[0], nn3[1], nn3[2], 1.0f}});
                for (int k = 0; k < 3; ++k) {
                    out.pos[k] = wp.v[k];
                    out.normal[k] = wn.v[k];
                }
                out.tex_coord[0] = uv ? uv[2 * v] : 0.0f;
                out.tex_coord[1] = uv ? uv[2 * v + 1] : 0.0f;
                out.material_index = material;
                model_.vertices.push_back(out);
            }
            const Json* pi = prim.get("indices");
            if (!pi) fail(path_ + ": primitive without indices");
            const Json& acc = array_of(root_, "accessors").items[index_of(pi, n_acc, "primitive.indices")];
            const double type = num(acc.get("componentType"), "accessor.componentType");
            const size_t n = count_of(acc, 4);
            if (type == 5125.0) {
                const uint8_t* b = data_of(acc, n * 4);
                for (size_t k = 0; k < n; ++k) {
                    uint32_t x;
                    memcpy(&x, b + 4 * k, 4);
                    model_.indices.push_back(x + vertex_start);
                }
            } else if (type == 5123.0) {
                const uint8_t* b = data_of(acc, n * 2);
                for (size_t k = 0; k < n; ++k) {
                    uint16_t x;
                    memcpy(&x, b + 2 * k, 2);
                    model_.indices.push_back((uint32_t)x + vertex_start);
                }
            } else if (type == 5121.0) {
                const uint8_t* b = data_of(acc, n);
                for (size_t k = 0; k < n; ++k) model_.indices.push_back((uint32_t)b[k] + vertex_start);
            } else {
                fail(path_ + ": index component type not supported");
            }
        }
    }
};

void set_err(char* err, size_t len, const std::string& m) {
    if (err && len) {
        strncpy(err, m.c_str(), len - 1);
        err[len - 1] = 0;
    }
}

}  // namespace

extern "C" {

int xrh_scene_gltf(const char* const* paths, uint32_t count, uint32_t flags, xrh_scene** out, char* err,
                   size_t err_len) {
    if (err && err_len) err[0] = 0;
    if (!paths || !out || count == 0) {
        set_err(err, err_len, "paths, count and out are required");
        return XR_ERR_INVALID_ARGUMENT;
    }
    if (flags & ~(uint32_t)XRH_GLTF_UNTEXTURED) {
        set_err(err, err_len, "unknown flags");
        return XR_ERR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    std::unique_ptr<xrh_scene> s;
    try {
        s.reset(new xrh_scene);
        for (uint32_t i = 0; i < count; ++i) {
            if (!paths[i]) fail("null path");
            Model m = GltfLoader(paths[i], (flags & XRH_GLTF_UNTEXTURED) != 0).load();
            const uint32_t index_offset = (uint32_t)s->indices.size();
            const uint32_t vertex_offset = (uint32_t)s->vertices.size();
            const int32_t material_offset = (int32_t)s->materials.size();
            s->offsets.push_back(index_offset);
            s->offsets.push_back(vertex_offset);
            for (xr_vertex& v : m.vertices) v.material_index += material_offset;
            s->vertices.insert(s->vertices.end(), m.vertices.begin(), m.vertices.end());
            s->indices.insert(s->indices.end(), m.indices.begin(), m.indices.end());
            s->materials.insert(s->materials.end(), m.materials.begin(), m.materials.end());
            // textures in file order, indices not offset (Scene.cpp:56-63)
            for (auto& t : m.textures) s->textures.push_back(std::move(t));
        }
        for (const auto& t : s->textures) s->texture_desc.push_back(xr_texture{t.width, t.height, t.rgba8.data()});
    } catch (const LoadError& e) {
        set_err(err, err_len, e.msg);
        return XR_ERR_INVALID_ARGUMENT;
    } catch (const std::bad_alloc&) {
        set_err(err, err_len, "out of host memory");
        return XR_ERR_OUT_OF_MEMORY;
    }
    *out = s.release();
    return XR_OK;
}

}  // extern "C"

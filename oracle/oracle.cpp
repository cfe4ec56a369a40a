// oracle.cpp -- CPU restatement of the reference's path-tracing / ReSTIR GI
// shaders.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the bench's
// CPU-baseline leg call it for reference results; the HIP product never links or calls
// it.
//
// The reference (0Setsuna0/XVK_Ray) cannot be built here: it is Vulkan +
// HLSL compiled by dxc and needs a window surface and a discrete GPU, none of
// which this image has.  This file therefore restates each shader function in
// C++, in the shaders' own shape (AoS structs, the same helper names), and
// evaluates every expression in the order the HLSL writes it.  Built with
// -ffp-contract=off and IEEE FP32 division/sqrt, so where the HLSL leaves a
// choice (mul() summation order, cos/sin) the choice is stated below and the
// GPU code must make the same one to match bit for bit.
//
// Source map (file:line in the reference):
//   InitRandomSeed/NextRandom/RandomFloat   shader/HLSL/random.hlsl:1-20
//   BranchlessONB, SampleHemisphere,        shader/HLSL/ReSTIRGI/Sampling.hlsl:4-56
//     WorldToScreenUV
//   SampleLambertian, SampleBSDF            shader/HLSL/ReSTIRGI/BSDF.hlsl:7-57
//   Update/MergeReservoir, p_hat,           shader/HLSL/ReSTIRGI/ReSTIRGICommon.hlsl:32-131
//     IsSimilar, ACESFilm
//   closest hit / miss                      shader/HLSL/ReSTIRGI/Ray.rchit.hlsl:26-61, Ray.rmiss.hlsl:4-8
//   initial sample pass                     shader/HLSL/ReSTIRGI/ReSTIRGIInitialSample.rgen.hlsl:35-214
//   temporal reuse pass                     shader/HLSL/ReSTIRGI/ReSTIRGITemporalReuse.rgen.hlsl:33-151
//   spatial reuse pass                      shader/HLSL/ReSTIRGI/ReSTIRGISpatialReuse.rgen.hlsl:35-183
//   plain path tracer                       shader/HLSL/rgen.hlsl:31-151, rhit.hlsl:29-59, sampling.hlsl:11-98
//   binding aliasing (prev == current       src/Vulkan/RayTracing/ReSTIRGIPipeline.cpp:141-199,
//     G-buffer, CheckVisibility true)         ApplicationReSTIRGI.cpp:367-372
//
// Choices where the HLSL leaves the arithmetic to the implementation:
//   * mul(M, v) for a column-major glm matrix: r[row] = sum over columns c of
//     M[c][row]*v[c], summed c = 0..3 left to right.
//   * cos/sin: Cody-Waite reduction by pi/2 (three FMA steps) and the Cephes
//     single-precision minimax polynomials (S. Moshier, Cephes sinf.c).
//   * normalize(v) = v * (1/sqrt(dot(v,v))), rsqrt(x) = 1/sqrt(x).
//   * TraceRay: the nearest triangle by the watertight test of Woop, Benthin
//     and Wald (JCGT 2013) in FP32, ties to the smaller canonical triangle id
//     (FP32 mirror).  A separate FP64 closest-hit search is the exact reference.
//   * float -> UNORM8: round(clamp(x) * 255).
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// HLSL-like vector types
// ---------------------------------------------------------------------------
struct float2 {
    float x, y;
};
struct float3 {
    float x, y, z;
};
struct float4 {
    float x, y, z, w;
};
inline float3 F3(float x, float y, float z) { return float3{x, y, z}; }
inline float3 operator+(float3 a, float3 b) { return F3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline float3 operator-(float3 a, float3 b) { return F3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline float3 operator*(float3 a, float3 b) { return F3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline float3 operator*(float3 a, float s) { return F3(a.x * s, a.y * s, a.z * s); }
inline float3 operator/(float3 a, float s) { return F3(a.x / s, a.y / s, a.z / s); }
inline float3 operator-(float3 a) { return F3(-a.x, -a.y, -a.z); }
inline float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(float3 a) { return sqrtf(dot(a, a)); }
inline float3 normalize(float3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
inline float rsqrt_(float x) { return 1.0f / sqrtf(x); }
inline float3 cross(float3 a, float3 b) {
    return F3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float hmax(float a, float b) { return a > b ? a : b; }
inline float hmin(float a, float b) { return a < b ? a : b; }
inline float hclamp(float x, float lo, float hi) { return hmin(hmax(x, lo), hi); }

const float PI = 3.14159265359f;
const float TWO_PI = 6.28318530718f;
const float INV_PI = (float)(1. / 3.14159265359);

// mul(M, v), M column-major (glm), rows summed over columns 0..3.
inline float4 mulv(const float* M, float4 v) {
    float r[4];
    for (int row = 0; row < 4; ++row) r[row] = M[row] * v.x + M[4 + row] * v.y + M[8 + row] * v.z + M[12 + row] * v.w;
    return float4{r[0], r[1], r[2], r[3]};
}

// cos/sin (see header).
void sincos_cephes(float x, float* s_out, float* c_out) {
    const float two_over_pi = 0.636619772367581343f;
    const float p1 = 1.5707963705062866211e+00f;
    const float p2 = -4.3711388286737928865e-08f;
    const float p3 = -1.7151245100059184123e-15f;
    float k = rintf(x * two_over_pi);
    float r = fmaf(-k, p1, x);
    r = fmaf(-k, p2, r);
    r = fmaf(-k, p3, r);
    float z = r * r;
    float sp = fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f);
    float sn = fmaf(sp * z, r, r);
    float cp = fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f);
    float cs = fmaf(cp * z, z, fmaf(-0.5f, z, 1.0f));
    switch (((int)k) & 3) {
        case 0: *s_out = sn; *c_out = cs; break;
        case 1: *s_out = cs; *c_out = -sn; break;
        case 2: *s_out = -sn; *c_out = -cs; break;
        default: *s_out = -cs; *c_out = sn; break;
    }
}
inline float hcos(float x) { float s, c; sincos_cephes(x, &s, &c); return c; }
inline float hsin(float x) { float s, c; sincos_cephes(x, &s, &c); return s; }

// ---------------------------------------------------------------------------
// random.hlsl:1-20
// ---------------------------------------------------------------------------
uint32_t InitRandomSeed(uint32_t px, uint32_t py, uint32_t resx, uint32_t frameNumber) {
    uint32_t seed = px * 1u + py * resx;  // dot(pixelCoords, uint2(1, resolution.x))
    seed = (seed ^ 61u) ^ (seed >> 16);
    seed *= 9u;
    seed = seed ^ (seed >> 4);
    seed *= 0x27d4eb2du;
    seed = seed ^ (seed >> 15);
    return seed * (frameNumber + 1u);
}
uint32_t NextRandom(uint32_t& state) {
    state = state * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
float RandomFloat(uint32_t& state) { return float(NextRandom(state) & 0x00FFFFFFu) / 16777216.0f; }

// ---------------------------------------------------------------------------
// Reference structs (host mirrors, ReSTIRGICommon.h:8-32, Vertex.h, Material.h,
// UniformBuffer.h)
// ---------------------------------------------------------------------------
struct Vertex {
    float Position[3];
    float Normal[3];
    float TexCoord[2];
    int32_t MaterialIndex;
};
struct BSDFMaterial {
    float baseColor[4];
    int32_t baseColorTextureIndex;
    float fuzziness;
    float refractionIndex;
    uint32_t materialModel;
};
enum { BSDFLambertian = 0, BSDFDiffuseLight = 4 };
struct UBO {
    float modelView[16], viewPrev[16], projection[16], projectionPrev[16], modelViewInverse[16],
        projectionInverse[16];
    uint32_t spp, totalNumberOfSamples, numberOfBounces, randomSeed, hasSkyBox, reuse, currentFrame;
};
static_assert(sizeof(UBO) == 412, "UBO");
struct ReSTIRGISample {
    float3 x_view;
    uint32_t materialID;
    float3 n_view;
    uint32_t randSeed;
    float3 x_sample;
    float p_q;
    float3 n_sample;
    uint32_t pad1;
    float3 L_out;
    uint32_t pad2;
    float3 f;
    uint32_t pad3;
};
struct ReSTIRGIReservoir {
    float w_sum;
    float W;
    uint32_t m;
    uint32_t pad;
    ReSTIRGISample s;
};
static_assert(sizeof(ReSTIRGISample) == 96 && sizeof(ReSTIRGIReservoir) == 112, "AoS");

// InitSample / InitReservoir (ReSTIRGICommon.hlsl:32-50).  p_q and f are left
// unset by the HLSL (out parameters); this restatement zeroes them.
void InitSample(ReSTIRGISample& s) {
    memset(&s, 0, sizeof(s));
    s.materialID = 0xFFFFFFFFu;
    s.randSeed = 0;
}
void InitReservoir(ReSTIRGIReservoir& r) {
    r.w_sum = 0;
    r.W = 0;
    r.m = 0;
    r.pad = 0;
    InitSample(r.s);
}
void UpdateReservoir(ReSTIRGIReservoir& r, const ReSTIRGISample& s, float w_new, uint32_t& seed) {
    r.w_sum += w_new;
    r.m++;
    if (RandomFloat(seed) < w_new / hmax(1e-10f, r.w_sum)) r.s = s;
}
void MergeReservoir(ReSTIRGIReservoir& target, const ReSTIRGIReservoir& source, float p_hat_, uint32_t& seed) {
    uint32_t M0 = target.m;
    float w_new = p_hat_ * source.W * float(source.m);
    UpdateReservoir(target, source.s, w_new, seed);
    target.m = M0 + source.m;
}
float p_hat(float3 f) { return length(f); }
bool IsSimilar(const ReSTIRGISample& current, const ReSTIRGISample& history) {
    float3 d = current.x_view - history.x_view;
    float dist2 = dot(d, d);
    if (dist2 > 0.1f) return false;
    if (dot(current.n_view, history.n_view) < 0.8f) return false;
    if (current.materialID != history.materialID) return false;
    return true;
}
float ACES1(float x) {
    float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return hclamp((x * (a * x + b)) / (x * (c * x + d) + e), 0.0f, 1.0f);
}
float3 ACESFilm(float3 x) { return F3(ACES1(x.x), ACES1(x.y), ACES1(x.z)); }
uint32_t unorm8(float x) { return (uint32_t)(hclamp(x, 0.0f, 1.0f) * 255.0f + 0.5f); }
// OutputImage[id] = float4(color, a) on a B8G8R8A8_UNORM image, after
// ACES and pow(1/2.2).
uint32_t store_bgra8(float3 c, float a) {
    float3 t = ACESFilm(c);
    const float g = (float)(1.0 / 2.2);
    float r = powf(t.x, g), gg = powf(t.y, g), b = powf(t.z, g);
    return unorm8(b) | (unorm8(gg) << 8) | (unorm8(r) << 16) | (unorm8(a) << 24);
}
// RG16F store (round to nearest even) and load.
uint16_t to_half(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u, ex = (x >> 23) & 0xFFu, man = x & 0x7FFFFFu;
    if (ex == 0xFFu) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u : 0u));
    int e = (int)ex - 112;
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        uint32_t sh = (uint32_t)(14 - e), h = man >> sh, rem = man & ((1u << sh) - 1u), mid = 1u << (sh - 1);
        if (rem > mid || (rem == mid && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = sign | ((uint32_t)e << 10) | (man >> 13), rem = man & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)h;
}
float from_half(uint16_t h) {
    // exact: every half is a float; compute by scaling
    int ex = (h >> 10) & 0x1F, man = h & 0x3FF;
    float v;
    if (ex == 0) v = ldexpf((float)man, -24);
    else if (ex == 31) v = man ? std::numeric_limits<float>::quiet_NaN() : std::numeric_limits<float>::infinity();
    else v = ldexpf((float)(man | 0x400), ex - 25);
    return (h & 0x8000) ? -v : v;
}

// ---------------------------------------------------------------------------
// Sampling.hlsl / BSDF.hlsl
// ---------------------------------------------------------------------------
void BranchlessONB(float3 n, float3& b1, float3& b2) {
    float sign = n.z >= 0.0f ? 1.0f : -1.0f;
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    b1 = F3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    b2 = F3(b, sign + n.y * n.y * a, -n.y);
}
float3 to_world(float3 v, float3 T, float3 B, float3 N) { return T * v.x + B * v.y + N * v.z; }
float3 SampleHemisphere(float2 xi) {
    float phi = TWO_PI * xi.x;
    float cos_theta = sqrtf(1.0f - xi.y);
    float sin_theta = sqrtf(xi.y);
    return F3(sin_theta * hcos(phi), sin_theta * hsin(phi), cos_theta);
}
// TextureSamplers[] (binding 8) with LinearSampler (BSDF.hlsl:4-5, 22):
// R8G8B8A8_UNORM images (TextureImage.cpp:25) sampled at level 0 with a
// LINEAR, CLAMP_TO_EDGE sampler (XVKSampler.h:9-26).  The texture-space
// position, the two taps per axis, their clamp to the edge and the weights
// (1-a)(1-b), a(1-b), (1-a)b, ab follow the Vulkan specification's linear
// filtering; the FP32 sum order and the clamp of x to [-1, width] before the
// floor (beyond it both taps are the edge texel) are this restatement's
// definition, which the device follows.  Parity with the reference's sampler
// hardware, which rounds its weights to a fixed-point grid, is unpinned.
struct TextureSet {
    std::vector<uint32_t> info;    // (first texel, width, height) per texture
    std::vector<uint8_t> texels;   // RGBA8
    uint32_t count() const { return (uint32_t)(info.size() / 3); }
};
float unorm8(uint8_t b) { return (float)b / 255.0f; }
float3 TextureSample(const TextureSet& ts, int t, float2 uv) {
    if ((uint32_t)t >= ts.count()) return F3(0, 0, 0);  // not reached: the device refuses such scenes
    const uint32_t first = ts.info[3 * t], w = ts.info[3 * t + 1], h = ts.info[3 * t + 2];
    const float x = fminf(fmaxf(uv.x * (float)w - 0.5f, -1.0f), (float)w);
    const float y = fminf(fmaxf(uv.y * (float)h - 0.5f, -1.0f), (float)h);
    const float fx = floorf(x), fy = floorf(y);
    const float a = x - fx, b = y - fy;
    const int i0 = (int)fx, j0 = (int)fy;
    auto cl = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
    const int x0 = cl(i0, (int)w - 1), x1 = cl(i0 + 1, (int)w - 1);
    const int y0 = cl(j0, (int)h - 1), y1 = cl(j0 + 1, (int)h - 1);
    auto tex = [&](int xx, int yy, int c) { return unorm8(ts.texels[4 * ((size_t)first + (size_t)yy * w + xx) + c]); };
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b), w01 = (1.0f - a) * b, w11 = a * b;
    float r[3];
    for (int c = 0; c < 3; ++c)
        r[c] = ((w00 * tex(x0, y0, c) + w10 * tex(x1, y0, c)) + w01 * tex(x0, y1, c)) + w11 * tex(x1, y1, c);
    return F3(r[0], r[1], r[2]);
}
// (mat.baseColorTextureIndex >= 0) ? texColor.rgb : mat.baseColor.rgb (BSDF.hlsl:22-23), the
// Lambertian albedo; emitters never reach it (the bounce loops break before SampleBSDF)
float3 MaterialColor(const TextureSet& ts, const BSDFMaterial& mat, float2 uv) {
    if (mat.baseColorTextureIndex >= 0) return TextureSample(ts, mat.baseColorTextureIndex, uv);
    return F3(mat.baseColor[0], mat.baseColor[1], mat.baseColor[2]);
}

float3 SampleLambertian(const TextureSet& ts, const BSDFMaterial& mat, float2 uv, float3& wi, float& pdf,
                        float& cos_theta, uint32_t& seed) {
    float x0 = RandomFloat(seed);
    float x1 = RandomFloat(seed);
    wi = SampleHemisphere(float2{x0, x1});
    cos_theta = wi.z;
    pdf = cos_theta * INV_PI;
    if (cos_theta <= 0) {
        pdf = 0;
        return F3(0, 0, 0);
    }
    float3 albedo = MaterialColor(ts, mat, uv);
    return albedo * INV_PI;
}
float3 SampleBSDF(const TextureSet& ts, float3 n, float2 uv, const BSDFMaterial& mat, float3& wi_world,
                  float& pdf_w, float& cos_theta, uint32_t& seed) {
    float3 f = F3(0, 0, 0);
    pdf_w = 0;
    cos_theta = 0;
    float3 T, B;
    BranchlessONB(n, T, B);
    float3 wi_local = F3(0, 0, 0);
    if (mat.materialModel == BSDFLambertian) f = SampleLambertian(ts, mat, uv, wi_local, pdf_w, cos_theta, seed);
    wi_world = to_world(wi_local, T, B, n);
    return f;
}
float2 WorldToScreenUV(float3 p, const float* view, const float* proj) {
    float4 v = mulv(view, float4{p.x, p.y, p.z, 1.0f});
    float4 c = mulv(proj, v);
    float nx = c.x / c.w, ny = c.y / c.w;
    return float2{nx * 0.5f + 0.5f, ny * 0.5f + 0.5f};
}

// ---------------------------------------------------------------------------
// Scene and TraceRay
// ---------------------------------------------------------------------------
struct Hit {
    float t;
    float u, v;
    uint32_t id, inst, prim;
    bool found;
};

struct Scene {
    std::vector<Vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<BSDFMaterial> materials;
    TextureSet textures;
    std::vector<uint32_t> offsets;      // 2 per instance
    std::vector<uint32_t> first;        // first canonical triangle per instance (+ total)
    std::vector<float> tv;              // 9 floats per canonical triangle (v0, v1, v2)
    std::vector<uint32_t> tinst, tprim;
    // BVH (double boxes) over canonical triangles
    struct Node {
        double lo[3], hi[3];
        int left, right;  // children, or -1
        int begin, end;   // triangle range into order[] when leaf
    };
    std::vector<Node> nodes;
    std::vector<uint32_t> order;
    // Box padding: the FP32 watertight test may report a hit a few ulps
    // outside a triangle's exact box, so every box is widened by
    // max(2^-16 * extent, 2^-20 * max|coordinate|) (far above that error for
    // origins inside the scene); brute force needs no boxes at all.
    double pad = 0.0;
};

void tri_verts(const Scene& s, uint32_t g, float3* p) {
    const float* q = &s.tv[9 * (size_t)g];
    p[0] = F3(q[0], q[1], q[2]);
    p[1] = F3(q[3], q[4], q[5]);
    p[2] = F3(q[6], q[7], q[8]);
}

int build_node(Scene& s, std::vector<double>& cent, int begin, int end) {
    Scene::Node nd;
    for (int a = 0; a < 3; ++a) { nd.lo[a] = 1e300; nd.hi[a] = -1e300; }
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    for (int i = begin; i < end; ++i) {
        uint32_t g = s.order[i];
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) {
                double v = s.tv[9 * (size_t)g + 3 * k + a];
                nd.lo[a] = std::min(nd.lo[a], v);
                nd.hi[a] = std::max(nd.hi[a], v);
            }
        for (int a = 0; a < 3; ++a) {
            clo[a] = std::min(clo[a], cent[3 * (size_t)g + a]);
            chi[a] = std::max(chi[a], cent[3 * (size_t)g + a]);
        }
    }
    for (int a = 0; a < 3; ++a) {
        nd.lo[a] -= s.pad;
        nd.hi[a] += s.pad;
    }
    nd.left = nd.right = -1;
    nd.begin = begin;
    nd.end = end;
    int idx = (int)s.nodes.size();
    s.nodes.push_back(nd);
    if (end - begin <= 4) return idx;
    int axis = 0;
    for (int a = 1; a < 3; ++a)
        if (chi[a] - clo[a] > chi[axis] - clo[axis]) axis = a;
    int mid = (begin + end) / 2;
    std::nth_element(s.order.begin() + begin, s.order.begin() + mid, s.order.begin() + end,
                     [&](uint32_t a, uint32_t b) { return cent[3 * (size_t)a + axis] < cent[3 * (size_t)b + axis]; });
    int l = build_node(s, cent, begin, mid);
    int r = build_node(s, cent, mid, end);
    s.nodes[idx].left = l;
    s.nodes[idx].right = r;
    return idx;
}

// Conservative FP64 box test: boxes are exact (vertex coordinates are FP32
// values), the interval is widened by 1e-9 relative + 1e-12 absolute.
bool box_hit(const Scene::Node& nd, const double* o, const double* d, double tmin, double tmax) {
    double tn = tmin, tf = tmax;
    for (int a = 0; a < 3; ++a) {
        if (d[a] == 0.0) {
            if (o[a] < nd.lo[a] - 1e-12 || o[a] > nd.hi[a] + 1e-12) return false;
            continue;
        }
        double t0 = (nd.lo[a] - o[a]) / d[a], t1 = (nd.hi[a] - o[a]) / d[a];
        if (t0 > t1) std::swap(t0, t1);
        tn = std::max(tn, t0 - 1e-9 * fabs(t0) - 1e-12);
        tf = std::min(tf, t1 + 1e-9 * fabs(t1) + 1e-12);
        if (tn > tf) return false;
    }
    return true;
}

// Watertight ray/triangle test in FP32 (Woop, Benthin, Wald 2013).
struct RayPre {
    float3 org;
    int kx, ky, kz;
    float Sx, Sy, Sz;
};
inline float cmp(float3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }
RayPre prepare(float3 org, float3 dir) {
    RayPre r;
    r.org = org;
    float ax = fabsf(dir.x), ay = fabsf(dir.y), az = fabsf(dir.z);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (cmp(dir, kz) < 0.0f) std::swap(kx, ky);
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = cmp(dir, kx) / cmp(dir, kz);
    r.Sy = cmp(dir, ky) / cmp(dir, kz);
    r.Sz = 1.0f / cmp(dir, kz);
    return r;
}
bool tri_fp32(const RayPre& r, const float3* p, float tmin, float tmax, float* t_out, float* u_out, float* v_out) {
    float3 A = p[0] - r.org, B = p[1] - r.org, C = p[2] - r.org;
    float Ax = fmaf(-r.Sx, cmp(A, r.kz), cmp(A, r.kx)), Ay = fmaf(-r.Sy, cmp(A, r.kz), cmp(A, r.ky));
    float Bx = fmaf(-r.Sx, cmp(B, r.kz), cmp(B, r.kx)), By = fmaf(-r.Sy, cmp(B, r.kz), cmp(B, r.ky));
    float Cx = fmaf(-r.Sx, cmp(C, r.kz), cmp(C, r.kx)), Cy = fmaf(-r.Sy, cmp(C, r.kz), cmp(C, r.ky));
    float U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float T = U * (r.Sz * cmp(A, r.kz)) + V * (r.Sz * cmp(B, r.kz)) + W * (r.Sz * cmp(C, r.kz));
    float t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    *t_out = t;
    *u_out = V / det;
    *v_out = W / det;
    return true;
}
// Watertight test in FP64 (same algorithm, all double): the exact reference.
bool tri_fp64(const double* o, const double* d, const float3* p, double tmin, double tmax, double* t_out) {
    double ad[3] = {fabs(d[0]), fabs(d[1]), fabs(d[2])};
    int kz = (ad[0] > ad[1]) ? ((ad[0] > ad[2]) ? 0 : 2) : ((ad[1] > ad[2]) ? 1 : 2);
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (d[kz] < 0) std::swap(kx, ky);
    double Sx = d[kx] / d[kz], Sy = d[ky] / d[kz], Sz = 1.0 / d[kz];
    double P[3][3];
    for (int k = 0; k < 3; ++k) {
        double q[3] = {(double)cmp(p[k], 0) - o[0], (double)cmp(p[k], 1) - o[1], (double)cmp(p[k], 2) - o[2]};
        P[k][0] = q[kx] - Sx * q[kz];
        P[k][1] = q[ky] - Sy * q[kz];
        P[k][2] = Sz * q[kz];
    }
    double U = P[2][0] * P[1][1] - P[2][1] * P[1][0];
    double V = P[0][0] * P[2][1] - P[0][1] * P[2][0];
    double W = P[1][0] * P[0][1] - P[1][1] * P[0][0];
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return false;
    double det = U + V + W;
    if (det == 0) return false;
    double t = (U * P[0][2] + V * P[1][2] + W * P[2][2]) / det;
    if (!(t >= tmin && t <= tmax)) return false;
    *t_out = t;
    return true;
}

// TraceRay (FP32 mirror): nearest hit, ties to the smaller canonical id.
// any_hit: stop at the first hit in [tmin, tmax].
Hit trace32(const Scene& s, float3 org, float3 dir, float tmin, float tmax, bool any_hit) {
    Hit h{};
    h.found = false;
    h.t = -1.0f;
    RayPre r = prepare(org, dir);
    double o[3] = {org.x, org.y, org.z}, d[3] = {dir.x, dir.y, dir.z};
    float best = tmax;
    uint32_t best_id = 0xFFFFFFFFu;
    int stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Scene::Node& nd = s.nodes[stack[--sp]];
        if (!box_hit(nd, o, d, (double)tmin, (double)best * (1.0 + 1e-6))) continue;
        if (nd.left < 0) {
            for (int i = nd.begin; i < nd.end; ++i) {
                uint32_t g = s.order[i];
                float3 p[3];
                tri_verts(s, g, p);
                float t, u, v;
                if (!tri_fp32(r, p, tmin, best, &t, &u, &v)) continue;
                if (any_hit) {
                    h.found = true; h.t = t; h.u = u; h.v = v; h.id = g; h.inst = s.tinst[g]; h.prim = s.tprim[g];
                    return h;
                }
                if (t < best || (t == best && g < best_id)) {
                    best = t; best_id = g;
                    h.found = true; h.t = t; h.u = u; h.v = v; h.id = g; h.inst = s.tinst[g]; h.prim = s.tprim[g];
                }
            }
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    return h;
}

// Closest hit in FP64 (exact reference): returns id and t, or id = ~0.
void trace64(const Scene& s, const double* o, const double* d, double tmin, double tmax, uint32_t* id_out,
             double* t_out) {
    double best = tmax;
    uint32_t best_id = 0xFFFFFFFFu;
    int stack[256];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Scene::Node& nd = s.nodes[stack[--sp]];
        if (!box_hit(nd, o, d, tmin, best * (1.0 + 1e-9))) continue;
        if (nd.left < 0) {
            for (int i = nd.begin; i < nd.end; ++i) {
                uint32_t g = s.order[i];
                float3 p[3];
                tri_verts(s, g, p);
                double t;
                if (!tri_fp64(o, d, p, tmin, best, &t)) continue;
                if (t < best || (t == best && g < best_id)) { best = t; best_id = g; }
            }
        } else {
            stack[sp++] = nd.left;
            stack[sp++] = nd.right;
        }
    }
    *id_out = best_id;
    *t_out = best_id == 0xFFFFFFFFu ? -1.0 : best;
}

// Closest-hit program (Ray.rchit.hlsl:26-61).  Identity instance transforms.
struct RayPayload {
    float3 Position, Normal, GeometryNormal;
    float2 UV;
    int MaterialIndex;
    float HitT;
};
RayPayload closest_hit(const Scene& s, const Hit& h) {
    RayPayload pl;
    uint32_t indexOffset = s.offsets[2 * h.inst], vertexOffset = s.offsets[2 * h.inst + 1];
    uint32_t i0 = s.indices[indexOffset + h.prim * 3 + 0];
    uint32_t i1 = s.indices[indexOffset + h.prim * 3 + 1];
    uint32_t i2 = s.indices[indexOffset + h.prim * 3 + 2];
    const Vertex& v0 = s.vertices[vertexOffset + i0];
    const Vertex& v1 = s.vertices[vertexOffset + i1];
    const Vertex& v2 = s.vertices[vertexOffset + i2];
    float3 bary = F3(1.0f - h.u - h.v, h.u, h.v);
    auto P = [](const Vertex& v) { return F3(v.Position[0], v.Position[1], v.Position[2]); };
    auto N = [](const Vertex& v) { return F3(v.Normal[0], v.Normal[1], v.Normal[2]); };
    float3 normal = normalize(N(v0) * bary.x + N(v1) * bary.y + N(v2) * bary.z);
    pl.UV = float2{v0.TexCoord[0] * bary.x + v1.TexCoord[0] * bary.y + v2.TexCoord[0] * bary.z,
                   v0.TexCoord[1] * bary.x + v1.TexCoord[1] * bary.y + v2.TexCoord[1] * bary.z};
    pl.Position = P(v0) * bary.x + P(v1) * bary.y + P(v2) * bary.z;
    pl.Normal = normalize(normal);  // worldNrm = normal (identity), then normalize again
    float3 e0 = P(v2) - P(v0), e1 = P(v1) - P(v0);
    pl.GeometryNormal = normalize(cross(e0, e1));
    pl.MaterialIndex = v0.MaterialIndex;
    pl.HitT = h.t;
    return pl;
}

// TraceRay + closest-hit / miss programs as the ray-gen shaders see them.
RayPayload TraceRayPayload(const Scene& s, float3 org, float3 dir, float tmin, float tmax) {
    Hit h = trace32(s, org, dir, tmin, tmax, false);
    if (!h.found) {
        RayPayload pl{};
        pl.MaterialIndex = -1;  // Ray.rmiss.hlsl:6-7
        pl.HitT = -1.0f;
        return pl;
    }
    return closest_hit(s, h);
}

// Camera ray (InitialSample.rgen:43-52).
void camera_ray(const UBO& ubo, uint32_t x, uint32_t y, uint32_t W, uint32_t H, uint32_t& rng, float3& origin,
                float3& direction) {
    float ox = RandomFloat(rng);
    float oy = RandomFloat(rng);
    float2 pixel{(float)x + ox, (float)y + oy};
    float2 uv{(pixel.x / (float)W) * 2.0f - 1.0f, (pixel.y / (float)H) * 2.0f - 1.0f};
    float4 o = mulv(ubo.modelViewInverse, float4{0, 0, 0, 1});
    float4 target = mulv(ubo.projectionInverse, float4{uv.x, uv.y, 1, 1});
    float3 tn = normalize(F3(target.x, target.y, target.z));
    float4 d = mulv(ubo.modelViewInverse, float4{tn.x, tn.y, tn.z, 0});
    origin = F3(o.x, o.y, o.z);
    direction = normalize(F3(d.x, d.y, d.z));
}

// ---------------------------------------------------------------------------
// Initial sample pass (ReSTIRGIInitialSample.rgen.hlsl:35-214), one pixel.
// ---------------------------------------------------------------------------
struct GOut {
    float pos[4], nrm[4], matuv[4];
    uint32_t motion;
};
// Radiance-cache path termination (no reference counterpart; Mueller et al.
// 2021, "Real-time Neural Radiance Caching", section 4).  A pixel whose
// path reaches a diffuse first secondary vertex x_s is a candidate.  Of the
// frame's C candidates exactly min(T, C) keep tracing as training samples:
// with r the candidate's rank in pixel order and r' = (r + o) mod C,
// o = (frame * 7919 mod 2^32) mod C, it trains when C <= T or
// floor((r' + 1) T / C) > floor(r' T / C).  Every other candidate stops into
// the cache when the spread t1^2 / (p_q |cos theta_s|) exceeds c times the
// primary footprint t0^2 / (4 pi |cos theta_v|), else traces on as a plain
// path.  Recorded in the spare words: pad1 = material at x_s, pad2 = 1
// (query) or 2 (training).
struct CacheOpt {
    bool enabled = false;
    bool probe = false;              // stop candidates at x_s with pad2 = 3 (the selection's pre-pass)
    const uint8_t* train = nullptr;  // per pixel: 1 = training sample
    float spread_c = 0.0f;
};

// The training choice above from the per-pixel candidate flags.
std::vector<uint8_t> select_training(const std::vector<uint8_t>& candidate, uint32_t frame, uint32_t target) {
    std::vector<uint8_t> train(candidate.size(), 0);
    uint64_t count = 0;
    for (uint8_t c : candidate) count += c;
    if (count == 0) return train;
    const uint64_t o = (uint64_t)(uint32_t)(frame * 7919u) % count;
    uint64_t r = 0;
    for (size_t p = 0; p < candidate.size(); ++p) {
        if (!candidate[p]) continue;
        const uint64_t rr = (r + o) % count;
        train[p] = count <= target || (rr + 1) * target / count > rr * target / count;
        ++r;
    }
    return train;
}
uint32_t initial_sample_pixel(const Scene& s, const UBO& ubo, uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                              ReSTIRGISample& out, GOut& g, const CacheOpt& cache = CacheOpt()) {
    uint32_t cache_kind = 0, xs_material = 0;
    float footprint0 = 0.0f;
    uint32_t rays = 0;
    uint32_t rngState = InitRandomSeed(x, y, W, ubo.currentFrame);
    float3 origin, direction;
    camera_ray(ubo, x, y, W, H, rngState, origin, direction);
    float3 throughput = F3(1, 1, 1);
    float3 x_view = F3(0, 0, 0), n_view = F3(0, 0, 0), x_sample = F3(0, 0, 0), n_sample = F3(0, 0, 0);
    float3 L_out = F3(0, 0, 0), f_ = F3(0, 0, 0);
    float p_q = 0;
    uint32_t mat_idx = 0xFFFFFFFFu, rand_seed = 0;
    float2 testUV{0, 0};
    bool primitiveHitValid = false;
    for (uint32_t bounce = 0;; bounce++) {
        if (bounce == ubo.numberOfBounces) break;
        ++rays;
        RayPayload pl = TraceRayPayload(s, origin, direction, 0.0001f, 10000.0f);
        if (pl.MaterialIndex == -1) {
            if (ubo.hasSkyBox) L_out = L_out + throughput * F3(1, 1, 1);
            if (bounce == 1) {
                x_sample = origin + direction * 1000.0f;
                n_sample = -direction;
            }
            break;
        }
        const BSDFMaterial& mat = s.materials[pl.MaterialIndex];
        // Emitters shade with the untextured mat.baseColor.rgb * 4.0
        // (InitialSample.rgen:117-120): the loop never calls SampleBSDF for them.
        if (mat.materialModel == BSDFDiffuseLight) {
            const float3 bc = F3(mat.baseColor[0], mat.baseColor[1], mat.baseColor[2]);
            L_out = L_out + throughput * bc * 4.0f;
            if (bounce == 1) { x_sample = pl.Position; n_sample = pl.Normal; }
            if (bounce == 0) { x_view = pl.Position; n_view = pl.Normal; }
            break;
        }
        float pdf = 0, cos_theta = 0;
        float3 wi = F3(0, 0, 0);
        if (bounce == 1) { x_sample = pl.Position; n_sample = pl.Normal; }
        if (cache.enabled) {
            const float cos_hit = std::fabs(dot(pl.Normal, direction));
            if (bounce == 0) {
                footprint0 = pl.HitT * pl.HitT / (4.0f * PI * cos_hit);
            } else if (bounce == 1) {
                xs_material = (uint32_t)pl.MaterialIndex;
                if (cache.probe) {
                    cache_kind = 3;
                    break;
                }
                if (cache.train[y * W + x]) {
                    cache_kind = 2;
                } else if (pl.HitT * pl.HitT / (p_q * cos_hit) > cache.spread_c * footprint0) {
                    cache_kind = 1;
                    break;
                }
            }
        }
        uint32_t seedForThisBounce = rngState;
        float3 f = SampleBSDF(s.textures, pl.Normal, pl.UV, mat, wi, pdf, cos_theta, rngState);
        if (bounce == 0) {
            primitiveHitValid = true;
            x_view = pl.Position;
            n_view = pl.Normal;
            rand_seed = seedForThisBounce;
            mat_idx = (uint32_t)pl.MaterialIndex;
            testUV = pl.UV;
            p_q = pdf;
            f_ = f;
        } else {
            throughput = throughput * (f * cos_theta / pdf);
        }
        if (bounce > 3) {
            float p = hmax(throughput.x, hmax(throughput.y, throughput.z));
            if (RandomFloat(rngState) > p) break;
            throughput = throughput * (1.0f / p);
        }
        float3 hitPosition = origin + direction * pl.HitT;
        float3 geoNormal = pl.GeometryNormal;
        float3 offsetDir = dot(wi, geoNormal) > 0 ? geoNormal : -geoNormal;
        origin = hitPosition + offsetDir * 0.01f;
        direction = wi;
    }
    out.x_view = x_view;
    out.n_view = n_view;
    out.x_sample = x_sample;
    out.n_sample = n_sample;
    out.L_out = L_out;
    out.materialID = mat_idx;
    out.randSeed = rand_seed;
    out.p_q = p_q;
    out.f = f_;
    out.pad1 = cache_kind ? xs_material : 0u;
    out.pad2 = cache_kind;
    out.pad3 = 0;
    if (primitiveHitValid) {
        float pos[4] = {x_view.x, x_view.y, x_view.z, 1.0f};
        float nrm[4] = {n_view.x, n_view.y, n_view.z, 0.0f};
        float mu[4] = {(float)mat_idx, testUV.x, testUV.y, 0.0f};
        memcpy(g.pos, pos, 16);
        memcpy(g.nrm, nrm, 16);
        memcpy(g.matuv, mu, 16);
        float2 currentUV{((float)x + 0.5f) / (float)W, ((float)y + 0.5f) / (float)H};
        float2 prevUV = WorldToScreenUV(x_view, ubo.viewPrev, ubo.projectionPrev);
        g.motion = (uint32_t)to_half(currentUV.x - prevUV.x) | ((uint32_t)to_half(currentUV.y - prevUV.y) << 16);
    } else {
        memset(g.pos, 0, 16);
        memset(g.nrm, 0, 16);
        float mu[4] = {-1.0f, 0, 0, 0};
        memcpy(g.matuv, mu, 16);
        g.motion = 0;
        out.L_out = F3(0, 0, 0);
        out.n_sample = F3(0, 0, 0);
    }
    return rays;
}

// ---------------------------------------------------------------------------
// Temporal reuse pass (ReSTIRGITemporalReuse.rgen.hlsl:33-151), one pixel.
// gpos_prev is what binding 16 reads: the current G-buffer position image
// (ReSTIRGIPipeline.cpp:171-199).
// ---------------------------------------------------------------------------
float EvaluateTargetFunction(const ReSTIRGISample& s, float3 x_curr, float3 n_curr) {
    float3 wi = normalize(s.x_sample - x_curr);
    float cos_theta = dot(n_curr, wi);
    if (cos_theta <= 0.0f) return 0.0f;
    return length(s.L_out);
}

void temporal_pixel(uint32_t x, uint32_t y, uint32_t W, uint32_t H, uint32_t frame, uint32_t m_max,
                    const ReSTIRGISample* samples, const uint32_t* motion, const float* gpos_prev,
                    ReSTIRGIReservoir* temporal, ReSTIRGIReservoir* spatial, uint32_t* output) {
    const uint32_t pixelIndex = y * W + x;
    uint32_t rngState = InitRandomSeed(x, y, W, frame);
    uint32_t totalPixels = W * H;
    uint32_t pingpong = frame % 2;
    uint32_t currentOffset = pingpong * totalPixels;
    uint32_t prevOffset = (1 - pingpong) * totalPixels;
    const ReSTIRGISample& s_new = samples[pixelIndex];
    ReSTIRGIReservoir r_new, r_prev;
    if (length(s_new.n_sample) == 0) {
        InitReservoir(r_new);
        temporal[currentOffset + pixelIndex] = r_new;
        spatial[currentOffset + pixelIndex] = r_new;
        return;
    }
    bool foundTemporal = false;
    uint32_t mb = motion[pixelIndex];
    float2 mv{from_half((uint16_t)(mb & 0xFFFF)), from_half((uint16_t)(mb >> 16))};
    float2 uv{((float)x + 0.5f) / (float)W, ((float)y + 0.5f) / (float)H};
    float2 prevUV{uv.x - mv.x, uv.y - mv.y};
    int px = (int)(prevUV.x * (float)W), py = (int)(prevUV.y * (float)H);
    if (px >= 0 && px < (int)W && py >= 0 && py < (int)H) {
        uint32_t prevPixelIndex = (uint32_t)py * W + (uint32_t)px;
        if (gpos_prev[4 * (size_t)prevPixelIndex + 3] != 0) {
            r_prev = temporal[prevOffset + prevPixelIndex];
            foundTemporal = true;
        } else {
            InitReservoir(r_prev);
            spatial[currentOffset + pixelIndex] = r_prev;
        }
    }
    float3 x_curr = s_new.x_view;
    float3 n_curr = normalize(s_new.n_view);
    ReSTIRGIReservoir r_final;
    InitReservoir(r_final);
    float target_p_new = EvaluateTargetFunction(s_new, x_curr, n_curr);
    float w_new = (s_new.p_q > 1e-6f) ? (target_p_new / s_new.p_q) : 0.0f;
    UpdateReservoir(r_final, s_new, w_new, rngState);
    if (foundTemporal) {
        float target_p_prev = EvaluateTargetFunction(r_prev.s, x_curr, n_curr);
        MergeReservoir(r_final, r_prev, target_p_prev, rngState);
    } else {
        r_final.m = 1;
    }
    if (r_final.m > m_max) {
        r_final.w_sum *= (float)m_max / (float)r_final.m;
        r_final.m = m_max;
    }
    float target_p_final = EvaluateTargetFunction(r_final.s, x_curr, n_curr);
    if (target_p_final > 1e-6f && r_final.m > 0) r_final.W = r_final.w_sum / (float(r_final.m) * target_p_final);
    else r_final.W = 0.0f;
    temporal[currentOffset + pixelIndex] = r_final;
    float3 finalColor = F3(0, 0, 0);
    if (r_final.W > 0) {
        float3 xc = s_new.x_view, nc = s_new.n_view;
        float3 wi = normalize(r_final.s.x_sample - xc);
        float cos_theta = hmax(0.0f, dot(nc, wi));
        finalColor = r_final.s.L_out * r_final.W * r_final.s.f * cos_theta;
    }
    output[pixelIndex] = store_bgra8(finalColor, 1.0f);
}

// ---------------------------------------------------------------------------
// Spatial reuse pass (ReSTIRGISpatialReuse.rgen.hlsl:35-183), one pixel.
// ---------------------------------------------------------------------------
uint32_t GetNeighborIndex(uint32_t x, uint32_t y, uint32_t W, uint32_t H, float max_r, uint32_t& seed) {
    float randa = RandomFloat(seed) * 2.0f * PI;
    float randr = sqrtf(RandomFloat(seed)) * max_r;
    int ox = (int)floorf(hcos(randa) * randr), oy = (int)floorf(hsin(randa) * randr);
    int cx = std::min(std::max((int)x + ox, 0), (int)W - 1);
    int cy = std::min(std::max((int)y + oy, 0), (int)H - 1);
    return (uint32_t)cy * W + (uint32_t)cx;
}

// Returns the number of shadow rays traced.
uint32_t spatial_pixel(const Scene* s, uint32_t x, uint32_t y, uint32_t W, uint32_t H, uint32_t frame,
                       uint32_t reuse, uint32_t num_samples, float radius, uint32_t m_max, int visibility,
                       const ReSTIRGISample* samples, const ReSTIRGIReservoir* temporal, ReSTIRGIReservoir* spatial,
                       uint32_t* output) {
    uint32_t rays = 0;
    const uint32_t pixelIndex = y * W + x;
    uint32_t rngState = InitRandomSeed(x, y, W, frame);
    uint32_t totalPixels = W * H;
    uint32_t pingpong = frame % 2;
    uint32_t readOffset = pingpong * totalPixels;
    uint32_t writeOffset = (1 - pingpong) * totalPixels;
    const ReSTIRGISample& s_new = samples[pixelIndex];
    if (length(s_new.n_view) == 0) {
        output[pixelIndex] = 0xFF000000u;  // float4(0,0,0,1), written without tone mapping
        return 0;
    }
    if (length(s_new.n_sample) == 0) {
        output[pixelIndex] = 0xFFFFFFFFu;  // float4(1,1,1,1)
        return 0;
    }
    const ReSTIRGIReservoir& centerRes = temporal[readOffset + pixelIndex];
    ReSTIRGIReservoir r_s;
    InitReservoir(r_s);
    MergeReservoir(r_s, centerRes, p_hat(centerRes.s.L_out), rngState);
    for (uint32_t i = 0; i < num_samples; i++) {
        uint32_t neighborIdx = GetNeighborIndex(x, y, W, H, radius, rngState);
        const ReSTIRGISample& s_neighbor = samples[neighborIdx];
        if (length(s_neighbor.n_view) == 0 || !IsSimilar(s_new, s_neighbor)) continue;
        const ReSTIRGIReservoir& r_neighbor = temporal[readOffset + neighborIdx];
        float3 x_sample = r_neighbor.s.x_sample, n_sample = r_neighbor.s.n_sample;
        float3 v_orig = s_neighbor.x_view - x_sample;
        float dist2_orig = dot(v_orig, v_orig);
        float3 v_curr = s_new.x_view - x_sample;
        float dist2_curr = dot(v_curr, v_curr);
        float jacobian = 1.0f;
        if (dist2_orig > 1e-6f && dist2_curr > 1e-6f) {
            float3 dir_orig = v_orig * rsqrt_(dist2_orig);
            float3 dir_curr = v_curr * rsqrt_(dist2_curr);
            float cos_phi_orig = hmax(0.0f, dot(n_sample, dir_orig));
            float cos_phi_curr = hmax(0.0f, dot(n_sample, dir_curr));
            if (cos_phi_orig > 1e-6f) jacobian = (cos_phi_curr * dist2_orig) / (cos_phi_orig * dist2_curr);
        }
        jacobian = hclamp(jacobian, 0.1f, 10.0f);
        float p_hat_n = p_hat(r_neighbor.s.L_out) * jacobian;
        if (visibility) {
            // CheckVisibility (Spatial.rgen:44-62) with a real any-hit ray.
            float3 from = r_s.s.x_view, to = x_sample, normal = r_s.s.n_view;
            float3 dir = to - from;
            float dist = length(dir);
            ++rays;
            Hit h = trace32(*s, from + normal * 0.01f, normalize(dir), 0.0f, dist - 0.01f, true);
            if (h.found) p_hat_n = 0;
        }
        MergeReservoir(r_s, r_neighbor, p_hat_n, rngState);
        if (p_hat_n == 0) r_s.m -= r_neighbor.m;
    }
    r_s.m = std::min(r_s.m, m_max);
    float p_hat_final = p_hat(r_s.s.L_out);
    if (p_hat_final > 1e-6f && r_s.m > 0) r_s.W = r_s.w_sum / (float(r_s.m) * p_hat_final);
    else r_s.W = 0.0f;
    float3 finalColor = F3(0, 0, 0);
    if (r_s.W > 0) {
        float3 xc = s_new.x_view, nc = s_new.n_view;
        float3 wi = normalize(r_s.s.x_sample - xc);
        float cos_theta = hmax(0.0f, dot(nc, wi));
        finalColor = r_s.s.L_out * r_s.W * r_s.s.f * cos_theta;
    }
    spatial[writeOffset + pixelIndex] = r_s;
    if (reuse) output[pixelIndex] = store_bgra8(finalColor, 1.0f);
    return rays;
}

// ---------------------------------------------------------------------------
// Plain path tracer (rgen.hlsl:31-151), one pixel.  max_depth 0 = unbounded.
// ---------------------------------------------------------------------------
uint32_t plain_pixel(const Scene& s, const UBO& ubo, uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                     uint32_t max_depth, float* accum, uint32_t* output) {
    uint32_t rays = 0;
    const uint32_t pix = y * W + x;
    uint32_t rngState = InitRandomSeed(x, y, W, ubo.currentFrame);
    float3 pixelColor = F3(0, 0, 0);
    for (uint32_t i = 0; i < ubo.spp; i++) {
        float3 origin, direction;
        camera_ray(ubo, x, y, W, H, rngState, origin, direction);
        float3 rayColor = F3(0, 0, 0), throughput = F3(1, 1, 1);
        for (uint32_t bounce = 0;; bounce++) {
            if (max_depth != 0 && bounce == max_depth) break;
            ++rays;
            RayPayload pl = TraceRayPayload(s, origin, direction, 0.0001f, 10000.0f);
            if (pl.MaterialIndex == -1) {
                float t = 0.5f * (direction.y + 1.0f);
                // lerp(a, b, t) = a + t*(b - a)
                float3 sky = F3(1.0f + t * (0.5f - 1.0f), 1.0f + t * (0.7f - 1.0f), 1.0f + t * (1.0f - 1.0f));
                if (ubo.hasSkyBox) rayColor = rayColor + throughput * sky;
                break;
            }
            const BSDFMaterial& mat = s.materials[pl.MaterialIndex];
            // untextured mat.baseColor.rgb * 4.0 (rgen.hlsl:96-99)
            if (mat.materialModel == BSDFDiffuseLight) {
                rayColor = rayColor + throughput * F3(mat.baseColor[0], mat.baseColor[1], mat.baseColor[2]) * 4.0f;
                break;
            }
            float pdf = 0, cos_theta = 0;
            float3 wi;
            float3 f = SampleBSDF(s.textures, pl.Normal, pl.UV, mat, wi, pdf, cos_theta, rngState);
            if (pdf <= 1e-6f) break;
            throughput = throughput * (f * cos_theta / pdf);
            if (bounce > 3) {
                float p = hmax(throughput.x, hmax(throughput.y, throughput.z));
                if (RandomFloat(rngState) > p) break;
                throughput = throughput * (1.0f / p);
            }
            float3 hitPosition = origin + direction * pl.HitT;
            float3 offsetDir = dot(wi, pl.Normal) > 0 ? pl.Normal : -pl.Normal;
            origin = hitPosition + offsetDir * 0.001f;
            direction = wi;
        }
        pixelColor = pixelColor + rayColor;
    }
    bool accumulate = ubo.spp != ubo.totalNumberOfSamples;
    float3 acc = accumulate ? F3(accum[4 * (size_t)pix], accum[4 * (size_t)pix + 1], accum[4 * (size_t)pix + 2])
                            : F3(0, 0, 0);
    acc = acc + pixelColor;
    float3 finalColor = acc / (float)ubo.totalNumberOfSamples;
    accum[4 * (size_t)pix] = acc.x;
    accum[4 * (size_t)pix + 1] = acc.y;
    accum[4 * (size_t)pix + 2] = acc.z;
    accum[4 * (size_t)pix + 3] = 0.0f;
    output[pix] = store_bgra8(finalColor, 0.0f);
    return rays;
}

}  // namespace

// ===========================================================================
// C ABI (called through ctypes by the tests / smoke / CPU baseline)
// ===========================================================================
extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

uint32_t orc_init_seed(uint32_t x, uint32_t y, uint32_t w, uint32_t frame) { return InitRandomSeed(x, y, w, frame); }
void orc_random_stream(uint32_t state, uint32_t n, uint32_t* words, float* floats) {
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t st = state;
        words[i] = NextRandom(state);
        floats[i] = RandomFloat(st);
    }
}
void orc_sincos(const float* x, uint32_t n, float* s, float* c) {
    for (uint32_t i = 0; i < n; ++i) sincos_cephes(x[i], &s[i], &c[i]);
}
void orc_onb(const float* n, float* b1, float* b2) {
    float3 t, b;
    BranchlessONB(F3(n[0], n[1], n[2]), t, b);
    b1[0] = t.x; b1[1] = t.y; b1[2] = t.z;
    b2[0] = b.x; b2[1] = b.y; b2[2] = b.z;
}
// SampleBSDF on a Lambertian of albedo a at normal n; returns draws used.
void orc_sample_lambert(const float* n, const float* albedo, uint32_t seed, float* wi, float* f, float* pdf,
                        float* cos_theta, uint32_t* seed_out) {
    BSDFMaterial m{};
    m.baseColor[0] = albedo[0]; m.baseColor[1] = albedo[1]; m.baseColor[2] = albedo[2]; m.baseColor[3] = 1;
    m.baseColorTextureIndex = -1;
    m.materialModel = BSDFLambertian;
    float3 w;
    const TextureSet none;
    float3 ff = SampleBSDF(none, F3(n[0], n[1], n[2]), float2{0.0f, 0.0f}, m, w, *pdf, *cos_theta, seed);
    wi[0] = w.x; wi[1] = w.y; wi[2] = w.z;
    f[0] = ff.x; f[1] = ff.y; f[2] = ff.z;
    *seed_out = seed;
}
uint32_t orc_resolve(const float* c, float alpha) { return store_bgra8(F3(c[0], c[1], c[2]), alpha); }
float orc_aces(float x) { return ACES1(x); }
uint16_t orc_to_half(float f) { return to_half(f); }
float orc_from_half(uint16_t h) { return from_half(h); }

// Reservoir primitives on reference-layout structs (112-byte reservoirs).
void orc_update_reservoir(void* r, const void* s, float w, uint32_t* seed) {
    UpdateReservoir(*(ReSTIRGIReservoir*)r, *(const ReSTIRGISample*)s, w, *seed);
}
void orc_merge_reservoir(void* r, const void* src, float p_hat_, uint32_t* seed) {
    MergeReservoir(*(ReSTIRGIReservoir*)r, *(const ReSTIRGIReservoir*)src, p_hat_, *seed);
}

// ---- scene ----
// Textures of the scene: info = (first texel, width, height) per texture,
// texels RGBA8.  Returns 0, or -1 when a texture runs past the texels.
int orc_scene_set_textures(void* scene, const uint32_t* info, uint32_t count, const uint8_t* texels,
                           size_t texel_bytes) {
    Scene* s = (Scene*)scene;
    for (uint32_t t = 0; t < count; ++t)
        if (((size_t)info[3 * t] + (size_t)info[3 * t + 1] * info[3 * t + 2]) * 4 > texel_bytes) return -1;
    s->textures.info.assign(info, info + 3 * (size_t)count);
    s->textures.texels.assign(texels, texels + texel_bytes);
    return 0;
}

// TextureSample of the scene's texture t at (u, v) (for tests).
void orc_texture_sample(const void* scene, int t, float u, float v, float* rgb) {
    const float3 c = TextureSample(((const Scene*)scene)->textures, t, float2{u, v});
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}

void* orc_scene_create(const void* vertices, uint32_t vcount, const uint32_t* indices, uint32_t icount,
                       const void* materials, uint32_t mcount, const uint32_t* offsets, uint32_t ninst) {
    // reject what the reference would read out of bounds
    for (uint32_t i = 0; i < ninst; ++i) {
        const uint32_t end = i + 1 < ninst ? offsets[2 * (i + 1)] : icount;
        for (uint32_t k = offsets[2 * i]; k < end; ++k)
            if ((uint64_t)offsets[2 * i + 1] + indices[k] >= vcount) return nullptr;
    }
    Scene* s = new Scene();
    s->vertices.assign((const Vertex*)vertices, (const Vertex*)vertices + vcount);
    s->indices.assign(indices, indices + icount);
    s->materials.assign((const BSDFMaterial*)materials, (const BSDFMaterial*)materials + mcount);
    s->offsets.assign(offsets, offsets + 2 * ninst);
    uint32_t ntri = icount / 3;
    s->first.resize(ninst + 1);
    for (uint32_t i = 0; i < ninst; ++i) s->first[i] = offsets[2 * i] / 3;
    s->first[ninst] = ntri;
    s->tv.resize(9 * (size_t)ntri);
    s->tinst.resize(ntri);
    s->tprim.resize(ntri);
    std::vector<double> cent(3 * (size_t)ntri);
    for (uint32_t i = 0; i < ninst; ++i) {
        for (uint32_t g = s->first[i]; g < s->first[i + 1]; ++g) {
            uint32_t prim = g - s->first[i];
            s->tinst[g] = i;
            s->tprim[g] = prim;
            for (int k = 0; k < 3; ++k) {
                const Vertex& v = s->vertices[offsets[2 * i + 1] + indices[offsets[2 * i] + prim * 3 + k]];
                for (int a = 0; a < 3; ++a) s->tv[9 * (size_t)g + 3 * k + a] = v.Position[a];
            }
            for (int a = 0; a < 3; ++a)
                cent[3 * (size_t)g + a] = ((double)s->tv[9 * (size_t)g + a] + s->tv[9 * (size_t)g + 3 + a] +
                                           s->tv[9 * (size_t)g + 6 + a]) / 3.0;
        }
    }
    {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (size_t k = 0; k < s->tv.size(); ++k) {
            lo[k % 3] = std::min(lo[k % 3], (double)s->tv[k]);
            hi[k % 3] = std::max(hi[k % 3], (double)s->tv[k]);
        }
        double ext = 0, mag = 0;
        for (int a = 0; a < 3; ++a) {
            ext = std::max(ext, hi[a] - lo[a]);
            mag = std::max(mag, std::max(fabs(lo[a]), fabs(hi[a])));
        }
        s->pad = std::max(ext / 65536.0, mag / 1048576.0);
    }
    s->order.resize(ntri);
    for (uint32_t g = 0; g < ntri; ++g) s->order[g] = g;
    s->nodes.reserve(2 * (size_t)ntri / 2 + 16);
    build_node(*s, cent, 0, (int)ntri);
    return s;
}
void orc_scene_destroy(void* h) { delete (Scene*)h; }

// Ray records as xr_ray (org, tmin, dir, tmax); hits as xr_hit.
struct OrcHit {
    float t, u, v;
    uint32_t pad0, inst, prim, id, pad1;
};
void orc_trace32(const void* h, const float* rays, uint32_t n, void* hits_out, int any_hit) {
    const Scene& s = *(const Scene*)h;
    OrcHit* out = (OrcHit*)hits_out;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* r = rays + 8 * i;
        Hit hh = trace32(s, F3(r[0], r[1], r[2]), F3(r[4], r[5], r[6]), r[3], r[7], any_hit != 0);
        OrcHit o{};
        if (hh.found) {
            o.t = hh.t; o.u = hh.u; o.v = hh.v; o.inst = hh.inst; o.prim = hh.prim; o.id = hh.id;
        } else {
            o.t = -1.0f; o.inst = o.prim = o.id = 0xFFFFFFFFu;
        }
        out[i] = o;
    }
}
// Brute force over every triangle (pins the oracle's own BVH on small scenes).
void orc_trace32_brute(const void* h, const float* rays, uint32_t n, void* hits_out) {
    const Scene& s = *(const Scene*)h;
    OrcHit* out = (OrcHit*)hits_out;
    const uint32_t ntri = (uint32_t)s.tinst.size();
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * i;
        RayPre rp = prepare(F3(r[0], r[1], r[2]), F3(r[4], r[5], r[6]));
        float best = r[7];
        OrcHit o{};
        o.t = -1.0f;
        o.inst = o.prim = o.id = 0xFFFFFFFFu;
        for (uint32_t g = 0; g < ntri; ++g) {
            float3 p[3];
            tri_verts(s, g, p);
            float t, u, v;
            if (tri_fp32(rp, p, r[3], best, &t, &u, &v) && (t < best || (t == best && g < o.id))) {
                best = t;
                o.t = t; o.u = u; o.v = v; o.id = g; o.inst = s.tinst[g]; o.prim = s.tprim[g];
            }
        }
        out[i] = o;
    }
}
// FP64 closest hit: ids (0xFFFFFFFF = miss) and t (-1 = miss).
void orc_trace64(const void* h, const double* rays, uint32_t n, uint32_t* ids, double* ts) {
    const Scene& s = *(const Scene*)h;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const double* r = rays + 8 * i;
        trace64(s, r, r + 4, r[3], r[7], &ids[i], &ts[i]);
    }
}

// ---- passes over whole images (reference layouts) ----
// pixels: list of pixel indices to compute (NULL = all W*H).
// train_target = 0: no cache.  Otherwise the training set is chosen from a
// pre-pass over every pixel (bounces 0 and 1), whatever `pixels` lists.
uint64_t orc_initial_sample_cache(const void* h, const void* ubo, uint32_t W, uint32_t H, const uint32_t* pixels,
                                  uint32_t npix, void* samples, float* gpos, float* gnrm, float* gmatuv,
                                  uint32_t* motion, uint32_t train_target, float spread_c, uint8_t* train_out) {
    const Scene& s = *(const Scene*)h;
    const UBO& u = *(const UBO*)ubo;
    CacheOpt cache;
    cache.enabled = train_target != 0;
    cache.spread_c = spread_c;
    std::vector<uint8_t> train;
    if (cache.enabled) {
        std::vector<uint8_t> candidate((size_t)W * H, 0);
        CacheOpt probe = cache;
        probe.probe = true;
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t p = 0; p < (int64_t)W * H; ++p) {
            ReSTIRGISample tmp;
            GOut g;
            initial_sample_pixel(s, u, (uint32_t)(p % W), (uint32_t)(p / W), W, H, tmp, g, probe);
            candidate[p] = tmp.pad2 == 3 ? 1 : 0;
        }
        train = select_training(candidate, u.currentFrame, train_target);
        cache.train = train.data();
        if (train_out) memcpy(train_out, train.data(), train.size());
    }
    ReSTIRGISample* out = (ReSTIRGISample*)samples;
    const int64_t count = pixels ? (int64_t)npix : (int64_t)W * H;
    uint64_t rays = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : rays)
    for (int64_t k = 0; k < count; ++k) {
        uint32_t p = pixels ? pixels[k] : (uint32_t)k;
        uint32_t x = p % W, y = p / W;
        GOut g;
        rays += initial_sample_pixel(s, u, x, y, W, H, out[p], g, cache);
        memcpy(gpos + 4 * (size_t)p, g.pos, 16);
        memcpy(gnrm + 4 * (size_t)p, g.nrm, 16);
        memcpy(gmatuv + 4 * (size_t)p, g.matuv, 16);
        motion[p] = g.motion;
    }
    return rays;
}

uint64_t orc_initial_sample(const void* h, const void* ubo, uint32_t W, uint32_t H, const uint32_t* pixels,
                            uint32_t npix, void* samples, float* gpos, float* gnrm, float* gmatuv,
                            uint32_t* motion) {
    return orc_initial_sample_cache(h, ubo, W, H, pixels, npix, samples, gpos, gnrm, gmatuv, motion, 0, 0.0f,
                                    nullptr);
}

void orc_temporal(uint32_t W, uint32_t H, uint32_t frame, uint32_t m_max, const uint32_t* pixels, uint32_t npix,
                  const void* samples, const uint32_t* motion, const float* gpos_prev, void* temporal, void* spatial,
                  uint32_t* output) {
    const int64_t count = pixels ? (int64_t)npix : (int64_t)W * H;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < count; ++k) {
        uint32_t p = pixels ? pixels[k] : (uint32_t)k;
        temporal_pixel(p % W, p / W, W, H, frame, m_max, (const ReSTIRGISample*)samples, motion, gpos_prev,
                       (ReSTIRGIReservoir*)temporal, (ReSTIRGIReservoir*)spatial, output);
    }
}

uint64_t orc_spatial(const void* h, uint32_t W, uint32_t H, uint32_t frame, uint32_t reuse, uint32_t num_samples,
                     float radius, uint32_t m_max, int visibility, const uint32_t* pixels, uint32_t npix,
                     const void* samples, const void* temporal, void* spatial, uint32_t* output) {
    const Scene* s = (const Scene*)h;
    const int64_t count = pixels ? (int64_t)npix : (int64_t)W * H;
    uint64_t rays = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : rays)
    for (int64_t k = 0; k < count; ++k) {
        uint32_t p = pixels ? pixels[k] : (uint32_t)k;
        rays += spatial_pixel(s, p % W, p / W, W, H, frame, reuse, num_samples, radius, m_max, visibility,
                              (const ReSTIRGISample*)samples, (const ReSTIRGIReservoir*)temporal,
                              (ReSTIRGIReservoir*)spatial, output);
    }
    return rays;
}

uint64_t orc_plain(const void* h, const void* ubo, uint32_t W, uint32_t H, uint32_t max_depth,
                   const uint32_t* pixels, uint32_t npix, float* accum, uint32_t* output) {
    const Scene& s = *(const Scene*)h;
    const UBO& u = *(const UBO*)ubo;
    const int64_t count = pixels ? (int64_t)npix : (int64_t)W * H;
    uint64_t rays = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : rays)
    for (int64_t k = 0; k < count; ++k) {
        uint32_t p = pixels ? pixels[k] : (uint32_t)k;
        rays += plain_pixel(s, u, p % W, p / W, W, H, max_depth, accum, output);
    }
    return rays;
}

}  // extern "C"

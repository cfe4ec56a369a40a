// Scalar/vector FP32 arithmetic shared by the HIP kernels and the C++ host.
//
// Every formula is written with an explicit evaluation order, and the whole
// library is compiled with -ffp-contract=off and correctly rounded FP32
// division/sqrt, so a given expression yields the same bits on gfx950 and on
// x86.  Where a fused multiply-add is wanted it is spelled out with fmaf().
//
// Reference semantics (HLSL, shader/HLSL/ReSTIRGI/*.hlsl):
//   dot(a,b)      = a.x*b.x + a.y*b.y + a.z*b.z           (left to right)
//   length(a)     = sqrt(dot(a,a))
//   normalize(a)  = a * (1 / sqrt(dot(a,a)))
#pragma once
#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define XR_HD __host__ __device__ __forceinline__
#else
#define XR_HD static inline
#endif

namespace xr {

struct v3 {
    float x, y, z;
};

XR_HD v3 mk3(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
XR_HD v3 add(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
XR_HD v3 sub(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
XR_HD v3 mul(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
XR_HD v3 muls(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
XR_HD v3 divs(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
XR_HD v3 neg(v3 a) { return mk3(-a.x, -a.y, -a.z); }
XR_HD float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
XR_HD float length(v3 a) { return sqrtf(dot(a, a)); }
XR_HD v3 normalize(v3 a) { float inv = 1.0f / sqrtf(dot(a, a)); return muls(a, inv); }
XR_HD v3 cross(v3 a, v3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
XR_HD float maxf(float a, float b) { return a > b ? a : b; }
XR_HD float minf(float a, float b) { return a < b ? a : b; }
XR_HD float clampf(float x, float lo, float hi) { return minf(maxf(x, lo), hi); }

// Constants exactly as the reference's #defines round to FP32
// (ReSTIRGICommon.hlsl:1-3, sampling.hlsl:3-5).
#define XR_PI 3.14159265359f
#define XR_TWO_PI 6.28318530718f
#define XR_INV_PI ((float)(1.0 / 3.14159265359))

// sin/cos of an FP32 argument: Cody-Waite reduction by pi/2 in three FMA
// steps, then the Cephes single-precision minimax polynomials on
// [-pi/4, pi/4] (S. Moshier, Cephes Math Library, sinf.c).  About 1 ulp for
// |x| < 1e4, which covers every argument the path produces (angles in
// [0, 2*pi)).  Written out here so that the GPU and the host produce the same
// bits.
XR_HD void sincos_xr(float x, float* s_out, float* c_out) {
    const float kTwoOverPi = 0.636619772367581343f;
    const float kPio2Hi = 1.5707963705062866211e+00f;
    const float kPio2Mid = -4.3711388286737928865e-08f;
    const float kPio2Lo = -1.7151245100059184123e-15f;
    float k = rintf(x * kTwoOverPi);
    float r = fmaf(-k, kPio2Hi, x);
    r = fmaf(-k, kPio2Mid, r);
    r = fmaf(-k, kPio2Lo, r);
    float z = r * r;
    float ps = fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f);
    float sn = fmaf(ps * z, r, r);
    float pc = fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f);
    float cs = fmaf(pc * z, z, fmaf(-0.5f, z, 1.0f));
    int q = ((int)k) & 3;
    float s, c;
    if (q == 0) { s = sn; c = cs; }
    else if (q == 1) { s = cs; c = -sn; }
    else if (q == 2) { s = -sn; c = -cs; }
    else { s = -cs; c = sn; }
    *s_out = s;
    *c_out = c;
}

// ---------------------------------------------------------------------------
// RNG (shader/HLSL/random.hlsl:1-20)
// ---------------------------------------------------------------------------
// InitRandomSeed: Wang-style hash of x + y*W, times (frame + 1).
XR_HD uint32_t init_seed(uint32_t x, uint32_t y, uint32_t width, uint32_t frame) {
    uint32_t s = x + y * width;
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s * (frame + 1u);
}
// NextRandom: PCG-RXS-M-XS 32.
XR_HD uint32_t next_u32(uint32_t* st) {
    uint32_t state = *st * 747796405u + 2891336453u;
    *st = state;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
// RandomFloat: 24 random bits / 2^24.
XR_HD float rand01(uint32_t* st) { return (float)(next_u32(st) & 0x00FFFFFFu) / 16777216.0f; }

// ---------------------------------------------------------------------------
// Tone map (ReSTIRGICommon.hlsl:123-131) and output conversion
// ---------------------------------------------------------------------------
XR_HD float aces1(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return clampf((x * (a * x + b)) / (x * (c * x + d) + e), 0.0f, 1.0f);
}
// float -> UNORM8 as a storage-image write performs it (round to nearest).
XR_HD uint32_t unorm8(float x) {
    float c = clampf(x, 0.0f, 1.0f);
    return (uint32_t)(c * 255.0f + 0.5f);
}
// finalColor = pow(ACES(x), 1/2.2) written to a B8G8R8A8_UNORM texel with alpha a.
XR_HD uint32_t resolve_bgra8(v3 color, float alpha) {
    const float g = (float)(1.0 / 2.2);
    float r = powf(aces1(color.x), g);
    float gg = powf(aces1(color.y), g);
    float b = powf(aces1(color.z), g);
    return unorm8(b) | (unorm8(gg) << 8) | (unorm8(r) << 16) | (unorm8(alpha) << 24);
}

// FP32 -> FP16 bits, round to nearest even (the RG16F motion-vector store).
XR_HD uint16_t f32_to_f16_bits(float f) {
    union { float f; uint32_t u; } v;
    v.f = f;
    uint32_t x = v.u;
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ex = (x >> 23) & 0xFFu;
    uint32_t man = x & 0x7FFFFFu;
    if (ex == 0xFFu) return (uint16_t)(sign | 0x7C00u | (man ? 0x200u : 0u));
    int e = (int)ex - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7C00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        man |= 0x800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t half = man >> shift;
        uint32_t rem = man & ((1u << shift) - 1u);
        uint32_t mid = 1u << (shift - 1);
        if (rem > mid || (rem == mid && (half & 1u))) half++;
        return (uint16_t)(sign | half);
    }
    uint32_t half = sign | ((uint32_t)e << 10) | (man >> 13);
    uint32_t rem = man & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (half & 1u))) half++;
    return (uint16_t)half;
}
XR_HD float f16_bits_to_f32(uint16_t h) {
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t ex = ((uint32_t)h >> 10) & 0x1Fu;
    uint32_t man = (uint32_t)h & 0x3FFu;
    uint32_t out;
    if (ex == 0) {
        if (man == 0) {
            out = sign;
        } else {
            int e = -1;
            do { man <<= 1; e++; } while ((man & 0x400u) == 0);
            man &= 0x3FFu;
            out = sign | ((uint32_t)(127 - 15 - e) << 23) | (man << 13);
        }
    } else if (ex == 0x1Fu) {
        out = sign | 0x7F800000u | (man << 13);
    } else {
        out = sign | ((ex - 15 + 127) << 23) | (man << 13);
    }
    union { uint32_t u; float f; } v;
    v.u = out;
    return v.f;
}

XR_HD float as_float(uint32_t u) { union { uint32_t u; float f; } v; v.u = u; return v.f; }
XR_HD uint32_t as_uint(float f) { union { float f; uint32_t u; } v; v.f = f; return v.u; }

}  // namespace xr

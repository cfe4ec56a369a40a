// Byte layouts shared by the C ABI, the HIP kernels and the host driver.
//
// The AoS structs are byte-identical to the reference's host mirrors so that
// dumps and fixtures can be exchanged with it:
//   AVertex            src/Asset/Vertex.h:11-16            36 B
//   AMaterial          src/Asset/Material.h:46-57          32 B
//   UniformBufferObject src/Asset/UniformBuffer.h:16-32    412 B
//   ReSTIRGISample     src/Vulkan/RayTracing/ReSTIRGICommon.h:8-20   96 B
//   ReSTIRGIReservoir  src/Vulkan/RayTracing/ReSTIRGICommon.h:22-31  112 B
//
// In HBM the kernels keep samples and reservoirs as six float4 planes each
// ("SoA by 16-byte field"), so a wavefront reads 64 consecutive float4s of one
// field per load instruction:
//   plane 0: x_view.xyz , materialID (bits)
//   plane 1: n_view.xyz , randSeed   (bits)
//   plane 2: x_sample.xyz, p_q
//   plane 3: n_sample.xyz, w_sum  (reservoir) / 0 (sample)
//   plane 4: L_out.xyz   , W      (reservoir) / 0 (sample)
//   plane 5: f.xyz       , M      (reservoir, bits) / 0 (sample)
// A reservoir therefore occupies 96 B per pixel with no padding, and copying
// a sample into a reservoir copies planes 0-2 whole and the xyz of 3-5.
#pragma once
#include <stdint.h>
#include <stddef.h>

namespace xr {

struct AVertex {
    float pos[3];
    float normal[3];
    float texCoord[2];
    int32_t materialIndex;
};
static_assert(sizeof(AVertex) == 36, "AVertex must be 36 bytes (Vertex.hlsl:12-33 reads 9 dwords)");

enum MaterialModel : uint32_t {
    kLambertian = 0,
    kMetallic = 1,
    kDielectric = 2,
    kIsotropic = 3,
    kDiffuseLight = 4,
};

struct alignas(16) AMaterial {
    float baseColor[4];
    int32_t baseColorTextureIndex;
    float fuzziness;
    float refractionIndex;
    uint32_t materialModel;
};
static_assert(sizeof(AMaterial) == 32, "AMaterial must be 32 bytes (Material.hlsl:7-14)");

struct UniformBufferObject {
    float view[16];               // column-major, glm convention
    float viewPrev[16];
    float projection[16];
    float projectionPrev[16];
    float viewInverse[16];
    float projectionInverse[16];
    uint32_t spp;
    uint32_t totalNumberOfSamples;
    uint32_t numberOfBounces;
    uint32_t randomSeed;
    uint32_t hasSkyBox;
    uint32_t reuse;
    uint32_t currentFrame;
};
static_assert(sizeof(UniformBufferObject) == 412, "UBO must be 412 bytes (UniformBuffer.h:16-32)");
static_assert(offsetof(UniformBufferObject, spp) == 384, "UBO scalar block starts after 6 mat4");
static_assert(offsetof(UniformBufferObject, currentFrame) == 408, "UBO currentFrame offset");

struct ReSTIRGISample {
    float x_view[3];
    uint32_t materialID;
    float n_view[3];
    uint32_t randSeed;
    float x_sample[3];
    float p_q;
    float n_sample[3];
    uint32_t pad1;
    float L_out[3];
    uint32_t pad2;
    float f[3];
    uint32_t pad3;
};
static_assert(sizeof(ReSTIRGISample) == 96, "ApplicationReSTIRGI.cpp:311");

struct ReSTIRGIReservoir {
    float w_sum;
    float W;
    uint32_t m;
    float pad;
    ReSTIRGISample s;
};
static_assert(sizeof(ReSTIRGIReservoir) == 112, "ApplicationReSTIRGI.cpp:326");

}  // namespace xr

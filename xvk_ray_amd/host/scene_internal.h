// The host-side scene: the reference's AScene buffers (src/Asset/Scene.h:51-68)
// as flat arrays, shared by the synthetic generators and the glTF loader.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "xr_host.h"

struct xrh_texture_data {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> rgba8;  // width * height * 4, rows top to bottom
};

struct xrh_scene {
    std::vector<xr_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<xr_material> materials;
    std::vector<uint32_t> offsets;  // (indexOffset, vertexOffset) per instance
    std::vector<xrh_texture_data> textures;
    std::vector<xr_texture> texture_desc;  // views of `textures` for xr_set_textures
};

namespace xrh_internal {
// PNG bytes -> RGBA8 (host/png_decode.cpp); false with a message on failure.
bool png_to_rgba(const uint8_t* data, size_t bytes, uint32_t* w, uint32_t* h, std::vector<uint8_t>* rgba,
                 std::string* err);
}  // namespace xrh_internal

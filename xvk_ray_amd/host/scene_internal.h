// The host-side scene: the reference's AScene buffers (src/Asset/Scene.h:51-68)
// as flat arrays, shared by the synthetic generators and the glTF loader.
#pragma once
#include <stdint.h>

#include <vector>

#include "xr_host.h"

struct xrh_scene {
    std::vector<xr_vertex> vertices;
    std::vector<uint32_t> indices;
    std::vector<xr_material> materials;
    std::vector<uint32_t> offsets;  // (indexOffset, vertexOffset) per instance
};

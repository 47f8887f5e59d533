// scene_encode.hpp — host encoding of a scene into the kernels' layout.
#pragma once
#include <cstdarg>
#include <string>
#include <vector>

#include "../../include/fnd_pt.h"
#include "bvh_build.hpp"
#include "pt_layout.hpp"

namespace fnd {

struct DeviceMaterial {     // 32 bytes, read by the shading kernels
    float albedo[3];
    uint32_t type;
    float emission[3];
    float roughness;
};

struct DeviceLight {        // 64 bytes
    float corner[3];
    float area;
    float edge0[3];
    float cdf;              // cumulative selection probability (power-proportional)
    float edge1[3];
    float pdf_select;
    float radiance[3];
    float pad;
};

struct HostScene {
    std::vector<WideNode> nodes;
    std::vector<LeafTri> tris;
    std::vector<InstanceRec> insts;
    std::vector<DeviceMaterial> materials;
    std::vector<DeviceLight> lights;
    uint32_t num_lights = 0;
    float background[3] = {0, 0, 0};
    uint32_t root = 0;
    bool two_level = false;
    fnd_scene_info info{};
};

fnd_status encode_scene(const fnd_scene_desc* d, HostScene& out, std::string& err);

}  // namespace fnd

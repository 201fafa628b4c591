// Host-side scene flattening of the core: material table, triangle records,
// light table and its alias table, BVH4.  Replaces BuildMaterialTable /
// BuildTriangles (Core/Source/Graphics/Cuda/PathTracing/Renderer.cpp:71-213),
// LightSampler::BuildSync (Core/Source/Graphics/Cuda/PathTracing/Memory/LightSampler.cpp:5-40)
// and AliasTable<T>::BuildSync
// (Core/Include/Core/Graphics/Cuda/PathTracing/Memory/AliasTable.hpp:22-84).
#pragma once

#include <array>
#include <string>
#include <vector>

#include "../../include/pt_core.h"
#include "bvh_build.hpp"
#include "pt_device.hpp"

namespace pt {

struct HostScene {
    std::vector<TriRecord> tris;          // BVH order
    std::vector<TriRecord> tris8;         // order of the 8-wide any-hit tree (empty without one)
    std::vector<ShadeRecord> shade;       // original order
    std::vector<MaterialRecord> materials;
    std::vector<LightRecord> lights;
    std::vector<AliasBin> bins;
    Bvh4Build bvh;
    uint32_t tri_count = 0;
    std::array<uint32_t, kQueueCount> tris_per_queue{};
};

// Alias table of AliasTable<T>::BuildSync: weights summed in double, p = w/sum,
// Vose-style pairing of the last under-full bin with the last over-full bin.
// Throws std::runtime_error if the weights do not sum to a positive value.
std::vector<AliasBin> build_alias_table(const std::vector<float>& weights, const std::vector<uint32_t>& values);

// Throws std::runtime_error with a message naming the bad input; the caller
// maps it to PT_ERROR_INVALID_ARGUMENT or PT_ERROR_UNSUPPORTED.
struct SceneError {
    pt_status status;
    std::string message;
};
// With build_bvh false the BVH and the BVH-ordered triangle records are left
// to the caller (the GPU builder).
bool build_host_scene(const pt_scene_desc& desc, const BvhBuildOptions& bvh_options, HostScene& out, SceneError& err,
                      bool build_bvh = true);

// MakeDeviceCamera (Renderer.cpp:237-257).
CameraBasis make_camera_basis(const pt_camera& camera, float aspect, const pt_motion_state& state);

}  // namespace pt

// GPU BVH4 builder (linear BVH over Morton codes, collapsed to BVH4 nodes).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_device.hpp"

namespace pt {

struct GpuBvhStats {
    uint32_t node_count = 0;   // BVH4 nodes written, root first
    uint32_t leaf_count = 0;
    uint32_t max_leaf = 0;
    uint32_t depth = 0;        // BVH4 levels, root included
    uint32_t bvh2_levels = 0;  // levels of the binary tree it was collapsed from
};

// Largest leaf the renderer asks of the GPU builder.
constexpr uint32_t kGpuBvhLeaf = 2;

// Room the builder needs for n triangles: one BVH4 node per binary internal node at most.
inline uint32_t gpu_bvh4_node_bound(uint32_t n) { return n > 1 ? n - 1 : 1; }

// Builds the BVH4 of n triangles (positions: 9 floats each, device memory) into
// d_nodes (at least gpu_bvh4_node_bound(n) nodes), the triangle records in
// BVH order into d_tris, and the original triangle of each BVH slot into
// d_order.  Leaves hold at most max_leaf (1..16) triangles.  Synchronises the
// stream (level sizes are read back); temporary memory is freed on return.
hipError_t gpu_build_bvh4(hipStream_t s, const float* d_positions, const uint32_t* d_material_ids,
                          const MaterialRecord* d_materials, uint32_t n, uint32_t max_leaf, Bvh4Node* d_nodes,
                          uint32_t node_capacity, TriRecord* d_tris, uint32_t* d_order, GpuBvhStats& stats);

// Quantized copies of n 4-wide nodes (the same bytes as quantize_bvh4_node on
// the host); asynchronous on stream s.
hipError_t gpu_quantize_bvh4(hipStream_t s, const Bvh4Node* d_nodes, uint32_t n, Bvh4QNode* d_out);

}  // namespace pt

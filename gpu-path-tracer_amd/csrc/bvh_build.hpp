// Host BVH builder of the core: binned-SAH BVH2, collapsed into 128-byte BVH4
// nodes.  Replaces HostBvh + DeviceBvh::BuildFlattenSync
// (Core/Source/Graphics/Cuda/Bvh/HostBvh.cpp:51-100, DeviceBvh.cpp:7-52),
// whose object-median BVH2 with 40-byte nodes is redesigned here for
// 128-byte cache lines and four-wide box tests.
#pragma once

#include <cstdint>
#include <vector>

#include "pt_device.hpp"

namespace pt {

struct Bvh4Build {
    std::vector<Bvh4Node> nodes;      // nodes[0] is the root; slots >= width stay empty
    std::vector<uint32_t> order;      // order[k] = original prim of BVH slot k
    uint32_t depth = 0;               // BVH4 levels below the root (root = 1)
    uint32_t leaf_count = 0;
    uint32_t max_leaf = 0;
    uint32_t width = 4;               // children per node: 4, or 2 (no collapse)
    uint32_t device_nodes = 0;        // node count of a tree built on the device (nodes and order stay empty)
    // the same binary tree collapsed into 8-wide quantized nodes (BvhBuildOptions::wide8),
    // breadth first, with the triangles of each node's leaf children consecutive:
    // order8[k] = original prim of position k of the any-hit tree's triangles
    std::vector<Bvh8QNode> qnodes8;
    std::vector<uint32_t> order8;
    uint32_t depth8 = 0;
};

// The nodes of a width-2 build in the 64-byte layout.
std::vector<Bvh2Node> to_bvh2(const Bvh4Build& b);

struct BvhBuildOptions {
    uint32_t max_leaf_size = 8;       // <= kMaxLeafSize
    uint32_t bins = 16;
    float traversal_cost = 1.0f;
    float intersection_cost = 1.0f;
    uint32_t threads = 0;             // 0 = hardware concurrency
    uint32_t width = 4;               // 4: collapse the BVH2 into BVH4 nodes; 2: keep it
    bool wide8 = false;               // also collapse it into 8-wide quantized nodes (width 4 only)
};

// positions: [T][3][3] floats.  Throws std::runtime_error on bad input.
Bvh4Build build_bvh4(const float* positions, uint32_t triangle_count, const BvhBuildOptions& options);

// Padding that the reference adds around every node box (HostBvh.cpp:37-48).
void pad_bounds(float lo[3], float hi[3]);

}  // namespace pt

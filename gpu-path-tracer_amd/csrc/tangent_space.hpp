// Tangent frames of indexed triangle meshes (host code).
#pragma once
#include <cstddef>
#include <cstdint>

namespace pt {

// Per-vertex tangents (x, y, z, sign) of the triangles `indices` by Mikkelsen's
// method with a 180-degree threshold; where faces sharing a vertex disagree,
// the last face in index order sets it.  Vertices no face uses get zeros.
void mikk_tangents(const float* positions, const float* normals, const float* uvs, uint32_t vertex_count,
                   const uint32_t* indices, size_t index_count, float* tangents);

}  // namespace pt

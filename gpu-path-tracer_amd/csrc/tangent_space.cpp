// Per-vertex tangent frames of an indexed triangle mesh by Mikkelsen's method
// (M. S. Mikkelsen, "Simulation of Wrinkled Surfaces Revisited", 2008; the
// method of the mikktspace library), with the settings the reference uses on
// every imported submesh: TangentSpace.cpp:66-89 runs genTangSpaceDefault
// (angular threshold 180 degrees) over triangles and writes each corner's
// (tangent, sign) into the shared vertex, so the last face in index order
// wins where faces disagree (TangentSpace.cpp:57-63).
//
// The steps and the float expressions follow the method as the vendored
// library evaluates it (External/mikktspace/mikktspace.c), because normal
// mapping must see the same frames:
//   1. corners with equal position, normal and uv are one vertex (exact
//      float equality);
//   2. a face with two equal corner positions is degenerate; it takes its
//      frames from good faces at the end;
//   3. each good face gets unit first-order uv derivatives dP/ds, dP/dt
//      (signed by the uv orientation), their magnitudes, an orientation bit
//      and a "group with any" bit when its uv area or a derivative is ~0;
//   4. faces are neighbours across an edge whose two welded corners they
//      share in opposite directions (the first unpaired candidate in face
//      order when more than two faces share the edge);
//   5. around every welded vertex, faces reachable over neighbouring edges
//      with the same orientation form a group (a "group with any" face takes
//      the orientation of the first group that reaches it);
//   6. a group's frame at the vertex is the angle-weighted mean of its faces'
//      derivatives projected into the vertex's tangent plane;
//   7. corners no group reached keep (1, 0, 0) with sign -1.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "pt_core.h"
#include "tangent_space.hpp"

namespace pt {
namespace {

struct V3 {
    float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(float s, V3 v) { return {s * v.x, s * v.y, s * v.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float length(V3 v) { return std::sqrt(dot(v, v)); }
inline bool nonzero(float x) { return std::fabs(x) > FLT_MIN; }
inline bool nonzero(V3 v) { return nonzero(v.x) || nonzero(v.y) || nonzero(v.z); }
inline V3 unit(V3 v) { return mul(1.0f / length(v), v); }
inline bool same(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
// v without its component along n, at unit length unless it vanished
inline V3 in_plane(V3 v, V3 n) {
    const V3 r = sub(v, mul(dot(n, v), n));
    return nonzero(r) ? unit(r) : r;
}

struct Frame {
    V3 s{1.0f, 0.0f, 0.0f};
    V3 t{0.0f, 1.0f, 0.0f};
    float mag_s = 1.0f, mag_t = 1.0f;
    bool orient = false;
};

struct Face {
    uint32_t corner;      // first corner (3 * original face index)
    uint32_t w[3];        // welded vertex ids of the corners
    V3 ds{0, 0, 0}, dt{0, 0, 0};
    float mag_s = 0.0f, mag_t = 0.0f;
    bool orient = false;
    bool any = true;      // "group with any": no usable uv derivatives
    int32_t nb[3] = {-1, -1, -1};      // neighbour across edge k (corner k -> k+1)
    int32_t group[3] = {-1, -1, -1};
};

struct Group {
    uint32_t vertex;      // welded vertex id
    bool orient;
    std::vector<int32_t> faces;
};

class Mesh {
public:
    Mesh(const float* p, const float* n, const float* uv, const uint32_t* idx) : p_(p), n_(n), uv_(uv), idx_(idx) {}
    V3 pos(uint32_t c) const { return load3(p_, idx_[c]); }
    V3 nrm(uint32_t c) const { return load3(n_, idx_[c]); }
    float u(uint32_t c) const { return uv_[2 * (size_t)idx_[c]]; }
    float v(uint32_t c) const { return uv_[2 * (size_t)idx_[c] + 1]; }

private:
    static V3 load3(const float* a, uint32_t i) { return {a[3 * (size_t)i], a[3 * (size_t)i + 1], a[3 * (size_t)i + 2]}; }
    const float *p_, *n_, *uv_;
    const uint32_t* idx_;
};

// Welded id per vertex: vertices with equal position, normal and uv share one
// (+0 and -0 compare equal; a vertex with a NaN equals nothing, not even its
// own other corners, so those get an id per corner).
std::vector<uint32_t> weld(const float* p, const float* n, const float* uv, uint32_t nv, const uint32_t* idx,
                           size_t corners) {
    struct Key {
        float a[8];
    };
    auto bits = [](float x) {
        uint32_t b;
        const float y = x == 0.0f ? 0.0f : x;
        std::memcpy(&b, &y, 4);
        return b;
    };
    struct Hash {
        decltype(bits) f;
        size_t operator()(const Key& k) const {
            uint64_t h = 1469598103934665603ULL;
            for (float x : k.a) h = (h ^ f(x)) * 1099511628211ULL;
            return (size_t)h;
        }
    };
    struct Eq {
        bool operator()(const Key& x, const Key& y) const {
            for (int i = 0; i < 8; i++)
                if (!(x.a[i] == y.a[i])) return false;
            return true;
        }
    };
    std::vector<uint32_t> vid(nv, UINT32_MAX);
    std::unordered_map<Key, uint32_t, Hash, Eq> seen(16, Hash{bits});
    uint32_t next = 0;
    std::vector<uint32_t> wc(corners);
    for (size_t c = 0; c < corners; c++) {
        const uint32_t i = idx[c];
        if (vid[i] == UINT32_MAX) {
            const Key k{{p[3 * (size_t)i], p[3 * (size_t)i + 1], p[3 * (size_t)i + 2], n[3 * (size_t)i],
                         n[3 * (size_t)i + 1], n[3 * (size_t)i + 2], uv[2 * (size_t)i], uv[2 * (size_t)i + 1]}};
            bool nan = false;
            for (float x : k.a) nan = nan || std::isnan(x);
            if (nan) {
                wc[c] = next++;
                continue;
            }
            vid[i] = seen.try_emplace(k, next).first->second;
            if (vid[i] == next) next++;
        }
        wc[c] = vid[i];
    }
    return wc;
}

// Step 3: derivatives, magnitudes and flags of one good face.
void face_derivatives(const Mesh& m, Face& f) {
    const uint32_t c = f.corner;
    const V3 p1 = m.pos(c), p2 = m.pos(c + 1), p3 = m.pos(c + 2);
    const float t21x = m.u(c + 1) - m.u(c), t21y = m.v(c + 1) - m.v(c);
    const float t31x = m.u(c + 2) - m.u(c), t31y = m.v(c + 2) - m.v(c);
    const V3 d1 = sub(p2, p1), d2 = sub(p3, p1);
    const float area2 = t21x * t31y - t21y * t31x;
    const V3 ds = sub(mul(t31y, d1), mul(t21y, d2));
    const V3 dt = add(mul(-t31x, d1), mul(t21x, d2));
    f.orient = area2 > 0.0f;
    if (!nonzero(area2)) return;
    const float a = std::fabs(area2), ls = length(ds), lt = length(dt);
    const float sign = f.orient ? 1.0f : -1.0f;
    if (nonzero(ls)) f.ds = mul(sign / ls, ds);
    if (nonzero(lt)) f.dt = mul(sign / lt, dt);
    f.mag_s = ls / a;
    f.mag_t = lt / a;
    if (nonzero(f.mag_s) && nonzero(f.mag_t)) f.any = false;
}

// Step 4: pair edges by welded endpoints, (lo, hi, face) order.
void find_neighbours(std::vector<Face>& faces) {
    struct Edge {
        uint32_t lo, hi;
        int32_t face;
        int32_t k;
    };
    std::vector<Edge> e;
    e.reserve(faces.size() * 3);
    for (int32_t f = 0; f < (int32_t)faces.size(); f++)
        for (int32_t k = 0; k < 3; k++) {
            const uint32_t a = faces[f].w[k], b = faces[f].w[(k + 1) % 3];
            e.push_back({std::min(a, b), std::max(a, b), f, k});
        }
    std::sort(e.begin(), e.end(), [](const Edge& x, const Edge& y) {
        if (x.lo != y.lo) return x.lo < y.lo;
        if (x.hi != y.hi) return x.hi < y.hi;
        return x.face < y.face;
    });
    for (size_t i = 0; i < e.size(); i++) {
        Face& fa = faces[e[i].face];
        const int32_t ka = e[i].k;
        if (fa.nb[ka] != -1) continue;
        const uint32_t a0 = fa.w[ka], a1 = fa.w[(ka + 1) % 3];
        for (size_t j = i + 1; j < e.size() && e[j].lo == e[i].lo && e[j].hi == e[i].hi; j++) {
            Face& fb = faces[e[j].face];
            const int32_t kb = e[j].k;
            if (fb.w[kb] == a1 && fb.w[(kb + 1) % 3] == a0 && fb.nb[kb] == -1) {
                fa.nb[ka] = e[j].face;
                fb.nb[kb] = e[i].face;
                break;
            }
        }
    }
}

int corner_of(const Face& f, uint32_t vertex) {
    for (int k = 0; k < 3; k++)
        if (f.w[k] == vertex) return k;
    return -1;
}

// Step 5: groups seeded in face-corner order, grown depth-first over the
// neighbour across the corner's outgoing edge, then its incoming edge.
std::vector<Group> build_groups(std::vector<Face>& faces) {
    std::vector<Group> groups;
    std::vector<int32_t> stack;
    for (int32_t f = 0; f < (int32_t)faces.size(); f++)
        for (int k = 0; k < 3; k++) {
            if (faces[f].any || faces[f].group[k] >= 0) continue;
            const int32_t g = (int32_t)groups.size();
            groups.push_back(Group{faces[f].w[k], faces[f].orient, {f}});
            faces[f].group[k] = g;
            Group& grp = groups.back();
            stack.clear();
            if (faces[f].nb[(k + 2) % 3] >= 0) stack.push_back(faces[f].nb[(k + 2) % 3]);
            if (faces[f].nb[k] >= 0) stack.push_back(faces[f].nb[k]);
            while (!stack.empty()) {
                const int32_t t = stack.back();
                stack.pop_back();
                Face& ft = faces[t];
                const int kt = corner_of(ft, grp.vertex);
                if (kt < 0 || ft.group[kt] >= 0) continue;
                if (ft.any && ft.group[0] < 0 && ft.group[1] < 0 && ft.group[2] < 0) ft.orient = grp.orient;
                if (ft.orient != grp.orient) continue;
                grp.faces.push_back(t);
                ft.group[kt] = g;
                if (ft.nb[(kt + 2) % 3] >= 0) stack.push_back(ft.nb[(kt + 2) % 3]);
                if (ft.nb[kt] >= 0) stack.push_back(ft.nb[kt]);
            }
        }
    return groups;
}

// Step 6: the frame of a set of faces (ascending) at one welded vertex.
Frame mean_frame(const Mesh& m, const std::vector<Face>& faces, const std::vector<int32_t>& members, uint32_t vertex) {
    Frame r;
    r.s = r.t = V3{0, 0, 0};
    r.mag_s = r.mag_t = 0.0f;
    float angle_sum = 0.0f;
    for (int32_t fi : members) {
        const Face& f = faces[fi];
        if (f.any) continue;
        const int k = corner_of(f, vertex);
        const V3 n = m.nrm(f.corner + k);
        const V3 s = in_plane(f.ds, n), t = in_plane(f.dt, n);
        const V3 p_prev = m.pos(f.corner + (k + 2) % 3), p = m.pos(f.corner + k), p_next = m.pos(f.corner + (k + 1) % 3);
        const V3 e1 = in_plane(sub(p_prev, p), n), e2 = in_plane(sub(p_next, p), n);
        const float c = std::clamp(dot(e1, e2), -1.0f, 1.0f);
        const float angle = (float)std::acos((double)c);
        r.s = add(r.s, mul(angle, s));
        r.t = add(r.t, mul(angle, t));
        r.mag_s += angle * f.mag_s;
        r.mag_t += angle * f.mag_t;
        angle_sum += angle;
    }
    if (nonzero(r.s)) r.s = unit(r.s);
    if (nonzero(r.t)) r.t = unit(r.t);
    if (angle_sum > 0.0f) {
        r.mag_s /= angle_sum;
        r.mag_t /= angle_sum;
    }
    return r;
}

}  // namespace

void mikk_tangents(const float* positions, const float* normals, const float* uvs, uint32_t vertex_count,
                   const uint32_t* indices, size_t index_count, float* tangents) {
    std::fill(tangents, tangents + 4 * (size_t)vertex_count, 0.0f);
    const size_t nf = index_count / 3;
    if (nf == 0) return;
    const Mesh m(positions, normals, uvs, indices);
    const std::vector<uint32_t> wc = weld(positions, normals, uvs, vertex_count, indices, 3 * nf);

    // step 2: good faces keep their order, degenerate ones are set aside
    std::vector<Face> faces;
    std::vector<uint32_t> degenerate;
    faces.reserve(nf);
    for (size_t f = 0; f < nf; f++) {
        const uint32_t c = (uint32_t)(3 * f);
        const V3 a = m.pos(c), b = m.pos(c + 1), d = m.pos(c + 2);
        if (same(a, b) || same(a, d) || same(b, d)) {
            degenerate.push_back(c);
            continue;
        }
        Face x;
        x.corner = c;
        for (int k = 0; k < 3; k++) x.w[k] = wc[c + k];
        face_derivatives(m, x);
        faces.push_back(x);
    }
    find_neighbours(faces);
    const std::vector<Group> groups = build_groups(faces);

    std::vector<Frame> frame(3 * nf);
    const float thres = (float)std::cos((double)((180.0f * (float)M_PI) / 180.0f));
    std::vector<std::vector<int32_t>> subsets;
    std::vector<Frame> subset_frame;
    std::vector<int32_t> members;
    for (size_t g = 0; g < groups.size(); g++) {
        const Group& grp = groups[g];
        subsets.clear();
        subset_frame.clear();
        for (int32_t fi : grp.faces) {
            const Face& f = faces[fi];
            int k = 0;
            while (f.group[k] != (int32_t)g) k++;
            const V3 n = m.nrm(f.corner + k);
            const V3 s = in_plane(f.ds, n), t = in_plane(f.dt, n);
            // faces of the group whose projected derivatives are within the threshold
            members.clear();
            for (int32_t ti : grp.faces) {
                const Face& o = faces[ti];
                const V3 s2 = in_plane(o.ds, n), t2 = in_plane(o.dt, n);
                if (f.any || o.any || ti == fi || (dot(s, s2) > thres && dot(t, t2) > thres)) members.push_back(ti);
            }
            std::sort(members.begin(), members.end());
            size_t l = 0;
            while (l < subsets.size() && subsets[l] != members) l++;
            if (l == subsets.size()) {
                subsets.push_back(members);
                subset_frame.push_back(mean_frame(m, faces, members, grp.vertex));
            }
            Frame& out = frame[f.corner + k];
            out = subset_frame[l];
            out.orient = grp.orient;
        }
    }

    // step 2, end: degenerate corners copy the frame of the first good corner
    // with the same welded vertex
    if (!degenerate.empty()) {
        std::unordered_map<uint32_t, uint32_t> first;
        for (const Face& f : faces)
            for (int k = 0; k < 3; k++) first.try_emplace(f.w[k], f.corner + k);
        for (uint32_t c : degenerate)
            for (int k = 0; k < 3; k++) {
                auto it = first.find(wc[c + k]);
                if (it != first.end()) frame[c + k] = frame[it->second];
            }
    }

    for (size_t c = 0; c < 3 * nf; c++) {
        float* t = tangents + 4 * (size_t)indices[c];
        t[0] = frame[c].s.x;
        t[1] = frame[c].s.y;
        t[2] = frame[c].s.z;
        t[3] = frame[c].orient ? 1.0f : -1.0f;
    }
}

}  // namespace pt

extern "C" pt_status pt_generate_tangents(const float* positions, const float* normals, const float* uvs,
                                          uint32_t vertex_count, const uint32_t* indices, uint32_t index_count,
                                          float* tangents) {
    if (index_count % 3 != 0) return PT_ERROR_INVALID_ARGUMENT;
    if (vertex_count > 0 && (!positions || !normals || !uvs || !tangents)) return PT_ERROR_INVALID_ARGUMENT;
    if (index_count > 0 && !indices) return PT_ERROR_INVALID_ARGUMENT;
    for (uint32_t i = 0; i < index_count; i++)
        if (indices[i] >= vertex_count) return PT_ERROR_INVALID_ARGUMENT;
    try {
        pt::mikk_tangents(positions, normals, uvs, vertex_count, indices, index_count, tangents);
    } catch (const std::bad_alloc&) {
        return PT_ERROR_OUT_OF_MEMORY;
    }
    return PT_OK;
}

#include "scene.hpp"

#include <cmath>
#include <stdexcept>
#include <string>

namespace pt {

std::vector<AliasBin> build_alias_table(const std::vector<float>& weights, const std::vector<uint32_t>& values) {
    const size_t n = weights.size();
    double acc = 0.0;
    for (float w : weights) acc = acc + w;
    const float sum = (float)acc;
    if (!(sum > 0.0f)) throw std::runtime_error("light weights must sum to a positive value");
    std::vector<AliasBin> bins(n);
    for (size_t i = 0; i < n; i++) bins[i] = AliasBin{0.0f, weights[i] / sum, 0u, values[i]};

    struct Outcome {
        float p_hat;
        uint32_t index;
    };
    std::vector<Outcome> under, over;
    for (uint32_t i = 0; i < (uint32_t)n; i++) {
        const float p_hat = bins[i].p * (float)n;
        if (p_hat < 1.0f) under.push_back(Outcome{p_hat, i});
        else over.push_back(Outcome{p_hat, i});
    }
    while (!under.empty() && !over.empty()) {
        const Outcome un = under.back(), ov = over.back();
        under.pop_back();
        over.pop_back();
        bins[un.index].q = un.p_hat;
        bins[un.index].alias = ov.index;
        const float excess = un.p_hat + ov.p_hat - 1.0f;
        if (excess < 1.0f) under.push_back(Outcome{excess, ov.index});
        else over.push_back(Outcome{excess, ov.index});
    }
    while (!over.empty()) {
        bins[over.back().index].q = 1.0f;
        bins[over.back().index].alias = kInvalid;
        over.pop_back();
    }
    while (!under.empty()) {
        bins[under.back().index].q = 1.0f;
        bins[under.back().index].alias = kInvalid;
        under.pop_back();
    }
    return bins;
}

namespace {

int queue_of_model(uint32_t model) {
    switch (model) {
        case PT_MODEL_DIFFUSE: return 0;
        case PT_MODEL_GGX: return 1;
        case PT_MODEL_MIRROR: return 2;
        case PT_MODEL_EMISSIVE: return 3;
        default: return -1;
    }
}

V3 load3(const float* p) { return v3(p[0], p[1], p[2]); }

bool finite3(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }

// Fallback tangent of ComputeWorldTangent (Renderer.cpp:116-142).
V3 fallback_tangent(V3 n) {
    const V3 axis = std::fabs(n.z) < 0.999f ? v3(0.0f, 0.0f, 1.0f) : v3(1.0f, 0.0f, 0.0f);
    return normalize(cross(axis, n));
}

}  // namespace

bool build_host_scene(const pt_scene_desc& d, const BvhBuildOptions& bvh_options, HostScene& out, SceneError& err,
                      bool build_bvh) {
    const uint32_t T = d.triangle_count;
    if (T > 0 && (!d.positions || !d.normals || !d.material_ids)) {
        err = {PT_ERROR_INVALID_ARGUMENT, "scene needs positions, normals and material ids"};
        return false;
    }
    if (d.material_count > 0 && !d.materials) {
        err = {PT_ERROR_INVALID_ARGUMENT, "material_count > 0 but materials is NULL"};
        return false;
    }
    out = HostScene();
    out.tri_count = T;

    // materials (BuildMaterialTable, Renderer.cpp:71-106)
    out.materials.resize(d.material_count);
    for (uint32_t m = 0; m < d.material_count; m++) {
        const pt_material& s = d.materials[m];
        const int q = queue_of_model(s.model);
        if (q < 0) {
            err = {s.model < PT_MODEL_COUNT ? PT_ERROR_UNSUPPORTED : PT_ERROR_INVALID_ARGUMENT,
                   "material " + std::to_string(m) + ": shading model " + std::to_string(s.model) +
                       " is not one of diffuse, ggx, mirror, emissive"};
            return false;
        }
        MaterialRecord& r = out.materials[m];
        r = MaterialRecord{};
        for (int k = 0; k < 3; k++) {
            r.color[k] = s.color[k];
            r.specular[k] = s.specular[k];
            r.emission[k] = s.emission[k];
        }
        r.queue = (uint32_t)q;
        r.model = s.model;
        r.alpha = fmaxf(s.roughness * s.roughness, 0.001f);
        r.metallic = s.metallic;
        r.ior = s.ior;
        const V3 nts = load3(s.normal_ts);
        if (!(dot(nts, nts) > 0.0f) || !finite3(nts)) {
            err = {PT_ERROR_INVALID_ARGUMENT, "material " + std::to_string(m) + ": normal_ts must be a nonzero vector"};
            return false;
        }
        const V3 nn = normalize(nts);
        r.nts[0] = nn.x;
        r.nts[1] = nn.y;
        r.nts[2] = nn.z;
    }

    // triangles (BuildTriangles, Renderer.cpp:159-213)
    out.shade.resize(T);
    std::vector<float> light_w;
    std::vector<uint32_t> light_idx;
    for (uint32_t t = 0; t < T; t++) {
        const uint32_t mid = d.material_ids[t];
        if (mid >= d.material_count) {
            err = {PT_ERROR_INVALID_ARGUMENT, "triangle " + std::to_string(t) + ": material id out of range"};
            return false;
        }
        const float* P = d.positions + 9 * (size_t)t;
        const V3 p0 = load3(P), p1 = load3(P + 3), p2 = load3(P + 6);
        if (!finite3(p0) || !finite3(p1) || !finite3(p2)) {
            err = {PT_ERROR_INVALID_ARGUMENT, "triangle " + std::to_string(t) + " has a non-finite vertex"};
            return false;
        }
        const V3 e1 = p1 - p0, e2 = p2 - p0;
        ShadeRecord& s = out.shade[t];
        s = ShadeRecord{};
        for (int k = 0; k < 3; k++) {
            const V3 n = load3(d.normals + 9 * (size_t)t + 3 * k);
            s.n[k][0] = n.x;
            s.n[k][1] = n.y;
            s.n[k][2] = n.z;
            s.n[k][3] = 0.0f;
            if (d.tangents) {
                const float* tg = d.tangents + 12 * (size_t)t + 4 * k;
                for (int c = 0; c < 4; c++) s.t[k][c] = tg[c];
            } else {
                const V3 tg = fallback_tangent(n);
                s.t[k][0] = tg.x;
                s.t[k][1] = tg.y;
                s.t[k][2] = tg.z;
                s.t[k][3] = 1.0f;
            }
        }
        const V3 gn = normalize(cross(e1, e2));   // ComputeGeometricNormal, Renderer.cpp:144-157
        s.gn[0] = gn.x;
        s.gn[1] = gn.y;
        s.gn[2] = gn.z;
        s.material = mid;
        const MaterialRecord& mr = out.materials[mid];
        out.tris_per_queue[mr.queue]++;
        if (mr.model == PT_MODEL_EMISSIVE) {
            // LightSampler::BuildSync, LightSampler.cpp:5-40
            const float area = 0.5f * length(cross(e1, e2));
            LightRecord L{};
            L.v0[0] = p0.x; L.v0[1] = p0.y; L.v0[2] = p0.z;
            L.e1[0] = e1.x; L.e1[1] = e1.y; L.e1[2] = e1.z;
            L.e2[0] = e2.x; L.e2[1] = e2.y; L.e2[2] = e2.z;
            L.gn[0] = gn.x; L.gn[1] = gn.y; L.gn[2] = gn.z;
            L.area = area;
            L.prim = t;
            for (int k = 0; k < 3; k++) L.emission[k] = mr.emission[k];
            const float w = d.light_weights ? d.light_weights[t] : area;
            if (!(w >= 0.0f) || !std::isfinite(w)) {
                err = {PT_ERROR_INVALID_ARGUMENT, "triangle " + std::to_string(t) + ": light weight must be finite and >= 0"};
                return false;
            }
            light_idx.push_back((uint32_t)out.lights.size());
            light_w.push_back(w);
            out.lights.push_back(L);
        }
    }
    if (!out.lights.empty()) {
        try {
            out.bins = build_alias_table(light_w, light_idx);
        } catch (const std::exception& e) {
            err = {PT_ERROR_INVALID_ARGUMENT, e.what()};
            return false;
        }
    }

    if (!build_bvh) return true;
    try {
        out.bvh = build_bvh4(d.positions, T, bvh_options);
    } catch (const std::exception& e) {
        err = {PT_ERROR_INVALID_ARGUMENT, std::string("BVH build failed: ") + e.what()};
        return false;
    }
    // triangle records in the order of a tree's leaves (BVH4, and the 8-wide
    // any-hit tree when it was built)
    auto records = [&](const std::vector<uint32_t>& order, std::vector<TriRecord>& recs) {
        recs.resize(order.size());
        for (size_t k = 0; k < order.size(); k++) {
            const uint32_t t = order[k];
            const float* P = d.positions + 9 * (size_t)t;
            const V3 p0 = load3(P), p1 = load3(P + 3), p2 = load3(P + 6);
            const V3 e1 = p1 - p0, e2 = p2 - p0;
            TriRecord& r = recs[k];
            r.v0[0] = p0.x; r.v0[1] = p0.y; r.v0[2] = p0.z;
            r.e1[0] = e1.x; r.e1[1] = e1.y; r.e1[2] = e1.z;
            r.e2[0] = e2.x; r.e2[1] = e2.y; r.e2[2] = e2.z;
            r.prim = t;
            r.material = d.material_ids[t];
            r.queue = out.materials[r.material].queue;
        }
    };
    records(out.bvh.order, out.tris);
    records(out.bvh.order8, out.tris8);
    return true;
}

namespace {
// glm-style rotation of v by unit quaternion q = (x, y, z, w).
V3 rotate(const float q[4], V3 v) {
    const V3 qv = v3(q[0], q[1], q[2]);
    const V3 uv = cross(qv, v);
    const V3 uuv = cross(qv, uv);
    return v + ((uv * q[3]) + uuv) * 2.0f;
}
}  // namespace

CameraBasis make_camera_basis(const pt_camera& camera, float aspect, const pt_motion_state& state) {
    const V3 forward = rotate(state.rotation, v3(0.0f, 0.0f, -1.0f));
    const V3 right = rotate(state.rotation, v3(1.0f, 0.0f, 0.0f));
    const V3 up = rotate(state.rotation, v3(0.0f, 1.0f, 0.0f));
    const float half_h = std::tan(camera.fov_y * 0.5f);
    const float half_w = half_h * aspect;
    const V3 origin = load3(state.position);
    const V3 horizontal = 2.0f * half_w * right;
    const V3 vertical = 2.0f * half_h * up;
    const V3 llc = origin + forward - half_w * right - half_h * up;
    CameraBasis b;
    const V3* src[4] = {&origin, &llc, &horizontal, &vertical};
    float* dst[4] = {b.origin, b.llc, b.horizontal, b.vertical};
    for (int i = 0; i < 4; i++) {
        dst[i][0] = src[i]->x;
        dst[i][1] = src[i]->y;
        dst[i][2] = src[i]->z;
    }
    return b;
}

}  // namespace pt

// Wavefront OBJ/MTL import into a pt_scene_desc, following the reference's
// import path: ObjLoader.cpp:67-234 (triangulated faces, per-vertex normals
// and uvs required, vertices deduplicated by (v, vt, vn), submeshes per shape
// and material) and Manager.cpp:311-412 (constant material channels stored as
// 1x1 textures: colour and specular as sRGB u8, the default RMA texel, fp32
// emission, the default normal-map texel), with tangents by Mikkelsen's
// method per submesh (TangentSpace.cpp:85-89).  The parser is written from
// the OBJ/MTL format description; the reference parses with tinyobjloader.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "pt_core.h"
#include "tangent_space.hpp"

struct pt_imported_scene {
    std::vector<float> positions, normals, tangents;
    std::vector<uint32_t> material_ids;
    std::vector<pt_material> materials;
    pt_scene_desc desc{};
};

namespace pt {
namespace {

struct ImportError {
    std::string message;
};

// Linear value -> sRGB u8 (Manager.cpp:64-79: clamp, IEC 61966-2-1 encode,
// lround) and the sRGB decode the texture unit applies when it is sampled.
uint8_t linear_to_srgb_u8(float x) {
    if (!std::isfinite(x)) x = 0.0f;
    x = std::clamp(x, 0.0f, 1.0f);
    const float s = x <= 0.0031308f ? 12.92f * x : 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f;
    return (uint8_t)std::lround(s * 255.0f);
}
float srgb_u8_to_linear(uint8_t u) {
    const double c = u / 255.0;
    return (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
}
float quantized_srgb(float x) { return srgb_u8_to_linear(linear_to_srgb_u8(x)); }

// `surface` key of the MTL (ObjLoader.cpp:51-64); Unlit, the default, and
// unknown values are shaded as Diffuse by the path tracer (Material.hpp:70).
uint32_t model_of_surface(const std::string& v, bool present) {
    if (!present) return PT_MODEL_DIFFUSE;
    if (v == "normal") return PT_MODEL_NORMAL;
    if (v == "diffuse" || v == "lambert" || v == "unlit") return PT_MODEL_DIFFUSE;
    if (v == "mirror") return PT_MODEL_MIRROR;
    if (v == "phong") return PT_MODEL_PHONG;
    if (v == "microfacet" || v == "ggx") return PT_MODEL_GGX;
    if (v == "emissive") return PT_MODEL_EMISSIVE;
    return PT_MODEL_DIFFUSE;
}

struct MtlEntry {
    std::string name;
    float kd[3] = {1.0f, 1.0f, 1.0f};   // tinyobj's defaults before Kd is read
    float ks[3] = {0.0f, 0.0f, 0.0f};
    float ke[3] = {0.0f, 0.0f, 0.0f};
    float ni = 1.0f;
    std::string surface;
    bool has_surface = false;
};

pt_material material_of(const MtlEntry& m) {
    pt_material r{};
    r.model = model_of_surface(m.surface, m.has_surface);
    for (int k = 0; k < 3; k++) {
        r.color[k] = quantized_srgb(m.kd[k]);
        r.specular[k] = quantized_srgb(m.ks[k]);
        r.emission[k] = m.ke[k];
    }
    // the RMA channel is the default texel (255, 0, 255) whatever the MTL says
    // (Manager.cpp:363-366 passes DefaultRma, Material.hpp:29)
    r.roughness = 1.0f;
    r.metallic = 0.0f;
    r.ior = m.ni;
    // DefaultNormal (128, 128, 255) decoded to [-1, 1] (Material.hpp:31)
    r.normal_ts[0] = 128.0f / 255.0f * 2.0f - 1.0f;
    r.normal_ts[1] = 128.0f / 255.0f * 2.0f - 1.0f;
    r.normal_ts[2] = 1.0f;
    return r;
}

// The default material (Manager.cpp ImportDefaultMaterial): white, Unlit.
pt_material default_material() {
    MtlEntry m;
    m.kd[0] = m.kd[1] = m.kd[2] = 1.0f;
    return material_of(m);
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

void read_mtl(const std::filesystem::path& file, std::vector<MtlEntry>& out) {
    std::ifstream in(file);
    if (!in) throw ImportError{"cannot open material library " + file.string()};
    std::string line;
    MtlEntry* cur = nullptr;
    while (std::getline(in, line)) {
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        std::istringstream ss(t);
        std::string key;
        ss >> key;
        if (key == "newmtl") {
            out.emplace_back();
            cur = &out.back();
            std::getline(ss, cur->name);
            cur->name = trim(cur->name);
            continue;
        }
        if (!cur) continue;
        auto read3 = [&](float v[3]) {
            ss >> v[0];
            if (!(ss >> v[1])) v[1] = v[2] = v[0];
            else ss >> v[2];
        };
        if (key == "Kd") read3(cur->kd);
        else if (key == "Ks") read3(cur->ks);
        else if (key == "Ke") read3(cur->ke);
        else if (key == "Ni") ss >> cur->ni;
        else if (key == "surface") {
            std::getline(ss, cur->surface);
            cur->surface = trim(cur->surface);
            cur->has_surface = true;
        }
    }
}

struct Corner {
    int v, t, n;
};
bool operator==(const Corner& a, const Corner& b) { return a.v == b.v && a.t == b.t && a.n == b.n; }
struct CornerHash {
    size_t operator()(const Corner& c) const {
        return std::hash<long long>()(((long long)c.v * 73856093LL) ^ ((long long)c.t * 19349663LL) ^
                                      ((long long)c.n * 83492791LL));
    }
};

// One OBJ index (1-based, or negative = relative to the end) to 0-based.
int resolve(long idx, size_t count, const char* what, size_t line_no) {
    long r = idx > 0 ? idx - 1 : (long)count + idx;
    if (idx == 0 || r < 0 || r >= (long)count)
        throw ImportError{"line " + std::to_string(line_no) + ": " + what + " index " + std::to_string(idx) +
                          " out of range"};
    return (int)r;
}

struct Face {
    Corner c[3];
    int material;   // index into the library, -1: none
};

struct Shape {
    std::vector<Face> faces;
};

void import_obj(const std::filesystem::path& path, pt_imported_scene& out) {
    std::ifstream in(path);
    if (!in) throw ImportError{"cannot open " + path.string()};
    std::vector<float> vp, vt, vn;
    std::vector<MtlEntry> lib;
    std::vector<Shape> shapes(1);
    int cur_mat = -1;
    std::string line;
    size_t line_no = 0;
    std::vector<std::string> pending_use;    // usemtl names before their library was read
    while (std::getline(in, line)) {
        line_no++;
        const std::string t = trim(line);
        if (t.empty() || t[0] == '#') continue;
        std::istringstream ss(t);
        std::string key;
        ss >> key;
        if (key == "v") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            vp.insert(vp.end(), {x, y, z});
        } else if (key == "vt") {
            float u = 0, v = 0;
            ss >> u;
            ss >> v;
            vt.insert(vt.end(), {u, v});
        } else if (key == "vn") {
            float x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            vn.insert(vn.end(), {x, y, z});
        } else if (key == "mtllib") {
            std::string name;
            while (ss >> name) read_mtl(path.parent_path() / name, lib);
        } else if (key == "usemtl") {
            std::string name;
            std::getline(ss, name);
            name = trim(name);
            cur_mat = -1;
            for (size_t m = 0; m < lib.size(); m++)
                if (lib[m].name == name) cur_mat = (int)m;
        } else if (key == "o" || key == "g") {
            if (!shapes.back().faces.empty()) shapes.emplace_back();
        } else if (key == "f") {
            std::vector<Corner> poly;
            std::string tok;
            while (ss >> tok) {
                long iv = 0, it = 0, in_ = 0;
                bool has_t = false, has_n = false;
                const size_t s1 = tok.find('/');
                iv = std::stol(tok.substr(0, s1));
                if (s1 != std::string::npos) {
                    const size_t s2 = tok.find('/', s1 + 1);
                    const std::string ts = tok.substr(s1 + 1, s2 == std::string::npos ? std::string::npos : s2 - s1 - 1);
                    if (!ts.empty()) {
                        it = std::stol(ts);
                        has_t = true;
                    }
                    if (s2 != std::string::npos && s2 + 1 < tok.size()) {
                        in_ = std::stol(tok.substr(s2 + 1));
                        has_n = true;
                    }
                }
                // ObjLoader.cpp:105-127: normals and uvs are required
                if (!has_n)
                    throw ImportError{"line " + std::to_string(line_no) + ": normal data is required but not found"};
                if (!has_t)
                    throw ImportError{"line " + std::to_string(line_no) + ": uv data is required but not found"};
                poly.push_back(Corner{resolve(iv, vp.size() / 3, "vertex", line_no),
                                      resolve(it, vt.size() / 2, "uv", line_no),
                                      resolve(in_, vn.size() / 3, "normal", line_no)});
            }
            if (poly.size() < 3) throw ImportError{"line " + std::to_string(line_no) + ": a face needs three corners"};
            for (size_t k = 1; k + 1 < poly.size(); k++)   // fan triangulation
                shapes.back().faces.push_back(Face{{poly[0], poly[k], poly[k + 1]}, cur_mat});
        }
    }

    // materials: the library's, then the default one for faces without
    out.materials.clear();
    for (const MtlEntry& m : lib) out.materials.push_back(material_of(m));
    const uint32_t default_id = (uint32_t)out.materials.size();
    bool need_default = false;

    // submeshes per shape and material (ascending id, none first), vertices
    // deduplicated by (v, vt, vn), normals normalised
    for (const Shape& sh : shapes) {
        std::map<int, std::vector<const Face*>> sub;
        for (const Face& f : sh.faces) sub[f.material].push_back(&f);
        for (const auto& [mat, faces] : sub) {
            std::unordered_map<Corner, uint32_t, CornerHash> index;
            std::vector<float> pos, uv, nrm;
            std::vector<uint32_t> idx;
            for (const Face* f : faces)
                for (const Corner& c : f->c) {
                    auto [it, fresh] = index.try_emplace(c, (uint32_t)(pos.size() / 3));
                    if (fresh) {
                        pos.insert(pos.end(), {vp[3 * c.v], vp[3 * c.v + 1], vp[3 * c.v + 2]});
                        uv.insert(uv.end(), {vt[2 * c.t], vt[2 * c.t + 1]});
                        // glm::normalize in float (ObjLoader.cpp:123): v * (1 / sqrt(v.v));
                        // a zero normal is kept as zero
                        const float x = vn[3 * c.n], y = vn[3 * c.n + 1], z = vn[3 * c.n + 2];
                        const float d = x * x + y * y + z * z;
                        const float r = d > 0.0f ? 1.0f / std::sqrt(d) : 0.0f;
                        nrm.insert(nrm.end(), {x * r, y * r, z * r});
                    }
                    idx.push_back(it->second);
                }
            // per-submesh tangents (ObjLoader.cpp:150-152)
            std::vector<float> tan4(pos.size() / 3 * 4);
            mikk_tangents(pos.data(), nrm.data(), uv.data(), (uint32_t)(pos.size() / 3), idx.data(), idx.size(),
                          tan4.data());
            const uint32_t mid = mat >= 0 ? (uint32_t)mat : default_id;
            need_default = need_default || mat < 0;
            for (size_t k = 0; k < idx.size(); k++) {
                const uint32_t i = idx[k];
                out.positions.insert(out.positions.end(), {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
                out.normals.insert(out.normals.end(), {nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]});
                out.tangents.insert(out.tangents.end(), {tan4[4 * i], tan4[4 * i + 1], tan4[4 * i + 2], tan4[4 * i + 3]});
            }
            for (size_t f = 0; f < idx.size() / 3; f++) out.material_ids.push_back(mid);
        }
    }
    if (need_default) out.materials.push_back(default_material());

    pt_scene_desc& d = out.desc;
    d = pt_scene_desc{};
    d.triangle_count = (uint32_t)out.material_ids.size();
    d.positions = out.positions.data();
    d.normals = out.normals.data();
    d.tangents = out.tangents.data();
    d.material_ids = out.material_ids.data();
    d.material_count = (uint32_t)out.materials.size();
    d.materials = out.materials.data();
    d.light_weights = nullptr;
}

}  // namespace
}  // namespace pt

extern "C" {

pt_status pt_import_obj(const char* path, pt_imported_scene** out, char* error, size_t error_capacity) {
    auto report = [&](const std::string& msg) {
        if (error && error_capacity) {
            std::strncpy(error, msg.c_str(), error_capacity - 1);
            error[error_capacity - 1] = '\0';
        }
    };
    if (!path || !out) {
        report("path and out must not be NULL");
        return PT_ERROR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    auto scene = new (std::nothrow) pt_imported_scene();
    if (!scene) {
        report("out of host memory");
        return PT_ERROR_OUT_OF_MEMORY;
    }
    try {
        pt::import_obj(std::filesystem::path(path), *scene);
    } catch (const pt::ImportError& e) {
        delete scene;
        report(e.message);
        return PT_ERROR_IO;
    } catch (const std::exception& e) {
        delete scene;
        report(std::string("malformed OBJ: ") + e.what());
        return PT_ERROR_IO;
    }
    *out = scene;
    return PT_OK;
}

const pt_scene_desc* pt_imported_scene_desc(const pt_imported_scene* scene) { return scene ? &scene->desc : nullptr; }

void pt_imported_scene_free(pt_imported_scene* scene) { delete scene; }

}  // extern "C"

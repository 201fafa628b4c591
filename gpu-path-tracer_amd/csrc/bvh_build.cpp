// Binned-SAH BVH2 build and BVH2 -> BVH4 collapse (host side of the core).
#include "bvh_build.hpp"

#include <algorithm>
#include <cmath>
#include <future>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>

namespace pt {
namespace {

constexpr float kInf = std::numeric_limits<float>::infinity();

struct Box {
    float lo[3] = {kInf, kInf, kInf};
    float hi[3] = {-kInf, -kInf, -kInf};
    void grow(const float* p) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[a]);
            hi[a] = std::max(hi[a], p[a]);
        }
    }
    void grow(const Box& b) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], b.lo[a]);
            hi[a] = std::max(hi[a], b.hi[a]);
        }
    }
    bool empty() const { return lo[0] > hi[0]; }
    float area() const {
        if (empty()) return 0.0f;
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct TNode {
    Box box;              // unpadded bounds of the triangles below
    uint32_t begin = 0;
    uint32_t count = 0;
    std::unique_ptr<TNode> left, right;
    bool leaf() const { return !left; }
};

struct Builder {
    const float* pos;
    std::vector<float> centroid;     // [T][3]
    std::vector<Box> tri_box;        // [T]
    std::vector<uint32_t> idx;
    BvhBuildOptions opt;
    uint32_t parallel_depth = 0;

    std::unique_ptr<TNode> build(uint32_t begin, uint32_t end, uint32_t depth) {
        auto node = std::make_unique<TNode>();
        node->begin = begin;
        node->count = end - begin;
        Box cbox;
        for (uint32_t i = begin; i < end; i++) {
            const uint32_t t = idx[i];
            node->box.grow(tri_box[t]);
            cbox.grow(&centroid[3 * (size_t)t]);
        }
        const uint32_t n = end - begin;
        if (n <= 1) return node;

        const uint32_t B = opt.bins;
        float best_cost = kInf;
        int best_axis = -1;
        uint32_t best_bin = 0;
        std::vector<uint32_t> cnt(B);
        std::vector<Box> bb(B);
        std::vector<float> right_area(B);
        std::vector<uint32_t> right_cnt(B);
        for (int a = 0; a < 3; a++) {
            const float ext = cbox.hi[a] - cbox.lo[a];
            if (!(ext > 0.0f)) continue;
            const float k = (float)B / ext;
            std::fill(cnt.begin(), cnt.end(), 0u);
            std::fill(bb.begin(), bb.end(), Box());
            for (uint32_t i = begin; i < end; i++) {
                const uint32_t t = idx[i];
                uint32_t b = bin_of(centroid[3 * (size_t)t + a], cbox.lo[a], k, B);
                cnt[b]++;
                bb[b].grow(tri_box[t]);
            }
            Box acc;
            uint32_t c = 0;
            for (int b = (int)B - 1; b >= 1; b--) {
                acc.grow(bb[b]);
                c += cnt[b];
                right_area[b] = acc.area();
                right_cnt[b] = c;
            }
            acc = Box();
            c = 0;
            for (uint32_t b = 1; b < B; b++) {   // split between bin b-1 and b
                acc.grow(bb[b - 1]);
                c += cnt[b - 1];
                if (c == 0 || right_cnt[b] == 0) continue;
                float cost = acc.area() * (float)c + right_area[b] * (float)right_cnt[b];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_bin = b;
                }
            }
        }
        const float area = node->box.area();
        const float leaf_cost = opt.intersection_cost * (float)n;
        float split_cost = kInf;
        if (best_axis >= 0 && area > 0.0f)
            split_cost = opt.traversal_cost + opt.intersection_cost * best_cost / area;
        if (n <= opt.max_leaf_size && leaf_cost <= split_cost) return node;

        uint32_t mid;
        if (best_axis < 0) {
            mid = begin + n / 2;   // all centroids coincide: split the list in half
        } else {
            const float ext = cbox.hi[best_axis] - cbox.lo[best_axis];
            const float k = (float)B / ext;
            const float lo = cbox.lo[best_axis];
            auto it = std::partition(idx.begin() + begin, idx.begin() + end, [&](uint32_t t) {
                return bin_of(centroid[3 * (size_t)t + best_axis], lo, k, B) < best_bin;
            });
            mid = (uint32_t)(it - idx.begin());
            if (mid == begin || mid == end) mid = begin + n / 2;
        }
        if (depth < parallel_depth && n > 65536) {
            auto fut = std::async(std::launch::async, [&, begin, mid, depth]() { return build(begin, mid, depth + 1); });
            node->right = build(mid, end, depth + 1);
            node->left = fut.get();
        } else {
            node->left = build(begin, mid, depth + 1);
            node->right = build(mid, end, depth + 1);
        }
        return node;
    }

    static uint32_t bin_of(float c, float lo, float k, uint32_t B) {
        float f = (c - lo) * k;
        int b = (int)f;
        if (b < 0) b = 0;
        if (b >= (int)B) b = (int)B - 1;
        return (uint32_t)b;
    }
};

void set_child(Bvh4Node& node, int slot, const Box& unpadded, uint32_t ref, uint32_t info) {
    float lo[3] = {unpadded.lo[0], unpadded.lo[1], unpadded.lo[2]};
    float hi[3] = {unpadded.hi[0], unpadded.hi[1], unpadded.hi[2]};
    pad_bounds(lo, hi);
    node.lox[slot] = lo[0];
    node.loy[slot] = lo[1];
    node.loz[slot] = lo[2];
    node.hix[slot] = hi[0];
    node.hiy[slot] = hi[1];
    node.hiz[slot] = hi[2];
    node.child[slot] = ref;
    node.info[slot] = info;
}

void clear_node(Bvh4Node& node) {
    for (int s = 0; s < 4; s++) {
        // an empty slot's box is [+inf, +inf]: every slab interval is empty
        node.lox[s] = node.loy[s] = node.loz[s] = kInf;
        node.hix[s] = node.hiy[s] = node.hiz[s] = kInf;
        node.child[s] = kInvalid;
        node.info[s] = 0;
    }
}

uint32_t leaf_ref_checked(const TNode* n, Bvh4Build& out) {
    if (n->count > kMaxLeafSize) throw std::runtime_error("BVH leaf larger than 16 triangles");
    if (n->begin > kMaxLeafFirst) throw std::runtime_error("scene exceeds 2^27 triangles");
    out.leaf_count++;
    out.max_leaf = std::max(out.max_leaf, n->count);
    return make_leaf_ref(n->begin, n->count);
}

// Emits the node that replaces BVH2 node `n` (internal) in DFS preorder: up to
// out.width children, by expanding the largest-area internal child.
uint32_t emit(const TNode* n, Bvh4Build& out, uint32_t level) {
    const uint32_t index = (uint32_t)out.nodes.size();
    out.nodes.emplace_back();
    clear_node(out.nodes[index]);
    out.depth = std::max(out.depth, level);

    const TNode* ch[4] = {n->left.get(), n->right.get(), nullptr, nullptr};
    int count = 2;
    while (count < (int)out.width) {
        int pick = -1;
        float best = -1.0f;
        for (int i = 0; i < count; i++) {
            if (!ch[i]->leaf() && ch[i]->box.area() > best) {
                best = ch[i]->box.area();
                pick = i;
            }
        }
        if (pick < 0) break;
        const TNode* c = ch[pick];
        ch[pick] = c->left.get();
        ch[count++] = c->right.get();
    }
    for (int i = 0; i < count; i++) {
        const TNode* c = ch[i];
        uint32_t ref, info;
        if (c->leaf()) {
            ref = leaf_ref_checked(c, out);
            info = c->count;
        } else {
            ref = emit(c, out, level + 1);
            info = 0;
        }
        set_child(out.nodes[index], i, c->box, ref, info);
    }
    return index;
}

// Slot of each of the `count` children of an 8-wide node (boxes lo/hi
// [axis][child]), in the manner of the compressed wide BVH of Ylitie, Karras
// and Laine (HPG 2017): slot s belongs to the ray octant s (bit a set = the
// direction's component a is negative), whose first child should be the one
// farthest upstream along the octant's diagonal d_s.  The cost of child i in
// slot s is (centre_i - mean centre) . d_s, and the slots are the assignment
// of least total cost, found by the Hungarian method (Kuhn 1955, in the
// potentials form of Jonker and Volgenant) over the 8 x 8 cost matrix whose
// rows past `count` are zero (empty slots).
void octant_slots(const float lo[3][8], const float hi[3][8], int count, int slot_of[8]) {
    float cen[8][3];
    float mid[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < count; i++)
        for (int a = 0; a < 3; a++) {
            cen[i][a] = 0.5f * (lo[a][i] + hi[a][i]);
            mid[a] += cen[i][a];
        }
    for (int a = 0; a < 3; a++) mid[a] /= (float)std::max(count, 1);
    constexpr int n = 8;
    double cost[n + 1][n + 1] = {};   // 1-based: cost[child][slot]
    for (int i = 0; i < count; i++)
        for (int s = 0; s < n; s++) {
            double c = 0.0;
            for (int a = 0; a < 3; a++) c += (s >> a & 1) ? mid[a] - cen[i][a] : cen[i][a] - mid[a];
            cost[i + 1][s + 1] = c;
        }
    double u[n + 1] = {}, v[n + 1] = {};
    int row_of[n + 1] = {}, way[n + 1] = {};   // row_of[slot]: child assigned to it
    for (int i = 1; i <= n; i++) {
        row_of[0] = i;
        int j0 = 0;
        double minv[n + 1];
        bool used[n + 1] = {};
        for (int j = 0; j <= n; j++) minv[j] = std::numeric_limits<double>::infinity();
        do {
            used[j0] = true;
            const int i0 = row_of[j0];
            double delta = std::numeric_limits<double>::infinity();
            int j1 = 0;
            for (int j = 1; j <= n; j++) {
                if (used[j]) continue;
                const double cur = cost[i0][j] - u[i0] - v[j];
                if (cur < minv[j]) {
                    minv[j] = cur;
                    way[j] = j0;
                }
                if (minv[j] < delta) {
                    delta = minv[j];
                    j1 = j;
                }
            }
            for (int j = 0; j <= n; j++) {
                if (used[j]) {
                    u[row_of[j]] += delta;
                    v[j] -= delta;
                } else {
                    minv[j] -= delta;
                }
            }
            j0 = j1;
        } while (row_of[j0] != 0);
        do {
            const int j1 = way[j0];
            row_of[j0] = row_of[j1];
            j0 = j1;
        } while (j0 != 0);
    }
    for (int i = 0; i < n; i++) slot_of[i] = -1;
    for (int j = 1; j <= n; j++)
        if (row_of[j] >= 1 && row_of[j] <= count) slot_of[row_of[j] - 1] = j - 1;
}

// The 8-wide quantized any-hit tree (Bvh8QNode) of the binary tree under
// `root`, numbered breadth first.  Each node takes up to eight children of its
// binary node by expanding the largest-area internal child (the rule of emit),
// child boxes padded as in set_child, puts them into octant slots
// (octant_slots), appends its inner children to the queue (consecutive node
// indices in slot order) and the triangles of its leaf children to order8
// (consecutive positions in slot order).  A leaf root is the only child of the
// root; an empty tree is a root without children.
void emit8_tree(const TNode* root, const std::vector<uint32_t>& idx, Bvh4Build& out) {
    struct Item {
        const TNode* n;
        uint32_t level;
    };
    std::vector<Item> queue{{root, 1u}};
    out.qnodes8.assign(1, Bvh8QNode{});
    out.order8.clear();
    out.depth8 = 0;
    for (size_t head = 0; head < queue.size(); head++) {
        const TNode* n = queue[head].n;
        const uint32_t level = queue[head].level;
        out.depth8 = std::max(out.depth8, level);
        const TNode* ch[8] = {};
        int count = 0;
        if (n->leaf()) {
            if (n->count > 0) ch[count++] = n;
        } else {
            ch[0] = n->left.get();
            ch[1] = n->right.get();
            count = 2;
            while (count < 8) {
                int pick = -1;
                float best = -1.0f;
                for (int i = 0; i < count; i++) {
                    if (!ch[i]->leaf() && ch[i]->box.area() > best) {
                        best = ch[i]->box.area();
                        pick = i;
                    }
                }
                if (pick < 0) break;
                const TNode* c = ch[pick];
                ch[pick] = c->left.get();
                ch[count++] = c->right.get();
            }
        }
        float lo[3][8] = {}, hi[3][8] = {};
        for (int i = 0; i < count; i++) {
            float l[3] = {ch[i]->box.lo[0], ch[i]->box.lo[1], ch[i]->box.lo[2]};
            float h[3] = {ch[i]->box.hi[0], ch[i]->box.hi[1], ch[i]->box.hi[2]};
            pad_bounds(l, h);
            for (int a = 0; a < 3; a++) {
                lo[a][i] = l[a];
                hi[a][i] = h[a];
            }
        }
        int slot_of[8];
        octant_slots(lo, hi, count, slot_of);
        const TNode* at[8] = {};
        uint32_t occupied[8];
        float slo[3][8] = {}, shi[3][8] = {};
        for (int s = 0; s < 8; s++) occupied[s] = kInvalid;
        for (int i = 0; i < count; i++) {
            const int s = slot_of[i];
            at[s] = ch[i];
            occupied[s] = 0u;
            for (int a = 0; a < 3; a++) {
                slo[a][s] = lo[a][i];
                shi[a][s] = hi[a][i];
            }
        }
        Bvh8QNode q{};
        quantize_children(slo, shi, occupied, 8, q.origin, q.scale, q.lo, q.hi);
        uint32_t imask = 0u;
        q.node_base = (uint32_t)out.qnodes8.size();
        for (int s = 0; s < 8; s++) {
            if (!at[s] || at[s]->leaf()) continue;
            imask |= 1u << s;
            queue.push_back({at[s], level + 1u});
            out.qnodes8.emplace_back();
        }
        if (out.qnodes8.size() > (size_t)kMaxNodeIndex) throw std::runtime_error("the 8-wide BVH has more than 2^30 nodes");
        q.scale |= imask << 24;
        q.tri_base = (uint32_t)out.order8.size();
        uint32_t end = 0u;
        for (int s = 0; s < 8; s++) {
            if (at[s] && at[s]->leaf()) {
                if (at[s]->count > kMaxLeafSize) throw std::runtime_error("BVH leaf larger than 16 triangles");
                for (uint32_t k = 0; k < at[s]->count; k++) out.order8.push_back(idx[at[s]->begin + k]);
                end += at[s]->count;
            }
            q.tri_end[s >> 2] |= end << (8 * (s & 3));
        }
        out.qnodes8[head] = q;
    }
}

}  // namespace

std::vector<Bvh2Node> to_bvh2(const Bvh4Build& b) {
    std::vector<Bvh2Node> out(b.nodes.size());
    for (size_t i = 0; i < b.nodes.size(); i++) {
        const Bvh4Node& s = b.nodes[i];
        Bvh2Node& d = out[i];
        for (int k = 0; k < 2; k++) {
            d.lox[k] = s.lox[k];
            d.hix[k] = s.hix[k];
            d.loy[k] = s.loy[k];
            d.hiy[k] = s.hiy[k];
            d.loz[k] = s.loz[k];
            d.hiz[k] = s.hiz[k];
            d.child[k] = s.child[k];
            d.info[k] = s.info[k];
        }
    }
    return out;
}

void pad_bounds(float lo[3], float hi[3]) {
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    const float largest = fmaxf(ex, fmaxf(ey, ez));
    const float padding = fmaxf(1e-5f * largest, 1e-5f);
    for (int a = 0; a < 3; a++) {
        lo[a] = lo[a] - padding;
        hi[a] = hi[a] + padding;
    }
}

Bvh4Build build_bvh4(const float* positions, uint32_t T, const BvhBuildOptions& options) {
    if (options.max_leaf_size < 1 || options.max_leaf_size > kMaxLeafSize)
        throw std::runtime_error("max_leaf_size must be in [1, 16]");
    if (options.bins < 2 || options.bins > 256) throw std::runtime_error("bins must be in [2, 256]");
    if (options.width != 2 && options.width != 4) throw std::runtime_error("BVH node width must be 2 or 4");
    if (options.wide8 && options.width != 4) throw std::runtime_error("8-wide nodes need a 4-wide build");
    if ((uint64_t)T > (uint64_t)kMaxLeafFirst) throw std::runtime_error("scene exceeds 2^27 triangles");

    Builder b;
    b.pos = positions;
    b.opt = options;
    unsigned hw = options.threads ? options.threads : std::max(1u, std::thread::hardware_concurrency());
    b.parallel_depth = 0;
    while ((1u << b.parallel_depth) < hw && b.parallel_depth < 6) b.parallel_depth++;
    b.centroid.resize(3 * (size_t)T);
    b.tri_box.resize(T);
    b.idx.resize(T);
    for (uint32_t t = 0; t < T; t++) {
        const float* p = positions + 9 * (size_t)t;
        for (int a = 0; a < 3; a++) {
            if (!std::isfinite(p[a]) || !std::isfinite(p[3 + a]) || !std::isfinite(p[6 + a]))
                throw std::runtime_error("triangle " + std::to_string(t) + " has a non-finite vertex");
            b.centroid[3 * (size_t)t + a] = (p[a] + p[3 + a] + p[6 + a]) / 3.0f;
        }
        b.tri_box[t].grow(p);
        b.tri_box[t].grow(p + 3);
        b.tri_box[t].grow(p + 6);
        b.idx[t] = t;
    }

    Bvh4Build out;
    out.width = options.width;
    if (T == 0) {
        out.nodes.emplace_back();
        clear_node(out.nodes[0]);
        out.depth = 1;
        if (options.wide8) {
            const TNode none;
            emit8_tree(&none, b.idx, out);
        }
        return out;
    }
    std::unique_ptr<TNode> root = b.build(0, T, 0);
    out.nodes.reserve(T / 2 + 8);
    if (root->leaf()) {
        out.nodes.emplace_back();
        clear_node(out.nodes[0]);
        out.depth = 1;
        set_child(out.nodes[0], 0, root->box, leaf_ref_checked(root.get(), out), root->count);
    } else {
        emit(root.get(), out, 1);
    }
    if (options.wide8) {
        out.qnodes8.reserve(T / 8 + 8);
        out.order8.reserve(T);
        emit8_tree(root.get(), b.idx, out);
    }
    out.order = std::move(b.idx);
    return out;
}

}  // namespace pt

// Host binned-SAH BVH builder producing the device node layout (16 words per
// node: both child boxes in the parent, see gy_internal.h).  It serves as the
// fallback builder (gy_bvh_params.builder = 1) and as the reference the GPU
// builder's SAH cost is compared with.  The split rule is the reference's SAH
// with bucketed centroids (accelerator/BVHAccel.cpp:296-349): cost
// 0.125 + (n0 A0 + n1 A1) / A, split on the centroid bounds' longest axis;
// leaves hold up to max_leaf triangles instead of the reference's one.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "gy_internal.h"

namespace gy {
namespace {

struct Box {
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], b.mn[k]);
            mx[k] = std::max(mx[k], b.mx[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], p[k]);
            mx[k] = std::max(mx[k], p[k]);
        }
    }
    double area() const {
        if (mn[0] > mx[0]) return 0.0;
        double dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        return 2.0 * (dx * dy + dx * dz + dy * dz);
    }
};

struct Prim {
    Box b;
    float c[3];
    uint32_t id;
};

struct Task {
    uint32_t begin, end;
    uint32_t node;  // node that owns this child slot
    int slot;       // 0 or 1
    int depth;
};

constexpr int kBins = 16;

}  // namespace

void BuildBVHHost(const float* pos, int32_t n_tri, int max_leaf, BVHBuild* out) {
    if (max_leaf < 1) max_leaf = 1;
    if (max_leaf > 15) max_leaf = 15;
    std::vector<Prim> prims(n_tri);
    for (int32_t i = 0; i < n_tri; ++i) {
        Prim& p = prims[i];
        for (int v = 0; v < 3; ++v) p.b.grow(pos + 9 * (size_t)i + 3 * v);
        for (int k = 0; k < 3; ++k) p.c[k] = 0.5f * p.b.mn[k] + 0.5f * p.b.mx[k];
        p.id = (uint32_t)i;
    }
    out->nodes.clear();
    out->tri_order.clear();
    out->tri_order.reserve(n_tri);
    out->max_depth = 0;
    out->n_leaves = 0;
    out->max_leaf = 0;
    auto new_node = [&]() {
        uint32_t idx = (uint32_t)(out->nodes.size() / 16);
        out->nodes.resize(out->nodes.size() + 16, 0u);
        uint32_t* n = &out->nodes[(size_t)idx * 16];
        const float inf = INFINITY, ninf = -INFINITY;
        for (int s = 0; s < 2; ++s) {
            for (int k = 0; k < 3; ++k) {
                std::memcpy(&n[s * 6 + k], &inf, 4);
                std::memcpy(&n[s * 6 + 3 + k], &ninf, 4);
            }
        }
        n[12] = n[13] = kEmptyRef;
        n[14] = n[15] = 0;
        return idx;
    };
    auto set_box = [&](uint32_t node, int slot, const Box& b) {
        uint32_t* n = &out->nodes[(size_t)node * 16];
        for (int k = 0; k < 3; ++k) {
            std::memcpy(&n[slot * 6 + k], &b.mn[k], 4);
            std::memcpy(&n[slot * 6 + 3 + k], &b.mx[k], 4);
        }
    };
    if (n_tri == 0) {
        new_node();
        return;
    }
    uint32_t root = new_node();
    // The root holds the whole scene in its slot 0; slot 1 stays empty unless
    // the root splits, in which case the two halves replace it.
    std::vector<Task> stack;
    stack.push_back({0, (uint32_t)n_tri, root, -1, 0});
    while (!stack.empty()) {
        Task t = stack.back();
        stack.pop_back();
        Box bounds, cb;
        for (uint32_t i = t.begin; i < t.end; ++i) {
            bounds.grow(prims[i].b);
            cb.grow(prims[i].c);
        }
        uint32_t n = t.end - t.begin;
        // Choose a split.
        int axis = 0;
        float ext[3] = {cb.mx[0] - cb.mn[0], cb.mx[1] - cb.mn[1], cb.mx[2] - cb.mn[2]};
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        uint32_t mid = t.begin;
        bool leaf = (int)n <= 1;
        if (!leaf) {
            if (ext[axis] <= 0.f) {
                // All centroids coincide: split by count if too many, else leaf.
                if ((int)n <= max_leaf) leaf = true;
                else mid = t.begin + n / 2;
            } else {
                int cnt[kBins] = {0};
                Box bb[kBins];
                const float scale = kBins / ext[axis];
                auto bin_of = [&](const Prim& p) {
                    int b = (int)((p.c[axis] - cb.mn[axis]) * scale);
                    return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
                };
                for (uint32_t i = t.begin; i < t.end; ++i) {
                    int b = bin_of(prims[i]);
                    cnt[b]++;
                    bb[b].grow(prims[i].b);
                }
                double best = INFINITY;
                int best_split = -1;
                Box left[kBins];
                int lc[kBins];
                Box acc;
                int ac = 0;
                for (int i = 0; i < kBins; ++i) {
                    acc.grow(bb[i]);
                    ac += cnt[i];
                    left[i] = acc;
                    lc[i] = ac;
                }
                Box racc;
                int rc = 0;
                const double area = bounds.area();
                for (int i = kBins - 1; i >= 1; --i) {
                    racc.grow(bb[i]);
                    rc += cnt[i];
                    if (lc[i - 1] == 0 || rc == 0) continue;
                    double cost = 0.125 + (lc[i - 1] * left[i - 1].area() + rc * racc.area()) / area;
                    if (cost < best) {
                        best = cost;
                        best_split = i - 1;
                    }
                }
                if ((int)n <= max_leaf && (best_split < 0 || best >= (double)n)) {
                    leaf = true;
                } else if (best_split < 0) {
                    mid = t.begin + n / 2;
                    std::nth_element(prims.begin() + t.begin, prims.begin() + mid, prims.begin() + t.end,
                                     [axis](const Prim& a, const Prim& b) { return a.c[axis] < b.c[axis]; });
                } else {
                    auto it = std::partition(prims.begin() + t.begin, prims.begin() + t.end,
                                             [&](const Prim& p) { return bin_of(p) <= best_split; });
                    mid = (uint32_t)(it - prims.begin());
                    if (mid == t.begin || mid == t.end) mid = t.begin + n / 2;
                }
            }
        }
        if (leaf) {
            uint32_t first = (uint32_t)out->tri_order.size();
            for (uint32_t i = t.begin; i < t.end; ++i) out->tri_order.push_back(prims[i].id);
            uint32_t owner = t.node;
            int slot = t.slot < 0 ? 0 : t.slot;
            set_box(owner, slot, bounds);
            uint32_t* nd = &out->nodes[(size_t)owner * 16];
            nd[12 + slot] = first;
            nd[14 + slot] = n;
            out->n_leaves++;
            out->max_leaf = std::max(out->max_leaf, (int32_t)n);
            out->max_depth = std::max(out->max_depth, t.depth);
            continue;
        }
        // Interior: a node whose two slots hold the halves.
        uint32_t node;
        if (t.slot < 0) {
            node = t.node;  // the root itself
        } else {
            node = new_node();
            set_box(t.node, t.slot, bounds);
            uint32_t* pd = &out->nodes[(size_t)t.node * 16];
            pd[12 + t.slot] = node;
            pd[14 + t.slot] = 0;
        }
        out->max_depth = std::max(out->max_depth, t.depth + 1);
        stack.push_back({mid, t.end, node, 1, t.depth + 1});
        stack.push_back({t.begin, mid, node, 0, t.depth + 1});
    }
    out->sah_cost = BVHSahCost(*out);
}

// SAH cost of a two-child BVH: sum over interior nodes of A(node)/A(root) * 1
// plus sum over leaves of A(leaf)/A(root) * n_tri, child areas from the parent.
double BVHSahCost(const BVHBuild& b) {
    if (b.nodes.empty()) return 0.0;
    auto box_area = [&](const uint32_t* n, int slot) {
        float v[6];
        std::memcpy(v, &n[slot * 6], 24);
        if (!(v[0] <= v[3])) return 0.0;
        double dx = v[3] - v[0], dy = v[4] - v[1], dz = v[5] - v[2];
        return 2.0 * (dx * dy + dx * dz + dy * dz);
    };
    const uint32_t* r = &b.nodes[0];
    double root = 0.0;
    {
        // root area: union of the two root slots
        float v0[6], v1[6];
        std::memcpy(v0, &r[0], 24);
        std::memcpy(v1, &r[6], 24);
        float mn[3], mx[3];
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(v0[k], v1[k]);
            mx[k] = std::max(v0[3 + k], v1[3 + k]);
        }
        double dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        root = 2.0 * (dx * dy + dx * dz + dy * dz);
    }
    if (root <= 0.0) return 0.0;
    double cost = 1.0;  // root traversal
    size_t nn = b.nodes.size() / 16;
    for (size_t i = 0; i < nn; ++i) {
        const uint32_t* n = &b.nodes[i * 16];
        for (int s = 0; s < 2; ++s) {
            if (n[12 + s] == kEmptyRef) continue;
            double a = box_area(n, s) / root;
            if (n[14 + s] > 0) cost += a * n[14 + s];
            else cost += a * 1.0;
        }
    }
    return cost;
}

}  // namespace gy

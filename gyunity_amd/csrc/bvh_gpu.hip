// GPU BVH build: top-down binned SAH, one level of the tree per round.
//
// The split rule is the reference's SAH over bucketed centroids
// (accelerator/BVHAccel.cpp:296-349: cost 0.125 + (n0 A0 + n1 A1) / A), here with
// 16 bins on each of the three axes; leaves hold up to max_leaf triangles.
// Output: the two-child node layout the traversal kernels read (16 words per
// node, both child boxes in the parent; see gy_internal.h) and the leaf order of
// the triangles.
//
// Per round, over the tasks (node ranges) of the current level:
//   k_bvh_bin       every primitive of a large task adds itself to its task's
//                   bins along the longest centroid axis (count, primitive
//                   bounds, centroid bounds) with atomics on order-preserving
//                   float encodings
//   k_bvh_split     one thread per task evaluates the 15 candidate planes,
//                   writes the two child slots of the task's node and emits
//                   child tasks (large -> next round, small -> finished later)
//   k_bvh_partition every position is scattered to the other index buffer:
//                   task members to their child's side, finished ranges as is
// Ranges of 2..kSmall triangles are finished by one thread each (k_bvh_small),
// sequentially and in place, after the last round, with the host builder's leaf
// rule (a leaf of at most max_leaf triangles only where no split is cheaper).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "runtime.h"

namespace gy {
namespace {

constexpr int kBins = 16;
constexpr int kSmall = 32;
constexpr int kBinWords = 13;  // count, prim min (3), prim max (3), centroid min (3), centroid max (3)
constexpr int kTaskBinWords = kBins * kBinWords;

struct BTask {
    uint32_t begin, end;
    uint32_t node;   // large task: its node; small task: the parent node
    uint32_t depth;  // large task: depth of its node; small task: depth of the slot
    uint32_t slot;   // small task: slot of the parent it fills
    float bmin[3], bmax[3];  // primitive bounds
    float cmin[3], cmax[3];  // centroid bounds
};

struct BuildDev {
    const float4* pmin;   // primitive bounds by primitive id
    const float4* pmax;
    uint32_t* idx[2];     // primitive id by position (ping-pong)
    int32_t* tid[2];      // task of each position this round (-1 = finished range)
    BTask* tasks[2];      // large tasks: this round / next round
    BTask* small;         // small tasks, finished at the end
    uint32_t* bins;       // per large task: kTaskBinWords
    uint32_t* cursor;     // per large task: left, right insertion cursors
    int32_t* child;       // per large task: next-round id of the left / right child (-1 none)
    uint32_t* split;      // per large task: axis, split bin (or ~0 for a median split), left count
    uint32_t* nodes;      // output, 16 words per node
    uint32_t* counters;   // 0 nodes, 1 next-round tasks, 2 small tasks, 3 max depth, 4 leaves, 5 overflow
    uint32_t max_nodes, max_large, max_small;
    int32_t max_leaf;
};

__device__ inline uint32_t enc(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float dec(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

__device__ inline float box_area(const float* mn, const float* mx) {
    if (!(mn[0] <= mx[0])) return 0.f;
    const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    return 2.f * (dx * dy + dx * dz + dy * dz);
}

__device__ inline void write_slot(const BuildDev& B, uint32_t node, int slot, const float* mn, const float* mx, uint32_t ref,
                                  uint32_t cnt) {
    uint32_t* n = B.nodes + (size_t)node * 16;
    for (int k = 0; k < 3; ++k) {
        n[slot * 6 + k] = __float_as_uint(mn[k]);
        n[slot * 6 + 3 + k] = __float_as_uint(mx[k]);
    }
    n[12 + slot] = ref;
    n[14 + slot] = cnt;
}

// Split axis of a range: the longest extent of its centroid bounds (the host
// builder's choice; choosing among all three axes by SAH with the 0.125
// traversal cost gives trees that trace slower on these meshes).
__device__ inline int long_axis(const float* cmn, const float* cmx) {
    int ax = 0;
    if (cmx[1] - cmn[1] > cmx[ax] - cmn[ax]) ax = 1;
    if (cmx[2] - cmn[2] > cmx[ax] - cmn[ax]) ax = 2;
    return ax;
}

__device__ inline int bin_of(float c, float lo, float hi) {
    const float ext = hi - lo;
    if (!(ext > 0.f)) return 0;
    int b = (int)((c - lo) * (kBins / ext));
    return b < 0 ? 0 : (b >= kBins ? kBins - 1 : b);
}

__global__ void k_bvh_prim(const float* __restrict__ pos, int32_t n, float4* pmin, float4* pmax, uint32_t* idx, int32_t* tid,
                           uint32_t* acc) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* p = pos + 9 * (size_t)i;
    float mn[3], mx[3], c[3];
    for (int k = 0; k < 3; ++k) {
        mn[k] = fminf(fminf(p[k], p[3 + k]), p[6 + k]);
        mx[k] = fmaxf(fmaxf(p[k], p[3 + k]), p[6 + k]);
        c[k] = 0.5f * mn[k] + 0.5f * mx[k];
    }
    pmin[i] = make_float4(mn[0], mn[1], mn[2], 0.f);
    pmax[i] = make_float4(mx[0], mx[1], mx[2], 0.f);
    idx[i] = (uint32_t)i;
    tid[i] = 0;
    for (int k = 0; k < 3; ++k) {
        atomicMin(&acc[k], enc(mn[k]));
        atomicMax(&acc[3 + k], enc(mx[k]));
        atomicMin(&acc[6 + k], enc(c[k]));
        atomicMax(&acc[9 + k], enc(c[k]));
    }
}

__global__ void k_bvh_root(BuildDev B, int32_t n, const uint32_t* acc) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    BTask t;
    t.begin = 0;
    t.end = (uint32_t)n;
    t.node = 0;
    t.depth = 0;
    for (int k = 0; k < 3; ++k) {
        t.bmin[k] = dec(acc[k]);
        t.bmax[k] = dec(acc[3 + k]);
        t.cmin[k] = dec(acc[6 + k]);
        t.cmax[k] = dec(acc[9 + k]);
    }
    B.tasks[0][0] = t;
    B.counters[0] = 1;  // node 0 is the root
    B.counters[1] = 0;
}

__global__ void k_bvh_bin_clear(BuildDev B, uint32_t ntask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntask * kTaskBinWords) return;
    const uint32_t w = (i % kTaskBinWords) % kBinWords;
    const bool is_min = (w >= 1 && w <= 3) || (w >= 7 && w <= 9);
    B.bins[i] = is_min ? 0xffffffffu : 0u;
    if (i < ntask * 2) B.cursor[i] = 0u;
}

// Bins of each primitive's task on the task's longest centroid axis.  Tasks
// are contiguous index ranges, so at the upper levels a whole workgroup
// belongs to one task (the task of its first index): those primitives are
// binned in LDS and the workgroup adds its bins to the task's with one global
// atomic per word.  Primitives of any other task go to global memory directly.
// Same-address global atomics serialise at the L2, and the upper levels put
// every primitive of the scene on one task's 16 bins.
__global__ void __launch_bounds__(256) k_bvh_bin(BuildDev B, int cur, int32_t n) {
    __shared__ uint32_t lb[kTaskBinWords];
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t first = blockIdx.x * blockDim.x;
    const int32_t t0 = B.tid[cur][first];  // first < n: the grid covers n
    for (int w = threadIdx.x; w < kTaskBinWords; w += blockDim.x) {
        const int k = w % kBinWords;
        lb[w] = k == 0 ? 0u : ((k >= 1 && k <= 3) || (k >= 7 && k <= 9) ? 0xffffffffu : 0u);
    }
    __syncthreads();
    const int32_t t = i < n ? B.tid[cur][i] : -1;
    if (t >= 0) {
        const BTask T = B.tasks[cur][t];
        const uint32_t p = B.idx[cur][i];
        const float4 a = B.pmin[p], b = B.pmax[p];
        const float mn[3] = {a.x, a.y, a.z}, mx[3] = {b.x, b.y, b.z};
        const float c[3] = {0.5f * a.x + 0.5f * b.x, 0.5f * a.y + 0.5f * b.y, 0.5f * a.z + 0.5f * b.z};
        const int ax = long_axis(T.cmin, T.cmax);
        const int bin = bin_of(c[ax], T.cmin[ax], T.cmax[ax]);
        uint32_t* w = (t == t0 ? lb : B.bins + (size_t)t * kTaskBinWords) + bin * kBinWords;
        atomicAdd(&w[0], 1u);
        for (int k = 0; k < 3; ++k) {
            atomicMin(&w[1 + k], enc(mn[k]));
            atomicMax(&w[4 + k], enc(mx[k]));
            atomicMin(&w[7 + k], enc(c[k]));
            atomicMax(&w[10 + k], enc(c[k]));
        }
    }
    __syncthreads();
    if (t0 < 0) return;
    for (int w = threadIdx.x; w < kTaskBinWords; w += blockDim.x) {
        const int bin = w / kBinWords, k = w % kBinWords;
        if (lb[bin * kBinWords] == 0u) continue;
        uint32_t* g = B.bins + (size_t)t0 * kTaskBinWords + w;
        if (k == 0) atomicAdd(g, lb[w]);
        else if ((k >= 1 && k <= 3) || (k >= 7 && k <= 9)) atomicMin(g, lb[w]);
        else atomicMax(g, lb[w]);
    }
}

// One child of a large task: a single-triangle leaf, a new node with a task for
// the next round (more than kSmall triangles), or a small task that fills the
// slot later.
__device__ int32_t emit_child(const BuildDev& B, BTask* next, const BTask& T, int slot, uint32_t begin, uint32_t end,
                              const float* bmn, const float* bmx, const float* cmn, const float* cmx) {
    const uint32_t nc = end - begin;
    if (nc == 1) {
        write_slot(B, T.node, slot, bmn, bmx, begin, 1u);
        atomicAdd(&B.counters[4], 1u);
        atomicMax(&B.counters[3], T.depth + 1);
        return -1;
    }
    BTask c;
    c.begin = begin;
    c.end = end;
    c.slot = (uint32_t)slot;
    for (int k = 0; k < 3; ++k) {
        c.bmin[k] = bmn[k];
        c.bmax[k] = bmx[k];
        c.cmin[k] = cmn[k];
        c.cmax[k] = cmx[k];
    }
    if (nc <= (uint32_t)kSmall) {
        c.node = T.node;
        c.depth = T.depth + 1;
        const uint32_t k = atomicAdd(&B.counters[2], 1u);
        if (k >= B.max_small) {
            atomicAdd(&B.counters[5], 1u);
            return -1;
        }
        B.small[k] = c;
        return -1;
    }
    const uint32_t node = atomicAdd(&B.counters[0], 1u);
    if (node >= B.max_nodes) {
        atomicAdd(&B.counters[5], 1u);
        return -1;
    }
    write_slot(B, T.node, slot, bmn, bmx, node, 0u);
    c.node = node;
    c.depth = T.depth + 1;
    const uint32_t k = atomicAdd(&B.counters[1], 1u);
    if (k >= B.max_large) {
        atomicAdd(&B.counters[5], 1u);
        return -1;
    }
    next[k] = c;
    return (int32_t)k;
}

__global__ void k_bvh_split(BuildDev B, int cur, uint32_t ntask) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntask) return;
    const BTask T = B.tasks[cur][t];
    BTask* next = B.tasks[cur ^ 1];
    const uint32_t* bins = B.bins + (size_t)t * kTaskBinWords;
    const uint32_t n = T.end - T.begin;
    const float A = box_area(T.bmin, T.bmax);
    float best = INFINITY;
    int best_ax = -1, best_b = -1;
    const int ax = long_axis(T.cmin, T.cmax);
    if (T.cmax[ax] > T.cmin[ax]) {
        // suffix areas / counts
        float rmn[3] = {INFINITY, INFINITY, INFINITY}, rmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        float rarea[kBins];
        uint32_t rcnt[kBins];
        uint32_t rc = 0;
        for (int b = kBins - 1; b >= 1; --b) {
            const uint32_t* w = bins + b * kBinWords;
            if (w[0]) {
                for (int k = 0; k < 3; ++k) {
                    rmn[k] = fminf(rmn[k], dec(w[1 + k]));
                    rmx[k] = fmaxf(rmx[k], dec(w[4 + k]));
                }
            }
            rc += w[0];
            rarea[b] = box_area(rmn, rmx);
            rcnt[b] = rc;
        }
        float lmn[3] = {INFINITY, INFINITY, INFINITY}, lmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        uint32_t lc = 0;
        for (int b = 0; b < kBins - 1; ++b) {
            const uint32_t* w = bins + b * kBinWords;
            if (w[0]) {
                for (int k = 0; k < 3; ++k) {
                    lmn[k] = fminf(lmn[k], dec(w[1 + k]));
                    lmx[k] = fmaxf(lmx[k], dec(w[4 + k]));
                }
            }
            lc += w[0];
            if (lc == 0 || rcnt[b + 1] == 0) continue;
            const float cost = 0.125f + (lc * box_area(lmn, lmx) + rcnt[b + 1] * rarea[b + 1]) / A;
            if (cost < best) {
                best = cost;
                best_ax = ax;
                best_b = b;
            }
        }
    }
    float lbmn[3], lbmx[3], lcmn[3], lcmx[3], rbmn[3], rbmx[3], rcmn[3], rcmx[3];
    uint32_t nl;
    if (best_ax >= 0) {
        for (int k = 0; k < 3; ++k) {
            lbmn[k] = lcmn[k] = rbmn[k] = rcmn[k] = INFINITY;
            lbmx[k] = lcmx[k] = rbmx[k] = rcmx[k] = -INFINITY;
        }
        nl = 0;
        for (int b = 0; b < kBins; ++b) {
            const uint32_t* w = bins + b * kBinWords;
            if (!w[0]) continue;
            float* bmn = b <= best_b ? lbmn : rbmn;
            float* bmx = b <= best_b ? lbmx : rbmx;
            float* cmn = b <= best_b ? lcmn : rcmn;
            float* cmx = b <= best_b ? lcmx : rcmx;
            if (b <= best_b) nl += w[0];
            for (int k = 0; k < 3; ++k) {
                bmn[k] = fminf(bmn[k], dec(w[1 + k]));
                bmx[k] = fmaxf(bmx[k], dec(w[4 + k]));
                cmn[k] = fminf(cmn[k], dec(w[7 + k]));
                cmx[k] = fmaxf(cmx[k], dec(w[10 + k]));
            }
        }
        B.split[3 * t] = (uint32_t)best_ax;
        B.split[3 * t + 1] = (uint32_t)best_b;
    } else {
        // All centroids coincide: split the range in the middle; both halves
        // keep the task's bounds, which contain them.
        nl = n / 2;
        for (int k = 0; k < 3; ++k) {
            lbmn[k] = rbmn[k] = T.bmin[k];
            lbmx[k] = rbmx[k] = T.bmax[k];
            lcmn[k] = rcmn[k] = T.cmin[k];
            lcmx[k] = rcmx[k] = T.cmax[k];
        }
        B.split[3 * t] = 0xffffffffu;
        B.split[3 * t + 1] = 0u;
    }
    B.split[3 * t + 2] = nl;
    B.child[2 * t] = emit_child(B, next, T, 0, T.begin, T.begin + nl, lbmn, lbmx, lcmn, lcmx);
    B.child[2 * t + 1] = emit_child(B, next, T, 1, T.begin + nl, T.end, rbmn, rbmx, rcmn, rcmx);
}

// Moves each primitive to its child's range.  As in k_bvh_bin, the primitives
// of the workgroup's first task take their places with two atomics per
// workgroup (left and right totals, ranked by wave ballots); the others claim
// theirs one atomic at a time.
__global__ void __launch_bounds__(256) k_bvh_partition(BuildDev B, int cur, int32_t n) {
    __shared__ uint32_t cnt[2][4];
    __shared__ uint32_t base[2];
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t t0 = B.tid[cur][blockIdx.x * blockDim.x];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t t = i < n ? B.tid[cur][i] : -1;
    const uint32_t p = i < n ? B.idx[cur][i] : 0u;
    bool left = false;
    if (t >= 0) {
        const BTask& T = B.tasks[cur][t];
        const uint32_t ax = B.split[3 * t], sb = B.split[3 * t + 1], nl = B.split[3 * t + 2];
        if (ax == 0xffffffffu) {
            left = (uint32_t)i - T.begin < nl;
        } else {
            const float4 a = B.pmin[p], b = B.pmax[p];
            const float c = ax == 0 ? 0.5f * a.x + 0.5f * b.x : (ax == 1 ? 0.5f * a.y + 0.5f * b.y : 0.5f * a.z + 0.5f * b.z);
            left = bin_of(c, T.cmin[ax], T.cmax[ax]) <= (int)sb;
        }
    }
    const bool mine = t >= 0 && t == t0;
    const uint64_t ml = __ballot(mine && left), mr = __ballot(mine && !left);
    if (lane == 0) {
        cnt[0][wave] = (uint32_t)__popcll(ml);
        cnt[1][wave] = (uint32_t)__popcll(mr);
    }
    __syncthreads();
    if (t0 >= 0 && threadIdx.x < 2) {
        const int s = threadIdx.x;
        const uint32_t total = cnt[s][0] + cnt[s][1] + cnt[s][2] + cnt[s][3];
        base[s] = total ? atomicAdd(&B.cursor[2 * t0 + s], total) : 0u;
    }
    __syncthreads();
    if (i >= n) return;
    if (t < 0) {
        B.idx[cur ^ 1][i] = p;
        B.tid[cur ^ 1][i] = -1;
        return;
    }
    const BTask& T = B.tasks[cur][t];
    const uint32_t nl = B.split[3 * t + 2];
    uint32_t rank;
    if (mine) {
        const int s = left ? 0 : 1;
        rank = base[s] + (uint32_t)__popcll((left ? ml : mr) & ((1ull << lane) - 1ull));
        for (int w = 0; w < wave; ++w) rank += cnt[s][w];
    } else {
        rank = atomicAdd(&B.cursor[2 * t + (left ? 0 : 1)], 1u);
    }
    const uint32_t dst = left ? T.begin + rank : T.begin + nl + rank;
    B.idx[cur ^ 1][dst] = p;
    B.tid[cur ^ 1][dst] = B.child[2 * t + (left ? 0 : 1)];
}

// A small task (2..kSmall triangles) finished by one thread, in place in its
// index range: the host builder's rule on each range (16-bin SAH on the longest
// centroid axis; a leaf when it holds at most max_leaf triangles and no split is
// cheaper, median split when the bins cannot separate the centroids).
__global__ void k_bvh_small(BuildDev B, uint32_t* idx, uint32_t nsmall) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsmall) return;
    struct Item {
        uint32_t begin, end, node, depth, slot;  // fills `slot` of `node`; depth of that slot
    };
    const BTask root = B.small[s];
    // Each pop pushes two ranges that are strictly smaller, so at most one
    // pending sibling per level: kSmall entries suffice.
    Item stack[kSmall + 2];
    int sp = 0;
    stack[sp++] = Item{root.begin, root.end, root.node, root.depth, root.slot};
    while (sp > 0) {
        const Item it = stack[--sp];
        const uint32_t n = it.end - it.begin;
        float bmn[3] = {INFINITY, INFINITY, INFINITY}, bmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        float cmn[3] = {INFINITY, INFINITY, INFINITY}, cmx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = it.begin; i < it.end; ++i) {
            const uint32_t p = idx[i];
            const float4 a = B.pmin[p], b = B.pmax[p];
            const float mn[3] = {a.x, a.y, a.z}, mx[3] = {b.x, b.y, b.z};
            for (int k = 0; k < 3; ++k) {
                bmn[k] = fminf(bmn[k], mn[k]);
                bmx[k] = fmaxf(bmx[k], mx[k]);
                const float c = 0.5f * mn[k] + 0.5f * mx[k];
                cmn[k] = fminf(cmn[k], c);
                cmx[k] = fmaxf(cmx[k], c);
            }
        }
        int ax = 0;
        if (cmx[1] - cmn[1] > cmx[ax] - cmn[ax]) ax = 1;
        if (cmx[2] - cmn[2] > cmx[ax] - cmn[ax]) ax = 2;
        bool leaf = n <= 1;
        int bs = -1;
        float scale = 0.f;
        if (!leaf && cmx[ax] > cmn[ax]) {
            uint32_t cnt[kBins] = {0};
            float amn[kBins][3], amx[kBins][3];
            for (int b = 0; b < kBins; ++b)
                for (int k = 0; k < 3; ++k) {
                    amn[b][k] = INFINITY;
                    amx[b][k] = -INFINITY;
                }
            scale = kBins / (cmx[ax] - cmn[ax]);
            for (uint32_t i = it.begin; i < it.end; ++i) {
                const uint32_t p = idx[i];
                const float4 a = B.pmin[p], b = B.pmax[p];
                const float c = ax == 0 ? 0.5f * a.x + 0.5f * b.x : (ax == 1 ? 0.5f * a.y + 0.5f * b.y : 0.5f * a.z + 0.5f * b.z);
                int bi = (int)((c - cmn[ax]) * scale);
                bi = bi < 0 ? 0 : (bi >= kBins ? kBins - 1 : bi);
                cnt[bi]++;
                const float mn[3] = {a.x, a.y, a.z}, mx[3] = {b.x, b.y, b.z};
                for (int k = 0; k < 3; ++k) {
                    amn[bi][k] = fminf(amn[bi][k], mn[k]);
                    amx[bi][k] = fmaxf(amx[bi][k], mx[k]);
                }
            }
            // suffix bounds, then a prefix sweep over the 15 planes
            float rarea[kBins];
            uint32_t rcnt[kBins];
            {
                float rmn[3] = {INFINITY, INFINITY, INFINITY}, rmx[3] = {-INFINITY, -INFINITY, -INFINITY};
                uint32_t rc = 0;
                for (int b = kBins - 1; b >= 1; --b) {
                    for (int k = 0; k < 3; ++k) {
                        rmn[k] = fminf(rmn[k], amn[b][k]);
                        rmx[k] = fmaxf(rmx[k], amx[b][k]);
                    }
                    rc += cnt[b];
                    rarea[b] = box_area(rmn, rmx);
                    rcnt[b] = rc;
                }
            }
            const float A = box_area(bmn, bmx);
            float best = INFINITY;
            float lmn[3] = {INFINITY, INFINITY, INFINITY}, lmx[3] = {-INFINITY, -INFINITY, -INFINITY};
            uint32_t l = 0;
            for (int sb = 0; sb < kBins - 1; ++sb) {
                for (int k = 0; k < 3; ++k) {
                    lmn[k] = fminf(lmn[k], amn[sb][k]);
                    lmx[k] = fmaxf(lmx[k], amx[sb][k]);
                }
                l += cnt[sb];
                if (!l || !rcnt[sb + 1]) continue;
                const float cost = 0.125f + (l * box_area(lmn, lmx) + rcnt[sb + 1] * rarea[sb + 1]) / A;
                if (cost < best) {
                    best = cost;
                    bs = sb;
                }
            }
            if ((int)n <= B.max_leaf && (bs < 0 || best >= (float)n)) leaf = true;
        } else if (!leaf && (int)n <= B.max_leaf) {
            leaf = true;  // coincident centroids, few enough for one leaf
        }
        if (leaf) {
            write_slot(B, it.node, (int)it.slot, bmn, bmx, it.begin, n);
            atomicAdd(&B.counters[4], 1u);
            atomicMax(&B.counters[3], it.depth);
            continue;
        }
        uint32_t mid = it.begin + n / 2;
        if (bs >= 0) {
            uint32_t lo = it.begin, hi = it.end;
            while (lo < hi) {
                const uint32_t p = idx[lo];
                const float4 a = B.pmin[p], b = B.pmax[p];
                const float c = ax == 0 ? 0.5f * a.x + 0.5f * b.x : (ax == 1 ? 0.5f * a.y + 0.5f * b.y : 0.5f * a.z + 0.5f * b.z);
                int bi = (int)((c - cmn[ax]) * scale);
                bi = bi < 0 ? 0 : (bi >= kBins ? kBins - 1 : bi);
                if (bi <= bs) {
                    ++lo;
                } else {
                    --hi;
                    idx[lo] = idx[hi];
                    idx[hi] = p;
                }
            }
            mid = lo;
            if (mid == it.begin || mid == it.end) mid = it.begin + n / 2;
        }
        const uint32_t node = atomicAdd(&B.counters[0], 1u);
        if (node >= B.max_nodes || sp + 2 > kSmall + 2) {
            atomicAdd(&B.counters[5], 1u);
            continue;
        }
        write_slot(B, it.node, (int)it.slot, bmn, bmx, node, 0u);
        atomicMax(&B.counters[3], it.depth + 1);
        stack[sp++] = Item{mid, it.end, node, it.depth + 1, 1u};
        stack[sp++] = Item{it.begin, mid, node, it.depth + 1, 0u};
    }
}

}  // namespace

// Build on the device, then copy nodes and leaf order back for the scene upload.
int BuildBVHDevice(Ctx* c, const float* pos, int32_t n_tri, int max_leaf, BVHBuild* out) {
    if (max_leaf < 1) max_leaf = 1;
    if (max_leaf > 15) max_leaf = 15;
    if (n_tri <= 2 * kSmall) {
        // tiny scenes: the host builder (one task, nothing to parallelise)
        BuildBVHHost(pos, n_tri, max_leaf, out);
        return GY_OK;
    }
    const int64_t n = n_tri;
    BuildDev B = {};
    B.max_nodes = (uint32_t)n + 1;
    B.max_large = (uint32_t)(n / (kSmall + 1) + 2);
    B.max_small = (uint32_t)(n / 2 + 2);
    B.max_leaf = max_leaf;
    // one allocation carved into the buffers
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_pos = take(n * 9 * 4), o_pmin = take(n * 16), o_pmax = take(n * 16), o_idx0 = take(n * 4),
                 o_idx1 = take(n * 4), o_tid0 = take(n * 4), o_tid1 = take(n * 4), o_t0 = take((size_t)B.max_large * sizeof(BTask)),
                 o_t1 = take((size_t)B.max_large * sizeof(BTask)), o_small = take((size_t)B.max_small * sizeof(BTask)),
                 o_bins = take((size_t)B.max_large * kTaskBinWords * 4), o_cur = take((size_t)B.max_large * 8),
                 o_child = take((size_t)B.max_large * 8), o_split = take((size_t)B.max_large * 12),
                 o_nodes = take((size_t)B.max_nodes * 64), o_ctr = take(64), o_acc = take(64);
    char* base = nullptr;
    int rc = DevAlloc(c, (void**)&base, off);
    if (rc != GY_OK) return rc;
    auto fail = [&](hipError_t e, const char* what) {
        DevFree(base);
        return HipFail(e, what);
    };
    hipError_t e = hipMemcpyAsync(base + o_pos, pos, n * 9 * 4, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return fail(e, "bvh upload");
    B.pmin = (const float4*)(base + o_pmin);
    B.pmax = (const float4*)(base + o_pmax);
    B.idx[0] = (uint32_t*)(base + o_idx0);
    B.idx[1] = (uint32_t*)(base + o_idx1);
    B.tid[0] = (int32_t*)(base + o_tid0);
    B.tid[1] = (int32_t*)(base + o_tid1);
    B.tasks[0] = (BTask*)(base + o_t0);
    B.tasks[1] = (BTask*)(base + o_t1);
    B.small = (BTask*)(base + o_small);
    B.bins = (uint32_t*)(base + o_bins);
    B.cursor = (uint32_t*)(base + o_cur);
    B.child = (int32_t*)(base + o_child);
    B.split = (uint32_t*)(base + o_split);
    B.nodes = (uint32_t*)(base + o_nodes);
    B.counters = (uint32_t*)(base + o_ctr);
    uint32_t* acc = (uint32_t*)(base + o_acc);
    const unsigned nb = (unsigned)((n + 255) / 256);
    // counters and bounds accumulators (every node the build creates has both
    // of its slots written, so the node array needs no initialisation)
    rc = Fill32(c, B.counters, 0u, 16);
    if (rc == GY_OK) rc = Fill32(c, acc, 0xffffffffu, 3);
    if (rc == GY_OK) rc = Fill32(c, acc + 3, 0u, 3);
    if (rc == GY_OK) rc = Fill32(c, acc + 6, 0xffffffffu, 3);
    if (rc == GY_OK) rc = Fill32(c, acc + 9, 0u, 3);
    if (rc != GY_OK) {
        DevFree(base);
        return rc;
    }
    hipLaunchKernelGGL(k_bvh_prim, dim3(nb), dim3(256), 0, c->stream, (const float*)(base + o_pos), (int32_t)n,
                       (float4*)B.pmin, (float4*)B.pmax, B.idx[0], B.tid[0], acc);
    hipLaunchKernelGGL(k_bvh_root, dim3(1), dim3(64), 0, c->stream, B, (int32_t)n, (const uint32_t*)acc);
    rc = CheckLaunch(c, "bvh root");
    if (rc != GY_OK) {
        DevFree(base);
        return rc;
    }
    int cur = 0;
    uint32_t ntask = 1;
    int rounds = 0;
    while (ntask > 0) {
        if (++rounds > 200) {
            DevFree(base);
            return Fail(GY_ERR_DEVICE, "bvh build did not converge");
        }
        const unsigned cb = (unsigned)(((size_t)ntask * kTaskBinWords + 255) / 256);
        hipLaunchKernelGGL(k_bvh_bin_clear, dim3(cb), dim3(256), 0, c->stream, B, ntask);
        hipLaunchKernelGGL(k_bvh_bin, dim3(nb), dim3(256), 0, c->stream, B, cur, (int32_t)n);
        rc = Fill32(c, &B.counters[1], 0u, 1);
        if (rc != GY_OK) break;
        hipLaunchKernelGGL(k_bvh_split, dim3((ntask + 63) / 64), dim3(64), 0, c->stream, B, cur, ntask);
        hipLaunchKernelGGL(k_bvh_partition, dim3(nb), dim3(256), 0, c->stream, B, cur, (int32_t)n);
        rc = CheckLaunch(c, "bvh round");
        if (rc != GY_OK) break;
        cur ^= 1;
        uint32_t h[2];
        e = hipMemcpyAsync(h, &B.counters[1], 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return fail(e, "bvh round counters");
        ntask = std::min(h[0], B.max_large);
    }
    if (rc != GY_OK) {
        DevFree(base);
        return rc;
    }
    uint32_t h[6];
    e = hipMemcpyAsync(h, B.counters, 24, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(e, "bvh counters");
    const uint32_t nsmall = std::min(h[2], B.max_small);
    if (nsmall > 0) {
        hipLaunchKernelGGL(k_bvh_small, dim3((nsmall + 63) / 64), dim3(64), 0, c->stream, B, B.idx[cur], nsmall);
        rc = CheckLaunch(c, "bvh small");
        if (rc != GY_OK) {
            DevFree(base);
            return rc;
        }
    }
    e = hipMemcpyAsync(h, B.counters, 24, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(e, "bvh counters");
    if (h[5] != 0) {
        DevFree(base);
        return Fail(GY_ERR_DEVICE, "bvh build overflowed its buffers");
    }
    const uint32_t nn = std::min(h[0], B.max_nodes);
    out->nodes.resize((size_t)nn * 16);
    out->tri_order.resize(n);
    e = hipMemcpyAsync(out->nodes.data(), B.nodes, (size_t)nn * 64, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out->tri_order.data(), B.idx[cur], n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(e, "bvh download");
    DevFree(base);
    out->max_depth = (int32_t)h[3];
    out->n_leaves = h[4];
    out->max_leaf = max_leaf;
    out->sah_cost = BVHSahCost(*out);
    return GY_OK;
}

}  // namespace gy

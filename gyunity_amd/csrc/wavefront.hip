// Wavefront path tracer: the same estimator as the megakernel (PathTracing::Li,
// integrator/PT.cpp:10-67, with Integrator::DirectIllumination,
// integrator/Integrator.cpp:9-107) split into per-bounce stages that run as
// separate kernels over queues of path slots:
//
//   raygen -> [ trace_closest (bins hits by material)
//               -> shade<material> (emission / NEE set-up / BSDF continuation)
//               -> trace_shadow_mis (shadow rays: any hit; then MIS rays:
//                  closest hit + emission; one persistent launch)
//               -> resolve (adds T_i (L1 w1 + L2 w2) / P(light), finishes paths) ] x depth
//
// Every FP32 operation is the megakernel's, in the same order, so a render is
// bit-identical to the megakernel's.  Queues are compacted with a wavefront
// ballot and one atomic per wave; launch grids are fixed by the pass capacity
// and kernels read their queue length from device memory, so a bounce needs no
// host round trip.
//
// The kernels are written once for two sources of per-path constants:
//   OneRender    one scene / camera / film, passed as kernel arguments;
//   ManyRenders  a batch: every lane carries the id of its entry (scene, view,
//                sample budget, film) in a device table, so paths of many
//                renders of uneven size share one lane space, one set of
//                queues and the same persistent trace waves, which steal queue
//                entries 64 at a time whatever render they belong to.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "device.h"
#include "runtime.h"
#include "integrator.h"

using namespace gyd;

namespace gy {

enum QueueId { Q_ACTIVE = 0, Q_NEXT = 1, Q_SHADE0 = 2, Q_SHADOW = 6, Q_MIS = 7, Q_COUNT = 8 };
// counts[] words 16..47: work-fetch counters of the persistent trace kernels,
// eight per trace MODE 0..3 (one per queue chunk, see k_wf_trace)
constexpr int C_FETCH0 = 16;
constexpr int kChunks = 8;  // XCDs of the device; workgroup b runs on XCD b % 8
constexpr int C_WORDS = 48;
// Each counter sits in its own 128-byte line, so that atomics on different
// queues do not serialise on one L2 channel.
constexpr int kCtrStride = 32;

// Path flags (PathState.seq.w)
constexpr uint32_t F_DEPTH_MASK = 0xffu;
constexpr uint32_t F_DELTA = 1u << 8;
// Resolve flags (pend3.w as bits): pending direct light of this bounce, and the
// path's end.  Zero for a slot with nothing to resolve.
constexpr uint32_t P_HAS_DI = 1u;
constexpr uint32_t P_OCCLUDED = 2u;
constexpr uint32_t P_L2 = 4u;
constexpr uint32_t P_W1_NF = 8u;   // w1 not finite: 0 * w1 is NaN
constexpr uint32_t P_W2_NF = 16u;  // w2 not finite: 0 * w2 is NaN
constexpr uint32_t P_DONE = 32u;   // the path ended this bounce: it goes to the film
constexpr uint32_t P_NEXT = 64u;   // the path goes on without pending direct light
constexpr uint32_t P_ESC = 128u;   // the MIS ray left the scene (k_wf_env adds an environment light's L2)
// A path goes on after this bounce iff its flags are non-zero without P_DONE.

struct WfDev {
    float4* o;      // ray origin (xyz)
    float4* d;      // ray direction (xyz)
    float4* T;      // throughput (xyz), film x in w
    float4* L;      // radiance (xyz), film y in w
    uint4* seq;     // sample index lo/hi, cursor, flags
    float4* hit;    // t, prim, b1, b2
    uint8_t* mtype;  // per slot: material type of the closest hit (0..3), 255 for a miss
    float4* pend0;  // L1 w1 before the shadow test (xyz), P(light)
    float4* pend1;  // MIS ray only: BSDF value of the MIS sample (xyz) / L2 w2 after trace_mis, cos
    float4* pend2;  // MIS ray only: pdf of the MIS sample, w2, -, light id
    float4* pend3;  // throughput at this bounce (xyz), resolve flags (P_*)
    float4* sh_o;   // shadow ray origin, tmax
    float4* sh_d;   // shadow ray direction
    float4* mis_o;  // MIS ray origin
    float4* mis_d;  // MIS ray direction
    uint32_t* ent;    // batch: entry of the lane
    uint32_t* pslot;  // batch: the path's slot in its entry's pass buffer
    uint32_t* q[Q_COUNT];  // Q_ACTIVE, Q_NEXT, 4 shade queues, shadow, mis
    uint32_t* sort_q;      // ray sort: the partitioned active queue
    const uint32_t* trace_q;  // queue of the closest-hit trace when it differs from the active queue (else null)
    uint32_t* sort_hist;   // ray sort: 8 x sort_blocks octant counts, then their offsets; 8 octant totals
    uint32_t sort_blocks;  // ray sort: workgroups of kSortPer queue entries over the capacity
    uint32_t* counts;          // Q_COUNT counters
    int64_t capacity;
    uint32_t* spill;       // traversal stack spill, 64 entries per lane
    int64_t spill_stride;  // lanes per launch (grid size)
};

// ---------------------------------------------------------------- render policies
struct OneRender {
    SceneDev S;
    PathParams P;
    FilmDev F;
    static constexpr bool kBatch = false;
    __device__ uint32_t entry(const WfDev&, uint32_t) const { return 0; }
    __device__ const SceneDev& scene(uint32_t) const { return S; }
    __device__ const PathParams& params(uint32_t) const { return P; }
    __device__ const FilmDev& film(uint32_t) const { return F; }
    __device__ uint32_t pass_slot(const WfDev&, uint32_t slot) const { return slot; }
};
// Light tracing: one scene, and the splat image its connections add to.
struct LightRender : OneRender {
    float* splat;  // W x H x 3
};
struct ManyRenders {
    const BatchEntryDev* __restrict__ E;
    int32_t n;
    static constexpr bool kBatch = true;
    __device__ uint32_t entry(const WfDev& W, uint32_t slot) const { return W.ent[slot]; }
    __device__ const SceneDev& scene(uint32_t e) const { return E[e].S; }
    __device__ const PathParams& params(uint32_t e) const { return E[e].P; }
    __device__ const FilmDev& film(uint32_t e) const { return E[e].F; }
    __device__ uint32_t pass_slot(const WfDev& W, uint32_t slot) const { return W.pslot[slot]; }
};

// NQ pushes for a 256-lane workgroup with one atomic per queue: each wave
// counts its pushing lanes, one thread per queue claims the workgroup's block,
// and each lane writes at its rank within the workgroup, so entries keep lane
// order.  Same-address atomics serialise at the L2, so one per workgroup
// instead of one per wave matters for kernels that push from every lane.
// Every thread of the workgroup must call it.
template <int NQ>
__device__ inline void block_push(uint32_t* const (&q)[NQ], uint32_t* const (&cnt)[NQ], uint32_t slot, const bool (&pred)[NQ]) {
    __shared__ uint32_t nw[NQ][4];
    __shared__ uint32_t base[NQ];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint64_t mask[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        mask[k] = __ballot(pred[k]);
        if (lane == 0) nw[k][wave] = (uint32_t)__popcll(mask[k]);
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        const int k = threadIdx.x;
        const uint32_t total = nw[k][0] + nw[k][1] + nw[k][2] + nw[k][3];
        base[k] = total ? atomicAdd(cnt[k], total) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        if (!pred[k]) continue;
        uint32_t off = base[k];
        for (int w = 0; w < wave; ++w) off += nw[k][w];
        q[k][off + __popcll(mask[k] & ((1ull << lane) - 1ull))] = slot;
    }
}

// One queue, K items per thread, one atomic per workgroup.  Item k of thread t
// ranks at (k, t): entries keep the order of k * 256 + t.
template <int K>
__device__ inline void block_push_items(uint32_t* q, uint32_t* cnt, const uint32_t (&item)[K], const bool (&pred)[K]) {
    __shared__ uint32_t nw[K][4];
    __shared__ uint32_t base;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint64_t mask[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        mask[k] = __ballot(pred[k]);
        if (lane == 0) nw[k][wave] = (uint32_t)__popcll(mask[k]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t total = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) total += nw[k][0] + nw[k][1] + nw[k][2] + nw[k][3];
        base = total ? atomicAdd(cnt, total) : 0u;
    }
    __syncthreads();
    uint32_t off = base;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        uint32_t o = off;
        for (int w = 0; w < wave; ++w) o += nw[k][w];
        if (pred[k]) q[o + __popcll(mask[k] & ((1ull << lane) - 1ull))] = item[k];
        off += nw[k][0] + nw[k][1] + nw[k][2] + nw[k][3];
    }
}

__device__ inline V3 f3(const float4 v) { return v3(v.x, v.y, v.z); }

// Camera sample of one path (TiledIntegrator::Render per sample,
// integrator/TiledIntegrator.cpp:40-45) into lane `slot`.
__device__ inline void start_path(const WfDev& W, const PathParams& P, uint32_t slot, int s, int px, int py) {
    Seq q;
    q.index = enum_index(P.en, (uint32_t)s, (uint32_t)px, (uint32_t)py);
    q.cursor = 0;
    const float u = P.en.sx * seq_next_wide(P.smp, q) - (float)px;
    const float v = P.en.sy * seq_next_wide(P.smp, q) - (float)py;
    V3 o, d;
    generate_ray(P.cam, px, py, u, v, &o, &d);
    W.o[slot] = make_float4(o.x, o.y, o.z, 0.f);
    W.d[slot] = make_float4(d.x, d.y, d.z, 0.f);
    W.T[slot] = make_float4(1.f, 1.f, 1.f, px + u);
    W.L[slot] = make_float4(0.f, 0.f, 0.f, py + v);
    W.seq[slot] = make_uint4((uint32_t)q.index, (uint32_t)(q.index >> 32), q.cursor, 0u);
    W.pend3[slot] = make_float4(0.f, 0.f, 0.f, __uint_as_float(P.max_depth == 0 ? P_DONE : 0u));
}

// ---------------------------------------------------------------- raygen
// One render: samples s0 .. s0+n_s-1 of every pixel, 8x8 pixel tiles per wave.
__global__ void __launch_bounds__(256) k_wf_raygen(WfDev W, PathParams P, int s0, int n_s, int width, int height) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int tilesX = (width + 7) >> 3, tilesY = (height + 7) >> 3;
    const int64_t per_s = (int64_t)tilesX * tilesY * 64;
    bool valid = t < per_s * n_s;
    int s_local = 0, px = 0, py = 0;
    if (valid) {
        s_local = (int)(t / per_s);
        const int64_t r = t - (int64_t)s_local * per_s;
        const int64_t tile = r >> 6;
        const int local = (int)(r & 63);
        px = (int)(tile % tilesX) * 8 + (local & 7);
        py = (int)(tile / tilesX) * 8 + (local >> 3);
        valid = px < width && py < height;
    }
    const uint32_t slot = (uint32_t)((int64_t)s_local * width * height + (int64_t)py * width + px);
    if (valid) start_path(W, P, slot, s0 + s_local, px, py);
    uint32_t* const qa[1] = {W.q[Q_ACTIVE]};
    uint32_t* const ca[1] = {&W.counts[Q_ACTIVE * kCtrStride]};
    const bool pa[1] = {valid};
    block_push<1>(qa, ca, slot, pa);
}

// Batch: lanes [0, count) take global path indices base + lane.  Entry e owns
// global indices [E[e].first, E[e].first + E[e].paths) in its own tile order
// (sample-major, 8x8 pixel tiles); the lane records its entry and pass slot.
__global__ void __launch_bounds__(256) k_wf_raygen_batch(WfDev W, ManyRenders R, int64_t base, int64_t count, int64_t lanes) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    bool valid = t < count;
    uint32_t e = 0;
    int s = 0, px = 0, py = 0;
    if (valid) {
        const int64_t g = base + t;
        int lo = 0, hi = R.n - 1;  // last entry with first <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (R.E[mid].first <= g) lo = mid;
            else hi = mid - 1;
        }
        e = (uint32_t)lo;
        const BatchEntryDev& B = R.E[e];
        const int64_t r = g - B.first;
        const int64_t per_s = (int64_t)B.tilesX * B.tilesY * 64;
        s = (int)(r / per_s);
        const int64_t rr = r - (int64_t)s * per_s;
        const int64_t tile = rr >> 6;
        const int local = (int)(rr & 63);
        px = (int)(tile % B.tilesX) * 8 + (local & 7);
        py = (int)(tile / B.tilesX) * 8 + (local >> 3);
        valid = r < B.paths && px < B.P.cam.width && py < B.P.cam.height;
    }
    const uint32_t slot = (uint32_t)t;
    if (valid) {
        const PathParams& P = R.E[e].P;
        W.ent[slot] = e;
        W.pslot[slot] = (uint32_t)((int64_t)s * P.cam.width * P.cam.height + (int64_t)py * P.cam.width + px);
        start_path(W, P, slot, s, px, py);
    } else if (t < lanes) {
        W.pend3[slot].w = __uint_as_float(0u);  // resolve walks every lane: one without a path has no flags
    }
    uint32_t* const qa[1] = {W.q[Q_ACTIVE]};
    uint32_t* const ca[1] = {&W.counts[Q_ACTIVE * kCtrStride]};
    const bool pa[1] = {valid};
    block_push<1>(qa, ca, slot, pa);
}

// ---------------------------------------------------------------- persistent trace
// MODE 0: closest hit of the active queue (hits binned by material),
// MODE 1: shadow rays (any hit), MODE 2: MIS rays (closest hit + emission),
// MODE 3: light tracing's camera connections (any hit on the shadow queue;
// an unoccluded one adds its splat, R = LightRender).
// A wave claims 64 queue entries per atomic.  Traversal (trav_run) returns to
// refilling once fewer than `refill` lanes of the wave are still traversing;
// idle lanes then take new entries.  A wave exits when the queue is drained
// and none of its lanes holds a ray.
constexpr uint32_t kBatch = 64;

template <int S, int MODE, class R>
__device__ inline void wf_trace_loop(const R& Rd, const WfDev& W, int refill, int chunks, uint32_t* stack) {
    constexpr bool kAny = MODE == 1 || MODE == 3;
    constexpr int qid = MODE == 0 ? Q_ACTIVE : (kAny ? Q_SHADOW : Q_MIS);
    const uint32_t n = W.counts[(qid) * kCtrStride];
    // The queue is cut into kChunks contiguous chunks, one per XCD.  Workgroups
    // are dealt to the XCDs round-robin, so a wave starts on its own XCD's
    // chunk, where its rays are neighbours of the rays the XCD's L2 already
    // holds nodes for, and steals from the next chunks once that one is empty.
    uint32_t* fetch = &W.counts[(C_FETCH0 + MODE * kChunks) * kCtrStride];
    int ck = (int)(blockIdx.x % kChunks) % chunks, tries = 0;  // wave-uniform
    const int lane = threadIdx.x & 63;
    const int64_t gl = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t* lds = stack + threadIdx.x;
    uint32_t* spill = W.spill + gl;
    int64_t my = -1;  // -1 wants work, -2 queue drained, >= 0 queue index
    uint32_t batch_base = 0, batch_left = 0;  // wave-uniform
    bool fresh = false;
    uint32_t slot = 0, e = 0;
    SceneDev Sc;  // traversal reads nodes and tri only
    if (!R::kBatch) Sc = Rd.scene(0);
    Trav t;
    while (true) {
        // Refill idle lanes from the wave's batch of queue entries; when the
        // batch runs dry, claim the next `kBatch` entries with one atomic.
        const bool need = my == -1;
        uint64_t nm = __ballot(need);
        while (nm) {
            if (batch_left == 0) {
                while (tries < chunks) {
                    const uint32_t lo = (uint32_t)(((uint64_t)n * ck) / chunks);
                    const uint32_t hi = (uint32_t)(((uint64_t)n * (ck + 1)) / chunks);
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&fetch[ck * kCtrStride], (uint32_t)kBatch);
                    base = __shfl(base, 0, 64);
                    if (base < hi - lo) {
                        batch_base = lo + base;
                        batch_left = min((uint32_t)kBatch, hi - batch_base);
                        break;
                    }
                    ck = ck + 1 == chunks ? 0 : ck + 1;
                    ++tries;
                }
                if (batch_left == 0) {
                    if (need && my == -1) my = -2;
                    break;
                }
            }
            const uint32_t rank = (uint32_t)__popcll(nm & ((1ull << lane) - 1ull));
            const uint32_t take = min((uint32_t)__popcll(nm), batch_left);
            if (need && my == -1 && rank < take) {
                my = batch_base + rank;
                fresh = true;
            }
            batch_base += take;
            batch_left -= take;
            nm = __ballot(need && my == -1);
        }
        if (__ballot(my >= 0) == 0) break;
        if (my >= 0) {
            if (fresh) {
                fresh = false;
                slot = (MODE == 0 && W.trace_q) ? W.trace_q[my] : W.q[qid][my];
                e = Rd.entry(W, slot);
                if (R::kBatch) {
                    Sc.nodes = Rd.scene(e).nodes;
                    Sc.tri = Rd.scene(e).tri;
                }
                if (MODE == 0) {
                    trav_init(t, f3(W.o[slot]), f3(W.d[slot]), 0.f, kInf);
                } else if (kAny) {
                    const float4 o = W.sh_o[slot];
                    trav_init(t, f3(o), f3(W.sh_d[slot]), 0.f, o.w);
                } else {
                    trav_init(t, f3(W.mis_o[slot]), f3(W.mis_d[slot]), 0.f, kInf);
                }
            }
            {
                const int r = trav_run<kAny, S>(Sc, t, lds, 256, spill, W.spill_stride, refill);
                if (r != 0) {
                    if (MODE == 0) {
                        if (t.best_prim >= 0) {
                            const SceneDev& Sf = Rd.scene(e);
                            W.hit[slot] = make_float4(t.best, __int_as_float(t.best_prim), t.bb1, t.bb2);
                            const uint32_t mat = __float_as_uint(Sf.shade[(size_t)t.best_prim * 8].w);
                            W.mtype[slot] = (uint8_t)Sf.mats[(size_t)mat * 32];
                        } else {
                            W.mtype[slot] = 255;
                            W.pend3[slot].w = __uint_as_float(P_DONE);  // the path ends (k_wf_env: environment)
                        }
                    } else if (MODE == 1) {
                        // atomic: the slot's MIS ray may be in flight in the same launch
                        if (r == 2) atomicOr(reinterpret_cast<uint32_t*>(&W.pend3[slot].w), P_OCCLUDED);
                    } else if (MODE == 3) {
                        if constexpr (MODE == 3) {
                            // Film::AddSplat (visual/Film.h:89-103) of an unoccluded connection
                            if (r == 1) {
                                const float4 p = W.pend0[slot];
                                float* px = Rd.splat + (size_t)__float_as_uint(p.w) * 3;
                                atomicAdd(px, p.x);
                                atomicAdd(px + 1, p.y);
                                atomicAdd(px + 2, p.z);
                            }
                        }
                    } else if (t.best_prim < 0) {
                        // a miss counts for an environment light (k_wf_env)
                        if (Rd.scene(e).env.w > 0) atomicOr(reinterpret_cast<uint32_t*>(&W.pend3[slot].w), P_ESC);
                    } else {
                        // MIS branch of DirectIllumination (integrator/Integrator.cpp:88-98)
                        const SceneDev& Sf = Rd.scene(e);
                        const float4 p2 = W.pend2[slot];
                        const int light = __float_as_int(p2.w);
                        const int hl = __float_as_int(Sf.shade[(size_t)t.best_prim * 8 + 1].w);
                        if (hl == light) {
                            Hit h;
                            h.t = t.best;
                            h.prim = t.best_prim;
                            h.b1 = t.bb1;
                            h.b2 = t.bb2;
                            const Isect li = query(Sf, t.o, t.d, h);
                            const V3 Le = v3(Sf.light_le[3 * light], Sf.light_le[3 * light + 1], Sf.light_le[3 * light + 2]);
                            const V3 Lem = dot(li.n, neg(t.d)) > 0.f ? Le : v3(0.f, 0.f, 0.f);
                            const float4 p1 = W.pend1[slot];
                            const V3 L2 = Lem * f3(p1) * v3s(1.f) * p1.w / p2.x;
                            const V3 L2w2 = L2 * p2.y;
                            W.pend1[slot] = make_float4(L2w2.x, L2w2.y, L2w2.z, p1.w);
                            atomicOr(reinterpret_cast<uint32_t*>(&W.pend3[slot].w), P_L2);
                        }
                    }
                    my = -1;
                }
            }
        }
    }
}

template <int S, int MODE, class R>
__global__ void __launch_bounds__(256) k_wf_trace(R Rd, WfDev W, int refill, int chunks) {
    __shared__ uint32_t stack[S * 256];
    wf_trace_loop<S, MODE, R>(Rd, W, refill, chunks, stack);
}

// Shadow rays, then MIS rays, in one persistent launch: a wave that finds the
// shadow queue drained and its lanes idle goes on to the MIS queue while other
// waves finish their last shadow rays, so the small MIS trace fills the shadow
// trace's tail instead of running as a launch of its own.  The two kinds set
// different flag bits of a slot with atomic ORs.
template <int S, class R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) k_wf_trace_shadow_mis(R Rd, WfDev W, int refill, int chunks_shadow, int chunks_mis) {
    __shared__ uint32_t stack[S * 256];
    wf_trace_loop<S, 1, R>(Rd, W, refill, chunks_shadow, stack);
    wf_trace_loop<S, 2, R>(Rd, W, refill, chunks_mis, stack);
}

// ---------------------------------------------------------------- ray sort
// Octant partition of the closest-hit queue: a stable counting sort of the
// active queue by the sign bits of each ray's direction.  Rays of one octant
// visit the BVH's children in the same order, so the lanes of a wave diverge
// less; within an octant the queue keeps its order (neighbouring paths stay
// together).  No result depends on queue order: each lane's path is its own,
// and resolve walks the slots.  Only the closest-hit trace takes its rays from
// the partitioned queue: binning, shading and the other traces keep the active
// queue's order, in which neighbouring lanes read neighbouring slots.  Three
// launches: per-workgroup octant counts (the octant of queue entry i is kept in
// mtype[i]; the closest-hit trace then writes mtype of every slot in the
// queue, the only entries read after it), an exclusive scan of each octant's
// counts over the workgroups, and the scatter (octant offset + workgroup
// offset + rank in the workgroup).
constexpr int kSortPer = 1024;  // queue entries per workgroup, 4 per thread

__device__ inline uint32_t ray_octant(const float4 d) {
    return (__float_as_uint(d.x) >> 31) | ((__float_as_uint(d.y) >> 31) << 1) | ((__float_as_uint(d.z) >> 31) << 2);
}

// Lanes of the wave with the same octant as this lane, among `valid` lanes.
__device__ inline uint64_t same_octant(uint32_t oc, bool valid) {
    const uint64_t v = __ballot(valid);
    const uint64_t b0 = __ballot(valid && (oc & 1u)), b1 = __ballot(valid && (oc & 2u)), b2 = __ballot(valid && (oc & 4u));
    return v & ((oc & 1u) ? b0 : ~b0) & ((oc & 2u) ? b1 : ~b1) & ((oc & 4u) ? b2 : ~b2);
}

__global__ void __launch_bounds__(256) k_wf_sort_count(WfDev W) {
    __shared__ uint32_t h[8];
    const uint32_t n = W.counts[Q_ACTIVE * kCtrStride];
    const uint32_t base = blockIdx.x * (uint32_t)kSortPer;
    if (base >= n) return;  // workgroup-uniform; the scan reads only the workgroups below n
    if (threadIdx.x < 8) h[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t i = base + k * 256 + threadIdx.x;
        const bool valid = i < n;
        uint32_t oc = 0;
        if (valid) {
            oc = ray_octant(W.d[W.q[Q_ACTIVE][i]]);
            W.mtype[i] = (uint8_t)oc;
        }
        const uint64_t same = same_octant(oc, valid);
        if (valid && __popcll(same & ((1ull << lane) - 1ull)) == 0) atomicAdd(&h[oc], (uint32_t)__popcll(same));
    }
    __syncthreads();
    if (threadIdx.x < 8) W.sort_hist[threadIdx.x * W.sort_blocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of one octant's counts over the workgroups below n, in place
// (workgroup o scans octant o); the octant's total goes to sort_hist[8 *
// sort_blocks + o].  Each thread scans a contiguous run of the counts.
__global__ void __launch_bounds__(1024) k_wf_sort_scan(WfDev W) {
    __shared__ uint32_t part[1024];
    const uint32_t n = W.counts[Q_ACTIVE * kCtrStride];
    const uint32_t nb = (n + kSortPer - 1) / kSortPer;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t lo = min(nb, threadIdx.x * per), hi = min(nb, lo + per);
    uint32_t* h = W.sort_hist + (size_t)blockIdx.x * W.sort_blocks;
    uint32_t sum = 0;
    for (uint32_t j = lo; j < hi; ++j) sum += h[j];
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - sum;
    for (uint32_t j = lo; j < hi; ++j) {
        const uint32_t v = h[j];
        h[j] = run;
        run += v;
    }
    if (threadIdx.x == 1023) W.sort_hist[(size_t)8 * W.sort_blocks + blockIdx.x] = part[1023];
}

// Entry p = k * 256 + t of a workgroup goes to its octant's offset plus the
// entries of that octant before it in the workgroup: the order of (k, wave, lane).
// LOCAL: the partition stays within each workgroup's kSortPer entries (octant
// offsets are the workgroup's own), so the queue keeps its coarse order and
// only the waves' contents change; the octants come from the directions.
template <bool LOCAL>
__global__ void __launch_bounds__(256) k_wf_sort_scatter(WfDev W) {
    __shared__ uint32_t cnt[8][4][4];  // octant, k, wave
    __shared__ uint32_t tot[8];
    const uint32_t n = W.counts[Q_ACTIVE * kCtrStride];
    const uint32_t base = blockIdx.x * (uint32_t)kSortPer;
    if (base >= n) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (threadIdx.x < 128) (&cnt[0][0][0])[threadIdx.x] = 0;
    __syncthreads();
    uint32_t oc[4], rank[4], slot[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t i = base + k * 256 + threadIdx.x;
        const bool valid = i < n;
        slot[k] = valid ? W.q[Q_ACTIVE][i] : 0u;
        if (LOCAL) oc[k] = valid ? ray_octant(W.d[slot[k]]) : 0u;
        else oc[k] = valid ? W.mtype[i] : 0u;
        const uint64_t same = same_octant(oc[k], valid);
        rank[k] = valid ? (uint32_t)__popcll(same & ((1ull << lane) - 1ull)) : ~0u;
        if (rank[k] == 0) cnt[oc[k]][k][wave] = (uint32_t)__popcll(same);
    }
    __syncthreads();
    if (LOCAL) {
        if (threadIdx.x < 8) {
            uint32_t t = 0;
            for (int j = 0; j < 16; ++j) t += (&cnt[threadIdx.x][0][0])[j];
            tot[threadIdx.x] = t;
        }
        __syncthreads();
    }
    if (threadIdx.x < 8) {  // exclusive scan of each octant's counts in (k, wave) order, from its offset
        uint32_t run;
        if (LOCAL) {
            run = base;
            for (uint32_t o = 0; o < threadIdx.x; ++o) run += tot[o];
        } else {
            const uint32_t* gt = W.sort_hist + (size_t)8 * W.sort_blocks;
            run = W.sort_hist[threadIdx.x * W.sort_blocks + blockIdx.x];
            for (uint32_t o = 0; o < threadIdx.x; ++o) run += gt[o];
        }
        uint32_t* c = &cnt[threadIdx.x][0][0];
        for (int j = 0; j < 16; ++j) {
            const uint32_t v = c[j];
            c[j] = run;
            run += v;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (rank[k] != ~0u) W.sort_q[cnt[oc[k]][k][wave] + rank[k]] = slot[k];
}

// ---------------------------------------------------------------- binning
// The closest-hit trace writes each hit's material type at the ray's slot;
// this pass bins the active queue into the four shade queues in queue order.
// The order of the shade queues (neighbouring paths together) then does not
// depend on when each ray finished, which lets the
// closest-hit waves refill early.  Each thread takes four consecutive entries
// and the workgroup claims its block of each queue with one atomic.
constexpr int kBinPerThread = 4;
__global__ void __launch_bounds__(256) k_wf_bin(WfDev W) {
    __shared__ uint64_t wsum[4];
    __shared__ uint32_t base[4];
    const uint32_t n = W.counts[Q_ACTIVE * kCtrStride];
    const uint32_t i0 = (blockIdx.x * 256 + threadIdx.x) * kBinPerThread;
    if (blockIdx.x * 256 * kBinPerThread >= n) return;
    uint32_t slot[kBinPerThread], m[kBinPerThread];
    uint64_t cnt = 0;  // four 16-bit counts, one per material type
#pragma unroll
    for (int k = 0; k < kBinPerThread; ++k) {
        const uint32_t i = i0 + k;
        slot[k] = i < n ? W.q[Q_ACTIVE][i] : 0u;
        m[k] = i < n ? W.mtype[slot[k]] : 255u;
        if (m[k] < 4u) cnt += 1ull << (16 * m[k]);
    }
    // inclusive scan of the packed counts over the wave, then over the waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t inc = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t v = __shfl_up(inc, d, 64);
        if (lane >= d) inc += v;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    if (threadIdx.x < 4) {
        const uint32_t mt = threadIdx.x;
        const uint64_t tot = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        const uint32_t c = (uint32_t)((tot >> (16 * mt)) & 0xffffu);
        base[mt] = c ? atomicAdd(&W.counts[(Q_SHADE0 + mt) * kCtrStride], c) : 0u;
    }
    __syncthreads();
    uint64_t before = inc - cnt;
    for (int w = 0; w < wave; ++w) before += wsum[w];
#pragma unroll
    for (int k = 0; k < kBinPerThread; ++k) {
        if (m[k] >= 4u) continue;
        const uint32_t mt = m[k];
        const uint32_t pos = base[mt] + (uint32_t)((before >> (16 * mt)) & 0xffffu);
        W.q[Q_SHADE0 + mt][pos] = slot[k];
        before += 1ull << (16 * mt);
    }
}

// ---------------------------------------------------------------- environment light
// Lookups of an environment light for rays that left the scene, in a kernel of
// their own so that the trace kernels keep their registers (the lat-long
// mapping evaluates atan2 and acos in double).
// MODE 0, after the closest-hit trace: a camera ray, or one after a delta
// bounce, returns T * Le(ray) in place of L (integrator/PT.cpp:18-21).
// MODE 1, after the MIS trace: an escaped MIS ray of a path that chose the
// environment light gives L2 = Le(wi) f cos / pdf (integrator/Integrator.cpp:97-99),
// stored as L2 w2 like a hit of an area light.
template <int MODE, class R>
__global__ void __launch_bounds__(256) k_wf_env(R Rd, WfDev W) {
    const uint32_t n = W.counts[(MODE == 0 ? Q_ACTIVE : Q_MIS) * kCtrStride];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t slot = W.q[MODE == 0 ? Q_ACTIVE : Q_MIS][i];
    const SceneDev& S = Rd.scene(Rd.entry(W, slot));
    if (S.env.w == 0) return;
    if (MODE == 0) {
        if (W.mtype[slot] != 255) return;
        const uint32_t sw = W.seq[slot].w;
        if ((sw & F_DEPTH_MASK) != 0u && !(sw & F_DELTA)) return;
        const float4 Lv = W.L[slot];
        const V3 L = f3(W.T[slot]) * env_emission(S.env, f3(W.d[slot]));
        W.L[slot] = make_float4(L.x, L.y, L.z, Lv.w);
    } else {
        float& fw = W.pend3[slot].w;
        const uint32_t pf = __float_as_uint(fw);
        if (!(pf & P_ESC)) return;
        const float4 p2 = W.pend2[slot];
        if (S.light_tri[__float_as_int(p2.w)] >= 0) return;
        const float4 p1 = W.pend1[slot];
        const V3 L2 = env_emission(S.env, f3(W.mis_d[slot])) * f3(p1) * v3s(1.f) * p1.w / p2.x;
        const V3 L2w2 = L2 * p2.y;
        W.pend1[slot] = make_float4(L2w2.x, L2w2.y, L2w2.z, p1.w);
        fw = __uint_as_float(pf | P_L2);
    }
}

// ---------------------------------------------------------------- shade
// ENV: some scene holds an environment light (light triangle -1); without it
// the env branches compile away.
template <int MTYPE, class R, bool ENV>
__global__ void __launch_bounds__(256) k_wf_shade(R Rd, WfDev W) {
    const uint32_t n = W.counts[(Q_SHADE0 + MTYPE) * kCtrStride];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x * 256 >= n) return;
    const bool valid = i < n;
    uint32_t slot = 0;
    bool push_shadow = false, push_mis = false;
    if (valid) {
        slot = W.q[Q_SHADE0 + MTYPE][i];
        const uint32_t e = Rd.entry(W, slot);
        const SceneDev& S = Rd.scene(e);
        const PathParams& P = Rd.params(e);
        const SampleSrc& M = P.smp;
        const V3 ro = f3(W.o[slot]), rd = f3(W.d[slot]);
        const float4 hv = W.hit[slot];
        Hit h;
        h.t = hv.x;
        h.prim = __float_as_int(hv.y);
        h.b1 = hv.z;
        h.b2 = hv.w;
        uint4 sq = W.seq[slot];
        Seq q;
        q.index = (uint64_t)sq.x | ((uint64_t)sq.y << 32);
        q.cursor = sq.z;
        const int depth = (int)(sq.w & F_DEPTH_MASK);
        const bool delta = (sq.w & F_DELTA) != 0;
        const float4 Tv = W.T[slot];
        V3 T = f3(Tv);
        const Isect it = query(S, ro, rd, h);
        BSDF b = make_bsdf(S, it);
        b.type = MTYPE;
        uint32_t pflags = 0;
        if ((depth == 0 || delta) && it.light >= 0) {
            const V3 Le = v3(S.light_le[3 * it.light], S.light_le[3 * it.light + 1], S.light_le[3 * it.light + 2]);
            const V3 em = dot(it.n, it.outDir) > 0.f ? Le : v3(0.f, 0.f, 0.f);
            // radiance changes only here: read and written on this branch alone
            const float4 Lv = W.L[slot];
            const V3 L = f3(Lv) + T * em;
            W.L[slot] = make_float4(L.x, L.y, L.z, Lv.w);
        } else {
            const float uL = seq_next_wide(M, q);
            const float u0 = seq_next_wide(M, q), u1 = seq_next_wide(M, q);
            const float v0 = seq_next_wide(M, q), v1 = seq_next_wide(M, q), v2 = seq_next_wide(M, q);
            // DirectIllumination set-up (integrator/Integrator.cpp:9-88)
            float lsp;
            const int light = sample_one_light(S, uL, &lsp);
            const int ltri = S.light_tri[light];  // < 0: the environment light
            const V3 Le = v3(S.light_le[3 * light], S.light_le[3 * light + 1], S.light_le[3 * light + 2]);
            V3 L1pre = v3s(0.f);
            float w1 = 1.f, w2 = 1.f;
            {
                float pdf;
                LightPoint lp;
                V3 wi, Li;
                if (ENV && ltri < 0) {
                    lp = env_light_point(S.env, it.pos, u0, u1, &wi, &Li, &pdf);
                } else {
                    lp = sample_light_point(S, ltri, it.pos, u0, u1, &pdf);
                    wi = norm(lp.pos - it.pos);
                    Li = dot(lp.n, neg(wi)) > 0.f ? Le : v3(0.f, 0.f, 0.f);
                }
                if (pdf > 0.f && !is_zero(Li)) {
                    const float sp = bsdf_pdf(b, it.outDir, wi);
                    const V3 f = bsdf_eval(b, it.outDir, wi);
                    const float c = fabsf(dot(it.n, wi));
                    if (!is_zero(f)) {
                        V3 so, sd;
                        float stmax;
                        spawn_to(it, lp, &so, &sd, &stmax);
                        W.sh_o[slot] = make_float4(so.x, so.y, so.z, stmax);
                        W.sh_d[slot] = make_float4(sd.x, sd.y, sd.z, 0.f);
                        push_shadow = true;
                        w1 = power_heuristic(pdf, sp);
                        L1pre = Li * v3s(1.f) * f * c / pdf;
                    }
                }
            }
            V3 wi2;
            float pdf2;
            const V3 f2 = bsdf_sample(b, it.outDir, &wi2, &pdf2, v0, v1, v2);
            const float c2 = fabsf(dot(it.n, wi2));
            if (pdf2 != 0.f && !is_zero(f2)) {
                const float lpdf = (ENV && ltri < 0) ? env_pdf_dir(S.env, wi2) : light_pdf(S, ltri, it, wi2);
                if (lpdf == 0.f) {
                    w2 = 0.f;
                } else {
                    w2 = power_heuristic(pdf2, lpdf);
                    V3 mo;
                    spawn_ray(it, wi2, &mo);
                    W.mis_o[slot] = make_float4(mo.x, mo.y, mo.z, 0.f);
                    W.mis_d[slot] = make_float4(wi2.x, wi2.y, wi2.z, 0.f);
                    push_mis = true;
                }
            }
            // L1 w1 here, L2 w2 in the MIS trace: the products (L1 w1 + L2 w2)
            // of DirectIllumination are formed in the same FP32 operations,
            // and a zero L1 (occluded) or L2 (no MIS ray) gives +0 for them, as
            // 0 * w does for a finite weight (NaN for an infinite or NaN one,
            // flagged below).  The MIS records are written only for a path
            // that has an MIS ray.
            const V3 L1w1 = L1pre * w1;
            W.pend0[slot] = make_float4(L1w1.x, L1w1.y, L1w1.z, lsp);
            if (push_mis) {
                W.pend1[slot] = make_float4(f2.x, f2.y, f2.z, c2);
                W.pend2[slot] = make_float4(pdf2, w2, lsp, __int_as_float(light));
            }
            pflags = P_HAS_DI | (isfinite(w1) ? 0u : P_W1_NF) | (isfinite(w2) ? 0u : P_W2_NF);
            W.pend3[slot] = make_float4(T.x, T.y, T.z, __uint_as_float(pflags));
        }
        // BSDF continuation and Russian roulette (integrator/PT.cpp:45-62)
        bool alive = depth + 1 < P.max_depth;
        uint32_t nflags = sq.w;
        if (alive) {
            const float r0 = seq_next_wide(M, q), r1 = seq_next_wide(M, q), r2 = seq_next_wide(M, q);
            V3 wi;
            float pdf;
            const V3 f = bsdf_sample(b, neg(rd), &wi, &pdf, r0, r1, r2);
            if (is_zero(f) || pdf == 0.f) {
                alive = false;
            } else {
                const V3 est = f * fabsf(dot(it.n, wi)) / pdf;
                const bool ndelta = bsdf_delta(b);
                const float p = rmin(1.0f, lumY(est * T) / lumY(T));
                if (p < 1.f && depth > P.rr_depth) {
                    if (seq_next_wide(M, q) < p) T = T / p;
                    else alive = false;
                }
                if (alive) {
                    T = T * est;
                    V3 o;
                    spawn_ray(it, wi, &o);
                    W.o[slot] = make_float4(o.x, o.y, o.z, 0.f);
                    W.d[slot] = make_float4(wi.x, wi.y, wi.z, 0.f);
                    nflags = (nflags & ~(F_DEPTH_MASK | F_DELTA)) | (uint32_t)(depth + 1) | (ndelta ? F_DELTA : 0u);
                }
            }
        }
        if (alive) {  // a path that ends keeps T unchanged and needs no sample state
            W.T[slot] = make_float4(T.x, T.y, T.z, Tv.w);
            W.seq[slot] = make_uint4(sq.x, sq.y, q.cursor, nflags);
        }
        // the flags word: written with the pending record above for a path that
        // goes on; otherwise here (a path that ends, or one without direct light
        // whose flags still hold the last bounce's)
        if (!alive || !(pflags & P_HAS_DI)) W.pend3[slot].w = __uint_as_float(alive ? P_NEXT : (pflags | P_DONE));
    }
    uint32_t* const qs[2] = {W.q[Q_SHADOW], W.q[Q_MIS]};
    uint32_t* const cs[2] = {&W.counts[Q_SHADOW * kCtrStride], &W.counts[Q_MIS * kCtrStride]};
    const bool ps[2] = {push_shadow, push_mis};
    block_push<2>(qs, cs, slot, ps);
}

// ---------------------------------------------------------------- resolve
// L += T_i * (L1 w1 + L2 w2) / P(light) (integrator/Integrator.cpp:106,
// integrator/PT.cpp:38); finished paths go to the film, and paths that go on
// form the next bounce's active queue.  Returns whether the path goes on.
template <class R>
__device__ inline bool resolve_slot(const R& Rd, const WfDev& W, uint32_t slot) {
    const float4 p3 = W.pend3[slot];
    const uint32_t pf = __float_as_uint(p3.w);
    if (pf == 0u) return false;  // no path in the slot
    const bool done = (pf & P_DONE) != 0u;
    if (!(pf & P_HAS_DI) && !done) return true;
    float4 Lv = W.L[slot];
    if (pf & P_HAS_DI) {
        const float4 p0 = W.pend0[slot];
        const float nan = __builtin_nanf("");
        const V3 L1w1 = (pf & P_OCCLUDED) ? v3s((pf & P_W1_NF) ? nan : 0.f) : f3(p0);
        const V3 L2w2 = (pf & P_L2) ? f3(W.pend1[slot]) : v3s((pf & P_W2_NF) ? nan : 0.f);
        const V3 DI = (L1w1 + L2w2) / p0.w;
        const V3 L = f3(Lv) + f3(p3) * DI;
        Lv = make_float4(L.x, L.y, L.z, Lv.w);
        if (!done) {
            W.L[slot] = Lv;
            return true;
        }
    }
    W.pend3[slot].w = __uint_as_float(0u);  // resolved: later bounces skip the slot
    // Film::AddSample (visual/Film.h:61-87) through the pass buffer
    const FilmDev& F = Rd.film(Rd.entry(W, slot));
    const uint32_t ps = Rd.pass_slot(W, slot);
    const int64_t WH = (int64_t)F.width * F.height;
    const int64_t pix = ps % WH;
    film_record(F, ps, (int)(pix % F.width), (int)(pix / F.width), W.T[slot].w, Lv.w, v3(Lv.x, Lv.y, Lv.z));
    return false;
}

// The kernel walks the slots [0, n) in order, kResolvePerThread per thread
// (slot = (block * K + k) * 256 + thread), so its loads are coalesced whatever
// order the queues left the paths in, and the next active queue comes out in
// slot order: shading, which works through queues, then reads and writes the
// path records of neighbouring slots together.  A slot with no flags costs one
// 16-byte load.  A path that goes on keeps its flags, which the next bounce
// rewrites (shading, or the closest-hit miss) before they are read again.
constexpr int kResolvePerThread = 4;
template <class R>
__global__ void __launch_bounds__(256) k_wf_resolve(R Rd, WfDev W, uint32_t n) {
    uint32_t slot[kResolvePerThread];
    bool next[kResolvePerThread];
#pragma unroll
    for (int k = 0; k < kResolvePerThread; ++k) {
        slot[k] = (blockIdx.x * kResolvePerThread + k) * 256 + threadIdx.x;
        next[k] = slot[k] < n && resolve_slot(Rd, W, slot[k]);
    }
    block_push_items<kResolvePerThread>(W.q[Q_NEXT], &W.counts[Q_NEXT * kCtrStride], slot, next);
}

// Queue bookkeeping between bounces: next -> active, counters reset.
__global__ void k_wf_advance(WfDev W, unsigned long long* ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        // rays traced this bounce: camera/continuation, shadow, MIS
        ctr[0] += W.counts[(Q_ACTIVE) * kCtrStride];
        ctr[1] += W.counts[(Q_SHADOW) * kCtrStride];
        ctr[2] += W.counts[(Q_MIS) * kCtrStride];
        W.counts[(Q_ACTIVE) * kCtrStride] = W.counts[(Q_NEXT) * kCtrStride];
        W.counts[(Q_NEXT) * kCtrStride] = 0;
        for (int q = Q_SHADE0; q < C_WORDS; ++q) W.counts[(q) * kCtrStride] = 0;
    }
}
__global__ void k_wf_reset(WfDev W) {
    if (threadIdx.x < C_WORDS && blockIdx.x == 0) W.counts[(threadIdx.x) * kCtrStride] = 0;
}

// ---------------------------------------------------------------- host side
static WfDev g_view(const WavefrontState& wf, bool batch, size_t* used) {
    WfDev W;
    char* base = (char*)wf.mem;
    const int64_t n = wf.capacity;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        void* p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    W.o = (float4*)take(n * 16);
    W.d = (float4*)take(n * 16);
    W.T = (float4*)take(n * 16);
    W.L = (float4*)take(n * 16);
    W.seq = (uint4*)take(n * 16);
    W.hit = (float4*)take(n * 16);
    W.mtype = (uint8_t*)take(n);
    W.pend0 = (float4*)take(n * 16);
    W.pend1 = (float4*)take(n * 16);
    W.pend2 = (float4*)take(n * 16);
    W.pend3 = (float4*)take(n * 16);
    W.sh_o = (float4*)take(n * 16);
    W.sh_d = (float4*)take(n * 16);
    W.mis_o = (float4*)take(n * 16);
    W.mis_d = (float4*)take(n * 16);
    W.ent = batch ? (uint32_t*)take(n * 4) : nullptr;
    W.pslot = batch ? (uint32_t*)take(n * 4) : nullptr;
    for (int q = 0; q < Q_COUNT; ++q) W.q[q] = (uint32_t*)take(n * 4);
    W.sort_blocks = (uint32_t)((n + kSortPer - 1) / kSortPer);
    W.sort_q = (uint32_t*)take(n * 4);
    W.trace_q = nullptr;
    W.sort_hist = (uint32_t*)take(((size_t)W.sort_blocks * 8 + 8) * 4);
    W.counts = (uint32_t*)take(C_WORDS * kCtrStride * 4);
    W.capacity = n;
    W.spill = (uint32_t*)wf.spill;
    W.spill_stride = 0;
    *used = off;
    return W;
}

static size_t WfBytes(int64_t n, bool batch) {
    const size_t per = 14 * 16 + 1 + (Q_COUNT + 1) * 4 + (batch ? 8 : 0);
    const size_t hist = ((size_t)((n + kSortPer - 1) / kSortPer) * 8 + 8) * 4;
    return (size_t)n * per + hist + 256 * (15 + Q_COUNT + 6) + C_WORDS * kCtrStride * 4 + 4096;
}

void WavefrontStateFree(WavefrontState* wf) {
    DevFree(wf->mem);
    DevFree(wf->spill);
    *wf = WavefrontState();
}
void WavefrontFree(Scene* sc) { WavefrontStateFree(&sc->wf); }

constexpr int kStackLds = 16;

// Persistent trace grids: as many 256-lane blocks as the device keeps resident
// for the most resident of the trace kernels (the spill buffer is sized for it).
static unsigned PersistBlocks(Ctx* c) {
    static int per_cu = 0;
    if (per_cu == 0) {
        int k[6] = {0, 0, 0, 0, 0, 0};
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&k[0], k_wf_trace<kStackLds, 0, OneRender>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&k[1], k_wf_trace_shadow_mis<kStackLds, OneRender>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&k[2], k_wf_trace<kStackLds, 0, ManyRenders>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&k[3], k_wf_trace_shadow_mis<kStackLds, ManyRenders>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&k[4], k_wf_trace<kStackLds, 0, LightRender>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&k[5], k_wf_trace<kStackLds, 3, LightRender>, 256, 0);
        per_cu = std::max(1, *std::max_element(k, k + 6));
    }
    return (unsigned)(per_cu * c->num_cu);
}
static int64_t WfLanes(Ctx* c) { return (int64_t)PersistBlocks(c) * 256; }

static int WfPrepareState(Ctx* c, WavefrontState& wf, int64_t n, bool batch) {
    if (wf.capacity >= n && wf.mem && wf.batch == batch) return GY_OK;
    WavefrontStateFree(&wf);
    const size_t bytes = WfBytes(n, batch);
    int rc = DevAlloc(c, &wf.mem, bytes);
    if (rc != GY_OK) return rc;
    const size_t sbytes = (size_t)WfLanes(c) * 64 * 4;
    rc = DevAlloc(c, &wf.spill, sbytes);
    if (rc != GY_OK) return rc;
    wf.capacity = n;
    wf.bytes = bytes;
    wf.spill_bytes = sbytes;
    wf.batch = batch;
    return GY_OK;
}

// The bounce loop of one pass: `depth_max` bounces over the queues.  The
// active / next queue pointers swap on the host between bounces, fixed per
// launch, so the whole sequence can be recorded in a graph.
// Launch shape of the three trace kinds.  Queue chunks: closest-hit rays are
// claimed from eight regions of the queue at once (one per XCD), shadow rays
// from four, MIS rays from one.  Automatic refill: a lane takes a new ray once
// fewer than 48 lanes of its wave are still traversing.  Refilling does not
// reorder the shade queues, which k_wf_bin builds in queue order.
static unsigned ResolveBlocks(uint32_t n) { return (unsigned)(((uint64_t)n + 256 * kResolvePerThread - 1) / (256 * kResolvePerThread)); }
constexpr int kTraceChunks[3] = {8, 4, 1};
constexpr int kAutoRefill = 48;
template <int STACK, class R>
static int WfBounceLoop(Ctx* c, const R& Rd, const WfDev& W0, int depth_max, int refill, int ray_sort, unsigned blocks,
                        uint32_t n_slots, bool env, unsigned long long* counters) {
    WfDev W = W0;
    const unsigned pblocks = PersistBlocks(c);
    for (int depth = 0; depth < depth_max; ++depth) {
        if (ray_sort != GY_RAY_SORT_NONE && depth > 0) {
            // camera rays keep their pixel-tile order; continuation rays are partitioned
            TimerBegin(c, kKernOther);
            if (ray_sort == GY_RAY_SORT_OCTANT) {
                hipLaunchKernelGGL(k_wf_sort_count, dim3(W.sort_blocks), dim3(256), 0, c->stream, W);
                hipLaunchKernelGGL(k_wf_sort_scan, dim3(8), dim3(1024), 0, c->stream, W);
                hipLaunchKernelGGL(k_wf_sort_scatter<false>, dim3(W.sort_blocks), dim3(256), 0, c->stream, W);
            } else {
                hipLaunchKernelGGL(k_wf_sort_scatter<true>, dim3(W.sort_blocks), dim3(256), 0, c->stream, W);
            }
            TimerEnd(c);
            W.trace_q = W.sort_q;
        }
        TimerBegin(c, kKernTraceClosest);
        hipLaunchKernelGGL((k_wf_trace<STACK, 0, R>), dim3(pblocks), dim3(256), 0, c->stream, Rd, W, refill > 0 ? refill : kAutoRefill, kTraceChunks[0]);
        TimerEnd(c);
        W.trace_q = nullptr;
        TimerBegin(c, kKernOther);  // queue bookkeeping
        hipLaunchKernelGGL(k_wf_bin, dim3((blocks + kBinPerThread - 1) / kBinPerThread), dim3(256), 0, c->stream, W);
        if (env) hipLaunchKernelGGL((k_wf_env<0, R>), dim3(blocks), dim3(256), 0, c->stream, Rd, W);
        TimerEnd(c);
        TimerBegin(c, kKernShade);
        if (env) {
            hipLaunchKernelGGL((k_wf_shade<0, R, true>), dim3(blocks), dim3(256), 0, c->stream, Rd, W);
            hipLaunchKernelGGL((k_wf_shade<1, R, true>), dim3(blocks), dim3(256), 0, c->stream, Rd, W);
            hipLaunchKernelGGL((k_wf_shade<2, R, true>), dim3(blocks), dim3(256), 0, c->stream, Rd, W);
            hipLaunchKernelGGL((k_wf_shade<3, R, true>), dim3(blocks), dim3(256), 0, c->stream, Rd, W);
        } else {
            hipLaunchKernelGGL((k_wf_shade<0, R, false>), dim3(blocks), dim3(256), 0, c->stream, Rd, W);
            hipLaunchKernelGGL((k_wf_shade<1, R, false>), dim3(blocks), dim3(256), 0, c->stream, Rd, W);
            hipLaunchKernelGGL((k_wf_shade<2, R, false>), dim3(blocks), dim3(256), 0, c->stream, Rd, W);
            hipLaunchKernelGGL((k_wf_shade<3, R, false>), dim3(blocks), dim3(256), 0, c->stream, Rd, W);
        }
        TimerEnd(c);
        TimerBegin(c, kKernTraceShadow);  // shadow and MIS rays
        hipLaunchKernelGGL((k_wf_trace_shadow_mis<STACK, R>), dim3(pblocks), dim3(256), 0, c->stream, Rd, W,
                           refill > 0 ? refill : kAutoRefill, kTraceChunks[1], kTraceChunks[2]);
        TimerEnd(c);
        TimerBegin(c, kKernOther);
        if (env) hipLaunchKernelGGL((k_wf_env<1, R>), dim3(blocks), dim3(256), 0, c->stream, Rd, W);
        hipLaunchKernelGGL((k_wf_resolve<R>), dim3(ResolveBlocks(n_slots)), dim3(256), 0, c->stream, Rd, W, n_slots);
        hipLaunchKernelGGL(k_wf_advance, dim3(1), dim3(64), 0, c->stream, W, counters);
        TimerEnd(c);
        int rc = CheckLaunch(c, "wavefront bounce");
        if (rc != GY_OK) return rc;
        // the next bounce works on this bounce's continuation queue
        uint32_t* t = W.q[Q_ACTIVE];
        W.q[Q_ACTIVE] = W.q[Q_NEXT];
        W.q[Q_NEXT] = t;
    }
    return GY_OK;
}

static int RefillOf(int r) { return (r > 0 && r <= 64) ? r : 0; }

int WavefrontPass(Ctx* c, Scene* sc, const PathParams& P, const gy_render_params& p, int s0, int n_s) {
    const int width = P.cam.width, height = P.cam.height;
    const int tilesX = (width + 7) >> 3, tilesY = (height + 7) >> 3;
    const int64_t threads = (int64_t)tilesX * tilesY * 64 * n_s;
    const int64_t cap = (int64_t)width * height * n_s;
    if (cap >= (1ll << 31)) return Fail(GY_ERR_UNSUPPORTED, "wavefront pass larger than 2^31 paths");
    int rc = WfPrepareState(c, sc->wf, std::max<int64_t>(cap, 1), false);
    if (rc != GY_OK) return rc;
    size_t used = 0;
    WfDev W = g_view(sc->wf, false, &used);
    W.spill_stride = WfLanes(c);
    if ((size_t)W.spill_stride * 64 * 4 > sc->wf.spill_bytes) return Fail(GY_ERR_ALLOC, "wavefront spill buffer too small");
    if (used > sc->wf.bytes) return Fail(GY_ERR_ALLOC, "wavefront state layout exceeds its allocation");
    OneRender Rd;
    Rd.S = sc->dev;
    Rd.P = P;
    Rd.F = FilmView(sc);
    hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, c->stream, W);
    TimerBegin(c, kKernRaygen);
    hipLaunchKernelGGL(k_wf_raygen, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, c->stream, W, P, s0, n_s, width, height);
    TimerEnd(c);
    rc = CheckLaunch(c, "k_wf_raygen");
    if (rc != GY_OK) return rc;
    const unsigned blocks = (unsigned)((cap + 255) / 256);
    if (p.max_depth == 0) {
        // no bounce: every path finishes with L = 0
        hipLaunchKernelGGL((k_wf_resolve<OneRender>), dim3(ResolveBlocks((uint32_t)cap)), dim3(256), 0, c->stream, Rd, W, (uint32_t)cap);
        return CheckLaunch(c, "k_wf_resolve");
    }
    if (sc->stats.max_depth + 1 > 64) return Fail(GY_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    return WfBounceLoop<kStackLds>(c, Rd, W, p.max_depth, RefillOf(p.refill_threshold), p.ray_sort, blocks, (uint32_t)cap,
                                   sc->dev.env.w > 0,
                                   sc->film.counters);
}

// ---------------------------------------------------------------- light tracing
// The wavefront form of LightTracing (integrator/LT.h:17-183), same values as
// lt_path: lanes hold light paths; each round traces the extension rays
// (k_wf_trace MODE 0), makes each hit a vertex (k_lt_shade: its camera
// connection and its next ray), then traces the connections (MODE 3, the
// shadow queue), which splats the unoccluded ones as they finish.  Per lane:
// o, d the extension ray, T the throughput, seq (index, cursor, bounce),
// sh_o / sh_d the connection ray (tmax in sh_o.w), pend0 its splat (L, pixel).
struct LtArgs {
    LightCamDev C;
    float total;  // light paths of the render (totalPhotons)
    int32_t max_depth;
};
// ConnectToCamera + WorldToScreen (integrator/LT.h:116-183) up to the
// visibility test: the connection ray and its splat, or false.
__device__ inline bool lt_prepare(const WfDev& W, const LtArgs& A, uint32_t slot, V3 pos, V3 n, V3 outDir, const BSDF* b,
                                  V3 thr, bool s0) {
    const LightCamDev& C = A.C;
    const V3 cpos = v3(C.pos[0], C.pos[1], C.pos[2]), cz = v3(C.cz[0], C.cz[1], C.cz[2]);
    V3 toCam = cpos - pos;
    const float len = length(toCam);
    toCam = norm(toCam);
    if (dot(cz, toCam * -1.f) < 0.f) return false;
    const V3 rp = xf_point(C.m, pos);
    const int ix = (int)rp.x, iy = (int)rp.y;
    if (!(ix >= 0 && ix <= C.width && iy >= 0 && iy <= C.height)) return false;
    V3 L;
    if (s0) {
        L = thr;
    } else {
        V3 wi = cpos - pos;
        const float distance = length(wi);
        wi = norm(wi);
        const float cosT = dot(cz, wi * -1.f);
        const float pdfW = 1.f * (distance * distance) / cosT;
        const V3 wdir = wi * -1.f;
        const float cw = dot(cz, wdir);
        const float c2 = cw * cw;
        const float We = C.dis * C.dis * 1.0f / (C.area * c2 * c2);
        const V3 f = bsdf_eval(*b, outDir, wi);
        L = v3s(We) * f * thr * fabsf(dot(wi, n)) / pdfW;
    }
    L = ((L * (float)C.width) * (float)C.height) / A.total;
    int X = ix < 0 ? 0 : (ix > C.width - 1 ? C.width - 1 : ix);
    int Y = iy < 0 ? 0 : (iy > C.height - 1 ? C.height - 1 : iy);
    const V3 so = pos + toCam * kRayEps;
    W.sh_o[slot] = make_float4(so.x, so.y, so.z, len);
    W.sh_d[slot] = make_float4(toCam.x, toCam.y, toCam.z, 0.f);
    W.pend0[slot] = make_float4(L.x, L.y, L.z, __uint_as_float((uint32_t)(Y * C.width + X)));
    return true;
}

// GenerateLightPath up to the first ray (integrator/LT.h:86-113), and the light
// vertex's connection.  Lane = light path base + lane.
template <bool ENV>
__global__ void __launch_bounds__(256) k_lt_start(SceneDev S, PathParams P, WfDev W, LtArgs A, int64_t base, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool valid = t < n;
    const uint32_t slot = (uint32_t)t;
    bool conn = false;
    if (valid) {
        const SampleSrc& M = P.smp;
        Seq q;
        q.index = (uint64_t)(base + t);
        q.cursor = 0;
        float pdfLight;
        const int light = sample_one_light(S, seq_next_wide(M, q), &pdfLight);
        const float u0 = seq_next_wide(M, q), u1 = seq_next_wide(M, q);
        const float w0 = seq_next_wide(M, q), w1 = seq_next_wide(M, q);
        V3 Le, lpos, ln, dir;
        float pdfA, pdfDir;
        if (!ENV || S.light_tri[light] >= 0) {
            Le = lt_area_origin(S, light, u0, u1, w0, w1, &lpos, &ln, &dir, &pdfA, &pdfDir);
            conn = lt_prepare(W, A, slot, lpos, ln, v3s(0.f), nullptr, Le, true);
        } else {
            // environment light: the light vertex's splat is 0 and is not made
            Le = v3s(0.f);
            lt_env_origin(S.env, A.C, u0, u1, w0, w1, &lpos, &dir, &pdfA, &pdfDir);
            ln = dir;
        }
        const V3 T = Le * fabsf(dot(ln, dir)) / pdfLight / pdfA / pdfDir;
        const V3 o = (lpos + dir * kRayEps) + dir * kRayEps;
        W.o[slot] = make_float4(o.x, o.y, o.z, 0.f);
        W.d[slot] = make_float4(dir.x, dir.y, dir.z, 0.f);
        W.T[slot] = make_float4(T.x, T.y, T.z, 0.f);
        W.seq[slot] = make_uint4((uint32_t)q.index, (uint32_t)(q.index >> 32), q.cursor, 1u);
    }
    uint32_t* const qs[2] = {W.q[Q_SHADOW], W.q[Q_NEXT]};
    uint32_t* const cs[2] = {&W.counts[Q_SHADOW * kCtrStride], &W.counts[Q_NEXT * kCtrStride]};
    const bool ps[2] = {conn, valid};
    block_push<2>(qs, cs, slot, ps);
}

// Trace's loop body after an intersection (integrator/LT.h:62-83) with the
// vertex's connection.
__global__ void __launch_bounds__(256) k_lt_shade(SceneDev S, PathParams P, WfDev W, LtArgs A) {
    const uint32_t n = W.counts[Q_ACTIVE * kCtrStride];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x * 256 >= n) return;
    uint32_t slot = 0;
    bool conn = false, next = false;
    if (i < n && W.mtype[W.q[Q_ACTIVE][i]] != 255) {
        slot = W.q[Q_ACTIVE][i];
        const V3 o = f3(W.o[slot]), d = f3(W.d[slot]);
        const float4 hv = W.hit[slot];
        Hit h;
        h.t = hv.x;
        h.prim = __float_as_int(hv.y);
        h.b1 = hv.z;
        h.b2 = hv.w;
        const Isect it = query(S, o, d, h);
        BSDF b = make_bsdf(S, it);
        b.importance = true;  // ComputeScatteringFunction(arena, TransportMode::Importance) (integrator/LT.h:69)
        V3 T = f3(W.T[slot]);
        conn = lt_prepare(W, A, slot, it.pos, it.n, it.outDir, &b, T, false);
        uint4 sq = W.seq[slot];
        const uint32_t bounce = sq.w + 1u;
        if ((int)bounce < A.max_depth) {
            Seq q;
            q.index = (uint64_t)sq.x | ((uint64_t)sq.y << 32);
            q.cursor = sq.z;
            const SampleSrc& M = P.smp;
            const float r0 = seq_next_wide(M, q), r1 = seq_next_wide(M, q), r2 = seq_next_wide(M, q);
            V3 wi;
            float pdfDir;
            const V3 f = bsdf_sample(b, neg(d), &wi, &pdfDir, r0, r1, r2);
            const V3 no = (it.pos + wi * kRayEps) + wi * kRayEps;
            T = T * f * fabsf(dot(wi, it.n)) / pdfDir;
            W.o[slot] = make_float4(no.x, no.y, no.z, 0.f);
            W.d[slot] = make_float4(wi.x, wi.y, wi.z, 0.f);
            W.T[slot] = make_float4(T.x, T.y, T.z, 0.f);
            W.seq[slot] = make_uint4(sq.x, sq.y, q.cursor, bounce);
            next = true;
        }
    }
    uint32_t* const qs[2] = {W.q[Q_SHADOW], W.q[Q_NEXT]};
    uint32_t* const cs[2] = {&W.counts[Q_SHADOW * kCtrStride], &W.counts[Q_NEXT * kCtrStride]};
    const bool ps[2] = {conn, next};
    block_push<2>(qs, cs, slot, ps);
}

int LightTracingWavefront(Ctx* c, Scene* sc, const PathParams& P, const LightCamDev& C, int64_t n, float total, float* splat) {
    const int64_t cap = std::min<int64_t>(n, (int64_t)1 << 26);
    int rc = WfPrepareState(c, sc->wf, std::max<int64_t>(cap, 1), false);
    if (rc != GY_OK) return rc;
    size_t used = 0;
    WfDev W = g_view(sc->wf, false, &used);
    W.spill_stride = WfLanes(c);
    if ((size_t)W.spill_stride * 64 * 4 > sc->wf.spill_bytes) return Fail(GY_ERR_ALLOC, "wavefront spill buffer too small");
    if (used > sc->wf.bytes) return Fail(GY_ERR_ALLOC, "wavefront state layout exceeds its allocation");
    LightRender Rd;
    Rd.S = sc->dev;
    Rd.P = P;
    Rd.F = FilmView(sc);
    Rd.splat = splat;
    LtArgs A;
    A.C = C;
    A.total = total;
    A.max_depth = P.max_depth;
    const unsigned pblocks = PersistBlocks(c);
    const unsigned blocks = (unsigned)((cap + 255) / 256);
    const int rounds = std::max(1, P.max_depth - 1);  // intersections of a path (integrator/LT.h:66-76)
    void* ctr = nullptr;  // ray counts of k_wf_advance (the scene's film may not exist yet)
    rc = DevAlloc(c, &ctr, 4 * sizeof(unsigned long long));
    if (rc == GY_OK) rc = Fill32(c, ctr, 0u, 8);
    if (rc != GY_OK) {
        DevFree(ctr);
        return rc;
    }
    for (int64_t base = 0; base < n && rc == GY_OK; base += cap) {
        const int64_t cnt = std::min(cap, n - base);
        hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, c->stream, W);
        hipLaunchKernelGGL(sc->dev.env.w > 0 ? k_lt_start<true> : k_lt_start<false>, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0,
                           c->stream, sc->dev, P, W, A, base, cnt);
        for (int r = 0; r <= rounds; ++r) {
            if (r > 0) {
                hipLaunchKernelGGL((k_wf_trace<kStackLds, 0, LightRender>), dim3(pblocks), dim3(256), 0, c->stream, Rd, W, kAutoRefill,
                                   kTraceChunks[0]);
                hipLaunchKernelGGL(k_lt_shade, dim3(blocks), dim3(256), 0, c->stream, sc->dev, P, W, A);
            }
            // connections: visibility and splat in one pass
            hipLaunchKernelGGL((k_wf_trace<kStackLds, 3, LightRender>), dim3(pblocks), dim3(256), 0, c->stream, Rd, W, kAutoRefill,
                               kTraceChunks[1]);
            hipLaunchKernelGGL(k_wf_advance, dim3(1), dim3(64), 0, c->stream, W, (unsigned long long*)ctr);
            rc = CheckLaunch(c, "light tracing round");
            if (rc != GY_OK) break;
            uint32_t* tq = W.q[Q_ACTIVE];
            W.q[Q_ACTIVE] = W.q[Q_NEXT];
            W.q[Q_NEXT] = tq;
        }
    }
    // the kernels are queued on the stream: release the counters after them
    const hipError_t e = hipStreamSynchronize(c->stream);
    DevFree(ctr);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "light tracing");
    return rc;
}

// ---------------------------------------------------------------- batch
// The whole batch: clear every entry's film, run the lane space in passes of
// `lanes` paths (raygen + bounce loop to the largest depth), resolve every
// entry's film.  No host round trip, so it is recorded into a graph once.
static int BatchSequence(Ctx* c, Batch* b) {
    const int64_t lanes = b->lanes;
    size_t used = 0;
    WfDev W = g_view(b->wf, true, &used);
    W.spill_stride = WfLanes(c);
    if ((size_t)W.spill_stride * 64 * 4 > b->wf.spill_bytes) return Fail(GY_ERR_ALLOC, "wavefront spill buffer too small");
    if (used > b->wf.bytes) return Fail(GY_ERR_ALLOC, "wavefront state layout exceeds its allocation");
    ManyRenders Rd;
    Rd.E = (const BatchEntryDev*)b->table;
    Rd.n = (int32_t)b->entries.size();
    int rc = GY_OK;
    for (auto& en : b->entries) {
        if (rc == GY_OK) rc = FillF4(c, en.film.film, (int64_t)en.film.width * en.film.height);
        if (rc == GY_OK) rc = Fill32(c, en.film.border, 0u, (int64_t)en.film.width * en.film.height + 1);
    }
    if (rc == GY_OK) rc = Fill32(c, b->counters, 0u, 32);
    if (rc != GY_OK) return rc;
    const unsigned blocks = (unsigned)((lanes + 255) / 256);
    bool env = false;
    for (auto& en : b->entries) env = env || en.scene->dev.env.w > 0;
    for (int64_t base = 0; base < b->total_paths && rc == GY_OK; base += lanes) {
        const int64_t count = std::min(lanes, b->total_paths - base);
        hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, c->stream, W);
        TimerBegin(c, kKernRaygen);
        hipLaunchKernelGGL(k_wf_raygen_batch, dim3(blocks), dim3(256), 0, c->stream, W, Rd, base, count, lanes);
        TimerEnd(c);
        rc = CheckLaunch(c, "k_wf_raygen_batch");
        if (rc == GY_OK) rc = WfBounceLoop<kStackLds>(c, Rd, W, b->depth_max, b->refill, b->ray_sort, blocks, (uint32_t)lanes, env, b->counters);
    }
    for (auto& en : b->entries)
        if (rc == GY_OK) rc = FilmResolveState(c, en.film, en.params.spp);
    return rc;
}

int BatchRender(Ctx* c, Batch* b) {
    int rc = WfPrepareState(c, b->wf, b->lanes, true);
    if (rc != GY_OK) return rc;
    if (!b->use_graph) return BatchSequence(c, b);
    if (!b->graph_exec) {
        // record the sequence once; replays reuse it (fill kernels only, no memset nodes)
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return HipFail(e, "hipStreamBeginCapture");
        c->capturing = true;
        rc = BatchSequence(c, b);
        c->capturing = false;
        e = hipStreamEndCapture(c->stream, &g);
        if (rc != GY_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return HipFail(e, "hipStreamEndCapture");
        e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            b->graph_exec = nullptr;
            return HipFail(e, "hipGraphInstantiate");
        }
    }
    TimerBegin(c, kKernOther);
    const hipError_t e = hipGraphLaunch(b->graph_exec, c->stream);
    TimerEnd(c);
    return e == hipSuccess ? GY_OK : HipFail(e, "hipGraphLaunch");
}

}  // namespace gy

// Wavefront path tracer: the same estimator as the megakernel (PathTracing::Li,
// integrator/PT.cpp:10-67, with Integrator::DirectIllumination,
// integrator/Integrator.cpp:9-107) split into per-bounce stages that run as
// separate kernels over queues of path slots:
//
//   raygen -> [ trace_closest (bins hits by material)
//               -> shade<material> (emission / NEE set-up / BSDF continuation)
//               -> trace_shadow (any hit) -> trace_mis (closest hit + emission)
//               -> resolve (adds T_i (L1 w1 + L2 w2) / P(light), finishes paths) ] x depth
//
// Every FP32 operation is the megakernel's, in the same order, so a render is
// bit-identical to the megakernel's.  Queues are compacted with a wavefront
// ballot and one atomic per wave; launch grids are fixed by the pass capacity
// and kernels read their queue length from device memory, so a bounce needs no
// host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "device.h"
#include "runtime.h"

using namespace gyd;

namespace gy {



enum QueueId { Q_ACTIVE = 0, Q_NEXT = 1, Q_SHADE0 = 2, Q_SHADOW = 6, Q_MIS = 7, Q_COUNT = 8 };
// counts[] words 8..10: work-fetch counters of the persistent trace kernels
constexpr int C_FETCH0 = 8;
constexpr int C_WORDS = 16;

// Path flags (PathState.seq.w)
constexpr uint32_t F_DEPTH_MASK = 0xffu;
constexpr uint32_t F_DELTA = 1u << 8;
constexpr uint32_t F_DONE = 1u << 9;
// Pending direct-light flags (pend3.w as bits)
constexpr uint32_t P_HAS_DI = 1u;
constexpr uint32_t P_OCCLUDED = 2u;
constexpr uint32_t P_L2 = 4u;

struct WfDev {
    float4* o;      // ray origin (xyz)
    float4* d;      // ray direction (xyz)
    float4* T;      // throughput (xyz), film x in w
    float4* L;      // radiance (xyz), film y in w
    uint4* seq;     // Sobol index lo/hi, cursor, flags
    float4* hit;    // t, prim, b1, b2
    float4* pend0;  // L1 before the shadow test (xyz), w1
    float4* pend1;  // BSDF value of the MIS sample (xyz) / L2 after trace_mis, cos
    float4* pend2;  // pdf of the MIS sample, w2, P(light), light id
    float4* pend3;  // throughput at this bounce (xyz), pending flags
    float4* sh_o;   // shadow ray origin, tmax
    float4* sh_d;   // shadow ray direction
    float4* mis_o;  // MIS ray origin
    float4* mis_d;  // MIS ray direction
    uint32_t* q[Q_COUNT];  // Q_ACTIVE, Q_NEXT, 4 shade queues, shadow, mis
    uint32_t* counts;          // Q_COUNT counters
    int64_t capacity;
    uint32_t* spill;       // traversal stack spill, 64 entries per lane
    int64_t spill_stride;  // lanes per launch (grid size)
};

// Queue push compacted per wavefront: one atomic per wave, lanes write at
// base + their rank among the wave's pushing lanes.  Every lane of the wave
// must call it (pred false for lanes that push nothing).
__device__ inline void wave_push(uint32_t* queue, uint32_t* count, uint32_t slot, bool pred) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, leader, 64);
    if (pred) queue[base + __popcll(mask & ((1ull << lane) - 1ull))] = slot;
}

__device__ inline V3 f3(const float4 v) { return v3(v.x, v.y, v.z); }

// ---------------------------------------------------------------- raygen
// TiledIntegrator::Render per sample (integrator/TiledIntegrator.cpp:40-45).
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
    if (valid) {
        Seq q;
        q.index = enum_index(P.en, (uint32_t)(s0 + s_local), (uint32_t)px, (uint32_t)py);
        q.cursor = 0;
        const float u = P.en.sx * seq_next(P.smp, q) - (float)px;
        const float v = P.en.sy * seq_next(P.smp, q) - (float)py;
        V3 o, d;
        generate_ray(P.cam, px, py, u, v, &o, &d);
        W.o[slot] = make_float4(o.x, o.y, o.z, 0.f);
        W.d[slot] = make_float4(d.x, d.y, d.z, 0.f);
        W.T[slot] = make_float4(1.f, 1.f, 1.f, px + u);
        W.L[slot] = make_float4(0.f, 0.f, 0.f, py + v);
        W.seq[slot] = make_uint4((uint32_t)q.index, (uint32_t)(q.index >> 32), q.cursor, P.max_depth == 0 ? F_DONE : 0u);
        W.pend3[slot] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    wave_push(W.q[Q_ACTIVE], &W.counts[Q_ACTIVE], slot, valid);
}

// ---------------------------------------------------------------- trace
template <int STACK>
__global__ void __launch_bounds__(256) k_wf_trace_closest(SceneDev S, WfDev W) {
    __shared__ uint32_t stack[STACK * 256];
    const uint32_t n = W.counts[Q_ACTIVE];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x * 256 >= n) return;  // whole block past the queue end
    const bool valid = i < n;
    uint32_t slot = 0;
    int mtype = -1;
    if (valid) {
        slot = W.q[Q_ACTIVE][i];
        const float4 o = W.o[slot], d = W.d[slot];
        Hit h;
        if (traverse_short<false, STACK>(S, f3(o), f3(d), 0.f, kInf, &h, stack + threadIdx.x, 256, W.spill + i, W.spill_stride)) {
            W.hit[slot] = make_float4(h.t, __int_as_float(h.prim), h.b1, h.b2);
            const uint32_t mat = __float_as_uint(S.shade[(size_t)h.prim * 8].w);
            mtype = (int)S.mats[(size_t)mat * 32];
        } else {
            uint4 sq = W.seq[slot];
            sq.w |= F_DONE;
            W.seq[slot] = sq;
        }
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) wave_push(W.q[Q_SHADE0 + m], &W.counts[Q_SHADE0 + m], slot, mtype == m);
}

template <int STACK>
__global__ void __launch_bounds__(256) k_wf_trace_shadow(SceneDev S, WfDev W) {
    __shared__ uint32_t stack[STACK * 256];
    const uint32_t n = W.counts[Q_SHADOW];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t slot = W.q[Q_SHADOW][i];
    const float4 o = W.sh_o[slot], d = W.sh_d[slot];
    Hit h;
    if (traverse_short<true, STACK>(S, f3(o), f3(d), 0.f, o.w, &h, stack + threadIdx.x, 256, W.spill + i, W.spill_stride)) {
        float4 p3 = W.pend3[slot];
        p3.w = __uint_as_float(__float_as_uint(p3.w) | P_OCCLUDED);
        W.pend3[slot] = p3;
    }
}

// MIS branch of DirectIllumination (integrator/Integrator.cpp:88-98): trace the
// BSDF-sampled ray; if it hits the sampled light, L2 = Le(front) f Tr cos / pdf.
template <int STACK>
__global__ void __launch_bounds__(256) k_wf_trace_mis(SceneDev S, WfDev W) {
    __shared__ uint32_t stack[STACK * 256];
    const uint32_t n = W.counts[Q_MIS];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t slot = W.q[Q_MIS][i];
    const V3 ro = f3(W.mis_o[slot]), wi = f3(W.mis_d[slot]);
    Hit h;
    if (!traverse_short<false, STACK>(S, ro, wi, 0.f, kInf, &h, stack + threadIdx.x, 256, W.spill + i, W.spill_stride)) return;
    const float4 p2 = W.pend2[slot];
    const int light = __float_as_int(p2.w);
    const int hl = __float_as_int(S.shade[(size_t)h.prim * 8 + 1].w);
    if (hl != light) return;
    const Isect li = query(S, ro, wi, h);
    const V3 Le = v3(S.light_le[3 * light], S.light_le[3 * light + 1], S.light_le[3 * light + 2]);
    const V3 Lem = dot(li.n, neg(wi)) > 0.f ? Le : v3(0.f, 0.f, 0.f);
    const float4 p1 = W.pend1[slot];
    const V3 L2 = Lem * f3(p1) * v3s(1.f) * p1.w / p2.x;
    W.pend1[slot] = make_float4(L2.x, L2.y, L2.z, p1.w);
    float4 p3 = W.pend3[slot];
    p3.w = __uint_as_float(__float_as_uint(p3.w) | P_L2);
    W.pend3[slot] = p3;
}

// ---------------------------------------------------------------- persistent trace
// MODE 0: closest hit of the active queue (hits binned by material),
// MODE 1: shadow rays (any hit), MODE 2: MIS rays (closest hit + emission).
// Each lane takes a new queue entry as soon as its ray finishes (one atomic per
// wave); the node loop returns to refilling once fewer than `refill` lanes of
// the wave are still traversing.  A wave exits when the queue is drained and
// none of its lanes holds a ray.
constexpr uint32_t kBatch = 64;

template <int S, int MODE>
__global__ void __launch_bounds__(256) k_wf_trace(SceneDev Sc, WfDev W, int refill) {
    __shared__ uint32_t stack[S * 256];
    constexpr int qid = MODE == 0 ? Q_ACTIVE : (MODE == 1 ? Q_SHADOW : Q_MIS);
    const uint32_t n = W.counts[qid];
    uint32_t* fetch = &W.counts[C_FETCH0 + MODE];
    const int lane = threadIdx.x & 63;
    const int64_t gl = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t* lds = stack + threadIdx.x;
    uint32_t* spill = W.spill + gl;
    int64_t my = -1;  // -1 wants work, -2 queue drained, >= 0 queue index
    uint32_t batch_base = 0, batch_left = 0;  // wave-uniform
    bool fresh = false, fin = false;
    int mt = -1;
    uint32_t slot = 0;
    Trav t;
    while (true) {
        if (MODE == 0) {
#pragma unroll
            for (int m = 0; m < 4; ++m) wave_push(W.q[Q_SHADE0 + m], &W.counts[Q_SHADE0 + m], slot, fin && mt == m);
        }
        fin = false;
        // Refill idle lanes from the wave's batch of queue entries; when the
        // batch runs dry, claim the next `kBatch` entries with one atomic.
        const bool need = my == -1;
        uint64_t nm = __ballot(need);
        while (nm) {
            if (batch_left == 0) {
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(fetch, (uint32_t)kBatch);
                base = __shfl(base, 0, 64);
                batch_base = base;
                batch_left = base < n ? min((uint32_t)kBatch, n - base) : 0u;
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
                slot = W.q[qid][my];
                if (MODE == 0) {
                    trav_init(t, f3(W.o[slot]), f3(W.d[slot]), 0.f, kInf);
                } else if (MODE == 1) {
                    const float4 o = W.sh_o[slot];
                    trav_init(t, f3(o), f3(W.sh_d[slot]), 0.f, o.w);
                } else {
                    trav_init(t, f3(W.mis_o[slot]), f3(W.mis_d[slot]), 0.f, kInf);
                }
            }
            while (true) {
                const int r = trav_step<MODE == 1, S>(Sc, t, lds, 256, spill, W.spill_stride);
                if (r != 0) {
                    if (MODE == 0) {
                        if (t.best_prim >= 0) {
                            W.hit[slot] = make_float4(t.best, __int_as_float(t.best_prim), t.bb1, t.bb2);
                            const uint32_t mat = __float_as_uint(Sc.shade[(size_t)t.best_prim * 8].w);
                            mt = (int)Sc.mats[(size_t)mat * 32];
                        } else {
                            mt = -1;
                            uint4 sq = W.seq[slot];
                            sq.w |= F_DONE;
                            W.seq[slot] = sq;
                        }
                    } else if (MODE == 1) {
                        if (r == 2) {
                            float4 p3 = W.pend3[slot];
                            p3.w = __uint_as_float(__float_as_uint(p3.w) | P_OCCLUDED);
                            W.pend3[slot] = p3;
                        }
                    } else if (t.best_prim >= 0) {
                        // MIS branch of DirectIllumination (integrator/Integrator.cpp:88-98)
                        const float4 p2 = W.pend2[slot];
                        const int light = __float_as_int(p2.w);
                        const int hl = __float_as_int(Sc.shade[(size_t)t.best_prim * 8 + 1].w);
                        if (hl == light) {
                            Hit h;
                            h.t = t.best;
                            h.prim = t.best_prim;
                            h.b1 = t.bb1;
                            h.b2 = t.bb2;
                            const Isect li = query(Sc, t.o, t.d, h);
                            const V3 Le = v3(Sc.light_le[3 * light], Sc.light_le[3 * light + 1], Sc.light_le[3 * light + 2]);
                            const V3 Lem = dot(li.n, neg(t.d)) > 0.f ? Le : v3(0.f, 0.f, 0.f);
                            const float4 p1 = W.pend1[slot];
                            const V3 L2 = Lem * f3(p1) * v3s(1.f) * p1.w / p2.x;
                            W.pend1[slot] = make_float4(L2.x, L2.y, L2.z, p1.w);
                            float4 p3 = W.pend3[slot];
                            p3.w = __uint_as_float(__float_as_uint(p3.w) | P_L2);
                            W.pend3[slot] = p3;
                        }
                    }
                    fin = true;
                    my = -1;
                    break;
                }
                if (__popcll(__ballot(true)) < refill) break;
            }
        }
    }
}

// ---------------------------------------------------------------- shade
template <int MTYPE>
__global__ void __launch_bounds__(256) k_wf_shade(SceneDev S, WfDev W, PathParams P) {
    const uint32_t n = W.counts[Q_SHADE0 + MTYPE];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x * 256 >= n) return;
    const bool valid = i < n;
    uint32_t slot = 0;
    bool push_shadow = false, push_mis = false, push_next = false;
    if (valid) {
        slot = W.q[Q_SHADE0 + MTYPE][i];
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
        float4 Tv = W.T[slot], Lv = W.L[slot];
        V3 T = f3(Tv), L = f3(Lv);
        const Isect it = query(S, ro, rd, h);
        BSDF b = make_bsdf(S, it);
        b.type = MTYPE;
        uint32_t pflags = 0;
        if ((depth == 0 || delta) && it.light >= 0) {
            const V3 Le = v3(S.light_le[3 * it.light], S.light_le[3 * it.light + 1], S.light_le[3 * it.light + 2]);
            const V3 e = dot(it.n, it.outDir) > 0.f ? Le : v3(0.f, 0.f, 0.f);
            L = L + T * e;
        } else {
            const float uL = seq_next(M, q);
            const float u0 = seq_next(M, q), u1 = seq_next(M, q);
            const float v0 = seq_next(M, q), v1 = seq_next(M, q), v2 = seq_next(M, q);
            // DirectIllumination set-up (integrator/Integrator.cpp:9-88)
            float lsp;
            const int light = sample_one_light(S, uL, &lsp);
            const int ltri = S.light_tri[light];
            const V3 Le = v3(S.light_le[3 * light], S.light_le[3 * light + 1], S.light_le[3 * light + 2]);
            V3 L1pre = v3s(0.f);
            float w1 = 1.f, w2 = 1.f;
            {
                float pdf;
                const LightPoint lp = sample_light_point(S, ltri, it.pos, u0, u1, &pdf);
                const V3 wi = norm(lp.pos - it.pos);
                const V3 Li = dot(lp.n, neg(wi)) > 0.f ? Le : v3(0.f, 0.f, 0.f);
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
                const float lpdf = light_pdf(S, ltri, it, wi2);
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
            W.pend0[slot] = make_float4(L1pre.x, L1pre.y, L1pre.z, w1);
            W.pend1[slot] = make_float4(f2.x, f2.y, f2.z, c2);
            W.pend2[slot] = make_float4(pdf2, w2, lsp, __int_as_float(light));
            pflags = P_HAS_DI;
        }
        W.pend3[slot] = make_float4(T.x, T.y, T.z, __uint_as_float(pflags));
        // BSDF continuation and Russian roulette (integrator/PT.cpp:45-62)
        bool alive = depth + 1 < P.max_depth;
        uint32_t nflags = sq.w;
        if (alive) {
            const float r0 = seq_next(M, q), r1 = seq_next(M, q), r2 = seq_next(M, q);
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
                    if (seq_next(M, q) < p) T = T / p;
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
        if (!alive) nflags |= F_DONE;
        push_next = alive;
        W.T[slot] = make_float4(T.x, T.y, T.z, Tv.w);
        W.L[slot] = make_float4(L.x, L.y, L.z, Lv.w);
        W.seq[slot] = make_uint4(sq.x, sq.y, q.cursor, nflags);
    }
    wave_push(W.q[Q_SHADOW], &W.counts[Q_SHADOW], slot, push_shadow);
    wave_push(W.q[Q_MIS], &W.counts[Q_MIS], slot, push_mis);
    wave_push(W.q[Q_NEXT], &W.counts[Q_NEXT], slot, push_next);
}

// ---------------------------------------------------------------- resolve
// L += T_i * (L1 w1 + L2 w2) / P(light) (integrator/Integrator.cpp:106,
// integrator/PT.cpp:38), then finished paths go to the film.
__global__ void __launch_bounds__(256) k_wf_resolve(WfDev W, FilmDev F) {
    const uint32_t n = W.counts[Q_ACTIVE];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t slot = W.q[Q_ACTIVE][i];
    float4 Lv = W.L[slot];
    const float4 p3 = W.pend3[slot];
    const uint32_t pf = __float_as_uint(p3.w);
    if (pf & P_HAS_DI) {
        const float4 p0 = W.pend0[slot], p1 = W.pend1[slot], p2 = W.pend2[slot];
        const V3 L1 = (pf & P_OCCLUDED) ? v3s(0.f) : f3(p0);
        const V3 L2 = (pf & P_L2) ? f3(p1) : v3(0.f, 0.f, 0.f);
        const V3 DI = (L1 * p0.w + L2 * p2.y) / p2.z;
        const V3 L = f3(Lv) + f3(p3) * DI;
        Lv = make_float4(L.x, L.y, L.z, Lv.w);
        W.L[slot] = Lv;
        W.pend3[slot] = make_float4(p3.x, p3.y, p3.z, 0.f);
    }
    const uint4 sq = W.seq[slot];
    if (!(sq.w & F_DONE)) return;
    // Film::AddSample (visual/Film.h:61-87) through the pass buffer / side list.
    const int WW = F.width, H = F.height;
    const float fx = W.T[slot].w, fy = Lv.w;
    const int64_t WH = (int64_t)WW * H;
    const int s_local = (int)(slot / WH);
    const int64_t pix = slot - (int64_t)s_local * WH;
    const int px = (int)(pix % WW), py = (int)(pix / WW);
    const float x = fx - 0.5f, y = fy - 0.5f;
    int x0 = (int)ceilf(x - 0.5f), x1 = (int)floorf(x + 0.5f);
    int y0 = (int)ceilf(y - 0.5f), y1 = (int)floorf(y + 0.5f);
    x0 = x0 < 0 ? 0 : x0;
    x1 = x1 > WW - 1 ? WW - 1 : x1;
    y0 = y0 < 0 ? 0 : y0;
    y1 = y1 > H - 1 ? H - 1 : y1;
    const bool own = (x0 == px && x1 == px && y0 == py && y1 == py);
    F.pass[slot] = make_float4(Lv.x, Lv.y, Lv.z, own ? 1.f : 0.f);
    if (own) return;
    for (int yy = y0; yy <= y1; ++yy)
        for (int xx = x0; xx <= x1; ++xx) {
            const uint32_t k = atomicAdd(F.side_count, 1u);
            if (k < F.side_cap) {
                SideEntry e;
                e.key = ((uint64_t)((uint32_t)(yy * WW + xx)) << 32) | (uint64_t)slot;
                e.L = make_float4(Lv.x, Lv.y, Lv.z, 0.f);
                F.side[k] = e;
            }
        }
}

// Queue bookkeeping between bounces: next -> active, counters reset.
__global__ void k_wf_advance(WfDev W, unsigned long long* ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        // rays traced this bounce: camera/continuation, shadow, MIS
        ctr[0] += W.counts[Q_ACTIVE];
        ctr[1] += W.counts[Q_SHADOW];
        ctr[2] += W.counts[Q_MIS];
        W.counts[Q_ACTIVE] = W.counts[Q_NEXT];
        W.counts[Q_NEXT] = 0;
        for (int q = Q_SHADE0; q < C_WORDS; ++q) W.counts[q] = 0;
    }
}
__global__ void k_wf_reset(WfDev W) {
    if (threadIdx.x < C_WORDS && blockIdx.x == 0) W.counts[threadIdx.x] = 0;
}

// ---------------------------------------------------------------- host side
static WfDev g_view(Scene* sc, size_t* used) {
    WfDev W;
    char* base = (char*)sc->wf.mem;
    const int64_t n = sc->wf.capacity;
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
    W.pend0 = (float4*)take(n * 16);
    W.pend1 = (float4*)take(n * 16);
    W.pend2 = (float4*)take(n * 16);
    W.pend3 = (float4*)take(n * 16);
    W.sh_o = (float4*)take(n * 16);
    W.sh_d = (float4*)take(n * 16);
    W.mis_o = (float4*)take(n * 16);
    W.mis_d = (float4*)take(n * 16);
    for (int q = 0; q < Q_COUNT; ++q) W.q[q] = (uint32_t*)take(n * 4);
    W.counts = (uint32_t*)take(64);
    W.capacity = n;
    *used = off;
    return W;
}

static size_t WfBytes(int64_t n) {
    const size_t per = 14 * 16 + Q_COUNT * 4;
    return (size_t)n * per + 256 * (14 + Q_COUNT + 2) + 4096;
}

void WavefrontFree(Scene* sc) {
    DevFree(sc->wf.mem);
    DevFree(sc->wf.spill);
    sc->wf = WavefrontState();
}

constexpr int kStackLds = 16;

// Persistent trace grids: as many 256-lane blocks as the device keeps resident.
static unsigned PersistBlocks(Ctx* c) {
    static int per_cu = 0;
    if (per_cu == 0) {
        int a = 0, b = 0, d = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_wf_trace<kStackLds, 0>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_wf_trace<kStackLds, 1>, 256, 0);
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&d, k_wf_trace<kStackLds, 2>, 256, 0);
        per_cu = std::max(1, std::max(a, std::max(b, d)));
    }
    return (unsigned)(per_cu * c->num_cu);
}
static int64_t WfLanes(Ctx* c) { return (int64_t)PersistBlocks(c) * 256; }

static int WfPrepare(Ctx* c, Scene* sc, int64_t n) {
    if (sc->wf.capacity >= n && sc->wf.mem) return GY_OK;
    WavefrontFree(sc);
    const size_t bytes = WfBytes(n);
    int rc = DevAlloc(c, &sc->wf.mem, bytes);
    if (rc != GY_OK) return rc;
    const size_t sbytes = (size_t)WfLanes(c) * 64 * 4;
    rc = DevAlloc(c, &sc->wf.spill, sbytes);
    if (rc != GY_OK) return rc;
    sc->wf.capacity = n;
    sc->wf.bytes = bytes;
    sc->wf.spill_bytes = sbytes;
    return GY_OK;
}

template <int STACK>
static int WfBounceLoop(Ctx* c, Scene* sc, const WfDev& W0, const PathParams& P, const gy_render_params& p, unsigned blocks) {
    const FilmDev F = FilmView(sc);
    WfDev W = W0;
    const unsigned pblocks = PersistBlocks(c);
    const int refill = (p.refill_threshold > 0 && p.refill_threshold <= 64) ? p.refill_threshold : 1;
    for (int depth = 0; depth < p.max_depth; ++depth) {
        TimerBegin(c, kKernTraceClosest);
        hipLaunchKernelGGL((k_wf_trace<STACK, 0>), dim3(pblocks), dim3(256), 0, c->stream, sc->dev, W, refill);
        TimerEnd(c);
        TimerBegin(c, kKernShade);
        hipLaunchKernelGGL(k_wf_shade<0>, dim3(blocks), dim3(256), 0, c->stream, sc->dev, W, P);
        hipLaunchKernelGGL(k_wf_shade<1>, dim3(blocks), dim3(256), 0, c->stream, sc->dev, W, P);
        hipLaunchKernelGGL(k_wf_shade<2>, dim3(blocks), dim3(256), 0, c->stream, sc->dev, W, P);
        hipLaunchKernelGGL(k_wf_shade<3>, dim3(blocks), dim3(256), 0, c->stream, sc->dev, W, P);
        TimerEnd(c);
        TimerBegin(c, kKernTraceShadow);
        hipLaunchKernelGGL((k_wf_trace<STACK, 1>), dim3(pblocks), dim3(256), 0, c->stream, sc->dev, W, refill);
        TimerEnd(c);
        TimerBegin(c, kKernTraceMis);
        hipLaunchKernelGGL((k_wf_trace<STACK, 2>), dim3(pblocks), dim3(256), 0, c->stream, sc->dev, W, refill);
        TimerEnd(c);
        TimerBegin(c, kKernOther);
        hipLaunchKernelGGL(k_wf_resolve, dim3(blocks), dim3(256), 0, c->stream, W, F);
        hipLaunchKernelGGL(k_wf_advance, dim3(1), dim3(64), 0, c->stream, W, sc->film.counters);
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

int WavefrontPass(Ctx* c, Scene* sc, const PathParams& P, const gy_render_params& p, int s0, int n_s) {
    const int width = P.cam.width, height = P.cam.height;
    const int tilesX = (width + 7) >> 3, tilesY = (height + 7) >> 3;
    const int64_t threads = (int64_t)tilesX * tilesY * 64 * n_s;
    const int64_t cap = (int64_t)width * height * n_s;
    if (cap >= (1ll << 31)) return Fail(GY_ERR_UNSUPPORTED, "wavefront pass larger than 2^31 paths");
    int rc = WfPrepare(c, sc, std::max<int64_t>(cap, 1));
    if (rc != GY_OK) return rc;
    size_t used = 0;
    WfDev W = g_view(sc, &used);
    W.spill = (uint32_t*)sc->wf.spill;
    W.spill_stride = WfLanes(c);
    if ((size_t)W.spill_stride * 64 * 4 > sc->wf.spill_bytes) return Fail(GY_ERR_ALLOC, "wavefront spill buffer too small");
    if (used > sc->wf.bytes) return Fail(GY_ERR_ALLOC, "wavefront state layout exceeds its allocation");
    hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, c->stream, W);
    TimerBegin(c, kKernRaygen);
    hipLaunchKernelGGL(k_wf_raygen, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, c->stream, W, P, s0, n_s, width, height);
    TimerEnd(c);
    rc = CheckLaunch(c, "k_wf_raygen");
    if (rc != GY_OK) return rc;
    const unsigned blocks = (unsigned)((cap + 255) / 256);
    if (p.max_depth == 0) {
        // no bounce: every path finishes with L = 0
        hipLaunchKernelGGL(k_wf_resolve, dim3(blocks), dim3(256), 0, c->stream, W, FilmView(sc));
        return CheckLaunch(c, "k_wf_resolve");
    }
    if (sc->stats.max_depth + 1 > 64) return Fail(GY_ERR_UNSUPPORTED, "BVH deeper than the traversal stack");
    return WfBounceLoop<kStackLds>(c, sc, W, P, p, blocks);
}

}  // namespace gy

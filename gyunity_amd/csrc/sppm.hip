// Stochastic progressive photon mapping (integrator/SPPM.cpp:11-277) on the
// GPU, for triangle scenes with area lights.  Each iteration:
//   1. eye pass, one thread per pixel (SPPM::TraceEyePath): delta bounces,
//      then a hit point (importance, position, |normal|, outgoing direction)
//      and its direct light (DirectIllumination, or emission on a camera /
//      after-delta hit of a light) added to the pixel's own sum;
//   2. the reference's spatial hash (integrator/HashGrid.h): one bucket per
//      hit point in use, cells of twice the search radius over the hit points'
//      box inflated by it, each hit point listed in the bucket of every cell
//      its radius box touches; built with per-bucket counts, a scan and a fill;
//   3. photons (GeneratePhoton + TracePhoton): at every non-delta vertex
//      after the first, each listed hit point within the iteration's starting
//      radius and facing the same way becomes a record (hit point, photon *
//      depth + vertex, contribution, distance^2).  A photon's path does not
//      depend on the hit points, so the paths of many iterations are traced
//      in one launch (one thread per photon; 300,000 photons alone are too few
//      threads to hide the traversal's latency), each depositing vertex kept
//      in its iteration's vertex list; the iteration's deposit pass then runs
//      one thread per vertex against that iteration's hash grid;
//   4. the records are sorted by (hit point, photon, vertex) and each hit
//      point applies its own in that order: the progressive radius/flux update
//      with the radius test at the current radius, or (batch shrink) the count
//      and flux sum followed by the per-iteration update.
// The reference lets photons update hit points concurrently without locks, so
// its result depends on thread timing; photon order is the canonical order
// here, the same arithmetic in one defined sequence.  At the end the image is
// flux / (pi r^2 iterations photons) + direct / iterations (Film::SetImage).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "integrator.h"

namespace gy {
namespace {

using gyd::V3;
using gyd::v3;
using gyd::v3s;

struct SppmDev {
    float4 *pos, *nrm, *out, *imp;  // hit points: pos.w = used
    float *r2, *N;
    float4 *flux, *direct;
    uint32_t* m;
    int W, H;
};
struct GridDev {
    const uint32_t* start;  // num + 1 bucket offsets
    uint32_t* cursor;
    float4* entries;        // 2 per entry: hit point position + radius^2, |normal| + hit point (pixel)
    float3 bmin;
    float inv;
    int64_t num;
};
struct RecDev {
    unsigned long long* keys;  // (hit point << 32) | (photon * max_depth + vertex)
    float4* val;               // contribution, distance^2
    unsigned int* count;
    unsigned int cap;
};
// Depositing photon vertices, one list of `cap` per iteration of a chunk:
// 4 float4 each: ray origin + hit t, ray direction + primitive, barycentrics
// + photon + vertex, photon flux.  The deposit pass rebuilds the surface
// interaction and BSDF from the ray and hit, as the tracing thread had them.
struct VtxDev {
    float4* v;
    unsigned int* count;  // per iteration of the chunk; counts past cap too
    unsigned int cap;
};

__device__ inline V3 f3(float4 a) { return v3(a.x, a.y, a.z); }
__device__ inline float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ inline V3 mul_m1(V3 d) { return d * -1.f; }  // -1 * v (a product, not 0 - v)

// HashGrid::hash: (ix * 73856093) ^ (iy * 19349663) ^ (iz * 83492791), int
// products wrapping, as uint32, modulo the bucket count.
__device__ inline uint32_t grid_hash(int ix, int iy, int iz, int64_t num) {
    const uint32_t h = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u) ^ ((uint32_t)iz * 83492791u);
    return (uint32_t)((int64_t)h % num);
}
// Appends for the lanes with pred set, one atomic per (wave, list): lane l of
// list r gets counters[r] + its rank among that list's lanes.  Every lane of
// the wave that reaches the call must make it.
__device__ inline uint32_t wave_append(unsigned int* counters, uint32_t list, uint32_t n, bool pred) {
    const int lane = threadIdx.x & 63;
    uint64_t pend = __ballot(pred && n > 0);
    uint32_t mine = 0;
    while (pend) {
        const int leader = __ffsll((unsigned long long)pend) - 1;
        const uint32_t lr = __shfl(list, leader, 64);
        const bool in = pred && n > 0 && list == lr;
        const uint64_t grp = __ballot(in);
        // exclusive prefix of n over the group's lanes
        uint32_t incl = in ? n : 0u;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        const uint32_t total = __shfl(incl, 63, 64);
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&counters[lr], total);
        base = __shfl(base, leader, 64);
        if (in) mine = base + incl - n;
        pend &= ~grp;
    }
    return mine;
}
__device__ inline uint32_t ordered(float f) {  // float -> uint32 with the same order
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// SPPM::TraceEyePath for pixel p of iteration iter (sample index
// SamplerEnum::GetIndex(iter, x, y), SPPM.cpp:212).
// Occupancy hints on the three per-thread kernels (eye 4, photon paths 5,
// deposit 6 waves per SIMD) trade a few spilled registers for waves: 4.6%
// faster per iteration on the reference's settings.
template <int STACK>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4))) k_sppm_eye(gyd::SceneDev S, PathParams P, SppmDev D, int iter) {
    __shared__ uint32_t stack[STACK * 128];
    const int64_t p = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if (p >= (int64_t)D.W * D.H) return;
    const int px = (int)(p % D.W), py = (int)(p / D.W);
    uint32_t* st = stack + threadIdx.x;
    const gyd::SampleSrc& M = P.smp;
    gyd::Seq q;
    q.index = gyd::enum_index(P.en, (uint32_t)iter, (uint32_t)px, (uint32_t)py);
    q.cursor = 0;
    const float u = P.en.sx * gyd::seq_next(M, q) - (float)px;
    const float v = P.en.sy * gyd::seq_next(M, q) - (float)py;
    V3 o, d;
    gyd::generate_ray(P.cam, px, py, u, v, &o, &d);
    bool delta = false;
    V3 importance = v3(1.f, 1.f, 1.f);
    float4 pos = D.pos[p];
    pos.w = 0.f;  // used = false
    V3 add = v3s(0.f);
    bool has_add = false;
    for (int i = 0; i < P.max_depth; ++i) {
        gyd::Hit h;
        if (!gyd::find_hit<false, STACK>(S, o, d, 0.f, gyd::kInf, &h, st, 128)) break;
        const gyd::Isect it = gyd::query(S, o, d, h);
        const gyd::BSDF b = gyd::make_bsdf(S, it);
        if (gyd::bsdf_delta(b)) {
            const float r0 = gyd::seq_next(M, q), r1 = gyd::seq_next(M, q), r2 = gyd::seq_next(M, q);
            V3 wi;
            float pdf;
            const V3 f = gyd::bsdf_sample(b, mul_m1(d), &wi, &pdf, r0, r1, r2);
            if (gyd::is_zero(f) || pdf == 0.f) break;
            importance = f * fabsf(gyd::dot(wi, it.n)) * importance / pdf;
            gyd::spawn_ray(it, wi, &o);
            d = wi;
            delta = true;
        } else {
            pos = f4(it.pos, 1.f);
            D.nrm[p] = f4(it.absN, 0.f);
            D.out[p] = f4(mul_m1(d), 0.f);
            D.imp[p] = f4(importance, 0.f);
            if ((i == 0 || delta) && it.light >= 0) {
                const V3 Le = v3(S.light_le[3 * it.light], S.light_le[3 * it.light + 1], S.light_le[3 * it.light + 2]);
                add = importance * (gyd::dot(it.n, it.outDir) > 0.f ? Le : v3s(0.f));
            } else {
                const float uL = gyd::seq_next(M, q);
                const float u0 = gyd::seq_next(M, q), u1 = gyd::seq_next(M, q);
                const float v0 = gyd::seq_next(M, q), v1 = gyd::seq_next(M, q), v2 = gyd::seq_next(M, q);
                PathCounters pc = {};
                add = importance * direct_illumination<STACK, false, false>(S, it, b, uL, u0, u1, v0, v1, v2, st, 128, pc);
            }
            has_add = true;
            break;
        }
    }
    D.pos[p] = pos;
    if (has_add) D.direct[p] = f4(f3(D.direct[p]) + add, 0.f);
}

__device__ inline uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = max(v, (uint32_t)__shfl_xor(v, m, 64));
    return v;
}
__device__ inline uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = min(v, (uint32_t)__shfl_xor(v, m, 64));
    return v;
}
// Hit points in use: count and the largest radius^2 (non-negative floats
// order as their bits).  A grid-stride loop over few workgroups, reduced over
// each wave: one atomic per wave and quantity.
constexpr unsigned kReduceBlocks = 256;
__global__ void __launch_bounds__(256) k_sppm_stats(SppmDev D, unsigned int* out) {
    uint32_t n = 0, mx = 0;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < (int64_t)D.W * D.H; p += (int64_t)gridDim.x * 256)
        if (D.pos[p].w != 0.f) {
            ++n;
            mx = max(mx, __float_as_uint(D.r2[p]));
        }
    mx = wave_max_u32(mx);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) n += (uint32_t)__shfl_xor(n, m, 64);
    if ((threadIdx.x & 63) == 0 && n > 0) {
        atomicAdd(&out[0], n);
        atomicMax(&out[1], mx);
    }
}
// The lower corner of the hit points' box inflated by the search radius.
__global__ void __launch_bounds__(256) k_sppm_bbox(SppmDev D, float search, unsigned int* out) {
    uint32_t x = 0xFFFFFFFFu, y = 0xFFFFFFFFu, z = 0xFFFFFFFFu;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < (int64_t)D.W * D.H; p += (int64_t)gridDim.x * 256) {
        const float4 a = D.pos[p];
        if (a.w != 0.f) {
            x = min(x, ordered(a.x - search));
            y = min(y, ordered(a.y - search));
            z = min(z, ordered(a.z - search));
        }
    }
    x = wave_min_u32(x);
    y = wave_min_u32(y);
    z = wave_min_u32(z);
    if ((threadIdx.x & 63) == 0 && x != 0xFFFFFFFFu) {
        atomicMin(&out[0], x);
        atomicMin(&out[1], y);
        atomicMin(&out[2], z);
    }
}
template <bool FILL>
__global__ void k_sppm_grid(SppmDev D, GridDev G, uint32_t* counts) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (int64_t)D.W * D.H || D.pos[p].w == 0.f) return;
    const V3 pos = f3(D.pos[p]);
    const float rad = sqrtf(D.r2[p]);
    const V3 mn = v3(G.bmin.x, G.bmin.y, G.bmin.z);
    const V3 BMin = ((pos - v3s(rad)) - mn) * G.inv;
    const V3 BMax = ((pos + v3s(rad)) - mn) * G.inv;
    for (int iz = abs(int(BMin.z)); iz <= abs(int(BMax.z)); iz++)
        for (int iy = abs(int(BMin.y)); iy <= abs(int(BMax.y)); iy++)
            for (int ix = abs(int(BMin.x)); ix <= abs(int(BMax.x)); ix++) {
                const uint32_t h = grid_hash(ix, iy, iz, G.num);
                if (FILL) {
                    const size_t k = atomicAdd(&G.cursor[h], 1u);
                    G.entries[2 * k] = make_float4(pos.x, pos.y, pos.z, D.r2[p]);
                    const float4 nv = D.nrm[p];
                    G.entries[2 * k + 1] = make_float4(nv.x, nv.y, nv.z, __uint_as_float((uint32_t)p));
                }
                else atomicAdd(&counts[h], 1u);
            }
}

// SPPM::GeneratePhoton + TracePhoton for photon j of iteration iter0 + r
// (thread r * photons + j; sampler instance iter * photons + j).  Each
// depositing vertex (non-delta, after the first) goes to list r.
template <int STACK>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(5))) k_sppm_trace(gyd::SceneDev S, PathParams P, VtxDev V, int iter0, int n_iter,
                                                    int64_t photons) {
    __shared__ uint32_t stack[STACK * 128];
    const int64_t t = (int64_t)blockIdx.x * 128 + threadIdx.x;
    if ((int64_t)blockIdx.x * 128 >= photons * n_iter) return;  // workgroup-uniform
    const bool valid = t < photons * n_iter;
    const uint32_t r = valid ? (uint32_t)(t / photons) : 0u;
    const int64_t j = valid ? t - (int64_t)r * photons : 0;
    uint32_t* st = stack + threadIdx.x;
    const gyd::SampleSrc& M = P.smp;
    gyd::Seq q;
    q.index = (uint64_t)(iter0 + (int64_t)r) * (uint64_t)photons + (uint64_t)j;
    q.cursor = 0;
    bool alive = valid;
    V3 o = v3s(0.f), d = v3s(0.f), flux = v3s(0.f);
    if (alive) {
        float lightPdf;
        const int light = gyd::sample_one_light(S, gyd::seq_next(M, q), &lightPdf);
        const float u0 = gyd::seq_next(M, q), u1 = gyd::seq_next(M, q);
        const float w0 = gyd::seq_next(M, q), w1 = gyd::seq_next(M, q);
        // AreaLight::SampleLight = SampleOnLight (light/AreaLight.h:60-71): Triangle::Sample, cosine direction
        const int ltri = S.light_tri[light];
        const float4* tq = S.shade + (size_t)ltri * 8;
        const V3 p0 = gyd::ld3(tq[0]), p1 = gyd::ld3(tq[1]), p2 = gyd::ld3(tq[2]);
        const V3 n0 = gyd::ld3(tq[3]), n1 = gyd::ld3(tq[4]), n2 = gyd::ld3(tq[5]), fn = gyd::ld3(tq[6]);
        const float su0 = sqrtf(u0);
        const float b0 = 1 - su0, b1 = u1 * su0;
        const V3 lpos = (b0 * p0 + b1 * p1) + (1 - b0 - b1) * p2;
        const V3 ln = (b0 * n0 + b1 * n1) + (1 - b0 - b1) * n2;
        const V3 lerr = gyd::gamma_n(6) * ((gyd::vabs(b0 * p0) + gyd::vabs(b1 * p1)) + gyd::vabs((1 - b0 - b1) * p2));
        const float pdfPos = 1 / tq[2].w;
        V3 ss, ts;
        gyd::coordinate_system(ln, &ss, &ts);
        const V3 dl = gyd::cosine_hemisphere(w0, w1);
        d = gyd::norm((ss * dl.x + ts * dl.y) + ln * dl.z);
        const float pdfDir = dl.z * gyd::kInvPI;
        o = gyd::offset_origin(lpos, lerr, fn, d) + d * gyd::kRayEps;  // lightPoint.SpawnRay
        const float cosTheta = fabsf(gyd::dot(ln, d));
        const V3 Le = v3(S.light_le[3 * light], S.light_le[3 * light + 1], S.light_le[3 * light + 2]);
        flux = (pdfPos == 0.f || pdfDir == 0.f) ? v3s(0.f) : Le * cosTheta / (pdfPos * pdfDir * lightPdf);
    }
    for (int i = 0; i < P.max_depth; ++i) {
        if (__ballot(alive) == 0) break;
        gyd::Hit h;
        bool deposit = false;
        V3 o_next = o, wi = d, flux_next = flux;
        if (alive) {
            if (!gyd::find_hit<false, STACK>(S, o, d, 0.f, gyd::kInf, &h, st, 128)) {
                alive = false;
            } else {
                const gyd::Isect it = gyd::query(S, o, d, h);
                gyd::BSDF b = gyd::make_bsdf(S, it);
                b.importance = true;  // ComputeScatteringFunction(arena, TransportMode::Importance)
                const float r0 = gyd::seq_next(M, q), r1 = gyd::seq_next(M, q), r2 = gyd::seq_next(M, q);
                float pdf;
                const V3 f = gyd::bsdf_sample(b, mul_m1(d), &wi, &pdf, r0, r1, r2);
                if (gyd::is_zero(f) || pdf == 0.f) {
                    alive = false;
                } else {
                    const V3 est = f * fabsf(gyd::dot(wi, it.n)) / pdf;
                    deposit = !gyd::bsdf_delta(b) && i > 0;
                    const float pr = gyd::rmin(1.0f, gyd::lumY(est * flux) / gyd::lumY(flux));
                    if (pr < 1.f) {
                        if (gyd::seq_next(M, q) < pr) flux_next = flux / pr;
                        else alive = false;
                    }
                    if (alive) {
                        flux_next = flux_next * est;
                        gyd::spawn_ray(it, wi, &o_next);
                    }
                }
            }
        }
        // the vertex deposits with the flux that arrived at it
        const uint32_t slot = wave_append(V.count, r, 1u, deposit);
        if (deposit && slot < V.cap) {
            float4* w = V.v + ((size_t)r * V.cap + slot) * 4;
            w[0] = f4(o, h.t);
            w[1] = make_float4(d.x, d.y, d.z, __int_as_float(h.prim));
            w[2] = make_float4(h.b1, h.b2, __uint_as_float((uint32_t)j), __uint_as_float((uint32_t)i));
            w[3] = f4(flux, 0.f);
        }
        o = o_next;
        d = wi;
        flux = flux_next;
    }
}

// The deposits of one iteration, one thread per photon vertex: each hit point
// listed in the vertex's bucket, within its radius and facing the same way,
// gets a record (SPPM.cpp:127-150 before the update).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) k_sppm_deposit(gyd::SceneDev S, int max_depth, SppmDev D, GridDev G, const float4* vtx,
                                                      uint32_t n, RecDev Rec) {
    const uint32_t t = blockIdx.x * 256 + threadIdx.x;
    const bool valid = t < n;
    gyd::Isect it = {};
    gyd::BSDF b = {};
    V3 d = v3s(0.f), flux = v3s(0.f);
    uint32_t e0 = 0, e1 = 0, key_lo = 0;
    if (valid) {
        const float4* w = vtx + (size_t)t * 4;
        const float4 a = w[0], c = w[1], bb = w[2], fl = w[3];
        gyd::Hit h;
        h.t = a.w;
        h.prim = __float_as_int(c.w);
        h.b1 = bb.x;
        h.b2 = bb.y;
        d = v3(c.x, c.y, c.z);
        it = gyd::query(S, f3(a), d, h);
        b = gyd::make_bsdf(S, it);
        b.importance = true;
        flux = f3(fl);
        key_lo = __float_as_uint(bb.z) * (uint32_t)max_depth + __float_as_uint(bb.w);
        const V3 hh = (it.pos - v3(G.bmin.x, G.bmin.y, G.bmin.z)) * G.inv;
        const uint32_t bk = grid_hash(abs(int(hh.x)), abs(int(hh.y)), abs(int(hh.z)), G.num);
        e0 = G.start[bk];
        e1 = G.start[bk + 1];
    }
    // Each entry carries its hit point's position, radius^2 and |normal|, so
    // a candidate costs one contiguous 32-byte read and only accepted ones
    // gather more.  Count, one atomic per wave, then write in bucket order.
    uint32_t cnt = 0;
    for (uint32_t e = e0; e < e1; ++e) {
        const float4 g0 = G.entries[2 * (size_t)e], g1 = G.entries[2 * (size_t)e + 1];
        const V3 vv = v3(g0.x, g0.y, g0.z) - it.pos;
        const float dd = gyd::dot(vv, vv);
        if ((double)gyd::dot(v3(g1.x, g1.y, g1.z), it.absN) > 0.0015 && dd <= g0.w) ++cnt;
    }
    uint32_t slot = wave_append(Rec.count, 0u, cnt, cnt > 0);
    for (uint32_t e = e0; e < e1 && cnt > 0; ++e) {
        const float4 g0 = G.entries[2 * (size_t)e], g1 = G.entries[2 * (size_t)e + 1];
        const V3 vv = v3(g0.x, g0.y, g0.z) - it.pos;
        const float dd = gyd::dot(vv, vv);
        if ((double)gyd::dot(v3(g1.x, g1.y, g1.z), it.absN) > 0.0015 && dd <= g0.w) {
            if (slot < Rec.cap) {
                const uint32_t hp = __float_as_uint(g1.w);
                const V3 c = f3(D.imp[hp]) * gyd::bsdf_eval(b, f3(D.out[hp]), mul_m1(d)) * flux;
                Rec.keys[slot] = ((unsigned long long)hp << 32) | (unsigned long long)key_lo;
                Rec.val[slot] = f4(c, dd);
            }
            ++slot;
        }
    }
}

// Each hit point's record range in the sorted order.
__global__ void k_sppm_segments(const unsigned long long* keys, unsigned int n, uint32_t* lo, uint32_t* hi) {
    const unsigned int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t hp = (uint32_t)(keys[i] >> 32);
    if (i == 0 || (uint32_t)(keys[i - 1] >> 32) != hp) lo[hp] = i;
    if (i + 1 == n || (uint32_t)(keys[i + 1] >> 32) != hp) hi[hp] = i + 1;
}
// The hit point's updates in (photon, vertex) order (SPPM.cpp:136-150), then
// with batchShrink the per-iteration update (SPPM.cpp:245-257).
__global__ void k_sppm_apply(SppmDev D, const unsigned int* order, const float4* val, const uint32_t* lo, const uint32_t* hi,
                             float alpha, int batch) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (int64_t)D.W * D.H) return;
    float r2 = D.r2[p], N = D.N[p];
    V3 flux = f3(D.flux[p]);
    uint32_t m = D.m[p];
    for (uint32_t k = lo[p]; k < hi[p]; ++k) {
        const float4 r = val[order[k]];
        if (!(r.w <= r2)) continue;
        const V3 c = f3(r);
        if (!batch) {
            const float g = (N * alpha + alpha) / (N * alpha + 1.f);
            r2 = r2 * g;
            N += 1;
            flux = (flux + c) * g;
        } else {
            m++;
            flux = flux + c;
        }
    }
    if (batch && m > 0) {
        const float g = (N + alpha * (float)m) / (N + (float)m);
        r2 = r2 * g;
        flux = flux * g;
        N = N + alpha * (float)m;
        m = 0;
    }
    D.r2[p] = r2;
    D.N[p] = N;
    D.flux[p] = f4(flux, 0.f);
    D.m[p] = m;
}
// Density estimate (SPPM.cpp:269-273).
__global__ void k_sppm_final(SppmDev D, int iterations, int64_t photons, float* rgb) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (int64_t)D.W * D.H) return;
    const float k = 1.f / (3.14159265358979f * D.r2[p] * (float)iterations * (float)photons);
    const V3 c = (v3s(0.f) + f3(D.flux[p]) * k) + f3(D.direct[p]) / (float)iterations;
    rgb[3 * p] = c.x;
    rgb[3 * p + 1] = c.y;
    rgb[3 * p + 2] = c.z;
}
__global__ void k_sppm_init(SppmDev D, float r2) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= (int64_t)D.W * D.H) return;
    D.pos[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    D.r2[p] = r2;
    D.N[p] = 0.f;
    D.flux[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    D.direct[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    D.m[p] = 0u;
}
__global__ void k_u32_fill(uint32_t* p, uint32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_u32_iota(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = i;
}

// Device buffers freed together (by reset() or at the end of the render).
struct Bufs {
    std::vector<void*> all;
    ~Bufs() { reset(); }
    void reset() {
        for (void* b : all) DevFree(b);
        all.clear();
    }
    template <class T>
    int get(Ctx* c, T** p, size_t count) {
        void* v = nullptr;
        const int rc = DevAlloc(c, &v, std::max<size_t>(count, 1) * sizeof(T));
        if (rc == GY_OK) all.push_back(v);
        *p = (T*)v;
        return rc;
    }
};
unsigned G1(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

int MakePathParams(Ctx* c, Scene* sc, const gy_camera_desc& cam, const gy_render_params& p, PathParams* P);

}  // namespace gy

using namespace gy;

// SPPM(iterations, nPhotonsPerStage, maxDepth, initialRadius, alpha,
// batchShrink, sampler, samplerEnum) (integrator/SPPM.h:48-49) through
// SPPM::Render: rgb = W x H x 3, the image Film::SetImage receives.
extern "C" int gy_render_sppm(gy_scene* s, const gy_camera_desc* cam, const gy_sppm_params* sp, float* rgb, float* radius2,
                              float* photon_count) {
    if (!s || !cam || !sp || !rgb) return Fail(GY_ERR_INVALID, "gy_render_sppm: null argument");
    if (sp->iterations <= 0 || sp->photons_per_iteration <= 0 || sp->max_depth <= 0 || sp->max_depth > 100 ||
        !(sp->initial_radius > 0.f) || !(sp->alpha > 0.f && sp->alpha <= 1.f))
        return Fail(GY_ERR_INVALID, "gy_render_sppm: iterations, photons, 0 < max_depth <= 100, radius > 0, 0 < alpha <= 1");
    if (sp->trace_glossy_ray) return Fail(GY_ERR_UNSUPPORTED, "gy_render_sppm: traceGlossyRay");
    if (sp->photons_per_launch < 0) return Fail(GY_ERR_INVALID, "gy_render_sppm: photons_per_launch >= 0");
    if ((uint64_t)sp->photons_per_iteration * (uint64_t)sp->max_depth >= (1ull << 32))
        return Fail(GY_ERR_UNSUPPORTED, "gy_render_sppm: photons_per_iteration * max_depth must stay below 2^32");
    Scene* sc = reinterpret_cast<Scene*>(s);
    if (sc->dev.n_sph > 0) return Fail(GY_ERR_UNSUPPORTED, "gy_render_sppm: scenes with spheres");
    if (sc->dev.env.w > 0) return Fail(GY_ERR_UNSUPPORTED, "gy_render_sppm: environment lights (photons from area lights only)");
    if (sc->area_tri.empty()) return Fail(GY_ERR_INVALID, "gy_render_sppm: the scene has no area light");
    const int64_t WH = (int64_t)cam->width * cam->height;
    if (WH <= 0 || WH >= (1ll << 31)) return Fail(GY_ERR_INVALID, "gy_render_sppm: image size");
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    gy_render_params rp = {};
    rp.spp = sp->iterations;  // eye pass sample index = iteration (SamplerEnum::GetIndex(iter, x, y))
    rp.max_depth = sp->max_depth;
    rp.rr_depth = sp->max_depth;
    rp.sampler = sp->sampler;
    PathParams P;
    int rc = MakePathParams(c, sc, *cam, rp, &P);
    if (rc != GY_OK) return rc;
    if (sp->sampler == GY_SAMPLER_SOBOL) {
        // photon paths draw 5 + 4 per bounce dimensions; photon indices reach iterations * photons
        if (5 + 4 * (int64_t)sp->max_depth > sobol::kDims) return Fail(GY_ERR_UNSUPPORTED, "gy_render_sppm: max_depth");
        const uint64_t last = (uint64_t)sp->iterations * (uint64_t)sp->photons_per_iteration;
        if (last >> sobol::kCols) return Fail(GY_ERR_UNSUPPORTED, "gy_render_sppm: photon index beyond 52 bits");
    }
    Bufs B;
    SppmDev D;
    D.W = cam->width;
    D.H = cam->height;
    rc = B.get(c, &D.pos, WH);
    if (rc == GY_OK) rc = B.get(c, &D.nrm, WH);
    if (rc == GY_OK) rc = B.get(c, &D.out, WH);
    if (rc == GY_OK) rc = B.get(c, &D.imp, WH);
    if (rc == GY_OK) rc = B.get(c, &D.r2, WH);
    if (rc == GY_OK) rc = B.get(c, &D.N, WH);
    if (rc == GY_OK) rc = B.get(c, &D.flux, WH);
    if (rc == GY_OK) rc = B.get(c, &D.direct, WH);
    if (rc == GY_OK) rc = B.get(c, &D.m, WH);
    uint32_t *stats = nullptr, *lo = nullptr, *hi = nullptr;
    float* d_rgb = nullptr;
    if (rc == GY_OK) rc = B.get(c, &stats, 8);
    if (rc == GY_OK) rc = B.get(c, &lo, WH);
    if (rc == GY_OK) rc = B.get(c, &hi, WH);
    if (rc == GY_OK) rc = B.get(c, &d_rgb, WH * 3);
    if (rc != GY_OK) return rc;
    const float r0 = sp->initial_radius;
    hipLaunchKernelGGL(k_sppm_init, dim3(G1(WH, 256)), dim3(256), 0, c->stream, D, r0 * r0);
    // grid, record and scratch buffers, grown as needed
    Bufs GB, EB, RB, TB;
    uint32_t *bcount = nullptr, *bstart = nullptr, *bcursor = nullptr;
    float4* entries = nullptr;
    size_t bucket_cap = 0, entry_cap = 0;
    unsigned long long *keys = nullptr, *keys2 = nullptr;
    float4* val = nullptr;
    unsigned int *order = nullptr, *order2 = nullptr, *rcount = nullptr;
    unsigned int rec_cap = 0;
    void* tmp = nullptr;
    size_t tmp_cap = 0;
    auto grow_tmp = [&](size_t need) -> int {
        if (need <= tmp_cap) return GY_OK;
        TB.reset();
        tmp_cap = need + need / 4;
        unsigned char* t = nullptr;
        const int r = TB.get(c, &t, tmp_cap);
        tmp = t;
        return r;
    };
    auto grow_records = [&](size_t want) -> int {
        RB.reset();
        rec_cap = (unsigned)want;
        int r = RB.get(c, &keys, rec_cap);
        if (r == GY_OK) r = RB.get(c, &keys2, rec_cap);
        if (r == GY_OK) r = RB.get(c, &val, rec_cap);
        if (r == GY_OK) r = RB.get(c, &order, rec_cap);
        if (r == GY_OK) r = RB.get(c, &order2, rec_cap);
        if (r == GY_OK) r = RB.get(c, &rcount, 1);
        return r;
    };
    int end_bit = 32;
    while ((1ll << (end_bit - 32)) < WH) ++end_bit;
    const bool stack64 = sc->stack > 32;
    // Photon vertices: one list per iteration of a chunk of iterations traced
    // together (photons_per_launch photon paths per launch, about 4M by
    // default), grown and traced again when a list overflows; the chunk
    // shrinks to keep the lists within 16 GiB.
    const int64_t photons = sp->photons_per_iteration;
    constexpr uint64_t kVtxBudget = 16ull << 30;
    const int64_t per_launch = sp->photons_per_launch > 0 ? sp->photons_per_launch : (1ll << 22);
    int chunk = (int)std::min<int64_t>(std::max<int64_t>((per_launch + photons - 1) / photons, 1), sp->iterations);
    {
        const int n_chunks = (sp->iterations + chunk - 1) / chunk;  // chunks of even size
        chunk = (sp->iterations + n_chunks - 1) / n_chunks;
    }
    Bufs VB;
    float4* vtx = nullptr;
    unsigned int* vcount = nullptr;
    uint32_t vcap = 0;
    std::vector<unsigned int> vn;
    auto grow_vtx = [&](uint64_t per_iter) -> int {
        VB.reset();
        vcap = 0;
        if (per_iter >= 0xFFFFFFFFull) return Fail(GY_ERR_UNSUPPORTED, "gy_render_sppm: 2^32 depositing photon vertices in one iteration");
        chunk = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)chunk, kVtxBudget / (per_iter * 64)));
        int r = VB.get(c, &vtx, (size_t)chunk * per_iter * 4);
        if (r == GY_OK) r = VB.get(c, &vcount, (size_t)chunk);
        if (r == GY_OK) vcap = (uint32_t)per_iter;
        return r;
    };
    int iter0 = 0, n_iter = 0;
    for (int iter = 0; iter < sp->iterations && rc == GY_OK; ++iter) {
        if (iter == iter0 + n_iter) {
            // 0. the photon paths of the next chunk of iterations
            iter0 = iter;
            for (;;) {
                if (vcap == 0) {
                    rc = grow_vtx((uint64_t)std::min<int64_t>(photons * (sp->max_depth - 1), std::max<int64_t>(photons * 4, 1 << 16)) + 1);
                    if (rc != GY_OK) break;
                }
                n_iter = std::min(chunk, sp->iterations - iter0);
                hipLaunchKernelGGL(k_u32_fill, dim3(1), dim3(256), 0, c->stream, vcount, 0u, (int64_t)n_iter);
                if (stack64)
                    hipLaunchKernelGGL(k_sppm_trace<64>, dim3(G1(photons * n_iter, 128)), dim3(128), 0, c->stream, sc->dev, P,
                                       VtxDev{vtx, vcount, vcap}, iter0, n_iter, photons);
                else
                    hipLaunchKernelGGL(k_sppm_trace<32>, dim3(G1(photons * n_iter, 128)), dim3(128), 0, c->stream, sc->dev, P,
                                       VtxDev{vtx, vcount, vcap}, iter0, n_iter, photons);
                rc = CheckLaunch(c, "k_sppm_trace");
                if (rc != GY_OK) break;
                vn.assign((size_t)n_iter, 0u);
                hipError_t e = hipMemcpyAsync(vn.data(), vcount, (size_t)n_iter * 4, hipMemcpyDeviceToHost, c->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
                if (e != hipSuccess) {
                    rc = HipFail(e, "sppm vertices");
                    break;
                }
                const uint32_t most = *std::max_element(vn.begin(), vn.end());
                if (most <= vcap) break;
                rc = grow_vtx((uint64_t)most + most / 4);
                if (rc != GY_OK) break;
            }
            if (rc != GY_OK) break;
        }
        // 1. eye pass
        if (stack64) hipLaunchKernelGGL(k_sppm_eye<64>, dim3(G1(WH, 128)), dim3(128), 0, c->stream, sc->dev, P, D, iter);
        else hipLaunchKernelGGL(k_sppm_eye<32>, dim3(G1(WH, 128)), dim3(128), 0, c->stream, sc->dev, P, D, iter);
        rc = CheckLaunch(c, "k_sppm_eye");
        if (rc != GY_OK) break;
        // 2. hash grid
        hipLaunchKernelGGL(k_u32_fill, dim3(1), dim3(256), 0, c->stream, stats, 0u, (int64_t)8);
        hipLaunchKernelGGL(k_sppm_stats, dim3(kReduceBlocks), dim3(256), 0, c->stream, D, stats);
        uint32_t hs[8];
        hipError_t e = hipMemcpyAsync(hs, stats, 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            rc = HipFail(e, "sppm stats");
            break;
        }
        const int64_t num = hs[0];
        GridDev Gd = {};
        Gd.num = num;
        if (num > 0) {
            float maxr2;
            std::memcpy(&maxr2, &hs[1], 4);
            const float search = std::sqrt(maxr2) + 1e-6f;  // Eps (common/Core.h:191)
            hipLaunchKernelGGL(k_u32_fill, dim3(1), dim3(256), 0, c->stream, stats, 0xFFFFFFFFu, (int64_t)3);
            hipLaunchKernelGGL(k_sppm_bbox, dim3(kReduceBlocks), dim3(256), 0, c->stream, D, search, stats);
            e = hipMemcpyAsync(hs, stats, 12, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) {
                rc = HipFail(e, "sppm bbox");
                break;
            }
            auto unord = [](uint32_t u) {
                const uint32_t b = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u;
                float f;
                std::memcpy(&f, &b, 4);
                return f;
            };
            Gd.bmin = make_float3(unord(hs[0]), unord(hs[1]), unord(hs[2]));
            Gd.inv = 1.f / (search * 2.f);
            if ((size_t)num + 1 > bucket_cap) {
                GB.reset();
                bucket_cap = (size_t)num + 1;
                rc = GB.get(c, &bcount, bucket_cap);
                if (rc == GY_OK) rc = GB.get(c, &bstart, bucket_cap);
                if (rc == GY_OK) rc = GB.get(c, &bcursor, bucket_cap);
                if (rc != GY_OK) break;
            }
            hipLaunchKernelGGL(k_u32_fill, dim3(G1(num + 1, 256)), dim3(256), 0, c->stream, bcount, 0u, num + 1);
            hipLaunchKernelGGL(k_sppm_grid<false>, dim3(G1(WH, 256)), dim3(256), 0, c->stream, D, Gd, bcount);
            size_t need = 0;
            e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, bcount, bstart, (int)(num + 1), c->stream);
            if (e == hipSuccess) rc = grow_tmp(need);
            if (rc != GY_OK) break;
            if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, need, bcount, bstart, (int)(num + 1), c->stream);
            uint32_t total = 0;
            if (e == hipSuccess) e = hipMemcpyAsync(&total, bstart + num, 4, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) {
                rc = HipFail(e, "sppm grid total");
                break;
            }
            if (total > entry_cap) {
                EB.reset();
                entry_cap = (size_t)total + total / 4 + 16;
                rc = EB.get(c, &entries, entry_cap * 2);
                if (rc != GY_OK) break;
            }
            e = hipMemcpyAsync(bcursor, bstart, (size_t)(num + 1) * 4, hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) {
                rc = HipFail(e, "sppm cursor");
                break;
            }
            Gd.start = bstart;
            Gd.cursor = bcursor;
            Gd.entries = entries;
            hipLaunchKernelGGL(k_sppm_grid<true>, dim3(G1(WH, 256)), dim3(256), 0, c->stream, D, Gd, bcount);
            rc = CheckLaunch(c, "k_sppm_grid");
            if (rc != GY_OK) break;
        }
        // 3. deposits of the iteration's photon vertices: records, grown and
        // deposited again when they overflow
        const uint32_t nv = std::min<uint32_t>(vn[iter - iter0], vcap);
        const float4* vlist = vtx + (size_t)(iter - iter0) * vcap * 4;
        unsigned int nrec = 0;
        for (;;) {
            if (rec_cap == 0) {
                // a first guess of 16 records per photon (40 bytes each); grown on overflow
                rc = grow_records((size_t)std::min<int64_t>(std::max<int64_t>(sp->photons_per_iteration * 16, 1 << 20), 0x7FFFFFFF));
                if (rc != GY_OK) break;
            }
            if (num == 0 || nv == 0) break;
            hipLaunchKernelGGL(k_u32_fill, dim3(1), dim3(256), 0, c->stream, rcount, 0u, (int64_t)1);
            RecDev R = {keys, val, rcount, rec_cap};
            hipLaunchKernelGGL(k_sppm_deposit, dim3(G1(nv, 256)), dim3(256), 0, c->stream, sc->dev, sp->max_depth, D, Gd, vlist, nv, R);
            rc = CheckLaunch(c, "k_sppm_deposit");
            if (rc != GY_OK) break;
            e = hipMemcpyAsync(&nrec, rcount, 4, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) {
                rc = HipFail(e, "sppm records");
                break;
            }
            if (nrec <= rec_cap) break;
            if ((uint64_t)nrec + nrec / 4 > 0x7FFFFFFFull) {
                rc = Fail(GY_ERR_UNSUPPORTED, "gy_render_sppm: more than 2^31 photon records in one iteration");
                break;
            }
            // too many records: room for all of them, and the iteration's vertices deposited again
            rc = grow_records((size_t)nrec + nrec / 4);
            if (rc != GY_OK) break;
        }
        if (rc != GY_OK) break;
        // 4. sort by (hit point, photon, vertex) and apply per hit point
        hipLaunchKernelGGL(k_u32_fill, dim3(G1(WH, 256)), dim3(256), 0, c->stream, lo, 0u, WH);
        hipLaunchKernelGGL(k_u32_fill, dim3(G1(WH, 256)), dim3(256), 0, c->stream, hi, 0u, WH);
        const unsigned int* sorted_order = order;
        if (nrec > 0) {
            hipLaunchKernelGGL(k_u32_iota, dim3(G1(nrec, 256)), dim3(256), 0, c->stream, order, nrec);
            size_t need = 0;
            hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, keys, keys2, order, order2, (int)nrec, 0, end_bit, c->stream);
            if (e == hipSuccess) rc = grow_tmp(need);
            if (rc != GY_OK) break;
            if (e == hipSuccess) e = hipcub::DeviceRadixSort::SortPairs(tmp, need, keys, keys2, order, order2, (int)nrec, 0, end_bit, c->stream);
            if (e != hipSuccess) {
                rc = HipFail(e, "sppm sort");
                break;
            }
            hipLaunchKernelGGL(k_sppm_segments, dim3(G1(nrec, 256)), dim3(256), 0, c->stream, keys2, nrec, lo, hi);
            sorted_order = order2;
        }
        hipLaunchKernelGGL(k_sppm_apply, dim3(G1(WH, 256)), dim3(256), 0, c->stream, D, sorted_order, val, lo, hi, sp->alpha,
                           sp->batch_shrink ? 1 : 0);
        rc = CheckLaunch(c, "k_sppm_apply");
    }
    if (rc != GY_OK) return rc;
    hipLaunchKernelGGL(k_sppm_final, dim3(G1(WH, 256)), dim3(256), 0, c->stream, D, sp->iterations,
                       (int64_t)sp->photons_per_iteration, d_rgb);
    rc = CheckLaunch(c, "k_sppm_final");
    if (rc != GY_OK) return rc;
    hipError_t e = hipMemcpyAsync(rgb, d_rgb, (size_t)WH * 12, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && radius2) e = hipMemcpyAsync(radius2, D.r2, (size_t)WH * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && photon_count) e = hipMemcpyAsync(photon_count, D.N, (size_t)WH * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return HipFail(e, "sppm D2H");
    return GY_OK;
}

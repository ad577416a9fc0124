// Per-path integrator on the device: PathTracing::Li with DirectIllumination
// for one camera sample (megakernel and path replay), and the film footprint.
#pragma once
#include "device.h"
#include "runtime.h"

namespace gy {
using namespace gyd;

// ------------------------------------------------------------------ path tracing (megakernel)
// The whole of TiledIntegrator's per-sample work (integrator/TiledIntegrator.cpp:40-47)
// and PathTracing::Li (integrator/PT.cpp:10-67) for one camera sample per thread.
struct PathCounters {
    uint32_t rays, shadow, mis, shade;
    uint32_t nodes[3], tris[3];  // by ray kind: camera/continuation, shadow, MIS
};

// Integrator::DirectIllumination (integrator/Integrator.cpp:9-107) in parts:
// light-sample term L1 with weight w1 towards wl, BSDF-sample (MIS) term L2
// with weight w2 towards wb, light-selection probability lsp.  The estimate is
// (L1 w1 + L2 w2) / lsp.
// ENV: the scene may hold an environment light (light triangle -1); without it
// the env branches compile away and the kernels keep their registers.
struct DIParts {
    V3 L1, L2, wl, wb;
    float w1, w2, lsp;
};
template <int STACK, bool COUNT, bool ENV = false>
__device__ DIParts di_parts(const SceneDev& S, const Isect& it, const BSDF& b, float uL, float u0, float u1, float v0,
                            float v1, float v2, uint32_t* stack, int stride, PathCounters& pc) {
    DIParts r;
    const int light = sample_one_light(S, uL, &r.lsp);
    const int ltri = S.light_tri[light];  // < 0: the environment light
    const V3 Le = v3(S.light_le[3 * light], S.light_le[3 * light + 1], S.light_le[3 * light + 2]);
    r.L1 = v3s(0.f);
    r.L2 = v3s(0.f);
    r.wl = r.wb = v3s(0.f);
    r.w1 = 1.f;
    r.w2 = 1.f;
    {
        float pdf;
        LightPoint lp;
        V3 wi, Li;
        if (ENV && ltri < 0) {
            lp = env_light_point_call(S.env, it.pos, u0, u1, &wi, &Li, &pdf);
        } else {
            lp = sample_light_point(S, ltri, it.pos, u0, u1, &pdf);
            wi = norm(lp.pos - it.pos);
            Li = dot(lp.n, neg(wi)) > 0.f ? Le : v3(0.f, 0.f, 0.f);
        }
        r.wl = wi;
        if (pdf > 0.f && !is_zero(Li)) {
            const float sp = bsdf_pdf(b, it.outDir, wi);
            const V3 f = bsdf_eval(b, it.outDir, wi);
            const float c = fabsf(dot(it.n, wi));
            if (!is_zero(f)) {
                V3 so, sd;
                float stmax;
                spawn_to(it, lp, &so, &sd, &stmax);
                Hit hs;
                uint32_t* nn = COUNT ? &pc.nodes[1] : nullptr;
                uint32_t* nt = COUNT ? &pc.tris[1] : nullptr;
                pc.shadow++;
                if (traverse<true, STACK>(S, so, sd, 0.f, stmax, &hs, stack, stride, nn, nt)) Li = v3s(0.f);
                r.w1 = power_heuristic(pdf, sp);
                r.L1 = Li * v3s(1.f) * f * c / pdf;
            }
        }
    }
    {
        V3 wi;
        float pdf;
        const V3 f = bsdf_sample(b, it.outDir, &wi, &pdf, v0, v1, v2);
        r.wb = wi;
        const float c = fabsf(dot(it.n, wi));
        if (pdf != 0.f && !is_zero(f)) {
            const float lpdf = (ENV && ltri < 0) ? env_pdf_dir_call(S.env, wi) : light_pdf(S, ltri, it, wi);
            if (lpdf == 0.f) {
                r.w2 = 0.f;
            } else {
                r.w2 = power_heuristic(pdf, lpdf);
                V3 ro;
                spawn_ray(it, wi, &ro);
                Hit h2;
                uint32_t* nn = COUNT ? &pc.nodes[2] : nullptr;
                uint32_t* nt = COUNT ? &pc.tris[2] : nullptr;
                pc.mis++;
                if (traverse<false, STACK>(S, ro, wi, 0.f, kInf, &h2, stack, stride, nn, nt)) {
                    const int hl = __float_as_int(S.shade[(size_t)h2.prim * 8 + 1].w);
                    if (hl == light) {
                        const Isect li = query(S, ro, wi, h2);
                        const V3 Lem = dot(li.n, neg(wi)) > 0.f ? Le : v3(0.f, 0.f, 0.f);
                        r.L2 = Lem * f * v3s(1.f) * c / pdf;
                    }
                } else if (ENV && ltri < 0) {  // integrator/Integrator.cpp:97-99
                    r.L2 = env_emission_call(S.env, wi) * f * v3s(1.f) * c / pdf;
                }
            }
        }
    }
    return r;
}

template <int STACK, bool COUNT, bool ENV = false>
__device__ V3 direct_illumination(const SceneDev& S, const Isect& it, const BSDF& b, float uL, float u0, float u1,
                                  float v0, float v1, float v2, uint32_t* stack, int stride, PathCounters& pc) {
    const DIParts d = di_parts<STACK, COUNT, ENV>(S, it, b, uL, u0, u1, v0, v1, v2, stack, stride, pc);
    return (d.L1 * d.w1 + d.L2 * d.w2) / d.lsp;
}

template <int STACK, bool COUNT, bool ENV = false>
__device__ V3 trace_path(const SceneDev& S, const PathParams& P, int px, int py, int s, uint32_t* stack, int stride,
                         float* fx, float* fy, uint32_t* dims, PathCounters& pc) {
    const SampleSrc& M = P.smp;
    Seq q;
    q.index = enum_index(P.en, (uint32_t)s, (uint32_t)px, (uint32_t)py);
    q.cursor = 0;
    // SampleX / SampleY (sampler/SobolEnum.h:30-36, sampler/HaltonEnum.h:30-46)
    const float u = P.en.sx * seq_next(M, q) - (float)px;
    const float v = P.en.sy * seq_next(M, q) - (float)py;
    V3 o, d;
    generate_ray(P.cam, px, py, u, v, &o, &d);
    bool delta = false;
    V3 T = v3(1.f, 1.f, 1.f);
    V3 L = v3s(0.f);
    for (int i = 0; i < P.max_depth; ++i) {
        Hit h;
        pc.rays++;
        uint32_t* nn = COUNT ? &pc.nodes[0] : nullptr;
        uint32_t* nt = COUNT ? &pc.tris[0] : nullptr;
        if (!traverse<false, STACK>(S, o, d, 0.f, kInf, &h, stack, stride, nn, nt)) {
            // integrator/PT.cpp:18-21: T * Le(ray) in place of L
            if (ENV && (i == 0 || delta) && S.env.w > 0) L = T * env_emission_call(S.env, d);
            break;
        }
        const Isect it = query(S, o, d, h);
        const BSDF b = make_bsdf(S, it);
        pc.shade++;
        if ((i == 0 || delta) && it.light >= 0) {
            const V3 Le = v3(S.light_le[3 * it.light], S.light_le[3 * it.light + 1], S.light_le[3 * it.light + 2]);
            const V3 e = dot(it.n, it.outDir) > 0.f ? Le : v3(0.f, 0.f, 0.f);
            L = L + T * e;
        } else {
            const float uL = seq_next(M, q);
            const float u0 = seq_next(M, q), u1 = seq_next(M, q);
            const float v0 = seq_next(M, q), v1 = seq_next(M, q), v2 = seq_next(M, q);
            L = L + T * direct_illumination<STACK, COUNT, ENV>(S, it, b, uL, u0, u1, v0, v1, v2, stack, stride, pc);
        }
        const float r0 = seq_next(M, q), r1 = seq_next(M, q), r2 = seq_next(M, q);
        V3 wi;
        float pdf;
        const V3 f = bsdf_sample(b, neg(d), &wi, &pdf, r0, r1, r2);
        if (is_zero(f) || pdf == 0.f) break;
        const V3 est = f * fabsf(dot(it.n, wi)) / pdf;
        delta = bsdf_delta(b);
        const float p = rmin(1.0f, lumY(est * T) / lumY(T));
        if (p < 1.f && i > P.rr_depth) {
            if (seq_next(M, q) < p) T = T / p;
            else break;
        }
        T = T * est;
        spawn_ray(it, wi, &o);
        d = wi;
    }
    *fx = px + u;
    *fy = py + v;
    if (dims) *dims = q.cursor;
    return L;
}

// Film::AddSample footprint with the radius-0.5 box filter (visual/Film.h:61-74).
__device__ inline void film_footprint(float fx, float fy, int W, int H, int* x0, int* x1, int* y0, int* y1) {
    const float x = fx - 0.5f, y = fy - 0.5f;
    int minX = (int)ceilf(x - 0.5f), maxX = (int)floorf(x + 0.5f);
    int minY = (int)ceilf(y - 0.5f), maxY = (int)floorf(y + 0.5f);
    *x0 = minX < 0 ? 0 : minX;
    *x1 = maxX > W - 1 ? W - 1 : maxX;
    *y0 = minY < 0 ? 0 : minY;
    *y1 = maxY > H - 1 ? H - 1 : maxY;
}

// Record one finished sample: own-pixel samples go to the pass buffer slot
// (owner adds them in film_accumulate); any sample whose box-filter footprint is
// not exactly its own pixel goes to the side list, one entry per covered pixel.
__device__ inline void record_sample(const FilmDev& F, int s_local, int px, int py, float fx, float fy, V3 L) {
    const int W = F.width, H = F.height;
    int x0, x1, y0, y1;
    film_footprint(fx, fy, W, H, &x0, &x1, &y0, &y1);
    const int64_t WH = (int64_t)W * H;
    const int64_t slot = (int64_t)s_local * WH + (int64_t)py * W + px;
    const bool own = (x0 == px && x1 == px && y0 == py && y1 == py);
    F.pass[slot] = make_float4(L.x, L.y, L.z, own ? 1.f : 0.f);
    if (own) return;
    for (int yy = y0; yy <= y1; ++yy)
        for (int xx = x0; xx <= x1; ++xx) {
            const uint32_t k = atomicAdd(F.side_count, 1u);
            if (k < F.side_cap) {
                SideEntry e;
                e.key = ((uint64_t)((uint32_t)(yy * W + xx)) << 32) | (uint64_t)(uint32_t)slot;
                e.L = make_float4(L.x, L.y, L.z, 0.f);
                F.side[k] = e;
            }
        }
}

__device__ inline void pass_pixel(const FilmDev& F, int64_t t, int* s_local, int* px, int* py) {
    // 8x8 pixel tiles, one tile per wavefront, tiles in row-major order.
    const int tilesX = (F.width + 7) >> 3, tilesY = (F.height + 7) >> 3;
    const int64_t per_s = (int64_t)tilesX * tilesY * 64;
    *s_local = (int)(t / per_s);
    const int64_t r = t - (int64_t)(*s_local) * per_s;
    const int64_t tile = r >> 6;
    const int local = (int)(r & 63);
    *px = (int)(tile % tilesX) * 8 + (local & 7);
    *py = (int)(tile / tilesX) * 8 + (local >> 3);
}


}  // namespace gy

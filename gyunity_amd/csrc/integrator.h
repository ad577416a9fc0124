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
    int light;   // the light chosen (SampleOneLight)
    V3 F1, F2;   // L1 and L2 per unit emitted radiance (the adjoint's emission terms)
};
template <int STACK, bool COUNT, bool ENV = false, bool SPH = false>
__device__ DIParts di_parts(const SceneDev& S, const Isect& it, const BSDF& b, float uL, float u0, float u1, float v0,
                            float v1, float v2, uint32_t* stack, int stride, PathCounters& pc) {
    DIParts r;
    const int light = sample_one_light(S, uL, &r.lsp);
    r.light = light;
    r.F1 = r.F2 = v3s(0.f);
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
            lp = sample_light_point<SPH>(S, ltri, it.pos, u0, u1, &pdf);
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
                const bool occluded = find_hit<true, STACK, COUNT, SPH>(S, so, sd, 0.f, stmax, &hs, stack, stride, nn, nt);
                if (occluded) Li = v3s(0.f);
                r.w1 = power_heuristic(pdf, sp);
                r.L1 = Li * v3s(1.f) * f * c / pdf;
                if (!occluded && !(ENV && ltri < 0)) r.F1 = v3s(1.f) * f * c / pdf;
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
            const float lpdf = (ENV && ltri < 0) ? env_pdf_dir_call(S.env, wi) : light_pdf<SPH>(S, ltri, it, wi);
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
                if (find_hit<false, STACK, COUNT, SPH>(S, ro, wi, 0.f, kInf, &h2, stack, stride, nn, nt)) {
                    const int hl = prim_light<SPH>(S, h2.prim);
                    if (hl == light) {
                        const Isect li = query<SPH>(S, ro, wi, h2);
                        const bool facing = dot(li.n, neg(wi)) > 0.f;
                        const V3 Lem = facing ? Le : v3(0.f, 0.f, 0.f);
                        r.L2 = Lem * f * v3s(1.f) * c / pdf;
                        if (facing) r.F2 = f * v3s(1.f) * c / pdf;
                    }
                } else if (ENV && ltri < 0) {  // integrator/Integrator.cpp:97-99
                    r.L2 = env_emission_call(S.env, wi) * f * v3s(1.f) * c / pdf;
                }
            }
        }
    }
    return r;
}

template <int STACK, bool COUNT, bool ENV = false, bool SPH = false>
__device__ V3 direct_illumination(const SceneDev& S, const Isect& it, const BSDF& b, float uL, float u0, float u1,
                                  float v0, float v1, float v2, uint32_t* stack, int stride, PathCounters& pc) {
    const DIParts d = di_parts<STACK, COUNT, ENV, SPH>(S, it, b, uL, u0, u1, v0, v1, v2, stack, stride, pc);
    return (d.L1 * d.w1 + d.L2 * d.w2) / d.lsp;
}

template <int STACK, bool COUNT, bool ENV = false, bool SPH = false>
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
        if (!find_hit<false, STACK, COUNT, SPH>(S, o, d, 0.f, kInf, &h, stack, stride, nn, nt)) {
            // integrator/PT.cpp:18-21: T * Le(ray) in place of L
            if (ENV && (i == 0 || delta) && S.env.w > 0) L = T * env_emission_call(S.env, d);
            break;
        }
        const Isect it = query<SPH>(S, o, d, h);
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
            L = L + T * direct_illumination<STACK, COUNT, ENV, SPH>(S, it, b, uL, u0, u1, v0, v1, v2, stack, stride, pc);
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

// ------------------------------------------------------------------ light tracing
// LightTracing (integrator/LT.h:17-241) for one light path: the sample index
// W H p + y W + x drives the sequence from dimension 0.  Each vertex is
// connected to the camera and splatted as soon as it exists: the reference
// connects after the whole path is built, but the only random numbers it draws
// there (PinHoleCamera::Sample_Wi's) are unused and drawn last, so the order
// changes nothing.  Area lights only; light-path BSDFs use the importance
// transport mode.
struct LightCamDev {
    float m[16];  // WorldToRaster, column-major
    float pos[3], cz[3];
    float dis, area;
    int32_t width, height;
    float env_c[3];  // the scene's bounding-sphere center (environment light paths)
};
// Transform::operator()(point) (math/Transform.h:106-115) with the column
// order of Matrix4 * Vector4 (math/Linagl.h:937-948).
__device__ inline V3 xf_point(const float* m, V3 p) {
    const float in[4] = {p.x, p.y, p.z, 1.0f};
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = m[12 + i] * in[3];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = m[8 + i] * in[2] + r[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = m[4 + i] * in[1] + r[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = m[0 + i] * in[0] + r[i];
    V3 q = v3(r[0], r[1], r[2]);
    if (!(r[3] == 1.f || r[3] == 0.f)) q = q / r[3];
    return q;
}
// ConnectToCamera (integrator/LT.h:135-183) + WorldToScreen (:116-129) +
// Film::AddSplat (visual/Film.h:89-103) of one vertex.  s0: the light vertex,
// whose splat is its throughput (Le) as the reference returns it.
template <int STACK>
__device__ void lt_connect(const SceneDev& S, const LightCamDev& C, float* splat, V3 pos, V3 n, V3 outDir, const BSDF* b,
                           V3 thr, bool s0, float total, uint32_t* stack, int stride) {
    const V3 cpos = v3(C.pos[0], C.pos[1], C.pos[2]), cz = v3(C.cz[0], C.cz[1], C.cz[2]);
    V3 toCam = cpos - pos;
    const float len = length(toCam);
    toCam = norm(toCam);
    if (dot(cz, toCam * -1.f) < 0.f) return;
    const V3 rp = xf_point(C.m, pos);
    const int ix = (int)rp.x, iy = (int)rp.y;
    if (!(ix >= 0 && ix <= C.width && iy >= 0 && iy <= C.height)) return;
    Hit hv;
    if (find_hit<true, STACK>(S, pos + toCam * kRayEps, toCam, 0.f, len, &hv, stack, stride)) return;
    V3 L;
    if (s0) {
        L = thr;
    } else {
        // PinHoleCamera::Sample_Wi / We (camera/PinholeCamera.h:31-49)
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
    L = ((L * (float)C.width) * (float)C.height) / total;  // L * resX * resY / totalPhotons (integrator/LT.h:40)
    int X = (int)floorf((float)ix), Y = (int)floorf((float)iy);
    X = X < 0 ? 0 : (X > C.width - 1 ? C.width - 1 : X);
    Y = Y < 0 ? 0 : (Y > C.height - 1 ? C.height - 1 : Y);
    float* px = splat + ((size_t)Y * C.width + X) * 3;
    atomicAdd(px, L.x);
    atomicAdd(px + 1, L.y);
    atomicAdd(px + 2, L.z);
}
// EnvironmentLight::SampleLight (light/EnvLight.cpp:22-39) for a light path: a
// direction from the map and a point on the bounding sphere's disk facing it.
// Its radiance is dropped: GenerateLightPath takes Le = Emission(), which the
// environment light leaves zero (light/EnvLight.h:20-22, integrator/LT.h:108-111),
// so these paths carry zero throughput (NaN where a pdf is zero).  ConcentricSampleDisk
// is the x, y of cosine_hemisphere.
__device__ __noinline__ void lt_env_origin(const EnvDev& E, const LightCamDev& C, float u0, float u1, float v0, float v1,
                                           V3* pos, V3* dir, float* pdfA, float* pdfDir) {
    V3 sd;
    (void)env_sample(E, u0, u1, &sd, pdfDir);
    const V3 d = neg(sd);
    V3 a1, a2;
    coordinate_system(neg(d), &a1, &a2);
    const V3 cd = cosine_hemisphere(v0, v1);
    const V3 ctr = v3(C.env_c[0], C.env_c[1], C.env_c[2]);
    const V3 pDisk = ctr + (a1 * cd.x + a2 * cd.y) * E.radius;
    *pos = pDisk + neg(d) * E.radius;
    *dir = d;
    *pdfA = 1.f / ((kPI * E.radius) * E.radius);
}
// AreaLight::SampleOnLight (light/AreaLight.h:60-71), Triangle::Sample
// (shape/Triangle.cpp:141-158); returns Le.
__device__ inline V3 lt_area_origin(const SceneDev& S, int light, float u0, float u1, float w0, float w1, V3* pos, V3* n,
                                    V3* dir, float* pdfA, float* pdfDir) {
    const int ltri = S.light_tri[light];
    const float4* tq = S.shade + (size_t)ltri * 8;
    const V3 p0 = ld3(tq[0]), p1 = ld3(tq[1]), p2 = ld3(tq[2]);
    const V3 n0 = ld3(tq[3]), n1 = ld3(tq[4]), n2 = ld3(tq[5]);
    const float su0 = sqrtf(u0);
    const float b0 = 1 - su0, b1 = u1 * su0;
    *pos = (b0 * p0 + b1 * p1) + (1 - b0 - b1) * p2;
    const V3 ln = (b0 * n0 + b1 * n1) + (1 - b0 - b1) * n2;
    *n = ln;
    *pdfA = 1 / tq[2].w;
    V3 ss, ts;
    coordinate_system(ln, &ss, &ts);
    const V3 dl = cosine_hemisphere(w0, w1);
    *dir = norm((ss * dl.x + ts * dl.y) + ln * dl.z);
    *pdfDir = dl.z * kInvPI;
    return v3(S.light_le[3 * light], S.light_le[3 * light + 1], S.light_le[3 * light + 2]);
}
// GenerateLightPath + Trace (integrator/LT.h:60-113) with the connections.
// ENV: the scene holds an environment light (its branch is compiled only then).
template <int STACK, bool ENV>
__device__ void lt_path(const SceneDev& S, const PathParams& P, const LightCamDev& C, uint64_t index, float total,
                        float* splat, uint32_t* stack, int stride) {
    const SampleSrc& M = P.smp;
    Seq q;
    q.index = index;
    q.cursor = 0;
    float pdfLight;
    const int light = sample_one_light(S, seq_next(M, q), &pdfLight);
    const float u0 = seq_next(M, q), u1 = seq_next(M, q);
    const float w0 = seq_next(M, q), w1 = seq_next(M, q);
    V3 Le, lpos, ln, dir;
    float pdfA, pdfDir;
    if (!ENV || S.light_tri[light] >= 0) {
        Le = lt_area_origin(S, light, u0, u1, w0, w1, &lpos, &ln, &dir, &pdfA, &pdfDir);
        lt_connect<STACK>(S, C, splat, lpos, ln, v3s(0.f), nullptr, Le, true, total, stack, stride);
    } else {
        // the light vertex's splat is Le = 0: adding it changes no pixel
        Le = v3s(0.f);
        lt_env_origin(S.env, C, u0, u1, w0, w1, &lpos, &dir, &pdfA, &pdfDir);
        ln = dir;
    }
    V3 T = Le * fabsf(dot(ln, dir)) / pdfLight / pdfA / pdfDir;
    V3 o = (lpos + dir * kRayEps) + dir * kRayEps;
    V3 d = dir;
    for (int bounce = 1;; ) {
        Hit h;
        if (!find_hit<false, STACK>(S, o, d, 0.f, kInf, &h, stack, stride)) break;
        const Isect it = query(S, o, d, h);
        BSDF b = make_bsdf(S, it);
        b.importance = true;  // ComputeScatteringFunction(arena, TransportMode::Importance) (integrator/LT.h:69)
        lt_connect<STACK>(S, C, splat, it.pos, it.n, it.outDir, &b, T, false, total, stack, stride);
        ++bounce;
        if (bounce >= P.max_depth) break;
        const float r0 = seq_next(M, q), r1 = seq_next(M, q), r2 = seq_next(M, q);
        V3 wi;
        const V3 f = bsdf_sample(b, neg(d), &wi, &pdfDir, r0, r1, r2);
        o = (it.pos + wi * kRayEps) + wi * kRayEps;
        d = wi;
        T = T * f * fabsf(dot(wi, it.n)) / pdfDir;
    }
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

// Film::AddSample (visual/Film.h:61-87) with ownership instead of locks: the
// sample is stored in its pass slot (pixel px, py of pass sample s_local) with
// a 9-bit mask of the pixels its box footprint covers, as offsets dx, dy in
// {-1, 0, 1} from its own pixel (bit (dy + 1) * 3 + dx + 1).  Offsets are taken
// around the image's edges: a sample whose u wrapped from 1 to 0 in the last
// column (sampler/Sampler.h:60-62) lands in column 0, offset +1.  A sample
// that covers any pixel but its own also flags its pixel in F.border, which
// tells the neighbours' owners to look at its slot (k_film_accumulate).  A
// footprint beyond the neighbouring pixels is counted in F.border[W * H] and
// fails the render.
constexpr uint32_t kFilmOwn = 1u << 4;
__device__ inline void film_record(const FilmDev& F, int64_t slot, int px, int py, float fx, float fy, V3 L) {
    const int W = F.width, H = F.height;
    int x0, x1, y0, y1;
    film_footprint(fx, fy, W, H, &x0, &x1, &y0, &y1);
    uint32_t mask = 0;
    bool far = false;
    for (int yy = y0; yy <= y1; ++yy)
        for (int xx = x0; xx <= x1; ++xx) {
            int dx = xx - px, dy = yy - py;
            if (dx > 1) dx -= W;
            if (dx < -1) dx += W;
            if (dy > 1) dy -= H;
            if (dy < -1) dy += H;
            if (dx < -1 || dx > 1 || dy < -1 || dy > 1) far = true;
            else mask |= 1u << ((dy + 1) * 3 + dx + 1);
        }
    F.pass[slot] = make_float4(L.x, L.y, L.z, (float)mask);
    if (mask != kFilmOwn) F.border[(int64_t)py * W + px] = 1u;
    if (far) atomicAdd(F.border + (int64_t)W * H, 1u);
}

__device__ inline void record_sample(const FilmDev& F, int s_local, int px, int py, float fx, float fy, V3 L) {
    const int64_t WH = (int64_t)F.width * F.height;
    film_record(F, (int64_t)s_local * WH + (int64_t)py * F.width + px, px, py, fx, fy, L);
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

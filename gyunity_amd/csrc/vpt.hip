// VolPathTracing (integrator/VPT.cpp:7-171) on the GPU, in two forms with
// equal films: one path per thread (k_vpt), and a wavefront form (k_vwf_*,
// below); homogeneous media (volume/Homogeneous.cpp:6-64), the Henyey-Greenstein
// phase function (visual/Medium.cpp:8-29), medium interfaces on triangles
// (visual/Medium.h:51-59, visual/Primitive.cpp:11-14) and material-less
// triangles that bound a medium.  Both of the reference's distance samplers:
// free-flight sampling with DirectIllumination at medium points, and
// equi-angular sampling about one light point with ConnectToLight.
//
// The estimator is the reference's FP32 arithmetic in its order, as for
// PathTracing; the reference's own rules are kept where they differ from a
// textbook volumetric path tracer:
//   - camera rays carry no medium;
//   - shadow rays see transmittance through material-less surfaces by
//     closest hits (VisibilityTester::Tr, visual/Visibility.cpp:13-35);
//   - a BSDF/phase sample's light hit is found by Scene::IntersectTr
//     (visual/Scene.cpp:14-26), which crosses a material-less surface without
//     querying it: the next segment starts at the hit point and takes the
//     interface's inside medium, whichever side it leaves by;
//   - ConnectToLight returns its light term without the MIS weight it forms;
//   - equi-angular sampling covers t in [0, 3] whatever the ray's extent;
//   - the roulette test min(1, Y(T) / Y(T)) < 1 never passes.
// A path ends once it has drawn more than 1008 dimensions (the reference reads
// past its 1024 Sobol dimensions there) or crossed 256 boundaries in one loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "integrator.h"

namespace gy {
namespace {

using gyd::V3;
using gyd::v3;
using gyd::v3s;
using gyd::Hit;
using gyd::Isect;
using gyd::LightPoint;
using gyd::SceneDev;

constexpr int kVptMaxCrossings = 256;
constexpr int kVptDimBudget = 1024 - 16;
constexpr float kInv4Pi = 0.07957747154594766788f;  // common/Core.h:188
constexpr float kMaxReal = 3.402823466e+38f;        // common/Core.h:198

// Media of a scene: per medium (sigma_a, g), (sigma_s, 0); per triangle one
// word: inside medium + 1 (bits 0-15), outside medium + 1 (bits 16-30), no
// material (bit 31).  tri == nullptr: no media, every triangle has a material.
struct MediaDev {
    const float4* __restrict__ med;
    const uint32_t* __restrict__ tri;
    int32_t n;
};
struct Med {
    V3 sa, ss, st;
    float g;
};
__device__ inline Med load_medium(const MediaDev& M, int m) {
    const float4 a = M.med[2 * m], b = M.med[2 * m + 1];
    Med r;
    r.sa = v3(a.x, a.y, a.z);
    r.ss = v3(b.x, b.y, b.z);
    r.st = r.ss + r.sa;  // sigma_t(sigma_s + sigma_a)
    r.g = a.w;
    return r;
}
__device__ inline bool no_material(const MediaDev& M, int32_t prim) { return M.tri && (M.tri[prim] >> 31) != 0u; }
// Primitive::Intersect: the primitive's interface when it is a transition
// (inside != outside), else the ray's medium on both sides.
__device__ inline void hit_media(const MediaDev& M, int32_t prim, int32_t ray_med, int32_t* in, int32_t* out) {
    const uint32_t w = M.tri ? M.tri[prim] : 0u;
    const int32_t a = (int32_t)(w & 0xFFFFu) - 1, b = (int32_t)((w >> 16) & 0x7FFFu) - 1;
    if (a != b) {
        *in = a;
        *out = b;
    } else {
        *in = *out = ray_med;
    }
}
// Intersection::GetMedium(w) (visual/Intersection.h:80-82)
__device__ inline int32_t isect_medium(const Isect& it, int32_t in, int32_t out, V3 w) { return gyd::dot(w, it.n) > 0.f ? out : in; }

__device__ inline float dexp(float x) { return (float)exp((double)x); }
__device__ inline float dtan(float x) { return (float)tan((double)x); }
__device__ inline V3 vexp(V3 a) { return v3(dexp(a.x), dexp(a.y), dexp(a.z)); }

// PhaseHG (visual/Medium.h:21-24)
__device__ inline float phase_hg(float cosTheta, float g) {
    const float denom = 1 + g * g + 2 * g * cosTheta;
    return kInv4Pi * (1 - g * g) / (denom * sqrtf(denom));
}
__device__ inline float hg_p(float g, V3 wo, V3 wi) { return phase_hg(gyd::dot(wo, wi), g); }
// HenyeyGreenstein::Sample_p: cos theta about -wo by inversion, the direction
// on CoordinateSystem(wo) (SphericalDirection, math/GeometryUtils.h:29-33).
__device__ inline float hg_sample(float g, V3 wo, V3* wi, float u0, float u1) {
    float cosTheta;
    if ((double)fabsf(g) < 1e-3) {
        cosTheta = 1 - 2 * u0;
    } else {
        const float sqrTerm = (1 - g * g) / (1 - g + 2 * g * u0);
        cosTheta = (1 + g * g - sqrTerm * sqrTerm) / (2 * g);
    }
    const float sinTheta = sqrtf(gyd::rmax(0.f, 1 - cosTheta * cosTheta));
    const float phi = 2 * 3.14159265358979f * u1;
    V3 v1, v2;
    gyd::coordinate_system(wo, &v1, &v2);
    *wi = ((sinTheta * gyd::dcos(phi)) * v1 + (sinTheta * gyd::dsin(phi)) * v2) + cosTheta * gyd::neg(wo);
    return phase_hg(-cosTheta, g);
}

// HomogeneousMedium::Tr over a segment of extent tmax along d
__device__ inline V3 medium_tr(const Med& m, V3 d, float tmax) {
    return vexp(gyd::neg(m.st) * gyd::rmin(tmax * gyd::length(d), kMaxReal));
}
// A scattering point (MediumIntersection): position, outgoing direction,
// medium; zero normals and error bound.
struct MedRec {
    V3 pos, wo;
    int32_t medium;
    bool valid;
};
__device__ inline Isect medium_isect(const MedRec& mi) {
    Isect it;
    it.pos = mi.pos;
    it.n = it.absN = it.ng = it.sdpdu = it.sdpdv = it.pErr = v3s(0.f);
    it.outDir = mi.wo;
    it.u = it.v = 0.f;
    it.prim = -1;
    it.mat = 0u;
    it.light = -1;
    return it;
}
// HomogeneousMedium::Sample: a channel, a free-flight distance in it, the
// weight over the channel-averaged density.
__device__ inline V3 medium_sample(const Med& m, int32_t med, V3 o, V3 d, float tmax, float u_ch, float u_d, MedRec* mi) {
    const int channel = min((int)(u_ch * 3), 2);
    const float dist = -gyd::dlog(1 - u_d) / gyd::comp(m.st, channel);
    const float t = gyd::rmin(dist / gyd::length(d), tmax);
    const bool sampled = t < tmax;
    if (sampled) {
        mi->valid = true;
        mi->pos = o + d * t;
        mi->wo = gyd::neg(d);
        mi->medium = med;
    }
    const V3 Tr = vexp((gyd::neg(m.st) * gyd::rmin(t, kMaxReal)) * gyd::length(d));
    const V3 density = sampled ? (m.st * Tr) : Tr;
    float pdf = 0.f;
    pdf += density.x;
    pdf += density.y;
    pdf += density.z;
    pdf *= 1 / (float)(3);
    if (pdf == 0.f) pdf = 1.f;
    return sampled ? (Tr * m.ss / pdf) : (Tr / pdf);
}
// HomogeneousMedium::EquiAngularSampling: t about the light point's foot on
// the ray over [0, 3]; past the surface, the probability of t beyond min(3, tmax).
__device__ inline V3 medium_equiangular(const Med& m, int32_t med, V3 o, V3 d, float tmax, float u, V3 lightPos, MedRec* mi) {
    const float tMax = 3;
    const float delta = gyd::dot(lightPos - o, d);
    const float D = gyd::length((o + delta * d) - lightPos);
    const float theta_a = gyd::datan2(0.f - delta, D);
    const float theta_b = gyd::datan2(tMax - delta, D);
    const float tau = D * dtan((1 - u) * theta_a + u * theta_b);
    const float t = delta + tau;
    const bool sampleMedia = t < tmax;
    const V3 Tr = sampleMedia ? vexp(gyd::neg(m.st) * t) : medium_tr(m, d, tmax);
    const float pdf = sampleMedia ? D / ((theta_b - theta_a) * (D * D + tau * tau))
                                  : 1 - (gyd::datan2(gyd::rmin(tMax, tmax) - delta, D) - theta_a) / (theta_b - theta_a);
    if (sampleMedia) {
        mi->valid = true;
        mi->pos = o + d * t;
        mi->wo = gyd::neg(d);
        mi->medium = med;
    }
    return sampleMedia ? Tr * m.ss / pdf : medium_tr(m, d, tmax) / pdf;
}

// Triangle::Sample(pdf, u) (shape/Triangle.cpp:141-158): the area sample
// AreaLight::SampleOnePoint returns, pdf 1 / area.
__device__ inline LightPoint tri_sample_area(const SceneDev& S, int tri, float u0, float u1, float* pdf) {
    const float4* q = S.shade + (size_t)tri * 8;
    const V3 p0 = gyd::ld3(q[0]), p1 = gyd::ld3(q[1]), p2 = gyd::ld3(q[2]);
    const V3 n0 = gyd::ld3(q[3]), n1 = gyd::ld3(q[4]), n2 = gyd::ld3(q[5]), fn = gyd::ld3(q[6]);
    const float su0 = sqrtf(u0);
    const float b0 = 1 - su0, b1 = u1 * su0, b2 = 1 - b0 - b1;
    LightPoint lp;
    lp.pos = (b0 * p0 + b1 * p1) + b2 * p2;
    lp.fn = fn;
    lp.n = (b0 * n0 + b1 * n1) + b2 * n2;
    lp.pErr = gyd::gamma_n(6) * ((gyd::vabs(b0 * p0) + gyd::vabs(b1 * p1)) + gyd::vabs(b2 * p2));
    *pdf = 1 / q[2].w;
    return lp;
}

// VisibilityTester::Tr: closest hits from p0 towards the light point; a
// surface with a material occludes; each segment's medium attenuates; a
// material-less surface is queried and the segment restarts from it.
template <int STACK>
__device__ V3 vis_tr(const SceneDev& S, const MediaDev& M, const Isect& p0, int32_t p0_in, int32_t p0_out, const LightPoint& lp,
                     uint32_t* stack, int stride) {
    V3 o, d;
    float tmax;
    gyd::spawn_to(p0, lp, &o, &d, &tmax);
    int32_t med = isect_medium(p0, p0_in, p0_out, d);
    V3 Tr = v3s(1.f);
    for (int k = 0; k < kVptMaxCrossings; ++k) {
        Hit h;
        const bool hit = gyd::find_hit<false, STACK>(S, o, d, 0.f, tmax, &h, stack, stride);
        if (hit && !no_material(M, h.prim)) return v3s(0.f);
        if (med >= 0) Tr = Tr * medium_tr(load_medium(M, med), d, hit ? h.t : tmax);
        if (!hit) return Tr;
        const Isect it = gyd::query(S, o, d, h);
        int32_t in, out;
        hit_media(M, h.prim, med, &in, &out);
        gyd::spawn_to(it, lp, &o, &d, &tmax);
        med = isect_medium(it, in, out, d);
    }
    return v3s(0.f);
}
// Scene::IntersectTr: closest hits with each segment's transmittance, crossing
// material-less surfaces unqueried (the next segment starts at the hit point,
// in the interface's inside medium).  o is updated to the last segment's origin.
template <int STACK>
__device__ bool intersect_tr(const SceneDev& S, const MediaDev& M, V3* o, V3 d, int32_t* med, Hit* h, V3* Tr, uint32_t* stack,
                             int stride) {
    *Tr = v3s(1.f);
    for (int k = 0; k < kVptMaxCrossings; ++k) {
        const bool hit = gyd::find_hit<false, STACK>(S, *o, d, 0.f, gyd::kInf, h, stack, stride);
        if (*med >= 0) *Tr = *Tr * medium_tr(load_medium(M, *med), d, hit ? h->t : gyd::kInf);
        if (!hit) return false;
        if (!no_material(M, h->prim)) return true;
        const float4* q = S.shade + (size_t)h->prim * 8;
        const V3 p0 = gyd::ld3(q[0]), p1 = gyd::ld3(q[1]), p2 = gyd::ld3(q[2]);
        const float b1 = h->b1, b2 = h->b2;
        const V3 pos = ((1 - b1 - b2) * p0 + b1 * p1) + b2 * p2;
        int32_t in, out;
        hit_media(M, h->prim, *med, &in, &out);
        *o = gyd::offset_origin(pos, v3s(0.f), v3s(0.f), d) + d * gyd::kRayEps;
        *med = in;
    }
    return false;
}

// Integrator::DirectIllumination with handleMedia (integrator/Integrator.cpp:9-107)
// at a surface (b given) or a medium point (phase g).
template <int STACK, bool ENV>
__device__ V3 di_vpt(const SceneDev& S, const MediaDev& M, const Isect& it, int32_t in, int32_t out, const gyd::BSDF* b, float g,
                     float uL, float u0, float u1, float v0, float v1, float v2, uint32_t* stack, int stride) {
    float lsp;
    const int light = gyd::sample_one_light(S, uL, &lsp);
    const int ltri = S.light_tri[light];  // < 0: the environment light
    const V3 Le = v3(S.light_le[3 * light], S.light_le[3 * light + 1], S.light_le[3 * light + 2]);
    V3 L1 = v3s(0.f), L2 = v3s(0.f);
    float w1 = 1.f, w2 = 1.f;
    {
        float pdf;
        LightPoint lp;
        V3 wi, Li;
        if (ENV && ltri < 0) {
            lp = gyd::env_light_point_call(S.env, it.pos, u0, u1, &wi, &Li, &pdf);
        } else {
            lp = gyd::sample_light_point(S, ltri, it.pos, u0, u1, &pdf);
            wi = gyd::norm(lp.pos - it.pos);
            Li = gyd::dot(lp.n, gyd::neg(wi)) > 0.f ? Le : v3s(0.f);
        }
        if (pdf > 0.f && !gyd::is_zero(Li)) {
            float sp, c = 1.f;
            V3 f;
            if (b) {
                sp = gyd::bsdf_pdf(*b, it.outDir, wi);
                f = gyd::bsdf_eval(*b, it.outDir, wi);
                c = fabsf(gyd::dot(it.n, wi));
            } else {
                const float p = hg_p(g, it.outDir, wi);
                f = v3s(p);
                sp = p;
            }
            if (!gyd::is_zero(f)) {
                const V3 Tr = vis_tr<STACK>(S, M, it, in, out, lp, stack, stride);
                w1 = gyd::power_heuristic(pdf, sp);
                L1 = Li * Tr * f * c / pdf;
            }
        }
    }
    {
        V3 wi;
        float pdf, c = 1.f;
        V3 f;
        if (b) {
            f = gyd::bsdf_sample(*b, it.outDir, &wi, &pdf, v0, v1, v2);
            c = fabsf(gyd::dot(it.n, wi));
        } else {
            const float p = hg_sample(g, it.outDir, &wi, v0, v1);
            f = v3s(p);
            pdf = p;
        }
        if (pdf != 0.f && !gyd::is_zero(f)) {
            const float lpdf = (ENV && ltri < 0) ? gyd::env_pdf_dir_call(S.env, wi) : gyd::light_pdf(S, ltri, it, wi);
            if (lpdf == 0.f) {
                w2 = 0.f;
            } else {
                w2 = gyd::power_heuristic(pdf, lpdf);
                V3 ro;
                gyd::spawn_ray(it, wi, &ro);
                int32_t med = isect_medium(it, in, out, wi);
                Hit h2;
                V3 Tr;
                if (intersect_tr<STACK>(S, M, &ro, wi, &med, &h2, &Tr, stack, stride)) {
                    if (gyd::prim_light(S, h2.prim) == light) {
                        const Isect li = gyd::query(S, ro, wi, h2);
                        const V3 Lem = gyd::dot(li.n, gyd::neg(wi)) > 0.f ? Le : v3s(0.f);
                        L2 = Lem * f * Tr * c / pdf;
                    }
                } else if (ENV && ltri < 0) {
                    L2 = gyd::env_emission_call(S.env, wi) * f * Tr * c / pdf;
                }
            }
        }
    }
    return (L1 * w1 + L2 * w2) / lsp;
}

// VolPathTracing::ConnectToLight: the light point alone, L1 without its MIS weight.
template <int STACK>
__device__ V3 connect_to_light(const SceneDev& S, const MediaDev& M, const MedRec& mi, float g, const LightPoint& lp, V3 Le,
                               float lightPdf, uint32_t* stack, int stride) {
    float cosTheta = 1.f;
    const V3 dir = mi.pos - lp.pos;
    const V3 wi = gyd::neg(gyd::norm(dir));
    if (!gyd::is_zero(lp.n)) {
        cosTheta = gyd::dot(gyd::norm(dir), lp.n);
        if (cosTheta < 0.f) return v3s(0.f);
    }
    const V3 Tr = vis_tr<STACK>(S, M, medium_isect(mi), mi.medium, mi.medium, lp, stack, stride);
    const float p = hg_p(g, mi.wo, wi);
    return Le * v3s(p) * Tr * cosTheta / gyd::dot(dir, dir) / lightPdf;
}

// VolPathTracing::Li for camera sample s of pixel (px, py).
template <int STACK, bool ENV, bool EQA>
__device__ V3 vpt_path(const SceneDev& S, const MediaDev& M, const PathParams& P, int px, int py, int s, uint32_t* stack,
                       int stride, float* fx, float* fy, uint32_t* dims) {
    const gyd::SampleSrc& Q = P.smp;
    gyd::Seq q;
    q.index = gyd::enum_index(P.en, (uint32_t)s, (uint32_t)px, (uint32_t)py);
    q.cursor = 0;
    const float u = P.en.sx * gyd::seq_next(Q, q) - (float)px;
    const float v = P.en.sy * gyd::seq_next(Q, q) - (float)py;
    V3 o, d;
    gyd::generate_ray(P.cam, px, py, u, v, &o, &d);
    int32_t med = -1;  // camera rays carry no medium
    bool delta = false;
    V3 T = v3(1.f, 1.f, 1.f);
    V3 L = v3s(0.f);
    int crossings = 0;
    for (int i = 0; i < P.max_depth; ++i) {
        if (q.cursor > kVptDimBudget) break;
        Hit h;
        const bool hit = gyd::find_hit<false, STACK>(S, o, d, 0.f, gyd::kInf, &h, stack, stride);
        const float tmax = hit ? h.t : gyd::kInf;  // Primitive::Intersect shrinks the ray's tMax
        MedRec mi;
        mi.valid = false;
        float eqProb = 0.f, eqPdf = 0.f;
        V3 eqLe = v3s(0.f);
        LightPoint eqLp;
        if (EQA) {
            const int light = gyd::sample_one_light(S, gyd::seq_next(Q, q), &eqProb);
            const float a = gyd::seq_next(Q, q), b = gyd::seq_next(Q, q);
            eqLp = tri_sample_area(S, S.light_tri[light], a, b, &eqPdf);
            eqLe = v3(S.light_le[3 * light], S.light_le[3 * light + 1], S.light_le[3 * light + 2]);
            if (med >= 0) T = T * medium_equiangular(load_medium(M, med), med, o, d, tmax, gyd::seq_next(Q, q), eqLp.pos, &mi);
        } else if (med >= 0) {
            const float a = gyd::seq_next(Q, q), b = gyd::seq_next(Q, q);
            T = T * medium_sample(load_medium(M, med), med, o, d, tmax, a, b, &mi);
        }
        if (gyd::is_zero(T)) break;
        if (mi.valid) {
            const float g = M.med[2 * mi.medium].w;
            if (EQA) {
                L = L + T * connect_to_light<STACK>(S, M, mi, g, eqLp, eqLe, eqPdf, stack, stride) / eqProb;
            } else {
                const float uL = gyd::seq_next(Q, q);
                const float u0 = gyd::seq_next(Q, q), u1 = gyd::seq_next(Q, q);
                const float v0 = gyd::seq_next(Q, q), v1 = gyd::seq_next(Q, q), v2 = gyd::seq_next(Q, q);
                L = L + T * di_vpt<STACK, ENV>(S, M, medium_isect(mi), mi.medium, mi.medium, nullptr, g, uL, u0, u1, v0, v1, v2,
                                               stack, stride);
            }
            const float a = gyd::seq_next(Q, q), b = gyd::seq_next(Q, q);
            V3 wi;
            hg_sample(g, mi.wo, &wi, a, b);
            gyd::spawn_ray(medium_isect(mi), wi, &o);
            d = wi;
            med = mi.medium;
            delta = false;
        } else {
            if (!hit) {
                // VPT.cpp:60-64: a camera ray, or one after a delta bounce,
                // that leaves the scene returns T * Le(ray) in place of L
                if (ENV && (i == 0 || delta) && S.env.w > 0) L = T * gyd::env_emission_call(S.env, d);
                break;
            }
            const Isect it = gyd::query(S, o, d, h);
            int32_t in, out;
            hit_media(M, h.prim, med, &in, &out);
            if (no_material(M, h.prim)) {
                if (++crossings > kVptMaxCrossings) break;
                gyd::spawn_ray(it, d, &o);
                med = isect_medium(it, in, out, d);
                --i;
                continue;
            }
            const gyd::BSDF b = gyd::make_bsdf(S, it);
            if ((i == 0 || delta) && it.light >= 0) {
                const V3 Le = v3(S.light_le[3 * it.light], S.light_le[3 * it.light + 1], S.light_le[3 * it.light + 2]);
                const V3 e = gyd::dot(it.n, it.outDir) > 0.f ? Le : v3s(0.f);
                L = L + T * e;
            } else {
                const float uL = gyd::seq_next(Q, q);
                const float u0 = gyd::seq_next(Q, q), u1 = gyd::seq_next(Q, q);
                const float v0 = gyd::seq_next(Q, q), v1 = gyd::seq_next(Q, q), v2 = gyd::seq_next(Q, q);
                L = L + T * di_vpt<STACK, ENV>(S, M, it, in, out, &b, 0.f, uL, u0, u1, v0, v1, v2, stack, stride);
            }
            const float r0 = gyd::seq_next(Q, q), r1 = gyd::seq_next(Q, q), r2 = gyd::seq_next(Q, q);
            V3 wi;
            float pdf;
            const V3 f = gyd::bsdf_sample(b, gyd::neg(d), &wi, &pdf, r0, r1, r2);
            if (gyd::is_zero(f) || pdf == 0.f) break;
            const V3 est = f * fabsf(gyd::dot(it.n, wi)) / pdf;
            delta = gyd::bsdf_delta(b);
            T = T * est;
            gyd::spawn_ray(it, wi, &o);
            d = wi;
            med = isect_medium(it, in, out, wi);
        }
    }
    *fx = px + u;
    *fy = py + v;
    if (dims) *dims = q.cursor;
    return L;
}

// Three waves per SIMD: the equi-angular kernel otherwise holds 256 VGPRs (one
// wave) and runs 2.3x slower; with the hint it keeps 168 and spills some.
template <int STACK, bool ENV, bool EQA>
__global__ void __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(3)))
k_vpt(SceneDev S, MediaDev M, PathParams P, FilmDev F, int s0, int n_s) {
    __shared__ uint32_t stack[STACK * 128];
    const int64_t t = (int64_t)blockIdx.x * 128 + threadIdx.x;
    int s_local, px, py;
    pass_pixel(F, t, &s_local, &px, &py);
    if (s_local < n_s && px < F.width && py < F.height) {
        float fx, fy;
        const V3 L = vpt_path<STACK, ENV, EQA>(S, M, P, px, py, s0 + s_local, stack + threadIdx.x, 128, &fx, &fy, nullptr);
        record_sample(F, s_local, px, py, fx, fy, L);
    }
}

template <int STACK, bool ENV, bool EQA>
__global__ void __launch_bounds__(128) k_vpt_li_samples(SceneDev S, MediaDev M, PathParams P, int n, const int32_t* __restrict__ xs,
                                                        const int32_t* __restrict__ ys, const int32_t* __restrict__ ss,
                                                        float* __restrict__ L, float* __restrict__ raster, int32_t* __restrict__ dims) {
    __shared__ uint32_t stack[STACK * 128];
    const int i = blockIdx.x * 128 + threadIdx.x;
    if (i >= n) return;
    float fx, fy;
    uint32_t nd;
    const V3 v = vpt_path<STACK, ENV, EQA>(S, M, P, xs[i], ys[i], ss[i], stack + threadIdx.x, 128, &fx, &fy, &nd);
    L[3 * i] = v.x;
    L[3 * i + 1] = v.y;
    L[3 * i + 2] = v.z;
    raster[2 * i] = fx;
    raster[2 * i + 1] = fy;
    dims[i] = (int32_t)nd;
}

// ---------------------------------------------------------------- wavefront form
// VolPathTracing::Li split into per-vertex stages over queues of path slots,
// with the same FP32 operations in the same order as vpt_path, so that its
// film equals the per-path kernel's bit for bit:
//
//   start -> [ trace (closest hit of the path ray)
//              -> vertex (medium sampling, the vertex's light set-up,
//                 the next ray: phase or BSDF sample, or a boundary crossing)
//              -> trace_shadow (VisibilityTester::Tr: closest hits through
//                 material-less surfaces, each segment's transmittance)
//              -> trace_mis (Scene::IntersectTr towards the chosen light)
//              -> resolve (adds the vertex's light in vpt_path's order,
//                 finishes paths, builds the next active queue in slot order) ]
//
// until no path is left: a boundary crossing takes a round without adding a
// bounce, so the number of rounds is found on the device.  The persistent
// trace waves run each shadow and MIS segment loop in the lane that holds it,
// refilling idle lanes from the queue.
enum VwfQueue { VQ_ACTIVE = 0, VQ_NEXT = 1, VQ_SHADOW = 2, VQ_MIS = 3, VQ_COUNT = 4 };
constexpr int kVwfCtr = 32;                  // counter stride in words (one 128-byte line each)
constexpr int kVwfChunks = 8;                // queue chunks a trace launch claims from, one per XCD
constexpr int kVwfWords = VQ_COUNT + 3 * kVwfChunks;  // queue lengths, then the fetch counters of each trace mode
// Path flags (pk): the slot holds a path; its vertex set up direct light (DI)
// or an equi-angular connection (CONN, CONN0 when the light faces away); a
// shadow or MIS segment was traced for it; the path ends after this round.
constexpr uint32_t VK_LIVE = 1u, VK_DI = 2u, VK_CONN = 4u, VK_CONN0 = 8u, VK_SH = 16u, VK_MIS = 32u, VK_DONE = 64u;

struct VwfDev {
    float4 *o, *d;        // path ray (xyz)
    float4 *T, *L;        // throughput (xyz) + film x, radiance (xyz) + film y
    uint4* st;            // sample index lo / hi, cursor, depth | delta << 8 | crossings << 9
    int32_t* med;         // the path ray's medium (-1: none)
    float4* hit;          // t, prim (-1: miss), b1, b2
    uint32_t* pk;         // VK_* flags
    float4 *pT;           // throughput at the vertex (xyz); P(light) (DI) or the light pick's probability (CONN)
    float4 *pA;           // DI: Li (xyz), |cos| of the light sample | CONN: Le * phase (xyz), cos at the light
    float4 *pB;           // DI: f of the light sample (xyz), its pdf | CONN: squared distance, the area pdf
    float4 *pC;           // DI: w1, w2, |cos| and pdf of the BSDF / phase sample
    float4 *pD;           // DI: f of the BSDF / phase sample (xyz), light (int)
    float4 *tr1;          // shadow transmittance
    float4 *e2;           // MIS: radiance reached (xyz), 1 when it counts
    float4 *tr2;          // MIS transmittance
    float4 *sh_o, *sh_d;  // shadow segment: origin + tmax, direction + medium (int)
    float4 *lp_pos, *lp_fn, *lp_err;  // the light point the shadow segments aim at
    float4 *mis_o, *mis_d;            // MIS segment: origin, direction + medium (int)
    uint32_t* q[VQ_COUNT];
    uint32_t* counts;
    int64_t capacity;
    uint32_t* spill;
    int64_t spill_stride;
};

__device__ inline float4 f4v(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }
__device__ inline V3 v3f(float4 a) { return v3(a.x, a.y, a.z); }

// Appends for a 256-lane workgroup, K items per lane, NQ queues, one atomic
// per queue per workgroup (a queue counter shared by every workgroup is the
// contended word).  Every lane of the workgroup must call it.
template <int NQ, int K>
__device__ inline void vwf_block_push(uint32_t* const (&q)[NQ], uint32_t* const (&cnt)[NQ], const uint32_t (&item)[K],
                                      const bool (&pred)[NQ][K]) {
    __shared__ uint32_t nw[NQ][K][4];
    __shared__ uint32_t base[NQ];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint64_t mask[NQ][K];
#pragma unroll
    for (int j = 0; j < NQ; ++j)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            mask[j][k] = __ballot(pred[j][k]);
            if (lane == 0) nw[j][k][wave] = (uint32_t)__popcll(mask[j][k]);
        }
    __syncthreads();
    if (threadIdx.x < NQ) {
        uint32_t total = 0;
        for (int k = 0; k < K; ++k) total += nw[threadIdx.x][k][0] + nw[threadIdx.x][k][1] + nw[threadIdx.x][k][2] + nw[threadIdx.x][k][3];
        base[threadIdx.x] = total ? atomicAdd(cnt[threadIdx.x], total) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
        uint32_t off = base[j];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            uint32_t o = off;
            for (int w = 0; w < wave; ++w) o += nw[j][k][w];
            if (pred[j][k]) q[j][o + __popcll(mask[j][k] & ((1ull << lane) - 1ull))] = item[k];
            off += nw[j][k][0] + nw[j][k][1] + nw[j][k][2] + nw[j][k][3];
        }
    }
}

__global__ void __launch_bounds__(256) k_vwf_start(VwfDev W, PathParams P, FilmDev F, int s0, int n_s, int64_t n_slots) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int s_local = 0, px = 0, py = 0;
    if (t < n_slots) pass_pixel(F, t, &s_local, &px, &py);
    const bool valid = t < n_slots && s_local < n_s && px < F.width && py < F.height;
    const uint32_t slot = (uint32_t)t;
    if (valid) {
        gyd::Seq q;
        q.index = gyd::enum_index(P.en, (uint32_t)(s0 + s_local), (uint32_t)px, (uint32_t)py);
        q.cursor = 0;
        const float u = P.en.sx * gyd::seq_next(P.smp, q) - (float)px;
        const float v = P.en.sy * gyd::seq_next(P.smp, q) - (float)py;
        V3 o, d;
        gyd::generate_ray(P.cam, px, py, u, v, &o, &d);
        W.o[slot] = f4v(o, 0.f);
        W.d[slot] = f4v(d, 0.f);
        W.T[slot] = make_float4(1.f, 1.f, 1.f, px + u);
        W.L[slot] = make_float4(0.f, 0.f, 0.f, py + v);
        W.st[slot] = make_uint4((uint32_t)q.index, (uint32_t)(q.index >> 32), q.cursor, 0u);
        W.med[slot] = -1;  // camera rays carry no medium
        W.pk[slot] = P.max_depth > 0 ? VK_LIVE : (VK_LIVE | VK_DONE);
    } else if (t < n_slots) {
        W.pk[slot] = 0u;
    }
    uint32_t* const qs[1] = {W.q[VQ_ACTIVE]};
    uint32_t* const cs[1] = {&W.counts[VQ_ACTIVE * kVwfCtr]};
    const uint32_t it[1] = {slot};
    const bool pr[1][1] = {{valid && P.max_depth > 0}};
    vwf_block_push<1, 1>(qs, cs, it, pr);
}

// Integrator::DirectIllumination with handleMedia set up at a surface (b) or
// a medium point (phase g): the light sample's segment and the BSDF / phase
// sample's segment go to their queues; the factors di_vpt multiplies are kept.
template <bool ENV>
__device__ inline void vwf_di_setup(const SceneDev& S, const VwfDev& W, uint32_t slot, const Isect& it, int32_t in, int32_t out,
                                    const gyd::BSDF* b, float g, V3 T, const float (&u)[6], uint32_t* pk, bool* push_sh,
                                    bool* push_mis) {
    float lsp;
    const int light = gyd::sample_one_light(S, u[0], &lsp);
    const int ltri = S.light_tri[light];  // < 0: the environment light
    const V3 Le = v3(S.light_le[3 * light], S.light_le[3 * light + 1], S.light_le[3 * light + 2]);
    float w1 = 1.f, w2 = 1.f;
    {
        float pdf;
        LightPoint lp;
        V3 wi, Li;
        if (ENV && ltri < 0) {
            lp = gyd::env_light_point_call(S.env, it.pos, u[1], u[2], &wi, &Li, &pdf);
        } else {
            lp = gyd::sample_light_point(S, ltri, it.pos, u[1], u[2], &pdf);
            wi = gyd::norm(lp.pos - it.pos);
            Li = gyd::dot(lp.n, gyd::neg(wi)) > 0.f ? Le : v3s(0.f);
        }
        if (pdf > 0.f && !gyd::is_zero(Li)) {
            float sp, c = 1.f;
            V3 f;
            if (b) {
                sp = gyd::bsdf_pdf(*b, it.outDir, wi);
                f = gyd::bsdf_eval(*b, it.outDir, wi);
                c = fabsf(gyd::dot(it.n, wi));
            } else {
                const float p = hg_p(g, it.outDir, wi);
                f = v3s(p);
                sp = p;
            }
            if (!gyd::is_zero(f)) {
                V3 so, sd;
                float stmax;
                gyd::spawn_to(it, lp, &so, &sd, &stmax);
                const int32_t smed = isect_medium(it, in, out, sd);
                W.sh_o[slot] = f4v(so, stmax);
                W.sh_d[slot] = f4v(sd, __int_as_float(smed));
                W.lp_pos[slot] = f4v(lp.pos, 0.f);
                W.lp_fn[slot] = f4v(lp.fn, 0.f);
                W.lp_err[slot] = f4v(lp.pErr, 0.f);
                w1 = gyd::power_heuristic(pdf, sp);
                W.pA[slot] = f4v(Li, c);
                W.pB[slot] = f4v(f, pdf);
                *pk |= VK_SH;
                *push_sh = true;
            }
        }
    }
    V3 wi;
    float pdf, c = 1.f;
    V3 f;
    if (b) {
        f = gyd::bsdf_sample(*b, it.outDir, &wi, &pdf, u[3], u[4], u[5]);
        c = fabsf(gyd::dot(it.n, wi));
    } else {
        const float p = hg_sample(g, it.outDir, &wi, u[3], u[4]);
        f = v3s(p);
        pdf = p;
    }
    if (pdf != 0.f && !gyd::is_zero(f)) {
        const float lpdf = (ENV && ltri < 0) ? gyd::env_pdf_dir_call(S.env, wi) : gyd::light_pdf(S, ltri, it, wi);
        if (lpdf == 0.f) {
            w2 = 0.f;
        } else {
            w2 = gyd::power_heuristic(pdf, lpdf);
            V3 ro;
            gyd::spawn_ray(it, wi, &ro);
            W.mis_o[slot] = f4v(ro, 0.f);
            W.mis_d[slot] = f4v(wi, __int_as_float(isect_medium(it, in, out, wi)));
            W.pD[slot] = f4v(f, __int_as_float(light));
            *pk |= VK_MIS;
            *push_mis = true;
        }
    }
    W.pC[slot] = make_float4(w1, w2, c, pdf);
    W.pT[slot] = f4v(T, lsp);
    *pk |= VK_DI;
}

// One vertex of vpt_path for each path of the active queue.
// Four waves per SIMD (128 VGPRs, some spilled; 132-145 and three waves
// without the hint): vertices 161 -> 144 ms a render; five waves took 213.
template <bool ENV, bool EQA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_vwf_vertex(SceneDev S, MediaDev M, PathParams P, VwfDev W) {
    const uint32_t n = W.counts[VQ_ACTIVE * kVwfCtr];
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x * 256 >= n) return;  // workgroup-uniform
    bool push_sh = false, push_mis = false;
    uint32_t slot = 0;
    if (i < n) {
        slot = W.q[VQ_ACTIVE][i];
        const gyd::SampleSrc& Q = P.smp;
        const uint4 sv = W.st[slot];
        gyd::Seq q;
        q.index = (uint64_t)sv.x | ((uint64_t)sv.y << 32);
        q.cursor = sv.z;
        int depth = (int)(sv.w & 0xFFu);
        bool delta = (sv.w & 0x100u) != 0u;
        int crossings = (int)(sv.w >> 9);
        uint32_t pk = VK_LIVE;
        bool ends = false, moved = false;  // moved: a new path ray (o, d, medium) is written
        const float4 Tv = W.T[slot];
        V3 T = v3f(Tv);
        V3 o = v3f(W.o[slot]), d = v3f(W.d[slot]);
        int32_t med = W.med[slot];
        if (q.cursor > (uint32_t)kVptDimBudget) {
            ends = true;
        } else {
            const float4 hv = W.hit[slot];
            Hit h;
            h.t = hv.x;
            h.prim = __float_as_int(hv.y);
            h.b1 = hv.z;
            h.b2 = hv.w;
            const bool hit = h.prim >= 0;
            const float tmax = hit ? h.t : gyd::kInf;  // Primitive::Intersect shrinks the ray's tMax
            MedRec mi;
            mi.valid = false;
            float eqProb = 0.f, eqPdf = 0.f;
            V3 eqLe = v3s(0.f);
            LightPoint eqLp;
            if (EQA) {
                const int light = gyd::sample_one_light(S, gyd::seq_next(Q, q), &eqProb);
                const float a = gyd::seq_next(Q, q), b = gyd::seq_next(Q, q);
                eqLp = tri_sample_area(S, S.light_tri[light], a, b, &eqPdf);
                eqLe = v3(S.light_le[3 * light], S.light_le[3 * light + 1], S.light_le[3 * light + 2]);
                if (med >= 0) T = T * medium_equiangular(load_medium(M, med), med, o, d, tmax, gyd::seq_next(Q, q), eqLp.pos, &mi);
            } else if (med >= 0) {
                const float a = gyd::seq_next(Q, q), b = gyd::seq_next(Q, q);
                T = T * medium_sample(load_medium(M, med), med, o, d, tmax, a, b, &mi);
            }
            if (gyd::is_zero(T)) {
                ends = true;
            } else if (mi.valid) {
                const float g = M.med[2 * mi.medium].w;
                const Isect mit = medium_isect(mi);
                if (EQA) {
                    // VolPathTracing::ConnectToLight, up to the transmittance
                    float cosTheta = 1.f;
                    const V3 dir = mi.pos - eqLp.pos;
                    const V3 wi = gyd::neg(gyd::norm(dir));
                    bool faces = true;
                    if (!gyd::is_zero(eqLp.n)) {
                        cosTheta = gyd::dot(gyd::norm(dir), eqLp.n);
                        faces = !(cosTheta < 0.f);
                    }
                    if (faces) {
                        V3 so, sd;
                        float stmax;
                        gyd::spawn_to(mit, eqLp, &so, &sd, &stmax);
                        const int32_t smed = isect_medium(mit, mi.medium, mi.medium, sd);
                        W.sh_o[slot] = f4v(so, stmax);
                        W.sh_d[slot] = f4v(sd, __int_as_float(smed));
                        W.lp_pos[slot] = f4v(eqLp.pos, 0.f);
                        W.lp_fn[slot] = f4v(eqLp.fn, 0.f);
                        W.lp_err[slot] = f4v(eqLp.pErr, 0.f);
                        const float p = hg_p(g, mi.wo, wi);
                        W.pA[slot] = f4v(eqLe * v3s(p), cosTheta);
                        W.pB[slot] = make_float4(gyd::dot(dir, dir), eqPdf, 0.f, 0.f);
                        pk |= VK_CONN | VK_SH;
                        push_sh = true;
                    } else {
                        pk |= VK_CONN0;
                    }
                    W.pT[slot] = f4v(T, eqProb);
                } else {
                    float u[6];
                    for (int k = 0; k < 6; ++k) u[k] = gyd::seq_next(Q, q);
                    vwf_di_setup<ENV>(S, W, slot, mit, mi.medium, mi.medium, nullptr, g, T, u, &pk, &push_sh, &push_mis);
                }
                const float a = gyd::seq_next(Q, q), b = gyd::seq_next(Q, q);
                V3 wi;
                hg_sample(g, mi.wo, &wi, a, b);
                gyd::spawn_ray(mit, wi, &o);
                d = wi;
                med = mi.medium;
                delta = false;
                ++depth;
                moved = true;
            } else if (!hit) {
                // VPT.cpp:60-64: a camera ray, or one after a delta bounce,
                // that leaves the scene returns T * Le(ray) in place of L
                if (ENV && (depth == 0 || delta) && S.env.w > 0) {
                    const V3 Le = T * gyd::env_emission_call(S.env, d);
                    W.L[slot] = f4v(Le, W.L[slot].w);
                }
                ends = true;
            } else {
                const Isect it = gyd::query(S, o, d, h);
                int32_t in, out;
                hit_media(M, h.prim, med, &in, &out);
                if (no_material(M, h.prim)) {
                    if (++crossings > kVptMaxCrossings) {
                        ends = true;
                    } else {
                        gyd::spawn_ray(it, d, &o);
                        med = isect_medium(it, in, out, d);
                        moved = true;  // the same vertex number again
                    }
                } else {
                    const gyd::BSDF b = gyd::make_bsdf(S, it);
                    if ((depth == 0 || delta) && it.light >= 0) {
                        const V3 Le = v3(S.light_le[3 * it.light], S.light_le[3 * it.light + 1], S.light_le[3 * it.light + 2]);
                        const V3 e = gyd::dot(it.n, it.outDir) > 0.f ? Le : v3s(0.f);
                        const float4 Lv = W.L[slot];
                        const V3 L = v3f(Lv) + T * e;
                        W.L[slot] = f4v(L, Lv.w);
                    } else {
                        float u[6];
                        for (int k = 0; k < 6; ++k) u[k] = gyd::seq_next(Q, q);
                        vwf_di_setup<ENV>(S, W, slot, it, in, out, &b, 0.f, T, u, &pk, &push_sh, &push_mis);
                    }
                    const float r0 = gyd::seq_next(Q, q), r1 = gyd::seq_next(Q, q), r2 = gyd::seq_next(Q, q);
                    V3 wi;
                    float pdf;
                    const V3 f = gyd::bsdf_sample(b, gyd::neg(d), &wi, &pdf, r0, r1, r2);
                    if (gyd::is_zero(f) || pdf == 0.f) {
                        ends = true;
                    } else {
                        const V3 est = f * fabsf(gyd::dot(it.n, wi)) / pdf;
                        delta = gyd::bsdf_delta(b);
                        T = T * est;
                        gyd::spawn_ray(it, wi, &o);
                        d = wi;
                        med = isect_medium(it, in, out, wi);
                        ++depth;
                        moved = true;
                    }
                }
            }
        }
        if (moved && depth >= P.max_depth) ends = true;
        if (!ends) {
            W.o[slot] = f4v(o, 0.f);
            W.d[slot] = f4v(d, 0.f);
            W.med[slot] = med;
            W.T[slot] = f4v(T, Tv.w);
            W.st[slot] = make_uint4(sv.x, sv.y, q.cursor, (uint32_t)depth | (delta ? 0x100u : 0u) | ((uint32_t)crossings << 9));
        }
        W.pk[slot] = pk | (ends ? VK_DONE : 0u);
    }
    uint32_t* const qs[2] = {W.q[VQ_SHADOW], W.q[VQ_MIS]};
    uint32_t* const cs[2] = {&W.counts[VQ_SHADOW * kVwfCtr], &W.counts[VQ_MIS * kVwfCtr]};
    const uint32_t it[1] = {slot};
    const bool pr[2][1] = {{push_sh}, {push_mis}};
    vwf_block_push<2, 1>(qs, cs, it, pr);
}

// Persistent trace waves.  MODE 0: closest hit of the path ray (active
// queue).  MODE 1: VisibilityTester::Tr of a shadow segment: a surface with
// a material occludes; each segment's medium attenuates; a material-less
// surface is queried and the segment restarts from it towards the light
// point (vis_tr).  MODE 2: Scene::IntersectTr of the MIS segment: closest
// hits that cross material-less surfaces unqueried; then the chosen light's
// emission if it is what the segment reaches (intersect_tr and di_vpt's tail).
// A lane keeps its segment loop until the loop ends; idle lanes take new
// entries, 64 per atomic.
template <int S, int MODE, bool ENV>
__device__ inline void vwf_trace_loop(const SceneDev& Sc, const MediaDev& M, const VwfDev& W, int refill, uint32_t* stack) {
    constexpr int qid = MODE == 0 ? VQ_ACTIVE : (MODE == 1 ? VQ_SHADOW : VQ_MIS);
    const uint32_t n = W.counts[qid * kVwfCtr];
    // The queue is cut into kVwfChunks regions with a fetch counter each; a
    // wave starts on the region of its XCD (workgroups are dealt to the XCDs
    // round-robin) and moves on when it runs dry, so no one counter takes
    // every claim.
    uint32_t* fetch = &W.counts[(VQ_COUNT + MODE * kVwfChunks) * kVwfCtr];
    int ck = (int)(blockIdx.x % kVwfChunks), tries = 0;  // wave-uniform
    const int lane = threadIdx.x & 63;
    const int64_t gl = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t* lds = stack + threadIdx.x;
    uint32_t* spill = W.spill + gl;
    int64_t my = -1;                          // -1 wants work, >= 0 queue index
    uint32_t batch_base = 0, batch_left = 0;  // wave-uniform
    bool fresh = false;
    uint32_t slot = 0;
    gyd::Trav t;
    V3 Tr = v3s(1.f);
    int32_t med = -1;
    int k = 0;
    float tmax = gyd::kInf;
    while (true) {
        const bool need = my == -1;
        uint64_t nm = __ballot(need);
        while (nm) {
            if (batch_left == 0) {
                while (tries < kVwfChunks) {
                    const uint32_t lo = (uint32_t)(((uint64_t)n * ck) / kVwfChunks);
                    const uint32_t hi = (uint32_t)(((uint64_t)n * (ck + 1)) / kVwfChunks);
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(&fetch[ck * kVwfCtr], 64u);
                    base = __shfl(base, 0, 64);
                    if (base < hi - lo) {
                        batch_base = lo + base;
                        batch_left = min(64u, hi - batch_base);
                        break;
                    }
                    ck = ck + 1 == kVwfChunks ? 0 : ck + 1;
                    ++tries;
                }
                if (batch_left == 0) break;  // every region is drained
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
        if (my < 0) continue;
        if (fresh) {
            fresh = false;
            slot = W.q[qid][my];
            Tr = v3s(1.f);
            k = 0;
            if (MODE == 0) {
                gyd::trav_init(t, v3f(W.o[slot]), v3f(W.d[slot]), 0.f, gyd::kInf);
            } else if (MODE == 1) {
                const float4 so = W.sh_o[slot], sd = W.sh_d[slot];
                tmax = so.w;
                med = __float_as_int(sd.w);
                gyd::trav_init(t, v3f(so), v3f(sd), 0.f, tmax);
            } else {
                const float4 sd = W.mis_d[slot];
                med = __float_as_int(sd.w);
                gyd::trav_init(t, v3f(W.mis_o[slot]), v3f(sd), 0.f, gyd::kInf);
            }
        }
        if (gyd::trav_run<false, S>(Sc, t, lds, 256, spill, W.spill_stride, refill) == 0) continue;
        const bool hit = t.best_prim >= 0;
        bool more = false;  // the lane's segment loop goes on
        if (MODE == 0) {
            W.hit[slot] = hit ? make_float4(t.best, __int_as_float(t.best_prim), t.bb1, t.bb2)
                              : make_float4(0.f, __int_as_float(-1), 0.f, 0.f);
        } else if (MODE == 1) {
            if (hit && !no_material(M, t.best_prim)) {
                Tr = v3s(0.f);
            } else {
                if (med >= 0) Tr = Tr * medium_tr(load_medium(M, med), t.d, hit ? t.best : tmax);
                if (hit) {
                    if (++k >= kVptMaxCrossings) {
                        Tr = v3s(0.f);
                    } else {
                        Hit h;
                        h.t = t.best;
                        h.prim = t.best_prim;
                        h.b1 = t.bb1;
                        h.b2 = t.bb2;
                        const Isect it = gyd::query(Sc, t.o, t.d, h);
                        int32_t in, out;
                        hit_media(M, h.prim, med, &in, &out);
                        LightPoint lp;
                        lp.pos = v3f(W.lp_pos[slot]);
                        lp.fn = v3f(W.lp_fn[slot]);
                        lp.pErr = v3f(W.lp_err[slot]);
                        lp.n = v3s(0.f);
                        V3 o2, d2;
                        gyd::spawn_to(it, lp, &o2, &d2, &tmax);
                        med = isect_medium(it, in, out, d2);
                        gyd::trav_init(t, o2, d2, 0.f, tmax);
                        more = true;
                    }
                }
            }
            if (!more) W.tr1[slot] = f4v(Tr, 0.f);
        } else {
            if (med >= 0) Tr = Tr * medium_tr(load_medium(M, med), t.d, hit ? t.best : gyd::kInf);
            bool reached = false;  // IntersectTr found a surface with a material
            if (hit) {
                if (!no_material(M, t.best_prim)) {
                    reached = true;
                } else if (++k < kVptMaxCrossings) {
                    const float4* qv = Sc.shade + (size_t)t.best_prim * 8;
                    const V3 p0 = gyd::ld3(qv[0]), p1 = gyd::ld3(qv[1]), p2 = gyd::ld3(qv[2]);
                    const float b1 = t.bb1, b2 = t.bb2;
                    const V3 pos = ((1 - b1 - b2) * p0 + b1 * p1) + b2 * p2;
                    int32_t in, out;
                    hit_media(M, t.best_prim, med, &in, &out);
                    const V3 o2 = gyd::offset_origin(pos, v3s(0.f), v3s(0.f), t.d) + t.d * gyd::kRayEps;
                    med = in;
                    gyd::trav_init(t, o2, t.d, 0.f, gyd::kInf);
                    more = true;
                }
            }
            if (!more) {
                const int light = __float_as_int(W.pD[slot].w);
                const int ltri = Sc.light_tri[light];
                float4 e = make_float4(0.f, 0.f, 0.f, 0.f);
                if (reached) {
                    if (gyd::prim_light(Sc, t.best_prim) == light) {
                        Hit h;
                        h.t = t.best;
                        h.prim = t.best_prim;
                        h.b1 = t.bb1;
                        h.b2 = t.bb2;
                        const Isect li = gyd::query(Sc, t.o, t.d, h);
                        const V3 Le = v3(Sc.light_le[3 * light], Sc.light_le[3 * light + 1], Sc.light_le[3 * light + 2]);
                        e = f4v(gyd::dot(li.n, gyd::neg(t.d)) > 0.f ? Le : v3s(0.f), 1.f);
                    }
                } else if (ENV && ltri < 0) {
                    e = f4v(gyd::env_emission_call(Sc.env, t.d), 1.f);
                }
                W.e2[slot] = e;
                W.tr2[slot] = f4v(Tr, 0.f);
            }
        }
        if (!more) my = -1;
    }
}
template <int S, int MODE, bool ENV>
__global__ void __launch_bounds__(256) k_vwf_trace(SceneDev Sc, MediaDev M, VwfDev W, int refill) {
    __shared__ uint32_t stack[S * 256];
    vwf_trace_loop<S, MODE, ENV>(Sc, M, W, refill, stack);
}
// Shadow segments, then MIS segments, in one persistent launch: a wave that
// finds the shadow queue drained and its lanes idle goes on to the MIS queue
// while other waves finish their last shadow loops.  The two write different
// records of a slot.
template <int S, bool ENV>
__global__ void __launch_bounds__(256) k_vwf_trace_shadow_mis(SceneDev Sc, MediaDev M, VwfDev W, int refill) {
    __shared__ uint32_t stack[S * 256];
    vwf_trace_loop<S, 1, ENV>(Sc, M, W, refill, stack);
    vwf_trace_loop<S, 2, ENV>(Sc, M, W, refill, stack);
}

// Adds each vertex's light in vpt_path's operation order and sends ended
// paths to the film; returns whether the path goes on.
__device__ inline bool vwf_resolve_slot(const VwfDev& W, const FilmDev& F, uint32_t slot) {
    const uint32_t pk = W.pk[slot];
    if (!(pk & VK_LIVE)) return false;
    if (pk & (VK_DI | VK_CONN | VK_CONN0)) {
        const float4 pt = W.pT[slot];
        const V3 Tv = v3f(pt);
        const float4 Lv = W.L[slot];
        V3 L = v3f(Lv);
        if (pk & VK_DI) {
            const float4 c = W.pC[slot];
            V3 L1 = v3s(0.f), L2 = v3s(0.f);
            if (pk & VK_SH) {
                const float4 a = W.pA[slot], b = W.pB[slot];
                L1 = v3f(a) * v3f(W.tr1[slot]) * v3f(b) * a.w / b.w;
            }
            if (pk & VK_MIS) {
                const float4 e = W.e2[slot];
                if (e.w != 0.f) L2 = v3f(e) * v3f(W.pD[slot]) * v3f(W.tr2[slot]) * c.z / c.w;
            }
            L = L + Tv * ((L1 * c.x + L2 * c.y) / pt.w);
        } else {
            V3 conn = v3s(0.f);
            if (pk & VK_CONN) {
                const float4 a = W.pA[slot], b = W.pB[slot];
                conn = v3f(a) * v3f(W.tr1[slot]) * a.w / b.x / b.y;
            }
            L = L + Tv * conn / pt.w;
        }
        W.L[slot] = f4v(L, Lv.w);
    }
    if (pk & VK_DONE) {
        const float4 Lv = W.L[slot];
        int s_local, px, py;
        pass_pixel(F, slot, &s_local, &px, &py);
        record_sample(F, s_local, px, py, W.T[slot].w, Lv.w, v3f(Lv));
        W.pk[slot] = 0u;
        return false;
    }
    W.pk[slot] = VK_LIVE;
    return true;
}
// Walks the slots in order, four per lane (slot = (block * 4 + k) * 256 +
// lane), and forms the next active queue with one atomic per workgroup.
constexpr int kVwfResolvePer = 4;
__global__ void __launch_bounds__(256) k_vwf_resolve(VwfDev W, FilmDev F, int64_t n_slots) {
    uint32_t slot[kVwfResolvePer];
    bool next[1][kVwfResolvePer];
#pragma unroll
    for (int k = 0; k < kVwfResolvePer; ++k) {
        const int64_t t = ((int64_t)blockIdx.x * kVwfResolvePer + k) * 256 + threadIdx.x;
        slot[k] = (uint32_t)t;
        next[0][k] = t < n_slots && vwf_resolve_slot(W, F, slot[k]);
    }
    uint32_t* const qs[1] = {W.q[VQ_NEXT]};
    uint32_t* const cs[1] = {&W.counts[VQ_NEXT * kVwfCtr]};
    vwf_block_push<1, kVwfResolvePer>(qs, cs, slot, next);
}

// Next round: the next queue becomes the active one; returns its length.
__global__ void k_vwf_advance(VwfDev W, uint32_t* out, unsigned long long* rays) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        rays[0] += W.counts[VQ_ACTIVE * kVwfCtr];
        rays[1] += W.counts[VQ_SHADOW * kVwfCtr];
        rays[2] += W.counts[VQ_MIS * kVwfCtr];
        const uint32_t nx = W.counts[VQ_NEXT * kVwfCtr];
        for (int w = 0; w < kVwfWords; ++w) W.counts[w * kVwfCtr] = 0u;
        W.counts[VQ_ACTIVE * kVwfCtr] = nx;
        *out = nx;
    }
}
__global__ void k_vwf_reset(VwfDev W) {
    if (threadIdx.x < kVwfWords && blockIdx.x == 0) W.counts[threadIdx.x * kVwfCtr] = 0u;
}

MediaDev MediaView(const Scene* sc) {
    MediaDev M;
    M.med = (const float4*)sc->d_media;
    M.tri = (const uint32_t*)sc->d_tri_media;
    M.n = sc->n_media;
    return M;
}

template <int STACK, bool ENV>
void LaunchVptT(Ctx* c, Scene* sc, const PathParams& P, const FilmDev& F, int s0, int n_s, bool eqa, unsigned blocks) {
    const MediaDev M = MediaView(sc);
    if (eqa)
        hipLaunchKernelGGL((k_vpt<STACK, ENV, true>), dim3(blocks), dim3(128), 0, c->stream, sc->dev, M, P, F, s0, n_s);
    else
        hipLaunchKernelGGL((k_vpt<STACK, ENV, false>), dim3(blocks), dim3(128), 0, c->stream, sc->dev, M, P, F, s0, n_s);
}

int LaunchVpt(Ctx* c, Scene* sc, const PathParams& P, int s0, int n_s, bool eqa) {
    const FilmDev F = FilmViewState(sc->film);
    const int tilesX = (F.width + 7) >> 3, tilesY = (F.height + 7) >> 3;
    const int64_t threads = (int64_t)tilesX * tilesY * 64 * n_s;
    const unsigned blocks = (unsigned)((threads + 127) / 128);
    const bool env = sc->dev.env.w > 0;
    if (sc->stack <= 32) {
        if (env) LaunchVptT<32, true>(c, sc, P, F, s0, n_s, eqa, blocks);
        else LaunchVptT<32, false>(c, sc, P, F, s0, n_s, eqa, blocks);
    } else {
        if (env) LaunchVptT<64, true>(c, sc, P, F, s0, n_s, eqa, blocks);
        else LaunchVptT<64, false>(c, sc, P, F, s0, n_s, eqa, blocks);
    }
    return CheckLaunch(c, "k_vpt");
}

// ---------------------------------------------------------------- wavefront host side
constexpr int kVwfStack = 16;  // LDS ring entries per lane; deeper stacks spill
constexpr int kVwfRefill = 48;

static size_t VwfBytes(int64_t n) {
    const size_t per = 24 * 16 + 4 + 4 + VQ_COUNT * 4;  // 24 float4 / uint4 arrays, med, pk, queues
    return (size_t)n * per + 256 * (24 + 2 + VQ_COUNT + 2) + kVwfWords * kVwfCtr * 4 + 4096;
}
static VwfDev VwfView(const WavefrontState& wf, size_t* used) {
    VwfDev W;
    char* base = (char*)wf.mem;
    const int64_t n = wf.capacity;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        void* p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    float4** f4s[] = {&W.o, &W.d, &W.T, &W.L, &W.hit, &W.pT, &W.pA, &W.pB, &W.pC, &W.pD, &W.tr1, &W.e2, &W.tr2,
                      &W.sh_o, &W.sh_d, &W.lp_pos, &W.lp_fn, &W.lp_err, &W.mis_o, &W.mis_d};
    for (float4** p : f4s) *p = (float4*)take(n * 16);
    W.st = (uint4*)take(n * 16);
    W.med = (int32_t*)take(n * 4);
    W.pk = (uint32_t*)take(n * 4);
    for (int q = 0; q < VQ_COUNT; ++q) W.q[q] = (uint32_t*)take(n * 4);
    W.counts = (uint32_t*)take((size_t)(kVwfWords + 1) * kVwfCtr * 4);  // + the round's active count for the host
    W.capacity = n;
    W.spill = (uint32_t*)wf.spill;
    W.spill_stride = 0;
    *used = off;
    return W;
}

template <bool ENV>
static unsigned VwfPersistBlocks(Ctx* c) {
    int k[2] = {0, 0};
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&k[0], k_vwf_trace<kVwfStack, 0, ENV>, 256, 0);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&k[1], k_vwf_trace_shadow_mis<kVwfStack, ENV>, 256, 0);
    return (unsigned)(std::max(1, std::max(k[0], k[1])) * c->num_cu);
}

// One pass (samples s0 .. s0 + n_s - 1 of every pixel) of the wavefront form.
template <bool ENV, bool EQA>
static int VwfPass(Ctx* c, Scene* sc, const PathParams& P, const FilmDev& F, int s0, int n_s) {
    const int tilesX = (F.width + 7) >> 3, tilesY = (F.height + 7) >> 3;
    const int64_t n_slots = (int64_t)tilesX * tilesY * 64 * n_s;
    if (n_slots >= (1ll << 31)) return Fail(GY_ERR_UNSUPPORTED, "VolPathTracing wavefront pass larger than 2^31 paths");
    const unsigned pblocks = VwfPersistBlocks<ENV>(c);
    WavefrontState& wf = sc->vwf;
    if (wf.capacity < n_slots || !wf.mem) {
        WavefrontStateFree(&wf);
        int rc = DevAlloc(c, &wf.mem, VwfBytes(n_slots));
        if (rc != GY_OK) return rc;
        wf.bytes = VwfBytes(n_slots);
        wf.capacity = n_slots;
    }
    const size_t sbytes = (size_t)pblocks * 256 * 64 * 4;
    if (wf.spill_bytes < sbytes) {
        DevFree(wf.spill);
        wf.spill = nullptr;
        wf.spill_bytes = 0;
        int rc = DevAlloc(c, &wf.spill, sbytes);
        if (rc != GY_OK) return rc;
        wf.spill_bytes = sbytes;
    }
    size_t used = 0;
    VwfDev W = VwfView(wf, &used);
    if (used > wf.bytes) return Fail(GY_ERR_ALLOC, "VolPathTracing wavefront layout exceeds its allocation");
    W.spill_stride = (int64_t)pblocks * 256;
    uint32_t* d_n = W.counts + kVwfWords * kVwfCtr;
    const MediaDev M = MediaView(sc);
    const unsigned blocks = (unsigned)((n_slots + 255) / 256);
    const unsigned rblocks = (unsigned)((n_slots + 256 * kVwfResolvePer - 1) / (256 * kVwfResolvePer));
    hipLaunchKernelGGL(k_vwf_reset, dim3(1), dim3(64), 0, c->stream, W);
    TimerBegin(c, kKernRaygen);
    hipLaunchKernelGGL(k_vwf_start, dim3(blocks), dim3(256), 0, c->stream, W, P, F, s0, n_s, n_slots);
    TimerEnd(c);
    int rc = CheckLaunch(c, "k_vwf_start");
    if (rc != GY_OK) return rc;
    if (P.max_depth == 0) {
        hipLaunchKernelGGL(k_vwf_resolve, dim3(rblocks), dim3(256), 0, c->stream, W, F, n_slots);
        return CheckLaunch(c, "k_vwf_resolve");
    }
    // rounds until no path is left (a boundary crossing adds a round but no
    // bounce; a path makes at most max_depth bounces and, per loop, 256
    // crossings, and draws a bounded number of dimensions)
    const int max_rounds = P.max_depth + kVptMaxCrossings + kVptDimBudget + 2;
    for (int round = 0;; ++round) {
        if (round > max_rounds) return Fail(GY_ERR_DEVICE, "VolPathTracing wavefront: paths left after the last possible round");
        TimerBegin(c, kKernTraceClosest);
        hipLaunchKernelGGL((k_vwf_trace<kVwfStack, 0, ENV>), dim3(pblocks), dim3(256), 0, c->stream, sc->dev, M, W, kVwfRefill);
        TimerEnd(c);
        TimerBegin(c, kKernShade);
        hipLaunchKernelGGL((k_vwf_vertex<ENV, EQA>), dim3(blocks), dim3(256), 0, c->stream, sc->dev, M, P, W);
        TimerEnd(c);
        TimerBegin(c, kKernTraceShadow);
        hipLaunchKernelGGL((k_vwf_trace_shadow_mis<kVwfStack, ENV>), dim3(pblocks), dim3(256), 0, c->stream, sc->dev, M, W, kVwfRefill);
        TimerEnd(c);
        TimerBegin(c, kKernOther);
        hipLaunchKernelGGL(k_vwf_resolve, dim3(rblocks), dim3(256), 0, c->stream, W, F, n_slots);
        hipLaunchKernelGGL(k_vwf_advance, dim3(1), dim3(64), 0, c->stream, W, d_n, sc->film.counters);
        TimerEnd(c);
        rc = CheckLaunch(c, "VolPathTracing wavefront round");
        if (rc != GY_OK) return rc;
        uint32_t left = 0;
        hipError_t e = hipMemcpyAsync(&left, d_n, 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return HipFail(e, "VolPathTracing wavefront round");
        if (left == 0) return GY_OK;
        uint32_t* tq = W.q[VQ_ACTIVE];
        W.q[VQ_ACTIVE] = W.q[VQ_NEXT];
        W.q[VQ_NEXT] = tq;
    }
}

int LaunchVwf(Ctx* c, Scene* sc, const PathParams& P, int s0, int n_s, bool eqa) {
    const FilmDev F = FilmViewState(sc->film);
    const bool env = sc->dev.env.w > 0;
    if (env) return eqa ? VwfPass<true, true>(c, sc, P, F, s0, n_s) : VwfPass<true, false>(c, sc, P, F, s0, n_s);
    return eqa ? VwfPass<false, true>(c, sc, P, F, s0, n_s) : VwfPass<false, false>(c, sc, P, F, s0, n_s);
}

int ValidateVpt(Scene* sc, int32_t equiangular) {
    if (sc->dev.n_sph > 0) return Fail(GY_ERR_UNSUPPORTED, "VolPathTracing: scenes with spheres");
    // AreaLight::SampleOnePoint is the only SampleOnePoint the reference
    // defines; for an environment light it reads an uninitialised pdf
    if (equiangular && sc->dev.env.w > 0) return Fail(GY_ERR_UNSUPPORTED, "VolPathTracing: equi-angular sampling with an environment light");
    return GY_OK;
}

}  // namespace

int MakePathParamsFor(Ctx* c, Scene* sc, const gy_camera_desc& cam, const gy_render_params& p, PathParams* P, bool vpt);
int ValidateRender(const gy_camera_desc* cam, const gy_render_params* p);
int FilmPrepare(Ctx* c, Scene* sc, int W, int H, int n_pass);
int FilmResolvePass(Ctx* c, Scene* sc, int n_s);
int SamplesPerPass(const gy_camera_desc& cam, const gy_render_params& p);
void TimersCollect(Ctx* c);

}  // namespace gy

using namespace gy;

// Scene::AddPrimitive(shape, material, MediumInterface) with HomogeneousMedium
// media (volume/Homogeneous.h:11-12) and material-less primitives.
extern "C" int gy_scene_set_media(gy_scene* s, int32_t n_media, const gy_medium* media, const gy_tri_medium* tri) {
    if (!s || n_media < 0 || (n_media > 0 && !media) || !tri) return Fail(GY_ERR_INVALID, "gy_scene_set_media: bad argument");
    Scene* sc = reinterpret_cast<Scene*>(s);
    if (n_media > 32767) return Fail(GY_ERR_UNSUPPORTED, "gy_scene_set_media: at most 32767 media");
    if (sc->batch_refs > 0) return Fail(GY_ERR_INVALID, "gy_scene_set_media: a batch holds the scene");
    std::vector<float4> med(2 * (size_t)std::max(n_media, 1), make_float4(0.f, 0.f, 0.f, 0.f));
    for (int32_t i = 0; i < n_media; ++i) {
        const gy_medium& m = media[i];
        for (int k = 0; k < 3; ++k)
            if (!(m.sigma_a[k] >= 0.f) || !(m.sigma_s[k] >= 0.f) || !std::isfinite(m.sigma_a[k]) || !std::isfinite(m.sigma_s[k]))
                return Fail(GY_ERR_INVALID, "gy_scene_set_media: sigma_a and sigma_s must be finite and >= 0");
        if (!(m.g > -1.f && m.g < 1.f)) return Fail(GY_ERR_INVALID, "gy_scene_set_media: g must lie in (-1, 1)");
        med[2 * i] = make_float4(m.sigma_a[0], m.sigma_a[1], m.sigma_a[2], m.g);
        med[2 * i + 1] = make_float4(m.sigma_s[0], m.sigma_s[1], m.sigma_s[2], 0.f);
    }
    std::vector<uint32_t> words((size_t)sc->n_tri);
    bool no_mat = false;
    for (int32_t i = 0; i < sc->n_tri; ++i) {
        const gy_tri_medium& t = tri[i];
        if (t.inside < -1 || t.inside >= n_media || t.outside < -1 || t.outside >= n_media)
            return Fail(GY_ERR_INVALID, "gy_scene_set_media: medium index out of range");
        words[i] = (uint32_t)(t.inside + 1) | ((uint32_t)(t.outside + 1) << 16) | (t.no_material ? 0x80000000u : 0u);
        no_mat = no_mat || t.no_material;
    }
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    void *dm = nullptr, *dt = nullptr;
    int rc = Upload(c, &dm, med.data(), med.size() * sizeof(float4));
    if (rc == GY_OK) rc = Upload(c, &dt, words.data(), std::max<size_t>(words.size(), 1) * sizeof(uint32_t));
    if (rc != GY_OK) {
        DevFree(dm);
        DevFree(dt);
        return rc;
    }
    (void)hipStreamSynchronize(c->stream);
    DevFree(sc->d_media);
    DevFree(sc->d_tri_media);
    sc->d_media = dm;
    sc->d_tri_media = dt;
    sc->n_media = n_media;
    sc->has_no_mat = no_mat;
    ++sc->revision;
    sc->film_valid = false;
    return GY_OK;
}

// VolPathTracing(spp, maxDepth, sampler, samplerEnum, equiAngular)
// (integrator/VPT.h:12-19) through TiledIntegrator::Render: the film as gy_render's.
extern "C" int gy_render_vpt(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p, int32_t equiangular, float* color,
                             float* weight) {
    if (!s) return Fail(GY_ERR_INVALID, "gy_render_vpt: null scene");
    int rc = ValidateRender(cam, p);
    if (rc != GY_OK) return rc;
    Scene* sc = reinterpret_cast<Scene*>(s);
    rc = ValidateVpt(sc, equiangular);
    if (rc != GY_OK) return rc;
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    PathParams P;
    rc = MakePathParamsFor(c, sc, *cam, *p, &P, true);
    if (rc != GY_OK) return rc;
    const int n_pass = SamplesPerPass(*cam, *p);
    rc = FilmPrepare(c, sc, cam->width, cam->height, n_pass);
    if (rc != GY_OK) return rc;
    std::memset(&c->stats, 0, sizeof(c->stats));
    c->ev_used = 0;
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    (void)hipEventRecord(t0, c->stream);
    const bool wavefront = p->mode == GY_RENDER_WAVEFRONT;
    for (int s0 = 0; s0 < p->spp && rc == GY_OK; s0 += n_pass) {
        const int n_s = std::min(n_pass, p->spp - s0);
        rc = wavefront ? LaunchVwf(c, sc, P, s0, n_s, equiangular != 0) : LaunchVpt(c, sc, P, s0, n_s, equiangular != 0);
        if (rc == GY_OK) rc = FilmResolvePass(c, sc, n_s);  // timed as kKernFilm inside
    }
    (void)hipEventRecord(t1, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "render_vpt sync");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    TimersCollect(c);
    c->stats.total_ms = ms;
    c->stats.paths = (int64_t)cam->width * cam->height * p->spp;
    if (rc == GY_OK && wavefront) {  // path, shadow and MIS segments started (crossings not counted)
        unsigned long long h[3] = {0, 0, 0};
        e = hipMemcpy(h, sc->film.counters, sizeof(h), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = HipFail(e, "render_vpt counters");
        c->stats.rays = (int64_t)(h[0] + h[1] + h[2]);
        c->stats.shadow_rays = (int64_t)h[1];
        c->stats.mis_rays = (int64_t)h[2];
    }
    sc->film_valid = false;  // the adjoint replays PathTracing renders only
    if (rc == GY_OK) rc = FilmCheck(sc->film);
    if (rc == GY_OK && (color || weight)) rc = gy_film_download(s, color, weight);
    return rc;
}

// Per-sample VolPathTracing::Li on given (x, y, sample) indices: radiance,
// film position and dimensions drawn (the parity tests' entry point).
extern "C" int gy_vpt_li_samples(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p, int32_t equiangular, int32_t n,
                                 const int32_t* x, const int32_t* y, const int32_t* smp, float* L, float* raster, int32_t* dims) {
    if (!s || n < 0 || (n > 0 && (!x || !y || !smp || !L || !raster || !dims)))
        return Fail(GY_ERR_INVALID, "gy_vpt_li_samples: bad argument");
    int rc = ValidateRender(cam, p);
    if (rc != GY_OK) return rc;
    Scene* sc = reinterpret_cast<Scene*>(s);
    rc = ValidateVpt(sc, equiangular);
    if (rc != GY_OK) return rc;
    if (n == 0) return GY_OK;
    for (int32_t i = 0; i < n; ++i)
        if (x[i] < 0 || x[i] >= cam->width || y[i] < 0 || y[i] >= cam->height || smp[i] < 0 || smp[i] >= p->spp)
            return Fail(GY_ERR_INVALID, "gy_vpt_li_samples: sample index outside the image or spp");
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    PathParams P;
    rc = MakePathParamsFor(c, sc, *cam, *p, &P, true);
    if (rc != GY_OK) return rc;
    void *dx = nullptr, *dy = nullptr, *ds = nullptr, *dL = nullptr, *dr = nullptr, *dd = nullptr;
    rc = Upload(c, &dx, x, (size_t)n * 4);
    if (rc == GY_OK) rc = Upload(c, &dy, y, (size_t)n * 4);
    if (rc == GY_OK) rc = Upload(c, &ds, smp, (size_t)n * 4);
    if (rc == GY_OK) rc = DevAlloc(c, &dL, (size_t)n * 12);
    if (rc == GY_OK) rc = DevAlloc(c, &dr, (size_t)n * 8);
    if (rc == GY_OK) rc = DevAlloc(c, &dd, (size_t)n * 4);
    if (rc == GY_OK) {
        const unsigned blocks = (unsigned)((n + 127) / 128);
        const bool env = sc->dev.env.w > 0, eqa = equiangular != 0;
        auto pick = [&](auto k32, auto k64) { return sc->stack <= 32 ? k32 : k64; };
        decltype(&k_vpt_li_samples<32, false, false>) kern;
        if (env) kern = eqa ? pick(k_vpt_li_samples<32, true, true>, k_vpt_li_samples<64, true, true>)
                            : pick(k_vpt_li_samples<32, true, false>, k_vpt_li_samples<64, true, false>);
        else kern = eqa ? pick(k_vpt_li_samples<32, false, true>, k_vpt_li_samples<64, false, true>)
                        : pick(k_vpt_li_samples<32, false, false>, k_vpt_li_samples<64, false, false>);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(128), 0, c->stream, sc->dev, MediaView(sc), P, n, (const int32_t*)dx,
                           (const int32_t*)dy, (const int32_t*)ds, (float*)dL, (float*)dr, (int32_t*)dd);
        rc = CheckLaunch(c, "k_vpt_li_samples");
    }
    if (rc == GY_OK) {
        hipError_t e = hipMemcpyAsync(L, dL, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(raster, dr, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dims, dd, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = HipFail(e, "vpt_li_samples D2H");
    }
    for (void* b : {dx, dy, ds, dL, dr, dd}) DevFree(b);
    return rc;
}

// Path-replay adjoint of the path tracer: gradients of an image loss with
// respect to texture texels (albedo of Lambert and GGX conductor materials,
// GGX roughness), by re-tracing every camera path with its own sample index.
//
// The render is the estimator of PathTracing::Li (integrator/PT.cpp:10-67)
// with DirectIllumination (integrator/Integrator.cpp:9-107).  For a path with
// vertices j, throughput T_j = prod_{k<j} f_k |cos_k| / (pdf_k p_rr,k) and
// radiance L = sum_j T_j (Le_j + DI_j), a parameter theta that scales the BSDF
// of vertex j changes every term that passes through that vertex:
//
//   dL/dtheta = (df/dtheta / f)(w_light) T_j L1_j w1_j / P_j
//             + (df/dtheta / f)(w_mis)   T_j L2_j w2_j / P_j
//             + (df/dtheta / f)(w_next)  (L - L_{<=j})
//
// with sampling pdfs, MIS weights and roulette probabilities held fixed (each
// is a valid choice of estimator for any value it takes, so the estimator of
// the derivative stays unbiased).  L - L_{<=j}, the radiance gathered after
// vertex j, comes from a first trace of the same path (memory per path: none
// beyond the path itself).  The image loss enters through dl/dI per pixel; a
// sample's share of pixel q is 1 / weight_q of the box filter
// (visual/Film.h:61-87), so its adjoint radiance is sum_q dl/dI_q / weight_q
// over its footprint.
//
//   albedo R (Lambert f = R/pi, conductor f proportional to R), texel i of
//     channel c through bilinear weight b_i:  (df/dtheta)/f = b_i / R_c;
//   GGX roughness (isotropic, alpha = texel value or RoughnessToAlpha of it):
//     (df/dtheta)/f = b_i dalpha/dr (dlog D/dalpha + dlog G1(wo)/dalpha
//                                    + dlog G1(wi)/dalpha).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "integrator.h"

namespace gy {
namespace {

// Bilinear taps of tex_bilinear (texture/ImageTexture.h:50-69): offsets of the
// four texels (channel 0) and their weights; clamped taps may coincide.
__device__ inline void tex_taps(const SceneDev& S, int id, float u, float v, int64_t off[4], float w[4], int* channels) {
    const int4 ti = S.tex_info[id];
    const float s = ti.x * u - 0.5f, t = ti.y * v - 0.5f;
    const int s0 = (int)floorf(s), t0 = (int)floorf(t);
    const float ds = s - s0, dt = t - t0;
    auto at = [&](int x, int y) {
        x = x < 0 ? 0 : (x > ti.x - 1 ? ti.x - 1 : x);
        y = y < 0 ? 0 : (y > ti.y - 1 ? ti.y - 1 : y);
        return (int64_t)ti.w + ((int64_t)y * ti.x + x) * ti.z;
    };
    off[0] = at(s0, t0);
    off[1] = at(s0 + 1, t0);
    off[2] = at(s0, t0 + 1);
    off[3] = at(s0 + 1, t0 + 1);
    w[0] = (1 - ds) * (1 - dt);
    w[1] = ds * (1 - dt);
    w[2] = (1 - ds) * dt;
    w[3] = ds * dt;
    *channels = ti.z;
}

// d log f / d alpha of the isotropic GGX reflection lobe (visual/Microfacet.cpp:
// 232-263) for local directions wo, wi; 0 where the lobe is zero.  Plain float
// arithmetic so the host can evaluate it too.
__host__ __device__ inline float ggx_dlogf_dalpha(float a, float ox, float oy, float oz, float ix, float iy, float iz) {
    if (!(oz * iz > 0.f)) return 0.f;
    float hx = ix + ox, hy = iy + oy, hz = iz + oz;
    const float hl = sqrtf(hx * hx + hy * hy + hz * hz);
    if (!(hl > 0.f)) return 0.f;
    hx /= hl;
    hy /= hl;
    hz /= hl;
    if (hz < 0.f) {
        hx = -hx;
        hy = -hy;
        hz = -hz;
    }
    if ((ox * hx + oy * hy + oz * hz) * oz <= 0.f || (ix * hx + iy * hy + iz * hz) * iz <= 0.f || hz <= 0.f) return 0.f;
    const float a2 = a * a;
    const float c2 = hz * hz;
    const float t = (1.f - c2) / c2;  // tan^2 of the half vector
    float r = 2.f * (t - a2) / (a * (a2 + t));
    const float zs[2] = {oz, iz};
    for (int k = 0; k < 2; ++k) {
        const float z2 = zs[k] * zs[k];
        const float tn2 = fmaxf(0.f, 1.f - z2) / z2;
        const float sq = sqrtf(1.f + a2 * tn2);
        r -= a * tn2 / (sq * (1.f + sq));
    }
    return r;
}
__device__ inline float ggx_dlogf_dalpha(float a, V3 wo, V3 wi) { return ggx_dlogf_dalpha(a, wo.x, wo.y, wo.z, wi.x, wi.y, wi.z); }

// d alpha / d r of RoughnessToAlpha (visual/Microfacet.h:49-54).
__host__ __device__ inline float dalpha_droughness(float r) {
    if (r < 1e-3f) return 0.f;
    const float x = logf(r);
    return (0.819955f + 2.f * 0.1734f * x + 3.f * 0.0171201f * x * x + 4.f * 0.000640711f * x * x * x) / r;
}

struct Grad {
    float* __restrict__ g;    // one float per texture float
    float* __restrict__ gle;  // n_light * 3 floats (emitted radiance per light), or null
};

// One camera sample replayed: the same random numbers, traversals and
// branches as trace_path, scattering gradients at every vertex.
template <int STACK>
__device__ void replay_path(const SceneDev& S, const PathParams& P, int px, int py, int s, uint32_t* stack, int stride,
                            V3 Ltot, V3 dL, Grad G) {
    PathCounters pc = {};
    const SampleSrc& M = P.smp;
    Seq q;
    q.index = enum_index(P.en, (uint32_t)s, (uint32_t)px, (uint32_t)py);
    q.cursor = 0;
    const float u = P.en.sx * seq_next(M, q) - (float)px;
    const float v = P.en.sy * seq_next(M, q) - (float)py;
    V3 o, d;
    generate_ray(P.cam, px, py, u, v, &o, &d);
    bool delta = false;
    V3 T = v3(1.f, 1.f, 1.f);
    V3 L = v3s(0.f);
    for (int i = 0; i < P.max_depth; ++i) {
        Hit h;
        if (!find_hit<false, STACK>(S, o, d, 0.f, kInf, &h, stack, stride)) break;
        const Isect it = query(S, o, d, h);
        const BSDF b = make_bsdf(S, it);
        const uint32_t* mw = S.mats + (size_t)it.mat * 32;
        const int32_t ta = (int32_t)mw[21], tr = (int32_t)mw[22];
        const bool g_alb = ta >= 0 && (b.type == MAT_LAMBERT || b.type == MAT_CONDUCTOR);
        const bool g_rough = tr >= 0 && b.type == MAT_CONDUCTOR;
        V3 Tdi = v3s(0.f);  // this vertex's direct light, T_j DI_j
        DIParts dp;
        bool has_di = false;
        if ((i == 0 || delta) && it.light >= 0) {
            const V3 Le = v3(S.light_le[3 * it.light], S.light_le[3 * it.light + 1], S.light_le[3 * it.light + 2]);
            const bool facing = dot(it.n, it.outDir) > 0.f;
            const V3 e = facing ? Le : v3(0.f, 0.f, 0.f);
            L = L + T * e;
            if (G.gle && facing) {
                // d(T Le) / d Le = T
                const V3 ge = dL * T;
                atomicAdd(&G.gle[3 * it.light], ge.x);
                atomicAdd(&G.gle[3 * it.light + 1], ge.y);
                atomicAdd(&G.gle[3 * it.light + 2], ge.z);
            }
        } else {
            const float uL = seq_next(M, q);
            const float u0 = seq_next(M, q), u1 = seq_next(M, q);
            const float v0 = seq_next(M, q), v1 = seq_next(M, q), v2 = seq_next(M, q);
            dp = di_parts<STACK, false>(S, it, b, uL, u0, u1, v0, v1, v2, stack, stride, pc);
            const V3 di = (dp.L1 * dp.w1 + dp.L2 * dp.w2) / dp.lsp;
            Tdi = T * di;
            L = L + Tdi;
            has_di = true;
            if (G.gle && !(is_zero(dp.F1) && is_zero(dp.F2))) {
                // the light sample and the MIS sample are linear in the chosen
                // light's Le; P(light) is a sampling probability, held fixed
                const V3 ge = dL * T * (dp.F1 * dp.w1 + dp.F2 * dp.w2) / dp.lsp;
                atomicAdd(&G.gle[3 * dp.light], ge.x);
                atomicAdd(&G.gle[3 * dp.light + 1], ge.y);
                atomicAdd(&G.gle[3 * dp.light + 2], ge.z);
            }
        }
        const float r0 = seq_next(M, q), r1 = seq_next(M, q), r2 = seq_next(M, q);
        V3 wi;
        float pdf;
        const V3 f = bsdf_sample(b, neg(d), &wi, &pdf, r0, r1, r2);
        const bool cont = !(is_zero(f) || pdf == 0.f);
        // radiance gathered after this vertex (through its continuation)
        const V3 Lrem = cont ? Ltot - L : v3s(0.f);
        if (g_alb && (has_di || cont)) {
            const V3 share = dL * (Tdi + Lrem);
            int64_t off[4];
            float bw[4];
            int ch;
            tex_taps(S, ta, it.u, it.v, off, bw, &ch);
            const float sc[3] = {b.Rc.x != 0.f ? share.x / b.Rc.x : 0.f, b.Rc.y != 0.f ? share.y / b.Rc.y : 0.f,
                                 b.Rc.z != 0.f ? share.z / b.Rc.z : 0.f};
            for (int k = 0; k < 4; ++k) {
                if (bw[k] == 0.f) continue;
                if (ch >= 3) {
                    for (int c = 0; c < 3; ++c) atomicAdd(&G.g[off[k] + c], bw[k] * sc[c]);
                } else {
                    // fewer channels: the last channel stands for the remaining ones
                    for (int c = 0; c < 3; ++c) atomicAdd(&G.g[off[k] + (c < ch ? c : ch - 1)], bw[k] * sc[c]);
                }
            }
        }
        if (g_rough && (has_di || cont)) {
            const V3 wo = to_local(b, it.outDir);
            const float a = b.dist.ax;
            float acc = 0.f;
            if (has_di) {
                const V3 t1 = dL * T * dp.L1 * dp.w1 / dp.lsp;
                const V3 t2 = dL * T * dp.L2 * dp.w2 / dp.lsp;
                if (!is_zero(t1)) acc += (t1.x + t1.y + t1.z) * ggx_dlogf_dalpha(a, wo, to_local(b, dp.wl));
                if (!is_zero(t2)) acc += (t2.x + t2.y + t2.z) * ggx_dlogf_dalpha(a, wo, to_local(b, dp.wb));
            }
            if (cont) {
                const V3 t3 = dL * Lrem;
                acc += (t3.x + t3.y + t3.z) * ggx_dlogf_dalpha(a, wo, to_local(b, wi));
            }
            if (acc != 0.f) {
                const uint32_t flags = mw[1];
                const float r = tex_bilinear(S, tr, it.u, it.v, 0);
                const float dadr = (flags & 1u) ? 1.f : dalpha_droughness(r);
                int64_t off[4];
                float bw[4];
                int ch;
                tex_taps(S, tr, it.u, it.v, off, bw, &ch);
                for (int k = 0; k < 4; ++k)
                    if (bw[k] != 0.f) atomicAdd(&G.g[off[k]], bw[k] * dadr * acc);
            }
        }
        if (!cont) break;
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
}

// One camera sample (pass sample s of pixel px, py) of a render on film F:
// its radiance, its adjoint radiance from dl/dI over its box-filter
// footprint, then the replay.  STORED: the render left every sample's
// radiance in its pass slot (one pass covered all samples), so the first
// trace is skipped.
template <int STACK, int BLOCK, bool STORED>
__device__ inline void adjoint_sample(const SceneDev& S, const PathParams& P, const FilmDev& F, const float* __restrict__ dI,
                                      const Grad& G, int s, int px, int py, uint32_t* stack) {
    float fx, fy;
    V3 Lt;
    if (STORED) {
        Seq q;
        q.index = enum_index(P.en, (uint32_t)s, (uint32_t)px, (uint32_t)py);
        q.cursor = 0;
        const float u = P.en.sx * seq_next(P.smp, q) - (float)px;
        const float v = P.en.sy * seq_next(P.smp, q) - (float)py;
        fx = px + u;
        fy = py + v;
        const float4 e = F.pass[((int64_t)s * F.height + py) * F.width + px];
        Lt = v3(e.x, e.y, e.z);
    } else {
        PathCounters pc = {};
        Lt = trace_path<STACK, false>(S, P, px, py, s, stack, BLOCK, &fx, &fy, nullptr, pc);
    }
    int x0, x1, y0, y1;
    film_footprint(fx, fy, F.width, F.height, &x0, &x1, &y0, &y1);
    V3 dL = v3s(0.f);
    for (int yy = y0; yy <= y1; ++yy)
        for (int xx = x0; xx <= x1; ++xx) {
            const int64_t q = (int64_t)yy * F.width + xx;
            const float w = F.film[q].w;
            if (w > 0.f) dL = dL + v3(dI[3 * q], dI[3 * q + 1], dI[3 * q + 2]) / w;
        }
    if (is_zero(dL)) return;
    replay_path<STACK>(S, P, px, py, s, stack, BLOCK, Lt, dL, G);
}

// One thread per camera sample of the render on the film (same mapping as the
// megakernel).
template <int STACK, int BLOCK, bool STORED>
__global__ void __launch_bounds__(BLOCK) k_pt_adjoint(SceneDev S, PathParams P, FilmDev F, const float* __restrict__ dI,
                                                      Grad G, int s0, int n_s) {
    __shared__ uint32_t stack[STACK * BLOCK];
    const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    int s_local, px, py;
    pass_pixel(F, t, &s_local, &px, &py);
    if (!(s_local < n_s && px < F.width && py < F.height)) return;
    adjoint_sample<STACK, BLOCK, STORED>(S, P, F, dI, G, s0 + s_local, px, py, stack + threadIdx.x);
}

// Every sample of every entry of a batch render in one launch.  Each entry owns
// whole workgroups (from E[e].adj_block0), so the entry, its camera and film
// are the same for the whole workgroup and their fields load as scalars; the
// scene, which every entry shares, is an argument.  Within an entry the lanes
// follow its sample-major 8x8-tile order, as k_wf_raygen_batch deals them.  A
// batch entry's pass buffer holds all its samples, so the radiance is always
// stored.  Entry e's dl/dI starts at dI + E[e].image_off.  Five waves per
// SIMD (96 VGPRs, some spilled) in place of three (165): the replay at scale
// takes 2.62 s per gradient instead of 3.10.
template <int STACK, int BLOCK>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5))) k_batch_adjoint(SceneDev S, const BatchEntryDev* __restrict__ E, int32_t n,
                                                         const float* __restrict__ dI, Grad G, int64_t block_base) {
    __shared__ uint32_t stack[STACK * BLOCK];
    const int64_t blk = block_base + blockIdx.x;
    int lo = 0, hi = n - 1;  // last entry with adj_block0 <= blk
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (E[mid].adj_block0 <= blk) lo = mid;
        else hi = mid - 1;
    }
    const BatchEntryDev& B = E[lo];
    const int64_t r = (blk - B.adj_block0) * BLOCK + threadIdx.x;
    if (r >= B.paths) return;
    const int64_t per_s = (int64_t)B.tilesX * B.tilesY * 64;
    const int s = (int)(r / per_s);
    const int64_t rr = r - (int64_t)s * per_s;
    const int64_t tile = rr >> 6;
    const int local = (int)(rr & 63);
    const int px = (int)(tile % B.tilesX) * 8 + (local & 7);
    const int py = (int)(tile / B.tilesX) * 8 + (local >> 3);
    if (px >= B.P.cam.width || py >= B.P.cam.height) return;
    adjoint_sample<STACK, BLOCK, true>(S, B.P, B.F, dI + B.image_off, G, s, px, py, stack + threadIdx.x);
}

// Pixel values colour / weight of the film (Film::WriteToPixelBuffer without
// splats, visual/Film.h:128-136).
__global__ void k_film_image(FilmDev F, float* __restrict__ out) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)F.width * F.height) return;
    const float4 a = F.film[p];
    out[3 * p] = a.x / a.w;
    out[3 * p + 1] = a.y / a.w;
    out[3 * p + 2] = a.z / a.w;
}

// Every entry's image colour / weight, concatenated (entry e at image_off).
__global__ void k_batch_images(const BatchEntryDev* __restrict__ E, int32_t e, float* __restrict__ out) {
    const BatchEntryDev& B = E[e];
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)B.F.width * B.F.height) return;
    const float4 a = B.F.film[p];
    float* o = out + B.image_off;
    o[3 * p] = a.x / a.w;
    o[3 * p + 1] = a.y / a.w;
    o[3 * p + 2] = a.z / a.w;
}

__global__ void k_zero_f32(float* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

}  // namespace

int MakePathParams(Ctx* c, Scene* sc, const gy_camera_desc& cam, const gy_render_params& p, PathParams* P);
int ValidateRender(const gy_camera_desc* cam, const gy_render_params* p);
int SamplesPerPass(const gy_camera_desc& cam, const gy_render_params& p);

// Adds the gradient of the loss to grad (device, one float per texture float)
// and, when gle is not null, to gle (device, three floats per area light).
static int AdjointDevice(Scene* sc, const gy_camera_desc* cam, const gy_render_params* p, const float* dI, float* grad, float* gle) {
    Ctx* c = sc->ctx;
    int rc = ValidateRender(cam, p);
    if (rc != GY_OK) return rc;
    if (!sc->film_valid || std::memcmp(&sc->film_cam, cam, sizeof(*cam)) != 0 || sc->film_params.spp != p->spp ||
        sc->film_params.max_depth != p->max_depth || sc->film_params.rr_depth != p->rr_depth ||
        sc->film_params.sampler != p->sampler)
        return Fail(GY_ERR_INVALID, "gy_adjoint: render this camera and these parameters with gy_render first");
    if (sc->tex_floats == 0) return Fail(GY_ERR_INVALID, "gy_adjoint: the scene has no textures");
    if (sc->dev.env.w > 0) return Fail(GY_ERR_UNSUPPORTED, "gy_adjoint: scenes with an environment light");
    if (sc->dev.n_sph > 0) return Fail(GY_ERR_UNSUPPORTED, "gy_adjoint: scenes with spheres");
    if (sc->stats.max_depth + 1 > 64) return Fail(GY_ERR_UNSUPPORTED, "gy_adjoint: BVH deeper than the traversal stack");
    PathParams P;
    rc = MakePathParams(c, sc, *cam, *p, &P);
    if (rc != GY_OK) return rc;
    const FilmDev F = FilmView(sc);
    const int n_pass = std::max(1, std::min(p->spp, SamplesPerPass(*cam, *p)));
    // the render's pass buffer still holds every sample's radiance when one
    // pass covered all samples
    const bool stored = sc->film_single_pass && sc->film.pass_slots >= (int64_t)F.width * F.height * p->spp;
    constexpr int kBlock = 128;
    const int tilesX = (F.width + 7) >> 3, tilesY = (F.height + 7) >> 3;
    TimerBegin(c, kKernOther);
    for (int s0 = 0; s0 < p->spp && rc == GY_OK; s0 += n_pass) {
        const int n_s = std::min(n_pass, p->spp - s0);
        const int64_t threads = (int64_t)tilesX * tilesY * 64 * n_s;
        const unsigned blocks = (unsigned)((threads + kBlock - 1) / kBlock);
        const Grad G{grad, gle};
        if (sc->stack <= 32) {
            if (stored) hipLaunchKernelGGL((k_pt_adjoint<32, kBlock, true>), dim3(blocks), dim3(kBlock), 0, c->stream, sc->dev, P, F, dI, G, s0, n_s);
            else hipLaunchKernelGGL((k_pt_adjoint<32, kBlock, false>), dim3(blocks), dim3(kBlock), 0, c->stream, sc->dev, P, F, dI, G, s0, n_s);
        } else {
            if (stored) hipLaunchKernelGGL((k_pt_adjoint<64, kBlock, true>), dim3(blocks), dim3(kBlock), 0, c->stream, sc->dev, P, F, dI, G, s0, n_s);
            else hipLaunchKernelGGL((k_pt_adjoint<64, kBlock, false>), dim3(blocks), dim3(kBlock), 0, c->stream, sc->dev, P, F, dI, G, s0, n_s);
        }
        rc = CheckLaunch(c, "k_pt_adjoint");
    }
    TimerEnd(c);
    return rc;
}

int BatchAdjoint(Ctx* c, Batch* b, const float* dI, float* grad, float* gle) {
    if (b->entries.empty() || b->dirty || !b->rendered) return Fail(GY_ERR_INVALID, "gy_batch_adjoint: render the batch first");
    Scene* sc = b->entries[0].scene;
    for (const BatchEntry& en : b->entries) {
        if (en.scene != sc) return Fail(GY_ERR_UNSUPPORTED, "gy_batch_adjoint: every entry must render the same scene");
        if (en.tex_revision != sc->tex_revision)
            return Fail(GY_ERR_INVALID, "gy_batch_adjoint: the textures changed since the batch render");
    }
    if (sc->tex_floats == 0) return Fail(GY_ERR_INVALID, "gy_batch_adjoint: the scene has no textures");
    if (sc->dev.env.w > 0) return Fail(GY_ERR_UNSUPPORTED, "gy_batch_adjoint: scenes with an environment light");
    if (sc->stats.max_depth + 1 > 64) return Fail(GY_ERR_UNSUPPORTED, "gy_batch_adjoint: BVH deeper than the traversal stack");
    constexpr int kBlock = kBatchAdjointBlock;
    const Grad G{grad, gle};
    // launches of at most 2^23 workgroups (2^30 samples)
    constexpr int64_t kChunk = (int64_t)1 << 23;
    int rc = GY_OK;
    TimerBegin(c, kKernOther);
    for (int64_t base = 0; base < b->adj_blocks && rc == GY_OK; base += kChunk) {
        const unsigned blocks = (unsigned)std::min(kChunk, b->adj_blocks - base);
        if (sc->stack <= 32)
            hipLaunchKernelGGL((k_batch_adjoint<32, kBlock>), dim3(blocks), dim3(kBlock), 0, c->stream, sc->dev,
                               (const BatchEntryDev*)b->table, (int32_t)b->entries.size(), dI, G, base);
        else
            hipLaunchKernelGGL((k_batch_adjoint<64, kBlock>), dim3(blocks), dim3(kBlock), 0, c->stream, sc->dev,
                               (const BatchEntryDev*)b->table, (int32_t)b->entries.size(), dI, G, base);
        rc = CheckLaunch(c, "k_batch_adjoint");
    }
    TimerEnd(c);
    return rc;
}

int BatchImages(Ctx* c, Batch* b, float* images) {
    if (b->entries.empty() || b->dirty || !b->rendered) return Fail(GY_ERR_INVALID, "gy_batch_images: render the batch first");
    for (size_t e = 0; e < b->entries.size(); ++e) {
        const int64_t n = (int64_t)b->entries[e].cam.width * b->entries[e].cam.height;
        hipLaunchKernelGGL(k_batch_images, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (const BatchEntryDev*)b->table,
                           (int32_t)e, images);
    }
    int rc = CheckLaunch(c, "k_batch_images");
    hipError_t e = hipStreamSynchronize(c->stream);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "batch images sync");
    return rc;
}

}  // namespace gy

using namespace gy;

extern "C" int gy_adjoint_emission_device(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p,
                                          const float* dimage_device, float* grad_device, float* grad_le_device) {
    if (!s || !dimage_device || !grad_device) return Fail(GY_ERR_INVALID, "gy_adjoint_device: null argument");
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    c->ev_used = 0;
    int rc = AdjointDevice(sc, cam, p, dimage_device, grad_device, grad_le_device);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "adjoint sync");
    TimersCollect(c);
    return rc;
}

extern "C" int gy_adjoint_device(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p, const float* dimage_device,
                                 float* grad_device) {
    return gy_adjoint_emission_device(s, cam, p, dimage_device, grad_device, nullptr);
}

extern "C" int gy_adjoint_emission(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p, const float* dimage,
                                   float* grad, float* grad_le) {
    if (!s || !cam || !dimage || !grad) return Fail(GY_ERR_INVALID, "gy_adjoint: null argument");
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    const int64_t npix = (int64_t)cam->width * cam->height;
    const int64_t n_le = (int64_t)sc->area_tri.size() * 3;
    void *dI = nullptr, *dg = nullptr, *dl = nullptr;
    int rc = DevAlloc(c, &dI, (size_t)npix * 3 * 4);
    if (rc == GY_OK) rc = DevAlloc(c, &dg, (size_t)std::max<int64_t>(sc->tex_floats, 1) * 4);
    if (rc == GY_OK && grad_le) rc = DevAlloc(c, &dl, (size_t)std::max<int64_t>(n_le, 1) * 4);
    if (rc == GY_OK) {
        hipError_t e = hipMemcpyAsync(dI, dimage, (size_t)npix * 3 * 4, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) rc = HipFail(e, "adjoint H2D");
    }
    if (rc == GY_OK) {
        hipLaunchKernelGGL(k_zero_f32, dim3((unsigned)((sc->tex_floats + 255) / 256 + 1)), dim3(256), 0, c->stream, (float*)dg,
                           sc->tex_floats);
        rc = CheckLaunch(c, "zero gradient");
    }
    if (rc == GY_OK && dl) rc = Fill32(c, dl, 0u, std::max<int64_t>(n_le, 1));
    c->ev_used = 0;
    if (rc == GY_OK) rc = AdjointDevice(sc, cam, p, (const float*)dI, (float*)dg, (float*)dl);
    if (rc == GY_OK) {
        hipError_t e = hipMemcpyAsync(grad, dg, (size_t)sc->tex_floats * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && dl) e = hipMemcpyAsync(grad_le, dl, (size_t)n_le * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = HipFail(e, "adjoint D2H");
    } else {
        (void)hipStreamSynchronize(c->stream);
    }
    TimersCollect(c);
    DevFree(dI);
    DevFree(dg);
    DevFree(dl);
    return rc;
}

extern "C" int gy_adjoint(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p, const float* dimage, float* grad) {
    return gy_adjoint_emission(s, cam, p, dimage, grad, nullptr);
}

// The roughness terms of the adjoint on the host, for tests: d log f / d alpha
// of the GGX lobe for local directions, and d alpha / d roughness.
extern "C" float gy_ggx_dlogf_dalpha(float alpha, const float* wo, const float* wi) {
    return gy::ggx_dlogf_dalpha(alpha, wo[0], wo[1], wo[2], wi[0], wi[1], wi[2]);
}
extern "C" float gy_dalpha_droughness(float r) { return gy::dalpha_droughness(r); }

extern "C" int gy_film_image_device(gy_scene* s, float* image_device) {
    if (!s || !image_device) return Fail(GY_ERR_INVALID, "gy_film_image_device: null argument");
    Scene* sc = reinterpret_cast<Scene*>(s);
    if (!sc->film.film) return Fail(GY_ERR_INVALID, "gy_film_image_device: nothing rendered");
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    const FilmDev F = FilmView(sc);
    const int64_t n = (int64_t)F.width * F.height;
    hipLaunchKernelGGL(k_film_image, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, F, image_device);
    int rc = CheckLaunch(c, "k_film_image");
    hipError_t e = hipStreamSynchronize(c->stream);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "film image sync");
    return rc;
}

extern "C" int64_t gy_scene_texture_floats(gy_scene* s) {
    return s ? reinterpret_cast<Scene*>(s)->tex_floats : -1;
}

// Replaces the texture texels (same layout as gy_scene_desc.tex_data); the next
// render sees the new values.  A render of the old film no longer matches.
static int SetTextures(Scene* sc, const float* data, int64_t n, hipMemcpyKind kind) {
    if (n != sc->tex_floats) return Fail(GY_ERR_INVALID, "gy_scene_set_textures: size differs from the scene's texture data");
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    hipError_t e = hipMemcpyAsync(sc->d_tex, data, (size_t)n * 4, kind, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return HipFail(e, "gy_scene_set_textures");
    sc->film_valid = false;
    ++sc->tex_revision;
    return GY_OK;
}
extern "C" int gy_scene_set_textures(gy_scene* s, const float* data, int64_t n) {
    if (!s || (!data && n)) return Fail(GY_ERR_INVALID, "gy_scene_set_textures: null argument");
    return SetTextures(reinterpret_cast<Scene*>(s), data, n, hipMemcpyHostToDevice);
}
extern "C" int gy_scene_set_textures_device(gy_scene* s, const float* data_device, int64_t n) {
    if (!s || (!data_device && n)) return Fail(GY_ERR_INVALID, "gy_scene_set_textures_device: null argument");
    return SetTextures(reinterpret_cast<Scene*>(s), data_device, n, hipMemcpyDeviceToDevice);
}

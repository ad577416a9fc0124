// libgyunity: context and scene management, ray-query kernels, the plain
// path-tracing megakernel, film accumulation with per-pixel ownership, and the
// C ABI declared in include/gyunity.h.  The wavefront integrator lives in
// wavefront.hip and shares everything declared in runtime.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "device.h"
#include "gy_internal.h"
#include "integrator.h"
#include "runtime.h"

using namespace gyd;

namespace gy {

static thread_local std::string t_error;
void SetError(const std::string& msg) { t_error = msg; }
int Fail(int code, const std::string& msg) {
    t_error = msg;
    return code;
}

int HipFail(hipError_t e, const char* what) {
    return Fail(GY_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ small kernels
__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_f4(float4* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

int Fill32(Ctx* c, void* p, uint32_t v, int64_t n) {
    if (n <= 0) return GY_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (uint32_t*)p, v, n);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GY_OK : HipFail(e, "fill kernel");
}

int FillF4(Ctx* c, float4* p, int64_t n) {
    if (n <= 0) return GY_OK;
    hipLaunchKernelGGL(k_fill_f4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, p, n);
    return CheckLaunch(c, "fill");
}

int DevAlloc(Ctx* c, void** p, size_t bytes) {
    if (c->fault == GY_FAULT_ALLOC) {
        c->fault = GY_FAULT_NONE;
        return Fail(GY_ERR_INJECTED, "injected allocation failure");
    }
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return Fail(GY_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    return GY_OK;
}

int Upload(Ctx* c, void** p, const void* src, size_t bytes) {
    int rc = DevAlloc(c, p, bytes);
    if (rc != GY_OK) return rc;
    if (bytes == 0) return GY_OK;
    hipError_t e = hipMemcpyAsync(*p, src, bytes, hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return HipFail(e, "hipMemcpyAsync H2D");
    e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? GY_OK : HipFail(e, "hipStreamSynchronize");
}

void DevFree(void* p) {
    if (p) (void)hipFree(p);
}

// Growable device scratch buffer.
int Grow(Ctx* c, DevBuf* b, size_t bytes) {
    if (b->bytes >= bytes && b->ptr) return GY_OK;
    DevFree(b->ptr);
    b->ptr = nullptr;
    b->bytes = 0;
    int rc = DevAlloc(c, &b->ptr, bytes);
    if (rc != GY_OK) return rc;
    b->bytes = bytes;
    return GY_OK;
}

int CheckLaunch(Ctx* c, const char* what) {
    if (c->fault == GY_FAULT_LAUNCH) {
        c->fault = GY_FAULT_NONE;
        return Fail(GY_ERR_INJECTED, std::string("injected launch failure in ") + what);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? GY_OK : HipFail(e, what);
}

// Timed region on the context stream: events recorded around each launch of a
// kernel category and summed after the stream synchronises.
void TimerBegin(Ctx* c, int cat) {
    if (c->capturing) return;
    if (c->ev_used + 2 > (int)c->events.size()) {
        hipEvent_t a, b;
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        c->events.push_back(a);
        c->events.push_back(b);
    }
    c->ev_cat.resize(c->events.size() / 2);
    c->ev_cat[c->ev_used / 2] = cat;
    (void)hipEventRecord(c->events[c->ev_used], c->stream);
}
// Timers do not nest: each TimerBegin is closed by one TimerEnd before the next.
void TimerEnd(Ctx* c) {
    if (c->capturing) return;
    if (c->ev_used + 1 >= (int)c->events.size()) return;  // an unmatched end records nothing
    (void)hipEventRecord(c->events[c->ev_used + 1], c->stream);
    c->ev_used += 2;
}
void TimersCollect(Ctx* c) {
    for (int i = 0; i + 1 < c->ev_used; i += 2) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->events[i], c->events[i + 1]);
        const int cat = c->ev_cat[i / 2];
        c->stats.kernel_ms[cat] += ms;
        c->stats.kernel_launches[cat] += 1;
    }
    c->ev_used = 0;
}

// ------------------------------------------------------------------ ray query kernels
// SPH: the scene has analytic spheres.  Each kernel is also built without the
// sphere code, which keeps the triangle-only instantiations' registers.
template <int STACK, bool SPH>
__global__ void __launch_bounds__(256) k_trace_closest(SceneDev S, const gy_ray* __restrict__ rays, int64_t n,
                                                       gy_hit* __restrict__ hits) {
    __shared__ uint32_t stack[STACK * 256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    const float4 a = rp[0], b = rp[1];
    Hit h;
    gy_hit out;
    if (traverse<false, STACK, SPH>(S, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w, &h, stack + threadIdx.x, blockDim.x)) {
        out.t = h.t;
        out.prim = h.prim;
        out.b1 = h.b1;
        out.b2 = h.b2;
    } else {
        out.t = 0.f;
        out.prim = -1;
        out.b1 = 0.f;
        out.b2 = 0.f;
    }
    *reinterpret_cast<float4*>(hits + i) = make_float4(out.t, __int_as_float(out.prim), out.b1, out.b2);
}

template <int STACK, bool SPH>
__global__ void __launch_bounds__(256) k_trace_any(SceneDev S, const gy_ray* __restrict__ rays, int64_t n,
                                                   int32_t* __restrict__ occ) {
    __shared__ uint32_t stack[STACK * 256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4* rp = reinterpret_cast<const float4*>(rays + i);
    const float4 a = rp[0], b = rp[1];
    Hit h;
    occ[i] = traverse<true, STACK, SPH>(S, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w, &h, stack + threadIdx.x, blockDim.x) ? 1 : 0;
}

__device__ inline void wave_add_u64(unsigned long long* dst, uint64_t v) {
    // reduce across the wavefront, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

template <int STACK, bool SPH>
__global__ void __launch_bounds__(256) k_count_visits(SceneDev S, const gy_ray* __restrict__ rays, int64_t n,
                                                      unsigned long long* counters) {
    __shared__ uint32_t stack[STACK * 256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nn = 0, nt = 0;
    if (i < n) {
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        const float4 a = rp[0], b = rp[1];
        Hit h;
        traverse<false, STACK, SPH>(S, v3(a.x, a.y, a.z), v3(b.x, b.y, b.z), a.w, b.w, &h, stack + threadIdx.x, blockDim.x, &nn, &nt);
    }
    wave_add_u64(&counters[0], nn);
    wave_add_u64(&counters[1], nt);
}

// Five waves per SIMD, the most its LDS stack allows at 128 lanes (166 VGPRs
// and three waves without the hint): 96 VGPRs with some spilled, 1.27x faster
// on the 1.31M-triangle icosphere (108.6 -> 138.3 Mpaths/s).
template <int STACK, int BLOCK, bool ENV, bool SPH>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5)))
k_pt_megakernel(SceneDev S, PathParams P, FilmDev F, int s0, int n_s,
                                                         unsigned long long* counters) {
    __shared__ uint32_t stack[STACK * BLOCK];
    const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    int s_local, px, py;
    pass_pixel(F, t, &s_local, &px, &py);
    PathCounters pc = {};
    if (s_local < n_s && px < F.width && py < F.height) {
        float fx, fy;
        const V3 L = trace_path<STACK, false, ENV, SPH>(S, P, px, py, s0 + s_local, stack + threadIdx.x, BLOCK, &fx, &fy, nullptr, pc);
        record_sample(F, s_local, px, py, fx, fy, L);
    }
    wave_add_u64(&counters[0], pc.rays);
    wave_add_u64(&counters[1], pc.shadow);
    wave_add_u64(&counters[2], pc.mis);
}

// Instrumented copy: counts node and triangle visits and shading points.
template <int STACK, int BLOCK, bool ENV, bool SPH>
__global__ void __launch_bounds__(BLOCK) k_pt_count(SceneDev S, PathParams P, FilmDev F, int s0, int n_s,
                                                    unsigned long long* counters) {
    __shared__ uint32_t stack[STACK * BLOCK];
    const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    int s_local, px, py;
    pass_pixel(F, t, &s_local, &px, &py);
    PathCounters pc = {};
    if (s_local < n_s && px < F.width && py < F.height) {
        float fx, fy;
        (void)trace_path<STACK, true, ENV, SPH>(S, P, px, py, s0 + s_local, stack + threadIdx.x, BLOCK, &fx, &fy, nullptr, pc);
    }
    wave_add_u64(&counters[0], pc.rays);
    wave_add_u64(&counters[1], pc.shadow);
    wave_add_u64(&counters[2], pc.mis);
    wave_add_u64(&counters[3], pc.shade);
    for (int k = 0; k < 3; ++k) {
        wave_add_u64(&counters[4 + k], pc.nodes[k]);
        wave_add_u64(&counters[7 + k], pc.tris[k]);
    }
}

template <int STACK, bool ENV, bool SPH>
__global__ void __launch_bounds__(128) k_li_samples(SceneDev S, PathParams P, int n, const int32_t* __restrict__ xs,
                                                    const int32_t* __restrict__ ys, const int32_t* __restrict__ ss,
                                                    float* __restrict__ L, float* __restrict__ raster, int32_t* __restrict__ dims) {
    __shared__ uint32_t stack[STACK * 128];
    const int i = blockIdx.x * 128 + threadIdx.x;
    if (i >= n) return;
    PathCounters pc = {};
    float fx, fy;
    uint32_t nd;
    const V3 v = trace_path<STACK, false, ENV, SPH>(S, P, xs[i], ys[i], ss[i], stack + threadIdx.x, 128, &fx, &fy, &nd, pc);
    L[3 * i] = v.x;
    L[3 * i + 1] = v.y;
    L[3 * i + 2] = v.z;
    raster[2 * i] = fx;
    raster[2 * i + 1] = fy;
    dims[i] = (int32_t)nd;
}

// LightTracing: one light path per thread, sample index = thread index.
// Five waves per SIMD (136 VGPRs and three without the hint): 9% faster.
template <int STACK, int BLOCK, bool ENV>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5))) k_lt(SceneDev S, PathParams P, LightCamDev C, int64_t n, float total,
                                              float* __restrict__ splat) {
    __shared__ uint32_t stack[STACK * BLOCK];
    const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (t >= n) return;
    lt_path<STACK, ENV>(S, P, C, (uint64_t)t, total, splat, stack + threadIdx.x, BLOCK);
}

// ------------------------------------------------------------------ film
// Each pixel's owner thread adds the pass's samples that cover it, in sample
// order (Film::AddSample with weight 1: weight += 1, color = color + L,
// visual/Film.h:80-85).  For each pass sample it takes its own slot when the
// slot's footprint mask holds its pixel, then the slots of the neighbouring
// pixels (offsets in bit order, around the image's edges) whose masks hold
// it.  Neighbours are looked at only when their border flag is set, so the
// common case reads one slot per sample.  The result does not depend on
// scheduling.
__global__ void k_film_accumulate(FilmDev F, int n_s) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int W = F.width, H = F.height;
    const int64_t WH = (int64_t)W * H;
    if (p >= WH) return;
    const int x = (int)(p % W), y = (int)(p / W);
    // bit b = (dy + 1) * 3 + dx + 1: the source pixel (x - dx, y - dy) has border samples
    uint32_t nb = 0;
    int64_t src[9];
    for (int b = 0; b < 9; ++b) {
        const int dx = b % 3 - 1, dy = b / 3 - 1;
        int sx = x - dx, sy = y - dy;
        sx = sx < 0 ? sx + W : (sx >= W ? sx - W : sx);
        sy = sy < 0 ? sy + H : (sy >= H ? sy - H : sy);
        src[b] = (int64_t)sy * W + sx;
        if (b != 4 && F.border[src[b]]) nb |= 1u << b;
    }
    const bool own_border = F.border[p] != 0u;
    float4 acc = F.film[p];
    for (int s = 0; s < n_s; ++s) {
        const int64_t base = (int64_t)s * WH;
        const float4 e = F.pass[base + p];
        if (own_border ? (((uint32_t)e.w & kFilmOwn) != 0u) : (e.w != 0.f)) {
            acc.w += 1.0f;
            acc.x = acc.x + e.x;
            acc.y = acc.y + e.y;
            acc.z = acc.z + e.z;
        }
        for (uint32_t m = nb; m != 0u; m &= m - 1u) {
            const int b = __builtin_ctz(m);
            const float4 q = F.pass[base + src[b]];
            if (((uint32_t)q.w >> b) & 1u) {
                acc.w += 1.0f;
                acc.x = acc.x + q.x;
                acc.y = acc.y + q.y;
                acc.z = acc.z + q.z;
            }
        }
    }
    F.film[p] = acc;
}

// ------------------------------------------------------------------ launch helpers
template <int STACK, bool SPH>
int LaunchTraceClosestT(Ctx* c, const SceneDev& S, const gy_ray* rays, int64_t n, gy_hit* hits) {
    hipLaunchKernelGGL((k_trace_closest<STACK, SPH>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, S, rays, n, hits);
    return CheckLaunch(c, "k_trace_closest");
}
template <int STACK, bool SPH>
int LaunchTraceAnyT(Ctx* c, const SceneDev& S, const gy_ray* rays, int64_t n, int32_t* occ) {
    hipLaunchKernelGGL((k_trace_any<STACK, SPH>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, S, rays, n, occ);
    return CheckLaunch(c, "k_trace_any");
}

int LaunchTraceClosest(Ctx* c, Scene* sc, const gy_ray* rays, int64_t n, gy_hit* hits) {
    if (n <= 0) return GY_OK;
    if (sc->dev.n_sph > 0)
        return sc->stack <= 32 ? LaunchTraceClosestT<32, true>(c, sc->dev, rays, n, hits)
                               : LaunchTraceClosestT<64, true>(c, sc->dev, rays, n, hits);
    return sc->stack <= 32 ? LaunchTraceClosestT<32, false>(c, sc->dev, rays, n, hits)
                           : LaunchTraceClosestT<64, false>(c, sc->dev, rays, n, hits);
}
int LaunchTraceAny(Ctx* c, Scene* sc, const gy_ray* rays, int64_t n, int32_t* occ) {
    if (n <= 0) return GY_OK;
    if (sc->dev.n_sph > 0)
        return sc->stack <= 32 ? LaunchTraceAnyT<32, true>(c, sc->dev, rays, n, occ) : LaunchTraceAnyT<64, true>(c, sc->dev, rays, n, occ);
    return sc->stack <= 32 ? LaunchTraceAnyT<32, false>(c, sc->dev, rays, n, occ) : LaunchTraceAnyT<64, false>(c, sc->dev, rays, n, occ);
}

// ------------------------------------------------------------------ scene upload
static float HostDot(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + (a[2] * b[2] + 0.f); }

constexpr float kInfHost = __builtin_huge_valf();
static float StdMin(float a, float b) { return (b < a) ? b : a; }  // std::min
static float StdMax(float a, float b) { return (a < b) ? b : a; }  // std::max

// Distribution1D over the light powers (visual/Sampling.h:16-45): CDF and integral.
static float LightCdf(const std::vector<float>& func, std::vector<float>* cdf) {
    const int n = (int)func.size();
    cdf->assign(n + 1, 0.f);
    std::vector<float>& c = *cdf;
    c[0] = 0.f;
    for (int i = 1; i < n + 1; ++i) c[i] = c[i - 1] + func[i - 1] / n;
    const float func_int = c[n];
    if (func_int == 0.f) {
        for (int i = 1; i < n + 1; ++i) c[i] = float(i) / float(n);
    } else {
        for (int i = 1; i < n + 1; ++i) c[i] /= func_int;
    }
    return func_int;
}

// The light table on the device: triangle (-1: the environment light), Le,
// power CDF and function.  Replaces the previous table.
static int UploadLights(Ctx* c, Scene* sc, const std::vector<int32_t>& tri, const std::vector<float>& le,
                        const std::vector<float>& func, const std::vector<float>& cdf) {
    void* bufs[4] = {nullptr, nullptr, nullptr, nullptr};
    const size_t n = tri.size();
    int rc = GY_OK;
    if (n > 0) {
        if (rc == GY_OK) rc = Upload(c, &bufs[0], tri.data(), n * 4);
        if (rc == GY_OK) rc = Upload(c, &bufs[1], le.data(), n * 12);
        if (rc == GY_OK) rc = Upload(c, &bufs[2], cdf.data(), cdf.size() * 4);
        if (rc == GY_OK) rc = Upload(c, &bufs[3], func.data(), func.size() * 4);
    }
    if (rc != GY_OK) {
        for (void* b : bufs) DevFree(b);
        return rc;
    }
    DevFree(sc->d_ltri);
    DevFree(sc->d_lle);
    DevFree(sc->d_lcdf);
    DevFree(sc->d_lfunc);
    sc->d_ltri = bufs[0];
    sc->d_lle = bufs[1];
    sc->d_lcdf = bufs[2];
    sc->d_lfunc = bufs[3];
    sc->dev.light_tri = (const int32_t*)bufs[0];
    sc->dev.light_le = (const float*)bufs[1];
    sc->dev.light_cdf = (const float*)bufs[2];
    sc->dev.light_func = (const float*)bufs[3];
    sc->dev.n_light = (int32_t)n;
    return GY_OK;
}

int SceneUpload(Ctx* c, const gy_scene_desc& d, const gy_bvh_params* bp, Scene** out) {
    if (c->fault == GY_FAULT_SCENE) {
        c->fault = GY_FAULT_NONE;
        return Fail(GY_ERR_INJECTED, "injected scene failure");
    }
    if (d.n_tri <= 0 || !d.pos || !d.nrm || !d.face_n || !d.mat || d.n_mat <= 0 || !d.mats)
        return Fail(GY_ERR_INVALID, "scene: need n_tri > 0, pos, nrm, face_n, mat, mats");
    if ((int64_t)d.n_tri > gyd::kTravMaxTriangles) return Fail(GY_ERR_UNSUPPORTED, "scene: more triangles than the traversal addresses (2^27 - 16)");
    if (d.n_light < 0 || (d.n_light > 0 && (!d.light_tri || !d.light_le)))
        return Fail(GY_ERR_INVALID, "scene: n_light >= 0 with light_tri and light_le");
    for (int i = 0; i < d.n_tri; ++i)
        if (d.mat[i] >= (uint32_t)d.n_mat) return Fail(GY_ERR_INVALID, "scene: material id out of range");
    for (int i = 0; i < d.n_light; ++i)
        if (d.light_tri[i] < 0 || d.light_tri[i] >= d.n_tri) return Fail(GY_ERR_INVALID, "scene: light triangle out of range");
    if (d.light)
        for (int i = 0; i < d.n_tri; ++i)
            if (d.light[i] >= d.n_light) return Fail(GY_ERR_INVALID, "scene: light id out of range");
    for (int i = 0; i < d.n_mat; ++i) {
        const uint32_t* w = d.mats + (size_t)i * GY_MAT_WORDS;
        if (w[0] > 3) return Fail(GY_ERR_INVALID, "scene: unknown material type");
        for (int k = 21; k <= 22; ++k)
            if ((int32_t)w[k] >= d.n_tex) return Fail(GY_ERR_INVALID, "scene: texture id out of range");
    }
    for (int i = 0; i < d.n_tex; ++i)
        if (d.tex_info[4 * i] <= 0 || d.tex_info[4 * i + 1] <= 0 || d.tex_info[4 * i + 2] <= 0 || d.tex_info[4 * i + 2] > 4)
            return Fail(GY_ERR_INVALID, "scene: bad texture dimensions");
    auto t0 = std::chrono::steady_clock::now();
    Scene* sc = new Scene();
    sc->ctx = c;
    sc->n_tri = d.n_tri;
    // Shading records by primitive id; area as Triangle::Area (shape/Triangle.h:67-69).
    std::vector<float4> shade((size_t)d.n_tri * 8);
    std::vector<float> area(d.n_tri);
    for (int i = 0; i < d.n_tri; ++i) {
        const float* p = d.pos + 9 * (size_t)i;
        const float* n = d.nrm + 9 * (size_t)i;
        const float* f = d.face_n + 3 * (size_t)i;
        const float e1[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]};
        const float e2[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
        const float cr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        area[i] = (float)(0.5 * std::sqrt(HostDot(cr, cr)));
        float uv[6] = {0, 0, 0, 0, 0, 0};
        if (d.uv) std::memcpy(uv, d.uv + 6 * (size_t)i, sizeof(uv));
        uint32_t mat = d.mat[i];
        int32_t light = d.light ? d.light[i] : -1;
        float matf, lightf;
        std::memcpy(&matf, &mat, 4);
        std::memcpy(&lightf, &light, 4);
        float4* q = &shade[(size_t)i * 8];
        q[0] = make_float4(p[0], p[1], p[2], matf);
        q[1] = make_float4(p[3], p[4], p[5], lightf);
        q[2] = make_float4(p[6], p[7], p[8], area[i]);
        q[3] = make_float4(n[0], n[1], n[2], uv[0]);
        q[4] = make_float4(n[3], n[4], n[5], uv[1]);
        q[5] = make_float4(n[6], n[7], n[8], uv[2]);
        q[6] = make_float4(f[0], f[1], f[2], uv[3]);
        q[7] = make_float4(uv[4], uv[5], 0.f, 0.f);
    }
    // World bound: the union of the triangles' boxes (BVHAccel::WorldBound,
    // Triangle::ObjectBound, math/AABB.cpp:29-60), for the environment light.
    {
        float lo[3] = {kInfHost, kInfHost, kInfHost}, hi[3] = {-kInfHost, -kInfHost, -kInfHost};
        for (int i = 0; i < d.n_tri; ++i) {
            const float* p = d.pos + 9 * (size_t)i;
            for (int k = 0; k < 3; ++k) {
                float a = StdMin(p[k], p[3 + k]), b = StdMax(p[k], p[3 + k]);
                a = StdMin(a, p[6 + k]);
                b = StdMax(b, p[6 + k]);
                lo[k] = StdMin(lo[k], a);
                hi[k] = StdMax(hi[k], b);
            }
        }
        for (int k = 0; k < 3; ++k) {
            sc->world_lo[k] = lo[k];
            sc->world_hi[k] = hi[k];
        }
    }
    // Light power distribution (visual/Scene.h:121-128, visual/Sampling.h:16-45,
    // light/AreaLight.h:45-48) in FP32.
    const int nl = d.n_light;
    const float PI = 3.14159265358979f;
    sc->area_func.resize(nl);
    for (int i = 0; i < nl; ++i) {
        const float* le = d.light_le + 3 * i;
        const float a = area[d.light_tri[i]];
        const float P[3] = {(le[0] * a) * PI, (le[1] * a) * PI, (le[2] * a) * PI};
        sc->area_func[i] = 0.212671f * P[0] + 0.715160f * P[1] + 0.072169f * P[2];
    }
    sc->area_tri.assign(d.light_tri, d.light_tri + nl);
    sc->area_le.assign(d.light_le, d.light_le + 3 * (size_t)nl);
    std::vector<float> func = sc->area_func, cdf;
    const float func_int = LightCdf(func, &cdf);
    sc->dev.light_func_int = func_int;
    sc->dev.n_light = nl;
    sc->dev.n_tri = d.n_tri;
    // BVH
    gy_bvh_params bvh = {};
    if (bp) bvh = *bp;
    const int leaf = bvh.max_leaf_size > 0 ? bvh.max_leaf_size : 4;
    int rc = GY_OK;
    BVHBuild b;
    auto tb0 = std::chrono::steady_clock::now();
    if (bvh.builder == 0) {
        rc = BuildBVHDevice(c, d.pos, d.n_tri, leaf, &b);
        if (rc != GY_OK) {
            delete sc;
            return rc;
        }
    } else {
        BuildBVHHost(d.pos, d.n_tri, leaf, &b);
    }
    auto tb1 = std::chrono::steady_clock::now();
    sc->stats.build_ms = std::chrono::duration<double, std::milli>(tb1 - tb0).count();
    sc->stats.n_nodes = (int64_t)(b.nodes.size() / 16);
    sc->stats.n_leaves = b.n_leaves;
    sc->stats.max_depth = b.max_depth;
    sc->stats.max_leaf_size = b.max_leaf;
    sc->stats.sah_cost = b.sah_cost;
    if (b.max_depth + 1 > 64) {
        delete sc;
        return Fail(GY_ERR_UNSUPPORTED, "BVH deeper than the 64-entry traversal stack");
    }
    sc->stack = (b.max_depth + 1 <= 32) ? 32 : 64;
    // Pad every child box outward so that the traversal never rejects a box
    // containing a triangle the triangle test accepts (rounding in that test
    // can place an accepted hit slightly outside the triangle).
    float ext = 0.f;
    for (size_t i = 0; i < (size_t)d.n_tri * 9; ++i) ext = std::max(ext, std::fabs(d.pos[i]));
    const float pad = 4e-6f * (1.f + ext);
    for (size_t nidx = 0; nidx < b.nodes.size() / 16; ++nidx) {
        uint32_t* nd = &b.nodes[nidx * 16];
        for (int s = 0; s < 2; ++s) {
            if (nd[12 + s] == kEmptyRef) continue;
            float v[6];
            std::memcpy(v, &nd[s * 6], 24);
            for (int k = 0; k < 3; ++k) {
                v[k] -= pad;
                v[3 + k] += pad;
            }
            std::memcpy(&nd[s * 6], v, 24);
        }
    }
    sc->tri_order = b.tri_order;
    std::vector<float4> tri((size_t)d.n_tri * 3);
    for (size_t k = 0; k < b.tri_order.size(); ++k) {
        const uint32_t id = b.tri_order[k];
        const float* p = d.pos + 9 * (size_t)id;
        int32_t idi = (int32_t)id;
        float idf;
        std::memcpy(&idf, &idi, 4);
        tri[k * 3 + 0] = make_float4(p[0], p[1], p[2], idf);
        tri[k * 3 + 1] = make_float4(p[3], p[4], p[5], 0.f);
        tri[k * 3 + 2] = make_float4(p[6], p[7], p[8], 0.f);
    }
    std::vector<float> texdata;
    std::vector<int4> texinfo(std::max(1, d.n_tex));
    size_t texn = 0;
    for (int i = 0; i < d.n_tex; ++i)
        texn = std::max(texn, (size_t)d.tex_info[4 * i + 3] + (size_t)d.tex_info[4 * i] * d.tex_info[4 * i + 1] * d.tex_info[4 * i + 2]);
    if (d.n_tex) texdata.assign(d.tex_data, d.tex_data + texn);
    for (int i = 0; i < d.n_tex; ++i)
        texinfo[i] = make_int4(d.tex_info[4 * i], d.tex_info[4 * i + 1], d.tex_info[4 * i + 2], d.tex_info[4 * i + 3]);
    void* p;
    if ((rc = Upload(c, &p, tri.data(), tri.size() * sizeof(float4))) != GY_OK) goto fail;
    sc->d_tri = p;
    sc->dev.tri = (const float4*)p;
    if ((rc = Upload(c, &p, shade.data(), shade.size() * sizeof(float4))) != GY_OK) goto fail;
    sc->d_shade = p;
    sc->dev.shade = (const float4*)p;
    if ((rc = Upload(c, &p, b.nodes.data(), b.nodes.size() * 4)) != GY_OK) goto fail;
    sc->d_nodes = p;
    sc->dev.nodes = (const uint4*)p;
    sc->dev.n_nodes = (int32_t)(b.nodes.size() / 16);
    if ((rc = Upload(c, &p, d.mats, (size_t)d.n_mat * GY_MAT_WORDS * 4)) != GY_OK) goto fail;
    sc->d_mats = p;
    sc->dev.mats = (const uint32_t*)p;
    if ((rc = Upload(c, &p, texdata.data(), texdata.size() * 4)) != GY_OK) goto fail;
    sc->d_tex = p;
    sc->dev.tex = (const float*)p;
    sc->tex_floats = (int64_t)texn;
    if ((rc = Upload(c, &p, texinfo.data(), texinfo.size() * sizeof(int4))) != GY_OK) goto fail;
    sc->d_texinfo = p;
    sc->dev.tex_info = (const int4*)p;
    if ((rc = UploadLights(c, sc, sc->area_tri, sc->area_le, func, cdf)) != GY_OK) goto fail;
    sc->host_nodes = std::move(b.nodes);
    {
        auto t1 = std::chrono::steady_clock::now();
        sc->upload_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    *out = sc;
    return GY_OK;
fail:
    SceneFree(sc);
    return rc;
}

void SceneFree(Scene* sc) {
    if (!sc) return;
    void* bufs[] = {sc->d_tri, sc->d_shade, sc->d_nodes, sc->d_mats, sc->d_tex, sc->d_texinfo, sc->d_ltri, sc->d_lle, sc->d_lcdf, sc->d_lfunc,
                    sc->d_env, sc->d_sph, sc->d_media, sc->d_tri_media};
    for (void* b : bufs) DevFree(b);
    FilmFree(&sc->film);
    WavefrontFree(sc);
    WavefrontStateFree(&sc->vwf);
    delete sc;
}

void FilmFree(FilmState* f) {
    DevFree(f->film);
    DevFree(f->pass);
    DevFree(f->border);
    DevFree(f->counters);
    *f = FilmState();
}

int FilmPrepareState(Ctx* c, FilmState& f, int W, int H, int64_t pass_slots) {
    const int64_t WH = (int64_t)W * H;
    if (f.width != W || f.height != H || !f.film) {
        DevFree(f.film);
        DevFree(f.border);
        f.film = nullptr;
        f.border = nullptr;
        int rc = DevAlloc(c, (void**)&f.film, (size_t)WH * sizeof(float4));
        if (rc == GY_OK) rc = DevAlloc(c, (void**)&f.border, (size_t)(WH + 1) * sizeof(uint32_t));
        if (rc != GY_OK) return rc;
        f.width = W;
        f.height = H;
    }
    if (f.pass_slots < pass_slots) {
        DevFree(f.pass);
        f.pass = nullptr;
        int rc = DevAlloc(c, (void**)&f.pass, (size_t)pass_slots * sizeof(float4));
        if (rc != GY_OK) return rc;
        f.pass_slots = pass_slots;
    }
    if (!f.counters) {
        int rc = DevAlloc(c, (void**)&f.counters, 16 * sizeof(unsigned long long));
        if (rc != GY_OK) return rc;
    }
    return FilmClear(c, f);
}

// Zero the film, its border flags and error count, and its ray counters
// (fill kernels only, so a recorded graph holds no memset node).
int FilmClear(Ctx* c, const FilmState& f) {
    const int64_t WH = (int64_t)f.width * f.height;
    hipLaunchKernelGGL(k_fill_f4, dim3((unsigned)((WH + 255) / 256)), dim3(256), 0, c->stream, f.film, WH);
    int rc = CheckLaunch(c, "film clear");
    if (rc == GY_OK) rc = Fill32(c, f.border, 0u, WH + 1);
    if (rc == GY_OK) rc = Fill32(c, (uint32_t*)f.counters, 0u, 32);
    return rc;
}

// GY_OK, or the error for samples whose footprint reached beyond the
// neighbouring pixels (counted by film_record).
int FilmCheck(const FilmState& f) {
    uint32_t far = 0;
    hipError_t e = hipMemcpy(&far, f.border + (int64_t)f.width * f.height, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return HipFail(e, "film error count");
    if (far != 0) return Fail(GY_ERR_UNSUPPORTED, "film: a sample's footprint reaches beyond its neighbouring pixels");
    return GY_OK;
}

int FilmPrepare(Ctx* c, Scene* sc, int W, int H, int n_pass) { return FilmPrepareState(c, sc->film, W, H, (int64_t)W * H * n_pass); }
FilmDev FilmViewState(const FilmState& f) {
    FilmDev F;
    F.film = f.film;
    F.pass = f.pass;
    F.border = f.border;
    F.width = f.width;
    F.height = f.height;
    return F;
}
FilmDev FilmView(Scene* sc) { return FilmViewState(sc->film); }
int FilmResolveState(Ctx* c, const FilmState& f, int n_s, bool timed) {
    FilmDev F = FilmViewState(f);
    const int64_t WH = (int64_t)F.width * F.height;
    if (timed) TimerBegin(c, kKernFilm);
    hipLaunchKernelGGL(k_film_accumulate, dim3((unsigned)((WH + 255) / 256)), dim3(256), 0, c->stream, F, n_s);
    if (timed) TimerEnd(c);
    int rc = CheckLaunch(c, "k_film_accumulate");
    if (rc != GY_OK) return rc;
    return Fill32(c, F.border, 0u, WH);  // border flags are per pass
}
int FilmResolvePass(Ctx* c, Scene* sc, int n_s) { return FilmResolveState(c, sc->film, n_s); }

// vpt: for gy_render_vpt, which also renders material-less medium boundaries.
int MakePathParamsFor(Ctx* c, Scene* sc, const gy_camera_desc& cam, const gy_render_params& p, PathParams* P, bool vpt) {
    CameraRT rt;
    int rc = BuildCamera(cam, &rt);
    if (rc != GY_OK) return rc;
    if (rt.log2res > sobol::kMaxLog2Res) return Fail(GY_ERR_UNSUPPORTED, "image too large for the Sobol enumeration");
    std::memcpy(P->cam.m, rt.raster_to_world, sizeof(P->cam.m));
    std::memcpy(P->cam.pos, rt.pos, sizeof(P->cam.pos));
    P->cam.width = rt.width;
    P->cam.height = rt.height;
    P->cam.res = rt.res;
    P->cam.log2res = rt.log2res;
    std::memset(&P->en, 0, sizeof(P->en));
    P->en.kind = (uint32_t)p.sampler;
    if (p.sampler == GY_SAMPLER_HALTON) {
        const halton::EnumParams h = halton::Enum((uint32_t)rt.width, (uint32_t)rt.height);
        P->en.p2 = h.p2;
        P->en.p3 = h.p3;
        P->en.hmx = h.mx;
        P->en.hmy = h.my;
        P->en.hinc = h.inc;
        P->en.sx = h.sx;
        P->en.sy = h.sy;
    } else {
        const sobol::EnumTables& E = sobol::Enum(rt.log2res);
        P->en.m = (uint32_t)rt.log2res;
        for (int i = 0; i < 32; ++i) {
            P->en.inv[i] = E.inv[i];
            P->en.frame[i] = E.frame[i];
        }
        P->en.sx = P->en.sy = (float)rt.res;
    }
    P->smp.sobol = c->d_sobol;
    P->smp.htab = (const uint16_t*)c->d_halton;
    P->smp.hdim = (const uint4*)((const char*)c->d_halton + c->halton_dim_off);
    P->smp.hscale = (const float*)((const char*)c->d_halton + c->halton_scale_off);
    P->smp.kind = p.sampler;
    P->max_depth = p.max_depth;
    P->rr_depth = p.rr_depth;
    if (p.sampler == GY_SAMPLER_SOBOL) {
        // dimensions a path may consume: 2 + depth * (6 + 3 + 1)
        const int64_t dims = 2 + (int64_t)p.max_depth * 10;
        if (dims > sobol::kDims) return Fail(GY_ERR_UNSUPPORTED, "max_depth needs more Sobol dimensions than the table holds");
        const int bits = 2 * rt.log2res + (p.spp > 1 ? 32 - __builtin_clz((unsigned)(p.spp - 1)) : 0);
        if (bits > sobol::kCols) return Fail(GY_ERR_UNSUPPORTED, "sample index beyond 52 bits");
    }
    // Halton: dimensions past 255 read 0 as HaltonSampler::sample returns
    if (sc->dev.n_light == 0) return Fail(GY_ERR_INVALID, "render: the scene has no light (area or environment)");
    // PathTracing dereferences every hit's BSDF; material-less medium
    // boundaries are for VolPathTracing only
    if (sc->has_no_mat && !vpt)
        return Fail(GY_ERR_UNSUPPORTED, "render: the scene has material-less triangles (medium boundaries): use gy_render_vpt");
    return GY_OK;
}
int MakePathParams(Ctx* c, Scene* sc, const gy_camera_desc& cam, const gy_render_params& p, PathParams* P) {
    return MakePathParamsFor(c, sc, cam, p, P, false);
}

int ValidateRender(const gy_camera_desc* cam, const gy_render_params* p) {
    if (!cam || !p) return Fail(GY_ERR_INVALID, "render: null camera or params");
    if (p->spp <= 0 || p->max_depth < 0 || p->max_depth > 100) return Fail(GY_ERR_INVALID, "render: spp > 0 and 0 <= max_depth <= 100");
    if ((int64_t)cam->width * cam->height > (1ll << 26)) return Fail(GY_ERR_UNSUPPORTED, "render: image larger than 2^26 pixels");
    if (p->sampler != GY_SAMPLER_SOBOL && p->sampler != GY_SAMPLER_HALTON) return Fail(GY_ERR_INVALID, "render: unknown sampler");
    if (p->ray_sort < GY_RAY_SORT_NONE || p->ray_sort > GY_RAY_SORT_OCTANT_LOCAL) return Fail(GY_ERR_INVALID, "render: unknown ray_sort");
    return GY_OK;
}

template <int STACK, int BLOCK, bool SPH>
int LaunchMegaT(Ctx* c, Scene* sc, const FilmState& film, const PathParams& P, int s0, int n_s, bool count, bool timed) {
    FilmDev F = FilmViewState(film);
    const int tilesX = (F.width + 7) >> 3, tilesY = (F.height + 7) >> 3;
    const int64_t threads = (int64_t)tilesX * tilesY * 64 * n_s;
    const unsigned blocks = (unsigned)((threads + BLOCK - 1) / BLOCK);
    unsigned long long* ctr = film.counters;
    if (count) {
        if (sc->dev.env.w > 0)
            hipLaunchKernelGGL((k_pt_count<STACK, BLOCK, true, SPH>), dim3(blocks), dim3(BLOCK), 0, c->stream, sc->dev, P, F, s0, n_s, ctr);
        else
            hipLaunchKernelGGL((k_pt_count<STACK, BLOCK, false, SPH>), dim3(blocks), dim3(BLOCK), 0, c->stream, sc->dev, P, F, s0, n_s, ctr);
        return CheckLaunch(c, "k_pt_count");
    }
    if (timed) TimerBegin(c, kKernMega);
    if (sc->dev.env.w > 0)
        hipLaunchKernelGGL((k_pt_megakernel<STACK, BLOCK, true, SPH>), dim3(blocks), dim3(BLOCK), 0, c->stream, sc->dev, P, F, s0, n_s, ctr);
    else
        hipLaunchKernelGGL((k_pt_megakernel<STACK, BLOCK, false, SPH>), dim3(blocks), dim3(BLOCK), 0, c->stream, sc->dev, P, F, s0, n_s, ctr);
    if (timed) TimerEnd(c);
    return CheckLaunch(c, "k_pt_megakernel");
}

// The megakernel over one pass into `film` (the scene's own film, or the
// no-alloc entry point's).  Scenes with spheres run the sphere instantiation
// at 128 threads a block.
int LaunchMegaFilm(Ctx* c, Scene* sc, const FilmState& film, const PathParams& P, int s0, int n_s, int block, bool count,
                   bool timed) {
    if (sc->dev.n_sph > 0)
        return sc->stack <= 32 ? LaunchMegaT<32, 128, true>(c, sc, film, P, s0, n_s, count, timed)
                               : LaunchMegaT<64, 128, true>(c, sc, film, P, s0, n_s, count, timed);
    if (sc->stack <= 32) {
        if (block == 64) return LaunchMegaT<32, 64, false>(c, sc, film, P, s0, n_s, count, timed);
        if (block == 256) return LaunchMegaT<32, 256, false>(c, sc, film, P, s0, n_s, count, timed);
        return LaunchMegaT<32, 128, false>(c, sc, film, P, s0, n_s, count, timed);
    }
    if (block == 64) return LaunchMegaT<64, 64, false>(c, sc, film, P, s0, n_s, count, timed);
    return LaunchMegaT<64, 128, false>(c, sc, film, P, s0, n_s, count, timed);
}

int LaunchMega(Ctx* c, Scene* sc, const PathParams& P, int s0, int n_s, int block, bool count) {
    return LaunchMegaFilm(c, sc, sc->film, P, s0, n_s, block, count, true);
}

int SamplesPerPass(const gy_camera_desc& cam, const gy_render_params& p) {
    if (p.samples_per_pass > 0) return std::min(p.samples_per_pass, p.spp);
    // Passes of up to 64M paths: the wavefront keeps every path of a pass
    // resident (~300 B each, ~20 GB at 64M of the 288 GB), which amortises
    // each bounce's kernel drain over more work (1024^2 x 64 spp in one pass
    // instead of two: 338 -> 350 Mpaths/s); the megakernel uses the same split
    // so both accumulate the film in the same order.
    const int64_t WH = (int64_t)cam.width * cam.height;
    int64_t s = (64ll << 20) / std::max<int64_t>(1, WH);
    if (s < 1) s = 1;
    if (s > p.spp) s = p.spp;
    return (int)s;
}

// ------------------------------------------------------------------ no-alloc scratch
// Device memory for gy_render_noalloc: the ray counters, the per-pixel border
// flags and the pass buffer of the megakernel render, in one file-scope array
// that every GPU the library runs on holds until the process ends.  Host code
// never passes the array's name to a kernel or a copy: it is a host-side
// stand-in.  gy_open looks up this device's own address once per device
// (hipGetSymbolAddress, with the device current) and keeps it per device;
// kernels get that address.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// This device's scratch address, looked up on the first call for the device
// (the current device must be `device`).  nullptr if the lookup fails.
static void* ScratchAddress(int device, hipError_t* err) {
    *err = hipSuccess;
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*err != hipSuccess) return nullptr;
        g_scratch_ptr[device] = p;
    }
    return g_scratch_ptr[device];
}

}  // namespace gy

// ====================================================================== C ABI
using namespace gy;

extern "C" const char* gy_last_error(void) { return t_error.c_str(); }
extern "C" const char* gy_version(void) { return "gyunity-amd 0.1 (gfx950)"; }

extern "C" int gy_open(int device, gy_context** out) {
    if (!out) return Fail(GY_ERR_INVALID, "gy_open: null output");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return Fail(GY_ERR_DEVICE, "gy_open: no HIP device available");
    if (device < 0 || device >= n) return Fail(GY_ERR_INVALID, "gy_open: device index out of range");
    e = hipSetDevice(device);
    if (e != hipSuccess) return HipFail(e, "hipSetDevice");
    Ctx* c = new Ctx();
    c->device = device;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
        if (c->num_cu <= 0) c->num_cu = 256;
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return HipFail(e, "hipStreamCreate");
    }
    {
        // the no-alloc scratch array's address on this device (gy_render_noalloc)
        hipError_t se = hipSuccess;
        c->scratch = ScratchAddress(device, &se);
        if (se != hipSuccess) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            return HipFail(se, "hipGetSymbolAddress(scratch)");
        }
    }
    void* p = nullptr;
    int rc = Upload(c, &p, sobol::ByteTables(), (size_t)sobol::kDims * sobol::kBytes * 256 * 4);
    if (rc != GY_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    c->d_sobol = (const uint32_t*)p;
    {
        // Halton tables: digit tables (u16), per-dimension info (uint4), scales
        const auto& tab = halton::Tables();
        const auto& dims = halton::Dims();
        const size_t tab_bytes = (tab.size() * 2 + 255) & ~(size_t)255;
        c->halton_dim_off = tab_bytes;
        c->halton_scale_off = tab_bytes + dims.size() * 16;
        std::vector<char> blob(c->halton_scale_off + dims.size() * 4);
        std::memcpy(blob.data(), tab.data(), tab.size() * 2);
        for (size_t d = 0; d < dims.size(); ++d) {
            const uint32_t info[4] = {dims[d].base, dims[d].tsize, dims[d].groups, dims[d].offset};
            std::memcpy(blob.data() + c->halton_dim_off + 16 * d, info, 16);
            std::memcpy(blob.data() + c->halton_scale_off + 4 * d, &dims[d].scale, 4);
        }
        rc = Upload(c, &c->d_halton, blob.data(), blob.size());
        if (rc != GY_OK) {
            DevFree((void*)c->d_sobol);
            (void)hipStreamDestroy(c->stream);
            delete c;
            return rc;
        }
    }
    *out = reinterpret_cast<gy_context*>(c);
    return GY_OK;
}

extern "C" int gy_close(gy_context* ctx) {
    if (!ctx) return GY_OK;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    DevFree((void*)c->d_sobol);
    DevFree(c->d_halton);
    DevFree(c->rays.ptr);
    DevFree(c->hits.ptr);
    DevFree(c->ctr.ptr);
    for (hipEvent_t ev : c->events) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return GY_OK;
}

extern "C" int gy_device_name(gy_context* ctx, char* buf, int len) {
    if (!ctx || !buf || len <= 0) return Fail(GY_ERR_INVALID, "gy_device_name: bad argument");
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, reinterpret_cast<Ctx*>(ctx)->device);
    if (e != hipSuccess) return HipFail(e, "hipGetDeviceProperties");
    std::snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return GY_OK;
}

extern "C" int gy_debug_inject_fault(gy_context* ctx, int kind) {
    if (!ctx) return Fail(GY_ERR_INVALID, "gy_debug_inject_fault: null context");
    if (kind < GY_FAULT_NONE || kind > GY_FAULT_SCENE) return Fail(GY_ERR_INVALID, "gy_debug_inject_fault: unknown kind");
    reinterpret_cast<Ctx*>(ctx)->fault = kind;
    return GY_OK;
}

extern "C" int gy_scene_create(gy_context* ctx, const gy_scene_desc* desc, const gy_bvh_params* bvh, gy_scene** out) {
    if (!ctx || !desc || !out) return Fail(GY_ERR_INVALID, "gy_scene_create: null argument");
    *out = nullptr;
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    (void)hipSetDevice(c->device);
    Scene* sc = nullptr;
    int rc = SceneUpload(c, *desc, bvh, &sc);
    if (rc != GY_OK) return rc;
    static std::atomic<uint64_t> next_uid{1};
    sc->uid = next_uid++;
    *out = reinterpret_cast<gy_scene*>(sc);
    return GY_OK;
}

int gy::SceneSignature(gy_scene* s, uint64_t* uid, uint64_t* revision) {
    if (!s) return Fail(GY_ERR_INVALID, "null scene");
    const Scene* sc = reinterpret_cast<const Scene*>(s);
    *uid = sc->uid;
    *revision = sc->revision;
    return GY_OK;
}

// EnvironmentLight(image, rotate, scale) added after the area lights, then
// Scene::Initialize: EnvironmentMap texels (ImageIO::LoadTexture row order),
// ConvertImageToPdf + Distribution2D (light/EnvMap.h:82-94, visual/Sampling.h:102-119),
// the bounding-sphere radius (visual/Scene.cpp:9-11) and the light power
// distribution with EnvironmentLight::Power (light/EnvLight.cpp:60-63), in FP32.
// Sphere(ObjectToWorld = WorldToObject = identity, radius, position, inside)
// primitives (shape/Sphere.h:21-27), after the triangles in primitive order;
// an emissive one also gets an AreaLight (light/AreaLight.h) after the
// existing area lights, with power Le * 4 pi r^2 * pi (light/AreaLight.h:45-48).
extern "C" int gy_scene_set_spheres(gy_scene* s, int32_t n, const gy_sphere* spheres) {
    if (!s || n < 0 || (n > 0 && !spheres)) return Fail(GY_ERR_INVALID, "gy_scene_set_spheres: bad argument");
    Scene* sc = reinterpret_cast<Scene*>(s);
    if (sc->dev.n_sph > 0) return Fail(GY_ERR_INVALID, "gy_scene_set_spheres: the scene already has spheres");
    if (sc->dev.env.w > 0) return Fail(GY_ERR_UNSUPPORTED, "gy_scene_set_spheres: add spheres before the environment light");
    if (n > 65536) return Fail(GY_ERR_UNSUPPORTED, "gy_scene_set_spheres: at most 65536 spheres (each ray tests every sphere)");
    if (n == 0) return GY_OK;
    if ((int64_t)sc->n_tri + n > gyd::kTravMaxTriangles) return Fail(GY_ERR_UNSUPPORTED, "gy_scene_set_spheres: too many primitives");
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    const float PI = 3.14159265358979f;
    std::vector<float4> rec(2 * (size_t)n);
    std::vector<int32_t> ltri = sc->area_tri;
    std::vector<float> lle = sc->area_le, func = sc->area_func, cdf;
    for (int32_t i = 0; i < n; ++i) {
        const gy_sphere& q = spheres[i];
        if (!(q.radius > 0.f)) return Fail(GY_ERR_INVALID, "gy_scene_set_spheres: radius must be positive");
        const float area = ((4.f * PI) * q.radius) * q.radius;
        int32_t light = -1;
        if (q.emissive) {
            light = (int32_t)ltri.size();
            ltri.push_back(sc->n_tri + i);
            lle.insert(lle.end(), {q.le[0], q.le[1], q.le[2]});
            const float P[3] = {(q.le[0] * area) * PI, (q.le[1] * area) * PI, (q.le[2] * area) * PI};
            func.push_back(0.212671f * P[0] + 0.715160f * P[1] + 0.072169f * P[2]);
        }
        rec[2 * i] = make_float4(q.center[0], q.center[1], q.center[2], q.radius);
        rec[2 * i + 1] = make_float4(__builtin_bit_cast(float, (uint32_t)(q.inside != 0)), __builtin_bit_cast(float, q.material),
                                     __builtin_bit_cast(float, light), area);
        for (int k = 0; k < 3; ++k) {
            sc->world_lo[k] = std::min(sc->world_lo[k], q.center[k] - q.radius);
            sc->world_hi[k] = std::max(sc->world_hi[k], q.center[k] + q.radius);
        }
    }
    void* d = nullptr;
    int rc = Upload(c, &d, rec.data(), rec.size() * sizeof(float4));
    if (rc != GY_OK) return rc;
    const float func_int = LightCdf(func, &cdf);
    rc = UploadLights(c, sc, ltri, lle, func, cdf);
    if (rc != GY_OK) {
        DevFree(d);
        return rc;
    }
    sc->area_tri = ltri;
    sc->area_le = lle;
    sc->area_func = func;
    sc->dev.light_func_int = func_int;
    sc->d_sph = d;
    sc->dev.sph = (const float4*)d;
    sc->dev.n_sph = n;
    ++sc->revision;
    sc->film_valid = false;
    return GY_OK;
}

extern "C" int gy_scene_set_environment(gy_scene* s, const float* rgb, int32_t width, int32_t height, float scale) {
    if (!s || !rgb || width <= 0 || height <= 0) return Fail(GY_ERR_INVALID, "gy_scene_set_environment: need an image and a positive size");
    if ((int64_t)width * height > (1ll << 26)) return Fail(GY_ERR_UNSUPPORTED, "gy_scene_set_environment: image larger than 2^26 texels");
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    const int w = width, h = height;
    const float PI = 3.14159265358979f;
    std::vector<float> tex((size_t)w * h * 3);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int k = 0; k < 3; ++k) tex[((size_t)y * w + x) * 3 + k] = rgb[((size_t)(h - 1 - y) * w + x) * 3 + k];
    std::vector<float> cfunc((size_t)w * h), ccdf((size_t)(w + 1) * h), cint(h), mcdf;
    std::vector<float> row(w), rcdf;
    for (int r = 0; r < h; ++r) {
        const float v = (float)(r + 0.5f) / (float)(h);
        const float st = (float)std::sin((double)(PI * v));
        for (int x = 0; x < w; ++x) {
            const float* t = &tex[((size_t)(h - r - 1) * w + x) * 3];  // ElementAt(x, r): image row r
            row[x] = st * (0.212671f * t[0] + 0.715160f * t[1] + 0.072169f * t[2]);
        }
        cint[r] = LightCdf(row, &rcdf);
        std::copy(row.begin(), row.end(), cfunc.begin() + (size_t)r * w);
        std::copy(rcdf.begin(), rcdf.end(), ccdf.begin() + (size_t)r * (w + 1));
    }
    const float mint = LightCdf(cint, &mcdf);
    // bounding-sphere radius: (max - min).Length() * 0.5, DPPS dot
    const float e[3] = {sc->world_hi[0] - sc->world_lo[0], sc->world_hi[1] - sc->world_lo[1], sc->world_hi[2] - sc->world_lo[2]};
    const float radius = std::sqrt(HostDot(e, e)) * 0.5f;
    // LookupRadiance(0.5, 0.5) * PI * r * r
    auto texel = [&](int u, int v, int k) {
        u = u < 0 ? 0 : (u > w - 1 ? w - 1 : u);
        v = v < 0 ? 0 : (v > h - 1 ? h - 1 : v);
        return tex[((size_t)v * w + u) * 3 + k];
    };
    const float sx = (float)w * 0.5f - 0.5f, ty = (float)h * (1.f - 0.5f) - 0.5f;
    const int s0 = (int)std::floor(sx), t0 = (int)std::floor(ty);
    const float ds = sx - (float)s0, dt = ty - (float)t0;
    float Pw[3];
    for (int k = 0; k < 3; ++k) {
        const float a = (1.f - ds) * texel(s0, t0, k) + ds * texel(s0 + 1, t0, k);
        const float b = (1.f - ds) * texel(s0, t0 + 1, k) + ds * texel(s0 + 1, t0 + 1, k);
        Pw[k] = ((((1.f - dt) * a + dt * b) * scale) * PI * radius) * radius;
    }
    std::vector<int32_t> ltri = sc->area_tri;
    std::vector<float> lle = sc->area_le, func = sc->area_func, cdf;
    ltri.push_back(-1);
    lle.insert(lle.end(), {0.f, 0.f, 0.f});
    func.push_back(0.212671f * Pw[0] + 0.715160f * Pw[1] + 0.072169f * Pw[2]);
    const float func_int = LightCdf(func, &cdf);
    // one device block: texels | conditional functions | CDFs | integrals | marginal CDF
    auto al = [](size_t bytes) { return (bytes + 255) & ~(size_t)255; };
    const size_t o_tex = 0, o_cf = o_tex + al(tex.size() * 4), o_cc = o_cf + al(cfunc.size() * 4),
                 o_ci = o_cc + al(ccdf.size() * 4), o_mc = o_ci + al(cint.size() * 4), total = o_mc + al(mcdf.size() * 4);
    std::vector<char> blk(total, 0);
    std::memcpy(blk.data() + o_tex, tex.data(), tex.size() * 4);
    std::memcpy(blk.data() + o_cf, cfunc.data(), cfunc.size() * 4);
    std::memcpy(blk.data() + o_cc, ccdf.data(), ccdf.size() * 4);
    std::memcpy(blk.data() + o_ci, cint.data(), cint.size() * 4);
    std::memcpy(blk.data() + o_mc, mcdf.data(), mcdf.size() * 4);
    void* d = nullptr;
    int rc = Upload(c, &d, blk.data(), total);
    if (rc != GY_OK) return rc;
    rc = UploadLights(c, sc, ltri, lle, func, cdf);
    if (rc != GY_OK) {
        DevFree(d);
        return rc;
    }
    DevFree(sc->d_env);
    sc->d_env = d;
    const char* base = (const char*)d;
    gyd::EnvDev& E = sc->dev.env;
    E.tex = (const float*)(base + o_tex);
    E.cfunc = (const float*)(base + o_cf);
    E.ccdf = (const float*)(base + o_cc);
    E.cint = (const float*)(base + o_ci);
    E.mcdf = (const float*)(base + o_mc);
    E.mint = mint;
    E.scale = scale;
    E.radius = radius;
    E.w = w;
    E.h = h;
    sc->dev.light_func_int = func_int;
    sc->film_valid = false;
    ++sc->revision;
    return GY_OK;
}

namespace gy {
int LightTracingWavefront(Ctx* c, Scene* sc, const PathParams& P, const LightCamDev& C, int64_t n, float total, float* splat);
}

// LightTracing::Render (integrator/LT.h:17-53): spp light paths per pixel
// index, rgb gets the film's splat sums (W x H x 3; the film has no samples,
// so they are its pixels).  The splats are added with float atomics, so the
// summation order, not the per-path values, varies between runs.
extern "C" int gy_render_light_tracing(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p, float* rgb) {
    if (!s || !rgb) return Fail(GY_ERR_INVALID, "gy_render_light_tracing: null argument");
    int rc = ValidateRender(cam, p);
    if (rc != GY_OK) return rc;
    if (p->max_depth < 1) return Fail(GY_ERR_INVALID, "gy_render_light_tracing: max_depth >= 1");
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    if (sc->area_tri.empty() && sc->dev.env.w <= 0) return Fail(GY_ERR_INVALID, "gy_render_light_tracing: the scene has no light");
    if (sc->has_no_mat) return Fail(GY_ERR_UNSUPPORTED, "gy_render_light_tracing: scenes with material-less triangles");
    if (sc->dev.n_sph > 0) return Fail(GY_ERR_UNSUPPORTED, "gy_render_light_tracing: scenes with spheres");
    const int64_t n = (int64_t)cam->width * cam->height * p->spp;
    if (n >= (1ll << 36)) return Fail(GY_ERR_UNSUPPORTED, "gy_render_light_tracing: more than 2^36 light paths");
    if (sc->stats.max_depth + 1 > 64) return Fail(GY_ERR_UNSUPPORTED, "gy_render_light_tracing: BVH deeper than the traversal stack");
    PathParams P;
    rc = MakePathParams(c, sc, *cam, *p, &P);
    if (rc != GY_OK) return rc;
    LightCameraRT lc;
    rc = BuildLightCamera(*cam, &lc);
    if (rc != GY_OK) return rc;
    LightCamDev C;
    for (int i = 0; i < 16; ++i) C.m[i] = lc.world_to_raster[i];
    for (int i = 0; i < 3; ++i) {
        C.pos[i] = lc.pos[i];
        C.cz[i] = lc.cz[i];
    }
    C.dis = lc.dis;
    C.area = lc.area;
    C.width = cam->width;
    C.height = cam->height;
    // AABB::GetBoundingSphere (math/AABB.h:69-72): (min + max) * 0.5
    for (int i = 0; i < 3; ++i) C.env_c[i] = (sc->world_lo[i] + sc->world_hi[i]) * 0.5f;
    const int64_t px = (int64_t)cam->width * cam->height * 3;
    void* d = nullptr;
    rc = DevAlloc(c, &d, (size_t)px * 4);
    if (rc == GY_OK) rc = Fill32(c, d, 0u, px);
    if (rc == GY_OK && p->mode == GY_RENDER_WAVEFRONT) {
        rc = gy::LightTracingWavefront(c, sc, P, C, n, (float)n, (float*)d);
    } else if (rc == GY_OK) {
        // block size: p->block_size (64, 128 or 256); 0 = 64, the fastest on the
        // 1.31M-triangle workload (86.6 Mpaths/s against 84.1 at 128, 81.4 at 256)
        const int bs = (p->block_size == 128 || p->block_size == 256) ? p->block_size : 64;
        const unsigned blocks = (unsigned)((n + bs - 1) / bs);
        void (*kern)(SceneDev, PathParams, LightCamDev, int64_t, float, float*);
        if (sc->dev.env.w > 0)
            kern = sc->stack <= 32 ? (bs == 64 ? k_lt<32, 64, true> : bs == 256 ? k_lt<32, 256, true> : k_lt<32, 128, true>)
                                   : (bs == 64 ? k_lt<64, 64, true> : bs == 256 ? k_lt<64, 256, true> : k_lt<64, 128, true>);
        else
            kern = sc->stack <= 32 ? (bs == 64 ? k_lt<32, 64, false> : bs == 256 ? k_lt<32, 256, false> : k_lt<32, 128, false>)
                                   : (bs == 64 ? k_lt<64, 64, false> : bs == 256 ? k_lt<64, 256, false> : k_lt<64, 128, false>);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(bs), 0, c->stream, sc->dev, P, C, n, (float)n, (float*)d);
        rc = CheckLaunch(c, "k_lt");
    }
    if (rc == GY_OK) {
        hipError_t e = hipMemcpyAsync(rgb, d, (size_t)px * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = HipFail(e, "gy_render_light_tracing: download");
    }
    DevFree(d);
    return rc;
}

extern "C" int gy_scene_destroy(gy_scene* s) {
    if (!s) return GY_OK;
    Scene* sc = reinterpret_cast<Scene*>(s);
    // a batch's entry table and recorded graph hold this scene's device pointers
    if (sc->batch_refs > 0) return Fail(GY_ERR_INVALID, "gy_scene_destroy: a batch still renders this scene (destroy the batch first)");
    (void)hipStreamSynchronize(sc->ctx->stream);
    SceneFree(sc);
    return GY_OK;
}

extern "C" int gy_scene_context(gy_scene* s, gy_context** ctx) {
    if (!s || !ctx) return Fail(GY_ERR_INVALID, "gy_scene_context: null argument");
    *ctx = reinterpret_cast<gy_context*>(reinterpret_cast<Scene*>(s)->ctx);
    return GY_OK;
}

extern "C" int gy_scene_stats_get(gy_scene* s, gy_scene_stats* out) {
    if (!s || !out) return Fail(GY_ERR_INVALID, "gy_scene_stats_get: null argument");
    *out = reinterpret_cast<Scene*>(s)->stats;
    return GY_OK;
}

extern "C" int gy_scene_bvh_download(gy_scene* s, uint32_t* nodes, int64_t max_nodes, int32_t* tri_order, int64_t max_tris) {
    if (!s) return Fail(GY_ERR_INVALID, "gy_scene_bvh_download: null scene");
    Scene* sc = reinterpret_cast<Scene*>(s);
    const int64_t nn = (int64_t)sc->host_nodes.size() / 16;
    if (nodes) {
        if (max_nodes < nn) return Fail(GY_ERR_INVALID, "gy_scene_bvh_download: node buffer too small");
        std::memcpy(nodes, sc->host_nodes.data(), sc->host_nodes.size() * 4);
    }
    if (tri_order) {
        if (max_tris < (int64_t)sc->tri_order.size()) return Fail(GY_ERR_INVALID, "gy_scene_bvh_download: order buffer too small");
        for (size_t i = 0; i < sc->tri_order.size(); ++i) tri_order[i] = (int32_t)sc->tri_order[i];
    }
    return GY_OK;
}

static int TraceHost(gy_scene* s, const gy_ray* rays, int64_t n, void* out, bool any) {
    if (!s || (n > 0 && (!rays || !out)) || n < 0) return Fail(GY_ERR_INVALID, "trace: bad argument");
    if (n == 0) return GY_OK;
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    int rc = Grow(c, &c->rays, (size_t)n * sizeof(gy_ray));
    if (rc == GY_OK) rc = Grow(c, &c->hits, (size_t)n * sizeof(gy_hit));
    if (rc != GY_OK) return rc;
    hipError_t e = hipMemcpyAsync(c->rays.ptr, rays, (size_t)n * sizeof(gy_ray), hipMemcpyHostToDevice, c->stream);
    if (e != hipSuccess) return HipFail(e, "trace H2D");
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, c->stream);
    rc = any ? LaunchTraceAny(c, sc, (const gy_ray*)c->rays.ptr, n, (int32_t*)c->hits.ptr)
             : LaunchTraceClosest(c, sc, (const gy_ray*)c->rays.ptr, n, (gy_hit*)c->hits.ptr);
    (void)hipEventRecord(b, c->stream);
    if (rc == GY_OK) {
        e = hipMemcpyAsync(out, c->hits.ptr, (size_t)n * (any ? 4 : sizeof(gy_hit)), hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) rc = HipFail(e, "trace D2H");
    }
    e = hipStreamSynchronize(c->stream);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "trace sync");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    c->last_kernel_ms = ms;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

extern "C" int gy_trace_closest(gy_scene* s, const gy_ray* rays, int64_t n, gy_hit* hits) { return TraceHost(s, rays, n, hits, false); }
extern "C" int gy_trace_any(gy_scene* s, const gy_ray* rays, int64_t n, int32_t* occ) { return TraceHost(s, rays, n, occ, true); }

static int TraceDevice(gy_scene* s, const gy_ray* rays, int64_t n, void* out, bool any) {
    if (!s || n < 0 || (n > 0 && (!rays || !out))) return Fail(GY_ERR_INVALID, "trace_device: bad argument");
    if (n == 0) return GY_OK;
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, c->stream);
    int rc = any ? LaunchTraceAny(c, sc, rays, n, (int32_t*)out) : LaunchTraceClosest(c, sc, rays, n, (gy_hit*)out);
    (void)hipEventRecord(b, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "trace sync");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    c->last_kernel_ms = ms;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}
extern "C" int gy_trace_closest_device(gy_scene* s, const gy_ray* rays, int64_t n, gy_hit* hits) {
    return TraceDevice(s, rays, n, hits, false);
}
extern "C" int gy_trace_any_device(gy_scene* s, const gy_ray* rays, int64_t n, int32_t* occ) { return TraceDevice(s, rays, n, occ, true); }

extern "C" double gy_last_kernel_ms(gy_context* ctx) { return ctx ? reinterpret_cast<Ctx*>(ctx)->last_kernel_ms : -1.0; }

extern "C" int gy_count_visits(gy_scene* s, const gy_ray* rays, int64_t n, int64_t* node_visits, int64_t* tri_visits) {
    if (!s || n < 0 || (n > 0 && !rays) || !node_visits || !tri_visits) return Fail(GY_ERR_INVALID, "gy_count_visits: bad argument");
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    int rc = Grow(c, &c->rays, (size_t)std::max<int64_t>(n, 1) * sizeof(gy_ray));
    if (rc == GY_OK) rc = Grow(c, &c->ctr, 64);
    if (rc != GY_OK) return rc;
    if (n > 0) {
        hipError_t e = hipMemcpyAsync(c->rays.ptr, rays, (size_t)n * sizeof(gy_ray), hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return HipFail(e, "count H2D");
    }
    rc = Fill32(c, c->ctr.ptr, 0u, 4);
    if (rc != GY_OK) return rc;
    if (n > 0) {
        const bool sph = sc->dev.n_sph > 0;
        auto kern = sc->stack <= 32 ? (sph ? k_count_visits<32, true> : k_count_visits<32, false>)
                                    : (sph ? k_count_visits<64, true> : k_count_visits<64, false>);
        hipLaunchKernelGGL(kern, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, sc->dev,
                           (const gy_ray*)c->rays.ptr, n, (unsigned long long*)c->ctr.ptr);
        rc = CheckLaunch(c, "k_count_visits");
        if (rc != GY_OK) return rc;
    }
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemcpyAsync(h, c->ctr.ptr, 16, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return HipFail(e, "count D2H");
    *node_visits = (int64_t)h[0];
    *tri_visits = (int64_t)h[1];
    return GY_OK;
}

extern "C" int gy_render(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p, float* color, float* weight) {
    if (!s) return Fail(GY_ERR_INVALID, "gy_render: null scene");
    int rc = ValidateRender(cam, p);
    if (rc != GY_OK) return rc;
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    PathParams P;
    rc = MakePathParams(c, sc, *cam, *p, &P);
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
    int block = p->block_size > 0 ? p->block_size : 128;
    // scenes with spheres render with the per-path kernels (same film)
    const bool wavefront = p->mode == GY_RENDER_WAVEFRONT && sc->dev.n_sph == 0;
    for (int s0 = 0; s0 < p->spp && rc == GY_OK; s0 += n_pass) {
        const int n_s = std::min(n_pass, p->spp - s0);
        if (wavefront) rc = WavefrontPass(c, sc, P, *p, s0, n_s);
        else rc = LaunchMega(c, sc, P, s0, n_s, block, false);
        if (rc == GY_OK) rc = FilmResolvePass(c, sc, n_s);
    }
    (void)hipEventRecord(t1, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "render sync");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    TimersCollect(c);
    c->stats.total_ms = ms;
    c->stats.paths = (int64_t)cam->width * cam->height * p->spp;
    if (rc == GY_OK) {
        unsigned long long h[8];
        e = hipMemcpy(h, sc->film.counters, sizeof(h), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return HipFail(e, "render counters");
        rc = FilmCheck(sc->film);
        if (rc != GY_OK) return rc;
        c->stats.rays = (int64_t)(h[0] + h[1] + h[2]);
        c->stats.shadow_rays = (int64_t)h[1];
        c->stats.mis_rays = (int64_t)h[2];
    }
    if (rc == GY_OK) {
        // what the film on the device holds (the adjoint replays exactly this render)
        sc->film_valid = true;
        sc->film_single_pass = n_pass >= p->spp;
        sc->film_cam = *cam;
        sc->film_params = *p;
    } else {
        sc->film_valid = false;
    }
    if (rc == GY_OK && (color || weight)) rc = gy_film_download(s, color, weight);
    return rc;
}

// gy_render without allocating: megakernel form, film into the caller's
// device buffer, counters, border flags and pass buffer in this device's
// scratch array.  The film equals gy_render's bit for bit: a pixel sums its
// samples in sample order whatever the pass split.
extern "C" int gy_render_noalloc(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p, float* film_device) {
    if (!s || !film_device) return Fail(GY_ERR_INVALID, "gy_render_noalloc: null scene or film");
    int rc = ValidateRender(cam, p);
    if (rc != GY_OK) return rc;
    if (p->mode != GY_RENDER_MEGAKERNEL)
        return Fail(GY_ERR_UNSUPPORTED, "gy_render_noalloc: megakernel form only (the wavefront's path state is allocated)");
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    if (!c->scratch) return Fail(GY_ERR_DEVICE, "gy_render_noalloc: no scratch address for this device");
    (void)hipSetDevice(c->device);
    PathParams P;
    rc = MakePathParams(c, sc, *cam, *p, &P);
    if (rc != GY_OK) return rc;
    // scratch layout: [0, 256) ray counters, then W*H + 1 border words (256-byte
    // aligned), then pass slots of W*H float4 per pass sample
    const int64_t WH = (int64_t)cam->width * cam->height;
    // 20 bytes per pixel (one pass sample and its border word) and at most
    // 515 of alignment and counters: max_pixels always fits
    const size_t max_pixels = (kNoAllocScratchBytes - 1024) / 20;
    const size_t pass_off = 256 + ((((size_t)WH + 1) * 4 + 255) & ~(size_t)255);
    if ((size_t)WH > max_pixels || pass_off + (size_t)WH * 16 > kNoAllocScratchBytes) {
        char msg[160];
        std::snprintf(msg, sizeof(msg), "gy_render_noalloc: %lld pixels exceed the %zu MiB scratch (at most %zu pixels)",
                      (long long)WH, kNoAllocScratchBytes >> 20, max_pixels);
        return Fail(GY_ERR_UNSUPPORTED, msg);
    }
    const int64_t cap = (int64_t)((kNoAllocScratchBytes - pass_off) / ((size_t)WH * 16));
    int n_pass;
    if (p->samples_per_pass > 0) {
        n_pass = std::min(p->samples_per_pass, p->spp);
        if (n_pass > cap) {
            char msg[160];
            std::snprintf(msg, sizeof(msg), "gy_render_noalloc: samples_per_pass %d exceeds the scratch (at most %lld at %lld pixels)",
                          n_pass, (long long)cap, (long long)WH);
            return Fail(GY_ERR_UNSUPPORTED, msg);
        }
    } else {
        n_pass = (int)std::min<int64_t>(SamplesPerPass(*cam, *p), cap);
    }
    char* base = (char*)c->scratch;
    FilmState f;
    f.film = (float4*)film_device;
    f.counters = (unsigned long long*)base;
    f.border = (uint32_t*)(base + 256);
    f.pass = (float4*)(base + pass_off);
    f.pass_slots = cap * WH;
    f.width = cam->width;
    f.height = cam->height;
    rc = FilmClear(c, f);
    const int block = p->block_size > 0 ? p->block_size : 128;
    for (int s0 = 0; s0 < p->spp && rc == GY_OK; s0 += n_pass) {
        const int n_s = std::min(n_pass, p->spp - s0);
        rc = LaunchMegaFilm(c, sc, f, P, s0, n_s, block, false, false);
        if (rc == GY_OK) rc = FilmResolveState(c, f, n_s, false);
    }
    hipError_t e = hipStreamSynchronize(c->stream);
    if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "gy_render_noalloc sync");
    if (rc == GY_OK) rc = FilmCheck(f);
    return rc;
}

extern "C" int gy_film_download(gy_scene* s, float* color, float* weight) {
    if (!s) return Fail(GY_ERR_INVALID, "gy_film_download: null scene");
    Scene* sc = reinterpret_cast<Scene*>(s);
    if (!sc->film.film) return Fail(GY_ERR_INVALID, "gy_film_download: nothing rendered");
    const int64_t WH = (int64_t)sc->film.width * sc->film.height;
    std::vector<float4> h(WH);
    hipError_t e = hipMemcpy(h.data(), sc->film.film, WH * sizeof(float4), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return HipFail(e, "film D2H");
    for (int64_t i = 0; i < WH; ++i) {
        if (color) {
            color[3 * i] = h[i].x;
            color[3 * i + 1] = h[i].y;
            color[3 * i + 2] = h[i].z;
        }
        if (weight) weight[i] = h[i].w;
    }
    return GY_OK;
}

// Film::SaveImage (visual/Film.cpp:6-9): pixel = color / weight + splat
// (visual/Film.h:128-136), where Vec3 / 0 is the zero vector (math/Linagl.h:
// 366-369) and the splat stays zero for path tracing.
extern "C" int gy_film_save(gy_scene* s, const char* path) {
    if (!s || !path) return Fail(GY_ERR_INVALID, "gy_film_save: null argument");
    Scene* sc = reinterpret_cast<Scene*>(s);
    if (!sc->film.film) return Fail(GY_ERR_INVALID, "gy_film_save: nothing rendered");
    const int64_t WH = (int64_t)sc->film.width * sc->film.height;
    std::vector<float4> h(WH);
    hipError_t e = hipMemcpy(h.data(), sc->film.film, WH * sizeof(float4), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return HipFail(e, "film D2H");
    std::vector<float> rgb((size_t)WH * 3);
    for (int64_t i = 0; i < WH; ++i) {
        const float w = h[i].w;
        const float c[3] = {h[i].x, h[i].y, h[i].z};
        for (int k = 0; k < 3; ++k) rgb[3 * i + k] = (w == 0.f ? 0.f : c[k] / w) + 0.f;
    }
    return WriteImage(path, rgb.data(), sc->film.width, sc->film.height);
}

extern "C" int gy_film_rmse(gy_scene* s, const float* reference, const int32_t* rows, int32_t n_rows, double* rmse) {
    if (!s || !reference || !rmse) return Fail(GY_ERR_INVALID, "gy_film_rmse: null argument");
    Scene* sc = reinterpret_cast<Scene*>(s);
    if (!sc->film.film) return Fail(GY_ERR_INVALID, "gy_film_rmse: nothing rendered");
    const int W = sc->film.width, H = sc->film.height;
    if (!rows) n_rows = H;
    if (n_rows <= 0) return Fail(GY_ERR_INVALID, "gy_film_rmse: no rows");
    for (int32_t i = 0; rows && i < n_rows; ++i)
        if (rows[i] < 0 || rows[i] >= H) return Fail(GY_ERR_INVALID, "gy_film_rmse: row outside the image");
    std::vector<float4> line(W);
    double sum = 0.0;
    for (int32_t i = 0; i < n_rows; ++i) {
        const int y = rows ? rows[i] : i;
        hipError_t e = hipMemcpy(line.data(), sc->film.film + (int64_t)y * W, W * sizeof(float4), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return HipFail(e, "film D2H");
        const float* ref = reference + (int64_t)i * W * 3;
        for (int x = 0; x < W; ++x) {
            const float w = line[x].w;
            const float c[3] = {line[x].x, line[x].y, line[x].z};
            for (int k = 0; k < 3; ++k) {
                const double d = (double)(w == 0.f ? 0.f : c[k] / w) - (double)ref[3 * x + k];
                sum += d * d;
            }
        }
    }
    *rmse = std::sqrt(sum / ((double)n_rows * W * 3));
    return GY_OK;
}

extern "C" int gy_render_stats_get(gy_context* ctx, gy_render_stats* out) {
    if (!ctx || !out) return Fail(GY_ERR_INVALID, "gy_render_stats_get: null argument");
    *out = reinterpret_cast<Ctx*>(ctx)->stats;
    return GY_OK;
}

extern "C" const char* gy_kernel_name(int i) {
    static const char* names[8] = {"pt_megakernel", "film_resolve", "wf_raygen", "wf_trace_closest", "wf_shade",
                                   "wf_trace_shadow_mis", "wf_trace_mis", "wf_other"};  // [6]: no longer timed apart
    return (i >= 0 && i < 8) ? names[i] : "";
}

extern "C" int gy_li_samples(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p, int32_t n, const int32_t* x,
                             const int32_t* y, const int32_t* smp, float* L, float* raster, int32_t* dims) {
    if (!s || n < 0 || (n > 0 && (!x || !y || !smp || !L || !raster || !dims))) return Fail(GY_ERR_INVALID, "gy_li_samples: bad argument");
    int rc = ValidateRender(cam, p);
    if (rc != GY_OK) return rc;
    if (n == 0) return GY_OK;
    for (int32_t i = 0; i < n; ++i)
        if (x[i] < 0 || x[i] >= cam->width || y[i] < 0 || y[i] >= cam->height || smp[i] < 0 || smp[i] >= p->spp)
            return Fail(GY_ERR_INVALID, "gy_li_samples: sample outside the image or spp");
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    PathParams P;
    rc = MakePathParams(c, sc, *cam, *p, &P);
    if (rc != GY_OK) return rc;
    void *dx = nullptr, *dy = nullptr, *ds = nullptr, *dL = nullptr, *dr = nullptr, *dd = nullptr;
    rc = Upload(c, &dx, x, n * 4);
    if (rc == GY_OK) rc = Upload(c, &dy, y, n * 4);
    if (rc == GY_OK) rc = Upload(c, &ds, smp, n * 4);
    if (rc == GY_OK) rc = DevAlloc(c, &dL, (size_t)n * 12);
    if (rc == GY_OK) rc = DevAlloc(c, &dr, (size_t)n * 8);
    if (rc == GY_OK) rc = DevAlloc(c, &dd, (size_t)n * 4);
    if (rc == GY_OK) {
        const unsigned blocks = (unsigned)((n + 127) / 128);
        const bool env = sc->dev.env.w > 0;
        auto kern = sc->stack <= 32 ? (env ? k_li_samples<32, true, false> : k_li_samples<32, false, false>)
                                    : (env ? k_li_samples<64, true, false> : k_li_samples<64, false, false>);
        if (sc->dev.n_sph > 0)
            kern = sc->stack <= 32 ? (env ? k_li_samples<32, true, true> : k_li_samples<32, false, true>)
                                   : (env ? k_li_samples<64, true, true> : k_li_samples<64, false, true>);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(128), 0, c->stream, sc->dev, P, n, (const int32_t*)dx, (const int32_t*)dy,
                           (const int32_t*)ds, (float*)dL, (float*)dr, (int32_t*)dd);
        rc = CheckLaunch(c, "k_li_samples");
    }
    if (rc == GY_OK) {
        hipError_t e = hipMemcpyAsync(L, dL, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(raster, dr, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(dims, dd, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = HipFail(e, "li_samples D2H");
    }
    DevFree(dx);
    DevFree(dy);
    DevFree(ds);
    DevFree(dL);
    DevFree(dr);
    DevFree(dd);
    return rc;
}

extern "C" int gy_count_traffic(gy_scene* s, const gy_camera_desc* cam, const gy_render_params* p, gy_render_stats* out) {
    if (!s || !out) return Fail(GY_ERR_INVALID, "gy_count_traffic: null argument");
    int rc = ValidateRender(cam, p);
    if (rc != GY_OK) return rc;
    Scene* sc = reinterpret_cast<Scene*>(s);
    Ctx* c = sc->ctx;
    (void)hipSetDevice(c->device);
    PathParams P;
    rc = MakePathParams(c, sc, *cam, *p, &P);
    if (rc != GY_OK) return rc;
    const int n_pass = SamplesPerPass(*cam, *p);
    sc->film_valid = false;  // the counting passes overwrite the film
    rc = FilmPrepare(c, sc, cam->width, cam->height, n_pass);
    if (rc != GY_OK) return rc;
    for (int s0 = 0; s0 < p->spp && rc == GY_OK; s0 += n_pass) rc = LaunchMega(c, sc, P, s0, std::min(n_pass, p->spp - s0), 128, true);
    if (rc != GY_OK) return rc;
    unsigned long long h[16];
    hipError_t e = hipMemcpyAsync(h, sc->film.counters, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return HipFail(e, "count D2H");
    std::memset(out, 0, sizeof(*out));
    out->paths = (int64_t)cam->width * cam->height * p->spp;
    out->rays = (int64_t)(h[0] + h[1] + h[2]);
    out->shadow_rays = (int64_t)h[1];
    out->mis_rays = (int64_t)h[2];
    out->shade_points = (int64_t)h[3];
    for (int k = 0; k < 3; ++k) {
        out->node_visits_kind[k] = (int64_t)h[4 + k];
        out->tri_visits_kind[k] = (int64_t)h[7 + k];
        out->node_visits += (int64_t)h[4 + k];
        out->tri_visits += (int64_t)h[7 + k];
    }
    return GY_OK;
}

// Plain device-to-device copy of a known size: the HBM reference for counter
// runs (FETCH_SIZE / WRITE_SIZE of this kernel against its byte count) and the
// copy bandwidth reachable on the device.
__global__ void k_copy_probe(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

extern "C" int gy_copy_probe(gy_context* ctx, int64_t bytes, double* ms_out) {
    if (!ctx || bytes <= 0 || (bytes & 15) || !ms_out) return Fail(GY_ERR_INVALID, "gy_copy_probe: bytes must be a positive multiple of 16");
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    (void)hipSetDevice(c->device);
    void *src = nullptr, *dst = nullptr;
    int rc = DevAlloc(c, &src, (size_t)bytes);
    if (rc == GY_OK) rc = DevAlloc(c, &dst, (size_t)bytes);
    if (rc == GY_OK) rc = Fill32(c, src, 0x3f800000u, bytes / 4);
    const int64_t n = bytes / 16;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, (int64_t)std::max(c->num_cu, 1) * 32);
    hipEvent_t a = nullptr, b = nullptr;
    if (rc == GY_OK) {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, c->stream, (const float4*)src, (float4*)dst, n);
        (void)hipEventRecord(a, c->stream);
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, c->stream, (const float4*)src, (float4*)dst, n);
        (void)hipEventRecord(b, c->stream);
        rc = CheckLaunch(c, "copy probe");
        hipError_t e = hipStreamSynchronize(c->stream);
        if (rc == GY_OK && e != hipSuccess) rc = HipFail(e, "copy probe sync");
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        *ms_out = ms;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    DevFree(src);
    DevFree(dst);
    return rc;
}

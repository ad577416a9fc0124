// Device building blocks of the path-tracing integrator (HIP, gfx950).
//
// FP32 arithmetic follows the reference's SSE build operation for operation
// (this translation unit is compiled with -ffp-contract=off and correctly
// rounded divide/sqrt): Dot = (x x' + y y') + (z z' + 0) (math/Linagl.h:496-500),
// Norm = v / sqrt(Dot) (Linagl.h:527-531), unary minus = 0 - v, Vec / s = 0 when
// s == 0 (Linagl.h:366-369).  Citations are file:line in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GY_D __device__ __forceinline__

namespace gyd {

constexpr float kPI = 3.14159265358979f;
constexpr float kInvPI = 0.31830988618379067154f;
constexpr float kPIOver2 = 1.57079632679489661923f;
constexpr float kPIOver4 = 0.78539816339744830961f;
constexpr float kRayEps = 1e-4f;       // common/Core.h:193-194
constexpr float kShadowRayEps = 1e-4f;
constexpr float kMachineEps = 5.96046448e-08f;  // FLT_EPSILON * 0.5
constexpr float kInf = __builtin_huge_valf();

struct V3 {
    float x, y, z;
};
GY_D V3 v3(float x, float y, float z) { return V3{x, y, z}; }
GY_D V3 v3s(float a) { return V3{a, a, a}; }
GY_D V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
GY_D V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
GY_D V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
GY_D V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
GY_D V3 operator*(float s, V3 a) { return a * s; }
GY_D V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
GY_D V3 operator/(V3 a, float s) {
    if (s == 0.f) return V3{0.f, 0.f, 0.f};
    return V3{a.x / s, a.y / s, a.z / s};
}
GY_D V3 neg(V3 a) { return V3{0.f - a.x, 0.f - a.y, 0.f - a.z}; }
GY_D bool eq(V3 a, V3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
GY_D bool is_zero(V3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
GY_D float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + (a.z * b.z + 0.f); }
GY_D V3 cross(V3 a, V3 b) { return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
GY_D float length(V3 a) { return sqrtf(dot(a, a)); }
GY_D V3 norm(V3 a) {
    const float l = length(a);
    return V3{a.x / l, a.y / l, a.z / l};
}
GY_D V3 vabs(V3 a) { return V3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
GY_D V3 vsqrt(V3 a) { return V3{sqrtf(a.x), sqrtf(a.y), sqrtf(a.z)}; }
GY_D float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
// Vector::Y (math/Linagl.h:569-572)
GY_D float lumY(V3 a) { return 0.212671f * a.x + 0.715160f * a.y + 0.072169f * a.z; }
GY_D float rmin(float a, float b) { return (b < a) ? b : a; }  // std::min
GY_D float rmax(float a, float b) { return (a < b) ? b : a; }  // std::max
GY_D float sgnf(float v) { return (v > 0.f) ? 1.f : -1.f; }     // math/MathUtils.h:130-132
GY_D float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
GY_D float gamma_n(int n) { return (n * kMachineEps) / (1 - n * kMachineEps); }  // numeric/NumericUtils.h:49-51
GY_D V3 ld3(const float4 q) { return V3{q.x, q.y, q.z}; }
// cos / sin / log evaluated in double and rounded once to float, which the
// oracle does too (the reference's CRT float versions are implementation-defined
// in the last ulp).
GY_D float dcos(float x) { return (float)cos((double)x); }
GY_D float dsin(float x) { return (float)sin((double)x); }
GY_D float dlog(float x) { return (float)log((double)x); }
GY_D float dacos(float x) { return (float)acos((double)x); }
GY_D float datan2(float y, float x) { return (float)atan2((double)y, (double)x); }

// ------------------------------------------------------------------ Sobol
// sobol::sample (sampler/Sobol.h:46-61) + RandomStateSequence (sampler/Sampler.h:53-64)
// T: byte tables, 7 x 256 words per dimension, entry [k][v] = XOR of the
// matrix columns 8k + j for the set bits j of v.  The XOR over the set bits of
// the index splits into one entry per index byte (indices stay below 2^52).
// The byte loop is kept rolled: unrolled, the compiler hoists the lookups of
// every sample of a path vertex at once, which costs long kernels (the adjoint
// replay) registers and occupancy.
GY_D float sobol_sample(const uint32_t* __restrict__ T, uint64_t index, uint32_t dim) {
    const uint32_t* t = T + (size_t)dim * (7 * 256);
    uint32_t r = 0;
#pragma unroll 1
    for (; index; index >>= 8, t += 256) r ^= t[index & 255u];
    return (float)r * 2.3283064365386963e-10f;
}
// The same sum with the four low index bytes looked up at once: entry 0 of a
// byte table is 0, so a zero byte adds nothing, and the four loads are
// independent.  One memory round trip per sample for indices below 2^32 in
// place of one per byte, for kernels whose samples sit on the critical path.
GY_D float sobol_sample_wide(const uint32_t* __restrict__ T, uint64_t index, uint32_t dim) {
    const uint32_t* t = T + (size_t)dim * (7 * 256);
    const uint32_t lo = (uint32_t)index;
    uint32_t r = (t[lo & 255u] ^ t[256 + ((lo >> 8) & 255u)]) ^ (t[512 + ((lo >> 16) & 255u)] ^ t[768 + (lo >> 24)]);
    uint32_t hi = (uint32_t)(index >> 32);
    t += 1024;
#pragma unroll 1
    for (; hi; hi >>= 8, t += 256) r ^= t[hi & 255u];
    return (float)r * 2.3283064365386963e-10f;
}
// HaltonSampler::sample with Faure permutations (sampler/Halton.h:648-910,
// 1441-3308): base 2 writes the bit-reversed index into a float mantissa;
// other bases sum digit-table entries over groups of k digits (host tables,
// halton.cpp) and scale the integer by float((1 - 2^-23) / tsize^groups).
struct SampleSrc {
    const uint32_t* __restrict__ sobol;  // Sobol byte tables, 7 x 256 words per dimension
    const uint16_t* __restrict__ htab;   // Halton digit tables
    const uint4* __restrict__ hdim;      // per dimension: base, table size, groups, table offset
    const float* __restrict__ hscale;    // per dimension: scale
    int32_t kind;                        // GY_SAMPLER_SOBOL / GY_SAMPLER_HALTON
};
GY_D float halton_sample(const SampleSrc& M, uint64_t index, uint32_t dim) {
    if (dim == 0) return __uint_as_float(0x3f800000u | (uint32_t)(__brevll(index) >> 41)) - 1.f;
    if (dim >= 256u) return 0.f;
    const uint4 di = M.hdim[dim];
    const uint16_t* t = M.htab + di.w;
    const uint32_t ts = di.y;
    uint32_t sum = 0;  // < 2^32 by construction of the groups
    if ((index >> 32) == 0) {
        uint32_t x = (uint32_t)index;
        for (uint32_t g = 0; g < di.z; ++g) {
            sum = sum * ts + t[x % ts];
            x /= ts;
        }
    } else {
        uint64_t x = index;
        for (uint32_t g = 0; g < di.z; ++g) {
            sum = sum * ts + t[x % ts];
            x /= ts;
        }
    }
    return (float)sum * M.hscale[dim];
}
struct Seq {
    uint64_t index;
    uint32_t cursor;
};
// RandomStateSequence::Sample (sampler/Sampler.h:53-64): values >= 1 wrap to 0.
GY_D float seq_next(const SampleSrc& M, Seq& s) {
    float v = M.kind == 1 ? halton_sample(M, s.index, s.cursor) : sobol_sample(M.sobol, s.index, s.cursor);
    ++s.cursor;
    if (v >= 1.f) v = 0.f;
    return v;
}
GY_D float seq_next_wide(const SampleSrc& M, Seq& s) {
    float v = M.kind == 1 ? halton_sample(M, s.index, s.cursor) : sobol_sample_wide(M.sobol, s.index, s.cursor);
    ++s.cursor;
    if (v >= 1.f) v = 0.f;
    return v;
}

// SobolEnum::GetIndex (sampler/SobolEnum.h:24-28) from derived tables
// (sobol.cpp), or HaltonEnum::get_index (sampler/HaltonEnum.h:96-103).  sx, sy
// scale a sample in [0,1) to the enumeration grid (SampleX / SampleY).
struct EnumDev {
    uint32_t m;
    uint32_t kind;
    uint64_t inv[32];
    uint64_t frame[32];
    uint32_t p2, p3;
    uint64_t hmx, hmy, hinc;
    float sx, sy;
};
GY_D uint64_t enum_index(const EnumDev& E, uint32_t frame, uint32_t px, uint32_t py) {
    if (E.kind == 1) {
        const uint64_t hx = E.p2 ? (__brevll((uint64_t)px) >> (64 - E.p2)) : 0ull;
        uint32_t hy = 0, y = py;
        for (uint32_t d = 0; d < E.p3; ++d) {
            hy = hy * 3 + y % 3;
            y /= 3;
        }
        return (hx * E.hmx + (uint64_t)hy * E.hmy) % E.hinc + (uint64_t)frame * E.hinc;
    }
    const uint32_t m = E.m;
    uint64_t delta = 0;
    for (uint32_t j = 0, f = frame; f; f >>= 1, ++j)
        if (f & 1u) delta ^= E.frame[j];
    uint64_t target = (((uint64_t)px << m) | py) ^ delta;
    uint64_t lo = 0;
    for (uint32_t b = 0; target; target >>= 1, ++b)
        if (target & 1u) lo ^= E.inv[b];
    return ((uint64_t)frame << (2 * m)) | lo;
}

// ------------------------------------------------------------------ camera
struct CamDev {
    float m[16];  // RasterToWorld, column-major
    float pos[3];
    int32_t width, height, res, log2res;
};
// PinHoleCamera::GenerateRay (camera/PinholeCamera.h:22-29) with
// Transform::operator()(point) (math/Transform.h:106-115) and the Ray
// constructor's RayEps origin shift (math/Ray.h:17-19).
GY_D void generate_ray(const CamDev& c, int px, int py, float u, float v, V3* o, V3* d) {
    const float in[4] = {px + u, py + v, -1.f, 1.0f};
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = c.m[12 + i] * in[3];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = c.m[8 + i] * in[2] + r[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = c.m[4 + i] * in[1] + r[i];
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = c.m[0 + i] * in[0] + r[i];
    V3 p = v3(r[0], r[1], r[2]);
    if (!(r[3] == 1.f || r[3] == 0.f)) p = p / r[3];
    const V3 pos = v3(c.pos[0], c.pos[1], c.pos[2]);
    const V3 dir = norm(p - pos);
    *o = (pos + dir * 0.0f) + dir * kRayEps;
    *d = dir;
}

// ------------------------------------------------------------------ scene
// Triangle records for intersection, in BVH leaf order: 3 x float4 =
// (p0, prim id), (p1, 0), (p2, 0).
// Shading records by primitive id: 8 x float4 =
// (p0, mat), (p1, light), (p2, area), (n0, u0), (n1, v0), (n2, u1), (fn, v1), (u2, v2, 0, 0).
// BVH nodes: 4 x uint4 (both child boxes in the parent, see gy_internal.h).
// EnvironmentLight over an EnvironmentMap (light/EnvLight.cpp, light/EnvMap.h).
struct EnvDev {
    const float* __restrict__ tex;    // RGB texels, ImageTexture rows: the image's last row first
    const float* __restrict__ cfunc;  // Distribution2D conditionals, one per image row (top first): w values
    const float* __restrict__ ccdf;   // ... w + 1 CDF values a row
    const float* __restrict__ cint;   // ... integral a row = the marginal's function
    const float* __restrict__ mcdf;   // marginal CDF, h + 1 values
    float mint;                       // marginal integral
    float scale;                      // EnvironmentMap::mScale
    float radius;                     // scene bounding-sphere radius
    int32_t w, h;                     // w == 0: no environment light
};

struct SceneDev {
    EnvDev env;
    const float4* __restrict__ tri;
    const float4* __restrict__ shade;
    const uint4* __restrict__ nodes;
    const uint32_t* __restrict__ mats;
    const float* __restrict__ tex;
    const int4* __restrict__ tex_info;
    const int32_t* __restrict__ light_tri;
    const float* __restrict__ light_le;    // 3 per light
    const float* __restrict__ light_cdf;   // n_light + 1
    const float* __restrict__ light_func;  // n_light
    float light_func_int;
    int32_t n_light;
    int32_t n_tri;
    int32_t n_nodes;
    // analytic spheres (primitive ids n_tri + i): two float4 each, centre and
    // radius; inside flag, material, light id (-1: none) as integer bits and
    // the area 4 pi r^2
    const float4* __restrict__ sph;
    int32_t n_sph;
};

struct Hit {
    float t, b1, b2;
    int32_t prim;
};

// Triangle::Intersect (shape/Triangle.cpp:71-103): accepts tmin <= t <= tmax.
GY_D bool tri_test(V3 p0, V3 e1, V3 e2, V3 o, V3 d, float tmin, float tmax, float* t, float* b1o, float* b2o) {
    const V3 s1 = cross(d, e2);
    const float divisor = dot(s1, e1);
    if (divisor == 0.f) return false;
    const float inv = 1.f / divisor;
    const V3 s = o - p0;
    const float b1 = dot(s, s1) * inv;
    if (b1 < 0.f || b1 > 1.f) return false;
    const V3 s2 = cross(s, e1);
    const float b2 = dot(d, s2) * inv;
    if (b2 < 0.f || b1 + b2 > 1.f) return false;
    const float th = dot(e2, s2) * inv;
    if (th < tmin || th > tmax) return false;
    *t = th;
    *b1o = b1;
    *b2o = b2;
    return true;
}

// Slab test against a padded child box.  NaN from 0 * inf (a ray parallel to
// and lying in a slab plane) is dropped by fminf/fmaxf, which keeps the test
// conservative.  Returns the entry distance or +inf on a miss.
GY_D float box_enter(const float* b, V3 o, V3 inv, float tmin, float tmax) {
    const float tx0 = (b[0] - o.x) * inv.x, tx1 = (b[3] - o.x) * inv.x;
    const float ty0 = (b[1] - o.y) * inv.y, ty1 = (b[4] - o.y) * inv.y;
    const float tz0 = (b[2] - o.z) * inv.z, tz1 = (b[5] - o.z) * inv.z;
    const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), tmin));
    const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tmax));
    return (tn <= tf * 1.0000004f || tn <= tf) ? tn : kInf;
}

// Entry distances of a node's two children (+inf on a miss or for an absent
// child, index 0xffffffff).  Both slab tests run whatever the child indices
// hold, and the indices only mask the results: a test guarded by its index
// let the compiler sink the first box's load behind the wait for that index,
// two memory round trips per node instead of one.
GY_D void node_enter(uint4 q0, uint4 q1, uint4 q2, uint4 q3, V3 o, V3 inv, float tmin, float tmax, float* t0, float* t1) {
    const float b0[6] = {__uint_as_float(q0.x), __uint_as_float(q0.y), __uint_as_float(q0.z),
                         __uint_as_float(q0.w), __uint_as_float(q1.x), __uint_as_float(q1.y)};
    const float b1[6] = {__uint_as_float(q1.z), __uint_as_float(q1.w), __uint_as_float(q2.x),
                         __uint_as_float(q2.y), __uint_as_float(q2.z), __uint_as_float(q2.w)};
    const float e0 = box_enter(b0, o, inv, tmin, tmax);
    const float e1 = box_enter(b1, o, inv, tmin, tmax);
    *t0 = q3.x != 0xffffffffu ? e0 : kInf;
    *t1 = q3.y != 0xffffffffu ? e1 : kInf;
}

// IntersectP (accelerator/BVHAccel.cpp:127-149) on a triangle's own bounds.
// The reference BVH keeps one primitive per leaf (maxPrimsInNode = 1), so this
// slab test, with its strict comparisons and NaN fall-through, gates every
// triangle test the reference makes; the traversal below applies it to each
// candidate triangle so that hit decisions at t == tMax (NEE shadow segments
// end on the light's own surface) come out as the reference's do.
GY_D bool ref_box(V3 mn, V3 mx, V3 o, V3 inv, float rtmin, float rtmax) {
    const bool nx = inv.x < 0.f, ny = inv.y < 0.f, nz = inv.z < 0.f;
    float tmin = ((nx ? mx.x : mn.x) - o.x) * inv.x;
    float tmax = ((nx ? mn.x : mx.x) - o.x) * inv.x;
    const float tymin = ((ny ? mx.y : mn.y) - o.y) * inv.y;
    const float tymax = ((ny ? mn.y : mx.y) - o.y) * inv.y;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    const float tzmin = ((nz ? mx.z : mn.z) - o.z) * inv.z;
    const float tzmax = ((nz ? mn.z : mx.z) - o.z) * inv.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return (tmin < rtmax) && (tmax > rtmin);
}

// One candidate triangle (p0, p1, p2) as the reference tests it: leaf bounds
// gate and Triangle::Intersect with e1 = p1 - p0, e2 = p2 - p0
// (shape/Triangle.h:107-111).  Both tests are pure and the hit needs both, so
// the triangle test runs first and the bounds gate only for its hits, which
// are few.  The bounds gate takes the ray's own tMax
// (box_tmax) and the triangle test the current closest distance (tmax).  In
// the reference both take the shrinking tMax, which makes the winner of an
// exact-t tie depend on visit order (a flat leaf box entered exactly at the
// hit fails `tmin < tMax` once tMax has shrunk to that t).  With the gate at
// the ray's own tMax the closest hit is the minimum of (t, -primId) over every
// triangle that passes both tests at the ray's own tMax, whatever order the
// traversal visits them in; for any-hit queries the two bounds are the same.
GY_D bool ref_triangle(V3 p0, V3 p1, V3 p2, V3 o, V3 d, V3 inv, float tmin, float box_tmax, float tmax, float* t,
                       float* b1, float* b2) {
    float th, u, v;
    if (!tri_test(p0, p1 - p0, p2 - p0, o, d, tmin, tmax, &th, &u, &v)) return false;
    const V3 mn = v3(rmin(rmin(p0.x, p1.x), p2.x), rmin(rmin(p0.y, p1.y), p2.y), rmin(rmin(p0.z, p1.z), p2.z));
    const V3 mx = v3(rmax(rmax(p0.x, p1.x), p2.x), rmax(rmax(p0.y, p1.y), p2.y), rmax(rmax(p0.z, p1.z), p2.z));
    if (!ref_box(mn, mx, o, inv, tmin, box_tmax)) return false;
    *t = th;
    *b1 = u;
    *b2 = v;
    return true;
}

// CoordinateSystem (math/GeometryUtils.h:11-17)
GY_D void coordinate_system(V3 v1, V3* v2, V3* v3o) {
    if (fabsf(v1.x) > fabsf(v1.y)) *v2 = v3(-v1.z, 0.f, v1.x) / sqrtf(v1.x * v1.x + v1.z * v1.z);
    else *v2 = v3(0.f, v1.z, -v1.y) / sqrtf(v1.y * v1.y + v1.z * v1.z);
    *v3o = cross(v1, *v2);
}

// ------------------------------------------------------------------ analytic spheres
// Sphere (shape/Sphere.h:10-351) with the identity object transform, which is
// what every reference scene gives its spheres.  The identity applied to a
// point or vector is exact except that -0 becomes +0.
GY_D V3 xf_identity(V3 p) { return v3(0.f + p.x, 0.f + p.y, 0.f + p.z); }
// WorldToObject(ray) (math/Transform.h:226-239: origin error bound, move dt
// along d, tMax - dt, tMin + dt, the Ray constructor's RayEps shift), then
// Sphere::Intersect (shape/Sphere.h:29-128): the hit distance and (u, v).  The
// reference builds the hit record and drops it; only uv survives.  Out of
// line: the per-path kernels keep their registers.
__device__ __noinline__ bool sph_hit(float4 s0, V3 o, V3 d, float tmin, float tmax, float* t, float* u, float* v) {
    const V3 c = v3(s0.x, s0.y, s0.z);
    const float rad = s0.w;
    const V3 oErr = gamma_n(3) * v3(fabsf(o.x), fabsf(o.y), fabsf(o.z));
    V3 ro = xf_identity(o);
    const V3 rd = xf_identity(d);
    const float length2 = dot(rd, rd);
    float tMax = tmax, tMin = tmin;
    if (length2 > 0.f) {
        const float dt = dot(vabs(rd), oErr) / length2;
        ro = ro + rd * dt;
        tMax = tMax - dt;
        tMin = tMin + dt;
    }
    ro = ro + rd * kRayEps;
    const V3 po = ro - c;
    const float vd = dot(po, rd), d2 = dot(rd, rd), po2 = dot(po, po);
    float det = vd * vd - d2 * po2 + d2 * rad * rad;
    if (det < 0.f) return false;
    det = sqrtf(det);
    float tHit = (-vd - det) / d2;
    if (tHit < tMin || tHit > tMax) {
        tHit = (-vd + det) / d2;
        if (tHit < tMin || tHit > tMax) return false;
    }
    V3 hit = ro + rd * tHit;
    const V3 scaledDir = (hit - c) * rad / length(hit - c);
    hit = c + scaledDir;
    const V3 pHit = hit - c;
    float phi = datan2(pHit.z, pHit.x);
    const float PI = 3.14159265358979f;
    if (phi < 0.f) phi += 2 * PI;
    *u = phi / (2 * PI);
    const float theta = dacos(clampf(pHit.y / rad, -1.f, 1.f));
    *v = 1.f - (theta - 0.f) / (PI - 0.f);
    *t = tHit;
    return true;
}
// The sphere's leaf box in the reference's tree: ObjectToWorld(AABB(c - r, c + r)).
GY_D void sph_box(float4 s0, V3* mn, V3* mx) {
    const V3 c = v3(s0.x, s0.y, s0.z), r = v3(s0.w, s0.w, s0.w);
    *mn = xf_identity(c - r);
    *mx = xf_identity(c + r);
}
// Spheres after the triangles of a closest-hit query: gate at the ray's own
// tMax, test at the closest distance so far, ties to the larger primitive id.
__device__ __noinline__ void sph_closest(const SceneDev& S, V3 o, V3 d, V3 inv, float tmin, float tmax0, float* best,
                                         int32_t* best_prim, float* bb1, float* bb2) {
    for (int i = 0; i < S.n_sph; ++i) {
        const float4 s0 = S.sph[2 * i];
        V3 mn, mx;
        sph_box(s0, &mn, &mx);
        if (!ref_box(mn, mx, o, inv, tmin, tmax0)) continue;
        float th, u, v;
        const int32_t prim = S.n_tri + i;
        if (sph_hit(s0, o, d, tmin, *best, &th, &u, &v) && (th < *best || prim > *best_prim)) {
            *best = th;
            *best_prim = prim;
            *bb1 = u;
            *bb2 = v;
        }
    }
}
// Any hit against the spheres: Sphere::Intersect(const Ray&) (shape/Sphere.h:
// 131-161) drops WorldToObject(ray) and tests a default Ray, whose tHit =
// 0 / 0 passes both range tests, so a sphere occludes every ray that reaches
// its leaf, i.e. whose segment passes its box.
__device__ __noinline__ bool sph_any(const SceneDev& S, V3 o, V3 inv, float tmin, float tmax) {
    for (int i = 0; i < S.n_sph; ++i) {
        V3 mn, mx;
        sph_box(S.sph[2 * i], &mn, &mx);
        if (ref_box(mn, mx, o, inv, tmin, tmax)) return true;
    }
    return false;
}

// Closest-hit / any-hit BVH traversal with a per-thread stack column in LDS
// (stack[i * stride]).  Node boxes are padded and tested conservatively, so
// every triangle the reference's traversal would test is tested here; each is
// gated by ref_triangle.  Closest hit: minimum t, ties to the larger primitive
// id, independent of visit order.  The stack never exceeds the BVH depth, which
// the host checks against STACK before launch.
template <bool ANY, int STACK, bool SPH = false>
GY_D bool traverse(const SceneDev& S, V3 o, V3 d, float tmin, float tmax, Hit* hit, uint32_t* stack, int stride,
                   uint32_t* n_node = nullptr, uint32_t* n_tri = nullptr) {
    const V3 inv = v3(1.f / d.x, 1.f / d.y, 1.f / d.z);
    float best = tmax;
    int32_t best_prim = -1;
    float bb1 = 0.f, bb2 = 0.f;
    int sp = 0;
    uint32_t node = 0;
    while (true) {
        const uint4* np = S.nodes + (size_t)node * 4;
        const uint4 q0 = np[0], q1 = np[1], q2 = np[2], q3 = np[3];
        if (n_node) ++*n_node;
        float t0, t1;
        node_enter(q0, q1, q2, q3, o, inv, tmin, best, &t0, &t1);
        uint32_t next[2];
        float tn[2];
        int nn = 0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const float tc = c == 0 ? t0 : t1;
            if (tc == kInf) continue;
            const uint32_t ref = c == 0 ? q3.x : q3.y;
            const uint32_t cnt = c == 0 ? q3.z : q3.w;
            if (cnt == 0) {
                next[nn] = ref;
                tn[nn] = tc;
                ++nn;
                continue;
            }
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4* tp = S.tri + (size_t)(ref + k) * 3;
                const float4 a = tp[0], pb = tp[1], pc = tp[2];
                if (n_tri) ++*n_tri;
                float th, u, v;
                if (ref_triangle(ld3(a), ld3(pb), ld3(pc), o, d, inv, tmin, tmax, best, &th, &u, &v)) {
                    const int32_t prim = __float_as_int(a.w);
                    if (ANY) {
                        hit->prim = prim;
                        return true;
                    }
                    if (th < best || prim > best_prim) {
                        best = th;
                        best_prim = prim;
                        bb1 = u;
                        bb2 = v;
                    }
                }
            }
        }
        if (nn == 2) {
            if (tn[1] < tn[0]) {
                const uint32_t tr = next[0];
                next[0] = next[1];
                next[1] = tr;
                const float tt = tn[0];
                tn[0] = tn[1];
                tn[1] = tt;
            }
            if (tn[1] <= best) stack[sp++ * stride] = next[1];
            if (tn[0] <= best) {
                node = next[0];
                continue;
            }
        } else if (nn == 1) {
            if (tn[0] <= best) {
                node = next[0];
                continue;
            }
        }
        if (sp == 0) break;
        node = stack[--sp * stride];
    }
    if (SPH && S.n_sph > 0) {
        if (ANY) return sph_any(S, o, inv, tmin, tmax);
        sph_closest(S, o, d, inv, tmin, tmax, &best, &best_prim, &bb1, &bb2);
    }
    if (ANY) return false;
    if (best_prim < 0) return false;
    hit->t = best;
    hit->prim = best_prim;
    hit->b1 = bb1;
    hit->b2 = bb2;
    return true;
}

// Resumable traversal state for the persistent trace kernels (trav_run), so
// that a lane can be refilled with a new ray when its ray finishes.  Same node
// tests, triangle gating and tie rule as traverse().
struct Trav {
    V3 o, d, inv;
    float tmin, tmax0, best, bb1, bb2;  // tmax0: the ray's own tMax
    int32_t best_prim;
    uint32_t node, head, nlds, nspill;
    uint32_t leaf;  // trav_run: postponed leaf entry, or kTravNone
};
GY_D void trav_init(Trav& t, V3 o, V3 d, float tmin, float tmax) {
    t.o = o;
    t.d = d;
    t.inv = v3(1.f / d.x, 1.f / d.y, 1.f / d.z);
    t.tmin = tmin;
    t.tmax0 = tmax;
    t.best = tmax;
    t.bb1 = t.bb2 = 0.f;
    t.best_prim = -1;
    t.node = 0;
    t.head = t.nlds = t.nspill = 0;
    t.leaf = 0xffffffffu;
}
// Traversal entries of trav_run: an interior node index (bit 31 clear), a
// leaf (bit 31 set, triangle count 1..15 in bits 27-30, first triangle slot in
// bits 0-26), or kTravNone.  Scenes stay below kTravMaxTriangles so that no
// leaf entry equals kTravNone.
constexpr uint32_t kTravNone = 0xffffffffu;
constexpr uint32_t kTravLeaf = 0x80000000u;
constexpr int64_t kTravMaxTriangles = (1ll << 27) - 16;
GY_D bool trav_inner(uint32_t e) { return (e & kTravLeaf) == 0u; }
GY_D bool trav_is_leaf(uint32_t e) { return e != kTravNone && (e & kTravLeaf) != 0u; }

template <int S>
GY_D void trav_push(Trav& t, uint32_t e, uint32_t* lds, int stride, uint32_t* __restrict__ spill, int64_t spill_stride) {
    if (t.nlds == (uint32_t)S) {
        spill[(int64_t)t.nspill * spill_stride] = lds[t.head * stride];
        ++t.nspill;
        t.head = t.head + 1 == (uint32_t)S ? 0 : t.head + 1;
        --t.nlds;
    }
    uint32_t pos = t.head + t.nlds;
    if (pos >= (uint32_t)S) pos -= S;
    lds[pos * stride] = e;
    ++t.nlds;
}
template <int S>
GY_D uint32_t trav_pop(Trav& t, uint32_t* lds, int stride, uint32_t* __restrict__ spill, int64_t spill_stride) {
    if (t.nlds > 0) {
        --t.nlds;
        uint32_t pos = t.head + t.nlds;
        if (pos >= (uint32_t)S) pos -= S;
        return lds[pos * stride];
    }
    if (t.nspill > 0) {
        --t.nspill;
        return spill[(int64_t)t.nspill * spill_stride];
    }
    return kTravNone;
}

// Traversal in two phases per round (Aila and Laine, "Understanding the
// efficiency of ray traversal on GPUs", 2009: while-while with postponed
// leaves).  Interior nodes are visited until at most 1/kSpecWait of the
// round's lanes still search for a leaf (a lane's first leaf is postponed and
// it keeps walking); then the lanes test their postponed leaves together, and
// a second leaf met on the way, while the few still searching wait and resume
// in the next round.  Waiting for every lane idles the lanes that hold a leaf
// behind the last few deep searches; the closest hit does not depend on the
// visit order, so the ratio moves time only.  The node tests, culling and triangle gate are
// traverse()'s (a leaf child is kept when its padded box is entered, an
// interior child when its entry distance is also <= best), and triangles are
// tested in the order their leaves are reached.  Returns 1 when the ray is
// finished (t.best_prim >= 0 on a hit), 2 when an any-hit query found a hit,
// and 0 when fewer than `refill` lanes remain, so that the caller can refill.
constexpr int kSpecWait = 8;
template <bool ANY, int S>
GY_D int trav_run(const SceneDev& Sc, Trav& t, uint32_t* lds, int stride, uint32_t* __restrict__ spill, int64_t spill_stride,
                  int refill) {
    while (true) {
        const int nact = __popcll(__ballot(true));
        while (trav_inner(t.node)) {
            const uint4* np = Sc.nodes + (size_t)t.node * 4;
            const uint4 q0 = np[0], q1 = np[1], q2 = np[2], q3 = np[3];
            float t0, t1;
            node_enter(q0, q1, q2, q3, t.o, t.inv, t.tmin, t.best, &t0, &t1);
            const bool keep0 = t0 != kInf && (q3.z != 0u || t0 <= t.best);
            const bool keep1 = t1 != kInf && (q3.w != 0u || t1 <= t.best);
            const uint32_t e0 = q3.z != 0u ? (kTravLeaf | (q3.z << 27) | q3.x) : q3.x;
            const uint32_t e1 = q3.w != 0u ? (kTravLeaf | (q3.w << 27) | q3.y) : q3.y;
            if (keep0 && keep1) {
                const bool swap = t1 < t0;
                trav_push<S>(t, swap ? e0 : e1, lds, stride, spill, spill_stride);
                t.node = swap ? e1 : e0;
            } else if (keep0) {
                t.node = e0;
            } else if (keep1) {
                t.node = e1;
            } else {
                t.node = trav_pop<S>(t, lds, stride, spill, spill_stride);
            }
            if (trav_is_leaf(t.node) && t.leaf == kTravNone) {
                t.leaf = t.node;
                t.node = trav_pop<S>(t, lds, stride, spill, spill_stride);
            }
            if (__popcll(__ballot(t.leaf == kTravNone)) * kSpecWait <= nact) break;
        }
        while (t.leaf != kTravNone) {
            const uint32_t ref = t.leaf & 0x07ffffffu, cnt = (t.leaf >> 27) & 15u;
            for (uint32_t k = 0; k < cnt; ++k) {
                const float4* tp = Sc.tri + (size_t)(ref + k) * 3;
                const float4 a = tp[0], pb = tp[1], pc = tp[2];
                float th, u, v;
                if (ref_triangle(ld3(a), ld3(pb), ld3(pc), t.o, t.d, t.inv, t.tmin, t.tmax0, t.best, &th, &u, &v)) {
                    const int32_t prim = __float_as_int(a.w);
                    if (ANY) {
                        t.best_prim = prim;
                        return 2;
                    }
                    if (th < t.best || prim > t.best_prim) {
                        t.best = th;
                        t.best_prim = prim;
                        t.bb1 = u;
                        t.bb2 = v;
                    }
                }
            }
            t.leaf = kTravNone;
            if (trav_is_leaf(t.node)) {
                t.leaf = t.node;
                t.node = trav_pop<S>(t, lds, stride, spill, spill_stride);
            }
        }
        if (t.node == kTravNone) return 1;
        if (__popcll(__ballot(true)) < refill) return 0;
    }
}

// One ray through trav_run to its end: the per-path kernels (megakernel,
// per-sample, light tracing, adjoint replay) use the wavefront's traversal.
// The lanes of a wave that call it together walk the tree with postponed
// leaves; with refill 0 it returns only when the ray is finished.  The stack
// never holds more entries than the BVH is deep, which the host keeps below
// STACK, so the ring never spills and no spill buffer is passed.  COUNT: the
// counting pass keeps the depth-first traverse(), whose visits are the
// algorithmic bytes' lower bound.
template <bool ANY, int STACK, bool COUNT = false, bool SPH = false>
GY_D bool find_hit(const SceneDev& S, V3 o, V3 d, float tmin, float tmax, Hit* hit, uint32_t* stack, int stride,
                   uint32_t* n_node = nullptr, uint32_t* n_tri = nullptr) {
    if (COUNT) return traverse<ANY, STACK, SPH>(S, o, d, tmin, tmax, hit, stack, stride, n_node, n_tri);
    Trav t;
    trav_init(t, o, d, tmin, tmax);
    const int r = trav_run<ANY, STACK>(S, t, stack, stride, nullptr, 0, 0);
    if (SPH && S.n_sph > 0) {
        if (ANY) return r == 2 || sph_any(S, t.o, t.inv, tmin, tmax);
        sph_closest(S, t.o, t.d, t.inv, tmin, tmax, &t.best, &t.best_prim, &t.bb1, &t.bb2);
    }
    if (ANY) return r == 2;
    if (t.best_prim < 0) return false;
    hit->t = t.best;
    hit->prim = t.best_prim;
    hit->b1 = t.bb1;
    hit->b2 = t.bb2;
    return true;
}

// ------------------------------------------------------------------ surface interaction
struct Isect {
    V3 pos, n, absN, ng, outDir, sdpdu, sdpdv, pErr;
    float u, v;
    int32_t prim;
    uint32_t mat;
    int32_t light;
};
// Triangle::QueryIntersectionInfo (shape/Triangle.cpp:6-69) + mPos
// (Triangle.cpp:97).  dpdu/dpdv are not used by the integrator and not formed.
template <bool SPH = false>
GY_D Isect query(const SceneDev& S, V3 ro, V3 rd, const Hit& h) {
    if (SPH && h.prim >= S.n_tri) {
        // a sphere hit: the default Intersection (zero position, normals,
        // frame, outgoing direction and error bound) with the hit's uv; the
        // reference drops the record Sphere::Intersect builds and
        // Scene::QueryIntersectionInfo adds nothing for spheres
        const uint4 w = __builtin_bit_cast(uint4, S.sph[2 * (h.prim - S.n_tri) + 1]);
        Isect it;
        it.pos = it.n = it.absN = it.ng = it.outDir = it.sdpdu = it.sdpdv = it.pErr = v3(0.f, 0.f, 0.f);
        it.u = h.b1;
        it.v = h.b2;
        it.prim = h.prim;
        it.mat = w.y;
        it.light = (int32_t)w.z;
        return it;
    }
    const float4* q = S.shade + (size_t)h.prim * 8;
    const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5], q6 = q[6], q7 = q[7];
    const V3 p0 = ld3(q0), p1 = ld3(q1), p2 = ld3(q2);
    const V3 n0 = ld3(q3), n1 = ld3(q4), n2 = ld3(q5), fn = ld3(q6);
    Isect it;
    const float b1 = h.b1, b2 = h.b2;
    it.pos = ((1 - b1 - b2) * p0 + b1 * p1) + b2 * p2;
    const float b0 = 1 - b1 - b2;
    it.u = (b0 * q3.w + b1 * q5.w) + b2 * q7.x;
    it.v = (b0 * q4.w + b1 * q6.w) + b2 * q7.y;
    const float xs = (fabsf(b0 * p0.x) + fabsf(b1 * p1.x)) + fabsf(b2 * p2.x);
    const float ys = (fabsf(b0 * p0.y) + fabsf(b1 * p1.y)) + fabsf(b2 * p2.y);
    const float zs = (fabsf(b0 * p0.z) + fabsf(b1 * p1.z)) + fabsf(b2 * p2.z);
    V3 ns = (b0 * n0 + b1 * n1) + b2 * n2;
    ns = norm(ns);
    it.ng = (dot(fn, ns) < 0.f) ? neg(fn) : fn;
    it.absN = dot(it.ng, rd) < 0.f ? it.ng : it.ng * -1.f;
    it.outDir = neg(rd);
    it.pErr = gamma_n(7) * v3(xs, ys, zs);
    const float sg = (dot(ro - p0, ns) > 0.f) ? 1.0f : -1.0f;
    V3 dpdus = norm(p1 - p0);
    const V3 dpdvs = norm(cross(sg * ns, dpdus));
    dpdus = norm(cross(dpdvs, ns));
    it.n = ns;
    it.sdpdu = dpdus;
    it.sdpdv = dpdvs;
    it.prim = h.prim;
    it.mat = __float_as_uint(q0.w);
    it.light = __float_as_int(q1.w);
    return it;
}

// numeric/NumericUtils.h:22-47
GY_D float next_up(float v) {
    if (isinf(v) && v > 0.f) return v;
    if (v == -0.f) v = 0.f;
    uint32_t ui = __float_as_uint(v);
    if (v >= 0.f) ++ui; else --ui;
    return __uint_as_float(ui);
}
GY_D float next_down(float v) {
    if (isinf(v) && v < 0.f) return v;
    if (v == 0.f) v = -0.f;
    uint32_t ui = __float_as_uint(v);
    if (v > 0.f) --ui; else ++ui;
    return __uint_as_float(ui);
}
// OffsetRayOrigin (math/GeometryUtils.h:39-64)
GY_D V3 offset_origin(V3 p, V3 pErr, V3 n, V3 w) {
    const float d = dot(vabs(n), pErr);
    V3 off = d * n;
    if (dot(w, n) < 0.f) off = neg(off);
    V3 po = p + off;
    po.x = off.x > 0.f ? next_up(po.x) : (off.x < 0.f ? next_down(po.x) : po.x);
    po.y = off.y > 0.f ? next_up(po.y) : (off.y < 0.f ? next_down(po.y) : po.y);
    po.z = off.z > 0.f ? next_up(po.z) : (off.z < 0.f ? next_down(po.z) : po.z);
    return po;
}
// Intersection::SpawnRay (visual/Intersection.h:74-78): origin + RayEps * d
GY_D void spawn_ray(const Isect& it, V3 d, V3* o) { *o = offset_origin(it.pos, it.pErr, it.absN, d) + d * kRayEps; }

// ------------------------------------------------------------------ warps
// visual/Sampling.cpp:5-36
GY_D V3 cosine_hemisphere(float u0, float u1) {
    const float ox = 2.0f * u0 - 1.f, oy = 2.0f * u1 - 1.f;
    float dx = 0.f, dy = 0.f;
    if (!(ox == 0.f && oy == 0.f)) {
        float theta, r;
        if (fabsf(ox) > fabsf(oy)) {
            r = ox;
            theta = kPIOver4 * (oy / ox);
        } else {
            r = oy;
            theta = kPIOver2 - kPIOver4 * (ox / oy);
        }
        dx = r * dcos(theta);
        dy = r * dsin(theta);
    }
    const float z = sqrtf(rmax(0.f, 1 - dx * dx - dy * dy));
    return v3(dx, dy, z);
}

// ------------------------------------------------------------------ textures
// ImageTexture::Triangle / Texel (texture/ImageTexture.h:50-69), CLAMP.
GY_D float texel(const SceneDev& S, int4 ti, int u, int v, int ch) {
    u = u < 0 ? 0 : (u > ti.x - 1 ? ti.x - 1 : u);
    v = v < 0 ? 0 : (v > ti.y - 1 ? ti.y - 1 : v);
    return S.tex[(size_t)ti.w + ((size_t)v * ti.x + u) * ti.z + ch];
}
GY_D float tex_bilinear(const SceneDev& S, int id, float u, float v, int ch) {
    const int4 ti = S.tex_info[id];
    if (ch >= ti.z) ch = ti.z - 1;
    const float s = ti.x * u - 0.5f, t = ti.y * v - 0.5f;
    const int s0 = (int)floorf(s), t0 = (int)floorf(t);
    const float ds = s - s0, dt = t - t0;
    const float a = (1 - ds) * texel(S, ti, s0, t0, ch) + ds * texel(S, ti, s0 + 1, t0, ch);
    const float b = (1 - ds) * texel(S, ti, s0, t0 + 1, ch) + ds * texel(S, ti, s0 + 1, t0 + 1, ch);
    return (1 - dt) * a + dt * b;
}

// ------------------------------------------------------------------ GGX
// GGXDistribution (visual/Microfacet.cpp:232-379) with visible-normal sampling.
struct GGX {
    float ax, ay;
};
GY_D float ggx_D(const GGX& g, V3 m) {
    if (m.z <= 0.f) return 0.0f;
    const float cos2 = m.z * m.z;
    const float be = ((m.x * m.x) / (g.ax * g.ax) + (m.y * m.y) / (g.ay * g.ay)) / cos2;
    const float root = (1.f + be) * cos2;
    float result = 1.f / (kPI * g.ax * g.ay * root * root);
    if (result * m.z < 1e-20f) result = 0.f;
    return result;
}
GY_D float ggx_G1(const GGX& g, V3 v, V3 m) {
    if (dot(v, m) * v.z <= 0.f) return 0.0f;
    const float sinT = sqrtf(rmax(0.f, 1.f - v.z * v.z));
    const float tanT = fabsf(sinT / v.z);
    if (tanT == 0.0f) return 1.0f;
    float alpha = g.ax;
    const float invSin2 = 1.f / rmax(0.f, 1.f - v.z * v.z);
    if (!(g.ax == g.ay || invSin2 <= 0.f)) {
        const float cp2 = v.x * v.x * invSin2, sp2 = v.y * v.y * invSin2;
        alpha = sqrtf(cp2 * g.ax * g.ax + sp2 * g.ay * g.ay);
    }
    const float root = alpha * tanT;
    return 2.0f / (1.0f + sqrtf(1 + root * root));
}
GY_D float ggx_G(const GGX& g, V3 wo, V3 wi, V3 wh) { return ggx_G1(g, wo, wh) * ggx_G1(g, wi, wh); }
GY_D float ggx_pdf(const GGX& g, V3 wo, V3 wh) {
    if (wo.z == 0.f) return 0.0f;
    return ggx_G1(g, wo, wh) * fabsf(dot(wo, wh)) * ggx_D(g, wh) / fabsf(wo.z);
}
// SampleGGXVNDF (Heitz 2018; visual/Microfacet.cpp:318-339)
GY_D V3 ggx_sample_wh(const GGX& g, V3 v_, float u0, float u1) {
    const V3 v = norm(v3(g.ax * v_.x, g.ay * v_.y, v_.z));
    const V3 T1 = ((double)v.z < 0.9999) ? norm(cross(v, v3(0.f, 0.f, 1.f))) : v3(1.f, 0.f, 0.f);
    const V3 T2 = cross(T1, v);
    const float a = 1.f / (1.f + v.z);
    const float r = sqrtf(u0);
    const float phi = (u1 < a) ? u1 / a * kPI : kPI + (u1 - a) / (1.f - a) * kPI;
    const float P1 = r * dcos(phi);
    const float P2 = r * dsin(phi) * ((u1 < a) ? 1.0f : v.z);
    const V3 dir = (P1 * T1 + P2 * T2) + sqrtf(rmax(0.f, 1 - P1 * P1 - P2 * P2)) * v;
    return norm(v3(g.ax * dir.x, g.ay * dir.y, rmax(0.f, dir.z)));
}
GY_D V3 reflect(V3 in, V3 n) { return 2.f * dot(in, n) * n - in; }  // math/GeometryUtils.h:20-22
GY_D bool same_hemi(V3 a, V3 b) { return a.z * b.z > 0.f; }
// FrDielectric (bsdf/BSDF.cpp:5-26)
GY_D float fr_dielectric(float cosI, float etaI, float etaT) {
    cosI = clampf(cosI, -1.f, 1.f);
    if (!(cosI > 0.f)) {
        const float t = etaI; etaI = etaT; etaT = t;
        cosI = fabsf(cosI);
    }
    const float sinI = sqrtf(rmax(0.f, 1 - cosI * cosI));
    const float sinT = etaI / etaT * sinI;
    if (sinT >= 1.f) return 1.f;
    const float cosT = sqrtf(rmax(0.f, 1 - sinT * sinT));
    const float Rparl = ((etaT * cosI) - (etaI * cosT)) / ((etaT * cosI) + (etaI * cosT));
    const float Rperp = ((etaI * cosI) - (etaT * cosT)) / ((etaI * cosI) + (etaT * cosT));
    return (Rparl * Rparl + Rperp * Rperp) / 2;
}
// FrConductor (bsdf/BSDF.cpp:29-51) with etai = 1
GY_D V3 fr_conductor(float cosI, V3 etat, V3 k) {
    cosI = clampf(cosI, -1.f, 1.f);
    const V3 eta = etat / v3s(1.f), etak = k / v3s(1.f);
    const float cos2 = cosI * cosI;
    const float sin2 = (float)(1. - (double)cos2);
    const V3 eta2 = eta * eta, etak2 = etak * etak;
    const V3 t0 = eta2 - etak2 - v3s(sin2);
    const V3 a2pb2 = vsqrt(t0 * t0 + 4.f * eta2 * etak2);
    const V3 t1 = a2pb2 + v3s(cos2);
    const V3 a = vsqrt(0.5f * (a2pb2 + t0));
    const V3 t2 = 2.f * cosI * a;
    const V3 Rs = (t1 - t2) / (t1 + t2);
    const V3 t3 = cos2 * a2pb2 + v3s(sin2 * sin2);
    const V3 t4 = t2 * sin2;
    const V3 Rp = Rs * (t3 - t4) / (t3 + t4);
    return 0.5f * (Rp + Rs);
}
// Schlick Fresnel from F0 (configuration extension, not in the reference).
GY_D V3 fr_schlick(float cosI, V3 f0) {
    const float c = 1.f - fabsf(cosI);
    const float c2 = c * c;
    const float c5 = c2 * c2 * c;
    return f0 + (v3s(1.f) - f0) * c5;
}
// TransmissionBSDF::Refract (bsdf/BSDF.h:292-304)
GY_D bool refract(V3 wi, V3 n, float eta, V3* wt) {
    const float cosI = dot(n, wi);
    const float sin2I = rmax(0.f, 1 - cosI * cosI);
    const float sin2T = eta * eta * sin2I;
    if (sin2T >= 1.f) return false;
    const float cosT = sqrtf(1 - sin2T);
    *wt = eta * (wi * -1.f) + (eta * cosI - cosT) * n;
    return true;
}
// visual/Microfacet.h:49-54
GY_D float roughness_to_alpha(float r) {
    r = rmax(r, 1e-3f);
    const float x = dlog(r);
    return 1.62142f + 0.819955f * x + 0.1734f * x * x + 0.0171201f * x * x * x + 0.000640711f * x * x * x * x;
}

// ------------------------------------------------------------------ BSDF
enum { MAT_LAMBERT = 0, MAT_CONDUCTOR = 1, MAT_GLASS = 2, MAT_ROUGH_DIELECTRIC = 3 };
struct BSDF {
    int type;
    bool schlick;
    V3 ng, X, Y, Z;
    V3 Rc, Tc;
    GGX dist;
    V3 eta, k;  // conductor eta/k, or F0 in eta when schlick
    float eta1, eta2;
    bool importance;  // TransportMode::Importance (light paths): no eta^2 scaling of transmission
};
GY_D bool bsdf_delta(const BSDF& b) { return b.type == MAT_GLASS; }
GY_D V3 to_local(const BSDF& b, V3 v) { return v3(dot(v, b.X), dot(v, b.Y), dot(v, b.Z)); }
GY_D V3 to_world(const BSDF& b, V3 v) { return (b.X * v.x + b.Y * v.y) + b.Z * v.z; }  // math/Frame.h:38-41
GY_D V3 cond_fresnel(const BSDF& b, float c) { return b.schlick ? fr_schlick(c, b.eta) : fr_conductor(fabsf(c), b.eta, b.k); }

// Material::ComputeScatteringFunction (material/Diffuse.h:27-31, Roughness.h:36-46,
// Glass.h:30-34, RoughDielectric.h:21-41) + BSDF(isect) frame (bsdf/BSDF.h:131-134)
GY_D BSDF make_bsdf(const SceneDev& S, const Isect& it) {
    const uint32_t* w = S.mats + (size_t)it.mat * 32;
    BSDF b;
    b.type = (int)w[0];
    b.importance = false;
    const uint32_t flags = w[1];
    b.schlick = (flags & 2u) != 0;
    b.ng = it.ng;
    b.X = it.sdpdu;
    b.Y = it.sdpdv;
    b.Z = it.n;
    b.Rc = v3(__uint_as_float(w[2]), __uint_as_float(w[3]), __uint_as_float(w[4]));
    b.Tc = v3(__uint_as_float(w[5]), __uint_as_float(w[6]), __uint_as_float(w[7]));
    const int32_t ta = (int32_t)w[21], tr = (int32_t)w[22];
    if (ta >= 0) b.Rc = v3(tex_bilinear(S, ta, it.u, it.v, 0), tex_bilinear(S, ta, it.u, it.v, 1), tex_bilinear(S, ta, it.u, it.v, 2));
    float ax = __uint_as_float(w[8]), ay = __uint_as_float(w[9]);
    if (tr >= 0) ax = ay = tex_bilinear(S, tr, it.u, it.v, 0);
    if (!(flags & 1u)) {
        ax = roughness_to_alpha(ax);
        ay = roughness_to_alpha(ay);
    }
    b.dist.ax = ax;
    b.dist.ay = ay;
    if (b.schlick) b.eta = v3(__uint_as_float(w[16]), __uint_as_float(w[17]), __uint_as_float(w[18]));
    else b.eta = v3(__uint_as_float(w[10]), __uint_as_float(w[11]), __uint_as_float(w[12]));
    b.k = v3(__uint_as_float(w[13]), __uint_as_float(w[14]), __uint_as_float(w[15]));
    b.eta1 = __uint_as_float(w[19]);
    b.eta2 = __uint_as_float(w[20]);
    return b;
}

GY_D float lobe_pdf(const BSDF& b, V3 wo, V3 wi) {
    if (b.type == MAT_LAMBERT) return same_hemi(wo, wi) ? fabsf(wi.z) * kInvPI : 0.f;  // bsdf/BSDF.h:170-173
    if (b.type == MAT_CONDUCTOR) {                                                      // bsdf/BSDF.h:323-333
        V3 wh = norm(wo + wi);
        const float dwh = 1.f / (4 * dot(wi, wh));
        wh = wh * sgnf(wh.z);
        return fabsf(ggx_pdf(b.dist, wo, wh) * dwh);
    }
    if (b.type == MAT_GLASS) return 0.f;
    // RoughDielectricBSDF::Pdf (bsdf/BSDF.h:513-543)
    const bool refl = wo.z * wi.z > 0.f;
    V3 wh;
    float dwh;
    if (refl) {
        wh = norm(wo + wi);
        dwh = 1.f / (4 * dot(wi, wh));
    } else {
        const float e = wo.z > 0.f ? (b.eta2 / b.eta1) : (b.eta1 / b.eta2);
        wh = neg(norm(wo + wi * e));
        const float sd = dot(wo, wh) + e * dot(wi, wh);
        dwh = fabsf((e * e * dot(wi, wh)) / (sd * sd));
    }
    wh = wh * sgnf(wh.z);
    float pdf = fabsf(ggx_pdf(b.dist, wo, wh) * dwh);
    const float F = fr_dielectric(dot(wo, wh), b.eta1, b.eta2);
    pdf *= refl ? F : (1 - F);
    return pdf;
}

GY_D V3 lobe_eval(const BSDF& b, V3 wo, V3 wi) {
    if (b.type == MAT_LAMBERT) return b.Rc * kInvPI;
    if (b.type == MAT_GLASS) return v3s(0.f);
    const bool refl = wo.z * wi.z > 0.f;
    if (b.type == MAT_CONDUCTOR || refl) {
        // MicrofacetReflectionBSDF::Evaluate (bsdf/BSDF.h:356-373) /
        // RoughDielectricBSDF::Evaluate reflection (bsdf/BSDF.h:471-488)
        if (b.type == MAT_CONDUCTOR && !same_hemi(wo, wi)) return v3s(0.f);
        const float cO = fabsf(wo.z), cI = fabsf(wi.z);
        V3 wh = wi + wo;
        if (cI == 0.f || cO == 0.f) return v3s(0.f);
        if (wh.x == 0.f && wh.y == 0.f && wh.z == 0.f) return v3s(0.f);
        wh = norm(wh);
        wh = sgnf(wh.z) * wh;
        const V3 F = b.type == MAT_CONDUCTOR ? cond_fresnel(b, dot(wo, wh)) : v3s(fr_dielectric(dot(wo, wh), b.eta1, b.eta2));
        const V3 D = v3s(ggx_D(b.dist, wh));
        const V3 G = v3s(ggx_G(b.dist, wo, wi, wh));
        return b.Rc * F * D * G / (4 * cI * cO);
    }
    // RoughDielectricBSDF::Evaluate transmission (bsdf/BSDF.h:489-510)
    const float cO = wo.z, cI = wi.z;
    if (cI == 0.f || cO == 0.f) return v3s(0.f);
    const float e = wo.z > 0.f ? (b.eta2 / b.eta1) : (b.eta1 / b.eta2);
    V3 wh = neg(norm(wo + wi * e));
    wh = sgnf(wh.z) * wh;
    const float F = fr_dielectric(dot(wo, wh), b.eta1, b.eta2);
    const float D = ggx_D(b.dist, wh);
    const float G = ggx_G(b.dist, wo, wi, wh);
    const float sd = dot(wo, wh) + e * dot(wi, wh);
    const float factor = b.importance ? 1.f : 1 / e;  // (bsdf/BSDF.h:505, 608)
    const float value = (1 - F) * D * G * e * e * fabsf(dot(wi, wh)) * fabsf(dot(wo, wh)) / fabsf(cI * cO * sd * sd);
    return b.Tc * fabsf(value * factor * factor);
}

GY_D V3 lobe_sample(const BSDF& b, V3 wo, V3* wi, float* pdf, float r0, float r1, float r2) {
    if (b.type == MAT_LAMBERT) {  // DiffuseBSDF::Sample (bsdf/BSDF.h:175-181)
        V3 w = cosine_hemisphere(r0, r1);
        if (wo.z < 0.f) w.z *= -1.f;
        *wi = w;
        *pdf = lobe_pdf(b, wo, w);
        return lobe_eval(b, wo, w);
    }
    if (b.type == MAT_CONDUCTOR) {  // MicrofacetReflectionBSDF::Sample (bsdf/BSDF.h:335-354)
        if (wo.z == 0.f) return v3s(0.f);
        V3 wh = ggx_sample_wh(b.dist, sgnf(wo.z) * wo, r0, r1);
        const float mpdf = ggx_pdf(b.dist, wo, wh);
        if (mpdf == 0.f) return v3s(0.f);
        const V3 w = reflect(wo, wh);
        *wi = w;
        if (!same_hemi(wo, w)) return v3s(0.f);
        wh = sgnf(wh.z) * wh;
        *pdf = fabsf(ggx_pdf(b.dist, wo, wh) / (4 * dot(w, wh)));
        return lobe_eval(b, wo, w);
    }
    if (b.type == MAT_GLASS) {  // TransmissionBSDF::Sample (bsdf/BSDF.h:238-270)
        const float Re = fr_dielectric(wo.z, b.eta1, b.eta2);
        const bool entering = wo.z > 0.f;
        const float etaI = entering ? b.eta1 : b.eta2, etaT = entering ? b.eta2 : b.eta1;
        const V3 nF = wo.z > 0.f ? v3(0.f, 0.f, 1.f) : v3(0.f, 0.f, -1.f);
        const float P = Re;
        if (r1 < P) {
            *wi = v3(-wo.x, -wo.y, wo.z);
            *pdf = P;
            return b.Rc * Re / fabsf(wi->z);
        }
        V3 w;
        if (!refract(wo, nF, etaI / etaT, &w)) return v3s(0.f);
        *wi = w;
        *pdf = 1 - P;
        const float e = etaI / etaT;
        if (b.importance) return b.Tc * (1.f - Re) / fabsf(w.z);
        return e * e * b.Tc * (1.f - Re) / fabsf(w.z);
    }
    // RoughDielectricBSDF::Sample (bsdf/BSDF.h:545-615)
    const V3 wh = ggx_sample_wh(b.dist, sgnf(wo.z) * wo, r0, r1);
    const float mpdf = ggx_pdf(b.dist, wo, wh);
    if (mpdf == 0.f) return v3s(0.f);
    const float F = fr_dielectric(dot(wo, wh), b.eta1, b.eta2);
    const float P = F;
    *pdf = 0.f;
    if (r2 < P) {
        if (wo.z == 0.f) return v3s(0.f);
        const V3 w = reflect(wo, wh);
        *wi = w;
        if (!same_hemi(wo, w)) return v3s(0.f);
        const float dwh = 1.f / (4 * dot(w, wh));
        *pdf = P * fabsf(ggx_pdf(b.dist, wo, wh) * dwh);
        const float cO = fabsf(wo.z), cI = fabsf(w.z);
        const V3 D = v3s(ggx_D(b.dist, wh));
        const V3 G = v3s(ggx_G(b.dist, wo, w, wh));
        return b.Rc * v3s(F) * D * G / (4 * cI * cO);
    }
    if (wo.z == 0.f) return v3s(0.f);
    const float etaR = wo.z > 0.f ? (b.eta1 / b.eta2) : (b.eta2 / b.eta1);
    const V3 ff = (dot(wh, wo) < 0.f) ? neg(wh) : wh;
    V3 w;
    if (!refract(wo, ff, etaR, &w)) {
        *pdf = 0.f;
        return v3s(0.f);
    }
    *wi = w;
    if (wo.z * w.z >= 0.f) return v3s(0.f);
    const float e = wo.z > 0.f ? (b.eta2 / b.eta1) : (b.eta1 / b.eta2);
    const float sd = dot(wo, wh) + e * dot(w, wh);
    const float dwh = fabsf((e * e * dot(w, wh)) / (sd * sd));
    *pdf = fabsf((1 - P) * ggx_pdf(b.dist, wo, wh) * dwh);
    const float cO = wo.z, cI = w.z;
    if (cI == 0.f || cO == 0.f) return v3s(0.f);
    const float D = ggx_D(b.dist, wh);
    const float G = ggx_G(b.dist, wo, w, wh);
    const float factor = b.importance ? 1.f : 1 / e;  // (bsdf/BSDF.h:505, 608)
    const float value = (1 - F) * D * G * e * e * fabsf(dot(w, wh)) * fabsf(dot(wo, wh)) / fabsf(cI * cO * sd * sd);
    return b.Tc * fabsf(value * factor * factor);
}

// BSDF::Pdf / Sample / Evaluate with one lobe (bsdf/BSDF.cpp:58-119)
GY_D float bsdf_pdf(const BSDF& b, V3 wo, V3 wi) {
    float pdf = 0.f;
    pdf += lobe_pdf(b, to_local(b, wo), to_local(b, wi));
    return pdf / 1;
}
GY_D V3 bsdf_sample(const BSDF& b, V3 wo, V3* wi, float* pdf, float r0, float r1, float r2) {
    const V3 woL = to_local(b, wo);
    V3 wiL = v3s(0.f);
    *pdf = 0.f;
    const V3 f = lobe_sample(b, woL, &wiL, pdf, r0, r1, r2);
    *wi = to_world(b, wiL);
    *pdf /= 1;
    return f;
}
GY_D V3 bsdf_eval(const BSDF& b, V3 wo, V3 wi) {
    V3 f = v3(0.f, 0.f, 0.f);
    const bool refl = dot(wi, b.ng) * dot(wo, b.ng) > 0.f;
    const bool has_trans = b.type == MAT_GLASS || b.type == MAT_ROUGH_DIELECTRIC;
    if (refl || has_trans) f = f + lobe_eval(b, to_local(b, wo), to_local(b, wi));
    return f;
}

// ------------------------------------------------------------------ lights
// Scene::SampleOneLight -> Distribution1D::SampleDiscrete (visual/Scene.h:97-109,
// visual/Sampling.h:86-97): upper_bound over the CDF.
GY_D int sample_one_light(const SceneDev& S, float u, float* pdf) {
    const int n = S.n_light;
    int lo = 0, hi = n + 1;  // first index with cdf > u in [0, n+1)
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (S.light_cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    int off = lo - 1;
    if (off < 0) off = 0;
    *pdf = S.light_func[off] / (S.light_func_int * n);
    return off;
}
struct LightPoint {
    V3 pos, n, fn, pErr;
};
// The area light of primitive `prim` (-1: none).
template <bool SPH = false>
GY_D int32_t prim_light(const SceneDev& S, int32_t prim) {
    if (SPH && prim >= S.n_tri) return __float_as_int(S.sph[2 * (prim - S.n_tri) + 1].z);
    return __float_as_int(S.shade[(size_t)prim * 8 + 1].w);
}
// Sphere::Sample(real*, const Vec2&) (shape/Sphere.h:163-193): a uniform point
// with its ObjectToWorld error bound; the absolute normal is the new record's
// own zero normal.
__device__ __noinline__ LightPoint sph_sample_area(float4 s0, uint32_t inside, float u0, float u1, float* pdf) {
    const V3 c = v3(s0.x, s0.y, s0.z);
    const float rad = s0.w, PI = 3.14159265358979f;
    *pdf = 1.f / (4.f * PI * rad * rad);
    const float z = 1 - 2 * u0;  // UniformSampleSphere (visual/Sampling.cpp:25-30)
    const float rr = sqrtf(rmax(0.f, 1.f - z * z));
    const float phi = 2 * PI * u1;
    const V3 pHit = v3(rr * dcos(phi), rr * dsin(phi), z) * rad + c;
    const V3 scaledDir = (pHit - c) * rad / length(pHit - c);
    const V3 hit = c + scaledDir;
    const V3 n = inside ? norm(c - hit) : norm(hit - c);
    const V3 pError = vabs(c) * gamma_n(1) + vabs(scaledDir) * gamma_n(6);
    const float g3 = gamma_n(3);
    LightPoint lp;
    lp.pErr = v3((g3 + 1.f) * ((pError.x + 0.f * pError.y) + 0.f * pError.z) + g3 * (((fabsf(hit.x) + 0.f) + 0.f) + 0.f),
                 (g3 + 1.f) * ((0.f * pError.x + pError.y) + 0.f * pError.z) + g3 * (((0.f + fabsf(hit.y)) + 0.f) + 0.f),
                 (g3 + 1.f) * ((0.f * pError.x + 0.f * pError.y) + pError.z) + g3 * (((0.f + 0.f) + fabsf(hit.z)) + 0.f));
    lp.pos = xf_identity(hit);
    lp.n = norm(xf_identity(n));
    lp.fn = v3(0.f, 0.f, 0.f);
    return lp;
}
// Sphere::Sample(const Intersection&, real*, const Vec2&) (shape/Sphere.h:243-291):
// inside the sphere an area sample whose pdf is multiplied by the distance
// (not its square) over the cosine; outside, a direction in the cone the
// sphere subtends, the point 1.5e-3 r beyond the surface (r * 1.5e-3 formed in
// double, a double literal), pdf 1 / cone solid angle.
__device__ __noinline__ LightPoint sph_sample_li(const SceneDev& S, int i, V3 ref, float u0, float u1, float* pdf) {
    const float4 s0 = S.sph[2 * i];
    const uint4 w = __builtin_bit_cast(uint4, S.sph[2 * i + 1]);
    const V3 pCenter = xf_identity(v3(s0.x, s0.y, s0.z));
    const float rad = s0.w, PI = 3.14159265358979f;
    if (dot(pCenter - ref, pCenter - ref) <= rad * rad) {
        LightPoint lp = sph_sample_area(s0, w.x, u0, u1, pdf);
        V3 wi = lp.pos - ref;
        if (dot(wi, wi) == 0.f) {
            *pdf = 0.f;
        } else {
            const float sd = length(wi);
            wi = norm(wi);
            *pdf = *pdf * (sd / fabsf(dot(norm(lp.pos - pCenter), wi * -1.f)));
        }
        if (isinf(*pdf)) *pdf = 0.f;
        return lp;
    }
    V3 localZ = pCenter - ref;
    const float dis = length(localZ);
    localZ = norm(localZ);
    V3 localX, localY;
    coordinate_system(localZ, &localX, &localY);
    const float sin2ThetaMax = rad * rad / (dis * dis);
    const float cosThetaMax = sqrtf(rmax(1.f - sin2ThetaMax, 0.f));
    const float cosTheta = (1.f - u0) + u0 * cosThetaMax;
    const float sinTheta = sqrtf(rmax(1.f - cosTheta * cosTheta, 0.f));
    const float phi = 2 * PI * u1;
    const float sdist = dis * cosTheta - sqrtf(rmax(rad * rad - dis * dis * sinTheta * sinTheta, 0.f));
    const float cosAlpha = (dis * dis + rad * rad - sdist * sdist) / (2 * dis * rad);
    const float sinAlpha = sqrtf(rmax(1.f - cosAlpha * cosAlpha, 0.f));
    const float cphi = dcos(phi), sphi = dsin(phi);
    V3 nWorld = ((-1.f * sinAlpha * cphi) * localX - (sinAlpha * sphi) * localY) - cosAlpha * localZ;
    nWorld = norm(nWorld);
    const float off = (float)((double)rad * 1.5e-3);
    const V3 pWorld = (pCenter + (rad * nWorld) / length(nWorld)) + off * norm(nWorld);
    *pdf = 1 / (2 * PI * (1 - cosThetaMax));
    LightPoint lp;
    lp.pos = pWorld;
    lp.n = nWorld;
    lp.fn = nWorld;
    lp.pErr = gamma_n(5) * vabs(pWorld);
    return lp;
}
// Triangle::Sample (shape/Triangle.cpp:141-158) + Shape::Sample (shape/Shape.cpp:6-19)
template <bool SPH = false>
GY_D LightPoint sample_light_point(const SceneDev& S, int tri, V3 ref, float u0, float u1, float* pdf) {
    if (SPH && tri >= S.n_tri) return sph_sample_li(S, tri - S.n_tri, ref, u0, u1, pdf);
    const float4* q = S.shade + (size_t)tri * 8;
    const V3 p0 = ld3(q[0]), p1 = ld3(q[1]), p2 = ld3(q[2]);
    const float area = q[2].w;
    const V3 n0 = ld3(q[3]), n1 = ld3(q[4]), n2 = ld3(q[5]), fn = ld3(q[6]);
    const float su0 = sqrtf(u0);
    const float b0 = 1 - su0, b1 = u1 * su0, b2 = 1 - b0 - b1;
    LightPoint lp;
    lp.pos = (b0 * p0 + b1 * p1) + b2 * p2;
    lp.fn = fn;
    lp.n = (b0 * n0 + b1 * n1) + b2 * n2;
    lp.pErr = gamma_n(6) * ((vabs(b0 * p0) + vabs(b1 * p1)) + vabs(b2 * p2));
    float p = 1 / area;
    V3 wi = ref - lp.pos;
    if (length(wi) == 0.f) {
        p = 0.f;
    } else {
        wi = norm(wi);
        const V3 dd = lp.pos - ref;
        p = p * (dot(dd, dd) / fabsf(dot(lp.n, neg(wi))));
        if (isinf(p)) p = 0.f;
    }
    *pdf = p;
    return lp;
}
// Shape::Pdf (shape/Shape.h:34-47): intersect only the light's own triangle.
// Shape::Pdf (shape/Shape.h:34-47) for a sphere light: the sphere's own hit
// record is the default one (zero position and normal), so the pdf is
// |d|^2 / 0: infinite, which returns 0, or NaN when the point is the origin.
__device__ __noinline__ float sph_light_pdf(const SceneDev& S, int i, const Isect& it, V3 wi) {
    V3 o;
    spawn_ray(it, wi, &o);
    const float4 s0 = S.sph[2 * i];
    float th, u, v;
    if (!sph_hit(s0, o, wi, 0.f, kInf, &th, &u, &v)) return 0.f;
    const V3 d = v3(0.f, 0.f, 0.f) - it.pos;
    float pdf = dot(d, d) / (fabsf(dot(v3(0.f, 0.f, 0.f), wi * -1.f)) * S.sph[2 * i + 1].w);
    if (isinf(pdf)) pdf = 0.f;
    return pdf;
}
template <bool SPH = false>
GY_D float light_pdf(const SceneDev& S, int tri, const Isect& it, V3 wi) {
    if (SPH && tri >= S.n_tri) return sph_light_pdf(S, tri - S.n_tri, it, wi);
    V3 o;
    spawn_ray(it, wi, &o);
    const float4* q = S.shade + (size_t)tri * 8;
    const V3 p0 = ld3(q[0]), p1 = ld3(q[1]), p2 = ld3(q[2]);
    float th, b1, b2;
    if (!tri_test(p0, p1 - p0, p2 - p0, o, wi, 0.f, kInf, &th, &b1, &b2)) return 0.f;
    Hit h;
    h.t = th; h.b1 = b1; h.b2 = b2; h.prim = tri;
    const Isect lp = query(S, o, wi, h);
    const V3 d = lp.pos - it.pos;
    float pdf = dot(d, d) / (fabsf(dot(lp.n, wi * -1.f)) * q[2].w);
    if (isinf(pdf)) pdf = 0.f;
    return pdf;
}
// ------------------------------------------------------------------ environment light
// ImageTexture::Triangle / Texel with CLAMP (texture/ImageTexture.h:50-69).
GY_D V3 env_texel(const EnvDev& E, int u, int v) {
    u = u < 0 ? 0 : (u > E.w - 1 ? E.w - 1 : u);
    v = v < 0 ? 0 : (v > E.h - 1 ? E.h - 1 : v);
    const float* p = E.tex + ((size_t)v * E.w + u) * 3;
    return v3(p[0], p[1], p[2]);
}
// EnvironmentMap::LookupRadiance = Sample(Vec2(u, 1 - v)) * mScale (light/EnvMap.h:70-72).
GY_D V3 env_radiance(const EnvDev& E, float u, float v) {
    const float tv = 1.f - v;
    const float sx = (float)E.w * u - 0.5f, ty = (float)E.h * tv - 0.5f;
    const int s0 = (int)floorf(sx), t0 = (int)floorf(ty);
    const float ds = sx - (float)s0, dt = ty - (float)t0;
    const V3 a = v3s(1.f - ds) * env_texel(E, s0, t0) + v3s(ds) * env_texel(E, s0 + 1, t0);
    const V3 b = v3s(1.f - ds) * env_texel(E, s0, t0 + 1) + v3s(ds) * env_texel(E, s0 + 1, t0 + 1);
    return (v3s(1.f - dt) * a + v3s(dt) * b) * E.scale;
}
// DirToLatLong (light/EnvMap.h:106-121).
GY_D void env_dir_to_uv(V3 d, float* u, float* v) {
    float phi = (d.x != 0.f || d.z != 0.f) ? datan2(d.z, d.x) : 0.f;
    const float theta = dacos(d.y);
    phi = (phi < 0.f) ? (phi + 2 * kPI) : phi;
    *u = phi * kInvPI * 0.5f;
    const float vv = theta * kInvPI;
    *v = vv < 0.f ? 0.f : (vv > 1.f ? 1.f : vv);
}
// LatLongToDir (light/EnvMap.h:96-104).
GY_D V3 env_uv_to_dir(float u, float v) {
    const float phi = u * 2 * kPI;
    const float theta = v * kPI;
    const float st = dsin(theta);
    return v3(st * dcos(phi), dcos(theta), st * dsin(phi));
}
// SinTheta (light/EnvMap.h:124-132).
GY_D float env_sin_theta(float v, float height) {
    if (v < 1.f) return dsin(kPI * (float)(int)((v * height) + 0.5f) / height);
    return dsin(kPI * ((height - 1) + 0.5f) / height);
}
// Distribution1D::SampleContinuous (visual/Sampling.h:52-83): upper_bound over
// the n + 1 CDF values; *off is the segment before the zero-length fix.
GY_D float dist1d_sample(const float* __restrict__ func, const float* __restrict__ cdf, int n, float func_int, float u,
                         float* pdf, int* off) {
    int lo = 0, hi = n + 1;  // first index with cdf > u
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid;
        else lo = mid + 1;
    }
    int offset = lo - 1;
    offset = offset < 0 ? 0 : (offset > n - 1 ? n - 1 : offset);
    *off = offset;
    if (cdf[offset] == cdf[offset + 1]) {  // only u >= 1 or NaN: u in [0, 1) lands on a rising segment
        do { offset--; } while (offset > 0 && cdf[offset] == cdf[offset + 1]);
        if (offset < 0) offset = 0;
    }
    const float du = (u - cdf[offset]) / (cdf[offset + 1] - cdf[offset]);
    *pdf = func[offset] / func_int;
    return ((float)offset + du) / (float)n;
}
// Distribution2D::Pdf (visual/Sampling.h:130-140); the marginal's function is cint.
GY_D float env_dist_pdf(const EnvDev& E, float u, float v) {
    int iu = (int)(u * (float)E.w), iv = (int)(v * (float)E.h);
    iu = iu < 0 ? 0 : (iu > E.w - 1 ? E.w - 1 : iu);
    iv = iv < 0 ? 0 : (iv > E.h - 1 ? E.h - 1 : iv);
    if (E.cint[iv] * E.mint == 0.f) return 0.f;
    return (E.cfunc[(size_t)iv * E.w + iu] * E.cint[iv]) / (E.cint[iv] * E.mint);
}
// EnvironmentMap::Sample (light/EnvMap.h:18-31), Distribution2D::SampleContinuous
// (visual/Sampling.h:121-128): marginal row with u1, column with u0.
GY_D V3 env_sample(const EnvDev& E, float u0, float u1, V3* dir, float* pdfW) {
    float p0, p1;
    int v, o;
    const float uv1 = dist1d_sample(E.cint, E.mcdf, E.h, E.mint, u1, &p1, &v);
    const float uv0 = dist1d_sample(E.cfunc + (size_t)v * E.w, E.ccdf + (size_t)v * (E.w + 1), E.w, E.cint[v], u0, &p0, &o);
    const float pdf = p0 * p1;
    const float st = env_sin_theta(uv1, (float)E.h);
    *pdfW = (st == 0.f) ? 0.f : pdf / (2 * kPI * kPI * st);
    *dir = env_uv_to_dir(uv0, uv1);
    return env_radiance(E, uv0, uv1);
}
// EnvironmentMap::Lookup / PdfDir (light/EnvMap.h:33-68).
GY_D V3 env_lookup(const EnvDev& E, V3 dir) {
    float u, v;
    env_dir_to_uv(norm(dir), &u, &v);
    return env_radiance(E, u, v);
}
GY_D float env_pdf_dir(const EnvDev& E, V3 dir) {
    float u, v;
    env_dir_to_uv(norm(dir), &u, &v);
    const float st = env_sin_theta(v, (float)E.h);
    if (st == 0.f) return 0.f;
    return env_dist_pdf(E, u, v) / (2 * kPI * kPI * st);
}
// EnvironmentLight::Emission(dir) and Emission(ray) = Lookup(dir.Norm()) (light/EnvLight.cpp:11-20).
GY_D V3 env_emission(const EnvDev& E, V3 dir) { return env_lookup(E, norm(dir)); }
// EnvironmentLight::Sample_Li (light/EnvLight.cpp:41-52): the light point lies
// 2 r along wi, its normals are wi and its position error is zero.
GY_D LightPoint env_light_point(const EnvDev& E, V3 pos, float u0, float u1, V3* wi, V3* Li, float* pdf) {
    V3 sd;
    *Li = env_sample(E, u0, u1, &sd, pdf);
    *wi = norm(sd);
    LightPoint lp;
    lp.pos = pos + *wi * (2 * E.radius);
    lp.n = lp.fn = *wi;
    lp.pErr = v3s(0.f);
    return lp;
}

// Out-of-line forms for the per-path integrator, whose registers are already
// near the limit: inlined there, the lat-long mapping spills.
__device__ __noinline__ V3 env_emission_call(const EnvDev& E, V3 dir) { return env_emission(E, dir); }
__device__ __noinline__ float env_pdf_dir_call(const EnvDev& E, V3 dir) { return env_pdf_dir(E, dir); }
__device__ __noinline__ LightPoint env_light_point_call(const EnvDev& E, V3 pos, float u0, float u1, V3* wi, V3* Li,
                                                        float* pdf) {
    return env_light_point(E, pos, u0, u1, wi, Li, pdf);
}

// SpawnTo (visual/Intersection.h:64-72): shadow ray from it toward lp.
GY_D void spawn_to(const Isect& it, const LightPoint& lp, V3* o, V3* d, float* tmax) {
    const V3 origin = offset_origin(it.pos, it.pErr, it.absN, lp.pos - it.pos);
    const V3 target = offset_origin(lp.pos, lp.pErr, lp.fn, origin - lp.pos);
    V3 dd = target - origin;
    const float dis = length(dd);
    dd = norm(dd);
    *o = origin + dd * kRayEps;
    *d = dd;
    *tmax = dis - kShadowRayEps;
}
GY_D float power_heuristic(float f, float g) {  // integrator/Integrator.h:11-15
    f = 1 * f;
    g = 1 * g;
    return (f * f) / (f * f + g * g);
}

}  // namespace gyd

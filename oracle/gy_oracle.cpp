// CPU oracle for the Gyunity path-tracing hot path.  TEST INFRASTRUCTURE ONLY:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
// as the checker or as the timed CPU baseline.  See gy_oracle.h.
//
// Everything here is a restatement of the reference CPU code, templated on the
// scalar type R: R = float reproduces the reference's FP32 rounding (SSE dot
// product order, no FMA, IEEE divide and sqrt; the Makefile builds with
// -ffp-contract=off), R = double runs the same algorithm in double for the
// converged reference image.  File:line citations are into the reference tree.
//
// Parity status of each part (what pins it) is recorded beside its tests.
#include "gy_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- constants
// common/Core.h:185-199
template <class R> struct K {
    static constexpr R PI = (R)3.14159265358979;
    static constexpr R INV_PI = (R)0.31830988618379067154;
    static constexpr R PI_Over2 = (R)1.57079632679489661923;
    static constexpr R PI_Over4 = (R)0.78539816339744830961;
    static constexpr R RayEps = (R)1e-4;
    static constexpr R ShadowRayEps = (R)1e-4;
    static constexpr R MachineEps = std::numeric_limits<R>::epsilon() * (R)0.5;
    static constexpr R Inf = std::numeric_limits<R>::infinity();
};

// Transcendentals: the reference calls the CRT's float cos/sin/log, whose
// last-ulp behaviour is implementation-defined.  Float mode evaluates them in
// double and rounds once (correctly rounded but for ~2^-29 double-rounding
// cases), which the device code does too; double mode uses double libm.
inline float rcos(float x) { return (float)std::cos((double)x); }
inline float rsin(float x) { return (float)std::sin((double)x); }
inline float rlog(float x) { return (float)std::log((double)x); }
inline double rcos(double x) { return std::cos(x); }
inline double rsin(double x) { return std::sin(x); }
inline double rlog(double x) { return std::log(x); }
inline float racos(float x) { return (float)std::acos((double)x); }
inline float ratan2(float y, float x) { return (float)std::atan2((double)y, (double)x); }
inline double racos(double x) { return std::acos(x); }
inline double ratan2(double y, double x) { return std::atan2(y, x); }

// ---------------------------------------------------------------- vectors
// Vec3 follows the SSE Vector<3,float> of math/Linagl.h: lane-wise + - * /,
// Dot = DPPS mask 0x71 = (x*x' + y*y') + (z*z' + 0) (Linagl.h:496-500),
// Norm = v / sqrt(Dot) with a true divide (Linagl.h:527-531), unary minus =
// 0 - v (Linagl.h:294-297), Vec / scalar returns 0 when the scalar is 0
// (Linagl.h:366-369).
template <class R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    explicit V3(R a) : x(a), y(a), z(a) {}
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> operator+(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V3<R> operator-(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V3<R> operator*(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> inline V3<R> operator/(const V3<R>& a, const V3<R>& b) { return V3<R>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class R> inline V3<R> operator*(const V3<R>& a, R s) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template <class R> inline V3<R> operator*(R s, const V3<R>& a) { return a * s; }
template <class R> inline V3<R> operator/(const V3<R>& a, R s) {
    if (s == 0) return V3<R>(0, 0, 0);
    return V3<R>(a.x / s, a.y / s, a.z / s);
}
template <class R> inline V3<R> operator-(const V3<R>& a) { return V3<R>(R(0) - a.x, R(0) - a.y, R(0) - a.z); }
template <class R> inline bool operator==(const V3<R>& a, const V3<R>& b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
template <class R> inline bool operator!=(const V3<R>& a, const V3<R>& b) { return !(a == b); }
template <class R> inline bool is_zero(const V3<R>& a) { return a.x == 0 && a.y == 0 && a.z == 0; }
template <class R> inline R dot(const V3<R>& a, const V3<R>& b) { return (a.x * b.x + a.y * b.y) + (a.z * b.z + R(0)); }
template <class R> inline V3<R> cross(const V3<R>& a, const V3<R>& b) {
    return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> inline R length(const V3<R>& a) { return std::sqrt(dot(a, a)); }
template <class R> inline V3<R> norm(const V3<R>& a) {
    R l = length(a);
    return V3<R>(a.x / l, a.y / l, a.z / l);
}
template <class R> inline V3<R> vabs(const V3<R>& a) { return V3<R>(std::abs(a.x), std::abs(a.y), std::abs(a.z)); }
template <class R> inline V3<R> vsqrt(const V3<R>& a) { return V3<R>(std::sqrt(a.x), std::sqrt(a.y), std::sqrt(a.z)); }
// Linagl.h:569-572
template <class R> inline R lumY(const V3<R>& a) {
    const R w0 = (R)0.212671, w1 = (R)0.715160, w2 = (R)0.072169f;
    return w0 * a.x + w1 * a.y + w2 * a.z;
}
template <class R> inline R rmin(R a, R b) { return (b < a) ? b : a; }  // std::min
template <class R> inline R rmax(R a, R b) { return (a < b) ? b : a; }  // std::max
template <class R> inline R sgn(R v) { return (v > 0) ? R(1) : R(-1); }  // math/MathUtils.h:130-132
template <class R> inline R clampr(R v, R lo, R hi) {                      // math/MathUtils.h:51-56
    if (v < lo) return lo;
    if (v > hi) return hi;
    return v;
}

// ---------------------------------------------------------------- 4x4 matrices
// Matrix<4,float> stores columns; (i,j) = d[j][i] (Linagl.h:834-840).  The SSE
// mat*vec is ((c3*w + c2*z) + c1*y) + c0*x (Linagl.h:937-948), mat*mat applies
// it per column (Linagl.h:950-952), the inverse is GLM's (Linagl.h:1178-1293).
template <class R> struct V4 { R v[4]; };
template <class R> struct M4 {
    V4<R> c[4];
    static M4 rows(R m00, R m01, R m02, R m03, R m10, R m11, R m12, R m13, R m20, R m21, R m22, R m23, R m30,
                   R m31, R m32, R m33) {
        M4 m;
        R a[4][4] = {{m00, m01, m02, m03}, {m10, m11, m12, m13}, {m20, m21, m22, m23}, {m30, m31, m32, m33}};
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) m.c[j].v[i] = a[i][j];
        return m;
    }
};
template <class R> inline V4<R> mul(const M4<R>& m, const V4<R>& a) {
    V4<R> r;
    for (int i = 0; i < 4; ++i) r.v[i] = m.c[3].v[i] * a.v[3];
    for (int i = 0; i < 4; ++i) r.v[i] = m.c[2].v[i] * a.v[2] + r.v[i];
    for (int i = 0; i < 4; ++i) r.v[i] = m.c[1].v[i] * a.v[1] + r.v[i];
    for (int i = 0; i < 4; ++i) r.v[i] = m.c[0].v[i] * a.v[0] + r.v[i];
    return r;
}
template <class R> inline M4<R> mul(const M4<R>& a, const M4<R>& b) {
    M4<R> r;
    for (int j = 0; j < 4; ++j) r.c[j] = mul(a, b.c[j]);
    return r;
}
template <class R> M4<R> glm_inverse(const M4<R>& M) {
    auto m = [&](int col, int row) { return M.c[col].v[row]; };
    R Coef00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3);
    R Coef02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3);
    R Coef03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3);
    R Coef04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3);
    R Coef06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3);
    R Coef07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3);
    R Coef08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2);
    R Coef10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2);
    R Coef11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2);
    R Coef12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3);
    R Coef14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3);
    R Coef15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3);
    R Coef16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2);
    R Coef18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2);
    R Coef19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2);
    R Coef20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1);
    R Coef22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1);
    R Coef23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
    R Fac0[4] = {Coef00, Coef00, Coef02, Coef03};
    R Fac1[4] = {Coef04, Coef04, Coef06, Coef07};
    R Fac2[4] = {Coef08, Coef08, Coef10, Coef11};
    R Fac3[4] = {Coef12, Coef12, Coef14, Coef15};
    R Fac4[4] = {Coef16, Coef16, Coef18, Coef19};
    R Fac5[4] = {Coef20, Coef20, Coef22, Coef23};
    R Vec0[4] = {m(1, 0), m(0, 0), m(0, 0), m(0, 0)};
    R Vec1[4] = {m(1, 1), m(0, 1), m(0, 1), m(0, 1)};
    R Vec2[4] = {m(1, 2), m(0, 2), m(0, 2), m(0, 2)};
    R Vec3[4] = {m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
    R Inv[4][4];
    for (int k = 0; k < 4; ++k) {
        Inv[0][k] = (Vec1[k] * Fac0[k] - Vec2[k] * Fac1[k]) + Vec3[k] * Fac2[k];
        Inv[1][k] = (Vec0[k] * Fac0[k] - Vec2[k] * Fac3[k]) + Vec3[k] * Fac4[k];
        Inv[2][k] = (Vec0[k] * Fac1[k] - Vec1[k] * Fac3[k]) + Vec3[k] * Fac5[k];
        Inv[3][k] = (Vec0[k] * Fac2[k] - Vec1[k] * Fac4[k]) + Vec2[k] * Fac5[k];
    }
    const R SignA[4] = {+1, -1, +1, -1}, SignB[4] = {-1, +1, -1, +1};
    M4<R> I;
    for (int k = 0; k < 4; ++k) {
        I.c[0].v[k] = Inv[0][k] * SignA[k];
        I.c[1].v[k] = Inv[1][k] * SignB[k];
        I.c[2].v[k] = Inv[2][k] * SignA[k];
        I.c[3].v[k] = Inv[3][k] * SignB[k];
    }
    R Row0[4] = {I.c[0].v[0], I.c[1].v[0], I.c[2].v[0], I.c[3].v[0]};
    R Dot0[4];
    for (int k = 0; k < 4; ++k) Dot0[k] = M.c[0].v[k] * Row0[k];
    R Dot1 = (Dot0[0] + Dot0[1]) + (Dot0[2] + Dot0[3]);
    R One = static_cast<R>(1) / Dot1;
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) I.c[j].v[k] = I.c[j].v[k] * One;
    return I;
}
template <class R> struct Xf {  // math/Transform.h: matrix + inverse
    M4<R> m, inv;
};
template <class R> Xf<R> xf_compose(const Xf<R>& a, const Xf<R>& b) {  // Transform.h:253-255
    return Xf<R>{mul(a.m, b.m), mul(b.inv, a.inv)};
}
template <class R> Xf<R> xf_scale(R sx, R sy, R sz) {  // Transform.cpp:18-28
    return Xf<R>{M4<R>::rows(sx, 0, 0, 0, 0, sy, 0, 0, 0, 0, sz, 0, 0, 0, 0, 1),
                 M4<R>::rows(1 / sx, 0, 0, 0, 0, 1 / sy, 0, 0, 0, 0, 1 / sz, 0, 0, 0, 0, 1)};
}
template <class R> Xf<R> xf_translate(R x, R y, R z) {  // Transform.cpp:6-16
    return Xf<R>{M4<R>::rows(1, 0, 0, x, 0, 1, 0, y, 0, 0, 1, z, 0, 0, 0, 1),
                 M4<R>::rows(1, 0, 0, -x, 0, 1, 0, -y, 0, 0, 1, -z, 0, 0, 0, 1)};
}
// Transform::operator()(point), math/Transform.h:106-115
template <class R> V3<R> xf_point(const M4<R>& m, const V3<R>& p) {
    V4<R> pp{{p.x, p.y, p.z, (R)1.0}};
    pp = mul(m, pp);
    if (pp.v[3] == 1 || pp.v[3] == 0) return V3<R>(pp.v[0], pp.v[1], pp.v[2]);
    return V3<R>(pp.v[0], pp.v[1], pp.v[2]) / pp.v[3];
}

// ---------------------------------------------------------------- camera
// ProjectiveCamera::Initialize (camera/Camera.cpp:6-38) with the lookAt
// constructor (Camera.cpp:68-72), Transform::Perspective (Transform.cpp:132-150).
template <class R> struct Camera {
    V3<R> pos;
    M4<R> raster_to_world;
};
template <class R> Camera<R> make_camera(const or_camera_desc& d) {
    Camera<R> c;
    V3<R> pos((R)d.pos[0], (R)d.pos[1], (R)d.pos[2]);
    V3<R> look((R)d.look_at[0], (R)d.look_at[1], (R)d.look_at[2]);
    V3<R> up((R)d.up[0], (R)d.up[1], (R)d.up[2]);
    V3<R> cz = -norm(pos - look);
    V3<R> cx = -norm(cross(up, cz));
    V3<R> cy = -cross(cz, cx);
    M4<R> c2w = M4<R>::rows(cx.x, cy.x, cz.x, pos.x, cx.y, cy.y, cz.y, pos.y, cx.z, cy.z, cz.z, pos.z, 0, 0, 0, (R)1.f);
    R fovy = (R)d.fovy, dis = (R)d.film_dist;
    R aspect = (R)d.width / (R)d.height;  // visual/Film.h:40
    R radians = (K<R>::PI / 180) * fovy;  // math/MathUtils.h:63
    R t = dis * std::tan(radians * 0.5);  // double tan, as written (Transform.cpp:134)
    R r = t * aspect;
    R f = (R)100000.f, n = dis;
    M4<R> persp = M4<R>::rows(dis / r, 0, 0, 0, 0, dis / t, 0, 0, 0, 0, (f + n) / (f - n), -2 * f * n / (f - n), 0, 0, 1, 0);
    M4<R> ndc_to_camera = glm_inverse(persp);
    Xf<R> ndc_to_raster = xf_compose(xf_compose(xf_scale((R)d.width, (R)d.height, (R)1), xf_scale((R)0.5f, (R)-0.5f, (R)1)),
                                     xf_translate((R)1, (R)-1, (R)0));
    M4<R> raster_to_ndc = ndc_to_raster.inv;
    M4<R> raster_to_camera = mul(ndc_to_camera, raster_to_ndc);
    c.raster_to_world = mul(c2w, raster_to_camera);
    c.pos = pos;
    return c;
}

template <class R> struct Ray {
    V3<R> o, d;
    R tmin, tmax;
};
// math/Ray.h:17-19: the constructor moves the origin by RayEps along dir.
template <class R> inline Ray<R> make_ray(const V3<R>& o, const V3<R>& d, R tmax, R tmin) {
    return Ray<R>{o + d * K<R>::RayEps, d, tmin, tmax};
}
// PinHoleCamera::GenerateRay (camera/PinholeCamera.h:22-29), offset = 0.
template <class R> Ray<R> generate_ray(const Camera<R>& c, int px, int py, R u, R v) {
    V3<R> sampled = xf_point(c.raster_to_world, V3<R>(px + u, py + v, (R)-1));
    V3<R> dir = norm(sampled - c.pos);
    return make_ray(c.pos + dir * (R)0.0, dir, K<R>::Inf, (R)0);
}

// ---------------------------------------------------------------- Sobol
// sobol::sample (sampler/Sobol.h:46-61) over the fixture copy of the
// reference's matrices; SobolEnum::look_up (sampler/SobolEnum.h:42-72).
static std::vector<uint32_t> g_sobol;  // ndim * 52
static uint32_t g_sobol_ndim = 0;
static uint64_t g_vdc[25][52], g_vdc_inv[26][52];
static bool g_sobol_ok = false;

inline float sobol_sample(uint64_t index, uint32_t dim) {
    if (dim >= g_sobol_ndim) {
        std::fprintf(stderr, "oracle: Sobol dimension %u beyond the %u loaded\n", dim, g_sobol_ndim);
        std::abort();
    }
    unsigned result = 0;
    for (unsigned i = dim * 52; index; index >>= 1, ++i)
        if (index & 1) result ^= g_sobol[i];
    return result * (1.f / (1ULL << 32));
}
inline uint64_t enum_look_up(uint32_t m, uint32_t frame, uint32_t px, uint32_t py) {
    if (m == 0) return 0;
    const uint32_t m2 = m << 1;
    uint64_t index = uint64_t(frame) << m2;
    uint64_t delta = 0;
    for (uint32_t c = 0; frame; frame >>= 1, ++c)
        if (frame & 1) delta ^= g_vdc[m - 1][c];
    uint64_t b = ((uint64_t(px) << m) | py) ^ delta;
    for (uint32_t c = 0; b; b >>= 1, ++c)
        if (b & 1) index ^= g_vdc_inv[m - 1][c];
    return index;
}
// math/MathUtils.h RoundUpPow2 / Log2Int, SobolEnum ctor (SobolEnum.h:17-21)
inline int round_up_pow2(int v) {
    v--;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    return v + 1;
}
inline int log2_int(int v) {
    int r = 0;
    while ((1 << (r + 1)) <= v) ++r;
    return r;
}
// RandomStateSequence::Sample (sampler/Sampler.h:53-64)
// ---------------------------------------------------------------- Halton
// HaltonSampler with Faure permutations (sampler/Halton.h:595-625, 648-923,
// 1441-3308), evaluated digit by digit: dimension d in base b = the (d+1)-th
// prime sums perm_b(digit_m) * b^(n-1-m) over the n = k*G lowest digits of the
// index (k digits per table entry with b^k <= 512, G entries with b^(kG) <=
// 2^32, as the reference's generated functions use) and scales the integer by
// float((1 - 2^-23) / b^n).  Base 2 writes the reversed index bits into a
// float mantissa.  HaltonEnum (sampler/HaltonEnum.h:61-143) by the CRT.
namespace hal {
inline uint64_t reverse64(uint64_t v) {
    uint64_t r = 0;
    for (int i = 0; i < 64; ++i, v >>= 1) r = (r << 1) | (v & 1);
    return r;
}
struct Tables {
    std::vector<uint32_t> primes;
    std::vector<std::vector<uint16_t>> perm;  // Faure permutation by base
    Tables() {
        for (uint32_t c = 2; primes.size() < 256; ++c) {
            bool is_p = true;
            for (uint32_t q = 2; q * q <= c; ++q)
                if (c % q == 0) { is_p = false; break; }
            if (is_p) primes.push_back(c);
        }
        const uint32_t maxb = primes.back();
        perm.resize(maxb + 1);
        for (uint32_t k = 1; k <= 3; ++k)
            for (uint32_t i = 0; i < k; ++i) perm[k].push_back((uint16_t)i);
        for (uint32_t base = 4; base <= maxb; ++base) {
            perm[base].resize(base);
            const uint32_t b = base / 2;
            if (base & 1) {
                for (uint32_t i = 0; i < base - 1; ++i)
                    perm[base][i + (i >= b)] = (uint16_t)(perm[base - 1][i] + (perm[base - 1][i] >= b));
                perm[base][b] = (uint16_t)b;
            } else {
                for (uint32_t i = 0; i < b; ++i) {
                    perm[base][i] = (uint16_t)(2 * perm[b][i]);
                    perm[base][b + i] = (uint16_t)(2 * perm[b][i] + 1);
                }
            }
        }
    }
};
inline const Tables& tables() {
    static const Tables t;
    return t;
}
inline float sample(uint32_t dim, uint64_t index) {
    if (dim == 0) {
        uint32_t u = 0x3f800000u | (uint32_t)(reverse64(index) >> 41);
        float f;
        std::memcpy(&f, &u, 4);
        return f - 1.f;
    }
    if (dim >= 256) return 0.f;
    const Tables& T = tables();
    const uint32_t b = T.primes[dim];
    int k = 1;
    while (std::pow((double)b, k + 1) <= 512.0) ++k;
    int G = 1;
    while (std::pow((double)b, k * (G + 1)) <= 4294967296.0) ++G;
    const int n = k * G;
    uint64_t sum = 0, x = index;
    for (int m = 0; m < n; ++m) {
        sum = sum * b + T.perm[b][x % b];
        x /= b;
    }
    const double denom = std::pow((double)b, n);
    return (float)sum * (float)(0x1.fffffcp-1 / denom);
}
struct Enum {
    uint32_t p2 = 0, p3 = 0, mx = 0, my = 0;
    uint64_t inc = 1;
    uint32_t w = 1, h = 1;
    static std::pair<int, int> euclid(int a, int b) {
        if (!b) return std::make_pair(1, 0);
        const std::pair<int, int> st = euclid(b, a % b);
        return std::make_pair(st.second, st.first - (a / b) * st.second);
    }
    Enum(uint32_t width, uint32_t height) {
        while (w < width) { w *= 2; ++p2; }
        while (h < height) { h *= 3; ++p3; }
        inc = (uint64_t)w * h;
        const std::pair<int, int> inv = euclid((int)h, (int)w);
        const uint32_t inv2 = inv.first < 0 ? (uint32_t)(inv.first + (int)w) : (uint32_t)(inv.first % (int)w);
        const uint32_t inv3 = inv.second < 0 ? (uint32_t)(inv.second + (int)h) : (uint32_t)(inv.second % (int)h);
        mx = h * inv2;
        my = w * inv3;
    }
    uint64_t index(uint64_t i, uint32_t x, uint32_t y) const {
        const uint64_t hx = p2 ? reverse64(x) >> (64 - p2) : 0;
        uint64_t hy = 0, yy = y;
        for (uint32_t d = 0; d < p3; ++d) {
            hy = hy * 3 + yy % 3;
            yy /= 3;
        }
        return (hx * mx + hy * my) % inc + i * inc;
    }
};
}  // namespace hal

// RandomStateSequence (sampler/Sampler.h:42-66) over the chosen sampler.
struct Seq {
    uint64_t instance;
    int cursor = 0;
    int kind = 0;  // 0 Sobol, 1 Halton
    float next() {
        float ret = kind == 1 ? hal::sample((uint32_t)cursor, instance) : sobol_sample(instance, (uint32_t)cursor);
        ++cursor;
        if (ret >= 1) ret = 0;
        return ret;
    }
};

// ---------------------------------------------------------------- numeric utils
// numeric/NumericUtils.h:22-51
inline float next_float_up(float v) {
    if (std::isinf(v) && v > 0.) return v;
    if (v == -0.f) v = 0.f;
    uint32_t ui;
    std::memcpy(&ui, &v, 4);
    if (v >= 0) ++ui; else --ui;
    std::memcpy(&v, &ui, 4);
    return v;
}
inline float next_float_down(float v) {
    if (std::isinf(v) && v < 0.) return v;
    if (v == 0.f) v = -0.f;
    uint32_t ui;
    std::memcpy(&ui, &v, 4);
    if (v > 0) --ui; else ++ui;
    std::memcpy(&v, &ui, 4);
    return v;
}
inline double next_float_up(double v) {
    if (std::isinf(v) && v > 0.) return v;
    if (v == -0.0) v = 0.0;
    uint64_t ui;
    std::memcpy(&ui, &v, 8);
    if (v >= 0) ++ui; else --ui;
    std::memcpy(&v, &ui, 8);
    return v;
}
inline double next_float_down(double v) {
    if (std::isinf(v) && v < 0.) return v;
    if (v == 0.0) v = -0.0;
    uint64_t ui;
    std::memcpy(&ui, &v, 8);
    if (v > 0) --ui; else ++ui;
    std::memcpy(&v, &ui, 8);
    return v;
}
template <class R> inline R gamma_n(int n) { return (n * K<R>::MachineEps) / (1 - n * K<R>::MachineEps); }

// math/GeometryUtils.h:39-64 (the float build: no USING_DOUBLE widening)
template <class R> V3<R> offset_ray_origin(const V3<R>& p, const V3<R>& pErr, const V3<R>& n, const V3<R>& w) {
    R d = dot(vabs(n), pErr);
    V3<R> offset = d * n;
    if (dot(w, n) < 0) offset = -offset;
    V3<R> po = p + offset;
    for (int i = 0; i < 3; ++i) {
        if (offset[i] > 0) po[i] = next_float_up(po[i]);
        else if (offset[i] < 0) po[i] = next_float_down(po[i]);
    }
    return po;
}

// ---------------------------------------------------------------- scene
template <class R> struct AABB {
    V3<R> mn{K<R>::Inf, K<R>::Inf, K<R>::Inf}, mx{-K<R>::Inf, -K<R>::Inf, -K<R>::Inf};
    const V3<R>& operator[](int i) const { return i == 0 ? mn : mx; }
};
template <class R> inline AABB<R> unite(const AABB<R>& a, const AABB<R>& b) {  // math/AABB.cpp:30-45
    AABB<R> r;
    r.mn = V3<R>(rmin(a.mn.x, b.mn.x), rmin(a.mn.y, b.mn.y), rmin(a.mn.z, b.mn.z));
    r.mx = V3<R>(rmax(a.mx.x, b.mx.x), rmax(a.mx.y, b.mx.y), rmax(a.mx.z, b.mx.z));
    return r;
}
template <class R> inline AABB<R> unite(const AABB<R>& a, const V3<R>& p) {  // math/AABB.cpp:47-60
    AABB<R> r;
    r.mn = V3<R>(rmin(a.mn.x, p.x), rmin(a.mn.y, p.y), rmin(a.mn.z, p.z));
    r.mx = V3<R>(rmax(a.mx.x, p.x), rmax(a.mx.y, p.y), rmax(a.mx.z, p.z));
    return r;
}
template <class R> inline R surface_area(const AABB<R>& b) {  // math/AABB.h:45-48
    V3<R> d = b.mx - b.mn;
    return 2 * (d.x * d.y + d.x * d.z + d.y * d.z);
}
template <class R> inline int max_extent(const AABB<R>& b) {  // math/AABB.h:55-62
    V3<R> d = b.mx - b.mn;
    if (d.x > d.y && d.x > d.z) return 0;
    else if (d.y > d.z) return 1;
    return 2;
}

struct Material {
    uint32_t type, flags;
    float R[3], T[3], ax, ay, eta[3], k[3], f0[3], eta1, eta2;
    int32_t tex_albedo, tex_rough;
};

template <class R> struct Texture {
    int w, h, c;
    std::vector<R> data;  // row-major rows of w texels, c channels
};

template <class R> struct Tri {
    V3<R> p0, p1, p2, n0, n1, n2, fn, e1, e2;
    R uv[3][2];
    uint32_t mat;
    int32_t light;
    R area;
};

template <class R> struct BVHNode {  // accelerator/BVHAccel.cpp:114-124
    AABB<R> bounds;
    uint32_t offset;  // primitivesOffset (leaf) / secondChildOffset (interior)
    uint8_t nPrims, axis;
};

// Distribution1D (visual/Sampling.h:8-98): piecewise-constant density on [0, 1).
template <class R> struct Dist1D {
    std::vector<R> func, cdf;
    R funcInt = 0;
    int count = 0;
    void build(const R* f, int n) {
        count = n;
        func.assign(f, f + n);
        cdf.assign(n + 1, 0);
        cdf[0] = 0.;
        for (int i = 1; i < n + 1; ++i) cdf[i] = cdf[i - 1] + func[i - 1] / n;
        funcInt = cdf[n];
        if (funcInt == 0.f) {
            for (int i = 1; i < n + 1; ++i) cdf[i] = R(i) / R(n);
        } else {
            for (int i = 1; i < n + 1; ++i) cdf[i] /= funcInt;
        }
    }
    // SampleContinuous (visual/Sampling.h:52-83); `off` is the segment before
    // the zero-length fix, as there.
    R sample_continuous(R u, R* pdf, int* off) const {
        auto it = std::upper_bound(cdf.begin(), cdf.begin() + count + 1, u);
        int offset = clampr(int(std::distance(cdf.begin(), it) - 1), 0, count - 1);
        if (off) *off = offset;
        if (cdf[offset] == cdf[offset + 1]) {  // only u >= 1 or NaN reach here
            do { offset--; } while (offset > 0 && cdf[offset] == cdf[offset + 1]);
            if (offset < 0) offset = 0;
        }
        R du = (u - cdf[offset]) / (cdf[offset + 1] - cdf[offset]);
        if (pdf) *pdf = func[offset] / funcInt;
        return (offset + du) / count;
    }
};

// EnvironmentMap + EnvironmentLight (light/EnvMap.h:7-134, light/EnvLight.cpp).
// tex holds ImageTexture rows: ImageIO::LoadTexture puts the image's last row
// first (image/ImageIO.cpp:100-114).
template <class R> struct Env {
    int w = 0, h = 0;  // w == 0: no environment light
    R scale = 1;
    std::vector<V3<R>> tex;
    std::vector<Dist1D<R>> cond;  // Distribution2D: one per image row, top row first
    Dist1D<R> marg;
    V3<R> center;
    R radius = 0;
};

// HomogeneousMedium (volume/Homogeneous.h:9-24): sigma_t = sigma_s + sigma_a.
template <class R> struct Medium {
    V3<R> sa, ss, st;
    R g = 0;
};
template <class R> struct Scene {
    Env<R> env;
    std::vector<Tri<R>> tris;
    std::vector<Material> mats;
    std::vector<int32_t> light_tri;
    std::vector<V3<R>> light_le;
    std::vector<R> light_cdf, light_func;
    R light_func_int = 0;
    std::vector<Texture<R>> tex;
    std::vector<BVHNode<R>> nodes;
    std::vector<uint32_t> ordered;  // BVH leaf order -> primitive id
    int max_depth = 0;
    // Test-only detached evaluation (or_set_detached): paths are sampled with
    // tex and light_le (directions, pdfs, MIS weights, light choice), while
    // BSDF values and emitted radiance come from tex_eval and light_le_eval.
    // Differences of renders in this mode are the derivative a path-replay
    // adjoint estimates, which holds every sampling decision fixed.
    bool detached = false;
    std::vector<Texture<R>> tex_eval;
    std::vector<V3<R>> light_le_eval;
    // Analytic spheres (or_scene_set_spheres): primitive ids after the
    // triangles, n_tri + i.
    struct Sph {
        V3<R> c;
        R r = 0, area = 0;
        bool inside = false;
        uint32_t mat = 0;
        int32_t light = -1;
        AABB<R> box;  // ObjectToWorld(AABB(c - r, c + r)) with the identity transform
    };
    std::vector<Sph> sph;
    // Participating media (or_scene_set_media): per triangle the medium
    // interface's inside and outside media (-1: none) and whether the
    // primitive has no material (a medium boundary).  Empty: no media.
    std::vector<Medium<R>> media;
    std::vector<int32_t> mi_in, mi_out;
    std::vector<uint8_t> no_mat;
};
template <class R> inline uint32_t prim_mat(const Scene<R>& s, int32_t prim) {
    return prim < (int32_t)s.tris.size() ? s.tris[prim].mat : s.sph[prim - s.tris.size()].mat;
}
template <class R> inline int32_t prim_light(const Scene<R>& s, int32_t prim) {
    return prim < (int32_t)s.tris.size() ? s.tris[prim].light : s.sph[prim - s.tris.size()].light;
}

// Triangle(..., normal) / Triangle(..., n0, n1, n2) + Initialize
// (shape/Triangle.h:13-55, 107-118); Area (Triangle.h:67-69).
template <class R> void build_tris(Scene<R>& s, const or_scene_desc& d) {
    s.tris.resize(d.n_tri);
    for (int i = 0; i < d.n_tri; ++i) {
        Tri<R>& t = s.tris[i];
        const float* p = d.pos + 9 * i;
        const float* n = d.nrm + 9 * i;
        t.p0 = V3<R>(p[0], p[1], p[2]);
        t.p1 = V3<R>(p[3], p[4], p[5]);
        t.p2 = V3<R>(p[6], p[7], p[8]);
        t.n0 = V3<R>(n[0], n[1], n[2]);
        t.n1 = V3<R>(n[3], n[4], n[5]);
        t.n2 = V3<R>(n[6], n[7], n[8]);
        t.fn = V3<R>(d.face_n[3 * i], d.face_n[3 * i + 1], d.face_n[3 * i + 2]);
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 2; ++c) t.uv[k][c] = d.uv ? (R)d.uv[6 * i + 2 * k + c] : (R)0;
        t.e1 = t.p1 - t.p0;
        t.e2 = t.p2 - t.p0;
        t.mat = d.mat[i];
        t.light = d.light ? d.light[i] : -1;
        t.area = (R)(0.5 * length(cross(t.p1 - t.p0, t.p2 - t.p0)));
    }
}

// ---------------------------------------------------------------- BVH build
// BVHAccel::BVHAccel / RecursiveBuild / FlattenBVHTree
// (accelerator/BVHAccel.cpp:39-48, 154-200, 208-393), SAH split, maxPrims = 1.
template <class R> struct PrimInfo {
    uint32_t prim;
    V3<R> centroid;
    AABB<R> bounds;
};
template <class R> struct BuildNode {
    AABB<R> bounds;
    BuildNode* child[2] = {nullptr, nullptr};
    uint32_t axis = 0, first = 0, n = 0;
};
template <class R> struct Builder {
    Scene<R>& s;
    std::vector<PrimInfo<R>> info;
    std::deque<BuildNode<R>>* pool;
    uint32_t total = 0;
    BuildNode<R>* alloc() {
        pool->emplace_back();
        return &pool->back();
    }
    BuildNode<R>* build(uint32_t start, uint32_t end, int depth) {
        total++;
        if (depth > s.max_depth) s.max_depth = depth;
        BuildNode<R>* node = alloc();
        AABB<R> box;
        for (uint32_t i = start; i < end; ++i) box = unite(box, info[i].bounds);
        uint32_t nP = end - start;
        auto make_leaf = [&]() {
            uint32_t first = (uint32_t)s.ordered.size();
            for (uint32_t i = start; i < end; ++i) s.ordered.push_back(info[i].prim);
            node->first = first;
            node->n = nP;
            node->bounds = box;
            return node;
        };
        if (nP == 1) return make_leaf();
        AABB<R> cb;
        for (uint32_t i = start; i < end; ++i) cb = unite(cb, info[i].centroid);
        int dim = max_extent(cb);
        uint32_t mid = (start + end) / 2;
        auto interior = [&](uint32_t m) {
            BuildNode<R>* c0 = build(start, m, depth + 1);
            BuildNode<R>* c1 = build(m, end, depth + 1);
            node->child[0] = c0;
            node->child[1] = c1;
            node->bounds = unite(c0->bounds, c1->bounds);
            node->axis = dim;
            node->n = 0;
            return node;
        };
        if (cb.mx[dim] == cb.mn[dim]) {
            // maxPrimsInNode = 1 < nP here: split at the midpoint
            return interior(mid);
        }
        if (nP <= 4) {
            mid = (start + end) / 2;
            std::nth_element(&info[start], &info[mid], &info[end - 1] + 1,
                             [dim](const PrimInfo<R>& a, const PrimInfo<R>& b) { return a.centroid[dim] < b.centroid[dim]; });
        } else {
            const int nB = 12;
            int count[nB] = {0};
            AABB<R> bb[nB];
            for (uint32_t i = start; i < end; ++i) {
                int b = nB * ((info[i].centroid[dim] - cb.mn[dim]) / (cb.mx[dim] - cb.mn[dim]));
                if (b == nB) b = nB - 1;
                count[b]++;
                bb[b] = unite(bb[b], info[i].bounds);
            }
            R cost[nB - 1];
            for (int i = 0; i < nB - 1; ++i) {
                AABB<R> b0, b1;
                int c0 = 0, c1 = 0;
                for (int j = 0; j <= i; ++j) { b0 = unite(b0, bb[j]); c0 += count[j]; }
                for (int j = i + 1; j < nB; ++j) { b1 = unite(b1, bb[j]); c1 += count[j]; }
                cost[i] = (R).125f + (c0 * surface_area(b0) + c1 * surface_area(b1)) / surface_area(box);
            }
            R minCost = cost[0];
            uint32_t split = 0;
            for (int i = 1; i < nB - 1; ++i)
                if (cost[i] < minCost) { minCost = cost[i]; split = i; }
            // nPrimitives > maxPrimsInNode (= 1) always holds here
            const AABB<R> cbc = cb;
            PrimInfo<R>* pmid = std::partition(&info[start], &info[end - 1] + 1, [&](const PrimInfo<R>& p) {
                int b = nB * ((p.centroid[dim] - cbc.mn[dim]) / (cbc.mx[dim] - cbc.mn[dim]));
                if (b == nB) b = nB - 1;
                return b <= (int)split;
            });
            mid = (uint32_t)(pmid - &info[0]);
        }
        return interior(mid);
    }
    uint32_t flatten(BuildNode<R>* node, uint32_t* offset) {
        BVHNode<R>& ln = s.nodes[*offset];
        ln.bounds = node->bounds;
        uint32_t my = (*offset)++;
        if (node->n > 0) {
            ln.offset = node->first;
            ln.nPrims = (uint8_t)node->n;
        } else {
            ln.axis = (uint8_t)node->axis;
            ln.nPrims = 0;
            flatten(node->child[0], offset);
            uint32_t second = flatten(node->child[1], offset);
            s.nodes[my].offset = second;
        }
        return my;
    }
};
template <class R> void build_bvh(Scene<R>& s) {
    Builder<R> b{s};
    std::deque<BuildNode<R>> pool;
    b.pool = &pool;
    b.info.resize(s.tris.size());
    for (size_t i = 0; i < s.tris.size(); ++i) {
        const Tri<R>& t = s.tris[i];
        AABB<R> box;
        box.mn = V3<R>(rmin(t.p0.x, t.p1.x), rmin(t.p0.y, t.p1.y), rmin(t.p0.z, t.p1.z));
        box.mx = V3<R>(rmax(t.p0.x, t.p1.x), rmax(t.p0.y, t.p1.y), rmax(t.p0.z, t.p1.z));
        box = unite(box, t.p2);
        b.info[i].prim = (uint32_t)i;
        b.info[i].bounds = box;
        b.info[i].centroid = (R).5f * box.mn + (R).5f * box.mx;
    }
    if (s.tris.empty()) return;
    s.ordered.reserve(s.tris.size());
    BuildNode<R>* root = b.build(0, (uint32_t)s.tris.size(), 0);
    s.nodes.resize(b.total);
    uint32_t off = 0;
    b.flatten(root, &off);
}

// ---------------------------------------------------------------- intersection
// Triangle::Intersect (shape/Triangle.cpp:71-103, 106-139)
template <class R> inline bool tri_hit(const Tri<R>& tr, const Ray<R>& ray, R* t, R* b1o, R* b2o) {
    V3<R> s1 = cross(ray.d, tr.e2);
    R divisor = dot(s1, tr.e1);
    if (divisor == 0.) return false;
    R inv = 1.f / divisor;
    V3<R> s = ray.o - tr.p0;
    R b1 = dot(s, s1) * inv;
    if (b1 < 0. || b1 > 1.) return false;
    V3<R> s2 = cross(s, tr.e1);
    R b2 = dot(ray.d, s2) * inv;
    if (b2 < 0. || b1 + b2 > 1.) return false;
    R th = dot(tr.e2, s2) * inv;
    if (th < ray.tmin || th > ray.tmax) return false;
    *t = th;
    *b1o = b1;
    *b2o = b2;
    return true;
}
// IntersectP (accelerator/BVHAccel.cpp:127-149)
template <class R> inline bool box_hit(const AABB<R>& bounds, const Ray<R>& ray, const V3<R>& invDir, const uint32_t neg[3]) {
    R tmin = (bounds[neg[0]].x - ray.o.x) * invDir.x;
    R tmax = (bounds[1 - neg[0]].x - ray.o.x) * invDir.x;
    R tymin = (bounds[neg[1]].y - ray.o.y) * invDir.y;
    R tymax = (bounds[1 - neg[1]].y - ray.o.y) * invDir.y;
    if ((tmin > tymax) || (tymin > tmax)) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    R tzmin = (bounds[neg[2]].z - ray.o.z) * invDir.z;
    R tzmax = (bounds[1 - neg[2]].z - ray.o.z) * invDir.z;
    if ((tmin > tzmax) || (tzmin > tmax)) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return (tmin < ray.tmax) && (tmax > ray.tmin);
}
template <class R> struct Hit {
    R t, b1, b2;
    int32_t prim = -1;
};
// BVHAccel::Intersect closest (BVHAccel.cpp:401-445) with Primitive::Intersect
// shrinking tMax (visual/Primitive.cpp:5-18); brute force in scene order
// (accelerator/BruteForceAccel.h:20-29).
// ---------------------------------------------------------------- analytic spheres
// Sphere (shape/Sphere.h:10-351) with the identity object transform, which is
// what every reference scene gives its spheres (Transform()).  The identity
// applied to a point or vector (math/Transform.h:106-115, 167-172; SSE column
// order ((c3 w + c2 z) + c1 y) + c0 x) is exact except that -0 becomes +0.
template <class R> inline V3<R> xf_identity(const V3<R>& p) { return V3<R>(R(0) + p.x, R(0) + p.y, R(0) + p.z); }
// WorldToObject(ray) (math/Transform.h:226-239): the origin with its gamma(3)
// error bound, moved dt along d, tMax - dt, tMin + dt, through the Ray
// constructor, which moves the origin RayEps along d again.
template <class R> inline Ray<R> sphere_object_ray(const Ray<R>& ray) {
    V3<R> oErr = gamma_n<R>(3) * V3<R>(std::abs(ray.o.x), std::abs(ray.o.y), std::abs(ray.o.z));
    V3<R> o = xf_identity(ray.o), d = xf_identity(ray.d);
    R length2 = dot(d, d), tMax = ray.tmax, tMin = ray.tmin;
    if (length2 > 0) {
        R dt = dot(vabs(d), oErr) / length2;
        o = o + d * dt;
        tMax = tMax - dt;
        tMin = tMin + dt;
    }
    return make_ray(o, d, tMax, tMin);
}
// Sphere::Intersect(const Ray&, Intersection*, real*) (shape/Sphere.h:29-128):
// the hit distance and (u, v).  Its intersection record is built and then
// dropped (`ObjectToWorld(it)` is evaluated and discarded, and *isect never
// receives it), so a hit leaves only uv, the shape and the primitive behind;
// sphere_isect restates that.
template <class R> bool sphere_hit(const typename Scene<R>::Sph& sp, const Ray<R>& ray, R* t, R* u, R* v) {
    const Ray<R> r = sphere_object_ray(ray);
    V3<R> po = r.o - sp.c;
    R vd = dot(po, r.d), d2 = dot(r.d, r.d), po2 = dot(po, po);
    R det = vd * vd - d2 * po2 + d2 * sp.r * sp.r;
    if (det < 0) return false;
    det = std::sqrt(det);
    R tHit = (-vd - det) / d2;
    if (tHit < r.tmin || tHit > r.tmax) {
        tHit = (-vd + det) / d2;
        if (tHit < r.tmin || tHit > r.tmax) return false;
    }
    V3<R> hit = r.o + r.d * tHit;
    V3<R> scaledDir = (hit - sp.c) * sp.r / length(hit - sp.c);
    hit = sp.c + scaledDir;
    V3<R> pHit = hit - sp.c;
    R phi = ratan2(pHit.z, pHit.x);
    if (phi < 0) phi += 2 * K<R>::PI;
    const R phiMax = 2 * K<R>::PI;
    *u = phi / phiMax;
    R theta = racos(clampr(pHit.y / sp.r, (R)-1, (R)1));
    R vv = (theta - 0) / (K<R>::PI - 0);
    *v = (R)1.f - vv;
    *t = tHit;
    return true;
}
template <class R> bool intersect(const Scene<R>& s, Ray<R>& ray, Hit<R>* h, bool use_bvh, int64_t* cnt_node = nullptr,
                                  int64_t* cnt_tri = nullptr) {
    bool hit = false;
    if (!use_bvh) {
        for (size_t i = 0; i < s.tris.size(); ++i) {
            R t, b1, b2;
            if (tri_hit(s.tris[i], ray, &t, &b1, &b2)) {
                ray.tmax = t;
                h->t = t; h->b1 = b1; h->b2 = b2; h->prim = (int32_t)i;
                hit = true;
            }
        }
        for (size_t i = 0; i < s.sph.size(); ++i) {
            R t, u, v;
            if (sphere_hit<R>(s.sph[i], ray, &t, &u, &v)) {
                ray.tmax = t;
                h->t = t; h->b1 = u; h->b2 = v; h->prim = (int32_t)(s.tris.size() + i);
                hit = true;
            }
        }
        return hit;
    }
    // spheres (few; one leaf each in the reference's tree): the leaf box
    // gate (BVHAccel::IntersectP) at the current tMax, then the sphere
    if (!s.sph.empty()) {
        V3<R> inv(1.f / ray.d.x, 1.f / ray.d.y, 1.f / ray.d.z);
        uint32_t ng[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
        for (size_t i = 0; i < s.sph.size(); ++i) {
            R t, u, v;
            if (box_hit(s.sph[i].box, ray, inv, ng) && sphere_hit<R>(s.sph[i], ray, &t, &u, &v)) {
                ray.tmax = t;
                h->t = t; h->b1 = u; h->b2 = v; h->prim = (int32_t)(s.tris.size() + i);
                hit = true;
            }
        }
    }
    if (s.nodes.empty()) return hit;
    V3<R> invDir(1.f / ray.d.x, 1.f / ray.d.y, 1.f / ray.d.z);
    uint32_t neg[3] = {invDir.x < 0, invDir.y < 0, invDir.z < 0};
    uint32_t todoOffset = 0, nodeNum = 0;
    uint32_t todo[256];
    while (true) {
        const BVHNode<R>& node = s.nodes[nodeNum];
        if (cnt_node) ++*cnt_node;
        if (box_hit(node.bounds, ray, invDir, neg)) {
            if (node.nPrims > 0) {
                for (uint32_t i = 0; i < node.nPrims; ++i) {
                    uint32_t prim = s.ordered[node.offset + i];
                    if (cnt_tri) ++*cnt_tri;
                    R t, b1, b2;
                    if (tri_hit(s.tris[prim], ray, &t, &b1, &b2)) {
                        ray.tmax = t;
                        h->t = t; h->b1 = b1; h->b2 = b2; h->prim = (int32_t)prim;
                        hit = true;
                    }
                }
                if (todoOffset == 0) break;
                nodeNum = todo[--todoOffset];
            } else {
                if (todoOffset >= 255) { std::fprintf(stderr, "oracle: BVH stack overflow\n"); std::abort(); }
                if (neg[node.axis]) {
                    todo[todoOffset++] = nodeNum + 1;
                    nodeNum = node.offset;
                } else {
                    todo[todoOffset++] = node.offset;
                    nodeNum = nodeNum + 1;
                }
            }
        } else {
            if (todoOffset == 0) break;
            nodeNum = todo[--todoOffset];
        }
    }
    return hit;
}
// BVHAccel::Intersect any-hit (BVHAccel.cpp:448-487)
template <class R> bool intersect_any(const Scene<R>& s, const Ray<R>& ray, bool use_bvh) {
    R t, b1, b2;
    // Sphere::Intersect(const Ray&) (shape/Sphere.h:131-161) drops
    // WorldToObject(ray) and tests a default Ray (zero origin and direction):
    // tHit = 0 / 0 passes both range tests, so any sphere the traversal
    // reaches occludes the ray.  The reference's tree reaches a sphere's leaf
    // when the ray passes its box (BVHAccel::IntersectP); brute force reaches
    // every sphere.
    if (!s.sph.empty()) {
        if (!use_bvh) return true;
        V3<R> inv(1.f / ray.d.x, 1.f / ray.d.y, 1.f / ray.d.z);
        uint32_t ng[3] = {inv.x < 0, inv.y < 0, inv.z < 0};
        for (const auto& sp : s.sph)
            if (box_hit(sp.box, ray, inv, ng)) return true;
    }
    if (!use_bvh) {
        for (size_t i = 0; i < s.tris.size(); ++i)
            if (tri_hit(s.tris[i], ray, &t, &b1, &b2)) return true;
        return false;
    }
    if (s.nodes.empty()) return false;
    V3<R> invDir(1.f / ray.d.x, 1.f / ray.d.y, 1.f / ray.d.z);
    uint32_t neg[3] = {invDir.x < 0, invDir.y < 0, invDir.z < 0};
    uint32_t todo[256];
    uint32_t todoOffset = 0, nodeNum = 0;
    while (true) {
        const BVHNode<R>& node = s.nodes[nodeNum];
        if (box_hit(node.bounds, ray, invDir, neg)) {
            if (node.nPrims > 0) {
                for (uint32_t i = 0; i < node.nPrims; ++i)
                    if (tri_hit(s.tris[s.ordered[node.offset + i]], ray, &t, &b1, &b2)) return true;
                if (todoOffset == 0) break;
                nodeNum = todo[--todoOffset];
            } else {
                if (todoOffset >= 255) { std::fprintf(stderr, "oracle: BVH stack overflow\n"); std::abort(); }
                if (neg[node.axis]) {
                    todo[todoOffset++] = nodeNum + 1;
                    nodeNum = node.offset;
                } else {
                    todo[todoOffset++] = node.offset;
                    nodeNum = nodeNum + 1;
                }
            }
        } else {
            if (todoOffset == 0) break;
            nodeNum = todo[--todoOffset];
        }
    }
    return false;
}

// ---------------------------------------------------------------- surface interaction
template <class R> struct Isect {
    V3<R> pos, n, absN, ng, outDir, sdpdu, sdpdv, pErr;
    R u = 0, v = 0;
    int32_t prim = -1;
};
// Triangle::QueryIntersectionInfo (shape/Triangle.cpp:6-69); mPos from
// Triangle::Intersect (Triangle.cpp:97).  dpdu/dpdv do not reach the
// integrator and are not formed.
// The intersection PathTracing::Li sees after a sphere hit: the default
// Intersection (zero position, normals, frame, outgoing direction and error
// bound) with the hit's uv; Scene::QueryIntersectionInfo adds nothing
// (Shape::QueryIntersectionInfo is empty for spheres, shape/Shape.h:54).
template <class R> Isect<R> sphere_isect(const Hit<R>& h) {
    Isect<R> it;
    it.pos = it.n = it.absN = it.ng = it.outDir = it.sdpdu = it.sdpdv = it.pErr = V3<R>(0, 0, 0);
    it.u = h.b1;
    it.v = h.b2;
    it.prim = h.prim;
    return it;
}

template <class R> Isect<R> query(const Scene<R>& s, const Ray<R>& ray, const Hit<R>& h) {
    if (h.prim >= (int32_t)s.tris.size()) return sphere_isect(h);
    const Tri<R>& t = s.tris[h.prim];
    Isect<R> it;
    R b1 = h.b1, b2 = h.b2;
    it.pos = ((1 - b1 - b2) * t.p0 + b1 * t.p1) + b2 * t.p2;
    R b0 = 1 - b1 - b2;
    it.u = (b0 * t.uv[0][0] + b1 * t.uv[1][0]) + b2 * t.uv[2][0];
    it.v = (b0 * t.uv[0][1] + b1 * t.uv[1][1]) + b2 * t.uv[2][1];
    R xs = (std::abs(b0 * t.p0.x) + std::abs(b1 * t.p1.x)) + std::abs(b2 * t.p2.x);
    R ys = (std::abs(b0 * t.p0.y) + std::abs(b1 * t.p1.y)) + std::abs(b2 * t.p2.y);
    R zs = (std::abs(b0 * t.p0.z) + std::abs(b1 * t.p1.z)) + std::abs(b2 * t.p2.z);
    V3<R> ns = (b0 * t.n0 + b1 * t.n1) + b2 * t.n2;
    ns = norm(ns);  // Normalize(): v / Length() lane-wise
    it.ng = (dot(t.fn, ns) < 0) ? -t.fn : t.fn;
    it.absN = dot(it.ng, ray.d) < 0 ? it.ng : it.ng * (R)-1;
    it.outDir = -ray.d;
    it.pErr = gamma_n<R>(7) * V3<R>(xs, ys, zs);
    R sg = (dot(ray.o - t.p0, ns) > 0) ? R(1.0f) : R(-1.0f);
    V3<R> dpdus = norm(t.e1);
    V3<R> dpdvs = norm(cross(sg * ns, dpdus));
    dpdus = norm(cross(dpdvs, ns));
    it.n = ns;
    it.sdpdu = dpdus;
    it.sdpdv = dpdvs;
    it.prim = h.prim;
    return it;
}
// Intersection::SpawnRay / SpawnTo (visual/Intersection.h:64-78)
template <class R> Ray<R> spawn_ray(const Isect<R>& it, const V3<R>& d) {
    V3<R> o = offset_ray_origin(it.pos, it.pErr, it.absN, d);
    return make_ray(o, d, K<R>::Inf, (R)0);
}
template <class R> struct LightPoint {
    V3<R> pos, n, fn, pErr;
};
// The Ray constructor moves the origin RayEps along d, which cancels the
// ShadowRayEps shortening: the segment ends on the offset light point, and the
// light's own triangle can occlude it depending on rounding.  g_shadow_short is
// a test-only switch that ends the segment ShadowRayEps before the target.
static int g_shadow_short = 0;
template <class R> Ray<R> spawn_to(const Isect<R>& a, const LightPoint<R>& b) {
    V3<R> origin = offset_ray_origin(a.pos, a.pErr, a.absN, b.pos - a.pos);
    V3<R> target = offset_ray_origin(b.pos, b.pErr, b.fn, origin - b.pos);
    V3<R> d = target - origin;
    R dis = length(d);
    d = norm(d);
    Ray<R> r = make_ray(origin, d, dis - K<R>::ShadowRayEps, (R)0);
    if (g_shadow_short) r.tmax = r.tmax - K<R>::RayEps;
    return r;
}

// ---------------------------------------------------------------- sampling warps
// visual/Sampling.cpp:5-41
template <class R> inline void concentric_disk(R u0, R u1, R* dx, R* dy) {
    R ox = (R)2.0 * u0 - 1, oy = (R)2.0 * u1 - 1;
    if (ox == 0 && oy == 0) { *dx = 0; *dy = 0; return; }
    R theta, r;
    if (std::abs(ox) > std::abs(oy)) {
        r = ox;
        theta = K<R>::PI_Over4 * (oy / ox);
    } else {
        r = oy;
        theta = K<R>::PI_Over2 - K<R>::PI_Over4 * (ox / oy);
    }
    *dx = r * rcos(theta);
    *dy = r * rsin(theta);
}
template <class R> inline V3<R> cosine_hemisphere(R u0, R u1) {
    R dx, dy;
    concentric_disk(u0, u1, &dx, &dy);
    R z = std::sqrt(rmax((R)0, 1 - dx * dx - dy * dy));
    return V3<R>(dx, dy, z);
}

// ---------------------------------------------------------------- textures
// ImageTexture::Triangle / Texel (texture/ImageTexture.h:50-69), CLAMP wrap.
template <class R> inline R texel(const Texture<R>& t, int u, int v, int ch) {
    u = u < 0 ? 0 : (u > t.w - 1 ? t.w - 1 : u);
    v = v < 0 ? 0 : (v > t.h - 1 ? t.h - 1 : v);
    return t.data[((size_t)v * t.w + u) * t.c + ch];
}
template <class R> inline R tex_bilinear(const Texture<R>& t, R u, R v, int ch) {
    R s = t.w * u - (R)0.5f, tt = t.h * v - (R)0.5f;
    int s0 = (int)std::floor(s), t0 = (int)std::floor(tt);
    R ds = s - s0, dt = tt - t0;
    R a = (1 - ds) * texel(t, s0, t0, ch) + ds * texel(t, s0 + 1, t0, ch);
    R b = (1 - ds) * texel(t, s0, t0 + 1, ch) + ds * texel(t, s0 + 1, t0 + 1, ch);
    return (1 - dt) * a + dt * b;
}

// ---------------------------------------------------------------- microfacet (GGX)
// GGXDistribution (visual/Microfacet.cpp:232-379), sample-visible-area mode.
template <class R> struct GGX {
    R ax, ay;
    R D(const V3<R>& m) const {
        if (m.z <= 0) return (R)0.0f;
        R cos2 = m.z * m.z;
        R be = ((m.x * m.x) / (ax * ax) + (m.y * m.y) / (ay * ay)) / cos2;
        R root = ((R)1 + be) * cos2;
        R result = (R)1 / (K<R>::PI * ax * ay * root * root);
        if (result * m.z < 1e-20f) result = 0;
        return result;
    }
    R project_roughness(const V3<R>& v) const {
        R invSin2 = 1 / rmax((R)0, (R)1 - v.z * v.z);
        if (ax == ay || invSin2 <= 0) return ax;
        R cosPhi2 = v.x * v.x * invSin2;
        R sinPhi2 = v.y * v.y * invSin2;
        return std::sqrt(cosPhi2 * ax * ax + sinPhi2 * ay * ay);
    }
    R smithG1(const V3<R>& v, const V3<R>& m) const {
        if (dot(v, m) * v.z <= 0) return (R)0.0f;
        R sinT = std::sqrt(rmax((R)0, (R)1 - v.z * v.z));
        R tanT = std::abs(sinT / v.z);
        if (tanT == 0.0f) return (R)1.0f;
        R alpha = project_roughness(v);
        R root = alpha * tanT;
        return (R)2.0f / ((R)1.0f + std::sqrt(1 + root * root));
    }
    R G(const V3<R>& wo, const V3<R>& wi, const V3<R>& wh) const { return smithG1(wo, wh) * smithG1(wi, wh); }
    R pdf(const V3<R>& wo, const V3<R>& wh) const {
        if (wo.z == 0) return (R)0.0f;
        return smithG1(wo, wh) * std::abs(dot(wo, wh)) * D(wh) / std::abs(wo.z);
    }
    // SampleGGXVNDF (Heitz 2018), Microfacet.cpp:318-339
    V3<R> sample_wh(const V3<R>& v_, R u0, R u1) const {
        V3<R> v = norm(V3<R>(ax * v_.x, ay * v_.y, v_.z));
        V3<R> T1 = (v.z < 0.9999) ? norm(cross(v, V3<R>(0, 0, 1))) : V3<R>(1, 0, 0);
        V3<R> T2 = cross(T1, v);
        R a = 1 / ((R)1.f + v.z);
        R r = std::sqrt(u0);
        R phi = (u1 < a) ? u1 / a * K<R>::PI : K<R>::PI + (u1 - a) / ((R)1.f - a) * K<R>::PI;
        R P1 = r * rcos(phi);
        R P2 = r * rsin(phi) * ((u1 < a) ? (R)1.0f : v.z);
        V3<R> dir = (P1 * T1 + P2 * T2) + std::sqrt(rmax((R)0, 1 - P1 * P1 - P2 * P2)) * v;
        return norm(V3<R>(ax * dir.x, ay * dir.y, rmax((R)0, dir.z)));
    }
};
template <class R> inline V3<R> reflect(const V3<R>& in, const V3<R>& n) {  // GeometryUtils.h:20-22
    return 2.f * dot(in, n) * n - in;
}
template <class R> inline bool same_hemi(const V3<R>& a, const V3<R>& b) { return a.z * b.z > 0; }
// FrDielectric (bsdf/BSDF.cpp:5-26)
template <class R> R fr_dielectric(R cosI, R etaI, R etaT) {
    cosI = clampr(cosI, (R)-1.f, (R)1.f);
    bool entering = cosI > 0.f;
    if (!entering) {
        std::swap(etaI, etaT);
        cosI = std::abs(cosI);
    }
    R sinI = std::sqrt(rmax((R)0, 1 - cosI * cosI));
    R sinT = etaI / etaT * sinI;
    if (sinT >= 1) return 1;
    R cosT = std::sqrt(rmax((R)0, 1 - sinT * sinT));
    R Rparl = ((etaT * cosI) - (etaI * cosT)) / ((etaT * cosI) + (etaI * cosT));
    R Rperp = ((etaI * cosI) - (etaT * cosT)) / ((etaI * cosI) + (etaT * cosT));
    return (Rparl * Rparl + Rperp * Rperp) / 2;
}
// FrConductor (bsdf/BSDF.cpp:29-51) with etai = (1,1,1)
template <class R> V3<R> fr_conductor(R cosI, const V3<R>& etai, const V3<R>& etat, const V3<R>& k) {
    cosI = clampr(cosI, (R)-1, (R)1);
    V3<R> eta = etat / etai;
    V3<R> etak = k / etai;
    R cos2 = cosI * cosI;
    R sin2 = (R)(1. - cos2);
    V3<R> eta2 = eta * eta;
    V3<R> etak2 = etak * etak;
    V3<R> t0 = eta2 - etak2 - V3<R>(sin2);
    V3<R> a2pb2 = vsqrt(t0 * t0 + (R)4 * eta2 * etak2);
    V3<R> t1 = a2pb2 + V3<R>(cos2);
    V3<R> a = vsqrt((R)0.5f * (a2pb2 + t0));
    V3<R> t2 = (R)2 * cosI * a;
    V3<R> Rs = (t1 - t2) / (t1 + t2);
    V3<R> t3 = cos2 * a2pb2 + V3<R>(sin2 * sin2);
    V3<R> t4 = t2 * sin2;
    V3<R> Rp = Rs * (t3 - t4) / (t3 + t4);
    return (R)0.5 * (Rp + Rs);
}
// Schlick Fresnel from F0: an extension the configurations require (the
// reference conductor takes eta/k only).  F = F0 + (1 - F0) (1 - |cos|)^5.
template <class R> V3<R> fr_schlick(R cosI, const V3<R>& f0) {
    R c = (R)1 - std::abs(cosI);
    R c2 = c * c;
    R c5 = c2 * c2 * c;
    return f0 + (V3<R>(1) - f0) * c5;
}
// TransmissionBSDF::Refract (bsdf/BSDF.h:292-304)
template <class R> bool refract(const V3<R>& wi, const V3<R>& n, R eta, V3<R>* wt) {
    R cosI = dot(n, wi);
    R sin2I = rmax(R(0), R(1 - cosI * cosI));
    R sin2T = eta * eta * sin2I;
    if (sin2T >= 1) return false;
    R cosT = std::sqrt(1 - sin2T);
    *wt = eta * (wi * (R)-1) + (eta * cosI - cosT) * n;
    return true;
}
// visual/Microfacet.h:49-54
template <class R> inline R roughness_to_alpha(R r) {
    r = rmax(r, (R)1e-3);
    R x = rlog(r);
    return (R)1.62142f + (R)0.819955f * x + (R)0.1734f * x * x + (R)0.0171201f * x * x * x + (R)0.000640711f * x * x * x * x;
}

// ---------------------------------------------------------------- BSDF
// BSDF with one lobe (material/*.h build exactly one BxDF), bsdf/BSDF.cpp:53-119.
enum { MAT_LAMBERT = 0, MAT_CONDUCTOR = 1, MAT_GLASS = 2, MAT_ROUGH_DIELECTRIC = 3 };
template <class R> struct BSDF {
    int type;
    V3<R> ng, X, Y, Z;
    V3<R> Rc, Tc;  // reflectance / transmittance
    GGX<R> dist;
    int fresnel_schlick;
    V3<R> eta, k, f0;
    R eta1, eta2;
    bool importance = false;  // TransportMode::Importance (light paths)
    bool delta() const { return type == MAT_GLASS; }
    V3<R> to_local(const V3<R>& v) const { return V3<R>(dot(v, X), dot(v, Y), dot(v, Z)); }
    V3<R> to_world(const V3<R>& v) const { return (X * v.x + Y * v.y) + Z * v.z; }
    V3<R> fresnel_cond(R c) const {
        return fresnel_schlick ? fr_schlick(c, f0) : fr_conductor(std::abs(c), V3<R>(1.0, 1.0, 1.0), eta, k);
    }
    // --- lobe functions in the local frame
    R lobe_pdf(const V3<R>& wo, const V3<R>& wi) const {
        switch (type) {
            case MAT_LAMBERT:  // DiffuseBSDF::Pdf (bsdf/BSDF.h:170-173)
                return same_hemi(wo, wi) ? std::abs(wi.z) * K<R>::INV_PI : (R)0.f;
            case MAT_CONDUCTOR: {  // MicrofacetReflectionBSDF::Pdf (bsdf/BSDF.h:323-333)
                V3<R> wh = norm(wo + wi);
                R dwh = 1.f / (4 * dot(wi, wh));
                wh = wh * sgn(wh.z);
                return std::abs(dist.pdf(wo, wh) * dwh);
            }
            case MAT_GLASS:
                return (R)0.0;
            default: {  // RoughDielectricBSDF::Pdf (bsdf/BSDF.h:513-543)
                bool refl = wo.z * wi.z > 0;
                V3<R> wh;
                R dwh;
                if (refl) {
                    wh = norm(wo + wi);
                    dwh = 1.f / (4 * dot(wi, wh));
                } else {
                    R e = wo.z > 0 ? (eta2 / eta1) : (eta1 / eta2);
                    wh = -(norm(wo + wi * e));
                    R sd = dot(wo, wh) + e * dot(wi, wh);
                    dwh = std::abs((e * e * dot(wi, wh)) / (sd * sd));
                }
                wh = wh * sgn(wh.z);
                R pdf = std::abs(dist.pdf(wo, wh) * dwh);
                R F = fr_dielectric(dot(wo, wh), eta1, eta2);
                pdf *= refl ? F : (1 - F);
                return pdf;
            }
        }
    }
    V3<R> lobe_eval(const V3<R>& wo, const V3<R>& wi) const {
        switch (type) {
            case MAT_LAMBERT:
                return Rc * K<R>::INV_PI;
            case MAT_CONDUCTOR: {  // MicrofacetReflectionBSDF::Evaluate (bsdf/BSDF.h:356-373)
                if (!same_hemi(wo, wi)) return V3<R>();
                R cO = std::abs(wo.z), cI = std::abs(wi.z);
                V3<R> wh = wi + wo;
                if (cI == 0 || cO == 0) return V3<R>(0, 0, 0);
                if (wh.x == 0 && wh.y == 0 && wh.z == 0) return V3<R>(0, 0, 0);
                wh = norm(wh);
                wh = sgn(wh.z) * wh;
                V3<R> F = fresnel_cond(dot(wo, wh));
                V3<R> D(dist.D(wh));
                V3<R> G(dist.G(wo, wi, wh));
                return Rc * F * D * G / (4 * cI * cO);
            }
            case MAT_GLASS:
                return V3<R>();
            default: {  // RoughDielectricBSDF::Evaluate (bsdf/BSDF.h:469-511)
                bool refl = wo.z * wi.z > 0;
                if (refl) {
                    R cO = std::abs(wo.z), cI = std::abs(wi.z);
                    V3<R> wh = wi + wo;
                    if (cI == 0 || cO == 0) return V3<R>(0, 0, 0);
                    if (wh.x == 0 && wh.y == 0 && wh.z == 0) return V3<R>(0, 0, 0);
                    wh = norm(wh);
                    wh = sgn(wh.z) * wh;
                    R F = fr_dielectric(dot(wo, wh), eta1, eta2);
                    V3<R> D(dist.D(wh));
                    V3<R> G(dist.G(wo, wi, wh));
                    return Rc * V3<R>(F) * D * G / (4 * cI * cO);
                }
                R cO = wo.z, cI = wi.z;
                if (cI == 0 || cO == 0) return V3<R>();
                R e = wo.z > 0 ? (eta2 / eta1) : (eta1 / eta2);
                V3<R> wh = (-norm(wo + wi * e));
                wh = sgn(wh.z) * wh;
                R F = fr_dielectric(dot(wo, wh), eta1, eta2);
                R D = dist.D(wh);
                R G = dist.G(wo, wi, wh);
                R sd = dot(wo, wh) + e * dot(wi, wh);
                R factor = importance ? (R)1 : 1 / e;  // (bsdf/BSDF.h:505)
                R value = (1 - F) * D * G * e * e * std::abs(dot(wi, wh)) * std::abs(dot(wo, wh)) /
                          std::abs(cI * cO * sd * sd);
                return Tc * std::abs(value * factor * factor);
            }
        }
    }
    V3<R> lobe_sample(const V3<R>& wo, V3<R>* wi, R* pdf, const V3<R>& rnd) const {
        switch (type) {
            case MAT_LAMBERT: {  // DiffuseBSDF::Sample (bsdf/BSDF.h:175-181)
                *wi = cosine_hemisphere(rnd.x, rnd.y);
                if (wo.z < 0) wi->z *= -1;
                *pdf = lobe_pdf(wo, *wi);
                return lobe_eval(wo, *wi);
            }
            case MAT_CONDUCTOR: {  // MicrofacetReflectionBSDF::Sample (bsdf/BSDF.h:335-354)
                if (wo.z == 0) return V3<R>(0.);
                V3<R> wh = dist.sample_wh(sgn(wo.z) * wo, rnd.x, rnd.y);
                R mpdf = dist.pdf(wo, wh);
                if (mpdf == 0) return V3<R>();
                *wi = reflect(wo, wh);
                if (!same_hemi(wo, *wi)) return V3<R>(0, 0, 0);
                wh = sgn(wh.z) * wh;
                *pdf = std::abs(dist.pdf(wo, wh) / (4 * dot(*wi, wh)));
                return lobe_eval(wo, *wi);
            }
            case MAT_GLASS: {  // TransmissionBSDF::Sample (bsdf/BSDF.h:238-270)
                R Re = fr_dielectric(wo.z, eta1, eta2);
                bool entering = wo.z > 0;
                R etaI = entering ? eta1 : eta2;
                R etaT = entering ? eta2 : eta1;
                V3<R> nF = wo.z > 0 ? V3<R>(0, 0, 1) : V3<R>(0, 0, -1);
                R P = Re;
                if (rnd.y < P) {
                    *wi = V3<R>(-wo.x, -wo.y, wo.z);
                    *pdf = P;
                    R c = std::abs(wi->z);
                    return Rc * Re / c;
                }
                if (!refract(wo, nF, etaI / etaT, wi)) return V3<R>();
                *pdf = 1 - P;
                R c = std::abs(wi->z);
                R e = etaI / etaT;
                if (importance) return Tc * ((R)1.f - Re) / c;  // TransportMode::Importance (bsdf/BSDF.h:266-271)
                return e * e * Tc * ((R)1.f - Re) / c;
            }
            default: {  // RoughDielectricBSDF::Sample (bsdf/BSDF.h:545-615)
                V3<R> wh = dist.sample_wh(sgn(wo.z) * wo, rnd.x, rnd.y);
                R mpdf = dist.pdf(wo, wh);
                if (mpdf == 0) return V3<R>();
                R F = fr_dielectric(dot(wo, wh), eta1, eta2);
                R P = F;
                *pdf = 0;
                if (rnd.z < P) {
                    if (wo.z == 0) return V3<R>();
                    *wi = reflect(wo, wh);
                    if (!same_hemi(wo, *wi)) return V3<R>(0, 0, 0);
                    R dwh = 1.f / (4 * dot(*wi, wh));
                    *pdf = P * std::abs(dist.pdf(wo, wh) * dwh);
                    R cO = std::abs(wo.z), cI = std::abs(wi->z);
                    V3<R> D(dist.D(wh));
                    V3<R> G(dist.G(wo, *wi, wh));
                    return Rc * V3<R>(F) * D * G / (4 * cI * cO);
                }
                if (wo.z == 0) return V3<R>();
                R etaR = wo.z > 0 ? (eta1 / eta2) : (eta2 / eta1);
                V3<R> ff = (dot(wh, wo) < 0.f) ? -wh : wh;
                if (!refract(wo, ff, etaR, wi)) {
                    *pdf = 0;
                    return V3<R>();
                }
                if (wo.z * wi->z >= 0) return V3<R>();
                R e = wo.z > 0 ? (eta2 / eta1) : (eta1 / eta2);
                R sd = dot(wo, wh) + e * dot(*wi, wh);
                R dwh = std::abs((e * e * dot(*wi, wh)) / (sd * sd));
                *pdf = std::abs((1 - P) * dist.pdf(wo, wh) * dwh);
                R cO = wo.z, cI = wi->z;
                if (cI == 0 || cO == 0) return V3<R>();
                R D = dist.D(wh);
                R G = dist.G(wo, *wi, wh);
                R factor = importance ? (R)1 : 1 / e;  // (bsdf/BSDF.h:608)
                R value = (1 - F) * D * G * e * e * std::abs(dot(*wi, wh)) * std::abs(dot(wo, wh)) /
                          std::abs(cI * cO * sd * sd);
                return Tc * std::abs(value * factor * factor);
            }
        }
    }
    // --- BSDF wrappers (bsdf/BSDF.cpp:58-119)
    R Pdf(const V3<R>& wo, const V3<R>& wi) const {
        V3<R> woL = to_local(wo), wiL = to_local(wi);
        R pdf = 0.f;
        pdf += lobe_pdf(woL, wiL);
        return pdf / 1;
    }
    V3<R> Sample(const V3<R>& wo, V3<R>* wi, R* pdf, const V3<R>& rnd) const {
        V3<R> woL = to_local(wo), wiL;
        *pdf = 0;
        V3<R> f = lobe_sample(woL, &wiL, pdf, rnd);
        *wi = to_world(wiL);
        *pdf /= 1;
        return f;
    }
    V3<R> Evaluate(const V3<R>& wo, const V3<R>& wi) const {
        V3<R> f(0, 0, 0);
        V3<R> woL = to_local(wo), wiL = to_local(wi);
        bool refl = dot(wi, ng) * dot(wo, ng) > 0;
        bool has_refl = true;                         // every lobe here has BSDF_REFLECTION
        bool has_trans = type == MAT_GLASS || type == MAT_ROUGH_DIELECTRIC;
        if ((refl && has_refl) || (!refl && has_trans)) f = f + lobe_eval(woL, wiL);
        return f;
    }
};
// Material::ComputeScatteringFunction (material/Diffuse.h:27-31, Roughness.h:36-46,
// Glass.h:30-34, RoughDielectric.h:21-41)
template <class R> BSDF<R> make_bsdf_tex(const Scene<R>& s, const Isect<R>& it, const std::vector<Texture<R>>& texs);
template <class R> BSDF<R> make_bsdf(const Scene<R>& s, const Isect<R>& it) { return make_bsdf_tex(s, it, s.tex); }
// The BSDF whose values the estimator takes: the sampled one, or in detached
// mode the same material with the evaluation textures.
template <class R> BSDF<R> eval_bsdf(const Scene<R>& s, const Isect<R>& it, const BSDF<R>& sampled) {
    return s.detached ? make_bsdf_tex(s, it, s.tex_eval) : sampled;
}
template <class R> BSDF<R> make_bsdf_tex(const Scene<R>& s, const Isect<R>& it, const std::vector<Texture<R>>& texs) {
    const Material& m = s.mats[prim_mat(s, it.prim)];
    BSDF<R> b;
    b.type = (int)m.type;
    b.ng = it.ng;
    b.X = it.sdpdu;
    b.Y = it.sdpdv;
    b.Z = it.n;
    b.Rc = V3<R>(m.R[0], m.R[1], m.R[2]);
    b.Tc = V3<R>(m.T[0], m.T[1], m.T[2]);
    if (m.tex_albedo >= 0) {
        const Texture<R>& t = texs[m.tex_albedo];
        // channels past the texture's last one read the last one
        b.Rc = V3<R>(tex_bilinear(t, it.u, it.v, 0), tex_bilinear(t, it.u, it.v, std::min(1, t.c - 1)),
                     tex_bilinear(t, it.u, it.v, std::min(2, t.c - 1)));
    }
    R ax = m.ax, ay = m.ay;
    if (m.tex_rough >= 0) {
        ax = ay = tex_bilinear(texs[m.tex_rough], it.u, it.v, 0);
    }
    if (!(m.flags & 1u)) {
        ax = roughness_to_alpha(ax);
        ay = roughness_to_alpha(ay);
    }
    b.dist.ax = ax;
    b.dist.ay = ay;
    b.fresnel_schlick = (m.flags & 2u) ? 1 : 0;
    b.eta = V3<R>(m.eta[0], m.eta[1], m.eta[2]);
    b.k = V3<R>(m.k[0], m.k[1], m.k[2]);
    b.f0 = V3<R>(m.f0[0], m.f0[1], m.f0[2]);
    b.eta1 = m.eta1;
    b.eta2 = m.eta2;
    return b;
}

// ---------------------------------------------------------------- lights
// Scene::ComputeLightPowerDistribution + Distribution1D (visual/Scene.h:121-128,
// visual/Sampling.h:16-97), AreaLight::Power (light/AreaLight.h:45-48).
template <class R> V3<R> env_radiance(const Env<R>& e, R u, R v);
template <class R> void build_light_distribution(Scene<R>& s) {
    int n = (int)s.light_tri.size();
    s.light_func.resize(n);
    for (int i = 0; i < n; ++i) {
        V3<R> P;
        if (s.light_tri[i] < 0)  // EnvironmentLight::Power (light/EnvLight.cpp:60-63)
            P = env_radiance(s.env, (R)0.5f, (R)0.5f) * K<R>::PI * s.env.radius * s.env.radius;
        else
            P = s.light_le[i] * (s.light_tri[i] < (int32_t)s.tris.size() ? s.tris[s.light_tri[i]].area
                                                                           : s.sph[s.light_tri[i] - s.tris.size()].area) * K<R>::PI;
        s.light_func[i] = lumY(P);
    }
    s.light_cdf.assign(n + 1, 0);
    s.light_cdf[0] = 0.;
    for (int i = 1; i < n + 1; ++i) s.light_cdf[i] = s.light_cdf[i - 1] + s.light_func[i - 1] / n;
    s.light_func_int = s.light_cdf[n];
    if (s.light_func_int == 0.f) {
        for (int i = 1; i < n + 1; ++i) s.light_cdf[i] = R(i) / R(n);
    } else {
        for (int i = 1; i < n + 1; ++i) s.light_cdf[i] /= s.light_func_int;
    }
}
template <class R> int sample_one_light(const Scene<R>& s, R u, R* pdf) {
    int n = (int)s.light_func.size();
    auto it = std::upper_bound(s.light_cdf.begin(), s.light_cdf.begin() + n + 1, u);
    int off = std::max(0, int(std::distance(s.light_cdf.begin(), it)) - 1);
    *pdf = s.light_func[off] / (s.light_func_int * n);
    return off;
}
// Triangle::Sample (shape/Triangle.cpp:141-158) + Shape::Sample (shape/Shape.cpp:6-19)
template <class R> LightPoint<R> sample_light_point(const Tri<R>& t, const Isect<R>& it, R u0, R u1, R* pdf) {
    R su0 = std::sqrt(u0);
    R b0 = 1 - su0, b1 = u1 * su0;
    LightPoint<R> lp;
    R b2 = 1 - b0 - b1;
    lp.pos = (b0 * t.p0 + b1 * t.p1) + b2 * t.p2;
    lp.fn = t.fn;
    lp.n = (b0 * t.n0 + b1 * t.n1) + b2 * t.n2;
    V3<R> pAbs = (vabs(b0 * t.p0) + vabs(b1 * t.p1)) + vabs(b2 * t.p2);
    lp.pErr = gamma_n<R>(6) * pAbs;
    *pdf = 1 / t.area;
    V3<R> wi = it.pos - lp.pos;
    if (length(wi) == 0) {
        *pdf = 0;
    } else {
        wi = norm(wi);
        V3<R> dd = lp.pos - it.pos;
        *pdf = *pdf * (dot(dd, dd) / std::abs(dot(lp.n, -wi)));
        if (std::isinf(*pdf)) *pdf = 0;
    }
    return lp;
}
// Shape::Pdf (shape/Shape.h:34-47): intersect only the light's triangle.
// Sphere::Sample(real*, const Vec2&) (shape/Sphere.h:163-193): a uniform
// point with its ObjectToWorld error bound; the absolute normal is copied
// from the new record's own zero normal.
template <class R> LightPoint<R> sphere_sample_area(const typename Scene<R>::Sph& sp, R u0, R u1, R* pdf) {
    *pdf = 1.f / (4.f * K<R>::PI * sp.r * sp.r);
    // UniformSampleSphere (visual/Sampling.cpp:25-30)
    R z = 1 - 2 * u0;
    R rr = std::sqrt(rmax((R)0, (R)1 - z * z));
    R phi = 2 * K<R>::PI * u1;
    V3<R> pHit = V3<R>(rr * rcos(phi), rr * rsin(phi), z) * sp.r + sp.c;
    V3<R> scaledDir = (pHit - sp.c) * sp.r / length(pHit - sp.c);
    V3<R> hit = sp.c + scaledDir;
    V3<R> n = sp.inside ? norm(sp.c - hit) : norm(hit - sp.c);  // GetNorm (Sphere.h:338-342)
    V3<R> pError = vabs(sp.c) * gamma_n<R>(1) + vabs(scaledDir) * gamma_n<R>(6);
    LightPoint<R> lp;
    // ObjectToWorld(hit, pError, &err) with the identity (math/Transform.h:137-164)
    const R g3 = gamma_n<R>(3);
    const V3<R> h = hit;
    lp.pErr = V3<R>((g3 + (R)1) * ((pError.x + 0 * pError.y) + 0 * pError.z) + g3 * (((std::abs(h.x) + 0) + 0) + 0),
                    (g3 + (R)1) * ((0 * pError.x + pError.y) + 0 * pError.z) + g3 * (((0 + std::abs(h.y)) + 0) + 0),
                    (g3 + (R)1) * ((0 * pError.x + 0 * pError.y) + pError.z) + g3 * (((0 + 0) + std::abs(h.z)) + 0));
    lp.pos = xf_identity(hit);
    lp.n = norm(xf_identity(n));  // TransformNormal(n).Norm()
    lp.fn = V3<R>(0, 0, 0);       // mAbsNormal = the new record's zero normal
    return lp;
}
// Sphere::Sample(const Intersection&, real*, const Vec2&) (shape/Sphere.h:243-291):
// inside the sphere, an area sample whose pdf is multiplied by the distance
// (not its square) over the cosine; outside, a direction in the cone the
// sphere subtends, the point 1.5e-3 r beyond the surface, pdf 1 / cone solid angle.
template <class R> LightPoint<R> sphere_sample_li(const typename Scene<R>::Sph& sp, const Isect<R>& it, R u0, R u1, R* pdf) {
    V3<R> pCenter = xf_identity(sp.c);
    if (dot(pCenter - it.pos, pCenter - it.pos) <= sp.r * sp.r) {
        LightPoint<R> lp = sphere_sample_area<R>(sp, u0, u1, pdf);
        V3<R> wi = lp.pos - it.pos;
        if (dot(wi, wi) == 0) {
            *pdf = 0;
        } else {
            R sd = length(wi);
            wi = norm(wi);
            *pdf = *pdf * (sd / std::abs(dot(norm(lp.pos - pCenter), wi * (R)-1)));
        }
        if (std::isinf(*pdf)) *pdf = 0.f;
        return lp;
    }
    V3<R> localZ = pCenter - it.pos;
    R dis = length(localZ);
    localZ = norm(localZ);
    V3<R> localX, localY;
    coordinate_system(localZ, &localX, &localY);
    R sin2ThetaMax = sp.r * sp.r / (dis * dis);
    R cosThetaMax = std::sqrt(rmax((R)1 - sin2ThetaMax, (R)0));
    R cosTheta = ((R)1 - u0) + u0 * cosThetaMax;
    R sinTheta = std::sqrt(rmax((R)1 - cosTheta * cosTheta, (R)0));
    R phi = 2 * K<R>::PI * u1;
    R sdist = dis * cosTheta - std::sqrt(rmax(sp.r * sp.r - dis * dis * sinTheta * sinTheta, (R)0));
    R cosAlpha = (dis * dis + sp.r * sp.r - sdist * sdist) / (2 * dis * sp.r);
    R sinAlpha = std::sqrt(rmax((R)1 - cosAlpha * cosAlpha, (R)0));
    R cphi = rcos(phi), sphi = rsin(phi);
    V3<R> nWorld = ((R)-1 * sinAlpha * cphi) * localX - (sinAlpha * sphi) * localY - cosAlpha * localZ;
    nWorld = norm(nWorld);
    // 1.5e-3 is a double literal: r * 1.5e-3 is formed in double and rounded
    const R off = (R)((double)sp.r * 1.5e-3);
    V3<R> pWorld = pCenter + (sp.r * nWorld) / length(nWorld) + off * norm(nWorld);
    *pdf = 1 / (2 * K<R>::PI * (1 - cosThetaMax));
    LightPoint<R> lp;
    lp.pos = pWorld;
    lp.n = nWorld;
    lp.fn = nWorld;
    lp.pErr = gamma_n<R>(5) * vabs(pWorld);
    return lp;
}
// Shape::Pdf (shape/Shape.h:34-47) for a sphere light: the sphere's own hit
// record is the default one (zero position and normal), so the pdf is
// |d|^2 / 0: infinite, which returns 0; NaN when the shading point is at the
// origin (d = 0).
template <class R> R sphere_light_pdf(const typename Scene<R>::Sph& sp, const Isect<R>& it, const V3<R>& wi) {
    Ray<R> ray = spawn_ray(it, wi);
    R t, u, v;
    if (!sphere_hit<R>(sp, ray, &t, &u, &v)) return 0;
    V3<R> d = V3<R>(0, 0, 0) - it.pos;
    R pdf = dot(d, d) / (std::abs(dot(V3<R>(0, 0, 0), wi * (R)-1)) * sp.area);
    if (std::isinf(pdf)) return 0;
    return pdf;
}

template <class R> R light_pdf(const Scene<R>& s, int tri, const Isect<R>& it, const V3<R>& wi) {
    if (tri >= (int)s.tris.size()) return sphere_light_pdf<R>(s.sph[tri - s.tris.size()], it, wi);
    Ray<R> ray = spawn_ray(it, wi);
    const Tri<R>& t = s.tris[tri];
    Hit<R> h;
    R th, b1, b2;
    if (!tri_hit(t, ray, &th, &b1, &b2)) return 0;
    h.t = th; h.b1 = b1; h.b2 = b2; h.prim = tri;
    Isect<R> lp = query(s, ray, h);
    V3<R> d = lp.pos - it.pos;
    R pdf = dot(d, d) / (std::abs(dot(lp.n, wi * (R)-1)) * t.area);
    if (std::isinf(pdf)) return 0;
    return pdf;
}
// ---------------------------------------------------------------- environment light
// ImageTexture::Triangle / Texel, CLAMP (texture/ImageTexture.h:50-69).
template <class R> inline V3<R> env_texel(const Env<R>& e, int u, int v) {
    u = u < 0 ? 0 : (u > e.w - 1 ? e.w - 1 : u);
    v = v < 0 ? 0 : (v > e.h - 1 ? e.h - 1 : v);
    return e.tex[(size_t)v * e.w + u];
}
// EnvironmentMap::LookupRadiance = mpImage->Sample(Vec2(u, 1 - v)) * mScale (EnvMap.h:70-72).
template <class R> V3<R> env_radiance(const Env<R>& e, R u, R v) {
    R tv = 1 - v;
    R sx = e.w * u - (R)0.5f, ty = e.h * tv - (R)0.5f;
    int s0 = (int)std::floor(sx), t0 = (int)std::floor(ty);
    R ds = sx - s0, dt = ty - t0;
    V3<R> a = (1 - ds) * env_texel(e, s0, t0) + ds * env_texel(e, s0 + 1, t0);
    V3<R> b = (1 - ds) * env_texel(e, s0, t0 + 1) + ds * env_texel(e, s0 + 1, t0 + 1);
    return ((1 - dt) * a + dt * b) * e.scale;
}
// DirToLatLong (EnvMap.h:106-121): u = phi / 2pi about +y from +x towards +z,
// v = theta / pi from +y.
template <class R> void env_dir_to_uv(const V3<R>& d, R* u, R* v) {
    R phi = (d.x != 0 || d.z != 0) ? ratan2(d.z, d.x) : (R)0;
    R theta = racos(d.y);
    phi = (phi < 0) ? (phi + 2 * K<R>::PI) : phi;
    *u = phi * K<R>::INV_PI * (R)0.5;
    *v = clampr(theta * K<R>::INV_PI, (R)0, (R)1);
}
// LatLongToDir (EnvMap.h:96-104).
template <class R> V3<R> env_uv_to_dir(R u, R v) {
    R phi = u * 2 * K<R>::PI;
    R theta = v * K<R>::PI;
    R st = rsin(theta);
    return V3<R>(st * rcos(phi), rcos(theta), st * rsin(phi));
}
// SinTheta (EnvMap.h:124-132): sin of the centre of the pixel row holding v.
template <class R> R env_sin_theta(R v, R height) {
    if (v < 1) return rsin(K<R>::PI * (int)((v * height) + (R)0.5f) / (R)(height));
    return rsin(K<R>::PI * (R)((height - 1) + (R)0.5f) / (R)(height));
}
// Distribution2D::Pdf (visual/Sampling.h:130-140).
template <class R> R env_dist_pdf(const Env<R>& e, R u, R v) {
    int iu = clampr((int)(u * e.cond[0].count), 0, e.cond[0].count - 1);
    int iv = clampr((int)(v * e.marg.count), 0, e.marg.count - 1);
    if (e.cond[iv].funcInt * e.marg.funcInt == 0.f) return 0.f;
    return (e.cond[iv].func[iu] * e.marg.func[iv]) / (e.cond[iv].funcInt * e.marg.funcInt);
}
// EnvironmentMap::Sample (EnvMap.h:18-31) with Distribution2D::SampleContinuous
// (visual/Sampling.h:121-128).
template <class R> V3<R> env_sample(const Env<R>& e, R u0, R u1, V3<R>* dir, R* pdfW) {
    R uv[2], pdfs[2];
    int v;
    uv[1] = e.marg.sample_continuous(u1, &pdfs[1], &v);
    uv[0] = e.cond[v].sample_continuous(u0, &pdfs[0], nullptr);
    R pdf = pdfs[0] * pdfs[1];
    R st = env_sin_theta(uv[1], (R)e.h);
    *pdfW = (st == 0) ? (R)0 : pdf / (2 * K<R>::PI * K<R>::PI * st);
    *dir = env_uv_to_dir(uv[0], uv[1]);
    return env_radiance(e, uv[0], uv[1]);
}
// EnvironmentMap::Lookup / PdfDir (EnvMap.h:33-68).
template <class R> V3<R> env_lookup(const Env<R>& e, const V3<R>& dir) {
    R u, v;
    env_dir_to_uv(norm(dir), &u, &v);
    return env_radiance(e, u, v);
}
template <class R> R env_pdf_dir(const Env<R>& e, const V3<R>& dir) {
    R u, v;
    env_dir_to_uv(norm(dir), &u, &v);
    R st = env_sin_theta(v, (R)e.h);
    if (st == 0) return 0;
    return env_dist_pdf(e, u, v) / (2 * K<R>::PI * K<R>::PI * st);
}
// EnvironmentLight::Emission(dir) = Lookup(dir.Norm()) (EnvLight.cpp:12-20).
template <class R> V3<R> env_emission(const Env<R>& e, const V3<R>& dir) { return env_lookup(e, norm(dir)); }
// EnvironmentMap(filename) after ImageIO::LoadTexture + ConvertImageToPdf
// (EnvMap.h:10-16, 82-94); rgb is the image top row first, 3 floats a texel.
template <class R> void build_env(Scene<R>& s, const float* rgb, int w, int h, float scale) {
    Env<R>& e = s.env;
    e.w = w;
    e.h = h;
    e.scale = scale;
    e.tex.resize((size_t)w * h);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const float* p = rgb + ((size_t)(h - 1 - y) * w + x) * 3;
            e.tex[(size_t)y * w + x] = V3<R>(p[0], p[1], p[2]);
        }
    std::vector<R> data((size_t)w * h);
    for (int r = 0; r < h; ++r) {
        R v = (R)(r + (R)0.5f) / (R)(h);
        R sinTheta = rsin(K<R>::PI * v);
        for (int c = 0; c < w; ++c) data[c + (size_t)r * w] = sinTheta * lumY(e.tex[(size_t)(h - r - 1) * w + c]);
    }
    e.cond.assign(h, Dist1D<R>());
    std::vector<R> marginal(h);
    for (int v = 0; v < h; ++v) {
        e.cond[v].build(&data[(size_t)v * w], w);
        marginal[v] = e.cond[v].funcInt;
    }
    e.marg.build(marginal.data(), h);
    // Scene::GetBoundingSphere from the BVH's world bound (visual/Scene.cpp:9-11,
    // math/AABB.h:69-72); the bound is the union of the triangles' boxes.
    V3<R> mn(K<R>::Inf, K<R>::Inf, K<R>::Inf), mx(-K<R>::Inf, -K<R>::Inf, -K<R>::Inf);
    for (const Tri<R>& t : s.tris) {
        V3<R> a(rmin(t.p0.x, t.p1.x), rmin(t.p0.y, t.p1.y), rmin(t.p0.z, t.p1.z));
        V3<R> b(rmax(t.p0.x, t.p1.x), rmax(t.p0.y, t.p1.y), rmax(t.p0.z, t.p1.z));
        a = V3<R>(rmin(a.x, t.p2.x), rmin(a.y, t.p2.y), rmin(a.z, t.p2.z));
        b = V3<R>(rmax(b.x, t.p2.x), rmax(b.y, t.p2.y), rmax(b.z, t.p2.z));
        mn = V3<R>(rmin(mn.x, a.x), rmin(mn.y, a.y), rmin(mn.z, a.z));
        mx = V3<R>(rmax(mx.x, b.x), rmax(mx.y, b.y), rmax(mx.z, b.z));
    }
    e.center = (mn + mx) * (R)0.5;
    e.radius = length(mx - mn) * (R)0.5;
}

template <class R> inline R power_heuristic(R f, R g) {  // integrator/Integrator.h:11-15
    f = 1 * f;
    g = 1 * g;
    return (f * f) / (f * f + g * g);
}

// Integrator::DirectIllumination (integrator/Integrator.cpp:9-107), surfaces only.
// The value a BSDF sample takes in detached mode: the evaluation BSDF's lobe
// at the sampled pair of directions, zero where the sample itself failed.
template <class R> V3<R> detached_value(const BSDF<R>& be, const V3<R>& wo, const V3<R>& wi, const V3<R>& f_sampled) {
    if (f_sampled == V3<R>()) return f_sampled;
    return be.lobe_eval(be.to_local(wo), be.to_local(wi));
}

template <class R> V3<R> direct_illumination(const Scene<R>& s, const Isect<R>& it, const BSDF<R>& bsdf, R uLight,
                                             R u0, R u1, const V3<R>& v, bool use_bvh) {
    R lightSamplingPdf;
    R weight1 = 1, weight2 = 1;
    V3<R> L1, L2;
    const BSDF<R> be = eval_bsdf(s, it, bsdf);
    int light = sample_one_light(s, uLight, &lightSamplingPdf);
    int ltri = s.light_tri[light];  // < 0: the environment light
    V3<R> Le = s.detached ? s.light_le_eval[light] : s.light_le[light];
    {
        R pdf;
        LightPoint<R> lp;
        V3<R> wi, Li;
        if (ltri < 0) {
            // EnvironmentLight::Sample_Li (light/EnvLight.cpp:41-52): the light
            // point sits 2 r along wi, with zero position error
            V3<R> sd;
            Li = env_sample(s.env, u0, u1, &sd, &pdf);
            wi = norm(sd);
            lp.pos = it.pos + wi * (2 * s.env.radius);
            lp.n = lp.fn = wi;
            lp.pErr = V3<R>(0, 0, 0);
        } else {
            lp = ltri < (int)s.tris.size() ? sample_light_point(s.tris[ltri], it, u0, u1, &pdf)
                                           : sphere_sample_li<R>(s.sph[ltri - s.tris.size()], it, u0, u1, &pdf);
            wi = norm(lp.pos - it.pos);
            Li = dot(lp.n, -wi) > 0 ? Le : V3<R>(0, 0, 0);
        }
        if (pdf > 0 && Li != V3<R>(0)) {
            R scatteringPdf = bsdf.Pdf(it.outDir, wi);
            V3<R> f = be.Evaluate(it.outDir, wi);
            R cosTheta = std::abs(dot(it.n, wi));
            if (f != V3<R>()) {
                V3<R> Tr(1.f);
                Ray<R> sr = spawn_to(it, lp);
                if (intersect_any(s, sr, use_bvh)) Li = V3<R>();
                weight1 = power_heuristic(pdf, scatteringPdf);
                L1 = Li * Tr * f * cosTheta / pdf;
            }
        }
    }
    {
        V3<R> wi;
        R pdf;
        V3<R> f = bsdf.Sample(it.outDir, &wi, &pdf, v);
        if (s.detached) f = detached_value(be, it.outDir, wi, f);
        R cosTheta = std::abs(dot(it.n, wi));
        if (pdf != 0 && f != V3<R>()) {
            R lpdf = ltri < 0 ? env_pdf_dir(s.env, wi) : light_pdf(s, ltri, it, wi);
            if (lpdf == 0) {
                weight2 = 0;
            } else {
                weight2 = power_heuristic(pdf, lpdf);
                Ray<R> ray = spawn_ray(it, wi);
                Hit<R> h;
                if (intersect(s, ray, &h, use_bvh)) {
                    if (prim_light(s, h.prim) == light) {
                        Isect<R> li = query(s, ray, h);
                        V3<R> Lemit = dot(li.n, -wi) > 0 ? Le : V3<R>(0, 0, 0);
                        V3<R> Tr(1.f);
                        L2 = Lemit * f * Tr * cosTheta / pdf;
                    }
                } else if (ltri < 0) {  // Integrator.cpp:97-99
                    V3<R> Tr(1.f);
                    L2 = env_emission(s.env, wi) * f * Tr * cosTheta / pdf;
                }
            }
        }
    }
    return (L1 * weight1 + L2 * weight2) / lightSamplingPdf;
}

// PathTracing::Li (integrator/PT.cpp:10-67).  The dimensions of the
// DirectIllumination(rand(), Vec2(rand(), rand()), Vec3(rand(), rand(), rand()))
// call (PT.cpp:38) and of Vec3(rand(), rand(), rand()) (PT.cpp:47) are drawn in
// the order written, left to right (the canonical order of this build; C++
// leaves argument evaluation order unspecified).
// Test-only path record: one row per bounce (see or_li_path).
constexpr int kRecWords = 16;
static thread_local double* t_rec = nullptr;
static thread_local int t_rec_max = 0, t_rec_n = 0;

template <class R> V3<R> Li(const Scene<R>& s, Ray<R> ray, Seq& rnd, int maxDepth, int rrDepth, bool use_bvh) {
    bool delta = false;
    V3<R> T(1, 1, 1);
    V3<R> L;
    for (int i = 0; i < maxDepth; ++i) {
        Hit<R> h;
        const bool hit = intersect(s, ray, &h, use_bvh);
        if (t_rec && t_rec_n < t_rec_max) {
            double* r = t_rec + (size_t)kRecWords * t_rec_n++;
            r[0] = ray.o.x; r[1] = ray.o.y; r[2] = ray.o.z;
            r[3] = ray.d.x; r[4] = ray.d.y; r[5] = ray.d.z;
            r[6] = ray.tmin; r[7] = ray.tmax;
            r[8] = hit ? h.prim : -1; r[9] = hit ? h.t : 0;
            r[10] = rnd.cursor;
            r[11] = T.x; r[12] = T.y; r[13] = T.z;
            r[14] = hit ? prim_mat(s, h.prim) : -1;
            r[15] = delta ? 1 : 0;
        }
        if (!hit) {
            // PT.cpp:18-21: a camera ray, or one after a delta bounce, that
            // leaves the scene returns T * Le(ray) in place of L
            if ((i == 0 || delta) && s.env.w > 0) return T * env_emission(s.env, ray.d);
            break;
        }
        Isect<R> it = query(s, ray, h);
        BSDF<R> bsdf = make_bsdf(s, it);
        int light = prim_light(s, it.prim);
        if ((i == 0 || delta) && light >= 0) {
            V3<R> Le = s.detached ? s.light_le_eval[light] : s.light_le[light];
            V3<R> e = dot(it.n, it.outDir) > 0 ? Le : V3<R>(0, 0, 0);
            L = L + T * e;
        } else {
            R uL = rnd.next();
            R u0 = rnd.next(), u1 = rnd.next();
            R v0 = rnd.next(), v1 = rnd.next(), v2 = rnd.next();
            L = L + T * direct_illumination(s, it, bsdf, uL, u0, u1, V3<R>(v0, v1, v2), use_bvh);
        }
        R r0 = rnd.next(), r1 = rnd.next(), r2 = rnd.next();
        V3<R> wi;
        R pdf;
        V3<R> f = bsdf.Sample(-ray.d, &wi, &pdf, V3<R>(r0, r1, r2));
        if (f == V3<R>() || pdf == 0) break;
        if (s.detached) f = detached_value(eval_bsdf(s, it, bsdf), -ray.d, wi, f);
        V3<R> est = f * std::abs(dot(it.n, wi)) / pdf;
        delta = bsdf.delta();
        R p = rmin((R)1.0, lumY(est * T) / lumY(T));
        if (p < 1 && i > rrDepth) {
            if ((R)rnd.next() < p) T = T / p;
            else break;
        }
        T = T * est;
        ray = spawn_ray(it, wi);
    }
    return L;
}

// One camera sample of TiledIntegrator::Render (integrator/TiledIntegrator.cpp:40-47).
template <class R> V3<R> Li_vpt(const Scene<R>& s, Ray<R> ray, Seq& rnd, int maxDepth, bool equiAngular, bool use_bvh);
// integrator 0: PathTracing; 1: VolPathTracing with free-flight sampling; 2:
// VolPathTracing with equi-angular sampling.
template <class R> V3<R> camera_sample(const Scene<R>& s, const Camera<R>& cam, int res, int m, int x, int y, int smp,
                                       int maxDepth, int rrDepth, bool use_bvh, R* fx, R* fy, int* dims,
                                       const hal::Enum* henum = nullptr, int integrator = 0) {
    Seq rnd;
    R u, v;
    if (henum) {
        // HaltonEnum::GetIndex / SampleX / SampleY (sampler/HaltonEnum.h:37-47)
        rnd.kind = 1;
        rnd.instance = henum->index((uint64_t)smp, (uint32_t)x, (uint32_t)y);
        u = (R)rnd.next() * (R)henum->w - R(x);
        v = (R)rnd.next() * (R)henum->h - R(y);
    } else {
        rnd.instance = enum_look_up((uint32_t)m, (uint32_t)smp, (uint32_t)x, (uint32_t)y);
        u = res * (R)rnd.next() - R(x);  // SobolEnum::SampleX (SobolEnum.h:30-32)
        v = res * (R)rnd.next() - R(y);
    }
    Ray<R> ray = generate_ray(cam, x, y, u, v);
    V3<R> L = integrator > 0 ? Li_vpt(s, ray, rnd, maxDepth, integrator == 2, use_bvh) : Li(s, ray, rnd, maxDepth, rrDepth, use_bvh);
    *fx = x + u;
    *fy = y + v;
    if (dims) *dims = rnd.cursor;
    return L;
}

// ---------------------------------------------------------------- light tracing
// LightTracing (integrator/LT.h:17-241): one light path per (x, y, p) sample
// index W H p + y W + x, every vertex connected to the pinhole camera and
// splatted with Film::AddSplat (visual/Film.h:89-103).  Area lights only;
// materials without transmission (their BSDFs do not depend on the transport
// mode).  The dimensions of the Vec2(rand(), rand()) / Vec3(rand(), rand(),
// rand()) arguments are drawn left to right, as for PathTracing.
template <class R> struct LTCamera {
    V3<R> pos, cz;
    M4<R> world_to_raster;  // CameraToRaster * WorldToCamera (integrator/LT.h:118-119)
    R dis, area;            // film distance, film area (camera/Camera.cpp:42-44)
    int w, h;
};
template <class R> LTCamera<R> make_lt_camera(const or_camera_desc& d) {
    LTCamera<R> c;
    V3<R> pos((R)d.pos[0], (R)d.pos[1], (R)d.pos[2]);
    V3<R> look((R)d.look_at[0], (R)d.look_at[1], (R)d.look_at[2]);
    V3<R> up((R)d.up[0], (R)d.up[1], (R)d.up[2]);
    V3<R> cz = -norm(pos - look);
    V3<R> cx = -norm(cross(up, cz));
    V3<R> cy = -cross(cz, cx);
    M4<R> c2w = M4<R>::rows(cx.x, cy.x, cz.x, pos.x, cx.y, cy.y, cz.y, pos.y, cx.z, cy.z, cz.z, pos.z, 0, 0, 0, (R)1.f);
    R fovy = (R)d.fovy, dis = (R)d.film_dist;
    R aspect = (R)d.width / (R)d.height;
    R radians = (K<R>::PI / 180) * fovy;
    R t = dis * std::tan(radians * 0.5);
    R r = t * aspect;
    R f = (R)100000.f, n = dis;
    M4<R> persp = M4<R>::rows(dis / r, 0, 0, 0, 0, dis / t, 0, 0, 0, 0, (f + n) / (f - n), -2 * f * n / (f - n), 0, 0, 1, 0);
    Xf<R> ndc_to_raster = xf_compose(xf_compose(xf_scale((R)d.width, (R)d.height, (R)1), xf_scale((R)0.5f, (R)-0.5f, (R)1)),
                                     xf_translate((R)1, (R)-1, (R)0));
    M4<R> cam_to_raster = mul(ndc_to_raster.m, persp);
    c.world_to_raster = mul(cam_to_raster, glm_inverse(c2w));
    c.pos = pos;
    c.cz = cz;
    c.dis = dis;
    // mHeight = d tan(fovy 0.5 PI / 180) 2, mWidth = mHeight aspect (camera/Camera.cpp:42-44)
    R height = dis * (R)std::tan((double)(fovy * (R)0.5f * K<R>::PI / 180)) * (R)2.f;
    R width = height * aspect;
    c.area = width * height;
    c.w = d.width;
    c.h = d.height;
    return c;
}
// PinHoleCamera::We (camera/PinholeCamera.h:31-39) for a ray leaving the camera along dir.
template <class R> inline R lt_we(const LTCamera<R>& c, const V3<R>& dir) {
    R pdfA = 1.0;
    R cosTheta = dot(c.cz, dir);
    R cos2Theta = cosTheta * cosTheta;
    return c.dis * c.dis * pdfA / (c.area * cos2Theta * cos2Theta);
}
// CoordinateSystem (math/GeometryUtils.h:11-17)
template <class R> inline void coordinate_system(const V3<R>& v1, V3<R>* v2, V3<R>* v3) {
    if (std::abs(v1.x) > std::abs(v1.y)) *v2 = V3<R>(-v1.z, 0, v1.x) / std::sqrt(v1.x * v1.x + v1.z * v1.z);
    else *v2 = V3<R>(0, v1.z, -v1.y) / std::sqrt(v1.y * v1.y + v1.z * v1.z);
    *v3 = cross(v1, *v2);
}
template <class R> struct LTVertex {
    V3<R> pos, n, outDir, throughput;
    BSDF<R> bsdf;
};
// GenerateLightPath + Trace + ConnectToCamera (integrator/LT.h:60-183).  Calls
// splat(x, y, L) for every vertex that reaches the screen.
template <class R, class F> void lt_path(const Scene<R>& s, const LTCamera<R>& c, Seq& rnd, int maxDepth, bool use_bvh,
                                         int64_t total, F splat) {
    std::vector<LTVertex<R>> path(maxDepth + 1);
    R pdfLight;
    const int light = sample_one_light(s, (R)rnd.next(), &pdfLight);
    R u0 = (R)rnd.next(), u1 = (R)rnd.next();
    R v0 = (R)rnd.next(), v1 = (R)rnd.next();
    V3<R> Le, lpos, ln, dir;
    R pdfA, pdfDir;
    if (s.light_tri[light] >= 0) {
        const Tri<R>& lt = s.tris[s.light_tri[light]];
        Le = s.light_le[light];
        // AreaLight::SampleOnLight (light/AreaLight.h:60-71), Triangle::Sample (shape/Triangle.cpp:141-158)
        R su0 = std::sqrt(u0);
        R b0 = 1 - su0, b1 = u1 * su0;
        lpos = (b0 * lt.p0 + b1 * lt.p1) + (1 - b0 - b1) * lt.p2;
        ln = (b0 * lt.n0 + b1 * lt.n1) + (1 - b0 - b1) * lt.n2;
        pdfA = 1 / lt.area;
        V3<R> ss, ts;
        coordinate_system(ln, &ss, &ts);
        V3<R> dl = cosine_hemisphere(v0, v1);
        dir = norm((ss * dl.x + ts * dl.y) + ln * dl.z);
        pdfDir = dl.z * K<R>::INV_PI;
    } else {
        // EnvironmentLight::SampleLight (light/EnvLight.cpp:22-39): a direction from
        // the map, a point on the bounding sphere's disk facing it.  Its radiance
        // is dropped: GenerateLightPath takes Le = Emission(), and
        // EnvironmentLight::Emission() is zero (light/EnvLight.h:20-22,
        // integrator/LT.h:108-111), so these paths carry zero throughput (NaN
        // where a pdf is zero).
        Le = V3<R>(0, 0, 0);
        V3<R> sd;
        (void)env_sample(s.env, u0, u1, &sd, &pdfDir);
        dir = -sd;
        ln = dir;
        V3<R> a1, a2;
        coordinate_system(-dir, &a1, &a2);
        R cx, cy;
        concentric_disk(v0, v1, &cx, &cy);
        const V3<R> pDisk = s.env.center + s.env.radius * (cx * a1 + cy * a2);
        lpos = pDisk + s.env.radius * (-dir);
        pdfA = 1 / (K<R>::PI * s.env.radius * s.env.radius);
    }
    path[0].pos = lpos;
    path[0].n = ln;
    path[0].throughput = Le;
    R cosTheta = std::abs(dot(ln, dir));
    V3<R> T = Le * cosTheta / pdfLight / pdfA / pdfDir;
    Ray<R> ray = make_ray(lpos + dir * K<R>::RayEps, dir, K<R>::Inf, (R)0);
    int bounce = 1;
    while (true) {
        Hit<R> h;
        if (!intersect(s, ray, &h, use_bvh)) break;
        Isect<R> it = query(s, ray, h);
        LTVertex<R>& vx = path[bounce];
        vx.pos = it.pos;
        vx.n = it.n;
        vx.outDir = it.outDir;
        vx.bsdf = make_bsdf(s, it);  // ComputeScatteringFunction(arena, Importance) (integrator/LT.h:69)
        vx.bsdf.importance = true;
        vx.throughput = T;
        bounce++;
        if (bounce >= maxDepth) break;
        R r0 = rnd.next(), r1 = rnd.next(), r2 = rnd.next();
        V3<R> wi;
        V3<R> f = vx.bsdf.Sample(-ray.d, &wi, &pdfDir, V3<R>(r0, r1, r2));
        ray = make_ray(it.pos + wi * K<R>::RayEps, wi, K<R>::Inf, (R)0);
        T = T * f * std::abs(dot(wi, it.n)) / pdfDir;
    }
    for (int sv = 0; sv < bounce; ++sv) {
        const LTVertex<R>& vx = path[sv];
        V3<R> toCam = c.pos - vx.pos;
        R len = length(toCam);
        toCam = norm(toCam);
        if (dot(c.cz, -toCam) < 0) continue;
        // WorldToScreen (integrator/LT.h:116-129): truncated raster position,
        // in screen for 0 <= x <= W, 0 <= y <= H
        V3<R> rp = xf_point(c.world_to_raster, vx.pos);
        int ix = int(rp.x), iy = int(rp.y);
        bool in = ix >= 0 && ix <= c.w && iy >= 0 && iy <= c.h;
        Ray<R> vis = make_ray(vx.pos, toCam, len, (R)0);
        if (intersect_any(s, vis, use_bvh)) in = false;
        V3<R> L;
        if (sv == 0) {
            L = vx.throughput;
        } else {
            // PinHoleCamera::Sample_Wi (camera/PinholeCamera.h:41-49); its
            // three random numbers are drawn but unused
            (void)rnd.next(); (void)rnd.next(); (void)rnd.next();
            V3<R> wi = c.pos - vx.pos;
            R distance = length(wi);
            wi = norm(wi);
            R cosT = dot(c.cz, -wi);
            R pdfW = (R)1.f * (distance * distance) / cosT;
            R We = lt_we(c, -wi);
            V3<R> f = vx.bsdf.Evaluate(vx.outDir, wi);
            L = V3<R>(We, We, We) * f * vx.throughput * std::abs(dot(wi, vx.n)) / pdfW;
        }
        if (in) splat((R)ix, (R)iy, L * (R)c.w * (R)c.h / (R)total);
    }
}

// Film::AddSample with the radius-0.5 box filter (visual/Film.h:61-87, math/Filter.h:26-38)
// ---------------------------------------------------------------- volumetric path tracing
// VolPathTracing (integrator/VPT.cpp:7-171) over HomogeneousMedium
// (volume/Homogeneous.cpp:6-64), HenyeyGreenstein (visual/Medium.cpp:8-29),
// MediumInterface (visual/Medium.h:51-59) and primitives without a material
// (medium boundaries).  Media are numbered; -1 is no medium (nullptr).
// Camera rays carry no medium (PinHoleCamera::GenerateRay sets none).
// Loops over medium boundaries are bounded (kVptMaxCrossings per loop), which
// no scene with closed boundaries reaches.  The reference reads past its 1024
// Sobol dimensions when a path draws more (an assert compiled out): here a
// path ends at the top of an iteration once it has drawn more than
// kVptDimBudget dimensions (an iteration draws at most 12).
constexpr int kVptMaxCrossings = 256;
constexpr int kVptDimBudget = 1024 - 16;
inline float rexp(float x) { return (float)std::exp((double)x); }
inline double rexp(double x) { return std::exp(x); }
inline float rtan(float x) { return (float)std::tan((double)x); }
inline double rtan(double x) { return std::tan(x); }
template <class R> inline V3<R> vexp(const V3<R>& a) { return V3<R>(rexp(a.x), rexp(a.y), rexp(a.z)); }
template <class R> constexpr R kInv4Pi = (R)0.07957747154594766788;  // common/Core.h:188
template <class R> constexpr R kMaxReal = (R)std::numeric_limits<float>::max();  // common/Core.h:198 (real = float)

// PhaseHG (visual/Medium.h:21-24)
template <class R> inline R phase_hg(R cosTheta, R g) {
    R denom = 1 + g * g + 2 * g * cosTheta;
    return kInv4Pi<R> * (1 - g * g) / (denom * std::sqrt(denom));
}
// HenyeyGreenstein::Sample_p: cos theta about -wo by inversion, the direction
// built on CoordinateSystem(wo) (SphericalDirection, math/GeometryUtils.h:29-33)
template <class R> R hg_sample(R g, const V3<R>& wo, V3<R>* wi, R u0, R u1) {
    R cosTheta;
    if ((double)std::abs(g) < 1e-3) {
        cosTheta = 1 - 2 * u0;
    } else {
        R sqrTerm = (1 - g * g) / (1 - g + 2 * g * u0);
        cosTheta = (1 + g * g - sqrTerm * sqrTerm) / (2 * g);
    }
    R sinTheta = std::sqrt(rmax((R)0, 1 - cosTheta * cosTheta));
    R phi = 2 * K<R>::PI * u1;
    V3<R> v1, v2;
    coordinate_system(wo, &v1, &v2);
    *wi = ((sinTheta * rcos(phi)) * v1 + (sinTheta * rsin(phi)) * v2) + cosTheta * (-wo);
    return phase_hg(-cosTheta, g);
}
template <class R> inline R hg_p(R g, const V3<R>& wo, const V3<R>& wi) { return phase_hg(dot(wo, wi), g); }

// HomogeneousMedium::Tr: exp(-sigma_t min(tMax |d|, MaxReal))
template <class R> inline V3<R> medium_tr(const Medium<R>& m, const Ray<R>& ray) {
    return vexp((-m.st) * rmin(ray.tmax * length(ray.d), kMaxReal<R>));
}
// A scattering point in a medium (MediumIntersection): position, outgoing
// direction, medium.  Its normals and error bound are zero.
template <class R> struct MediumRec {
    bool valid = false;
    V3<R> pos, wo;
    int32_t medium = -1;
};
template <class R> inline Isect<R> medium_isect(const MediumRec<R>& mi) {
    Isect<R> it;
    it.pos = mi.pos;
    it.n = it.absN = it.ng = it.sdpdu = it.sdpdv = it.pErr = V3<R>(0, 0, 0);
    it.outDir = mi.wo;
    return it;
}
// HomogeneousMedium::Sample: a channel, then a free-flight distance in it;
// the weight over the channel-averaged density.
template <class R> V3<R> medium_sample(const Medium<R>& m, int32_t med, const Ray<R>& ray, R u_ch, R u_d, MediumRec<R>* mi,
                                       R* t_out = nullptr) {
    int channel = std::min((int)(u_ch * 3), 2);
    R dist = -rlog(1 - u_d) / m.st[channel];
    R t = rmin(dist / length(ray.d), ray.tmax);
    bool sampled = t < ray.tmax;
    if (t_out) *t_out = t;
    if (sampled) {
        mi->valid = true;
        mi->pos = ray.o + ray.d * t;
        mi->wo = -ray.d;
        mi->medium = med;
    }
    V3<R> Tr = vexp(((-m.st) * rmin(t, kMaxReal<R>)) * length(ray.d));
    V3<R> density = sampled ? (m.st * Tr) : Tr;
    R pdf = 0;
    for (int i = 0; i < 3; ++i) pdf += density[i];
    pdf *= 1 / (R)(3);
    if (pdf == 0) pdf = 1;
    return sampled ? (Tr * m.ss / pdf) : (Tr / pdf);
}
// HomogeneousMedium::EquiAngularSampling: t about the light point's foot on
// the ray over [0, 3] (tMax = 3 is fixed), pdf D / ((theta_b - theta_a)(D^2 + tau^2));
// past the surface, the probability of t beyond min(3, tMax).
template <class R> V3<R> medium_equiangular(const Medium<R>& m, int32_t med, const Ray<R>& ray, R u, const V3<R>& lightPos,
                                            MediumRec<R>* mi, R* t_out = nullptr, R* pdf_out = nullptr) {
    const R tMax = 3;
    R delta = dot(lightPos - ray.o, ray.d);
    R D = length((ray.o + delta * ray.d) - lightPos);
    R theta_a = ratan2((R)0.f - delta, D);
    R theta_b = ratan2(tMax - delta, D);
    R tau = D * rtan((1 - u) * theta_a + u * theta_b);
    R t = delta + tau;
    bool sampleMedia = t < ray.tmax;
    V3<R> Tr = sampleMedia ? vexp((-m.st) * t) : medium_tr(m, ray);
    R pdf = sampleMedia ? D / ((theta_b - theta_a) * (D * D + tau * tau))
                        : 1 - (ratan2(rmin(tMax, ray.tmax) - delta, D) - theta_a) / (theta_b - theta_a);
    if (t_out) *t_out = t;
    if (pdf_out) *pdf_out = pdf;
    if (sampleMedia) {
        mi->valid = true;
        mi->pos = ray.o + ray.d * t;
        mi->wo = -ray.d;
        mi->medium = med;
    }
    return sampleMedia ? Tr * m.ss / pdf : medium_tr(m, ray) / pdf;
}

// The medium interface a closest hit leaves in its Intersection
// (Primitive::Intersect, visual/Primitive.cpp:11-14): the primitive's own when
// it is a transition (inside != outside), else the ray's medium on both sides.
template <class R> inline void hit_media(const Scene<R>& s, int32_t prim, int32_t ray_medium, int32_t* in, int32_t* out) {
    const int32_t a = s.mi_in.empty() ? -1 : s.mi_in[prim], b = s.mi_out.empty() ? -1 : s.mi_out[prim];
    if (a != b) {
        *in = a;
        *out = b;
    } else {
        *in = *out = ray_medium;
    }
}
// Intersection::GetMedium(w) (visual/Intersection.h:80-82): outside when w
// leaves along the shading normal, else inside.
template <class R> inline int32_t isect_medium(const Isect<R>& it, int32_t in, int32_t out, const V3<R>& w) {
    return dot(w, it.n) > 0 ? out : in;
}
template <class R> inline bool has_material(const Scene<R>& s, int32_t prim) { return s.no_mat.empty() || !s.no_mat[prim]; }

// VisibilityTester::Tr (visual/Visibility.cpp:13-35): closest hits from p0
// to the light point; a surface with a material occludes; transmittance of
// each segment's medium; a material-less surface is crossed (queried, then
// SpawnTo the light point again).
template <class R> V3<R> vis_tr(const Scene<R>& s, const Isect<R>& p0, int32_t p0_in, int32_t p0_out, const LightPoint<R>& p1,
                                bool use_bvh) {
    Ray<R> ray = spawn_to(p0, p1);
    int32_t med = isect_medium(p0, p0_in, p0_out, ray.d);
    V3<R> Tr(1.f);
    for (int k = 0; k < kVptMaxCrossings; ++k) {
        Hit<R> h;
        const bool hit = intersect(s, ray, &h, use_bvh);
        if (hit && has_material(s, h.prim)) return V3<R>();
        if (med >= 0) Tr = Tr * medium_tr(s.media[med], ray);
        if (!hit) return Tr;
        Isect<R> it = query(s, ray, h);
        int32_t in, out;
        hit_media(s, h.prim, med, &in, &out);
        ray = spawn_to(it, p1);
        med = isect_medium(it, in, out, ray.d);
    }
    return V3<R>();
}
// Scene::IntersectTr (visual/Scene.cpp:14-26): closest hits with the
// transmittance of each segment, crossing material-less surfaces.  The crossed
// record is never queried: it holds Triangle::Intersect's position and no
// normal or error bound, so the next segment starts at the hit point itself
// and takes the interface's inside medium.
// g_vpt_query_crossings is a test-only switch that queries the crossed
// surface and takes the medium on the far side of its normal instead.
static int g_vpt_query_crossings = 0;
template <class R> bool intersect_tr(const Scene<R>& s, Ray<R>& ray, int32_t* med, Hit<R>* h, V3<R>* Tr, bool use_bvh) {
    *Tr = V3<R>(1.f);
    for (int k = 0; k < kVptMaxCrossings; ++k) {
        const bool hit = intersect(s, ray, h, use_bvh);
        if (*med >= 0) *Tr = *Tr * medium_tr(s.media[*med], ray);
        if (!hit) return false;
        if (has_material(s, h->prim)) return true;
        const Tri<R>& t = s.tris[h->prim];
        const R b1 = h->b1, b2 = h->b2;
        const V3<R> pos = ((1 - b1 - b2) * t.p0 + b1 * t.p1) + b2 * t.p2;
        int32_t in, out;
        hit_media(s, h->prim, *med, &in, &out);
        if (g_vpt_query_crossings) {
            const Isect<R> it = query(s, ray, *h);
            *med = isect_medium(it, in, out, ray.d);
            ray = spawn_ray(it, ray.d);
            continue;
        }
        ray = make_ray(offset_ray_origin(pos, V3<R>(0, 0, 0), V3<R>(0, 0, 0), ray.d), ray.d, K<R>::Inf, (R)0);
        *med = in;
    }
    return false;
}

// Integrator::DirectIllumination with handleMedia (integrator/Integrator.cpp:9-107):
// at a surface (bsdf given) or a medium point (phase g), visibility by
// VisibilityTester::Tr and the BSDF/phase sample's light hit by IntersectTr.
template <class R> V3<R> di_vpt(const Scene<R>& s, const Isect<R>& it, int32_t in, int32_t out, const BSDF<R>* bsdf, R g,
                                R uLight, R u0, R u1, const V3<R>& v, bool use_bvh) {
    R lightSamplingPdf;
    R weight1 = 1, weight2 = 1;
    V3<R> L1, L2;
    int light = sample_one_light(s, uLight, &lightSamplingPdf);
    int ltri = s.light_tri[light];  // < 0: the environment light
    V3<R> Le = s.light_le[light];
    {
        R pdf;
        LightPoint<R> lp;
        V3<R> wi, Li;
        if (ltri < 0) {
            V3<R> sd;
            Li = env_sample(s.env, u0, u1, &sd, &pdf);
            wi = norm(sd);
            lp.pos = it.pos + wi * (2 * s.env.radius);
            lp.n = lp.fn = wi;
            lp.pErr = V3<R>(0, 0, 0);
        } else {
            lp = sample_light_point(s.tris[ltri], it, u0, u1, &pdf);
            wi = norm(lp.pos - it.pos);
            Li = dot(lp.n, -wi) > 0 ? Le : V3<R>(0, 0, 0);
        }
        if (pdf > 0 && Li != V3<R>(0)) {
            R scatteringPdf, cosTheta = 1;
            V3<R> f;
            if (bsdf) {
                scatteringPdf = bsdf->Pdf(it.outDir, wi);
                f = bsdf->Evaluate(it.outDir, wi);
                cosTheta = std::abs(dot(it.n, wi));
            } else {
                R p = hg_p(g, it.outDir, wi);
                f = V3<R>(p);
                scatteringPdf = p;
            }
            if (f != V3<R>()) {
                V3<R> Tr = vis_tr(s, it, in, out, lp, use_bvh);
                weight1 = power_heuristic(pdf, scatteringPdf);
                L1 = Li * Tr * f * cosTheta / pdf;
            }
        }
    }
    {
        V3<R> wi;
        R pdf, cosTheta = 1;
        V3<R> f;
        if (bsdf) {
            f = bsdf->Sample(it.outDir, &wi, &pdf, v);
            cosTheta = std::abs(dot(it.n, wi));
        } else {
            R p = hg_sample(g, it.outDir, &wi, v.x, v.y);
            f = V3<R>(p);
            pdf = p;
        }
        if (pdf != 0 && f != V3<R>()) {
            R lpdf = ltri < 0 ? env_pdf_dir(s.env, wi) : light_pdf(s, ltri, it, wi);
            if (lpdf == 0) {
                weight2 = 0;
            } else {
                weight2 = power_heuristic(pdf, lpdf);
                Ray<R> ray = spawn_ray(it, wi);
                int32_t med = isect_medium(it, in, out, wi);
                Hit<R> h;
                V3<R> Tr;
                if (intersect_tr(s, ray, &med, &h, &Tr, use_bvh)) {
                    if (prim_light(s, h.prim) == light) {
                        Isect<R> li = query(s, ray, h);
                        V3<R> Lemit = dot(li.n, -wi) > 0 ? Le : V3<R>(0, 0, 0);
                        L2 = Lemit * f * Tr * cosTheta / pdf;
                    }
                } else if (ltri < 0) {
                    L2 = env_emission(s.env, wi) * f * Tr * cosTheta / pdf;
                }
            }
        }
    }
    return (L1 * weight1 + L2 * weight2) / lightSamplingPdf;
}

// VolPathTracing::ConnectToLight (integrator/VPT.cpp:108-171): the light
// point alone, L1 without its MIS weight.
template <class R> V3<R> connect_to_light(const Scene<R>& s, const MediumRec<R>& mi, const LightPoint<R>& lp, const V3<R>& Le,
                                          R lightPdf, bool use_bvh) {
    R cosTheta = 1;
    V3<R> dir = mi.pos - lp.pos;
    V3<R> wi = -norm(dir);
    if (lp.n != V3<R>()) {
        cosTheta = dot(norm(dir), lp.n);
        if (cosTheta < 0) return V3<R>();
    }
    const Isect<R> it = medium_isect(mi);
    V3<R> Tr = vis_tr(s, it, mi.medium, mi.medium, lp, use_bvh);
    R p = hg_p(s.media[mi.medium].g, mi.wo, wi);
    V3<R> f = V3<R>(p);
    return Le * f * Tr * cosTheta / dot(dir, dir) / lightPdf;
}
// Triangle::Sample(pdf, u) (shape/Triangle.cpp:141-158): the area sample
// AreaLight::SampleOnePoint returns, pdf 1 / area.
template <class R> LightPoint<R> tri_sample_area(const Tri<R>& t, R u0, R u1, R* pdf) {
    R su0 = std::sqrt(u0);
    R b0 = 1 - su0, b1 = u1 * su0;
    R b2 = 1 - b0 - b1;
    LightPoint<R> lp;
    lp.pos = (b0 * t.p0 + b1 * t.p1) + b2 * t.p2;
    lp.fn = t.fn;
    lp.n = (b0 * t.n0 + b1 * t.n1) + b2 * t.n2;
    lp.pErr = gamma_n<R>(6) * ((vabs(b0 * t.p0) + vabs(b1 * t.p1)) + vabs(b2 * t.p2));
    *pdf = 1 / t.area;
    return lp;
}

// VolPathTracing::Li.  The dimensions of Vec2(rand(), rand()) /
// Vec3(rand(), rand(), rand()) arguments are drawn left to right, as for
// PathTracing.  The roulette test min(1, Y(T) / Y(T)) < 1 never passes.
template <class R> V3<R> Li_vpt(const Scene<R>& s, Ray<R> ray, Seq& rnd, int maxDepth, bool equiAngular, bool use_bvh) {
    int32_t medium = -1;
    bool delta = false;
    V3<R> T(1, 1, 1);
    V3<R> L;
    int crossings = 0;
    for (int i = 0; i < maxDepth; ++i) {
        if (rnd.cursor > kVptDimBudget) break;
        Hit<R> h;
        const bool hit = intersect(s, ray, &h, use_bvh);
        if (t_rec && t_rec_n < t_rec_max) {  // the path record of or_li_path, as Li's
            double* r = t_rec + (size_t)kRecWords * t_rec_n++;
            r[0] = ray.o.x; r[1] = ray.o.y; r[2] = ray.o.z;
            r[3] = ray.d.x; r[4] = ray.d.y; r[5] = ray.d.z;
            r[6] = ray.tmin; r[7] = ray.tmax;
            r[8] = hit ? h.prim : -1; r[9] = hit ? h.t : 0;
            r[10] = rnd.cursor;
            r[11] = T.x; r[12] = T.y; r[13] = T.z;
            r[14] = medium;
            r[15] = delta ? 1 : 0;
        }
        MediumRec<R> mi;
        R eqProb = 0, eqPdf = 0;
        V3<R> eqLe;
        LightPoint<R> eqLp;
        if (equiAngular) {
            const int light = sample_one_light(s, (R)rnd.next(), &eqProb);
            R a = rnd.next(), b = rnd.next();
            eqLp = tri_sample_area(s.tris[s.light_tri[light]], a, b, &eqPdf);
            eqLe = s.light_le[light];
            if (medium >= 0) T = T * medium_equiangular(s.media[medium], medium, ray, (R)rnd.next(), eqLp.pos, &mi);
        } else if (medium >= 0) {
            R a = rnd.next(), b = rnd.next();
            T = T * medium_sample(s.media[medium], medium, ray, a, b, &mi);
        }
        if (T == V3<R>()) break;
        if (mi.valid) {
            if (equiAngular) {
                L = L + T * connect_to_light(s, mi, eqLp, eqLe, eqPdf, use_bvh) / eqProb;
            } else {
                R uL = rnd.next();
                R u0 = rnd.next(), u1 = rnd.next();
                R v0 = rnd.next(), v1 = rnd.next(), v2 = rnd.next();
                L = L + T * di_vpt(s, medium_isect(mi), mi.medium, mi.medium, (const BSDF<R>*)nullptr, s.media[mi.medium].g, uL,
                                   u0, u1, V3<R>(v0, v1, v2), use_bvh);
            }
            R a = rnd.next(), b = rnd.next();
            V3<R> wi;
            hg_sample(s.media[mi.medium].g, mi.wo, &wi, a, b);
            ray = spawn_ray(medium_isect(mi), wi);
            medium = mi.medium;
            delta = false;
        } else {
            if (!hit) {
                if ((i == 0 || delta) && s.env.w > 0) return T * env_emission(s.env, ray.d);
                break;
            }
            Isect<R> it = query(s, ray, h);
            int32_t in, out;
            hit_media(s, h.prim, medium, &in, &out);
            if (!has_material(s, h.prim)) {
                if (++crossings > kVptMaxCrossings) break;
                ray = spawn_ray(it, ray.d);
                medium = isect_medium(it, in, out, ray.d);
                --i;
                continue;
            }
            BSDF<R> bsdf = make_bsdf(s, it);
            int light = prim_light(s, it.prim);
            if ((i == 0 || delta) && light >= 0) {
                V3<R> e = dot(it.n, it.outDir) > 0 ? s.light_le[light] : V3<R>(0, 0, 0);
                L = L + T * e;
            } else {
                R uL = rnd.next();
                R u0 = rnd.next(), u1 = rnd.next();
                R v0 = rnd.next(), v1 = rnd.next(), v2 = rnd.next();
                L = L + T * di_vpt(s, it, in, out, &bsdf, (R)0, uL, u0, u1, V3<R>(v0, v1, v2), use_bvh);
            }
            R r0 = rnd.next(), r1 = rnd.next(), r2 = rnd.next();
            V3<R> wi;
            R pdf;
            V3<R> f = bsdf.Sample(-ray.d, &wi, &pdf, V3<R>(r0, r1, r2));
            if (f == V3<R>() || pdf == 0) break;
            V3<R> est = f * std::abs(dot(it.n, wi)) / pdf;
            delta = bsdf.delta();
            T = T * est;
            ray = spawn_ray(it, wi);
            medium = isect_medium(it, in, out, wi);
        }
    }
    return L;
}

template <class R> struct Film {
    int W, H;
    std::vector<R> color, weight;
    std::vector<std::atomic_flag> locks;
    Film(int w, int h) : W(w), H(h), color((size_t)w * h * 3, 0), weight((size_t)w * h, 0), locks((size_t)w * h) {
        for (auto& l : locks) l.clear();
    }
    void add(R x, R y, const V3<R>& L) {
        x -= 0.5;
        y -= 0.5;
        const R rad = (R)0.5;
        int minX = (int)std::ceil(x - rad), maxX = (int)std::floor(x + rad);
        int minY = (int)std::ceil(y - rad), maxY = (int)std::floor(y + rad);
        minX = std::max(0, minX);
        maxX = std::min(maxX, W - 1);
        minY = std::max(0, minY);
        maxY = std::min(maxY, H - 1);
        for (int i = minY; i <= maxY; i++)
            for (int j = minX; j <= maxX; j++) {
                size_t p = (size_t)i * W + j;
                while (locks[p].test_and_set(std::memory_order_acquire)) {
                }
                R w = 1.0;
                weight[p] += w;
                color[3 * p + 0] = color[3 * p + 0] + L.x * w;
                color[3 * p + 1] = color[3 * p + 1] + L.y * w;
                color[3 * p + 2] = color[3 * p + 2] + L.z * w;
                locks[p].clear(std::memory_order_release);
            }
    }
};

struct SceneHandle {
    or_scene_desc desc;  // points into the owned copies below
    std::vector<float> pos, nrm, fn, uv, le, texd;
    std::vector<uint32_t> mat, mats;
    std::vector<int32_t> light, ltri, texi;
    Scene<float> sf;
    std::unique_ptr<Scene<double>> sd;
    std::mutex mu;
    std::vector<float> env;  // environment image, top row first (empty: none)
    int env_w = 0, env_h = 0;
    float env_scale = 1.f;
    std::vector<or_medium> media;  // or_scene_set_media (empty: none)
    std::vector<or_tri_medium> tri_media;
};

template <class R> void attach_media(Scene<R>& s, const SceneHandle* h) {
    if (h->tri_media.empty()) return;
    s.media.resize(h->media.size());
    for (size_t i = 0; i < h->media.size(); ++i) {
        const or_medium& m = h->media[i];
        Medium<R>& d = s.media[i];
        d.sa = V3<R>(m.sigma_a[0], m.sigma_a[1], m.sigma_a[2]);
        d.ss = V3<R>(m.sigma_s[0], m.sigma_s[1], m.sigma_s[2]);
        d.st = d.ss + d.sa;
        d.g = m.g;
    }
    const size_t n = h->tri_media.size();
    s.mi_in.resize(n);
    s.mi_out.resize(n);
    s.no_mat.resize(n);
    for (size_t i = 0; i < n; ++i) {
        s.mi_in[i] = h->tri_media[i].inside;
        s.mi_out[i] = h->tri_media[i].outside;
        s.no_mat[i] = h->tri_media[i].no_material ? 1 : 0;
    }
}

// Scene::AddLight of an EnvironmentLight after the area lights, then
// Scene::Initialize: light init (bounding sphere), power distribution.
template <class R> void attach_env(Scene<R>& s, const SceneHandle* h) {
    if (h->env.empty()) return;
    build_env(s, h->env.data(), h->env_w, h->env_h, h->env_scale);
    if (s.light_tri.empty() || s.light_tri.back() >= 0) {
        s.light_tri.push_back(-1);
        s.light_le.push_back(V3<R>(0, 0, 0));
    }
    build_light_distribution(s);
}

template <class R> void fill_scene(Scene<R>& s, const or_scene_desc& d) {
    build_tris(s, d);
    s.mats.resize(d.n_mat);
    for (int i = 0; i < d.n_mat; ++i) {
        const uint32_t* w = d.mats + OR_MAT_WORDS * i;
        Material& m = s.mats[i];
        auto f = [&](int k) { float v; std::memcpy(&v, w + k, 4); return v; };
        m.type = w[0];
        m.flags = w[1];
        for (int c = 0; c < 3; ++c) {
            m.R[c] = f(2 + c); m.T[c] = f(5 + c); m.eta[c] = f(10 + c); m.k[c] = f(13 + c); m.f0[c] = f(16 + c);
        }
        m.ax = f(8); m.ay = f(9); m.eta1 = f(19); m.eta2 = f(20);
        m.tex_albedo = (int32_t)w[21];
        m.tex_rough = (int32_t)w[22];
    }
    s.light_tri.assign(d.light_tri, d.light_tri + d.n_light);
    s.light_le.resize(d.n_light);
    for (int i = 0; i < d.n_light; ++i) s.light_le[i] = V3<R>(d.light_le[3 * i], d.light_le[3 * i + 1], d.light_le[3 * i + 2]);
    s.tex.resize(d.n_tex);
    for (int i = 0; i < d.n_tex; ++i) {
        Texture<R>& t = s.tex[i];
        t.w = d.tex_info[4 * i]; t.h = d.tex_info[4 * i + 1]; t.c = d.tex_info[4 * i + 2];
        const float* src = d.tex_data + d.tex_info[4 * i + 3];
        t.data.assign(src, src + (size_t)t.w * t.h * t.c);
    }
    if (!s.light_tri.empty()) build_light_distribution(s);
    build_bvh(s);
}

template <class R> Scene<R>& scene_of(SceneHandle* h);
template <> Scene<float>& scene_of<float>(SceneHandle* h) { return h->sf; }
template <> Scene<double>& scene_of<double>(SceneHandle* h) {
    std::lock_guard<std::mutex> g(h->mu);
    if (!h->sd) {
        h->sd.reset(new Scene<double>());
        fill_scene(*h->sd, h->desc);
        attach_media(*h->sd, h);
        attach_env(*h->sd, h);
    }
    return *h->sd;
}

// ---------------------------------------------------------------- SPPM
// Stochastic progressive photon mapping (integrator/SPPM.cpp:11-277) with its
// spatial hash (integrator/HashGrid.h:11-125), for triangle scenes with area
// lights.  The reference runs eye paths and photons in parallel and updates
// hit points from concurrent photons without locks; here photons update the
// hit points one at a time in photon-index order (and, within a photon, in
// vertex order), the canonical order of the same arithmetic.
template <class R> struct SppmState {
    int W = 0, H = 0;
    std::vector<R> r2, N;
    std::vector<V3<R>> flux, direct;
    std::vector<V3<R>> imp, pos, nrm, out;  // HPoint: importance, pos, nrm (mAbsNormal), outDir
    std::vector<uint8_t> used;
    std::vector<int64_t> m;
};
// SPPM::TraceEyePath (SPPM.cpp:45-108) for pixel (x, y) of iteration iter; the
// sample index is SamplerEnum::GetIndex(iter, x, y) (SPPM.cpp:212).
template <class R> void sppm_eye(const Scene<R>& s, const Camera<R>& cam, int res, int m, const hal::Enum* henum, int x, int y,
                                 int iter, int maxDepth, SppmState<R>& st) {
    const int64_t px = x + (int64_t)y * st.W;
    st.used[px] = 0;
    Seq rnd;
    R u, v;
    if (henum) {
        rnd.kind = 1;
        rnd.instance = henum->index((uint64_t)iter, (uint32_t)x, (uint32_t)y);
        u = (R)rnd.next() * (R)henum->w - R(x);
        v = (R)rnd.next() * (R)henum->h - R(y);
    } else {
        rnd.instance = enum_look_up((uint32_t)m, (uint32_t)iter, (uint32_t)x, (uint32_t)y);
        u = res * (R)rnd.next() - R(x);
        v = res * (R)rnd.next() - R(y);
    }
    Ray<R> r = generate_ray(cam, x, y, u, v);
    bool delta = false;
    V3<R> importance(1.0, 1.0, 1.0);
    for (int i = 0; i < maxDepth; ++i) {
        Hit<R> h;
        if (!intersect(s, r, &h, true)) {
            if ((i == 0 || delta) && s.env.w > 0) st.direct[px] = st.direct[px] + importance * env_emission(s.env, r.d);
            return;
        }
        Isect<R> it = query(s, r, h);
        BSDF<R> bsdf = make_bsdf(s, it);
        if (bsdf.delta()) {
            R a = rnd.next(), b = rnd.next(), c = rnd.next();
            V3<R> wi;
            R pdf;
            V3<R> f = bsdf.Sample((R)-1 * r.d, &wi, &pdf, V3<R>(a, b, c));
            if (f == V3<R>() || pdf == 0) return;
            importance = f * std::abs(dot(wi, it.n)) * importance / pdf;
            r = spawn_ray(it, wi);
            delta = true;
        } else {
            st.used[px] = 1;
            st.imp[px] = importance;
            st.pos[px] = it.pos;
            st.nrm[px] = it.absN;
            st.out[px] = (R)-1 * r.d;
            const int light = prim_light(s, it.prim);
            if ((i == 0 || delta) && light >= 0) {
                V3<R> e = dot(it.n, it.outDir) > 0 ? s.light_le[light] : V3<R>(0, 0, 0);
                st.direct[px] = st.direct[px] + importance * e;
            } else {
                R uL = rnd.next();
                R u0 = rnd.next(), u1 = rnd.next();
                R v0 = rnd.next(), v1 = rnd.next(), v2 = rnd.next();
                V3<R> Ld = importance * direct_illumination(s, it, bsdf, uL, u0, u1, V3<R>(v0, v1, v2), true);
                st.direct[px] = st.direct[px] + Ld;
            }
            return;
        }
    }
}
// HashGrid: buckets = used hit points, cells of twice the search radius over
// the hit points' box inflated by it; a hit point is listed in the bucket of
// every cell its own radius box touches (a bucket may list it more than once).
template <class R> struct SppmGrid {
    std::vector<std::vector<int64_t>> buckets;
    int64_t num = 0;
    R inv = 0;
    V3<R> bmin;
};
inline uint32_t sppm_hash(int ix, int iy, int iz, int64_t num) {  // HashGrid::hash: int products, wrapping
    const uint32_t h = ((uint32_t)ix * 73856093u) ^ ((uint32_t)iy * 19349663u) ^ ((uint32_t)iz * 83492791u);
    return (uint32_t)((int64_t)h % num);
}
template <class R> void sppm_grid_build(const SppmState<R>& st, SppmGrid<R>* g) {
    R maxR2 = 0;
    int64_t n = 0;
    for (size_t i = 0; i < st.used.size(); ++i)
        if (st.used[i]) {
            maxR2 = rmax(maxR2, st.r2[i]);
            ++n;
        }
    const R search = std::sqrt(maxR2) + (R)1e-6f;  // Eps (common/Core.h:191)
    V3<R> mn(K<R>::Inf, K<R>::Inf, K<R>::Inf);
    for (size_t i = 0; i < st.used.size(); ++i)
        if (st.used[i]) {
            const V3<R> a = st.pos[i] - V3<R>(search), b = st.pos[i] + V3<R>(search);
            for (int k = 0; k < 3; ++k) {
                if (a[k] < mn[k]) mn[k] = a[k];
                if (b[k] < mn[k]) mn[k] = b[k];
            }
        }
    g->bmin = mn;
    g->inv = 1.f / (search * 2.f);
    g->num = n;
    g->buckets.assign((size_t)n, {});
    if (n == 0) return;
    for (size_t i = 0; i < st.used.size(); ++i) {
        if (!st.used[i]) continue;
        const R rad = std::sqrt(st.r2[i]);
        const V3<R> BMin = ((st.pos[i] - V3<R>(rad)) - mn) * g->inv;
        const V3<R> BMax = ((st.pos[i] + V3<R>(rad)) - mn) * g->inv;
        for (int iz = std::abs(int(BMin.z)); iz <= std::abs(int(BMax.z)); iz++)
            for (int iy = std::abs(int(BMin.y)); iy <= std::abs(int(BMax.y)); iy++)
                for (int ix = std::abs(int(BMin.x)); ix <= std::abs(int(BMax.x)); ix++)
                    g->buckets[sppm_hash(ix, iy, iz, n)].push_back((int64_t)i);
    }
}
template <class R> const std::vector<int64_t>& sppm_grid_bucket(const SppmGrid<R>& g, const V3<R>& p) {
    const V3<R> hh = (p - g.bmin) * g.inv;
    return g.buckets[sppm_hash(std::abs(int(hh.x)), std::abs(int(hh.y)), std::abs(int(hh.z)), g.num)];
}
// SPPM::GeneratePhoton + TracePhoton (SPPM.cpp:28-43, 110-167) for photon
// `index` (the sampler instance iter * photons + j, SPPM.cpp:238).
template <class R> void sppm_photon(const Scene<R>& s, uint64_t index, int sampler, int maxDepth, R alpha, bool batch,
                                    SppmState<R>& st, const SppmGrid<R>& g) {
    Seq rnd;
    rnd.kind = sampler;
    rnd.instance = index;
    R lightPdf;
    const int light = sample_one_light(s, (R)rnd.next(), &lightPdf);
    R u0 = rnd.next(), u1 = rnd.next();
    R w0 = rnd.next(), w1 = rnd.next();
    // AreaLight::SampleLight = SampleOnLight (light/AreaLight.h:60-71) + Emission()
    const Tri<R>& t = s.tris[s.light_tri[light]];
    R su0 = std::sqrt(u0);
    R b0 = 1 - su0, b1 = u1 * su0;
    const V3<R> lpos = (b0 * t.p0 + b1 * t.p1) + (1 - b0 - b1) * t.p2;
    const V3<R> ln = (b0 * t.n0 + b1 * t.n1) + (1 - b0 - b1) * t.n2;
    const V3<R> lerr = gamma_n<R>(6) * ((vabs(b0 * t.p0) + vabs(b1 * t.p1)) + vabs((1 - b0 - b1) * t.p2));
    const R pdfPos = 1 / t.area;
    V3<R> ss, ts;
    coordinate_system(ln, &ss, &ts);
    const V3<R> dl = cosine_hemisphere(w0, w1);
    const V3<R> dir = norm((ss * dl.x + ts * dl.y) + ln * dl.z);
    const R pdfDir = dl.z * K<R>::INV_PI;
    Ray<R> r = make_ray(offset_ray_origin(lpos, lerr, t.fn, dir), dir, K<R>::Inf, (R)0);  // lightPoint.SpawnRay
    const R cosTheta = std::abs(dot(ln, dir));
    V3<R> flux = (pdfPos == 0 || pdfDir == 0) ? V3<R>(0, 0, 0) : s.light_le[light] * cosTheta / (pdfPos * pdfDir * lightPdf);
    for (int i = 0; i < maxDepth; ++i) {
        Hit<R> h;
        if (!intersect(s, r, &h, true)) return;
        Isect<R> it = query(s, r, h);
        BSDF<R> bsdf = make_bsdf(s, it);
        bsdf.importance = true;  // ComputeScatteringFunction(arena, TransportMode::Importance)
        R a = rnd.next(), b = rnd.next(), c = rnd.next();
        V3<R> wi;
        R pdf;
        V3<R> f = bsdf.Sample((R)-1 * r.d, &wi, &pdf, V3<R>(a, b, c));
        if (f == V3<R>() || pdf == 0) break;
        V3<R> est = f * std::abs(dot(wi, it.n)) / pdf;
        if (!bsdf.delta() && i > 0 && g.num > 0) {
            for (int64_t hp : sppm_grid_bucket(g, it.pos)) {
                const V3<R> v = st.pos[hp] - it.pos;
                if ((double)dot(st.nrm[hp], it.absN) > 0.0015 && dot(v, v) <= st.r2[hp]) {
                    const V3<R> contribution = st.imp[hp] * bsdf.Evaluate(st.out[hp], (R)-1 * r.d) * flux;
                    if (!batch) {
                        R gg = (st.N[hp] * alpha + alpha) / (st.N[hp] * alpha + 1.f);
                        st.r2[hp] = st.r2[hp] * gg;
                        st.N[hp] += 1;
                        st.flux[hp] = (st.flux[hp] + contribution) * gg;
                    } else {
                        st.m[hp]++;
                        st.flux[hp] = st.flux[hp] + contribution;
                    }
                }
            }
        }
        R p = rmin((R)1.0, lumY(est * flux) / lumY(flux));
        if (p < 1) {
            if ((R)rnd.next() < p) flux = flux / p;
            else break;
        }
        flux = flux * est;
        r = spawn_ray(it, wi);
    }
}
// SPPM::Render (SPPM.cpp:198-277): iterations of eye pass, grid, photons and
// (batchShrink) the per-hit-point update; then the density estimate.
template <class R> void sppm_render(SceneHandle* h, const or_camera_desc* cd, const or_sppm_params* p, double* color, double* r2_out,
                                    double* n_out) {
    Scene<R>& s = scene_of<R>(h);
    Camera<R> cam = make_camera<R>(*cd);
    const int W = cd->width, H = cd->height;
    const int res = round_up_pow2(std::max(W, H));
    const int m = log2_int(res);
    const hal::Enum henum((uint32_t)W, (uint32_t)H);
    SppmState<R> st;
    st.W = W;
    st.H = H;
    const size_t n = (size_t)W * H;
    const R r0 = (R)p->initial_radius, alpha = (R)p->alpha;
    st.r2.assign(n, r0 * r0);
    st.N.assign(n, 0);
    st.flux.assign(n, V3<R>());
    st.direct.assign(n, V3<R>());
    st.imp.assign(n, V3<R>());
    st.pos.assign(n, V3<R>());
    st.nrm.assign(n, V3<R>());
    st.out.assign(n, V3<R>());
    st.used.assign(n, 0);
    st.m.assign(n, 0);
    const int nth = p->nthreads > 0 ? p->nthreads : (int)std::thread::hardware_concurrency();
    for (int iter = 0; iter < p->iterations; ++iter) {
        std::atomic<int> next(0);
        auto work = [&]() {
            for (int y; (y = next.fetch_add(1)) < H;)
                for (int x = 0; x < W; ++x)
                    sppm_eye(s, cam, res, m, p->sampler == 1 ? &henum : nullptr, x, y, iter, p->max_depth, st);
        };
        std::vector<std::thread> th;
        for (int t = 0; t < nth; ++t) th.emplace_back(work);
        for (auto& t : th) t.join();
        SppmGrid<R> grid;
        sppm_grid_build(st, &grid);
        for (int64_t j = 0; j < p->photons; ++j)
            sppm_photon(s, (uint64_t)iter * (uint64_t)p->photons + (uint64_t)j, p->sampler, p->max_depth, alpha, p->batch_shrink != 0,
                        st, grid);
        if (p->batch_shrink) {
            for (size_t i = 0; i < n; ++i) {
                if (st.m[i] > 0) {
                    R gg = (st.N[i] + alpha * (R)st.m[i]) / (st.N[i] + (R)st.m[i]);
                    st.r2[i] = st.r2[i] * gg;
                    st.flux[i] = st.flux[i] * gg;
                    st.N[i] = st.N[i] + alpha * (R)st.m[i];
                    st.m[i] = 0;
                }
            }
        }
    }
    for (size_t i = 0; i < n; ++i) {
        const R k = 1.f / (K<R>::PI * st.r2[i] * (R)p->iterations * (R)p->photons);
        const V3<R> c = (V3<R>() + st.flux[i] * k) + st.direct[i] / (R)p->iterations;
        color[3 * i] = c.x;
        color[3 * i + 1] = c.y;
        color[3 * i + 2] = c.z;
        if (r2_out) r2_out[i] = st.r2[i];
        if (n_out) n_out[i] = st.N[i];
    }
}

template <class R> int li_samples_t(SceneHandle* h, const or_camera_desc* cd, const or_render_params* p, int32_t n,
                                    const int32_t* x, const int32_t* y, const int32_t* s, double* out_L, float* out_raster,
                                    int32_t* out_dims) {
    Scene<R>& sc = scene_of<R>(h);
    Camera<R> cam = make_camera<R>(*cd);
    int res = round_up_pow2(std::max(cd->width, cd->height));
    int m = log2_int(res);
    const hal::Enum henum((uint32_t)cd->width, (uint32_t)cd->height);
    int nth = p->nthreads > 0 ? p->nthreads : (int)std::thread::hardware_concurrency();
    std::atomic<int32_t> next(0);
    auto work = [&]() {
        for (;;) {
            int32_t i = next.fetch_add(64);
            if (i >= n) break;
            for (int32_t k = i; k < std::min(n, i + 64); ++k) {
                R fx, fy;
                int dims;
                V3<R> L = camera_sample(sc, cam, res, m, x[k], y[k], s[k], p->max_depth, p->rr_depth, p->use_bvh != 0, &fx, &fy,
                                        &dims, p->sampler == 1 ? &henum : nullptr, p->integrator);
                out_L[3 * k] = L.x; out_L[3 * k + 1] = L.y; out_L[3 * k + 2] = L.z;
                if (out_raster) { out_raster[2 * k] = (float)fx; out_raster[2 * k + 1] = (float)fy; }
                if (out_dims) out_dims[k] = dims;
            }
        }
    };
    if (nth <= 1) {
        work();  // on the calling thread (the path recorder is thread-local)
        return 0;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t) th.emplace_back(work);
    for (auto& t : th) t.join();
    return 0;
}

template <class R> int64_t render_t(SceneHandle* h, const or_camera_desc* cd, const or_render_params* p, int32_t row0,
                                    int32_t row1, double* color, double* weight, double* seconds) {
    Scene<R>& sc = scene_of<R>(h);
    Camera<R> cam = make_camera<R>(*cd);
    int W = cd->width, H = cd->height;
    if (row1 <= 0 || row1 > H) row1 = H;
    if (row0 < 0) row0 = 0;
    int res = round_up_pow2(std::max(W, H));
    int m = log2_int(res);
    const hal::Enum henum((uint32_t)W, (uint32_t)H);
    Film<R> film(W, H);
    // 32x32 tiles in row-major order (integrator/TiledIntegrator.cpp:13-25)
    struct Tile { int x0, y0, x1, y1; };
    std::vector<Tile> tiles;
    for (int ty = 0; ty < H; ty += 32)
        for (int tx = 0; tx < W; tx += 32) {
            Tile t{tx, std::max(ty, row0), std::min(W, tx + 32), std::min(std::min(H, ty + 32), row1)};
            if (t.y0 < t.y1) tiles.push_back(t);
        }
    int nth = p->nthreads > 0 ? p->nthreads : (int)std::thread::hardware_concurrency();
    std::atomic<int> next(0);
    std::atomic<int64_t> paths(0);
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&]() {
        int64_t local = 0;
        for (;;) {
            int ti = next.fetch_add(1);
            if (ti >= (int)tiles.size()) break;
            const Tile& t = tiles[ti];
            for (int y = t.y0; y < t.y1; ++y)
                for (int x = t.x0; x < t.x1; ++x)
                    for (int smp = 0; smp < p->spp; ++smp) {
                        R fx, fy;
                        V3<R> L = camera_sample(sc, cam, res, m, x, y, smp, p->max_depth, p->rr_depth, p->use_bvh != 0, &fx, &fy,
                                                (int*)nullptr, p->sampler == 1 ? &henum : nullptr, p->integrator);
                        film.add(fx, fy, L);
                        ++local;
                    }
        }
        paths += local;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < nth; ++t) th.emplace_back(work);
    for (auto& t : th) t.join();
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (color)
        for (size_t i = 0; i < film.color.size(); ++i) color[i] = film.color[i];
    if (weight)
        for (size_t i = 0; i < film.weight.size(); ++i) weight[i] = film.weight[i];
    return paths.load();
}

// Test-only: switch a scene to detached evaluation (Scene::detached) with
// texels tex_eval (the layout of the scene's tex_data; null: the scene's own)
// and per-light emitted radiance le_eval (3 per light; null: the scene's
// own); both null switches it off.
template <class R> void set_detached(Scene<R>& s, const or_scene_desc& d, const float* tex, const float* le) {
    if (!tex && !le) {
        s.detached = false;
        s.tex_eval.clear();
        s.light_le_eval.clear();
        return;
    }
    s.detached = true;
    s.tex_eval = s.tex;
    if (tex)
        for (size_t i = 0; i < s.tex_eval.size(); ++i) {
            Texture<R>& t = s.tex_eval[i];
            const float* src = tex + d.tex_info[4 * i + 3];
            for (size_t k = 0; k < t.data.size(); ++k) t.data[k] = (R)src[k];
        }
    s.light_le_eval = s.light_le;
    if (le)
        for (size_t i = 0; i < s.light_le_eval.size(); ++i) s.light_le_eval[i] = V3<R>(le[3 * i], le[3 * i + 1], le[3 * i + 2]);
}

// Spheres appended to a scene: 9 floats each (center xyz, radius, inside,
// material, emissive, Le rgb, as gy_sphere); an emissive sphere's AreaLight
// joins the light list after the existing area lights.
template <class R> void add_spheres(Scene<R>& s, int32_t n, const or_sphere* sp) {
    for (int32_t i = 0; i < n; ++i) {
        typename Scene<R>::Sph q;
        q.c = V3<R>(sp[i].center[0], sp[i].center[1], sp[i].center[2]);
        q.r = sp[i].radius;
        q.inside = sp[i].inside != 0;
        q.mat = sp[i].material;
        const V3<R> rv(q.r, q.r, q.r);
        // AABB(c - r, c + r), ObjectToWorld of its eight corners
        q.box.mn = xf_identity(q.c - rv);
        q.box.mx = xf_identity(q.c + rv);
        q.area = 4.f * K<R>::PI * q.r * q.r;
        if (sp[i].emissive) {
            q.light = (int32_t)s.light_tri.size();
            s.light_tri.push_back((int32_t)(s.tris.size() + s.sph.size()));
            s.light_le.push_back(V3<R>(sp[i].le[0], sp[i].le[1], sp[i].le[2]));
        }
        s.sph.push_back(q);
    }
    build_light_distribution(s);
}

}  // namespace orc

using namespace orc;

// Test hook: sphere i's Sample(Intersection, pdf, u) and Pdf_Li from a
// shading point at ref (zero error bound, normal along wi for the spawned ray).
template <class R> static int sphere_light_probe(Scene<R>& s, int32_t i, const double* ref, const double* u,
                                                 const double* wi, double* out) {
    if (i < 0 || i >= (int32_t)s.sph.size()) return -1;
    Isect<R> it{};
    it.pos = V3<R>((R)ref[0], (R)ref[1], (R)ref[2]);
    it.pErr = V3<R>(0, 0, 0);
    R pdf;
    const LightPoint<R> lp = sphere_sample_li<R>(s.sph[i], it, (R)u[0], (R)u[1], &pdf);
    out[0] = lp.pos.x; out[1] = lp.pos.y; out[2] = lp.pos.z;
    out[3] = lp.n.x; out[4] = lp.n.y; out[5] = lp.n.z;
    out[6] = pdf;
    const V3<R> w((R)wi[0], (R)wi[1], (R)wi[2]);
    it.absN = w;
    out[7] = sphere_light_pdf<R>(s.sph[i], it, w);
    return 0;
}

template <class R> static int hg_probe_t(double g, const double* wo, const double* u, double* wi, double* p) {
    V3<R> w((R)wo[0], (R)wo[1], (R)wo[2]), o;
    R v = hg_sample<R>((R)g, w, &o, (R)u[0], (R)u[1]);
    wi[0] = o.x; wi[1] = o.y; wi[2] = o.z;
    p[0] = v;
    p[1] = hg_p<R>((R)g, w, o);
    return 0;
}
template <class R> static int medium_probe_t(int kind, const double* sa, const double* ss, const double* o, const double* d,
                                             double tmax, const double* u, const double* lp, double* out) {
    Medium<R> m;
    m.sa = V3<R>((R)sa[0], (R)sa[1], (R)sa[2]);
    m.ss = V3<R>((R)ss[0], (R)ss[1], (R)ss[2]);
    m.st = m.ss + m.sa;
    Ray<R> ray{V3<R>((R)o[0], (R)o[1], (R)o[2]), V3<R>((R)d[0], (R)d[1], (R)d[2]), (R)0, (R)tmax};
    MediumRec<R> mi;
    R t = 0, pdf = 0;
    V3<R> w;
    if (kind == 0) w = medium_sample<R>(m, 0, ray, (R)u[0], (R)u[1], &mi, &t);
    else if (kind == 1) w = medium_equiangular<R>(m, 0, ray, (R)u[0], V3<R>((R)lp[0], (R)lp[1], (R)lp[2]), &mi, &t, &pdf);
    else w = medium_tr<R>(m, ray);
    out[0] = t;
    out[1] = mi.valid ? 1 : 0;
    out[2] = w.x; out[3] = w.y; out[4] = w.z;
    out[5] = pdf;
    return 0;
}
// Test hook: Scene::IntersectTr from o along d (through the Ray constructor)
// in the given medium: out = hit primitive (-1: none), Tr[3].
template <class R> static int vpt_tr_probe(Scene<R>& s, const double* o, const double* d, int32_t medium, double* out) {
    Ray<R> ray = make_ray(V3<R>((R)o[0], (R)o[1], (R)o[2]), V3<R>((R)d[0], (R)d[1], (R)d[2]), K<R>::Inf, (R)0);
    int32_t med = medium;
    Hit<R> h;
    V3<R> Tr;
    const bool hit = intersect_tr(s, ray, &med, &h, &Tr, true);
    out[0] = hit ? h.prim : -1;
    out[1] = Tr.x; out[2] = Tr.y; out[3] = Tr.z;
    return 0;
}
// Test hook: the hash grid of n points (all in use) with radii^2 r2; for each
// query point the bucket it looks up (point indices, duplicates kept).
static int sppm_grid_query_t(int32_t n, const float* pos, const float* r2, int32_t nq, const float* q, int32_t* counts,
                             int64_t* idx, int64_t cap) {
    SppmState<float> st;
    st.W = n;
    st.H = 1;
    st.used.assign(n, 1);
    st.r2.assign(r2, r2 + n);
    st.pos.resize(n);
    for (int32_t i = 0; i < n; ++i) st.pos[i] = V3<float>(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    SppmGrid<float> g;
    sppm_grid_build(st, &g);
    int64_t w = 0;
    for (int32_t k = 0; k < nq; ++k) {
        const auto& b = sppm_grid_bucket(g, V3<float>(q[3 * k], q[3 * k + 1], q[3 * k + 2]));
        counts[k] = (int32_t)b.size();
        if (!idx) continue;
        for (int64_t e : b) {
            if (w >= cap) return -1;
            idx[w++] = e;
        }
    }
    return 0;
}
extern "C" {

int or_load_sobol(const char* dir) {
    std::string d(dir);
    FILE* f = std::fopen((d + "/sobol_matrices.u32").c_str(), "rb");
    if (!f) return -1;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    g_sobol.resize(sz / 4);
    size_t got = std::fread(g_sobol.data(), 4, g_sobol.size(), f);
    std::fclose(f);
    if (got != g_sobol.size()) return -2;
    g_sobol_ndim = (uint32_t)(g_sobol.size() / 52);
    f = std::fopen((d + "/sobol_enum_vdc.u64").c_str(), "rb");
    if (!f) return -3;
    got = std::fread(g_vdc, 8, 25 * 52, f);
    std::fclose(f);
    if (got != 25 * 52) return -4;
    f = std::fopen((d + "/sobol_enum_vdc_inv.u64").c_str(), "rb");
    if (!f) return -5;
    got = std::fread(g_vdc_inv, 8, 26 * 52, f);
    std::fclose(f);
    if (got != 26 * 52) return -6;
    g_sobol_ok = true;
    return (int)g_sobol_ndim;
}

float or_sobol_sample(uint64_t index, uint32_t dim) { return sobol_sample(index, dim); }
uint64_t or_sobol_enum_index(uint32_t m, uint32_t frame, uint32_t px, uint32_t py) { return enum_look_up(m, frame, px, py); }

int or_camera_matrix(const or_camera_desc* cam, float* out) {
    Camera<float> c = make_camera<float>(*cam);
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) out[j * 4 + i] = c.raster_to_world.c[j].v[i];
    return 0;
}

int or_camera_rays(const or_camera_desc* cam, int32_t n, const int32_t* px, const int32_t* py, const float* u,
                   const float* v, float* org, float* dir) {
    Camera<float> c = make_camera<float>(*cam);
    for (int32_t i = 0; i < n; ++i) {
        Ray<float> r = generate_ray(c, px[i], py[i], u[i], v[i]);
        org[3 * i] = r.o.x; org[3 * i + 1] = r.o.y; org[3 * i + 2] = r.o.z;
        dir[3 * i] = r.d.x; dir[3 * i + 1] = r.d.y; dir[3 * i + 2] = r.d.z;
    }
    return 0;
}

void* or_scene_create(const or_scene_desc* desc) {
    SceneHandle* h = new SceneHandle();
    const or_scene_desc& d = *desc;
    int n = d.n_tri;
    h->pos.assign(d.pos, d.pos + 9 * (size_t)n);
    h->nrm.assign(d.nrm, d.nrm + 9 * (size_t)n);
    h->fn.assign(d.face_n, d.face_n + 3 * (size_t)n);
    if (d.uv) h->uv.assign(d.uv, d.uv + 6 * (size_t)n);
    h->mat.assign(d.mat, d.mat + n);
    if (d.light) h->light.assign(d.light, d.light + n);
    h->mats.assign(d.mats, d.mats + OR_MAT_WORDS * (size_t)d.n_mat);
    h->ltri.assign(d.light_tri, d.light_tri + d.n_light);
    h->le.assign(d.light_le, d.light_le + 3 * (size_t)d.n_light);
    size_t ntexf = 0;
    for (int i = 0; i < d.n_tex; ++i)
        ntexf = std::max(ntexf, (size_t)d.tex_info[4 * i + 3] + (size_t)d.tex_info[4 * i] * d.tex_info[4 * i + 1] * d.tex_info[4 * i + 2]);
    if (d.n_tex) {
        h->texi.assign(d.tex_info, d.tex_info + 4 * d.n_tex);
        h->texd.assign(d.tex_data, d.tex_data + ntexf);
    }
    h->desc = d;
    h->desc.pos = h->pos.data();
    h->desc.nrm = h->nrm.data();
    h->desc.face_n = h->fn.data();
    h->desc.uv = d.uv ? h->uv.data() : nullptr;
    h->desc.mat = h->mat.data();
    h->desc.light = d.light ? h->light.data() : nullptr;
    h->desc.mats = h->mats.data();
    h->desc.light_tri = h->ltri.data();
    h->desc.light_le = h->le.data();
    h->desc.tex_info = d.n_tex ? h->texi.data() : nullptr;
    h->desc.tex_data = d.n_tex ? h->texd.data() : nullptr;
    fill_scene(h->sf, h->desc);
    return h;
}

void or_scene_destroy(void* s) { delete (SceneHandle*)s; }

int or_scene_set_environment(void* s, const float* rgb, int32_t w, int32_t h, float scale) {
    SceneHandle* sh = (SceneHandle*)s;
    if (!rgb || w <= 0 || h <= 0) return -1;
    std::lock_guard<std::mutex> g(sh->mu);
    sh->env.assign(rgb, rgb + (size_t)w * h * 3);
    sh->env_w = w;
    sh->env_h = h;
    sh->env_scale = scale;
    attach_env(sh->sf, sh);
    if (sh->sd) attach_env(*sh->sd, sh);
    return 0;
}

// EnvironmentMap queries for tests (float mode): radiance at (u, v), sample at
// (u0, u1) -> dir + pdf, pdf of a direction, Le of a direction.
int or_env_query(void* s, int32_t n, const float* a, const float* b, const float* c, float* out_rgb, float* out_dir,
                 float* out_pdf, int kind) {
    SceneHandle* sh = (SceneHandle*)s;
    const Env<float>& e = sh->sf.env;
    if (e.w == 0) return -1;
    for (int32_t i = 0; i < n; ++i) {
        V3<float> L, d;
        float pdf = 0;
        if (kind == 0) {
            L = env_radiance(e, a[i], b[i]);
        } else if (kind == 1) {
            L = env_sample(e, a[i], b[i], &d, &pdf);
        } else if (kind == 2) {
            pdf = env_pdf_dir(e, V3<float>(a[i], b[i], c[i]));
        } else {
            L = env_emission(e, V3<float>(a[i], b[i], c[i]));
        }
        if (out_rgb) { out_rgb[3 * i] = L.x; out_rgb[3 * i + 1] = L.y; out_rgb[3 * i + 2] = L.z; }
        if (out_dir) { out_dir[3 * i] = d.x; out_dir[3 * i + 1] = d.y; out_dir[3 * i + 2] = d.z; }
        if (out_pdf) out_pdf[i] = pdf;
    }
    return 0;
}
int or_env_info(void* s, float* radius, float* light_func, int32_t max_lights) {
    SceneHandle* sh = (SceneHandle*)s;
    *radius = sh->sf.env.radius;
    const int n = (int)sh->sf.light_func.size();
    for (int i = 0; i < n && i < max_lights; ++i) light_func[i] = sh->sf.light_func[i];
    return n;
}

int or_bvh_stats(void* s, int64_t* n_nodes, int32_t* max_depth) {
    SceneHandle* h = (SceneHandle*)s;
    *n_nodes = (int64_t)h->sf.nodes.size();
    *max_depth = h->sf.max_depth;
    return 0;
}

int or_trace_closest(void* s, int32_t n, const float* org, const float* dir, const float* tmin, const float* tmax,
                     int use_bvh, float* out_t, int32_t* out_prim, float* out_b1, float* out_b2) {
    SceneHandle* h = (SceneHandle*)s;
    for (int32_t i = 0; i < n; ++i) {
        Ray<float> r{V3<float>(org[3 * i], org[3 * i + 1], org[3 * i + 2]), V3<float>(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]),
                     tmin[i], tmax[i]};
        Hit<float> hit;
        bool ok = intersect(h->sf, r, &hit, use_bvh != 0);
        out_prim[i] = ok ? hit.prim : -1;
        out_t[i] = ok ? hit.t : 0.f;
        if (out_b1) out_b1[i] = ok ? hit.b1 : 0.f;
        if (out_b2) out_b2[i] = ok ? hit.b2 : 0.f;
    }
    return 0;
}

int or_trace_any(void* s, int32_t n, const float* org, const float* dir, const float* tmin, const float* tmax, int use_bvh,
                 int32_t* out_hit) {
    SceneHandle* h = (SceneHandle*)s;
    for (int32_t i = 0; i < n; ++i) {
        Ray<float> r{V3<float>(org[3 * i], org[3 * i + 1], org[3 * i + 2]), V3<float>(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]),
                     tmin[i], tmax[i]};
        out_hit[i] = intersect_any(h->sf, r, use_bvh != 0) ? 1 : 0;
    }
    return 0;
}

int or_trace_counts(void* s, int32_t n, const float* org, const float* dir, const float* tmin, const float* tmax,
                    int64_t* n_node, int64_t* n_tri) {
    SceneHandle* h = (SceneHandle*)s;
    *n_node = 0;
    *n_tri = 0;
    for (int32_t i = 0; i < n; ++i) {
        Ray<float> r{V3<float>(org[3 * i], org[3 * i + 1], org[3 * i + 2]), V3<float>(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]),
                     tmin[i], tmax[i]};
        Hit<float> hit;
        intersect(h->sf, r, &hit, true, n_node, n_tri);
    }
    return 0;
}

int or_li_samples(void* s, const or_camera_desc* cam, const or_render_params* p, int32_t n, const int32_t* x, const int32_t* y,
                  const int32_t* smp, double* out_L, float* out_raster, int32_t* out_dims) {
    if (!g_sobol_ok) return -1;
    SceneHandle* h = (SceneHandle*)s;
    if (p->precision == 1) return li_samples_t<double>(h, cam, p, n, x, y, smp, out_L, out_raster, out_dims);
    return li_samples_t<float>(h, cam, p, n, x, y, smp, out_L, out_raster, out_dims);
}

// Test-only: the oracle's path for one camera sample, one row of kRecWords
// doubles per bounce: ray origin, direction, tmin, tmax, hit primitive (-1
// miss), t, sample dimensions used before the bounce, throughput, material id,
// delta flag.  Returns the number of rows written.
int or_li_path(void* hv, const or_camera_desc* cam, const or_render_params* p, int x, int y, int smp, int max_rows,
               double* rows) {
    t_rec = rows;
    t_rec_max = max_rows;
    t_rec_n = 0;
    double L[3];
    float raster[2];
    int32_t dims;
    or_render_params q = *p;
    q.nthreads = 1;
    int32_t xs = x, ys = y, ss = smp;
    or_li_samples(hv, cam, &q, 1, &xs, &ys, &ss, L, raster, &dims);
    const int n = t_rec_n;
    t_rec = nullptr;
    t_rec_max = t_rec_n = 0;
    return n;
}
float or_halton_sample(uint64_t index, uint32_t dim) { return hal::sample(dim, index); }
uint64_t or_halton_enum_index(uint32_t w, uint32_t h, uint64_t i, uint32_t x, uint32_t y) { return hal::Enum(w, h).index(i, x, y); }
int or_faure_perm(uint32_t base, uint16_t* out) {
    const hal::Tables& T = hal::tables();
    if (base < 1 || base >= T.perm.size()) return -1;
    for (uint32_t i = 0; i < base; ++i) out[i] = T.perm[base][i];
    return 0;
}
int or_scene_set_spheres(void* hv, int32_t n, const or_sphere* spheres) {
    SceneHandle* h = (SceneHandle*)hv;
    if (n < 0 || (n > 0 && !spheres) || !h->env.empty()) return -1;
    add_spheres(h->sf, n, spheres);
    add_spheres(scene_of<double>(h), n, spheres);
    return 0;
}

int or_sphere_light_probe(void* hv, int32_t i, int precision, const double* ref, const double* u, const double* wi,
                          double* out) {
    SceneHandle* h = (SceneHandle*)hv;
    return precision ? sphere_light_probe(scene_of<double>(h), i, ref, u, wi, out)
                     : sphere_light_probe(h->sf, i, ref, u, wi, out);
}

int or_scene_set_media(void* hv, int32_t n_media, const or_medium* media, int32_t n_tri, const or_tri_medium* tri) {
    SceneHandle* h = (SceneHandle*)hv;
    if (n_media < 0 || (n_media > 0 && !media) || n_tri != (int32_t)h->sf.tris.size() || !tri) return -1;
    for (int32_t i = 0; i < n_tri; ++i)
        if (tri[i].inside < -1 || tri[i].inside >= n_media || tri[i].outside < -1 || tri[i].outside >= n_media) return -1;
    h->media.assign(media, media + n_media);
    h->tri_media.assign(tri, tri + n_tri);
    attach_media(h->sf, h);
    if (h->sd) attach_media(*h->sd, h);
    return 0;
}

int or_hg_probe(int precision, double g, const double* wo, const double* u, double* wi, double* p) {
    return precision ? hg_probe_t<double>(g, wo, u, wi, p) : hg_probe_t<float>(g, wo, u, wi, p);
}
int or_medium_probe(int precision, int kind, const double* sigma_a, const double* sigma_s, const double* o, const double* d,
                    double tmax, const double* u, const double* light_pos, double* out) {
    return precision ? medium_probe_t<double>(kind, sigma_a, sigma_s, o, d, tmax, u, light_pos, out)
                     : medium_probe_t<float>(kind, sigma_a, sigma_s, o, d, tmax, u, light_pos, out);
}

int or_set_detached(void* hv, const float* tex_eval, const float* le_eval) {
    SceneHandle* h = (SceneHandle*)hv;
    set_detached(h->sf, h->desc, tex_eval, le_eval);
    set_detached(scene_of<double>(h), h->desc, tex_eval, le_eval);
    return 0;
}

int or_vpt_intersect_tr(void* hv, int precision, const double* o, const double* d, int32_t medium, double* out) {
    SceneHandle* h = (SceneHandle*)hv;
    return precision ? vpt_tr_probe(scene_of<double>(h), o, d, medium, out) : vpt_tr_probe(h->sf, o, d, medium, out);
}

int or_render_sppm(void* hv, const or_camera_desc* cd, const or_sppm_params* p, double* color, double* radius2, double* photons) {
    SceneHandle* h = (SceneHandle*)hv;
    if (!p || p->iterations <= 0 || p->photons <= 0 || p->max_depth <= 0) return -1;
    if (!h->env.empty() || !h->sf.sph.empty() || h->sf.light_tri.empty()) return -2;  // area lights, triangles
    if (p->precision) sppm_render<double>(h, cd, p, color, radius2, photons);
    else sppm_render<float>(h, cd, p, color, radius2, photons);
    return 0;
}
int or_sppm_grid_query(int32_t n, const float* pos, const float* r2, int32_t nq, const float* q, int32_t* counts, int64_t* idx,
                       int64_t cap) {
    return sppm_grid_query_t(n, pos, r2, nq, q, counts, idx, cap);
}

int or_set_vpt_query_crossings(int on) {
    g_vpt_query_crossings = on;
    return 0;
}

int or_set_shadow_short(int on) {
    g_shadow_short = on;
    return 0;
}

// LightTracing::Render (integrator/LT.h:17-53) in float mode, one thread,
// rows in order: splat (W x H x 3) receives Film::AddSplat's per-pixel sums.
int or_render_lt(void* hv, const or_camera_desc* cd, const or_render_params* p, float* splat) {
    SceneHandle* h = (SceneHandle*)hv;
    const Scene<float>& sc = h->sf;
    if (sc.light_tri.empty()) return -1;
    if (!sc.sph.empty()) return -2;  // light tracing of sphere scenes is not restated
    LTCamera<float> cam = make_lt_camera<float>(*cd);
    const int W = cd->width, H = cd->height;
    const int64_t total = (int64_t)p->spp * W * H;
    const size_t npx = (size_t)W * H * 3;
    for (size_t i = 0; i < npx; ++i) splat[i] = 0.f;
    // ParallelFor over image rows (integrator/LT.h:25-47): thread t takes rows
    // t, t + n, ...  and adds its splats into a buffer of its own; the buffers
    // are then summed in thread order (one thread: straight into splat).
    int nth = p->nthreads > 0 ? p->nthreads : (int)std::thread::hardware_concurrency();
    nth = std::max(1, std::min(nth, H));
    std::vector<std::vector<float>> bufs(nth - 1, std::vector<float>(npx, 0.f));
    auto worker = [&](int t) {
        float* buf = t == 0 ? splat : bufs[t - 1].data();
        auto add = [&](float x, float y, const V3<float>& L) {
            int X = (int)std::floor(x), Y = (int)std::floor(y);
            X = X < 0 ? 0 : (X > W - 1 ? W - 1 : X);
            Y = Y < 0 ? 0 : (Y > H - 1 ? H - 1 : Y);
            float* px = buf + ((size_t)Y * W + X) * 3;
            px[0] = px[0] + L.x;
            px[1] = px[1] + L.y;
            px[2] = px[2] + L.z;
        };
        for (int y = t; y < H; y += nth)
            for (int x = 0; x < W; ++x)
                for (int smp = 0; smp < p->spp; ++smp) {
                    Seq rnd;
                    rnd.kind = p->sampler == 1 ? 1 : 0;
                    rnd.instance = (uint64_t)W * H * smp + (uint64_t)y * W + x;
                    lt_path(sc, cam, rnd, p->max_depth, p->use_bvh != 0, total, add);
                }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nth; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto& x : th) x.join();
    for (int t = 1; t < nth; ++t)
        for (size_t i = 0; i < npx; ++i) splat[i] = splat[i] + bufs[t - 1][i];
    return 0;
}

int64_t or_render(void* s, const or_camera_desc* cam, const or_render_params* p, int32_t row0, int32_t row1, double* color,
                  double* weight, double* seconds) {
    if (!g_sobol_ok) return -1;
    SceneHandle* h = (SceneHandle*)s;
    if (p->precision == 1) return render_t<double>(h, cam, p, row0, row1, color, weight, seconds);
    return render_t<float>(h, cam, p, row0, row1, color, weight, seconds);
}

int or_mat4_inverse(const float* rows16, float* out_rows16) {
    M4<float> m = M4<float>::rows(rows16[0], rows16[1], rows16[2], rows16[3], rows16[4], rows16[5], rows16[6], rows16[7],
                                  rows16[8], rows16[9], rows16[10], rows16[11], rows16[12], rows16[13], rows16[14], rows16[15]);
    M4<float> inv = glm_inverse(m);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out_rows16[i * 4 + j] = inv.c[j].v[i];
    return 0;
}

int or_mat4_mul_vec(const float* rows16, const float* v4, float* out4) {
    M4<float> m = M4<float>::rows(rows16[0], rows16[1], rows16[2], rows16[3], rows16[4], rows16[5], rows16[6], rows16[7],
                                  rows16[8], rows16[9], rows16[10], rows16[11], rows16[12], rows16[13], rows16[14], rows16[15]);
    V4<float> a{{v4[0], v4[1], v4[2], v4[3]}};
    V4<float> r = mul(m, a);
    for (int i = 0; i < 4; ++i) out4[i] = r.v[i];
    return 0;
}

// BSDF of one material record in the local shading frame (X, Y, Z axes, ng = Z).
static BSDF<double> local_bsdf(const uint32_t* w) {
    Material m;
    auto f = [&](int k) { float v; std::memcpy(&v, w + k, 4); return v; };
    BSDF<double> b;
    b.type = (int)w[0];
    b.ng = V3<double>(0, 0, 1);
    b.X = V3<double>(1, 0, 0);
    b.Y = V3<double>(0, 1, 0);
    b.Z = V3<double>(0, 0, 1);
    b.Rc = V3<double>(f(2), f(3), f(4));
    b.Tc = V3<double>(f(5), f(6), f(7));
    double ax = f(8), ay = f(9);
    if (!(w[1] & 1u)) { ax = roughness_to_alpha(ax); ay = roughness_to_alpha(ay); }
    b.dist.ax = ax;
    b.dist.ay = ay;
    b.fresnel_schlick = (w[1] & 2u) ? 1 : 0;
    b.eta = V3<double>(f(10), f(11), f(12));
    b.k = V3<double>(f(13), f(14), f(15));
    b.f0 = V3<double>(f(16), f(17), f(18));
    b.eta1 = f(19);
    b.eta2 = f(20);
    (void)m;
    return b;
}

// Evaluate f (3) and pdf for n direction pairs (wo, wi as 3n doubles each), in double.
int or_bsdf_eval(const uint32_t* mat_words, int32_t n, const double* wo, const double* wi, double* f, double* pdf) {
    BSDF<double> b = local_bsdf(mat_words);
    for (int32_t i = 0; i < n; ++i) {
        V3<double> o(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), d(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        V3<double> v = b.Evaluate(o, d);
        f[3 * i] = v.x; f[3 * i + 1] = v.y; f[3 * i + 2] = v.z;
        pdf[i] = b.Pdf(o, d);
    }
    return 0;
}

// Sample n directions for wo with random numbers u (3n); outputs wi (3n), f (3n), pdf (n).
int or_bsdf_sample_mode(const uint32_t* mat_words, int32_t n, const double* wo, const double* u, double* wi, double* f,
                        double* pdf, int importance);
int or_bsdf_sample(const uint32_t* mat_words, int32_t n, const double* wo, const double* u, double* wi, double* f, double* pdf) {
    return or_bsdf_sample_mode(mat_words, n, wo, u, wi, f, pdf, 0);
}
int or_bsdf_sample_mode(const uint32_t* mat_words, int32_t n, const double* wo, const double* u, double* wi, double* f,
                        double* pdf, int importance) {
    BSDF<double> b = local_bsdf(mat_words);
    b.importance = importance != 0;
    for (int32_t i = 0; i < n; ++i) {
        V3<double> o(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), w;
        double p;
        V3<double> v = b.Sample(o, &w, &p, V3<double>(u[3 * i], u[3 * i + 1], u[3 * i + 2]));
        wi[3 * i] = w.x; wi[3 * i + 1] = w.y; wi[3 * i + 2] = w.z;
        f[3 * i] = v.x; f[3 * i + 1] = v.y; f[3 * i + 2] = v.z;
        pdf[i] = p;
    }
    return 0;
}

double or_fr_dielectric(double cosI, double etaI, double etaT) { return fr_dielectric(cosI, etaI, etaT); }
int or_fr_conductor(double cosI, const double* eta, const double* k, double* out) {
    V3<double> r = fr_conductor(cosI, V3<double>(1, 1, 1), V3<double>(eta[0], eta[1], eta[2]), V3<double>(k[0], k[1], k[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
    return 0;
}
double or_roughness_to_alpha(double r) { return roughness_to_alpha(r); }
double or_ggx_D(double ax, double ay, const double* m) {
    GGX<double> g{ax, ay};
    return g.D(V3<double>(m[0], m[1], m[2]));
}
// ImageTexture bilinear lookup in FP32 (texture/ImageTexture.h:50-69).
int or_tex_lookup(int w, int h, int c, const float* data, int n, const float* u, const float* v, int ch, float* out) {
    Texture<float> t;
    t.w = w;
    t.h = h;
    t.c = c;
    t.data.assign(data, data + (size_t)w * h * c);
    for (int i = 0; i < n; ++i) out[i] = tex_bilinear(t, u[i], v[i], ch);
    return 0;
}

}  // extern "C"

// Pinhole camera set-up on the host: the RasterToWorld matrix the raygen
// kernels apply.  It is formed in FP32 with the reference's operation order so
// that device camera rays match the reference's bit for bit:
//   ProjectiveCamera::Initialize   camera/Camera.cpp:6-38 (lookAt constructor :68-72)
//   Transform::Perspective         math/Transform.cpp:132-150
//   Transform::Scale / Translate   math/Transform.cpp:6-28, Transform::operator* Transform.h:253-255
//   Matrix4 * Vector4 (SSE order)  math/Linagl.h:937-948
//   Inversed (GLM 4x4 inverse)     math/Linagl.h:1178-1293
#include <cmath>

#include "gy_internal.h"

namespace gy {
namespace {

struct M4 {
    float c[4][4];  // c[col][row]
};

M4 Rows(const float r[16]) {
    M4 m;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m.c[j][i] = r[i * 4 + j];
    return m;
}

// ((c3*w + c2*z) + c1*y) + c0*x, lane by lane.
void MulVec(const M4& m, const float a[4], float out[4]) {
    float r[4];
    for (int i = 0; i < 4; ++i) r[i] = m.c[3][i] * a[3];
    for (int i = 0; i < 4; ++i) r[i] = m.c[2][i] * a[2] + r[i];
    for (int i = 0; i < 4; ++i) r[i] = m.c[1][i] * a[1] + r[i];
    for (int i = 0; i < 4; ++i) r[i] = m.c[0][i] * a[0] + r[i];
    for (int i = 0; i < 4; ++i) out[i] = r[i];
}

M4 Mul(const M4& a, const M4& b) {
    M4 r;
    for (int j = 0; j < 4; ++j) MulVec(a, b.c[j], r.c[j]);
    return r;
}

M4 Inverse(const M4& M) {
    auto m = [&](int col, int row) { return M.c[col][row]; };
    const float c00 = m(2, 2) * m(3, 3) - m(3, 2) * m(2, 3), c02 = m(1, 2) * m(3, 3) - m(3, 2) * m(1, 3),
                c03 = m(1, 2) * m(2, 3) - m(2, 2) * m(1, 3), c04 = m(2, 1) * m(3, 3) - m(3, 1) * m(2, 3),
                c06 = m(1, 1) * m(3, 3) - m(3, 1) * m(1, 3), c07 = m(1, 1) * m(2, 3) - m(2, 1) * m(1, 3),
                c08 = m(2, 1) * m(3, 2) - m(3, 1) * m(2, 2), c10 = m(1, 1) * m(3, 2) - m(3, 1) * m(1, 2),
                c11 = m(1, 1) * m(2, 2) - m(2, 1) * m(1, 2), c12 = m(2, 0) * m(3, 3) - m(3, 0) * m(2, 3),
                c14 = m(1, 0) * m(3, 3) - m(3, 0) * m(1, 3), c15 = m(1, 0) * m(2, 3) - m(2, 0) * m(1, 3),
                c16 = m(2, 0) * m(3, 2) - m(3, 0) * m(2, 2), c18 = m(1, 0) * m(3, 2) - m(3, 0) * m(1, 2),
                c19 = m(1, 0) * m(2, 2) - m(2, 0) * m(1, 2), c20 = m(2, 0) * m(3, 1) - m(3, 0) * m(2, 1),
                c22 = m(1, 0) * m(3, 1) - m(3, 0) * m(1, 1), c23 = m(1, 0) * m(2, 1) - m(2, 0) * m(1, 1);
    const float f0[4] = {c00, c00, c02, c03}, f1[4] = {c04, c04, c06, c07}, f2[4] = {c08, c08, c10, c11};
    const float f3[4] = {c12, c12, c14, c15}, f4[4] = {c16, c16, c18, c19}, f5[4] = {c20, c20, c22, c23};
    const float v0[4] = {m(1, 0), m(0, 0), m(0, 0), m(0, 0)}, v1[4] = {m(1, 1), m(0, 1), m(0, 1), m(0, 1)};
    const float v2[4] = {m(1, 2), m(0, 2), m(0, 2), m(0, 2)}, v3[4] = {m(1, 3), m(0, 3), m(0, 3), m(0, 3)};
    const float sa[4] = {+1, -1, +1, -1}, sb[4] = {-1, +1, -1, +1};
    M4 I;
    for (int k = 0; k < 4; ++k) {
        I.c[0][k] = ((v1[k] * f0[k] - v2[k] * f1[k]) + v3[k] * f2[k]) * sa[k];
        I.c[1][k] = ((v0[k] * f0[k] - v2[k] * f3[k]) + v3[k] * f4[k]) * sb[k];
        I.c[2][k] = ((v0[k] * f1[k] - v1[k] * f3[k]) + v3[k] * f5[k]) * sa[k];
        I.c[3][k] = ((v0[k] * f2[k] - v1[k] * f4[k]) + v2[k] * f5[k]) * sb[k];
    }
    float d0[4];
    const float row0[4] = {I.c[0][0], I.c[1][0], I.c[2][0], I.c[3][0]};
    for (int k = 0; k < 4; ++k) d0[k] = M.c[0][k] * row0[k];
    const float det = (d0[0] + d0[1]) + (d0[2] + d0[3]);
    const float inv = 1.0f / det;
    for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) I.c[j][k] = I.c[j][k] * inv;
    return I;
}

struct V3 {
    float x, y, z;
};
inline V3 Sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 Neg(V3 a) { return {0.f - a.x, 0.f - a.y, 0.f - a.z}; }
inline float Dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + (a.z * b.z + 0.f); }
inline V3 Cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 Norm(V3 a) {
    const float l = std::sqrt(Dot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}

}  // namespace

int BuildCamera(const gy_camera_desc& d, CameraRT* out) {
    if (d.width <= 0 || d.height <= 0 || !(d.fovy > 0.f && d.fovy < 180.f) || !(d.film_dist > 0.f))
        return Fail(GY_ERR_INVALID, "camera: width/height must be positive, 0 < fovy < 180, film_dist > 0");
    const V3 pos{d.pos[0], d.pos[1], d.pos[2]}, look{d.look_at[0], d.look_at[1], d.look_at[2]},
        up{d.up[0], d.up[1], d.up[2]};
    const V3 cz = Neg(Norm(Sub(pos, look)));
    const V3 cx = Neg(Norm(Cross(up, cz)));
    const V3 cy = Neg(Cross(cz, cx));
    const float c2w_r[16] = {cx.x, cy.x, cz.x, pos.x, cx.y, cy.y, cz.y, pos.y, cx.z, cy.z, cz.z, pos.z, 0, 0, 0, 1.f};
    const M4 c2w = Rows(c2w_r);
    const float aspect = (float)d.width / (float)d.height;
    const float PI = 3.14159265358979f;
    const float radians = (PI / 180) * d.fovy;
    const float dis = d.film_dist;
    const float t = dis * std::tan(radians * 0.5);  // double-precision tan, as written
    const float r = t * aspect;
    const float f = 100000.f, n = dis;
    const float persp_r[16] = {dis / r, 0, 0, 0, 0, dis / t, 0, 0, 0, 0, (f + n) / (f - n), -2 * f * n / (f - n), 0, 0, 1, 0};
    const M4 ndc_to_camera = Inverse(Rows(persp_r));
    // NDCToRaster = Scale(W, H, 1) * Scale(0.5, -0.5, 1) * Translate(1, -1, 0); its
    // inverse is composed from the factors' stored inverses.
    const float W = (float)d.width, H = (float)d.height;
    const float s1inv_r[16] = {1 / W, 0, 0, 0, 0, 1 / H, 0, 0, 0, 0, 1 / 1.f, 0, 0, 0, 0, 1};
    const float s2inv_r[16] = {1 / 0.5f, 0, 0, 0, 0, 1 / -0.5f, 0, 0, 0, 0, 1 / 1.f, 0, 0, 0, 0, 1};
    const float tinv_r[16] = {1, 0, 0, -1.f, 0, 1, 0, -(-1.f), 0, 0, 1, -0.f, 0, 0, 0, 1};
    const M4 raster_to_ndc = Mul(Rows(tinv_r), Mul(Rows(s2inv_r), Rows(s1inv_r)));
    const M4 raster_to_camera = Mul(ndc_to_camera, raster_to_ndc);
    const M4 rtw = Mul(c2w, raster_to_camera);
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) out->raster_to_world[j * 4 + i] = rtw.c[j][i];
    out->pos[0] = pos.x;
    out->pos[1] = pos.y;
    out->pos[2] = pos.z;
    out->width = d.width;
    out->height = d.height;
    int res = 1;
    while (res < (d.width > d.height ? d.width : d.height)) res <<= 1;
    out->res = res;
    int lg = 0;
    while ((1 << lg) < res) ++lg;
    out->log2res = lg;
    return GY_OK;
}

}  // namespace gy

extern "C" int gy_camera_matrix(const gy_camera_desc* cam, float* out16) {
    if (!cam || !out16) return gy::Fail(GY_ERR_INVALID, "gy_camera_matrix: null argument");
    gy::CameraRT rt;
    int rc = gy::BuildCamera(*cam, &rt);
    if (rc != GY_OK) return rc;
    for (int i = 0; i < 16; ++i) out16[i] = rt.raster_to_world[i];
    return GY_OK;
}

// FP32 host math in the reference's operation order, shared by the camera
// set-up and the mesh transforms:
//   Matrix4 * Vector4 (SSE order)  math/Linagl.h:937-948
//   Inversed (GLM 4x4 inverse)     math/Linagl.h:1178-1293
//   Dot (DPPS order), Cross, Norm  math/Linagl.h:496-500, 527-531, 564-567
// Every function is host-only and built without FMA contraction.
#pragma once
#include <cmath>

namespace gy {
namespace hm {

struct M4 {
    float c[4][4];  // c[col][row]
};

inline M4 Rows(const float r[16]) {
    M4 m;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) m.c[j][i] = r[i * 4 + j];
    return m;
}

// ((c3*w + c2*z) + c1*y) + c0*x, lane by lane.
inline void MulVec(const M4& m, const float a[4], float out[4]) {
    float r[4];
    for (int i = 0; i < 4; ++i) r[i] = m.c[3][i] * a[3];
    for (int i = 0; i < 4; ++i) r[i] = m.c[2][i] * a[2] + r[i];
    for (int i = 0; i < 4; ++i) r[i] = m.c[1][i] * a[1] + r[i];
    for (int i = 0; i < 4; ++i) r[i] = m.c[0][i] * a[0] + r[i];
    for (int i = 0; i < 4; ++i) out[i] = r[i];
}

inline M4 Mul(const M4& a, const M4& b) {
    M4 r;
    for (int j = 0; j < 4; ++j) MulVec(a, b.c[j], r.c[j]);
    return r;
}

inline M4 Inverse(const M4& M) {
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

}  // namespace hm
}  // namespace gy

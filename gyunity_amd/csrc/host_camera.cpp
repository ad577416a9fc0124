// Pinhole camera set-up on the host: the RasterToWorld matrix the raygen
// kernels apply.  It is formed in FP32 with the reference's operation order so
// that device camera rays match the reference's bit for bit:
//   ProjectiveCamera::Initialize   camera/Camera.cpp:6-38 (lookAt constructor :68-72)
//   Transform::Perspective         math/Transform.cpp:132-150
//   Transform::Scale / Translate   math/Transform.cpp:6-28, Transform::operator* Transform.h:253-255
//   Matrix4 * Vector4, Inversed    host_math.h
#include <cmath>

#include "gy_internal.h"
#include "host_math.h"

namespace gy {
using namespace hm;

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

int BuildLightCamera(const gy_camera_desc& d, LightCameraRT* out) {
    CameraRT rt;
    int rc = BuildCamera(d, &rt);  // validates the description
    if (rc != GY_OK) return rc;
    const V3 pos{d.pos[0], d.pos[1], d.pos[2]}, look{d.look_at[0], d.look_at[1], d.look_at[2]},
        up{d.up[0], d.up[1], d.up[2]};
    const V3 cz = Neg(Norm(Sub(pos, look)));
    const V3 cx = Neg(Norm(Cross(up, cz)));
    const V3 cy = Neg(Cross(cz, cx));
    const float c2w_r[16] = {cx.x, cy.x, cz.x, pos.x, cx.y, cy.y, cz.y, pos.y, cx.z, cy.z, cz.z, pos.z, 0, 0, 0, 1.f};
    const float aspect = (float)d.width / (float)d.height;
    const float PI = 3.14159265358979f;
    const float radians = (PI / 180) * d.fovy;
    const float dis = d.film_dist;
    const float t = dis * std::tan(radians * 0.5);
    const float r = t * aspect;
    const float f = 100000.f, n = dis;
    const float persp_r[16] = {dis / r, 0, 0, 0, 0, dis / t, 0, 0, 0, 0, (f + n) / (f - n), -2 * f * n / (f - n), 0, 0, 1, 0};
    // NDCToRaster = Scale(W, H, 1) * Scale(0.5, -0.5, 1) * Translate(1, -1, 0)
    const float W = (float)d.width, H = (float)d.height;
    const float s1_r[16] = {W, 0, 0, 0, 0, H, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const float s2_r[16] = {0.5f, 0, 0, 0, 0, -0.5f, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    const float tr_r[16] = {1, 0, 0, 1, 0, 1, 0, -1, 0, 0, 1, 0, 0, 0, 0, 1};
    const M4 ndc_to_raster = Mul(Mul(Rows(s1_r), Rows(s2_r)), Rows(tr_r));
    const M4 cam_to_raster = Mul(ndc_to_raster, Rows(persp_r));
    const M4 w2r = Mul(cam_to_raster, Inverse(Rows(c2w_r)));
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) out->world_to_raster[j * 4 + i] = w2r.c[j][i];
    out->pos[0] = pos.x;
    out->pos[1] = pos.y;
    out->pos[2] = pos.z;
    out->cz[0] = cz.x;
    out->cz[1] = cz.y;
    out->cz[2] = cz.z;
    out->dis = dis;
    // mHeight = d * tan(fovy * 0.5 * PI / 180) * 2 (tan evaluated in double), mWidth = mHeight * aspect
    const float height = dis * (float)std::tan((double)(d.fovy * 0.5f * PI / 180)) * 2.f;
    const float width = height * aspect;
    out->area = width * height;
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

"""The path-replay adjoint against finite differences of the oracle (FP32 and FP64).

Replay holds every sampling decision of a path fixed (directions, pdfs, MIS
weights, light choice) and differentiates the values the estimator takes.
The oracle's detached mode (OracleScene.set_detached) evaluates exactly that
estimator at other parameter values on the same paths, so its central
differences are the derivative replay estimates: texel albedo, GGX roughness
and each light's emitted radiance.  In FP32 (the GPU's estimator) all agree
within 1e-2; in FP64 albedo and roughness within 1e-2 and emission within
2e-2 (a few paths branch differently in FP64).

CPU tests pin the detached mode itself: with the scene's own values it is the
ordinary render, it is linear in Le with the light choice fixed, and for
albedo (which no sampling decision reads) it equals ordinary differences.
"""
import numpy as np
import pytest

import gyunity_amd as g
from oracle import oracle as O


def _setup(tex_res=16, res=32, spp=16, seed=7):
    sa, cam, prm = g.generate_scene(g.GY_SCENE_ADJOINT, 0, tex_res)
    cam.width = cam.height = res
    prm.spp, prm.rr_depth, prm.mode = spp, 100, g.GY_RENDER_WAVEFRONT
    gt = sa.tex_data.copy()
    n_alb = int(sa.tex_info[5, 3])  # the five wall textures come first, roughness last
    rng = np.random.default_rng(seed)
    theta = gt.copy()
    theta[:n_alb] = np.clip(gt[:n_alb] + rng.uniform(-0.2, 0.2, n_alb), 0.05, 0.95)
    theta[n_alb:] = np.clip(gt[n_alb:] + rng.uniform(-0.2, 0.2, gt.size - n_alb), 0.05, 0.8)
    return sa, cam, prm, gt, theta.astype(np.float32), n_alb


def _with_textures(sa, tex):
    s2 = g.SceneArrays(sa.pos, sa.nrm, sa.face_n, sa.mat, sa.light, sa.mats, sa.light_tri, sa.light_le, uv=sa.uv,
                       tex_info=sa.tex_info, tex_data=np.ascontiguousarray(tex, np.float32))
    return s2


def _image(c, w):
    return c / np.maximum(w, 1e-300)[..., None]


def _oracle_image(os_, cam, prm, prec=1):
    c, w, _, _ = os_.render(cam, O.params(prm.spp, prm.max_depth, prm.rr_depth, precision=prec, nthreads=8))
    return _image(c, w)


def test_detached_with_own_values_is_the_render():
    sa, cam, prm, gt, theta, n_alb = _setup()
    os_ = O.OracleScene(_with_textures(sa, theta))
    # the detached value is the lobe at the sampled directions taken back to
    # the local frame, so it differs from the sample's own value by rounding
    for prec, tol in ((0, 1e-6), (1, 1e-14)):
        ref = _oracle_image(os_, cam, prm, prec)
        os_.set_detached(theta, sa.light_le)
        got = _oracle_image(os_, cam, prm, prec)
        os_.set_detached()
        assert np.allclose(got, ref, rtol=tol, atol=0), prec


def test_detached_is_linear_in_emission():
    # Le enters the estimator linearly once the light choice is held fixed
    sa, cam, prm, gt, theta, n_alb = _setup()
    os_ = O.OracleScene(_with_textures(sa, theta))
    ref = _oracle_image(os_, cam, prm)
    os_.set_detached(None, 2.0 * sa.light_le)
    got = _oracle_image(os_, cam, prm)
    os_.set_detached()
    assert np.allclose(got, 2.0 * ref, rtol=1e-12, atol=0)


def test_detached_albedo_differences_equal_ordinary_differences():
    # albedo is read by no sampling decision: moving the paths' sampling
    # textures along changes nothing, so both differences agree
    sa, cam, prm, gt, theta, n_alb = _setup()
    os_ = O.OracleScene(_with_textures(sa, theta))
    k = int(np.argmax(theta[:n_alb]))
    h = 1e-3
    th_p, th_m = theta.copy(), theta.copy()
    th_p[k] += h
    th_m[k] -= h
    os_.set_detached(th_p)
    dp = _oracle_image(os_, cam, prm)
    os_.set_detached(th_m)
    dm = _oracle_image(os_, cam, prm)
    os_.set_detached()
    op = _oracle_image(O.OracleScene(_with_textures(sa, th_p)), cam, prm)
    om = _oracle_image(O.OracleScene(_with_textures(sa, th_m)), cam, prm)
    assert np.allclose(dp - dm, op - om, rtol=0, atol=1e-12)
    assert np.abs(dp - dm).max() > 0


def test_roughness_to_alpha_known_values_and_derivative():
    # RoughnessToAlpha (visual/Microfacet.h:49-54): the polynomial in
    # x = log r is 1.62142 at r = 1, and r is clamped to 1e-3 below.  The
    # library's d alpha / d r (adjoint of roughness textures) against central
    # differences of the oracle's FP64 value.
    import ctypes as C

    assert O.roughness_to_alpha(1.0) == pytest.approx(1.62142, abs=1e-6)
    assert O.roughness_to_alpha(1e-4) == O.roughness_to_alpha(1e-3)
    lib = g.lib()
    for r in np.geomspace(2e-3, 1.0, 40):
        h = 1e-6 * r
        fd = (O.roughness_to_alpha(r + h) - O.roughness_to_alpha(r - h)) / (2 * h)
        assert abs(lib.gy_dalpha_droughness(C.c_float(r)) - fd) <= 1e-4 * abs(fd), r
    assert lib.gy_dalpha_droughness(C.c_float(5e-4)) == 0.0


def test_ggx_dlogf_dalpha_matches_oracle_lobe():
    # d log f / d alpha of the GGX conductor lobe (replay's roughness term)
    # against central differences of the oracle's FP64 lobe value at alpha +- h.
    import ctypes as C

    from tests.scene_util import material

    rng = np.random.default_rng(4)
    lib = g.lib()
    n = 0
    for _ in range(300):
        a = float(rng.uniform(0.05, 0.8))
        wo = rng.normal(size=3)
        wi = rng.normal(size=3)
        wo[2], wi[2] = abs(wo[2]) + 0.05, abs(wi[2]) + 0.05
        wo /= np.linalg.norm(wo)
        wi /= np.linalg.norm(wi)
        h = 1e-3 * a
        fs = []
        for ah in (a + h, a - h):
            m = material(1, R=(0.8, 0.8, 0.8), alpha=(ah, ah), alpha_direct=True, f0=(0.9, 0.6, 0.5))
            f, _ = O.bsdf_eval(m, wo, wi)
            fs.append(f[0, 0])
        if min(fs) <= 0:
            continue
        fd = (np.log(fs[0]) - np.log(fs[1])) / (2 * h)
        wo32 = np.ascontiguousarray(wo, np.float32)
        wi32 = np.ascontiguousarray(wi, np.float32)
        got = lib.gy_ggx_dlogf_dalpha(C.c_float(a), wo32.ctypes.data_as(C.POINTER(C.c_float)),
                                      wi32.ctypes.data_as(C.POINTER(C.c_float)))
        assert abs(got - fd) <= 2e-3 * max(1.0, abs(fd)), (a, wo, wi, got, fd)
        n += 1
    assert n > 250


@pytest.mark.gpu
def test_adjoint_matches_oracle_detached_differences(gpu_ctx):
    # GPU replay gradients (FP32) of an L2 image loss against central
    # differences of the oracle's detached renders on the same sample indices:
    # 8 albedo texels, 8 roughness texels, every light's Le.
    # - FP32 (the oracle's bit-faithful mode, the GPU's estimator): all within 1e-2.
    # - FP64: albedo and roughness within 1e-2; Le within 2e-2.  A few paths
    #   branch differently in FP64 (56 of 3072 pixel values of this render move
    #   by more than 1e-4), which moves the image-wide emission gradient by up
    #   to 1.2%.
    # Target and dl/dI come from the oracle's images of the same precision.
    sa, cam, prm, gt, theta, n_alb = _setup()
    ot = O.OracleScene(_with_textures(sa, gt))
    target = {0: _oracle_image(ot, cam, prm, 0), 1: _oracle_image(ot, cam, prm, 1)}
    os_ = O.OracleScene(_with_textures(sa, theta))
    img32 = _oracle_image(os_, cam, prm, 0)
    dI = (2.0 * (img32 - target[0]) / img32.size).astype(np.float32)
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_textures(theta)
        sc.render(cam, prm)
        grad, gle = sc.adjoint(cam, prm, dI, emission=True)

    def loss(prec, tex=None, le=None):
        os_.set_detached(theta if tex is None else tex, sa.light_le if le is None else le)
        v = float(((_oracle_image(os_, cam, prm, prec) - target[prec]) ** 2).mean())
        os_.set_detached()
        return v

    def central(prec, k=None, h=None, light=None):
        if light is None:
            tp, tm = theta.astype(np.float64), theta.astype(np.float64)
            tp[k] += h
            tm[k] -= h
            return (loss(prec, tex=tp) - loss(prec, tex=tm)) / (2 * h)
        l, c = light
        lp, lm = sa.light_le.astype(np.float64), sa.light_le.astype(np.float64)
        lp[l, c] += h
        lm[l, c] -= h
        return (loss(prec, le=lp) - loss(prec, le=lm)) / (2 * h)

    ka = np.argsort(-np.abs(grad[:n_alb]))[:8]
    kr = n_alb + np.argsort(-np.abs(grad[n_alb:]))[:8]
    rows = []  # kind, index, replay, FD32, FD64
    for k in ka:
        rows.append(("albedo", int(k), float(grad[k]), central(0, int(k), 1e-2), central(1, int(k), 1e-3)))
    for k in kr:
        rows.append(("roughness", int(k), float(grad[k]), central(0, int(k), 1e-3), central(1, int(k), 1e-4)))
    for l in range(sa.light_tri.shape[0]):
        for c in range(3):
            le = float(sa.light_le[l, c])
            rows.append(("Le", 3 * l + c, float(gle[l, c]), central(0, h=1e-2 * le, light=(l, c)),
                         central(1, h=1e-3 * le, light=(l, c))))
    for kind, k, gr, f32, f64 in rows:
        print("%-9s %6d  replay %+.6e  FD32 %+.6e (rel %.1e)  FD64 %+.6e (rel %.1e)"
              % (kind, k, gr, f32, abs(gr - f32) / abs(f32), f64, abs(gr - f64) / abs(f64)))
    for kind, k, gr, f32, f64 in rows:
        assert abs(gr - f32) <= 1e-2 * abs(f32), (kind, k, gr, f32)
        assert abs(gr - f64) <= (2e-2 if kind == "Le" else 1e-2) * abs(f64), (kind, k, gr, f64)

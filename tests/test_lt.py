"""Light tracing: LightTracing (integrator/LT.h) through gy_render_light_tracing.

CPU tests pin the oracle's restatement (the reference build needs TBB and
MSVC):
- a scene holding only a light that faces the camera: every light path's own
  vertex reaches the film unoccluded and nothing else does, so the splat sum
  is Le * W * H exactly (up to float summation), which pins the light-vertex
  splat and the L * resX * resY / totalPhotons scale;
- light tracing and path tracing estimate the same image on a diffuse scene
  (PinHoleCamera::We / Sample_Wi against the camera ray's measurement): their
  means over interior blocks agree within the noise of 256 spp;
- light paths scatter in TransportMode::Importance: smooth-glass
  transmission weights f |cos| / pdf are Tc exactly (radiance mode: e^2 Tc),
  rough-dielectric transmission values are e^2 times the radiance mode's
  (bsdf/BSDF.h:266-271, 505, 608);
- an environment light's paths carry zero throughput: GenerateLightPath takes
  Le = Emission(), which EnvironmentLight leaves zero (light/EnvLight.h:20-22,
  integrator/LT.h:108-111), so with a map added the light vertices' splats
  sum to the area lights' share of the paths (their pdfLight) times Le W H.
GPU tests compare the device's splat image with the oracle's, for both
forms of the integrator (one light path per thread, and the wavefront that
traces every path's rays of one bounce together): each light path's splats
are the same values, their float-atomic summation order is not (tolerance
below).
"""
import numpy as np
import pytest

import gyunity_amd as g
from oracle import oracle as O
from tests.scene_util import env_image, material, mixed_material_scene

S_REL = 1e-4  # per pixel, relative to max(|L|, 1e-3 of the image's mean)
MODES = pytest.mark.parametrize("mode", [g.GY_RENDER_MEGAKERNEL, g.GY_RENDER_WAVEFRONT], ids=["megakernel", "wavefront"])


def _light_only(le=(3.0, 2.0, 1.0)):
    """One quad light at z = 0 facing +z; the camera looks at it from z = 2."""
    pos = np.array([[-0.3, -0.3, 0, 0.3, -0.3, 0, 0.3, 0.3, 0], [-0.3, -0.3, 0, 0.3, 0.3, 0, -0.3, 0.3, 0]], np.float32)
    fn = np.array([[0, 0, 1], [0, 0, 1]], np.float32)
    nrm = np.repeat(fn, 3, axis=0).reshape(-1, 9)
    sa = g.SceneArrays(pos, nrm, fn, np.zeros(2, np.uint32), np.array([0, 1], np.int32), material(0, R=(0, 0, 0))[None, :],
                       np.array([0, 1], np.int32), np.array([le, le], np.float32))
    cam = g.camera((0.0, 0.0, 2.0), (0.0, 0.0, 0.0), (0, 1, 0), 40.0, 40, 32)
    return sa, cam


def _cornell(res=32):
    sa, cam, prm = g.generate_scene(g.GY_SCENE_CORNELL)
    cam.width = cam.height = res
    return sa, cam, prm


def test_lt_light_only_closed_form():
    le = (3.0, 2.0, 1.0)
    sa, cam = _light_only(le)
    os_ = O.OracleScene(sa)
    img = os_.render_lt(cam, O.params(16, 4, 5, 0, 1, 1))
    total = img.astype(np.float64).sum(axis=(0, 1))
    assert np.allclose(total, np.array(le) * cam.width * cam.height, rtol=1e-4), total
    # the splats land on the light's footprint only
    assert img[0, 0].sum() == 0 and img[cam.height // 2, cam.width // 2].sum() > 0


def test_lt_converges_to_path_tracing():
    sa, cam, _ = _cornell(32)
    os_ = O.OracleScene(sa)
    lt = os_.render_lt(cam, O.params(256, 5, 100, 0, 1, 1)).astype(np.float64)
    c, w, _, _ = os_.render(cam, O.params(256, 5, 100, 0, 8, 1))
    pt = c / w[..., None]
    # interior blocks away from the ceiling light and the border pixels (which
    # collect splats truncated from just outside the image, integrator/LT.h:122-127)
    for (y0, y1, x0, x1) in [(10, 22, 8, 24), (22, 30, 4, 28), (12, 28, 2, 10)]:
        r = lt[y0:y1, x0:x1].mean() / pt[y0:y1, x0:x1].mean()
        assert abs(r - 1) < 0.06, ((y0, y1, x0, x1), r)


def test_lt_importance_mode_closed_forms():
    rng = np.random.default_rng(3)
    n = 4000
    wo = rng.normal(size=(n, 3))
    wo[:, 2] = np.abs(wo[:, 2]) + 0.05  # from outside (eta 1 -> 1.5)
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    u = rng.random((n, 3))
    glass = material(2, R=(1, 1, 1), T=(1, 1, 1), eta1=1.0, eta2=1.5)
    wi_r, f_r, p_r = O.bsdf_sample(glass, wo, u)
    wi_i, f_i, p_i = O.bsdf_sample_importance(glass, wo, u)
    assert np.array_equal(wi_r, wi_i) and np.array_equal(p_r, p_i)
    trans = (wi_i[:, 2] < 0) & (p_i > 0)
    assert trans.sum() > n // 2
    w_i = f_i[trans, 0] * np.abs(wi_i[trans, 2]) / p_i[trans]
    w_r = f_r[trans, 0] * np.abs(wi_r[trans, 2]) / p_r[trans]
    assert np.allclose(w_i, 1.0, rtol=1e-12) and np.allclose(w_r, (1 / 1.5) ** 2, rtol=1e-12)
    rough = material(3, R=(1, 1, 1), T=(1, 1, 1), alpha=(0.3, 0.3), eta1=1.0, eta2=1.5)
    wi_r, f_r, p_r = O.bsdf_sample(rough, wo, u)
    wi_i, f_i, p_i = O.bsdf_sample_importance(rough, wo, u)
    trans = (wi_i[:, 2] < 0) & (f_i[:, 0] > 0)
    assert trans.sum() > n // 4
    assert np.allclose(f_r[trans, 0] / f_i[trans, 0], 1 / 1.5 ** 2, rtol=1e-12)


def test_lt_oracle_threads_sum_the_same_splats():
    # rows dealt to threads, per-thread buffers summed afterwards: the same
    # splats in another summation order
    sa, cam, _ = _cornell(24)
    os_ = O.OracleScene(sa)
    one = os_.render_lt(cam, O.params(8, 5, 5, 0, 1, 1))
    four = os_.render_lt(cam, O.params(8, 5, 5, 0, 4, 1))
    assert np.allclose(four, one, rtol=1e-5, atol=1e-6 * np.abs(one).max())
    assert not np.array_equal(one, 0 * one)


ENV_SCALE = 4.0  # the light-only scene's map power, near its quad's


def _p_area(os_):
    _, f = os_.env_info()  # light powers, the environment light last
    return float(f[:-1].astype(np.float64).sum() / f.astype(np.float64).sum())


def test_lt_environment_paths_carry_zero():
    le = (3.0, 2.0, 1.0)
    sa, cam = _light_only(le)
    os_ = O.OracleScene(sa)
    os_.set_environment(env_image(16, 8), ENV_SCALE)
    p_area = _p_area(os_)
    assert 0.2 < p_area < 0.8, p_area
    img = os_.render_lt(cam, O.params(16, 4, 5, 0, 1, 1))
    total = img.astype(np.float64).sum(axis=(0, 1))
    # Sobol's first dimension is stratified: the area lights' share of the
    # 20480 paths is p_area to within a few paths
    assert np.allclose(total, p_area * np.array(le) * cam.width * cam.height, rtol=2e-3), (total, p_area)


# ---------------------------------------------------------------- GPU


def _compare(gpu, orc):
    assert np.array_equal(np.isnan(gpu), np.isnan(orc))
    ok = ~np.isnan(orc)
    scale = max(np.abs(orc[ok]).mean(), 1e-30)
    err = np.abs(gpu[ok] - orc[ok]) / np.maximum(np.abs(orc[ok]), 1e-3 * scale)
    assert np.quantile(err, 0.999) <= S_REL, np.quantile(err, [0.5, 0.99, 0.999, 1.0])
    return err


@pytest.mark.gpu
@MODES
def test_lt_gpu_matches_oracle_cornell(gpu_ctx, mode):
    sa, cam, prm = _cornell(48)
    prm.spp, prm.max_depth, prm.mode = 16, 5, mode
    orc = O.OracleScene(sa).render_lt(cam, O.params(16, 5, 5, 0, 1, 1))
    with g.Scene(gpu_ctx, sa) as sc:
        gpu = sc.render_light_tracing(cam, prm)
        gpu2 = sc.render_light_tracing(cam, prm)
    _compare(gpu, orc)
    assert np.allclose(gpu2, gpu, rtol=1e-5, atol=1e-7 * np.abs(gpu).max())


@pytest.mark.gpu
@MODES
def test_lt_gpu_matches_oracle_ggx_and_halton(gpu_ctx, mode):
    sa, cam, prm = g.generate_scene(g.GY_SCENE_ICOSPHERE_GGX, 0, 3)
    cam.width = cam.height = 40
    prm.spp, prm.max_depth, prm.sampler, prm.mode = 8, 6, g.GY_SAMPLER_HALTON, mode
    orc = O.OracleScene(sa).render_lt(cam, O.params(8, 6, 5, 0, 1, 1, sampler=1))
    with g.Scene(gpu_ctx, sa) as sc:
        gpu = sc.render_light_tracing(cam, prm)
    _compare(gpu, orc)


@pytest.mark.gpu
@MODES
def test_lt_gpu_matches_oracle_dielectrics(gpu_ctx, mode):
    # glass, rough dielectric, textured Lambert, rough conductor: light paths
    # scatter through the dielectrics in the importance transport mode
    sa, cam, prm = mixed_material_scene(res=40)
    prm.spp, prm.max_depth, prm.mode = 8, 6, mode
    orc = O.OracleScene(sa).render_lt(cam, O.params(8, 6, 5, 0, 1, 1))
    with g.Scene(gpu_ctx, sa) as sc:
        gpu = sc.render_light_tracing(cam, prm)
    _compare(gpu, orc)


@pytest.mark.gpu
@MODES
def test_lt_gpu_light_only_closed_form_and_errors(gpu_ctx, mode):
    le = (3.0, 2.0, 1.0)
    sa, cam = _light_only(le)
    prm = g.render_params(spp=16, max_depth=4, rr_depth=5, mode=mode)
    with g.Scene(gpu_ctx, sa) as sc:
        img = sc.render_light_tracing(cam, prm)
        prm.max_depth = 0
        with pytest.raises(g.GyError):
            sc.render_light_tracing(cam, prm)
    total = img.astype(np.float64).sum(axis=(0, 1))
    assert np.allclose(total, np.array(le) * cam.width * cam.height, rtol=1e-4), total


@pytest.mark.gpu
@MODES
def test_lt_gpu_matches_oracle_environment(gpu_ctx, mode):
    # environment-light paths: the disk point and direction are sampled and
    # traced with zero throughput (NaN where a pdf is zero), as in the oracle
    le = (3.0, 2.0, 1.0)
    sa, cam = _light_only(le)
    os_ = O.OracleScene(sa)
    os_.set_environment(env_image(16, 8), ENV_SCALE)
    prm = g.render_params(spp=16, max_depth=4, rr_depth=5, mode=mode)
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_environment(env_image(16, 8), ENV_SCALE)
        img = sc.render_light_tracing(cam, prm)
    _compare(img, os_.render_lt(cam, O.params(16, 4, 5, 0, 1, 1)))
    total = img.astype(np.float64).sum(axis=(0, 1))
    assert np.allclose(total, _p_area(os_) * np.array(le) * cam.width * cam.height, rtol=2e-3), total
    sa, cam, prm = mixed_material_scene(res=40, closed=False)
    prm.spp, prm.max_depth, prm.mode = 8, 6, mode
    os_ = O.OracleScene(sa)
    os_.set_environment(env_image(32, 16))
    orc = os_.render_lt(cam, O.params(8, 6, 5, 0, 1, 1))
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_environment(env_image(32, 16))
        gpu = sc.render_light_tracing(cam, prm)
    _compare(gpu, orc)


@pytest.mark.gpu
def test_lt_gpu_wavefront_matches_per_path_form(gpu_ctx):
    # the two forms on one scene, depth 1 to 8, then a path tracing render and
    # light tracing again on the same scene (the forms share its wavefront state)
    sa, cam, prm = mixed_material_scene(res=32)
    with g.Scene(gpu_ctx, sa) as sc:
        for depth in (1, 2, 3, 8):
            prm.spp, prm.max_depth = 4, depth
            prm.mode = g.GY_RENDER_MEGAKERNEL
            mk = sc.render_light_tracing(cam, prm)
            prm.mode = g.GY_RENDER_WAVEFRONT
            wf = sc.render_light_tracing(cam, prm)
            assert np.array_equal(np.isnan(mk), np.isnan(wf)), depth
            ok = ~np.isnan(mk)
            assert np.allclose(wf[ok], mk[ok], rtol=1e-5, atol=1e-6 * np.abs(mk[ok]).max()), depth
        prm.max_depth = 5
        sc.render(cam, prm)
        wf2 = sc.render_light_tracing(cam, prm)
        prm.mode = g.GY_RENDER_MEGAKERNEL
        mk2 = sc.render_light_tracing(cam, prm)
    ok = ~np.isnan(mk2)
    assert np.allclose(wf2[ok], mk2[ok], rtol=1e-5, atol=1e-6 * np.abs(mk2[ok]).max())

"""VolPathTracing (integrator/VPT.cpp:7-171) with homogeneous media.

CPU tests pin the oracle's restatement against closed forms:
- Henyey-Greenstein: p(wo, wi) equals the textbook phase function of the
  angle to the forward direction -wo; sampled directions follow its CDF
  (KS), mean cosine g, uniform azimuth; Sample_p returns p(wo, wi);
- free-flight sampling: E[weight] over the sampler equals
  sigma_s / sigma_t (1 - e^{-sigma_t tmax}) + e^{-sigma_t tmax} per channel,
  and the sampled distance follows the truncated exponential of its channel;
- equi-angular sampling: t follows the Cauchy-like CDF about the light
  point's foot on the ray over [0, 3], the pdf it returns is that CDF's
  derivative, and its weights are unbiased for tmax < 3 (for tmax > 3 they
  estimate the integral over [0, 3] only, the reference's fixed range);
- transmittance exp(-sigma_t tmax |d|);
- Scene::IntersectTr crosses a material-less boundary into the interface's
  inside medium (the reference's rule: the crossed record is never queried);
  with the test switch that queries it, into the medium beyond;
- a furnace (black emitting walls, Le = 1, around a purely scattering
  medium cube): every pixel's mean is 1 for both distance samplers, in FP32
  and FP64 (shadow-ray quirk and crossing rule switched to the textbook
  forms, depth large enough that the truncated tail is below 1e-6);
- an absorbing-only cube: the mean radiance of camera samples through it is
  the mean of exp(-sigma_a chord) over their rays.
GPU tests: per-sample radiance and dimensions against the oracle (level b)
for both samplers, the film against the oracle's, and the entry points that
refuse material-less triangles.
"""
import numpy as np
import pytest

import gyunity_amd as g
from gyunity_amd._lib import GY_ERR_UNSUPPORTED
from oracle import oracle as O
from tests.scene_util import env_image, furnace_medium_scene, medium_cornell

RAY_EPS = 1e-4


def _hg_textbook(cos_fwd, gg):
    return (1 - gg * gg) / (4 * np.pi * (1 + gg * gg - 2 * gg * cos_fwd) ** 1.5)


def _hg_cdf(mu, gg):
    # P(cos_fwd <= mu) of the Henyey-Greenstein distribution
    if abs(gg) < 1e-6:
        return (mu + 1) / 2
    return (1 - gg * gg) / (2 * gg) * (1 / np.sqrt(1 + gg * gg - 2 * gg * mu) - 1 / (1 + gg))


@pytest.mark.parametrize("gg", [0.0, 0.6, -0.45, 0.95])
@pytest.mark.parametrize("prec", [0, 1])
def test_henyey_greenstein_sampling_and_value(gg, prec):
    from scipy import stats

    rng = np.random.default_rng(5)
    wo = rng.normal(size=3)
    wo /= np.linalg.norm(wo)
    n = 4000
    cos_f, azi = np.zeros(n), np.zeros(n)
    a = np.cross(wo, (1.0, 0.0, 0.0) if abs(wo[0]) < 0.9 else (0.0, 1.0, 0.0))
    a /= np.linalg.norm(a)
    b = np.cross(wo, a)
    tol = 1e-5 if prec == 0 else 1e-12
    for k in range(n):
        wi, ret, pval = O.hg_probe(gg, wo, rng.random(2), precision=prec)
        assert abs(np.linalg.norm(wi) - 1) <= tol * 10
        c = float(np.dot(wi, -wo))
        assert pval == pytest.approx(_hg_textbook(c, gg), rel=tol * 100)
        assert ret == pytest.approx(pval, rel=tol * 100)
        cos_f[k] = c
        azi[k] = np.arctan2(wi @ b, wi @ a)
    assert stats.kstest(cos_f, lambda m: _hg_cdf(np.clip(m, -1, 1), gg)).pvalue > 1e-3
    assert abs(cos_f.mean() - gg) < 4 * cos_f.std() / np.sqrt(n)
    assert stats.kstest(np.mod(azi, 2 * np.pi), stats.uniform(0, 2 * np.pi).cdf).pvalue > 1e-3


@pytest.mark.parametrize("prec", [0, 1])
def test_free_flight_weights_are_unbiased(prec):
    sa_, ss_ = np.array([0.3, 0.1, 1.2]), np.array([0.9, 2.0, 0.4])
    st = sa_ + ss_
    tmax = 0.8
    rng = np.random.default_rng(7)
    n = 40000
    W = np.zeros((n, 3))
    ts = []
    for k in range(n):
        u = rng.random(2)
        t, sampled, W[k], _ = O.medium_probe(0, sa_, ss_, (0, 0, 0), (0, 0, 1), tmax, u, precision=prec)
        if sampled and int(u[0] * 3) == 1:
            ts.append(t)
    # g(t) = 1 inside, h = 1 at the surface: E[weight] is the exact integral
    exact = ss_ / st * (1 - np.exp(-st * tmax)) + np.exp(-st * tmax)
    se = W.std(0) / np.sqrt(n)
    assert np.all(np.abs(W.mean(0) - exact) < 5 * se), (W.mean(0), exact, se)
    from scipy import stats

    c = st[1]
    cdf = lambda x: (1 - np.exp(-c * x)) / (1 - np.exp(-c * tmax))
    assert stats.kstest(np.array(ts), cdf).pvalue > 1e-3


@pytest.mark.parametrize("prec", [0, 1])
def test_equiangular_sampling_distribution_pdf_and_weights(prec):
    from scipy import stats

    sa_, ss_ = np.array([0.2, 0.5, 0.1]), np.array([1.0, 0.6, 2.0])
    st = sa_ + ss_
    o, d = np.array([0.0, 0.0, 0.0]), np.array([0.0, 0.0, 1.0])
    lp = np.array([0.3, -0.2, 0.9])
    delta = lp @ d
    D = np.linalg.norm(o + delta * d - lp)
    th_a, th_b = np.arctan2(-delta, D), np.arctan2(3 - delta, D)
    cdf = lambda t: (np.arctan2(t - delta, D) - th_a) / (th_b - th_a)
    rng = np.random.default_rng(9)
    for tmax in (1.7, 5.0):
        n = 40000
        W = np.zeros((n, 3))
        ts = []
        for k in range(n):
            u = rng.random(2)
            t, sampled, W[k], pdf = O.medium_probe(1, sa_, ss_, o, d, tmax, u, lp, precision=prec)
            if sampled:
                ts.append(t)
                assert pdf == pytest.approx(D / ((th_b - th_a) * (D * D + (t - delta) ** 2)), rel=1e-4)
            else:
                assert pdf == pytest.approx(1 - cdf(min(3.0, tmax)), rel=1e-4)
        ts = np.array(ts)
        assert stats.kstest(ts, lambda x: np.clip(cdf(x) / cdf(min(3.0, tmax)), 0, 1)).pvalue > 1e-3
        if tmax < 3:
            exact = ss_ / st * (1 - np.exp(-st * tmax)) + np.exp(-st * tmax)
        else:  # the fixed range [0, 3]: the medium past t = 3 and the surface are never reached
            exact = ss_ / st * (1 - np.exp(-st * 3.0))
        se = W.std(0) / np.sqrt(n)
        assert np.all(np.abs(W.mean(0) - exact) < 5 * se), (tmax, W.mean(0), exact, se)


def test_transmittance_closed_form():
    sa_, ss_ = np.array([0.2, 0.5, 0.1]), np.array([1.0, 0.6, 2.0])
    for prec, tol in ((0, 1e-6), (1, 1e-14)):
        _, _, tr, _ = O.medium_probe(2, sa_, ss_, (0, 0, 0), (0.0, 0.6, 0.8), 1.3, precision=prec)
        assert np.allclose(tr, np.exp(-(sa_ + ss_) * 1.3), rtol=tol)


def _furnace(**kw):
    sa, cam, prm, media, tri = furnace_medium_scene(**kw)
    os_ = O.OracleScene(sa)
    os_.set_media(media, tri)
    return sa, cam, prm, media, tri, os_


def test_intersect_tr_crossing_rule():
    # a ray from inside the cube (medium 0) to the far wall: the reference's
    # IntersectTr keeps medium 0 past the boundary, the queried crossing does not
    sa, cam, prm, media, tri, os_ = _furnace(sigma_s=(0.8, 0.4, 1.6), sigma_a=(0.1, 0.2, 0.0))
    st = media["sigma_s"][0].astype(np.float64) + media["sigma_a"][0]
    o, d = np.array([0.5, 0.5, 0.5]), np.array([0.0, 0.0, -1.0])
    inside = 0.5 - 0.25 - RAY_EPS      # to the cube face z = 0.25, from the shifted origin
    outside = 0.25 - RAY_EPS           # from the face (plus RayEps again) to the wall z = 0
    try:
        for q, length in ((0, inside + outside), (1, inside)):
            O.lib().or_set_vpt_query_crossings(q)
            prim, tr = os_.intersect_tr(o, d, 0)
            assert prim >= 0 and prim < 12  # a wall
            assert np.allclose(tr, np.exp(-st * length), rtol=1e-9), (q, tr, np.exp(-st * length))
    finally:
        O.lib().or_set_vpt_query_crossings(0)


@pytest.mark.parametrize("integrator", [1, 2])
@pytest.mark.parametrize("prec", [0, 1])
def test_vpt_furnace_mean_is_one(integrator, prec):
    sa, cam, prm, media, tri, os_ = _furnace()
    O.lib().or_set_shadow_short(1)
    O.lib().or_set_vpt_query_crossings(1)
    try:
        c, w, _, _ = os_.render(cam, O.params(256, prm.max_depth, prm.rr_depth, prec, 8, integrator=integrator))
    finally:
        O.lib().or_set_shadow_short(0)
        O.lib().or_set_vpt_query_crossings(0)
    img = c / w[..., None]
    se = img.std() / np.sqrt(img[..., 0].size)
    assert abs(img.mean() - 1) < 5 * se + 1e-6, (img.mean(), se)


def test_vpt_furnace_mean_is_one_chromatic_anisotropic():
    sa, cam, prm, media, tri, os_ = _furnace(sigma_s=(2.0, 1.0, 0.5), g=-0.4, depth=16)
    O.lib().or_set_shadow_short(1)
    O.lib().or_set_vpt_query_crossings(1)
    try:
        c, w, _, _ = os_.render(cam, O.params(256, prm.max_depth, prm.rr_depth, 1, 8, integrator=1))
    finally:
        O.lib().or_set_shadow_short(0)
        O.lib().or_set_vpt_query_crossings(0)
    img = c / w[..., None]
    for k in range(3):
        se = img[..., k].std() / np.sqrt(img[..., k].size)
        assert abs(img[..., k].mean() - 1) < 5 * se + 1e-6, (k, img[..., k].mean(), se)


def _chords(org, d, lo, hi):
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo - org) / d
        t1 = (hi - org) / d
    tn = np.minimum(t0, t1).max(-1)
    tf = np.maximum(t0, t1).min(-1)
    return np.where(tf > np.maximum(tn, 0), tf - np.maximum(tn, 0), 0.0), tn, tf


@pytest.mark.parametrize("integrator", [1, 2])
def test_vpt_absorbing_cube_transmittance(integrator):
    sig_a = np.array([1.0, 2.0, 3.0])
    sa, cam, prm, media, tri, os_ = _furnace(sigma_s=(0.0, 0.0, 0.0), sigma_a=tuple(sig_a), depth=3)
    rng = np.random.default_rng(2)
    n = 20000
    x = rng.integers(0, cam.width, n).astype(np.int32)
    y = rng.integers(0, cam.height, n).astype(np.int32)
    s = rng.integers(0, 64, n).astype(np.int32)
    L, raster, _ = os_.li_samples(cam, O.params(64, prm.max_depth, prm.rr_depth, 1, 8, integrator=integrator), x, y, s)
    px = np.floor(raster[:, 0]).astype(np.int32)
    py = np.floor(raster[:, 1]).astype(np.int32)
    org, d = O.camera_rays(cam, px, py, (raster[:, 0] - px).astype(np.float32), (raster[:, 1] - py).astype(np.float32))
    chord, tn, tf = _chords(org.astype(np.float64), d.astype(np.float64), 0.25, 0.75)
    through = (chord > 1e-3) & (tn > 1e-3)
    assert through.sum() > 3000
    expect = np.exp(-chord[through, None] * sig_a[None, :])
    got = L[through]
    err = np.abs(got.mean(0) - expect.mean(0))
    se = got.std(0) / np.sqrt(through.sum())
    assert np.all(err < 5 * se + 1e-9), (got.mean(0), expect.mean(0), se)
    # rays that miss the cube see the wall's Le = 1 exactly
    assert np.all(L[chord == 0] == 1.0)


# ----------------------------------------------------------------------------- GPU


def _vpt_parity(gpu_ctx, sa, cam, prm, media, tri, equiangular, n, seed):
    from tests.parity_util import assert_tie_explains

    rng = np.random.default_rng(seed)
    x = rng.integers(0, cam.width, n).astype(np.int32)
    y = rng.integers(0, cam.height, n).astype(np.int32)
    s = rng.integers(0, prm.spp, n).astype(np.int32)
    os_ = O.OracleScene(sa)
    os_.set_media(media, tri)
    prm_o = O.params(prm.spp, prm.max_depth, prm.rr_depth, 0, 8, 1, integrator=2 if equiangular else 1)
    Lo, ro, do = os_.li_samples(cam, prm_o, x, y, s)
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_media(media, tri)
        Lg, rg, dg = sc.vpt_li_samples(cam, prm, x, y, s, equiangular=equiangular)
        assert np.array_equal(rg, ro)
        with np.errstate(invalid="ignore"):
            err = np.abs(Lg - Lo).max(axis=1) / np.maximum(np.abs(Lo).max(axis=1), 1e-3)
            ok = (dg == do) & (err <= 1e-4)
        assert (~ok).mean() < 2e-3, (~ok).mean()
        for i in np.nonzero(~ok)[0]:
            assert_tie_explains(sc, os_, cam, prm_o, int(x[i]), int(y[i]), int(s[i]))
    return err[ok]


@pytest.mark.gpu
@pytest.mark.parametrize("equiangular", [False, True])
def test_vpt_level_b_medium_cornell(gpu_ctx, equiangular):
    sa, cam, prm, media, tri = medium_cornell(res=64, spp=16, depth=8)
    err = _vpt_parity(gpu_ctx, sa, cam, prm, media, tri, equiangular, 20000, 4)
    assert err.size > 19900 and err.max() <= 1e-4


@pytest.mark.gpu
def test_vpt_level_b_furnace_chromatic(gpu_ctx):
    sa, cam, prm, media, tri = furnace_medium_scene(sigma_s=(2.0, 1.0, 0.5), sigma_a=(0.1, 0.3, 0.0), g=0.6, res=32, spp=16,
                                                    depth=10)
    err = _vpt_parity(gpu_ctx, sa, cam, prm, media, tri, False, 10000, 6)
    assert err.size > 9950


@pytest.mark.gpu
@pytest.mark.parametrize("equiangular", [False, True])
def test_vpt_film_matches_oracle(gpu_ctx, equiangular):
    sa, cam, prm, media, tri = medium_cornell(res=48, spp=8, depth=6)
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_media(media, tri)
        color, weight = sc.render_vpt(cam, prm, equiangular=equiangular)
    os_ = O.OracleScene(sa)
    os_.set_media(media, tri)
    oc, ow, _, _ = os_.render(cam, O.params(prm.spp, prm.max_depth, prm.rr_depth, 0, 8, integrator=2 if equiangular else 1))
    assert np.array_equal(weight, ow.astype(np.float32))
    err = (np.abs(color - oc) / np.maximum(np.abs(oc), 1e-2)).max(axis=2)
    assert (err < 1e-4).mean() > 0.99, (err >= 1e-4).sum()
    assert np.abs(color - oc).sum() / np.abs(oc).sum() < 2e-3


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["cornell-free-flight", "cornell-equiangular", "furnace-chromatic", "cornell-env"])
def test_vpt_wavefront_film_equals_per_path(gpu_ctx, case):
    # The wavefront form (trace, vertex, shadow and MIS segment loops, resolve)
    # performs vpt_path's FP32 operations in its order: the films are equal bit
    # for bit, in one pass and in ragged passes, at every depth limit.
    if case == "furnace-chromatic":
        sa, cam, prm, media, tri = furnace_medium_scene(sigma_s=(2.0, 1.0, 0.5), sigma_a=(0.1, 0.3, 0.0), g=0.6, res=32, spp=8,
                                                        depth=10)
    else:
        sa, cam, prm, media, tri = medium_cornell(res=48, spp=8, depth=6)
    eqa = case == "cornell-equiangular"
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_media(media, tri)
        if case == "cornell-env":
            sc.set_environment(env_image())
        for spp_pass, depth in ((0, prm.max_depth), (3, prm.max_depth), (0, 1), (0, 0)):
            prm.samples_per_pass, prm.max_depth = spp_pass, depth
            prm.mode = g.GY_RENDER_MEGAKERNEL
            c0, w0 = sc.render_vpt(cam, prm, equiangular=eqa)
            prm.mode = g.GY_RENDER_WAVEFRONT
            c1, w1 = sc.render_vpt(cam, prm, equiangular=eqa)
            assert np.array_equal(w0, w1), (spp_pass, depth)
            assert np.array_equal(c0, c1, equal_nan=True), (spp_pass, depth, np.abs(c0 - c1).max())
            if depth > 0:
                assert c0.sum() > 0


@pytest.mark.gpu
def test_vpt_refusals(gpu_ctx):
    sa, cam, prm, media, tri = medium_cornell(res=16, spp=2, depth=4)
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_media(media, tri)
        for call in (lambda: sc.render(cam, prm), lambda: sc.render_light_tracing(cam, prm),
                     lambda: sc.li_samples(cam, prm, [0], [0], [0])):
            with pytest.raises(g.GyError) as e:
                call()
            assert e.value.code == GY_ERR_UNSUPPORTED
        with g.Batch(gpu_ctx) as b:
            with pytest.raises(g.GyError):
                b.add(sc, cam, prm)
        bad = media.copy()
        bad["g"] = 1.0
        with pytest.raises(g.GyError):
            sc.set_media(bad, tri)
        t2 = tri.copy()
        t2["inside"][-1] = 5
        with pytest.raises(g.GyError):
            sc.set_media(media, t2)
        # media on triangles that all keep their materials: PathTracing ignores them
        t3 = tri.copy()
        t3["no_material"] = 0
        sc.set_media(media, t3)
        c, w = sc.render(cam, prm)
        assert np.isfinite(c).all()

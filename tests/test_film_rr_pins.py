"""Pins of the oracle's film footprint and Russian roulette (CPU only).

- Film::AddSample with the radius-0.5 box filter (visual/Film.h:61-87,
  math/Filter.h:26-38): per-pixel weights predicted by a brute-force scan of
  every Sobol sample index (tests/scene_util.film_weight_scan), which finds
  each sample's pixels as the closed unit squares holding its raster position.
  The scan shares no code with the oracle's Film::add.
- PathTracing::Li's roulette (integrator/PT.cpp:52-60): in the furnace box the
  estimate must keep its closed-form mean for every roulette start depth, and
  the number of sample dimensions a path consumes must follow the closed-form
  distribution of where roulette ends it (it draws one dimension per roulette
  test, only when p < 1 and i > rrDepth).
"""
import numpy as np
import pytest

import gyunity_amd as g
from oracle import oracle as O
from tests.scene_util import film_weight_scan, furnace_expected, furnace_scene

LUM = np.array([0.212671, 0.715160, 0.072169])  # Vec3::Y (math/Linagl.h:569-572)


# ---------------------------------------------------------------- film footprint
@pytest.mark.parametrize("W,H,spp", [(256, 256, 16), (1024, 1024, 64), (200, 120, 8), (64, 96, 4)])
def test_film_weights_equal_brute_force_scan(W, H, spp):
    # Depth 0: the camera samples reach the film with L = 0; only the footprint
    # counts.  The 1024^2, 64-spp case holds 2304 samples on a left pixel edge
    # (u == 0), 1278 on a right edge (u rounds to 1) and 2 whose u rounds to
    # 1.0 and wraps to 0, which lands them at raster x = 0.
    sa, cam, _ = g.generate_scene(g.GY_SCENE_CORNELL, 0, 64)
    cam.width, cam.height = W, H
    want, counts = film_weight_scan(W, H, spp)
    os_ = O.OracleScene(sa)
    _, wgt, n, _ = os_.render(cam, O.params(spp, 0, 100, nthreads=8))
    assert n == W * H * spp
    assert np.array_equal(wgt, want.astype(np.float64))
    border = want != spp
    if (W, H, spp) == (1024, 1024, 64):
        assert counts == {"u0": 2304, "u1": 1278, "wrap": 2}
        # the wrapped samples come from column 1023 and land in column 0
        assert border[:, 0].any() and want.sum() > W * H * spp
    if (W, H, spp) == (256, 256, 16):
        assert counts["u0"] == 256 and border.any()


def test_film_colour_follows_the_footprint():
    # Furnace box at depth 1: every camera sample returns exactly L = Le = 1
    # (the emission branch, PT.cpp:33-36), so each channel of the film equals
    # the predicted weight, in FP32 and FP64.
    sa = furnace_scene(0.5)
    cam = g.camera((0.5, 0.5, 0.5), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 60.0, 128, 128)
    want, _ = film_weight_scan(128, 128, 16)
    os_ = O.OracleScene(sa)
    for prec in (0, 1):
        c, wgt, _, _ = os_.render(cam, O.params(16, 1, 100, precision=prec, nthreads=8))
        assert np.array_equal(wgt, want.astype(np.float64))
        for ch in range(3):
            assert np.array_equal(c[..., ch], want.astype(np.float64)), (prec, ch)


# ---------------------------------------------------------------- Russian roulette
def roulette_dims_distribution(albedo, depth, rr_depth):
    """Closed-form distribution of the dimensions a furnace-box path consumes.

    Every camera ray hits an emitter, so bounce 0 takes the emission branch
    (no light-sample dimensions); later bounces draw 6 for the light sample.
    Every bounce draws 3 for the BSDF sample.  A Lambert BSDF sampled by its
    cosine gives est = albedo exactly, so the throughput T and the roulette
    probability p = min(1, Y(est T) / Y(T)) follow a fixed sequence; a test
    at bounce i draws one more dimension when p < 1 and i > rr_depth, and ends
    the path with probability 1 - p (PT.cpp:52-60)."""
    a = np.asarray(albedo, np.float64)
    T = np.ones(3)
    dims, alive = 2, 1.0
    dist = {}
    for i in range(depth):
        dims += (6 if i > 0 else 0) + 3
        p = min(1.0, float(LUM @ (a * T)) / float(LUM @ T))
        if p < 1 and i > rr_depth:
            dims += 1
            dist[dims] = dist.get(dims, 0.0) + alive * (1 - p)
            alive *= p
            T = T / p
        T = T * a
    dist[dims] = dist.get(dims, 0.0) + alive
    return dist


def _furnace_samples(sa, depth, rr, prec, res=32, spp=64):
    os_ = O.OracleScene(sa)
    cam = g.camera((0.5, 0.5, 0.5), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 60.0, res, res)
    yy, xx, ss = np.meshgrid(np.arange(res), np.arange(res), np.arange(spp), indexing="ij")
    L, _, dims = os_.li_samples(cam, O.params(spp, depth, rr, precision=prec), xx.ravel(), yy.ravel(), ss.ravel())
    return L, dims


@pytest.mark.parametrize("albedo", [0.5, (0.9, 0.5, 0.2)])
@pytest.mark.parametrize("rr", [0, 1, 2, 3])
def test_russian_roulette_keeps_the_furnace_mean(albedo, rr):
    # Roulette divides the surviving throughput by p; a missing or misplaced
    # division moves the mean below the closed form.  Per channel, with
    # per-channel albedo the roulette probability mixes the channels through
    # Y() and the throughput after a survival is no longer grey.
    depth = 6
    expect = furnace_expected(np.asarray(albedo, np.float64) * np.ones(3), depth)
    sa = furnace_scene(albedo)
    try:
        O.set_shadow_short(True)
        for prec in (0, 1):
            L, _ = _furnace_samples(sa, depth, rr, prec, spp=128)
            mean = L.mean(axis=0)
            se = L.std(axis=0) / np.sqrt(L.shape[0])
            assert np.all(np.abs(mean - expect) < 5 * se + 1e-4), (prec, mean, expect, se)
    finally:
        O.set_shadow_short(False)


@pytest.mark.parametrize("albedo", [0.5, (0.9, 0.5, 0.2)])
@pytest.mark.parametrize("rr", [0, 1, 2, 3, 7])
def test_russian_roulette_path_length_distribution(albedo, rr):
    # The roulette start rule (i > rrDepth, not >=) and p decide which
    # dimension counts occur and how often.
    depth = 6
    sa = furnace_scene(albedo)
    want = roulette_dims_distribution(albedo, depth, rr)
    for prec in (0, 1):
        _, dims = _furnace_samples(sa, depth, rr, prec)
        n = dims.shape[0]
        vals, cnt = np.unique(dims, return_counts=True)
        # A few rays per 10^5 leave the box through a shared edge and end the
        # path before its roulette test (also without roulette).
        extra = ~np.isin(vals, list(want))
        assert cnt[extra].sum() < 1e-3 * n, (prec, vals, sorted(want))
        for d, p in want.items():
            got = cnt[vals == d].sum() / n
            assert abs(got - p) < 5 * np.sqrt(p * (1 - p) / n) + 1e-3, (prec, d, got, p)


def test_roulette_distribution_formula_separates_the_start_rules():
    # The closed form distinguishes i > rrDepth from i >= rrDepth and p from
    # 1 - p, so the test above would catch either slip.
    a = (0.9, 0.5, 0.2)
    assert roulette_dims_distribution(a, 6, 2) != roulette_dims_distribution(a, 6, 1)
    assert roulette_dims_distribution(0.3, 6, 0) != roulette_dims_distribution(0.7, 6, 0)

"""Analytic spheres (shape/Sphere.h:10-351) with the reference's defects.

CPU tests pin the oracle's restatement against geometry:
- closest hit: the nearest root of |o + s d - c| = r, less the RayEps origin
  shift (Ray's constructor), with (u, v) from the hit point's spherical angles;
- any hit (Sphere::Intersect(const Ray&), :131-161, drops its ray): occluded
  exactly when the segment passes the sphere's box (the BVH's leaf test; brute
  force reaches every sphere, so every ray is occluded);
- light sampling from outside: points on the cap visible from the shading
  point, 1.5e-3 r beyond it, directions uniform in the subtended cone, pdf one
  over its solid angle; from inside: uniform points with the pdf's distance
  factor not squared (:243-262); Pdf_Li 0 (the default hit record);
- a camera ray that hits a sphere carries no radiance (the default zero
  intersection record leaves every term zero).
GPU tests: level (a) closest/any hit, level (b) per-sample radiance, the
wavefront form's film equal to the megakernel's, and the forms that refuse
sphere scenes.
"""
import numpy as np
import pytest

import gyunity_amd as g
from gyunity_amd import RAY_DTYPE
from oracle import oracle as O
from tests.scene_util import camera_rays, random_rays, reference_sphere_cornell, sphere_scene

RAY_EPS = 1e-4


def _oracle(sa, sp=None):
    os_ = O.OracleScene(sa)
    if sp is not None:
        os_.set_spheres(sp)
    return os_


def _roots(o, d, c, r):
    """Both roots of |o + s d - c| = r in FP64 (NaN where the line misses)."""
    o, d = o.astype(np.float64), d.astype(np.float64)
    po = o - c
    a = (d * d).sum(-1)
    b = (po * d).sum(-1)
    q = (po * po).sum(-1) - r * r
    disc = b * b - a * q
    with np.errstate(invalid="ignore"):
        sq = np.sqrt(disc)
    return (-b - sq) / a, (-b + sq) / a


def _rays_at_sphere(n, seed, c, r, lo=(0.03, 0.03, 0.03), hi=(0.97, 0.97, 0.97)):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, RAY_DTYPE)
    org = rng.uniform(lo, hi, size=(n, 3))
    tgt = np.asarray(c) + rng.uniform(-1.3 * r, 1.3 * r, size=(n, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["org"], rays["dir"], rays["tmin"], rays["tmax"] = org, d, 0.0, np.inf
    return rays


def test_sphere_closest_hit_is_the_nearest_root_with_spherical_uv():
    sa, cam, prm, sp = sphere_scene()
    os_s, os_t = _oracle(sa, sp), _oracle(sa)
    n_tri = sa.pos.shape[0]
    checked = 0
    for i in range(sp.shape[0]):
        c, r = sp["center"][i].astype(np.float64), float(sp["radius"][i])
        rays = _rays_at_sphere(4000, 10 + i, c, r)
        for use_bvh in (True, False):
            t, prim, u, v = os_s.trace_closest(rays, use_bvh)
            t_tri, _, _, _ = os_t.trace_closest(rays, use_bvh)
            t_tri = np.where(t_tri > 0, t_tri, np.inf)
            o = rays["org"].astype(np.float64)
            d = rays["dir"].astype(np.float64)
            s_near = np.full(rays.shape[0], np.inf)
            k_near = np.full(rays.shape[0], -1)
            for j in range(sp.shape[0]):
                a0, a1 = _roots(o, d, sp["center"][j].astype(np.float64), float(sp["radius"][j]))
                s = np.where(a0 > 0, a0, np.where(a1 > 0, a1, np.inf))
                s = np.where(np.isnan(s), np.inf, s)
                closer = s < s_near
                s_near[closer], k_near[closer] = s[closer], j
            # clear cases: sphere i nearest by a margin over the triangles and the other spheres
            sel = (k_near == i) & (s_near < t_tri - 1e-3) & (s_near > 1e-3)
            assert sel.sum() > 1000
            assert np.all(prim[sel] == n_tri + i)
            # the Ray constructor moves the origin RayEps along d; t is measured
            # from there.  FP32 roots of grazing rays (closest approach near r)
            # lose digits to the discriminant's cancellation: those within 1e-4.
            po = o - c
            perp = np.linalg.norm(po - (po * d).sum(-1)[:, None] * d, axis=1)
            err = np.abs(t - (s_near - RAY_EPS))
            assert err[sel & (perp < 0.95 * r)].max() <= 2e-6
            assert err[sel].max() <= 1e-4
            p = o[sel] + d[sel] * s_near[sel, None] - c
            phi = np.mod(np.arctan2(p[:, 2], p[:, 0]), 2 * np.pi)
            theta = np.arccos(np.clip(p[:, 1] / r, -1, 1))
            assert np.abs(u[sel] - phi / (2 * np.pi)).max() <= 2e-4
            assert np.abs(v[sel] - (1 - theta / np.pi)).max() <= 2e-4
            # rays whose nearest hit is a triangle by a margin keep the triangle
            tri = (t_tri < s_near - 1e-3)
            assert np.all(prim[tri] < n_tri)
            checked += int(sel.sum())
    assert checked > 6000


def _box_passes(o, d, tmin, tmax, lo, hi):
    """Slab test of the segment o + [tmin, tmax] d against [lo, hi] in FP64 and
    the distance of its decision from the nearest slab boundary (t units)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t0 = (lo - o) * inv
        t1 = (hi - o) * inv
    tn = np.minimum(t0, t1).max(-1)
    tf = np.maximum(t0, t1).min(-1)
    lo_t = np.maximum(tn, tmin)
    hi_t = np.minimum(tf, tmax)
    return lo_t <= hi_t, np.abs(hi_t - lo_t)


def test_sphere_any_hit_occludes_every_segment_through_its_box():
    sa, cam, prm, sp = sphere_scene()
    os_s, os_t = _oracle(sa, sp), _oracle(sa)
    rays = random_rays(40000, 5, lo=(0.02, 0.02, 0.02), hi=(0.98, 0.98, 0.98))
    rays["tmax"] = np.random.default_rng(6).uniform(0.02, 1.2, rays.shape[0]).astype(np.float32)
    occ = os_s.trace_any(rays, True)
    occ_bf = os_s.trace_any(rays, False)
    tri = os_t.trace_any(rays, True).astype(bool)
    o = rays["org"].astype(np.float64)
    d = rays["dir"].astype(np.float64)
    tmin = rays["tmin"].astype(np.float64)
    tmax = rays["tmax"].astype(np.float64)
    box = np.zeros(rays.shape[0], bool)
    margin = np.full(rays.shape[0], np.inf)
    sphere_hit = np.zeros(rays.shape[0], bool)
    for j in range(sp.shape[0]):
        c, r = sp["center"][j].astype(np.float64), float(sp["radius"][j])
        b, m = _box_passes(o, d, tmin, tmax, c - r, c + r)
        box |= b
        margin = np.minimum(margin, m)
        a0, a1 = _roots(o, d, c, r)
        sphere_hit |= ((a0 >= tmin) & (a0 <= tmax)) | ((a1 >= tmin) & (a1 <= tmax))
    clear = margin > 1e-5
    exp = tri | box
    assert np.array_equal(occ[clear].astype(bool), exp[clear])
    # brute force (BruteForceAccel) has no leaf boxes: every sphere is reached
    assert occ_bf.all()
    # the defect is exercised: segments through a box that miss its sphere
    assert (clear & box & ~sphere_hit & ~tri).sum() > 300


def _cone_setup(i=2, seed=3, n=4000):
    sa, cam, prm, sp = sphere_scene()
    os_ = _oracle(sa, sp)
    c, r = sp["center"][i].astype(np.float64), float(sp["radius"][i])
    rng = np.random.default_rng(seed)
    return os_, c, r, rng, n


def test_sphere_light_sample_from_outside_is_uniform_in_the_cone():
    from scipy import stats

    os_, c, r, rng, n = _cone_setup()
    for ref in ((0.5, 0.1, 0.5), (0.2, 0.6, 0.8), (0.9, 0.9, 0.1)):
        ref = np.asarray(ref, np.float64)
        dvec = c - ref
        dist = np.linalg.norm(dvec)
        z = dvec / dist
        cos_max = np.sqrt(1 - r * r / (dist * dist))
        cos_t, azi = [], []
        x_axis = np.cross(z, (1.0, 0.0, 0.0) if abs(z[0]) < 0.9 else (0.0, 1.0, 0.0))
        x_axis /= np.linalg.norm(x_axis)
        y_axis = np.cross(z, x_axis)
        for _ in range(n):
            u = rng.random(2)
            pos, nrm, pdf, pdf_li = os_.sphere_light_probe(2, ref, u, z)
            # 1.5e-3 r beyond the surface, along the outward normal
            assert abs(np.linalg.norm(pos - c) - r * (1 + 1.5e-3)) <= 1e-12
            assert np.allclose(nrm, (pos - c) / np.linalg.norm(pos - c), atol=1e-12)
            s = c + r * nrm
            w = (s - ref) / np.linalg.norm(s - ref)
            # the surface point is the one the direction sees first
            a0, _ = _roots(ref[None], w[None], c, r)
            assert abs(a0[0] - np.linalg.norm(s - ref)) <= 1e-9
            # pdf = 1 / solid angle of the cone
            assert pdf == pytest.approx(1 / (2 * np.pi * (1 - cos_max)), rel=1e-12)
            assert pdf_li == 0.0  # the default hit record: |d|^2 / 0 -> inf -> 0
            cos_t.append(w @ z)
            azi.append(np.arctan2(w @ y_axis, w @ x_axis))
        cos_t = np.asarray(cos_t)
        assert cos_t.min() >= cos_max - 1e-12
        assert stats.kstest(cos_t, stats.uniform(cos_max, 1 - cos_max).cdf).pvalue > 1e-3
        assert stats.kstest(np.mod(azi, 2 * np.pi), stats.uniform(0, 2 * np.pi).cdf).pvalue > 1e-3


def test_sphere_light_sample_from_inside_is_uniform_with_unsquared_distance():
    os_, c, r, rng, n = _cone_setup(i=1, n=6000)
    ref = c + np.array([0.3, -0.2, 0.1]) * r
    pts = []
    for _ in range(n):
        u = rng.random(2)
        pos, nrm, pdf, pdf_li = os_.sphere_light_probe(1, ref, u, (0.0, 1.0, 0.0))
        assert abs(np.linalg.norm(pos - c) - r) <= 1e-12
        w = (pos - ref) / np.linalg.norm(pos - ref)
        cos = abs(((pos - c) / r) @ w)
        dist = np.linalg.norm(pos - ref)
        area_pdf_solid = dist * dist / (4 * np.pi * r * r * cos)  # the textbook conversion
        # Sphere::Sample multiplies by the distance, not its square
        assert pdf * dist == pytest.approx(area_pdf_solid, rel=1e-9)
        pts.append((pos - c) / r)
    pts = np.asarray(pts)
    # uniform on the sphere: zero mean, second moments 1/3
    assert np.abs(pts.mean(0)).max() < 4 / np.sqrt(3 * n)
    assert np.abs((pts ** 2).mean(0) - 1 / 3).max() < 4 * np.sqrt(4 / 45 / n)


def test_sphere_hits_carry_no_radiance():
    sa, cam, prm, sp = sphere_scene(res=48)
    os_ = _oracle(sa, sp)
    color, weight, _, _ = os_.render(cam, O.params(8, prm.max_depth, prm.rr_depth, precision=1, nthreads=8))
    img = color / weight[..., None]
    W, H = cam.width, cam.height
    n_tri = sa.pos.shape[0]
    covered = np.ones((H, W), bool)
    # the film footprint of pixel (x, y) spans raster [x - 0.5, x + 1.5]: a 5 x 5
    # grid of camera rays over it must all see a sphere first, 5% inside its rim
    ys, xs = np.mgrid[0:H, 0:W]
    for fx in np.linspace(-0.5, 1.5, 5):
        for fy in np.linspace(-0.5, 1.5, 5):
            rx = np.clip(xs.ravel() + fx, 0, W - 1e-3)
            ry = np.clip(ys.ravel() + fy, 0, H - 1e-3)
            px, py = np.floor(rx).astype(np.int32), np.floor(ry).astype(np.int32)
            org, d = O.camera_rays(cam, px, py, (rx - px).astype(np.float32), (ry - py).astype(np.float32))
            rays = np.zeros(px.shape[0], RAY_DTYPE)
            rays["org"], rays["dir"], rays["tmin"], rays["tmax"] = org, d, 0.0, np.inf
            _, prim, _, _ = os_.trace_closest(rays, True)
            deep = np.zeros(px.shape, bool)
            for j in range(sp.shape[0]):
                c, r = sp["center"][j].astype(np.float64), float(sp["radius"][j])
                a0, _ = _roots(org, d, c, 0.95 * r)
                deep |= (prim == n_tri + j) & np.isfinite(a0) & (a0 > 0)
            covered &= deep.reshape(H, W)
    covered[0, :] = covered[-1, :] = covered[:, 0] = covered[:, -1] = False
    assert covered.sum() > 40
    assert np.all(img[covered] == 0.0)
    assert img[~covered].max() > 0.0


def test_reference_sphere_cornell_renders_black():
    # The reference's all-sphere Cornell box: every camera ray hits a sphere
    # (inside-out walls enclose the camera), so with the default intersection
    # record every pixel is exactly 0 while every sample still adds its weight.
    sa, cam, prm, sp = reference_sphere_cornell(32, 24, 4)
    os_ = _oracle(sa, sp)
    _, prim, _, _ = os_.trace_closest(camera_rays(cam, 4000, 1), True)
    assert np.all(prim >= sa.n_tri)
    color, weight, _, _ = os_.render(cam, O.params(4, prm.max_depth, prm.rr_depth, precision=1, nthreads=8))
    assert np.all(color == 0.0)
    assert weight.sum() >= 32 * 24 * 4


# ----------------------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_sphere_level_a_closest_and_any_hit(gpu_ctx):
    from tests.parity_util import check_closest

    sa, cam, prm, sp = sphere_scene()
    os_ = _oracle(sa, sp)
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_spheres(sp)
        for seed in range(2):
            check_closest(sc, os_, camera_rays(cam, 20000, seed))
            check_closest(sc, os_, random_rays(20000, 100 + seed))
            for i in range(sp.shape[0]):
                c, r = sp["center"][i], float(sp["radius"][i])
                frac = check_closest(sc, os_, _rays_at_sphere(8000, 30 + 3 * seed + i, c, r))
                assert frac > 0.999
            rays = random_rays(20000, 200 + seed)
            rays["tmax"] = np.random.default_rng(seed).uniform(0.05, 1.5, rays.shape[0]).astype(np.float32)
            occ = sc.trace_any(rays)
            ref = os_.trace_any(rays, True)
            assert np.array_equal(occ, ref), (occ != ref).sum()


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2])
def test_sphere_level_b_per_sample(gpu_ctx, seed):
    from tests.parity_util import li_parity

    sa, cam, prm, sp = sphere_scene(res=64, spp=16)
    err = li_parity(gpu_ctx, sa, cam, prm, 20000, seed, spheres=sp)
    assert err.size > 19000 and err.max() <= 1e-4


@pytest.mark.gpu
def test_reference_sphere_cornell_level_b_and_film(gpu_ctx):
    # the all-sphere scene through the library: per-sample radiance (all 0)
    # and dimensions consumed equal the oracle's; the film is black with the
    # oracle's weights
    from tests.parity_util import li_parity

    sa, cam, prm, sp = reference_sphere_cornell(64, 48, 8)
    err = li_parity(gpu_ctx, sa, cam, prm, 10000, 3, spheres=sp)
    assert err.size >= 9980
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_spheres(sp)
        color, weight = sc.render(cam, prm)
    _, w, _, _ = _oracle(sa, sp).render(cam, O.params(prm.spp, prm.max_depth, prm.rr_depth, 0, 8))
    assert np.all(color == 0.0) and np.array_equal(weight, w.astype(np.float32))


@pytest.mark.gpu
def test_sphere_scene_film_equal_in_both_forms_and_matches_oracle(gpu_ctx):
    sa, cam, prm, sp = sphere_scene(res=64, spp=16)
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_spheres(sp)
        prm.mode = g.GY_RENDER_MEGAKERNEL
        cm, wm = sc.render(cam, prm)
        prm.mode = g.GY_RENDER_WAVEFRONT  # runs the megakernel for sphere scenes
        cw, ww = sc.render(cam, prm)
        assert np.array_equal(cm, cw) and np.array_equal(wm, ww)
    os_ = _oracle(sa, sp)
    c, w, _, _ = os_.render(cam, O.params(prm.spp, prm.max_depth, prm.rr_depth, 0, 8))
    assert np.array_equal(wm, w.astype(np.float32))
    # Pixels agree to 1e-4 except those holding a sample whose path met an
    # exact-t tie (each of the 16 samples of a pixel's 3 x 3 neighbourhood can
    # land in its footprint, so a ~1.5e-4 tie rate per sample marks ~0.5% of
    # pixels); the image as a whole within 2e-3.
    err = (np.abs(cm - c) / np.maximum(np.abs(c), 1e-2)).max(axis=2)
    assert (err < 1e-4).mean() > 0.99, (err >= 1e-4).sum()
    assert np.abs(cm - c).sum() / np.abs(c).sum() < 2e-3


@pytest.mark.gpu
def test_sphere_scene_refused_by_batch_light_tracing_and_adjoint(gpu_ctx):
    sa, cam, prm, sp = sphere_scene(res=32, spp=4)
    with g.Scene(gpu_ctx, sa) as sc:
        sc.set_spheres(sp)
        with pytest.raises(g.GyError) as e:
            sc.set_spheres(sp)
        assert "already" in str(e.value)
        with pytest.raises(g.GyError):
            sc.render_light_tracing(cam, prm)
        with pytest.raises(g.GyError):
            sc.adjoint(cam, prm, np.zeros((cam.height, cam.width, 3), np.float32))
        b = g.Batch(gpu_ctx)
        try:
            with pytest.raises(g.GyError):
                b.add(sc, cam, prm)
        finally:
            b.close()

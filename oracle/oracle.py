"""ctypes wrapper of the CPU oracle (oracle/libgy_oracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load it, as the checker or the timed CPU baseline.  The
product path (gyunity_amd) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgy_oracle.so")
GOLDEN = os.path.join(os.path.dirname(HERE), "tests", "golden")


class ORSceneDesc(C.Structure):
    _fields_ = [
        ("n_tri", C.c_int32),
        ("pos", C.c_void_p),
        ("nrm", C.c_void_p),
        ("face_n", C.c_void_p),
        ("uv", C.c_void_p),
        ("mat", C.c_void_p),
        ("light", C.c_void_p),
        ("n_mat", C.c_int32),
        ("mats", C.c_void_p),
        ("n_light", C.c_int32),
        ("light_tri", C.c_void_p),
        ("light_le", C.c_void_p),
        ("n_tex", C.c_int32),
        ("tex_info", C.c_void_p),
        ("tex_data", C.c_void_p),
    ]


class ORCamera(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3),
        ("look_at", C.c_float * 3),
        ("up", C.c_float * 3),
        ("fovy", C.c_float),
        ("film_dist", C.c_float),
        ("width", C.c_int32),
        ("height", C.c_int32),
    ]


class ORParams(C.Structure):
    _fields_ = [
        ("spp", C.c_int32),
        ("max_depth", C.c_int32),
        ("rr_depth", C.c_int32),
        ("precision", C.c_int32),
        ("nthreads", C.c_int32),
        ("use_bvh", C.c_int32),
        ("sampler", C.c_int32),
        ("integrator", C.c_int32),
    ]


class ORSppmParams(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("max_depth", C.c_int32),
        ("photons", C.c_int64),
        ("initial_radius", C.c_float),
        ("alpha", C.c_float),
        ("batch_shrink", C.c_int32),
        ("sampler", C.c_int32),
        ("precision", C.c_int32),
        ("nthreads", C.c_int32),
    ]


def sppm_params(iterations, photons, max_depth, initial_radius, alpha, batch_shrink=0, sampler=0, precision=0, nthreads=0):
    p = ORSppmParams()
    p.iterations, p.photons, p.max_depth = iterations, photons, max_depth
    p.initial_radius, p.alpha, p.batch_shrink = initial_radius, alpha, batch_shrink
    p.sampler, p.precision, p.nthreads = sampler, precision, nthreads
    return p


def sppm_grid_query(pos, r2, q):
    """The HashGrid bucket each query point looks up (point indices, duplicates kept)."""
    pos = np.ascontiguousarray(pos, np.float32)
    r2 = np.ascontiguousarray(r2, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    counts = np.zeros(q.shape[0], np.int32)
    lib().or_sppm_grid_query(pos.shape[0], _p(pos), _p(r2), q.shape[0], _p(q), _p(counts), None, 0)
    cap = int(counts.sum())
    idx = np.zeros(max(cap, 1), np.int64)
    rc = lib().or_sppm_grid_query(pos.shape[0], _p(pos), _p(r2), q.shape[0], _p(q), _p(counts), _p(idx), cap)
    assert rc == 0
    return np.split(idx[:counts.sum()], np.cumsum(counts)[:-1])


# or_medium / or_tri_medium
MEDIUM_DTYPE = np.dtype([("sigma_a", "<f4", 3), ("sigma_s", "<f4", 3), ("g", "<f4")])
TRI_MEDIUM_DTYPE = np.dtype([("inside", "<i4"), ("outside", "<i4"), ("no_material", "<i4")])


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u32, u64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_float
    sig = {
        "or_load_sobol": (C.c_int, [C.c_char_p]),
        "or_sobol_sample": (f32, [u64, u32]),
        "or_sobol_enum_index": (u64, [u32, u32, u32, u32]),
        "or_camera_matrix": (C.c_int, [vp, vp]),
        "or_camera_rays": (C.c_int, [vp, i32, vp, vp, vp, vp, vp, vp]),
        "or_scene_create": (vp, [vp]),
        "or_scene_destroy": (None, [vp]),
        "or_bvh_stats": (C.c_int, [vp, vp, vp]),
        "or_trace_closest": (C.c_int, [vp, i32, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp]),
        "or_trace_any": (C.c_int, [vp, i32, vp, vp, vp, vp, C.c_int, vp]),
        "or_trace_counts": (C.c_int, [vp, i32, vp, vp, vp, vp, vp, vp]),
        "or_li_samples": (C.c_int, [vp, vp, vp, i32, vp, vp, vp, vp, vp, vp]),
        "or_render": (i64, [vp, vp, vp, i32, i32, vp, vp, vp]),
        "or_mat4_inverse": (C.c_int, [vp, vp]),
        "or_mat4_mul_vec": (C.c_int, [vp, vp, vp]),
        "or_bsdf_eval": (C.c_int, [vp, i32, vp, vp, vp, vp]),
        "or_bsdf_sample": (C.c_int, [vp, i32, vp, vp, vp, vp, vp]),
        "or_fr_dielectric": (C.c_double, [C.c_double, C.c_double, C.c_double]),
        "or_fr_conductor": (C.c_int, [C.c_double, vp, vp, vp]),
        "or_ggx_D": (C.c_double, [C.c_double, C.c_double, vp]),
        "or_roughness_to_alpha": (C.c_double, [C.c_double]),
        "or_li_path": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp]),
        "or_halton_sample": (C.c_float, [C.c_uint64, C.c_uint32]),
        "or_halton_enum_index": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32]),
        "or_faure_perm": (C.c_int, [C.c_uint32, vp]),
        "or_tex_lookup": (C.c_int, [C.c_int, C.c_int, C.c_int, vp, C.c_int, vp, vp, C.c_int, vp]),
        "or_set_shadow_short": (C.c_int, [C.c_int]),
        "or_set_detached": (C.c_int, [vp, vp, vp]),
        "or_scene_set_environment": (C.c_int, [vp, vp, i32, i32, f32]),
        "or_scene_set_spheres": (C.c_int, [vp, i32, vp]),
        "or_sphere_light_probe": (C.c_int, [vp, i32, C.c_int, vp, vp, vp, vp]),
        "or_scene_set_media": (C.c_int, [vp, i32, vp, i32, vp]),
        "or_set_vpt_query_crossings": (C.c_int, [C.c_int]),
        "or_render_sppm": (C.c_int, [vp, vp, vp, vp, vp, vp]),
        "or_sppm_grid_query": (C.c_int, [i32, vp, vp, i32, vp, vp, vp, C.c_int64]),
        "or_vpt_intersect_tr": (C.c_int, [vp, C.c_int, vp, vp, i32, vp]),
        "or_hg_probe": (C.c_int, [C.c_int, C.c_double, vp, vp, vp, vp]),
        "or_medium_probe": (C.c_int, [C.c_int, C.c_int, vp, vp, vp, vp, C.c_double, vp, vp, vp]),
        "or_env_query": (C.c_int, [vp, i32, vp, vp, vp, vp, vp, vp, C.c_int]),
        "or_env_info": (C.c_int, [vp, vp, vp, i32]),
        "or_render_lt": (C.c_int, [vp, vp, vp, vp]),
        "or_bsdf_sample_mode": (C.c_int, [vp, i32, vp, vp, vp, vp, vp, C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    rc = L.or_load_sobol(GOLDEN.encode())
    if rc <= 0:
        raise RuntimeError("oracle: cannot load Sobol fixtures from %s (rc %d)" % (GOLDEN, rc))
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def camera_from(cam):
    """Copy a gyunity_amd CameraDesc (same layout) into the oracle's struct."""
    o = ORCamera()
    o.pos[:], o.look_at[:], o.up[:] = list(cam.pos), list(cam.look_at), list(cam.up)
    o.fovy, o.film_dist, o.width, o.height = cam.fovy, cam.film_dist, cam.width, cam.height
    return o


def params(spp, max_depth, rr_depth, precision=0, nthreads=0, use_bvh=1, sampler=0, integrator=0):
    """integrator 0: PathTracing; 1: VolPathTracing, free-flight sampling;
    2: VolPathTracing, equi-angular sampling."""
    p = ORParams()
    p.spp, p.max_depth, p.rr_depth, p.precision, p.nthreads, p.use_bvh = spp, max_depth, rr_depth, precision, nthreads, use_bvh
    p.sampler, p.integrator = sampler, integrator
    return p


def hg_probe(g, wo, u, precision=1):
    """HenyeyGreenstein::Sample_p(wo, u): (wi, returned value, p(wo, wi))."""
    wo = np.ascontiguousarray(wo, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    wi = np.zeros(3, np.float64)
    pv = np.zeros(2, np.float64)
    lib().or_hg_probe(int(precision), float(g), _p(wo), _p(u), _p(wi), _p(pv))
    return wi, pv[0], pv[1]


def medium_probe(kind, sigma_a, sigma_s, o, d, tmax, u=(0.0, 0.0), light_pos=(0.0, 0.0, 0.0), precision=1):
    """kind 0: HomogeneousMedium::Sample (u = channel, distance); 1:
    EquiAngularSampling towards light_pos (u[0]); 2: Tr.  Returns (t, sampled,
    weight[3], pdf)."""
    a = [np.ascontiguousarray(x, np.float64) for x in (sigma_a, sigma_s, o, d, u, light_pos)]
    out = np.zeros(6, np.float64)
    lib().or_medium_probe(int(precision), int(kind), _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), float(tmax), _p(a[4]),
                          _p(a[5]), _p(out))
    return out[0], bool(out[1]), out[2:5], out[5]


def sobol_sample(index, dim):
    return lib().or_sobol_sample(index, dim)


def sobol_enum_index(m, frame, px, py):
    return lib().or_sobol_enum_index(m, frame, px, py)


def camera_matrix(cam):
    out = np.zeros(16, dtype=np.float32)
    c = camera_from(cam)
    lib().or_camera_matrix(C.byref(c), out.ctypes.data)
    return out


def camera_rays(cam, px, py, u, v):
    c = camera_from(cam)
    px = np.ascontiguousarray(px, np.int32)
    py = np.ascontiguousarray(py, np.int32)
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    n = px.shape[0]
    org = np.zeros((n, 3), np.float32)
    d = np.zeros((n, 3), np.float32)
    lib().or_camera_rays(C.byref(c), n, _p(px), _p(py), _p(u), _p(v), _p(org), _p(d))
    return org, d


class OracleScene:
    def __init__(self, arrays):
        self.arrays = arrays
        d = ORSceneDesc()
        a = arrays
        d.n_tri = a.n_tri
        d.pos, d.nrm, d.face_n = _p(a.pos), _p(a.nrm), _p(a.face_n)
        d.uv = _p(a.uv)
        d.mat, d.light = _p(a.mat), _p(a.light)
        d.n_mat, d.mats = a.mats.shape[0], _p(a.mats)
        d.n_light, d.light_tri, d.light_le = a.light_tri.shape[0], _p(a.light_tri), _p(a.light_le)
        if a.tex_info is not None:
            d.n_tex, d.tex_info, d.tex_data = a.tex_info.shape[0], _p(a.tex_info), _p(a.tex_data)
        self.h = lib().or_scene_create(C.byref(d))

    def close(self):
        if self.h:
            lib().or_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def bvh_stats(self):
        n, dmax = C.c_int64(), C.c_int32()
        lib().or_bvh_stats(self.h, C.byref(n), C.byref(dmax))
        return n.value, dmax.value

    def _rays(self, rays):
        org = np.ascontiguousarray(rays["org"], np.float32)
        d = np.ascontiguousarray(rays["dir"], np.float32)
        tmin = np.ascontiguousarray(rays["tmin"], np.float32)
        tmax = np.ascontiguousarray(rays["tmax"], np.float32)
        return org, d, tmin, tmax

    def trace_closest(self, rays, use_bvh=True):
        org, d, tmin, tmax = self._rays(rays)
        n = org.shape[0]
        t = np.zeros(n, np.float32)
        prim = np.zeros(n, np.int32)
        b1 = np.zeros(n, np.float32)
        b2 = np.zeros(n, np.float32)
        lib().or_trace_closest(self.h, n, _p(org), _p(d), _p(tmin), _p(tmax), int(use_bvh), _p(t), _p(prim), _p(b1), _p(b2))
        return t, prim, b1, b2

    def trace_any(self, rays, use_bvh=True):
        org, d, tmin, tmax = self._rays(rays)
        n = org.shape[0]
        occ = np.zeros(n, np.int32)
        lib().or_trace_any(self.h, n, _p(org), _p(d), _p(tmin), _p(tmax), int(use_bvh), _p(occ))
        return occ

    def trace_counts(self, rays):
        org, d, tmin, tmax = self._rays(rays)
        a, b = C.c_int64(), C.c_int64()
        lib().or_trace_counts(self.h, org.shape[0], _p(org), _p(d), _p(tmin), _p(tmax), C.byref(a), C.byref(b))
        return a.value, b.value

    def li_samples(self, cam, prm, x, y, s):
        c = camera_from(cam)
        x = np.ascontiguousarray(x, np.int32)
        y = np.ascontiguousarray(y, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        n = x.shape[0]
        L = np.zeros((n, 3), np.float64)
        raster = np.zeros((n, 2), np.float32)
        dims = np.zeros(n, np.int32)
        rc = lib().or_li_samples(self.h, C.byref(c), C.byref(prm), n, _p(x), _p(y), _p(s), _p(L), _p(raster), _p(dims))
        assert rc == 0
        return L, raster, dims

    def li_path(self, cam, prm, x, y, s, max_rows=64):
        """Test-only: per-bounce record of one sample's path (16 doubles per
        bounce: origin, direction, tmin, tmax, hit prim or -1, t, dims used
        before the bounce, throughput, material id, delta flag)."""
        c = camera_from(cam)
        rows = np.zeros((max_rows, 16), np.float64)
        n = lib().or_li_path(self.h, C.byref(c), C.byref(prm), int(x), int(y), int(s), max_rows, _p(rows))
        return rows[:n]

    def set_spheres(self, spheres):
        """Sphere primitives after the triangles (shape/Sphere.h:21-27), emissive
        ones with an AreaLight after the existing lights; spheres is a
        structured array with gy_sphere's fields (or_sphere has the same layout)."""
        self._sph = np.ascontiguousarray(spheres)
        assert self._sph.dtype.itemsize == 40
        rc = lib().or_scene_set_spheres(self.h, self._sph.shape[0], _p(self._sph))
        assert rc == 0

    def set_media(self, media, tri):
        """HomogeneousMedium list (MEDIUM_DTYPE) and per-triangle interface
        (TRI_MEDIUM_DTYPE: inside, outside, -1 none; no_material)."""
        self._media = np.ascontiguousarray(media, MEDIUM_DTYPE)
        self._tri_media = np.ascontiguousarray(tri, TRI_MEDIUM_DTYPE)
        rc = lib().or_scene_set_media(self.h, self._media.shape[0], _p(self._media), self._tri_media.shape[0],
                                      _p(self._tri_media))
        assert rc == 0, rc

    def render_sppm(self, cam, sp):
        """SPPM (or_render_sppm): (image (H, W, 3), radius^2 (H, W), photon count N (H, W))."""
        c = camera_from(cam)
        img = np.zeros((cam.height, cam.width, 3), np.float64)
        r2 = np.zeros((cam.height, cam.width), np.float64)
        nn = np.zeros((cam.height, cam.width), np.float64)
        rc = lib().or_render_sppm(self.h, C.byref(c), C.byref(sp), _p(img), _p(r2), _p(nn))
        assert rc == 0, rc
        return img, r2, nn

    def intersect_tr(self, o, d, medium, precision=1):
        """Scene::IntersectTr from o along d in medium: (hit primitive or -1, Tr)."""
        o = np.ascontiguousarray(o, np.float64)
        d = np.ascontiguousarray(d, np.float64)
        out = np.zeros(4, np.float64)
        lib().or_vpt_intersect_tr(self.h, int(precision), _p(o), _p(d), int(medium), _p(out))
        return int(out[0]), out[1:4]

    def sphere_light_probe(self, i, ref, u, wi, precision=1):
        """Sphere i's Sample(Intersection, pdf, u) from a shading point at ref
        and its Pdf_Li towards wi: (pos, n, pdf, pdf_li)."""
        ref = np.ascontiguousarray(ref, np.float64)
        u = np.ascontiguousarray(u, np.float64)
        wi = np.ascontiguousarray(wi, np.float64)
        out = np.zeros(8, np.float64)
        rc = lib().or_sphere_light_probe(self.h, int(i), int(precision), _p(ref), _p(u), _p(wi), _p(out))
        assert rc == 0
        return out[0:3], out[3:6], out[6], out[7]

    def set_environment(self, rgb, scale=1.0):
        """EnvironmentLight(image, rotate, scale) added after the area lights;
        rgb is (h, w, 3) float32, top row first."""
        self._env = np.ascontiguousarray(rgb, np.float32)
        h, w = self._env.shape[:2]
        rc = lib().or_scene_set_environment(self.h, _p(self._env), w, h, float(scale))
        assert rc == 0

    def env_query(self, kind, a, b, c=None):
        """kind 0: LookupRadiance(a, b); 1: Sample(a, b) -> (rgb, dir, pdf);
        2: PdfDir((a, b, c)); 3: Emission((a, b, c))."""
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        c = np.ascontiguousarray(c if c is not None else np.zeros_like(a), np.float32)
        n = a.shape[0]
        rgb = np.zeros((n, 3), np.float32)
        d = np.zeros((n, 3), np.float32)
        pdf = np.zeros(n, np.float32)
        rc = lib().or_env_query(self.h, n, _p(a), _p(b), _p(c), _p(rgb), _p(d), _p(pdf), kind)
        assert rc == 0
        return rgb, d, pdf

    def env_info(self):
        """(bounding-sphere radius, light power function values)."""
        r = C.c_float()
        f = np.zeros(64, np.float32)
        n = lib().or_env_info(self.h, C.byref(r), _p(f), 64)
        return r.value, f[:n].copy()

    def set_detached(self, tex_eval=None, le_eval=None):
        """Test-only detached evaluation: paths sampled with the scene's own
        texels and Le, BSDF values and emission from tex_eval / le_eval
        (None: the scene's own; both None: off)."""
        self._tex_eval = None if tex_eval is None else np.ascontiguousarray(tex_eval, np.float32)
        self._le_eval = None if le_eval is None else np.ascontiguousarray(le_eval, np.float32)
        lib().or_set_detached(self.h, _p(self._tex_eval), _p(self._le_eval))

    def render_lt(self, cam, prm):
        """LightTracing (integrator/LT.h) in float mode: the (H, W, 3) splat sums."""
        c = camera_from(cam)
        out = np.zeros((cam.height, cam.width, 3), np.float32)
        rc = lib().or_render_lt(self.h, C.byref(c), C.byref(prm), _p(out))
        if rc != 0:
            raise ValueError("or_render_lt: unsupported scene (%d)" % rc)
        return out

    def render(self, cam, prm, row0=0, row1=0):
        c = camera_from(cam)
        color = np.zeros((cam.height, cam.width, 3), np.float64)
        weight = np.zeros((cam.height, cam.width), np.float64)
        secs = C.c_double()
        paths = lib().or_render(self.h, C.byref(c), C.byref(prm), row0, row1, _p(color), _p(weight), C.byref(secs))
        return color, weight, paths, secs.value


def mat4_inverse(rows):
    rows = np.ascontiguousarray(rows, np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    lib().or_mat4_inverse(rows.ctypes.data, out.ctypes.data)
    return out.reshape(4, 4)


def mat4_mul_vec(rows, v):
    rows = np.ascontiguousarray(rows, np.float32).reshape(16)
    v = np.ascontiguousarray(v, np.float32)
    out = np.zeros(4, np.float32)
    lib().or_mat4_mul_vec(rows.ctypes.data, v.ctypes.data, out.ctypes.data)
    return out


def bsdf_eval(mat_words, wo, wi):
    mat_words = np.ascontiguousarray(mat_words, np.uint32)
    wo = np.ascontiguousarray(wo, np.float64).reshape(-1, 3)
    wi = np.ascontiguousarray(wi, np.float64).reshape(-1, 3)
    n = wo.shape[0]
    f = np.zeros((n, 3))
    pdf = np.zeros(n)
    lib().or_bsdf_eval(mat_words.ctypes.data, n, wo.ctypes.data, wi.ctypes.data, f.ctypes.data, pdf.ctypes.data)
    return f, pdf


def bsdf_sample(mat_words, wo, u):
    mat_words = np.ascontiguousarray(mat_words, np.uint32)
    wo = np.ascontiguousarray(wo, np.float64).reshape(-1, 3)
    u = np.ascontiguousarray(u, np.float64).reshape(-1, 3)
    n = wo.shape[0]
    wi = np.zeros((n, 3))
    f = np.zeros((n, 3))
    pdf = np.zeros(n)
    lib().or_bsdf_sample(mat_words.ctypes.data, n, wo.ctypes.data, u.ctypes.data, wi.ctypes.data, f.ctypes.data, pdf.ctypes.data)
    return wi, f, pdf


def bsdf_sample_importance(mat_words, wo, u):
    """bsdf_sample in TransportMode::Importance."""
    wo = np.ascontiguousarray(wo, np.float64)
    u = np.ascontiguousarray(u, np.float64)
    n = wo.shape[0]
    wi = np.zeros((n, 3))
    f = np.zeros((n, 3))
    pdf = np.zeros(n)
    w = np.ascontiguousarray(mat_words, np.uint32)
    lib().or_bsdf_sample_mode(_p(w), n, _p(wo), _p(u), _p(wi), _p(f), _p(pdf), 1)
    return wi, f, pdf


def fr_dielectric(c, eta_i, eta_t):
    return lib().or_fr_dielectric(c, eta_i, eta_t)


def fr_conductor(c, eta, k):
    eta = np.ascontiguousarray(eta, np.float64)
    k = np.ascontiguousarray(k, np.float64)
    out = np.zeros(3)
    lib().or_fr_conductor(c, eta.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def roughness_to_alpha(r):
    return lib().or_roughness_to_alpha(float(r))


def ggx_D(ax, ay, m):
    m = np.ascontiguousarray(m, np.float64)
    return lib().or_ggx_D(ax, ay, m.ctypes.data)


def halton_sample(index, dim):
    return lib().or_halton_sample(index, dim)


def halton_enum_index(w, h, i, x, y):
    return lib().or_halton_enum_index(w, h, i, x, y)


def faure_perm(base):
    out = np.zeros(base, np.uint16)
    assert lib().or_faure_perm(base, out.ctypes.data) == 0
    return out


def tex_lookup(data, u, v, ch):
    """Bilinear lookup (CLAMP) of an (h, w, c) float32 texture at (u, v)."""
    data = np.ascontiguousarray(data, np.float32)
    h, w, c = data.shape
    u = np.ascontiguousarray(u, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    out = np.zeros(u.shape[0], np.float32)
    lib().or_tex_lookup(w, h, c, data.ctypes.data, u.shape[0], u.ctypes.data, v.ctypes.data, ch, out.ctypes.data)
    return out


def set_shadow_short(on):
    """Test-only: end NEE shadow segments ShadowRayEps before the light point."""
    lib().or_set_shadow_short(1 if on else 0)

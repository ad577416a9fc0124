"""ctypes binding of libgyunity.so (the C ABI in include/gyunity.h).

Loads the in-tree library next to this file and raises if it is missing:
there is no fallback path.
"""
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgyunity.so")

GY_OK = 0
GY_ERR_INVALID = 1
GY_ERR_DEVICE = 2
GY_ERR_ALLOC = 3
GY_ERR_UNSUPPORTED = 4
GY_ERR_INJECTED = 5
MAT_WORDS = 32
GY_SCENE_CORNELL = 0
GY_SCENE_ICOSPHERE_GGX = 1
GY_RENDER_MEGAKERNEL = 0
GY_RENDER_WAVEFRONT = 1
GY_SAMPLER_SOBOL = 0
GY_SCENE_BATCH_MEMBER = 2
GY_SCENE_ADJOINT = 3
GY_SCENE_ICOSPHERE_TEXTURED = 4
GY_SAMPLER_HALTON = 1
GY_RAY_SORT_NONE = 0
GY_RAY_SORT_OCTANT = 1
GY_RAY_SORT_OCTANT_LOCAL = 2
FAULT_NONE, FAULT_ALLOC, FAULT_LAUNCH, FAULT_SCENE = 0, 1, 2, 3


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_tri", C.c_int32),
        ("pos", C.POINTER(C.c_float)),
        ("nrm", C.POINTER(C.c_float)),
        ("face_n", C.POINTER(C.c_float)),
        ("uv", C.POINTER(C.c_float)),
        ("mat", C.POINTER(C.c_uint32)),
        ("light", C.POINTER(C.c_int32)),
        ("n_mat", C.c_int32),
        ("mats", C.POINTER(C.c_uint32)),
        ("n_light", C.c_int32),
        ("light_tri", C.POINTER(C.c_int32)),
        ("light_le", C.POINTER(C.c_float)),
        ("n_tex", C.c_int32),
        ("tex_info", C.POINTER(C.c_int32)),
        ("tex_data", C.POINTER(C.c_float)),
    ]


class CameraDesc(C.Structure):
    _fields_ = [
        ("pos", C.c_float * 3),
        ("look_at", C.c_float * 3),
        ("up", C.c_float * 3),
        ("fovy", C.c_float),
        ("film_dist", C.c_float),
        ("width", C.c_int32),
        ("height", C.c_int32),
    ]


class RenderParams(C.Structure):
    _fields_ = [
        ("spp", C.c_int32),
        ("max_depth", C.c_int32),
        ("rr_depth", C.c_int32),
        ("mode", C.c_int32),
        ("samples_per_pass", C.c_int32),
        ("block_size", C.c_int32),
        ("refill_threshold", C.c_int32),
        ("sampler", C.c_int32),
        ("ray_sort", C.c_int32),
    ]


class BatchParams(C.Structure):
    _fields_ = [("lanes_per_pass", C.c_int64), ("refill_threshold", C.c_int32), ("use_graph", C.c_int32),
                ("ray_sort", C.c_int32), ("reserved", C.c_int32)]


class BVHParams(C.Structure):
    _fields_ = [("max_leaf_size", C.c_int32), ("builder", C.c_int32), ("reserved", C.c_int32 * 2)]


class SceneStats(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int64),
        ("n_leaves", C.c_int64),
        ("max_depth", C.c_int32),
        ("max_leaf_size", C.c_int32),
        ("build_ms", C.c_double),
        ("sah_cost", C.c_double),
    ]


class SppmParams(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("max_depth", C.c_int32),
        ("photons_per_iteration", C.c_int64),
        ("initial_radius", C.c_float),
        ("alpha", C.c_float),
        ("batch_shrink", C.c_int32),
        ("sampler", C.c_int32),
        ("trace_glossy_ray", C.c_int32),
        ("photons_per_launch", C.c_int64),
    ]


def sppm_params(iterations, photons_per_iteration, max_depth, initial_radius, alpha, batch_shrink=False, sampler=0,
                photons_per_launch=0):
    p = SppmParams()
    p.iterations, p.photons_per_iteration, p.max_depth = iterations, photons_per_iteration, max_depth
    p.initial_radius, p.alpha, p.batch_shrink, p.sampler = initial_radius, alpha, 1 if batch_shrink else 0, sampler
    p.photons_per_launch = photons_per_launch
    return p


class RenderStats(C.Structure):
    _fields_ = [
        ("paths", C.c_int64),
        ("rays", C.c_int64),
        ("shadow_rays", C.c_int64),
        ("mis_rays", C.c_int64),
        ("total_ms", C.c_double),
        ("kernel_ms", C.c_double * 8),
        ("kernel_launches", C.c_int64 * 8),
        ("node_visits", C.c_int64),
        ("tri_visits", C.c_int64),
        ("shade_points", C.c_int64),
        ("node_visits_kind", C.c_int64 * 3),
        ("tri_visits_kind", C.c_int64 * 3),
    ]


RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim", "<i4"), ("b1", "<f4"), ("b2", "<f4")])
# gy_medium (HomogeneousMedium) and gy_tri_medium (per-triangle MediumInterface,
# -1: none, and the no-material flag of a medium boundary)
MEDIUM_DTYPE = np.dtype([("sigma_a", "<f4", 3), ("sigma_s", "<f4", 3), ("g", "<f4")])
TRI_MEDIUM_DTYPE = np.dtype([("inside", "<i4"), ("outside", "<i4"), ("no_material", "<i4")])
# gy_sphere: centre, radius, inside flag, material, emissive flag, emitted radiance
SPHERE_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("inside", "<i4"), ("material", "<u4"),
                         ("emissive", "<i4"), ("le", "<f4", 3)])

_lib = None
_torch_first = False


def lib():
    """Load libgyunity.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libgyunity.so is not built (expected at %s); run __graft_entry__.build() or make -C gyunity_amd" % LIB_PATH
        )
    global _torch_first
    # libgyunity.so and torch both need libamdhip64.so.7; whichever loads first
    # provides the HIP runtime for the process (see torch_render.py)
    _torch_first = "torch" in sys.modules
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, u32, u64, f32 = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_float
    P = C.POINTER
    sig = {
        "gy_open": (C.c_int, [C.c_int, P(vp)]),
        "gy_close": (C.c_int, [vp]),
        "gy_last_error": (C.c_char_p, []),
        "gy_version": (C.c_char_p, []),
        "gy_device_name": (C.c_int, [vp, C.c_char_p, C.c_int]),
        "gy_scene_create": (C.c_int, [vp, P(SceneDesc), P(BVHParams), P(vp)]),
        "gy_scene_destroy": (C.c_int, [vp]),
        "gy_scene_stats_get": (C.c_int, [vp, P(SceneStats)]),
        "gy_scene_bvh_download": (C.c_int, [vp, P(u32), i64, P(i32), i64]),
        "gy_trace_closest": (C.c_int, [vp, vp, i64, vp]),
        "gy_trace_any": (C.c_int, [vp, vp, i64, P(i32)]),
        "gy_trace_closest_device": (C.c_int, [vp, vp, i64, vp]),
        "gy_trace_any_device": (C.c_int, [vp, vp, i64, vp]),
        "gy_last_kernel_ms": (C.c_double, [vp]),
        "gy_count_visits": (C.c_int, [vp, vp, i64, P(i64), P(i64)]),
        "gy_render": (C.c_int, [vp, P(CameraDesc), P(RenderParams), P(f32), P(f32)]),
        "gy_render_noalloc": (C.c_int, [vp, P(CameraDesc), P(RenderParams), vp]),
        "gy_film_rmse": (C.c_int, [vp, P(f32), P(i32), i32, P(C.c_double)]),
        "gy_film_download": (C.c_int, [vp, P(f32), P(f32)]),
        "gy_scene_texture_floats": (i64, [vp]),
        "gy_scene_context": (C.c_int, [vp, P(vp)]),
        "gy_autotune": (C.c_int, [vp, P(CameraDesc), P(RenderParams), i32, P(RenderParams), P(C.c_double)]),
        "gy_autotune_clear_cache": (C.c_int, []),
        "gy_ggx_dlogf_dalpha": (f32, [f32, P(f32), P(f32)]),
        "gy_dalpha_droughness": (f32, [f32]),
        "gy_scene_set_textures": (C.c_int, [vp, P(f32), i64]),
        "gy_scene_set_textures_device": (C.c_int, [vp, vp, i64]),
        "gy_adjoint": (C.c_int, [vp, P(CameraDesc), P(RenderParams), P(f32), P(f32)]),
        "gy_adjoint_device": (C.c_int, [vp, P(CameraDesc), P(RenderParams), vp, vp]),
        "gy_adjoint_emission": (C.c_int, [vp, P(CameraDesc), P(RenderParams), P(f32), P(f32), P(f32)]),
        "gy_adjoint_emission_device": (C.c_int, [vp, P(CameraDesc), P(RenderParams), vp, vp, vp]),
        "gy_batch_image_floats": (i64, [vp]),
        "gy_batch_images_device": (C.c_int, [vp, vp]),
        "gy_batch_adjoint": (C.c_int, [vp, P(f32), P(f32), P(f32)]),
        "gy_batch_adjoint_device": (C.c_int, [vp, vp, vp, vp]),
        "gy_batch_create": (C.c_int, [vp, P(BatchParams), P(vp)]),
        "gy_batch_add": (C.c_int, [vp, vp, P(CameraDesc), P(RenderParams), P(i32)]),
        "gy_batch_render": (C.c_int, [vp]),
        "gy_batch_film_download": (C.c_int, [vp, i32, P(f32), P(f32)]),
        "gy_batch_destroy": (C.c_int, [vp]),
        "gy_batch_autotune": (C.c_int, [vp, P(BatchParams), P(C.c_double)]),
        "gy_render_stats_get": (C.c_int, [vp, P(RenderStats)]),
        "gy_kernel_name": (C.c_char_p, [C.c_int]),
        "gy_li_samples": (C.c_int, [vp, P(CameraDesc), P(RenderParams), i32, P(i32), P(i32), P(i32), P(f32), P(f32), P(i32)]),
        "gy_scene_set_media": (C.c_int, [vp, i32, vp, vp]),
        "gy_render_sppm": (C.c_int, [vp, P(CameraDesc), P(SppmParams), vp, vp, vp]),
        "gy_render_vpt": (C.c_int, [vp, P(CameraDesc), P(RenderParams), i32, vp, vp]),
        "gy_vpt_li_samples": (C.c_int, [vp, P(CameraDesc), P(RenderParams), i32, i32, P(i32), P(i32), P(i32), P(f32), P(f32),
                                        P(i32)]),
        "gy_count_traffic": (C.c_int, [vp, P(CameraDesc), P(RenderParams), P(RenderStats)]),
        "gy_copy_probe": (C.c_int, [vp, i64, P(C.c_double)]),
        "gy_sobol_sample": (f32, [u64, u32]),
        "gy_halton_sample": (f32, [u64, u32]),
        "gy_halton_enum_index": (u64, [u32, u32, u64, u32, u32]),
        "gy_sobol_enum_index": (u64, [u32, u32, u32, u32]),
        "gy_camera_matrix": (C.c_int, [P(CameraDesc), P(f32)]),
        "gy_generate_scene": (C.c_int, [C.c_int, u32, i32, P(vp)]),
        "gy_generated_desc": (C.c_int, [vp, P(SceneDesc), P(CameraDesc), P(RenderParams)]),
        "gy_generated_free": (C.c_int, [vp]),
        "gy_debug_inject_fault": (C.c_int, [vp, C.c_int]),
        "gy_image_write": (C.c_int, [C.c_char_p, P(f32), i32, i32]),
        "gy_image_to_bytes": (C.c_int, [P(f32), i64, P(C.c_uint8)]),
        "gy_image_read": (C.c_int, [C.c_char_p, P(i32), P(i32), vp, i64]),
        "gy_film_save": (C.c_int, [vp, C.c_char_p]),
        "gy_scene_set_environment": (C.c_int, [vp, vp, C.c_int32, C.c_int32, C.c_float]),
        "gy_scene_set_spheres": (C.c_int, [vp, C.c_int32, vp]),
        "gy_render_light_tracing": (C.c_int, [vp, P(CameraDesc), P(RenderParams), vp]),
        "gy_mesh_load_obj": (C.c_int, [C.c_char_p, i32, P(f32), P(f32), P(vp)]),
        "gy_mesh_triangles": (i64, [vp]),
        "gy_mesh_arrays": (C.c_int, [vp, P(f32), P(f32), P(f32), P(f32)]),
        "gy_mesh_free": (C.c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class GyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gyunity error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc != GY_OK:
        raise GyError(rc, lib().gy_last_error().decode())
    return rc


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class SceneArrays:
    """Host copy of a flattened scene (the arrays a gy_scene_desc points at)."""

    def __init__(self, pos, nrm, face_n, mat, light, mats, light_tri, light_le, uv=None, tex_info=None, tex_data=None):
        self.pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 9)
        self.nrm = np.ascontiguousarray(nrm, dtype=np.float32).reshape(-1, 9)
        self.face_n = np.ascontiguousarray(face_n, dtype=np.float32).reshape(-1, 3)
        self.mat = np.ascontiguousarray(mat, dtype=np.uint32)
        self.light = np.ascontiguousarray(light, dtype=np.int32)
        self.mats = np.ascontiguousarray(mats, dtype=np.uint32).reshape(-1, MAT_WORDS)
        self.light_tri = np.ascontiguousarray(light_tri, dtype=np.int32)
        self.light_le = np.ascontiguousarray(light_le, dtype=np.float32).reshape(-1, 3)
        self.uv = None if uv is None else np.ascontiguousarray(uv, dtype=np.float32).reshape(-1, 6)
        self.tex_info = None if tex_info is None else np.ascontiguousarray(tex_info, dtype=np.int32).reshape(-1, 4)
        self.tex_data = None if tex_data is None else np.ascontiguousarray(tex_data, dtype=np.float32)

    @property
    def n_tri(self):
        return self.pos.shape[0]

    def desc(self):
        d = SceneDesc()
        d.n_tri = self.n_tri
        d.pos, d.nrm, d.face_n = _fp(self.pos), _fp(self.nrm), _fp(self.face_n)
        d.uv = _fp(self.uv) if self.uv is not None else C.POINTER(C.c_float)()
        d.mat, d.light = _up(self.mat), _ip(self.light)
        d.n_mat, d.mats = self.mats.shape[0], _up(self.mats)
        d.n_light, d.light_tri, d.light_le = self.light_tri.shape[0], _ip(self.light_tri), _fp(self.light_le)
        if self.tex_info is not None:
            d.n_tex, d.tex_info, d.tex_data = self.tex_info.shape[0], _ip(self.tex_info), _fp(self.tex_data)
        else:
            d.n_tex = 0
        return d


def generate_scene(kind, seed=0, param=0):
    """Procedural scene -> (SceneArrays, CameraDesc, RenderParams)."""
    L = lib()
    h = C.c_void_p()
    check(L.gy_generate_scene(kind, seed, param, C.byref(h)))
    try:
        d, cam, prm = SceneDesc(), CameraDesc(), RenderParams()
        check(L.gy_generated_desc(h, C.byref(d), C.byref(cam), C.byref(prm)))
        n = d.n_tri
        arr = lambda p, cnt, dt: np.ctypeslib.as_array(p, shape=(cnt,)).astype(dt, copy=True)
        sa = SceneArrays(
            arr(d.pos, 9 * n, np.float32),
            arr(d.nrm, 9 * n, np.float32),
            arr(d.face_n, 3 * n, np.float32),
            arr(d.mat, n, np.uint32),
            arr(d.light, n, np.int32),
            arr(d.mats, MAT_WORDS * d.n_mat, np.uint32),
            arr(d.light_tri, d.n_light, np.int32),
            arr(d.light_le, 3 * d.n_light, np.float32),
            uv=arr(d.uv, 6 * n, np.float32) if bool(d.uv) else None,
            tex_info=arr(d.tex_info, 4 * d.n_tex, np.int32) if d.n_tex > 0 else None,
        )
        if d.n_tex > 0:
            ti = sa.tex_info
            total = int((ti[:, 3] + ti[:, 0] * ti[:, 1] * ti[:, 2]).max())
            sa.tex_data = arr(d.tex_data, total, np.float32)
    finally:
        L.gy_generated_free(h)
    return sa, cam, prm


def camera(pos, look_at, up, fovy, width, height, film_dist=1.0):
    c = CameraDesc()
    c.pos[:], c.look_at[:], c.up[:] = pos, look_at, up
    c.fovy, c.film_dist, c.width, c.height = fovy, film_dist, width, height
    return c


def orbit_views(n, res, seed=3, centre=(0.5, 0.4, 0.5), radius=0.45, fovy=60.0):
    """n seeded pinhole views of res x res looking at `centre` from points on a
    sphere of `radius` about it that stay inside the unit room."""
    rng = np.random.default_rng(seed)
    c = np.asarray(centre, np.float64)
    views = []
    while len(views) < n:
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        pos = c + radius * d
        if (pos < 0.04).any() or (pos > 0.96).any():
            continue
        up = (0.0, 1.0, 0.0) if abs(d[1]) < 0.9 else (1.0, 0.0, 0.0)
        views.append(camera(tuple(pos), tuple(c), up, fovy, res, res))
    return views


def render_params(spp, max_depth, rr_depth=5, mode=GY_RENDER_MEGAKERNEL, samples_per_pass=0, block_size=0,
                  sampler=GY_SAMPLER_SOBOL, ray_sort=GY_RAY_SORT_NONE):
    p = RenderParams()
    p.spp, p.max_depth, p.rr_depth, p.mode = spp, max_depth, rr_depth, mode
    p.samples_per_pass, p.block_size, p.sampler = samples_per_pass, block_size, sampler
    p.ray_sort = ray_sort
    return p


def autotune_clear_cache():
    """Forget the autotuner's kept choices (gy_autotune_clear_cache)."""
    check(lib().gy_autotune_clear_cache())


def halton_sample(index, dim):
    return lib().gy_halton_sample(index, dim)


def halton_enum_index(width, height, i, x, y):
    return lib().gy_halton_enum_index(width, height, i, x, y)


def sobol_sample(index, dim):
    return lib().gy_sobol_sample(index, dim)


def sobol_enum_index(m, frame, px, py):
    return lib().gy_sobol_enum_index(m, frame, px, py)


def load_obj(path, reverse=False, object_to_world=None, world_to_object=None):
    """OBJ file -> dict of per-triangle float32 arrays in the gy_scene_desc
    layout: pos (n, 9), nrm (n, 9), face_n (n, 3), uv (n, 6).  object_to_world
    and world_to_object are row-major 4x4 (None: identity, and the inverse)."""
    L = lib()
    m = None if object_to_world is None else np.ascontiguousarray(object_to_world, dtype=np.float32).reshape(16)
    mi = None if world_to_object is None else np.ascontiguousarray(world_to_object, dtype=np.float32).reshape(16)
    h = C.c_void_p()
    check(L.gy_mesh_load_obj(os.fsencode(path), int(reverse), None if m is None else _fp(m), None if mi is None else _fp(mi),
                             C.byref(h)))
    try:
        n = L.gy_mesh_triangles(h)
        out = {"pos": np.zeros((n, 9), np.float32), "nrm": np.zeros((n, 9), np.float32),
               "face_n": np.zeros((n, 3), np.float32), "uv": np.zeros((n, 6), np.float32)}
        check(L.gy_mesh_arrays(h, _fp(out["pos"]), _fp(out["nrm"]), _fp(out["face_n"]), _fp(out["uv"])))
    finally:
        L.gy_mesh_free(h)
    return out


def image_write(path, rgb):
    """Write an (H, W, 3) float image as .hdr, .bmp or .png (by suffix), the way
    the reference's Film::SaveImage does."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    h, w = rgb.shape[:2]
    check(lib().gy_image_write(os.fsencode(path), _fp(rgb), w, h))


def image_read(path):
    """A Radiance .hdr file as (height, width, 3) float32, top row first, as
    stbi_loadf reads it for ImageIO::LoadTexture (gy_image_read)."""
    w, h = C.c_int32(), C.c_int32()
    check(lib().gy_image_read(os.fsencode(path), C.byref(w), C.byref(h), None, 0))
    out = np.zeros((h.value, w.value, 3), np.float32)
    check(lib().gy_image_read(os.fsencode(path), C.byref(w), C.byref(h), out.ctypes.data, out.size))
    return out


def image_to_bytes(rgb):
    """8-bit values of the reference's ImageIO::toInt (gamma 2.2)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    out = np.zeros(rgb.shape, np.uint8)
    check(lib().gy_image_to_bytes(_fp(rgb), rgb.size, out.ctypes.data_as(C.POINTER(C.c_uint8))))
    return out


def camera_matrix(cam):
    out = np.zeros(16, dtype=np.float32)
    check(lib().gy_camera_matrix(C.byref(cam), _fp(out)))
    return out


class Context:
    """One HIP device and stream (gy_open / gy_close)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(lib().gy_open(device, C.byref(self.h)))

    def close(self):
        if self.h:
            lib().gy_close(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_name(self):
        buf = C.create_string_buffer(256)
        check(lib().gy_device_name(self.h, buf, 256))
        return buf.value.decode()

    def stats(self):
        s = RenderStats()
        check(lib().gy_render_stats_get(self.h, C.byref(s)))
        return s

    def copy_probe(self, nbytes):
        """Device time (ms) of one plain copy of nbytes between device buffers."""
        ms = C.c_double()
        check(lib().gy_copy_probe(self.h, int(nbytes), C.byref(ms)))
        return ms.value

    def last_kernel_ms(self):
        return lib().gy_last_kernel_ms(self.h)

    def inject_fault(self, kind):
        check(lib().gy_debug_inject_fault(self.h, kind))


class Batch:
    """Many renders traced together (gy_batch_*): add (scene, camera, params)
    entries, render them all in one lane space, download each entry's film."""

    def __init__(self, ctx, lanes_per_pass=0, refill_threshold=0, use_graph=True, ray_sort=GY_RAY_SORT_NONE):
        self.ctx = ctx
        bp = BatchParams()
        bp.lanes_per_pass, bp.refill_threshold, bp.use_graph = lanes_per_pass, refill_threshold, 1 if use_graph else 0
        bp.ray_sort = ray_sort
        self.h = C.c_void_p()
        self.cams = []
        check(lib().gy_batch_create(ctx.h, C.byref(bp), C.byref(self.h)))

    def add(self, scene, cam, params):
        e = C.c_int32()
        check(lib().gy_batch_add(self.h, scene.h, C.byref(cam), C.byref(params), C.byref(e)))
        self.cams.append((cam.width, cam.height))
        return e.value

    def render(self):
        check(lib().gy_batch_render(self.h))

    def autotune(self):
        """Fastest refill threshold and ray-sort key for the whole batch
        (gy_batch_autotune), kept in the batch: (BatchParams, ms per render)."""
        best = BatchParams()
        ms = C.c_double()
        check(lib().gy_batch_autotune(self.h, C.byref(best), C.byref(ms)))
        return best, ms.value

    def film(self, entry):
        W, H = self.cams[entry]
        color = np.zeros((H, W, 3), dtype=np.float32)
        weight = np.zeros((H, W), dtype=np.float32)
        check(lib().gy_batch_film_download(self.h, entry, _fp(color), _fp(weight)))
        return color, weight

    def adjoint(self, dimages, n_tex, n_light=0):
        """Path-replay adjoint of every entry (all views of one scene) in one
        launch: dimages is the list of the entries' dl/dI (H, W, 3); returns the
        texture gradient (n_tex floats) and, for n_light > 0, the emission
        gradient (n_light, 3)."""
        flat = np.ascontiguousarray(np.concatenate([np.asarray(d, np.float32).ravel() for d in dimages]))
        assert flat.size == lib().gy_batch_image_floats(self.h)
        grad = np.zeros(n_tex, dtype=np.float32)
        gle = np.zeros((max(n_light, 0), 3), dtype=np.float32)
        check(lib().gy_batch_adjoint(self.h, _fp(flat), _fp(grad), _fp(gle) if n_light > 0 else C.POINTER(C.c_float)()))
        return (grad, gle) if n_light > 0 else grad

    def close(self):
        if self.h:
            lib().gy_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Scene:
    """A device scene with its BVH (gy_scene_create)."""

    def __init__(self, ctx, arrays, max_leaf_size=0, builder=0):
        self.ctx = ctx
        self.arrays = arrays
        self._desc = arrays.desc()
        bp = BVHParams()
        bp.max_leaf_size, bp.builder = max_leaf_size, builder
        self.h = C.c_void_p()
        self.n_light = int(arrays.light_tri.shape[0])  # area lights
        check(lib().gy_scene_create(ctx.h, C.byref(self._desc), C.byref(bp), C.byref(self.h)))

    def close(self):
        """gy_scene_destroy; raises (and keeps the scene) while a batch holds it."""
        if self.h:
            check(lib().gy_scene_destroy(self.h))
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def stats(self):
        s = SceneStats()
        check(lib().gy_scene_stats_get(self.h, C.byref(s)))
        return s

    def bvh(self):
        st = self.stats()
        nodes = np.zeros((st.n_nodes, 16), dtype=np.uint32)
        order = np.zeros(self.arrays.n_tri, dtype=np.int32)
        check(lib().gy_scene_bvh_download(self.h, _up(nodes), st.n_nodes, _ip(order), order.size))
        return nodes, order

    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        hits = np.zeros(rays.shape[0], dtype=HIT_DTYPE)
        check(lib().gy_trace_closest(self.h, rays.ctypes.data, rays.shape[0], hits.ctypes.data))
        return hits

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        occ = np.zeros(rays.shape[0], dtype=np.int32)
        check(lib().gy_trace_any(self.h, rays.ctypes.data, rays.shape[0], _ip(occ)))
        return occ

    def count_visits(self, rays):
        rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
        a, b = C.c_int64(), C.c_int64()
        check(lib().gy_count_visits(self.h, rays.ctypes.data, rays.shape[0], C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_environment(self, rgb, scale=1.0):
        """EnvironmentLight over a latitude-longitude map, added after the area
        lights (gy_scene_set_environment); rgb is (height, width, 3), top row first."""
        rgb = np.ascontiguousarray(rgb, dtype=np.float32)
        if rgb.ndim != 3 or rgb.shape[2] != 3:
            raise GyError(1, "set_environment: rgb must be (height, width, 3)")
        check(lib().gy_scene_set_environment(self.h, rgb.ctypes.data, rgb.shape[1], rgb.shape[0], float(scale)))

    def set_spheres(self, spheres):
        """Analytic spheres after the triangles (gy_scene_set_spheres); spheres
        is an array of SPHERE_DTYPE.  Emissive ones become area lights after the
        scene's own."""
        sp = np.ascontiguousarray(spheres, dtype=SPHERE_DTYPE)
        check(lib().gy_scene_set_spheres(self.h, sp.shape[0], sp.ctypes.data))
        self.n_light += int(np.count_nonzero(sp["emissive"]))

    def render_light_tracing(self, cam, params):
        """LightTracing::Render (gy_render_light_tracing): the (H, W, 3) splat sums."""
        out = np.zeros((cam.height, cam.width, 3), dtype=np.float32)
        check(lib().gy_render_light_tracing(self.h, C.byref(cam), C.byref(params), out.ctypes.data))
        return out

    def save_film(self, path):
        """Film::SaveImage of the last render: .hdr, .bmp or .png by suffix."""
        check(lib().gy_film_save(self.h, os.fsencode(path)))

    def render(self, cam, params, download=True):
        W, H = cam.width, cam.height
        if download:
            color = np.zeros((H, W, 3), dtype=np.float32)
            weight = np.zeros((H, W), dtype=np.float32)
            check(lib().gy_render(self.h, C.byref(cam), C.byref(params), _fp(color), _fp(weight)))
            return color, weight
        check(lib().gy_render(self.h, C.byref(cam), C.byref(params), C.POINTER(C.c_float)(), C.POINTER(C.c_float)()))
        return None

    def film_rmse(self, reference, rows=None):
        """RMSE of the last render's pixels against `reference` (the given
        image rows, or the whole image) over every channel (gy_film_rmse)."""
        ref = np.ascontiguousarray(reference, dtype=np.float32)
        out = C.c_double()
        if rows is None:
            check(lib().gy_film_rmse(self.h, _fp(ref), C.POINTER(C.c_int32)(), 0, C.byref(out)))
        else:
            r = np.ascontiguousarray(rows, dtype=np.int32)
            check(lib().gy_film_rmse(self.h, _fp(ref), _ip(r), r.shape[0], C.byref(out)))
        return out.value

    def autotune(self, cam, params, pilot_spp=0):
        """Fastest launch choice for this view (gy_autotune): (params, ms per Mpath)."""
        best = RenderParams()
        ms = C.c_double()
        check(lib().gy_autotune(self.h, C.byref(cam), C.byref(params), pilot_spp, C.byref(best), C.byref(ms)))
        return best, ms.value

    def texture_floats(self):
        return lib().gy_scene_texture_floats(self.h)

    def set_textures(self, data):
        data = np.ascontiguousarray(data, dtype=np.float32)
        check(lib().gy_scene_set_textures(self.h, _fp(data), data.size))

    def adjoint(self, cam, params, dimage, emission=False):
        """Gradient of a loss w.r.t. the texture floats, given dl/dI (H, W, 3)
        for the film of the last render of this camera and these parameters.
        emission=True: also w.r.t. each area light's Le, (n_light, 3)."""
        dimage = np.ascontiguousarray(dimage, dtype=np.float32)
        assert dimage.shape == (cam.height, cam.width, 3)
        grad = np.zeros(self.texture_floats(), dtype=np.float32)
        if not emission:
            check(lib().gy_adjoint(self.h, C.byref(cam), C.byref(params), _fp(dimage), _fp(grad)))
            return grad
        gle = np.zeros((self.n_light, 3), dtype=np.float32)
        check(lib().gy_adjoint_emission(self.h, C.byref(cam), C.byref(params), _fp(dimage), _fp(grad), _fp(gle)))
        return grad, gle

    def film(self, W, H):
        color = np.zeros((H, W, 3), dtype=np.float32)
        weight = np.zeros((H, W), dtype=np.float32)
        check(lib().gy_film_download(self.h, _fp(color), _fp(weight)))
        return color, weight

    def li_samples(self, cam, params, x, y, s):
        x = np.ascontiguousarray(x, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        s = np.ascontiguousarray(s, dtype=np.int32)
        n = x.shape[0]
        L = np.zeros((n, 3), dtype=np.float32)
        raster = np.zeros((n, 2), dtype=np.float32)
        dims = np.zeros(n, dtype=np.int32)
        check(lib().gy_li_samples(self.h, C.byref(cam), C.byref(params), n, _ip(x), _ip(y), _ip(s), _fp(L), _fp(raster), _ip(dims)))
        return L, raster, dims

    def set_media(self, media, tri):
        """gy_scene_set_media: media (MEDIUM_DTYPE) and one TRI_MEDIUM_DTYPE
        record per triangle."""
        self._media = np.ascontiguousarray(media, dtype=MEDIUM_DTYPE)
        self._tri_media = np.ascontiguousarray(tri, dtype=TRI_MEDIUM_DTYPE)
        if self._tri_media.shape[0] != self.arrays.n_tri:
            raise GyError(1, "set_media: one record per triangle")
        check(lib().gy_scene_set_media(self.h, self._media.shape[0], self._media.ctypes.data, self._tri_media.ctypes.data))

    def render_vpt(self, cam, params, equiangular=True):
        """VolPathTracing (gy_render_vpt): (color (H, W, 3), weight (H, W))."""
        W, H = cam.width, cam.height
        color = np.zeros((H, W, 3), dtype=np.float32)
        weight = np.zeros((H, W), dtype=np.float32)
        check(lib().gy_render_vpt(self.h, C.byref(cam), C.byref(params), 1 if equiangular else 0, color.ctypes.data,
                                  weight.ctypes.data))
        return color, weight

    def render_sppm(self, cam, sp):
        """SPPM (gy_render_sppm): (image (H, W, 3), radius^2 (H, W), photon count N (H, W))."""
        W, H = cam.width, cam.height
        img = np.zeros((H, W, 3), dtype=np.float32)
        r2 = np.zeros((H, W), dtype=np.float32)
        nn = np.zeros((H, W), dtype=np.float32)
        check(lib().gy_render_sppm(self.h, C.byref(cam), C.byref(sp), img.ctypes.data, r2.ctypes.data, nn.ctypes.data))
        return img, r2, nn

    def vpt_li_samples(self, cam, params, x, y, s, equiangular=True):
        x = np.ascontiguousarray(x, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.int32)
        s = np.ascontiguousarray(s, dtype=np.int32)
        n = x.shape[0]
        L = np.zeros((n, 3), dtype=np.float32)
        raster = np.zeros((n, 2), dtype=np.float32)
        dims = np.zeros(n, dtype=np.int32)
        check(lib().gy_vpt_li_samples(self.h, C.byref(cam), C.byref(params), 1 if equiangular else 0, n, _ip(x), _ip(y), _ip(s),
                                      _fp(L), _fp(raster), _ip(dims)))
        return L, raster, dims

    def count_traffic(self, cam, params):
        st = RenderStats()
        check(lib().gy_count_traffic(self.h, C.byref(cam), C.byref(params), C.byref(st)))
        return st

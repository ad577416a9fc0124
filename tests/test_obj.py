"""OBJ ingest (gy_mesh_load_obj): the triangles a Gyunity scene holds for an
OBJ file after Mesh::LoadFromFile (mesh/Mesh.cpp:7-84) and Scene::AddMesh
(visual/Scene.h:68-77).

- The parse is compared value for value with the reference's own loader,
  tinyobjloader 2.0.0 (called at mesh/Mesh.cpp:13). oracle/Makefile builds it
  from the reference tree into oracle/_ref/obj_ref. The comparison covers the
  reference's five meshes (resources/meshs) and a file of number formats.
  These tests skip where the reference tree or the build is absent.
- The per-face expansion (mesh/Mesh.cpp:30-80) is restated in float32 numpy
  on the loader's output.
- The transforms (shape/Triangle.h:78-88) are checked against closed forms.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import gyunity_amd as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_MESHES = "/root/reference/resources/meshs"
OBJ_REF = os.path.join(ROOT, "oracle", "_ref", "obj_ref")
f32 = np.float32


def _ref_parse(path, tmpdir):
    out = os.path.join(tmpdir, "ref.bin")
    subprocess.check_call([OBJ_REF, path, out])
    raw = open(out, "rb").read()
    nv, nn, nt, ntri = np.frombuffer(raw, np.int64, 4)
    o = 32
    v = np.frombuffer(raw, np.float32, 3 * nv, o).reshape(-1, 3); o += 12 * nv
    vn = np.frombuffer(raw, np.float32, 3 * nn, o).reshape(-1, 3); o += 12 * nn
    vt = np.frombuffer(raw, np.float32, 2 * nt, o).reshape(-1, 2); o += 8 * nt
    corners = np.frombuffer(raw, np.int32, 9 * ntri, o).reshape(-1, 3, 3)  # (tri, corner, [v, vt, vn])
    return v, vn, vt, corners


def _dot(a, b):  # DPPS order (x x' + y y') + (z z' + 0)
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + (a[..., 2] * b[..., 2] + f32(0))


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)


def _norm(a):
    with np.errstate(invalid="ignore", divide="ignore"):
        return a / np.sqrt(_dot(a, a))[..., None]


def _ident(a, w=1.0):
    """The identity matrix applied to (a, w) in the SSE column order
    ((c3 w + c2 z) + c1 y) + c0 x, lane by lane (math/Linagl.h:937-948)."""
    vec = [a[..., 0], a[..., 1], a[..., 2], np.full(a.shape[:-1], w, f32)]
    out = []
    for i in range(3):
        c = [f32(1) if j == i else f32(0) for j in range(4)]  # row i of columns 0..3
        r = c[3] * vec[3]
        r = c[2] * vec[2] + r
        r = c[1] * vec[1] + r
        r = c[0] * vec[0] + r
        out.append(r)
    return np.stack(out, -1)


def _expected(v, vn, vt, corners):
    """Mesh::LoadFromFile + Triangle::SetTransform(identity) in float32."""
    p = v[corners[..., 0]]
    has_n = corners[..., 2] >= 0
    assert np.all(has_n.all(axis=1) | (~has_n).all(axis=1))
    gen = _cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    with np.errstate(invalid="ignore", divide="ignore"):
        ln = np.sqrt(_dot(gen, gen))
        gen = np.where((ln > f32(1e-6))[:, None], gen / ln[:, None], gen)
    n = np.where(has_n[..., None], vn[np.maximum(corners[..., 2], 0)] if len(vn) else f32(0), gen[:, None, :])
    uv = np.zeros(corners.shape[:2] + (2,), f32)
    if len(vt):
        uv = np.where((corners[..., 1] >= 0)[..., None], vt[np.maximum(corners[..., 1], 0)], uv)
    q = _ident(p)
    nn = _norm(_ident(n, 0.0))
    fn = _norm(_cross(q[:, 1] - q[:, 0], q[:, 2] - q[:, 1]))
    return {"pos": q.reshape(-1, 9), "nrm": nn.reshape(-1, 9), "face_n": fn, "uv": uv.astype(f32).reshape(-1, 6)}


def _same_bits(a, b):
    a, b = np.asarray(a, f32), np.asarray(b, f32)
    na, nb = np.isnan(a), np.isnan(b)
    return a.shape == b.shape and np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


@pytest.mark.parametrize("name", ["cube", "bunny", "bunny2", "water", "venusm"])
def test_obj_matches_reference_loader(name, tmp_path):
    path = os.path.join(REF_MESHES, name + ".obj")
    if not (os.path.exists(path) and os.path.exists(OBJ_REF)):
        pytest.skip("the reference tree (and oracle/_ref/obj_ref built from it) is not present")
    v, vn, vt, corners = _ref_parse(path, str(tmp_path))
    got = g.load_obj(path)
    exp = _expected(v, vn, vt, corners)
    assert got["pos"].shape[0] == corners.shape[0] > 0
    for k in ("pos", "uv", "nrm", "face_n"):
        assert _same_bits(got[k], exp[k]), k


def test_obj_number_formats_match_reference_loader(tmp_path):
    if not os.path.exists(OBJ_REF):
        pytest.skip("oracle/_ref/obj_ref is built only where the reference tree is present")
    rng = np.random.default_rng(7)
    vals = []
    for _ in range(3000):
        x = float(rng.standard_normal() * 10.0 ** rng.integers(-8, 8))
        form = rng.integers(0, 8)
        s = ["%.7e" % x, "%.9g" % x, "%.17g" % x, "%f" % x, "%.3E" % x, "%d" % int(x), "%.12f" % x, "%g" % x][form]
        vals.append(s)
    vals += ["+1.5", "-.25", ".5", "7.", "1e3", "1E-7", "-0", "0.000000000001234", "123456789012.5", "2.5e+0",
             "-3.40282e+38", "1e-45", "3"]
    while len(vals) % 9:
        vals.append("1")
    lines = []
    for i in range(0, len(vals), 3):
        lines.append("v %s %s %s" % tuple(vals[i:i + 3]))
    nvert = len(vals) // 3
    for i in range(0, nvert, 3):
        lines.append("f %d %d %d" % (i + 1, i + 2, i + 3))
    path = tmp_path / "numbers.obj"
    path.write_text("\n".join(lines) + "\n")
    v, vn, vt, corners = _ref_parse(str(path), str(tmp_path))
    got = g.load_obj(str(path))
    assert _same_bits(got["pos"], _ident(v[corners[..., 0]]).reshape(-1, 9))


def _write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_obj_quad_indices_uv_normals(tmp_path):
    # a convex quad with uv and normals, indexed v/t/n: ear clipping emits
    # (0, 1, 2) then (0, 2, 3); negative indices count back from the end
    path = _write(tmp_path, """# quad
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 0 1
f 1/1/1 2/2/1 3/3/1 4/4/1
f -4/-4/-1 -3/-3/-1 -2/-2/-1
""")
    m = g.load_obj(path)
    assert m["pos"].shape == (3, 9)
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], f32)
    UV = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], f32)
    for t, idx in enumerate([(0, 1, 2), (0, 2, 3), (0, 1, 2)]):
        assert np.array_equal(m["pos"][t].reshape(3, 3), P[list(idx)])
        assert np.array_equal(m["uv"][t].reshape(3, 2), UV[list(idx)])
        assert np.array_equal(m["nrm"][t].reshape(3, 3), np.tile([0, 0, 1], (3, 1)).astype(f32))
        assert np.array_equal(m["face_n"][t], np.array([0, 0, 1], f32))


def test_obj_generated_normals_and_reverse(tmp_path):
    path = _write(tmp_path, "v 0 0 0\nv 2 0 0\nv 0 3 0\nf 1 2 3\n")
    m = g.load_obj(path)
    assert np.array_equal(m["nrm"][0], np.tile([0, 0, 1], 3).astype(f32))  # Cross((2,0,0),(0,3,0)) = (0,0,6)
    assert np.array_equal(m["uv"][0], np.zeros(6, f32))
    r = g.load_obj(path, reverse=True)
    assert np.array_equal(r["pos"][0].reshape(3, 3), m["pos"][0].reshape(3, 3)[[0, 2, 1]])
    assert np.array_equal(r["nrm"][0], m["nrm"][0])           # generated before the reversal
    assert np.array_equal(r["face_n"][0], np.array([0, 0, -1], f32))


def test_obj_transform_closed_form(tmp_path):
    # a uniform power-of-two scale and a representable translation are exact;
    # normals go through the inverse transpose (n / 2), normalised: unchanged
    path = _write(tmp_path, "v 0.25 0.5 0\nv 1.5 0.5 0\nv 0.25 2 0\nvn 0 0 1\nf 1//1 2//1 3//1\n")
    M = np.array([[2, 0, 0, 1], [0, 2, 0, -0.5], [0, 0, 2, 4], [0, 0, 0, 1]], f32)
    m = g.load_obj(path, object_to_world=M)
    P = np.array([[0.25, 0.5, 0], [1.5, 0.5, 0], [0.25, 2, 0]], f32)
    assert np.array_equal(m["pos"][0].reshape(3, 3), 2 * P + np.array([1, -0.5, 4], f32))
    assert np.array_equal(m["nrm"][0], np.tile([0, 0, 1], 3).astype(f32))
    assert np.array_equal(m["face_n"][0], np.array([0, 0, 1], f32))
    # an explicit inverse is used as given (Transform(m, invM), math/Transform.h:47-50)
    Minv = np.linalg.inv(M.astype(np.float64)).astype(f32)
    m2 = g.load_obj(path, object_to_world=M, world_to_object=Minv)
    assert np.array_equal(m2["pos"], m["pos"]) and np.allclose(m2["nrm"], m["nrm"], atol=1e-7)


@pytest.mark.parametrize("text", ["v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",               # index 0
                                  "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nf 1//1 2 3\n",  # mixed normals
                                  "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 9\n"])               # out of range
def test_obj_rejects_bad_files(tmp_path, text):
    with pytest.raises(g.GyError):
        g.load_obj(_write(tmp_path, text))


def test_obj_missing_file():
    with pytest.raises(g.GyError):
        g.load_obj(os.path.join(tempfile.gettempdir(), "no_such_mesh_file.obj"))

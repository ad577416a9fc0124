"""Small hand-built scenes and ray batches shared by the tests."""
import os
import struct

import numpy as np

from gyunity_amd import RAY_DTYPE, SceneArrays

MAT_WORDS = 32


def f2w(x):
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


def material(mtype, R=(1, 1, 1), T=(1, 1, 1), alpha=(0.2, 0.2), alpha_direct=True, eta=(1, 1, 1), k=(0, 0, 0),
             f0=None, eta1=1.0, eta2=1.5, tex_albedo=-1, tex_rough=-1):
    w = np.zeros(MAT_WORDS, np.uint32)
    w[0] = mtype
    w[1] = (1 if alpha_direct else 0) | (2 if f0 is not None else 0)
    for i in range(3):
        w[2 + i] = f2w(R[i])
        w[5 + i] = f2w(T[i])
        w[10 + i] = f2w(eta[i])
        w[13 + i] = f2w(k[i])
        w[16 + i] = f2w(0 if f0 is None else f0[i])
    w[8], w[9] = f2w(alpha[0]), f2w(alpha[1])
    w[19], w[20] = f2w(eta1), f2w(eta2)
    w[21] = np.uint32(tex_albedo & 0xFFFFFFFF)
    w[22] = np.uint32(tex_rough & 0xFFFFFFFF)
    return w


def box_triangles(lo, hi, inward):
    """12 triangles of an axis-aligned box with face normals (inward or outward)."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    quads = [
        ((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1), (0, -1, 0)),
        ((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (0, 1, 0)),
        ((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0), (-1, 0, 0)),
        ((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1), (1, 0, 0)),
        ((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0), (0, 0, -1)),
        ((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1), (0, 0, 1)),
    ]
    tris, normals = [], []
    for a, b, c, d, n in quads:
        n = tuple(-v for v in n) if inward else n
        tris += [(a, b, c), (a, c, d)]
        normals += [n, n]
    return np.array(tris, np.float32).reshape(-1, 9), np.array(normals, np.float32)


def furnace_scene(albedo=0.5, le=1.0):
    """Closed unit box; every triangle is an inward-facing emitter with a Lambert
    material of the given albedo (a scalar, or one value per channel)."""
    pos, fn = box_triangles((0, 0, 0), (1, 1, 1), inward=True)
    n = pos.shape[0]
    nrm = np.repeat(fn, 3, axis=0).reshape(-1, 9)
    R = tuple(albedo) if np.ndim(albedo) else (albedo,) * 3
    mats = material(0, R=R)[None, :]
    return SceneArrays(pos, nrm, fn, np.zeros(n, np.uint32), np.arange(n, dtype=np.int32), mats,
                       np.arange(n, dtype=np.int32), np.full((n, 3), le, np.float32))


def furnace_expected(albedo, depth):
    """E[L] of PathTracing::Li in the furnace box: Le at the first hit (emission
    branch, no NEE there), then NEE + MIS estimates albedo*Le at every later
    vertex k >= 1, weighted by the throughput albedo^k."""
    return 1.0 + sum(albedo ** k for k in range(2, depth + 1))


def random_rays(n, seed, lo=(-0.2, -0.2, -0.2), hi=(1.2, 1.2, 1.2), tmax=np.inf):
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, RAY_DTYPE)
    rays["org"] = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["dir"] = d.astype(np.float32)
    rays["tmin"] = 0.0
    rays["tmax"] = tmax
    return rays


def camera_rays(cam, n, seed):
    """Camera rays exactly as the oracle's GenerateRay forms them."""
    from oracle import oracle as O

    rng = np.random.default_rng(seed)
    px = rng.integers(0, cam.width, n).astype(np.int32)
    py = rng.integers(0, cam.height, n).astype(np.int32)
    u = rng.random(n).astype(np.float32)
    v = rng.random(n).astype(np.float32)
    org, d = O.camera_rays(cam, px, py, u, v)
    rays = np.zeros(n, RAY_DTYPE)
    rays["org"], rays["dir"], rays["tmin"], rays["tmax"] = org, d, 0.0, np.inf
    return rays


def check_bvh(nodes, order, pos, max_leaf):
    """Structural invariants of a downloaded BVH (16 words per node: two child
    boxes, two references, two counts).  Every triangle sits in exactly one
    leaf, every leaf box contains its triangles, every child box contains the
    boxes of the node it references, every node is reached once, leaves hold at
    most max_leaf triangles.  Returns the deepest leaf's depth (root children
    are at depth 1)."""
    n = pos.shape[0]
    assert np.array_equal(np.sort(order), np.arange(n)), "leaf order is not a permutation"
    f = nodes.view(np.float32)
    tri = pos.reshape(n, 3, 3)
    vmin, vmax = tri.min(axis=1), tri.max(axis=1)
    covered = np.zeros(n, np.int64)
    reached = np.zeros(nodes.shape[0], np.int64)
    reached[0] = 1
    frontier = np.array([0])
    depth, leaf_depth = 0, 0
    while frontier.size:
        depth += 1
        assert depth < 128, "cycle in the node graph"
        nd, fd = nodes[frontier], f[frontier]
        nxt = []
        for s in (0, 1):
            ref, cnt = nd[:, 12 + s], nd[:, 14 + s]
            bmn, bmx = fd[:, 6 * s:6 * s + 3], fd[:, 6 * s + 3:6 * s + 6]
            valid = ref != 0xFFFFFFFF
            leaf = valid & (cnt > 0)
            inner = valid & (cnt == 0)
            if leaf.any():
                leaf_depth = depth
                st, ct = ref[leaf].astype(np.int64), cnt[leaf].astype(np.int64)
                assert ct.max() <= max_leaf
                assert (st + ct).max() <= n
                rep = np.repeat(np.arange(st.size), ct)
                posi = st[rep] + (np.arange(ct.sum()) - (np.cumsum(ct) - ct)[rep])
                np.add.at(covered, posi, 1)
                t = order[posi]
                assert (vmin[t] >= bmn[leaf][rep]).all() and (vmax[t] <= bmx[leaf][rep]).all(), "leaf box misses a triangle"
            if inner.any():
                ch = ref[inner].astype(np.int64)
                assert ch.max() < nodes.shape[0]
                np.add.at(reached, ch, 1)
                for cs in (0, 1):
                    cv = nodes[ch, 12 + cs] != 0xFFFFFFFF
                    cmn, cmx = f[ch, 6 * cs:6 * cs + 3][cv], f[ch, 6 * cs + 3:6 * cs + 6][cv]
                    assert (cmn >= bmn[inner][cv]).all() and (cmx <= bmx[inner][cv]).all(), "child box outside parent"
                nxt.append(ch)
        frontier = np.concatenate(nxt) if nxt else np.array([], np.int64)
    assert (covered == 1).all(), "triangle leaf counts range %d..%d" % (covered.min(), covered.max())
    assert (reached == 1).all(), "nodes reached %d..%d times" % (reached.min(), reached.max())
    return leaf_depth


def mixed_material_scene(seed=0, res=64, closed=True, area_light=True):
    """Cornell-style room exercising every material path of the device shader:
    textured Lambert walls (RGB and one-channel albedo textures with explicit
    uvs), a smooth glass box (GlassMaterial), a rough dielectric box
    (RoughDielectricMaterial, roughness -> alpha), a rough conductor box with
    FrConductor(eta, k) and a roughness texture, and a quad area light.
    closed=False leaves out the ceiling and the side walls (a scene open to an
    environment light); area_light=False leaves out the quad light."""
    import gyunity_amd as g

    rng = np.random.default_rng(seed)
    tex_rgb = rng.uniform(0.1, 0.9, (16, 16, 3)).astype(np.float32)
    tex_one = rng.uniform(0.2, 0.8, (8, 8, 1)).astype(np.float32)
    tex_rough = rng.uniform(0.05, 0.8, (8, 8, 1)).astype(np.float32)
    texs = [tex_rgb, tex_one, tex_rough]
    tex_info, off = [], 0
    for t in texs:
        tex_info.append((t.shape[1], t.shape[0], t.shape[2], off))
        off += t.size
    tex_data = np.concatenate([t.reshape(-1) for t in texs])
    mats = [
        material(0, R=(0.73, 0.73, 0.73)),                          # 0 white
        material(0, R=(1, 1, 1), tex_albedo=0),                     # 1 RGB texture
        material(0, R=(1, 1, 1), tex_albedo=1),                     # 2 one-channel texture
        material(0, R=(0.14, 0.45, 0.091)),                         # 3 green
        material(2, R=(1, 1, 1), T=(1, 1, 1), eta1=1.0, eta2=1.5),  # 4 glass
        material(3, R=(1, 1, 1), T=(1, 1, 1), alpha=(0.3, 0.3), alpha_direct=False, eta1=1.0, eta2=1.5),  # 5 rough dielectric
        material(1, R=(1, 1, 1), alpha=(0.3, 0.3), alpha_direct=False, eta=(0.143, 0.374, 1.442),
                 k=(3.983, 2.385, 1.603), tex_rough=2),             # 6 gold-like conductor, roughness texture
        material(0, R=(0, 0, 0)),                                   # 7 light (black)
    ]
    quads = []  # (a, b, c, d, normal, material, emissive)

    def quad(a, b, c, d, n, m, emit=False):
        quads.append((a, b, c, d, n, m, emit))

    quad((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), 0)        # floor
    if closed:
        quad((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0), (0, -1, 0), 0)   # ceiling
    quad((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 1), 2)        # back, one-channel texture
    if closed:
        quad((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 0, 0), 1)    # left, RGB texture
        quad((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (-1, 0, 0), 3)   # right, green
    if area_light:
        quad((0.375, 0.99, 0.375), (0.625, 0.99, 0.375), (0.625, 0.99, 0.625), (0.375, 0.99, 0.625), (0, -1, 0), 7, True)

    def box(lo, hi, m):
        x0, y0, z0 = lo
        x1, y1, z1 = hi
        quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1), (0, -1, 0), m)
        quad((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (0, 1, 0), m)
        quad((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0), (-1, 0, 0), m)
        quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1), (1, 0, 0), m)
        quad((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0), (0, 0, -1), m)
        quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1), (0, 0, 1), m)

    # The boxes float 1 mm above the floor: a face coplanar with the floor
    # would make every ray leaving a dielectric box downwards an exact-t tie,
    # which the reference resolves in its BVH's visit order.
    box((0.12, 0.001, 0.55), (0.36, 0.3, 0.8), 4)   # glass
    box((0.42, 0.001, 0.2), (0.62, 0.45, 0.42), 5)  # rough dielectric
    box((0.66, 0.001, 0.55), (0.88, 0.25, 0.78), 6)  # rough conductor
    pos, fn, uv, mat, light, ltri = [], [], [], [], [], []
    for a, b, c, d, n, m, emit in quads:
        for tri, tuv in (((a, b, c), (0, 0, 1, 0, 1, 1)), ((a, c, d), (0, 0, 1, 1, 0, 1))):
            if emit:
                light.append(len(ltri))
                ltri.append(len(pos))
            else:
                light.append(-1)
            pos.append(np.array(tri, np.float32).reshape(9))
            fn.append(n)
            uv.append(tuv)
            mat.append(m)
    pos = np.array(pos, np.float32)
    fn = np.array(fn, np.float32)
    nrm = np.repeat(fn, 3, axis=0).reshape(-1, 9)
    sa = SceneArrays(pos, nrm, fn, np.array(mat, np.uint32), np.array(light, np.int32), np.stack(mats),
                     np.array(ltri, np.int32).reshape(-1), np.tile(np.array([17, 12, 4], np.float32), (len(ltri), 1)),
                     uv=np.array(uv, np.float32), tex_info=np.array(tex_info, np.int32), tex_data=tex_data)
    cam = g.camera((0.5, 0.5, 2.35), (0.5, 0.5, 0.0), (0, 1, 0), 40.0, res, res)
    prm = g.render_params(spp=16, max_depth=6, rr_depth=3)
    return sa, cam, prm


def env_image(w=64, h=32, seed=0):
    """A latitude-longitude sky, top row first: a horizon gradient, a bright
    small sun and seeded texel noise, all positive."""
    rng = np.random.default_rng(seed)
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    sky = np.stack([0.4 + 0.6 * (1 - v), 0.5 + 0.4 * (1 - v), 0.8 + 0.2 * (1 - v)], -1)[:, None, :] * np.ones((1, w, 1))
    du = np.minimum(np.abs(u[None, :] - 0.3), 1 - np.abs(u[None, :] - 0.3))
    sun = np.exp(-((du / 0.04) ** 2 + ((v[:, None] - 0.25) / 0.06) ** 2)) * 40.0
    img = sky * rng.uniform(0.8, 1.2, (h, w, 3)) + sun[..., None] * np.array([1.0, 0.9, 0.7])
    return img.astype(np.float32)


# ---------------------------------------------------------------- film footprint by brute force
_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sobol_fixture_matrices():
    """The reference's Sobol generator matrices (sampler/Sobol.cpp:38-10689),
    first 256 dimensions x 52 columns, from the committed fixture."""
    return np.fromfile(os.path.join(_GOLDEN, "sobol_matrices.u32"), dtype="<u4").reshape(256, 52)


def sobol_values_u32(index, cols):
    """XOR of the matrix columns selected by the set bits of index: the 32-bit
    fixed-point value of sobol::sample (sampler/Sobol.h:46-61) before the
    conversion to float."""
    r = np.zeros(index.shape, np.uint32)
    nbytes = (max(1, int(index.max()).bit_length()) + 7) // 8 if index.size else 1
    for b in range(nbytes):
        # XOR of the columns 8b..8b+7 for every value of index byte b
        tab = np.zeros(256, np.uint32)
        for v in range(1, 256):
            low = v & -v
            tab[v] = tab[v ^ low] ^ cols[8 * b + low.bit_length() - 1]
        r ^= tab[((index >> np.uint64(8 * b)) & np.uint64(255)).astype(np.intp)]
    return r


def _sample_positions(W, H, spp, chunk=1 << 22):
    """Every camera sample of TiledIntegrator::Render with the Sobol sampler
    (scramble 0), in chunks: (index, px, py, u, v, u wrapped, v wrapped).

    - Frame f of every pixel of the 2^m grid (res = 2^m >= max(W, H)) uses
      the indices [f 4^m, (f + 1) 4^m), so frames 0..spp-1 are exactly the
      indices [0, spp 4^m), and the frame of index i is i >> 2m.
    - An index belongs to the pixel given by the top m bits of the exact 32-bit
      values of dimensions 0 and 1; pixels outside W x H are not rendered.
    - u = res * s - x in FP32 with s the FP32 sample, which RandomStateSequence
      maps from >= 1 to 0 (sampler/Sampler.h:60-62; SobolEnum::SampleX,
      sampler/SobolEnum.h:30-32)."""
    M = sobol_fixture_matrices()
    res = 1
    while res < max(W, H):
        res *= 2
    m = res.bit_length() - 1
    n = spp * res * res
    sh = np.uint32(32 - m)
    for a in range(0, n, chunk):
        idx = np.arange(a, min(n, a + chunk), dtype=np.uint64)
        r0 = sobol_values_u32(idx, M[0])
        r1 = sobol_values_u32(idx, M[1])
        px = (r0 >> sh).astype(np.int64) if m > 0 else np.zeros(idx.shape, np.int64)
        py = (r1 >> sh).astype(np.int64) if m > 0 else np.zeros(idx.shape, np.int64)
        keep = (px < W) & (py < H)
        idx, px, py, r0, r1 = idx[keep], px[keep], py[keep], r0[keep], r1[keep]
        s0 = r0.astype(np.float32) * np.float32(2.0 ** -32)
        s1 = r1.astype(np.float32) * np.float32(2.0 ** -32)
        w0, w1 = s0 >= 1, s1 >= 1
        s0[w0] = 0
        s1[w1] = 0
        u = np.float32(res) * s0 - px.astype(np.float32)
        v = np.float32(res) * s1 - py.astype(np.float32)
        yield idx >> np.uint64(2 * m), px, py, u, v, w0, w1


def film_weight_scan(W, H, spp):
    """Per-pixel film weights of TiledIntegrator::Render with the Sobol sampler
    and the radius-0.5 box filter, predicted by brute force over every sample
    index (_sample_positions) rather than by Film::AddSample's arithmetic: a
    box filter of radius 0.5 about the pixel centres counts a sample once in
    every pixel whose closed unit square [j, j+1] x [i, i+1] holds its raster
    position x + u, y + v (FP32), inside the image.

    Returns (weights (H, W) int64, counts) where counts tallies the samples
    with u == 0, u == 1 and u wrapped from 1 to 0."""
    wgt = np.zeros(H * W, np.int64)
    counts = {"u0": 0, "u1": 0, "wrap": 0}
    for _, px, py, u, v, w0, _ in _sample_positions(W, H, spp):
        fx = (px.astype(np.float32) + u).astype(np.float64)
        fy = (py.astype(np.float32) + v).astype(np.float64)
        counts["u0"] += int(np.count_nonzero(u == 0))
        counts["u1"] += int(np.count_nonzero(u == 1))
        counts["wrap"] += int(np.count_nonzero(w0))
        # closed squares holding the point: floor(p), and p - 1 when p is an integer
        xs = [np.floor(fx), np.where(fx == np.floor(fx), fx - 1, np.nan)]
        ys = [np.floor(fy), np.where(fy == np.floor(fy), fy - 1, np.nan)]
        for X in xs:
            for Y in ys:
                ok = (X >= 0) & (X < W) & (Y >= 0) & (Y < H)
                wgt += np.bincount((Y[ok] * W + X[ok]).astype(np.int64), minlength=H * W)
    return wgt.reshape(H, W), counts


def film_edge_samples(W, H, spp):
    """(x, y, sample) of every camera sample that lies on a pixel edge (u or v
    equal to 0 or 1) or whose u or v wrapped from 1 to 0."""
    out = []
    for frame, px, py, u, v, w0, w1 in _sample_positions(W, H, spp):
        e = (u == 0) | (u == 1) | (v == 0) | (v == 1) | w0 | w1
        out.append(np.stack([px[e], py[e], frame[e].astype(np.int64)], 1))
    xys = np.concatenate(out).astype(np.int32)
    return xys[:, 0], xys[:, 1], xys[:, 2]


def sphere_scene(res=64, spp=16):
    """The Cornell room (36 triangles, two area-light triangles) with three
    analytic spheres after the triangles: white, coloured and a small emissive
    one whose AreaLight follows the room's two."""
    import gyunity_amd as g

    sa, cam, prm = g.generate_scene(g.GY_SCENE_CORNELL)
    cam.width = cam.height = res
    prm.spp = spp
    sp = np.zeros(3, g.SPHERE_DTYPE)
    sp["center"] = [(0.3, 0.18, 0.55), (0.68, 0.25, 0.4), (0.5, 0.72, 0.45)]
    sp["radius"] = [0.15, 0.2, 0.07]
    sp["material"] = [0, 3, 0]
    sp["emissive"] = [0, 0, 1]
    sp["le"][2] = (6.0, 6.0, 6.0)
    return sa, cam, prm, sp


def reference_sphere_cornell(width=64, height=48, spp=16):
    """The reference's sphere Cornell box (Gyunity_exe/CornellBox.h:22-100,
    camera of Gyunity_exe/main.cpp:236-243): five inside-out walls of radius
    1e5, a diffuse ball, three glass balls and a light ball (Le = 30), all
    analytic spheres.  gy_scene_create needs a triangle, so the scene holds one
    degenerate triangle (three equal vertices, which no ray hits) far below."""
    import gyunity_amd as g

    tri = np.tile(np.array([50.0, -1000.0, 50.0], np.float32), 3)[None, :]
    n = np.tile(np.array([0.0, 1.0, 0.0], np.float32), 3)[None, :]
    red, blue, white, full = (0.75, 0.25, 0.25), (0.25, 0.25, 0.75), (0.75, 0.75, 0.75), (1.0, 1.0, 1.0)
    mats = np.stack([material(0, R=red), material(0, R=blue), material(0, R=white), material(0, R=full),
                     material(2, R=full, T=full), material(0, R=(0.0, 0.0, 0.0))])
    sa = SceneArrays(tri, n, np.array([[0.0, 1.0, 0.0]], np.float32), np.array([2], np.uint32),
                     np.array([-1], np.int32), mats, np.zeros(0, np.int32), np.zeros((0, 3), np.float32))
    rows = [  # centre, radius, material, inside, Le
        ((1e5 + 1, 40.8, 81.6), 1e5, 0, 1, None),
        ((-1e5 + 99, 40.8, 81.6), 1e5, 1, 1, None),
        ((50, 40.8, 1e5), 1e5, 2, 1, None),
        ((50, 1e5, 81.6), 1e5, 2, 1, None),
        ((50, -1e5 + 81.6, 81.6), 1e5, 2, 1, None),
        ((23, 0.0, 98), 9.5, 3, 0, None),
        ((73, 26.5, 78), 16.5, 4, 0, None),
        ((27, 16.5, 47), 16.5, 4, 0, None),
        ((53, 9.5, 88), 9.5, 4, 0, None),
        ((50, 81.6 - 16.5, 81.6), 8.0, 5, 0, (30.0, 30.0, 30.0)),
    ]
    sp = np.zeros(len(rows), g.SPHERE_DTYPE)
    for i, (c, r, m, inside, le) in enumerate(rows):
        sp[i]["center"], sp[i]["radius"], sp[i]["material"], sp[i]["inside"] = c, r, m, inside
        if le is not None:
            sp[i]["emissive"], sp[i]["le"] = 1, le
    cz = np.array([0.0, -0.042612, -1.0])
    pos = np.array([50.0, 52.0, 295.6])
    cx = np.array([1.0, 0.0, 0.0])
    cy = np.cross(cx, cz)
    cy /= np.linalg.norm(cy)
    cam = g.camera(tuple(pos), tuple(pos + cz), tuple(cy), 28.7993, width, height, float(np.linalg.norm(cz)))
    prm = g.render_params(spp=spp, max_depth=5, rr_depth=3)
    return sa, cam, prm, sp


def _medium_arrays(n_tri_room, n_cube, medium_id=0):
    from oracle.oracle import TRI_MEDIUM_DTYPE

    tri = np.zeros(n_tri_room + n_cube, TRI_MEDIUM_DTYPE)
    tri["inside"], tri["outside"] = -1, -1
    tri["inside"][n_tri_room:] = medium_id
    tri["no_material"][n_tri_room:] = 1
    return tri


def furnace_medium_scene(sigma_s=(0.5, 0.5, 0.5), sigma_a=(0.0, 0.0, 0.0), g=0.0, res=16, spp=64, depth=12):
    """The furnace box (12 inward-facing black emitters, Le = 1) holding a
    material-less cube [0.25, 0.75]^3 (outward normals) filled with a
    homogeneous medium; the pinhole camera sits inside the box, outside the
    cube, looking through it.  Returns (arrays, camera, params, media, tri)."""
    import gyunity_amd as g_
    from oracle.oracle import MEDIUM_DTYPE

    room, room_n = box_triangles((0, 0, 0), (1, 1, 1), inward=True)
    cube, cube_n = box_triangles((0.25, 0.25, 0.25), (0.75, 0.75, 0.75), inward=False)
    pos = np.concatenate([room, cube])
    fn = np.concatenate([room_n, cube_n])
    nrm = np.repeat(fn, 3, axis=0).reshape(-1, 9)
    n_room, n = room.shape[0], pos.shape[0]
    mats = material(0, R=(0.0, 0.0, 0.0))[None, :]
    light = np.full(n, -1, np.int32)
    light[:n_room] = np.arange(n_room)
    sa = SceneArrays(pos, nrm, fn, np.zeros(n, np.uint32), light, mats, np.arange(n_room, dtype=np.int32),
                     np.ones((n_room, 3), np.float32))
    media = np.zeros(1, MEDIUM_DTYPE)
    media["sigma_a"], media["sigma_s"], media["g"] = sigma_a, sigma_s, g
    tri = _medium_arrays(n_room, n - n_room)
    cam = g_.camera((0.5, 0.5, 0.97), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 70.0, res, res)
    prm = g_.render_params(spp=spp, max_depth=depth, rr_depth=depth)
    return sa, cam, prm, media, tri


def medium_cornell(res=64, spp=16, depth=8):
    """The Cornell room (36 triangles) with a material-less slab
    [0.15, 0.85] x [0.62, 0.92] x [0.1, 0.9] (outward normals) of coloured,
    forward-scattering medium between the boxes and the light."""
    import gyunity_amd as g_
    from oracle.oracle import MEDIUM_DTYPE

    sa0, cam, prm = g_.generate_scene(g_.GY_SCENE_CORNELL)
    cube, cube_n = box_triangles((0.15, 0.62, 0.1), (0.85, 0.92, 0.9), inward=False)
    n0, nc = sa0.pos.shape[0], cube.shape[0]
    pos = np.concatenate([sa0.pos, cube])
    nrm = np.concatenate([sa0.nrm, np.repeat(cube_n, 3, axis=0).reshape(-1, 9)])
    fn = np.concatenate([sa0.face_n, cube_n])
    mat = np.concatenate([sa0.mat, np.zeros(nc, np.uint32)])
    light = np.concatenate([sa0.light, np.full(nc, -1, np.int32)])
    sa = SceneArrays(pos, nrm, fn, mat, light, sa0.mats, sa0.light_tri, sa0.light_le)
    media = np.zeros(1, MEDIUM_DTYPE)
    media["sigma_a"], media["sigma_s"], media["g"] = (0.3, 0.6, 0.9), (2.0, 1.5, 1.0), 0.5
    tri = _medium_arrays(n0, nc)
    cam.width = cam.height = res
    prm.spp, prm.max_depth, prm.rr_depth = spp, depth, depth
    return sa, cam, prm, media, tri

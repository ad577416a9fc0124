"""Benchmark: path tracing of the 1.31M-triangle displaced-icosphere GGX scene
at 1024x1024, 64 spp, max depth 8 (NEE + MIS) on one MI355X.

One step = one full render (67,108,864 camera paths) of that image with the
scene and BVH already resident in HBM.  Prints one JSON line on stdout.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode wavefront|megakernel]
                  [--no-cpu-baseline] [--level 8] [--workload icosphere|env|batch64]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(name, st, paths):
    """Algorithmic bytes one render moves through a kernel category: 64 per BVH
    node record and 48 per triangle record the traversal visits (counted by an
    instrumented copy of the integrator over the same render), plus the
    per-ray / per-path records the kernel reads and writes."""
    nv, tv = list(st.node_visits_kind), list(st.tri_visits_kind)
    primary = st.rays - st.shadow_rays - st.mis_rays
    if name == "pt_megakernel":
        # all traversals, material records, the radiance record of each sample
        return 64.0 * st.node_visits + 48.0 * st.tri_visits + 16.0 * st.shade_points + 16.0 * paths
    if name == "wf_trace_closest":
        # + ray (32 B) read, hit (16 B) written, queue entry read and pushed
        return 64.0 * nv[0] + 48.0 * tv[0] + 56.0 * primary
    if name == "wf_trace_shadow_mis":
        # shadow rays, then MIS rays, in one launch
        return 64.0 * (nv[1] + nv[2]) + 48.0 * (tv[1] + tv[2]) + 36.0 * (st.shadow_rays + st.mis_rays)
    if name == "wf_shade":
        # shading record 128 B, material 128 B, path state ~200 B per shading point
        return 456.0 * st.shade_points
    return float("nan")


def host_cpu():
    """The host's CPU model, its logical CPU count and the CPUs this process may
    run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return model, os.cpu_count() or 1, allowed


# Worker threads for the CPU baseline: the host CPU share a one-GPU job is
# given (16 cores per GPU), or fewer where the process may use fewer.
CPU_SHARE = 16


def cpu_baseline(sa, cam, prm, rows, env=None, rows_1core=8):
    """The oracle's FP32 restatement of the reference CPU integrator
    (TiledIntegrator + PathTracing, reference BVH) on the host cores, over a
    band of `rows` image rows of the same workload; and the same on one core
    over `rows_1core` rows."""
    from oracle import oracle as O

    model, logical, allowed = host_cpu()
    threads = max(1, min(CPU_SHARE, allowed))
    os_ = O.OracleScene(sa)
    if env is not None:
        os_.set_environment(env)
    row0 = cam.height // 2 - rows // 2
    p = O.params(prm.spp, prm.max_depth, prm.rr_depth, precision=0, nthreads=threads, use_bvh=1)
    _, _, paths, secs = os_.render(cam, p, row0, row0 + rows)
    r1 = cam.height // 2 - rows_1core // 2
    p1 = O.params(prm.spp, prm.max_depth, prm.rr_depth, precision=0, nthreads=1, use_bvh=1)
    _, _, paths1, secs1 = os_.render(cam, p1, r1, r1 + rows_1core)
    os_.close()
    return {
        "value": round(paths / secs / 1e6, 4),
        "unit": "Mpaths/s",
        "cores": threads,
        "kind": "port",
        "sample": "rows %d-%d of the %dx%d image (%d paths, %.1f s), same scene, spp and depth"
        % (row0, row0 + rows - 1, cam.width, cam.height, paths, secs),
        "one_core": {"value": round(paths1 / secs1 / 1e6, 4), "unit": "Mpaths/s",
                     "sample": "rows %d-%d (%d paths, %.1f s)" % (r1, r1 + rows_1core - 1, paths1, secs1)},
        "cpu_model": model,
        "host_logical_cpus": logical,
        "cpus_allowed": allowed,
        "threads_note": "%d threads: the host CPU share of a one-GPU job (%d cores per GPU)" % (threads, CPU_SHARE),
    }


def open_room_sky(g, sa, w=1024, h=512):
    """The icosphere workload lit by an environment light: the room keeps its
    floor and back wall (triangles 0-1 and 4-5; the ceiling, side walls and
    ceiling light go), and a procedural latitude-longitude sky with a sun
    replaces the light."""
    keep = np.r_[0:2, 4:6, 12:sa.n_tri]
    out = g.SceneArrays(sa.pos[keep], sa.nrm[keep], sa.face_n[keep], sa.mat[keep], np.full(len(keep), -1, np.int32),
                        sa.mats, np.zeros(0, np.int32), np.zeros((0, 3), np.float32), uv=sa.uv[keep] if sa.uv is not None else None)
    v = (np.arange(h) + 0.5) / h
    u = (np.arange(w) + 0.5) / w
    sky = np.stack([0.3 + 0.5 * (1 - v), 0.4 + 0.4 * (1 - v), 0.7 + 0.3 * (1 - v)], -1)[:, None, :] * np.ones((1, w, 1))
    du = np.minimum(np.abs(u[None, :] - 0.3), 1 - np.abs(u[None, :] - 0.3))
    sun = np.exp(-((du / 0.02) ** 2 + ((v[:, None] - 0.2) / 0.03) ** 2)) * 200.0
    img = (sky + sun[..., None] * np.array([1.0, 0.9, 0.75])).astype(np.float32)
    return out, img


def lt_main(args):
    """Light tracing (gy_render_light_tracing) on the icosphere workload: the
    same scene, camera, spp and depth; one light path per (pixel, sample)
    index, every vertex connected to the camera and splatted."""
    import gyunity_amd as g

    sa, cam, prm = g.generate_scene(g.GY_SCENE_ICOSPHERE_GGX, 0, args.level)
    prm.mode = g.GY_RENDER_WAVEFRONT if args.mode == "wavefront" else g.GY_RENDER_MEGAKERNEL
    paths = cam.width * cam.height * prm.spp
    ctx = g.Context(0)
    log("device:", ctx.device_name())
    sc = g.Scene(ctx, sa)
    for i in range(args.warmup):
        sc.render_light_tracing(cam, prm)
        log("lt warmup %d done" % i)
    t0 = time.perf_counter()
    for i in range(args.steps):
        img = sc.render_light_tracing(cam, prm)
        log("lt step %d done" % i)
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    value = paths / (ms / 1e3) / 1e6
    out = {
        "metric": "Mpaths/s at fixed RMSE: batched wavefront vs plain megakernel; adjoint s/grad",
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: procedural displaced icosphere (seed 0) in the Cornell room",
        "config": {
            "workload": "light tracing, icosphere-L%d GGX conductor, %d triangles, %dx%d, %d light paths per pixel, depth %d"
            % (args.level, sa.n_tri, cam.width, cam.height, prm.spp, prm.max_depth),
            "integrator": "light tracing, %s form, float-atomic splats"
            % ("wavefront (one bounce of every path per launch)" if args.mode == "wavefront" else "per-path (one path per thread)"),
            "paths_per_step": paths,
            "finite_fraction": float(np.isfinite(img).mean()),
            "parallelism": "single GPU",
        },
    }
    sc.close()
    ctx.close()
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        log("cpu baseline ...")
        # the oracle's light tracer over rows on the host's cores, as the
        # reference's ParallelFor runs it; a smaller image, same scene and depth
        model, logical, allowed = host_cpu()
        cores = max(1, min(CPU_SHARE, allowed))
        res, spp = 512, 16
        small = g.camera(tuple(cam.pos), tuple(cam.look_at), tuple(cam.up), cam.fovy, res, res, cam.film_dist)
        os_ = O.OracleScene(sa)
        t0 = time.perf_counter()
        os_.render_lt(small, O.params(spp, prm.max_depth, prm.rr_depth, 0, cores, 1))
        secs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(res * res * spp / secs / 1e6, 4), "unit": "Mpaths/s", "cores": cores, "kind": "port",
                               "sample": "%dx%d image, %d light paths per pixel (%d paths, %.1f s), same scene and depth"
                               % (res, res, spp, res * res * spp, secs),
                               "cpu_model": model, "host_logical_cpus": logical, "cpus_allowed": allowed}
    print(json.dumps(out), flush=True)


def fog_box(g, sa, lo=(0.06, 0.01, 0.06), hi=(0.94, 0.9, 0.94), sigma_a=(0.05, 0.05, 0.08), sigma_s=(0.8, 0.7, 0.6),
            gg=0.3):
    """The scene with a material-less box of homogeneous fog (outward normals,
    inside = medium 0, outside = none) around the icosphere, below the light."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    quads = [((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1), (0, -1, 0)),
             ((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (0, 1, 0)),
             ((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0), (-1, 0, 0)),
             ((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1), (1, 0, 0)),
             ((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0), (0, 0, -1)),
             ((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1), (0, 0, 1))]
    tris, normals = [], []
    for a, b, c, d, n in quads:
        tris += [(a, b, c), (a, c, d)]
        normals += [n, n]
    box = np.array(tris, np.float32).reshape(-1, 9)
    bn = np.array(normals, np.float32)
    nb = box.shape[0]
    out = g.SceneArrays(np.concatenate([sa.pos, box]), np.concatenate([sa.nrm, np.repeat(bn, 3, axis=0).reshape(-1, 9)]),
                        np.concatenate([sa.face_n, bn]), np.concatenate([sa.mat, np.zeros(nb, np.uint32)]),
                        np.concatenate([sa.light, np.full(nb, -1, np.int32)]), sa.mats, sa.light_tri, sa.light_le,
                        uv=np.concatenate([sa.uv, np.zeros((nb, 6), np.float32)]) if sa.uv is not None else None)
    media = np.zeros(1, g.MEDIUM_DTYPE)
    media["sigma_a"], media["sigma_s"], media["g"] = sigma_a, sigma_s, gg
    tri = np.zeros(out.n_tri, g.TRI_MEDIUM_DTYPE)
    tri["inside"], tri["outside"] = -1, -1
    tri["inside"][sa.n_tri:] = 0
    tri["no_material"][sa.n_tri:] = 1
    return out, media, tri


def vpt_main(args):
    """VolPathTracing (gy_render_vpt) on the icosphere workload inside a box of
    homogeneous fog: the same camera, spp and depth; equi-angular sampling (the
    reference's default) and free-flight sampling."""
    import gyunity_amd as g

    sa0, cam, prm = g.generate_scene(g.GY_SCENE_ICOSPHERE_GGX, 0, args.level)
    sa, media, tri = fog_box(g, sa0)
    paths = cam.width * cam.height * prm.spp
    ctx = g.Context(0)
    log("device:", ctx.device_name())
    sc = g.Scene(ctx, sa)
    sc.set_media(media, tri)
    res = {}
    forms = [("wavefront", g.GY_RENDER_WAVEFRONT)] + ([] if args.single else [("per_path", g.GY_RENDER_MEGAKERNEL)])
    for form, mode in forms:
        prm.mode = mode
        for eqa in (True, False):
            name = ("equiangular" if eqa else "free_flight") + ("" if form == "wavefront" else "_per_path")
            for i in range(args.warmup):
                sc.render_vpt(cam, prm, equiangular=eqa)
                log("vpt %s warmup %d done" % (name, i))
            t0 = time.perf_counter()
            for i in range(args.steps):
                c, w = sc.render_vpt(cam, prm, equiangular=eqa)
                log("vpt %s step %d done" % (name, i))
            ms = (time.perf_counter() - t0) * 1e3 / args.steps
            st = ctx.stats()
            log("vpt %s per-kernel ms/step: %s" % (name, ", ".join(
                "%s %.1f" % (g.lib().gy_kernel_name(k).decode(), st.kernel_ms[k]) for k in range(8) if st.kernel_ms[k] > 0)))
            res[name] = {"ms_per_step": round(ms, 3), "mpaths_per_s": round(paths / (ms / 1e3) / 1e6, 3),
                         "finite_fraction": float(np.isfinite(c).mean()), "mean": float((c / np.maximum(w, 1)[..., None]).mean())}
    value = res["equiangular"]["mpaths_per_s"]
    out = {
        "metric": "Mpaths/s at fixed RMSE: batched wavefront vs plain megakernel; adjoint s/grad",
        "value": value,
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["equiangular"]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: procedural displaced icosphere (seed 0) in the Cornell room, in a box of homogeneous fog",
        "config": {
            "workload": "VolPathTracing, icosphere-L%d GGX conductor, %d triangles (12 material-less fog-box walls), %dx%d, %d spp, "
                        "depth %d, fog sigma_a %s sigma_s %s g %.2f"
                        % (args.level, sa.n_tri, cam.width, cam.height, prm.spp, prm.max_depth,
                           [round(float(v), 3) for v in media["sigma_a"][0]], [round(float(v), 3) for v in media["sigma_s"][0]],
                           float(media["g"][0])),
            "integrator": "VolPathTracing, wavefront form (per-path form alongside unless --single; equal films)",
            "forms": res,
            "paths_per_step": paths,
            "parallelism": "single GPU",
        },
    }
    sc.close()
    ctx.close()
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        log("cpu baseline ...")
        model, logical, allowed = host_cpu()
        cores = max(1, min(CPU_SHARE, allowed))
        rows = 32
        r0 = cam.height // 2 - rows // 2
        os_ = O.OracleScene(sa)
        os_.set_media(media, tri)
        t0 = time.perf_counter()
        _, _, npaths, _ = os_.render(cam, O.params(prm.spp, prm.max_depth, prm.rr_depth, 0, cores, 1, integrator=2), r0, r0 + rows)
        secs = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(npaths / secs / 1e6, 4), "unit": "Mpaths/s", "cores": cores, "kind": "port",
                               "sample": "rows %d-%d, equi-angular (%d paths, %.1f s), same scene, spp and depth"
                               % (r0, r0 + rows - 1, npaths, secs),
                               "cpu_model": model, "host_logical_cpus": logical, "cpus_allowed": allowed}
    print(json.dumps(out), flush=True)


def sppm_main(args):
    """SPPM (gy_render_sppm) on the icosphere workload with the reference's own
    SPPM settings (Gyunity_exe/main.cpp:54-56, 247-249): 300,000 photons per
    iteration, alpha 0.75, depth 20, the Halton sampler; initial radius 0.01
    (1.0 in the reference's 100-unit Cornell box)."""
    import gyunity_amd as g

    sa, cam, prm = g.generate_scene(g.GY_SCENE_ICOSPHERE_GGX, 0, args.level)
    iters, photons = args.sppm_iterations, 300000
    sp = g.sppm_params(iters, photons, 20, 0.01, 0.75, batch_shrink=False, sampler=g.GY_SAMPLER_HALTON)
    ctx = g.Context(0)
    log("device:", ctx.device_name())
    sc = g.Scene(ctx, sa)
    for i in range(args.warmup):
        sc.render_sppm(cam, g.sppm_params(2, photons, 20, 0.01, 0.75, sampler=g.GY_SAMPLER_HALTON))
        log("sppm warmup %d done" % i)
    t0 = time.perf_counter()
    img, r2, N = sc.render_sppm(cam, sp)
    secs = time.perf_counter() - t0
    log("sppm: %d iterations in %.2f s" % (iters, secs))
    out = {
        "metric": "Mpaths/s at fixed RMSE: batched wavefront vs plain megakernel; adjoint s/grad",
        "value": round(iters * photons / secs / 1e6, 3),
        "unit": "Mphotons/s",
        "n_gpus": 1,
        "steps": iters,
        "warmup": args.warmup,
        "ms_per_step": round(secs * 1e3 / iters, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: procedural displaced icosphere (seed 0) in the Cornell room",
        "config": {
            "workload": "SPPM, icosphere-L%d GGX conductor, %d triangles, %dx%d eye paths and %d photons per iteration, depth 20, "
                        "alpha 0.75, initial radius 0.01, Halton" % (args.level, sa.n_tri, cam.width, cam.height, photons),
            "step": "one SPPM iteration: eye pass, hash grid, photon pass, per-hit-point updates in photon order",
            "iterations": iters,
            "eye_paths_per_s_millions": round(iters * cam.width * cam.height / secs / 1e6, 3),
            "mean_photons_per_pixel": float(N.mean()),
            "mean_radius": float(np.sqrt(r2).mean()),
            "finite_fraction": float(np.isfinite(img).mean()),
            "parallelism": "single GPU",
        },
    }
    sc.close()
    ctx.close()
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        log("cpu baseline ...")
        model, logical, allowed = host_cpu()
        cores = max(1, min(CPU_SHARE, allowed))
        res, it_c, ph_c = 256, 2, 300000
        small = g.camera(tuple(cam.pos), tuple(cam.look_at), tuple(cam.up), cam.fovy, res, res, cam.film_dist)
        os_ = O.OracleScene(sa)
        t0 = time.perf_counter()
        os_.render_sppm(small, O.sppm_params(it_c, ph_c, 20, 0.01, 0.75, 0, 1, 0, cores))
        secs_c = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(it_c * ph_c / secs_c / 1e6, 4), "unit": "Mphotons/s", "cores": cores, "kind": "port",
                               "sample": "%dx%d, %d iterations of %d photons (%.1f s): eye pass on %d threads, photons one at a "
                                         "time in index order" % (res, res, it_c, ph_c, secs_c, cores),
                               "cpu_model": model, "host_logical_cpus": logical, "cpus_allowed": allowed}
    print(json.dumps(out), flush=True)


def batch_main(args):
    """The 64-scene batch workload: every member traced in one lane space
    (gy_batch_render, recorded into a hipGraph once and replayed), against the
    same members rendered one call each with the wavefront and the megakernel."""
    import gyunity_amd as g

    ctx = g.Context(0)
    log("device:", ctx.device_name())
    t0 = time.perf_counter()
    members = [g.generate_scene(g.GY_SCENE_BATCH_MEMBER, s, 0) for s in range(args.batch_size)]
    log("generated %d members in %.1f s" % (len(members), time.perf_counter() - t0))
    scenes = [g.Scene(ctx, sa) for sa, _, _ in members]
    paths = sum(cam.width * cam.height * prm.spp for _, cam, prm in members)
    tris = sum(sa.n_tri for sa, _, _ in members)
    log("scenes resident: %d triangles, %d paths per batch render" % (tris, paths))
    batch = g.Batch(ctx, lanes_per_pass=args.batch_lanes, refill_threshold=args.refill, use_graph=not args.batch_no_graph,
                    ray_sort=RAY_SORT[args.ray_sort])
    for sc, (_, cam, prm) in zip(scenes, members):
        batch.add(sc, cam, prm)
    for i in range(args.warmup):
        batch.render()
        log("batch warmup %d done" % i)
    rays = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        batch.render()
        st = ctx.stats()
        rays = st.rays
        log("batch step %d: %.1f ms, %.1f Mrays" % (i, st.total_ms, st.rays / 1e6))
    ms_batch = (time.perf_counter() - t0) * 1e3 / args.steps
    per_scene = {}
    # per-member launch choice by the autotuner (its automatic pilot size)
    tuned = []
    t0 = time.perf_counter()
    for sc, (_, cam, prm) in zip(scenes, members):
        tuned.append(sc.autotune(cam, prm, 0)[0])
    log("autotuned %d members in %.1f s: %d megakernel, %d wavefront" % (
        len(tuned), time.perf_counter() - t0, sum(t.mode == g.GY_RENDER_MEGAKERNEL for t in tuned),
        sum(t.mode == g.GY_RENDER_WAVEFRONT for t in tuned)))
    for sc, (_, cam, _), prm in zip(scenes, members, tuned):
        sc.render(cam, prm, download=False)
    t0 = time.perf_counter()
    for sc, (_, cam, _), prm in zip(scenes, members, tuned):
        sc.render(cam, prm, download=False)
    per_scene["autotuned"] = (time.perf_counter() - t0) * 1e3
    log("one call per member, autotuned launch: %.1f ms" % per_scene["autotuned"])
    for mode, name in ((g.GY_RENDER_WAVEFRONT, "wavefront"), (g.GY_RENDER_MEGAKERNEL, "megakernel")):
        for sc, (_, cam, prm) in zip(scenes, members):  # warm-up: allocations, first launches
            prm.mode = mode
            sc.render(cam, prm, download=False)
        t0 = time.perf_counter()
        for sc, (_, cam, prm) in zip(scenes, members):
            prm.mode = mode
            sc.render(cam, prm, download=False)
        per_scene[name] = (time.perf_counter() - t0) * 1e3
        log("one call per member, %s: %.1f ms" % (name, per_scene[name]))
    batch.close()
    for sc in scenes:
        sc.close()
    ctx.close()
    out = {
        "metric": "Mpaths/s at fixed RMSE: batched wavefront vs plain megakernel; adjoint s/grad",
        "value": round(paths / (ms_batch / 1e3) / 1e6, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_batch, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: %d procedural batch members (seeds 0..%d)" % (args.batch_size, args.batch_size - 1),
        "config": {
            "workload": "batch of %d scenes: displaced icospheres (1e3..2e6 triangles, log-uniform) + 0-20 boxes, "
                        "GGX conductor / 25%% rough dielectric, 128^2..1024^2, 4..64 spp, depth 4..12" % args.batch_size,
            "integrator": "batched wavefront (one lane space, %s)" % ("eager launches" if args.batch_no_graph else "hipGraph replay"),
            "paths_per_step": paths,
            "triangles": tris,
            "mrays_per_s": round(rays / (ms_batch / 1e3) / 1e6, 2),
            "per_member_wavefront_mpaths_per_s": round(paths / (per_scene["wavefront"] / 1e3) / 1e6, 3),
            "per_member_megakernel_mpaths_per_s": round(paths / (per_scene["megakernel"] / 1e3) / 1e6, 3),
            "per_member_autotuned_mpaths_per_s": round(paths / (per_scene["autotuned"] / 1e3) / 1e6, 3),
            "same_estimator": "each member's batch film equals its own render bit for bit (tests/test_gpu_parity.py)",
            "parallelism": "single GPU",
        },
    }
    print(json.dumps(out), flush=True)


def _image(color, weight):
    return color.astype(np.float64) / np.maximum(weight, 1e-30)[..., None]


def adjoint_main(args):
    """Adjoint, small: the textured Cornell room (64^2 RGB albedo per wall, GGX
    roughness texture on a 20,480-triangle sphere), 256^2, 64 spp, L2 loss to a
    render of the ground-truth textures.  Seconds per gradient (render +
    path-replay adjoint) from a seeded uniform start, then path-replay against
    central finite differences on random texels (common random numbers)."""
    import gyunity_amd as g

    ctx = g.Context(0)
    log("device:", ctx.device_name())
    sa, cam, prm = g.generate_scene(g.GY_SCENE_ADJOINT, 0, 64)
    prm.mode = g.GY_RENDER_WAVEFRONT
    gt = sa.tex_data.copy()
    n_alb = int(sa.tex_info[5, 3])
    rng = np.random.default_rng(1)
    theta = np.concatenate([rng.uniform(0.2, 0.9, n_alb), rng.uniform(0.05, 0.8, gt.size - n_alb)]).astype(np.float32)
    sc = g.Scene(ctx, sa)
    target = _image(*sc.render(cam, prm))
    sc.set_textures(theta)

    def gradient(p):
        img = _image(*sc.render(cam, p))
        dI = (2.0 * (img - target_of[p.spp]) / img.size).astype(np.float32)
        return float(((img - target_of[p.spp]) ** 2).mean()), sc.adjoint(cam, p, dI)

    target_of = {prm.spp: target}
    for i in range(args.warmup):
        gradient(prm)
    t0 = time.perf_counter()
    for i in range(args.steps):
        loss, grad = gradient(prm)
    s_per_grad = (time.perf_counter() - t0) / args.steps
    log("adjoint: %.4f s per gradient (%d paths rendered + replayed), loss %.6f" % (s_per_grad, cam.width * cam.height * prm.spp, loss))
    fd_report = None
    if args.fd_texels > 0:
        import copy

        pf = copy.copy(prm)
        pf.spp = args.fd_spp
        sc.set_textures(gt)
        target_of[pf.spp] = _image(*sc.render(cam, pf))
        sc.set_textures(theta)
        _, gfd = gradient(pf)
        ks = np.sort(np.random.default_rng(2).choice(gt.size, args.fd_texels, replace=False))
        fd = np.zeros(len(ks))
        h = args.fd_h
        for j, k in enumerate(ks):
            vals = []
            for sgn in (1, -1):
                th = theta.copy()
                th[k] += sgn * h
                sc.set_textures(th)
                vals.append(float(((_image(*sc.render(cam, pf)) - target_of[pf.spp]) ** 2).mean()))
            fd[j] = (vals[0] - vals[1]) / (2 * h)
            log("  texel float %6d (%s): replay %+.5e  FD %+.5e" % (k, "albedo" if k < n_alb else "roughness", gfd[k], fd[j]))
        sc.set_textures(theta)
        gr = gfd[ks].astype(np.float64)

        def rel(mask):
            return float(np.linalg.norm(gr[mask] - fd[mask]) / max(np.linalg.norm(fd[mask]), 1e-30)) if mask.any() else None

        alb = ks < n_alb
        fd_report = {"texels": int(len(ks)), "spp": pf.spp, "h": h, "rel_l2_all": rel(np.ones_like(alb)),
                     "rel_l2_albedo": rel(alb), "rel_l2_roughness": rel(~alb),
                     "albedo_texels": int(alb.sum()), "roughness_texels": int((~alb).sum())}
        log("FD check:", fd_report)
    sc.close()
    ctx.close()
    out = {
        "metric": "Mpaths/s at fixed RMSE: batched wavefront vs plain megakernel; adjoint s/grad",
        "value": round(s_per_grad, 4),
        "unit": "s/grad",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(s_per_grad * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded textures (ground truth seed 0, start seed 1)",
        "config": {
            "workload": "adjoint small: textured Cornell room + 20,480-triangle GGX sphere, %d texture floats, 256^2, %d spp, depth %d"
                        % (gt.size, prm.spp, prm.max_depth),
            "gradient": "render + path-replay adjoint, albedo and roughness texels",
            "paths_per_gradient": cam.width * cam.height * prm.spp,
            "finite_difference_check": fd_report,
            "parallelism": "single GPU",
        },
    }
    print(json.dumps(out), flush=True)


def adjoint_scale_main(args):
    """Adjoint at scale: the 1.31M-triangle GGX scene with 1024^2 RGB albedo and
    roughness textures on the sphere (4.2M parameters), 32 seeded views of
    512^2 at 32 spp, depth 6; Adam (lr 1e-2) from a seeded uniform start
    towards renders of the ground-truth textures, through the LibTorch
    autograd.Function (gy_autograd).  Reports loss per step and s/grad."""
    import torch  # before the library: one HIP runtime for both

    import gyunity_amd as g
    from gyunity_amd.torch_render import render_views

    ctx = g.Context(0)
    log("device:", ctx.device_name())
    sa, cam0, prm = g.generate_scene(g.GY_SCENE_ICOSPHERE_TEXTURED, 0, args.tex_res)
    prm.mode = g.GY_RENDER_WAVEFRONT
    gt = sa.tex_data.copy()
    n_alb = int(sa.tex_info[1, 3])
    sc = g.Scene(ctx, sa)
    rng = np.random.default_rng(3)
    views = g.orbit_views(args.views, args.view_res)
    # every view is an entry of one batch: one lane space forward (recorded
    # graph), one replay launch backward
    batch = g.Batch(ctx)
    for c in views:
        batch.add(sc, c, prm)
    tex_gt = torch.from_numpy(gt).cuda()
    with torch.no_grad():
        targets = render_views(sc, batch, tex_gt).clone()
    init = np.concatenate([rng.uniform(0.2, 0.8, n_alb), rng.uniform(0.05, 0.8, gt.size - n_alb)]).astype(np.float32)
    tex = torch.from_numpy(init).cuda().requires_grad_(True)
    opt = torch.optim.Adam([tex], lr=1e-2)
    lo = torch.cat([torch.full((n_alb,), 0.01), torch.full((gt.size - n_alb,), 0.05)]).cuda()
    hi = torch.cat([torch.full((n_alb,), 0.99), torch.full((gt.size - n_alb,), 0.8)]).cuda()
    losses, times = [], []
    for step in range(args.adam_steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        opt.zero_grad()
        imgs = render_views(sc, batch, tex)
        loss = ((imgs - targets) ** 2).mean()
        loss.backward()
        total = float(loss)
        opt.step()
        with torch.no_grad():
            tex.copy_(torch.minimum(torch.maximum(tex, lo), hi))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        losses.append(total)
        err = float((tex.detach() - tex_gt).abs().mean())
        log("adam step %3d: loss %.6e, mean |texel - ground truth| %.4f, %.2f s" % (step, total, err, times[-1]))
    s_per_grad = float(np.mean(times[1:])) if len(times) > 1 else times[0]
    paths = len(views) * args.view_res * args.view_res * prm.spp
    batch.close()
    sc.close()
    ctx.close()
    out = {
        "metric": "Mpaths/s at fixed RMSE: batched wavefront vs plain megakernel; adjoint s/grad",
        "value": round(s_per_grad, 4),
        "unit": "s/grad",
        "n_gpus": 1,
        "steps": args.adam_steps,
        "warmup": 1,
        "ms_per_step": round(s_per_grad * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: smooth seeded ground-truth textures, seeded uniform start, seeded views",
        "config": {
            "workload": "adjoint at scale: 1.31M-triangle GGX icosphere, %d texture floats (%d^2 RGB albedo + roughness), "
                        "%d views %d^2, %d spp, depth %d" % (gt.size, args.tex_res, len(views), args.view_res, prm.spp, prm.max_depth),
            "gradient": "per step: the views rendered as one batch and replayed in one launch (gy_autograd render_views), "
                        "Adam lr 1e-2",
            "paths_per_gradient": paths,
            "loss_per_step": [float("%.6e" % v) for v in losses],
            "parallelism": "single GPU",
        },
    }
    print(json.dumps(out), flush=True)


def batch_leg(g, ctx, steps=2):
    """The 64-scene batch (seeds 0..63), compact: one graph-replayed batch
    render warmed up and `steps` timed, against one wavefront and one
    megakernel call per member (each warmed up once)."""
    t0 = time.perf_counter()
    members = [g.generate_scene(g.GY_SCENE_BATCH_MEMBER, s, 0) for s in range(64)]
    scenes = [g.Scene(ctx, sa) for sa, _, _ in members]
    paths = sum(cam.width * cam.height * prm.spp for _, cam, prm in members)
    log("batch64 leg: 64 members resident in %.1f s" % (time.perf_counter() - t0))
    batch = g.Batch(ctx)
    for sc, (_, cam, prm) in zip(scenes, members):
        batch.add(sc, cam, prm)
    batch.render()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.render()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    per = {}
    for mode, name in ((g.GY_RENDER_WAVEFRONT, "wavefront"), (g.GY_RENDER_MEGAKERNEL, "megakernel")):
        for sc, (_, cam, prm) in zip(scenes, members):
            prm.mode = mode
            sc.render(cam, prm, download=False)
        t0 = time.perf_counter()
        for sc, (_, cam, prm) in zip(scenes, members):
            sc.render(cam, prm, download=False)
        per[name] = (time.perf_counter() - t0) * 1e3
    log("batch64 leg: batch %.1f ms, per member wavefront %.1f ms, megakernel %.1f ms" % (ms, per["wavefront"], per["megakernel"]))
    batch.close()
    for sc in scenes:
        sc.close()
    return {"workload": "64-scene batch, seeds 0..63 (1.2e3..1.9e6 triangles, 128^2..1024^2, 4..64 spp, depth 4..12)",
            "paths_per_render": paths, "batch_mpaths_per_s": round(paths / ms / 1e3, 3), "batch_ms": round(ms, 2),
            "per_member_wavefront_mpaths_per_s": round(paths / per["wavefront"] / 1e3, 3),
            "per_member_megakernel_mpaths_per_s": round(paths / per["megakernel"] / 1e3, 3),
            "form": "one lane space, recorded hipGraph replayed", "steps": steps}


def adjoint_leg(g, ctx, steps_small=5, steps_scale=2):
    """Seconds per gradient: the small adjoint (textured Cornell room, 64^2
    textures, 256^2 at 64 spp: set texels, render, replay) and the adjoint at
    scale (1.31M triangles, 1024^2 albedo + roughness textures, 32 views of
    512^2 at 32 spp, depth 6: one batch render and one replay launch per
    gradient, images and dl/dI on the host)."""
    out = {}
    sa, cam, prm = g.generate_scene(g.GY_SCENE_ADJOINT, 0, 64)
    cam.width = cam.height = 256
    prm.spp, prm.rr_depth, prm.mode = 64, 100, g.GY_RENDER_WAVEFRONT
    gt = sa.tex_data.copy()
    theta = np.clip(gt + np.random.default_rng(7).uniform(-0.2, 0.2, gt.size), 0.05, 0.8).astype(np.float32)
    with g.Scene(ctx, sa) as sc:
        target = _image(*sc.render(cam, prm))
        times = []
        for i in range(steps_small + 1):
            t0 = time.perf_counter()
            sc.set_textures(theta)
            img = _image(*sc.render(cam, prm))
            sc.adjoint(cam, prm, (2.0 * (img - target) / img.size).astype(np.float32))
            times.append(time.perf_counter() - t0)
    out["small"] = {"s_per_grad": round(float(np.mean(times[1:])), 4), "paths_per_grad": 256 * 256 * 64,
                    "workload": "textured Cornell room, 64^2 textures, 256^2, 64 spp"}
    log("adjoint leg small: %.4f s/grad" % out["small"]["s_per_grad"])
    sa, _, prm = g.generate_scene(g.GY_SCENE_ICOSPHERE_TEXTURED, 0, 1024)
    prm.mode = g.GY_RENDER_WAVEFRONT
    gt = sa.tex_data.copy()
    n_alb = int(sa.tex_info[1, 3])
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(0.2, 0.8, n_alb), rng.uniform(0.05, 0.8, gt.size - n_alb)]).astype(np.float32)
    views = g.orbit_views(32, 512)
    with g.Scene(ctx, sa) as sc, g.Batch(ctx) as b:
        for c in views:
            b.add(sc, c, prm)
        b.render()
        targets = [_image(*b.film(e)) for e in range(len(views))]
        times = []
        for i in range(steps_scale + 1):
            t0 = time.perf_counter()
            sc.set_textures(x)
            b.render()
            imgs = [_image(*b.film(e)) for e in range(len(views))]
            n = sum(im.size for im in imgs)
            grad = b.adjoint([(2.0 * (im - t) / n).astype(np.float32) for im, t in zip(imgs, targets)], sc.texture_floats())
            times.append(time.perf_counter() - t0)
        assert np.isfinite(grad).all()
    out["scale"] = {"s_per_grad": round(float(np.mean(times[1:])), 4), "paths_per_grad": 32 * 512 * 512 * 32,
                    "workload": "1.31M-triangle GGX icosphere, 4.2M texture floats, 32 views 512^2, 32 spp, depth 6, "
                                "one batch render + one replay launch"}
    log("adjoint leg scale: %.3f s/grad" % out["scale"]["s_per_grad"])
    return out


RAY_SORT = {"none": 0, "octant": 1, "octant-local": 2}  # GY_RAY_SORT_*


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="wavefront", choices=["wavefront", "megakernel"])
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--cpu-rows", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="default workload: skip the batch64 and adjoint legs")
    ap.add_argument("--single", action="store_true", help="time only --mode, not the other integrator")
    ap.add_argument("--refill", type=int, default=0, help="wavefront refill threshold (0 = library default)")
    ap.add_argument("--ray-sort", default="none", choices=["none", "octant", "octant-local"],
                    help="wavefront closest-hit queue order from the second bounce on")
    ap.add_argument("--spp-pass", type=int, default=0, help="samples per pass (0 = library default)")
    ap.add_argument("--leaf", type=int, default=0, help="BVH leaf size (0 = library default, 4)")
    ap.add_argument("--workload", default="icosphere",
                    choices=["icosphere", "env", "lt", "vpt", "sppm", "batch64", "adjoint", "adjoint-scale"],
                    help="icosphere: the 1.31M-triangle GGX render; batch64: the 64-scene batch; adjoint: "
                         "gradient of the textured Cornell room; adjoint-scale: Adam on 4.2M texels over 32 views")
    ap.add_argument("--sppm-iterations", type=int, default=16)
    ap.add_argument("--fd-texels", type=int, default=0, help="adjoint: texels checked against finite differences")
    ap.add_argument("--fd-spp", type=int, default=4096)
    ap.add_argument("--fd-h", type=float, default=1e-3)
    ap.add_argument("--adam-steps", type=int, default=20)
    ap.add_argument("--views", type=int, default=32)
    ap.add_argument("--view-res", type=int, default=512)
    ap.add_argument("--tex-res", type=int, default=1024)
    ap.add_argument("--batch-size", type=int, default=64)
    ap.add_argument("--batch-lanes", type=int, default=0, help="batch paths per pass (0 = library default)")
    ap.add_argument("--batch-no-graph", action="store_true", help="launch the batch sequence eagerly instead of replaying its graph")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="after the run, copy MIB MiB with the library's plain copy kernel (k_copy_probe), "
                         "the known-size reference for HBM counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one GPU: this build runs on a single MI355X; --gpus %d is not supported" % args.gpus)
        sys.exit(2)
    if args.workload == "batch64":
        batch_main(args)
        return
    if args.workload == "adjoint":
        adjoint_main(args)
        return
    if args.workload == "adjoint-scale":
        adjoint_scale_main(args)
        return
    if args.workload == "lt":
        lt_main(args)
        return
    if args.workload == "vpt":
        vpt_main(args)
        return
    if args.workload == "sppm":
        sppm_main(args)
        return

    import gyunity_amd as g

    sa, cam, prm = g.generate_scene(g.GY_SCENE_ICOSPHERE_GGX, 0, args.level)
    env = None
    if args.workload == "env":
        sa, env = open_room_sky(g, sa)
    prm.mode = g.GY_RENDER_WAVEFRONT if args.mode == "wavefront" else g.GY_RENDER_MEGAKERNEL
    prm.refill_threshold = args.refill
    prm.ray_sort = RAY_SORT[args.ray_sort]
    prm.samples_per_pass = args.spp_pass
    paths = cam.width * cam.height * prm.spp
    ctx = g.Context(0)
    log("device:", ctx.device_name())
    t0 = time.perf_counter()
    sc = g.Scene(ctx, sa, max_leaf_size=args.leaf)
    if env is not None:
        sc.set_environment(env)
    sst = sc.stats()
    log("scene: %d triangles, BVH %d nodes depth %d, build %.1f ms (scene create %.1f ms)"
        % (sa.n_tri, sst.n_nodes, sst.max_depth, sst.build_ms, (time.perf_counter() - t0) * 1e3))
    # FP64 16384-spp reference of five rows of this image (tests/golden): the
    # "at fixed RMSE" of the metric, measured on each integrator's last render
    ref_rows, ref_img = None, None
    gpath = os.path.join(ROOT, "tests", "golden", "level_c.json")
    if args.level == 8 and env is None and os.path.exists(gpath):
        with open(gpath) as f:
            ref_rows = json.load(f)["ico_rows"]
        ref_img = np.load(os.path.join(ROOT, "tests", "golden", "icosphere_rows_fp64_16384spp.npy"))
    rmse = {}

    def run(mode, steps, warmup):
        prm.mode = g.GY_RENDER_WAVEFRONT if mode == "wavefront" else g.GY_RENDER_MEGAKERNEL
        for i in range(warmup):
            sc.render(cam, prm, download=False)
            log("%s warmup %d done" % (mode, i))
        kms, kn = np.zeros(8), np.zeros(8)
        t0 = time.perf_counter()
        for i in range(steps):
            sc.render(cam, prm, download=False)
            st = ctx.stats()
            kms += np.array(st.kernel_ms[:])
            kn += np.array(st.kernel_launches[:])
            log("%s step %d: %.1f ms, %.1f Mrays" % (mode, i, st.total_ms, st.rays / 1e6))
        ms = (time.perf_counter() - t0) * 1e3 / steps
        if ref_img is not None:
            rmse[mode] = sc.film_rmse(ref_img, ref_rows)
            log("%s rmse against the FP64 16384-spp reference rows: %.6f" % (mode, rmse[mode]))
        log("%s per-kernel ms/step: %s" % (mode, ", ".join(
            "%s %.1f" % (g.lib().gy_kernel_name(k).decode(), kms[k] / steps) for k in range(8) if kn[k] > 0)))
        return ms, kms, kn, st

    other = "megakernel" if args.mode == "wavefront" else "wavefront"
    other_ms = None
    if not args.single:
        other_ms = run(other, max(1, args.steps - 1), 1)[0]
    ms_per_step, kern_ms, kern_n, st = run(args.mode, args.steps, args.warmup)
    prm.mode = g.GY_RENDER_WAVEFRONT if args.mode == "wavefront" else g.GY_RENDER_MEGAKERNEL
    rays = st.rays
    value = paths / (ms_per_step / 1e3) / 1e6

    # Dominant kernel and its algorithmic bytes per launch (counted by an
    # instrumented copy of the integrator over the same render).
    cnt = sc.count_traffic(cam, prm)
    dom = int(np.argmax(kern_ms))
    dom_name = g.lib().gy_kernel_name(dom).decode()
    launches_per_step = kern_n[dom] / args.steps
    avg_launch_ms = kern_ms[dom] / kern_n[dom]
    bytes_total = kernel_bytes(dom_name, cnt, paths)
    share = kern_ms[dom] / max(kern_ms.sum(), 1e-9)
    bytes_per_launch = bytes_total / launches_per_step
    achieved = bytes_per_launch / (avg_launch_ms / 1e3) / 1e9
    # HBM bytes per launch from the counter runs (FETCH_SIZE, WRITE_SIZE, scaled by
    # the same counters on a plain copy of known size), when they cover this workload.
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_icosphere-L%d-1024-64spp.txt" % args.level)
    if os.path.exists(tpath) and env is None:
        with open(tpath) as f:
            for line in f:
                w = line.split()
                if len(w) == 4 and w[0] == "traffic" and w[1] == dom_name and w[2] == "icosphere-L%d-1024-64spp" % args.level:
                    traffic = float(w[3])
    roof = {
        "bound": "hbm",
        "kernel": dom_name,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        # HBM bytes the counters measured per launch over the live launch time
        "measured_hbm_gbs": round(traffic / (avg_launch_ms / 1e3) / 1e9, 2) if traffic else None,
        "measured_hbm_frac": round(traffic / (avg_launch_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
        "limiter": "VALU with divergent lanes, 0.41 busy per instruction (profiles/counters_icosphere-L8-1024-64spp.txt): "
                   "the upper BVH levels are re-read from L2 and the Infinity Cache, so counter-measured HBM bytes "
                   "are about half the algorithmic bytes, and the algorithmic rate (frac) can pass HBM's peak",
        "algorithmic_bytes_per_launch": bytes_per_launch,
        "avg_launch_ms": round(avg_launch_ms, 4),
        "launches_per_step": launches_per_step,
        "kernel_time_share": round(share, 4),
    }
    out = {
        "metric": "Mpaths/s at fixed RMSE: batched wavefront vs plain megakernel; adjoint s/grad",
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: procedural displaced icosphere (seed 0) in the Cornell room"
        + ("; open room, procedural 1024x512 sky" if env is not None else ""),
        "config": {
            "workload": ("icosphere-L%d GGX conductor, %d triangles, %dx%d, %d spp, depth %d, NEE+MIS"
                         % (args.level, sa.n_tri, cam.width, cam.height, prm.spp, prm.max_depth))
            + (", environment light (open room, 1024x512 sky)" if env is not None else ""),
            "integrator": args.mode,
            "megakernel_mpaths_per_s": round(paths / ((ms_per_step if args.mode == "megakernel" else other_ms) / 1e3) / 1e6, 3)
            if (args.mode == "megakernel" or other_ms) else None,
            "wavefront_mpaths_per_s": round(paths / ((ms_per_step if args.mode == "wavefront" else other_ms) / 1e3) / 1e6, 3)
            if (args.mode == "wavefront" or other_ms) else None,
            "same_estimator": "wavefront and megakernel films are bit-identical (tests/test_gpu_parity.py), so RMSE is equal at equal spp",
            "rmse_vs_fp64_reference": ({k: round(v, 6) for k, v in rmse.items()} if rmse else None),
            "rmse_reference": ("FP64 16384-spp oracle render of rows %s (tests/golden), all channels" % ref_rows)
            if rmse else None,
            "paths_per_step": paths,
            "mrays_per_s": round(rays / (ms_per_step / 1e3) / 1e6, 2),
            "rays_per_step": int(rays),
            "bvh_nodes": int(sst.n_nodes),
            "bvh_build_ms": round(sst.build_ms, 1),
            "parallelism": "single GPU",
        },
        "roofline": roof,
    }
    sc.close()
    if not args.no_legs and env is None and args.level == 8:
        out["legs"] = {"batch64": batch_leg(g, ctx), "adjoint": adjoint_leg(g, ctx)}
    if args.calibrate_copy > 0:
        nbytes = args.calibrate_copy << 20
        cms = ctx.copy_probe(nbytes)
        log("calibration copy: k_copy_probe read %d and wrote %d bytes in %.3f ms (%.1f GB/s)"
            % (nbytes, nbytes, cms, 2 * nbytes / cms / 1e6))
        out["calibration_copy"] = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": round(cms, 4),
                                   "gbs": round(2 * nbytes / cms / 1e6, 1)}
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(sa, cam, prm, args.cpu_rows, env)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

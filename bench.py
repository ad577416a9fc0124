"""Benchmark: path tracing of the 1.31M-triangle displaced-icosphere GGX scene
at 1024x1024, 64 spp, max depth 8 (NEE + MIS) on one MI355X.

One step = one full render (67,108,864 camera paths) of that image with the
scene and BVH already resident in HBM.  Prints one JSON line on stdout.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode wavefront|megakernel]
                  [--no-cpu-baseline] [--level 8]
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


def algorithmic_bytes(st, paths):
    """Bytes the dominant kernels must touch: 64 per BVH node record, 48 per
    triangle record, 16 per shading point (material record) and 16 per path
    for the radiance record each sample writes."""
    return 64.0 * st.node_visits + 48.0 * st.tri_visits + 16.0 * st.shade_points + 16.0 * paths


def cpu_baseline(sa, cam, prm, rows):
    """The oracle's FP32 restatement of the reference CPU integrator
    (TiledIntegrator + PathTracing, reference BVH) on the host cores, over a
    band of `rows` image rows of the same workload."""
    from oracle import oracle as O

    threads = min(16, os.cpu_count() or 1)
    os_ = O.OracleScene(sa)
    row0 = cam.height // 2 - rows // 2
    p = O.params(prm.spp, prm.max_depth, prm.rr_depth, precision=0, nthreads=threads, use_bvh=1)
    _, _, paths, secs = os_.render(cam, p, row0, row0 + rows)
    os_.close()
    return {
        "value": round(paths / secs / 1e6, 4),
        "unit": "Mpaths/s",
        "cores": threads,
        "kind": "port",
        "sample": "rows %d-%d of the 1024^2 image (%d paths, %.1f s), same scene, spp and depth"
        % (row0, row0 + rows - 1, paths, secs),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", default="wavefront", choices=["wavefront", "megakernel"])
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--cpu-rows", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        log("bench.py measures one GPU: this build runs on a single MI355X; --gpus %d is not supported" % args.gpus)
        sys.exit(2)

    import gyunity_amd as g

    sa, cam, prm = g.generate_scene(g.GY_SCENE_ICOSPHERE_GGX, 0, args.level)
    prm.mode = g.GY_RENDER_WAVEFRONT if args.mode == "wavefront" else g.GY_RENDER_MEGAKERNEL
    paths = cam.width * cam.height * prm.spp
    ctx = g.Context(0)
    log("device:", ctx.device_name())
    t0 = time.perf_counter()
    sc = g.Scene(ctx, sa)
    sst = sc.stats()
    log("scene: %d triangles, BVH %d nodes depth %d, build %.1f ms (scene create %.1f ms)"
        % (sa.n_tri, sst.n_nodes, sst.max_depth, sst.build_ms, (time.perf_counter() - t0) * 1e3))
    for i in range(args.warmup):
        sc.render(cam, prm, download=False)
        log("warmup %d done" % i)
    kern_ms = np.zeros(8)
    kern_n = np.zeros(8)
    t0 = time.perf_counter()
    for i in range(args.steps):
        sc.render(cam, prm, download=False)
        st = ctx.stats()
        kern_ms += np.array(st.kernel_ms[:])
        kern_n += np.array(st.kernel_launches[:])
        log("step %d: %.1f ms, %.1f Mrays" % (i, st.total_ms, st.rays / 1e6))
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    rays = st.rays
    value = paths / (ms_per_step / 1e3) / 1e6

    # Dominant kernel and its algorithmic bytes per launch (counted by an
    # instrumented copy of the integrator over the same render).
    cnt = sc.count_traffic(cam, prm)
    dom = int(np.argmax(kern_ms))
    dom_name = g.lib().gy_kernel_name(dom).decode()
    launches_per_step = kern_n[dom] / args.steps
    avg_launch_ms = kern_ms[dom] / kern_n[dom]
    bytes_total = algorithmic_bytes(cnt, paths)
    share = kern_ms[dom] / max(kern_ms.sum(), 1e-9)
    bytes_per_launch = bytes_total / launches_per_step
    achieved = bytes_per_launch / (avg_launch_ms / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_%s.json" % dom_name)
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("workload") == "icosphere-L%d-1024-64spp" % args.level:
            traffic = t.get("bytes_per_launch")
    roof = {
        "bound": "hbm",
        "kernel": dom_name,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
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
        "data": "synthetic: procedural displaced icosphere (seed 0) in the Cornell room",
        "config": {
            "workload": "icosphere-L%d GGX conductor, %d triangles, %dx%d, %d spp, depth %d, NEE+MIS"
            % (args.level, sa.n_tri, cam.width, cam.height, prm.spp, prm.max_depth),
            "integrator": args.mode,
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
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(sa, cam, prm, args.cpu_rows)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Writes the converged-image fixtures of parity level (c).

FP64 renders at 16384 spp by the oracle's double-precision restatement of the
reference integrator (TiledIntegrator::Render + PathTracing::Li,
integrator/TiledIntegrator.cpp:7-60, integrator/PT.cpp:10-67):

  cornell64_fp64_16384spp.npy      the Cornell box at 64x64 (max depth 5,
                                   roulette from bounce 3): float32 (64, 64, 3)
  icosphere_rows_fp64_16384spp.npy rows ICO_ROWS of the 1024x1024 image of the
                                   1,310,720-triangle displaced icosphere (GGX
                                   conductor, depth 8, roulette from bounce 5):
                                   float32 (len(ICO_ROWS), 1024, 3)
  level_c.json                     the rows, the settings, and the RMSE against
                                   those references of the oracle's FP32 mode at
                                   the spp the GPU tests render with (the stated
                                   bounds: the GPU integrators compute the same
                                   FP32 estimator)

Pixel value = colour sum / weight.  Run from the repository root:
  python tests/golden/make_fp64_reference.py
(about 4 minutes on 8 cores).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import gyunity_amd as g  # noqa: E402
from oracle import oracle as O  # noqa: E402

REF_SPP = 16384
CORNELL_RES = 64
CORNELL_SPP = (16, 256, 4096)
ICO_ROWS = (128, 320, 512, 704, 896)
ICO_SPP = 64


def cornell_camera():
    sa, cam, prm = g.generate_scene(g.GY_SCENE_CORNELL)
    cam = g.camera(cam.pos[:], cam.look_at[:], cam.up[:], cam.fovy, CORNELL_RES, CORNELL_RES)
    return sa, cam, prm


def image(color, weight):
    return (color / np.maximum(weight, 1e-30)[..., None]).astype(np.float64)


def rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)))


def main():
    threads = min(16, os.cpu_count() or 1)
    meta = {"ref_spp": REF_SPP, "cornell_res": CORNELL_RES, "ico_rows": list(ICO_ROWS), "ico_spp": ICO_SPP,
            "cornell_fp32_rmse": {}, "ico_fp32_rmse": None}

    sa, cam, prm = cornell_camera()
    os_ = O.OracleScene(sa)
    t = time.time()
    c, w, _, _ = os_.render(cam, O.params(REF_SPP, prm.max_depth, prm.rr_depth, precision=1, nthreads=threads))
    ref = image(c, w)
    print("cornell FP64 %d spp: %.1f s" % (REF_SPP, time.time() - t), flush=True)
    np.save(os.path.join(HERE, "cornell64_fp64_16384spp.npy"), ref.astype(np.float32))
    for spp in CORNELL_SPP:
        c, w, _, _ = os_.render(cam, O.params(spp, prm.max_depth, prm.rr_depth, precision=0, nthreads=threads))
        meta["cornell_fp32_rmse"][str(spp)] = rmse(image(c, w), ref)
        print("cornell FP32 %d spp rmse %.6f" % (spp, meta["cornell_fp32_rmse"][str(spp)]), flush=True)
    os_.close()

    sa, cam, prm = g.generate_scene(g.GY_SCENE_ICOSPHERE_GGX, 0, 8)
    os_ = O.OracleScene(sa)
    rows_ref, rows_32 = [], []
    for y in ICO_ROWS:
        t = time.time()
        c, w, _, _ = os_.render(cam, O.params(REF_SPP, prm.max_depth, prm.rr_depth, precision=1, nthreads=threads), y, y + 1)
        rows_ref.append(image(c, w)[y])
        c, w, _, _ = os_.render(cam, O.params(ICO_SPP, prm.max_depth, prm.rr_depth, precision=0, nthreads=threads), y, y + 1)
        rows_32.append(image(c, w)[y])
        print("icosphere row %d FP64 %d spp: %.1f s" % (y, REF_SPP, time.time() - t), flush=True)
    ref = np.stack(rows_ref)
    np.save(os.path.join(HERE, "icosphere_rows_fp64_16384spp.npy"), ref.astype(np.float32))
    meta["ico_fp32_rmse"] = rmse(np.stack(rows_32), ref)
    print("icosphere FP32 %d spp rmse %.6f" % (ICO_SPP, meta["ico_fp32_rmse"]), flush=True)
    os_.close()
    with open(os.path.join(HERE, "level_c.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()

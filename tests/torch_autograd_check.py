"""Child process of test_torch_autograd_function_matches_adjoint: torch is
imported before the library, so both use the HIP runtime torch ships.  Prints
one JSON line."""
import json
import os
import sys

import torch  # first: libgyunity.so then binds to the HIP runtime torch loaded

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import gyunity_amd as g  # noqa: E402
from gyunity_amd.torch_render import render_textures, render_views  # noqa: E402
from tests.test_gpu_parity import _adjoint_setup, _image  # noqa: E402


def main():
    sa, cam, prm, gt, theta, n_alb = _adjoint_setup(spp=64)
    with g.Context(0) as ctx, g.Scene(ctx, sa) as sc:
        target = torch.from_numpy(_image(*sc.render(cam, prm)).astype(np.float32)).cuda()
        tex = torch.from_numpy(theta).cuda().requires_grad_(True)
        img = render_textures(sc, cam, prm, tex)
        loss = ((img - target) ** 2).mean()
        loss.backward()
        img_np = img.detach().cpu().numpy()
        color, weight = sc.film(cam.width, cam.height)
        image_equal = bool(np.array_equal(img_np, color / weight[..., None]))
        sc.set_textures(theta)
        ref_img = _image(*sc.render(cam, prm))
        dI = (2.0 * (ref_img - target.cpu().numpy()) / ref_img.size).astype(np.float32)
        ref = sc.adjoint(cam, prm, dI)
        got = tex.grad.cpu().numpy()
        # three views in one batch through the batch Function: forward equals
        # the batch films, backward equals the batch adjoint
        views = g.orbit_views(3, 40, seed=9, radius=0.3)
        with g.Batch(ctx) as b:
            for v in views:
                b.add(sc, v, prm)
            tex_v = torch.from_numpy(theta).cuda().requires_grad_(True)
            imgs = render_views(sc, b, tex_v)
            tgt = torch.rand_like(imgs)
            lv = ((imgs - tgt) ** 2).mean()
            lv.backward()
            films = np.concatenate([_image(*b.film(e)).astype(np.float32).ravel() for e in range(len(views))])
            views_equal = bool(np.array_equal(imgs.detach().cpu().numpy(), films))
            dIs = (2.0 * (films - tgt.cpu().numpy()) / films.size).astype(np.float32)
            sizes = [v.width * v.height * 3 for v in views]
            parts = np.split(dIs, np.cumsum(sizes)[:-1])
            ref_v = b.adjoint([p.reshape(v.height, v.width, 3) for p, v in zip(parts, views)], sc.texture_floats())
            got_v = tex_v.grad.cpu().numpy()
    print(json.dumps({"image_equal": image_equal, "loss": float(loss), "grad_max_abs": float(np.abs(ref).max()),
                      "grad_max_abs_diff": float(np.abs(got - ref).max()), "views_equal": views_equal,
                      "views_grad_max_abs": float(np.abs(ref_v).max()),
                      "views_grad_max_abs_diff": float(np.abs(got_v - ref_v).max())}), flush=True)


if __name__ == "__main__":
    main()

"""Differentiable render for PyTorch: image = render_textures(scene, cam,
params, textures) with textures a float32 CUDA tensor holding the scene's
texture floats (the layout of gy_scene_desc.tex_data).  Backward runs the
path-replay adjoint of libgyunity.so through the LibTorch
torch::autograd::Function in gy_autograd (torch_ext/gy_autograd.cpp).

Import torch before gyunity_amd loads libgyunity.so: torch ships its own HIP
runtime under the same soname (libamdhip64.so.7), and the one loaded first is
the one the process uses; loaded the other way round, torch finds no GPU."""
import os

_ext = None


def _load():
    global _ext
    if _ext is None:
        here = os.path.dirname(os.path.abspath(__file__))
        if not any(f.startswith("gy_autograd") and f.endswith(".so") for f in os.listdir(here)):
            raise RuntimeError("gy_autograd is not built; run python gyunity_amd/torch_ext/build.py")
        import torch  # noqa: F401  (loads libtorch before the extension)

        from . import _lib

        if _lib._lib is not None and not _lib._torch_first:
            raise RuntimeError("libgyunity.so was loaded before torch: import torch first so that both use one HIP runtime")
        from . import gy_autograd

        _ext = gy_autograd
    return _ext


def render_textures(scene, cam, params, textures):
    camera = [*cam.pos, *cam.look_at, *cam.up, cam.fovy, cam.film_dist, cam.width, cam.height]
    prm = [params.spp, params.max_depth, params.rr_depth, params.mode, params.samples_per_pass, params.block_size,
           params.refill_threshold, params.sampler, params.ray_sort]
    return _load().render_textures(textures, scene.h.value, [float(v) for v in camera], [int(v) for v in prm])


def render_views(scene, batch, textures):
    """Every entry of `batch` (views of `scene`) rendered with `textures` in
    one lane space; returns the entries' images (H, W, 3 each) concatenated
    into one flat tensor.  Backward replays every view in one launch."""
    return _load().render_views(textures, batch.h.value, scene.h.value)

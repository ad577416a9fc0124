"""gy_render_noalloc against gy_render, through the C ABI (GPU).

The allocation-free entry point renders with the megakernel into a device
film the caller owns, using the library's per-device scratch array; its film
must equal gy_render's bit for bit.  Two calls in a row share the scratch,
the second with a smaller image and fewer samples.  The film buffer is the
caller's, allocated here with the HIP runtime the library uses.
"""
import ctypes as C

import numpy as np
import pytest

import gyunity_amd as g
from gyunity_amd._lib import GY_ERR_UNSUPPORTED

pytestmark = pytest.mark.gpu


def _hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return h


def _noalloc_film(scene, cam, prm, hip, buf):
    W, H = cam.width, cam.height
    rc = g.lib().gy_render_noalloc(scene.h, C.byref(cam), C.byref(prm), buf)
    if rc != 0:
        return rc, None, None
    out = np.zeros((H, W, 4), np.float32)
    assert hip.hipMemcpy(out.ctypes.data, buf, out.nbytes, 2) == 0  # device to host
    return rc, out[..., :3].copy(), out[..., 3].copy()


def test_noalloc_film_equals_render(gpu_ctx):
    hip = _hip()
    buf = C.c_void_p()
    nbytes = 1024 * 1024 * 16
    assert hip.hipMalloc(C.byref(buf), nbytes) == 0
    try:
        # garbage in the caller's buffer: the entry point writes the film whole
        assert hip.hipMemset(buf, 0x7f, nbytes) == 0
        sa, cam, prm = g.generate_scene(g.GY_SCENE_CORNELL)
        with g.Scene(gpu_ctx, sa) as sc:
            cases = [
                # Cornell box, 256^2, 16 spp: two passes of the scratch
                (256, 256, 16, 0),
                # 1024^2 x 16 spp: passes of 3 samples (what fits), against
                # gy_render's single pass
                (1024, 1024, 16, 0),
                # smaller image and fewer samples, same scratch
                (96, 64, 4, 0),
                # an explicit pass size
                (128, 128, 8, 3),
            ]
            for W, H, spp, spp_pass in cases:
                cam.width, cam.height = W, H
                prm.spp = spp
                prm.mode = g.GY_RENDER_WAVEFRONT
                prm.samples_per_pass = 0
                color, weight = sc.render(cam, prm)
                prm.mode = g.GY_RENDER_MEGAKERNEL
                prm.samples_per_pass = spp_pass
                rc, nc, nw = _noalloc_film(sc, cam, prm, hip, buf)
                assert rc == 0, g.lib().gy_last_error()
                assert np.array_equal(nw, weight), (W, H, spp)
                assert np.array_equal(nc, color), (W, H, spp)
                assert weight.sum() >= W * H * spp
            # beyond the scratch: too many pixels, an explicit pass that does
            # not fit, the wavefront form
            cam.width, cam.height = 2048, 2048
            prm.spp, prm.samples_per_pass = 1, 0
            assert g.lib().gy_render_noalloc(sc.h, C.byref(cam), C.byref(prm), buf) == GY_ERR_UNSUPPORTED
            assert b"scratch" in g.lib().gy_last_error()
            cam.width, cam.height = 1024, 1024
            prm.spp, prm.samples_per_pass = 16, 8
            assert g.lib().gy_render_noalloc(sc.h, C.byref(cam), C.byref(prm), buf) == GY_ERR_UNSUPPORTED
            prm.samples_per_pass = 0
            prm.mode = g.GY_RENDER_WAVEFRONT
            assert g.lib().gy_render_noalloc(sc.h, C.byref(cam), C.byref(prm), buf) == GY_ERR_UNSUPPORTED
    finally:
        hip.hipFree(buf)

"""mb3d_hip_march_noalloc, the march that allocates nothing (its buffers in
the library's scratch array), against mb3d_hip_march on the same inputs:
two calls in a row, the second smaller, equal bit for bit; a picture too
large for the array is refused, and marched in bands of rows instead."""
import numpy as np
import pytest

import mb3d_hip as H
from test_gpu_batch import device_out, host

pytestmark = pytest.mark.gpu


def noalloc(lib, p, out, row0, rows, opts=None):
    return lib.mb3d_hip_march_noalloc(H.C.byref(p), row0, rows, H.C.byref(out), H.C.byref(opts) if opts else None)


def test_no_alloc_march_equals_the_march(hip):
    lib = hip.load()
    for w, h, seed, cone in ((640, 360, 1, False), (320, 180, 2, True)):
        p = H.params_from_scene(H.scene_bulb(w, h, seed=seed))
        opts = H.options(cone=cone)
        want = hip.march(p, opts=opts)
        bufs, out = device_out(w * h)
        H.check(noalloc(lib, p, out, 0, h, opts), "mb3d_hip_march_noalloc")
        got = host(bufs)
        assert np.array_equal(got["depth"], want["depth"]), (w, h)
        assert np.array_equal(got["depth64"], want["depth64"]), (w, h)
        assert np.array_equal(got["iters"].astype(np.uint16), want["iters"]), (w, h)
        assert np.array_equal(got["normals"].reshape(-1, 3), want["normals"]), (w, h)
        assert np.array_equal(got["gbuf"].view(np.uint32).reshape(-1, 5), want["gbuf"]), (w, h)
        assert H.counters()["hits"] == want["counters"]["hits"]


def test_too_large_is_refused_and_bands_work(hip):
    lib = hip.load()
    p = H.params_from_scene(H.scene_bulb(1920, 1080, seed=3))
    bufs, out = device_out(1920 * 1080)
    st = noalloc(lib, p, out, 0, 1080)
    assert st == H.ERR_INVALID and b"bands of rows" in H.load().mb3d_hip_last_error()
    # two bands of 540 rows: each band's outputs start at its own first pixel
    import torch

    for r0 in (0, 540):
        band = {k: v[r0 * 1920 * (v.numel() // (1920 * 1080)):] for k, v in bufs.items()}
        o = H.Out(*(band[k].data_ptr() for k in ("depth", "depth64", "iters", "normals", "gbuf")), 1, 0)
        H.check(noalloc(lib, p, o, r0, 540), "mb3d_hip_march_noalloc band")
    got = host(bufs)
    want = hip.march(p, outputs=("depth64", "iters"))
    assert np.array_equal(got["depth64"], want["depth64"])
    assert np.array_equal(got["iters"].astype(np.uint16), want["iters"])

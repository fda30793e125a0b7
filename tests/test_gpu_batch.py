"""The batched path: frames of uneven sizes in one persistent launch, the
HIP-graph plan and the autotuner, against single-frame calls (which are
checked against the oracle in test_gpu_parity.py) and against the oracle."""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O
import parity

pytestmark = pytest.mark.gpu

SIZES = ((96, 72), (160, 120), (64, 48), (200, 150))


def device_out(n):
    import torch

    dev = torch.device("cuda:0")
    bufs = {
        "depth": torch.zeros(n, dtype=torch.float32, device=dev),
        "depth64": torch.zeros(n, dtype=torch.float64, device=dev),
        "iters": torch.zeros(n, dtype=torch.int16, device=dev),
        "normals": torch.zeros(n * 3, dtype=torch.float32, device=dev),
        "gbuf": torch.zeros(n * 5, dtype=torch.int32, device=dev),
    }
    out = H.Out(*(bufs[k].data_ptr() for k in ("depth", "depth64", "iters", "normals", "gbuf")), 1, 0)
    return bufs, out


def host(bufs):
    import torch

    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in bufs.items()}


def frames(n=6, seed=3):
    return H.animation(n, seed=seed, sizes=SIZES)


def test_batch_equals_single_frames(hip):
    scenes = frames()
    arr = H.params_array(scenes)
    n = sum(s.width * s.height for s in scenes)
    bufs, out = device_out(n)
    H.check(hip.load().mb3d_hip_march_batch(arr, len(arr), H.C.byref(out), None), "batch")
    got = host(bufs)
    off = 0
    for i, s in enumerate(scenes):
        m = s.width * s.height
        one = hip.march(arr[i])
        assert np.array_equal(got["depth"][off:off + m], one["depth"]), i
        assert np.array_equal(got["iters"][off:off + m].astype(np.uint16), one["iters"]), i
        assert np.array_equal(got["gbuf"][off * 5:(off + m) * 5].view(np.uint32), one["gbuf"].ravel()), i
        off += m


def test_batch_matches_the_oracle(hip):
    scenes = frames(4, seed=5)
    arr = H.params_array(scenes)
    n = sum(s.width * s.height for s in scenes)
    bufs, out = device_out(n)
    H.check(hip.load().mb3d_hip_march_batch(arr, len(arr), H.C.byref(out), None), "batch")
    got = host(bufs)
    off = 0
    for i, s in enumerate(scenes):
        m = s.width * s.height
        r = O.render(O.params(s), threads=0)
        g = {"depth64": got["depth64"][off:off + m], "iters": got["iters"][off:off + m].astype(np.uint16),
             "normals": got["normals"][3 * off:3 * (off + m)].reshape(-1, 3)}
        res = parity.compare(g, r, s.width, s.height, arr[i].step_width, 2.0)
        assert res["mask_agree"] >= 0.999 and res["depth_ok_frac"] >= 0.999 and res["normal_ok_frac"] >= 0.999, res
        off += m


def test_graph_replays_equal_the_eager_batch(hip):
    """The captured graph, replayed several times with eager work and
    synchronisation between the replays, gives the eager batch's results."""
    import torch

    scenes = frames(5, seed=7)
    arr = H.params_array(scenes)
    n = sum(s.width * s.height for s in scenes)
    bufs, out = device_out(n)
    H.check(hip.load().mb3d_hip_march_batch(arr, len(arr), H.C.byref(out), None), "batch")
    eager = host(bufs)
    plan = H.Plan(arr, out)
    try:
        for rep in range(4):
            for v in bufs.values():
                v.fill_(7)  # eager work on the same buffers between replays
            torch.cuda.synchronize()
            c = plan.launch(1)
            got = host(bufs)
            for k in eager:
                assert np.array_equal(got[k], eager[k]), (rep, k)
            assert c["pixels"] == n
        c = plan.launch(3)
        assert np.array_equal(host(bufs)["depth"], eager["depth"])
    finally:
        plan.close()


def test_autotune_returns_a_working_choice(hip):
    scenes = frames(3, seed=9)
    arr = H.params_array(scenes)
    n = sum(s.width * s.height for s in scenes)
    bufs, out = device_out(n)
    best, ms = H.autotune(arr, out)
    assert ms > 0 and best.tile_w * best.tile_h >= 32 and best.waves_per_cu >= 8
    H.check(hip.load().mb3d_hip_march_batch(arr, len(arr), H.C.byref(out), H.C.byref(best)), "batch")
    a = host(bufs)["depth"].copy()
    H.check(hip.load().mb3d_hip_march_batch(arr, len(arr), H.C.byref(out), None), "batch")
    assert np.array_equal(a, host(bufs)["depth"])


def test_batch_rejects_mixed_formulas(hip):
    scenes = [H.scene_bulb(64, 48, seed=1), H.scene_box(64, 48, seed=1)]
    arr = H.params_array(scenes)
    bufs, out = device_out(64 * 48 * 2)
    assert hip.load().mb3d_hip_march_batch(arr, 2, H.C.byref(out), None) == -1
    assert "share" in H.last_error()


def test_iteration_cap_split_keeps_the_bar(hip):
    """it_split hands rays still marching after that many FP32 steps to the
    f64 finish, which resumes them: the pictures stay within the bar."""
    sc = H.scene_bulb(320, 240, seed=4)
    p = hip.params_from_scene(sc)
    r = O.render(O.params(sc), threads=0)
    for sp in (8, 64):
        g = hip.march(p, opts=H.options(it_split=sp))
        m = parity.compare(g, r, p.width, p.height, p.step_width, 2.0)
        assert m["mask_agree"] >= 0.999 and m["depth_ok_frac"] >= 0.999, (sp, m)
        assert m["normal_ok_frac"] >= 0.999, (sp, m)


def test_animation_of_64_frames_in_one_plan(hip):
    """The batched configuration at its full size: 64 frames with sizes from
    {320x240, 640x480, 1280x720, 1920x1080} along a seeded camera path, one
    graph-captured plan after the autotuner's choice; a sampled frame is
    checked against the oracle."""
    import time

    import torch

    scenes = H.animation(64, seed=11)
    arr = H.params_array(scenes)
    n = sum(s.width * s.height for s in scenes)
    dev = torch.device("cuda:0")
    depth = torch.zeros(n, dtype=torch.float32, device=dev)
    d64 = torch.zeros(n, dtype=torch.float64, device=dev)
    iters = torch.zeros(n, dtype=torch.int16, device=dev)
    normals = torch.zeros(3 * n, dtype=torch.float32, device=dev)
    out = H.Out(depth.data_ptr(), d64.data_ptr(), iters.data_ptr(), normals.data_ptr(), 0, 1, 0)
    best, ms = H.autotune(arr, out)
    plan = H.Plan(arr, out, best)
    try:
        plan.launch(1)
        torch.cuda.synchronize()
        t = time.perf_counter()
        plan.launch(3)
        wall = (time.perf_counter() - t) / 3
    finally:
        plan.close()
    print(f"64 frames, {n / 1e6:.1f} Mpx: {wall * 1e3:.1f} ms a launch, {n / wall / 1e6:.0f} Mpix/s "
          f"(tile {best.tile_w}x{best.tile_h}, {best.waves_per_cu} waves/CU, split {best.it_split})")
    off = 0
    for i, s in enumerate(scenes):
        m = s.width * s.height
        if i == 5:
            r = O.render(O.params(s), threads=0)
            g = {"depth64": d64[off:off + m].cpu().numpy(), "iters": iters[off:off + m].cpu().numpy().astype(np.uint16),
                 "normals": normals[3 * off:3 * (off + m)].cpu().numpy().reshape(-1, 3)}
            mm = parity.compare(g, r, s.width, s.height, arr[i].step_width, 2.0)
            assert mm["mask_agree"] >= 0.999 and mm["depth_ok_frac"] >= 0.999, mm
        off += m

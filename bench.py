#!/usr/bin/env python3
"""Benchmark of the HIP DE ray marcher on one MI355X: learned + batched
against plain.

Workload: the power-8 Mandelbulb Z-buffer at 1920 x 1080 (the small
configuration's scene -- 60 iterations, bailout r^2 = 16, analytic DE
0.5 r ln r / dr, camera start plane 2.5 from the origin, FOV 30, seeded yaw
and pitch, 6 bisection steps, 500-step cap -- rendered at 1080p).  Inputs
are parameters; the outputs stay in device memory (HBM).

- plain: one frame per call of mb3d_hip_march, the plain FP32 march and its
  f64 finish.
- plain + batched: the same frames in one HIP-graph plan without the
  network (what batching alone gives).
- learned + batched (the value): the warm-start network, trained first
  (untimed) on the plain marcher's (ray, hit distance) pairs of the same
  view at 480 x 270, and `--batch` copies of the frame in one HIP-graph
  plan (one persistent launch of each kernel for all of them); one step is
  one replay of the plan.

Prints one JSON line: Mpixels/s of the three paths, the roofline figure of
the learned + batched path's dominant kernel (its algorithmic FLOPs per
launch over its HIP-event time on the library's stream), the CPU baseline
(the f64 oracle on the host cores) and the accuracy of the plain and the
learned march against the oracle.  Beside the metric: the painter's kernel
on the frame's G-buffer (device time, bytes, the oracle's painter on the
host cores), and the frame's first-bounce reflection rays through the
secondary-ray kernels (device time, the oracle on a sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mb3d-rs_amd"))

import numpy as np  # noqa: E402

import mb3d_hip as H  # noqa: E402

PEAK_FP32 = 157.3e12  # MI355X FP32 vector, FLOP/s (published)
PEAK_FP64 = 78.6e12   # MI355X FP64 vector, FLOP/s (published)
PEAK_HBM = 8.0e12     # MI355X HBM3E, bytes/s (published)
# algorithmic FLOPs per unit of work (the reference's formulas, counted by op)
F_P8 = 60.0     # int_pow power 8 (src/formulas.rs:837-858) + the rout update
F_BOX = 31.0    # amazing_box (src/formulas.rs:707-736) + the rout update
F_STEP = 25.0   # one march step of march_from (src/calc.rs:2555-2640)
F_DE = 8.0      # DE tail: sqrt, log, divide, scale, clamp


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed replays of the batch (and plain frames x batch)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=48, help="frames per launch of the batched paths (at most 64)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and accuracy legs")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also copy a 512 MiB device buffer once (the HBM counter scale of a profile)")
    return ap.parse_args()


def flops_f32(c: dict) -> float:
    """The FP32 march's: its iterations, DEs and steps (the normals pass's
    FP32 stencil -- its_normals_f32, six DEs per hit -- counts with the
    finish)."""
    its = c["its_p8_f32"] - c["its_normals_f32"]
    return its * F_P8 + c["its_box_f32"] * F_BOX + (c["de_evals_f32"] - 6 * c["hits"]) * F_DE + c["steps"] * F_STEP


def flops_f64(c: dict) -> float:
    """The finish and normals passes': the f64 iterations and DEs, and the
    normals' FP32 stencil (unpacked FP32 issues at the FP64 rate)."""
    return (c["its_p8_f64"] * F_P8 + c["its_box_f64"] * F_BOX + c["de_evals_f64"] * F_DE +
            c["its_normals_f32"] * F_P8 + 6 * c["hits"] * F_DE)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): every box of this build has one GPU, "
              "and the marcher shards by pixel bands or frames per GPU (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    lib = H.load()
    H.check(lib.mb3d_hip_init(0), "mb3d_hip_init")
    sc = H.scene_bulb(a.width, a.height, seed=a.seed)
    p = H.params_from_scene(sc)
    n = a.width * a.height
    dev = torch.device("cuda:0")
    depth = torch.empty(n * a.batch, dtype=torch.float32, device=dev)
    iters = torch.empty(n * a.batch, dtype=torch.int16, device=dev)
    normals = torch.empty(n * a.batch * 3, dtype=torch.float32, device=dev)
    out = H.Out(depth.data_ptr(), None, iters.data_ptr(), normals.data_ptr(), None, 1, 0)
    plain_opts = H.options()

    if a.copy_calibration:
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # --- the network (untimed): the plain marcher's pairs of the same view
    t = time.perf_counter()
    tw, th = 480, a.height * 480 // a.width
    net, curve = H.train_warm_start([H.params_from_scene(H.scene_bulb(tw, th, seed=a.seed))],
                                    steps=a.train_steps, batch=8192, lr=3e-3, seed=0)
    train_s = time.perf_counter() - t
    # hit jumps (verified) and miss jumps (opt-in: the network's view is the
    # benchmark's own, trained at 480 x 270; its mask error is reported)
    learn_opts = H.options(mlp=net, conf_miss=0.02)

    def plain_frame():
        st = lib.mb3d_hip_march(H.C.byref(p), None, 0, a.height, H.C.byref(out), H.C.byref(plain_opts),
                                H.PROGRESS_FN(0), H.CANCEL_FN(0), None)
        H.check(st, "mb3d_hip_march")
        return H.counters()

    # --- plain: one frame per call, the same number of frames as the batch leg
    for _ in range(a.warmup):
        plain_frame()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps * a.batch):
        cp = plain_frame()
    torch.cuda.synchronize()
    plain_wall = time.perf_counter() - t0
    plain_value = n * a.steps * a.batch / plain_wall / 1e6

    # --- plain + batched: the same plan without the network (what the
    # batching alone gives); both batched legs take the autotuner's tile,
    # grid and iteration-cap split for their own options (untimed)
    frames = H.params_array([sc] * a.batch)
    pb_opts, _ = H.autotune(frames, out, H.options())
    lb_opts, _ = H.autotune(frames, out, H.options(mlp=net, conf_miss=0.02))
    plan = H.Plan(frames, out, pb_opts)
    for _ in range(a.warmup):
        plan.launch(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        plan.launch(1)
    torch.cuda.synchronize()
    pb_wall = time.perf_counter() - t0
    plan.close()
    plain_batched_value = n * a.batch * a.steps / pb_wall / 1e6

    # --- learned + batched: the frame `batch` times in one graph plan
    plan = H.Plan(frames, out, lb_opts)
    for _ in range(a.warmup):
        plan.launch(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        cl = plan.launch(1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    plan.close()
    value = n * a.batch * a.steps / wall / 1e6

    # --- the dominant kernel of the learned + batched path: per-kernel
    # HIP-event times of the same batch launched eagerly (events bracket the
    # march -- with the inference ahead of it -- and the finish on the
    # library's stream; a captured graph does not time its nodes), averaged
    ms_m, ms_f, cf = [], [], None
    for _ in range(max(3, a.steps)):
        H.check(lib.mb3d_hip_march_batch(frames, a.batch, H.C.byref(out), H.C.byref(lb_opts)), "march_batch")
        cf = H.counters()
        ms_m.append(cf["march_f32_ms"])
        ms_f.append(cf["finish_f64_ms"])
    t_march, t_finish = float(np.mean(ms_m)), float(np.mean(ms_f))
    k_march = {"kernel": "k_mlp_propose + k_march32 (batch)", "bound": "valu-fp32", "kernel_ms": t_march,
               "flops_per_launch": flops_f32(cf), "peak": PEAK_FP32 / 1e12}
    k_finish = {"kernel": "k_finish64 + k_normals (batch)", "bound": "valu-fp64", "kernel_ms": t_finish,
                "flops_per_launch": flops_f64(cf), "peak": PEAK_FP64 / 1e12}
    for k in (k_march, k_finish):
        k["achieved"] = k["flops_per_launch"] / (k["kernel_ms"] * 1e-3) / 1e12
        k["frac"] = k["achieved"] / k["peak"]
    dom = k_finish if t_finish >= t_march else k_march
    roofline = {
        "bound": dom["bound"],
        "kernel": dom["kernel"],
        "achieved": dom["achieved"],
        "peak": dom["peak"],
        "unit": "TFLOP/s",
        "frac": dom["frac"],
        "traffic": None,
        "kernel_ms": dom["kernel_ms"],
        "flops_per_launch": dom["flops_per_launch"],
        "kernels": [k_march, k_finish],
    }
    # --- the painter (beside the metric): the frame's G-buffer, marched once
    # into device memory, lit and coloured on the device by k_paint (the
    # reference's default lighting); its HIP-event time per launch
    gbuf = torch.empty(n * 5, dtype=torch.int32, device=dev)
    rgb = torch.empty(n * 3, dtype=torch.uint8, device=dev)
    out_g = H.Out(None, None, None, None, gbuf.data_ptr(), 1, 0)
    H.check(lib.mb3d_hip_march(H.C.byref(p), None, 0, a.height, H.C.byref(out_g), H.C.byref(plain_opts),
                               H.PROGRESS_FN(0), H.CANCEL_FN(0), None), "mb3d_hip_march")
    cam, lv = H.paint_setup(sc, p, H.lighting_default())
    pms = []
    for i in range(a.warmup + max(5, a.steps)):
        H.check(lib.mb3d_hip_paint(H.C.byref(cam), H.C.byref(lv), H.C.c_void_p(gbuf.data_ptr()), 0, 0, a.width,
                                   a.height, H.PAINT_U8, H.C.c_void_p(rgb.data_ptr()), 1), "mb3d_hip_paint")
        if i >= a.warmup:
            pms.append(H.counters()["paint_ms"])
    t_paint = float(np.mean(pms))
    paint_bytes = n * (20 + 3)  # one 20-byte record read, 3 bytes written per pixel
    paint = {"kernel": "k_paint<u8>", "kernel_ms": t_paint, "mpix_s": n / (t_paint * 1e-3) / 1e6,
             "bytes_per_launch": paint_bytes, "achieved_gbs": paint_bytes / (t_paint * 1e-3) / 1e9,
             "peak_gbs": PEAK_HBM / 1e9, "frac_hbm": paint_bytes / (t_paint * 1e-3) / PEAK_HBM}
    paint_gbuf = gbuf.cpu().numpy().view(np.uint32).reshape(n, 5)

    # --- the post passes (beside the metric): the frame marched with one hard
    # shadow and quality-1 DE AO (k_shadow, k_deao); their HIP-event time
    post = H.post_job(lights=[(0.6, -0.5, -0.6)], deao_quality=1, params=p)
    post_ms = []
    for i in range(a.warmup + max(3, a.steps // 2)):
        H.check(lib.mb3d_hip_march(H.C.byref(p), H.C.byref(post), 0, a.height, H.C.byref(out_g), H.C.byref(plain_opts),
                                   H.PROGRESS_FN(0), H.CANCEL_FN(0), None), "mb3d_hip_march")
        if i >= a.warmup:
            post_ms.append(H.counters()["post_ms"])
    post_leg = {"kernels": "k_shadow + k_deao", "post_ms": float(np.mean(post_ms)),
                "mpix_s": n / (float(np.mean(post_ms)) * 1e-3) / 1e6, "job": "1 hard shadow, DE AO quality 1 (8 rays)"}

    # --- secondary rays (beside the metric): the frame's first-bounce
    # reflection rays from every surface pixel, marched on the device with
    # one hard shadow and quality-2 DE AO at their hits (tasks and hits in
    # device memory); the ray kernels' HIP-event time per batch
    g_host = H.march(p, outputs=("depth64", "normals"))
    tasks = H.reflection_tasks(p, g_host)
    tasks["ao_q"] = 2
    ray_opts = H.ray_opts(p, lights=[(0.6, -0.5, -0.6)], ao_max_len=1.0)
    d_tasks = torch.from_numpy(tasks.view(np.uint8).copy()).to(dev)
    d_hits = torch.empty(len(tasks) * H.RAY_HIT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    rms = []
    for i in range(a.warmup + max(3, a.steps // 2)):
        H.check(lib.mb3d_hip_march_rays(H.C.byref(p), H.C.c_void_p(d_tasks.data_ptr()), len(tasks), 0,
                                        H.C.byref(ray_opts), H.C.c_void_p(d_hits.data_ptr()), 1), "mb3d_hip_march_rays")
        if i >= a.warmup:
            rms.append(H.counters()["kernel_ms"])
    t_rays = float(np.mean(rms))
    hits_np = d_hits.cpu().numpy().view(H.RAY_HIT_DTYPE)
    rays = {"kernel": "k_rays + k_ray_post", "rays": len(tasks),
            "hits": int(((hits_np["flags"] & H.RAY_HIT) != 0).sum()), "kernel_ms": t_rays,
            "mrays_s": len(tasks) / (t_rays * 1e-3) / 1e6, "hit_pass": "1 hard shadow, DE AO quality 2 (7 rays)"}

    # --- the Monte Carlo renderer's marches (beside the metric): every pixel's
    # primary ray with the inside probe, then a shadow ray towards one light
    # from each surface it met (tasks and results in device memory); the k_mc
    # batches' HIP-event time
    def mc_time(t, reps):
        d_t = torch.from_numpy(t.view(np.uint8).copy()).to(dev)
        d_r = torch.empty(len(t) * H.MC_RES_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        ms = []
        for i in range(a.warmup + reps):
            H.check(lib.mb3d_hip_mc_batch(H.C.byref(p), H.C.c_void_p(d_t.data_ptr()), len(t), 0,
                                          H.C.c_void_p(d_r.data_ptr()), 1), "mb3d_hip_mc_batch")
            if i >= a.warmup:
                ms.append(H.counters()["kernel_ms"])
        return float(np.mean(ms)), d_r.cpu().numpy().view(H.MC_RES_DTYPE)

    mc_light = (0.6, -0.5, -0.6)
    mc_prim = H.mc_primary_tasks(p, seed=a.seed)
    t_mc0, mc_res = mc_time(mc_prim, max(3, a.steps // 2))
    mc_shad, _ = H.mc_shadow_tasks(p, mc_res, mc_prim["dir"], mc_light)
    t_mc1, mc_sres = mc_time(mc_shad, max(3, a.steps // 2))
    mc = {"kernel": "k_mc",
          "primary": {"rays": len(mc_prim), "surfaces": int((mc_res["kind"] == 1).sum()), "kernel_ms": t_mc0,
                      "mrays_s": len(mc_prim) / (t_mc0 * 1e-3) / 1e6},
          "shadow": {"rays": len(mc_shad), "blocked": int((mc_sres["len"] > 0.01).sum()), "kernel_ms": t_mc1,
                     "mrays_s": len(mc_shad) / (t_mc1 * 1e-3) / 1e6, "length": 300.0}}

    result = {
        "metric": "Mpixels/s 1080p Mandelbulb-p8 Z-buffer; learned vs plain at equal depth error",
        "value": value,
        "unit": "Mpixels/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic (seeded camera, random-initialised network trained on the plain marcher's pairs)",
        "config": {
            "workload": f"mandelbulb-p8 z-buffer {a.width}x{a.height}, analytic DE, 60 its, bailout r2=16, "
                        f"fov 30, seed {a.seed}; learned warm start + {a.batch} frames per graph launch",
            "path": "learned+batched",
            "batch": a.batch,
            "autotuned": {"learned": [lb_opts.tile_w, lb_opts.tile_h, lb_opts.waves_per_cu, lb_opts.it_split],
                          "plain_batched": [pb_opts.tile_w, pb_opts.tile_h, pb_opts.waves_per_cu, pb_opts.it_split]},
        },
        "plain": {"value": plain_value, "unit": "Mpixels/s", "ms_per_frame": plain_wall * 1e3 / (a.steps * a.batch),
                  "kernel_ms_per_frame": cp["kernel_ms"], "march_f32_ms": cp["march_f32_ms"],
                  "finish_f64_ms": cp["finish_f64_ms"], "steps": cp["steps"], "its_p8_f32": cp["its_p8_f32"]},
        "plain_batched": {"value": plain_batched_value, "unit": "Mpixels/s",
                          "ms_per_frame": pb_wall * 1e3 / (a.steps * a.batch)},
        "speedup_vs_plain": value / plain_value,
        "speedup_vs_plain_batched": value / plain_batched_value,
        "learned": {"ms_per_frame": wall * 1e3 / (a.steps * a.batch), "jumps": cl["jumps"] // a.batch,
                    "rollbacks": cl["rollbacks"] // a.batch, "steps": cl["steps"] // a.batch,
                    "its_p8_f32": cl["its_p8_f32"] // a.batch, "train_s": train_s,
                    "train_loss": [float(curve[:20].mean()), float(curve[-100:].mean())]},
        "roofline": roofline,
        "paint": paint,
        "post": post_leg,
        "rays": rays,
        "mc": mc,
    }
    if not a.no_cpu:
        result["cpu_baseline"], result["accuracy"] = cpu_leg(sc, p, a, learn_opts)
        paint["cpu"] = paint_cpu_leg(cam, lv, paint_gbuf, result["cpu_baseline"]["cores"])
        rays["cpu"] = rays_cpu_leg(p, tasks, ray_opts, result["cpu_baseline"]["cores"])
        mc["cpu"] = mc_cpu_leg(p, mc_prim, mc_shad, result["cpu_baseline"]["cores"])
    net.close()
    print(json.dumps(result))


def cpu_leg(sc, p, a, learn_opts):
    """The f64 oracle (a restatement of the reference's CPU marcher) on the
    host cores, on the same frame, and both paths' accuracy against it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    import parity

    threads = a.cpu_threads or min(os.cpu_count() or 1, 16)
    po = O.params(sc)
    t = time.perf_counter()
    ref = O.render(po, threads=threads)
    dt = time.perf_counter() - t
    keys = ("mask_agree", "depth_rel_p99", "depth_rel_max", "depth_ok_frac", "normal_ok_frac", "normal_deg_p99",
            "iters_equal_inner")
    acc = {}
    for name, o in (("plain", None), ("learned", learn_opts)):
        g = H.march(p, opts=o)
        m = parity.compare(g, ref, p.width, p.height, p.step_width, 2.0)
        acc[name] = {k: m[k] for k in keys}
    cpu = {
        "value": p.width * p.height / dt / 1e6,
        "unit": "Mpixels/s",
        "cores": threads,
        "kind": "port",
        "sample": f"one full {p.width}x{p.height} frame of the same workload, {dt:.2f} s, "
                  f"{threads} threads, rows interleaved as src/render.rs:757-803",
    }
    return cpu, acc


def paint_cpu_leg(cam, lv, g, threads):
    """The oracle's painter (a restatement of the reference's
    pixel_color_s) on the host cores, on the same G-buffer."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.paint(cam, lv, g, threads=threads)  # the diffuse tables, once
    t = time.perf_counter()
    O.paint(cam, lv, g, threads=threads)
    dt = time.perf_counter() - t
    return {"value": g.shape[0] / dt / 1e6, "unit": "Mpixels/s", "cores": threads, "kind": "port",
            "sample": f"one full {cam.width}x{cam.height} G-buffer, {dt:.3f} s"}


def rays_cpu_leg(p, tasks, opts, threads, sample=1_000_000):
    """The oracle's secondary-ray march and hit pass (a restatement of the
    reference's SrTracer::march_ray, calc_hs and deao_1pos) on the host
    cores, on every k-th task of the same batch (about `sample` of them, so
    that the sample covers the whole picture)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    k = max(1, len(tasks) // sample)
    t_s = np.ascontiguousarray(tasks[::k])
    t = time.perf_counter()
    O.march_rays(p, t_s, opts=opts, threads=threads)
    dt = time.perf_counter() - t
    return {"value": len(t_s) / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"every {k}th ray of the batch, {len(t_s)} rays, {dt:.2f} s"}


def mc_cpu_leg(p, prim, shad, threads, sample=4_000_000):
    """The oracle's Monte Carlo marches (a restatement of the reference's
    McMarch::run) on the host cores, on every k-th task of each batch."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    out = {"unit": "Mrays/s", "cores": threads, "kind": "port"}
    for name, t in (("primary", prim), ("shadow", shad)):
        k = max(1, len(t) // sample)
        t_s = np.ascontiguousarray(t[::k])
        t0 = time.perf_counter()
        O.mc_batch(p, t_s, threads=threads)
        dt = time.perf_counter() - t0
        out[name] = len(t_s) / dt / 1e6
        out[name + "_sample"] = f"every {k}th task, {len(t_s)} tasks, {dt:.2f} s"
    return out


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the HIP DE ray marcher on one MI355X: Mpixels/s of the
1080p Mandelbulb Z-buffer at the plain marcher's depth error, the learned
warm start and batching against the plain FP32 marcher.

Workload: the power-8 Mandelbulb Z-buffer at 1920 x 1080 (the small
configuration's scene -- 60 iterations, bailout r^2 = 16, analytic DE
0.5 r ln r / dr, camera start plane 2.5 from the origin, FOV 30, seeded yaw
and pitch, 6 bisection steps, 500-step cap -- rendered at 1080p), over
`--batch` frames of distinct seeded cameras (seeds 1 ..).  Inputs are
parameters; the outputs (depth, iterations, normals) stay in device memory.

The warm start's network is trained first (untimed) on the plain marcher's
(ray, hit distance) pairs of 48 other seeded cameras at 240 x 135 (seeds
10001 .., about 1.5 M pairs), none of them a benchmark frame.

Legs (every leg marches the same frames):
- plain: one frame per call of mb3d_hip_march, the plain FP32 march and its
  f64 finish;
- plain + batched: the frames in one HIP-graph plan (one persistent launch
  of each kernel for all of them), autotuned;
- cone + batched: plus the cone ray per tile (the reference's cone march):
  a tile whose cone ray passes the end of the view is background, every
  other ray marches plainly -- the plain picture bit for bit (depths,
  iterations and normals; the background's fog counts are the cone's, as
  in the reference's cone march); autotuned;
- gated cone + batched: the network as the cone pass's gate -- it runs on
  each tile's middle ray first, and a tile it calls a hit (--gate) marches
  no cone ray (its rays march plainly, as a cone ray that meets the surface
  would leave them); hit jumps off, so the plain picture bit for bit (as
  the cone leg's: the background's fog counts from the cone where it
  decided), and a wrong call costs time, never a pixel; the cone leg's
  launch options;
- learned + batched: the network's hit jumps, each checked by the exact DE
  at its landing and rolled back to the plain march on failure;
- learned + cone + batched: both.
One step is one replay of the plan.

The value is the fastest leg whose picture meets the plain marcher's bar
against the f64 oracle: mask, depth within 1e-4 of the scene scale and
normals within 0.5 degrees, each on >= 99.9 % of the pixels, on three of
the frames.  A leg that adds the network counts when it beats the same leg
without it by more than 1 % (the run-to-run spread).  `config.path` names
the leg.  Without the oracle legs (--no-cpu) the candidates are the two
legs whose picture is the plain one by construction: cone + batched, and
gated cone + batched.  The learned legs are reported beside the value with
their accuracy: a jump moves the march's sample points, and on a fractal
the DE-stop decision depends on where they fall, so about 0.7 % of the
surface pixels move by more than 1e-4 (a median of 2 pixels of depth)
whatever the jump's margin.

Prints one JSON line: Mpixels/s of the legs, each path's accuracy against
the f64 oracle, the roofline figures of the value's kernels (the FP32 march
against the FP32 vector peak, the cone rays, finish and normals against the
FP64 vector peak) and of the network (against the BF16 MFMA peak, on the
learned + cone leg); algorithmic FLOPs per launch over HIP-event time on
the library's stream; the CPU baseline (the f64 oracle on the host cores,
and on one core), and the 64-frame animation of mixed sizes.  Beside the
metric: the painter, the post passes, the frame's reflection rays and the
Monte Carlo batches.
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
PEAK_BF16 = 2.5e15    # MI355X dense BF16 MFMA, FLOP/s (published)
PEAK_HBM = 8.0e12     # MI355X HBM3E, bytes/s (published)
# algorithmic FLOPs per unit of work (the reference's formulas, counted by op)
F_P8 = 60.0     # int_pow power 8 (src/formulas.rs:837-858) + the rout update
F_BOX = 31.0    # amazing_box (src/formulas.rs:707-736) + the rout update
F_STEP = 25.0   # one march step of march_from (src/calc.rs:2555-2640)
F_DE = 8.0      # DE tail: sqrt, log, divide, scale, clamp
# the network's forward pass per ray: 80 -> 64 -> 64 -> 64 -> 2, two FLOPs a multiply-add
F_MLP = 2.0 * (80 * 64 + 64 * 64 + 64 * 64 + 64 * 2)
TRAIN_SEEDS = tuple(range(10001, 10049))  # the network's 48 cameras: none is a benchmark frame
ACC_FRAMES = (0, 16, 32)  # the frames whose accuracy is measured against the oracle
TRAFFIC = os.path.join(ROOT, "profiles", "traffic_mandelbulb_p8_1080p.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed replays of each batched leg (plain: frames x steps)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=48, help="frames per launch of the batched paths (at most 64)")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=1, help="the first frame's camera seed")
    ap.add_argument("--train-steps", type=int, default=2000)
    ap.add_argument("--fixed", type=str, default="",
                    help="tile_w,tile_h,waves_per_cu,it_split,refill_min for every batched leg instead of the "
                         "autotuner's choice (profiler runs)")
    ap.add_argument("--gate", type=float, default=0.9,
                    help="the network's hit probability above which a tile skips its cone ray (the gated leg)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and accuracy legs")
    ap.add_argument("--no-anim", action="store_true", help="skip the 64-frame animation leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also copy a 512 MiB device buffer once (the HBM counter scale of a profile)")
    return ap.parse_args()


def flops_march(c: dict) -> float:
    """k_march32's: the FP32 iterations, DEs and steps (less the normals
    pass's FP32 stencil -- its_normals_f32, six DEs per hit -- and the f64
    finish's steps)."""
    its = c["its_p8_f32"] - c["its_normals_f32"]
    return (its * F_P8 + c["its_box_f32"] * F_BOX + (c["de_evals_f32"] - 6 * c["hits"]) * F_DE +
            (c["steps"] - c["steps_f64"]) * F_STEP)


def flops_cone(c: dict) -> float:
    """k_cone's: the cone rays' f64 iterations and DEs."""
    return c["its_cone_f64"] * F_P8 + c["de_cone_f64"] * F_DE


def flops_finish(c: dict) -> float:
    """k_finish64's: the f64 iterations, DEs and steps, less the cone rays'."""
    return ((c["its_p8_f64"] - c["its_cone_f64"]) * F_P8 + c["its_box_f64"] * F_BOX +
            (c["de_evals_f64"] - c["de_cone_f64"]) * F_DE + c["steps_f64"] * F_STEP)


def flops_normals(c: dict) -> float:
    """k_normals's FP32 stencil: six DEs per surface pixel (without a
    G-buffer output the centre DE is not made)."""
    return c["its_normals_f32"] * F_P8 + 6 * c["hits"] * F_DE


def with_opts(base, mlp=None, cone=False, gate=0.0):
    """A copy of `base` (tile, grid, split, refill) with the network and the
    cone rays set; `gate` > 0: the network as the cone pass's gate alone
    (hit jumps off)."""
    o = H.Options()
    H.C.pointer(o)[0] = base
    o.mlp = mlp.h if mlp is not None else None
    o.mlp_cone = 1 if cone else 0
    o.mlp_cone_gate = gate
    o.mlp_conf_hit = 2.0 if gate > 0.0 else base.mlp_conf_hit
    return o


def traffic_of(kernel: str, batch: int, width: int, height: int):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by a copy kernel's
    counts), when it was measured on this workload; None otherwise."""
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if (t.get("batch"), t.get("width"), t.get("height")) != (batch, width, height):
        return None
    k = t.get("kernels", {}).get(kernel)
    return float(k["fetch_bytes"] + k["write_bytes"]) if k else None


def note(what: str) -> None:
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"bench: {what}", file=sys.stderr, flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU (asked for {a.gpus}): the marcher shards by pixel bands or whole "
              "frames per GPU, each GPU marching its own share with no exchange between GPUs",
              file=sys.stderr)
        sys.exit(2)
    import torch

    lib = H.load()
    H.check(lib.mb3d_hip_init(0), "mb3d_hip_init")
    scenes = [H.scene_bulb(a.width, a.height, seed=a.seed + i) for i in range(a.batch)]
    params = [H.params_from_scene(s) for s in scenes]
    frames = H.params_array(scenes)
    n = a.width * a.height
    dev = torch.device("cuda:0")
    depth = torch.empty(n * a.batch, dtype=torch.float32, device=dev)
    iters = torch.empty(n * a.batch, dtype=torch.int16, device=dev)
    normals = torch.empty(n * a.batch * 3, dtype=torch.float32, device=dev)
    out = H.Out(depth.data_ptr(), None, iters.data_ptr(), normals.data_ptr(), None, 1, 0)

    if a.copy_calibration:
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    note("training the network")
    # --- the network (untimed): the plain marcher's pairs of other cameras
    t = time.perf_counter()
    tw, th = 240, a.height * 240 // a.width
    net, curve = H.train_warm_start([H.params_from_scene(H.scene_bulb(tw, th, seed=s)) for s in TRAIN_SEEDS],
                                    steps=a.train_steps, batch=8192, lr=3e-3, seed=0)
    train_s = time.perf_counter() - t

    note("the plain leg")
    # --- plain: one frame per call, every frame `steps` times
    plain_opts = H.options()

    def march_one(p, o):
        H.check(lib.mb3d_hip_march(H.C.byref(p), None, 0, a.height, H.C.byref(out), H.C.byref(o),
                                   H.PROGRESS_FN(0), H.CANCEL_FN(0), None), "mb3d_hip_march")

    for i in range(a.warmup):
        march_one(params[i % a.batch], plain_opts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        for p in params:
            march_one(p, plain_opts)
    torch.cuda.synchronize()
    plain_wall = time.perf_counter() - t0
    cp = H.counters()
    plain_value = n * a.steps * a.batch / plain_wall / 1e6

    note("the batched legs")
    # --- the batched legs: the autotuner's options for the plain batch, the
    # cone batch and the learned + cone batch (untimed); the learned-only leg
    # takes the latter's launch options
    if a.fixed:
        tw_, th_, wpc_, split_, refill_ = (int(v) for v in a.fixed.split(","))

        def tuned(o):
            o.tile_w, o.tile_h, o.waves_per_cu, o.it_split, o.refill_min = tw_, th_, wpc_, split_, refill_
            return o, 0.0
    else:
        def tuned(o):
            return H.autotune(frames, out, o)
    note("autotuning the plain batch")
    pb_opts, _ = tuned(H.options())
    note("autotuning the cone batch")
    c_opts, _ = tuned(H.options(cone=True))
    note("autotuning the learned + cone batch")
    lc_opts, _ = tuned(H.options(mlp=net, cone=True))
    legs = {"plain_batched": with_opts(pb_opts), "cone_batched": with_opts(c_opts, cone=True),
            "gated_cone_batched": with_opts(c_opts, mlp=net, cone=True, gate=a.gate),
            "learned_batched": with_opts(lc_opts, mlp=net), "learned_cone_batched": with_opts(lc_opts, mlp=net, cone=True)}
    res = {}
    for name, o in legs.items():
        note(f"the {name} leg")
        plan = H.Plan(frames, out, o)
        for _ in range(a.warmup):
            plan.launch(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            c = plan.launch(1)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        plan.close()
        res[name] = {"value": n * a.batch * a.steps / wall / 1e6, "unit": "Mpixels/s",
                     "ms_per_frame": wall * 1e3 / (a.steps * a.batch), "ms_per_step": wall * 1e3 / a.steps,
                     "steps": c["steps"] // a.batch, "its_p8_f32": c["its_p8_f32"] // a.batch,
                     "jumps": c["jumps"] // a.batch, "rollbacks": c["rollbacks"] // a.batch,
                     "cone_background": c["cone_background"] // a.batch,
                     "options": {"tile": [o.tile_w, o.tile_h], "waves_per_cu": o.waves_per_cu,
                                 "it_split": o.it_split, "refill_min": o.refill_min}}
    note("accuracy and the CPU baseline")
    # --- each path's accuracy against the f64 oracle (three frames), and
    # the oracle's own rate on the host cores; the value is the fastest leg
    # that meets the plain marcher's bar
    cpu_base = accuracy = None
    if not a.no_cpu:
        cpu_base, accuracy = cpu_leg(scenes, params, a, legs, net)
    at_bar = {"plain_batched": "plain", "cone_batched": "cone", "gated_cone_batched": "gated_cone",
              "learned_batched": "learned", "learned_cone_batched": "learned_cone"}
    if accuracy is not None:
        ok = [k for k in legs if accuracy[at_bar[k]]["meets_bar"]]
    else:
        # the plain picture by construction (tests/test_gpu_cone.py,
        # tests/test_gpu_learned.py)
        ok = ["cone_batched", "gated_cone_batched"]
    # a leg that adds the network counts only when it beats the same leg
    # without it by more than the run-to-run spread (1 %)
    without = {"gated_cone_batched": "cone_batched", "learned_cone_batched": "cone_batched",
               "learned_batched": "plain_batched"}
    ok = [k for k in ok if k not in without or res[k]["value"] > 1.01 * res[without[k]]["value"]] or ["cone_batched"]
    head = max(ok, key=lambda k: res[k]["value"])
    lead = res[head]
    value = lead["value"]

    note("kernel times")
    # --- the kernels of the value's leg: HIP-event times of the same batch
    # launched eagerly on the library's stream (a captured graph does not
    # time its nodes), averaged; the network's on the learned + cone leg
    def kernel_times(o):
        tms = {"mlp": [], "cone": [], "march": [], "finish": [], "normals": []}
        c = None
        for _ in range(max(3, a.steps)):
            H.check(lib.mb3d_hip_march_batch(frames, a.batch, H.C.byref(out), H.C.byref(o)), "march_batch")
            c = H.counters()
            for k, f in (("mlp", "mlp_ms"), ("cone", "cone_ms"), ("march", "march_f32_ms"),
                         ("finish", "finish_f64_ms"), ("normals", "normals_ms")):
                tms[k].append(c[f])
        t = {k: float(np.mean(v)) for k, v in tms.items()}
        t["finish"] -= t["normals"]  # (finish_f64_ms holds both passes)
        return t, c

    lo = legs[head]
    t_k, cf = kernel_times(lo)
    kernels = [
        # (with the cone rays the FP32 march is two kernels: k_cone_bg, the
        # background tiles' start DEs, then k_march32 on the other tiles)
        {"kernel": "k_cone_bg + k_march32 (batch)" if lo.mlp_cone else "k_march32 (batch)", "bound": "valu-fp32",
         "kernel_ms": t_k["march"],
         "flops_per_launch": flops_march(cf), "peak": PEAK_FP32 / 1e12},
        {"kernel": "k_finish64 (batch)", "bound": "valu-fp64", "kernel_ms": t_k["finish"],
         "flops_per_launch": flops_finish(cf), "peak": PEAK_FP64 / 1e12},
        {"kernel": "k_normals (batch)", "bound": "valu-fp32", "kernel_ms": t_k["normals"],
         "flops_per_launch": flops_normals(cf), "peak": PEAK_FP32 / 1e12},
    ]
    if lo.mlp_cone:
        kernels.append({"kernel": "k_cone (batch)", "bound": "valu-fp64", "kernel_ms": t_k["cone"],
                        "flops_per_launch": flops_cone(cf), "peak": PEAK_FP64 / 1e12, "rays": cf["cone_rays"]})
        if lo.mlp_cone_gate > 0.0:
            kernels[-1]["note"] = ("kernel_ms includes the gate's network pass over the tiles' middle rays "
                                   f"({cf['mlp_rays']} rays, k_mlp_propose ahead of k_cone)")
    ll = legs["learned_cone_batched"]
    t_l, cl = kernel_times(ll)
    items = sum(((s.width + ll.tile_w - 1) // ll.tile_w) * ll.tile_w * ((s.height + ll.tile_h - 1) // ll.tile_h) *
                ll.tile_h for s in scenes)
    net_kernel = {"kernel": "k_mlp_propose (batch)", "bound": "mfma-bf16", "kernel_ms": t_l["mlp"],
                  "flops_per_launch": F_MLP * cl["mlp_rays"], "peak": PEAK_BF16 / 1e12, "rays": cl["mlp_rays"],
                  "items": items, "leg": "learned_cone_batched"}
    for k in kernels + [net_kernel]:
        k["achieved"] = k["flops_per_launch"] / (k["kernel_ms"] * 1e-3) / 1e12
        k["frac"] = k["achieved"] / k["peak"]
    dom = max(kernels, key=lambda k: k["kernel_ms"])
    roofline = {
        "bound": dom["bound"],
        "kernel": dom["kernel"],
        "achieved": dom["achieved"],
        "peak": dom["peak"],
        "unit": "TFLOP/s",
        "frac": dom["frac"],
        "traffic": traffic_of(dom["kernel"], a.batch, a.width, a.height),
        "kernel_ms": dom["kernel_ms"],
        "flops_per_launch": dom["flops_per_launch"],
        "leg": head,
        "kernels": kernels,
        "network": net_kernel,
    }
    for k in kernels + [net_kernel]:
        k["traffic"] = traffic_of(k["kernel"], a.batch, a.width, a.height)

    note("the painter")
    # --- the painter (beside the metric): the first frame's G-buffer,
    # marched once into device memory, lit and coloured on the device by
    # k_paint (the reference's default lighting); its HIP-event time
    sc, p = scenes[0], params[0]
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

    note("the post passes")
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

    note("the secondary rays")
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

    note("the Monte Carlo batches")
    # --- the Monte Carlo renderer's marches (beside the metric): every pixel's
    # primary ray with the inside probe, then a shadow ray towards one light
    # from each surface it met (tasks and results in device memory); the
    # k_mc batches' HIP-event time
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
    del depth, iters, normals, gbuf, rgb, d_tasks, d_hits

    note("the animation")
    # --- the batched animation (the mixed-size configuration, beside the
    # metric): 64 frames of {320x240 .. 1920x1080} along a seeded camera path
    # around the bulb, in one plan, with the batch legs' launch options
    anim = None
    if not a.no_anim:
        anim = animation_leg(a, lib, pb_opts, legs[head], head)

    result = {
        "metric": "Mpixels/s 1080p Mandelbulb-p8 Z-buffer; learned vs plain at equal depth error",
        "value": value,
        "unit": "Mpixels/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": lead["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic (seeded cameras; network random-initialised, trained on the plain marcher's pairs of 48 "
                "other cameras)",
        "config": {
            "workload": f"mandelbulb-p8 z-buffer {a.width}x{a.height}, analytic DE, 60 its, bailout r2=16, fov 30; "
                        f"{a.batch} frames of distinct seeded cameras (seeds {a.seed}..{a.seed + a.batch - 1}) per "
                        "graph launch; the fastest leg at the plain marcher's bar against the f64 oracle",
            "path": head.replace("_", "+"),
            "batch": a.batch,
            "train_cameras": list(TRAIN_SEEDS),
        },
        "plain": {"value": plain_value, "unit": "Mpixels/s", "ms_per_frame": plain_wall * 1e3 / (a.steps * a.batch),
                  "kernel_ms_per_frame": cp["kernel_ms"], "march_f32_ms": cp["march_f32_ms"],
                  "finish_f64_ms": cp["finish_f64_ms"], "steps": cp["steps"], "its_p8_f32": cp["its_p8_f32"]},
        **res,
        "speedup_vs_plain": value / plain_value,
        "speedup_vs_plain_batched": value / res["plain_batched"]["value"],
        "learned": {
            leg: {
                "value": res[leg]["value"],
                "speedup_vs_plain": res[leg]["value"] / plain_value,
                "speedup_vs_plain_batched": res[leg]["value"] / res["plain_batched"]["value"],
                "speedup_of_the_network": res[leg]["value"] / res["cone_batched"]["value"],
                "meets_bar": None if accuracy is None else accuracy[at_bar[leg]]["meets_bar"],
                "role": role,
            }
            for leg, role in (("gated_cone_batched", f"the cone pass's gate (p >= {a.gate}), hit jumps off"),
                              ("learned_cone_batched", "hit jumps, checked at the landing"))
        },
        "network": {"train_s": train_s, "train_pairs": int(len(TRAIN_SEEDS) * tw * th),
                    "train_loss": [float(curve[:20].mean()), float(curve[-100:].mean())]},
        "roofline": roofline,
        "paint": paint,
        "post": post_leg,
        "rays": rays,
        "mc": mc,
    }
    if anim is not None:
        result["animation"] = anim
    if not a.no_cpu:
        result["cpu_baseline"], result["accuracy"] = cpu_base, accuracy
        cores = result["cpu_baseline"]["cores"]
        paint["cpu"] = paint_cpu_leg(cam, lv, paint_gbuf, cores)
        rays["cpu"] = rays_cpu_leg(p, tasks, ray_opts, cores)
        mc["cpu"] = mc_cpu_leg(p, mc_prim, mc_shad, cores)
        result["at_equal_depth_error"] = {k: v["meets_bar"] for k, v in result["accuracy"].items()}
    net.close()
    print(json.dumps(result))


def animation_leg(a, lib, pb_opts, head_opts, head):
    """The 64-frame animation of mixed sizes in one plan: plain + batched and
    the value's leg, Mpixels/s of a replay."""
    import torch

    scenes = H.animation(64, seed=a.seed)
    frames = H.params_array(scenes)
    n = sum(s.width * s.height for s in scenes)
    dev = torch.device("cuda:0")
    depth = torch.empty(n, dtype=torch.float32, device=dev)
    iters = torch.empty(n, dtype=torch.int16, device=dev)
    normals = torch.empty(n * 3, dtype=torch.float32, device=dev)
    out = H.Out(depth.data_ptr(), None, iters.data_ptr(), normals.data_ptr(), None, 1, 0)
    res = {"frames": 64, "pixels": n, "sizes": sorted({(s.width, s.height) for s in scenes})}
    for name, o in (("plain_batched", pb_opts), (head, head_opts)):
        plan = H.Plan(frames, out, o)
        plan.launch(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = max(2, a.steps // 3)
        for _ in range(reps):
            plan.launch(1)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / reps
        plan.close()
        res[name] = {"value": n / wall / 1e6, "unit": "Mpixels/s", "ms_per_launch": wall * 1e3}
    res["leg"] = head
    res["speedup"] = res[head]["value"] / res["plain_batched"]["value"]
    return res


def cpu_leg(scenes, params, a, legs, net):
    """The f64 oracle (a restatement of the reference's CPU marcher) on the
    host cores, on a frame of the benchmark, and on one core on a band of
    its rows; each path's accuracy against it on three frames."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    import parity

    host_threads = os.cpu_count() or 1
    share = len(os.sched_getaffinity(0))
    # a host's CPU share for one GPU: at most 16 cores
    threads = a.cpu_threads or min(share, 16)
    keys = ("mask_agree", "depth_rel_p99", "depth_rel_max", "depth_ok_frac", "normal_ok_frac", "normal_deg_p99",
            "iters_equal_inner")
    paths = {"plain": H.options(), "cone": H.options(cone=True),
             "gated_cone": H.options(mlp=net, cone=True, cone_gate=a.gate, conf_hit=2.0),
             "learned": H.options(mlp=net), "learned_cone": H.options(mlp=net, cone=True)}
    of_leg = {"plain": "plain_batched", "cone": "cone_batched", "gated_cone": "gated_cone_batched",
              "learned": "learned_batched", "learned_cone": "learned_cone_batched"}
    for k, o in paths.items():
        o.tile_w, o.tile_h = legs[of_leg[k]].tile_w, legs[of_leg[k]].tile_h
    acc = {k: {m: [] for m in keys} for k in paths}
    t_full = []
    for f in ACC_FRAMES:
        if f >= len(scenes):
            continue
        t = time.perf_counter()
        ref = O.render(O.params(scenes[f]), threads=threads)
        t_full.append(time.perf_counter() - t)
        p = params[f]
        for name, o in paths.items():
            m = parity.compare(H.march(p, opts=o), ref, p.width, p.height, p.step_width, 2.0)
            for k in keys:
                acc[name][k].append(m[k])
    worst = {"mask_agree": min, "depth_rel_p99": max, "depth_rel_max": max, "depth_ok_frac": min,
             "normal_ok_frac": min, "normal_deg_p99": max, "iters_equal_inner": min}
    out = {}
    for name in paths:
        r = {k: worst[k](v) for k, v in acc[name].items()}
        r["frames"] = [f for f in ACC_FRAMES if f < len(scenes)]
        r["meets_bar"] = bool(r["mask_agree"] >= 0.999 and r["depth_ok_frac"] >= 0.999 and r["normal_ok_frac"] >= 0.999)
        out[name] = r
    # one core: a band of rows of the first frame
    p0 = O.params(scenes[0])
    rows = max(1, a.height // 16)
    r0 = (a.height - rows) // 2
    t = time.perf_counter()
    O.render(p0, row0=r0, rows=rows, threads=1)
    dt1 = time.perf_counter() - t
    dt = float(np.min(t_full))
    cpu = {
        "value": a.width * a.height / dt / 1e6,
        "unit": "Mpixels/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full {a.width}x{a.height} frames of the workload ({len(t_full)} frames, the fastest "
                  f"{dt:.2f} s), {threads} threads, rows interleaved as src/render.rs:757-803",
        "host_threads": host_threads,
        "affinity_threads": share,
        "one_core": {"value": a.width * rows / dt1 / 1e6, "unit": "Mpixels/s", "cores": 1,
                     "sample": f"rows {r0}..{r0 + rows - 1} of the first frame, {dt1:.2f} s"},
    }
    return cpu, out


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

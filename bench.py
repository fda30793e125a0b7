#!/usr/bin/env python3
"""Benchmark of the HIP DE ray marcher on one MI355X.

Workload: the power-8 Mandelbulb Z-buffer at 1920 x 1080 (the small
configuration's scene -- 60 iterations, bailout r^2 = 16, analytic DE
0.5 r ln r / dr, camera start plane 2.5 from the origin, FOV 30, seeded yaw
and pitch, 6 bisection steps, 500-step cap -- rendered at 1080p).  One step
is one full 1080p frame marched through the C ABI into device buffers (the
inputs are parameters; the outputs stay in HBM).

Prints one JSON line: Mpixels/s, the roofline figure of the dominant kernel
(the FP32 march, its algorithmic FP32 FLOPs over its HIP-event time), the CPU
baseline (the f64 oracle on the host cores) and the accuracy of the frame
against the oracle.
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

PEAK_FP32 = 157.3e12  # MI355X dense FP32 vector, FLOP/s
PEAK_FP64 = 78.6e12
# algorithmic FLOPs per unit of work (the reference's formulas, counted by op)
F_P8 = 60.0     # int_pow power 8 (src/formulas.rs:837-858) + the rout update
F_BOX = 31.0    # amazing_box (src/formulas.rs:707-736) + the rout update
F_STEP = 25.0   # one march step of march_from (src/calc.rs:2555-2640)
F_DE = 8.0      # DE tail: sqrt, log, divide, scale, clamp


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--de", choices=["analytic", "numeric"], default="analytic")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also copy a 512 MiB device buffer once (the HBM counter scale of a profile)")
    return ap.parse_args()


def flops_f32(c: dict) -> float:
    return c["its_p8_f32"] * F_P8 + c["its_box_f32"] * F_BOX + c["de_evals_f32"] * F_DE + c["steps"] * F_STEP


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}): every box of this build has one GPU, "
              "and the marcher shards by pixel bands per GPU (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch

    lib = H.load()
    H.check(lib.mb3d_hip_init(0), "mb3d_hip_init")
    de_mode = H.DE_ANALYTIC_P8 if a.de == "analytic" else H.DE_REFERENCE
    sc = H.scene_bulb(a.width, a.height, seed=a.seed, de_mode=de_mode)
    p = H.params_from_scene(sc)
    n = a.width * a.height
    dev = torch.device("cuda:0")
    depth = torch.empty(n, dtype=torch.float32, device=dev)
    iters = torch.empty(n, dtype=torch.int16, device=dev)
    normals = torch.empty(n * 3, dtype=torch.float32, device=dev)
    out = H.Out(depth.data_ptr(), None, iters.data_ptr(), normals.data_ptr(), None, 1, 0)
    opts = H.options()

    def step():
        st = lib.mb3d_hip_march(H.C.byref(p), None, 0, a.height, H.C.byref(out), H.C.byref(opts),
                                H.PROGRESS_FN(0), H.CANCEL_FN(0), None)
        H.check(st, "mb3d_hip_march")
        return H.counters()

    if a.copy_calibration:
        src = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    k_ms, f32_ms, cs = [], [], []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        c = step()
        k_ms.append(c["kernel_ms"])
        f32_ms.append(c["march_f32_ms"])
        cs.append(c)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / a.steps
    value = n * a.steps / wall / 1e6
    c = cs[-1]
    f32_avg = float(np.mean(f32_ms))
    achieved = flops_f32(c) / (f32_avg * 1e-3)
    roofline = {
        "bound": "valu-fp32",
        "kernel": "k_march32",
        "achieved": achieved / 1e12,
        "peak": PEAK_FP32 / 1e12,
        "unit": "TFLOP/s",
        "frac": achieved / PEAK_FP32,
        "traffic": None,
        "kernel_ms": f32_avg,
        "flops_per_launch": flops_f32(c),
    }
    result = {
        "metric": "Mpixels/s 1080p Mandelbulb-p8 Z-buffer; learned vs plain at equal depth error",
        "value": value,
        "unit": "Mpixels/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 iterate, fp64 ray/stop/bisection/normals",
        "data": "synthetic (seeded camera)",
        "config": {
            "workload": f"mandelbulb-p8 z-buffer {a.width}x{a.height}, {a.de} DE, 60 its, bailout r2=16, "
                        f"fov 30, seed {a.seed}, plain FP32 marcher",
            "path": "plain",
        },
        "kernel_ms_per_frame": float(np.mean(k_ms)),
        "counters": {k: c[k] for k in ("hits", "steps", "de_evals_f32", "de_evals_f64", "its_p8_f32",
                                        "its_p8_f64", "overflow")},
        "roofline": roofline,
    }
    if not a.no_cpu:
        result["cpu_baseline"], result["accuracy"] = cpu_leg(sc, p, a)
    print(json.dumps(result))


def cpu_leg(sc, p, a):
    """The f64 oracle (a restatement of the reference's CPU marcher) on the
    host cores, on the same frame, and the frame's accuracy against it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle as O
    import parity

    threads = a.cpu_threads or os.cpu_count() or 1
    po = O.params(sc)
    t = time.perf_counter()
    ref = O.render(po, threads=threads)
    dt = time.perf_counter() - t
    g = H.march(p)
    acc = parity.compare(g, ref, p.width, p.height, p.step_width, 2.0)
    cpu = {
        "value": p.width * p.height / dt / 1e6,
        "unit": "Mpixels/s",
        "cores": threads,
        "kind": "port",
        "sample": f"one full {p.width}x{p.height} frame of the same workload, {dt:.2f} s, "
                  f"{threads} threads, rows interleaved as src/render.rs:757-803",
    }
    return cpu, acc


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/: small pictures of every
workload rendered by the oracle (the f64 C++ restatement of the reference's
CPU marcher), after its own tests pass (tests/test_oracle.py).

Each fixture is one .npz: the march parameters and post job as the C ABI
takes them (their fields by name, as JSON), and the oracle's outputs -- depth (f64 march
units, +inf background), iterations, normals, the 5-word G-buffer records.
tests/test_golden.py checks that the oracle still reproduces them bit for
bit and that the HIP backend meets the parity bar against them.

    python tests/golden/make_golden.py
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("mb3d-rs_amd", "oracle"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import json  # noqa: E402

import numpy as np  # noqa: E402

import mb3d_hip as H  # noqa: E402
import oracle as O  # noqa: E402


def cases():
    """name -> (scene, post job builder or None, scene scale)."""
    return {
        "bulb_analytic_96x72_s1": (H.scene_bulb(96, 72, seed=1), None, 2.0),
        "bulb_analytic_96x72_s2_shadow": (H.scene_bulb(96, 72, seed=2),
                                          lambda p: H.post_job(lights=[(0.6, -0.5, -0.6)], params=p), 2.0),
        "bulb_numeric_96x72_s3": (H.scene_bulb(96, 72, seed=3, de_mode=H.DE_REFERENCE), None, 2.0),
        "box_64x48_s1_ao8": (H.scene_box(64, 48, seed=1), lambda p: H.post_job(deao_quality=4, params=p), 4.0),
        "hybrid_64x48_s1": (H.scene_hybrid(64, 48, seed=1), None, 2.0),
    }


def main():
    for name, (sc, post_fn, scale) in cases().items():
        p = O.params(sc)
        job = post_fn(p) if post_fn else None
        r = O.render(p, job, threads=0)
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            params_json=np.array(json.dumps(H.struct_to_dict(p))),
            post_json=np.array(json.dumps(H.struct_to_dict(job) if job is not None else None)),
            scale=np.float64(scale),
            depth=r["depth"], iters=r["iters"], normals=r["normals"], gbuf=r["gbuf"])
        hit = np.isfinite(r["depth"]) & (r["depth"] > 0)
        print(f"{name}: {int(hit.sum())} surface pixels of {hit.size}")


if __name__ == "__main__":
    main()

"""The golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle): the oracle must still
reproduce them bit for bit (CPU), and the HIP backend must meet the parity
bar against them (GPU) -- the same pictures, without running the oracle."""
import glob
import json
import os

import numpy as np
import pytest

import mb3d_hip as H
import parity

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
NAMES = [os.path.basename(f)[:-4] for f in FIXTURES]


def load(name):
    z = np.load(os.path.join(HERE, "golden", name + ".npz"))
    p = H.struct_from_dict(H.Params, json.loads(str(z["params_json"])))
    # the fixtures carry the ABI version they were written with; versions 2
    # and 3 share the parameter record (3 grew mb3d_hip_options only)
    assert p.abi_version in (2, H.ABI_VERSION), p.abi_version
    p.abi_version = H.ABI_VERSION
    post = json.loads(str(z["post_json"]))
    # the post job's last word, reserved (0) when the fixtures were written,
    # is now the DE AO's dithering (0: off)
    if post and "reserved" in post:
        assert post.pop("reserved") == 0
    job = H.struct_from_dict(H.PostJob, post) if post else None
    return p, job, float(z["scale"]), {k: z[k] for k in ("depth", "iters", "normals", "gbuf")}


def test_every_workload_has_a_fixture():
    assert len(FIXTURES) >= 5, FIXTURES


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_the_fixture(name):
    import oracle as O

    p, job, _, want = load(name)
    got = O.render(p, job, threads=0)
    for k in ("iters", "gbuf"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(got["depth"], want["depth"])
    assert np.array_equal(got["normals"], want["normals"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_hip_meets_the_bar_on_the_fixture(hip, name):
    p, job, scale, want = load(name)
    g = hip.march(p, post=job)
    m = parity.compare(g, want, p.width, p.height, p.step_width, scale)
    assert m["mask_agree"] >= 0.999, m
    assert m["depth_ok_frac"] >= 0.999 and m["depth_rel_p99"] <= 1e-4, m
    assert m["normal_ok_frac"] >= 0.99 and m["normal_deg_p99"] < 0.5, m
    assert m["iters_equal_inner"] >= 0.99, m
    surf = np.isfinite(want["depth"]) & (want["depth"] > 0) & np.isfinite(g["depth64"]) & (g["depth64"] > 0)
    if job is not None and job.n_lights:
        sg = (g["gbuf"][surf, 1] >> 16) & 0xFC00
        so = (want["gbuf"][surf, 1] >> 16) & 0xFC00
        assert (sg == so).mean() >= 0.99
    if job is not None and job.deao:
        ag = (g["gbuf"][surf, 4] & 0xFFFF).astype(np.float64) / 16383.0
        ao = (want["gbuf"][surf, 4] & 0xFFFF).astype(np.float64) / 16383.0
        assert np.abs(ag - ao).mean() < 0.01

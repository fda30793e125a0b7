"""The backend's `gpu` interface in C++ (mb3d-rs_amd/host/gpu.hpp, the
reference's crate::gpu: march, march_bands, post_on_card, unsupported, the
switch and the status texts), driven by a C++ caller (tests/cpp)."""
import os
import subprocess

import numpy as np
import pytest

import mb3d_hip as H

HERE = os.path.dirname(os.path.abspath(__file__))
CHECK = os.path.join(HERE, "cpp", "gpu_api_check")


def run(*args):
    assert os.path.exists(CHECK), "tests/cpp/gpu_api_check is not built: run __graft_entry__.build()"
    r = subprocess.run([CHECK, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr


def test_cpu_side_of_the_gpu_interface():
    """Record packing, the switch (no device is touched when off), status
    texts and unsupported scenes."""
    run("cpu")


@pytest.mark.gpu
def test_march_bands_and_post_on_card_agree_with_the_c_abi(hip, tmp_path):
    p = hip.params_from_scene(H.scene_bulb(160, 120, seed=2))
    job = H.post_job(lights=[(0.6, -0.5, -0.6)], params=p)
    (tmp_path / "p.bin").write_bytes(bytes(p))
    (tmp_path / "j.bin").write_bytes(bytes(job))
    run("gpu", str(tmp_path / "p.bin"), str(tmp_path / "j.bin"), str(tmp_path))
    read = lambda n: np.fromfile(tmp_path / n, np.uint32).reshape(-1, 5)
    direct = read("direct.bin")
    assert np.array_equal(read("march.bin"), direct)
    assert np.array_equal(read("bands.bin"), direct)
    # the post passes on a G-buffer marched without them give the march's
    # own shadow bits
    post = read("post.bin")
    surf = (direct[:, 2] >> 16) < 32768
    assert (((post[surf, 1] >> 16) & 0xFC00) == ((direct[surf, 1] >> 16) & 0xFC00)).mean() >= 0.995
    ref = hip.march(p, post=job)["gbuf"]
    assert np.array_equal(ref, direct)

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


def _mixed(hip, tmp_path, threads, delay_us, stop_after=0):
    import oracle as O

    sc = H.scene_preset("Integer Power", 120, 90)
    p = hip.params_from_scene(sc)
    (tmp_path / "p.bin").write_bytes(bytes(p))
    cpu = O.render(O.params(sc), None, threads=0)["gbuf"]
    cpu.astype(np.uint32).tofile(tmp_path / "cpu.bin")
    r = subprocess.run([CHECK, "mixed", str(tmp_path / "p.bin"), str(tmp_path / "cpu.bin"), str(tmp_path),
                        str(threads), str(delay_us), str(stop_after)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
    info = r.stdout.split()
    card_rows, cpu_rows = int(info[1]), int(info[3])
    got = np.fromfile(tmp_path / "mixed.bin", np.uint32).reshape(p.height, p.width, 5)
    sinks = np.fromfile(tmp_path / "sinks.bin", np.int32)
    card = hip.march(p)["gbuf"].reshape(p.height, p.width, 5)
    return p, got, sinks, card, cpu.reshape(p.height, p.width, 5), card_rows, cpu_rows, r.stdout


@pytest.mark.gpu
def test_march_mixed_shares_the_rows(hip, tmp_path):
    """gpu::march_mixed (render::march_mixed, src/render.rs:1500-1700): four
    CPU threads (the oracle's rows, 2 ms each) and the card split the
    picture; every row but the checkpoint's is delivered once, by one side,
    and is that side's row."""
    p, got, sinks, card, cpu, card_rows, cpu_rows, out = _mixed(hip, tmp_path, 4, 2000)
    ck = np.arange(p.height) % 7 == 3
    assert np.all(sinks[ck] == 0) and np.all(sinks[~ck] == 1), out
    assert np.all(got[ck, :, 2] == 0x5A5A5A5A)  # the checkpoint's rows are kept
    from_card = np.all(got == card, axis=(1, 2)) & ~ck
    from_cpu = np.all(got == cpu, axis=(1, 2)) & ~ck
    assert np.all(from_card | from_cpu | ck), out
    assert card_rows > 0 and cpu_rows > 0 and card_rows + cpu_rows == int((~ck).sum()), out


@pytest.mark.gpu
def test_march_mixed_stops_when_asked(hip, tmp_path):
    p, got, sinks, card, cpu, card_rows, cpu_rows, out = _mixed(hip, tmp_path, 2, 3000, stop_after=20)
    assert "cancelled" in out and sinks.sum() < p.height - 12, out
    assert np.all(sinks <= 1)


def _mc_share(tmp_path, p, threads, delay_us, chunk, env=None):
    import oracle as O

    tasks = H.mc_primary_tasks(p, seed=3) if env is None else _cpu_tasks(p)
    (tmp_path / "p.bin").write_bytes(bytes(p))
    tasks.tofile(tmp_path / "tasks.bin")
    want = O.mc_batch(p, tasks)
    want.tofile(tmp_path / "want.bin")
    r = subprocess.run([CHECK, "mcshare", str(tmp_path / "p.bin"), str(tmp_path / "tasks.bin"),
                        str(tmp_path / "want.bin"), str(tmp_path), str(threads), str(delay_us), str(chunk)],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
    info = r.stdout.split()
    got = np.fromfile(tmp_path / "mc_res.bin", H.MC_RES_DTYPE)
    side = np.fromfile(tmp_path / "mc_side.bin", np.uint8)
    return tasks, want, got, side, int(info[1]), int(info[3]), r.stdout


def _cpu_tasks(p):
    """View rays of the picture from its start plane, built on the host
    (no device call)."""
    from test_oracle_mc import view_tasks

    t = view_tasks(p)
    t["flags"] = H.MC_PROBE
    return t


def test_mc_batch_shared_without_the_card(tmp_path):
    """gpu::mc_batch_shared with the card switched off (MB3D_GPU=0): the
    CPU threads march every task once, from the end."""
    import os

    p = H.params_from_scene(H.scene_preset("Integer Power", 40, 30))
    env = dict(os.environ, MB3D_GPU="0")
    tasks, want, got, side, card, cpu, out = _mc_share(tmp_path, p, 3, 0, 256, env=env)
    assert card == 0 and cpu == len(tasks) and np.all(side == 2), out
    assert "switched off" in out
    assert got.tobytes() == want.tobytes()


@pytest.mark.gpu
def test_mc_batch_shared_with_the_card(hip, tmp_path):
    """gpu::mc_batch_shared (gpu::mc_batch with RaySharing,
    src/gpu.rs:3013-3190): the card from the front in chunks of 1024, four
    CPU threads (the oracle's results, 3 ms a chunk of 64) from the end;
    every task once, by one side, equal to that side's result."""
    p = hip.params_from_scene(H.scene_bulb(160, 120, seed=1))
    tasks, want, got, side, card, cpu, out = _mc_share(tmp_path, p, 4, 3000, 1024)
    assert card > 0 and cpu > 0 and card + cpu == len(tasks), out
    assert np.all((side == 1) | (side == 2)) and (side == 1).sum() == card, out
    # the card's tasks are a prefix, the CPU's a suffix
    first_cpu = int(np.argmax(side == 2))
    assert np.all(side[:first_cpu] == 1) and np.all(side[first_cpu:] == 2), out
    dev = hip.mc_batch(p, tasks)
    assert got[side == 1].tobytes() == dev[side == 1].tobytes()
    assert got[side == 2].tobytes() == want[side == 2].tobytes()

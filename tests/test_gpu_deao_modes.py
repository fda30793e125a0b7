"""The DE AO's dithering and random first step (deao_row, src/deao.rs:
262-351) on the GPU against the oracle, and the post passes at the
workloads' own sizes.

On the oracle's G-buffer with the f64 iterate the backend's AO pass sees
the records the oracle saw: its amb words are the oracle's, the random
first steps drawn in the CPU's order (each surface pixel of a row takes one
number per ray from the row's generator, left to right; the backend counts
the surface pixels left of each by a scan of the row)."""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O

pytestmark = pytest.mark.gpu


def _amb(g):
    return (g[:, 4] & 0xFFFF).astype(np.int64)


def _object(g):
    return ((g[:, 2] >> 16) < 32768) & ((g[:, 3] & 0xFFFF) < 32768)


@pytest.mark.parametrize("quality,dither,rnd", [(1, 1, False), (2, 2, False), (1, 0, True), (2, 1, True),
                                               (0, 2, True), (3, 1, True)])
def test_dither_and_random_first_step_equal_the_oracle(hip, quality, dither, rnd):
    sc = H.scene_preset("Integer Power", 160, 120)
    sc.first_step_random = 0
    p = hip.params_from_scene(sc)
    po = O.params(sc)
    r = O.render(po, None, threads=0)
    job = H.post_job(deao_quality=quality, deao_dither=dither, deao_first_step_random=rnd)
    want = O.post(po, job, r["gbuf"])
    got = hip.post(p, job, r["gbuf"], opts=H.options(precision=H.PREC_F64))
    obj = _object(r["gbuf"])
    assert obj.sum() > 2000
    a, b = _amb(got)[obj], _amb(want)[obj]
    assert (a == b).mean() >= 0.999, (a != b).sum()
    assert np.abs(a - b).max() <= 2
    assert np.array_equal(got[:, :4], want[:, :4])


def test_the_random_first_step_is_the_rows_not_the_pixels(hip):
    """The same picture in two bands: the rows' generators start at their
    own rows, so the bands give the whole picture's words."""
    sc = H.scene_preset("Integer Power", 160, 120)
    sc.first_step_random = 0
    p = hip.params_from_scene(sc)
    r = O.render(O.params(sc), None, threads=0)
    job = H.post_job(deao_quality=2, deao_dither=1, deao_first_step_random=True)
    o = H.options(precision=H.PREC_F64)
    whole = hip.post(p, job, r["gbuf"], opts=o)
    top = hip.post(p, job, r["gbuf"][:60 * 160], row0=0, rows=60, opts=o)
    bottom = hip.post(p, job, r["gbuf"][60 * 160:], row0=60, rows=60, opts=o)
    assert np.array_equal(np.concatenate([top, bottom]), whole)


def test_march_with_dithered_random_ao_meets_the_card_bar(hip):
    """The FP32 march and its post pass, with dithering and the random first
    step, against the oracle's picture: the reference's card bar for the AO
    (mean |dAO| < 0.01, under 1 % of the pixels off by more than 0.1)."""
    sc = H.scene_bulb(320, 240, seed=2)
    p = hip.params_from_scene(sc)
    job = H.post_job(deao_quality=2, deao_dither=2, deao_first_step_random=True)
    g = hip.march(p, post=job)
    r = O.render(O.params(sc), job, threads=0)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0) & np.isfinite(g["depth64"]) & (g["depth64"] > 0)
    ag = _amb(g["gbuf"])[surf] / 16383.0
    ao = _amb(r["gbuf"])[surf] / 16383.0
    assert np.abs(ag - ao).mean() < 0.01
    assert (np.abs(ag - ao) > 0.1).mean() < 0.01


def test_bad_dithering_is_refused(hip):
    sc = H.scene_bulb(64, 48, seed=1)
    p = hip.params_from_scene(sc)
    with pytest.raises(RuntimeError):
        hip.march(p, post=H.post_job(deao_quality=1, deao_dither=3))
    with pytest.raises(RuntimeError, match="8-ray"):
        hip.march(p, post=H.post_job(deao_quality=4, deao_dither=1))


def test_shadows_at_the_plain_workloads_size(hip):
    """The hard shadow of the 640 x 480 power-8 bulb (one directional light)
    against the oracle on the whole picture."""
    sc = H.scene_bulb(640, 480, seed=1)
    p = hip.params_from_scene(sc)
    job = H.post_job(lights=[(0.6, -0.5, -0.6)], params=p)
    g = hip.march(p, post=job)
    r = O.render(O.params(sc), job, threads=0)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0) & np.isfinite(g["depth64"]) & (g["depth64"] > 0)
    sg = (g["gbuf"][surf, 1] >> 16) & 0x400
    so = (r["gbuf"][surf, 1] >> 16) & 0x400
    assert (so != 0).mean() > 0.05
    assert (sg == so).mean() >= 0.995


def test_box_band_with_16x16_tiles_and_8_ray_ao(hip):
    """The Amazing Box workload's launch: 1920-wide rows in 16 x 16 tiles of
    the persistent queue, 8-ray AO per hit; the tiles change nothing (the
    8 x 8 picture's words), and the AO meets the card bar against the
    oracle."""
    sc = H.scene_box(1920, 1080, seed=1)
    p = hip.params_from_scene(sc)
    job = H.post_job(deao_quality=4)
    g16 = hip.march(p, post=job, row0=480, rows=96, opts=H.options(tile=(16, 16)))
    g8 = hip.march(p, post=job, row0=480, rows=96)
    assert np.array_equal(g16["gbuf"], g8["gbuf"]) and np.array_equal(g16["depth"], g8["depth"])
    r = O.render(O.params(sc), job, row0=480, rows=96, threads=0)
    surf = np.isfinite(r["depth"]) & (r["depth"] > 0) & np.isfinite(g16["depth64"]) & (g16["depth64"] > 0)
    assert surf.mean() > 0.2
    ag = _amb(g16["gbuf"])[surf] / 16383.0
    ao = _amb(r["gbuf"])[surf] / 16383.0
    assert np.abs(ag - ao).mean() < 0.01
    assert (np.abs(ag - ao) > 0.1).mean() < 0.01

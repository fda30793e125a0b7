"""The oracle's cone march (test infrastructure), pinned by the reference's
own cone tests (src/cone.rs:176-220) re-expressed for this backend's one
level of blocks, the work queue's tiles (CPU only):
  - a block whose coarse ray passed the end of the view holds background
    pixels only, in the plain march (every pixel);
  - with the cone's background verdict alone (mode 1) the picture is the
    plain march's, bit for bit (depths, iteration counts, normals);
  - with the march from the block's depth (mode 2) few jumps pass the
    pixel's surface (< 1/200) and few pixels change between background and
    surface (< 1/100);
  - the dynamic fog's count comes with a background block's depth.
"""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O


def scenes():
    box = H.scene_box(96, 64, seed=3)
    box.step_cap = 500
    return {
        "bulb": H.scene_bulb(96, 64, seed=4),
        "bulb_numeric": H.scene_bulb(96, 64, seed=5, de_mode=H.DE_REFERENCE),
        "box": box,
        "hybrid": H.scene_hybrid(96, 64, seed=2),
    }


def blocks_to_pixels(g, bw, bh, w, h):
    return np.repeat(np.repeat(g, bh, 0), bw, 1)[:h, :w].ravel()


@pytest.mark.parametrize("name", ["bulb", "bulb_numeric", "box", "hybrid"])
@pytest.mark.parametrize("block", [(8, 8), (16, 4)])
def test_background_blocks_hold_background_only(name, block):
    sc = scenes()[name]
    p = O.params(sc)
    bw, bh = block
    t, _ = O.cone_grid(p, bw, bh)
    plain = O.render(p, threads=0)
    bg_px = blocks_to_pixels(np.isinf(t), bw, bh, p.width, p.height)
    assert not np.any(bg_px & np.isfinite(plain["depth"])), name
    if name != "box" and block == (8, 8):  # (the box camera sits inside the object's hull; at 96 x 64 a
        assert bg_px.sum() > 100, bg_px.sum()  # 16-pixel-wide cone spans a quarter radian)
    # mode 1: the plain picture, bit for bit
    one = O.render(p, threads=0, cone=(bw, bh, 1))
    for k in ("depth", "iters", "normals"):
        assert np.array_equal(one[k], plain[k]), k
    surf = np.isfinite(plain["depth"])
    assert np.array_equal(one["gbuf"][surf], plain["gbuf"][surf])
    assert one["counters"]["steps"] <= plain["counters"]["steps"]


@pytest.mark.parametrize("name", ["bulb", "bulb_numeric", "hybrid"])
def test_block_depths_lie_before_the_pixels_surfaces(name):
    sc = scenes()[name]
    p = O.params(sc)
    t, _ = O.cone_grid(p, 8, 8)
    tp = blocks_to_pixels(t, 8, 8, p.width, p.height)
    plain = O.render(p, threads=0)
    jumps = np.isfinite(tp) & (tp > 0)
    surf = np.isfinite(plain["depth"]) & (plain["depth"] > 0)
    behind = jumps & surf & (plain["depth"] < tp)
    assert jumps.sum() > 1000 and behind.sum() * 200 < jumps.sum(), (behind.sum(), jumps.sum())
    two = O.render(p, threads=0, cone=(8, 8, 2))
    edge = np.isfinite(two["depth"]) != np.isfinite(plain["depth"])
    assert edge.sum() * 100 < p.width * p.height, edge.sum()


def test_the_dynamic_fogs_count_comes_with_the_depth():
    sc = H.scene_bulb(96, 64, seed=4)  # (every step counts: dfog_on_it 0)
    p = O.params(sc)
    t, cnt = O.cone_grid(p, 8, 8)
    assert np.isinf(t[0, 0]) and cnt[0, 0] > 0.0
    one = O.render(p, threads=0, cone=(8, 8, 1))
    assert not np.isfinite(one["depth"][0]) and (one["gbuf"][0, 1] >> 16) & 0x3FF > 0

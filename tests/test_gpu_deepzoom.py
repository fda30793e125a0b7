"""Deep zooms: the camera 1e-9 scene units from a surface, where a pixel is
~1e-14 units wide and FP32 cannot tell neighbouring pixels' points apart.
The backend keeps the ray origin and the depth in f64 in every mode, and
the automatic precision switches the march to FP64 there (the reference's
needs_double rule, src/gpu.rs:569-576).  Each case reports its per-pixel
depth error against the f64 oracle, with and without that switch."""
import numpy as np
import pytest

import mb3d_hip as H
import oracle as O
import parity

pytestmark = pytest.mark.gpu


def oracle_march(p):
    return O.render(p, threads=0)["depth"]


def oracle_de(p, pts):
    out = [O.de_at(p, q) for q in np.asarray(pts)]
    return np.array([a for a, _ in out]), np.array([b for _, b in out])


def error_map(g, r, p):
    """Per-pixel relative depth error (scene scale 2) where both hit; NaN
    elsewhere; and the mask agreement."""
    dg = np.asarray(g["depth64"], np.float64)
    dr = np.asarray(r["depth"], np.float64)
    both = np.isfinite(dg) & (dg > 0) & np.isfinite(dr) & (dr > 0)
    err = np.full(dg.shape, np.nan)
    err[both] = np.abs(dg[both] - dr[both]) * p.step_width / 2.0
    return err, float((np.isfinite(dg) == np.isfinite(dr)).mean())


def report(name, err, mask):
    e = err[np.isfinite(err)]
    q = (lambda a: float(np.percentile(e, a))) if e.size else (lambda a: 0.0)
    print(f"{name}: mask {mask:.6f}, {e.size} pixels hit, rel depth error p50 {q(50):.3e} p99 {q(99):.3e} "
          f"max {q(100):.3e}")


def test_bulb_deep_zoom_marches_in_fp64_and_matches_the_oracle(hip):
    sc = H.scene_bulb_deepzoom(320, 180, 1, oracle_march, oracle_de)
    p = hip.params_from_scene(sc)
    r = O.render(O.params(sc), threads=0)
    assert np.isfinite(r["depth"]).mean() > 0.5  # the surface fills the picture
    g = hip.march(p)
    assert g["counters"]["de_evals_f32"] == 0  # the automatic precision took FP64
    err, mask = error_map(g, r, p)
    report("bulb 1e-9, automatic (FP64)", err, mask)
    assert mask == 1.0 and np.nanmax(err) < 1e-9
    m = parity.compare(g, r, p.width, p.height, p.step_width, 2.0)
    assert m["iters_equal_inner"] == 1.0 and m["normal_deg_max"] < 1e-3, m
    # FP32 iterate forced: its per-pixel error, reported (the reason the
    # automatic precision does not choose it here)
    g32 = hip.march(p, opts=H.options(precision=H.PREC_F32))
    assert g32["counters"]["de_evals_f32"] > 0
    err32, mask32 = error_map(g32, r, p)
    report("bulb 1e-9, FP32 iterate forced", err32, mask32)


def test_hybrid_deep_zoom_config(hip):
    """The hybrid chain's deep zoom as configured: the camera 1e-9 in front
    of a seeded coarse march's surface point.  That point is a DE surface of
    the coarse picture (a pixel there is 0.07 units), and the chain's set is
    dust at the zoomed scale, so the oracle's picture is background there --
    and the backend's must be the same, pixel for pixel."""
    sc = H.scene_deepzoom(192, 108, 1, oracle_march)
    p = hip.params_from_scene(sc)
    r = O.render(O.params(sc), threads=0)
    g = hip.march(p)
    assert g["counters"]["de_evals_f32"] == 0
    err, mask = error_map(g, r, p)
    report("hybrid 1e-9, automatic (FP64)", err, mask)
    assert mask == 1.0
    assert np.array_equal(np.isfinite(g["depth64"]), np.isfinite(r["depth"]))
    if np.isfinite(err).any():
        assert np.nanmax(err) < 1e-9

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
    return O.de_many(p, np.asarray(pts, np.float64))


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


def px_error(g, r, p):
    """Depth error in step widths (pixels) where both hit."""
    dg = np.asarray(g["depth64"], np.float64)
    dr = np.asarray(r["depth"], np.float64)
    both = np.isfinite(dg) & (dg > 0) & np.isfinite(dr) & (dr > 0)
    return np.abs(dg[both] - dr[both])


def chain_case(hip, width, height, seed, bands):
    """Config 2 at width x height: the backend's whole picture, the oracle's
    on `bands` (row0, rows); both in the automatic precision (FP64 here)."""
    sc = H.scene_deepzoom(width, height, seed, oracle_de)
    p = hip.params_from_scene(sc)
    po = O.params(sc)
    g = hip.march(p)
    assert g["counters"]["de_evals_f32"] == 0  # the automatic precision took FP64
    g32 = hip.march(p, opts=H.options(precision=H.PREC_F32))
    assert g32["counters"]["de_evals_f32"] > 0
    for row0, rows in bands:
        r = O.render(po, row0=row0, rows=rows, threads=0)
        sl = slice(row0 * width, (row0 + rows) * width)
        gb = {k: (v[sl] if isinstance(v, np.ndarray) else v) for k, v in g.items()}
        g32b = {k: (v[sl] if isinstance(v, np.ndarray) else v) for k, v in g32.items()}
        yield p, r, gb, g32b, (row0, rows)


@pytest.mark.parametrize("seed", [1, 2])
def test_hybrid_deep_zoom_config(hip, seed):
    """Config 2: the hybrid chain at 1920 x 1080, the camera 1e-9 outside
    the chain's surface (the boundary of its bounded orbits, found by a
    seeded coarse march and bisected), FOV 0.5 degrees.  The surface fills
    the view; the automatic precision marches it in FP64 (the reference's
    needs_double rule) and matches the oracle: every mask bit, depths within
    1e-6 of a step width (8.1e-21 units, 4e-12 relative to the depth),
    iteration counts equal.  The FP32 iterate, forced, is reported per
    pixel: at this zoom a pixel is 1/58,000 of an FP32 ulp of the
    coordinates, so it cannot tell the pixels' points apart."""
    bands = ((0, 6), (537, 6), (1074, 6))
    for p, r, g, g32, band in chain_case(hip, 1920, 1080, seed, bands):
        err, mask = error_map(g, r, p)
        report(f"hybrid 1e-9 seed {seed} rows {band}, automatic (FP64)", err, mask)
        assert np.isfinite(r["depth"]).mean() > 0.5  # the surface fills the view
        assert mask == 1.0
        assert np.nanmax(err) < 1e-9
        px = px_error(g, r, p)
        print(f"  depth error in step widths: max {px.max():.3e}")
        assert px.max() < 1e-6
        m = parity.compare(g, r, p.width, band[1], p.step_width, 2.0)
        assert m["iters_equal_inner"] == 1.0, m
        err32, mask32 = error_map(g32, r, p)
        report(f"hybrid 1e-9 seed {seed} rows {band}, FP32 iterate forced", err32, mask32)
        px32 = px_error(g32, r, p)
        if px32.size:
            print(f"  FP32 depth error in step widths: p50 {np.percentile(px32, 50):.3e} "
                  f"p99 {np.percentile(px32, 99):.3e} max {px32.max():.3e}")


def test_hybrid_deep_zoom_whole_picture(hip):
    """Config 2's scene at 192 x 108, every pixel against the oracle."""
    for p, r, g, g32, band in chain_case(hip, 192, 108, 3, ((0, 108),)):
        err, mask = error_map(g, r, p)
        report("hybrid 1e-9 192x108 seed 3, automatic (FP64)", err, mask)
        assert np.isfinite(r["depth"]).mean() > 0.5
        assert mask == 1.0 and px_error(g, r, p).max() < 1e-6
        m = parity.compare(g, r, p.width, p.height, p.step_width, 2.0)
        assert m["iters_equal_inner"] == 1.0 and m["normal_deg_max"] < 1e-3, m
        err32, mask32 = error_map(g32, r, p)
        report("hybrid 1e-9 192x108 seed 3, FP32 iterate forced", err32, mask32)

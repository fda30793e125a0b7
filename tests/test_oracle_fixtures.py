"""The reference's scene fixtures as the oracle renders them (CPU): their
keys transcribed map to the modes the reference's loader gives them (4D
hybrid loop with the 4D rotation, the interpolation hybrid with the
analytic box DE, smooth normals), and each renders with the reference's
coverage window (tests/render.rs:24-31: between 5 % and 99.5 % object)."""
import math

import numpy as np
import pytest

import mb3d_hip as H
import oracle as O


def test_quaternion_julia_fixture_is_the_4d_loop():
    p = O.params(H.scene_m3s_quaternion_julia(160, 120))
    assert p.four_d and not p.ipol and p.do_julia
    r = O.render(p, threads=0)
    assert 0.05 < np.isfinite(r["depth"]).mean() < 0.995


def test_interpolation_boxes_fixture_is_the_blended_box():
    p = O.params(H.scene_m3s_interpolation_boxes(160, 120))
    assert p.ipol and not p.four_d and p.de_option == 11 and p.sm_normals == 2
    assert math.isclose(p.ipol_w[0], 0.6, rel_tol=1e-6) and math.isclose(p.ipol_w[1], 0.4, rel_tol=1e-6)
    r = O.render(p, threads=0)
    assert 0.05 < np.isfinite(r["depth"]).mean() < 0.995

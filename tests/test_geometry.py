"""Host mesh -> SDF builder (numpy), pinned by closed forms: the winding number of a
closed sphere mesh is 1 inside and 0 outside, and the signed distance of an icosphere
matches the analytic ball within the polyhedral approximation plus the 0.4-cell surface
sampling (mesh_sdf.py:328-366 of the reference builds the same way)."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd.geometry import build_sdf, icosphere, winding_number


def _ball_error(sdf, r):
    ax = [sdf.origin[d] + sdf.cell * np.arange(sdf.res) for d in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    return np.abs(sdf.values - (np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - r)).max()


def test_winding_number_of_a_closed_sphere():
    v, f = icosphere(0.1, 2)
    w = winding_number(np.array([[0.0, 0.0, 0.0], [0.05, -0.03, 0.02], [0.2, 0.0, 0.0], [0.0, 0.0, -0.5]]), v, f)
    assert np.allclose(w, [1.0, 1.0, 0.0, 0.0], atol=1e-9)
    assert np.allclose(winding_number(np.zeros((1, 3)), v, f[:, ::-1]), -1.0)  # flipped faces


def test_icosphere_sdf_matches_the_ball():
    v, f = icosphere(0.1, 2)
    sdf = build_sdf(v, f, res=20)
    # icosphere faces sit inside the ball by at most r (1 - cos(edge angle / 2)) ~ 1.2 % r
    assert _ball_error(sdf, 0.1) < 0.25 * sdf.cell + 0.015 * 0.1
    assert sdf.values[10, 10, 10] < 0 and sdf.values[0, 0, 0] > 0
    assert sdf.grads.shape == (20, 20, 20, 3)
    # the gradient of a distance field has unit length away from the centre and the medial axis
    g = np.linalg.norm(sdf.grads[2:-2, 2:-2, 2:-2], axis=-1)
    assert np.median(g) == pytest.approx(1.0, abs=0.05)


def test_probe_fixes_inverted_orientation():
    v, f = icosphere(0.1, 1)
    a = build_sdf(v, f, res=16)
    b = build_sdf(v, f[:, ::-1], res=16)       # inward-facing faces
    assert np.array_equal(np.sign(a.values), np.sign(b.values))


def test_bad_faces_rejected():
    v, f = icosphere(0.1, 0)
    with pytest.raises(ValueError, match="index"):
        build_sdf(v, f + 100, res=12)

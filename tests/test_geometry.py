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


def test_cached_builders_hit_miss_and_keys(tmp_path, monkeypatch):
    """The cached builders (mesh_sdf.py:496-563): a second call with the same geometry and
    grid loads the npz instead of building (bitwise-equal fields); another resolution,
    builder or sign probe is another file; a bad probe raises ValueError; no cache_dir
    builds without touching the disk."""
    import mpm_amd.geometry as geo

    v, f = geo.icosphere(0.2, 1)
    calls = {"sdf": 0, "cdf": 0}
    real_sdf, real_cdf = geo.build_sdf, geo.build_surface_cdf

    def count_sdf(*a, **k):
        calls["sdf"] += 1
        return real_sdf(*a, **k)

    def count_cdf(*a, **k):
        calls["cdf"] += 1
        return real_cdf(*a, **k)

    monkeypatch.setattr(geo, "build_sdf", count_sdf)
    monkeypatch.setattr(geo, "build_surface_cdf", count_cdf)
    a = geo.build_sdf_cached(v, f, res=16, cache_dir=tmp_path)
    b = geo.build_sdf_cached(v, f, res=16, cache_dir=tmp_path)
    assert calls["sdf"] == 1
    for fld in ("values", "grads", "origin"):
        assert np.array_equal(getattr(a, fld), getattr(b, fld))
    assert a.cell == b.cell and a.sdf_max == b.sdf_max
    geo.build_sdf_cached(v, f, res=20, cache_dir=tmp_path)
    geo.build_sdf_cached(v, f, res=16, cache_dir=tmp_path, interior_probe=(0.0, 0.0, 0.01))
    assert calls["sdf"] == 3
    assert len(list(tmp_path.glob("sdf_*.npz"))) == 3
    with pytest.raises(ValueError, match="interior_probe"):
        geo.build_sdf_cached(v, f, res=16, cache_dir=tmp_path, interior_probe=(0.0, np.nan, 0.0))
    sv, sf = geo.square_sheet(0.3)
    c = geo.build_surface_cdf_cached(sv, sf, res=16, cache_dir=tmp_path / "cdf")
    d = geo.build_surface_cdf_cached(sv, sf, res=16, cache_dir=tmp_path / "cdf")
    assert calls["cdf"] == 1
    assert np.array_equal(c.values, d.values) and np.array_equal(c.valid, d.valid)
    assert geo._field_key(1, sv, sf, [16, 3.0, 4.0], "cpu") != geo._field_key(1, sv, sf, [16, 3.0, 4.0], "hip")
    geo.build_surface_cdf_cached(sv, sf, res=16)
    assert calls["cdf"] == 2 and len(list((tmp_path / "cdf").glob("cdf_*.npz"))) == 1

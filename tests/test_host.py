"""Host-side logic: synthetic scenes, shortwave windows, meteorology, OBJ
loader, optics builders (no GPU)."""
import math

import numpy as np
import pytest

from streamsim_amd import faceteb, optics, scenes


def _check_facets(sc):
    v = sc.vertices
    p = v[:, 0:3].reshape(-1, 3, 3).astype(np.float64)
    n = v[0::3, 3:6].astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)
    assert np.all(v[0::3, 3:6] == v[1::3, 3:6]) and np.all(v[0::3, 3:6] == v[2::3, 3:6])
    c = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert np.all(np.sum(c * n, axis=1) > 0)  # normal follows the winding


def test_row_crop_counts():
    sc = scenes.row_crop()
    assert sc.facet_count == 4800 and sc.leaf_count == 1600
    _check_facets(sc)
    ground_n = sc.facet_normals()[sc.leaf_count:]
    np.testing.assert_allclose(ground_n, [[0, 1, 0]] * 3200, atol=1e-6)
    leaf_n = sc.facet_normals()[:sc.leaf_count]
    assert np.all(leaf_n[:, 1] >= -1e-6)  # upper hemisphere
    side = np.linalg.norm(sc.vertices[1:3 * 1600:3, :3] - sc.vertices[0:3 * 1600:3, :3], axis=1)
    np.testing.assert_allclose(side, 0.08, rtol=1e-4)


def test_forest_counts():
    sc = scenes.forest()
    assert sc.facet_count == 60000 and sc.leaf_count == 40000
    _check_facets(sc)


def test_windows_flat_spectrum():
    w = faceteb.ShortwaveWindows.for_bands(8)
    assert w.count == 8 and w.width == 250
    dni, dif = w.irradiance(600.0, 0.5)
    # the direct integral over all windows equals directNormalScale * I_d
    assert dni.sum() == pytest.approx(min(1400.0, 600.0 * 0.8 / 0.5), rel=1e-5)
    assert dif.sum() == pytest.approx(600.0 * 0.2, rel=1e-5)
    w48 = faceteb.ShortwaveWindows.for_bands(48)
    assert w48.count == 48
    assert w48.last[-1] == 2000


def test_solar_position_noon_and_night():
    zen, az = faceteb.solar_position(172, 12.0, 40.0)
    assert zen == pytest.approx(40.0 - 23.45, abs=0.1)
    assert az == pytest.approx(180.0, abs=0.5)
    zen, _ = faceteb.solar_position(172, 0.0, 40.0)
    assert zen > 90
    s = faceteb.sun_vector(30.0, 135.0)
    assert s[1] == pytest.approx(math.cos(math.radians(30)))


def test_diurnal_meteo_shape():
    m = faceteb.diurnal_meteo(24)
    assert len(m) == 24
    ta = [x.air_temperature_k for x in m]
    assert max(ta) == pytest.approx(304.0, abs=0.1) and min(ta) == pytest.approx(292.0, abs=0.1)
    assert all(x.shortwave == 0.0 for x in m if x.zenith >= 90)


def test_make_optics_matches_reference_constants():
    sc = scenes.row_crop(leaves_per_row=5, cells=2)
    sun = faceteb.sun_vector(30.0, 135.0)
    sl = np.ones(sc.surface_count, np.float32)
    o = optics.make_optics(sc, sl, sun)
    leaf = np.repeat(np.arange(sc.facet_count) < sc.leaf_count, 2)
    assert np.all(o[leaf, 0] == np.float32(0.10)) and np.all(o[leaf, 1] == np.float32(0.05))
    assert np.all(o[~leaf, 0] == np.float32(0.20)) and np.all(o[~leaf, 1] == 0)
    up = ~leaf & (np.arange(sc.surface_count) % 2 == 0)
    np.testing.assert_allclose(o[up, 3], 700.0 * math.cos(math.radians(30)), rtol=1e-6)


def test_planck():
    # Wien peak near 2898/T um; Planck at 10 um, 300 K ~ 9.92 W m-2 sr-1 um-1
    assert optics.planck_radiance(10000.0, 300.0) == pytest.approx(9.92, rel=0.01)

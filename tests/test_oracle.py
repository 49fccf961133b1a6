"""Pins of the CPU oracle against closed forms, invariants and brute force.

Each test names the oracle function it pins and the independent fact that
pins it (a textbook view factor, an exact invariant of the reference's
graph, brute force over every triangle, an FP64 direct solve, a bisection
root, finite differences).
"""
import math

import numpy as np
import pytest

from streamsim_amd import scenes


def _square_grid(n, y=0.0, down=False):
    tri = scenes.grid_ground(0.0, 1.0, 0.0, 1.0, n, n)
    tri[:, :, 1] = y
    if down:
        tri = tri[:, [0, 2, 1], :]
    return tri


def _vertical_square(n):
    # square in the x = 0 plane (y, z in [0, 1]) with normal +x
    tri = scenes.grid_ground(0.0, 1.0, 0.0, 1.0, n, n)
    out = np.empty_like(tri)
    out[:, :, 0] = 0.0
    out[:, :, 1] = tri[:, :, 0]
    out[:, :, 2] = tri[:, :, 2]
    v = scenes.triangles_to_vertices(out)
    flip = v[0::3, 3] < 0  # make the normal point to +x
    out[flip] = out[flip][:, [0, 2, 1], :]
    return out


def _vf_parallel_squares(a=1.0, c=1.0):
    X = Y = a / c
    t = 0.5 * math.log((1 + X * X) * (1 + Y * Y) / (1 + X * X + Y * Y))
    t += X * math.sqrt(1 + Y * Y) * math.atan(X / math.sqrt(1 + Y * Y))
    t += Y * math.sqrt(1 + X * X) * math.atan(Y / math.sqrt(1 + X * X))
    t -= X * math.atan(X) + Y * math.atan(Y)
    return 2.0 / (math.pi * X * Y) * t


def _vf_perpendicular_squares():
    H = W = 1.0
    s = math.sqrt(H * H + W * W)
    a = W * math.atan(1 / W) + H * math.atan(1 / H) - s * math.atan(1 / s)
    l1 = (1 + W * W) * (1 + H * H) / (1 + W * W + H * H)
    l2 = (W * W * (1 + W * W + H * H) / ((1 + W * W) * (W * W + H * H))) ** (W * W)
    l3 = (H * H * (1 + H * H + W * W) / ((1 + H * H) * (H * H + W * W))) ** (H * H)
    return (a + 0.25 * math.log(l1 * l2 * l3)) / (math.pi * W)


def test_hemisphere_directions(oracle):
    # core:475-489: y_i = (i + 1/2) / 40 (equal solid angle), unit length
    d = oracle.hemisphere_directions(40)
    assert d.shape == (40, 3)
    np.testing.assert_allclose(d[:, 1], (np.arange(40) + 0.5) / 40, rtol=0, atol=1e-7)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    az = np.arctan2(d[:, 2], d[:, 0])
    step = np.mod(np.diff(az), 2 * np.pi)
    np.testing.assert_allclose(step, np.mod(2.39996322972865332, 2 * np.pi), atol=1e-5)


def test_sun_direction(oracle):
    # core:491-498: +X north, +Z east, +Y up
    np.testing.assert_allclose(oracle.sun_direction(0.0, 0.0), [0, 1, 0], atol=1e-7)
    np.testing.assert_allclose(oracle.sun_direction(90.0, 0.0), [1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(oracle.sun_direction(90.0, 90.0), [0, 0, 1], atol=1e-6)


def _graph_invariants(g, ground_sides_below=False):
    S = g.S
    # closure: sky + sum(count) == den on every side (vk:1042-1056)
    tot = g.skyCounts.astype(np.int64).copy()
    rows = np.repeat(np.arange(S), np.diff(g.rowOffsets))
    np.add.at(tot, rows, g.counts)
    assert np.array_equal(tot, g.denominators)
    # no same-facet entries (visibility_resolve.comp:54-56)
    assert not np.any((rows >> 1) == (g.neighbors >> 1))
    return rows


def test_raster_graph_invariants(oracle):
    sc = scenes.row_crop(leaves_per_row=30, cells=8)
    g = oracle.raster_graph(sc.vertices, sc.bmin, sc.bmax, raster=256, layers=64)
    rows = _graph_invariants(g)
    # den[2f] == den[2f+1] (both sides counted per fragment, resolve.comp:104-108)
    assert np.array_equal(g.denominators[0::2], g.denominators[1::2])
    # symmetric counts: count(a, b) == count(b, a) (vk:1031-1036)
    fwd = dict(zip(zip(rows.tolist(), g.neighbors.tolist()), g.counts.tolist()))
    for (a, b), c in fwd.items():
        assert fwd[(b, a)] == c
    assert g.closure == 0.0


def test_raster_single_square_sees_sky(oracle):
    sc = scenes._finish("sq", None, scenes.grid_ground(0.0, 1.0, 0.0, 1.0, 1, 1))
    g = oracle.raster_graph(sc.vertices, sc.bmin, sc.bmax, raster=128, layers=8)
    assert g.nnz == 0
    assert np.all(g.denominators > 0)
    assert np.array_equal(g.skyCounts, g.denominators)


def test_ray_graph_brute_equals_bvh(oracle):
    # the closest hit is defined over all triangles; the oracle's BVH must
    # reproduce brute force exactly
    sc = scenes.row_crop(leaves_per_row=25, cells=5)
    g1, h1 = oracle.ray_graph(sc.vertices, 11, 48, brute=True, hits=True)
    g2, h2 = oracle.ray_graph(sc.vertices, 11, 48, brute=False, hits=True)
    assert np.array_equal(h1, h2)
    assert np.array_equal(g1.neighbors, g2.neighbors) and np.array_equal(g1.counts, g2.counts)
    _graph_invariants(g1)
    assert np.all(g1.denominators == 48)


def test_ray_graph_periodic_brute_equals_bvh_and_closes_the_gap(oracle):
    # periodic copies (sh/visibility.vert:15-35): a 1 x 1 ground tile under
    # a 1 x 1 roof 0.1 above it.  Alone, the ground sees the roof with the
    # parallel-squares view factor; tiled with the 20 neighbours of the
    # reference's list, rays leaving the tile hit the copies and the view
    # factor goes to 1 (only rays flatter than ~87 deg get past 2 tiles)
    ground = _square_grid(4)
    roof = _square_grid(4, y=0.1, down=True)
    v = scenes.triangles_to_vertices(np.concatenate([roof, ground]))
    nr, R = roof.shape[0], 256
    kw = dict(periodic=20, size=(1.0, 1.0))
    gb, hb = oracle.ray_graph(v, 5, R, brute=True, hits=True, **kw)
    gv, hv = oracle.ray_graph(v, 5, R, hits=True, **kw)
    assert np.array_equal(hb, hv)
    g0 = oracle.ray_graph(v, 5, R)

    def roof_fraction(g):
        rows = np.repeat(np.arange(g.S), np.diff(g.rowOffsets))
        src = (rows % 2 == 0) & (rows // 2 >= nr)  # upward sides of the ground
        return g.counts[src & (g.neighbors // 2 < nr)].sum() / (ground.shape[0] * R)

    f0, fp = roof_fraction(g0), roof_fraction(gv)
    sigma = math.sqrt(0.25 / (ground.shape[0] * R))
    assert f0 == pytest.approx(_vf_parallel_squares(1.0, 0.1), abs=5 * sigma)
    assert fp > 0.99
    # a low sun from +X shades the whole ground only with the periodic roof
    sun = np.array([0.5, 0.1, 0.0], np.float32)
    sun /= np.linalg.norm(sun)
    s0 = oracle.ray_sunlit(v, sun, 3, 32)
    sp = oracle.ray_sunlit(v, sun, 3, 32, **kw)
    assert np.array_equal(sp, oracle.ray_sunlit(v, sun, 3, 32, brute=True, **kw))
    ground_up = np.arange(2 * nr, 2 * (nr + ground.shape[0]), 2)
    assert s0[ground_up].mean() > 0.3 and sp[ground_up].max() == 0.0


def test_ray_trace_one_hits_plane(oracle):
    v = scenes.triangles_to_vertices(_square_grid(1, y=1.0, down=True))
    side, t = oracle.trace_one(v, [0.3, 0.0, 0.4], [0.0, 2.0, 0.0], 0xFFFFFFFF)
    assert side // 2 in (0, 1) and side % 2 == 0  # ray going up hits the downward normal side
    assert t == pytest.approx(0.5)
    side, _ = oracle.trace_one(v, [0.3, 2.0, 0.4], [0.0, -1.0, 0.0], 0xFFFFFFFF)
    assert side % 2 == 1


def test_ray_cosine_sampling_direction_law(oracle):
    # Malley's method: cos(theta) of the local direction has density 2 cos
    # sin, i.e. E[cos] = 2/3 and E[cos^2] = 1/2
    v = scenes.triangles_to_vertices(_square_grid(1))
    c = []
    for ray in range(4000):
        _, d = oracle.ray_generate(v, 0, ray, 5)
        c.append(d[1] / np.linalg.norm(d))
    c = np.asarray(c)
    assert c.min() > 0
    assert c.mean() == pytest.approx(2 / 3, abs=0.01)
    assert (c * c).mean() == pytest.approx(0.5, abs=0.01)


def test_ray_view_factor_parallel_squares(oracle):
    n = 6
    bottom = _square_grid(n)
    top = _square_grid(n, y=1.0, down=True)
    v = scenes.triangles_to_vertices(np.concatenate([bottom, top]))
    nb = bottom.shape[0]
    R = 1024
    g = oracle.ray_graph(v, 3, R)
    rows = np.repeat(np.arange(g.S), np.diff(g.rowOffsets))
    src = (rows % 2 == 0) & (rows // 2 < nb)        # upward sides of the bottom square
    hit_top = (g.neighbors // 2) >= nb
    F = g.counts[src & hit_top].sum() / (nb * R)
    sigma = math.sqrt(0.2 * 0.8 / (nb * R))
    assert F == pytest.approx(_vf_parallel_squares(), abs=5 * sigma)


def test_ray_view_factor_perpendicular_squares(oracle):
    n = 6
    a = _square_grid(n)
    b = _vertical_square(n)
    v = scenes.triangles_to_vertices(np.concatenate([a, b]))
    na = a.shape[0]
    R = 1024
    g = oracle.ray_graph(v, 4, R)
    rows = np.repeat(np.arange(g.S), np.diff(g.rowOffsets))
    src = (rows % 2 == 0) & (rows // 2 < na)
    F = g.counts[src & ((g.neighbors // 2) >= na)].sum() / (na * R)
    sigma = math.sqrt(0.2 * 0.8 / (na * R))
    assert F == pytest.approx(_vf_perpendicular_squares(), abs=5 * sigma)


def test_ray_sunlit_brute_equals_bvh_and_shadow(oracle):
    # a small square floating over a ground grid casts a shadow: the shaded
    # ground sides see fraction < 1, the rest 1, down-facing sides 0
    ground = scenes.grid_ground(-2.0, 2.0, -2.0, 2.0, 8, 8)
    roof = scenes.grid_ground(-0.5, 0.5, -0.5, 0.5, 1, 1)
    roof[:, :, 1] = 1.0
    sc = scenes._finish("shadow", roof, ground)
    sun = oracle.sun_direction(0.0, 0.0)
    f1 = oracle.ray_sunlit(sc.vertices, sun, 9, 64, brute=True)
    f2 = oracle.ray_sunlit(sc.vertices, sun, 9, 64, brute=False)
    assert np.array_equal(f1, f2)
    assert np.all(f1[1::2] == 0.0)
    g_up = f1[2 * sc.leaf_count::2]
    assert g_up.min() == 0.0 and g_up.max() == 1.0
    # fully shaded ground fraction ~ (1 m^2 / 16 m^2)
    area = sc.facet_areas()[sc.leaf_count:]
    shade = np.sum(area * (1.0 - g_up)) / np.sum(area)
    assert shade == pytest.approx(1.0 / 16.0, abs=0.01)


def test_solve_isolated_facet_closed_form(oracle):
    # one facet alone: in(s) = S (all sky), so B0 = e0 + r0 (S + D0) + t0 (S + D1)
    sc = scenes._finish("one", None, scenes.grid_ground(0.0, 1.0, 0.0, 1.0, 1, 1)[:1])
    g = oracle.raster_graph(sc.vertices, sc.bmin, sc.bmax, raster=64, layers=4)
    o = np.array([[0.3, 0.2, 5.0, 400.0], [0.1, 0.2, 2.0, 0.0]], np.float32)
    B, _ = oracle.solve(g, o, 100.0, 64)
    e, r, t, D, S = o[:, 2], o[:, 0], o[:, 1], o[:, 3], 100.0
    B0 = e[0] + r[0] * (S + D[0]) + t[0] * (S + D[1])
    B1 = e[1] + r[1] * (S + D[1]) + t[1] * (S + D[0])
    np.testing.assert_allclose(B, [B0, B1], rtol=1e-6)


def test_solve_matches_direct_lu(oracle):
    sc = scenes.row_crop(leaves_per_row=25, cells=5)
    g = oracle.ray_graph(sc.vertices, 2, 64)
    rng = np.random.default_rng(0)
    S = g.S
    o = np.zeros((S, 4), np.float32)
    o[:, 0] = rng.uniform(0.0, 0.5, S)
    o[:, 1] = rng.uniform(0.0, 0.4, S)
    o[:, 2] = rng.uniform(0.0, 10.0, S)
    o[:, 3] = rng.uniform(0.0, 300.0, S)
    Bd = oracle.solve_direct(g, o, 80.0)
    Bj, _ = oracle.solve(g, o, 80.0, 400)
    np.testing.assert_allclose(Bj, Bd, rtol=2e-6)
    # the multi-band FP64 Jacobi converges to the same fixed point
    r = o[:, :1].repeat(2, 1)
    t = o[:, 1:2].repeat(2, 1)
    e = o[:, 2:3].repeat(2, 1)
    D = o[:, 3:4].repeat(2, 1)
    Bb, it, last = oracle.solve_bands(g, r, t, e, D, [80.0, 80.0], tol_rel=1e-9)
    assert last <= 1e-9
    np.testing.assert_allclose(Bb[:, 0], Bd, rtol=1e-7)
    np.testing.assert_allclose(Bb[:, 1], Bd, rtol=1e-7)


def test_solve_bands_thread_count_independent(oracle):
    # the all-core variant (OpenMP over facets) computes every value as the
    # single-threaded restatement does
    sc = scenes.row_crop(leaves_per_row=25, cells=5)
    g = oracle.ray_graph(sc.vertices, 3, 64)
    rng = np.random.default_rng(1)
    S, nb = g.S, 5
    r = rng.uniform(0.0, 0.5, (S, nb)).astype(np.float32)
    t = (rng.uniform(0.0, 1.0, (S, nb)) * (1.0 - r) * 0.8).astype(np.float32)
    e = rng.uniform(0.0, 5.0, (S, nb)).astype(np.float32)
    D = rng.uniform(0.0, 300.0, (S, nb)).astype(np.float32)
    sky = rng.uniform(10.0, 80.0, nb).astype(np.float32)
    n = oracle.max_threads()
    try:
        oracle.set_threads(1)
        B1, it1, l1 = oracle.solve_bands(g, r, t, e, D, sky, tol_rel=1e-8)
        oracle.set_threads(max(n, 4))
        Bn, itn, ln = oracle.solve_bands(g, r, t, e, D, sky, tol_rel=1e-8)
    finally:
        oracle.set_threads(n)
    assert it1 == itn and l1 == ln
    assert np.array_equal(B1, Bn)


def test_incident_definition(oracle):
    # H = max(0, (sky S + sum count B) / den + D)
    sc = scenes.row_crop(leaves_per_row=10, cells=3)
    g = oracle.ray_graph(sc.vertices, 2, 32)
    rng = np.random.default_rng(1)
    B = rng.uniform(0, 100, g.S).astype(np.float32)
    o = np.zeros((g.S, 4), np.float32)
    o[:, 3] = rng.uniform(0, 50, g.S)
    H = oracle.incident(g, B, o, 70.0)
    rows = np.repeat(np.arange(g.S), np.diff(g.rowOffsets))
    acc = g.skyCounts.astype(np.float64) * 70.0
    np.add.at(acc, rows, g.counts * B[g.neighbors].astype(np.float64))
    ref = np.maximum(0.0, acc / g.denominators + o[:, 3])
    np.testing.assert_allclose(H, ref, rtol=1e-6)


def _soil_material(oracle, **kw):
    m = oracle.OrEbMaterial()
    vals = dict(tirReflectance=0.05, tirTransmittance=0.0, surfaceResistance=120.0, specificHeat=1180.0,
                density=1800.0, thermalConductivity=1.55, soilMoisture=0.0, heatCapacityPerArea=60000.0,
                convectiveScale=1.0, soilTemperatureMethod=0, vegetation=0)
    vals.update(kw)
    for k, v in vals.items():
        setattr(m, k, v)
    return m


def test_eb_isolated_ground_matches_bisection(oracle):
    # an isolated horizontal ground facet, dry soil (LE = 0), G = 0.35 Rn:
    # 0.65 (a + eps L - eps sigma T^4) - h_a (T - Ta) = 0 per side
    S = 2
    g = oracle.Graph(np.zeros(S + 1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                     np.full(S, 64, np.uint32), np.full(S, 64, np.uint32))
    p = oracle.OrEbParams()
    p.airTemperatureK = 298.0
    p.vaporPressure = 15.0
    p.pressure = 1000.0
    p.windSpeed = 2.0
    p.skyLongwave = 350.0
    p.dTime = 3600.0
    p.tolerance = 1e-4  # T is stored in FP32: its ulp near 300 K is 3e-5 K
    p.relaxation = 0.5
    p.maxCoupling = 500
    p.vegetationMethod = 0
    p.lwMaxSweeps = 50
    p.lwTolRel = 1e-9
    absorbed = np.array([500.0, 20.0], np.float32)
    T0 = np.full(S, 298.0, np.float32)
    T, _, it, conv, _, _ = oracle.eb_node(g, 1, 0, [0], [_soil_material(oracle)], absorbed, p, T0, T0)
    assert conv
    sigma, eps, L, Ta = 5.670374419e-8, 0.95, 350.0, 298.0
    ha = 5.8 + 4.1 * math.sqrt(2.0)
    for s in range(2):
        f = lambda x: 0.65 * (absorbed[s] + eps * L - eps * sigma * x ** 4) - ha * (x - Ta)
        lo, hi = 220.0, 360.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if f(lo) * f(mid) <= 0:
                hi = mid
            else:
                lo = mid
        assert T[s] == pytest.approx(0.5 * (lo + hi), abs=2e-3)


def test_mlp_gradient_matches_finite_differences(oracle):
    rng = np.random.default_rng(3)
    H, n = 8, 6
    npar = H * 9 + H + H * H + H + H + 1
    P = rng.normal(0, 0.4, npar).astype(np.float32)
    x = rng.normal(0, 1, (n, 9)).astype(np.float32)
    t = rng.normal(0, 1, n).astype(np.float32)
    loss, g = oracle.mlp_loss_grad(P, x, t, H)
    y = oracle.mlp_forward(P, x, H)
    assert loss == pytest.approx(np.mean((y - t) ** 2), rel=1e-12)
    idx = rng.choice(npar, 25, replace=False)
    for i in idx:
        h = np.float32(2.0 ** -8)
        Pp = P.copy()
        Pm = P.copy()
        Pp[i] += h
        Pm[i] -= h
        lp = np.mean((oracle.mlp_forward(Pp, x, H) - t) ** 2)
        lm = np.mean((oracle.mlp_forward(Pm, x, H) - t) ** 2)
        fd = (lp - lm) / (float(Pp[i]) - float(Pm[i]))
        assert g[i] == pytest.approx(fd, rel=2e-3, abs=2e-5)


def test_hapke_background_closed_forms(oracle):
    # core:546-605.  Pins: strength 0 gives the Lambertian value (shortwave
    # clamped to 0.999); at normal incidence and exit (mu0 = mu = 1, phase 0,
    # full surge) the kernel reduces to w (1 + H(1)^2) / (8 pi), so the
    # reflectance is base (1 + H^2) / H^2 with H = 3 / (1 + 2 sqrt(1 - w)),
    # w = 4 b / (1 + b)^2; with the sun in the plane of the side mu0 hits its
    # floor and the directional value vanishes; the longwave value is the
    # Lambertian one at cosine 1 for every strength and follows the closed
    # form of the angular factor elsewhere.
    n = np.array([0.0, 1.0, 0.0], np.float32)
    for b in (0.05, 0.2, 0.45):
        assert oracle.hapke_reflectance(b, [0.3, 0.9, 0.1], n, n, 0.0) == np.float32(b)
        w = 4 * b / (1 + b) ** 2
        H = 3.0 / (1.0 + 2.0 * np.sqrt(1.0 - w))
        want = min(1.0, b * (1 + H * H) / (H * H))
        assert oracle.hapke_reflectance(b, n, n, n, 1.0) == pytest.approx(want, rel=2e-6)
        # half strength: halfway between the Lambertian and the directional value
        assert oracle.hapke_reflectance(b, n, n, n, 0.5) == pytest.approx(0.5 * (b + want), rel=2e-6)
        # the sun below the side's plane is mirrored onto it (same value)
        assert oracle.hapke_reflectance(b, -n, n, n, 1.0) == oracle.hapke_reflectance(b, n, n, n, 1.0)
        assert oracle.hapke_reflectance(b, [1.0, 0.0, 0.0], n, n, 1.0) < 1e-3 * b
    assert oracle.hapke_reflectance(1.0, n, n, n, 0.0) == np.float32(0.999)
    for e in (0.9, 0.95, 0.98):
        for s in (0.0, 0.5, 1.0):
            assert oracle.hapke_emissivity(e, 1.0, s) == pytest.approx(e, rel=1e-7)
        w = 1 - e * e
        root = np.sqrt(1 - w)
        f = lambda mu: root * (1 + 2 * mu / 1.5) / (1 + 2 * mu / 1.5 * root)  # noqa: E731
        for mu in (0.2, 0.7):
            want = min(1.0, e * f(mu) / f(1.0))
            assert oracle.hapke_emissivity(e, mu, 1.0) == pytest.approx(want, rel=2e-6)
            assert oracle.hapke_emissivity(e, mu, 0.3) == pytest.approx(e + 0.3 * (want - e), rel=2e-6)

"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Integer / index work (form-factor hit counts, CSR indices, sunlit counts) is
bit-exact against the oracle's host replay of the same seeded rays.
Radiosities agree within 1e-5 relative and temperatures within 0.01 K.
"""
import math

import numpy as np
import pytest

from streamsim_amd import faceteb, optics, scenes
from streamsim_amd._lib import FH_PRECISION_BF16_DC, FH_PRECISION_BF16_F, FH_PRECISION_FP8_INC, FH_PRECISION_FP32

pytestmark = pytest.mark.gpu

RTOL = 1e-5  # radiosity parity bound


def _engine(R=64, K=16, seed=7):
    from streamsim_amd import FacetEngine
    return FacetEngine(rays_per_side=R, sun_rays_per_side=K, seed=seed)


def _as_oracle_graph(oracle, g):
    return oracle.Graph(g.rowOffsets, g.neighbors, g.counts, g.denominators, g.skyCounts)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))


@pytest.mark.parametrize("leaves,cells,R,seed", [(20, 6, 64, 7), (400, 40, 256, 1)])
def test_form_factor_graph_bitexact(oracle, leaves, cells, R, seed):
    print(f"graph case leaves={leaves} cells={cells} R={R}", flush=True)
    sc = scenes.row_crop(leaves_per_row=leaves, cells=cells)
    eng = _engine(R=R, seed=seed)
    eng.setGeometry(sc.vertices)
    d = eng.buildVisibilityGraph()
    g = eng.getGraph()
    og = oracle.ray_graph(sc.vertices, seed, R)
    assert d.directedEdgeCount == og.nnz
    assert d.maxClosureError == 0.0
    for k in ("rowOffsets", "neighbors", "counts", "denominators", "skyCounts"):
        assert np.array_equal(getattr(g, k), getattr(og, k)), k


def test_form_factor_graph_forest_crown(oracle):
    # one forest crown over a ground patch: full-sphere leaf normals
    print("graph case crown", flush=True)
    rng = np.random.default_rng(5)
    crown = scenes.crown_leaves((2.0, 4.0, 2.0), (2.0, 3.0, 2.0), 400, 0.1, rng)
    sc = scenes._finish("crown", crown, scenes.grid_ground(0.0, 4.0, 0.0, 4.0, 10, 10))
    eng = _engine(R=96, seed=3)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    g = eng.getGraph()
    og = oracle.ray_graph(sc.vertices, 3, 96)
    assert np.array_equal(g.rowOffsets, og.rowOffsets)
    assert np.array_equal(g.neighbors, og.neighbors)
    assert np.array_equal(g.counts, og.counts)
    assert np.array_equal(g.skyCounts, og.skyCounts)


# 32 and 16 rays per side: one packet of 2 or 4 sides per wave; 24: one
# traversal per ray (24 does not divide the wave)
@pytest.mark.parametrize("K", [32, 16, 24])
def test_sunlit_bitexact(oracle, K):
    print(f"sunlit case K={K}", flush=True)
    sc = scenes.row_crop(leaves_per_row=60, cells=10)
    eng = _engine(R=32, K=K, seed=9)
    eng.setGeometry(sc.vertices)
    for zen, az in ((30.0, 135.0), (60.0, 250.0), (0.0, 0.0), (85.0, 20.0)):
        sun = oracle.sun_direction(zen, az)
        f = eng.computeSunlitFraction(sun)
        of = oracle.ray_sunlit(sc.vertices, sun, 9, K)
        assert np.array_equal(f, of), (zen, az)


@pytest.mark.parametrize("neighbors", [4, 20])
def test_periodic_graph_and_sunlit_bitexact(oracle, neighbors):
    # periodic copies of a row crop (10 x 10 m tile): form-factor CSR and
    # sunlit fractions equal the oracle's replay of the same tiles
    print(f"periodic case neighbors={neighbors}", flush=True)
    from streamsim_amd import FacetEngine
    sc = scenes.row_crop(leaves_per_row=60, cells=8)
    eng = FacetEngine(rays_per_side=64, sun_rays_per_side=32, seed=4, periodic_neighbors=neighbors,
                      periodic_size=(10.0, 10.0))
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    g = eng.getGraph()
    og = oracle.ray_graph(sc.vertices, 4, 64, periodic=neighbors, size=(10.0, 10.0))
    for k in ("rowOffsets", "neighbors", "counts", "denominators", "skyCounts"):
        assert np.array_equal(getattr(g, k), getattr(og, k)), k
    # the neighbours close part of the sky seen by the ground
    g0 = oracle.ray_graph(sc.vertices, 4, 64)
    ground = np.arange(2 * sc.leaf_count, sc.surface_count, 2)
    assert og.skyCounts[ground].sum() < g0.skyCounts[ground].sum()
    for zen, az in ((60.0, 100.0), (75.0, 270.0)):
        sun = oracle.sun_direction(zen, az)
        f = eng.computeSunlitFraction(sun)
        of = oracle.ray_sunlit(sc.vertices, sun, 4, 32, periodic=neighbors, size=(10.0, 10.0))
        assert np.array_equal(f, of), (zen, az)


def _vertex_bounds(sc):
    p = np.asarray(sc.vertices, np.float32)[:, :3]
    return p.min(axis=0), p.max(axis=0)


# the reference's raster graph on the device: CSR, denominators and sky
# counts equal the oracle's A-buffer restatement (the fragments of a pixel in
# depth order, ties in submission order) for the same directions; 384 pixels
# across with 8 periodic copies, 512 without
@pytest.mark.parametrize("neighbors,size", [(0, 512), (8, 384)])
def test_raster_graph_and_sunlit_bitexact(oracle, neighbors, size):
    print(f"raster graph case neighbors={neighbors} size={size}", flush=True)
    from streamsim_amd import FacetEngine
    sc = scenes.row_crop(leaves_per_row=100, cells=12)
    eng = FacetEngine(rays_per_side=32, sun_rays_per_side=16, seed=1, periodic_neighbors=neighbors,
                      periodic_size=(10.0, 10.0))
    eng.setGeometry(sc.vertices)
    dirs = oracle.hemisphere_directions(24)
    d = eng.buildRasterGraph(dirs, size, size, 64)
    g = eng.getGraph()
    lo, hi = _vertex_bounds(sc)
    og = oracle.raster_graph(sc.vertices, lo, hi, raster=size, layers=64, periodic=neighbors, px=10.0, pz=10.0,
                             dirs=dirs)
    assert d.fragmentOverflow == 0 and d.directedEdgeCount == og.nnz > 0
    for k in ("rowOffsets", "neighbors", "counts", "denominators", "skyCounts"):
        assert np.array_equal(getattr(g, k), getattr(og, k)), k
    assert abs(d.maxClosureError - og.closure) < 1e-6
    for zen, az in ((30.0, 135.0), (70.0, 260.0)):
        sun = oracle.sun_direction(zen, az)
        f = eng.computeRasterSunlitFraction(sun, size, size, 64)
        of = oracle.raster_sunlit(sc.vertices, lo, hi, sun, raster=size, layers=64, periodic=neighbors, px=10.0,
                                  pz=10.0)
        assert np.array_equal(f, of), (zen, az)


def test_raster_graph_wide_keys_and_regrown_buffers(oracle):
    # 38 000 facets: 76 000 sides need 17 bits, so the directed pairs are
    # sorted as 64-bit keys.  The first build (128 pixels across) leaves
    # fragment and pair capacities sized for itself; the second (384 across)
    # outgrows both, is measured and runs again at the measured sizes.  Both
    # graphs equal the oracle's.
    print("raster graph wide keys case", flush=True)
    from streamsim_amd import FacetEngine
    sc = scenes.row_crop(rows=4, leaves_per_row=4500, cells=100)
    assert 2 * sc.facet_count > 65536
    eng = FacetEngine(rays_per_side=32, sun_rays_per_side=16, seed=1)
    eng.setGeometry(sc.vertices)
    dirs = oracle.hemisphere_directions(6)
    lo, hi = _vertex_bounds(sc)
    for size in (128, 384):
        d = eng.buildRasterGraph(dirs, size, size, 64)
        g = eng.getGraph()
        og = oracle.raster_graph(sc.vertices, lo, hi, raster=size, layers=64, dirs=dirs)
        assert d.fragmentOverflow == 0 and d.directedEdgeCount == og.nnz > 0, size
        for k in ("rowOffsets", "neighbors", "counts", "denominators", "skyCounts"):
            assert np.array_equal(getattr(g, k), getattr(og, k)), (size, k)
    sun = oracle.sun_direction(40.0, 200.0)
    f = eng.computeRasterSunlitFraction(sun, 384, 384, 64)
    assert np.array_equal(f, oracle.raster_sunlit(sc.vertices, lo, hi, sun, raster=384, layers=64))


def test_raster_graph_overflow_reported(oracle):
    # two fragments allowed per pixel: the leaves over the ground overflow,
    # the call fails and leaves no graph
    print("raster overflow case", flush=True)
    from streamsim_amd import FacetEngine
    from streamsim_amd._lib import LogicError
    sc = scenes.row_crop(leaves_per_row=60, cells=8)
    eng = FacetEngine(rays_per_side=32, sun_rays_per_side=16, seed=1)
    eng.setGeometry(sc.vertices)
    with pytest.raises(RuntimeError, match="A-buffer overflow"):
        eng.buildRasterGraph(oracle.hemisphere_directions(8), 256, 256, 1)
    with pytest.raises(LogicError):
        eng.getGraph()
    with pytest.raises(RuntimeError, match="A-buffer overflow"):
        eng.computeRasterSunlitFraction(oracle.sun_direction(10.0, 0.0), 256, 256, 1)


def test_reference_solve_on_raster_graph(oracle):
    # the reference's own graph (A-buffer restatement) uploaded through
    # fh_set_graph; fixed 64 sweeps from zero vs solveCpu (FP64 gather)
    print("reference solve case", flush=True)
    sc = scenes.row_crop()
    g = oracle.raster_graph(sc.vertices, sc.bmin, sc.bmax, raster=512, layers=64)
    eng = _engine()
    eng.setGeometry(sc.vertices)
    d = eng.setGraph(g.rowOffsets, g.neighbors, g.counts, g.denominators, g.skyCounts)
    assert d.directedEdgeCount == g.nnz and d.maxClosureError == 0.0
    sun = oracle.sun_direction(30.0, 135.0)
    sl = oracle.raster_sunlit(sc.vertices, sc.bmin, sc.bmax, sun, raster=512, layers=64)
    o = optics.make_optics(sc, sl, sun)
    res = eng.solve(o, 100.0, 64, 1.0)
    ref, md = oracle.solve(g, o, 100.0, 64, 1.0)
    assert res.iterations == 64
    assert _rel(res.radiosity, ref) < RTOL
    H = eng.incidentIrradiance(res, o, 100.0)
    oH = oracle.incident(g, ref, o, 100.0)
    assert _rel(H, oH) < RTOL
    # relaxation path
    res2 = eng.solve(o, 100.0, 40, 0.7)
    ref2, _ = oracle.solve(g, o, 100.0, 40, 0.7)
    assert _rel(res2.radiosity, ref2) < RTOL


def test_solve_accelerated_batches(oracle):
    print("accelerated case", flush=True)
    sc = scenes.row_crop(leaves_per_row=100, cells=12)
    eng = _engine(R=64)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    og = _as_oracle_graph(oracle, eng.getGraph())
    o = optics.make_optics(sc, np.ones(sc.surface_count, np.float32), oracle.sun_direction(30.0, 135.0))
    res = eng.solveAccelerated(o, 100.0, None, 64, 8, 1.0e-4, 1.0)
    assert res.iterations % 8 == 0 and res.iterations <= 64
    assert res.maxDelta <= 1.0e-4 or res.iterations == 64
    ref, _ = oracle.solve(og, o, 100.0, res.iterations, 1.0)
    assert _rel(res.radiosity, ref) < RTOL
    # warm start from the converged answer stops after one batch
    res2 = eng.solveAccelerated(o, 100.0, res.radiosity, 64, 8, 1.0e-4, 1.0)
    assert res2.iterations == 8


def _c1_bands(oracle, sc, sl, sun):
    tir = dict(leaf_r=0.02, leaf_t=0.0, soil_r=0.05, leaf_sunlit_T=300.0, leaf_shaded_T=296.0,
               soil_sunlit_T=305.0, soil_shaded_T=295.0, sky_T=250.0)
    return optics.spectral_optics(sc, sl, sun, [650.0, 850.0, 10000.0], [0.05, 0.45, 0.02], [0.05, 0.45, 0.0],
                                  [0.15, 0.25, 0.05], tir)


@pytest.mark.parametrize("precision", [FH_PRECISION_FP32, FH_PRECISION_BF16_DC])
def test_row_crop_three_bands_converged(oracle, precision):
    # configuration 1: 4800 facets, 3 bands, 256 rays, relative change < 1e-6
    print(f"row crop bands precision={precision}", flush=True)
    sc = scenes.row_crop()
    eng = _engine(R=256, K=64, seed=1)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    og = _as_oracle_graph(oracle, eng.getGraph())
    sun = oracle.sun_direction(30.0, 135.0)
    sl = eng.computeSunlitFraction(sun)
    r, t, e, D, sky = _c1_bands(oracle, sc, sl, sun)
    res = eng.solveBands(r, t, e, D, sky, tolRel=1e-6, precision=precision)
    assert res.maxRelDelta <= 1e-6
    ref, it, last = oracle.solve_bands(og, r, t, e, D, sky, tol_rel=1e-9)
    assert _rel(res.radiosity, ref) < RTOL
    H = eng.incidentBands(res.radiosity, D, sky)
    oH = oracle.incident_bands(og, ref, D, sky)
    assert _rel(H, oH) < RTOL
    if precision == FH_PRECISION_BF16_DC:
        assert res.refreshes >= 1


def test_direct_lu_small_graph(oracle):
    print("direct LU case", flush=True)
    sc = scenes.row_crop(leaves_per_row=40, cells=6)
    eng = _engine(R=128)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    og = _as_oracle_graph(oracle, eng.getGraph())
    rng = np.random.default_rng(2)
    S = sc.surface_count
    o = np.zeros((S, 4), np.float32)
    o[:, 0] = rng.uniform(0.0, 0.6, S)
    o[:, 1] = rng.uniform(0.0, 0.35, S)
    o[:, 2] = rng.uniform(0.0, 20.0, S)
    o[:, 3] = rng.uniform(0.0, 400.0, S)
    Bd = oracle.solve_direct(og, o, 60.0)
    for prec in (FH_PRECISION_FP32, FH_PRECISION_BF16_DC):
        res = eng.solveBands(o[:, :1], o[:, 1:2], o[:, 2:3], o[:, 3:4], [60.0], tolRel=1e-7, precision=prec)
        assert _rel(res.radiosity[:, 0], Bd) < RTOL


def test_error_paths():
    print("error paths", flush=True)
    from streamsim_amd._lib import LogicError
    eng = _engine()
    with pytest.raises(LogicError):
        eng.buildVisibilityGraph()
    sc = scenes.row_crop(leaves_per_row=5, cells=2)
    bad = sc.vertices.copy()
    bad[0, 3:6] = [0.0, 2.0, 0.0]
    with pytest.raises(ValueError):
        eng.setGeometry(bad)
    eng.setGeometry(sc.vertices)
    o = np.zeros((sc.surface_count, 4), np.float32)
    with pytest.raises(LogicError):
        eng.solve(o, 100.0)
    eng.buildVisibilityGraph()
    o[0, 0] = 0.8
    o[0, 1] = 0.5
    with pytest.raises(ValueError):
        eng.solve(o, 100.0)
    o[0, :2] = 0.1
    with pytest.raises(ValueError):
        eng.solve(o, 100.0, 64, 1.5)
    with pytest.raises(ValueError):
        eng.solve(o, -1.0)
    res = eng.solve(o, 100.0, 8)
    assert res.iterations == 8


def test_eb_state_follows_the_scene():
    # a new geometry drops the energy-balance scene (its buffers are sized for
    # the old facet count): the node fails cleanly before any kernel runs;
    # non-finite sun directions, and a zero one under the Hapke ground, are
    # rejected
    print("eb scene guard", flush=True)
    from streamsim_amd._lib import LogicError
    sc = scenes.row_crop(leaves_per_row=20, cells=4)
    eng = _engine(R=64, K=16, seed=2)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    win = faceteb.ShortwaveWindows.for_bands(4)
    faceteb.setup_scene(eng, sc, win)
    cfg = faceteb.EbConfig()
    met = faceteb.diurnal_meteo(24)[12]
    dni, dif = win.irradiance(met.shortwave, math.cos(math.radians(met.zenith)))
    st = eng.ebRunNode(faceteb.node_params(met, cfg, first=True), dni, dif)
    assert st.converged
    bigger = scenes.row_crop(leaves_per_row=30, cells=5)
    eng.setGeometry(bigger.vertices)
    with pytest.raises(LogicError):
        eng.ebRunNode(faceteb.node_params(met, cfg, first=False), dni, dif)
    eng.buildVisibilityGraph()
    with pytest.raises(LogicError):
        eng.ebRunNode(faceteb.node_params(met, cfg, first=False), dni, dif)
    with pytest.raises(LogicError):
        eng.ebGetState()
    faceteb.setup_scene(eng, bigger, win)
    with pytest.raises(LogicError):  # no node on this scene yet
        eng.ebDiffuseRadiosity()
    p = faceteb.node_params(met, cfg, first=True)
    p.sun.x = float("nan")
    with pytest.raises(ValueError):
        eng.ebRunNode(p, dni, dif)
    eng.ebSetBackgroundAngularStrength(0.5)
    p = faceteb.node_params(met, cfg, first=True)
    p.sun.x = p.sun.y = p.sun.z = 0.0
    with pytest.raises(ValueError):
        eng.ebRunNode(p, dni, dif)
    st = eng.ebRunNode(faceteb.node_params(met, cfg, first=True), dni, dif)
    assert st.converged
    assert np.all(np.isfinite(eng.ebGetState()["temperature"]))
    # a night node solves no window: its diffuse-only radiosity is zero
    night = faceteb.diurnal_meteo(24)[0]
    dn, fn = win.irradiance(night.shortwave, math.cos(math.radians(night.zenith)))
    assert night.shortwave == 0 and not np.any(dn) and not np.any(fn)
    st = eng.ebRunNode(faceteb.node_params(night, cfg, first=False), dn, fn)
    total, per, ds = eng.ebDiffuseRadiosity()
    assert not np.any(total) and not np.any(per) and ds.iterations == 0 and st.swSweeps == 0


def _eb_case(oracle, leaves=40, cells=6, R=128, K=32, nb=4):
    sc = scenes.row_crop(leaves_per_row=leaves, cells=cells)
    eng = _engine(R=R, K=K, seed=4)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    win = faceteb.ShortwaveWindows.for_bands(nb)
    faceteb.setup_scene(eng, sc, win)
    return sc, eng, win


def _check_diffuse_solve(oracle, eng, og, r, t, dif, s):
    # the node's windows solved again without the direct irradiance
    # (withoutDirectLight, facetrtcore.cpp:1561-1569, 3533-3536): the same
    # optics and sky radiance, against the oracle's solve with D = 0; the
    # node's own state is left as it was
    total, per, ds = eng.ebDiffuseRadiosity()
    want, _, _ = oracle.solve_bands(og, r, t, np.zeros_like(r), np.zeros_like(r), dif, tol_rel=1e-9)
    assert _rel(per[want > 1e-3], want[want > 1e-3]) < RTOL
    assert _rel(total[want.sum(1) > 1e-3], want.sum(1)[want.sum(1) > 1e-3]) < RTOL
    assert ds.iterations > 0
    # light enhancement is the direct light's share: non-negative up to the
    # solves' tolerance, and positive where the sun shines
    enh = s["shortwaveRadiosity"].sum(1) - total
    assert enh.min() > -1e-4 * total.max() and enh[s["sunlit"] > 0.9].mean() > 1.0
    after = eng.ebGetState()
    for k in ("shortwaveRadiosity", "absorbedShortwave", "temperature", "longwaveRadiosity"):
        np.testing.assert_array_equal(after[k], s[k])


# 4 windows: entry-parallel sweeps; 24 and 47: band-parallel sweeps with the
# windows sorted internally, per-band retirement and (BF16) BF16 gathers
@pytest.mark.parametrize("nb", [4, 24, 47])
def test_eb_node_matches_oracle(oracle, nb):
    print(f"eb node case nb={nb}", flush=True)
    sc, eng, win = _eb_case(oracle, nb=nb)
    og = _as_oracle_graph(oracle, eng.getGraph())
    cfg = faceteb.EbConfig(tolerance=1e-3, radiosity_tolerance=1e-7, max_coupling=200)
    met = faceteb.diurnal_meteo(24)[12]
    dni, dif = win.irradiance(met.shortwave, math.cos(math.radians(met.zenith)))
    p = faceteb.node_params(met, cfg, first=True)
    st = eng.ebRunNode(p, dni, dif)
    assert st.converged
    s = eng.ebGetState()
    # shortwave: oracle multi-band solve on the same graph, sunlit and optics
    r_tab, t_tab = faceteb.material_tables(["leaf", "soil", "building"], win)
    mat = np.repeat(sc.material, 2)
    r = r_tab[mat]
    t = np.minimum(np.clip(t_tab[mat], 0, 1 - r), 1 - r)
    sun = faceteb.sun_vector(met.zenith, met.azimuth)
    cos = np.maximum(0.0, optics.side_normals(sc) @ sun).astype(np.float32)
    D = (np.clip(s["sunlit"], 0, 1) * cos)[:, None] * dni[None, :]
    sw, _, _ = oracle.solve_bands(og, r, t, np.zeros_like(r), D, dif, tol_rel=1e-9)
    # radiosity returned per window in window order
    assert _rel(s["shortwaveRadiosity"][sw > 1e-3], sw[sw > 1e-3]) < RTOL
    H = oracle.incident_bands(og, sw, D, dif)
    absorbed = np.sum(np.clip(1 - r - t, 0, 1) * H, axis=1)
    assert _rel(s["absorbedShortwave"][absorbed > 1], absorbed[absorbed > 1]) < RTOL
    assert st.swSweeps > 0 and st.lwFinalSweeps > 0 and 0 <= st.swMaxDelta < 1 and 0 <= st.lwMaxDelta < 1
    _check_diffuse_solve(oracle, eng, og, r, t, dif, s)
    # coupled energy balance in FP64 with the oracle from the same start
    mats = []
    for m in (faceteb.LEAF, faceteb.SOIL, faceteb.BUILDING):
        om = oracle.OrEbMaterial()
        for k, v in m.items():
            setattr(om, k, v)
        mats.append(om)
    op = oracle.OrEbParams()
    op.airTemperatureK = met.air_temperature_k
    op.vaporPressure = met.vapor_pressure
    op.pressure = met.pressure
    op.windSpeed = met.wind
    op.skyLongwave = met.longwave
    op.dTime = cfg.dtime
    op.tolerance = 1e-3
    op.relaxation = cfg.relaxation
    op.maxCoupling = 200
    op.lwMaxSweeps = 200
    op.lwTolRel = 1e-9
    T0 = np.full(sc.surface_count, met.air_temperature_k, np.float32)
    T, Blw, it, conv, _, _ = oracle.eb_node(og, sc.facet_count, sc.leaf_count, sc.material, mats,
                                            s["absorbedShortwave"], op, T0, T0)
    assert conv
    assert np.max(np.abs(s["temperature"] - T)) < 0.01
    assert _rel(s["longwaveRadiosity"], Blw) < RTOL


# Reflectance-variance texture factors (core:848-851, 965-972) on the
# energy balance's windows: per facet r = clamp(r_material * factor, 0, 1),
# t limited to 1 - r, on the per-side optics path; against the oracle's
# solve with those optics; removing the factors restores the plain node
@pytest.mark.parametrize("nb", [4, 24])
def test_eb_node_texture_factors_match_oracle(oracle, nb):
    print(f"eb texture factors nb={nb}", flush=True)
    sc, eng, win = _eb_case(oracle, nb=nb)
    og = _as_oracle_graph(oracle, eng.getGraph())
    rng = np.random.default_rng(11)
    factor = rng.uniform(0.5, 1.6, sc.facet_count).astype(np.float32)
    factor[:5] = 0.0
    eng.ebSetReflectanceFactors(factor)
    with pytest.raises(ValueError):
        eng.ebSetReflectanceFactors(factor[:-1])
    cfg = faceteb.EbConfig(tolerance=1e-3, radiosity_tolerance=1e-7, max_coupling=200)
    met = faceteb.diurnal_meteo(24)[12]
    dni, dif = win.irradiance(met.shortwave, math.cos(math.radians(met.zenith)))
    st = eng.ebRunNode(faceteb.node_params(met, cfg, first=True), dni, dif)
    assert st.converged
    s = eng.ebGetState()
    r_tab, t_tab = faceteb.material_tables(["leaf", "soil", "building"], win)
    mat = np.repeat(sc.material, 2)
    r = np.clip(r_tab[mat] * np.repeat(factor, 2)[:, None], 0, 1)
    t = np.minimum(np.clip(t_tab[mat], 0, 1), 1 - r)
    ground = np.repeat(np.arange(sc.facet_count) >= sc.leaf_count, 2)
    r[ground] = np.minimum(r[ground], 0.999)
    t = np.minimum(t, 1 - r)
    sun = faceteb.sun_vector(met.zenith, met.azimuth)
    cos = np.maximum(0.0, optics.side_normals(sc) @ sun).astype(np.float32)
    D = (np.clip(s["sunlit"], 0, 1) * cos)[:, None] * dni[None, :]
    sw, _, _ = oracle.solve_bands(og, r, t, np.zeros_like(r), D, dif, tol_rel=1e-9)
    assert _rel(s["shortwaveRadiosity"][sw > 1e-3], sw[sw > 1e-3]) < RTOL
    H = oracle.incident_bands(og, sw, D, dif)
    absorbed = np.sum(np.clip(1 - r - t, 0, 1) * H, axis=1)
    assert _rel(s["absorbedShortwave"][absorbed > 1], absorbed[absorbed > 1]) < RTOL
    _check_diffuse_solve(oracle, eng, og, r, t, dif, s)
    # without the factors: the material's own reflectance again
    eng.ebSetReflectanceFactors(None)
    eng.ebRunNode(faceteb.node_params(met, cfg, first=True), dni, dif)
    plain = eng.ebGetState()["shortwaveRadiosity"]
    r0 = r_tab[mat].copy()
    r0[ground] = np.minimum(r0[ground], 0.999)
    t0 = np.minimum(np.clip(t_tab[mat], 0, 1), 1 - r0)
    sw0, _, _ = oracle.solve_bands(og, r0, t0, np.zeros_like(r0), D, dif, tol_rel=1e-9)
    assert _rel(plain[sw0 > 1e-3], sw0[sw0 > 1e-3]) < RTOL


# Hapke ground (core:546-605) at strength 0.7: the ground facets' normals
# tilted so that |n_y| < 1 (the longwave emissivity then varies); shortwave
# reflectance per ground side and band from the sun, longwave emissivity
# per ground facet, against the oracle's restatement
@pytest.mark.parametrize("nb", [4, 24])
def test_eb_node_hapke_ground_matches_oracle(oracle, nb):
    print(f"eb hapke case nb={nb}", flush=True)
    strength = 0.7
    sc = scenes.row_crop(leaves_per_row=40, cells=6)
    v = np.array(sc.vertices, np.float32, copy=True)
    ground = np.arange(sc.leaf_count, sc.facet_count)
    n = np.stack([np.where(ground % 2 == 0, 0.35, -0.5), np.ones(ground.size), np.full(ground.size, 0.2)], 1)
    n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    for k in range(3):
        v[3 * ground + k, 3:6] = n
    sc.vertices = v
    eng = _engine(R=128, K=32, seed=4)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    win = faceteb.ShortwaveWindows.for_bands(nb)
    faceteb.setup_scene(eng, sc, win)
    eng.ebSetBackgroundAngularStrength(strength)
    og = _as_oracle_graph(oracle, eng.getGraph())
    cfg = faceteb.EbConfig(tolerance=1e-3, radiosity_tolerance=1e-7, max_coupling=200)
    met = faceteb.diurnal_meteo(24)[10]
    dni, dif = win.irradiance(met.shortwave, math.cos(math.radians(met.zenith)))
    st = eng.ebRunNode(faceteb.node_params(met, cfg, first=True), dni, dif)
    assert st.converged
    s = eng.ebGetState()
    r_tab, t_tab = faceteb.material_tables(["leaf", "soil", "building"], win)
    mat = np.repeat(sc.material, 2)
    r = r_tab[mat].copy()
    t = np.minimum(np.clip(t_tab[mat], 0, 1 - r), 1 - r)
    sun = faceteb.sun_vector(met.zenith, met.azimuth)
    sn = optics.side_normals(sc)
    for side in range(2 * sc.leaf_count, sc.surface_count):
        for b in range(nb):
            r[side, b] = oracle.hapke_reflectance(r[side, b], sun, sn[side], sn[side], strength)
    t = np.minimum(t, 1 - r)
    assert np.ptp(r[2 * sc.leaf_count:, 0]) > 1e-3  # the sides' reflectance follows their normal
    cos = np.maximum(0.0, sn @ sun).astype(np.float32)
    D = (np.clip(s["sunlit"], 0, 1) * cos)[:, None] * dni[None, :]
    sw, _, _ = oracle.solve_bands(og, r, t, np.zeros_like(r), D, dif, tol_rel=1e-9)
    assert _rel(s["shortwaveRadiosity"][sw > 1e-3], sw[sw > 1e-3]) < RTOL
    H = oracle.incident_bands(og, sw, D, dif)
    absorbed = np.sum(np.clip(1 - r - t, 0, 1) * H, axis=1)
    assert _rel(s["absorbedShortwave"][absorbed > 1], absorbed[absorbed > 1]) < RTOL
    _check_diffuse_solve(oracle, eng, og, r, t, dif, s)
    mats = []
    for m in (faceteb.LEAF, faceteb.SOIL, faceteb.BUILDING):
        om = oracle.OrEbMaterial()
        for k, val in m.items():
            setattr(om, k, val)
        mats.append(om)
    op = oracle.OrEbParams()
    op.airTemperatureK = met.air_temperature_k
    op.vaporPressure = met.vapor_pressure
    op.pressure = met.pressure
    op.windSpeed = met.wind
    op.skyLongwave = met.longwave
    op.dTime = cfg.dtime
    op.tolerance = 1e-3
    op.relaxation = cfg.relaxation
    op.maxCoupling = 200
    op.lwMaxSweeps = 200
    op.lwTolRel = 1e-9
    op.backgroundAngularStrength = strength
    T0 = np.full(sc.surface_count, met.air_temperature_k, np.float32)
    T, Blw, it, conv, _, _ = oracle.eb_node(og, sc.facet_count, sc.leaf_count, sc.material, mats,
                                            s["absorbedShortwave"], op, T0, T0,
                                            normal_y=sc.facet_normals()[:, 1])
    assert conv
    assert np.max(np.abs(s["temperature"] - T)) < 0.01
    assert _rel(s["longwaveRadiosity"], Blw) < RTOL
    # strength 0 again: the Lambertian ground's longwave radiosity differs
    op.backgroundAngularStrength = 0.0
    _, Blw0, _, _, _, _ = oracle.eb_node(og, sc.facet_count, sc.leaf_count, sc.material, mats,
                                         s["absorbedShortwave"], op, T0, T0)
    assert _rel(Blw0, Blw) > 1e-4


@pytest.mark.parametrize("nb", [4, 24])
def test_eb_node_photovoltaic_matches_oracle(oracle, nb):
    # photovoltaic object facets (core:3655-3663, 3772-3778): the leaves as
    # bifacial panels; the coupled balance with their electric power against
    # the FP64 oracle given the same incident and absorbed shortwave
    print(f"eb photovoltaic case nb={nb}", flush=True)
    sc, eng, win = _eb_case(oracle, nb=nb)
    pv_leaf = dict(faceteb.LEAF, vegetation=0, heatCapacityPerArea=12000.0, photovoltaic=1, pvEta25=0.22,
                   pvGamma=-0.0035, pvBifaciality=0.3)
    r_tab, t_tab = faceteb.material_tables(["leaf", "soil", "building"], win)
    eng.ebSetScene(sc.leaf_count, sc.material, [pv_leaf, faceteb.SOIL, faceteb.BUILDING], r_tab, t_tab)
    og = _as_oracle_graph(oracle, eng.getGraph())
    cfg = faceteb.EbConfig(tolerance=1e-3, radiosity_tolerance=1e-7, max_coupling=200)
    met = faceteb.diurnal_meteo(24)[12]
    dni, dif = win.irradiance(met.shortwave, math.cos(math.radians(met.zenith)))
    st = eng.ebRunNode(faceteb.node_params(met, cfg, first=True), dni, dif)
    assert st.converged
    s = eng.ebGetState()
    pv = eng.ebGetPhotovoltaic()
    np.testing.assert_array_equal(pv["absorbedShortwave"], s["absorbedShortwave"])
    mats = []
    for m in (pv_leaf, faceteb.SOIL, faceteb.BUILDING):
        om = oracle.OrEbMaterial()
        for k, v in m.items():
            setattr(om, k, v)
        mats.append(om)
    op = oracle.OrEbParams()
    op.airTemperatureK = met.air_temperature_k
    op.vaporPressure = met.vapor_pressure
    op.pressure = met.pressure
    op.windSpeed = met.wind
    op.skyLongwave = met.longwave
    op.dTime = cfg.dtime
    op.tolerance = 1e-3
    op.relaxation = cfg.relaxation
    op.maxCoupling = 200
    op.lwMaxSweeps = 200
    op.lwTolRel = 1e-9
    T0 = np.full(sc.surface_count, met.air_temperature_k, np.float32)
    T, Blw, it, conv, _, _ = oracle.eb_node(og, sc.facet_count, sc.leaf_count, sc.material, mats,
                                            s["absorbedShortwave"], op, T0, T0, incident=pv["incidentShortwave"])
    assert conv
    assert np.max(np.abs(s["temperature"] - T)) < 0.01
    assert _rel(s["longwaveRadiosity"], Blw) < RTOL
    # power of every leaf: the model at the oracle's temperatures
    L = sc.leaf_count
    inc = pv["incidentShortwave"].astype(np.float64).reshape(-1, 2)[:L]
    ab = pv["absorbedShortwave"].astype(np.float64).reshape(-1, 2)[:L]
    Tl = T.astype(np.float64).reshape(-1, 2)[:L, 0]
    want = np.clip(0.22 * (1 - 0.0035 * (Tl - 298.15)) * (inc[:, 0] + 0.3 * inc[:, 1]), 0, ab.sum(axis=1))
    got = pv["power"].reshape(-1, 2)
    np.testing.assert_array_equal(got[:L, 0], got[:L, 1])
    assert np.all(got[L:] == 0)
    assert want.max() > 10.0
    np.testing.assert_allclose(got[:L, 0], want, rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("nb", [4, 24, 47])
def test_eb_diurnal_plain_vs_bf16(oracle, nb):
    # the BF16 path reaches the same converged state (shortwave BF16
    # increments from 33 windows on; FP32 below)
    print(f"eb diurnal case nb={nb}", flush=True)
    sc, eng, win = _eb_case(oracle, leaves=30, cells=5, R=64, K=16, nb=nb)
    met = faceteb.diurnal_meteo(6)
    cfg = faceteb.EbConfig(tolerance=0.01)
    s1 = faceteb.run_diurnal(eng, met, win, cfg, precision=FH_PRECISION_FP32)
    T1 = eng.ebGetState()["temperature"]
    s2 = faceteb.run_diurnal(eng, met, win, cfg, precision=FH_PRECISION_BF16_DC)
    T2 = eng.ebGetState()["temperature"]
    assert all(s1.converged) and all(s2.converged)
    assert np.max(np.abs(T1 - T2)) < 0.02
    # FP8 increments (from 33 windows on) reach the same state too
    s3 = faceteb.run_diurnal(eng, met, win, cfg, precision=FH_PRECISION_FP8_INC)
    T3 = eng.ebGetState()["temperature"]
    assert all(s3.converged)
    assert np.max(np.abs(T1 - T3)) < 0.02
    # BF16 F with FP32 defect correction (from 33 windows on) as well
    s4 = faceteb.run_diurnal(eng, met, win, cfg, precision=FH_PRECISION_BF16_F)
    T4 = eng.ebGetState()["temperature"]
    assert all(s4.converged)
    assert np.max(np.abs(T1 - T4)) < 0.02


def test_mlp_gradient_matches_oracle(oracle):
    print("mlp gradient case", flush=True)
    eng = _engine()
    rng = np.random.default_rng(8)
    H = 32
    n_par = eng.mlpParamCount(H)
    P = rng.normal(0, 0.3, n_par).astype(np.float32)
    eng.mlpSetParams(H, P)
    x = rng.normal(0, 1, (64, 9)).astype(np.float32)
    t = rng.normal(0, 1, 64).astype(np.float32)
    loss, g = eng.mlpLossGrad(x, t)
    oloss, og = oracle.mlp_loss_grad(P, x, t, H)
    assert loss == pytest.approx(oloss, rel=1e-5)
    np.testing.assert_allclose(g, og, rtol=1e-3, atol=1e-5 * np.max(np.abs(og)))
    y = eng.mlpForward(x)
    np.testing.assert_allclose(y, oracle.mlp_forward(P, x, H), rtol=1e-4, atol=1e-5)


def test_mlp_forward_width128_matrix_cores(oracle):
    # width 128 runs layer 2 on the matrix cores (v_mfma_f32_32x32x2f32):
    # against the FP64-accumulating oracle on a unit count that leaves a
    # partial block
    print("mlp forward width 128 case", flush=True)
    eng = _engine()
    rng = np.random.default_rng(11)
    H = 128
    P = rng.normal(0, 0.15, eng.mlpParamCount(H)).astype(np.float32)
    eng.mlpSetParams(H, P)
    x = rng.normal(0, 1, (1000, 9)).astype(np.float32)
    y = eng.mlpForward(x)
    ref = oracle.mlp_forward(P, x, H)
    np.testing.assert_allclose(y, ref, rtol=1e-4, atol=1e-4 * np.max(np.abs(ref)))


def test_surrogate_warm_start_gain():
    # the committed network (trained on the device by fh_mlp_train on two
    # forests, the urban block and a terrain patch) warm-starts a held-out
    # forest through a held-out diurnal driver: the coupled iteration still
    # runs to 0.01 K, so the converged states agree with the plain run, and
    # it saves coupling iterations (131 -> 123 when the network was trained)
    print("surrogate gain case", flush=True)
    from pathlib import Path
    from streamsim_amd._lib import FH_WARM_PREVIOUS, FH_WARM_SURROGATE
    path = Path(__file__).resolve().parents[1] / "streamsim_amd" / "data" / "surrogate.npz"
    z = np.load(path, allow_pickle=False)
    sc = scenes.forest(seed=77)
    eng = _engine(R=1024, K=16, seed=1)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    win = faceteb.ShortwaveWindows.for_bands(8)
    faceteb.setup_scene(eng, sc, win)
    eng.mlpSetParams(int(z["hidden"]), z["params"], z["feature_mean"], z["feature_std"], float(z["target_mean"]),
                     float(z["target_std"]))
    met = faceteb.diurnal_meteo(24, t_mean=295.0, t_amp=7.0, wind=3.0, day_of_year=200)
    cfg = faceteb.EbConfig()
    s0 = faceteb.run_diurnal(eng, met, win, cfg, warm=FH_WARM_PREVIOUS)
    T0 = eng.ebGetState()["temperature"]
    s1 = faceteb.run_diurnal(eng, met, win, cfg, warm=FH_WARM_SURROGATE)
    T1 = eng.ebGetState()["temperature"]
    assert all(s0.converged) and all(s1.converged)
    assert np.max(np.abs(T0 - T1)) < 0.05
    print(f"coupling iterations plain {sum(s0.coupling)} surrogate {sum(s1.coupling)}", flush=True)
    assert sum(s1.coupling) < sum(s0.coupling)


def test_mlp_gradient_width128_matches_oracle(oracle):
    # the shipped width: the device backward pass (its layer-2 products and
    # reductions over a batch that leaves partial blocks) against the FP64 oracle
    print("mlp gradient width 128 case", flush=True)
    eng = _engine()
    rng = np.random.default_rng(12)
    H = 128
    P = rng.normal(0, 0.1, eng.mlpParamCount(H)).astype(np.float32)
    eng.mlpSetParams(H, P)
    x = rng.normal(0, 1, (1000, 9)).astype(np.float32)
    t = rng.normal(0, 1, 1000).astype(np.float32)
    loss, g = eng.mlpLossGrad(x, t)
    oloss, og = oracle.mlp_loss_grad(P, x, t, H)
    assert loss == pytest.approx(oloss, rel=1e-5)
    np.testing.assert_allclose(g, og, rtol=1e-3, atol=1e-5 * np.max(np.abs(og)))


def _facing_pairs_scene(pairs=120, gap=0.02, side=0.3, seed=11):
    """Leaves in parallel, facing pairs `gap` apart over a ground patch: the
    inner sides of a pair see each other with weights near 0.8."""
    rng = np.random.default_rng(seed)
    c = rng.uniform([0.5, 0.6, 0.5], [3.5, 2.0, 3.5], (pairs, 3))
    n = scenes.sphere_normals(pairs, np.random.default_rng(seed + 1))
    a = scenes.equilateral_leaves(c, n, side, np.random.default_rng(seed + 2))
    b = scenes.equilateral_leaves(c + gap * n, n, side, np.random.default_rng(seed + 2))  # same rotation
    leaves = np.concatenate([a, b])
    return scenes._finish("pairs", leaves, scenes.grid_ground(0.0, 4.0, 0.0, 4.0, 8, 8))


@pytest.mark.parametrize("nb", [24, 47])
def test_pair_blocks_match_oracle(oracle, nb):
    # strongly coupled facet pairs are solved as 2-facet blocks inside the
    # band-parallel sweeps: same fixed point as the oracle's Jacobi solve,
    # far fewer sweeps than Jacobi needs for the pairs' slow modes
    print(f"pair blocks nb={nb}", flush=True)
    sc = _facing_pairs_scene()
    eng = _engine(R=128, seed=5)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    g = eng.getGraph()
    og = _as_oracle_graph(oracle, g)
    # precondition: most leaves have a partner facet holding > 0.2 of a
    # side's row (the inner sides of the pairs)
    S, F = sc.surface_count, sc.facet_count
    rows = np.repeat(np.arange(S), np.diff(g.rowOffsets.astype(np.int64)))
    w = g.counts / np.maximum(g.denominators[rows], 1)
    key = rows * F + g.neighbors // 2
    uk, inv = np.unique(key, return_inverse=True)
    strongest = np.zeros(S)
    np.maximum.at(strongest, uk // F, np.bincount(inv, weights=w))
    assert np.mean(strongest.reshape(-1, 2).max(axis=1)[: sc.leaf_count] > 0.2) > 0.8
    rng = np.random.default_rng(nb)
    r = rng.uniform(0.35, 0.5, (S, nb)).astype(np.float32)
    t = (0.95 - r) * rng.uniform(0.8, 1.0, (S, nb)).astype(np.float32)
    e = rng.uniform(0.0, 5.0, (S, nb)).astype(np.float32)
    D = rng.uniform(0.0, 300.0, (S, nb)).astype(np.float32)
    sky = rng.uniform(10.0, 80.0, nb).astype(np.float32)
    ref, _, _ = oracle.solve_bands(og, r, t, e, D, sky, tol_rel=1e-9)
    _, it_jacobi, _ = oracle.solve_bands(og, r, t, e, D, sky, tol_rel=1e-7)
    for prec in (FH_PRECISION_FP32, FH_PRECISION_BF16_DC, FH_PRECISION_FP8_INC):
        res = eng.solveBands(r, t, e, D, sky, tolRel=1e-7, precision=prec)
        print(f"  precision {prec}: {res.iterations} sweeps, Jacobi {it_jacobi}", flush=True)
        assert _rel(res.radiosity, ref) < RTOL, prec
        assert res.iterations < 0.85 * it_jacobi, prec


@pytest.mark.parametrize("nb", [5, 12, 32, 47, 77])
def test_many_bands_match_oracle(oracle, nb):
    # every band layout of the SpMM kernels (band stride 8, 12, 32, 48 and a
    # partial 64-band chunk at 80) against the FP64 oracle
    print(f"many bands nb={nb}", flush=True)
    sc = scenes.row_crop(leaves_per_row=60, cells=8)
    eng = _engine(R=64, seed=6)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    og = _as_oracle_graph(oracle, eng.getGraph())
    rng = np.random.default_rng(nb)
    S = sc.surface_count
    r = rng.uniform(0.0, 0.5, (S, nb)).astype(np.float32)
    t = (rng.uniform(0.0, 1.0, (S, nb)) * (1.0 - r) * 0.8).astype(np.float32)
    e = rng.uniform(0.0, 5.0, (S, nb)).astype(np.float32)
    D = rng.uniform(0.0, 300.0, (S, nb)).astype(np.float32)
    sky = rng.uniform(10.0, 80.0, nb).astype(np.float32)
    ref, _, _ = oracle.solve_bands(og, r, t, e, D, sky, tol_rel=1e-9)
    for prec in (FH_PRECISION_FP32, FH_PRECISION_BF16_DC):
        res = eng.solveBands(r, t, e, D, sky, tolRel=1e-7, precision=prec)
        assert _rel(res.radiosity, ref) < RTOL, prec
    H = eng.incidentBands(ref, D, sky)
    assert _rel(H, oracle.incident_bands(og, ref, D, sky)) < RTOL


@pytest.mark.parametrize("nb", [24, 47])
def test_bf16_increments_follow_fp32(oracle, nb):
    # BF16 solves at band-parallel strides gather BF16 increments of B and
    # carry what rounding drops into the next increment: from the same warm
    # start they take the FP32 solve's sweeps and reach its fixed point (their
    # stopping test resolves relative changes down to 3e-7, 2.5 FP32 ulps)
    print(f"bf16 increments nb={nb}", flush=True)
    sc = scenes.row_crop(leaves_per_row=60, cells=8)
    eng = _engine(R=64, seed=9)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    og = _as_oracle_graph(oracle, eng.getGraph())
    rng = np.random.default_rng(100 + nb)
    S = sc.surface_count
    r = rng.uniform(0.0, 0.5, (S, nb)).astype(np.float32)
    t = (rng.uniform(0.0, 1.0, (S, nb)) * (1.0 - r) * 0.8).astype(np.float32)
    e = np.zeros((S, nb), np.float32)
    D = rng.uniform(0.0, 300.0, (S, nb)).astype(np.float32)
    sky = rng.uniform(10.0, 80.0, nb).astype(np.float32)
    ref, _, _ = oracle.solve_bands(og, r, t, e, D, sky, tol_rel=1e-9)
    init = (ref * rng.uniform(0.5, 1.5, ref.shape)).astype(np.float32)
    f32 = eng.solveBands(r, t, e, D, sky, init=init, tolRel=1e-7, precision=FH_PRECISION_FP32)
    b16 = eng.solveBands(r, t, e, D, sky, init=init, tolRel=1e-7, precision=FH_PRECISION_BF16_DC)
    assert f32.iterations - 2 <= b16.iterations <= f32.iterations, (b16.iterations, f32.iterations)
    assert _rel(b16.radiosity, f32.radiosity) < 2e-6
    assert _rel(b16.radiosity, ref) < RTOL
    # FP8 increments (12-band chunks, shared exponent): the same fixed point,
    # at most two more sweeps than FP32 (each increment loses up to 1/16 to
    # rounding, sent in the next sweep)
    q8 = eng.solveBands(r, t, e, D, sky, init=init, tolRel=1e-7, precision=FH_PRECISION_FP8_INC)
    print(f"  sweeps: fp32 {f32.iterations}, bf16 {b16.iterations}, fp8 {q8.iterations}", flush=True)
    assert f32.iterations - 2 <= q8.iterations <= f32.iterations + 2, (q8.iterations, f32.iterations)
    assert _rel(q8.radiosity, f32.radiosity) < 2e-6
    assert _rel(q8.radiosity, ref) < RTOL
    # from a zero start too (the first sweep gathers the FP8 rounding of 0)
    z32 = eng.solveBands(r, t, e, D, sky, tolRel=1e-7, precision=FH_PRECISION_FP32)
    z8 = eng.solveBands(r, t, e, D, sky, tolRel=1e-7, precision=FH_PRECISION_FP8_INC)
    assert z8.iterations <= z32.iterations + 2 and _rel(z8.radiosity, ref) < RTOL
    # BF16 F and a BF16 copy of B, then FP32 sweeps (defect correction): the
    # FP32 fixed point, after the BF16 phase's sweeps
    bf = eng.solveBands(r, t, e, D, sky, init=init, tolRel=1e-7, precision=FH_PRECISION_BF16_F)
    print(f"  bf16 F + defect correction: {bf.iterations} sweeps", flush=True)
    assert _rel(bf.radiosity, ref) < RTOL and _rel(bf.radiosity, f32.radiosity) < 2e-6


@pytest.mark.parametrize("scene,R,nb", [("forest", 256, 47), ("row_crop", 64, 40)])
def test_staged_gathers_match_l1_gathers(oracle, monkeypatch, scene, R, nb):
    # The increment sweeps of the single facets gather each workgroup's
    # distinct columns once through the LDS (chunk-major copy of the CSR);
    # FH_STAGED=0 gathers every entry through the vector L1.  Only the order
    # of the FP32 sums differs: three sweeps from the same start agree to a
    # few FP32 ulps, and converged solves take the same sweeps to the same
    # fixed point.
    print(f"staged gathers {scene} R={R} nb={nb}", flush=True)
    sc = scenes.forest() if scene == "forest" else scenes.row_crop(leaves_per_row=200, cells=20)
    eng = _engine(R=R, seed=3)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    rng = np.random.default_rng(7 + nb)
    S = sc.surface_count
    r = rng.uniform(0.0, 0.5, (S, nb)).astype(np.float32)
    t = (rng.uniform(0.0, 1.0, (S, nb)) * (1.0 - r) * 0.8).astype(np.float32)
    e = np.zeros((S, nb), np.float32)
    D = rng.uniform(0.0, 300.0, (S, nb)).astype(np.float32)
    sky = rng.uniform(10.0, 80.0, nb).astype(np.float32)
    init = rng.uniform(10.0, 200.0, (S, nb)).astype(np.float32)
    out = {}
    for staged in ("1", "0"):
        monkeypatch.setenv("FH_STAGED", staged)
        few = eng.solveBands(r, t, e, D, sky, init=init, maxIter=3, tolRel=0.0, precision=FH_PRECISION_BF16_DC)
        full = eng.solveBands(r, t, e, D, sky, init=init, tolRel=1e-7, precision=FH_PRECISION_BF16_DC)
        out[staged] = (few, full)
    (f1, c1), (f0, c0) = out["1"], out["0"]
    print(f"  sweeps staged {c1.iterations}, L1 {c0.iterations}", flush=True)
    assert f1.iterations == f0.iterations == 3
    assert _rel(f1.radiosity, f0.radiosity) < 1e-6
    assert abs(c1.iterations - c0.iterations) <= 1
    assert _rel(c1.radiosity, c0.radiosity) < 2e-6
    if scene == "row_crop":
        og = _as_oracle_graph(oracle, eng.getGraph())
        ref, _, _ = oracle.solve_bands(og, r, t, e, D, sky, tol_rel=1e-9)
        assert _rel(c1.radiosity, ref) < RTOL
    eng.destroy()


@pytest.mark.parametrize("scale", [64, 70000])
def test_staged_gathers_split_large_counts(oracle, monkeypatch, scale):
    # The staged layout holds ray counts of at most 63 per entry (6 bits in
    # the 2-byte keys, exact in BF16); a larger count is split into parts of
    # the same slot.  An uploaded graph with every count scaled up (64 c + a
    # random 0..63: every entry split) must sweep as the L1 gathers and the
    # FP64 oracle do; at scale 70000 the parts overflow the segment limit
    # (255 groups of 4), the layout is left unbuilt and the L1 gathers run
    # (bit for bit the same sweeps with FH_STAGED 1 and 0, where the staged
    # sums, in another order, differ in the last bits).
    print(f"staged gathers, scaled counts x{scale}", flush=True)
    sc = scenes.row_crop(leaves_per_row=120, cells=12)
    eng = _engine(R=64, seed=5)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    g = eng.getGraph()
    rng = np.random.default_rng(scale)
    cnt = g.counts.astype(np.int64) * scale + rng.integers(0, scale, g.counts.size)
    sky = g.skyCounts.astype(np.int64) * scale
    rows = np.repeat(np.arange(g.denominators.size), np.diff(g.rowOffsets.astype(np.int64)))
    den = np.bincount(rows, weights=cnt, minlength=g.denominators.size).astype(np.int64) + sky
    assert np.all(cnt > 63)
    d = eng.setGraph(g.rowOffsets, g.neighbors, cnt.astype(np.uint32), den.astype(np.uint32), sky.astype(np.uint32))
    assert d.maxClosureError == 0.0
    nb = 40
    S = sc.surface_count
    r = rng.uniform(0.0, 0.5, (S, nb)).astype(np.float32)
    t = (rng.uniform(0.0, 1.0, (S, nb)) * (1.0 - r) * 0.8).astype(np.float32)
    e = np.zeros((S, nb), np.float32)
    D = rng.uniform(0.0, 300.0, (S, nb)).astype(np.float32)
    skyv = rng.uniform(10.0, 80.0, nb).astype(np.float32)
    init = rng.uniform(10.0, 200.0, (S, nb)).astype(np.float32)
    out = {}
    for staged in ("1", "0"):
        monkeypatch.setenv("FH_STAGED", staged)
        few = eng.solveBands(r, t, e, D, skyv, init=init, maxIter=3, tolRel=0.0, precision=FH_PRECISION_BF16_DC)
        full = eng.solveBands(r, t, e, D, skyv, init=init, tolRel=1e-7, precision=FH_PRECISION_BF16_DC)
        out[staged] = (few, full)
    (f1, c1), (f0, c0) = out["1"], out["0"]
    print(f"  sweeps staged {c1.iterations}, L1 {c0.iterations}", flush=True)
    # the sums round differently (count x value summed, then times 1 / den,
    # against FP32 weights count / den): a few ulps per sweep, amplified at a
    # handful of strongly coupled facets (5 values above 1e-6 of 61 k after
    # three sweeps)
    d = np.abs(f1.radiosity.astype(np.float64) - f0.radiosity) / np.abs(f0.radiosity.astype(np.float64))
    assert np.quantile(d, 0.999) < 1e-6 and d.max() < 2e-5
    assert np.array_equal(f1.radiosity, f0.radiosity) == (scale > 65535)
    assert abs(c1.iterations - c0.iterations) <= 1
    og = _as_oracle_graph(oracle, eng.getGraph())
    ref, _, _ = oracle.solve_bands(og, r, t, e, D, skyv, tol_rel=1e-9)
    assert _rel(c1.radiosity, ref) < RTOL
    assert _rel(c0.radiosity, ref) < RTOL
    eng.destroy()


def test_sw_increment_state_continues_across_nodes(oracle, monkeypatch):
    # From the previous node's final radiosity, a node's BF16 increment
    # solve continues that node's increment state (accumulated gathers,
    # carries, the increments written last folded into the carries) with a
    # first sweep that gathers nothing, instead of a base sweep gathering
    # bf16(B_0).  Both reach the FP32 fixed point of the node's operator: the
    # same sweeps (+-1) and radiosities within the solve tolerance; the first
    # node (no state yet) is bit for bit the same, the continued ones not.
    print("sw increment continuation", flush=True)
    sc, eng, win = _eb_case(oracle, leaves=120, cells=12, R=64, K=16, nb=47)
    met = faceteb.diurnal_meteo(12)[4:8]  # 09:00 .. 15:00
    cfg = faceteb.EbConfig(tolerance=0.01)
    out = {}
    for cont in ("1", "0"):
        monkeypatch.setenv("FH_SW_CONTINUE", cont)
        faceteb.setup_scene(eng, sc, win)  # the same cold start for both runs
        caps = []
        st = faceteb.run_diurnal(eng, met, win, cfg, precision=FH_PRECISION_BF16_DC,
                                 record=lambda e, k: caps.append(e.ebGetState()["shortwaveRadiosity"].copy()))
        out[cont] = (st, caps)
    (s1, b1), (s0, b0) = out["1"], out["0"]
    print(f"  sweeps continued {s1.sw_sweeps}, base {s0.sw_sweeps}", flush=True)
    assert all(s1.converged) and all(s0.converged)
    assert all(abs(x - y) <= 1 for x, y in zip(s1.sw_sweeps, s0.sw_sweeps))
    assert np.array_equal(b1[0], b0[0])
    for x, y in zip(b1[1:], b0[1:]):
        assert not np.array_equal(x, y)
        big = np.abs(y) > 1e-3 * np.abs(y).max()
        assert _rel(x[big], y[big]) < 5e-5
    eng.destroy()


def test_staged_single_band_sweeps_match_team_kernel(oracle, monkeypatch):
    # Single-band FP32 solves sweep the staged facets with k_sweep1_staged
    # (the coupled pairs' facets by the team kernel beside it); FH_STAGED=0
    # sweeps every facet with the team kernel.  Only the order and rounding
    # of the sums differ (count x value summed, times 1 / den): 64 fixed
    # sweeps agree to a few ulps, and both match the FP64 oracle.
    print("staged single-band sweeps", flush=True)
    sc = scenes.row_crop(leaves_per_row=120, cells=12)
    eng = _engine(R=64, seed=5)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    sun = oracle.sun_direction(35.0, 120.0)
    o = optics.make_optics(sc, eng.computeSunlitFraction(sun), sun)
    out = {}
    for staged in ("1", "0"):
        monkeypatch.setenv("FH_STAGED", staged)
        out[staged] = eng.solve(o, 100.0, 64, 0.9)
    r1, r0 = out["1"], out["0"]
    assert r1.iterations == r0.iterations == 64
    assert not np.array_equal(r1.radiosity, r0.radiosity)
    assert _rel(r1.radiosity, r0.radiosity) < 1e-6
    og = _as_oracle_graph(oracle, eng.getGraph())
    ref, _ = oracle.solve(og, o, 100.0, 64, 0.9)
    assert _rel(r1.radiosity, ref) < RTOL
    eng.destroy()


def test_staged_coupled_gather_matches_team_gather(oracle, monkeypatch):
    # The coupled longwave sweep gathers each workgroup's distinct columns
    # once through the LDS (staged layout; the pair blocks' facets by the
    # team kernel); FH_STAGED=0 runs the team kernel over every facet.  Only
    # summation order and the weight's last bit (count * (1 / den)) differ:
    # both reach the converged state within the coupling tolerance.
    print("staged coupled gather", flush=True)
    sc, eng, win = _eb_case(oracle, leaves=120, cells=12, R=64, K=16, nb=47)
    met = faceteb.diurnal_meteo(6)
    cfg = faceteb.EbConfig(tolerance=0.01)
    out = {}
    for staged in ("1", "0"):
        monkeypatch.setenv("FH_STAGED", staged)
        faceteb.setup_scene(eng, sc, win)  # the same cold start for both runs
        st = faceteb.run_diurnal(eng, met, win, cfg, precision=FH_PRECISION_FP32)
        out[staged] = (st, eng.ebGetState()["temperature"])
    (s1, T1), (s0, T0) = out["1"], out["0"]
    print(f"  coupling iterations staged {s1.coupling}, team {s0.coupling}", flush=True)
    assert all(s1.converged) and all(s0.converged)
    assert all(abs(x - y) <= 1 for x, y in zip(s1.coupling, s0.coupling))
    assert np.max(np.abs(T1 - T0)) < 0.02
    eng.destroy()


class _Hip:
    """hipMalloc / hipMemcpy / hipFree through ctypes, for the device-pointer entry points."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.held = []

    def put(self, a):
        C = self.C
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        self.held.append(p)
        return p

    def get(self, p, shape):
        out = np.empty(shape, np.float32)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.held:
            self.lib.hipFree(p)


@pytest.mark.parametrize("nb", [1, 8, 47])
def test_device_pointer_entry_points(oracle, nb):
    # fh_sweep_device: one Jacobi sweep on device arrays [S][ldb] equals
    # c + r G_s + t G_o with G = F B from the graph's own weights;
    # fh_solve_bands_device equals fh_solve_bands on the same inputs
    print(f"device pointers nb={nb}", flush=True)
    import ctypes as C
    from streamsim_amd import _lib
    sc = scenes.row_crop(leaves_per_row=60, cells=8)
    eng = _engine(R=64, seed=2)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    g = eng.getGraph()
    S = sc.surface_count
    lib = _lib.load()
    ldb = int(lib.fh_band_stride(nb))
    rng = np.random.default_rng(nb)
    c = rng.uniform(0.0, 50.0, (S, ldb)).astype(np.float32)
    r = rng.uniform(0.0, 0.5, (S, ldb)).astype(np.float32)
    t = (rng.uniform(0.0, 0.4, (S, ldb))).astype(np.float32)
    B = rng.uniform(0.0, 80.0, (S, ldb)).astype(np.float32)
    hip = _Hip()
    try:
        dc, dr, dt, dB, dO = hip.put(c), hip.put(r), hip.put(t), hip.put(B), hip.put(np.zeros((S, ldb)))
        st = lib.fh_sweep_device(eng.handle, nb, dc, dr, dt, dB, dO, _lib.FH_PRECISION_FP32)
        assert st == 0
        out = hip.get(dO, (S, ldb))
        rows = np.repeat(np.arange(S), np.diff(g.rowOffsets))
        w = (g.counts / np.maximum(g.denominators[rows], 1)).astype(np.float32).astype(np.float64)
        G = np.zeros((S, ldb))
        np.add.at(G, rows, w[:, None] * B[g.neighbors].astype(np.float64))
        Go = G[np.arange(S) ^ 1]
        ref = c + r * G + t * Go
        np.testing.assert_allclose(out[:, :nb], ref[:, :nb], rtol=2e-5, atol=1e-3)
        # the multi-band solve from device arrays [S][nb] equals the host-array call
        rr = r[:, :nb].copy()
        tt = (t[:, :nb] * (1.0 - rr)).astype(np.float32)
        e = rng.uniform(0.0, 5.0, (S, nb)).astype(np.float32)
        D = rng.uniform(0.0, 300.0, (S, nb)).astype(np.float32)
        sky = rng.uniform(10.0, 80.0, nb).astype(np.float32)
        host = eng.solveBands(rr, tt, e, D, sky, tolRel=1e-6)
        dout = hip.put(np.zeros((S, nb)))
        stats = _lib.FhSolveStats()
        st = lib.fh_solve_bands_device(eng.handle, nb, hip.put(rr), hip.put(tt), hip.put(e), hip.put(D),
                                       sky.ctypes.data, None, 1000, 1e-6, 1.0,
                                       _lib.FH_PRECISION_FP32, dout, C.byref(stats))
        assert st == 0
        dev = hip.get(dout, (S, nb))
        np.testing.assert_array_equal(dev, np.asarray(host.radiosity).reshape(S, nb))
        assert stats.iterations == host.iterations
    finally:
        hip.free()



def test_benchmark_entry_points():
    # the timing entry points run in both forms and both precisions
    print("benchmark entry points", flush=True)
    sc = scenes.row_crop(leaves_per_row=20, cells=4)
    eng = _engine(R=32, seed=3)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    for nb in (1, 24, 47):
        for prec in (FH_PRECISION_FP32, FH_PRECISION_BF16_DC):
            for tables in (False, True):
                assert eng.benchmarkSweep(nb, prec, 3, tables=tables) > 0.0
    assert eng.benchmarkCopy(1 << 20, 3) > 0.0


# makeSpectralOptics / makeStateSpectralOptics (facetrtcore.cpp:1057-1199)
# built on the device against the same formulas on the host with the
# oracle's Hapke functions: optical and thermal bands (nm and micrometre
# wavelengths), texture factors, a Hapke ground on tilted facets, both
# modes; then fh_solve_spectral against fh_solve_bands on those arrays
@pytest.mark.parametrize("mode", [0, 1])
def test_spectral_optics_on_device(oracle, mode):
    print(f"spectral optics mode={mode}", flush=True)
    from streamsim_amd.optics import planck_radiance
    sc = scenes.row_crop(leaves_per_row=20, cells=4)
    v = np.array(sc.vertices, np.float32, copy=True)
    ground = np.arange(sc.leaf_count, sc.facet_count)
    n = np.stack([np.where(ground % 2 == 0, 0.3, -0.4), np.ones(ground.size), np.full(ground.size, 0.25)], 1)
    n = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
    for k in range(3):
        v[3 * ground + k, 3:6] = n
    sc.vertices = v
    eng = _engine(R=64, K=16, seed=3)
    eng.setGeometry(sc.vertices)
    eng.buildVisibilityGraph()
    F, S, L = sc.facet_count, sc.surface_count, sc.leaf_count
    rng = np.random.default_rng(5)
    wl = np.array([450.0, 0.55, 865.0, 2400.0, 2.6, 10500.0], np.float32)
    nb = wl.size
    nm = 3
    mat = np.asarray(sc.material, np.uint32)
    R = rng.uniform(0.02, 0.6, (nm, nb)).astype(np.float32)
    T = rng.uniform(0.0, 0.4, (nm, nb)).astype(np.float32)
    R[1, 0] = 1.2  # clamped
    factor = rng.uniform(0.4, 1.8, F).astype(np.float32)
    sun = oracle.sun_direction(35.0, 140.0)
    sunlit = rng.uniform(-0.2, 1.2, S).astype(np.float32)
    dni = rng.uniform(0.1, 2.0, nb).astype(np.float32)
    temp = rng.uniform(270.0, 330.0, S).astype(np.float32)
    tsun, tshade = np.array([300.0, 310.0, 305.0], np.float32), np.array([296.0, 300.0, 295.0], np.float32)
    strength = 0.6
    kw = dict(strength=strength, factor=factor)
    kw.update(temperature=temp) if mode == 1 else kw.update(sunlit_temperature=tsun, shaded_temperature=tshade)
    r, t, e, D = eng.spectralOptics(mode, mat, R, T, wl, sun, sunlit, dni, L, **kw)
    # host restatement
    nrm = sc.facet_normals().astype(np.float32)
    want = [np.zeros((S, nb), np.float32) for _ in range(4)]
    for s in range(S):
        f = s // 2
        sg = np.float32(1.0 if s % 2 == 0 else -1.0)
        sn = nrm[f] * sg
        for b in range(nb):
            w = float(wl[b])
            nmw = w * 1000.0 if w <= 2.5 else w
            thermal = (w > 2500.0) if mode == 0 else not (nmw <= 2500.0)
            rr = float(np.clip(np.float32(R[mat[f], b]) * (np.float32(1.0) if thermal else factor[f]), 0, 1))
            tt = float(np.clip(T[mat[f], b], 0, 1 - rr))
            eps = max(0.0, 1 - rr - tt)
            if f >= L and thermal:
                eps = oracle.hapke_emissivity(eps, abs(float(sn[1])), strength)
                rr = float(np.clip(1 - tt - eps, 0, 1))
            elif f >= L:
                rr = oracle.hapke_reflectance(rr, sun, sn, sn, strength)
            want[0][s, b] = rr
            want[1][s, b] = min(tt, 1 - rr)
            if thermal:
                if mode == 0:
                    ill = min(max(float(sunlit[s]), 0.0), 1.0)
                    want[2][s, b] = eps * (ill * planck_radiance(w, tsun[mat[f]]) +
                                           (1 - ill) * planck_radiance(w, tshade[mat[f]]))
                else:
                    want[2][s, b] = eps * planck_radiance(w, min(max(float(temp[s]), 1.0), 2000.0))
            else:
                lit = float(sunlit[s]) if mode == 0 else min(max(float(sunlit[s]), 0.0), 1.0)
                want[3][s, b] = lit * max(0.0, float(sn @ np.asarray(sun, np.float32))) * float(dni[b])
    for got, exp, name in zip((r, t, e, D), want, "rteD"):
        np.testing.assert_allclose(got, exp, rtol=2e-6, atol=1e-7, err_msg=name)
    assert np.any(e > 0) and np.any(D > 0) and np.ptp(r[2 * L:, 1]) > 1e-3
    # the fused solve equals fh_solve_bands on the built arrays (sunlit in
    # [0, 1], as the sunlit pass gives it: the direct term is non-negative)
    sky = np.full(nb, 3.0, np.float32)
    lit = np.clip(sunlit, 0, 1)
    r, t, e, D = eng.spectralOptics(mode, mat, R, T, wl, sun, lit, dni, L, **kw)
    a = eng.solveSpectral(mode, mat, R, T, wl, sun, lit, dni, L, sky, maxIter=64, **kw)
    b = eng.solveBands(r, t, e, D, sky, maxIter=64, tolRel=-1.0)
    np.testing.assert_array_equal(a.radiosity, b.radiosity)

"""Benchmark: converged facet-band-steps/s of the coupled radiosity + energy
balance diurnal run on one MI355X, surrogate+BF16 vs plain FP32, plus the
achieved HBM GB/s of the SpMM sweep against the 8 TB/s peak.

A step is one meteorological node: sunlit rays, the multi-band shortwave
solve, absorbed shortwave, and the longwave / energy-balance coupling run
until max dT <= 0.01 K and every radiosity solve's relative change <= 1e-5.
facet-band-steps = facets x (shortwave bands + 1 longwave band) x steps.

  python bench.py [--gpus 1] [--steps 24] [--warmup 2] [--workload terrain]

Prints one JSON line.  The CPU baseline is the oracle's restatement of the
reference CPU path on one node of the same workload, timed on 1 host core.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # forest scene with coupled energy balance: 60k facets, 1024 rays, 8 SW
    # bands + LW, 24 hourly steps
    "forest": dict(builder="forest", rays=1024, sw_bands=8, sun_rays=32, steps=24),
    # large terrain forest: ~1.0 M facets, 512 rays, 47 SW bands + LW
    "terrain": dict(builder="terrain", rays=512, sw_bands=47, sun_rays=16, steps=24),
    # urban block + trees: ~262k facets, 512 rays, 31 SW bands + LW
    "urban": dict(builder="urban", rays=512, sw_bands=31, sun_rays=16, steps=24),
}


WORKLOAD_SCENES = {
    "forest": "25 ellipsoid crowns (2 x 2 x 3 m, 1600 leaves of 0.1 m) on 50 x 50 m ground of 20k triangles",
    "terrain": "512 x 512 DEM at 1 m: sum of 4 sines of 20 m amplitude each (wavelengths 50-300 m) + N(0, 0.5 m) "
               "noise; 300 crowns (2 x 2 x 3 m, 1600 leaves of 0.1 m)",
    "urban": "40 box buildings (5-30 m) meshed at 0.5 m + 100 crowns",
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="terrain", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fast-precision", default="bf16", choices=["bf16", "fp8"],
                    help="increments the fast path's shortwave solves gather (BF16, or FP8 in 12-band chunks)")
    ap.add_argument("--surrogate", default=str(ROOT / "streamsim_amd" / "data" / "surrogate.npz"))
    ap.add_argument("--roofline-reps", type=int, default=50)
    ap.add_argument("--mode", default="run", choices=["run", "sweep", "precision", "raster"],
                    help="sweep: only the roofline kernels (form factors, SpMM sweeps, copy calibration); "
                         "precision: per-band error of the FP32 and BF16 shortwave solves against the FP64 oracle; "
                         "raster: the reference's raster graph and sunlit fraction on the urban validation scene")
    return ap.parse_args()


def sweep_bytes(S: int, F: int, nnz: int, nb: int, tables: bool, entry: int = 8) -> int:
    """Algorithmic HBM bytes of one SpMM sweep (compulsory traffic, gathers
    served from cache not counted): row offsets 4 (S + 1), `entry` bytes per
    CSR entry (8: column + FP32 weight; 4: packed column | count), fsky 4 S,
    and per side-band c, r, t read + B read + B write (20 S nb).  In the
    energy balance's form (`tables`) r and t come from per-material tables
    (4 F: the material index per facet) and c, D from per-side direct
    factors (4 S), so the per side-band part is B read + B write (8 S nb)."""
    b = 4 * (S + 1) + entry * nnz + 4 * S
    if tables:
        return b + 4 * F + 4 * S + 8 * S * nb
    return b + 20 * S * nb


def scheme_bytes(S: int, nb: int, increments: str | None, chebyshev: bool) -> int:
    """Bytes per sweep that the solver's own schemes add to the algorithmic
    ones: an increment sweep reads and writes the FP32 accumulated gathers
    and carries (16 bytes per side-band) and writes the next increments and
    reads the gathered ones once (BF16: 2 + 2; FP8: 1 + 1, plus the chunk
    exponents, 1 byte per 12 bands); the BF16 phase of the BF16-F scheme
    writes and reads its BF16 copy of B (2 + 2); Chebyshev reads the iterate
    before B (4)."""
    per = {None: 0.0, "bf16": 20.0, "fp8": 18.0 + 2.0 / 12.0, "bf16f": 4.0}[increments]
    return int(S * nb * (per + (4 if chebyshev else 0)))


def time_sweep(eng, nb: int, precision: int, reps: int, tables: bool = True) -> float:
    """Average duration (s) of the SpMM sweep kernel: HIP events recorded on
    the engine's stream around `reps` back-to-back launches, inputs in HBM;
    `tables`: the form the energy balance's shortwave solves run."""
    return eng.benchmarkSweep(nb, precision, reps, tables=tables)


def cpu_baseline(eng, scene, windows, meteo, cfg, nb_total: int, budget_s: float = 25.0):
    """The oracle (its current thread count) on one daytime node of the same graph: multi-band
    SW solve, incident, absorbed, then the coupled FP64 energy balance, both
    to the benchmark's tolerances.  Returns facet-band-steps/s."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_ref as O
    from streamsim_amd import faceteb, optics

    g = eng.getGraph()
    og = O.Graph(g.rowOffsets, g.neighbors, g.counts, g.denominators, g.skyCounts)
    k = max(range(len(meteo)), key=lambda i: meteo[i].shortwave)
    met = meteo[k]
    cz = math.cos(math.radians(met.zenith))
    dni, dif = windows.irradiance(met.shortwave, cz)
    sun = faceteb.sun_vector(met.zenith, met.azimuth)
    sl = eng.computeSunlitFraction(sun)
    r_tab, t_tab = faceteb.material_tables(["leaf", "soil", "building"], windows)
    mat = np.repeat(scene.material, 2)
    r = r_tab[mat]
    t = np.minimum(np.clip(t_tab[mat], 0, 1 - r), 1 - r)
    cos = np.maximum(0.0, optics.side_normals(scene) @ sun).astype(np.float32)
    D = (np.clip(sl, 0, 1) * cos)[:, None] * dni[None, :]
    t0 = time.perf_counter()
    B, it, _ = O.solve_bands(og, r, t, np.zeros_like(r), D, dif, tol_rel=cfg.radiosity_tolerance)
    H = O.incident_bands(og, B, D, dif)
    absorbed = np.sum(np.clip(1 - r - t, 0, 1) * H, axis=1).astype(np.float32)
    mats = []
    for m in (faceteb.LEAF, faceteb.SOIL, faceteb.BUILDING):
        om = O.OrEbMaterial()
        for kk, v in m.items():
            setattr(om, kk, v)
        mats.append(om)
    p = O.OrEbParams()
    p.airTemperatureK = met.air_temperature_k
    p.vaporPressure = met.vapor_pressure
    p.pressure = met.pressure
    p.windSpeed = met.wind
    p.skyLongwave = met.longwave
    p.dTime = cfg.dtime
    p.tolerance = cfg.tolerance
    p.relaxation = cfg.relaxation
    p.maxCoupling = cfg.max_coupling
    p.lwMaxSweeps = cfg.max_sweeps
    p.lwTolRel = cfg.radiosity_tolerance
    T0 = np.full(scene.surface_count, met.air_temperature_k, np.float32)
    O.eb_node(og, scene.facet_count, scene.leaf_count, scene.material, mats, absorbed, p, T0, T0)
    dt = time.perf_counter() - t0
    return scene.facet_count * nb_total / dt, dt, f"1 converged daytime node (hour {met.hour:.1f}, {it} SW sweeps)"


def precision_report(eng, scene, windows, meteo, cfg) -> dict:
    """Per-band error of the GPU shortwave multi-band solve on the brightest
    node against the FP64 oracle on the same graph and inputs: FP32 sweeps,
    the BF16 solve (BF16 increments gathered, FP32 weights and
    accumulation), the FP8 increments and BF16 F with FP32 defect
    correction, each to relative change 1e-7 and to the run's 1e-5; the
    oracle to 1e-9 (FP64 accumulation, its threads).  Error per band: max
    |B - B_ref| / B_ref over the sides with B_ref above 1e-3 of the band's
    maximum."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_ref as O
    from streamsim_amd import faceteb, optics
    from streamsim_amd._lib import FH_PRECISION_BF16_DC, FH_PRECISION_BF16_F, FH_PRECISION_FP8_INC, FH_PRECISION_FP32

    g = eng.getGraph()
    og = O.Graph(g.rowOffsets, g.neighbors, g.counts, g.denominators, g.skyCounts)
    k = max(range(len(meteo)), key=lambda i: meteo[i].shortwave)
    met = meteo[k]
    cz = math.cos(math.radians(met.zenith))
    dni, dif = windows.irradiance(met.shortwave, cz)
    sun = faceteb.sun_vector(met.zenith, met.azimuth)
    sl = eng.computeSunlitFraction(sun)
    r_tab, t_tab = faceteb.material_tables(["leaf", "soil", "building"], windows)
    mat = np.repeat(scene.material, 2)
    r = r_tab[mat].astype(np.float32)
    t = np.minimum(np.clip(t_tab[mat], 0, 1 - r), 1 - r).astype(np.float32)
    cos = np.maximum(0.0, optics.side_normals(scene) @ sun).astype(np.float32)
    D = ((np.clip(sl, 0, 1) * cos)[:, None] * dni[None, :]).astype(np.float32)
    e = np.zeros_like(r)
    t0 = time.perf_counter()
    ref, it_ref, _ = O.solve_bands(og, r, t, e, D, dif, max_iter=400, tol_rel=1e-9)
    t_ref = time.perf_counter() - t0
    ref = ref.astype(np.float64)
    mask = ref > 1e-3 * ref.max(axis=0, keepdims=True)
    out = {"node_hour": met.hour, "bands": int(windows.count), "sides": int(scene.surface_count),
           "oracle": {"iterations": it_ref, "seconds": t_ref, "threads": O.max_threads()}}
    for name, prec in (("fp32", FH_PRECISION_FP32), ("bf16", FH_PRECISION_BF16_DC), ("fp8", FH_PRECISION_FP8_INC),
                       ("bf16f", FH_PRECISION_BF16_F)):
        for tol in (1e-7, cfg.radiosity_tolerance):
            eng.synchronize()
            t1 = time.perf_counter()
            res = eng.solveBands(r, t, e, D, dif, tolRel=tol, precision=prec)
            eng.synchronize()
            dt = time.perf_counter() - t1
            got = res.radiosity.astype(np.float64)
            err = [float(np.max(np.abs(got[mask[:, b], b] - ref[mask[:, b], b]) / ref[mask[:, b], b]))
                   for b in range(windows.count)]
            out[f"{name}_tol{tol:g}"] = {"iterations": res.iterations, "seconds": dt, "max_rel_error": max(err),
                                         "last_rel_change": res.maxRelDelta, "max_rel_error_per_band": err}
    return out


def cpu_baseline_sampled(eng, scene, windows, nb_total: int, st_plain, reps: int = 3):
    """The metric's criterion on the CPU for graphs too large to converge a
    node within the bench: the oracle (its current thread count) times `reps`
    full-scene FP64 sweeps of all shortwave bands and `reps` longwave sweeps;
    a node is then priced with the plain GPU run's sweep counts per node (the
    oracle's Jacobi sweeps would need more of them).  Returns
    facet-band-steps/s."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_ref as O

    g = eng.getGraph()
    og = O.Graph(g.rowOffsets, g.neighbors, g.counts, g.denominators, g.skyCounts)
    S = scene.surface_count
    nb = windows.count
    rng = np.random.default_rng(0)
    r = rng.uniform(0.0, 0.4, (S, nb)).astype(np.float32)
    t = (0.5 * r).astype(np.float32)
    z = np.zeros((S, nb), np.float32)
    D = rng.uniform(0.0, 300.0, (S, nb)).astype(np.float32)
    sky = np.full(nb, 20.0, np.float32)
    t0 = time.perf_counter()
    O.solve_bands(og, r, t, z, D, sky, max_iter=reps, tol_rel=-1.0)
    t_sw = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    O.solve_bands(og, r[:, :1], t[:, :1], z[:, :1], D[:, :1], sky[:1], max_iter=reps, tol_rel=-1.0)
    t_lw = (time.perf_counter() - t0) / reps
    nodes = st_plain.nodes
    per_node = (sum(st_plain.sw_sweeps) * t_sw + sum(st_plain.lw_sweeps) * t_lw) / nodes
    sample = (f"{reps} FP64 sweeps of {nb} bands ({t_sw:.2f}s each) + {reps} LW sweeps ({t_lw:.2f}s each) over the full "
              f"graph, priced at the plain run's {sum(st_plain.sw_sweeps) / nodes:.1f} SW + "
              f"{sum(st_plain.lw_sweeps) / nodes:.1f} LW sweeps per node")
    return scene.facet_count * nb_total / per_node, reps * (t_sw + t_lw), sample


# the reference's fixed schedule per node (core:3495-3565, 3745-3846)
REF_SWEEPS = 64            # kIterations, core:81
REF_COUPLING_BUDGET = 20   # couplingIterations default, core:3206-3208


def cpu_baseline_reference(eng, scene, windows, nb_total: int, day: list, budget_s: float = 12.0):
    """The reference's own CPU route (facetBackend=cpu) at its own schedule:
    the oracle's restatement of solveCpu (core:1476-1529: one band, FP64
    accumulation, single-threaded as in the reference) timed on full-graph
    sweeps of the benchmark graph; a node is priced at the reference's fixed
    schedule: per shortwave window with light a 64-sweep solve, an incident
    pass and a 64-sweep diffuse-only solve (core:3528-3536); per coupling
    iteration a 64-sweep longwave solve and an incident pass, for the
    default budget of 20 iterations (core:3745-3832); the final 64-sweep
    longwave solve and incident pass (core:3834-3846).  An incident pass is
    counted as one sweep (both are one pass over the CSR).  Returns
    facet-band-steps/s, the CPU seconds spent and the sample description."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_ref as O

    g = eng.getGraph()
    og = O.Graph(g.rowOffsets, g.neighbors, g.counts, g.denominators, g.skyCounts)
    S = scene.surface_count
    rng = np.random.default_rng(1)
    optics = np.zeros((S, 4), np.float32)
    optics[:, 0] = rng.uniform(0.0, 0.4, S)
    optics[:, 1] = 0.5 * optics[:, 0]
    optics[:, 3] = rng.uniform(0.0, 300.0, S)
    t0 = time.perf_counter()
    O.solve(og, optics, 20.0, 1)
    t1 = time.perf_counter() - t0
    reps = int(max(1, min(64, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    O.solve(og, optics, 20.0, reps)
    t_sweep = (time.perf_counter() - t0) / reps
    lit = sum(1 for m in day if m.shortwave > 0.0)
    per_window = 2 * REF_SWEEPS + 1
    per_coupling = REF_SWEEPS + 1
    sweeps = lit * windows.count * per_window + len(day) * (REF_COUPLING_BUDGET * per_coupling + per_coupling)
    seconds = sweeps * t_sweep
    sample = (f"{reps + 1} full-graph single-band FP64 sweeps of the solveCpu restatement ({t_sweep:.3f}s each, "
              f"1 core); {len(day)} nodes ({lit} lit) priced at the reference's schedule: per lit window "
              f"2 x {REF_SWEEPS} sweeps + incident ({windows.count} windows), per node "
              f"{REF_COUPLING_BUDGET} coupling iterations (its default budget) x ({REF_SWEEPS} LW sweeps + incident) "
              f"+ the final LW solve; {sweeps} sweeps, {seconds:.0f} s")
    return scene.facet_count * nb_total * len(day) / seconds, (reps + 1) * t_sweep, sample


def load_surrogate(eng, path: str) -> bool:
    if not os.path.exists(path):
        return False
    z = np.load(path, allow_pickle=False)
    eng.mlpSetParams(int(z["hidden"]), z["params"], z["feature_mean"], z["feature_std"], float(z["target_mean"]),
                     float(z["target_std"]))
    return True


def raster_report(reps: int) -> None:
    """The raster visibility graph (1024 x 1024 pixels, 64 fragments, the 40
    hemisphere directions) and the raster sunlit fraction of the urban
    validation scene (4746 facets), host to host, averaged over `reps`
    calls after one warm-up call."""
    from streamsim_amd import FacetEngine, scenes
    sc = scenes.load_tree_with_soil(ROOT / "tests" / "golden" / "urban_tree_building.obj")
    eng = FacetEngine()
    eng.setGeometry(sc.vertices)
    n = 40  # makeHemisphereDirections: golden-angle spiral over the upper hemisphere
    i = np.arange(n, dtype=np.float32)
    y = (i + np.float32(0.5)) / np.float32(n)
    rad = np.sqrt(np.maximum(np.float32(0.0), np.float32(1.0) - y * y))
    az = np.float32(2.39996322972865332) * i
    dirs = np.stack([rad * np.cos(az), y, rad * np.sin(az)], 1).astype(np.float32)
    sun = np.array([np.sin(np.radians(30.0)) * np.cos(np.radians(135.0)), np.cos(np.radians(30.0)),
                    np.sin(np.radians(30.0)) * np.sin(np.radians(135.0))], np.float32)
    d = eng.buildRasterGraph(dirs, 1024, 1024, 64)
    eng.computeRasterSunlitFraction(sun, 1024, 1024, 64)
    t = time.perf_counter()
    for _ in range(reps):
        eng.buildRasterGraph(dirs, 1024, 1024, 64)
    tg = (time.perf_counter() - t) / reps
    t = time.perf_counter()
    for _ in range(reps):
        eng.computeRasterSunlitFraction(sun, 1024, 1024, 64)
    ts = (time.perf_counter() - t) / reps
    print(json.dumps({"workload": "urban validation scene, 4746 facets, raster 1024 x 1024 x 64, 40 directions",
                      "directed_edges": d.directedEdgeCount, "graph_ms": tg * 1e3, "sunlit_ms": ts * 1e3,
                      "reps": reps}), flush=True)


def main() -> None:
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (no multi-GPU path is built)"}))
        sys.exit(2)
    if a.mode == "raster":
        raster_report(a.roofline_reps)
        return
    wl = WORKLOADS[a.workload]
    steps = a.steps if a.steps is not None else wl["steps"]
    from streamsim_amd import FacetEngine, faceteb, scenes
    from streamsim_amd._lib import (FH_PRECISION_BF16_DC, FH_PRECISION_BF16_F, FH_PRECISION_FP8_INC, FH_PRECISION_FP32,
                                    FH_WARM_PREVIOUS)

    log(f"[bench] workload={a.workload} steps={steps} warmup={a.warmup}")
    t = time.perf_counter()
    scene = scenes.BUILDERS[wl["builder"]]()
    log(f"[bench] scene {scene.facet_count} facets ({scene.leaf_count} objects) in {time.perf_counter() - t:.1f}s")
    eng = FacetEngine(rays_per_side=wl["rays"], sun_rays_per_side=wl["sun_rays"], seed=1)
    t = time.perf_counter()
    eng.setGeometry(scene.vertices)
    t_bvh = time.perf_counter() - t
    eng.synchronize()
    t = time.perf_counter()
    diag = eng.buildVisibilityGraph()
    eng.synchronize()
    t_graph = time.perf_counter() - t
    S = scene.surface_count
    rays_per_s = S * wl["rays"] / t_graph
    log(f"[bench] BVH {t_bvh:.2f}s, form factors {t_graph:.3f}s ({rays_per_s:.3g} rays/s), "
        f"nnz {diag.directedEdgeCount} ({diag.directedEdgeCount / S:.1f}/side)")

    windows = faceteb.ShortwaveWindows.for_bands(wl["sw_bands"])
    nb_total = windows.count + 1
    if a.mode == "precision":
        rep = precision_report(eng, scene, windows, faceteb.diurnal_meteo(24), faceteb.EbConfig())
        rep["workload"] = a.workload
        print(json.dumps(rep), flush=True)
        return
    if a.mode == "sweep":
        # counter runs: the SW sweeps of the energy balance (material tables)
        # in FP32 and as the increment sweep of BF16 solves, the same with
        # streamed c, r, t arrays, the LW sweep, and a 1 GiB copy as the byte
        # calibration of the traffic counters
        nb = windows.count
        F, nnz = scene.facet_count, diag.directedEdgeCount
        out = {"workload": a.workload, "sides": S, "facets": F, "nnz": nnz}
        for name, nbk, prec, tab in (("sw_fp32", nb, FH_PRECISION_FP32, True), ("sw_inc", nb, FH_PRECISION_BF16_DC, True),
                                     ("sw_inc_l1", nb, FH_PRECISION_BF16_DC, True),
                                     ("sw_q8", nb, FH_PRECISION_FP8_INC, True),
                                     ("sw_bf16f", nb, FH_PRECISION_BF16_F, True),
                                     ("sw_fp32_arrays", nb, FH_PRECISION_FP32, False),
                                     ("sw_inc_arrays", nb, FH_PRECISION_BF16_DC, False),
                                     ("lw_fp32", 1, FH_PRECISION_FP32, False)):
            # sw_inc_l1: the increment sweep without the staged gathers (every
            # entry gathered through the vector L1), for comparison
            os.environ["FH_STAGED"] = "0" if name == "sw_inc_l1" else "1"
            sec = time_sweep(eng, nbk, prec, a.roofline_reps, tables=tab)
            # single-band sweeps read the packed entries
            alg = sweep_bytes(S, F, nnz, nbk, tables=tab and nbk >= 16, entry=8 if nbk >= 16 else 4)
            inc = (None if prec == FH_PRECISION_FP32 or nbk < 16 else
                   {FH_PRECISION_FP8_INC: "fp8", FH_PRECISION_BF16_F: "bf16f"}.get(prec, "bf16"))
            extra = scheme_bytes(S, nbk, increments=inc, chebyshev=tab and nbk >= 16)
            out[name] = {"us": sec * 1e6, "alg_bytes": alg, "scheme_bytes": extra, "GBps": alg / sec / 1e9,
                         "frac": alg / sec / 1e9 / HBM_PEAK_GBS}
        os.environ.pop("FH_STAGED", None)
        cs = eng.benchmarkCopy(1 << 30, 20)
        out["copy_1GiB"] = {"us": cs * 1e6, "GBps": 2.0 * (1 << 30) / cs / 1e9}
        print(json.dumps(out), flush=True)
        return
    faceteb.setup_scene(eng, scene, windows)
    cfg = faceteb.EbConfig()
    meteo = faceteb.diurnal_meteo(24)
    day = [meteo[k % 24] for k in range(steps)]
    have_sur = load_surrogate(eng, a.surrogate)
    fast_warm = 1 if have_sur else FH_WARM_PREVIOUS

    def run(precision, warm):
        if a.warmup:
            faceteb.run_diurnal(eng, meteo[:a.warmup], windows, cfg, precision=precision, warm=warm)
        eng.synchronize()
        t0 = time.perf_counter()
        st = faceteb.run_diurnal(eng, day, windows, cfg, precision=precision, warm=warm)
        eng.synchronize()
        return time.perf_counter() - t0, st

    fast_prec = FH_PRECISION_FP8_INC if a.fast_precision == "fp8" else FH_PRECISION_BF16_DC
    t_plain, st_plain = run(FH_PRECISION_FP32, FH_WARM_PREVIOUS)
    log(f"[bench] plain FP32: {t_plain:.3f}s {st_plain.summary()}")
    t_fast, st_fast = run(fast_prec, fast_warm)
    log(f"[bench] {a.fast_precision.upper()}{'+surrogate' if have_sur else ''}: {t_fast:.3f}s {st_fast.summary()}")
    work = scene.facet_count * nb_total * steps
    value = work / t_fast
    plain_value = work / t_plain
    # ablations: each change alone
    ablation = {}
    # (the BF16-F scheme: BF16 F and a BF16 copy of B until the BF16
    # operator's fixed point, then FP32 sweeps as the defect correction)
    for name, prec in (("bf16_increments_only", FH_PRECISION_BF16_DC), ("fp8_increments_only", FH_PRECISION_FP8_INC),
                       ("bf16_weights_defect_correction", FH_PRECISION_BF16_F)):
        t_b, st_b = run(prec, FH_WARM_PREVIOUS)
        ablation[name] = {"value": work / t_b, "seconds": t_b, **st_b.summary()}
    if have_sur:
        t_s, st_s = run(FH_PRECISION_FP32, 1)
        ablation["surrogate_only"] = {"value": work / t_s, "seconds": t_s, **st_s.summary()}
    log(f"[bench] ablation {json.dumps({k: round(v['seconds'], 4) for k, v in ablation.items()})}")

    # roofline of the fast path's dominant kernel, the increment sweep of
    # the BF16 shortwave solves (all bands active, the energy balance's form:
    # material tables, direct factors, Chebyshev), and of the plain path's,
    # the FP32 sweep in the same form
    nb = windows.count
    nnz = diag.directedEdgeCount
    F = scene.facet_count
    sec16 = time_sweep(eng, nb, fast_prec, a.roofline_reps)
    sec32 = time_sweep(eng, nb, FH_PRECISION_FP32, a.roofline_reps)
    alg = sweep_bytes(S, F, nnz, nb, tables=True)
    extra16 = scheme_bytes(S, nb, increments=a.fast_precision, chebyshev=True)
    extra32 = scheme_bytes(S, nb, increments=None, chebyshev=True)
    achieved = alg / sec16 / 1e9
    log(f"[bench] SW sweep nb={nb}: increments {sec16 * 1e6:.1f} us ({achieved:.0f} GB/s), "
        f"fp32 {sec32 * 1e6:.1f} us ({alg / sec32 / 1e9:.0f} GB/s)")
    traffic32 = traffic16 = None
    tf = ROOT / "profiles" / f"sweep_traffic_{a.workload}.json"
    if tf.exists():
        tj = json.loads(tf.read_text())
        traffic32 = tj.get("sw_fp32", {}).get("bytes_per_launch")
        traffic16 = tj.get("sw_inc" if a.fast_precision == "bf16" else "sw_q8", {}).get("bytes_per_launch")

    cpu = cpu_metric = None
    if not a.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_ref as O
        ncores = O.max_threads()
        v, dt, sample = cpu_baseline_reference(eng, scene, windows, nb_total, day)
        cpu = {"value": v, "unit": "facet-band-steps/s", "cores": 1, "kind": "port",
               "sample": f"{sample}; {dt:.1f}s timed"}
        log(f"[bench] cpu baseline, reference schedule (1 core) {v:.4g} facet-band-steps/s ({dt:.1f}s)")
        O.set_threads(ncores)
        if nnz <= 50_000_000:
            v, dt, sample = cpu_baseline(eng, scene, windows, meteo, cfg, nb_total)
        else:
            v, dt, sample = cpu_baseline_sampled(eng, scene, windows, nb_total, st_plain)
        cpu_metric = {"value": v, "unit": "facet-band-steps/s", "cores": ncores, "kind": "port",
                      "sample": f"{sample}; {dt:.1f}s on {ncores} cores"}
        log(f"[bench] cpu baseline, metric criterion ({ncores} cores) {v:.4g} facet-band-steps/s ({dt:.1f}s)")

    # all-in: the form-factor build (host BVH + device rays) charged to the run
    t_setup = t_graph + t_bvh
    out = {
        "metric": "converged facet-band-steps/s at 0.01 K: surrogate+BF16 vs plain FP32; HBM GB/s",
        "value": value,
        "unit": "facet-band-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": t_fast / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": a.workload,
            "facets": scene.facet_count,
            "sides": S,
            "rays_per_side": wl["rays"],
            "bands": nb_total,
            "shortwave_bands": nb,
            "time_steps": steps,
            "tolerance_K": cfg.tolerance,
            "radiosity_rel_tol": cfg.radiosity_tolerance,
            "fast_path": f"shortwave: {a.fast_precision.upper()} increments of B gathered, FP32 weights and "
                         "accumulation (FP32 fixed point)" + (" + surrogate warm start" if have_sur else "") +
                         "; both paths: Chebyshev, per-band retirement, coupled facet pairs as 2-facet blocks",
            "parallelism": "single GPU",
            "scene": WORKLOAD_SCENES[a.workload],
        },
        "plain_fp32": {"value": plain_value, "seconds": t_plain, **st_plain.summary()},
        "fast": {"value": value, "seconds": t_fast, **st_fast.summary()},
        "speedup_vs_plain": t_plain / t_fast,
        "all_in": {"value": work / (t_fast + t_setup), "plain_value": work / (t_plain + t_setup),
                   "speedup_vs_plain": (t_plain + t_setup) / (t_fast + t_setup), "setup_seconds": t_setup,
                   "note": "form factors (host BVH + device rays) charged to the run"},
        "ablation": ablation,
        "form_factors": {"seconds": t_graph, "rays_per_s": rays_per_s, "directed_edges": nnz,
                         "bvh_host_seconds": t_bvh},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic16,
                     "kernel": (f"k_sweep_staged<2> (single facets, gathers staged through the LDS) + "
                                f"k_sweep_hv<2, true> (coupled pairs, second stream, side by side): increment sweep "
                                f"of the BF16 shortwave solves"
                                if a.fast_precision == "bf16" else
                                "k_sweep_q8_single (single facets) + k_sweep_q8_pair (coupled pairs, second stream, side "
                                "by side): increment sweep of the FP8 shortwave solves") +
                               f", {nb} bands, material tables; bytes and time per sweep",
                     "bytes_per_launch": alg, "us_per_launch": sec16 * 1e6,
                     "formula": "4(S+1) + 8 nnz + 4S + 4F + 4S + 8 S nb (row offsets, column + FP32 weight, fsky, "
                                "material per facet, direct factor per side, B read + write)",
                     "scheme_bytes_per_launch": extra16,
                     "scheme_note": "increment scheme's accumulated gathers and carries (16 B), increments "
                                    "written and gathered (BF16 2 + 2 B, FP8 1 + 1 B + exponents), Chebyshev's previous "
                                    "iterate (4 B) per side-band: not algorithmic, not in achieved",
                     "fp32_sweep": {"kernel": f"k_sweep_v<0, false, false> + k_sweep_v<0, false, true>: FP32 "
                                              f"shortwave sweep of the plain path, {nb} bands, material tables",
                                    "achieved": alg / sec32 / 1e9, "frac": alg / sec32 / 1e9 / HBM_PEAK_GBS,
                                    "bytes_per_launch": alg, "scheme_bytes_per_launch": extra32,
                                    "us_per_launch": sec32 * 1e6, "traffic": traffic32}},
        "cpu_baseline": cpu,
        "cpu_baseline_metric": cpu_metric,
    }
    print(json.dumps(out), flush=True)
    eng.destroy()


if __name__ == "__main__":
    main()

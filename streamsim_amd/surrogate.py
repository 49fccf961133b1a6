"""Learned warm start: dataset generation, on-GPU training, evaluation.

The plain coupled solver (device FacetEB node step) generates (features,
converged temperature) pairs on the forest and urban scenes under random
conditions (sun zenith U(0, 70 deg), azimuth U(0, 360), air temperature
U(280, 310) K, wind U(0.5, 6) m/s, a fifth of the nodes at night) -- two
forests, the urban block and a 160 m terrain patch; the MLP
9 -> 128 -> 128 -> 1 (GELU) is trained by the engine's own HIP forward /
backward / Adam kernels (fh_mlp_train) on device-resident data.  The warm
start only seeds the fixed point: every node still iterates to 0.01 K, so
the converged answer is unchanged within tolerance.

  python -m streamsim_amd.surrogate --out gpurun_out/surrogate.npz
"""
from __future__ import annotations

import argparse
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import faceteb, scenes
from ._lib import FH_PRECISION_FP32, FH_WARM_PREVIOUS, FH_WARM_SURROGATE
from .engine import FacetEngine

HIDDEN = 128


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def random_meteo(rng, n: int) -> list[faceteb.Meteo]:
    out = []
    for _ in range(n):
        night = rng.random() < 0.2
        zen = 100.0 if night else float(rng.uniform(0.0, 70.0))
        az = float(rng.uniform(0.0, 360.0))
        ta = float(rng.uniform(280.0, 310.0))
        wind = float(rng.uniform(0.5, 6.0))
        cz = max(0.0, math.cos(math.radians(zen)))
        sw = 0.75 * 1000.0 * cz
        lw = float(rng.uniform(0.65, 0.85)) * faceteb.SIGMA * ta ** 4
        out.append(faceteb.Meteo(12.0, ta, float(rng.uniform(8.0, 20.0)), 1000.0, wind, sw, lw, zen, az))
    return out


def engine_for(scene, rays: int, sun_rays: int, bands: int, seed: int = 1):
    eng = FacetEngine(rays_per_side=rays, sun_rays_per_side=sun_rays, seed=seed)
    eng.setGeometry(scene.vertices)
    eng.buildVisibilityGraph()
    win = faceteb.ShortwaveWindows.for_bands(bands)
    faceteb.setup_scene(eng, scene, win)
    return eng, win


def collect(eng, win, meteo, cfg, sequence: int = 4):
    """Runs the nodes in sequences of `sequence` consecutive nodes (the first
    of each starts from the air temperature) and returns stacked features and
    converged unit temperatures."""
    X, Y = [], []
    for k, m in enumerate(meteo):
        cz = math.cos(math.radians(m.zenith))
        dni, dif = win.irradiance(m.shortwave, cz)
        p = faceteb.node_params(m, cfg, first=(k % sequence == 0), precision=FH_PRECISION_FP32,
                                warm=FH_WARM_PREVIOUS, record=True)
        st = eng.ebRunNode(p, dni, dif)
        if not st.converged:
            continue
        f, T = eng.ebGetFeatures()
        X.append(f)
        Y.append(T - np.float32(m.air_temperature_k))
    return np.concatenate(X), np.concatenate(Y)


def train(eng, X: np.ndarray, Y: np.ndarray, epochs: int, batch: int, lr: float, seed: int):
    mean = X.mean(axis=0).astype(np.float32)
    std = X.std(axis=0).astype(np.float32)
    std[std < 1e-6] = 1.0
    tmean = float(Y.mean())
    tstd = float(Y.std()) or 1.0
    Xn = ((X - mean) / std).astype(np.float32)
    Yn = ((Y - tmean) / tstd).astype(np.float32)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(Xn.shape[0])
    Xn, Yn = Xn[perm], Yn[perm]
    n_par = FacetEngine.mlpParamCount(HIDDEN)
    P = np.zeros(n_par, np.float32)
    # He-style init for the two hidden layers, small output layer
    o = 0
    P[o:o + HIDDEN * 9] = rng.normal(0, math.sqrt(2.0 / 9), HIDDEN * 9)
    o += HIDDEN * 9 + HIDDEN
    P[o:o + HIDDEN * HIDDEN] = rng.normal(0, math.sqrt(2.0 / HIDDEN), HIDDEN * HIDDEN)
    o += HIDDEN * HIDDEN + HIDDEN
    P[o:o + HIDDEN] = rng.normal(0, math.sqrt(1.0 / HIDDEN), HIDDEN)
    eng.mlpSetParams(HIDDEN, P, mean, std, tmean, tstd)
    t0 = time.perf_counter()
    losses = []
    for e in range(0, epochs, 5):
        k = min(5, epochs - e)
        losses.append(eng.mlpTrain(Xn, Yn, batch, k, lr * (0.6 ** (e // 5)), 1.0, seed + e))
    dt = time.perf_counter() - t0
    return dict(hidden=HIDDEN, params=eng.mlpGetParams(), feature_mean=mean, feature_std=std,
                target_mean=np.float32(tmean), target_std=np.float32(tstd), losses=np.asarray(losses, np.float32),
                train_seconds=np.float32(dt), samples=np.int64(Xn.shape[0]))


def predict(eng, X: np.ndarray, model) -> np.ndarray:
    Xn = ((X - model["feature_mean"]) / model["feature_std"]).astype(np.float32)
    y = eng.mlpForward(Xn)
    return float(model["target_mean"]) + float(model["target_std"]) * y


def evaluate(eng, win, meteo, cfg, model):
    """Surrogate error and coupling iterations with / without the warm start
    on the same node sequence."""
    res = {}
    for name, warm in (("plain", FH_WARM_PREVIOUS), ("surrogate", FH_WARM_SURROGATE)):
        its, errs, times = [], [], []
        for k, m in enumerate(meteo):
            cz = math.cos(math.radians(m.zenith))
            dni, dif = win.irradiance(m.shortwave, cz)
            p = faceteb.node_params(m, cfg, first=(k == 0), precision=FH_PRECISION_FP32, warm=warm,
                                    record=(warm == FH_WARM_PREVIOUS))
            t0 = time.perf_counter()
            st = eng.ebRunNode(p, dni, dif)
            times.append(time.perf_counter() - t0)
            its.append(st.couplingIterations)
            if warm == FH_WARM_PREVIOUS:
                f, T = eng.ebGetFeatures()
                pred = predict(eng, f, model) + m.air_temperature_k
                errs.append(float(np.sqrt(np.mean((pred - T) ** 2))))
        res[name] = dict(coupling=int(np.sum(its)), seconds=float(np.sum(times)), per_node=its,
                         first_node=int(its[0]))
        if errs:
            res["surrogate_rmse_K"] = float(np.mean(errs))
    return res


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="gpurun_out/surrogate.npz")
    ap.add_argument("--nodes", type=int, default=16, help="training nodes per scene")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--quick", action="store_true", help="small scenes (test run)")
    a = ap.parse_args()
    rng = np.random.default_rng(2024)
    cfg = faceteb.EbConfig()
    Xs, Ys = [], []
    if a.quick:
        plan = [("row_crop", dict(seed=11), 128, 4)]
    else:
        plan = [("forest", dict(seed=12), 1024, 8), ("forest", dict(seed=13), 1024, 8),
                ("urban", dict(seed=14), 512, 31), ("terrain", dict(seed=31, n=160, trees=30), 512, 47)]
    eng = None
    for name, kw, rays, bands in plan:
        t0 = time.perf_counter()
        sc = scenes.BUILDERS[name](**kw)
        eng, win = engine_for(sc, rays, 16, bands)
        X, Y = collect(eng, win, random_meteo(rng, a.nodes), cfg)
        Xs.append(X)
        Ys.append(Y)
        log(f"[surrogate] {name}{kw}: {X.shape[0]} samples in {time.perf_counter() - t0:.1f}s")
        if name != plan[-1][0] or kw != plan[-1][1]:
            eng.destroy()
    X = np.concatenate(Xs)
    Y = np.concatenate(Ys)
    log(f"[surrogate] dataset {X.shape}, target mean {Y.mean():.2f} K sd {Y.std():.2f} K")
    model = train(eng, X, Y, a.epochs, a.batch, a.lr, 7)
    log(f"[surrogate] trained in {float(model['train_seconds']):.1f}s, loss {model['losses'][-1]:.4f}")
    # held-out geometry and conditions: a forest seed and the 1 M-facet
    # terrain forest of the benchmark (neither in the training data)
    held = [("row_crop", scenes.row_crop(seed=99, leaves_per_row=100, cells=16), 128, 4)] if a.quick else \
        [("forest", scenes.forest(seed=77), 1024, 8), ("terrain", scenes.terrain(), 512, 47)]
    eng.destroy()
    evals = {}
    for name, sc, rays, bands in held:
        eng, win = engine_for(sc, rays, 16, bands)
        eng.mlpSetParams(int(model["hidden"]), model["params"], model["feature_mean"], model["feature_std"],
                         float(model["target_mean"]), float(model["target_std"]))
        evals[name] = evaluate(eng, win, faceteb.diurnal_meteo(24, t_mean=295.0, t_amp=7.0, wind=3.0,
                                                               day_of_year=200), cfg, model)
        log(f"[surrogate] held-out {name}: {evals[name]}")
        eng.destroy()
    ev = evals[held[0][0]]
    Path(a.out).parent.mkdir(parents=True, exist_ok=True)
    extra = {}
    for name, e in evals.items():
        extra[f"eval_{name}_plain_coupling"] = e["plain"]["coupling"]
        extra[f"eval_{name}_surrogate_coupling"] = e["surrogate"]["coupling"]
        extra[f"eval_{name}_plain_first_node"] = e["plain"]["first_node"]
        extra[f"eval_{name}_surrogate_first_node"] = e["surrogate"]["first_node"]
        extra[f"eval_{name}_rmse_K"] = e.get("surrogate_rmse_K", float("nan"))
    np.savez(a.out, **model, eval_plain_coupling=ev["plain"]["coupling"],
             eval_surrogate_coupling=ev["surrogate"]["coupling"],
             eval_rmse_K=ev.get("surrogate_rmse_K", float("nan")), **extra)
    print(f"saved {a.out}", flush=True)


if __name__ == "__main__":
    main()

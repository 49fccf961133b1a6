"""Host-side mirror of the reference's device backend class.

``FacetEngine`` has the operations of ``facetvk::FacetrtVulkan``
(models/histream/src/facetrt/facetrt_vulkan.h:70-104) with the same argument
meaning and error classes (ValueError = std::invalid_argument,
LogicError = std::logic_error, RuntimeError, OverflowError); every call goes
through the C-ABI of ``include/facet_hip.h`` into HIP kernels on one GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (FhConfig, FhDirection, FhEbMaterial, FhEbNodeParams, FhEbNodeStats, FhGraphDiagnostics,
                   FhSolveStats, raise_for)


@dataclass
class GraphDiagnostics:
    """facetvk::GraphDiagnostics (facetrt_vulkan.h:49-56)."""
    facetCount: int
    surfaceCount: int
    directedEdgeCount: int
    fragmentOverflow: int
    hashOverflow: int
    maxClosureError: float


@dataclass
class SolveResult:
    """facetvk::SolveResult (facetrt_vulkan.h:58-62) plus the relative delta."""
    radiosity: np.ndarray
    iterations: int
    maxDelta: float
    maxRelDelta: float = 0.0
    refreshes: int = 0


@dataclass
class Graph:
    rowOffsets: np.ndarray
    neighbors: np.ndarray
    counts: np.ndarray
    denominators: np.ndarray
    skyCounts: np.ndarray


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a, shape=None) -> np.ndarray:
    x = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None and x.shape != shape:
        x = x.reshape(shape)
    return x


def _diag(d: FhGraphDiagnostics) -> GraphDiagnostics:
    return GraphDiagnostics(d.facetCount, d.surfaceCount, d.directedEdgeCount, d.fragmentOverflow,
                            d.hashOverflow, d.maxClosureError)


class FacetEngine:
    """One GPU context: geometry, form factors, radiosity solves, sunlit fraction."""

    def __init__(self, rays_per_side: int = 256, sun_rays_per_side: int = 64, seed: int = 1,
                 gpu_index: int = 0, ray_epsilon: float = 1.0e-5, periodic_neighbors: int = 0,
                 periodic_size: tuple[float, float] = (1.0, 1.0)):
        self._lib = _lib.load()
        cfg = FhConfig()
        cfg.raysPerSide = rays_per_side
        cfg.sunRaysPerSide = sun_rays_per_side
        cfg.seed = seed
        cfg.gpuIndex = gpu_index
        cfg.rayEpsilon = ray_epsilon
        cfg.periodicNeighborCount = periodic_neighbors
        cfg.periodicSizeX, cfg.periodicSizeZ = periodic_size
        self.config = cfg
        h = C.c_void_p()
        st = self._lib.fh_create(C.byref(cfg), C.byref(h))
        if st != _lib.FH_OK:
            raise_for(st, None)
        self._h = h

    # -- lifecycle --------------------------------------------------------
    def destroy(self) -> None:
        if getattr(self, "_h", None):
            self._lib.fh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def _check(self, st: int) -> None:
        raise_for(st, self._h)

    @property
    def handle(self):
        return self._h

    def facetCount(self) -> int:
        return int(self._lib.fh_facet_count(self._h))

    def surfaceCount(self) -> int:
        return int(self._lib.fh_surface_count(self._h))

    # -- geometry and form factors -----------------------------------------
    def setGeometry(self, vertices) -> None:
        v = _f32(vertices)
        if v.ndim != 2 or v.shape[1] != 6:
            raise ValueError("vertices must be an (N, 6) array of position + normal")
        self._check(self._lib.fh_set_geometry(self._h, _ptr(v), v.shape[0]))

    def buildVisibilityGraph(self) -> GraphDiagnostics:
        d = FhGraphDiagnostics()
        self._check(self._lib.fh_build_form_factors(self._h, C.byref(d)))
        return _diag(d)

    def buildRasterGraph(self, directions, rasterWidth=512, rasterHeight=512,
                         maxFragmentsPerPixel=32) -> GraphDiagnostics:
        """buildVisibilityGraph(directions) by the reference's raster method
        (facetrt_vulkan.cpp:890-1061); the Config defaults (vkh:38-40)."""
        d3 = _f32(directions).reshape(-1, 3)
        dirs = (FhDirection * d3.shape[0])(*[FhDirection(float(x), float(y), float(z)) for x, y, z in d3])
        d = FhGraphDiagnostics()
        self._check(self._lib.fh_build_raster_graph(self._h, dirs, d3.shape[0], rasterWidth, rasterHeight,
                                                    maxFragmentsPerPixel, C.byref(d)))
        return _diag(d)

    def setGraph(self, rowOffsets, neighbors, counts, denominators, skyCounts) -> GraphDiagnostics:
        ro = np.ascontiguousarray(rowOffsets, dtype=np.uint32)
        nb = np.ascontiguousarray(neighbors, dtype=np.uint32)
        ct = np.ascontiguousarray(counts, dtype=np.uint32)
        de = np.ascontiguousarray(denominators, dtype=np.uint32)
        sk = np.ascontiguousarray(skyCounts, dtype=np.uint32)
        S = self.surfaceCount()
        if ro.size != S + 1 or de.size != S or sk.size != S or nb.size != ct.size:
            raise ValueError("graph array sizes do not match the geometry")
        d = FhGraphDiagnostics()
        self._check(self._lib.fh_set_graph(self._h, nb.size, _ptr(ro), _ptr(nb), _ptr(ct), _ptr(de), _ptr(sk),
                                           C.byref(d)))
        return _diag(d)

    def getGraph(self) -> Graph:
        S = self.surfaceCount()
        nnz = int(self._lib.fh_directed_edge_count(self._h))
        g = Graph(np.zeros(S + 1, np.uint32), np.zeros(nnz, np.uint32), np.zeros(nnz, np.uint32),
                  np.zeros(S, np.uint32), np.zeros(S, np.uint32))
        self._check(self._lib.fh_get_graph(self._h, _ptr(g.rowOffsets), _ptr(g.neighbors), _ptr(g.counts),
                                           _ptr(g.denominators), _ptr(g.skyCounts)))
        return g

    def computeSunlitFraction(self, direction) -> np.ndarray:
        S = self.surfaceCount()
        out = np.zeros(S, np.float32)
        d = FhDirection(float(direction[0]), float(direction[1]), float(direction[2]))
        self._check(self._lib.fh_compute_sunlit_fraction(self._h, d, _ptr(out)))
        return out

    def computeRasterSunlitFraction(self, direction, rasterWidth=512, rasterHeight=512,
                                    maxFragmentsPerPixel=32) -> np.ndarray:
        """computeSunlitFraction(direction) by the raster method
        (facetrt_vulkan.cpp:1064-1157)."""
        S = self.surfaceCount()
        out = np.zeros(S, np.float32)
        d = FhDirection(float(direction[0]), float(direction[1]), float(direction[2]))
        self._check(self._lib.fh_compute_raster_sunlit_fraction(self._h, d, rasterWidth, rasterHeight,
                                                                maxFragmentsPerPixel, _ptr(out)))
        return out

    # -- radiosity -----------------------------------------------------------
    def _optics(self, optics) -> np.ndarray:
        o = _f32(optics)
        S = self.surfaceCount()
        if o.shape != (S, 4):
            raise ValueError("Optics array must contain exactly two sides per facet")
        return o

    def solve(self, optics, skyRadiosity: float, iterations: int = 64, relaxation: float = 1.0) -> SolveResult:
        o = self._optics(optics)
        out = np.zeros(self.surfaceCount(), np.float32)
        st = FhSolveStats()
        self._check(self._lib.fh_solve(self._h, _ptr(o), skyRadiosity, iterations, relaxation, _ptr(out),
                                       C.byref(st)))
        return SolveResult(out, st.iterations, st.maxDelta, st.maxRelDelta, st.refreshes)

    def solveAccelerated(self, optics, skyRadiosity: float, initialRadiosity=None, maximumIterations: int = 64,
                         batchIterations: int = 8, tolerance: float = 1.0e-4,
                         relaxation: float = 1.0) -> SolveResult:
        o = self._optics(optics)
        init = None if initialRadiosity is None else _f32(initialRadiosity)
        if init is not None and init.size != self.surfaceCount():
            raise ValueError("Initial radiosity must contain one finite non-negative value per surface")
        out = np.zeros(self.surfaceCount(), np.float32)
        st = FhSolveStats()
        self._check(self._lib.fh_solve_accelerated(self._h, _ptr(o), skyRadiosity, _ptr(init), maximumIterations,
                                                   batchIterations, tolerance, relaxation, _ptr(out), C.byref(st)))
        return SolveResult(out, st.iterations, st.maxDelta, st.maxRelDelta, st.refreshes)

    def incidentIrradiance(self, solution, optics, skyRadiosity: float) -> np.ndarray:
        B = solution.radiosity if isinstance(solution, SolveResult) else solution
        B = _f32(B)
        o = self._optics(optics)
        if B.size != self.surfaceCount():
            raise ValueError("Incident irradiance requires a completed solve and visibility graph")
        out = np.zeros(self.surfaceCount(), np.float32)
        self._check(self._lib.fh_incident_irradiance(self._h, _ptr(B), _ptr(o), skyRadiosity, _ptr(out)))
        return out

    def solveBands(self, reflectance, transmittance, emission, direct, sky, init=None, maxIter: int = 1000,
                   tolRel: float = 1.0e-6, relaxation: float = 1.0,
                   precision: int = _lib.FH_PRECISION_FP32) -> SolveResult:
        S = self.surfaceCount()
        r = _f32(reflectance)
        nb = r.size // S
        shape = (S, nb)
        r = r.reshape(shape)
        t = _f32(transmittance, shape)
        e = _f32(emission, shape)
        D = _f32(direct, shape)
        sk = _f32(sky).reshape(-1)
        if sk.size != nb:
            raise ValueError("sky must hold one value per band")
        i0 = None if init is None else _f32(init, shape)
        out = np.zeros(shape, np.float32)
        st = FhSolveStats()
        self._check(self._lib.fh_solve_bands(self._h, nb, _ptr(r), _ptr(t), _ptr(e), _ptr(D), _ptr(sk), _ptr(i0),
                                             maxIter, tolRel, relaxation, precision, _ptr(out), C.byref(st)))
        return SolveResult(out, st.iterations, st.maxDelta, st.maxRelDelta, st.refreshes)

    def _spectral(self, mode, material, reflectance, transmittance, wavelength, sun, sunlit, direct_normal,
                  leaf_count, strength=0.0, factor=None, temperature=None, sunlit_temperature=None,
                  shaded_temperature=None):
        F, S = self.facetCount(), self.surfaceCount()
        R = _f32(reflectance)
        nm, nb = R.shape
        keep = {
            "mat": np.ascontiguousarray(material, np.uint32).reshape(-1),
            "r": R, "t": _f32(transmittance, (nm, nb)), "wl": _f32(wavelength).reshape(-1),
            "dni": _f32(direct_normal).reshape(-1), "sunlit": _f32(sunlit).reshape(-1),
            "k": None if factor is None else _f32(factor).reshape(-1),
            "T": None if temperature is None else _f32(temperature).reshape(-1),
            "ts": None if sunlit_temperature is None else _f32(sunlit_temperature).reshape(-1),
            "th": None if shaded_temperature is None else _f32(shaded_temperature).reshape(-1),
        }
        if keep["mat"].size != F or keep["wl"].size != nb or keep["dni"].size != nb or keep["sunlit"].size != S:
            raise ValueError("optics inputs have the wrong sizes")
        o = _lib.FhSpectralOptics()
        ad = {k: (None if v is None else v.ctypes.data) for k, v in keep.items()}
        o.mode, o.bands, o.materialCount, o.leafFacetCount = mode, nb, nm, leaf_count
        o.facetMaterial, o.reflectance, o.transmittance = ad["mat"], ad["r"], ad["t"]
        o.wavelength, o.reflectanceFactor = ad["wl"], ad["k"]
        o.backgroundAngularStrength = strength
        o.sun = FhDirection(*[float(x) for x in sun])
        o.sunlit, o.directNormal, o.temperature = ad["sunlit"], ad["dni"], ad["T"]
        o.sunlitTemperature, o.shadedTemperature = ad["ts"], ad["th"]
        return o, keep, nb

    def spectralOptics(self, mode, material, reflectance, transmittance, wavelength, sun, sunlit, direct_normal,
                       leaf_count, **kw):
        """Per-side, per-band optics built on the device (makeSpectralOptics /
        makeStateSpectralOptics, facetrtcore.cpp:1057-1199): (r, t, e, D),
        each [surfaceCount][bands]."""
        o, keep, nb = self._spectral(mode, material, reflectance, transmittance, wavelength, sun, sunlit,
                                     direct_normal, leaf_count, **kw)
        out = [np.zeros((self.surfaceCount(), nb), np.float32) for _ in range(4)]
        self._check(self._lib.fh_spectral_optics_build(self._h, C.byref(o), *[_ptr(a) for a in out]))
        return tuple(out)

    def solveSpectral(self, mode, material, reflectance, transmittance, wavelength, sun, sunlit, direct_normal,
                      leaf_count, sky, maxIter: int = 64, tolRel: float = -1.0, relaxation: float = 1.0,
                      precision: int = _lib.FH_PRECISION_FP32, **kw) -> SolveResult:
        o, keep, nb = self._spectral(mode, material, reflectance, transmittance, wavelength, sun, sunlit,
                                     direct_normal, leaf_count, **kw)
        sk = _f32(sky).reshape(-1)
        if sk.size != nb:
            raise ValueError("sky must hold one value per band")
        out = np.zeros((self.surfaceCount(), nb), np.float32)
        st = FhSolveStats()
        self._check(self._lib.fh_solve_spectral(self._h, C.byref(o), _ptr(sk), None, maxIter, tolRel, relaxation,
                                                precision, _ptr(out), C.byref(st)))
        return SolveResult(out, st.iterations, st.maxDelta, st.maxRelDelta, st.refreshes)

    def incidentBands(self, radiosity, direct, sky) -> np.ndarray:
        B = _f32(radiosity)
        S = self.surfaceCount()
        nb = B.size // S
        B = B.reshape(S, nb)
        D = _f32(direct, (S, nb))
        sk = _f32(sky).reshape(-1)
        out = np.zeros((S, nb), np.float32)
        self._check(self._lib.fh_incident_bands(self._h, nb, _ptr(B), _ptr(D), _ptr(sk), _ptr(out)))
        return out

    # -- FacetEB ---------------------------------------------------------------
    def ebSetBackgroundAngularStrength(self, strength: float) -> None:
        """scene.backgroundAngularStrength (facetrtcore.cpp:2904-2910): the
        Hapke ground of the energy balance, 0 = Lambertian."""
        self._check(self._lib.fh_eb_set_background_angular_strength(self._h, float(strength)))

    def ebSetScene(self, leaf_count: int, facet_material, materials, sw_reflectance=None, sw_transmittance=None):
        """materials: sequence of dicts with the FhEbMaterial fields; sw tables
        [materialCount][swBands] (window-centre optics)."""
        mat = np.ascontiguousarray(facet_material, dtype=np.uint32)
        arr = (FhEbMaterial * len(materials))()
        for i, m in enumerate(materials):
            for k, _ in FhEbMaterial._fields_:
                if k in m:
                    setattr(arr[i], k, m[k])
        nb = 0
        r = t = None
        if sw_reflectance is not None:
            r = _f32(sw_reflectance)
            t = _f32(sw_transmittance)
            nb = r.shape[1]
        self._check(self._lib.fh_eb_set_scene(self._h, leaf_count, _ptr(mat), len(materials),
                                              C.cast(arr, C.c_void_p), nb, _ptr(r), _ptr(t)))
        self._eb_bands = nb

    def ebRunNode(self, params: FhEbNodeParams, sw_direct_normal=None, sw_diffuse=None) -> FhEbNodeStats:
        d = None if sw_direct_normal is None else _f32(sw_direct_normal)
        f = None if sw_diffuse is None else _f32(sw_diffuse)
        st = FhEbNodeStats()
        self._check(self._lib.fh_eb_run_node(self._h, C.byref(params), _ptr(d), _ptr(f), C.byref(st)))
        return st

    def ebGetState(self) -> dict:
        S = self.surfaceCount()
        names = ("temperature", "absorbedShortwave", "longwaveRadiosity", "sunlit", "netRadiation",
                 "sensibleHeat", "latentHeat", "storageHeat")
        out = {k: np.zeros(S, np.float32) for k in names}
        nb = getattr(self, "_eb_bands", 0)
        sw = np.zeros((S, max(nb, 1)), np.float32) if nb else None
        self._check(self._lib.fh_eb_get_state(self._h, *[_ptr(out[k]) for k in names], _ptr(sw)))
        if sw is not None:
            out["shortwaveRadiosity"] = sw
        return out

    def ebGetPhotovoltaic(self) -> dict:
        """Per side: incident and absorbed shortwave, photovoltaic power."""
        S = self.surfaceCount()
        out = {k: np.zeros(S, np.float32) for k in ("incidentShortwave", "absorbedShortwave", "power")}
        self._check(self._lib.fh_eb_get_photovoltaic(self._h, _ptr(out["incidentShortwave"]),
                                                     _ptr(out["absorbedShortwave"]), _ptr(out["power"])))
        return out

    def ebSetReflectanceFactors(self, factors=None) -> None:
        """Per-facet factors on the shortwave reflectance (texture factors,
        facetrtcore.cpp:848-851, 965-972); None removes them."""
        f = None if factors is None else _f32(factors)
        if f is not None and f.size != self.facetCount():
            raise ValueError("one reflectance factor per facet is required")
        self._check(self._lib.fh_eb_set_reflectance_factors(self._h, _ptr(f)))

    def ebDiffuseRadiosity(self) -> tuple[np.ndarray, np.ndarray, FhSolveStats]:
        """The last node's windows solved without the direct irradiance
        (withoutDirectLight, facetrtcore.cpp:1561-1569, 3533-3536): per side
        the sum over the windows, per side and window, and the solve stats."""
        S = self.surfaceCount()
        nb = getattr(self, "_eb_bands", 0)
        total = np.zeros(S, np.float32)
        bands = np.zeros((S, max(nb, 1)), np.float32)
        st = FhSolveStats()
        self._check(self._lib.fh_eb_diffuse_radiosity(self._h, _ptr(total), _ptr(bands) if nb else None,
                                                      C.byref(st)))
        return total, bands[:, :nb], st

    def ebSetTemperature(self, T) -> None:
        T = _f32(T)
        self._check(self._lib.fh_eb_set_temperature(self._h, _ptr(T)))

    def ebUnitCount(self) -> int:
        return int(self._lib.fh_eb_unit_count(self._h))

    def ebGetFeatures(self):
        n = self.ebUnitCount()
        feat = np.zeros((n, 9), np.float32)
        T = np.zeros(n, np.float32)
        self._check(self._lib.fh_eb_get_features(self._h, _ptr(feat), _ptr(T)))
        return feat, T

    # -- surrogate MLP -----------------------------------------------------------
    @staticmethod
    def mlpParamCount(hidden: int) -> int:
        return int(_lib.load().fh_mlp_param_count(hidden))

    def mlpSetParams(self, hidden: int, params, feature_mean=None, feature_std=None, target_mean: float = 0.0,
                     target_std: float = 1.0) -> None:
        p = _f32(params)
        if p.size != self.mlpParamCount(hidden):
            raise ValueError("parameter count does not match the hidden width")
        fm = None if feature_mean is None else _f32(feature_mean)
        fs = None if feature_std is None else _f32(feature_std)
        self._check(self._lib.fh_mlp_set_params(self._h, hidden, _ptr(p), _ptr(fm), _ptr(fs), target_mean,
                                                target_std))
        self._mlp_hidden = hidden

    def mlpGetParams(self) -> np.ndarray:
        out = np.zeros(self.mlpParamCount(self._mlp_hidden), np.float32)
        self._check(self._lib.fh_mlp_get_params(self._h, _ptr(out)))
        return out

    def mlpForward(self, x) -> np.ndarray:
        x = _f32(x)
        y = np.zeros(x.shape[0], np.float32)
        self._check(self._lib.fh_mlp_forward(self._h, _ptr(x), x.shape[0], _ptr(y)))
        return y

    def mlpLossGrad(self, x, target):
        x = _f32(x)
        t = _f32(target)
        g = np.zeros(self.mlpParamCount(self._mlp_hidden), np.float32)
        loss = C.c_float(0.0)
        self._check(self._lib.fh_mlp_loss_grad(self._h, _ptr(x), _ptr(t), x.shape[0], _ptr(g), C.byref(loss)))
        return float(loss.value), g

    def mlpTrain(self, x, target, batch: int, epochs: int, lr: float, lr_decay: float = 1.0, seed: int = 0) -> float:
        x = _f32(x)
        t = _f32(target)
        loss = C.c_float(0.0)
        self._check(self._lib.fh_mlp_train(self._h, _ptr(x), _ptr(t), x.shape[0], batch, epochs, lr, lr_decay, seed,
                                           C.byref(loss)))
        return float(loss.value)

    # -- timing ------------------------------------------------------------------
    def benchmarkSweep(self, nb: int, precision: int, reps: int = 50, tables: bool = False) -> float:
        """Average SpMM sweep kernel duration in seconds (HIP events); tables:
        the energy balance's form (material tables, direct factors,
        Chebyshev weights) at band-parallel strides."""
        us = C.c_float(0.0)
        self._check(self._lib.fh_benchmark_sweep(self._h, nb, precision, 1 if tables else 0, reps, C.byref(us)))
        return us.value * 1e-6

    def benchmarkCopy(self, nbytes: int, reps: int = 20) -> float:
        """Average duration in seconds of a plain copy of nbytes (read + write)."""
        us = C.c_float(0.0)
        self._check(self._lib.fh_benchmark_copy(self._h, nbytes, reps, C.byref(us)))
        return us.value * 1e-6

    def synchronize(self) -> None:
        self._check(self._lib.fh_synchronize(self._h))

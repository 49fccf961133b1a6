"""ctypes access to the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "liboracle.so"

_P = C.c_void_p
_U = C.c_uint32
_F = C.c_float
_D = C.c_double
_I = C.c_int


class OrGraph(C.Structure):
    _fields_ = [("S", _U), ("nnz", _U), ("rowptr", _P), ("col", _P), ("cnt", _P), ("den", _P), ("sky", _P),
                ("overflow", _U), ("closure", _F)]


class OrEbMaterial(C.Structure):
    _fields_ = [("tirReflectance", _F), ("tirTransmittance", _F), ("surfaceResistance", _F), ("specificHeat", _F),
                ("density", _F), ("thermalConductivity", _F), ("soilMoisture", _F), ("heatCapacityPerArea", _F),
                ("convectiveScale", _F), ("soilTemperatureMethod", C.c_int32), ("vegetation", C.c_int32),
                ("photovoltaic", C.c_int32), ("pvEta25", _F), ("pvGamma", _F), ("pvBifaciality", _F)]


class OrEbParams(C.Structure):
    _fields_ = [("airTemperatureK", _F), ("vaporPressure", _F), ("pressure", _F), ("windSpeed", _F),
                ("skyLongwave", _F), ("dTime", _F), ("tolerance", _F), ("relaxation", _F), ("maxCoupling", _U),
                ("vegetationMethod", _U), ("lwMaxSweeps", _U), ("lwTolRel", _F),
                ("backgroundAngularStrength", _F)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)
    lib = C.CDLL(str(LIB))
    sig = {
        "or_graph_free": (None, [C.POINTER(OrGraph)]),
        "or_graph_copy": (None, [C.POINTER(OrGraph), _P, _P, _P, _P, _P]),
        "or_graph_from_arrays": (_I, [C.POINTER(OrGraph), _U, _U, _P, _P, _P, _P, _P]),
        "or_hemisphere_directions": (None, [_U, _P]),
        "or_hapke_reflectance": (_F, [_F, _P, _P, _P, _F]),
        "or_hapke_emissivity": (_F, [_F, _F, _F]),
        "or_sun_direction": (None, [_F, _F, _P]),
        "or_raster_graph": (_I, [_P, _U, _P, _P, _U, _U, _U, _F, _F, _P, _U, C.POINTER(OrGraph)]),
        "or_raster_sunlit": (_I, [_P, _U, _P, _P, _U, _U, _U, _F, _F, _P, _P]),
        "or_ray_graph": (_I, [_P, _U, _U, _U, _F, _I, C.POINTER(OrGraph), _P]),
        "or_ray_sunlit": (_I, [_P, _U, _P, _U, _U, _F, _I, _P]),
        "or_ray_graph_periodic": (_I, [_P, _U, _U, _U, _F, _I, _U, _F, _F, C.POINTER(OrGraph), _P]),
        "or_ray_sunlit_periodic": (_I, [_P, _U, _P, _U, _U, _F, _I, _U, _F, _F, _P]),
        "or_trace_one": (_U, [_P, _U, _P, _P, _U, _F, _P]),
        "or_ray_generate": (None, [_P, _U, _U, _U, _P, _P]),
        "or_solve": (None, [C.POINTER(OrGraph), _P, _F, _U, _F, _P, _P]),
        "or_solve_bands": (_U, [C.POINTER(OrGraph), _U, _P, _P, _P, _P, _P, _P, _U, _D, _F, _P, _P]),
        "or_set_threads": (None, [_I]),
        "or_max_threads": (_I, []),
        "or_incident": (None, [C.POINTER(OrGraph), _P, _P, _F, _P]),
        "or_incident_bands": (None, [C.POINTER(OrGraph), _U, _P, _P, _P, _P]),
        "or_solve_direct": (_I, [C.POINTER(OrGraph), _P, _F, _P]),
        "or_eb_node": (_U, [C.POINTER(OrGraph), _U, _U, _P, _P, _P, _P, _P, C.POINTER(OrEbParams), _P, _P, _P, _P,
                            _P, _P]),
        "or_sat_vp": (_D, [_D]),
        "or_sat_vp_slope": (_D, [_D]),
        "or_latent": (_D, [_D, _D, _D, _D, _D, _D]),
        "or_latent_slope": (_D, [_D, _D, _D, _D, _D, _D]),
        "or_soil_storage": (_D, [C.POINTER(OrEbMaterial), _D]),
        "or_mlp_forward": (None, [_P, _U, _U, _P, _U, _P]),
        "or_mlp_loss_grad": (_D, [_P, _U, _U, _P, _P, _U, _P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_P)


@dataclass
class Graph:
    rowOffsets: np.ndarray
    neighbors: np.ndarray
    counts: np.ndarray
    denominators: np.ndarray
    skyCounts: np.ndarray
    closure: float = 0.0
    overflow: int = 0

    @property
    def S(self) -> int:
        return self.denominators.size

    @property
    def nnz(self) -> int:
        return self.neighbors.size


def _take(g: OrGraph) -> Graph:
    lib = load()
    out = Graph(np.zeros(g.S + 1, np.uint32), np.zeros(g.nnz, np.uint32), np.zeros(g.nnz, np.uint32),
                np.zeros(g.S, np.uint32), np.zeros(g.S, np.uint32), g.closure, g.overflow)
    lib.or_graph_copy(C.byref(g), _p(out.rowOffsets), _p(out.neighbors), _p(out.counts), _p(out.denominators),
                      _p(out.skyCounts))
    lib.or_graph_free(C.byref(g))
    return out


class _Held:
    """An OrGraph whose arrays live in numpy buffers kept alive by this object."""

    def __init__(self, g: Graph):
        self.arrays = [np.ascontiguousarray(x, dtype=np.uint32) for x in
                       (g.rowOffsets, g.neighbors, g.counts, g.denominators, g.skyCounts)]
        self.s = OrGraph()
        self.s.S = g.S
        self.s.nnz = g.nnz
        self.s.rowptr, self.s.col, self.s.cnt, self.s.den, self.s.sky = [a.ctypes.data for a in self.arrays]

    def ref(self):
        return C.byref(self.s)


def hemisphere_directions(n: int = 40) -> np.ndarray:
    out = np.zeros((n, 3), np.float32)
    load().or_hemisphere_directions(n, _p(out))
    return out


def sun_direction(zenith: float, azimuth: float) -> np.ndarray:
    out = np.zeros(3, np.float32)
    load().or_sun_direction(zenith, azimuth, _p(out))
    return out


def raster_graph(vertices, bmin, bmax, raster=1024, layers=64, periodic=0, px=1.0, pz=1.0, dirs=None) -> Graph:
    v = np.ascontiguousarray(vertices, np.float32)
    F = v.shape[0] // 3
    d = hemisphere_directions(40) if dirs is None else np.ascontiguousarray(dirs, np.float32)
    g = OrGraph()
    rc = load().or_raster_graph(_p(v), F, _p(np.asarray(bmin, np.float32)), _p(np.asarray(bmax, np.float32)),
                                raster, layers, periodic, px, pz, _p(d), d.shape[0], C.byref(g))
    if rc == -2:
        raise MemoryError("oracle allocation failed")
    out = _take(g)
    if rc == -1:
        raise RuntimeError("CPU A-buffer overflow")
    return out


def raster_sunlit(vertices, bmin, bmax, sun, raster=1024, layers=64, periodic=0, px=1.0, pz=1.0) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32)
    F = v.shape[0] // 3
    out = np.zeros(2 * F, np.float32)
    rc = load().or_raster_sunlit(_p(v), F, _p(np.asarray(bmin, np.float32)), _p(np.asarray(bmax, np.float32)),
                                 raster, layers, periodic, px, pz, _p(np.asarray(sun, np.float32)), _p(out))
    if rc:
        raise RuntimeError("CPU sunlight A-buffer overflow")
    return out


def ray_graph(vertices, seed: int, rays: int, tmin: float = 1.0e-5, brute: bool = False, hits: bool = False,
              periodic: int = 0, size=(1.0, 1.0)):
    v = np.ascontiguousarray(vertices, np.float32)
    F = v.shape[0] // 3
    h = np.zeros(2 * F * rays, np.uint32) if hits else None
    g = OrGraph()
    rc = load().or_ray_graph_periodic(_p(v), F, seed, rays, tmin, int(brute), periodic, size[0], size[1], C.byref(g),
                                      _p(h))
    if rc:
        raise MemoryError("oracle allocation failed")
    out = _take(g)
    return (out, h) if hits else out


def ray_sunlit(vertices, sun, seed: int, K: int, tmin: float = 1.0e-5, brute: bool = False, periodic: int = 0,
               size=(1.0, 1.0)) -> np.ndarray:
    v = np.ascontiguousarray(vertices, np.float32)
    F = v.shape[0] // 3
    out = np.zeros(2 * F, np.float32)
    load().or_ray_sunlit_periodic(_p(v), F, _p(np.asarray(sun, np.float32)), seed, K, tmin, int(brute), periodic,
                                  size[0], size[1], _p(out))
    return out


def trace_one(vertices, o, d, skip: int, tmin: float = 1.0e-5):
    v = np.ascontiguousarray(vertices, np.float32)
    t = np.zeros(1, np.float32)
    side = load().or_trace_one(_p(v), v.shape[0] // 3, _p(np.asarray(o, np.float32)), _p(np.asarray(d, np.float32)),
                               skip, tmin, _p(t))
    return int(side), float(t[0])


def ray_generate(vertices, side: int, ray: int, seed: int):
    v = np.ascontiguousarray(vertices, np.float32)
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    load().or_ray_generate(_p(v), side, ray, seed, _p(o), _p(d))
    return o, d


def solve(g: Graph, optics, sky: float, iterations: int = 64, relaxation: float = 1.0):
    h = _Held(g)
    o = np.ascontiguousarray(optics, np.float32)
    out = np.zeros(g.S, np.float32)
    md = np.zeros(1, np.float32)
    load().or_solve(h.ref(), _p(o), sky, iterations, relaxation, _p(out), _p(md))
    return out, float(md[0])


def solve_bands(g: Graph, r, t, e, D, sky, init=None, max_iter=10000, tol_rel=1e-6, relaxation=1.0):
    h = _Held(g)
    r = np.ascontiguousarray(r, np.float32)
    nb = r.size // g.S
    arr = [np.ascontiguousarray(x, np.float32).reshape(g.S, nb) for x in (r, t, e, D)]
    sk = np.ascontiguousarray(sky, np.float32).reshape(-1)
    i0 = None if init is None else np.ascontiguousarray(init, np.float32).reshape(g.S, nb)
    out = np.zeros((g.S, nb), np.float32)
    last = np.zeros(1, np.float64)
    it = load().or_solve_bands(h.ref(), nb, *[_p(a) for a in arr], _p(sk), _p(i0), max_iter, tol_rel, relaxation,
                               _p(out), _p(last))
    return out, int(it), float(last[0])


def incident(g: Graph, B, optics, sky: float) -> np.ndarray:
    h = _Held(g)
    out = np.zeros(g.S, np.float32)
    load().or_incident(h.ref(), _p(np.ascontiguousarray(B, np.float32)), _p(np.ascontiguousarray(optics, np.float32)),
                       sky, _p(out))
    return out


def incident_bands(g: Graph, B, D, sky) -> np.ndarray:
    h = _Held(g)
    B = np.ascontiguousarray(B, np.float32)
    nb = B.size // g.S
    out = np.zeros((g.S, nb), np.float32)
    load().or_incident_bands(h.ref(), nb, _p(B), _p(np.ascontiguousarray(D, np.float32)),
                             _p(np.ascontiguousarray(sky, np.float32)), _p(out))
    return out


def solve_direct(g: Graph, optics, sky: float) -> np.ndarray:
    h = _Held(g)
    out = np.zeros(g.S, np.float64)
    rc = load().or_solve_direct(h.ref(), _p(np.ascontiguousarray(optics, np.float32)), sky, _p(out))
    if rc:
        raise RuntimeError("direct solve failed")
    return out


def hapke_reflectance(reflectance, incoming, outgoing, normal, strength) -> float:
    f = lambda v: np.ascontiguousarray(v, np.float32)  # noqa: E731
    return float(load().or_hapke_reflectance(reflectance, _p(f(incoming)), _p(f(outgoing)), _p(f(normal)),
                                             strength))


def hapke_emissivity(emissivity, cosine, strength) -> float:
    return float(load().or_hapke_emissivity(emissivity, cosine, strength))


def eb_node(g: Graph, F: int, leaf_count: int, material, mats, absorbed, params: OrEbParams, T, Tprev,
            incident=None, normal_y=None):
    h = _Held(g)
    mat = np.ascontiguousarray(material, np.uint32)
    marr = (OrEbMaterial * len(mats))(*mats)
    T = np.ascontiguousarray(T, np.float32).copy()
    Tp = np.ascontiguousarray(Tprev, np.float32)
    Blw = np.zeros(2 * F, np.float32)
    conv = C.c_int(0)
    mdt = np.zeros(1, np.float32)
    meq = np.zeros(1, np.float32)
    inc = None if incident is None else np.ascontiguousarray(incident, np.float32)
    ny = None if normal_y is None else np.ascontiguousarray(normal_y, np.float32)
    if params.backgroundAngularStrength > 0 and ny is None:
        raise ValueError("the Hapke ground needs the facet normals' y components")
    it = load().or_eb_node(h.ref(), F, leaf_count, _p(mat), C.cast(marr, _P), _p(ny),
                           _p(np.ascontiguousarray(absorbed, np.float32)), _p(inc), C.byref(params), _p(T), _p(Tp),
                           _p(Blw),
                           C.cast(C.pointer(conv), _P), _p(mdt), _p(meq))
    return T, Blw, int(it), bool(conv.value), float(mdt[0]), float(meq[0])


def sat_vp(T: float) -> float:
    """Saturation vapour pressure (hPa) at T (K), core:854-858."""
    return float(load().or_sat_vp(T))


def sat_vp_slope(T: float) -> float:
    return float(load().or_sat_vp_slope(T))


def latent(T, rho, P, ea, ra, rs) -> float:
    """Latent heat flux (W m-2), core:867-883."""
    return float(load().or_latent(T, rho, P, ea, ra, rs))


def latent_slope(T, rho, P, ea, ra, rs) -> float:
    return float(load().or_latent_slope(T, rho, P, ea, ra, rs))


def soil_storage(material: OrEbMaterial, dtime: float) -> float:
    """Soil storage coefficient (W m-2 K-1), core:909-915."""
    return float(load().or_soil_storage(C.byref(material), dtime))


def mlp_forward(params, x, hidden: int) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    n, fin = x.shape
    y = np.zeros(n, np.float64)
    load().or_mlp_forward(_p(np.ascontiguousarray(params, np.float32)), fin, hidden, _p(x), n, _p(y))
    return y


def mlp_loss_grad(params, x, target, hidden: int):
    x = np.ascontiguousarray(x, np.float32)
    n, fin = x.shape
    p = np.ascontiguousarray(params, np.float32)
    g = np.zeros(p.size, np.float64)
    loss = load().or_mlp_loss_grad(_p(p), fin, hidden, _p(x), _p(np.ascontiguousarray(target, np.float32)), n, _p(g))
    return float(loss), g


def set_threads(n: int) -> None:
    """Threads of the oracle's multi-band solve (results do not depend on it)."""
    load().or_set_threads(int(n))


def max_threads() -> int:
    return int(load().or_max_threads())

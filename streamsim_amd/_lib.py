"""ctypes binding of ``libfacet_hip.so`` (the C-ABI in ``include/facet_hip.h``).

The library is built in-tree by ``python -m streamsim_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback for the engine.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libfacet_hip.so"

FH_OK = 0
FH_INVALID_ARGUMENT = 1
FH_LOGIC_ERROR = 2
FH_RUNTIME_ERROR = 3
FH_OVERFLOW_ERROR = 4

FH_PRECISION_FP32 = 0
FH_PRECISION_BF16_DC = 1
FH_PRECISION_FP8_INC = 2
FH_PRECISION_BF16_F = 3

FH_WARM_PREVIOUS = 0
FH_WARM_SURROGATE = 1


class FhConfig(C.Structure):
    _fields_ = [
        ("raysPerSide", C.c_uint32),
        ("sunRaysPerSide", C.c_uint32),
        ("seed", C.c_uint32),
        ("gpuIndex", C.c_uint32),
        ("rayEpsilon", C.c_float),
        ("periodicNeighborCount", C.c_uint32),
        ("periodicSizeX", C.c_float),
        ("periodicSizeZ", C.c_float),
    ]


class FhDirection(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


FH_OPTICS_SCENE = 0
FH_OPTICS_STATE = 1


class FhSpectralOptics(C.Structure):
    _fields_ = [
        ("mode", C.c_uint32), ("bands", C.c_uint32), ("materialCount", C.c_uint32), ("leafFacetCount", C.c_uint32),
        ("facetMaterial", C.c_void_p), ("reflectance", C.c_void_p), ("transmittance", C.c_void_p),
        ("wavelength", C.c_void_p), ("reflectanceFactor", C.c_void_p), ("backgroundAngularStrength", C.c_float),
        ("sun", FhDirection), ("sunlit", C.c_void_p), ("directNormal", C.c_void_p), ("temperature", C.c_void_p),
        ("sunlitTemperature", C.c_void_p), ("shadedTemperature", C.c_void_p),
    ]


class FhGraphDiagnostics(C.Structure):
    _fields_ = [
        ("facetCount", C.c_uint32),
        ("surfaceCount", C.c_uint32),
        ("directedEdgeCount", C.c_uint32),
        ("fragmentOverflow", C.c_uint32),
        ("hashOverflow", C.c_uint32),
        ("maxClosureError", C.c_float),
    ]


class FhSolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_uint32),
        ("maxDelta", C.c_float),
        ("maxRelDelta", C.c_float),
        ("refreshes", C.c_uint32),
    ]


class FhEbMaterial(C.Structure):
    _fields_ = [
        ("tirReflectance", C.c_float), ("tirTransmittance", C.c_float), ("surfaceResistance", C.c_float),
        ("specificHeat", C.c_float), ("density", C.c_float), ("thermalConductivity", C.c_float),
        ("soilMoisture", C.c_float), ("heatCapacityPerArea", C.c_float), ("convectiveScale", C.c_float),
        ("soilTemperatureMethod", C.c_int32), ("vegetation", C.c_int32), ("photovoltaic", C.c_int32),
        ("pvEta25", C.c_float), ("pvGamma", C.c_float), ("pvBifaciality", C.c_float),
    ]


class FhEbNodeParams(C.Structure):
    _fields_ = [
        ("sun", FhDirection), ("airTemperatureK", C.c_float), ("vaporPressure", C.c_float),
        ("pressure", C.c_float), ("windSpeed", C.c_float), ("skyLongwave", C.c_float), ("dTime", C.c_float),
        ("temperatureTolerance", C.c_float), ("temperatureRelaxation", C.c_float), ("maxCoupling", C.c_uint32),
        ("vegetationMethod", C.c_uint32), ("radiosityTolerance", C.c_float), ("maxSweeps", C.c_uint32),
        ("precision", C.c_uint32), ("warmStart", C.c_uint32), ("firstNode", C.c_uint32),
        ("zeroStart", C.c_uint32), ("recordFeatures", C.c_uint32), ("adaptiveRelaxation", C.c_uint32),
    ]


class FhEbNodeStats(C.Structure):
    _fields_ = [
        ("couplingIterations", C.c_uint32), ("converged", C.c_uint32), ("swSweeps", C.c_uint32),
        ("lwSweeps", C.c_uint32), ("maxTemperatureDelta", C.c_float), ("maxResidualEquivalent", C.c_float),
        ("swRefreshes", C.c_uint32), ("maxEnergyResidual", C.c_float), ("swInput", C.c_double),
        ("swAbsorbed", C.c_double), ("swEscaped", C.c_double), ("netRadiation", C.c_double),
        ("swMaxDelta", C.c_float), ("lwFinalSweeps", C.c_uint32), ("lwMaxDelta", C.c_float),
    ]


_P = C.c_void_p
_U32 = C.c_uint32
_F = C.c_float
_I = C.c_int

# name -> (restype, argtypes).  Every symbol include/facet_hip.h declares.
SIGNATURES = {
    "fh_status_string": (C.c_char_p, [_I]),
    "fh_last_error": (C.c_char_p, [_P]),
    "fh_create": (_I, [C.POINTER(FhConfig), C.POINTER(_P)]),
    "fh_destroy": (None, [_P]),
    "fh_set_geometry": (_I, [_P, _P, _U32]),
    "fh_build_form_factors": (_I, [_P, C.POINTER(FhGraphDiagnostics)]),
    "fh_set_graph": (_I, [_P, _U32, _P, _P, _P, _P, _P, C.POINTER(FhGraphDiagnostics)]),
    "fh_directed_edge_count": (_U32, [_P]),
    "fh_get_graph": (_I, [_P, _P, _P, _P, _P, _P]),
    "fh_solve": (_I, [_P, _P, _F, _U32, _F, _P, C.POINTER(FhSolveStats)]),
    "fh_solve_accelerated": (_I, [_P, _P, _F, _P, _U32, _U32, _F, _F, _P, C.POINTER(FhSolveStats)]),
    "fh_incident_irradiance": (_I, [_P, _P, _P, _F, _P]),
    "fh_compute_sunlit_fraction": (_I, [_P, FhDirection, _P]),
    "fh_build_raster_graph": (_I, [_P, _P, _U32, _U32, _U32, _U32, C.POINTER(FhGraphDiagnostics)]),
    "fh_compute_raster_sunlit_fraction": (_I, [_P, FhDirection, _U32, _U32, _U32, _P]),
    "fh_facet_count": (_U32, [_P]),
    "fh_surface_count": (_U32, [_P]),
    "fh_solve_bands": (_I, [_P, _U32, _P, _P, _P, _P, _P, _P, _U32, _F, _F, _U32, _P, C.POINTER(FhSolveStats)]),
    "fh_spectral_optics_build": (_I, [_P, C.POINTER(FhSpectralOptics), _P, _P, _P, _P]),
    "fh_solve_spectral": (_I, [_P, C.POINTER(FhSpectralOptics), _P, _P, _U32, _F, _F, _U32, _P,
                               C.POINTER(FhSolveStats)]),
    "fh_solve_bands_device": (_I, [_P, _U32, _P, _P, _P, _P, _P, _P, _U32, _F, _F, _U32, _P,
                                   C.POINTER(FhSolveStats)]),
    "fh_incident_bands": (_I, [_P, _U32, _P, _P, _P, _P]),
    "fh_sweep_device": (_I, [_P, _U32, _P, _P, _P, _P, _P, _U32]),
    "fh_sweep_device_noalloc": (_I, [_P, _U32, _P, _P, _P, _P, _P, _U32]),
    "fh_band_stride": (_U32, [_U32]),
    "fh_stream": (_P, [_P]),
    "fh_eb_set_background_angular_strength": (_I, [_P, _F]),
    "fh_eb_set_scene": (_I, [_P, _U32, _P, _U32, _P, _U32, _P, _P]),
    "fh_eb_run_node": (_I, [_P, C.POINTER(FhEbNodeParams), _P, _P, C.POINTER(FhEbNodeStats)]),
    "fh_eb_get_state": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "fh_eb_get_photovoltaic": (_I, [_P, _P, _P, _P]),
    "fh_eb_set_temperature": (_I, [_P, _P]),
    "fh_eb_diffuse_radiosity": (_I, [_P, _P, _P, C.POINTER(FhSolveStats)]),
    "fh_eb_set_reflectance_factors": (_I, [_P, _P]),
    "fh_mlp_param_count": (_U32, [_U32]),
    "fh_mlp_set_params": (_I, [_P, _U32, _P, _P, _P, _F, _F]),
    "fh_mlp_get_params": (_I, [_P, _P]),
    "fh_mlp_forward": (_I, [_P, _P, _U32, _P]),
    "fh_mlp_loss_grad": (_I, [_P, _P, _P, _U32, _P, _P]),
    "fh_mlp_train": (_I, [_P, _P, _P, _U32, _U32, _U32, _F, _F, _U32, _P]),
    "fh_benchmark_sweep": (_I, [_P, _U32, _U32, _U32, _U32, _P]),
    "fh_synchronize": (_I, [_P]),
    "fh_benchmark_copy": (_I, [_P, C.c_uint64, _U32, _P]),
    "fh_eb_unit_count": (_U32, [_P]),
    "fh_eb_get_features": (_I, [_P, _P, _P]),
}

_lib: C.CDLL | None = None


def load() -> C.CDLL:
    """Load the engine library (raises ImportError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"HIP engine library missing: {LIB_PATH} (build it with `python -m streamsim_amd.build`)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class LogicError(RuntimeError):
    """Call-order error (std::logic_error in the reference)."""


def raise_for(status: int, ctx) -> None:
    if status == FH_OK:
        return
    lib = load()
    msg = lib.fh_last_error(ctx)
    text = msg.decode() if msg else lib.fh_status_string(status).decode()
    if status == FH_INVALID_ARGUMENT:
        raise ValueError(text)
    if status == FH_LOGIC_ERROR:
        raise LogicError(text)
    if status == FH_OVERFLOW_ERROR:
        raise OverflowError(text)
    raise RuntimeError(text)

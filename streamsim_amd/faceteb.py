"""FacetEB host driver: materials, shortwave windows, meteorology, diurnal run.

Restates the host side of runFacetEBCore (models/histream/src/facetrt/
facetrtcore.cpp, "core") around the device node step ``fh_eb_run_node``:

* material defaults: SurfaceParameters (core:113-143) -- leaf and the ground
  fallback {0.20}, {0.0}, tir 0.05 (core:615-619);
* shortwave windows: the flat per-micrometre spectrum the reference falls
  back to when no Esun/Esky files are given (core:2522-2529), split into
  windows of ``spectralAccelerationWidth`` 1-nm samples (core:3495-3565):
  window DNI = directNormalScale * integral(direct), diffuse = SW *
  integral(diffuse), directNormalScale = min(1400, SW * I_d / cos z) / I_d;
* energy-balance air constants clamp(Tair, 220, 340) etc. (core:3567-3578).

Synthetic diurnal meteorology (the benchmark has no meteorological file):
Tair = T0 + A sin(2 pi (h - 9) / 24), solar geometry from the declination /
hour-angle formulas (Cooper 1969 declination), global SW = 0.75 * 1000 *
cos z, sky LW = 0.75 sigma Tair^4.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ._lib import FhDirection, FhEbNodeParams

SIGMA = 5.670374419e-8
N_SAMPLES = 2001  # 400..2400 nm at 1 nm (N1 of the reference's spectra)

# SurfaceParameters defaults (core:113-143)
LEAF = dict(tirReflectance=0.03, tirTransmittance=0.0, surfaceResistance=120.0, specificHeat=1180.0,
            density=1800.0, thermalConductivity=1.55, soilMoisture=0.25, heatCapacityPerArea=60000.0,
            convectiveScale=1.0, soilTemperatureMethod=1, vegetation=1)
SOIL = dict(LEAF, tirReflectance=0.05)
BUILDING = dict(LEAF, tirReflectance=0.10, vegetation=0, heatCapacityPerArea=200000.0)


# Broad-band leaf / soil / facade spectra (piecewise linear in nm); window
# optics are sampled at the window centre (makeEnergyShortwaveOptics takes
# the centre sample, core:3512-3527).
_WL = np.array([400, 500, 680, 720, 800, 1100, 1300, 1450, 1600, 1900, 2000, 2200, 2400], float)
_LEAF_R = np.array([0.06, 0.08, 0.05, 0.20, 0.45, 0.45, 0.40, 0.25, 0.32, 0.12, 0.10, 0.15, 0.08])
_LEAF_T = np.array([0.03, 0.06, 0.03, 0.20, 0.45, 0.45, 0.40, 0.25, 0.32, 0.10, 0.08, 0.13, 0.06])
_SOIL_R = np.array([0.08, 0.10, 0.14, 0.15, 0.18, 0.24, 0.27, 0.26, 0.30, 0.28, 0.27, 0.29, 0.25])
_BLDG_R = np.full(_WL.shape, 0.30)


def spectrum(kind: str, wavelengths_nm) -> tuple[np.ndarray, np.ndarray]:
    w = np.asarray(wavelengths_nm, float)
    if kind == "leaf":
        return np.interp(w, _WL, _LEAF_R), np.interp(w, _WL, _LEAF_T)
    if kind == "soil":
        return np.interp(w, _WL, _SOIL_R), np.zeros_like(w)
    if kind == "building":
        return np.interp(w, _WL, _BLDG_R), np.zeros_like(w)
    raise ValueError(kind)


@dataclass
class ShortwaveWindows:
    """Windows of the flat fallback spectrum (core:2522-2529, 3495-3565)."""
    width: int
    first: np.ndarray
    last: np.ndarray
    direct_fraction: float = 0.8

    @classmethod
    def for_bands(cls, bands: int, direct_fraction: float = 0.8) -> "ShortwaveWindows":
        width = int(math.ceil((N_SAMPLES - 1) / bands))
        first, last = [], []
        f = 0
        while f + 1 < N_SAMPLES:
            first.append(f)
            last.append(min(f + width, N_SAMPLES - 1))
            f += width
        return cls(width, np.array(first), np.array(last), direct_fraction)

    @property
    def count(self) -> int:
        return self.first.size

    @property
    def centre_nm(self) -> np.ndarray:
        return 400.0 + (self.first + self.last) // 2

    def _integral(self, value: float, a: int, b: int) -> float:
        # trapezoid over 1-nm samples of a constant per-micrometre spectrum
        return float(np.float32(value * (b - a) * 0.001))

    def irradiance(self, shortwave: float, cos_zenith: float) -> tuple[np.ndarray, np.ndarray]:
        d = 0.5 * self.direct_fraction
        q = 0.5 * (1.0 - self.direct_fraction)
        Id = self._integral(d, 0, N_SAMPLES - 1)
        scale = (min(1400.0, shortwave * Id / cos_zenith) / Id) if (cos_zenith > 1.0e-4 and Id > 1e-8) else 0.0
        dni = np.array([scale * self._integral(d, a, b) for a, b in zip(self.first, self.last)], np.float32)
        dif = np.array([shortwave * self._integral(q, a, b) for a, b in zip(self.first, self.last)], np.float32)
        if shortwave <= 0.0:
            dni[:] = 0.0
            dif[:] = 0.0
        return dni, dif


def material_tables(kinds: list[str], windows: ShortwaveWindows):
    r = np.zeros((len(kinds), windows.count), np.float32)
    t = np.zeros_like(r)
    for i, k in enumerate(kinds):
        rr, tt = spectrum(k, windows.centre_nm)
        r[i] = rr
        t[i] = tt
    return r, t


def solar_position(day_of_year: int, solar_hour: float, latitude_deg: float) -> tuple[float, float]:
    """(zenith, azimuth clockwise from north) in degrees."""
    decl = math.radians(23.45) * math.sin(2.0 * math.pi * (284 + day_of_year) / 365.0)
    lat = math.radians(latitude_deg)
    h = math.radians(15.0 * (solar_hour - 12.0))
    cz = math.sin(lat) * math.sin(decl) + math.cos(lat) * math.cos(decl) * math.cos(h)
    cz = max(-1.0, min(1.0, cz))
    zen = math.acos(cz)
    az = math.atan2(math.sin(h), math.cos(h) * math.sin(lat) - math.tan(decl) * math.cos(lat)) + math.pi
    return math.degrees(zen), math.degrees(az) % 360.0


def sun_vector(zenith_deg: float, azimuth_deg: float) -> np.ndarray:
    """+X north, +Z east, +Y up (sunDirection, core:491-498)."""
    z = math.radians(zenith_deg)
    a = math.radians(azimuth_deg)
    return np.array([math.sin(z) * math.cos(a), math.cos(z), math.sin(z) * math.sin(a)], np.float32)


@dataclass
class Meteo:
    hour: float
    air_temperature_k: float
    vapor_pressure: float
    pressure: float
    wind: float
    shortwave: float
    longwave: float
    zenith: float
    azimuth: float


def diurnal_meteo(steps: int = 24, t_mean: float = 298.0, t_amp: float = 6.0, wind: float = 2.0,
                  day_of_year: int = 180, latitude: float = 40.0) -> list[Meteo]:
    out = []
    for k in range(steps):
        h = (k + 0.5) * 24.0 / steps
        zen, az = solar_position(day_of_year, h, latitude)
        cz = math.cos(math.radians(zen))
        ta = t_mean + t_amp * math.sin(2.0 * math.pi * (h - 9.0) / 24.0)
        ta = min(max(ta, 220.0), 340.0)
        sw = 0.75 * 1000.0 * max(0.0, cz)
        lw = 0.75 * SIGMA * ta ** 4
        out.append(Meteo(h, ta, 15.0, 1000.0, wind, sw, lw, zen, az))
    return out


@dataclass
class EbConfig:
    tolerance: float = 0.01
    relaxation: float = 0.5
    max_coupling: int = 60
    vegetation_method: int = 0
    radiosity_tolerance: float = 1.0e-5
    max_sweeps: int = 2000
    dtime: float = 3600.0
    adaptive_relaxation: bool = True


def node_params(m: Meteo, cfg: EbConfig, first: bool, precision: int = 0, warm: int = 0, zero_start: bool = False,
                record: bool = False) -> FhEbNodeParams:
    p = FhEbNodeParams()
    s = sun_vector(m.zenith, m.azimuth)
    p.sun = FhDirection(float(s[0]), float(s[1]), float(s[2]))
    p.airTemperatureK = m.air_temperature_k
    p.vaporPressure = m.vapor_pressure
    p.pressure = m.pressure
    p.windSpeed = m.wind
    p.skyLongwave = m.longwave
    p.dTime = cfg.dtime
    p.temperatureTolerance = cfg.tolerance
    p.temperatureRelaxation = cfg.relaxation
    p.maxCoupling = cfg.max_coupling
    p.vegetationMethod = cfg.vegetation_method
    p.radiosityTolerance = cfg.radiosity_tolerance
    p.maxSweeps = cfg.max_sweeps
    p.precision = precision
    p.warmStart = warm
    p.firstNode = 1 if first else 0
    p.zeroStart = 1 if zero_start else 0
    p.recordFeatures = 1 if record else 0
    p.adaptiveRelaxation = 1 if cfg.adaptive_relaxation else 0
    return p


@dataclass
class RunStats:
    nodes: int = 0
    coupling: list = field(default_factory=list)
    sw_sweeps: list = field(default_factory=list)
    lw_sweeps: list = field(default_factory=list)
    converged: list = field(default_factory=list)
    max_dt: list = field(default_factory=list)
    closure: list = field(default_factory=list)

    def add(self, st) -> None:
        self.nodes += 1
        self.coupling.append(st.couplingIterations)
        self.sw_sweeps.append(st.swSweeps)
        self.lw_sweeps.append(st.lwSweeps)
        self.converged.append(bool(st.converged))
        self.max_dt.append(st.maxTemperatureDelta)
        if st.swInput > 0:
            self.closure.append((st.swInput - st.swAbsorbed - st.swEscaped) / st.swInput)

    def summary(self) -> dict:
        return dict(nodes=self.nodes, coupling_iterations=int(np.sum(self.coupling)),
                    coupling_per_node=[int(x) for x in self.coupling],
                    sw_sweeps=int(np.sum(self.sw_sweeps)), lw_sweeps=int(np.sum(self.lw_sweeps)),
                    all_converged=bool(all(self.converged)),
                    max_sw_closure=float(np.max(np.abs(self.closure))) if self.closure else 0.0)


def setup_scene(engine, scene, windows: ShortwaveWindows, kinds=("leaf", "soil", "building")) -> None:
    mats = {"leaf": LEAF, "soil": SOIL, "building": BUILDING}
    r, t = material_tables(list(kinds), windows)
    engine.ebSetScene(scene.leaf_count, scene.material, [mats[k] for k in kinds], r, t)


def run_diurnal(engine, meteo: list[Meteo], windows: ShortwaveWindows, cfg: EbConfig, precision: int = 0,
                warm: int = 0, zero_start: bool = False, record=None) -> RunStats:
    """Runs the nodes in order on the device; ``record(engine, k)`` is called
    after every node when given (dataset capture)."""
    stats = RunStats()
    for k, m in enumerate(meteo):
        cz = math.cos(math.radians(m.zenith))
        dni, dif = windows.irradiance(m.shortwave, cz)
        p = node_params(m, cfg, first=(k == 0), precision=precision, warm=warm, zero_start=zero_start,
                        record=record is not None)
        st = engine.ebRunNode(p, dni, dif)
        stats.add(st)
        if record is not None:
            record(engine, k)
    return stats

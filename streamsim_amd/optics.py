"""Per-side optics builders for FacetRT runs (host side).

Restates makeOptics (models/histream/src/facetrt/facetrtcore.cpp:500-522)
and the spectral band optics of makeSpectralOptics (core:1057-1126) for
explicit per-band material values, background angular strength 0 and no
reflectance texture, plus planckRadiance (core:524-536).  Arrays are
[surfaceCount][bands]; side 2f carries the facet normal, side 2f+1 its
negation.
"""
from __future__ import annotations

import numpy as np

BEAM = np.float32(700.0)     # kBeamNormalIrradiance (core:84)
SKY = np.float32(100.0)      # kSkyDiffuseIrradiance (core:85)


def side_normals(scene) -> np.ndarray:
    n = scene.facet_normals().astype(np.float32)
    out = np.empty((2 * n.shape[0], 3), np.float32)
    out[0::2] = n
    out[1::2] = -n
    return out


def _cos_sun(scene, sun) -> np.ndarray:
    n = side_normals(scene)
    s = np.asarray(sun, np.float32)
    d = n[:, 0] * s[0] + n[:, 1] * s[1] + n[:, 2] * s[2]
    return np.maximum(np.float32(0.0), d)


def make_optics(scene, sunlit, sun) -> np.ndarray:
    """makeOptics: leaf r 0.10 t 0.05, ground r 0.20; D = sunlit max(0, n.s) 700."""
    S = scene.surface_count
    leaf = np.repeat(np.arange(scene.facet_count) < scene.leaf_count, 2)
    o = np.zeros((S, 4), np.float32)
    o[:, 0] = np.where(leaf, np.float32(0.10), np.float32(0.20))
    o[:, 1] = np.where(leaf, np.float32(0.05), np.float32(0.0))
    o[:, 3] = np.asarray(sunlit, np.float32) * _cos_sun(scene, sun) * BEAM
    return o


def planck_radiance(wavelength_nm: float, temperature: float) -> float:
    c1 = 1.1910439340652e8
    c2 = 14388.291040407
    wl = wavelength_nm / 1000.0 if wavelength_nm > 50.0 else wavelength_nm
    ex = c2 / (max(1.0, temperature) * max(1.0e-6, wl))
    return float(np.float32(c1 / (wl ** 5 * np.expm1(ex))))


def spectral_optics(scene, sunlit, sun, wavelengths_nm, leaf_r, leaf_t, soil_r, tir=None):
    """Band optics as makeSpectralOptics for explicit per-band values.

    tir (thermal bands, wavelength > 2500 nm): dict with leaf/soil
    tirReflectance, tirTransmittance and sunlit/shaded temperatures.
    Returns r, t, e, D ([S][nb]) and the per-band sky radiance (50 for
    optical bands, Planck(sky T) for thermal bands, core:1705-1708)."""
    S = scene.surface_count
    nb = len(wavelengths_nm)
    leaf = np.repeat(np.arange(scene.facet_count) < scene.leaf_count, 2)
    cos = _cos_sun(scene, sun)
    sl = np.asarray(sunlit, np.float32)
    r = np.zeros((S, nb), np.float32)
    t = np.zeros_like(r)
    e = np.zeros_like(r)
    D = np.zeros_like(r)
    sky = np.zeros(nb, np.float32)
    for b, wl in enumerate(wavelengths_nm):
        thermal = wl > 2500.0
        if not thermal:
            rr = np.where(leaf, np.float32(leaf_r[b]), np.float32(soil_r[b])).astype(np.float32)
            tt = np.where(leaf, np.float32(leaf_t[b]), np.float32(0.0)).astype(np.float32)
            rr = np.clip(rr, 0.0, 1.0)
            tt = np.minimum(np.clip(tt, 0.0, 1.0 - rr), 1.0 - rr)
            r[:, b] = rr
            t[:, b] = tt
            D[:, b] = sl * cos * BEAM
            sky[b] = np.float32(SKY * np.float32(0.5))
        else:
            lr = np.float32(tir["leaf_r"]); lt = np.float32(tir["leaf_t"]); gr = np.float32(tir["soil_r"])
            rr = np.where(leaf, lr, gr).astype(np.float32)
            tt = np.where(leaf, lt, np.float32(0.0)).astype(np.float32)
            tt = np.clip(tt, 0.0, 1.0 - rr)
            eps = np.maximum(np.float32(0.0), np.float32(1.0) - rr - tt)
            ill = np.clip(sl, 0.0, 1.0)
            ps_l = planck_radiance(wl, tir["leaf_sunlit_T"]); pd_l = planck_radiance(wl, tir["leaf_shaded_T"])
            ps_g = planck_radiance(wl, tir["soil_sunlit_T"]); pd_g = planck_radiance(wl, tir["soil_shaded_T"])
            ps = np.where(leaf, np.float32(ps_l), np.float32(ps_g))
            pd = np.where(leaf, np.float32(pd_l), np.float32(pd_g))
            r[:, b] = np.clip(np.float32(1.0) - tt - eps, 0.0, 1.0)
            t[:, b] = tt
            e[:, b] = eps * (ill * ps + (np.float32(1.0) - ill) * pd)
            sky[b] = planck_radiance(wl, tir["sky_T"])
    return r, t, e, D, sky

"""Independent pins of the oracle's FacetEB node (oracle/facet_oracle.c,
or_eb_node and its flux helpers; reference core:854-915, 3567-3846).

The flux helpers are checked against published values and against what the
mathematics fixes (finite differences, the step response of a semi-infinite
conductor); each branch of the node's surface balance is then checked by
bisecting that branch's scalar balance on an isolated facet, whose sides see
only the sky, so the longwave incident is the sky's.  A dropped or doubled
term, a wrong sign or a T / Tprev mix-up in the node loop moves its fixed
point away from the bisection root.
"""
import math

import numpy as np
import pytest

SIGMA = 5.670374419e-8


def _material(oracle, **kw):
    m = oracle.OrEbMaterial()
    vals = dict(tirReflectance=0.05, tirTransmittance=0.0, surfaceResistance=120.0, specificHeat=1180.0,
                density=1800.0, thermalConductivity=1.55, soilMoisture=0.0, heatCapacityPerArea=60000.0,
                convectiveScale=1.0, soilTemperatureMethod=0, vegetation=0)
    vals.update(kw)
    for k, v in vals.items():
        setattr(m, k, v)
    return m


def _params(oracle, Ta=298.0, ea=15.0, P=1000.0, u=2.0, L=350.0, dt=3600.0, veg_method=0, relax=0.5):
    p = oracle.OrEbParams()
    p.airTemperatureK = Ta
    p.vaporPressure = ea
    p.pressure = P
    p.windSpeed = u
    p.skyLongwave = L
    p.dTime = dt
    p.tolerance = 1e-4  # T is stored in FP32: its ulp near 300 K is 3e-5 K
    p.relaxation = relax
    p.maxCoupling = 2000
    p.vegetationMethod = veg_method
    p.lwMaxSweeps = 50
    p.lwTolRel = 1e-9
    return p


def _isolated_graph(oracle):
    # one facet whose two sides see only the sky (64 of 64 samples escape)
    S = 2
    return oracle.Graph(np.zeros(S + 1, np.uint32), np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                        np.full(S, 64, np.uint32), np.full(S, 64, np.uint32))


def _air(p):
    """Air constants of the node (core:3567-3578)."""
    ha = 5.8 + 4.1 * math.sqrt(max(0.1, p.windSpeed))
    P = max(100.0, p.pressure)
    rho = P * 100.0 / (287.05 * p.airTemperatureK)
    ra = rho * 1005.0 / max(0.1, ha)
    return ha, P, rho, ra


def _bisect(f, lo=220.0, hi=360.0):
    flo = f(lo)
    assert flo * f(hi) < 0.0, "no root in the bracket"
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        fm = f(mid)
        if flo * fm <= 0.0:
            hi = mid
        else:
            lo, flo = mid, fm
    return 0.5 * (lo + hi)


# --- flux helpers ---------------------------------------------------------

def test_saturation_vapour_pressure_published_values(oracle):
    # Magnus-Tetens over water (Tetens 1930; Murray 1967 tables):
    # 6.11 hPa at 0 C, 23.37 hPa at 20 C, 42.43 hPa at 30 C, 12.27 at 10 C
    for tc, want in ((0.0, 6.11), (10.0, 12.27), (20.0, 23.37), (30.0, 42.43)):
        assert oracle.sat_vp(273.15 + tc) == pytest.approx(want, rel=3e-3)
    # the temperature is clamped to [-80, 80] C (core:856)
    assert oracle.sat_vp(273.15 + 120.0) == oracle.sat_vp(273.15 + 80.0)
    assert oracle.sat_vp(273.15 - 120.0) == oracle.sat_vp(273.15 - 80.0)


def test_saturation_slope_is_the_derivative(oracle):
    # central differences of sat_vp; the reference writes ln 10 as 2.3026,
    # 6e-6 relative from the exact derivative
    for T in (263.15, 283.15, 298.15, 313.15):
        h = 1e-3
        fd = (oracle.sat_vp(T + h) - oracle.sat_vp(T - h)) / (2 * h)
        assert oracle.sat_vp_slope(T) == pytest.approx(fd, rel=2e-5)
    # 20 C: 1.447 hPa / K (FAO-56 table 2.4 gives 0.1447 kPa / C)
    assert oracle.sat_vp_slope(293.15) == pytest.approx(1.447, rel=2e-3)


def test_latent_heat_flux_bulk_formula(oracle):
    # LE = rho_a lambda (q_s(T) - q_a) / (r_a + r_s) with specific humidity
    # q = 0.622 e / P and the latent heat of vaporisation at 20 C,
    # 2.454 MJ / kg (Smithsonian tables, 2.4535e6 J / kg)
    T, rho, P, ea, ra, rs = 293.15, 1.18, 1000.0, 12.0, 45.0, 120.0
    lam20 = 2.4535e6
    es20 = 23.37
    want = rho * lam20 * 0.622 * (es20 - ea) / P / (ra + rs)
    assert oracle.latent(T, rho, P, ea, ra, rs) == pytest.approx(want, rel=2e-3)
    # saturated air transpires nothing (max(0, e_s - e_a), core:877-878); the
    # slope keeps its saturation term (core:899-902), the right-hand
    # derivative where e_s(T) = e_a: a finite difference taken from there
    assert oracle.latent(T, rho, P, 30.0, ra, rs) == 0.0
    Tsat = _bisect(lambda t: oracle.sat_vp(t) - 30.0, 250.0, 330.0)
    h = 1e-4
    right = (oracle.latent(Tsat + h, rho, P, 30.0, ra, rs) - oracle.latent(Tsat, rho, P, 30.0, ra, rs)) / h
    assert oracle.latent_slope(Tsat, rho, P, 30.0, ra, rs) == pytest.approx(right, rel=1e-3)
    # the resistance sum is floored at 1 s/m and the pressure at 100 hPa
    assert oracle.latent(T, rho, P, ea, 0.2, 0.3) == pytest.approx(oracle.latent(T, rho, P, ea, 0.0, 1.0), rel=1e-12)
    assert oracle.latent(T, rho, 50.0, ea, ra, rs) == pytest.approx(oracle.latent(T, rho, 100.0, ea, ra, rs),
                                                                    rel=1e-12)


def test_latent_heat_slope_is_the_derivative(oracle):
    rho, P, ea, ra, rs = 1.18, 1000.0, 12.0, 45.0, 120.0
    for T in (290.15, 300.15, 315.15):
        h = 1e-3
        fd = (oracle.latent(T + h, rho, P, ea, ra, rs) - oracle.latent(T - h, rho, P, ea, ra, rs)) / (2 * h)
        assert oracle.latent_slope(T, rho, P, ea, ra, rs) == pytest.approx(fd, rel=1e-4)


def test_soil_storage_is_the_mean_flux_of_a_semi_infinite_soil(oracle):
    # A semi-infinite conductor whose surface temperature steps by dT at t=0
    # draws q(t) = I dT / sqrt(pi t) (I = sqrt(k rho c), Carslaw & Jaeger
    # 2.4); its mean over a time step dt is 2 I dT / sqrt(pi dt), the storage
    # coefficient per kelvin.  Checked against an implicit finite-difference
    # solution of the heat equation, independent of the closed form.
    c, rho, k, dt = 1180.0, 1800.0, 1.55, 3600.0
    m = _material(oracle, specificHeat=c, density=rho, thermalConductivity=k)
    alpha = k / (rho * c)
    depth = 12.0 * math.sqrt(alpha * dt)
    n, steps = 600, 3600
    dz, tau = depth / n, dt / steps
    r = alpha * tau / dz ** 2
    # implicit Euler on nodes 1..n-1, surface node held at 1 K, far end at 0
    A = np.zeros((3, n - 1))
    A[0, 1:] = -r
    A[1, :] = 1 + 2 * r
    A[2, :-1] = -r
    from scipy.linalg import solve_banded
    u = np.zeros(n - 1)
    flux_sum = 0.0
    for _ in range(steps):
        rhs = u.copy()
        rhs[0] += r * 1.0
        u = solve_banded((1, 1), A, rhs)
        # surface flux from a second-order one-sided difference
        flux_sum += k * (3 * 1.0 - 4 * u[0] + u[1]) / (2 * dz) * tau
    mean_flux = flux_sum / dt
    assert oracle.soil_storage(m, dt) == pytest.approx(mean_flux, rel=1e-2)
    # the time step is floored at 1 s (core:914)
    assert oracle.soil_storage(m, 0.25) == oracle.soil_storage(m, 1.0)


# --- node branches by bisection -----------------------------------------

@pytest.mark.parametrize("veg_method", [0, 1])
def test_eb_isolated_transpiring_leaf_matches_bisection(oracle, veg_method):
    # a leaf (one temperature for both sides): both sides' net radiation,
    # 2 h_a convectiveScale (T - Ta), 2 LE(T) through the surface
    # resistance (vegetationMethod 1: opening from the absorbed shortwave and
    # the temperature response, core:3625-3640), C_A / dt (T - Tprev)
    # (core:3611-3680); Tprev differs from Ta so the storage term matters
    g = _isolated_graph(oracle)
    leaf = _material(oracle, tirReflectance=0.03, vegetation=1, heatCapacityPerArea=9000.0, convectiveScale=1.7,
                     surfaceResistance=150.0)
    p = _params(oracle, Ta=300.0, ea=14.0, veg_method=veg_method)
    absorbed = np.array([420.0, 65.0], np.float32)
    Tprev = np.full(2, 296.5, np.float32)
    T, _, it, conv, _, _ = oracle.eb_node(g, 1, 1, [0], [leaf], absorbed, p, Tprev, Tprev)
    assert conv
    assert T[0] == T[1]
    ha, P, rho, ra = _air(p)
    eps = 1.0 - 0.03
    par = float(absorbed.sum())

    def rs_of(t):
        if veg_method == 0:
            return 150.0
        opening = 0.25 + 0.75 * par / (par + 180.0)
        response = math.exp(-(((t - 273.15) - 25.0) / 18.0) ** 2)
        return min(max(150.0 / max(0.08, opening * response), 40.0), 100000.0)

    def f(t):
        rn = par + 2 * eps * 350.0 - 2 * eps * SIGMA * t ** 4
        h = 2 * ha * 1.7 * (t - 300.0)
        le = 2 * oracle.latent(t, rho, P, 14.0, ra, rs_of(t))
        gs = 9000.0 / 3600.0 * (t - 296.5)
        return rn - h - le - gs

    root = _bisect(f)
    assert T[0] == pytest.approx(root, abs=2e-3)
    # the latent term is a real part of the balance here
    assert 2 * oracle.latent(root, rho, P, 14.0, ra, rs_of(root)) > 20.0


@pytest.mark.parametrize("method", [1, 2])
def test_eb_isolated_wet_soil_matches_bisection(oracle, method):
    # a ground facet of moist soil: per side h_a (T - Ta), LE(T) through the
    # soil's surface resistance, G = k (T - Tprev) with k the storage
    # coefficient (method 1) or 1.5 k (method 2) (core:3681-3718)
    g = _isolated_graph(oracle)
    soil = _material(oracle, soilMoisture=0.25, soilTemperatureMethod=method, surfaceResistance=80.0)
    p = _params(oracle, Ta=295.0, ea=11.0, u=3.5)
    absorbed = np.array([610.0, 12.0], np.float32)
    Tprev = np.array([301.0, 289.0], np.float32)
    T, _, it, conv, _, _ = oracle.eb_node(g, 1, 0, [0], [soil], absorbed, p, Tprev.copy(), Tprev)
    assert conv
    ha, P, rho, ra = _air(p)
    k = oracle.soil_storage(soil, 3600.0) * (1.5 if method == 2 else 1.0)
    eps = 0.95
    for s in range(2):
        def f(t, s=s):
            rn = absorbed[s] + eps * 350.0 - eps * SIGMA * t ** 4
            return rn - ha * (t - 295.0) - oracle.latent(t, rho, P, 11.0, ra, 80.0) - k * (t - Tprev[s])

        assert T[s] == pytest.approx(_bisect(f), abs=2e-3)
    # dry soil (moisture below 0.03) has no latent term: a warmer side 0
    dry = _material(oracle, soilMoisture=0.02, soilTemperatureMethod=method, surfaceResistance=80.0)
    Td, _, _, conv, _, _ = oracle.eb_node(g, 1, 0, [0], [dry], absorbed, p, Tprev.copy(), Tprev)
    assert conv and Td[0] > T[0] + 0.5


@pytest.mark.parametrize("clamped", [False, True])
def test_eb_isolated_photovoltaic_panel_matches_bisection(oracle, clamped):
    # a bifacial panel (an object facet, no transpiration): electric power
    # P = clamp(eta25 (1 + gamma (T - 298.15)) (I_front + b I_back), 0,
    # A_front + A_back) leaves the balance (core:3655-3663, 3772-3778); with
    # little absorbed and a strong incident the clamp holds P at the absorbed
    g = _isolated_graph(oracle)
    panel = _material(oracle, tirReflectance=0.10, vegetation=0, heatCapacityPerArea=12000.0, photovoltaic=1,
                      pvEta25=0.22, pvGamma=-0.0035, pvBifaciality=0.3)
    p = _params(oracle, Ta=302.0, ea=16.0, u=1.5)
    if clamped:
        absorbed = np.array([60.0, 20.0], np.float32)
        incident = np.array([900.0, 300.0], np.float32)
    else:
        absorbed = np.array([700.0, 90.0], np.float32)
        incident = np.array([800.0, 120.0], np.float32)
    Tprev = np.full(2, 300.0, np.float32)
    T, _, it, conv, _, _ = oracle.eb_node(g, 1, 1, [0], [panel], absorbed, p, Tprev, Tprev, incident=incident)
    assert conv
    ha, _, _, _ = _air(p)
    eps = 0.90
    avail = float(absorbed.sum())
    inc = float(incident[0]) + 0.3 * float(incident[1])

    def power(t):
        return min(max(0.22 * (1.0 - 0.0035 * (t - 298.15)) * inc, 0.0), avail)

    def f(t):
        rn = avail + 2 * eps * 350.0 - 2 * eps * SIGMA * t ** 4
        return rn - 2 * ha * (t - 302.0) - 12000.0 / 3600.0 * (t - 300.0) - power(t)

    root = _bisect(f)
    assert T[0] == pytest.approx(root, abs=2e-3)
    raw = 0.22 * (1.0 - 0.0035 * (root - 298.15)) * inc
    assert (raw > avail) == clamped


def test_eb_relaxation_changes_only_the_path(oracle):
    # the relaxed update T + w (candidate - T) has the same fixed point for
    # any w in (0, 1] (core:3758-3827): only the iteration count changes
    g = _isolated_graph(oracle)
    leaf = _material(oracle, tirReflectance=0.03, vegetation=1, heatCapacityPerArea=20000.0)
    absorbed = np.array([350.0, 40.0], np.float32)
    T0 = np.full(2, 298.0, np.float32)
    out = []
    for w in (1.0, 0.5, 0.2):
        T, _, it, conv, _, _ = oracle.eb_node(g, 1, 1, [0], [leaf], absorbed, _params(oracle, relax=w), T0, T0)
        assert conv
        out.append((T[0], it))
    assert abs(out[0][0] - out[1][0]) < 2e-3 and abs(out[0][0] - out[2][0]) < 2e-3
    assert out[2][1] > out[1][1] > out[0][1]

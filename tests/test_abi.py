"""The C-ABI library loads on a host without a GPU and exports every symbol
include/facet_hip.h declares; the ctypes binding covers all of them."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "facet_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(fh_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_reference_operations():
    names = declared_symbols()
    # one entry per facetvk::FacetrtVulkan operation (facetrt_vulkan.h:78-104)
    for op in ("fh_create", "fh_destroy", "fh_set_geometry", "fh_build_form_factors", "fh_solve",
               "fh_solve_accelerated", "fh_incident_irradiance", "fh_compute_sunlit_fraction",
               "fh_facet_count", "fh_surface_count"):
        assert op in names


def test_library_exports_every_declared_symbol():
    from streamsim_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.fail("libfacet_hip.so is not built (python -m streamsim_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    from streamsim_amd import _lib
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_status_strings_without_gpu():
    from streamsim_amd import _lib
    lib = _lib.load()
    assert lib.fh_status_string(0) == b"ok"
    assert lib.fh_status_string(1) == b"invalid argument"
    assert lib.fh_band_stride(1) == 1 and lib.fh_band_stride(3) == 4 and lib.fh_band_stride(48) == 48
    assert lib.fh_mlp_param_count(128) == 128 * 9 + 128 + 128 * 128 + 128 + 128 + 1


def test_cli_usage_and_input_errors():
    # main.cpp:54-79 exit codes: 2 usage; 1 failure (no GPU needed)
    import subprocess
    exe = ROOT / "streamsim_amd" / "facet_engine"
    if not exe.exists():
        pytest.fail("facet_engine is not built (python -m streamsim_amd.build)")
    assert subprocess.run([str(exe)], capture_output=True).returncode == 2
    assert subprocess.run([str(exe), "eVoxelRT", "x.obj"], capture_output=True).returncode == 2
    r = subprocess.run([str(exe), "eFacetRT", "project.json"], capture_output=True, text=True)
    assert r.returncode == 1 and "facetrt:" in r.stderr
    r = subprocess.run([str(exe), "eFacetEB", "/nonexistent/scene.obj"], capture_output=True, text=True)
    assert r.returncode == 1 and "faceteb:" in r.stderr


def test_cli_solar_position_matches_spa_example():
    # the published example of NREL's Solar Position Algorithm (Reda &
    # Andreas, NREL/TP-560-34302): 2003-10-17 12:30:30 local, timezone -7,
    # 105.1786 W, 39.742476 N, 820 hPa, 11 C -> zenith 50.111622 deg,
    # azimuth 194.340241 deg.  The low-precision form is within 0.001 deg.
    import subprocess
    exe = ROOT / "streamsim_amd" / "facet_engine"
    r = subprocess.run([str(exe), "--solar-position", "2003", "10", "17", "12", "30", "30", "-7", "-105.1786",
                        "39.742476", "820", "11"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    zen, az = map(float, r.stdout.split())
    assert abs(zen - 50.111622) < 1e-3 and abs(az - 194.340241) < 1e-3
    # sun below the horizon at local midnight, near the zenith at a tropical noon
    r = subprocess.run([str(exe), "--solar-position", "2019", "1", "1", "0", "0", "0", "8", "116", "40", "839", "25"],
                       capture_output=True, text=True)
    assert float(r.stdout.split()[0]) > 90.0
    r = subprocess.run([str(exe), "--solar-position", "2019", "6", "21", "12", "0", "0", "0", "0", "23.44", "1013",
                        "15"], capture_output=True, text=True)
    assert float(r.stdout.split()[0]) < 1.0


def test_periodic_config_validation():
    # fh_create rejects the periodic fields before it looks for a device
    from streamsim_amd import FacetEngine
    with pytest.raises(ValueError, match="periodicNeighborCount"):
        FacetEngine(periodic_neighbors=21, periodic_size=(1.0, 1.0))
    with pytest.raises(ValueError, match="periodic tile sizes"):
        FacetEngine(periodic_neighbors=4, periodic_size=(0.0, 1.0))



def test_backend_compiles_at_the_reference_call_sites():
    # include/facet_backend.hpp after the type swap: the reference's Config
    # fields, defaults and method calls compile unchanged
    import subprocess
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", f"-I{ROOT / 'include'}",
                        str(ROOT / "tests" / "backend_callsites.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

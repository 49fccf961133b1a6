"""project.json scenes (facetInputView + loadProjectScene,
core:1888-2043, 2553-3025): the assembled scene of the reference's
validation project (tests/golden/urban_tree_project: its project.json,
models and position files) against the reference's own assembled scene
(tests/golden/urban_tree_building.obj, the object facets of its
radiosity_scene.obj), plus the loader's rules on small synthetic projects."""
import json
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HERE = Path(__file__).resolve().parent
GOLDEN = HERE / "golden"
EXE = ROOT / "streamsim_amd" / "facet_engine"


def _need_exe():
    if not EXE.exists():
        pytest.fail("facet_engine is not built (python -m streamsim_amd.build)")


def _load_obj(path):
    v, faces, group = [], [], None
    for line in open(path):
        t = line.split()
        if not t:
            continue
        if t[0] == "v":
            v.append(tuple(map(float, t[1:4])))
        elif t[0] in ("o", "g"):
            group = t[1]
        elif t[0] == "f":
            faces.append(([int(x.split("/")[0]) - 1 for x in t[1:4]], group))
    v = np.array(v)
    return [(v[idx], g) for idx, g in faces]


def _assemble(project, tmp_path):
    out = tmp_path / "scene.obj"
    r = subprocess.run([str(EXE), "--scene-obj", str(project), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return _load_obj(out), r


def test_validation_project_matches_the_reference_scene(tmp_path):
    _need_exe()
    tris, r = _assemble(GOLDEN / "urban_tree_project" / "project.json", tmp_path)
    assert "facets 4138 objects 138 materials 4" in r.stdout
    objects = np.array([t for t, g in tris if g == "object"])
    ground = np.array([t for t, g in tris if g == "STREAMSIM_TERRAIN"])
    # the reference's radiosity_scene.obj (an earlier build without the
    # centring of the scene on the origin): 7 trees and a building, 138 facets
    ref = np.array([t for t, _ in _load_obj(GOLDEN / "urban_tree_building.obj")])
    assert objects.shape == ref.shape == (138, 3, 3)
    shifted = objects + np.array([50.0 / 2, 0.0, 40.0 / 2])
    assert np.max(np.abs(shifted - ref)) < 1e-5
    # flat ground: 50 x 40 one-metre cells, two triangles each, 1 cm below 0
    assert ground.shape == (4000, 3, 3)
    assert np.allclose(ground[:, :, 1], -0.01, atol=1e-8, rtol=0)
    assert ground[:, :, 0].min() == -25.0 and ground[:, :, 0].max() == 25.0
    assert ground[:, :, 2].min() == -20.0 and ground[:, :, 2].max() == 20.0
    # ground triangles face up (cross(b - a, c - a) . y > 0)
    n = np.cross(ground[:, 1] - ground[:, 0], ground[:, 2] - ground[:, 0])
    assert np.all(n[:, 1] > 0)


def _write_project(tmp_path, objects, materials=None, scene=None, spectra=None, extra=None):
    cfg = {
        "scene": {"x": 10, "y": 8, "height": 5, "voxel": 2, "terrain": False,
                  "background": {"spectralName": "soil", "materialName": "soilset"}},
        "sensor": {"bands": "550 850"},
        "objects": {"items": objects},
        "spectra": spectra if spectra is not None else [
            {"name": "soil", "model": "custom", "reflectance": "0.25", "transmittance": "0"},
            {"name": "leafy", "model": "custom", "reflectance": "0.1,0.5", "transmittance": "0.05,0.4",
             "refTir": 0.02}],
        "materials": materials if materials is not None else [
            {"name": "soilset", "type": "Soil", "params": {"SMC": 30}},
            {"name": "leaf_c3", "type": "Vegetation"}],
    }
    if scene:
        cfg["scene"].update(scene)
    if extra:
        cfg.update(extra)
    d = tmp_path / "prj"
    d.mkdir(exist_ok=True)
    p = d / "project.json"
    p.write_text(json.dumps({"schemaVersion": 6, "format": "streamsim-project", "mode": "eFacetEB",
                             "configuration": cfg}))
    return p


def test_box_objects_placements_and_clipping(tmp_path):
    _need_exe()
    d = tmp_path / "prj"
    d.mkdir()
    (d / "pos.txt").write_text("2 2 0\n5 4 0 2 90\nnot a row\n9.8 7.8 0 1 0\n")
    p = _write_project(tmp_path, [
        {"name": "box", "type": "Building", "dimensions": [1, 1, 1], "positionFile": "pos.txt"},
        {"name": "smoke", "type": "Fog", "dimensions": [1, 1, 1]},
    ])
    tris, r = _assemble(p, tmp_path)
    objects = np.array([t for t, g in tris if g == "object"])
    # 12 box triangles per placement; the third box (x 4.3..5.3, z 3.3..4.3
    # after centring) sticks out of the scene box and is clipped to it
    assert len(objects) > 24
    assert objects[:, :, 0].max() <= 5.0 + 1e-5 and objects[:, :, 2].max() <= 4.0 + 1e-5
    assert abs(objects[:, :, 0].max() - 5.0) < 1e-5 and abs(objects[:, :, 2].max() - 4.0) < 1e-5
    # the second box: scale 2, so 2 m tall, centred on (5, 4) -> (0, 0)
    second = objects[12:24]
    assert abs(second[:, :, 1].max() - 2.0) < 1e-6
    assert np.allclose(second[:, :, [0, 2]].mean(axis=(0, 1)), 0.0, atol=1e-5)
    # the Fog object is skipped; 5 x 4 ground cells of 2 m
    assert sum(1 for _, g in tris if g == "STREAMSIM_TERRAIN") == 40


def test_obj_groups_pick_their_mesh_material(tmp_path):
    _need_exe()
    d = tmp_path / "prj"
    d.mkdir()
    (d / "m.obj").write_text("v 0 1 0\nv 1 1 0\nv 0 1 1\nv 0 2 0\ng Roof\nf 1 2 3\ng wall\nf 1 2 4\ng other\nf 1 3 4\n")
    p = _write_project(tmp_path, [
        {"name": "house", "type": "Building", "fileName": "m.obj",
         "meshes": [{"name": "wall", "spectralName": "soil"}, {"name": "roof", "spectralName": "leafy"}]},
    ])
    out = tmp_path / "s.obj"
    r = subprocess.run([str(EXE), "--scene-obj", str(p), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    text = out.read_text().split("o STREAMSIM_TERRAIN")[0]
    # Roof -> the "roof" binding (case-insensitive, material 1), wall -> 0,
    # an unmatched group -> the first binding (0)
    assert [l for l in text.splitlines() if l.startswith("usemtl")] == ["usemtl m1", "usemtl m0"]
    tris = [t for t, g in _load_obj(out) if g == "object"]
    assert len(tris) == 3
    # the OBJ's lowest vertex is moved to y = 0; placed at the scene centre
    assert min(t[:, 1].min() for t in tris) == 0.0


_PANEL = [{"name": "panel", "type": "Building", "dimensions": [1, 1, 0.1], "spectralName": "pv_flat",
           "materialName": "pv_solid"}]
_PV_SPECTRA = [{"name": "soil", "model": "custom", "reflectance": "0.25", "transmittance": "0"},
               {"name": "pv_flat", "model": "custom", "reflectance": "0.06", "transmittance": "0", "refTir": 0.05}]


def test_photovoltaic_material_loads(tmp_path):
    # a photovoltaic surface-energy entry is accepted on a scene object
    # (core:2015-2026, 2953-2973): component 4 for its facets
    _need_exe()
    p = _write_project(tmp_path, _PANEL, spectra=_PV_SPECTRA,
                       materials=[{"name": "soilset", "type": "Soil"},
                                  {"name": "pv_solid", "type": "Building", "energyModel": "photovoltaic",
                                   "params": {"eta25": 0.2, "gamma": -0.004, "bifaciality": 0.5}}])
    r = subprocess.run([str(EXE), "--scene-obj", str(p), str(tmp_path / "o.obj")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("case, message", [
    ("pv_ground", "Photovoltaic panels must be scene objects"),
    ("pv_params", "Invalid photovoltaic material parameters: pv_solid"),
    ("pv_opaque", "Photovoltaic surface must be opaque: pv_solid"),
    ("dem", "Cannot open DEM"),
    ("dem_packed", "compressed images are not read"),
    ("dem_small", "DEM must contain at least 2 x 2 cells"),
    ("dem_nodata", "DEM contains no valid elevations"),
    ("missing", "Cannot open scene OBJ"),
    ("name", "accepts project.json only"),
])
def test_project_errors(tmp_path, case, message):
    # the reference's rules (core:1895-1900, 2963-2969) and this engine's limits
    _need_exe()
    if case == "pv_ground":
        p = _write_project(tmp_path, _PANEL, spectra=_PV_SPECTRA,
                           materials=[{"name": "soilset", "type": "Soil", "energyModel": "photovoltaic"},
                                      {"name": "pv_solid", "type": "Building", "energyModel": "photovoltaic"}])
    elif case == "pv_params":
        p = _write_project(tmp_path, _PANEL, spectra=_PV_SPECTRA,
                           materials=[{"name": "soilset", "type": "Soil"},
                                      {"name": "pv_solid", "type": "Building", "energyModel": "photovoltaic",
                                       "params": {"eta25": 0.9}}])
    elif case == "pv_opaque":
        spectra = [dict(_PV_SPECTRA[0]), dict(_PV_SPECTRA[1], tauTir=0.1)]
        p = _write_project(tmp_path, _PANEL, spectra=spectra,
                           materials=[{"name": "soilset", "type": "Soil"},
                                      {"name": "pv_solid", "type": "Building", "energyModel": "photovoltaic"}])
    elif case == "dem":
        p = _write_project(tmp_path, [], scene={"terrain": True, "demFile": "terrain/dem.tif"})
    elif case == "dem_packed":
        p = _write_project(tmp_path, [], scene={"terrain": True, "demFile": "dem.tif"})
        (p.parent / "dem.tif").write_bytes(_tiff(np.ones((3, 3), np.float32), compression=5))
    elif case == "dem_small":
        p = _write_project(tmp_path, [], scene={"terrain": True, "demFile": "dem.asc"})
        (p.parent / "dem.asc").write_text(_asc(np.ones((1, 4))))
    elif case == "dem_nodata":
        p = _write_project(tmp_path, [], scene={"terrain": True, "demFile": "dem.asc"})
        (p.parent / "dem.asc").write_text(_asc(np.full((2, 2), -9999.0)))
    elif case == "missing":
        p = _write_project(tmp_path, [{"name": "tree", "fileName": "models/none.obj"}])
    else:
        p = _write_project(tmp_path, [])
        p = p.rename(p.with_name("scene.json"))
    r = subprocess.run([str(EXE), "--scene-obj", str(p), str(tmp_path / "o.obj")], capture_output=True, text=True)
    assert r.returncode == 1 and message in r.stderr, r.stderr


def _asc(grid, nodata=-9999.0):
    """An ESRI ASCII grid, rows from north to south."""
    h, w = grid.shape
    rows = "\n".join(" ".join(f"{x:.6g}" for x in row) for row in grid)
    return f"ncols {w}\nNROWS {h}\nxllcorner 0\nyllcorner 0\ncellsize 1\nNODATA_value {nodata:g}\n{rows}\n"


def _tiff(grid, big=False, compression=1, tile=None, nodata=None):
    """A baseline TIFF of one band (grid's dtype: float32 or int16), in
    strips of 2 rows or in tiles of tile x tile pixels."""
    e = ">" if big else "<"
    h, w = grid.shape
    fmt = 3 if grid.dtype == np.float32 else 2
    bits = grid.dtype.itemsize * 8
    data, offsets, counts = b"", [], []
    head = 8
    if tile:
        for ty in range(0, h, tile):
            for tx in range(0, w, tile):
                block = np.zeros((tile, tile), grid.dtype)
                part = grid[ty:ty + tile, tx:tx + tile]
                block[:part.shape[0], :part.shape[1]] = part
                offsets.append(head + len(data))
                raw = block.astype(grid.dtype.newbyteorder(e)).tobytes()
                counts.append(len(raw))
                data += raw
    else:
        for y in range(0, h, 2):
            offsets.append(head + len(data))
            raw = grid[y:y + 2].astype(grid.dtype.newbyteorder(e)).tobytes()
            counts.append(len(raw))
            data += raw
    extra = b""
    ifd_at = head + len(data)
    n_entries = 8 + (4 if tile else 3) + (1 if nodata is not None else 0)
    after_ifd = ifd_at + 2 + 12 * n_entries + 4

    def arr(values):
        nonlocal extra
        if len(values) == 1:
            return 4, 1, struct.pack(e + "I", values[0])
        off = after_ifd + len(extra)
        extra += struct.pack(e + f"{len(values)}I", *values)
        return 4, len(values), struct.pack(e + "I", off)

    def short(v):
        return 3, 1, struct.pack(e + "HH", v, 0)

    entries = [(256, 4, 1, struct.pack(e + "I", w)), (257, 4, 1, struct.pack(e + "I", h)),
               (258,) + short(bits), (259,) + short(compression), (262,) + short(1), (277,) + short(1),
               (284,) + short(1), (339,) + short(fmt)]
    if tile:
        entries += [(322, 4, 1, struct.pack(e + "I", tile)), (323, 4, 1, struct.pack(e + "I", tile)),
                    (324,) + arr(offsets), (325,) + arr(counts)]
    else:
        entries += [(273,) + arr(offsets), (278, 4, 1, struct.pack(e + "I", 2)), (279,) + arr(counts)]
    if nodata is not None:
        text = f"{nodata:g}".encode() + b"\0"
        off = after_ifd + len(extra)
        extra += text
        entries.append((42113, 2, len(text), struct.pack(e + "I", off) if len(text) > 4 else text.ljust(4, b"\0")))
    entries.sort()
    assert len(entries) == n_entries
    ifd = struct.pack(e + "H", len(entries))
    for tag, typ, cnt, val in entries:
        ifd += struct.pack(e + "HHI", tag, typ, cnt) + val
    ifd += struct.pack(e + "I", 0)
    return (b"II" if not big else b"MM") + struct.pack(e + "HI", 42, ifd_at) + data + ifd + extra


def _ground_heights(grid, nodata, W, D, x, z):
    """Bilinear DEM heights above the valid minimum at scene-centred x
    (north), z (east): rows run from north (x = W / 2) to south, columns
    from west (z = -D / 2) to east."""
    from scipy.interpolate import RegularGridInterpolator
    g = np.array(grid, np.float64)
    valid = np.isfinite(g) & (g != nodata)
    lo = g[valid].min()
    g[~valid] = lo
    h, w = g.shape
    f = RegularGridInterpolator((np.arange(h), np.arange(w)), g)
    v = (1 - np.clip((x + W / 2) / W, 0, 1)) * (h - 1)
    u = np.clip((z + D / 2) / D, 0, 1) * (w - 1)
    return np.maximum(0.0, f(np.stack([v, u], -1)) - lo)


_GRID = (np.arange(20, dtype=np.float64).reshape(5, 4) * 0.7 + np.array([0.0, 1.5, -0.5, 3.0])
         + np.array([[0], [2], [5], [9], [14]]))  # rows 5 (north to south), columns 4 (west to east)


@pytest.mark.parametrize("fmt", ["asc", "tif_le", "tif_be", "tif_tiles", "tif_int16"])
def test_dem_ground_follows_the_grid(tmp_path, fmt):
    # a DEM stretched over the scene (readDemHeightField / heightAboveMinimum
    # / appendProjectGround, core:2617-2697, 2830-2878): every ground vertex
    # 1 cm below the bilinear height above the valid minimum (a no-data cell
    # takes the minimum), the north-west corner on the grid's first cell and
    # the south-east one on its last; a box stands on the terrain at its
    # placement (core:2991-2992); the same scene from every encoding
    _need_exe()
    grid = _GRID.copy()
    grid[2, 1] = -9999.0
    name = "dem.asc" if fmt == "asc" else "dem.tif"
    d = tmp_path / "prj"
    d.mkdir()
    if fmt == "asc":
        (d / name).write_text(_asc(grid))
    elif fmt == "tif_int16":
        grid = np.round(grid).astype(np.int16)
        (d / name).write_bytes(_tiff(grid, nodata=-9999))
    else:
        (d / name).write_bytes(_tiff(grid.astype(np.float32), big=fmt == "tif_be",
                                     tile=16 if fmt == "tif_tiles" else None, nodata=-9999))
    (d / "pos.txt").write_text("3 6 0\n")
    p = _write_project(tmp_path, [{"name": "box", "type": "Building", "dimensions": [1, 1, 1],
                                   "positionFile": "pos.txt"}],
                       scene={"terrain": True, "demFile": name, "height": 50})
    tris, r = _assemble(p, tmp_path)
    ground = np.array([t for t, g in tris if g == "STREAMSIM_TERRAIN"])
    box = np.array([t for t, g in tris if g == "object"])
    assert ground.shape == (40, 3, 3)  # 5 x 4 cells of 2 m
    pts = ground.reshape(-1, 3)
    want = _ground_heights(grid, -9999.0, 10.0, 8.0, pts[:, 0], pts[:, 2]) - 0.01
    np.testing.assert_allclose(pts[:, 1], want, rtol=0, atol=2e-5)
    g = np.where(grid == -9999, np.inf, grid).astype(np.float64)
    lo = g.min()
    nw = pts[(pts[:, 0] == 5.0) & (pts[:, 2] == -4.0)]
    se = pts[(pts[:, 0] == -5.0) & (pts[:, 2] == 4.0)]
    assert np.allclose(nw[:, 1], grid[0, 0] - lo - 0.01, atol=1e-5)
    assert np.allclose(se[:, 1], grid[-1, -1] - lo - 0.01, atol=1e-5)
    # the box's base at its placement (x 3, y 6 -> centred -2, 2)
    base = _ground_heights(grid, -9999.0, 10.0, 8.0, np.array([3.0 - 5.0]), np.array([6.0 - 4.0]))[0]
    assert abs(box[:, :, 1].min() - base) < 1e-5 and abs(box[:, :, 1].max() - base - 1.0) < 1e-5


def test_dem_library_grids(tmp_path):
    # the reference's DEM library: central_hill_65x65.asc (65 x 65 cells,
    # 0..20 m) and single_slope_60x60m.tif (60 x 60 float32 strips, GDAL_NODATA
    # -9999), their elevations read here with numpy
    _need_exe()
    hill = np.loadtxt(GOLDEN / "dem" / "central_hill_65x65.asc", skiprows=6)
    raw = (GOLDEN / "dem" / "single_slope_60x60m.tif").read_bytes()
    slope = np.frombuffer(raw, "<f4", count=3600, offset=8).reshape(60, 60).astype(np.float64)
    for name, grid in (("central_hill_65x65.asc", hill), ("single_slope_60x60m.tif", slope)):
        d = tmp_path / name
        d.mkdir()
        p = _write_project(d, [], scene={"terrain": True, "demFile": name, "x": 30, "y": 20, "voxel": 1})
        shutil.copy(GOLDEN / "dem" / name, p.parent / name)
        tris, r = _assemble(p, d)
        pts = np.array([t for t, g in tris if g == "STREAMSIM_TERRAIN"]).reshape(-1, 3)
        assert len(pts) == 30 * 20 * 6
        want = _ground_heights(grid, -9999.0, 30.0, 20.0, pts[:, 0], pts[:, 2]) - 0.01
        np.testing.assert_allclose(pts[:, 1], want, rtol=0, atol=1e-4)
        assert np.ptp(pts[:, 1]) > 1.0


def test_dem_of_the_validation_project(tmp_path):
    # the reference's urban_dem_building terrain (dem_test_input.tif: 32 x 24
    # big-endian float32 strips) against its stored FacetRT run.  That run
    # predates the current ground rule: its ground is the DEM's own pixel grid
    # (31 x 23 cells over the 50 x 40 m scene, columns along x) at heights
    # above the DEM mean, where the current source resamples the DEM to the
    # scene's voxel grid at heights above the minimum, rows along x from the
    # north (core:2630-2650, 2844-2850).  On a 23 x 31 m scene at 1 m voxels
    # the scene grid's vertices fall on the DEM's pixels, so the ground read
    # here must carry the stored run's pixel values: ours + 1 cm equals the
    # stored height minus the stored minimum, at the pixel the two
    # orientations share
    _need_exe()
    stored = np.load(GOLDEN / "dem" / "urban_dem_building_ground.npy")  # [z 24][x 32]
    d = tmp_path / "pixels"
    d.mkdir()
    p = _write_project(d, [], scene={"terrain": True, "demFile": "dem.tif", "x": 23, "y": 31, "voxel": 1})
    shutil.copy(GOLDEN / "dem" / "dem_test_input.tif", p.parent / "dem.tif")
    tris, _ = _assemble(p, d)
    pts = np.array([t for t, g in tris if g == "STREAMSIM_TERRAIN"]).reshape(-1, 3)
    assert len(pts) == 23 * 31 * 6
    row = np.rint(11.5 - pts[:, 0]).astype(int)  # DEM row from the north edge (x = 11.5)
    col = np.rint(pts[:, 2] + 15.5).astype(int)  # DEM column from the west edge (z = -15.5)
    assert np.allclose(pts[:, 0], 11.5 - row, atol=1e-5) and np.allclose(pts[:, 2], col - 15.5, atol=1e-5)
    np.testing.assert_allclose(pts[:, 1] + 0.01, stored[row, col] - stored.min(), rtol=0, atol=2e-5)
    # the validation project's own 50 x 40 m scene: 50 x 40 ground cells by
    # the current rule (the stored run has 31 x 23)
    d = tmp_path / "scene"
    d.mkdir()
    p = _write_project(d, [], scene={"terrain": True, "demFile": "dem.tif", "x": 50, "y": 40, "voxel": 1})
    shutil.copy(GOLDEN / "dem" / "dem_test_input.tif", p.parent / "dem.tif")
    tris, _ = _assemble(p, d)
    assert sum(g == "STREAMSIM_TERRAIN" for _, g in tris) == 50 * 40 * 2


def test_large_scene_raster_and_adaptive_layers(tmp_path):
    # scenes of 1000 m or more rasterise at 2048 pixels and 16 layers, and
    # their ground grid at up to 512 cells a side (core:2895-2901,
    # 2843-2849); the layers then double while 16 times the mean fragment
    # depth over the raster, periodic copies included, exceeds them
    # (adaptiveVisibilityRasterLayers, core:168-188)
    _need_exe()
    p = _write_project(tmp_path, [], scene={"x": 1200, "y": 1000, "voxel": 4})
    tris, r = _assemble(p, tmp_path)
    assert "raster 2048 layers 16" in r.stdout
    assert len(tris) == 300 * 250 * 2  # 300 x 250 cells: above the small raster's 256 cap
    p = _write_project(tmp_path, [], scene={"x": 100, "y": 100, "voxel": 0.25})
    tris, r = _assemble(p, tmp_path)
    assert "raster 1024 layers 64" in r.stdout and len(tris) == 256 * 256 * 2
    # 104 135 cells -> 208 270 facets, 21 copies: 16 x 1.04 mean depth -> 17 > 16 layers -> 32
    p = _write_project(tmp_path, [], scene={"x": 1200, "y": 1000, "voxel": 3.4},
                       extra={"control": {"periodicBoundary": True, "periodicTraversalCount": 20}})
    tris, r = _assemble(p, tmp_path)
    assert len(tris) == 353 * 295 * 2
    assert "automatically increased A-buffer layers from 16 to 32 for 208270 facets" in r.stdout
    assert "raster 2048 layers 32" in r.stdout


def _pillow():
    return pytest.importorskip("PIL.Image")


def _decode(path, tmp_path):
    out = tmp_path / "img.raw"
    r = subprocess.run([str(EXE), "--decode-image", str(path), str(out)], capture_output=True, text=True)
    if r.returncode:
        return None, r.stderr
    w, h, c = map(int, r.stdout.split())
    return np.fromfile(out, np.uint8).reshape(h, w, c), ""


def test_texture_images_decode_like_libjpeg(tmp_path):
    # the texture reader's samples against Pillow's (libjpeg-turbo, libpng)
    # decode of the same files: the library's concrete texture (baseline,
    # 4:4:4), synthetic JPEGs at 4:4:4, 4:2:2 and 4:2:0 (fancy upsampling),
    # greyscale and with restart markers, and PNGs of every colour type
    # (palette: indices; 16-bit: clamped to 255, as a byte read gives them);
    # progressive JPEGs are refused
    _need_exe()
    Image = _pillow()
    a, _ = _decode(GOLDEN / "textures" / "concrete_variance.jpg", tmp_path)
    np.testing.assert_array_equal(a, np.asarray(Image.open(GOLDEN / "textures" / "concrete_variance.jpg")))
    rng = np.random.default_rng(3)
    img = np.clip(np.cumsum(rng.integers(0, 40, (37, 53, 3)), axis=1) // 3, 0, 255).astype(np.uint8)
    cases = [("s0.jpg", dict(quality=90, subsampling=0)), ("s1.jpg", dict(quality=60, subsampling=1)),
             ("s2.jpg", dict(quality=75, subsampling=2)), ("rst.jpg", dict(quality=80, restart_marker_blocks=3))]
    for name, kw in cases:
        Image.fromarray(img).save(tmp_path / name, **kw)
        a, err = _decode(tmp_path / name, tmp_path)
        np.testing.assert_array_equal(a, np.asarray(Image.open(tmp_path / name).convert("RGB")), err_msg=name)
    Image.fromarray(img[:, :, 1]).save(tmp_path / "g.jpg", quality=85)
    a, _ = _decode(tmp_path / "g.jpg", tmp_path)
    np.testing.assert_array_equal(a[:, :, 0], np.asarray(Image.open(tmp_path / "g.jpg")))
    Image.fromarray(img).save(tmp_path / "p.jpg", progressive=True)
    a, err = _decode(tmp_path / "p.jpg", tmp_path)
    assert a is None and "progressive" in err
    for mode in ("RGB", "RGBA", "L", "LA", "P"):
        Image.fromarray(img).convert(mode).save(tmp_path / "x.png")
        a, _ = _decode(tmp_path / "x.png", tmp_path)
        want = np.asarray(Image.open(tmp_path / "x.png"))
        want = want[:, :, None] if want.ndim == 2 else want[:, :, :3]
        np.testing.assert_array_equal(a, want, err_msg=mode)
    Image.fromarray(img[:, :, 0].astype(np.uint16) * 3).save(tmp_path / "w.png")
    a, _ = _decode(tmp_path / "w.png", tmp_path)
    np.testing.assert_array_equal(a[:, :, 0], np.minimum(img[:, :, 0].astype(np.uint16) * 3, 255))


def _textured(tmp_path, image, strength=0.2, repeat=2.0, reflectance="0.25", objects=()):
    spectra = [{"name": "soil", "model": "custom", "reflectance": reflectance, "transmittance": "0",
                "physicalTexture": {"enabled": True, "fileName": image, "strength": strength, "repeatSize": repeat}},
               {"name": "plain", "model": "custom", "reflectance": "0.3", "transmittance": "0"}]
    p = _write_project(tmp_path, list(objects), spectra=spectra, scene={"voxel": 1, "height": 10})
    return p


def _factors(p, tmp_path):
    k = tmp_path / "k.bin"
    r = subprocess.run([str(EXE), "--scene-obj", str(p), str(tmp_path / "s.obj"), str(k)], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    tris = _load_obj(tmp_path / "s.obj")
    return np.fromfile(k, np.float32), tris, r


def test_texture_factors(tmp_path):
    # finalizeReflectanceTextureFactors (core:718-846) on the 10 x 8 ground
    # (160 one-metre triangles of equal area): the area-weighted mean factor
    # is one; the deviations scale with the strength; a ramp across the
    # texture's columns varies the ground's factors along x only, and one
    # down its rows along z only (the ground samples the (x, z) plane); at
    # strength 1 and r = 0.9 the largest factor is capped at (1 - t) / r
    _need_exe()
    Image = _pillow()
    d = tmp_path / "prj"
    d.mkdir()
    ramp = np.tile(np.arange(64, dtype=np.uint8)[None, :] * 2, (8, 1))
    Image.fromarray(ramp).save(d / "cols.png")
    Image.fromarray(np.ascontiguousarray(ramp.T)).save(d / "rows.png")
    p = _textured(tmp_path, "cols.png", strength=0.1, repeat=40.0)
    k1, tris, r = _factors(p, tmp_path)
    assert "physical texture material 0, area-weighted factor mean=1" in r.stdout
    ground = np.array([t for t, g in tris if g == "STREAMSIM_TERRAIN"])
    assert k1.size == len(ground) == 160
    assert abs(k1.astype(np.float64).mean() - 1.0) < 1e-6
    cx, cz = ground[:, :, 0].mean(1), ground[:, :, 2].mean(1)
    groups = {}
    for f, x in zip(k1, np.round(cx, 4)):
        groups.setdefault(x, []).append(f)
    assert len(groups) > 10 and max(np.ptp(v) for v in groups.values()) < 1e-6  # constant along z
    assert np.ptp(k1) > 0.01
    p = _textured(tmp_path, "cols.png", strength=0.05, repeat=40.0)
    k2, _, _ = _factors(p, tmp_path)
    np.testing.assert_allclose(k1 - 1.0, 2.0 * (k2 - 1.0), rtol=1e-4, atol=1e-7)
    p = _textured(tmp_path, "rows.png", strength=0.1, repeat=40.0)
    k3, _, _ = _factors(p, tmp_path)
    groups = {}
    for f, z in zip(k3, np.round(cz, 4)):
        groups.setdefault(z, []).append(f)
    assert len(groups) > 8 and max(np.ptp(v) for v in groups.values()) < 1e-6  # constant along x
    p = _textured(tmp_path, "cols.png", strength=1.0, repeat=40.0, reflectance="0.9")
    k4, _, _ = _factors(p, tmp_path)
    assert abs(float(k4.max()) - 1.0 / 0.9) < 1e-5 and k4.min() > 0.0
    assert abs(k4.astype(np.float64).mean() - 1.0) < 1e-6
    # no texture: every factor is one
    p = _write_project(tmp_path, [])
    k0, _, _ = _factors(p, tmp_path)
    assert np.all(k0 == 1.0)


def test_texture_errors(tmp_path):
    _need_exe()
    Image = _pillow()
    p = _textured(tmp_path, "")
    r = subprocess.run([str(EXE), "--scene-obj", str(p), str(tmp_path / "o.obj")], capture_output=True, text=True)
    assert r.returncode == 1 and "Physical texture is enabled but has no file: soil" in r.stderr
    Image.fromarray(np.full((4, 4), 90, np.uint8)).save(p.parent / "flat.png")
    p = _textured(tmp_path, "flat.png")
    r = subprocess.run([str(EXE), "--scene-obj", str(p), str(tmp_path / "o.obj")], capture_output=True, text=True)
    assert r.returncode == 1 and "effectively constant" in r.stderr

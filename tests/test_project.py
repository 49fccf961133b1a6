"""project.json scenes (facetInputView + loadProjectScene,
core:1888-2043, 2553-3025): the assembled scene of the reference's
validation project (tests/golden/urban_tree_project: its project.json,
models and position files) against the reference's own assembled scene
(tests/golden/urban_tree_building.obj, the object facets of its
radiosity_scene.obj), plus the loader's rules on small synthetic projects."""
import json
import shutil
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
    ("dem", "DEM terrain is not read"),
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
    elif case == "missing":
        p = _write_project(tmp_path, [{"name": "tree", "fileName": "models/none.obj"}])
    else:
        p = _write_project(tmp_path, [])
        p = p.rename(p.with_name("scene.json"))
    r = subprocess.run([str(EXE), "--scene-obj", str(p), str(tmp_path / "o.obj")], capture_output=True, text=True)
    assert r.returncode == 1 and message in r.stderr, r.stderr

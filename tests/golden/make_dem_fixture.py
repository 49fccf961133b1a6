"""Writes the DEM validation fixture from the reference's urban_dem_building
case: its terrain file (terrain/dem_test_input.tif, 32 x 24 big-endian
float32 strips) as given, and the ground of its stored FacetRT run
(output/radiosity_scene.obj: the STREAMSIM_TERRAIN facets after the 138
object facets) as a grid of vertex heights, rows along z (-20 .. 20),
columns along x (-25 .. 25).  Data only; run in a
container where the reference is mounted:

    python tests/golden/make_dem_fixture.py /root/reference
"""
import shutil
import sys
from pathlib import Path

import numpy as np

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
case = ref / "validation" / "urban_dem_building"
here = Path(__file__).resolve().parent / "dem"
shutil.copyfile(case / "terrain" / "dem_test_input.tif", here / "dem_test_input.tif")
verts, faces = [], []
for line in open(case / "output" / "radiosity_scene.obj"):
    t = line.split()
    if t and t[0] == "v":
        verts.append([float(x) for x in t[1:4]])
    elif t and t[0] == "f":
        faces.append([int(x.split("/")[0]) - 1 for x in t[1:4]])
v = np.array(verts)
ground = v[np.unique(np.array(faces[138:]).ravel())]
xs = np.unique(np.round(ground[:, 0], 4))
zs = np.unique(np.round(ground[:, 2], 4))
grid = np.full((len(zs), len(xs)), np.nan)
for x, y, z in ground:
    grid[np.searchsorted(zs, round(z, 4)), np.searchsorted(xs, round(x, 4))] = y
assert grid.shape == (24, 32) and np.isfinite(grid).all()
np.save(here / "urban_dem_building_ground.npy", grid)
print("wrote", here / "dem_test_input.tif", here / "urban_dem_building_ground.npy")

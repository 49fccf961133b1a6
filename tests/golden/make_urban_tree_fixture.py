"""Writes the urban-tree validation fixture from the reference's stored
FacetRT GPU output (validation/urban_tree_building/output): the scene OBJ as
given, and the per-side sunlit / radiosity / light-enhancement arrays plus
the graph summary of radiosity_gpu.json.  Data only; run in a container
where the reference is mounted:

    python tests/golden/make_urban_tree_fixture.py /root/reference
"""
import json
import shutil
import sys
from pathlib import Path

import numpy as np

ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
src = ref / "validation" / "urban_tree_building" / "output"
here = Path(__file__).resolve().parent
shutil.copyfile(src / "radiosity_scene.obj", here / "urban_tree_building.obj")
d = json.loads((src / "radiosity_gpu.json").read_text())
np.savez_compressed(here / "urban_tree_building_gpu.npz",
                    facetCount=np.int64(d["facetCount"]), leafFacetCount=np.int64(d["leafFacetCount"]),
                    directedEdges=np.int64(d["graph"]["directedEdges"]), iterations=np.int64(d["iterations"]),
                    sunlit=np.asarray(d["sunlit"], np.float32), radiosity=np.asarray(d["radiosity"], np.float32),
                    lightEnhancement=np.asarray(d["lightEnhancement"], np.float32))
print("wrote", here / "urban_tree_building.obj", here / "urban_tree_building_gpu.npz")

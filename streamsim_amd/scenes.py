"""Seeded synthetic scenes for the benchmark configurations.

Geometry follows the reference's conventions (models/histream/src/facetrt/
facetrtcore.cpp): +Y is up, a facet is three consecutive vertices sharing the
unit normal of cross(b - a, c - a) (appendTriangle, core:288-306), leaf
(object) facets come first and ground facets last (leafFacetCount).

  row_crop     4800 facets: 10 x 10 m ground of 40 x 40 cells x 2 triangles,
               4 rows (2.5 m apart) of 400 equilateral leaves (0.08 m side),
               centres uniform in a 0.4 m wide x 1.0 m high box along the row,
               normals uniform on the upper hemisphere.
  forest       60 000 facets: 50 x 50 m ground of 100 x 100 cells x 2,
               25 ellipsoid crowns (semi-axes 2, 3, 2 m; x, y, z) on a 5 x 5
               grid, 1600 leaves (0.1 m side) each, normals uniform on the sphere.
  urban        ~262 000 facets: 40 box buildings (footprint U(3, 5) m, height
               U(5, 30) m) meshed at 0.5 m, 100 forest crowns, 140 x 100 m
               ground meshed at 2 m.
  terrain      ~1.0 M facets: 512 x 512 DEM at 1 m (sum of 4 random sines with
               a 20 m total amplitude plus N(0, 0.5 m) noise) + 300 crowns.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Scene:
    name: str
    vertices: np.ndarray          # (3F, 6) float32: position xyz, normal xyz
    leaf_count: int               # facets [0, leaf_count) are objects
    material: np.ndarray          # (F,) uint32 material index
    bmin: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))
    bmax: np.ndarray = field(default_factory=lambda: np.zeros(3, np.float32))

    @property
    def facet_count(self) -> int:
        return self.vertices.shape[0] // 3

    @property
    def surface_count(self) -> int:
        return 2 * self.facet_count

    def facet_normals(self) -> np.ndarray:
        return self.vertices[0::3, 3:6]

    def facet_areas(self) -> np.ndarray:
        p = self.vertices[:, 0:3].reshape(-1, 3, 3).astype(np.float64)
        return 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)


def triangles_to_vertices(tri: np.ndarray) -> np.ndarray:
    """(F, 3, 3) float32 corners -> (3F, 6) vertices with float32 normals.

    The normal is cross(b - a, c - a) scaled by 1/length in float32, the
    arithmetic of appendTriangle (core:288-306)."""
    tri = np.ascontiguousarray(tri, dtype=np.float32)
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    u = b - a
    v = c - a
    n = np.empty_like(u)
    n[:, 0] = u[:, 1] * v[:, 2] - u[:, 2] * v[:, 1]
    n[:, 1] = u[:, 2] * v[:, 0] - u[:, 0] * v[:, 2]
    n[:, 2] = u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0]
    dd = n[:, 0] * n[:, 0] + n[:, 1] * n[:, 1] + n[:, 2] * n[:, 2]
    keep = dd > np.float32(1.0e-16)
    length = np.sqrt(dd[keep])
    n = n[keep] * (np.float32(1.0) / length)[:, None]
    tri = tri[keep]
    out = np.empty((tri.shape[0], 3, 6), np.float32)
    out[:, :, 0:3] = tri
    out[:, :, 3:6] = n[:, None, :]
    return out.reshape(-1, 6)


def _basis(n: np.ndarray):
    """Two unit vectors t1, t2 with t1 x t2 = n (rows of n unit)."""
    ref = np.where(np.abs(n[:, 1:2]) < 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t1 = np.cross(ref, n)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(n, t1)
    return t1, t2


def equilateral_leaves(centres: np.ndarray, normals: np.ndarray, side: float, rng) -> np.ndarray:
    """(F, 3, 3) equilateral triangles, counter-clockwise about their normal."""
    t1, t2 = _basis(normals)
    psi = rng.uniform(0.0, 2.0 * np.pi, size=centres.shape[0])
    rc = side / np.sqrt(3.0)
    tri = np.empty((centres.shape[0], 3, 3))
    for k in range(3):
        ang = psi + 2.0 * np.pi * k / 3.0
        tri[:, k] = centres + rc * (np.cos(ang)[:, None] * t1 + np.sin(ang)[:, None] * t2)
    return tri


def hemisphere_normals(n: int, rng) -> np.ndarray:
    y = rng.uniform(0.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    return np.stack([r * np.cos(phi), y, r * np.sin(phi)], axis=1)


def sphere_normals(n: int, rng) -> np.ndarray:
    y = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    r = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    return np.stack([r * np.cos(phi), y, r * np.sin(phi)], axis=1)


def grid_ground(x0, x1, z0, z1, nx, nz, height=None) -> np.ndarray:
    """(2 nx nz, 3, 3) upward-facing triangles of a regular grid.  height(x, z)
    gives the elevation (default 0)."""
    xs = np.linspace(x0, x1, nx + 1)
    zs = np.linspace(z0, z1, nz + 1)
    X, Z = np.meshgrid(xs, zs, indexing="ij")
    Y = np.zeros_like(X) if height is None else height(X, Z)
    P = np.stack([X, Y, Z], axis=-1)
    a = P[:-1, :-1].reshape(-1, 3)  # (x0, z0)
    b = P[:-1, 1:].reshape(-1, 3)   # (x0, z1)
    c = P[1:, :-1].reshape(-1, 3)   # (x1, z0)
    d = P[1:, 1:].reshape(-1, 3)    # (x1, z1)
    t1 = np.stack([a, b, c], axis=1)
    t2 = np.stack([c, b, d], axis=1)
    return np.stack([t1, t2], axis=1).reshape(-1, 3, 3)


def crown_leaves(centre, semi, count: int, side: float, rng) -> np.ndarray:
    pts = []
    need = count
    while need > 0:
        cand = rng.uniform(-1.0, 1.0, size=(2 * need + 16, 3))
        cand = cand[np.sum(cand * cand, axis=1) <= 1.0][:need]
        pts.append(cand)
        need -= cand.shape[0]
    p = np.concatenate(pts)[:count] * np.asarray(semi)[None, :] + np.asarray(centre)[None, :]
    return equilateral_leaves(p, sphere_normals(count, rng), side, rng)


def _finish(name, leaf_tri, ground_tri, leaf_mat=None, ground_mat=None) -> Scene:
    lv = triangles_to_vertices(leaf_tri) if leaf_tri is not None and len(leaf_tri) else np.zeros((0, 6), np.float32)
    gv = triangles_to_vertices(ground_tri)
    nl = lv.shape[0] // 3
    ng = gv.shape[0] // 3
    v = np.concatenate([lv, gv])
    mat = np.concatenate([
        np.zeros(nl, np.uint32) if leaf_mat is None else leaf_mat[:nl].astype(np.uint32),
        np.ones(ng, np.uint32) if ground_mat is None else ground_mat[:ng].astype(np.uint32),
    ])
    p = v[:, 0:3]
    return Scene(name, np.ascontiguousarray(v), nl, mat, p.min(axis=0).astype(np.float32),
                 p.max(axis=0).astype(np.float32))


def row_crop(seed: int = 1, rows: int = 4, leaves_per_row: int = 400, cells: int = 40) -> Scene:
    """Plain radiosity row crop: 1600 leaves + 3200 ground facets at defaults."""
    rng = np.random.default_rng(seed)
    tris = []
    for k in range(rows):
        xr = 1.25 + 2.5 * k
        c = np.stack([rng.uniform(xr - 0.2, xr + 0.2, leaves_per_row),
                      rng.uniform(0.05, 1.05, leaves_per_row),
                      rng.uniform(0.05, 9.95, leaves_per_row)], axis=1)
        tris.append(equilateral_leaves(c, hemisphere_normals(leaves_per_row, rng), 0.08, rng))
    ground = grid_ground(0.0, 10.0, 0.0, 10.0, cells, cells)
    return _finish("row_crop", np.concatenate(tris), ground)


def forest(seed: int = 2, trees_x: int = 5, leaves: int = 1600, cells: int = 100, size: float = 50.0) -> Scene:
    """Forest: 25 ellipsoid crowns of 1600 leaves on 20 000 ground triangles."""
    rng = np.random.default_rng(seed)
    sp = size / trees_x
    tris = []
    for i in range(trees_x):
        for j in range(trees_x):
            centre = (sp * (i + 0.5), 4.0, sp * (j + 0.5))
            tris.append(crown_leaves(centre, (2.0, 3.0, 2.0), leaves, 0.1, rng))
    ground = grid_ground(0.0, size, 0.0, size, cells, cells)
    return _finish("forest", np.concatenate(tris), ground)


def _box_building(x0, z0, w, d, h, step=0.5):
    """Outward walls and roof of a box meshed at `step` (no floor)."""
    out = []

    def wall(p0, ea, eb, na, nb_):
        # quad p0 + i/na*ea + j/nb*eb, split in two triangles with normal ea x eb
        i = np.arange(na)
        j = np.arange(nb_)
        I, J = np.meshgrid(i, j, indexing="ij")
        I = I.reshape(-1, 1)
        J = J.reshape(-1, 1)
        ua, ub = ea / na, eb / nb_
        a = p0 + I * ua + J * ub
        b = a + ua
        c = a + ub
        dd = a + ua + ub
        out.append(np.stack([a, b, c], axis=1))
        out.append(np.stack([b, dd, c], axis=1))

    nh = max(1, int(np.ceil(h / step)))
    nw = max(1, int(np.ceil(w / step)))
    nd = max(1, int(np.ceil(d / step)))
    X = np.array([1.0, 0.0, 0.0])
    Y = np.array([0.0, 1.0, 0.0])
    Zv = np.array([0.0, 0.0, 1.0])
    p = np.array([x0, 0.0, z0])
    # normal = ea x eb must point outward
    wall(p, Y * h, X * w, nh, nw)                       # -z face: Y x X = -Z
    wall(p + Zv * d, X * w, Y * h, nw, nh)              # +z face: X x Y = +Z
    wall(p, Zv * d, Y * h, nd, nh)                      # -x face: Z x Y = -X
    wall(p + X * w, Y * h, Zv * d, nh, nd)              # +x face: Y x Z = +X
    wall(p + Y * h, Zv * d, X * w, nd, nw)              # roof: Z x X = +Y
    return np.concatenate(out)


def urban(seed: int = 3, buildings: int = 40, trees: int = 100, leaves: int = 1600) -> Scene:
    """Urban block + trees (~262k facets)."""
    rng = np.random.default_rng(seed)
    W, Dz = 140.0, 100.0
    # buildings on an 8 x 5 grid of 17.5 x 20 m lots
    bl = []
    lots = [(i, j) for i in range(8) for j in range(5)][:buildings]
    occupied = []
    for (i, j) in lots:
        w, d = rng.uniform(3.0, 5.0, 2)
        h = rng.uniform(5.0, 30.0)
        x0 = 17.5 * i + 2.0
        z0 = 20.0 * j + 2.0
        bl.append(_box_building(x0, z0, w, d, h))
        occupied.append((x0 - 2.5, x0 + w + 2.5, z0 - 2.5, z0 + d + 2.5))
    tris = []
    placed = 0
    while placed < trees:
        x, z = rng.uniform(3.0, W - 3.0), rng.uniform(3.0, Dz - 3.0)
        if any(a <= x <= b and c <= z <= d for (a, b, c, d) in occupied):
            continue
        tris.append(crown_leaves((x, 4.0, z), (2.0, 3.0, 2.0), leaves, 0.1, rng))
        placed += 1
    objects = np.concatenate(tris + bl)
    nleaf = trees * leaves
    ground = grid_ground(0.0, W, 0.0, Dz, 70, 50)
    mat = np.concatenate([np.zeros(nleaf, np.uint32), np.full(objects.shape[0] - nleaf, 2, np.uint32)])
    return _finish("urban", objects, ground, leaf_mat=mat)


def terrain(seed: int = 5, n: int = 512, trees: int = 300, leaves: int = 1600, amplitude: float = 20.0) -> Scene:
    """Large terrain forest: 512 x 512 DEM at 1 m + 300 crowns (~1.0 M facets).

    Heights: the sum of 4 sines of `amplitude` metres each (random
    wavevectors, wavelengths 50-300 m along each axis, random phases) plus
    Gaussian noise of 0.5 m.  Crowns (semi-axes 2 x 3 x 2 m) are centred
    4 m above the highest ground point under their footprint."""
    rng = np.random.default_rng(seed)
    kx = rng.uniform(2 * np.pi / 300.0, 2 * np.pi / 50.0, 4)
    kz = rng.uniform(2 * np.pi / 300.0, 2 * np.pi / 50.0, 4)
    ph = rng.uniform(0, 2 * np.pi, 4)
    xs = np.arange(n, dtype=np.float64)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    H = sum(amplitude * np.sin(kx[k] * X + kz[k] * Z + ph[k]) for k in range(4)) + rng.normal(0.0, 0.5, X.shape)
    H = H - H.min()

    def height(x, z):
        return H

    ground = grid_ground(0.0, n - 1.0, 0.0, n - 1.0, n - 1, n - 1, height)
    tris = []
    for _ in range(trees):
        i, j = rng.integers(8, n - 8, 2)
        top = float(H[i - 2:i + 3, j - 2:j + 3].max())
        tris.append(crown_leaves((float(i), top + 4.0, float(j)), (2.0, 3.0, 2.0), leaves, 0.1, rng))
    return _finish("terrain", np.concatenate(tris), ground)


def tiny_pair(gap: float = 1.0, size: float = 1.0) -> Scene:
    """Two parallel coaxial squares (2 facets each) facing each other; the
    lower one is 'ground'.  Used by closed-form view-factor tests."""
    s = size
    top = np.array([[[0, gap, 0], [s, gap, 0], [0, gap, s]], [[s, gap, 0], [s, gap, s], [0, gap, s]]], float)
    bottom = grid_ground(0.0, s, 0.0, s, 1, 1)
    return _finish("pair", top, bottom)


def load_tree_with_soil(path, soil_resolution: int = 48) -> Scene:
    """OBJ scene as loadTreeWithSoil (core:362-473): faces fan-triangulated,
    y shifted so the lowest vertex sits at 0, every face an object facet,
    plus a soil grid of soil_resolution^2 x 2 triangles at y = -0.01 over the
    object's footprint widened by half its larger horizontal extent."""
    positions = []
    faces = []
    for line in open(path):
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "v":
            positions.append([float(x) for x in tok[1:4]])
        elif tok[0] == "f":
            idx = []
            for t in tok[1:]:
                raw = int(t.split("/")[0])
                idx.append(raw - 1 if raw > 0 else len(positions) + raw)
            for i in range(1, len(idx) - 1):
                faces.append((idx[0], idx[i], idx[i + 1]))
    P = np.asarray(positions, np.float32)
    lo = P.min(axis=0)
    hi = P.max(axis=0)
    P[:, 1] -= lo[1]
    tri = P[np.asarray(faces)]
    x0, x1, z0, z1 = float(lo[0]), float(hi[0]), float(lo[2]), float(hi[2])
    margin = np.float32(0.5) * max(np.float32(hi[0] - lo[0]), np.float32(hi[2] - lo[2]))
    x0, x1, z0, z1 = [np.float32(v) for v in (x0 - margin, x1 + margin, z0 - margin, z1 + margin)]
    n = soil_resolution
    soil = []
    for row in range(n):
        za = z0 + (z1 - z0) * np.float32(row) / np.float32(n)
        zb = z0 + (z1 - z0) * np.float32(row + 1) / np.float32(n)
        for col in range(n):
            xa = x0 + (x1 - x0) * np.float32(col) / np.float32(n)
            xb = x0 + (x1 - x0) * np.float32(col + 1) / np.float32(n)
            soil.append([[xa, -0.01, za], [xb, -0.01, zb], [xb, -0.01, za]])
            soil.append([[xa, -0.01, za], [xa, -0.01, zb], [xb, -0.01, zb]])
    sc = _finish("obj", tri, np.asarray(soil, np.float32))
    sc.bmin = np.array([x0, -0.01, z0], np.float32)
    sc.bmax = np.array([x1, hi[1] - lo[1], z1], np.float32)
    return sc


BUILDERS = {"row_crop": row_crop, "forest": forest, "urban": urban, "terrain": terrain}

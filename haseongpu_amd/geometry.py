"""Extruded triangle-prism mesh: ``Grid``, ``MeshTopology`` and ``GainMedium``.

Host-side mirror of the reference's prism-mesh builder
(pyInclude/geometry/core.py): ``Grid`` (core.py:788-863), the regular grid
triangulation (core.py:630-668), ``MeshTopology.fromGrid/fromPoints``
(core.py:866-1077) and the derived walk arrays of ``_derived_topology``
(core.py:671-774): triangle centres, surfaces, outward edge normals, normal
points, neighbours and forbidden edges.  Arrays that the device consumes are
flattened column-major (entry (t, k) at t + k*T, core.py:769-773).

``GainMedium`` keeps the legacy prism layouts of the physical properties:
``betaVolume`` (T, L-1) and ``betaCells`` (Np, L) stored Fortran-ordered,
``claddingCellTypes`` (T,), per-triangle ``reflectivities`` (2T, index
t*2 + {0 bottom, 1 top}, include/core/mesh.hpp:301-310) and
``refractiveIndices`` {bottom in, bottom out, top in, top out}
(include/core/mesh.hpp:312-332).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

__all__ = ["Grid", "MeshTopology", "GainMedium", "PropertyHandle"]


class Grid:
    """Axis-aligned box sampled by a regular lattice: a (nx x ny) point grid in
    x and y, split two triangles per cell, extruded over the level planes of
    the z axis.  The API is that of pyInclude/geometry/core.py:788-863, so a
    Grid built with the same arguments yields the same mesh."""

    # spellings the reference also accepts, mapped onto the canonical keywords
    _KEYWORD_SPELLINGS = {"x": "xExtent", "y": "yExtent", "z": "zExtent",
                          "tilesSizeX": "tileSizeX", "tilesSizeY": "tileSizeY", "tilesSizeZ": "tileSizeZ"}

    def __init__(self, xExtent=None, yExtent=None, zExtent=None, tileSizeX=1.0,
                 tileSizeY=None, tileSizeZ=None, origin=(0.0, 0.0), **extra):
        given = dict(xExtent=xExtent, yExtent=yExtent, zExtent=zExtent, tileSizeX=tileSizeX,
                     tileSizeY=tileSizeY, tileSizeZ=tileSizeZ)
        # a short extent name may only fill an extent left unset; the old
        # tile-size spellings always replace the tile size
        stray = [k for k in extra if k not in self._KEYWORD_SPELLINGS
                 or (self._KEYWORD_SPELLINGS[k].endswith("Extent") and given[self._KEYWORD_SPELLINGS[k]] is not None)]
        if stray:
            raise TypeError(f"Grid() got unexpected keyword(s): {', '.join(stray)}")
        for spelling, value in extra.items():
            given[self._KEYWORD_SPELLINGS[spelling]] = value
        missing = [k for k in ("xExtent", "yExtent", "zExtent") if given[k] is None]
        if missing:
            raise TypeError(f"Grid() needs an extent for every axis; missing {', '.join(missing)}")
        self.xExtent, self.yExtent, self.zExtent = (float(given[k]) for k in ("xExtent", "yExtent", "zExtent"))
        self.tileSizeX = float(given["tileSizeX"])
        self.tileSizeY = float(self.tileSizeX if given["tileSizeY"] is None else given["tileSizeY"])
        self.tileSizeZ = float(self.tileSizeX if given["tileSizeZ"] is None else given["tileSizeZ"])
        self.origin = tuple(float(v) for v in origin)
        if not min(self.xExtent, self.yExtent, self.zExtent) > 0.0:
            raise ValueError(f"Grid extents must be > 0, got {self.xExtent}, {self.yExtent}, {self.zExtent}")
        if not min(self.tileSizeX, self.tileSizeY, self.tileSizeZ) > 0.0:
            raise ValueError(f"Grid tile sizes must be > 0, got {self.tileSizeX}, {self.tileSizeY}, {self.tileSizeZ}")

    @staticmethod
    def _axisValues(extent, tileSize):
        # lattice 0, h, 2h, ... up to the extent; the arange bound reaches half
        # a tile past it so that an extent that is a whole number of tiles
        # keeps its last plane despite rounding in the multiples of h
        values = np.arange(0.0, extent + tileSize * 0.5, tileSize)
        return values[values <= extent + 1e-12]

    @property
    def numberOfLevels(self):
        return int(len(self._axisValues(self.zExtent, self.tileSizeZ)))

    @property
    def thickness(self):
        return self.tileSizeZ

    def constructPoints(self):
        ox, oy = self.origin
        xs = ox + self._axisValues(self.xExtent, self.tileSizeX)
        ys = oy + self._axisValues(self.yExtent, self.tileSizeY)
        return np.column_stack([np.tile(xs, ys.size), np.repeat(ys, xs.size)])


def _grid_triangles(nx, ny):
    """Two counter-clockwise triangles per grid cell, row by row; the first
    half of each row holds (p00, p10, p01), the second (p01, p10, p11)
    (pyInclude/geometry/core.py:630-668)."""
    cx, cy = nx - 1, ny - 1
    jj, ii = np.meshgrid(np.arange(cy), np.arange(cx), indexing="ij")
    p00 = jj * nx + ii
    p10 = p00 + 1
    p01 = p00 + nx
    p11 = p01 + 1
    first = np.stack([p00, p10, p01], axis=-1)
    second = np.stack([p01, p10, p11], axis=-1)
    rows = np.concatenate([first, second], axis=1)
    return rows.reshape(-1, 3).astype(np.uint32)


def _distinct_points(points):
    """(N, 2) float64 points with repeated coordinates removed, each kept at
    its first position (core.py:555-562)."""
    arr = np.asarray(points, dtype=np.float64)
    if arr.ndim != 2 or arr.shape[1] != 2:
        raise ValueError(f"points must have shape (N, 2), got {arr.shape}")
    if arr.shape[0] < 3:
        raise ValueError("at least three points are required")
    seen, keep = set(), []
    for i, (x, y) in enumerate(arr.tolist()):
        # -0.0 and 0.0 hash and compare equal, as they do for the rows of np.unique
        if (x, y) not in seen:
            seen.add((x, y))
            keep.append(i)
    return arr[np.asarray(keep, dtype=np.int64)]


def bowyerWatson(points):
    """Incremental Delaunay triangulation with the insertion rules of
    core.py:583-627, vectorised over the current triangle list.

    The result depends on more than the point set: triangles whose
    circumcircle holds a new point are removed from the list in place, the
    cavity's boundary edges (counted in the order the removed triangles list
    them) are re-joined to the point as (low id, high id, point), swapped to
    counter-clockwise, and appended; the three enclosing vertices start as
    one triangle in the iteration order of the set of their ids.  Every
    in-circle decision uses the reference's determinant expression and its
    1e-12 margin on float64 values, term for term, so the decisions - and
    with them triangle order and rotation - are identical.
    Returns (points without duplicates, (T, 3) uint32 triangles)."""
    pts = _distinct_points(points)
    n = pts.shape[0]
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    span = max(hi - lo)
    if span <= 0.0:
        raise ValueError("points must span a two-dimensional area")
    mid = (lo + hi) / 2.0
    box = np.array([[mid[0] - 20.0 * span, mid[1] - span],
                    [mid[0], mid[1] + 20.0 * span],
                    [mid[0] + 20.0 * span, mid[1] - span]], dtype=np.float64)
    allp = np.vstack([pts, box])
    X, Y = allp[:, 0], allp[:, 1]
    # the first triangle takes its vertices in the set's iteration order
    tri = np.asarray([tuple({n, n + 1, n + 2})], dtype=np.int64)

    def orientation(t):
        ax, ay = X[t[:, 0]], Y[t[:, 0]]
        return (X[t[:, 1]] - ax) * (Y[t[:, 2]] - ay) - (Y[t[:, 1]] - ay) * (X[t[:, 2]] - ax)

    ccw = orientation(tri) > 0.0
    for p in range(n):
        px, py = X[p], Y[p]
        ax, ay = X[tri[:, 0]] - px, Y[tri[:, 0]] - py
        bx, by = X[tri[:, 1]] - px, Y[tri[:, 1]] - py
        cx, cy = X[tri[:, 2]] - px, Y[tri[:, 2]] - py
        det = ((ax * ax + ay * ay) * (bx * cy - cx * by)
               - (bx * bx + by * by) * (ax * cy - cx * ay)
               + (cx * cx + cy * cy) * (ax * by - bx * ay))
        inside = np.where(ccw, det > 1e-12, det < -1e-12)
        cavity = tri[inside]
        count = {}
        for a, b, c in cavity.tolist():
            for e in ((a, b), (b, c), (c, a)):
                key = (e[0], e[1]) if e[0] < e[1] else (e[1], e[0])
                count[key] = count.get(key, 0) + 1
        rim = [key for key, k in count.items() if k == 1]
        new = np.empty((len(rim), 3), dtype=np.int64)
        for i, (a, b) in enumerate(rim):
            # counter-clockwise test on the candidate (a, b, p)
            turn = (X[b] - X[a]) * (py - Y[a]) - (Y[b] - Y[a]) * (px - X[a])
            new[i] = (b, a, p) if turn < 0.0 else (a, b, p)
        tri = np.concatenate([tri[~inside], new])
        ccw = np.concatenate([ccw[~inside], orientation(new) > 0.0])
    keep = np.all(tri < n, axis=1)
    if not np.any(keep):
        raise ValueError("could not construct triangles; points may be collinear")
    return pts, tri[keep].astype(np.uint32)


def _orient_ccw(points, triangles):
    tri = np.asarray(triangles, dtype=np.int64).copy()
    a, b, c = points[tri[:, 0]], points[tri[:, 1]], points[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    flip = cross < 0.0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    if np.any(cross == 0.0):
        raise ValueError("degenerate (zero-area) triangle in the base mesh")
    return tri.astype(np.uint32)


def derivedTopology(points, triangles):
    """Walk arrays of a counter-clockwise triangle mesh, column-major.

    Edge k of triangle t joins vertices k and (k+1)%3; its outward unit normal
    is (e_y, -e_x)/|e| and its normal point is vertex k; neighbours across a
    shared edge and the edge's index inside that neighbour ("forbidden edge")
    come from matching sorted vertex pairs (core.py:671-774)."""
    pts = np.asarray(points, dtype=np.float64)
    tri = np.asarray(triangles, dtype=np.int64)
    T = tri.shape[0]
    x, y = pts[:, 0], pts[:, 1]
    cx = x[tri].mean(axis=1)
    cy = y[tri].mean(axis=1)
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    cross = (x[b] - x[a]) * (y[c] - y[a]) - (y[b] - y[a]) * (x[c] - x[a])
    surfaces = (0.5 * np.abs(cross)).astype(np.float32)
    start = tri
    end = np.roll(tri, -1, axis=1)
    ex = x[end] - x[start]
    ey = y[end] - y[start]
    norm = np.sqrt(ex * ex + ey * ey)
    nx = ey / norm
    ny = -ex / norm
    lo = np.minimum(start, end).reshape(-1)
    hi = np.maximum(start, end).reshape(-1)
    owner = np.repeat(np.arange(T), 3)
    local = np.tile(np.arange(3), T)
    order = np.lexsort((hi, lo))
    lo_s, hi_s = lo[order], hi[order]
    same = (lo_s[1:] == lo_s[:-1]) & (hi_s[1:] == hi_s[:-1])
    if np.any(same[1:] & same[:-1]):
        raise ValueError("non-manifold base mesh: an edge is shared by more than two triangles")
    left = order[:-1][same]
    right = order[1:][same]
    neighbors = np.full((T, 3), -1, dtype=np.int32)
    forbidden = np.full((T, 3), -1, dtype=np.int32)
    neighbors[owner[left], local[left]] = owner[right]
    forbidden[owner[left], local[left]] = local[right]
    neighbors[owner[right], local[right]] = owner[left]
    forbidden[owner[right], local[right]] = local[left]
    return {
        "triangleCenterX": cx,
        "triangleCenterY": cy,
        "triangleSurfaces": surfaces,
        "triangleNormalsX": nx.reshape(-1, order="F"),
        "triangleNormalsY": ny.reshape(-1, order="F"),
        "triangleNormalPoint": start.astype(np.uint32).reshape(-1, order="F"),
        "triangleNeighbors": neighbors.reshape(-1, order="F"),
        "forbiddenEdge": forbidden.reshape(-1, order="F"),
    }


@dataclass
class MeshTopology:
    """Planar triangle mesh extruded over ``levels`` z planes ``thickness`` apart
    (pyInclude/geometry/core.py:866-1077)."""

    points: np.ndarray
    trianglePointIndices: np.ndarray
    levels: int | None = None
    thickness: float | None = None
    metadata: dict = field(default_factory=dict)

    def __post_init__(self):
        self.points = np.asarray(self.points, dtype=np.float64)
        if self.points.ndim != 2 or self.points.shape[1] != 2 or self.points.shape[0] < 3:
            raise ValueError(f"points must have shape (N>=3, 2), got {self.points.shape}")
        tri = np.asarray(self.trianglePointIndices)
        if tri.ndim != 2 or tri.shape[1] != 3 or tri.shape[0] < 1:
            raise ValueError("trianglePointIndices must have shape (T, 3)")
        if tri.min() < 0 or tri.max() >= self.points.shape[0]:
            raise ValueError("triangle point index out of range")
        self.trianglePointIndices = _orient_ccw(self.points, tri)
        self._derived = None

    @classmethod
    def fromPoints(cls, points, numberOfLevels=None, thickness=None):
        """Delaunay-triangulate transverse points by incremental insertion
        (pyInclude/geometry/core.py:583-627, called at core.py:886-889).
        Duplicate points are dropped (first occurrence kept); the triangle
        list, its order and each triangle's vertex rotation are those of the
        reference's insertion, so per-triangle arrays built against a
        reference topology index the same triangles here."""
        pts, tri = bowyerWatson(points)
        return cls(pts, tri, levels=numberOfLevels, thickness=thickness,
                   metadata={"source": "delaunay"})

    @classmethod
    def fromGrid(cls, grid):
        if not isinstance(grid, Grid):
            raise TypeError("fromGrid expects a Grid instance")
        pts = grid.constructPoints()
        nx = len(Grid._axisValues(grid.xExtent, grid.tileSizeX))
        ny = len(Grid._axisValues(grid.yExtent, grid.tileSizeY))
        if nx < 2 or ny < 2:
            raise ValueError("grid must contain at least two points along x and y")
        topo = cls(pts, _grid_triangles(nx, ny), levels=grid.numberOfLevels,
                   thickness=grid.thickness)
        topo.metadata.update({"source": "grid", "grid": {
            "x": grid.xExtent, "y": grid.yExtent, "z": grid.zExtent,
            "tileSizeX": grid.tileSizeX, "tileSizeY": grid.tileSizeY,
            "tileSizeZ": grid.tileSizeZ}})
        return topo

    def numberOfLevels(self, levels):
        levels = int(levels)
        if levels < 2:
            raise ValueError("numberOfLevels must be at least 2")
        self.levels = levels
        return self

    def withThickness(self, thickness):
        thickness = float(thickness)
        if thickness <= 0.0:
            raise ValueError("thickness must be positive")
        self.thickness = thickness
        return self

    @property
    def numberOfTriangles(self):
        return int(self.trianglePointIndices.shape[0])

    @property
    def numberOfPoints(self):
        return int(self.points.shape[0])

    @property
    def numberOfPrisms(self):
        self._require_levels()
        return self.numberOfTriangles * (int(self.levels) - 1)

    @property
    def numberOfSamples(self):
        self._require_levels()
        return self.numberOfPoints * int(self.levels)

    def levelCoordinates(self):
        self._require_levels()
        self._require_thickness()
        return np.arange(int(self.levels), dtype=np.float64) * float(self.thickness)

    def samplePoints(self):
        """(S, 3) coordinates of sample s = p + l * Np."""
        z = self.levelCoordinates()
        xy = np.tile(self.points, (int(self.levels), 1))
        return np.column_stack([xy, np.repeat(z, self.numberOfPoints)])

    def prismCentroids(self):
        """(P, 3) centroids of prism j = t + l * T."""
        d = self.topology()
        z = (np.arange(int(self.levels) - 1) + 0.5) * float(self.thickness)
        T = self.numberOfTriangles
        return np.column_stack([np.tile(d["triangleCenterX"], int(self.levels) - 1),
                                np.tile(d["triangleCenterY"], int(self.levels) - 1),
                                np.repeat(z, T)])

    def prismVolumes(self):
        d = self.topology()
        return np.tile(d["triangleSurfaces"].astype(np.float64), int(self.levels) - 1) * float(self.thickness)

    def topology(self):
        if self._derived is None:
            self._derived = derivedTopology(self.points, self.trianglePointIndices)
        return self._derived

    def boundaryTriangles(self):
        """Mask of triangles with an exterior edge (the cladding-ring rule of
        tests/data/laserPumpCladding/true_cladding_no_reflection_phiase_reference/
        generate_reference.py:81-105)."""
        nb = self.topology()["triangleNeighbors"].reshape(3, -1)
        return np.any(nb < 0, axis=0)

    def _require_levels(self):
        if self.levels is None or int(self.levels) < 2:
            raise ValueError("numberOfLevels (>= 2) is required")

    def _require_thickness(self):
        if self.thickness is None or float(self.thickness) <= 0.0:
            raise ValueError("a positive thickness is required")


_SCALARS = {
    "nTot": (np.float64, 0.0),
    "crystalTFluo": (np.float64, 0.0),
    "claddingNumber": (np.uint32, 1),
    "claddingAbsorption": (np.float64, 0.0),
}

_ALIASES = {
    "cladding_cell_types": "claddingCellTypes",
    "cladding_absorption": "claddingAbsorption",
    "cladding_number": "claddingNumber",
    "crystal_t_fluo": "crystalTFluo",
    "beta_volume": "betaVolume",
    "beta_cells": "betaCells",
}


class PropertyHandle:
    """``GainMedium.get(name)``: expected shape, dtype and value of a property."""

    def __init__(self, medium, name):
        self._medium = medium
        self.name = name

    @property
    def expectedShape(self):
        return self._medium._shape(self.name)

    @property
    def dtype(self):
        return self._medium._dtype(self.name)

    @property
    def value(self):
        return self._medium.physical.get(self.name, self._medium._default(self.name))

    @value.setter
    def value(self, values):
        self._medium.set(self.name, values)


@dataclass
class GainMedium:
    """Mesh plus the physical state the ASE estimator reads
    (pyInclude/geometry/core.py:1080-1613, property specs core.py:495-552)."""

    topology: MeshTopology
    physical: dict = field(default_factory=dict)

    _ARRAYS = ("betaVolume", "betaCells", "claddingCellTypes", "reflectivities", "refractiveIndices")

    def _canonical(self, name):
        name = _ALIASES.get(name, name)
        if name not in self._ARRAYS and name not in _SCALARS:
            raise KeyError(f"unknown gain-medium property '{name}'")
        return name

    def _shape(self, name):
        topo = self.topology
        if name == "betaVolume":
            return (topo.numberOfTriangles, int(topo.levels) - 1)
        if name == "betaCells":
            return (topo.numberOfPoints, int(topo.levels))
        if name == "claddingCellTypes":
            return (topo.numberOfTriangles,)
        if name == "reflectivities":
            return (2 * topo.numberOfTriangles,)
        if name == "refractiveIndices":
            return (4,)
        return ()

    def _dtype(self, name):
        if name == "claddingCellTypes":
            return np.uint32
        if name in ("reflectivities", "refractiveIndices"):
            return np.float32
        if name in _SCALARS:
            return _SCALARS[name][0]
        return np.float64

    def _default(self, name):
        if name in _SCALARS:
            return _SCALARS[name][0](_SCALARS[name][1])
        if name == "refractiveIndices":
            return np.array([1.0, 1.0, 1.0, 1.0], dtype=np.float32)
        return np.zeros(int(np.prod(self._shape(name))), dtype=self._dtype(name))

    def get(self, name):
        return PropertyHandle(self, self._canonical(name))

    def set(self, name, value):
        name = self._canonical(name)
        if name in _SCALARS:
            self.physical[name] = _SCALARS[name][0](value)
            return self
        shape = self._shape(name)
        arr = np.asarray(value, dtype=self._dtype(name))
        # matrix-shaped input is stored Fortran-ordered; flat input is taken as is
        flat = arr.reshape(-1, order="F") if (len(shape) > 1 and arr.shape == shape) else arr.reshape(-1)
        if flat.size != int(np.prod(shape)):
            raise ValueError(f"{name} must have {int(np.prod(shape))} values (shape {shape}), got {flat.size}")
        if name in ("betaVolume", "betaCells") and not np.all(np.isfinite(flat)):
            raise ValueError(f"{name} must be finite")
        self.physical[name] = np.ascontiguousarray(flat)
        return self

    def withPhysicalProperties(self, **properties):
        for name, value in properties.items():
            self.set(name, value)
        return self

    def value(self, name):
        return self.get(name).value

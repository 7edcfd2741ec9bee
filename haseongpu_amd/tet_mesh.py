"""Tetrahedral (Tet4) gain-medium meshes for the forward estimator.

``TetMesh`` holds what the forward path reads (include/core/mesh.hpp:143-190:
points SoA, cellPointIndices, claddingCellTypes, betaVolume, nTot,
crystalTFluo, claddingNumber, claddingAbsorption).  ``TetMesh.fromPrisms``
splits every prism of an extruded triangle mesh into three tetrahedra the
way utils/convert_vtk_topology.py:38-53 does: the base corners sorted by
global point index (a < b < c, tops d, e, f) give [a,b,c,d], [b,e,c,d],
[c,e,f,d], so both prisms beside a quadrilateral face cut it along the same
diagonal and the tetrahedra conform.  Point p of level l is point
p + l * Np; prism j = t + l * T becomes cells 3 j, 3 j + 1, 3 j + 2 and
keeps its beta and its triangle's cladding type.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["TetMesh"]


@dataclass
class TetMesh:
    points: np.ndarray                 # (Np, 3) float64
    cells: np.ndarray                  # (C, 4) uint32 point indices
    beta: np.ndarray                   # (C,) float64 excited fraction per cell
    cell_types: np.ndarray | None = None   # (C,) uint32; == cladding_number marks cladding
    cladding_number: int = 3
    cladding_absorption: float = 0.0
    n_tot: float = 2.76e20
    tau: float = 9.5e-4

    def __post_init__(self):
        self.points = np.ascontiguousarray(self.points, dtype=np.float64).reshape(-1, 3)
        self.cells = np.ascontiguousarray(self.cells, dtype=np.uint32).reshape(-1, 4)
        self.beta = np.ascontiguousarray(self.beta, dtype=np.float64).reshape(-1)
        if self.cell_types is not None:
            self.cell_types = np.ascontiguousarray(self.cell_types, dtype=np.uint32).reshape(-1)
        if self.beta.size != self.n_cells:
            raise ValueError(f"beta needs {self.n_cells} values, got {self.beta.size}")
        if self.cell_types is not None and self.cell_types.size != self.n_cells:
            raise ValueError(f"cell_types needs {self.n_cells} values, got {self.cell_types.size}")
        if self.cells.size and int(self.cells.max()) >= self.n_points:
            raise ValueError("cell point index out of range")

    @property
    def n_points(self):
        return int(self.points.shape[0])

    @property
    def n_cells(self):
        return int(self.cells.shape[0])

    def cladding_mask(self):
        if self.cell_types is None:
            return np.zeros(self.n_cells, dtype=bool)
        return self.cell_types == np.uint32(self.cladding_number)

    def volumes(self):
        p = self.points[self.cells]
        a = p[:, 1] - p[:, 0]
        b = p[:, 2] - p[:, 0]
        c = p[:, 3] - p[:, 0]
        return np.abs(np.einsum("ij,ij->i", np.cross(a, b), c)) / 6.0

    def centroids(self):
        return self.points[self.cells].mean(axis=1)

    def faceSurfacesByPlane(self, planes, *, atol=1e-12):
        """Surface ids of the cell faces (C, 4), face k opposite vertex k: a
        face whose three vertices lie on the plane z = z0 gets planes[z0]
        (e.g. {0.0: 1, thickness_total: 2} for the bottom and top of an
        extruded mesh), every other face 0 (no surface)."""
        out = np.zeros((self.n_cells, 4), dtype=np.int32)
        z = self.points[self.cells][:, :, 2]                     # (C, 4)
        for z0, sid in planes.items():
            on = np.abs(z - float(z0)) <= atol                   # vertex on the plane
            for k in range(4):
                others = [j for j in range(4) if j != k]
                out[:, k] = np.where(on[:, others].all(axis=1), int(sid), out[:, k])
        return out

    @classmethod
    def fromPrisms(cls, topology, beta_volume, *, cladding_types=None, cladding_number=3, cladding_absorption=0.0,
                   n_tot=2.76e20, tau=9.5e-4):
        """Three tetrahedra per prism of a ``MeshTopology`` with levels and
        thickness; ``beta_volume`` has one value per prism (flat j = t + l T,
        or shape (T, L-1)); ``cladding_types`` one per triangle."""
        topology._require_levels()
        topology._require_thickness()
        Np, T, L = topology.numberOfPoints, topology.numberOfTriangles, int(topology.levels)
        xy = np.asarray(topology.points, dtype=np.float64).reshape(-1, 2)
        z = topology.levelCoordinates()
        pts = np.column_stack([np.tile(xy, (L, 1)), np.repeat(z, Np)])
        tri = np.sort(np.asarray(topology.trianglePointIndices, dtype=np.int64).reshape(-1, 3), axis=1)
        a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
        cells = []
        for level in range(L - 1):
            lo, hi = level * Np, (level + 1) * Np
            d, e, f = a + hi, b + hi, c + hi
            tets = np.stack([np.column_stack([a + lo, b + lo, c + lo, d]),
                             np.column_stack([b + lo, e, c + lo, d]),
                             np.column_stack([c + lo, e, f, d])], axis=1)   # (T, 3, 4)
            cells.append(tets.reshape(-1, 4))
        cells = np.concatenate(cells)
        bv = np.asarray(beta_volume, dtype=np.float64)
        bv = bv.reshape(-1, order="F") if bv.ndim == 2 else bv.reshape(-1)
        if bv.size != T * (L - 1):
            raise ValueError(f"beta_volume needs {T * (L - 1)} values (one per prism), got {bv.size}")
        types = None
        if cladding_types is not None:
            ct = np.asarray(cladding_types, dtype=np.uint32).reshape(-1)
            if ct.size != T:
                raise ValueError(f"cladding_types needs {T} values (one per triangle), got {ct.size}")
            types = np.repeat(np.tile(ct, L - 1), 3)
        return cls(points=pts, cells=cells, beta=np.repeat(bv, 3), cell_types=types, cladding_number=cladding_number,
                   cladding_absorption=cladding_absorption, n_tot=n_tot, tau=tau)

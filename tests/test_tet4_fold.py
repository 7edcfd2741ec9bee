"""hase_fold_tet4_mesh: the prism arrays hase_mesh_t takes, from the Tet4
arrays of a 2.2.0 HostMesh (CPU; the fold is host code).

The Tet4 input is made the way utils/convert_vtk_topology.py:38-53 splits a
wedge mesh (``TetMesh.fromPrisms``), with the surface tables of
include/core/mesh.hpp:300-332.  The fold must give back the prism mesh it came
from, bit for bit, once each triangle of the original is written
counter-clockwise from its lowest point index (the rotation the split keeps).
"""
import numpy as np
import pytest

from haseongpu_amd import _native
from haseongpu_amd.geometry import MeshTopology, derivedTopology
from haseongpu_amd.tet_mesh import TetMesh
from haseongpu_amd.workloads import disk_points


def _canonical(tri):
    """Each row rotated to start at its lowest entry (orientation kept)."""
    tri = np.asarray(tri, dtype=np.int64)
    k = np.argmin(tri, axis=1)
    return np.stack([tri[np.arange(len(tri)), (k + i) % 3] for i in range(3)], axis=1)


def _slab(levels=5, seed=3):
    topo = MeshTopology.fromPoints(disk_points(2.0, 260, seed=seed), numberOfLevels=levels, thickness=0.05)
    T, L = topo.numberOfTriangles, levels
    rng = np.random.default_rng(seed)
    beta = rng.uniform(0.0, 0.3, size=T * (L - 1))
    clad = topo.boundaryTriangles().astype(np.uint32)
    refl = rng.uniform(0.0, 0.1, size=2 * T).astype(np.float32)
    return topo, beta, clad, refl


def _host_mesh(topo, beta, clad, refl, n_in=1.83, n_out=1.0):
    """2.2.0 HostMesh-like arrays: points SoA, cells, per-cell beta and types,
    and face surface ids 1 + t (bottom) / 1 + T + t (top) into surface tables."""
    T, L = topo.numberOfTriangles, int(topo.levels)
    tm = TetMesh.fromPrisms(topo, beta, cladding_types=clad, cladding_number=1)
    z_top = float(topo.levelCoordinates()[-1])
    fs = tm.faceSurfacesByPlane({0.0: -1, z_top: -2})
    tri = (np.arange(tm.n_cells) // 3) % T
    fs = np.where(fs == -1, 1 + tri[:, None], np.where(fs == -2, 1 + T + tri[:, None], fs))
    table = np.concatenate([[0.0], refl[0::2], refl[1::2]]).astype(np.float32)
    n = 1 + 2 * T
    return tm, dict(points=tm.points.T.reshape(-1), cells=tm.cells.reshape(-1), beta_volume=tm.beta,
                    number_of_points=topo.numberOfPoints, number_of_levels=L, cladding_cell_types=tm.cell_types,
                    cell_face_boundaries=fs, surface_reflectivities=table,
                    surface_refractive_index_inside=np.full(n, n_in, np.float32),
                    surface_refractive_index_outside=np.full(n, n_out, np.float32))


def test_fold_gives_back_the_prism_mesh_bit_for_bit():
    topo, beta, clad, refl = _slab()
    T, L = topo.numberOfTriangles, int(topo.levels)
    tm, host = _host_mesh(topo, beta, clad, refl)
    out = _native.fold_tet4_mesh(**host)
    assert out["number_of_triangles"] == T
    assert out["thickness"] == topo.thickness
    np.testing.assert_array_equal(out["points"], np.concatenate([topo.points[:, 0], topo.points[:, 1]]))
    tri = _canonical(topo.trianglePointIndices)
    np.testing.assert_array_equal(out["triangles"].reshape(3, T).T, tri)
    d = derivedTopology(topo.points, tri)
    np.testing.assert_array_equal(out["neighbors"], d["triangleNeighbors"])
    np.testing.assert_array_equal(out["forbidden"], d["forbiddenEdge"])
    np.testing.assert_array_equal(out["beta"], beta)
    np.testing.assert_array_equal(out["cladding"], clad)
    np.testing.assert_array_equal(out["reflectivities"], refl)
    np.testing.assert_array_equal(out["refractive_indices"], np.array([1.83, 1.0, 1.83, 1.0], np.float32))
    # the canonical rotation is still the same mesh: the original's neighbour
    # relation, up to the renumbered edges
    orig = derivedTopology(topo.points, topo.trianglePointIndices.astype(np.int64))
    assert sorted(orig["triangleNeighbors"].tolist()) == sorted(out["neighbors"].tolist())


def test_legacy_tables_and_mixed_cell_beta():
    """Without surface ids the legacy entries are read: index cell*2 + face%2
    for reflectivity and 2*(face%2) + {0 in, 1 out} for refractive indices
    (mesh.hpp:300-332).  The bottom face of a wedge is face 3 of its first
    cell [a,b,c,d] (odd: entry 1), the top face is face 0 of its third cell
    [c,e,f,d] (even: entry 0).  Cells of one wedge with different beta fold to
    their volume-weighted mean."""
    topo, beta, clad, refl = _slab(levels=3, seed=5)
    T, L = topo.numberOfTriangles, 3
    tm, host = _host_mesh(topo, beta, clad, refl)
    C_ = tm.n_cells
    legacy = np.full(2 * C_, -1.0, np.float32)
    for t in range(T):
        legacy[2 * (3 * t) + 1] = refl[2 * t]                         # bottom: first cell of wedge t, face 3
        legacy[2 * (3 * (t + (L - 2) * T) + 2) + 0] = refl[2 * t + 1]   # top: third cell of the top wedge, face 0
    host.pop("cell_face_boundaries")
    for k in ("surface_reflectivities", "surface_refractive_index_inside", "surface_refractive_index_outside"):
        host.pop(k)
    cell_beta = tm.beta.copy()
    cell_beta[0:3] = [0.1, 0.2, 0.4]                                     # wedge 0 = prism 0
    host["beta_volume"] = cell_beta
    out = _native.fold_tet4_mesh(**host, reflectivities=legacy, refractive_indices=[1.7, 1.1, 1.6, 1.2])
    np.testing.assert_array_equal(out["reflectivities"], refl)
    # base face 3 (odd) reads entries 2, 3; top face 0 (even) reads entries 0, 1
    np.testing.assert_array_equal(out["refractive_indices"], np.array([1.6, 1.2, 1.7, 1.1], np.float32))
    vol = np.array([abs(np.linalg.det(tm.points[c[1:]] - tm.points[c[0]])) / 6 for c in tm.cells[:3]])
    assert out["beta"][0] == pytest.approx(float(np.dot(vol, [0.1, 0.2, 0.4]) / vol.sum()), rel=1e-14)
    np.testing.assert_array_equal(out["beta"][1:], beta[1:])


def test_fold_rejects_what_a_prism_mesh_cannot_hold():
    topo, beta, clad, refl = _slab(levels=3, seed=7)
    tm, host = _host_mesh(topo, beta, clad, refl)
    bad = dict(host)
    types = tm.cell_types.copy()
    types[3 * topo.numberOfTriangles + 1] ^= 1                          # one cell of level 1 changes type
    bad["cladding_cell_types"] = types
    with pytest.raises(_native.HaseError, match="cladding"):
        _native.fold_tet4_mesh(**bad)
    bad = dict(host)
    bad["cells"] = np.roll(tm.cells, 3, axis=0).reshape(-1)              # wedges out of t + l*T order
    with pytest.raises(_native.HaseError):
        _native.fold_tet4_mesh(**bad)
    bad = dict(host)
    bad["number_of_points"] = topo.numberOfPoints + 1
    with pytest.raises(_native.HaseError, match="number_of_mesh_points"):
        _native.fold_tet4_mesh(**bad)

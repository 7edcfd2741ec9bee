"""Legacy VTK writers on the wedge mesh (CPU).

The first five cases restate the reference's own writer tests
(tests/python/io/test_vtkWedge.py:40-193: cell data from a state through the
callback, several and aliased fields, fields written directly from a state's
topology, mixed point and cell data).  The grid itself is pinned by geometry
rather than by the writer's own formulas: every cell's six points are the
triangle's corners on two adjacent level planes, and the wedge volume
computed from the written coordinates equals the prism volume of the mesh.
"""
import numpy as np
import pytest

from haseongpu_amd import Grid, MeshTopology, TimeStepState, compareVtk, vtkWedge, writePointsToVtk, writePrismToVtk
from haseongpu_amd.vtk_wedge import readVtkScalars


def _topology():
    return MeshTopology.fromGrid(Grid(xExtent=1, yExtent=1, zExtent=0.5, tileSizeZ=0.25))


def _state(topology, step, with_topology=False):
    v = np.arange(topology.numberOfPrisms, dtype=np.float64).reshape(
        (topology.numberOfTriangles, topology.levels - 1), order="F")
    return TimeStepState(step=step, time=step * 1e-5, betaVolume=v, phiAse=v + 1.0, dndtAse=v + 2.0,
                         dndtPump=v + 3.0, aseResult=None, standardError=np.zeros_like(v),
                         relativeStandardError=np.zeros_like(v), totalRays=np.zeros_like(v, dtype=np.uint32),
                         topology=topology if with_topology else None)


def _grid(path):
    """(points (n, 3), cells (m, 6), cell types) read back from a file."""
    lines = path.read_text(encoding="utf-8").splitlines()
    i = next(k for k, s in enumerate(lines) if s.startswith("POINTS "))
    n = int(lines[i].split()[1])
    pts = np.array([s.split() for s in lines[i + 1:i + 1 + n]], dtype=np.float64)
    i = next(k for k, s in enumerate(lines) if s.startswith("CELLS "))
    m = int(lines[i].split()[1])
    cells = np.array([s.split() for s in lines[i + 1:i + 1 + m]], dtype=np.int64)
    i = next(k for k, s in enumerate(lines) if s.startswith("CELL_TYPES "))
    types = np.array(lines[i + 1:i + 1 + m], dtype=np.int64)
    assert np.all(cells[:, 0] == 6)
    return pts, cells[:, 1:], types


def test_cell_data_from_state_through_callback(tmp_path):
    topo = _topology()
    path = vtkWedge(tmp_path / "phi_{step:03d}", topo)(_state(topo, 2))
    text = path.read_text(encoding="utf-8")
    assert path.name == "phi_002.vtk"
    assert "DATASET UNSTRUCTURED_GRID" in text
    assert f"POINTS {topo.numberOfPoints * topo.levels} float" in text
    assert f"CELLS {topo.numberOfPrisms} {topo.numberOfPrisms * 7}" in text
    assert f"CELL_TYPES {topo.numberOfPrisms}" in text and f"CELL_DATA {topo.numberOfPrisms}" in text
    kind, values = readVtkScalars(path)["phiAse"]
    assert kind == "CELL_DATA"
    np.testing.assert_array_equal(values, np.arange(topo.numberOfPrisms) + 1.0)


def test_several_and_aliased_fields(tmp_path):
    topo = _topology()
    st = _state(topo, 3)
    path = vtkWedge(tmp_path / "out_{field}_{step:03d}", topo, field=["phiAse", "dndtAse"])(st)
    assert path.name == "out_phiAse_dndtAse_003.vtk"
    assert path.read_text(encoding="utf-8").count(f"CELL_DATA {topo.numberOfPrisms}") == 1
    got = readVtkScalars(path)
    np.testing.assert_array_equal(got["dndtAse"][1], st.dndtAse.reshape(-1, order="F"))
    path = vtkWedge(tmp_path / "named_{step:03d}", topo, field={"phi": "phiAse", "dn": "dndtAse"})(_state(topo, 4))
    assert path.name == "named_004.vtk" and set(readVtkScalars(path)) == {"phi", "dn"}


def test_fields_written_directly_from_state_topology(tmp_path):
    topo = _topology()
    st = _state(topo, 5, with_topology=True)
    path = vtkWedge(tmp_path / "state_{step:03d}", st,
                    fields={"betaVolume": st.betaVolume, "phiASE": st.phiAse, "dndtAse": "dndtAse",
                            "cladAbs": st.phiAse * 5.5})
    got = readVtkScalars(path)
    assert path.name == "state_005.vtk"
    assert {k: v[0] for k, v in got.items()} == {"betaVolume": "CELL_DATA", "phiASE": "CELL_DATA",
                                                "dndtAse": "CELL_DATA", "cladAbs": "CELL_DATA"}
    np.testing.assert_array_equal(got["cladAbs"][1], (st.phiAse * 5.5).reshape(-1, order="F"))


def test_mixed_point_and_cell_fields(tmp_path):
    topo = _topology()
    pv = np.arange(topo.numberOfPoints * topo.levels, dtype=np.float64).reshape(
        (topo.numberOfPoints, topo.levels), order="F")
    cv = np.arange(topo.numberOfPrisms, dtype=np.float64).reshape((topo.numberOfTriangles, topo.levels - 1),
                                                                  order="F")
    path = vtkWedge(tmp_path / "mixed", geometry=topo, fields={"phiAse": pv, "betaVolume": cv})
    text = path.read_text(encoding="utf-8")
    assert f"POINT_DATA {topo.numberOfPoints * topo.levels}" in text and f"CELL_DATA {topo.numberOfPrisms}" in text
    assert text.index("POINT_DATA") < text.index("CELL_DATA")
    got = readVtkScalars(path)
    assert got["phiAse"][0] == "POINT_DATA" and got["betaVolume"][0] == "CELL_DATA"
    # sample s = p + l * Np is VTK point s
    np.testing.assert_array_equal(got["phiAse"][1], np.arange(topo.numberOfPoints * topo.levels))


def test_grid_geometry_and_full_precision(tmp_path):
    rng = np.random.default_rng(11)
    pts = rng.random((30, 2)) * 2.0
    topo = MeshTopology.fromPoints(pts, numberOfLevels=4, thickness=0.3)
    beta = rng.random((topo.numberOfTriangles, 3))
    path = vtkWedge(tmp_path / "g", beta, geometry=topo, scalar_name="beta")
    xyz, cells, types = _grid(path)
    assert np.all(types == 13) and cells.shape == (topo.numberOfPrisms, 6)
    # values read back bit for bit (17 significant digits)
    np.testing.assert_array_equal(readVtkScalars(path)["beta"][1], beta.reshape(-1, order="F"))
    # each cell: a triangle on level plane l, the same triangle on plane l + 1
    lo, hi = xyz[cells[:, :3]], xyz[cells[:, 3:]]
    np.testing.assert_array_equal(lo[:, :, :2], hi[:, :, :2])
    assert np.allclose(hi[:, :, 2] - lo[:, :, 2], 0.3, rtol=0, atol=1e-15)
    assert np.all(np.ptp(lo[:, :, 2], axis=1) == 0)
    # wedge volume from the written coordinates equals the prism volume (the
    # mesh keeps triangle surfaces in FP32, hence the tolerance)
    e1, e2 = lo[:, 1, :2] - lo[:, 0, :2], lo[:, 2, :2] - lo[:, 0, :2]
    area = 0.5 * np.abs(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    np.testing.assert_allclose(area * (hi[:, 0, 2] - lo[:, 0, 2]), topo.prismVolumes(), rtol=2e-7)
    # cell j = t + l * T sits on level l
    levels = np.repeat(np.arange(3), topo.numberOfTriangles)
    np.testing.assert_allclose(lo[:, 0, 2], levels * 0.3, rtol=0, atol=1e-15)


def test_callback_period_and_shape_errors(tmp_path):
    topo = _topology()
    cb = vtkWedge(tmp_path / "p_{step}", topo, field="betaVolume", every=2)
    assert cb(_state(topo, 3)) is None
    assert cb(_state(topo, 4)).name == "p_4.vtk"
    with pytest.raises(ValueError):
        vtkWedge(tmp_path / "p", topo, every=0)(_state(topo, 1))
    with pytest.raises(ValueError):
        vtkWedge(tmp_path / "bad", np.zeros(7), geometry=topo)
    with pytest.raises(TypeError):
        vtkWedge(tmp_path / "bad", np.zeros(topo.numberOfPrisms))
    with pytest.raises(ValueError):
        vtkWedge(tmp_path / "none", _state(topo, 1), geometry=topo, field="aseResult")


def test_run_level_writers_header_and_compare(tmp_path):
    topo = _topology()
    S = topo.numberOfPoints * topo.levels
    ase = np.linspace(1.0, 2.0, S)
    assert writePointsToVtk(topo, ase, tmp_path / "phiase.vtk", 123456, 10 ** 7, 0.005, True, 1.5) == 0
    lines = (tmp_path / "phiase.vtk").read_text(encoding="utf-8").splitlines()
    assert lines[0] == "# vtk DataFile Version 2.0"
    assert lines[1] == "RAYS=123456 MAXRAYS=10000000 REFLECTIONS=0 EXPECTATION=0.005 RUNTIME=1.5 REQUESTED_REFLECTIONS=1"
    assert f"POINT_DATA {S}" in lines and "SCALARS scalars float 1" in lines
    k = lines.index("LOOKUP_TABLE default")
    assert lines[k + 1] == "1.000000" and lines[k + S] == "2.000000"
    # relative difference in percent: |larger/smaller| - 1
    cmp = ase.copy()
    cmp[0], cmp[1], cmp[2] = 2.0, -ase[1], ase[2] / 4.0
    d = compareVtk(cmp, tmp_path / "phiase.vtk")
    assert d.shape == (S,)
    np.testing.assert_allclose(d[:3], [100.0, 0.0, 300.0], rtol=1e-5, atol=1e-4)   # six decimals
    assert np.all(np.abs(d[3:]) < 1e-3)            # six decimals in the file
    assert compareVtk(cmp, "").size == 0
    with pytest.raises(FileNotFoundError):
        compareVtk(cmp, tmp_path / "missing.vtk")
    prism = np.arange(topo.numberOfPrisms, dtype=np.float64)
    assert writePrismToVtk(topo, prism, tmp_path / "beta.vtk", 1, 2, 0.1, False, 0.25) == 0
    kind, values = readVtkScalars(tmp_path / "beta.vtk")["scalars"]
    assert kind == "CELL_DATA"
    np.testing.assert_array_equal(values, prism)
    with pytest.raises(ValueError):
        writePrismToVtk(topo, ase, tmp_path / "wrong.vtk", 1, 2, 0.1, False, 0.25)

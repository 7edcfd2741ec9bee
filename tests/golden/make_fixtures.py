"""Write the golden fixtures under tests/golden/ from the reference's data files.

Run once in a container that has the reference checkout at /root/reference:

    python tests/golden/make_fixtures.py [--reference /root/reference]

Everything written is data the reference's own tests hold (inputs and
expected outputs), converted to the prism layout this library consumes:

* compareSerial_{cuboid,cylindrical}.npz - legacy serial backward solver,
  1e6 rays per sample (tests/data/compareSerial/phiase_reference.npz,
  tests/data/compareSerial/README.md).  Geometry, beta and constants come
  from example/data/{cuboid,cylindrical}.vtk: each group of three Tet4 cells
  is one wedge split as [a,b,c,d],[b,e,c,d],[c,e,f,d]
  (utils/convert_vtk_topology.py:38-53) and is folded back to its prism.
* trueCladding.npz - legacy importance-sampled phiASE / mse / totalRays at
  2000 rays per sample for cladding absorption 0, 5.5 and 55 cm^-1
  (tests/data/laserPumpCladding/true_cladding_no_reflection_phiase_reference/).
* legacyPump3Wedge.npz - the six-step pumped legacy time series without
  reflections: phiASE at every wedge point after each step and the wedge
  integrals the archive's metadata records
  (tests/data/laserPumpCladding/fixed_legacy_no_reflection_pump3_wedge_reference/).
"""
from __future__ import annotations

import argparse
import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _tokens(path):
    with open(path, "r", encoding="ascii") as handle:
        for line in handle:
            for token in line.split():
                yield token


def parse_vtk(path):
    """Minimal ASCII legacy-VTK reader for the reference's Tet4 inputs."""
    it = _tokens(path)
    out = {"field": {}, "point_data": {}, "cell_data": {}}
    section = None
    for tok in it:
        if tok == "POINTS":
            n = int(next(it))
            next(it)
            out["points"] = np.array([float(next(it)) for _ in range(3 * n)]).reshape(n, 3)
        elif tok == "CELLS":
            n = int(next(it))
            next(it)
            cells = []
            for _ in range(n):
                k = int(next(it))
                cells.append([int(next(it)) for _ in range(k)])
            out["cells"] = np.array(cells, dtype=np.int64)
        elif tok == "CELL_TYPES":
            n = int(next(it))
            out["cell_types"] = np.array([int(next(it)) for _ in range(n)])
        elif tok == "FIELD":
            next(it)
            count = int(next(it))
            for _ in range(count):
                name = next(it)
                comps = int(next(it))
                n = int(next(it))
                next(it)
                out["field"][name] = np.array([float(next(it)) for _ in range(comps * n)])
        elif tok in ("POINT_DATA", "CELL_DATA"):
            section = "point_data" if tok == "POINT_DATA" else "cell_data"
            next(it)
        elif tok == "SCALARS":
            name = next(it)
            next(it)
            comps = int(next(it))
            assert next(it) == "LOOKUP_TABLE"
            next(it)
            n = len(out["points"]) if section == "point_data" else len(out["cells"])
            out[section][name] = np.array([float(next(it)) for _ in range(comps * n)])
    return out


def wedges_from_tets(points, tets, levels):
    """Fold consecutive Tet4 triples back into wedges; return base triangles
    (T, 3) as transverse point ids and the wedge level of every wedge.  The
    three Tets of a wedge together use exactly its six vertices, three on
    level l and their copies on level l + 1, whatever diagonal split the file
    uses."""
    np_per_level = points.shape[0] // levels
    groups = tets.reshape(-1, 3, 4)
    tri, lev = [], []
    for g in groups:
        verts = np.unique(g.reshape(-1))
        if verts.size != 6:
            raise ValueError("a Tet triple does not span exactly six wedge vertices")
        level = int(verts.min() // np_per_level)
        base = verts[verts // np_per_level == level]
        top = verts[verts // np_per_level == level + 1]
        if base.size != 3 or top.size != 3 or not np.all(np.sort(top - np_per_level) == base):
            raise ValueError("wedge top is not the next level of its base")
        tri.append(base - level * np_per_level)
        lev.append(level)
    return np.array(tri), np.array(lev)


def prism_fixture_from_vtk(path):
    vtk = parse_vtk(path)
    f = vtk["field"]
    levels = int(f["structuredNumberOfLevels"][0])
    npts = int(f["structuredNumberOfPoints"][0])
    thickness = float(f["structuredThickness"][0])
    points = vtk["points"]
    assert points.shape[0] == npts * levels
    for lv in range(levels):
        sl = points[lv * npts:(lv + 1) * npts]
        assert np.allclose(sl[:, :2], points[:npts, :2])
        assert np.allclose(sl[:, 2], lv * thickness)
    tri, lev = wedges_from_tets(points, vtk["cells"], levels)
    T = tri.shape[0] // (levels - 1)
    base = tri[:T]
    for lv in range(levels - 1):
        assert np.all(tri[lv * T:(lv + 1) * T] == base)
        assert np.all(lev[lv * T:(lv + 1) * T] == lv)
    beta_tet = vtk["cell_data"]["betaVolume"].reshape(-1, 3)
    assert np.all(beta_tet == beta_tet[:, :1])
    clad_tet = f["claddingCellTypes"].reshape(-1, 3)
    refl_tet = f["reflectivities"].reshape(-1, 3, 2)
    return {
        "points": points[:npts, :2].copy(),
        "triangles": base.astype(np.uint32),
        "levels": np.int64(levels),
        "thickness": np.float64(thickness),
        "betaVolume": beta_tet[:, 0].copy(),
        "betaCells": vtk["point_data"]["betaCells"].copy(),
        "claddingCellTypes": clad_tet[:T, 0].astype(np.uint32),
        "claddingNumber": np.int64(f["claddingNumber"][0]),
        "claddingAbsorption": np.float64(f["claddingAbsorption"][0]),
        "nTot": np.float64(f["nTot"][0]),
        "crystalTFluo": np.float64(f["crystalTFluo"][0]),
        "refractiveIndices": f["refractiveIndices"].astype(np.float64),
        "reflectivitiesMax": np.float64(np.abs(refl_tet).max()),
    }


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__)
    parser.add_argument("--reference", type=Path, default=Path("/root/reference"))
    args = parser.parse_args(argv)
    ref = args.reference

    serial = np.load(ref / "tests/data/compareSerial/phiase_reference.npz")
    meta = json.loads(str(serial["metadata"]))
    for name in ("cuboid", "cylindrical"):
        fx = prism_fixture_from_vtk(ref / f"example/data/{name}.vtk")
        spectra = meta["spectra"][name]
        fx["sigmaA"] = np.float64(spectra["sigmaA"][0])
        fx["sigmaE"] = np.float64(spectra["sigmaE"][0])
        fx["phiASE"] = serial[f"{name}_phiASE"]
        fx["dndtAse"] = serial[f"{name}_dndtAse"]
        fx["raysPerSample"] = np.int64(meta["parameters"]["experiment"]["minRays"])
        np.savez_compressed(HERE / f"compareSerial_{name}.npz", **fx)
        print(f"wrote compareSerial_{name}.npz: T={fx['triangles'].shape[0]} L={int(fx['levels'])}")

    clad_dir = ref / "tests/data/laserPumpCladding/true_cladding_no_reflection_phiase_reference"
    d = np.load(clad_dir / "phiase_reference.npz")
    m = json.loads(str(d["metadata"]))
    pts = d["points"]
    levels = int(m["geometry"]["numberOfLevels"])
    npts = int(m["geometry"]["numberOfBasePoints"])
    T = int(m["geometry"]["numberOfBaseTriangles"])
    cells = d["cells"].astype(np.int64)
    base = cells[:T, :3]
    for lv in range(levels - 1):
        assert np.all(cells[lv * T:(lv + 1) * T, :3] == base + lv * npts)
        assert np.all(cells[lv * T:(lv + 1) * T, 3:] == base + (lv + 1) * npts)
    thickness = float(pts[:, 2].max()) / (levels - 1)
    assert np.allclose(pts[:, 2].reshape(levels, npts), (np.arange(levels) * thickness)[:, None])
    np.savez_compressed(
        HERE / "trueCladding.npz",
        points=pts[:npts, :2].copy(),
        triangles=base.astype(np.uint32),
        levels=np.int64(levels),
        thickness=np.float64(thickness),
        betaVolume=d["wedgeBetaVolume"],
        claddingCellTypes=d["baseCladdingCellTypes"].astype(np.uint32),
        claddingNumber=np.int64(m["material"]["claddingNumber"]),
        claddingAbsorptions=np.asarray(m["material"]["claddingAbsorptions"], dtype=np.float64),
        sigmaA=np.float64(m["material"]["crossSectionAbsorption"]),
        sigmaE=np.float64(m["material"]["crossSectionEmission"]),
        betaCells=np.float64(m["material"]["betaCells"]),
        # nTot and tau are not stored in the archive; these are the values of
        # example/laserPumpCladdingApi.py:240-241 that produced it
        nTot=np.float64(2 * 1.388e20),
        crystalTFluo=np.float64(9.41e-4),
        phiASE=d["phiASE"],
        mse=d["mse"],
        totalRays=d["totalRays"],
        raysPerSample=np.int64(m["parameters"]["raysPerSample"]),
        rngSeed=np.int64(m["parameters"]["rngSeed"]),
    )
    print(f"wrote trueCladding.npz: T={T} L={levels}")

    wedge_dir = ref / "tests/data/laserPumpCladding/fixed_legacy_no_reflection_pump3_wedge_reference"
    d = np.load(wedge_dir / "phiase_reference.npz")
    m = json.loads(str(d["metadata"]))
    pts = d["points"]
    cells = d["cells"].astype(np.int64)
    levels = int(np.unique(pts[:, 2]).size)
    npts = pts.shape[0] // levels
    T = cells.shape[0] // (levels - 1)
    base = cells[:T, :3]
    for lv in range(levels - 1):
        assert np.all(cells[lv * T:(lv + 1) * T, :3] == base + lv * npts)
        assert np.all(cells[lv * T:(lv + 1) * T, 3:] == base + (lv + 1) * npts)
    thickness = float(m["parameters"]["thickness"])
    assert np.allclose(pts[:, 2].reshape(levels, npts), (np.arange(levels) * thickness)[:, None])
    np.savez_compressed(
        HERE / "legacyPump3Wedge.npz",
        points=pts[:npts, :2].copy(),
        triangles=base.astype(np.uint32),
        levels=np.int64(levels),
        thickness=np.float64(thickness),
        phiASE=d["phiASE"],
        integrals=np.asarray(m["integrals"], dtype=np.float64),
        meanPhi=np.asarray(m["meanPhi"], dtype=np.float64),
        totalWedgeVolume=np.float64(m["totalWedgeVolume"]),
        timeStepSeconds=np.float64(m["parameters"]["timeStepSeconds"]),
    )
    print(f"wrote legacyPump3Wedge.npz: T={T} L={levels}")


if __name__ == "__main__":
    main()

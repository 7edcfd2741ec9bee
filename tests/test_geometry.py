"""Mesh builder, gain medium, spectra and front-end checks (CPU)."""
import numpy as np
import pytest

from haseongpu_amd import CrossSectionData, GainMedium, Grid, MeshTopology, PhiASE
from haseongpu_amd.geometry import bowyerWatson
from haseongpu_amd.laser import interpolateLinear
from haseongpu_amd.workloads import combined_disk, depletion_field, disk_points, plain_small, production_disk


def _check_topology(topo, area):
    d = topo.topology()
    T = topo.numberOfTriangles
    tri = topo.trianglePointIndices.astype(np.int64)
    nb = d["triangleNeighbors"].reshape(3, T)
    fe = d["forbiddenEdge"].reshape(3, T)
    P = topo.points
    # neighbour symmetry and forbidden-edge reciprocity (core.py:733-774)
    for t in range(T):
        for k in range(3):
            n = nb[k, t]
            if n < 0:
                assert fe[k, t] == -1
                continue
            assert nb[fe[k, t], n] == t and fe[fe[k, t], n] == k
            # the shared edge is the same vertex pair, traversed the other way
            assert {tri[t, k], tri[t, (k + 1) % 3]} == {tri[n, fe[k, t]], tri[n, (fe[k, t] + 1) % 3]}
    # unit outward normals from the normal point (vertex k)
    nx = d["triangleNormalsX"].reshape(3, T)
    ny = d["triangleNormalsY"].reshape(3, T)
    npnt = d["triangleNormalPoint"].reshape(3, T)
    assert np.allclose(nx ** 2 + ny ** 2, 1.0)
    cx, cy = d["triangleCenterX"], d["triangleCenterY"]
    for k in range(3):
        assert np.all(npnt[k] == tri[:, k])
        to_center = np.stack([cx - P[npnt[k], 0], cy - P[npnt[k], 1]])
        assert np.all(nx[k] * to_center[0] + ny[k] * to_center[1] < 0)
    # counter-clockwise, areas sum to the domain
    a, b, c = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    assert np.all(cross > 0)
    assert d["triangleSurfaces"].astype(np.float64).sum() == pytest.approx(area, rel=1e-6)


def test_grid_mesh_small_plain_slab():
    w = plain_small()
    topo = w.topology
    assert (topo.numberOfPoints, topo.numberOfTriangles, topo.levels, topo.numberOfPrisms) == (121, 200, 10, 1800)
    assert topo.thickness == pytest.approx(0.7 / 9)
    _check_topology(topo, 1.0)
    assert np.sum(topo.topology()["triangleNeighbors"] < 0) == 40   # perimeter edges


def test_delaunay_disk_is_delaunay_and_consistent():
    pts = disk_points(3.0, 600, seed=4)
    topo = MeshTopology.fromPoints(pts, numberOfLevels=3, thickness=0.1)
    hull_area = np.pi * 9.0 * (1 - (np.pi ** 2) / (6 * 90 ** 2))  # polygon inscribed in the circle
    _check_topology(topo, topo.topology()["triangleSurfaces"].astype(np.float64).sum())
    assert topo.topology()["triangleSurfaces"].sum() == pytest.approx(np.pi * 9.0, rel=0.01)
    # empty circumcircle: no point strictly inside any triangle's circumcircle
    tri = topo.trianglePointIndices.astype(np.int64)
    P = topo.points
    a, b, c = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]
    d = 2 * (a[:, 0] * (b[:, 1] - c[:, 1]) + b[:, 0] * (c[:, 1] - a[:, 1]) + c[:, 0] * (a[:, 1] - b[:, 1]))
    ux = ((a ** 2).sum(1) * (b[:, 1] - c[:, 1]) + (b ** 2).sum(1) * (c[:, 1] - a[:, 1]) + (c ** 2).sum(1) * (a[:, 1] - b[:, 1])) / d
    uy = ((a ** 2).sum(1) * (c[:, 0] - b[:, 0]) + (b ** 2).sum(1) * (a[:, 0] - c[:, 0]) + (c ** 2).sum(1) * (b[:, 0] - a[:, 0])) / d
    r2 = (a[:, 0] - ux) ** 2 + (a[:, 1] - uy) ** 2
    dist2 = (P[None, :, 0] - ux[:, None]) ** 2 + (P[None, :, 1] - uy[:, None]) ** 2
    assert np.all(dist2 >= r2[:, None] * (1 - 1e-9))
    assert hull_area > 0


def test_insertion_order_of_a_fan_derived_by_hand():
    """Points 0..2 form one triangle and point 3 lies inside it.  Before 3 is
    inserted the only triangle without an enclosing vertex is the one made
    when 2 arrived, on the sorted edge (0, 1), already counter-clockwise:
    (0, 1, 2).  Point 3 lies in its circumcircle (centre (2, 2), r^2 = 8) and
    in no circle of a triangle through a far enclosing vertex, so the cavity
    is (0, 1, 2) alone; its edges in rotation order (0, 1), (1, 2), (2, 0)
    become (0, 1, 3), (1, 2, 3) and (0, 2, 3) -> swapped to (2, 0, 3), the
    last being clockwise as written (core.py:583-627)."""
    pts, tri = bowyerWatson([[0.0, 0.0], [4.0, 0.0], [0.0, 4.0], [1.0, 1.0]])
    assert tri.tolist() == [[0, 1, 3], [1, 2, 3], [2, 0, 3]]
    # a repeated point is dropped where it repeats, not where it first appears
    pts2, tri2 = bowyerWatson([[0.0, 0.0], [4.0, 0.0], [0.0, 4.0], [4.0, 0.0], [1.0, 1.0]])
    np.testing.assert_array_equal(pts2, pts)
    np.testing.assert_array_equal(tri2, tri)
    with pytest.raises(ValueError):
        bowyerWatson([[0.0, 0.0], [1.0, 1.0], [2.0, 2.0]])


def test_insertion_makes_the_last_point_the_third_vertex():
    """Unit square with an interior point inserted last: all four triangles
    are made by that insertion, so each ends with point 4 and starts with
    its rim edge in ascending order unless that order is clockwise:
    (0,1,4) and (1,3,4) stay, (2,3,4) -> (3,2,4), (0,2,4) -> (2,0,4)."""
    pts, tri = bowyerWatson([[0, 0], [1, 0], [0, 1], [1, 1], [0.4, 0.6]])
    assert sorted(tri.tolist()) == [[0, 1, 4], [1, 3, 4], [2, 0, 4], [3, 2, 4]]
    # triangles are appended when made: the list runs in order of their last vertex
    topo = MeshTopology.fromPoints(disk_points(3.0, 400, seed=7), numberOfLevels=2, thickness=0.1)
    t = topo.trianglePointIndices.astype(np.int64)
    assert np.all(t[:, 2] == t.max(axis=1))
    assert np.all(np.diff(t[:, 2]) >= 0)
    # the rim edge keeps ascending order exactly when that is counter-clockwise
    P = topo.points
    a, b, c = P[t[:, 0]], P[t[:, 1]], P[t[:, 2]]
    lo, hi = np.minimum(t[:, 0], t[:, 1]), np.maximum(t[:, 0], t[:, 1])
    A, B = P[lo], P[hi]
    ascending_ccw = (B[:, 0] - A[:, 0]) * (c[:, 1] - A[:, 1]) - (B[:, 1] - A[:, 1]) * (c[:, 0] - A[:, 0]) >= 0
    assert np.array_equal(t[:, 0] < t[:, 1], ascending_ccw)


def test_production_disk_workload():
    w = production_disk()
    topo = w.topology
    assert 3800 <= topo.numberOfTriangles <= 4300
    assert topo.levels == 20 and topo.numberOfSamples > 38000
    beta = w.medium.value("betaVolume").reshape(topo.levels - 1, -1)
    clad = topo.boundaryTriangles()
    assert np.all(beta[:, clad] == 0.0)
    assert 0.0 <= beta.min() and beta.max() <= 1.0
    sa, se, msa, mse = w.spectra.spectralPairs(False)
    assert sa.size == 64 and 0.9e-21 < msa < 1.2e-21 and 2.4e-20 < mse < 2.7e-20


def test_combined_disk_workload_and_depletion_field():
    w = combined_disk()
    topo = w.topology
    assert topo.numberOfTriangles == 20682 and topo.levels == 30
    assert topo.numberOfPrisms == 599778 and topo.numberOfSamples == 315300
    r = np.hypot(topo.points[:, 0], topo.points[:, 1])
    assert abs(r.max() - 5.0) < 1e-9 and abs(topo.thickness * (topo.levels - 1) - 1.0) < 1e-12
    sa, se, msa, mse = w.spectra.spectralPairs(False)
    assert sa.size == 128
    s = w.settings
    assert (s["minRaysPerSample"], s["maxRaysPerSample"], s["mseThreshold"]) == (100_000, 100_000_000, 0.005)
    assert s["useReflections"] and float(w.medium.value("claddingAbsorption")) == 5.5
    beta = w.medium.value("betaVolume").reshape(topo.levels - 1, -1)
    assert np.all(beta[:, topo.boundaryTriangles()] == 0.0)
    f = depletion_field(topo)
    # seeded, smooth, a depletion (factors in (0, 1]) strongest near the spots and weaker deeper in z
    np.testing.assert_array_equal(f, depletion_field(topo))
    assert not np.array_equal(f, depletion_field(topo, seed=6))
    assert f.shape == (topo.numberOfPrisms,) and np.all((f > 0.95) & (f <= 1.0)) and f.min() < 0.995
    layers = f.reshape(topo.levels - 1, -1)
    assert np.all(layers[0] <= layers[-1] + 1e-15)


def test_gain_medium_layouts():
    topo = MeshTopology.fromGrid(Grid(x=1, y=1, z=0.5, tileSizeX=0.5, tileSizeZ=0.25))
    med = GainMedium(topo)
    T, L = topo.numberOfTriangles, topo.levels
    assert med.get("betaVolume").expectedShape == (T, L - 1)
    m = np.arange(T * (L - 1), dtype=np.float64).reshape(T, L - 1)
    med.set("betaVolume", m)
    # matrix input is stored Fortran-ordered: prism j = t + l * T
    assert np.array_equal(med.value("betaVolume"), m.reshape(-1, order="F"))
    with pytest.raises(ValueError):
        med.set("betaVolume", np.zeros(3))
    with pytest.raises(ValueError):
        med.set("betaVolume", np.full(T * (L - 1), np.nan))
    with pytest.raises(KeyError):
        med.get("unknownProperty")
    med.set("beta_volume", np.zeros(T * (L - 1)))   # alias
    assert med.value("claddingNumber") == 1
    assert np.array_equal(med.value("refractiveIndices"), [1, 1, 1, 1])


def test_interpolate_linear_exact_for_linear_functions():
    """tests/interpolation.cpp: exact for linear data, end points pinned."""
    rng = np.random.default_rng(0)
    x = np.sort(rng.uniform(-1000, 1000, 100))
    y = 3.5 * x - 7.0
    yi = interpolateLinear(y, x, 1000)
    xi = x[0] + np.arange(1000) * (x[-1] - x[0]) / 999
    np.testing.assert_allclose(yi, 3.5 * xi - 7.0, rtol=1e-10, atol=1e-9)
    assert yi[-1] == y[-1] and yi[0] == pytest.approx(y[0])
    assert np.all(interpolateLinear([2.0], [5.0], 4) == 2.0)
    with pytest.raises(ValueError):
        interpolateLinear(y, x, 10)


def test_cross_sections_pairs():
    cs = CrossSectionData([900, 1000, 1100], [1e-21, 3e-21, 2e-21], [900, 1000, 1100], [1e-20, 2e-20, 5e-20])
    sa, se, msa, mse = cs.spectralPairs(monochromatic=True)
    assert sa.tolist() == [1e-21] and se.tolist() == [1e-20] and (msa, mse) == (1e-21, 1e-20)
    sa, se, msa, mse = cs.spectralPairs()
    assert sa.size == 3 and msa == 3e-21 and mse == 5e-20
    cs.resolution = 5
    sa, se, _, _ = cs.spectralPairs()
    assert sa.size == 5 and sa[2] == pytest.approx(3e-21)
    with pytest.raises(ValueError):
        CrossSectionData([1, 2], [1e-21], [1], [1e-20])


def test_phiase_config_accepts_legacy_keys(tmp_path):
    cfg = tmp_path / "phiASE.yaml"
    cfg.write_text("experiment:\n  minRaysPerSample: 2000\n  maxRaysPerSample: 1000000\n"
                   "  mseThreshold: 0.01087\n  adaptiveSteps: 8\n  useReflections: true\n"
                   "compute:\n  repetitions: 3\n  rngSeed: 5489\n")
    p = PhiASE(str(cfg))
    assert (p.minRaysPerSample, p.maxRaysPerSample, p.mseThreshold, p.adaptiveSteps, p.useReflections,
            p.repetitions, p.rngSeed) == (2000, 1000000, 0.01087, 8, True, 3, 5489)
    with pytest.raises(ValueError, match="maxRays"):
        PhiASE({"minRays": 10, "maxRays": 5})
    w = plain_small()
    with pytest.raises(RuntimeError):
        PhiASE().getResults()
    with pytest.raises(ValueError, match="crossSections"):
        PhiASE().run(w.medium)


def test_phiase_takes_the_present_keyword_surface():
    """The 2.2.0 constructor keywords (pyInclude/simulation.py:130-175):
    minRays / maxRays are each sample point's ray counts (the reference maps
    minRaysPerSample onto minRays, simulation.py:318-323), a relative
    standard error threshold r stops a sample at mse = RSE^2 <= r^2, and the
    settings this one-GPU path cannot honour are refused by name."""
    w = plain_small()
    p = PhiASE(crossSections=w.spectra, minRays=1000, maxRays=2000, relativeStandardErrorThreshold=0.1)
    c = p.controls()
    assert (c.min_rays, c.max_rays, c.mse_threshold) == (1000, 2000, pytest.approx(0.01))
    mesh, exp, comp = p.nativeArguments(w.medium)
    e = exp.struct
    assert (e.min_rays_per_sample, e.max_rays_per_sample) == (1000, 2000)
    assert e.mse_threshold == pytest.approx(0.01, rel=1e-15)
    # the legacy names mean the same; both at once must agree
    legacy = PhiASE(crossSections=w.spectra, minRaysPerSample=1000, maxRaysPerSample=2000, mseThreshold=0.01)
    lc = legacy.controls()
    assert (lc.min_rays, lc.max_rays, lc.adaptive_steps) == (c.min_rays, c.max_rays, c.adaptive_steps)
    assert lc.mse_threshold == pytest.approx(c.mse_threshold, rel=1e-15)
    assert PhiASE(minRays=1000, minRaysPerSample=1000).controls().min_rays == 1000
    with pytest.raises(ValueError, match="contradicts"):
        PhiASE(minRays=1000, minRaysPerSample=2000)
    with pytest.raises(ValueError, match="contradicts"):
        PhiASE(mseThreshold=0.005, relativeStandardErrorThreshold=0.1)
    # a fixed ray count disables escalation
    c = PhiASE(forwardRayCount=4096, minRays=100, maxRays=10 ** 6).controls()
    assert (c.min_rays, c.max_rays) == (4096, 4096)
    # YAML and mapping configs under the present snake-case keys
    c = PhiASE({"phi_ase": {"min_rays": 500, "max_rays": 5000, "relative_standard_error_threshold": 0.05,
                            "adaptive_steps": 2, "devices": [0]}}).controls()
    assert (c.min_rays, c.max_rays, c.adaptive_steps, c.device) == (500, 5000, 2, 0)
    assert c.mse_threshold == pytest.approx(0.0025)
    for bad in (dict(propagationMode="forward"), dict(parallelMode="mpi"), dict(numDevices=2),
                dict(devices=[0, 1]), dict(backend="Host_Cpu_CpuOmpBlocks")):
        with pytest.raises(ValueError):
            PhiASE(**bad)
    assert PhiASE(propagationMode="backward", backend="Hip_AmdGpu_GpuHip").controls().min_rays == 100000
    # defaults when nothing is given: 1e5 rays per sample, mse 0.005
    c = PhiASE().controls()
    assert (c.min_rays, c.max_rays, c.mse_threshold) == (100000, 100000, 0.005)

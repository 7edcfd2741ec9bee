"""Pins of the forward Tet4 oracle (oracle/forward_oracle.c), on the CPU.

* Traversal known answers of the reference (tests/forwardRayTraversal.cpp:
  156-255): a ray through a vertex shared by 30 tetrahedra ties three faces,
  crosses at the vertex and enters cell 23; a ray crossing a plane at
  perpendicular incidence exits face 0 after 0.4, enters cell 1 through its
  face 0 and has 0.6 to go.
* Mesh set-up: face planes are the barycentric coordinates (1 at the opposite
  vertex, 0 on the face), neighbours are symmetric, volumes add up to the box.
* Walk: the scores a ray leaves in the vertices add up to the closed forms
  of its track -- the ray-box exit distance for zero gain, expm1(g L)/g for a
  uniform gain g -- for rays from random interior points.
* Sampling: folded-cube points are uniform in the tetrahedron (barycentric
  coordinate moments and tail of Beta(1, 3)), directions are unit and
  isotropic, beta*V stratification puts each cell within +-8 rays of its
  share.
* Estimator: with uniform beta and gain in a box, the cells around the
  centre carry the expectation computed from the angular quadrature
  (tests/quadrature.py: the point flux averaged by the vertex hat functions)
  within 4 sigma; the reported standard error matches the spread over seeds.
"""
import math

import numpy as np
import pytest

from haseongpu_amd import Grid, MeshTopology
from haseongpu_amd.tet_mesh import TetMesh
from oracle import forward as F
from tests.quadrature import box_point_flux, hat_smoothed_cell_flux

SA, SE = 1.16e-21, 2.48e-20
NTOT, TAU = 2.76e20, 9.5e-4


def _box(nx=4, levels=5, z=0.5, beta=0.2, seed=None):
    topo = MeshTopology.fromGrid(Grid(xExtent=1.0, yExtent=1.0, zExtent=z, tileSizeX=1.0 / nx,
                                      tileSizeZ=z / (levels - 1)))
    P = topo.numberOfPrisms
    b = np.full(P, beta) if seed is None else np.random.default_rng(seed).uniform(0.0, 0.3, P)
    return TetMesh.fromPrisms(topo, b, n_tot=NTOT, tau=TAU), topo


def _fan():
    pts = [[0, 0, 0], [0, 0, 1], [0, 0, -1]]
    for k in range(15):
        a = 2 * math.pi * k / 15
        pts.append([math.cos(a), math.sin(a), 0.0])
    cells = [[0, 1, 3 + c, 3 + (c + 1) % 15] for c in range(15)] + [[0, 2, 3 + (c + 1) % 15, 3 + c] for c in range(15)]
    return np.array(pts), cells


def test_ray_through_a_vertex_shared_by_30_tetrahedra():
    pts, cells = _fan()
    o = F.ForwardOracle(TetMesh(points=pts, cells=cells, beta=np.full(30, 0.1)))
    assert np.count_nonzero(np.asarray(cells) == 0) == 30
    origin = pts[1] * 0.2 + pts[3] * 0.1 + pts[4] * 0.3
    u = -origin / np.linalg.norm(origin)
    h = o.next_face(0, origin, u)
    assert h.tied_mask == (1 << 1) | (1 << 2) | (1 << 3)
    p = origin + u * h.length
    np.testing.assert_allclose(p, 0.0, atol=1e-14)
    t = o.cross(0, h, p, u)
    assert t.status == F.ENTERED and t.cell == 23
    h2 = o.next_face(t.cell, p, u, t.forbidden)
    assert h2.face >= 0 and h2.length > 0.0


def test_ray_crossing_a_plane_at_perpendicular_incidence():
    pts = np.array([[-1.0, 0, 0], [0, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 0]])
    o = F.ForwardOracle(TetMesh(points=pts, cells=[[0, 1, 2, 3], [4, 1, 3, 2]], beta=[0.1, 0.1]))
    u = np.array([1.0, 0, 0])
    assert np.dot(u, pts[2] - pts[1]) == 0.0 and np.dot(u, pts[3] - pts[1]) == 0.0
    origin = np.array([-0.4, 0.2, 0.2])
    h = o.next_face(0, origin, u)
    assert h.face == 0 and h.tied_mask == 1 and h.length == pytest.approx(0.4)
    p = origin + u * h.length
    np.testing.assert_allclose(p, [0.0, 0.2, 0.2], atol=1e-14)
    t = o.cross(0, h, p, u)
    assert t.status == F.ENTERED and t.cell == 1 and t.forbidden == 0
    h2 = o.next_face(t.cell, p, u, t.forbidden)
    assert h2.face >= 0 and h2.length == pytest.approx(0.6)


def test_mesh_setup_planes_neighbours_volumes():
    mesh, topo = _box(nx=3, levels=4, z=0.6)
    o = F.ForwardOracle(mesh)
    C = mesh.n_cells
    np.testing.assert_allclose(o.volume.sum(), 1.0 * 1.0 * 0.6, rtol=1e-12)
    np.testing.assert_allclose(o.volume, mesh.volumes(), rtol=1e-12)
    planes = o.planes.reshape(C, 4, 4)
    verts = mesh.points[mesh.cells]                       # (C, 4, 3)
    lam = np.einsum("ckx,cjx->ckj", planes[:, :, :3], verts) + planes[:, :, 3][:, :, None]
    np.testing.assert_allclose(lam, np.broadcast_to(np.eye(4), lam.shape), atol=1e-12)
    nbr, nf = o.nbr.reshape(C, 4), o.nbr_face.reshape(C, 4)
    inner = nbr >= 0
    cc, kk = np.nonzero(inner)
    assert np.all(nbr[nbr[cc, kk], nf[cc, kk]] == cc)
    # boundary faces: two per bottom/top triangle, two per lateral quad per level
    T, L = topo.numberOfTriangles, topo.levels
    lateral_edges = int(np.sum(topo.topology()["triangleNeighbors"] < 0))
    assert np.count_nonzero(~inner) == 2 * T + 2 * lateral_edges * (L - 1)
    np.testing.assert_allclose(o.lumped[:mesh.n_points].sum(), o.volume.sum(), rtol=1e-12)


def _box_exit(p, u, hi):
    t = np.inf
    for a in range(3):
        if u[a] > 0:
            t = min(t, (hi[a] - p[a]) / u[a])
        elif u[a] < 0:
            t = min(t, -p[a] / u[a])
    return t


@pytest.mark.parametrize("beta", [SA / (SA + SE), 0.3])
def test_walk_scores_add_up_to_the_track_integral(beta):
    mesh, _ = _box(nx=5, levels=6, z=0.5, beta=beta)
    o = F.ForwardOracle(mesh)
    g = NTOT * (beta * (SE + SA) - SA)
    rng = np.random.default_rng(7)
    for trial in range(60):
        cell = int(rng.integers(mesh.n_cells))
        w = rng.dirichlet(np.ones(4))
        p = w @ mesh.points[mesh.cells[cell]]
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        scores, visits, dropped, seg, status = o.walk(cell, p, u, SA, SE)
        assert status == 0 and dropped.sum() == 0 and seg == visits.sum() > 0
        length = _box_exit(p, u, (1.0, 1.0, 0.5))
        expect = length if g == 0.0 else math.expm1(g * length) / g
        assert scores.sum() == pytest.approx(expect, rel=1e-11), (trial, g)
        assert scores[mesh.n_points:].sum() == 0.0     # no cladding


def test_folded_cube_points_are_uniform_in_the_tetrahedron():
    U = np.random.default_rng(4).random((150_000, 3))
    B = np.array([F.point_in_tet(u) for u in U])
    assert np.all(B >= -1e-15) and np.allclose(B.sum(1), 1.0, atol=1e-15)
    # barycentric coordinates of a uniform point are Beta(1, 3): mean 1/4,
    # second moment 1/10, P(lambda > 1/2) = 1/8 (150k points: 4 sigma bounds)
    np.testing.assert_allclose(B.mean(0), 0.25, atol=2.2e-3)
    np.testing.assert_allclose((B ** 2).mean(0), 0.1, atol=1.6e-3)
    np.testing.assert_allclose((B > 0.5).mean(0), 0.125, atol=3.5e-3)


def test_directions_unit_and_isotropic():
    U = np.array([F.direction(11, 0, r)[0] for r in range(20000)])
    np.testing.assert_allclose(np.linalg.norm(U, axis=1), 1.0, atol=1e-15)
    np.testing.assert_allclose(U.mean(0), 0.0, atol=0.02)
    np.testing.assert_allclose((U ** 2).mean(0), 1.0 / 3.0, atol=0.01)
    # the z component is uniform on [-1, 1] (Archimedes)
    hist, _ = np.histogram(U[:, 2], bins=10, range=(-1, 1))
    assert np.all(np.abs(hist / len(U) - 0.1) < 0.012)


def test_stratified_sources_follow_beta_volume():
    mesh, _ = _box(nx=3, levels=3, z=0.4, seed=5)
    o = F.ForwardOracle(mesh)
    rays = 8 * 3001
    cells = np.array([o.ray_start(9, 2, rays, r, 16).cell for r in range(rays)])
    counts = np.bincount(cells, minlength=mesh.n_cells)
    share = rays * o.beta * o.volume / np.sum(o.beta * o.volume)
    assert np.all(np.abs(counts - share) <= 8.0 + 1e-9)
    firsts = [F.batch_range(rays + 5, b) for b in range(8)]
    assert [c for _, c in firsts] == [3002] * 5 + [3001] * 3
    assert firsts[0][0] == 0 and all(firsts[b + 1][0] == firsts[b][0] + firsts[b][1] for b in range(7))


@pytest.mark.parametrize("beta", [SA / (SA + SE), 0.25])
def test_centre_cells_carry_the_quadrature_flux(beta):
    """Zero gain (beta = sigma_a / (sigma_a + sigma_e)) and 1.47 / cm gain:
    the centre cells' flux equals its expectation from the angular
    quadrature, the point flux averaged by the vertex hat functions (1%
    below the point values at the centre, where the flux peaks)."""
    nx, levels, z = 10, 11, 0.7
    mesh, _ = _box(nx=nx, levels=levels, z=z, beta=beta)
    o = F.ForwardOracle(mesh)
    g = NTOT * (beta * (SE + SA) - SA)
    res = o.evaluate([SA], [SE], min_rays=400_000, max_rays=400_000, steps=0, seed=3)
    cen = mesh.centroids()
    near = np.nonzero(np.linalg.norm((cen - [0.5, 0.5, z / 2]) / [1, 1, z], axis=1) < 0.12)[0]
    assert near.size >= 12

    def flux(p):
        return box_point_flux(p, (0.0, 1.0, 0.0, 1.0), z, beta, g, NTOT / TAU, n_mu=200, n_phi=400)

    expect = hat_smoothed_cell_flux(mesh.points, mesh.cells, o.volume, near, flux)
    point = np.array([flux(c) for c in cen[near]])
    phi = res["phiAse"][near].astype(np.float64)
    se = res["standardError"][near]
    assert 0.985 < expect.sum() / point.sum() < 0.995
    assert abs(phi.sum() / expect.sum() - 1.0) < 0.006, phi.sum() / expect.sum()
    assert np.all(np.abs(phi - expect) <= 4.0 * se + 0.002 * expect)


def test_standard_error_matches_the_spread_over_seeds():
    mesh, _ = _box(nx=4, levels=4, z=0.4, beta=0.2)
    o = F.ForwardOracle(mesh)
    runs = [o.evaluate([SA], [SE], min_rays=20_000, max_rays=20_000, steps=0, seed=s) for s in range(12)]
    phi = np.array([r["phiAse"] for r in runs], dtype=np.float64)
    se = np.array([r["standardError"] for r in runs])
    ratio = phi.std(axis=0, ddof=1).mean() / se.mean()
    assert 0.75 < ratio < 1.3, ratio


def test_adaptive_loop_targets_and_convergence_record():
    mesh, _ = _box(nx=3, levels=3, z=0.3, beta=0.2)
    o = F.ForwardOracle(mesh)
    res = o.evaluate([SA], [SE], min_rays=1000, max_rays=64_000, steps=3, threshold=0.02 ** 2, seed=1)
    targets = [1000, 4000, 16000, 64000]
    assert res["rays"] in targets and res["launches"] == targets.index(res["rays"]) + 1
    conv = res["convergenceRays"]
    assert set(np.unique(conv)) <= set(targets) | {0}
    rse = res["relativeStandardError"]
    ok = np.isfinite(rse) & (rse ** 2 <= 0.02 ** 2)
    assert np.all(conv[ok] > 0)
    if res["rays"] < 64000:
        assert np.all(ok)
    # a fixed ray count (min == max) is a single launch
    one = o.evaluate([SA], [SE], min_rays=3000, max_rays=3000, steps=4, seed=1)
    assert one["rays"] == 3000 and one["launches"] == 1


# --------------------------------------------------------------- reflections
def _slab(nx=8, levels=9, z=0.5, beta=SA / (SA + SE)):
    mesh, topo = _box(nx=nx, levels=levels, z=z, beta=beta)
    return mesh, mesh.faceSurfacesByPlane({0.0: 1, z: 2})


def test_zero_reflectivity_reservoir_equals_plain():
    mesh, fs = _slab()
    o = F.ForwardOracle(mesh)
    refl = F.reflecting(o, fs, [0.0, 0.0, 0.0])
    a = F.trace_srm(o, refl, [SA], [SE], 3, 0, 40_000)
    b = o.trace([SA], [SE], 3, 0, 40_000, threads=1)
    assert refl.n_faces == 2 * 2 * 8 * 8              # two triangles per grid cell, top and bottom
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[0], b[0])
    assert refl.status()[0] == [1] * 8 and refl.status()[1] == [0] * 8   # converged without a pass


def test_first_order_reflections_match_the_image_method():
    """One reservoir pass with reflectivity 0.4 (bottom) and 0.2 (top), zero
    gain: the centre cells carry the first-order image-method flux of the
    angular quadrature averaged by the vertex hat functions (the reflected
    rays restart at face centroids, which the tolerance allows for)."""
    z, beta = 0.5, SA / (SA + SE)
    mesh, fs = _slab(z=z, beta=beta)
    o = F.ForwardOracle(mesh)
    refl = F.reflecting(o, fs, [0.0, 0.4, 0.2], max_passes=1)
    res = F.evaluate_srm(o, refl, [SA], [SE], min_rays=400_000, max_rays=400_000, steps=0, seed=3)
    assert refl.status()[0] == [3] * 8 and refl.status()[1] == [1] * 8     # stopped at the pass limit
    cen = mesh.centroids()
    near = np.nonzero(np.linalg.norm((cen - [0.5, 0.5, z / 2]) / [1, 1, z], axis=1) < 0.12)[0]

    def flux(p):
        return box_point_flux(p, (0.0, 1.0, 0.0, 1.0), z, beta, 0.0, NTOT / TAU, reflect=True, refl_bottom=0.4,
                              refl_top=0.2, n_mu=200, n_phi=400)

    with np.errstate(invalid="ignore"):
        expect = hat_smoothed_cell_flux(mesh.points, mesh.cells, o.volume, near, flux)
    phi = res["phiAse"][near].astype(np.float64)
    assert abs(phi.sum() / expect.sum() - 1.0) < 0.008, phi.sum() / expect.sum()


def test_reservoir_results_depend_on_the_deposit_order():
    """The surface reservoir (depositReflection, accumulation.hpp:409-455)
    keeps the first `slots` deposits of a face unconditionally and replaces
    later ones with probability min(1, slots w / W), so what it keeps depends
    on the order of the deposits, not only on the rays.  In ray order the
    rays of neighbouring source cells arrive together; in any permuted order
    (the device's atomic arrival order is one) the reflected flux comes out
    0.1-0.4% higher on this slab (gain 1.47 / cm, reflectivity 0.5 / 0.3,
    passes to convergence), while permuted orders agree among themselves
    within their reservoir noise (pairwise mean z within +-0.2).  This is the
    offset a device run shows against the oracle's ray order; the GPU test
    compares against a permuted order."""
    mesh, fs = _slab(nx=6, levels=7, z=0.4, beta=0.25)
    o = F.ForwardOracle(mesh)
    kw = dict(min_rays=120_000, max_rays=120_000, steps=0, seed=21)
    vol = o.volume

    def run(stride):
        refl = F.reflecting(o, fs, [0.0, 0.5, 0.3], tolerance=1e-4, order_stride=stride)
        return F.evaluate_srm(o, refl, [SA], [SE], **kw)

    def mean_z(a, b):
        return float(np.mean((a["phiAse"].astype(np.float64) - b["phiAse"]) /
                             np.hypot(a["standardError"], b["standardError"])))

    ray_order = run(0)
    permuted = [run(s) for s in (1000003, 7919, 104729)]
    assert all(p["rays"] == ray_order["rays"] for p in permuted)
    energy = [float(np.sum(p["phiAse"] * vol) / np.sum(ray_order["phiAse"] * vol)) for p in permuted]
    assert all(1.0005 < e < 1.006 for e in energy), energy
    assert all(0.05 < mean_z(p, ray_order) < 0.4 for p in permuted)
    assert all(abs(mean_z(a, b)) < 0.2 for i, a in enumerate(permuted) for b in permuted[i + 1:])
    # a stride sharing a factor with the ray count keeps the ray order
    same = run(2)
    np.testing.assert_array_equal(same["phiAse"], ray_order["phiAse"])


def test_reflection_passes_converge_and_add_flux():
    mesh, fs = _slab(nx=4, levels=5, z=0.4)
    o = F.ForwardOracle(mesh)
    kw = dict(min_rays=40_000, max_rays=40_000, steps=0, seed=5)
    plain = o.evaluate([SA], [SE], **kw)["phiAse"].astype(np.float64)
    one = F.evaluate_srm(o, F.reflecting(o, fs, [0.0, 0.6, 0.6], max_passes=1), [SA], [SE], **kw)
    many_r = F.reflecting(o, fs, [0.0, 0.6, 0.6], max_passes=40, tolerance=1e-3)
    many = F.evaluate_srm(o, many_r, [SA], [SE], **kw)
    status, passes, remaining = many_r.status()
    assert all(s in (1, 2) for s in status) and min(passes) > 1
    assert all(r < 1e-3 for s, r in zip(status, remaining) if s == 1)
    assert one["phiAse"].sum() > plain.sum() and many["phiAse"].sum() > one["phiAse"].sum()
    # the primary rays are the plain run's: their visits are all there
    assert np.all(many["rayVisits"] >= o.evaluate([SA], [SE], **kw)["rayVisits"])


def test_total_internal_reflection_only_beyond_the_critical_angle():
    mesh, fs = _slab(nx=2, levels=3, z=0.2)
    o = F.ForwardOracle(mesh)
    refl = F.reflecting(o, fs, [0.0, 0.0, 0.0], n_in=[1.0, 1.83, 1.83], n_out=[1.0, 1.0, 1.0])
    crit = math.asin(1.0 / 1.83)
    f = 0
    normal = refl.surfaces.normal[0:3]
    for theta, expect in ((crit - 0.05, 0.0), (crit + 0.05, 1.0), (0.0, 0.0), (1.5, 1.0)):
        # a direction at angle theta to the outward normal of face 0
        n = np.array(normal)
        t = np.cross(n, [1.0, 0.0, 0.0] if abs(n[0]) < 0.9 else [0.0, 1.0, 0.0])
        t /= np.linalg.norm(t)
        d = np.ascontiguousarray(math.cos(theta) * n + math.sin(theta) * t)
        got = F._load().orf_reflectance(F.C.byref(refl.surfaces), f, d.ctypes.data_as(F._f64p))
        assert got == expect, (theta, got)


def test_single_tetrahedron_flux_is_the_mean_exit_distance():
    """One tetrahedron, zero gain: each ray scores its exit distance and the
    lumped vertex volumes of a single cell add up to its volume, so the cell
    flux is (nTot/tau) * beta * (mean exit distance of the rays), the exit
    distances computed here by brute force from the oracle's own ray starts
    (ray-plane distances to the four faces)."""
    pts = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    beta = SA / (SA + SE)
    mesh = TetMesh(points=pts, cells=[[0, 1, 2, 3]], beta=[beta], n_tot=NTOT, tau=TAU)
    o = F.ForwardOracle(mesh)
    rays = 4000
    res = o.evaluate([SA], [SE], min_rays=rays, max_rays=rays, steps=0, seed=8)
    n = np.array([1.0, 1.0, 1.0]) / np.sqrt(3.0)
    total = 0.0
    for r in range(rays):
        ray = o.ray_start(8, 0, rays, r, 1)
        p, u = np.array(ray.p), np.array(ray.u)
        ts = [-p[k] / u[k] for k in range(3) if u[k] < 0]      # planes x_k = 0
        if u @ n > 0:                                          # plane x + y + z = 1
            ts.append((1.0 / np.sqrt(3.0) - p @ n) / (u @ n))
        total += min(ts)
    expect = (NTOT / TAU) * beta * total / rays
    assert res["phiAse"][0] == pytest.approx(expect, rel=1e-6)

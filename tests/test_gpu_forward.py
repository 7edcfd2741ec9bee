"""GPU parity of the forward Tet4 estimator (hase_forward_*) with the FP64
oracle (oracle/forward_oracle.c).

The device walks a ray through exactly the cells the oracle walks (same
Philox streams, same FP64 arithmetic with contraction off), so per-cell visit
and drop counts are equal; vertex scores differ only by the order of the
FP64 atomic sums and by exp/expm1 in their last bits (relative 1e-9).  Cases:
a box with 16 wavelengths and random beta; a disk split from prisms with a
cladding ring (alpha 5.5 / cm); the adaptive evaluation end to end; the
box-centre flux against its expectation from the angular quadrature (the
point flux averaged by the vertex hat functions) at 4e6 rays, for zero gain
and 1.47 / cm; beta refresh; host-side rejections.
"""
import numpy as np
import pytest

from haseongpu_amd import Grid, MeshTopology, _native
from haseongpu_amd.tet_mesh import TetMesh
from haseongpu_amd.workloads import production_disk
from oracle import forward as F
from tests.quadrature import box_point_flux, hat_smoothed_cell_flux

pytestmark = pytest.mark.gpu

SA, SE = 1.16e-21, 2.48e-20
NTOT, TAU = 2.76e20, 9.5e-4


def _box(nx=6, levels=6, z=0.5, beta=None, seed=1):
    topo = MeshTopology.fromGrid(Grid(xExtent=1.0, yExtent=1.0, zExtent=z, tileSizeX=1.0 / nx,
                                      tileSizeZ=z / (levels - 1)))
    b = np.random.default_rng(seed).uniform(0.0, 0.3, topo.numberOfPrisms) if beta is None else \
        np.full(topo.numberOfPrisms, beta)
    return TetMesh.fromPrisms(topo, b, n_tot=NTOT, tau=TAU)


def _disk():
    w = production_disk(levels=5, n_points=300, wavelengths=8, min_rays=1000, max_rays=1000)
    topo = w.topology
    beta = np.asarray(w.medium.value("betaVolume"), dtype=np.float64)
    types = np.asarray(w.medium.value("claddingCellTypes"), dtype=np.uint32)
    mesh = TetMesh.fromPrisms(topo, beta, cladding_types=types, cladding_number=int(w.medium.value("claddingNumber")),
                              cladding_absorption=5.5, n_tot=float(w.medium.value("nTot")),
                              tau=float(w.medium.value("crystalTFluo")))
    sa, se, _, _ = w.spectra.spectralPairs(False)
    return mesh, np.asarray(sa, dtype=np.float64), np.asarray(se, dtype=np.float64)


def _spectra(n, seed=2):
    rng = np.random.default_rng(seed)
    return SA * rng.uniform(0.5, 1.5, n), SE * rng.uniform(0.5, 1.5, n)


def _context(mesh, sa, se, *, min_rays, max_rays, steps=0, threshold=0.005, seed=5489):
    exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=min_rays, max_rays=max_rays,
                                   mse_threshold=threshold)
    comp = _native.make_compute(n_samples=1, adaptive_steps=steps, rng_seed=seed)
    return _native.ForwardContext(mesh, exp, comp)


def _check_raw(mesh, sa, se, seed, launch, rays):
    o = F.ForwardOracle(mesh)
    ref_scores, ref_visits, ref_dropped, counts, ref_seg = o.trace(sa, se, seed, launch, rays)
    with _context(mesh, sa, se, min_rays=rays, max_rays=rays) as ctx:
        scores, visits, dropped, seg = ctx.trace(seed, launch, rays)
    np.testing.assert_array_equal(visits, ref_visits)
    np.testing.assert_array_equal(dropped, ref_dropped)
    assert seg == ref_seg
    np.testing.assert_allclose(scores, ref_scores, rtol=1e-9, atol=1e-12 * np.abs(ref_scores).max())
    return ref_visits


def test_trace_matches_oracle_box_16_wavelengths():
    mesh = _box()
    sa, se = _spectra(16)
    print(f"[forward] box {mesh.n_cells} cells, 200k rays, 16 wavelengths", flush=True)
    visits = _check_raw(mesh, sa, se, seed=77, launch=3, rays=200_003)
    assert visits.sum() > 200_003


def test_trace_matches_oracle_disk_with_cladding():
    mesh, sa, se = _disk()
    assert mesh.cladding_mask().any()
    print(f"[forward] disk {mesh.n_cells} cells with cladding ring, 150k rays", flush=True)
    _check_raw(mesh, sa, se, seed=5, launch=0, rays=150_000)


def test_adaptive_evaluation_matches_oracle():
    mesh = _box(nx=4, levels=5, z=0.4)
    sa, se = _spectra(4, seed=9)
    kw = dict(min_rays=20_000, max_rays=320_000, steps=4, threshold=0.03 ** 2, seed=123)
    o = F.ForwardOracle(mesh)
    ref = o.evaluate(sa, se, min_rays=kw["min_rays"], max_rays=kw["max_rays"], steps=kw["steps"],
                     threshold=kw["threshold"], seed=kw["seed"])
    print(f"[forward] adaptive: oracle {ref['rays']} rays in {ref['launches']} launches", flush=True)
    with _context(mesh, sa, se, **kw) as ctx:
        res = ctx.evaluate()
        st = ctx.stats()
    assert st["rays"] == ref["rays"] and st["launches"] == ref["launches"]
    np.testing.assert_array_equal(res.rayVisits, ref["rayVisits"])
    np.testing.assert_array_equal(res.convergenceRays, ref["convergenceRays"])
    np.testing.assert_allclose(res.phiAse, ref["phiAse"], rtol=2e-7)
    np.testing.assert_allclose(res.relativeStandardError, ref["relativeStandardError"], rtol=1e-6)
    np.testing.assert_allclose(res.dndtAse, ref["dndtAse"], rtol=2e-7, atol=1e-300)
    assert st["segments"] == int(res.rayVisits.sum())


@pytest.mark.parametrize("beta", [SA / (SA + SE), 0.25])
def test_box_centre_flux_against_quadrature_at_4e6_rays(beta):
    z = 0.7
    mesh = _box(nx=10, levels=11, z=z, beta=beta)
    g = NTOT * (beta * (SE + SA) - SA)
    print(f"[forward] box centre flux, beta {beta:.3f}, 4e6 rays", flush=True)
    with _context(mesh, [SA], [SE], min_rays=4_000_000, max_rays=4_000_000) as ctx:
        res = ctx.evaluate()
    cen = mesh.centroids()
    near = np.nonzero(np.linalg.norm((cen - [0.5, 0.5, z / 2]) / [1, 1, z], axis=1) < 0.12)[0]

    def flux(p):
        return box_point_flux(p, (0.0, 1.0, 0.0, 1.0), z, beta, g, NTOT / TAU, n_mu=200, n_phi=400)

    expect = hat_smoothed_cell_flux(mesh.points, mesh.cells, mesh.volumes(), near, flux)
    phi = res.phiAse[near].astype(np.float64)
    print(f"[forward] mean ratio to the smoothed quadrature {phi.sum() / expect.sum():.5f}", flush=True)
    assert abs(phi.sum() / expect.sum() - 1.0) < 0.004, phi.sum() / expect.sum()
    assert np.all(np.abs(phi - expect) <= 4.0 * res.standardError[near] + 0.002 * expect)


def test_update_beta_and_one_shot_call():
    mesh = _box(nx=3, levels=4, z=0.3)
    sa, se = _spectra(2, seed=4)
    beta2 = np.random.default_rng(8).uniform(0.0, 0.3, mesh.n_cells)
    mesh2 = TetMesh(points=mesh.points, cells=mesh.cells, beta=beta2, n_tot=NTOT, tau=TAU)
    with _context(mesh, sa, se, min_rays=50_000, max_rays=50_000) as ctx:
        ctx.update_beta(beta2)
        a = ctx.evaluate()
    exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=50_000, max_rays=50_000)
    b = _native.calc_forward_phi_ase(mesh2, exp, _native.make_compute(n_samples=1, adaptive_steps=0))
    np.testing.assert_array_equal(a.rayVisits, b.rayVisits)
    np.testing.assert_allclose(a.phiAse, b.phiAse, rtol=1e-9)


def test_rejections():
    mesh = _box(nx=2, levels=3, z=0.2)
    flat = TetMesh(points=mesh.points * [1, 1, 0], cells=mesh.cells, beta=mesh.beta)
    with pytest.raises(_native.HaseError, match="degenerate"):
        _context(flat, [SA], [SE], min_rays=10, max_rays=10)
    exp = _native.NativeExperiment(sigma_a=[SA], sigma_e=[SE], min_rays=10, max_rays=10, use_reflections=True)
    with pytest.raises(_native.HaseError, match="surface"):
        _native.ForwardContext(mesh, exp, _native.make_compute(n_samples=1))

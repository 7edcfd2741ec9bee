"""The drop-in boundary on the GPU: the one-shot ``hase_calc_phi_ase`` (the
C form of calcPhiAse / startSimulation, include/core/simulation.hpp:42-333)
against the FP64 oracle on identical random streams, and a ``hase_mesh_t``
built by ``hase_fold_tet4_mesh`` from the Tet4 arrays of a 2.2.0 HostMesh
against the prism mesh it came from.

Tolerances as in test_gpu_parity.py: same streams, median per-sample
|phi_gpu/phi_cpu - 1| <= 1e-5 and 99th percentile <= 1e-4; the folded mesh
must give bit-identical outputs (same arrays, same rays)."""
import numpy as np
import pytest

from haseongpu_amd import _native
from haseongpu_amd.geometry import GainMedium, MeshTopology
from haseongpu_amd.tet_mesh import TetMesh
from haseongpu_amd.workloads import plain_small, production_disk
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_one_shot_entry_matches_the_oracle():
    w = plain_small()
    sa, se, msa, mse_ = w.spectra.spectralPairs(True)
    first, last = 300, 659                     # levels 2-5: an interior block of the sample range
    mesh = _native.NativeMesh(w.topology, w.medium)
    exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=4096, max_rays=65536, max_sigma_a=msa,
                                   max_sigma_e=mse_, mse_threshold=2e-4)
    comp = _native.make_compute(n_samples=w.samples, adaptive_steps=4, min_sample_range=first,
                                max_sample_range=last, rng_seed=91)
    print(f"one-shot hase_calc_phi_ase: plain_small samples {first}..{last}", flush=True)
    res = _native.calc_phi_ase(mesh, exp, comp)
    om = O.OracleMesh(w.topology, w.medium)
    op = O.OracleParams(sigma_a=sa, sigma_e=se, max_sigma_a=msa, max_sigma_e=mse_,
                        n_tot=float(w.medium.value("nTot")), tau=float(w.medium.value("crystalTFluo")),
                        beta_cells=w.medium.value("betaCells"), min_rays=4096, max_rays=65536, adaptive_steps=4,
                        mse_threshold=2e-4, seed=91, sample_min=first, sample_max=last, n_samples=w.samples)
    ref = O.calc_phi_ase(om, op)
    sl = slice(first, last + 1)
    # outside the range: zeros (include/hase_ase.h, hase_result_t)
    assert np.all(res.phiAse[:first] == 0) and np.all(res.totalRays[last + 1:] == 0)
    # the same adaptive decisions on (almost) every sample, hence the same streams
    same = res.totalRays[sl] == ref["totalRays"][sl]
    assert same.mean() >= 0.97, same.mean()
    assert len(np.unique(res.totalRays[sl])) > 1               # the schedule did escalate
    rel = np.abs(res.phiAse[sl][same].astype(np.float64) / ref["phiAse"][sl][same] - 1.0)
    assert np.median(rel) <= 1e-5 and np.quantile(rel, 0.99) <= 1e-4, np.sort(rel)[-5:]
    np.testing.assert_allclose(res.mse[sl][same], ref["mse"][sl][same], rtol=2e-3)
    np.testing.assert_allclose(res.dndtAse[sl][same], ref["dndtAse"][sl][same], rtol=1e-4)
    assert np.all((res.mse[sl] <= 2e-4) | (res.totalRays[sl] == 65536))
    assert np.all(res.droppedRays[sl] == 0)


def _canonical(tri):
    tri = np.asarray(tri, dtype=np.int64)
    k = np.argmin(tri, axis=1)
    return np.stack([tri[np.arange(len(tri)), (k + i) % 3] for i in range(3)], axis=1)


def test_mesh_folded_from_tet4_cells_gives_identical_results():
    """A HostMesh of 2.2.0 shape (three Tet4 cells per wedge, face surface ids
    into reflectivity and refractive-index tables) folded by the library
    evaluates bit for bit like the prism mesh, reflections and cladding on."""
    w = production_disk(levels=5, n_points=320, wavelengths=8, min_rays=4096, max_rays=4096)
    topo, med = w.topology, w.medium
    T, L = topo.numberOfTriangles, int(topo.levels)
    med.set("refractiveIndices", [1.83, 1.0, 1.83, 1.0])
    beta = med.value("betaVolume")
    clad = med.value("claddingCellTypes")
    refl = med.value("reflectivities")
    tm = TetMesh.fromPrisms(topo, beta, cladding_types=clad, cladding_number=1)
    fs = tm.faceSurfacesByPlane({0.0: -1, float(topo.levelCoordinates()[-1]): -2})
    tri_of = (np.arange(tm.n_cells) // 3) % T
    fs = np.where(fs == -1, 1 + tri_of[:, None], np.where(fs == -2, 1 + T + tri_of[:, None], fs))
    n = 1 + 2 * T
    folded = _native.fold_tet4_mesh(
        tm.points.T.reshape(-1), tm.cells.reshape(-1), tm.beta, topo.numberOfPoints, L,
        cladding_cell_types=tm.cell_types, cell_face_boundaries=fs,
        surface_reflectivities=np.concatenate([[0.0], refl[0::2], refl[1::2]]),
        surface_refractive_index_inside=np.full(n, 1.83), surface_refractive_index_outside=np.full(n, 1.0))
    scalars = dict(cladding_number=int(med.value("claddingNumber")),
                   cladding_absorption=float(med.value("claddingAbsorption")), n_tot=float(med.value("nTot")),
                   crystal_t_fluo=float(med.value("crystalTFluo")))
    mesh_a = _native.NativeMesh.fromFolded(folded, beta_cells=med.value("betaCells"), number_of_levels=L, **scalars)
    canon = MeshTopology(topo.points, _canonical(topo.trianglePointIndices), levels=L, thickness=topo.thickness)
    mesh_b = _native.NativeMesh(canon, GainMedium(canon, dict(med.physical)))
    sa, se, msa, mse_ = w.spectra.spectralPairs(False)
    exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=4096, max_rays=4096, max_sigma_a=msa,
                                   max_sigma_e=mse_, use_reflections=True, max_reflections=1)
    comp = _native.make_compute(n_samples=w.samples, rng_seed=13)
    print(f"folded Tet4 mesh vs prism mesh: {T} triangles, {L} levels", flush=True)
    a = _native.calc_phi_ase(mesh_a, exp, comp)
    b = _native.calc_phi_ase(mesh_b, exp, comp)
    for name in ("phiAse", "mse", "totalRays", "dndtAse", "standardError", "droppedRays"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)
    assert np.all(a.totalRays == 4096) and a.droppedRays.sum() == 0

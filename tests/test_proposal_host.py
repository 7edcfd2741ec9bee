"""Host-side pieces of the learned proposal and the batching workload: the
saved network round-trips through a pickle-free .npz, seeded pump fields are
reproducible and physical, and the batch workload honours its ranges."""
import numpy as np

from haseongpu_amd import proposal as prop
from haseongpu_amd.workloads import batch_problems, production_disk


def test_proposal_round_trips_through_npz(tmp_path):
    p = prop.LearnedProposal(prop.init_params(3), np.arange(10, dtype=np.float32),
                             np.linspace(1, 2, 10).astype(np.float32), 8, 2, False, {"rows": 12, "note": "x"})
    path = tmp_path / "net.npz"
    p.save(path)
    with np.load(path, allow_pickle=False) as z:
        assert set(z.files) == {"params", "shift", "scale", "layout", "meta"}
    q = prop.LearnedProposal.load(path)
    assert np.array_equal(q.params, p.params) and np.array_equal(q.shift, p.shift)
    assert np.array_equal(q.scale, p.scale)
    assert (q.cluster_triangles, q.cluster_levels, q.fallback) == (8, 2, False)
    assert q.meta == p.meta
    r = q.with_clusters(1, 1, fallback=True)
    assert (r.cluster_triangles, r.cluster_levels, r.fallback) == (1, 1, True)
    assert r.params is q.params


def test_random_pump_fields_are_seeded_and_physical():
    w = production_disk(levels=5, n_points=300, wavelengths=2)
    clad = np.asarray(w.medium.value("claddingCellTypes")) == 1
    a = prop.random_pump_beta(w.topology, np.random.default_rng(4), cladding=clad)
    b = prop.random_pump_beta(w.topology, np.random.default_rng(4), cladding=clad)
    c = prop.random_pump_beta(w.topology, np.random.default_rng(5), cladding=clad)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.shape == (w.topology.numberOfPrisms,)
    assert a.min() >= 0.0 and a.max() <= 1.0 and a.max() > 0.05
    T = w.topology.numberOfTriangles
    assert np.all(a.reshape(-1, T)[:, clad] == 0.0)


def test_batch_workload_ranges():
    ws = batch_problems(12, seed=3, triangles=(200, 1200), levels=(5, 9), wavelengths=(1, 4))
    T = np.array([w.topology.numberOfTriangles for w in ws])
    L = np.array([w.topology.levels for w in ws])
    W = np.array([w.spectra.spectralPairs(False)[0].size for w in ws])
    assert T.min() >= 150 and T.max() <= 1400 and len(set(T)) > 6
    assert L.min() >= 5 and L.max() <= 9
    assert W.min() >= 1 and W.max() <= 4
    again = batch_problems(12, seed=3, triangles=(200, 1200), levels=(5, 9), wavelengths=(1, 4))
    assert all(np.array_equal(a.medium.value("betaVolume"), b.medium.value("betaVolume")) for a, b in zip(ws, again))
    for w in ws:
        beta = np.asarray(w.medium.value("betaVolume"))
        assert beta.min() >= 0 and beta.max() <= 1 and w.settings["mseThreshold"] == 0.005

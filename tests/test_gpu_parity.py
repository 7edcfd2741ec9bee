"""GPU parity: the HIP library against the FP64 CPU oracle and the legacy fixtures.

All calls go through the C ABI (ctypes).  Tolerances:
* single-ray gain: |G_gpu / G_oracle - 1| <= 1e-5 on identical rays;
* importance map: FP32 vs FP64, rtol 1e-4 per entry (tiny entries by atol);
* estimator on identical random streams: median per-sample |phi_gpu/phi_cpu - 1|
  <= 1e-5 and the 99th percentile <= 1e-4 (rare systematic-sampling boundary
  flips between the FP32/FP64 proposals make the tail);
* legacy fixtures: statistical agreement (see each test).
"""
import numpy as np
import pytest

from haseongpu_amd import _native
from haseongpu_amd.workloads import plain_small, production_disk
from oracle import oracle as O
from tests.helpers import load_compare_serial, load_true_cladding

pytestmark = pytest.mark.gpu


def _context(w, *, min_rays=1000, max_rays=None, **kw):
    sa, se, msa, mse_ = w.spectra.spectralPairs(bool(w.settings.get("monochromatic", False)))
    mesh = _native.NativeMesh(w.topology, w.medium)
    exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=min_rays, max_rays=max_rays or min_rays,
                                   max_sigma_a=msa, max_sigma_e=mse_,
                                   mse_threshold=kw.pop("mse_threshold", 0.005),
                                   use_reflections=kw.pop("use_reflections", False),
                                   max_reflections=kw.pop("max_reflections", 1),
                                   defensive_fraction=kw.pop("defensive", 0.05))
    comp = _native.make_compute(n_samples=w.samples, **kw)
    return _native.Context(mesh, exp, comp), (sa, se, msa, mse_)


def _oracle(w, spectra, **kw):
    sa, se, msa, mse_ = spectra
    om = O.OracleMesh(w.topology, w.medium)
    op = O.OracleParams(sigma_a=sa, sigma_e=se, max_sigma_a=msa, max_sigma_e=mse_,
                        n_tot=float(w.medium.value("nTot")), tau=float(w.medium.value("crystalTFluo")),
                        beta_cells=w.medium.value("betaCells"), n_samples=w.samples, **kw)
    return om, op


def _f32_exact(w):
    """Same workload on a mesh whose coordinates and thickness are exactly
    representable in FP32, so the FP32 kernel and the FP64 oracle walk
    identical geometry and identical rays."""
    from haseongpu_amd.geometry import GainMedium, MeshTopology
    from haseongpu_amd.workloads import Workload

    t = w.topology
    topo = MeshTopology(t.points.astype(np.float32).astype(np.float64), t.trianglePointIndices,
                        levels=t.levels, thickness=float(np.float32(t.thickness)))
    med = GainMedium(topo, dict(w.medium.physical))
    return Workload(w.name, topo, med, w.spectra, w.settings)


def _random_rays(w, om, n, seed, reflections=0):
    """Start points inside random prisms, end points at random sample points
    (or their mirror images)."""
    rng = np.random.default_rng(seed)
    T, L, Np = w.topology.numberOfTriangles, int(w.topology.levels), w.topology.numberOfPoints
    P = T * (L - 1)
    prisms = rng.integers(0, P, n)
    starts = np.array([O.ray_start(om, seed, 0, 0, i, int(j)) for i, j in enumerate(prisms)])
    samples = rng.integers(0, Np * L, n)
    sp = w.topology.samplePoints()[samples]
    nr = rng.integers(0, reflections + 1, n) if reflections else np.zeros(n, np.int64)
    fp = rng.integers(0, 2, n)
    ends = np.array([O.image_point(om, sp[i], int(nr[i]), int(fp[i])) for i in range(n)])
    # identical inputs on both sides: the FP32 values the GPU receives
    return (starts.astype(np.float32).astype(np.float64), prisms, ends.astype(np.float32).astype(np.float64),
            nr, fp)


@pytest.mark.parametrize("case", ["plain_small", "true_cladding", "disk_reflect"])
def test_single_ray_gain_matches_oracle(case):
    if case == "plain_small":
        w = plain_small()
        refl = 0
    elif case == "true_cladding":
        topo, med, sp, ref = load_true_cladding(1)
        from haseongpu_amd.laser import CrossSectionData
        from haseongpu_amd.workloads import Workload
        w = Workload("tc", topo, med, CrossSectionData.monochromatic(
            wavelength=1030e-9, crossSectionAbsorption=sp["sigmaA"], crossSectionEmission=sp["sigmaE"]),
            {"monochromatic": True})
        refl = 0
    else:
        w = production_disk(levels=6, n_points=400, wavelengths=4)
        w.medium.set("refractiveIndices", [1.83, 1.0, 1.83, 1.0])
        refl = 3
    w = _f32_exact(w)
    ctx, spectra = _context(w)
    om, _ = _oracle(w, spectra, min_rays=1)
    starts, prisms, ends, nr, fp = _random_rays(w, om, 4000, seed=11, reflections=refl)
    with ctx:
        out = ctx.trace_rays(starts, prisms, ends, nr, fp)
    sa, se = spectra[0][0], spectra[1][0]
    nt = float(w.medium.value("nTot"))
    alpha = float(w.medium.value("claddingAbsorption"))
    rel = []
    for i in range(len(prisms)):
        o = O.walk_ray(om, starts[i], int(prisms[i]), ends[i], int(nr[i]), int(fp[i]))
        assert o["status"] == 0
        assert out["status"][i] == 0, f"ray {i} dropped on the GPU"
        Eg = nt * (se + sa) * out["beta_len"][i] - nt * sa * out["gain_len"][i] - alpha * out["clad_len"][i]
        Eo = nt * (se + sa) * o["beta_len"] - nt * sa * o["gain_len"] - alpha * o["clad_len"]
        rel.append(abs(np.exp(Eg - Eo) - 1.0))
        assert out["refl"][i] == pytest.approx(o["refl"], rel=1e-6, abs=1e-7)
        assert out["gain_len"][i] + out["clad_len"][i] == pytest.approx(o["length"], rel=2e-5)
    rel = np.array(rel)
    assert rel.max() <= 1e-5, f"max single-ray gain error {rel.max():.3e}"


@pytest.mark.parametrize("reflect", [False, True])
def test_importance_matches_oracle(reflect):
    w = _f32_exact(production_disk(levels=5, n_points=300, wavelengths=8) if reflect else plain_small())
    ctx, spectra = _context(w, use_reflections=reflect, max_reflections=2)
    om, op = _oracle(w, spectra, min_rays=1, use_reflections=reflect, max_reflections=2)
    rng = np.random.default_rng(3)
    with ctx:
        for s in rng.integers(0, w.samples, 6):
            g = ctx.importance(int(s)).astype(np.float64)
            c = O.importance(om, op, int(s))
            np.testing.assert_allclose(g, c, rtol=1e-4, atol=1e-6 * c.max())


def test_estimator_same_streams_plain_small():
    w = plain_small()
    ctx, spectra = _context(w, min_rays=20000, rng_seed=77)
    om, op = _oracle(w, spectra, min_rays=20000, seed=77, mse_threshold=0.005)
    with ctx:
        res = ctx.evaluate()
        st = ctx.stats()
    ref = O.calc_phi_ase(om, op)
    assert np.all(res.totalRays == 20000)
    assert np.all(res.droppedRays == 0)
    rel = np.abs(res.phiAse.astype(np.float64) / ref["phiAse"] - 1.0)
    assert np.median(rel) <= 1e-5 and np.quantile(rel, 0.99) <= 1e-4, np.sort(rel)[-5:]
    np.testing.assert_allclose(res.mse, ref["mse"], rtol=2e-3)
    np.testing.assert_allclose(res.dndtAse, ref["dndtAse"], rtol=1e-4)
    assert st["rays"] == 20000 * w.samples
    assert st["ray_steps"] > st["rays"]


def test_estimator_same_streams_reflections_cladding():
    w = production_disk(levels=6, n_points=500, wavelengths=8)
    ctx, spectra = _context(w, min_rays=8192, use_reflections=True, max_reflections=1, rng_seed=5)
    om, op = _oracle(w, spectra, min_rays=8192, seed=5, use_reflections=True, max_reflections=1)
    with ctx:
        res = ctx.evaluate()
    ref = O.calc_phi_ase(om, op)
    assert res.droppedRays.sum() == 0
    rel = np.abs(res.phiAse.astype(np.float64) / ref["phiAse"] - 1.0)
    assert np.median(rel) <= 2e-5 and np.quantile(rel, 0.99) <= 2e-4, np.sort(rel)[-5:]


def test_adaptive_escalation_matches_oracle():
    w = plain_small()
    kw = dict(min_rays=2000, max_rays=64000)
    ctx, spectra = _context(w, **kw, mse_threshold=2e-4, adaptive_steps=5, max_repetitions=2, rng_seed=9)
    om, op = _oracle(w, spectra, min_rays=2000, max_rays=64000, mse_threshold=2e-4, adaptive_steps=5,
                     max_repetitions=2, seed=9)
    with ctx:
        res = ctx.evaluate()
    ref = O.calc_phi_ase(om, op)
    # identical convergence decisions except where mse sits on the threshold
    agree = np.mean(res.totalRays == ref["totalRays"])
    assert agree > 0.98
    assert set(np.unique(res.totalRays)) <= {2000, 4000, 8000, 16000, 32000, 64000, 128000}
    conv = res.mse <= 2e-4
    assert np.all(conv | (res.totalRays == 128000))


def test_true_cladding_fixture_statistics():
    """Legacy importance-sampled phiASE at 2000 rays/sample (F2): the GPU at
    2000 rays must match its sample-mean level, and at 100k rays the legacy
    values must scatter around the GPU estimate like a 2000-ray estimator."""
    for idx in range(3):
        topo, med, sp, ref = load_true_cladding(idx)
        from haseongpu_amd.laser import CrossSectionData
        from haseongpu_amd.workloads import Workload
        w = Workload("tc", topo, med, CrossSectionData.monochromatic(
            wavelength=1030e-9, crossSectionAbsorption=sp["sigmaA"], crossSectionEmission=sp["sigmaE"]),
            {"monochromatic": True})
        ctx, _ = _context(w, min_rays=100000, rng_seed=5489)
        with ctx:
            res = ctx.evaluate()
        phi = res.phiAse.astype(np.float64)
        ratio = ref["phiASE"] / phi
        assert abs(np.median(ratio) - 1.0) < 0.03
        assert abs(ref["phiASE"].mean() / phi.mean() - 1.0) < 0.02
        assert res.droppedRays.sum() == 0


@pytest.mark.parametrize("name", ["cuboid", "cylindrical"])
def test_compare_serial_fixture(name):
    """Legacy serial solver, 1e6 rays/sample (F1): per-sample agreement."""
    topo, med, sp, ref = load_compare_serial(name)
    from haseongpu_amd.laser import CrossSectionData
    from haseongpu_amd.workloads import Workload
    w = Workload(name, topo, med, CrossSectionData.monochromatic(
        wavelength=905e-9, crossSectionAbsorption=sp["sigmaA"], crossSectionEmission=sp["sigmaE"]),
        {"monochromatic": True})
    ctx, _ = _context(w, min_rays=200000, rng_seed=1)
    with ctx:
        res = ctx.evaluate()
    phi = res.phiAse.astype(np.float64)
    ratio = ref["phiASE"] / phi
    assert abs(np.median(ratio) - 1.0) < 0.005
    # the point estimator's 1/r^2 integrand has a heavy tail: a handful of samples
    # carry a ray that landed next to the sample point (their own RSE reports it),
    # so the level is compared on the 1%-trimmed ratios
    lo, hi = np.quantile(ratio, [0.005, 0.995])
    core = (ratio >= lo) & (ratio <= hi)
    assert abs(ref["phiASE"][core].mean() / phi[core].mean() - 1.0) < 0.005
    assert np.quantile(np.abs(ratio - 1.0), 0.9) < 0.03
    # dndtAse = (beta (sigma_e + sigma_a) - sigma_a) * phiAse (compareSerial README)
    gain = med.value("betaCells") * (sp["sigmaE"] + sp["sigmaA"]) - sp["sigmaA"]
    np.testing.assert_allclose(res.dndtAse, gain * phi, rtol=1e-12)


def test_beta_volume_proposal_same_streams():
    """HASE_PROPOSAL_BETA_VOLUME (no map: rays in proportion to beta*V, each
    reflection slice 1/R) against the oracle fed an all-zero external
    importance, whose defensive mixture is then beta*V / (R sum beta*V)
    alone: identical rays, the same-stream tolerance, with reflections and
    cladding on and an escalating schedule."""
    w = production_disk(levels=6, n_points=500, wavelengths=8)
    first, last = 500, 899
    ctx, spectra = _context(w, min_rays=4096, max_rays=32768, mse_threshold=1e-3, adaptive_steps=3,
                            use_reflections=True, max_reflections=1, rng_seed=23, min_sample_range=first,
                            max_sample_range=last)
    with ctx:
        ctx.set_beta_volume_proposal()
        res = ctx.evaluate()
        st = ctx.stats()
    R = 3
    imp = np.zeros((last - first + 1, R * w.topology.numberOfPrisms))
    om, op = _oracle(w, spectra, min_rays=4096, max_rays=32768, mse_threshold=1e-3, adaptive_steps=3,
                     use_reflections=True, max_reflections=1, seed=23, sample_min=first, sample_max=last,
                     importance=imp)
    ref = O.calc_phi_ase(om, op)
    sl = slice(first, last + 1)
    assert st["importance_walks"] == 0                       # no map was walked
    same = res.totalRays[sl] == ref["totalRays"][sl]
    assert same.mean() > 0.97 and len(np.unique(res.totalRays[sl])) > 1
    rel = np.abs(res.phiAse[sl][same].astype(np.float64) / ref["phiAse"][sl][same] - 1.0)
    assert np.median(rel) <= 2e-5 and np.quantile(rel, 0.99) <= 2e-4, np.sort(rel)[-5:]
    assert res.droppedRays[sl].sum() == 0

"""Combined case at full scale, one share of its sample points.

The 20,682-triangle, 30-level disk (599,778 prisms, 315,300 sample points,
128 wavelengths, first-order reflections, cladding ring, adaptive
1e5 -> 1e8 rays, mse 0.005) after two seeded pump-depletion beta updates,
evaluated on 8 contiguous mid-slab sample points (the unit the path shards
by; the mesh and beta are whole and resident).

* plain form on the oracle's random streams: per-sample relative difference
  to the FP64 oracle with median <= 5e-4 and max <= 3e-3, and every
  difference within one standard error of the estimate (FP32 importance
  against FP64 over 1.8M proposal entries per sample moves rays across
  systematic-sampling boundaries more often than on small meshes: 1.5e-4
  median, 7.7e-4 max, at most 0.62 standard errors, on one MI355X), mse
  within 10% (the boundary flips move the heaviest rays), identical ray
  counts, nothing dropped;
* learned form (an untrained network, so the walked-map fallback is
  exercised) within 4 combined standard errors of the oracle per sample.

On the same mesh with FP32-representable coordinates (both sides walk
identical geometry), 2000 random rays with up to one reflection (paths up
to ~10 cm) keep the single-ray gain within 1e-5 of the FP64 walk at the
largest emission cross section, with no mean drift (|mean| <= 1e-6); the
walked importance map of a sample agrees entry by entry within 1e-4 except
for at most 4 of its ~1.1M entries above 1e-6 of its maximum (a
centroid-to-sample path that grazes a cell edge takes the other side in
FP32: one entry per sample seen, -27% and -7%), none zero on one side
only, and its total within 1e-6.
"""
import numpy as np
import pytest

from haseongpu_amd.geometry import GainMedium
from haseongpu_amd import proposal as prop
from haseongpu_amd.workloads import Workload, combined_disk, depletion_field
from oracle import oracle as O
from tests.test_gpu_parity import _context, _f32_exact, _oracle, _random_rays

pytestmark = pytest.mark.gpu


def test_combined_scale_share_matches_oracle():
    w = combined_disk()
    topo = w.topology
    assert topo.numberOfPrisms == 599_778 and w.spectra.spectralPairs(False)[0].size == 128
    first = (int(topo.levels) // 2) * topo.numberOfPoints
    last = first + 7
    part = slice(first, last + 1)
    beta = np.asarray(w.medium.value("betaVolume"), dtype=np.float64)
    factor = depletion_field(topo)
    s = w.settings
    kw = dict(min_rays=s["minRaysPerSample"], max_rays=s["maxRaysPerSample"], mse_threshold=s["mseThreshold"],
              adaptive_steps=s["adaptiveSteps"], use_reflections=True, max_reflections=1)
    ctx, spectra = _context(w, rng_seed=s["rngSeed"], min_sample_range=first, max_sample_range=last, **kw)
    print(f"combined share {first}..{last}", flush=True)
    with ctx:
        for _ in range(2):
            beta = beta * factor
            ctx.update_beta(beta)
        plain = ctx.evaluate()
        net = prop.LearnedProposal(prop.init_params(2), np.zeros(10, np.float32),
                                   np.array([0.5, 0.5, 2.0, 0.5, 0.5, 2.0, 0.2, 3.0, 0.5, 0.5], np.float32),
                                   16, 4, True)
        ctx.set_proposal(net)
        learned = ctx.evaluate()
    phys = dict(w.medium.physical)
    phys["betaVolume"] = beta
    wl = Workload(w.name, topo, GainMedium(topo, phys), w.spectra, w.settings)
    om, op = _oracle(wl, spectra, seed=s["rngSeed"], sample_min=first, sample_max=last,
                     **{k: v for k, v in kw.items()})
    ref = O.calc_phi_ase(om, op)
    phi_r = ref["phiAse"][part]
    assert np.all(ref["droppedRays"][part] == 0)
    assert np.all(plain.droppedRays[part] == 0) and np.all(learned.droppedRays[part] == 0)
    assert np.all(plain.mse[part] <= s["mseThreshold"]) and np.all(learned.mse[part] <= s["mseThreshold"])
    np.testing.assert_array_equal(plain.totalRays[part], ref["totalRays"][part])
    diff = plain.phiAse[part].astype(np.float64) - phi_r
    print("signed relative differences", diff / phi_r, "in standard errors", diff / plain.standardError[part])
    rel = np.abs(diff / phi_r)
    assert np.median(rel) <= 5e-4 and rel.max() <= 3e-3, rel
    assert np.all(np.abs(diff) <= plain.standardError[part]), diff / plain.standardError[part]
    np.testing.assert_allclose(plain.mse[part], ref["mse"][part], rtol=0.1)
    se_r = np.sqrt(ref["mse"][part]) * phi_r
    z = (learned.phiAse[part].astype(np.float64) - phi_r) / np.sqrt(learned.standardError[part] ** 2 + se_r ** 2)
    assert np.max(np.abs(z)) < 4, z


def test_combined_scale_single_rays_and_importance_match_oracle():
    w = _f32_exact(combined_disk())
    topo = w.topology
    first = (int(topo.levels) // 2) * topo.numberOfPoints
    kw = dict(min_rays=100_000, max_rays=100_000_000, mse_threshold=0.005, use_reflections=True,
              max_reflections=1)
    ctx, spectra = _context(w, min_sample_range=first, max_sample_range=first, **kw)
    om, op = _oracle(w, spectra, sample_min=first, sample_max=first, **kw)
    starts, prisms, ends, nr, fp = _random_rays(w, om, 2000, seed=11, reflections=1)
    print(f"combined single rays and importance of sample {first}", flush=True)
    with ctx:
        out = ctx.trace_rays(starts, prisms, ends, nr, fp)
        imp = ctx.importance(first).astype(np.float64)
    sa, se = spectra[0], spectra[1]
    k = int(np.argmax(se))
    nt = float(w.medium.value("nTot"))
    alpha = float(w.medium.value("claddingAbsorption"))
    err = []
    for i in range(prisms.size):
        o = O.walk_ray(om, starts[i], int(prisms[i]), ends[i], int(nr[i]), int(fp[i]))
        assert o["status"] == 0 and out["status"][i] == 0, i
        Eg = nt * (se[k] + sa[k]) * out["beta_len"][i] - nt * sa[k] * out["gain_len"][i] - alpha * out["clad_len"][i]
        Eo = nt * (se[k] + sa[k]) * o["beta_len"] - nt * sa[k] * o["gain_len"] - alpha * o["clad_len"]
        err.append(np.exp(Eg - Eo) - 1.0)
    err = np.array(err)
    assert np.abs(err).max() <= 1e-5 and abs(err.mean()) <= 1e-6, (np.abs(err).max(), err.mean())
    ref = O.importance(om, op, first)
    assert not np.any((imp > 0) != (ref > 0))
    big = ref > 1e-6 * ref.max()
    rel = np.abs(imp[big] / ref[big] - 1.0)
    assert np.count_nonzero(rel > 1e-4) <= 4, np.sort(rel)[-6:]
    assert abs(imp.sum() / ref.sum() - 1.0) <= 1e-6

"""GPU tests of the learned proposal on the remaining configurations.

* Small slab (the 1 x 1 cm, 200-triangle, 10-level mesh): a network trained on
  2.05e6 rows from 500 seeded random pump fields of that mesh, evaluated on
  held-out fields with
  the adaptive schedule to mse <= 0.005.  Its estimates agree with the walked
  map's (at most 2% of the samples beyond 3 combined standard errors, |mean
  z| < 0.3); it traces fewer rays than beta*V sampling (the estimator without
  importance information) to reach the threshold, and at a fixed 4096 rays its
  median relative error is below 0.7 of beta*V's and within 2.5x of the
  walked map's.
* Combined: reflections, cladding and 16 wavelengths on a small disk, the
  time-stepped loop (frozen-flux RK4, two steps) with the learned proposal
  against the same loop with the walked map: the last ASE evaluation agrees
  statistically and beta differs by a small fraction of what ASE changes.
"""
import numpy as np
import pytest

from haseongpu_amd import _native
from haseongpu_amd import proposal as prop
from haseongpu_amd.simulation import native_problem
from haseongpu_amd.workloads import local_pump, plain_small, production_disk

pytestmark = pytest.mark.gpu


def _z(a_phi, a_se, b_phi, b_se):
    ok = (a_phi > 0) & (b_phi > 0)
    return (a_phi[ok].astype(np.float64) - b_phi[ok]) / np.sqrt(a_se[ok] ** 2 + b_se[ok] ** 2)


def test_small_slab_learned_proposal_on_held_out_fields():
    w = plain_small()
    center = (0.5, 0.5)
    # the configuration's training set: ~2e6 (sample, prism, transfer) rows
    # from 500 seeded random pump fields of the small slab
    print("[learned-small] training on 500 random pump fields of the small slab", flush=True)
    net, (X, _) = prop.train_on_workload(w, fields=500, samples_per_field=16, rows_per_sample=256, seed=3,
                                         epochs=10, batch=4096, cluster_triangles=1, cluster_levels=1,
                                         center=center)
    assert X.shape[0] == 500 * 16 * 256
    rng = np.random.default_rng(2024)
    rays = {"walked": 0, "learned": 0, "beta_v": 0}
    rse = {"walked": [], "learned": [], "beta_v": []}
    zs = []
    runs = (("walked", 0.05, False), ("learned", 0.05, True), ("beta_v", 0.999, False))
    for field in range(3):
        beta = prop.random_pump_beta(w.topology, rng, center=center)
        out = {}
        # adaptive 256 -> 2^20 rays to mse <= 0.005; every ray traced counts
        # (the learned form's restarted fallback samples included)
        w.settings.update(minRaysPerSample=256, maxRaysPerSample=1 << 20, adaptiveSteps=12, rngSeed=17 + field)
        for name, defensive, learned in runs:
            mesh, exp, comp, _ = native_problem(w, defensiveFraction=defensive)
            with _native.Context(mesh, exp, comp) as ctx:
                ctx.update_beta(beta)
                if learned:
                    ctx.set_proposal(net)
                out[name] = ctx.evaluate()
                rays[name] += int(ctx.stats()["rays"])
            assert np.all(out[name].mse <= 0.005), name
        zs.append(_z(out["learned"].phiAse, out["learned"].standardError, out["walked"].phiAse,
                     out["walked"].standardError))
        # fixed 4096 rays per sample: relative standard error of each proposal
        w.settings.update(minRaysPerSample=4096, maxRaysPerSample=4096, adaptiveSteps=1)
        for name, defensive, learned in runs:
            mesh, exp, comp, _ = native_problem(w, defensiveFraction=defensive)
            with _native.Context(mesh, exp, comp) as ctx:
                ctx.update_beta(beta)
                if learned:
                    ctx.set_proposal(net.with_clusters(1, 1, fallback=False))
                r = ctx.evaluate()
            rse[name].append(np.sqrt(r.mse[r.phiAse > 0]))
    z = np.concatenate(zs)
    med = {k: float(np.median(np.concatenate(v))) for k, v in rse.items()}
    print(f"[learned-small] rays to mse 0.005 {rays}; median rse at 4096 rays {med}; "
          f"mean z {z.mean():.3f}, beyond 3 sigma {np.mean(np.abs(z) > 3):.4f}", flush=True)
    assert np.mean(np.abs(z) > 3) <= 0.02 and abs(z.mean()) < 0.3
    assert rays["learned"] < 0.8 * rays["beta_v"], rays
    assert med["learned"] < 0.7 * med["beta_v"] and med["learned"] < 2.5 * med["walked"], med


def test_combined_time_loop_with_learned_proposal():
    w = production_disk(levels=6, n_points=400, wavelengths=16, min_rays=4096, max_rays=1 << 20)
    print("[combined] training on 8 random pump fields", flush=True)
    net, _ = prop.train_on_workload(w, fields=8, samples_per_field=32, rows_per_sample=512, seed=4, epochs=8,
                                    batch=2048, cluster_triangles=4, cluster_levels=2)
    pump = local_pump(w.topology, peak_rate=2500.0, width=1.0, absorption=1.0)
    beta0 = np.asarray(w.medium.value("betaVolume"), dtype=np.float64)
    h = 2e-4
    states = {}
    for name in ("walked", "learned", "dark"):
        mesh, exp, comp, _ = native_problem(w)
        with _native.Context(mesh, exp, comp) as ctx:
            if name == "learned":
                ctx.set_proposal(net)
            with _native.NativeSimulation(ctx, time_step=h, integrator="frozen-phi-ase-runge-kutta-4",
                                          beta_volume=beta0, pump_rate=pump.rate, pump_saturation=pump.saturation,
                                          ase_steps=0 if name == "dark" else 5, pump_steps=5) as sim:
                sim.step(2)
                states[name] = sim.state()
    lw, ll = states["walked"], states["learned"]
    assert ll["ase_evaluations"] == lw["ase_evaluations"] == 2
    assert np.all(ll["mse"] <= 0.005) and np.all(lw["mse"] <= 0.005)
    z = _z(ll["phi_samples"], ll["standard_error"], lw["phi_samples"], lw["standard_error"])
    change = np.abs(lw["beta_volume"] - states["dark"]["beta_volume"]).max()
    diff = np.abs(ll["beta_volume"] - lw["beta_volume"]).max()
    print(f"[combined] mean z {z.mean():.3f}, beyond 3 sigma {np.mean(np.abs(z) > 3):.4f}, "
          f"beta difference {diff:.2e} of ASE change {change:.2e}", flush=True)
    assert np.mean(np.abs(z) > 3) <= 0.02 and abs(z.mean()) < 0.3
    assert change > 1e-4 and diff < 0.05 * change

"""GPU parity of the time-stepped beta loop (hase_simulation_*) with the
oracle's restatement (oracle/time_stepping.py).

* ASE off: every integrator, cladding, pump with saturation - the device loop
  and the oracle run the same FP64 operations in the same order (the device
  kernels are built without contraction), so beta, dndt and the derivative
  must be bit-identical after several steps;
* ASE on (small grid, identical random streams): the ASE flux of each stage
  comes from the FP32 walk on the device and the FP64 walk in the oracle, so
  the prism flux agrees to the estimator's identical-stream tolerance (median
  1e-5, 99th percentile 1e-4), and beta differs from the oracle's by the same
  fractions of the largest change that ASE itself makes to beta (median 1e-5,
  99th percentile 1e-4, at most 1e-3);
* production disk, one explicit Euler step: the prism flux is the six-vertex
  mean of the returned sample flux and dndt_ase / dndt_pump follow from it
  and beta exactly as the oracle's composition computes them.
"""
import numpy as np
import pytest

from haseongpu_amd import _native
from haseongpu_amd.workloads import local_pump, plain_small, production_disk
from oracle import oracle as O
from oracle import time_stepping as TS

pytestmark = pytest.mark.gpu

ALL = ("explicit-euler", "heun", "midpoint", "runge-kutta-4", "frozen-phi-ase-runge-kutta-4", "implicit-euler",
       "exponential-euler")


def _context(w, min_rays, seed, **kw):
    sa, se, msa, mse_ = w.spectra.spectralPairs(bool(w.settings.get("monochromatic", False)))
    mesh = _native.NativeMesh(w.topology, w.medium)
    exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=min_rays, max_rays=min_rays,
                                   max_sigma_a=msa, max_sigma_e=mse_, mse_threshold=0.005,
                                   use_reflections=kw.pop("use_reflections", False), max_reflections=1,
                                   defensive_fraction=0.05)
    comp = _native.make_compute(n_samples=w.samples, rng_seed=seed, **kw)
    return _native.Context(mesh, exp, comp), (sa, se, msa, mse_)


def _oracle_loop(w, spectra, integrator, beta, pump, *, ase, h, **run):
    sa, se, msa, mse_ = spectra
    topo, med = w.topology, w.medium
    clad = np.asarray(med.value("claddingCellTypes")) == int(med.value("claddingNumber"))
    return TS.TimeLoop(tri=topo.trianglePointIndices, n_points=topo.numberOfPoints, layers=int(topo.levels) - 1,
                       beta=beta, pump_rate=pump.rate, pump_saturation=pump.saturation,
                       cladding=np.tile(clad, int(topo.levels) - 1), sigma_e=mse_,
                       sigma_a=TS.absorption_at_emission_peak(sa, se, msa), tau=float(med.value("crystalTFluo")),
                       time_step=h, integrator=integrator, ase=ase, **run)


def _initial_beta(w, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 0.6, size=w.topology.numberOfPrisms)


@pytest.mark.parametrize("integrator", ALL)
def test_integrators_bit_identical_without_ase(integrator):
    w = production_disk(levels=5, n_points=300, wavelengths=4)
    beta0 = _initial_beta(w, 11)
    pump = local_pump(w.topology, pump_steps=4)
    ctx, spectra = _context(w, 1000, 3)
    run = dict(implicit_iterations=3, ase_steps=0, pump_steps=4)
    h = 1.5e-4
    print(f"[time] {integrator}: {w.topology.numberOfPrisms} prisms, ASE off", flush=True)
    with ctx, _native.NativeSimulation(ctx, time_step=h, integrator=integrator, beta_volume=beta0,
                                       pump_rate=pump.rate, pump_saturation=pump.saturation, **run) as sim:
        sim.step(3)
        mid = sim.state()
        sim.step(3)
        st = sim.state()
    ref = _oracle_loop(w, spectra, integrator, beta0, pump, ase=None, h=h, **run)
    ref.step(3)
    np.testing.assert_array_equal(mid["beta_volume"], ref.beta)
    ref.step(3)
    np.testing.assert_array_equal(st["beta_volume"], ref.beta)
    np.testing.assert_array_equal(st["dndt_pump"], ref.dndt_pump)
    np.testing.assert_array_equal(st["dndt_ase"], ref.dndt_ase)
    assert st["step"] == 6 and st["ase_evaluations"] == 0
    np.testing.assert_allclose(st["time"], 6 * h, rtol=1e-15)
    assert np.all(st["phi_ase"] == 0.0)
    # the pump stopped after step 4, and cladding prisms kept beta except
    # under exponential Euler, whose decay the reference applies everywhere
    assert np.all(st["dndt_pump"] == 0.0) and np.any(mid["dndt_pump"] != 0.0)
    clad = np.tile(w.topology.boundaryTriangles(), int(w.topology.levels) - 1)
    if integrator != "exponential-euler":
        np.testing.assert_array_equal(st["beta_volume"][clad], np.clip(beta0[clad], 0, 1))
    assert np.all((st["beta_volume"] >= 0) & (st["beta_volume"] <= 1))


@pytest.mark.parametrize("integrator", ["explicit-euler", "heun", "runge-kutta-4", "frozen-phi-ase-runge-kutta-4",
                                        "implicit-euler", "exponential-euler"])
def test_ase_coupled_steps_match_oracle(integrator):
    w = plain_small(rays=2000)
    beta0 = _initial_beta(w, 5)
    pump = local_pump(w.topology, peak_rate=2000.0, width=0.4, center=(0.5, 0.5))
    rays, seed, h = 2000, 21, 2e-4
    run = dict(implicit_iterations=2, ase_steps=10, pump_steps=10)
    print(f"[time] {integrator}: plain grid, ASE on, {rays} rays", flush=True)
    ctx, spectra = _context(w, rays, seed)
    with ctx, _native.NativeSimulation(ctx, time_step=h, integrator=integrator, beta_volume=beta0,
                                       pump_rate=pump.rate, pump_saturation=pump.saturation, **run) as sim:
        sim.step(2)
        st = sim.state()
    sa, se, msa, mse_ = spectra
    om = O.OracleMesh(w.topology, w.medium)
    op = O.OracleParams(sigma_a=sa, sigma_e=se, max_sigma_a=msa, max_sigma_e=mse_,
                        n_tot=float(w.medium.value("nTot")), tau=float(w.medium.value("crystalTFluo")),
                        beta_cells=w.medium.value("betaCells"), min_rays=rays, seed=seed, n_samples=w.samples)
    ref = _oracle_loop(w, spectra, integrator, beta0, pump, ase=TS.oracle_ase(om, op), h=h, **run).step(2)
    assert st["ase_evaluations"] == ref.ase_evaluations
    assert np.all(st["total_rays"] == rays)
    rel = np.abs(st["phi_ase"] / ref.phi - 1.0)
    assert np.median(rel) <= 1e-5 and np.quantile(rel, 0.99) <= 1e-4, np.sort(rel)[-5:]
    # beta differs only through the flux: on the scale of the largest change
    # that ASE makes to beta, the difference has the flux's own distribution
    dark = _oracle_loop(w, spectra, integrator, beta0, pump, ase=None, h=h,
                        **{**run, "ase_steps": 0}).step(2)
    scale = np.abs(ref.beta - dark.beta).max()
    assert scale > 1e-2
    err = np.abs(st["beta_volume"] - ref.beta) / scale
    print(f"[time] {integrator}: beta error / ASE change: median {np.median(err):.1e}, "
          f"q99 {np.quantile(err, 0.99):.1e}, max {err.max():.1e}", flush=True)
    assert np.median(err) <= 1e-5 and np.quantile(err, 0.99) <= 1e-4 and err.max() <= 1e-3, np.sort(err)[-5:]


def test_production_disk_stage_composition():
    w = production_disk()
    beta0 = np.asarray(w.medium.value("betaVolume"), dtype=np.float64)
    pump = local_pump(w.topology, peak_rate=1500.0, width=1.2, absorption=1.0)
    ctx, spectra = _context(w, 20000, 7, use_reflections=True)
    print(f"[time] production disk: {w.topology.numberOfPrisms} prisms, {w.samples} samples", flush=True)
    h = 1e-4
    with ctx, _native.NativeSimulation(ctx, time_step=h, integrator="explicit-euler", beta_volume=beta0,
                                       pump_rate=pump.rate, pump_saturation=pump.saturation,
                                       ase_steps=5, pump_steps=5) as sim:
        sim.step(1)
        st = sim.state()
        times = sim.times()
    assert st["ase_evaluations"] == 1
    assert np.all(st["total_rays"] == 20000)
    topo = w.topology
    layers = int(topo.levels) - 1
    phi_p = TS.prism_flux(st["phi_samples"], topo.trianglePointIndices, topo.numberOfPoints, layers)
    np.testing.assert_array_equal(st["phi_ase"], phi_p)
    # the one stage of explicit Euler composes at beta0 with this flux
    sa, se, msa, mse_ = spectra
    sig_a = TS.absorption_at_emission_peak(sa, se, msa)
    clad = np.tile(topo.boundaryTriangles(), layers)
    dp, da, d = TS.compose(beta0, phi_p, pump.rate, pump.saturation, clad, mse_ + sig_a, sig_a,
                           float(w.medium.value("crystalTFluo")), True)
    np.testing.assert_array_equal(st["dndt_pump"], dp)
    np.testing.assert_array_equal(st["dndt_ase"], da)
    np.testing.assert_array_equal(st["beta_volume"], TS.clip(TS.add_scaled(beta0, d, h)))
    np.testing.assert_array_equal(st["beta_volume"][clad], beta0[clad])
    assert np.all(st["phi_ase"][~clad] > 0.0)
    assert times["ase_ms"] > 0.0 and times["integrator_ms"] > 0.0
    assert times["rays"] == 20000 * w.samples


def test_simulation_front_end_schedules_pumps_outputs_and_control():
    from haseongpu_amd import LocalPump, PhiASE, RungeKutta4, Simulation

    w = plain_small(rays=1000)
    topo = w.topology
    beta0 = _initial_beta(w, 9)
    w.medium.set("betaVolume", beta0)
    p1 = local_pump(topo, peak_rate=1200.0, width=0.3, center=(0.3, 0.3), pump_steps=2)
    p2 = LocalPump(rate=np.full(topo.numberOfPrisms, 400.0), saturation=None, pump_steps=4)
    phi = PhiASE(crossSections=w.spectra, **{**w.settings, "rngSeed": 4}, ase_steps=3)
    seen = []
    sim = Simulation(gain_medium=w.medium, phi_ase=phi, time_integrator=RungeKutta4(), time_step_size=1e-4,
                     output_steps=(2, 4))
    sim.add_pump(p1).add_pump(p2).on_step(lambda st, tag: seen.append((tag, st.step, st.betaVolume.copy())), "x")
    print("[time] front end: RK4, two pumps, outputs at 2 and 4", flush=True)
    sim.step(4)
    st = sim.get_last_state()
    assert [s[1] for s in seen] == [2, 4] and seen[0][0] == "x"
    assert st.betaVolume.shape == (topo.numberOfTriangles, topo.levels - 1)
    assert st.totalRays.shape == (topo.numberOfPoints, topo.levels)
    assert sim.current_step == 4 and abs(sim.current_time - 4e-4) < 1e-18
    np.testing.assert_array_equal(w.medium.value("betaVolume"), st.betaVolume.reshape(-1, order="F"))
    sim.close()

    # the same schedule driven through the C ABI directly
    w.medium.set("betaVolume", beta0)
    mesh, exp, comp = phi.nativeArguments(w.medium, w.spectra)
    r1, s1 = p1.arrays(topo.numberOfPrisms)
    r2, _ = p2.arrays(topo.numberOfPrisms)
    with _native.Context(mesh, exp, comp) as ctx, _native.NativeSimulation(
            ctx, time_step=1e-4, integrator="runge-kutta-4", beta_volume=beta0, pump_rate=r1 + r2,
            pump_saturation=s1, ase_steps=3, pump_steps=0xFFFFFFFF) as nat:
        nat.step(2)
        b2 = nat.state()["beta_volume"]
        nat.set_pump(r2, None)
        nat.step(2)
        b4 = nat.state()["beta_volume"]
    np.testing.assert_array_equal(seen[0][2].reshape(-1, order="F"), b2)
    np.testing.assert_array_equal(seen[1][2].reshape(-1, order="F"), b4)

    # synchronized-debug: before_step halves beta after every nonfinal step
    w.medium.set("betaVolume", beta0)
    dbg = Simulation(gain_medium=w.medium, phi_ase=phi, time_integrator="explicit-euler", time_step_size=1e-4,
                     execution_mode="synchronized-debug", control_fields=("beta_volume",))
    dbg.add_pump(LocalPump(rate=np.full(topo.numberOfPrisms, 400.0), pump_steps=3))
    dbg.before_step(lambda s: s.gain_medium.set("betaVolume", 0.5 * s.gain_medium.value("betaVolume")))
    steps = []
    dbg.on_step(lambda st: steps.append(st.betaVolume.reshape(-1, order="F").copy()))
    dbg.step(2)
    dbg.close()
    mesh, exp, comp = phi.nativeArguments(w.medium, w.spectra)
    with _native.Context(mesh, exp, comp) as ctx, _native.NativeSimulation(
            ctx, time_step=1e-4, integrator="explicit-euler", beta_volume=beta0,
            pump_rate=np.full(topo.numberOfPrisms, 400.0), ase_steps=3, pump_steps=0xFFFFFFFF) as nat:
        nat.step(1)
        first = nat.state()["beta_volume"]
        nat.set_beta(0.5 * first)
        nat.step(1)
        second = nat.state()["beta_volume"]
    assert len(steps) == 2
    np.testing.assert_array_equal(steps[0], first)
    np.testing.assert_array_equal(steps[1], second)

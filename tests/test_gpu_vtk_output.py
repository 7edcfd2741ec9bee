"""GPU runs that write the legacy VTK outputs.

* ``PhiASE(writeVtk=True)`` writes the five per-sample files of the C++ entry
  (include/core/simulation.hpp:161-224) under <outputPath>/vtk; the phiase
  file holds the run's phiAse to its six written decimals and its header the
  traced ray count.
* A ``vtkWedge`` callback on ``Simulation.on_step`` writes every second step;
  the beta written equals the state's beta bit for bit.
"""
import numpy as np
import pytest

from haseongpu_amd import PhiASE, RungeKutta4, Simulation, vtkWedge
from haseongpu_amd.vtk_wedge import readVtkScalars
from haseongpu_amd.workloads import local_pump, plain_small

pytestmark = pytest.mark.gpu


def test_phiase_writes_run_files(tmp_path):
    w = plain_small(rays=1000)
    print("[vtk] PhiASE writeVtk on the small slab", flush=True)
    phi = PhiASE(crossSections=w.spectra, **{**w.settings, "rngSeed": 7}, writeVtk=True, outputPath=tmp_path)
    res = phi.run(w.medium)
    files = sorted(p.name.rsplit("_", 1)[0] for p in (tmp_path / "vtk").glob("*.vtk"))
    assert files == ["dndt", "phiase", "relative_standard_error", "standard_error", "total_rays"]
    path = next((tmp_path / "vtk").glob("phiase_*.vtk"))
    header = path.read_text(encoding="utf-8").splitlines()[1]
    assert header.startswith(f"RAYS={int(res.stats['rays'])} MAXRAYS={int(w.settings['maxRaysPerSample'])} ")
    kind, values = readVtkScalars(path)["scalars"]
    assert kind == "POINT_DATA" and values.size == w.topology.numberOfSamples
    np.testing.assert_allclose(values, res.phiAse.astype(np.float64), rtol=1e-12, atol=5.1e-7)
    rays = readVtkScalars(next((tmp_path / "vtk").glob("total_rays_*.vtk")))["scalars"][1]
    np.testing.assert_array_equal(rays, res.totalRays.astype(np.float64))


def test_simulation_step_callback_writes_every_second_step(tmp_path):
    w = plain_small(rays=1000)
    topo = w.topology
    pump = local_pump(topo, peak_rate=1200.0, width=0.3, center=(0.5, 0.5), pump_steps=4)
    phi = PhiASE(crossSections=w.spectra, **{**w.settings, "rngSeed": 5}, ase_steps=4)
    sim = Simulation(gain_medium=w.medium, phi_ase=phi, time_integrator=RungeKutta4(), time_step_size=1e-4)
    states = {}
    sim.add_pump(pump).on_step(lambda st: states.setdefault(st.step, st.betaVolume.copy()))
    sim.on_step(vtkWedge(tmp_path / "beta_{step:02d}", topo, field=["betaVolume", "phiAse"], every=2))
    print("[vtk] Simulation on_step vtkWedge, 4 RK4 steps", flush=True)
    sim.step(4)
    sim.close()
    assert sorted(p.name for p in tmp_path.glob("*.vtk")) == ["beta_02.vtk", "beta_04.vtk"]
    for step in (2, 4):
        got = readVtkScalars(tmp_path / f"beta_{step:02d}.vtk")
        assert got["betaVolume"][0] == "CELL_DATA" and got["phiAse"][0] == "CELL_DATA"
        np.testing.assert_array_equal(got["betaVolume"][1], states[step].reshape(-1, order="F"))

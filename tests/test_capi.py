"""The C-ABI library: it loads, exports every declared symbol, its struct
layouts match include/hase_ase.h, and host-side validation rejects bad
arguments before any device work (no compute calls here)."""
import ctypes as C
import json
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from haseongpu_amd import _native
from haseongpu_amd.workloads import plain_small

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "hase_ase.h"


def _declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(hase_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    declared = _declared_functions()
    assert len(declared) >= 10
    assert set(declared) == set(_native.EXPORTED_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.hase_abi_version() == 2
    assert b"gfx950" in lib.hase_build_info()


def test_struct_layouts_match_header(tmp_path):
    structs = {"hase_mesh_t": _native.Mesh, "hase_experiment_t": _native.Experiment,
               "hase_compute_t": _native.Compute, "hase_result_t": _native.ResultC, "hase_stats_t": _native.Stats,
               "hase_train_t": _native.Train, "hase_proposal_t": _native.Proposal,
               "hase_proposal_stats_t": _native.ProposalStats, "hase_tuning_t": _native.Tuning,
               "hase_run_control_t": _native.RunControl, "hase_time_state_t": _native.TimeState,
               "hase_tet_mesh_t": _native.TetMeshC, "hase_forward_result_t": _native.ForwardResultC,
               "hase_forward_stats_t": _native.ForwardStats,
               "hase_forward_reflections_t": _native.ForwardReflections,
               "hase_tet4_host_mesh_t": _native.Tet4HostMesh, "hase_prism_arrays_t": _native.PrismArrays}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){", 'printf("{");']
    first = True
    for cname, cls in structs.items():
        for field, _ in cls._fields_:
            sep = "" if first else ","
            first = False
            lines.append(f'printf("{sep}\\"{cname}.{field}\\": %zu", offsetof({cname}, {field}));')
        lines.append(f'printf(",\\"{cname}\\": %zu", sizeof({cname}));')
    lines += ['printf("}");', "return 0;}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    layout = json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    for cname, cls in structs.items():
        assert layout[cname] == C.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert layout[f"{cname}.{field}"] == getattr(cls, field).offset, f"{cname}.{field}"


def _args(**overrides):
    w = plain_small(rays=1000)
    mesh = _native.NativeMesh(w.topology, w.medium)
    exp = _native.NativeExperiment(sigma_a=[1.16e-21], sigma_e=[2.48e-20], min_rays=1000, max_rays=1000)
    comp = _native.make_compute(n_samples=w.samples)
    for key, value in overrides.items():
        for obj in (mesh.struct, exp.struct, comp):
            if any(f == key for f, _ in obj._fields_):
                setattr(obj, key, value)
    return w, mesh, exp, comp


@pytest.mark.parametrize("overrides,message", [
    ({"min_rays_per_sample": 0}, "min_rays_per_sample"),
    ({"max_rays_per_sample": 10}, "max_rays_per_sample"),
    ({"number_of_levels": 1}, "number_of_levels"),
    ({"mse_threshold": -1.0}, "mse_threshold"),
    ({"max_sample_range": 10**6}, "sample range"),
    ({"defensive_fraction": 1.0}, "defensive_fraction"),
    ({"max_repetitions": 0}, "max_repetitions"),
    ({"n_tot": 0.0}, "n_tot"),
])
def test_invalid_arguments_are_rejected_on_the_host(overrides, message):
    lib = _native.load()
    w, mesh, exp, comp = _args(**overrides)
    res = _native.ResultArrays(w.samples)
    code = lib.hase_calc_phi_ase(C.byref(mesh.struct), C.byref(exp.struct), C.byref(comp), C.byref(res.struct))
    assert code == -1
    assert message in lib.hase_last_error().decode()
    # the allocation-free entry checks the same preconditions before any device call
    code = lib.hase_calc_phi_ase_noalloc(C.byref(mesh.struct), C.byref(exp.struct), C.byref(comp),
                                         C.byref(res.struct))
    assert code == -1
    assert message in lib.hase_last_error().decode()


def test_null_arguments_are_rejected():
    lib = _native.load()
    assert lib.hase_calc_phi_ase(None, None, None, None) == -1
    assert b"NULL" in lib.hase_last_error()
    assert lib.hase_calc_phi_ase_noalloc(None, None, None, None) == -1
    assert b"NULL" in lib.hase_last_error()
    assert lib.hase_context_evaluate(None, None) == -1
    assert lib.hase_context_stats(None, None) == -1
    lib.hase_context_destroy(None)
    with pytest.raises(_native.HaseError):
        _native.check(-1)


def test_network_layout_constants():
    lib = _native.load()
    # W1 b1 (10 -> 128), three 128 x 128 layers, w5 b5
    assert lib.hase_mlp_param_count() == 128 * 10 + 128 + 3 * (128 * 128 + 128) + 128 + 1
    text = HEADER.read_text()
    assert f"#define HASE_MLP_INPUTS {_native.MLP_INPUTS}" in text
    assert f"#define HASE_MLP_HIDDEN {_native.MLP_HIDDEN}" in text


def test_network_entry_points_reject_bad_arguments_on_the_host():
    lib = _native.load()
    p = np.zeros(lib.hase_mlp_param_count(), np.float32)
    pp = p.ctypes.data_as(C.POINTER(C.c_float))
    y = np.zeros(4, np.float32)
    yp = y.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.hase_mlp_forward(pp, 4, None, yp, 0) == -1
    assert lib.hase_mlp_forward_bf16(pp, 0, yp, yp, 0) == -1
    assert lib.hase_mlp_loss_gradient(pp, 4, yp, None, None, None, None, 0) == -1
    cfg = _native.Train(0, 16, 1e-3, 0.9, 0.999, 1e-8, 0.0, 0)
    x = np.zeros((4, 10), np.float32)
    assert lib.hase_mlp_train(pp, 4, x.ctypes.data_as(C.POINTER(C.c_float)), yp, None, C.byref(cfg), None, 0) == -1
    assert b"epochs" in lib.hase_last_error()
    assert lib.hase_context_set_proposal(None, None) == -1
    assert lib.hase_context_transfer_map(None, 0, 1, yp, 4) == -1
    assert lib.hase_context_proposal_logits(None, 0, 1, yp, 4, None, None) == -1
    assert lib.hase_context_proposal_stats(None, None) == -1


def test_tuning_entry_points_reject_bad_arguments_on_the_host():
    lib = _native.load()
    t = _native.Tuning(0, 256, 1024, 2, 16, 8, 1, 0, 0.0)
    assert lib.hase_context_set_tuning(None, C.byref(t)) == -1
    assert lib.hase_context_get_tuning(None, C.byref(t)) == -1
    assert lib.hase_context_autotune(None, 0, C.byref(t)) == -1
    text = HEADER.read_text()
    assert "#define HASE_TRACE_QUEUE 0" in text and "#define HASE_TRACE_ROUNDS 1" in text


def test_batch_entry_points_reject_bad_arguments_on_the_host():
    lib = _native.load()
    out = C.c_void_p()
    assert lib.hase_batch_create(0, None, C.byref(out)) == -1
    assert lib.hase_batch_create(1, None, None) == -1
    assert lib.hase_batch_evaluate(None, None) == -1
    assert lib.hase_batch_stats(None, None) == -1
    lib.hase_batch_destroy(None)


def test_time_loop_entry_points_reject_bad_arguments_on_the_host():
    lib = _native.load()
    out = C.c_void_p()
    run = _native.RunControl(1e-5, 0, 8, 0, 1, 1, 0)
    rate = np.zeros(4)
    rp = rate.ctypes.data_as(C.POINTER(C.c_double))
    assert lib.hase_simulation_create(None, C.byref(run), None, rp, None, C.byref(out)) == -1
    assert b"context" in lib.hase_last_error()
    assert lib.hase_simulation_create(None, C.byref(run), None, rp, None, None) == -1
    assert lib.hase_simulation_set_pump(None, rp, None) == -1
    assert lib.hase_simulation_set_beta(None, rp) == -1
    assert lib.hase_simulation_step(None, 1) == -1
    assert lib.hase_simulation_state(None, None) == -1
    assert lib.hase_simulation_times(None, None, None, None) == -1
    lib.hase_simulation_destroy(None)
    text = HEADER.read_text()
    for name, code in _native.INTEGRATORS.items():
        macro = "HASE_INTEGRATOR_" + name.upper().replace("-", "_")
        assert f"#define {macro} {code}" in text, macro


def test_time_loop_front_end_validation():
    from haseongpu_amd import ExplicitEuler, LocalPump, PhiASE, Simulation
    from haseongpu_amd.workloads import plain_small

    w = plain_small(rays=1000)
    phi = PhiASE(crossSections=w.spectra, **w.settings)
    ok = dict(gain_medium=w.medium, phi_ase=phi, time_integrator=ExplicitEuler(), time_step_size=1e-5)
    Simulation(**ok)
    for bad, msg in [({"time_integrator": None}, "time_integrator"), ({"time_step_size": 0.0}, "positive"),
                     ({"time_integrator": "leapfrog"}, "unsupported"),
                     ({"simulation_steps": 2, "max_time": 1.0}, "either"),
                     ({"output_steps": (2, 1)}, "increasing"), ({"output_fields": ("beta",)}, "unsupported"),
                     ({"control_fields": ("beta_volume",)}, "synchronized-debug"),
                     ({"execution_mode": "synchronized-debug", "output_steps": (1,)}, "output_steps")]:
        with pytest.raises(ValueError, match=msg):
            Simulation(**{**ok, **bad})
    sim = Simulation(**ok)
    with pytest.raises(TypeError):
        sim.add_pump(object())
    with pytest.raises(ValueError, match="simulation_steps is required"):
        sim.step()
    with pytest.raises(ValueError, match="one per prism"):
        LocalPump(rate=np.zeros(3)).arrays(w.topology.numberOfPrisms)
    with pytest.raises(TypeError):
        PhiASE(crossSections=w.spectra, ase_steps=1.5)


def test_forward_host_rejections():
    """The forward entry points validate on the host before touching a device."""
    import numpy as np

    from haseongpu_amd.tet_mesh import TetMesh

    pts = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    mesh = TetMesh(points=pts, cells=[[0, 1, 2, 3]], beta=[0.1])
    comp = _native.make_compute(n_samples=1)
    ok = _native.NativeExperiment(sigma_a=[1e-21], sigma_e=[2e-20], min_rays=10, max_rays=10)
    bad = _native.NativeTetMesh(mesh)
    bad.cells[3] = 9
    with pytest.raises(_native.HaseError, match="out of range"):
        _native.ForwardContext(bad, ok, comp)
    zero = _native.NativeExperiment(sigma_a=[1e-21], sigma_e=[2e-20], min_rays=0, max_rays=10)
    with pytest.raises(_native.HaseError, match="min rays"):
        _native.ForwardContext(mesh, zero, comp)
    with pytest.raises(ValueError):
        TetMesh(points=pts, cells=[[0, 1, 2, 4]], beta=[0.1])

"""ctypes binding of the HIP library ``lib/libhase_ase.so`` (include/hase_ase.h).

The library is loaded from the package's ``lib/`` folder; there is no other
implementation of the estimator behind this module, so a missing or broken
library raises instead of falling back.  ctypes loading follows the
reference's own C-ABI loader (pyInclude/alpakaUtils/AlpakaBackends.py:9-53).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"
LIB_PATH = LIB_DIR / "libhase_ase.so"

HASE_OK = 0
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)

# every symbol include/hase_ase.h declares
EXPORTED_SYMBOLS = (
    "hase_calc_phi_ase",
    "hase_calc_phi_ase_noalloc",
    "hase_context_create",
    "hase_context_update_beta",
    "hase_context_evaluate",
    "hase_context_stats",
    "hase_context_destroy",
    "hase_context_trace_rays",
    "hase_context_importance",
    "hase_mlp_param_count",
    "hase_mlp_forward",
    "hase_mlp_loss_gradient",
    "hase_mlp_train",
    "hase_mlp_forward_bf16",
    "hase_context_set_proposal",
    "hase_context_transfer_map",
    "hase_context_proposal_logits",
    "hase_context_proposal_stats",
    "hase_context_set_tuning",
    "hase_context_get_tuning",
    "hase_context_autotune",
    "hase_batch_create",
    "hase_batch_evaluate",
    "hase_batch_stats",
    "hase_batch_destroy",
    "hase_simulation_create",
    "hase_simulation_set_pump",
    "hase_simulation_set_beta",
    "hase_simulation_step",
    "hase_simulation_state",
    "hase_simulation_times",
    "hase_simulation_destroy",
    "hase_forward_create",
    "hase_forward_update_beta",
    "hase_forward_evaluate",
    "hase_forward_trace",
    "hase_forward_stats",
    "hase_forward_set_reflections",
    "hase_forward_destroy",
    "hase_calc_forward_phi_ase",
    "hase_fold_tet4_mesh",
    "hase_last_error",
    "hase_abi_version",
    "hase_build_info",
)


class HaseError(RuntimeError):
    """Error returned by the HIP library (negative status code)."""

    def __init__(self, code, message):
        super().__init__(f"hase_ase error {code}: {message}")
        self.code = code


class Mesh(C.Structure):
    _fields_ = [
        ("number_of_points", C.c_uint32),
        ("number_of_triangles", C.c_uint32),
        ("number_of_levels", C.c_uint32),
        ("thickness", C.c_double),
        ("points", _f64p),
        ("triangle_point_indices", _u32p),
        ("triangle_neighbors", _i32p),
        ("forbidden_edge", _i32p),
        ("beta_volume", _f64p),
        ("beta_cells", _f64p),
        ("cladding_cell_types", _u32p),
        ("cladding_number", C.c_uint32),
        ("cladding_absorption", C.c_double),
        ("n_tot", C.c_double),
        ("crystal_t_fluo", C.c_double),
        ("reflectivities", _f32p),
        ("refractive_indices", _f32p),
    ]


class Experiment(C.Structure):
    _fields_ = [
        ("min_rays_per_sample", C.c_uint32),
        ("max_rays_per_sample", C.c_uint32),
        ("number_of_wavelengths", C.c_uint32),
        ("sigma_a", _f64p),
        ("sigma_e", _f64p),
        ("max_sigma_a", C.c_double),
        ("max_sigma_e", C.c_double),
        ("mse_threshold", C.c_double),
        ("use_reflections", C.c_int32),
        ("max_reflections", C.c_uint32),
        ("defensive_fraction", C.c_double),
    ]


class Compute(C.Structure):
    _fields_ = [
        ("max_repetitions", C.c_uint32),
        ("adaptive_steps", C.c_uint32),
        ("min_sample_range", C.c_uint32),
        ("max_sample_range", C.c_uint32),
        ("rng_seed", C.c_uint32),
        ("device", C.c_int32),
        ("rays_per_thread", C.c_uint32),
        ("block_size", C.c_uint32),
        ("sample_batch", C.c_uint32),
    ]


class ResultC(C.Structure):
    _fields_ = [
        ("phi_ase", _f32p),
        ("mse", _f64p),
        ("total_rays", _u32p),
        ("dndt_ase", _f64p),
        ("standard_error", _f64p),
        ("dropped_rays", _u32p),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("importance_ms", C.c_double),
        ("trace_ms", C.c_double),
        ("finalize_ms", C.c_double),
        ("total_ms", C.c_double),
        ("rays", C.c_uint64),
        ("ray_steps", C.c_uint64),
        ("importance_walks", C.c_uint64),
        ("importance_steps", C.c_uint64),
        ("dropped_rays", C.c_uint64),
        ("trace_launches", C.c_uint32),
        ("importance_launches", C.c_uint32),
    ]

    def asdict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Train(C.Structure):
    _fields_ = [
        ("epochs", C.c_uint32),
        ("batch", C.c_uint32),
        ("learning_rate", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("epsilon", C.c_float),
        ("weight_decay", C.c_float),
        ("seed", C.c_uint32),
    ]


class Proposal(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("params", _f32p),
        ("input_shift", _f32p),
        ("input_scale", _f32p),
        ("cluster_triangles", C.c_uint32),
        ("cluster_levels", C.c_uint32),
        ("fallback", C.c_int32),
    ]


class ProposalStats(C.Structure):
    _fields_ = [
        ("inference_ms", C.c_double),
        ("proposal_ms", C.c_double),
        ("evaluations", C.c_uint64),
        ("fallback_samples", C.c_uint32),
        ("clusters", C.c_uint32),
    ]

    def asdict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [
        ("mode", C.c_uint32),
        ("block_size", C.c_uint32),
        ("rays_per_item", C.c_uint32),
        ("queue_chunk", C.c_uint32),
        ("refill_lanes", C.c_uint32),
        ("step_budget", C.c_uint32),
        ("use_graph", C.c_uint32),
        ("stage_lds", C.c_uint32),
        ("probe_ms", C.c_double),
    ]

    def asdict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class RunControl(C.Structure):
    _fields_ = [
        ("time_step", C.c_double),
        ("integrator", C.c_uint32),
        ("implicit_iterations", C.c_uint32),
        ("first_simulation_step", C.c_uint32),
        ("ase_steps", C.c_uint32),
        ("pump_steps", C.c_uint32),
        ("pre_pump", C.c_uint32),
    ]


class TimeState(C.Structure):
    _fields_ = [
        ("step", C.c_uint32),
        ("ase_evaluations", C.c_uint32),
        ("time", C.c_double),
        ("beta_volume", _f64p),
        ("phi_ase", _f64p),
        ("dndt_ase", _f64p),
        ("dndt_pump", _f64p),
        ("phi_samples", _f32p),
        ("mse", _f64p),
        ("total_rays", _u32p),
        ("standard_error", _f64p),
    ]


# integrator codes of include/hase_ase.h, keyed by the reference's names
# (include/core/simulationRunControl.hpp:65-74)
INTEGRATORS = {
    "explicit-euler": 0,
    "heun": 1,
    "midpoint": 2,
    "runge-kutta-4": 3,
    "frozen-phi-ase-runge-kutta-4": 4,
    "implicit-euler": 5,
    "exponential-euler": 6,
}

TRACE_QUEUE = 0
TRACE_ROUNDS = 1
MLP_INPUTS = 10
MLP_HIDDEN = 128
PROPOSAL_PLAIN = 0
PROPOSAL_LEARNED = 1
PROPOSAL_BETA_VOLUME = 2

_lib = None


class TetMeshC(C.Structure):
    _fields_ = [
        ("number_of_points", C.c_uint32), ("number_of_cells", C.c_uint32), ("points", _f64p),
        ("cells", _u32p), ("beta_volume", _f64p), ("cladding_cell_types", _u32p),
        ("cladding_number", C.c_uint32), ("reserved", C.c_uint32), ("cladding_absorption", C.c_double),
        ("n_tot", C.c_double), ("crystal_t_fluo", C.c_double),
    ]


class ForwardResultC(C.Structure):
    _fields_ = [
        ("phi_ase", _f32p), ("standard_error", _f64p), ("relative_standard_error", _f64p), ("dndt_ase", _f64p),
        ("ray_visits", _u32p), ("dropped_rays", _u32p), ("convergence_rays", _u32p),
    ]


class ForwardStats(C.Structure):
    _fields_ = [
        ("trace_ms", C.c_double), ("total_ms", C.c_double), ("rays", C.c_uint64), ("segments", C.c_uint64),
        ("dropped_rays", C.c_uint64), ("launches", C.c_uint32), ("open_cells", C.c_uint32),
        ("srm_status", C.c_uint32), ("srm_passes", C.c_uint32), ("srm_remaining", C.c_double),
    ]


class ForwardReflections(C.Structure):
    _fields_ = [
        ("face_boundaries", _i32p), ("number_of_surfaces", C.c_uint32), ("reservoir_slots", C.c_uint32),
        ("reflectivity", _f32p), ("n_inside", _f32p), ("n_outside", _f32p), ("max_passes", C.c_uint32),
        ("divergence_streak", C.c_uint32), ("tolerance", C.c_double),
    ]


SRM_STATUS = {0: "DISABLED", 1: "CONVERGED", 2: "STABLE", 3: "MAX_PASSES", 4: "DIVERGED"}


class Tet4HostMesh(C.Structure):
    _fields_ = [
        ("number_of_mesh_points", C.c_uint32),
        ("number_of_cells", C.c_uint32),
        ("number_of_points", C.c_uint32),
        ("number_of_levels", C.c_uint32),
        ("points", _f64p),
        ("cell_point_indices", _u32p),
        ("beta_volume", _f64p),
        ("cladding_cell_types", _u32p),
        ("cell_face_boundaries", _i32p),
        ("number_of_surfaces", C.c_uint32),
        ("surface_reflectivities", _f32p),
        ("surface_refractive_index_inside", _f32p),
        ("surface_refractive_index_outside", _f32p),
        ("number_of_reflectivities", C.c_uint32),
        ("reflectivities", _f32p),
        ("refractive_indices", _f32p),
    ]


class PrismArrays(C.Structure):
    _fields_ = [
        ("points", _f64p),
        ("triangle_point_indices", _u32p),
        ("triangle_neighbors", _i32p),
        ("forbidden_edge", _i32p),
        ("beta_volume", _f64p),
        ("cladding_cell_types", _u32p),
        ("reflectivities", _f32p),
        ("refractive_indices", C.c_float * 4),
        ("thickness", C.c_double),
        ("number_of_triangles", C.c_uint32),
    ]


def load():
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(hipcc --offload-arch=gfx950)")
    L = C.CDLL(str(LIB_PATH))
    ctx_p = C.c_void_p
    L.hase_calc_phi_ase.argtypes = [C.POINTER(Mesh), C.POINTER(Experiment), C.POINTER(Compute), C.POINTER(ResultC)]
    L.hase_calc_phi_ase_noalloc.argtypes = [C.POINTER(Mesh), C.POINTER(Experiment), C.POINTER(Compute),
                                            C.POINTER(ResultC)]
    L.hase_context_create.argtypes = [C.POINTER(Mesh), C.POINTER(Experiment), C.POINTER(Compute),
                                      C.POINTER(ctx_p)]
    L.hase_context_update_beta.argtypes = [ctx_p, _f64p, _f64p]
    L.hase_context_evaluate.argtypes = [ctx_p, C.POINTER(ResultC)]
    L.hase_context_stats.argtypes = [ctx_p, C.POINTER(Stats)]
    L.hase_context_destroy.argtypes = [ctx_p]
    L.hase_context_destroy.restype = None
    L.hase_context_trace_rays.argtypes = [ctx_p, C.c_uint32, _f32p, _u32p, _f32p, _u32p, _u32p,
                                          _f32p, _f32p, _f32p, _f32p, _u32p, _i32p]
    L.hase_context_importance.argtypes = [ctx_p, C.c_uint32, _f32p, C.c_uint32]
    L.hase_mlp_param_count.argtypes = []
    L.hase_mlp_param_count.restype = C.c_uint32
    L.hase_mlp_forward.argtypes = [_f32p, C.c_uint32, _f32p, _f32p, C.c_int32]
    L.hase_mlp_loss_gradient.argtypes = [_f32p, C.c_uint32, _f32p, _f32p, _f32p, _f64p, _f32p, C.c_int32]
    L.hase_mlp_train.argtypes = [_f32p, C.c_uint32, _f32p, _f32p, _f32p, C.POINTER(Train), _f32p, C.c_int32]
    L.hase_mlp_forward_bf16.argtypes = [_f32p, C.c_uint32, _f32p, _f32p, C.c_int32]
    L.hase_context_set_proposal.argtypes = [ctx_p, C.POINTER(Proposal)]
    L.hase_context_transfer_map.argtypes = [ctx_p, C.c_uint32, C.c_uint32, _f32p, C.c_size_t]
    L.hase_context_proposal_logits.argtypes = [ctx_p, C.c_uint32, C.c_uint32, _f32p, C.c_size_t, _f32p, _u32p]
    L.hase_context_proposal_stats.argtypes = [ctx_p, C.POINTER(ProposalStats)]
    L.hase_context_set_tuning.argtypes = [ctx_p, C.POINTER(Tuning)]
    L.hase_context_get_tuning.argtypes = [ctx_p, C.POINTER(Tuning)]
    L.hase_context_autotune.argtypes = [ctx_p, C.c_uint32, C.POINTER(Tuning)]
    L.hase_batch_create.argtypes = [C.c_uint32, C.POINTER(ctx_p), C.POINTER(ctx_p)]
    L.hase_batch_evaluate.argtypes = [ctx_p, C.POINTER(C.POINTER(ResultC))]
    L.hase_batch_stats.argtypes = [ctx_p, C.POINTER(Stats)]
    L.hase_batch_destroy.argtypes = [ctx_p]
    L.hase_batch_destroy.restype = None
    L.hase_simulation_create.argtypes = [ctx_p, C.POINTER(RunControl), _f64p, _f64p, _f64p, C.POINTER(ctx_p)]
    L.hase_simulation_set_pump.argtypes = [ctx_p, _f64p, _f64p]
    L.hase_simulation_set_beta.argtypes = [ctx_p, _f64p]
    L.hase_simulation_step.argtypes = [ctx_p, C.c_uint32]
    L.hase_simulation_state.argtypes = [ctx_p, C.POINTER(TimeState)]
    L.hase_simulation_times.argtypes = [ctx_p, _f64p, _f64p, C.POINTER(C.c_uint64)]
    L.hase_simulation_destroy.argtypes = [ctx_p]
    L.hase_simulation_destroy.restype = None
    L.hase_forward_create.argtypes = [C.POINTER(TetMeshC), C.POINTER(Experiment), C.POINTER(Compute),
                                      C.POINTER(ctx_p)]
    L.hase_forward_update_beta.argtypes = [ctx_p, _f64p]
    L.hase_forward_evaluate.argtypes = [ctx_p, C.POINTER(ForwardResultC)]
    L.hase_forward_trace.argtypes = [ctx_p, C.c_uint32, C.c_uint32, C.c_uint32, _f64p, _u32p, _u32p,
                                     C.POINTER(C.c_uint64)]
    L.hase_forward_stats.argtypes = [ctx_p, C.POINTER(ForwardStats)]
    L.hase_forward_set_reflections.argtypes = [ctx_p, C.POINTER(ForwardReflections)]
    L.hase_forward_destroy.argtypes = [ctx_p]
    L.hase_calc_forward_phi_ase.argtypes = [C.POINTER(TetMeshC), C.POINTER(Experiment), C.POINTER(Compute),
                                            C.POINTER(ForwardResultC)]
    L.hase_fold_tet4_mesh.argtypes = [C.POINTER(Tet4HostMesh), C.POINTER(PrismArrays)]
    L.hase_last_error.restype = C.c_char_p
    L.hase_abi_version.restype = C.c_int
    L.hase_build_info.restype = C.c_char_p
    for name in ("hase_calc_phi_ase", "hase_calc_phi_ase_noalloc", "hase_context_create", "hase_context_update_beta",
                 "hase_context_evaluate", "hase_context_stats", "hase_context_trace_rays",
                              "hase_context_importance", "hase_mlp_forward", "hase_mlp_loss_gradient", "hase_mlp_train",
                 "hase_mlp_forward_bf16", "hase_context_set_proposal", "hase_context_transfer_map",
                 "hase_context_proposal_logits", "hase_context_proposal_stats", "hase_context_set_tuning",
                 "hase_context_get_tuning", "hase_context_autotune", "hase_batch_create",
                 "hase_batch_evaluate", "hase_batch_stats", "hase_simulation_create", "hase_simulation_set_pump",
                 "hase_simulation_set_beta", "hase_simulation_step", "hase_simulation_state",
                 "hase_simulation_times", "hase_forward_create", "hase_forward_update_beta",
                 "hase_forward_evaluate", "hase_forward_trace", "hase_forward_stats", "hase_forward_destroy",
                 "hase_forward_set_reflections",
                 "hase_calc_forward_phi_ase", "hase_fold_tet4_mesh"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(code):
    if code != HASE_OK:
        raise HaseError(code, load().hase_last_error().decode())


def _ptr(arr, ctype):
    return None if arr is None else arr.ctypes.data_as(C.POINTER(ctype))


class NativeMesh:
    """Owns the arrays a ``hase_mesh_t`` points into."""

    def __init__(self, topology, medium):
        d = topology.topology()
        self.points = np.ascontiguousarray(np.concatenate([topology.points[:, 0], topology.points[:, 1]]),
                                           dtype=np.float64)
        self.tri = np.ascontiguousarray(topology.trianglePointIndices.reshape(-1, order="F"), dtype=np.uint32)
        self.nbr = np.ascontiguousarray(d["triangleNeighbors"], dtype=np.int32)
        self.fe = np.ascontiguousarray(d["forbiddenEdge"], dtype=np.int32)
        self.beta = np.ascontiguousarray(medium.value("betaVolume"), dtype=np.float64)
        self.beta_cells = np.ascontiguousarray(medium.value("betaCells"), dtype=np.float64)
        self.clad = np.ascontiguousarray(medium.value("claddingCellTypes"), dtype=np.uint32)
        self.refl = np.ascontiguousarray(medium.value("reflectivities"), dtype=np.float32)
        self.nidx = np.ascontiguousarray(medium.value("refractiveIndices"), dtype=np.float32)
        self.struct = Mesh(
            topology.numberOfPoints, topology.numberOfTriangles, int(topology.levels), float(topology.thickness),
            _ptr(self.points, C.c_double), _ptr(self.tri, C.c_uint32), _ptr(self.nbr, C.c_int32),
            _ptr(self.fe, C.c_int32), _ptr(self.beta, C.c_double), _ptr(self.beta_cells, C.c_double),
            _ptr(self.clad, C.c_uint32), int(medium.value("claddingNumber")),
            float(medium.value("claddingAbsorption")), float(medium.value("nTot")),
            float(medium.value("crystalTFluo")), _ptr(self.refl, C.c_float), _ptr(self.nidx, C.c_float))


    @classmethod
    def fromFolded(cls, folded, *, beta_cells, number_of_levels, cladding_number, cladding_absorption, n_tot,
                   crystal_t_fluo):
        """hase_mesh_t straight from the arrays of ``fold_tet4_mesh`` (no
        topology recomputed on the host)."""
        self = cls.__new__(cls)
        self.points = np.ascontiguousarray(folded["points"], dtype=np.float64)
        self.tri = np.ascontiguousarray(folded["triangles"], dtype=np.uint32)
        self.nbr = np.ascontiguousarray(folded["neighbors"], dtype=np.int32)
        self.fe = np.ascontiguousarray(folded["forbidden"], dtype=np.int32)
        self.beta = np.ascontiguousarray(folded["beta"], dtype=np.float64)
        self.beta_cells = np.ascontiguousarray(beta_cells, dtype=np.float64)
        self.clad = np.ascontiguousarray(folded["cladding"], dtype=np.uint32)
        self.refl = np.ascontiguousarray(folded["reflectivities"], dtype=np.float32)
        self.nidx = np.ascontiguousarray(folded["refractive_indices"], dtype=np.float32)
        self.struct = Mesh(
            self.points.size // 2, int(folded["number_of_triangles"]), int(number_of_levels),
            float(folded["thickness"]), _ptr(self.points, C.c_double), _ptr(self.tri, C.c_uint32),
            _ptr(self.nbr, C.c_int32), _ptr(self.fe, C.c_int32), _ptr(self.beta, C.c_double),
            _ptr(self.beta_cells, C.c_double), _ptr(self.clad, C.c_uint32), int(cladding_number),
            float(cladding_absorption), float(n_tot), float(crystal_t_fluo), _ptr(self.refl, C.c_float),
            _ptr(self.nidx, C.c_float))
        return self


class NativeExperiment:
    def __init__(self, *, sigma_a, sigma_e, min_rays, max_rays, max_sigma_a=None, max_sigma_e=None,
                 mse_threshold=0.005, use_reflections=False, max_reflections=1, defensive_fraction=0.05):
        self.sa = np.ascontiguousarray(np.atleast_1d(sigma_a), dtype=np.float64)
        self.se = np.ascontiguousarray(np.atleast_1d(sigma_e), dtype=np.float64)
        if self.sa.size != self.se.size or self.sa.size == 0:
            raise ValueError("sigma_a and sigma_e must be non-empty and of equal length")
        msa = float(self.sa.max()) if max_sigma_a is None else float(max_sigma_a)
        mse_ = float(self.se.max()) if max_sigma_e is None else float(max_sigma_e)
        self.struct = Experiment(int(min_rays), int(max_rays), self.sa.size, _ptr(self.sa, C.c_double),
                                 _ptr(self.se, C.c_double), msa, mse_, float(mse_threshold),
                                 int(bool(use_reflections)), int(max_reflections), float(defensive_fraction))


def make_compute(*, n_samples, max_repetitions=1, adaptive_steps=1, min_sample_range=0, max_sample_range=None,
                 rng_seed=5489, device=0, rays_per_thread=0, block_size=0, sample_batch=0):
    last = n_samples - 1 if max_sample_range is None else int(max_sample_range)
    return Compute(int(max_repetitions), int(adaptive_steps), int(min_sample_range), last, int(rng_seed) & 0xFFFFFFFF,
                   int(device), int(rays_per_thread), int(block_size), int(sample_batch))


class ResultArrays:
    def __init__(self, n_samples):
        self.phiAse = np.zeros(n_samples, dtype=np.float32)
        self.mse = np.zeros(n_samples, dtype=np.float64)
        self.totalRays = np.zeros(n_samples, dtype=np.uint32)
        self.dndtAse = np.zeros(n_samples, dtype=np.float64)
        self.standardError = np.zeros(n_samples, dtype=np.float64)
        self.droppedRays = np.zeros(n_samples, dtype=np.uint32)
        self.struct = ResultC(_ptr(self.phiAse, C.c_float), _ptr(self.mse, C.c_double),
                              _ptr(self.totalRays, C.c_uint32), _ptr(self.dndtAse, C.c_double),
                              _ptr(self.standardError, C.c_double), _ptr(self.droppedRays, C.c_uint32))


class Context:
    """Persistent device context (mesh resident between evaluations)."""

    def __init__(self, mesh: NativeMesh, experiment: NativeExperiment, compute: Compute):
        L = load()
        self._mesh, self._exp, self._comp = mesh, experiment, compute
        self.n_samples = mesh.struct.number_of_points * mesh.struct.number_of_levels
        handle = C.c_void_p()
        check(L.hase_context_create(C.byref(mesh.struct), C.byref(experiment.struct), C.byref(compute),
                                    C.byref(handle)))
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            load().hase_context_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def evaluate(self):
        res = ResultArrays(self.n_samples)
        check(load().hase_context_evaluate(self._h, C.byref(res.struct)))
        return res

    def update_beta(self, beta_volume, beta_cells=None):
        bv = np.ascontiguousarray(beta_volume, dtype=np.float64)
        bc = None if beta_cells is None else np.ascontiguousarray(beta_cells, dtype=np.float64)
        check(load().hase_context_update_beta(self._h, _ptr(bv, C.c_double), _ptr(bc, C.c_double)))

    def stats(self):
        st = Stats()
        check(load().hase_context_stats(self._h, C.byref(st)))
        return st.asdict()

    def trace_rays(self, start, prism, end, reflections=None, first_plane=None):
        start = np.ascontiguousarray(start, dtype=np.float32).reshape(-1)
        end = np.ascontiguousarray(end, dtype=np.float32).reshape(-1)
        prism = np.ascontiguousarray(prism, dtype=np.uint32)
        n = prism.size
        nr = np.zeros(n, np.uint32) if reflections is None else np.ascontiguousarray(reflections, dtype=np.uint32)
        fp = np.zeros(n, np.uint32) if first_plane is None else np.ascontiguousarray(first_plane, dtype=np.uint32)
        if start.size != 3 * n or end.size != 3 * n or nr.size != n or fp.size != n:
            raise ValueError("start/end must be (n, 3) and prism/reflections/first_plane (n,)")
        out = {k: np.zeros(n, np.float32) for k in ("beta_len", "gain_len", "clad_len", "refl")}
        steps = np.zeros(n, np.uint32)
        status = np.zeros(n, np.int32)
        check(load().hase_context_trace_rays(
            self._h, n, _ptr(start, C.c_float), _ptr(prism, C.c_uint32), _ptr(end, C.c_float),
            _ptr(nr, C.c_uint32), _ptr(fp, C.c_uint32), _ptr(out["beta_len"], C.c_float),
            _ptr(out["gain_len"], C.c_float), _ptr(out["clad_len"], C.c_float), _ptr(out["refl"], C.c_float),
            _ptr(steps, C.c_uint32), _ptr(status, C.c_int32)))
        out["steps"] = steps
        out["status"] = status
        return out

    @property
    def slices(self):
        return 1 + 2 * ((self._exp.struct.max_reflections or 1) if self._exp.struct.use_reflections else 0)

    @property
    def prisms(self):
        return self._mesh.struct.number_of_triangles * (self._mesh.struct.number_of_levels - 1)

    def set_proposal(self, proposal=None):
        """Select the learned proposal (a ``LearnedProposal``) or, with None, the walked map."""
        if proposal is None:
            st = Proposal(PROPOSAL_PLAIN, None, None, None, 1, 1, 0)
            check(load().hase_context_set_proposal(self._h, C.byref(st)))
            self._proposal = None
            return
        params = np.ascontiguousarray(proposal.params, dtype=np.float32)
        shift = np.ascontiguousarray(proposal.shift, dtype=np.float32)
        scale = np.ascontiguousarray(proposal.scale, dtype=np.float32)
        if params.size != mlp_param_count() or shift.size != MLP_INPUTS or scale.size != MLP_INPUTS:
            raise ValueError("proposal arrays do not match the network layout")
        st = Proposal(PROPOSAL_LEARNED, _ptr(params, C.c_float), _ptr(shift, C.c_float), _ptr(scale, C.c_float),
                      int(proposal.cluster_triangles), int(proposal.cluster_levels), int(bool(proposal.fallback)))
        check(load().hase_context_set_proposal(self._h, C.byref(st)))
        self._proposal = (params, shift, scale)

    def set_beta_volume_proposal(self):
        """Rays in proportion to beta * V for every sample (no importance map)."""
        st = Proposal(PROPOSAL_BETA_VOLUME, None, None, None, 1, 1, 0)
        check(load().hase_context_set_proposal(self._h, C.byref(st)))
        self._proposal = None

    def transfer_map(self, first, count):
        """(count, R, P) transfer factors of the centroid walks, caller prism order."""
        R, P = self.slices, self.prisms
        out = np.zeros(count * R * P, dtype=np.float32)
        check(load().hase_context_transfer_map(self._h, int(first), int(count), _ptr(out, C.c_float), out.size))
        return out.reshape(count, R, P)

    def proposal_logits(self, first, count):
        """Network outputs (count, R, C) and the cluster representatives (C, 4)."""
        R = self.slices
        Cn = int(self.proposal_stats()["clusters"])
        if Cn == 0:
            raise RuntimeError("no learned proposal is set")
        n_cl = C.c_uint32(0)
        out = np.zeros(count * R * Cn, dtype=np.float32)
        clusters = np.zeros(Cn * 4, dtype=np.float32)
        check(load().hase_context_proposal_logits(self._h, int(first), int(count), _ptr(out, C.c_float), out.size,
                                                  _ptr(clusters, C.c_float), C.byref(n_cl)))
        return out.reshape(count, R, Cn), clusters.reshape(Cn, 4)

    def tuning(self):
        t = Tuning()
        check(load().hase_context_get_tuning(self._h, C.byref(t)))
        return t.asdict()

    def set_tuning(self, **fields):
        """Change fields of the ray-scheduling settings (see hase_tuning_t)."""
        t = Tuning()
        check(load().hase_context_get_tuning(self._h, C.byref(t)))
        for key, value in fields.items():
            setattr(t, key, value)
        check(load().hase_context_set_tuning(self._h, C.byref(t)))
        return t.asdict()

    def autotune(self, probe_samples=0):
        """Pick the queue kernel's launch shape by timing probes; returns the choice."""
        t = Tuning()
        check(load().hase_context_autotune(self._h, int(probe_samples), C.byref(t)))
        return t.asdict()

    def proposal_stats(self):
        st = ProposalStats()
        check(load().hase_context_proposal_stats(self._h, C.byref(st)))
        return st.asdict()

    def importance(self, sample):
        R = 1 + 2 * ((self._exp.struct.max_reflections or 1) if self._exp.struct.use_reflections else 0)
        P = self._mesh.struct.number_of_triangles * (self._mesh.struct.number_of_levels - 1)
        out = np.zeros(R * P, dtype=np.float32)
        check(load().hase_context_importance(self._h, int(sample), _ptr(out, C.c_float), out.size))
        return out


class Batch:
    """Independent problems (contexts) evaluated together: their samples share
    one persistent work-queue kernel per adaptive round."""

    def __init__(self, contexts):
        L = load()
        self._contexts = list(contexts)
        arr = (C.c_void_p * len(self._contexts))(*[c._h for c in self._contexts])
        handle = C.c_void_p()
        check(L.hase_batch_create(len(self._contexts), arr, C.byref(handle)))
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            load().hase_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def evaluate(self):
        res = [ResultArrays(c.n_samples) for c in self._contexts]
        arr = (C.POINTER(ResultC) * len(res))(*[C.pointer(r.struct) for r in res])
        check(load().hase_batch_evaluate(self._h, arr))
        return res

    def stats(self):
        st = Stats()
        check(load().hase_batch_stats(self._h, C.byref(st)))
        return st.asdict()


class NativeSimulation:
    """Time-stepped beta loop on a context (hase_simulation_*): beta per
    prism stays on the device between steps."""

    def __init__(self, context: Context, *, time_step, integrator="explicit-euler", implicit_iterations=8,
                 first_simulation_step=0, ase_steps=0, pump_steps=0, pre_pump=False, beta_volume=None,
                 pump_rate=None, pump_saturation=None):
        L = load()
        if isinstance(integrator, str):
            if integrator not in INTEGRATORS:
                raise ValueError(f"unsupported time integrator '{integrator}'")
            integrator = INTEGRATORS[integrator]
        self._ctx = context
        self.n_samples = context.n_samples
        self.n_prisms = context.prisms
        self.run = RunControl(float(time_step), int(integrator), int(implicit_iterations),
                              int(first_simulation_step), int(ase_steps), int(pump_steps), int(bool(pre_pump)))
        beta = None if beta_volume is None else self._prism_array(beta_volume, "beta_volume")
        rate = self._prism_array(np.zeros(self.n_prisms) if pump_rate is None else pump_rate, "pump_rate")
        sat = None if pump_saturation is None else self._prism_array(pump_saturation, "pump_saturation")
        handle = C.c_void_p()
        check(L.hase_simulation_create(context._h, C.byref(self.run), _ptr(beta, C.c_double),
                                       _ptr(rate, C.c_double), _ptr(sat, C.c_double), C.byref(handle)))
        self._h = handle

    def _prism_array(self, values, name):
        arr = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1, order="F"))
        if arr.size != self.n_prisms:
            raise ValueError(f"{name} needs {self.n_prisms} values (one per prism), got {arr.size}")
        return arr

    def close(self):
        if getattr(self, "_h", None):
            load().hase_simulation_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_pump(self, pump_rate, pump_saturation=None):
        rate = self._prism_array(pump_rate, "pump_rate")
        sat = None if pump_saturation is None else self._prism_array(pump_saturation, "pump_saturation")
        check(load().hase_simulation_set_pump(self._h, _ptr(rate, C.c_double), _ptr(sat, C.c_double)))

    def set_beta(self, beta_volume):
        beta = self._prism_array(beta_volume, "beta_volume")
        check(load().hase_simulation_set_beta(self._h, _ptr(beta, C.c_double)))

    def step(self, n_steps=1):
        check(load().hase_simulation_step(self._h, int(n_steps)))

    def state(self):
        """dict of the state after the last completed step (prism arrays in
        the caller's order j = t + l * T)."""
        P, S = self.n_prisms, self.n_samples
        out = {"beta_volume": np.zeros(P), "phi_ase": np.zeros(P), "dndt_ase": np.zeros(P),
               "dndt_pump": np.zeros(P), "phi_samples": np.zeros(S, np.float32), "mse": np.zeros(S),
               "total_rays": np.zeros(S, np.uint32), "standard_error": np.zeros(S)}
        st = TimeState(0, 0, 0.0, _ptr(out["beta_volume"], C.c_double), _ptr(out["phi_ase"], C.c_double),
                       _ptr(out["dndt_ase"], C.c_double), _ptr(out["dndt_pump"], C.c_double),
                       _ptr(out["phi_samples"], C.c_float), _ptr(out["mse"], C.c_double),
                       _ptr(out["total_rays"], C.c_uint32), _ptr(out["standard_error"], C.c_double))
        check(load().hase_simulation_state(self._h, C.byref(st)))
        out.update(step=int(st.step), ase_evaluations=int(st.ase_evaluations), time=float(st.time))
        return out

    def times(self):
        a, b, r = C.c_double(), C.c_double(), C.c_uint64()
        check(load().hase_simulation_times(self._h, C.byref(a), C.byref(b), C.byref(r)))
        return {"ase_ms": a.value, "integrator_ms": b.value, "rays": int(r.value)}


def mlp_param_count():
    return int(load().hase_mlp_param_count())


def _rows(inputs):
    x = np.ascontiguousarray(inputs, dtype=np.float32)
    if x.ndim != 2 or x.shape[1] != MLP_INPUTS:
        raise ValueError(f"inputs must be (n, {MLP_INPUTS})")
    return x


def mlp_forward(params, inputs, device=0):
    """FP32 device forward pass of the proposal network on normalised inputs."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = _rows(inputs)
    y = np.zeros(x.shape[0], dtype=np.float32)
    check(load().hase_mlp_forward(_ptr(p, C.c_float), x.shape[0], _ptr(x, C.c_float), _ptr(y, C.c_float), device))
    return y


def mlp_forward_bf16(params, inputs, device=0):
    """BF16-MFMA device inference of the proposal network on normalised inputs."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = _rows(inputs)
    y = np.zeros(x.shape[0], dtype=np.float32)
    check(load().hase_mlp_forward_bf16(_ptr(p, C.c_float), x.shape[0], _ptr(x, C.c_float), _ptr(y, C.c_float),
                                       device))
    return y


def mlp_loss_gradient(params, inputs, targets, weights=None, device=0):
    """Weighted squared error and its gradient from the device backward kernels."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = _rows(inputs)
    t = np.ascontiguousarray(targets, dtype=np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    g = np.zeros_like(p)
    loss = C.c_double(0.0)
    check(load().hase_mlp_loss_gradient(_ptr(p, C.c_float), x.shape[0], _ptr(x, C.c_float), _ptr(t, C.c_float),
                                        _ptr(w, C.c_float), C.byref(loss), _ptr(g, C.c_float), device))
    return loss.value, g


def mlp_train(params, inputs, targets, weights=None, *, epochs=10, batch=4096, learning_rate=1e-3, beta1=0.9,
              beta2=0.999, epsilon=1e-8, weight_decay=0.0, seed=0, device=0):
    """AdamW training on the device; returns (params, per-epoch loss)."""
    p = np.array(params, dtype=np.float32, copy=True)
    x = _rows(inputs)
    t = np.ascontiguousarray(targets, dtype=np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
    hist = np.zeros(int(epochs), dtype=np.float32)
    cfg = Train(int(epochs), int(batch), float(learning_rate), float(beta1), float(beta2), float(epsilon),
                float(weight_decay), int(seed) & 0xFFFFFFFF)
    check(load().hase_mlp_train(_ptr(p, C.c_float), x.shape[0], _ptr(x, C.c_float), _ptr(t, C.c_float),
                                _ptr(w, C.c_float), C.byref(cfg), _ptr(hist, C.c_float), device))
    return p, hist


def calc_phi_ase(mesh: NativeMesh, experiment: NativeExperiment, compute: Compute):
    """One-shot ``hase_calc_phi_ase`` call."""
    n = mesh.struct.number_of_points * mesh.struct.number_of_levels
    res = ResultArrays(n)
    check(load().hase_calc_phi_ase(C.byref(mesh.struct), C.byref(experiment.struct), C.byref(compute),
                                   C.byref(res.struct)))
    return res


def calc_phi_ase_noalloc(mesh: NativeMesh, experiment: NativeExperiment, compute: Compute):
    """One ``hase_calc_phi_ase_noalloc`` call: the one-shot evaluation with
    every device buffer taken from the library's fixed device scratch."""
    n = mesh.struct.number_of_points * mesh.struct.number_of_levels
    res = ResultArrays(n)
    check(load().hase_calc_phi_ase_noalloc(C.byref(mesh.struct), C.byref(experiment.struct), C.byref(compute),
                                           C.byref(res.struct)))
    return res


def fold_tet4_mesh(points, cells, beta_volume, number_of_points, number_of_levels, *, cladding_cell_types=None,
                   cell_face_boundaries=None, surface_reflectivities=None, surface_refractive_index_inside=None,
                   surface_refractive_index_outside=None, reflectivities=None, refractive_indices=None):
    """``hase_fold_tet4_mesh``: prism arrays from the Tet4 arrays of a 2.2.0
    HostMesh (points [3 Nm] SoA, cells [4 C], per-cell beta / types / face
    surface ids).  Returns a dict of the hase_mesh_t arrays: points [2 Np],
    triangles [3T] / neighbours / forbidden edges column-major, beta [P],
    cladding types [T], reflectivities [2T], refractive indices [4],
    thickness and T."""
    Np, L = int(number_of_points), int(number_of_levels)
    keep = []

    def arr(x, dtype):
        if x is None:
            return None
        a = np.ascontiguousarray(np.asarray(x).reshape(-1), dtype=dtype)
        keep.append(a)
        return a

    P3, cl, b = arr(points, np.float64), arr(cells, np.uint32), arr(beta_volume, np.float64)
    ct, fb = arr(cladding_cell_types, np.uint32), arr(cell_face_boundaries, np.int32)
    sr, si, so = (arr(v, np.float32) for v in (surface_reflectivities, surface_refractive_index_inside,
                                                surface_refractive_index_outside))
    lr, ri = arr(reflectivities, np.float32), arr(refractive_indices, np.float32)
    n_surf = max((v.size for v in (sr, si, so) if v is not None), default=0)
    m = Tet4HostMesh(P3.size // 3, cl.size // 4, Np, L, _ptr(P3, C.c_double), _ptr(cl, C.c_uint32),
                     _ptr(b, C.c_double), _ptr(ct, C.c_uint32), _ptr(fb, C.c_int32), n_surf, _ptr(sr, C.c_float),
                     _ptr(si, C.c_float), _ptr(so, C.c_float), 0 if lr is None else lr.size, _ptr(lr, C.c_float),
                     _ptr(ri, C.c_float))
    T = (cl.size // 4) // (3 * max(L - 1, 1))
    out = {"points": np.zeros(2 * Np), "triangles": np.zeros(3 * T, np.uint32),
           "neighbors": np.zeros(3 * T, np.int32), "forbidden": np.zeros(3 * T, np.int32),
           "beta": np.zeros(T * (L - 1)), "cladding": np.zeros(T, np.uint32), "reflectivities": np.zeros(2 * T, np.float32)}
    o = PrismArrays(_ptr(out["points"], C.c_double), _ptr(out["triangles"], C.c_uint32),
                    _ptr(out["neighbors"], C.c_int32), _ptr(out["forbidden"], C.c_int32),
                    _ptr(out["beta"], C.c_double), _ptr(out["cladding"], C.c_uint32),
                    _ptr(out["reflectivities"], C.c_float))
    check(load().hase_fold_tet4_mesh(C.byref(m), C.byref(o)))
    out["refractive_indices"] = np.array(o.refractive_indices[:], dtype=np.float32)
    out["thickness"] = float(o.thickness)
    out["number_of_triangles"] = int(o.number_of_triangles)
    return out


class NativeTetMesh:
    """Owns the arrays a ``hase_tet_mesh_t`` points into (from a ``TetMesh``)."""

    def __init__(self, mesh):
        self.points = np.ascontiguousarray(mesh.points.T.reshape(-1), dtype=np.float64)
        self.cells = np.array(mesh.cells.reshape(-1), dtype=np.uint32)
        self.beta = np.array(mesh.beta, dtype=np.float64)
        self.types = None if mesh.cell_types is None else np.array(mesh.cell_types, dtype=np.uint32)
        self.struct = TetMeshC(mesh.n_points, mesh.n_cells, _ptr(self.points, C.c_double),
                               _ptr(self.cells, C.c_uint32), _ptr(self.beta, C.c_double),
                               _ptr(self.types, C.c_uint32), int(mesh.cladding_number), 0,
                               float(mesh.cladding_absorption), float(mesh.n_tot), float(mesh.tau))


class ForwardResultArrays:
    def __init__(self, n_cells):
        self.phiAse = np.zeros(n_cells, dtype=np.float32)
        self.standardError = np.zeros(n_cells)
        self.relativeStandardError = np.zeros(n_cells)
        self.dndtAse = np.zeros(n_cells)
        self.rayVisits = np.zeros(n_cells, dtype=np.uint32)
        self.droppedRays = np.zeros(n_cells, dtype=np.uint32)
        self.convergenceRays = np.zeros(n_cells, dtype=np.uint32)
        self.struct = ForwardResultC(_ptr(self.phiAse, C.c_float), _ptr(self.standardError, C.c_double),
                                     _ptr(self.relativeStandardError, C.c_double), _ptr(self.dndtAse, C.c_double),
                                     _ptr(self.rayVisits, C.c_uint32), _ptr(self.droppedRays, C.c_uint32),
                                     _ptr(self.convergenceRays, C.c_uint32))


class ForwardContext:
    """Persistent forward Tet4 estimator (``hase_forward_*``)."""

    BATCHES = 8

    def __init__(self, mesh, experiment: NativeExperiment, compute: Compute):
        self._mesh = mesh if isinstance(mesh, NativeTetMesh) else NativeTetMesh(mesh)
        self._exp, self._comp = experiment, compute
        self.n_cells = self._mesh.struct.number_of_cells
        self.n_points = self._mesh.struct.number_of_points
        handle = C.c_void_p()
        check(load().hase_forward_create(C.byref(self._mesh.struct), C.byref(experiment.struct), C.byref(compute),
                                         C.byref(handle)))
        self._h = handle

    def update_beta(self, beta):
        b = np.ascontiguousarray(beta, dtype=np.float64).reshape(-1)
        if b.size != self.n_cells:
            raise ValueError(f"beta needs {self.n_cells} values, got {b.size}")
        check(load().hase_forward_update_beta(self._h, _ptr(b, C.c_double)))

    def evaluate(self):
        res = ForwardResultArrays(self.n_cells)
        check(load().hase_forward_evaluate(self._h, C.byref(res.struct)))
        return res

    def trace(self, seed, launch, rays):
        """Raw accumulators of one launch: scores (8, 2 Np), visits, dropped, segments."""
        scores = np.zeros(self.BATCHES * 2 * self.n_points)
        visits = np.zeros(self.n_cells, np.uint32)
        dropped = np.zeros(self.n_cells, np.uint32)
        seg = C.c_uint64(0)
        check(load().hase_forward_trace(self._h, int(seed) & 0xFFFFFFFF, int(launch), int(rays),
                                        _ptr(scores, C.c_double), _ptr(visits, C.c_uint32),
                                        _ptr(dropped, C.c_uint32), C.byref(seg)))
        return scores.reshape(self.BATCHES, -1), visits, dropped, int(seg.value)

    def set_reflections(self, face_boundaries, reflectivity, n_inside=None, n_outside=None, *, slots=32,
                        max_passes=40, divergence_streak=3, tolerance=1e-4):
        """Reflective surfaces for the surface reservoir passes: face_boundaries
        (C, 4) surface ids (face k opposite vertex k, 0 = none), per-surface
        reflectivity and optional refractive indices (index 0 unused)."""
        fb = np.array(face_boundaries, dtype=np.int32).reshape(-1)
        if fb.size != 4 * self.n_cells:
            raise ValueError(f"face_boundaries needs {4 * self.n_cells} values, got {fb.size}")
        refl = np.array(reflectivity, dtype=np.float32).reshape(-1)
        nin = None if n_inside is None else np.array(n_inside, dtype=np.float32).reshape(-1)
        nout = None if n_outside is None else np.array(n_outside, dtype=np.float32).reshape(-1)
        for name, arr in (("n_inside", nin), ("n_outside", nout)):
            if arr is not None and arr.size != refl.size:
                raise ValueError(f"{name} needs {refl.size} values")
        self._refl_arrays = (fb, refl, nin, nout)
        st = ForwardReflections(_ptr(fb, C.c_int32), refl.size, int(slots), _ptr(refl, C.c_float),
                                _ptr(nin, C.c_float), _ptr(nout, C.c_float), int(max_passes),
                                int(divergence_streak), float(tolerance))
        check(load().hase_forward_set_reflections(self._h, C.byref(st)))

    def stats(self):
        st = ForwardStats()
        check(load().hase_forward_stats(self._h, C.byref(st)))
        out = {name: getattr(st, name) for name, _ in ForwardStats._fields_}
        out["srm_status_name"] = SRM_STATUS.get(out["srm_status"], "?")
        return out

    def close(self):
        if self._h:
            load().hase_forward_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def calc_forward_phi_ase(mesh, experiment: NativeExperiment, compute: Compute):
    """One-shot ``hase_calc_forward_phi_ase`` call."""
    m = mesh if isinstance(mesh, NativeTetMesh) else NativeTetMesh(mesh)
    res = ForwardResultArrays(m.struct.number_of_cells)
    check(load().hase_calc_forward_phi_ase(C.byref(m.struct), C.byref(experiment.struct), C.byref(compute),
                                           C.byref(res.struct)))
    return res

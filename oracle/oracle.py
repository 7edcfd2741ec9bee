"""ctypes front of the FP64 CPU oracle (oracle/hase_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py`` - as the checker, never as the
thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "libhase_oracle.so"
# the same restatement built with FP32 ray geometry and FP64 sums (CPU baseline)
LIB_F32_PATH = HERE / "build" / "libhase_oracle_f32.so"

_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class _Mesh(C.Structure):
    _fields_ = [
        ("np", C.c_uint32), ("nt", C.c_uint32), ("nl", C.c_uint32), ("h", C.c_double),
        ("px", _f64p), ("py", _f64p), ("tri", _u32p), ("nbr", _i32p), ("fe", _i32p),
        ("beta", _f64p), ("clad", _u32p), ("clad_number", C.c_uint32),
        ("clad_alpha", C.c_double), ("refl", _f64p), ("nidx", _f64p),
    ]


class _Walk(C.Structure):
    _fields_ = [
        ("beta_len", C.c_double), ("gain_len", C.c_double), ("clad_len", C.c_double),
        ("refl", C.c_double), ("length", C.c_double), ("steps", C.c_uint32),
        ("status", C.c_int32),
    ]


class _Params(C.Structure):
    _fields_ = [
        ("min_rays", C.c_uint32), ("max_rays", C.c_uint32), ("n_lambda", C.c_uint32),
        ("sigma_a", _f64p), ("sigma_e", _f64p), ("max_sigma_a", C.c_double),
        ("max_sigma_e", C.c_double), ("mse_threshold", C.c_double),
        ("use_reflections", C.c_int32), ("max_reflections", C.c_uint32),
        ("defensive", C.c_double), ("max_repetitions", C.c_uint32),
        ("adaptive_steps", C.c_uint32), ("sample_min", C.c_uint32),
        ("sample_max", C.c_uint32), ("seed", C.c_uint32), ("n_tot", C.c_double),
        ("tau", C.c_double), ("beta_cells", _f64p), ("threads", C.c_int32),
        ("importance", _f64p), ("fallback", C.c_int32),
        ("sample_list", _u32p), ("n_list", C.c_uint32),
    ]


_lib = None
_lib_f32 = None


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib(geometry="f64"):
    """The oracle library: FP64 throughout (default), or with FP32 ray
    geometry (`geometry="f32"`)."""
    global _lib, _lib_f32
    if geometry == "f32":
        if _lib_f32 is None:
            if not LIB_F32_PATH.exists():
                build()
            _lib_f32 = _declare(C.CDLL(str(LIB_F32_PATH)))
        return _lib_f32
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = _declare(C.CDLL(str(LIB_PATH)))
    return _lib


def _declare(L):
    if True:
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_u32_to_unit.argtypes = [C.c_uint32]
        L.orc_u32_to_unit.restype = C.c_double
        L.orc_image_point.argtypes = [C.POINTER(_Mesh), _f64p, C.c_uint32, C.c_uint32, _f64p]
        L.orc_walk_ray.argtypes = [C.POINTER(_Mesh), _f64p, C.c_uint32, _f64p, C.c_uint32,
                                   C.c_uint32, C.POINTER(_Walk)]
        L.orc_walk_ray.restype = C.c_int
        L.orc_adaptive_target.argtypes = [C.c_uint32] * 4
        L.orc_adaptive_target.restype = C.c_uint32
        L.orc_relative_standard_error.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_relative_standard_error.restype = C.c_double
        L.orc_importance.argtypes = [C.POINTER(_Mesh), C.POINTER(_Params), C.c_uint32, _f64p]
        L.orc_triangle_order.argtypes = [C.POINTER(_Mesh), _u32p]
        L.orc_proposal_prefix.argtypes = [C.POINTER(_Mesh), C.POINTER(_Params), _f64p, _u32p, _f64p]
        L.orc_proposal_prefix.restype = C.c_int
        L.orc_ray_start.argtypes = [C.POINTER(_Mesh), C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_uint32, _f64p]
        L.orc_systematic_offset.argtypes = [C.c_uint32] * 3
        L.orc_systematic_offset.restype = C.c_double
        L.orc_calc_phi_ase.argtypes = [C.POINTER(_Mesh), C.POINTER(_Params), _f64p, _f64p, _u32p,
                                       _f64p, _u32p, _f64p, _f64p]
        L.orc_calc_phi_ase.restype = C.c_int
    return L


def _p(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


class OracleMesh:
    """Holds FP64/u32 copies of a prism mesh and its ctypes view."""

    def __init__(self, topology, medium):
        d = topology.topology()
        self.np_, self.nt, self.nl = topology.numberOfPoints, topology.numberOfTriangles, int(topology.levels)
        self.h = float(topology.thickness)
        self.px = np.ascontiguousarray(topology.points[:, 0], dtype=np.float64)
        self.py = np.ascontiguousarray(topology.points[:, 1], dtype=np.float64)
        self.tri = np.ascontiguousarray(topology.trianglePointIndices.reshape(-1, order="F"), dtype=np.uint32)
        self.nbr = np.ascontiguousarray(d["triangleNeighbors"], dtype=np.int32)
        self.fe = np.ascontiguousarray(d["forbiddenEdge"], dtype=np.int32)
        self.beta = np.ascontiguousarray(medium.value("betaVolume"), dtype=np.float64)
        self.clad = np.ascontiguousarray(medium.value("claddingCellTypes"), dtype=np.uint32)
        self.refl = np.ascontiguousarray(medium.value("reflectivities"), dtype=np.float64)
        self.nidx = np.ascontiguousarray(medium.value("refractiveIndices"), dtype=np.float64)
        self.s = _Mesh(self.np_, self.nt, self.nl, self.h, _p(self.px, C.c_double), _p(self.py, C.c_double),
                       _p(self.tri, C.c_uint32), _p(self.nbr, C.c_int32), _p(self.fe, C.c_int32),
                       _p(self.beta, C.c_double), _p(self.clad, C.c_uint32),
                       int(medium.value("claddingNumber")), float(medium.value("claddingAbsorption")),
                       _p(self.refl, C.c_double), _p(self.nidx, C.c_double))

    def set_beta(self, beta):
        self.beta = np.ascontiguousarray(beta, dtype=np.float64)
        self.s.beta = _p(self.beta, C.c_double)


class OracleParams:
    def __init__(self, *, sigma_a, sigma_e, n_tot, tau, beta_cells, min_rays, max_rays=None,
                 max_sigma_a=None, max_sigma_e=None, mse_threshold=0.005, use_reflections=False,
                 max_reflections=1, defensive=0.05, max_repetitions=1, adaptive_steps=1,
                 sample_min=0, sample_max=None, seed=5489, threads=0, n_samples=None, importance=None,
                 fallback=False, samples=None):
        self.sa = np.ascontiguousarray(np.atleast_1d(sigma_a), dtype=np.float64)
        self.se = np.ascontiguousarray(np.atleast_1d(sigma_e), dtype=np.float64)
        self.bc = np.ascontiguousarray(beta_cells, dtype=np.float64)
        max_rays = min_rays if max_rays is None else max_rays
        msa = float(self.sa.max()) if max_sigma_a is None else float(max_sigma_a)
        mse_ = float(self.se.max()) if max_sigma_e is None else float(max_sigma_e)
        n_samples = self.bc.size if n_samples is None else n_samples
        smax = n_samples - 1 if sample_max is None else sample_max
        self.imp = None if importance is None else np.ascontiguousarray(importance, dtype=np.float64)
        self.s = _Params(min_rays, max_rays, self.sa.size, _p(self.sa, C.c_double), _p(self.se, C.c_double),
                         msa, mse_, mse_threshold, int(bool(use_reflections)), max_reflections, defensive,
                         max_repetitions, adaptive_steps, sample_min, smax, seed, n_tot, tau,
                         _p(self.bc, C.c_double), threads,
                         None if self.imp is None else _p(self.imp, C.c_double), int(bool(fallback)),
                         None, 0)
        # an explicit sample list (any order) instead of the contiguous range
        self.list = None if samples is None else np.ascontiguousarray(samples, dtype=np.uint32)
        if self.list is not None:
            self.s.sample_list = _p(self.list, C.c_uint32)
            self.s.n_list = self.list.size


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return out


def u32_to_unit(x):
    return lib().orc_u32_to_unit(int(x))


def image_point(mesh, p, reflections, first_plane):
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_image_point(C.byref(mesh.s), _p(p, C.c_double), reflections, first_plane, _p(out, C.c_double))
    return out


def walk_ray(mesh, a, prism, e, reflections=0, first_plane=0):
    a = np.ascontiguousarray(a, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    w = _Walk()
    lib().orc_walk_ray(C.byref(mesh.s), _p(a, C.c_double), int(prism), _p(e, C.c_double),
                       int(reflections), int(first_plane), C.byref(w))
    return {f: getattr(w, f) for f, _ in _Walk._fields_}


def adaptive_target(min_rays, max_rays, steps, k):
    return lib().orc_adaptive_target(min_rays, max_rays, steps, k)


def relative_standard_error(s1, s2, n):
    return lib().orc_relative_standard_error(s1, s2, n)


def importance(mesh, params, sample):
    R = 1 + 2 * (max(1, params.s.max_reflections) if params.s.use_reflections else 0)
    out = np.zeros(R * mesh.nt * (mesh.nl - 1))
    lib().orc_importance(C.byref(mesh.s), C.byref(params.s), sample, _p(out, C.c_double))
    return out


def triangle_order(mesh):
    out = np.zeros(mesh.nt, dtype=np.uint32)
    lib().orc_triangle_order(C.byref(mesh.s), _p(out, C.c_uint32))
    return out


def proposal_prefix(mesh, params, imp):
    imp = np.ascontiguousarray(imp, dtype=np.float64)
    order = triangle_order(mesh)
    out = np.zeros_like(imp)
    rc = lib().orc_proposal_prefix(C.byref(mesh.s), C.byref(params.s), _p(imp, C.c_double),
                                   _p(order, C.c_uint32), _p(out, C.c_double))
    return out if rc == 0 else None


def ray_start(mesh, seed, sample, launch, ray, prism):
    out = np.zeros(3)
    lib().orc_ray_start(C.byref(mesh.s), seed, sample, launch, ray, prism, _p(out, C.c_double))
    return out


def systematic_offset(seed, sample, launch):
    return lib().orc_systematic_offset(seed, sample, launch)


def calc_phi_ase(mesh, params, geometry="f64"):
    S = mesh.np_ * mesh.nl
    phi = np.zeros(S)
    mse = np.zeros(S)
    rays = np.zeros(S, dtype=np.uint32)
    dndt = np.zeros(S)
    dropped = np.zeros(S, dtype=np.uint32)
    s1 = np.zeros(S)
    s2 = np.zeros(S)
    rc = lib(geometry).orc_calc_phi_ase(C.byref(mesh.s), C.byref(params.s), _p(phi, C.c_double),
                                         _p(mse, C.c_double),
                                _p(rays, C.c_uint32), _p(dndt, C.c_double), _p(dropped, C.c_uint32),
                                _p(s1, C.c_double), _p(s2, C.c_double))
    if rc != 0:
        raise ValueError(f"oracle rejected the parameters (code {rc})")
    return {"phiAse": phi, "mse": mse, "totalRays": rays, "dndtAse": dndt, "droppedRays": dropped,
            "s1": s1, "s2": s2}

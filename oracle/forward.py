"""ctypes front of the forward Tet4 oracle (oracle/forward_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker of the HIP forward path, imported by
tests/ and bench.py's CPU baseline, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import oracle as _base

BATCHES = 8
ENTERED, BOUNDARY, FAILED = 0, 1, 2

_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_u64p = C.POINTER(C.c_uint64)


class _Mesh(C.Structure):
    _fields_ = [
        ("n_points", C.c_uint32), ("n_cells", C.c_uint32), ("px", _f64p), ("py", _f64p), ("pz", _f64p),
        ("cells", _u32p), ("cell_types", _u32p), ("cladding_number", C.c_uint32),
        ("cladding_absorption", C.c_double), ("beta", _f64p), ("n_tot", C.c_double), ("tau", C.c_double),
        ("planes", _f64p), ("nbr", _i32p), ("nbr_face", _i32p), ("volume", _f64p), ("bv_prefix", _f64p),
        ("lumped", _f64p),
    ]


class _Surfaces(C.Structure):
    _fields_ = [("n_faces", C.c_uint32), ("face", _u32p), ("face_index", _i32p), ("centroid", _f64p),
                ("normal", _f64p), ("reflectivity", _f64p), ("n_in", _f64p), ("n_out", _f64p)]


class Srm(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("max_passes", C.c_uint32), ("divergence_streak", C.c_uint32),
                ("tolerance", C.c_double), ("status", C.c_int32 * BATCHES), ("passes", C.c_uint32 * BATCHES),
                ("remaining", C.c_double * BATCHES), ("order_stride", C.c_uint32)]


class Hit(C.Structure):
    _fields_ = [("face", C.c_int32), ("length", C.c_double), ("tied_mask", C.c_uint32)]


class Transition(C.Structure):
    _fields_ = [("cell", C.c_uint32), ("forbidden", C.c_int32), ("boundary_face", C.c_int32),
                ("status", C.c_int32)]


class Ray(C.Structure):
    _fields_ = [("cell", C.c_uint32), ("p", C.c_double * 3), ("u", C.c_double * 3), ("lambda_", C.c_uint32),
                ("batch", C.c_uint32), ("batch_index", C.c_uint32), ("batch_count", C.c_uint32)]


_lib = None


def _load():
    global _lib
    if _lib is None:
        lib = _base.lib()
        lib.orf_prepare.argtypes = [C.POINTER(_Mesh)]
        lib.orf_prepare.restype = C.c_int
        lib.orf_set_beta.argtypes = [C.POINTER(_Mesh), _f64p]
        lib.orf_next_face.argtypes = [C.POINTER(_Mesh), C.c_uint32, _f64p, _f64p, C.c_int32]
        lib.orf_next_face.restype = Hit
        lib.orf_recover.argtypes = [C.POINTER(_Mesh), C.c_uint32, C.c_int32, _f64p, _f64p]
        lib.orf_recover.restype = Transition
        lib.orf_cross.argtypes = [C.POINTER(_Mesh), C.c_uint32, Hit, _f64p, _f64p]
        lib.orf_cross.restype = Transition
        lib.orf_batch_range.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p]
        lib.orf_point_in_tet.argtypes = [_f64p, _f64p]
        lib.orf_direction.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, _f64p]
        lib.orf_direction.restype = C.c_int
        lib.orf_ray_start.argtypes = [C.POINTER(_Mesh), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                      C.POINTER(Ray)]
        lib.orf_walk.argtypes = [C.POINTER(_Mesh), C.c_double, C.c_double, C.c_uint32, _f64p, _f64p, _f64p, _u32p,
                                 _u32p, _u64p]
        lib.orf_walk.restype = C.c_int
        lib.orf_trace.argtypes = [C.POINTER(_Mesh), _f64p, _f64p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                  _f64p, _u32p, _u32p, _u32p, _u64p, C.c_int]
        lib.orf_finalize.argtypes = [C.POINTER(_Mesh), _f64p, _u32p, C.c_uint64, _u32p, C.c_double, C.c_double,
                                     _f32p, _f64p, _f64p, _f64p]
        lib.orf_evaluate.argtypes = [C.POINTER(_Mesh), _f64p, _f64p, C.c_uint32, C.c_double, C.c_double, C.c_uint32,
                                     C.c_uint32, C.c_uint32, C.c_double, C.c_uint32, _f32p, _f64p, _f64p, _f64p,
                                     _u32p, _u32p, _u32p, _u64p, C.c_int]
        lib.orf_evaluate.restype = C.c_int
        _f32p_ = C.POINTER(C.c_float)
        lib.orf_surfaces_build.argtypes = [C.POINTER(_Mesh), _i32p, C.c_uint32, _f32p_, _f32p_, _f32p_,
                                           C.POINTER(_Surfaces)]
        lib.orf_surfaces_build.restype = C.c_int
        lib.orf_surfaces_free.argtypes = [C.POINTER(_Surfaces)]
        lib.orf_reflectance.argtypes = [C.POINTER(_Surfaces), C.c_uint32, _f64p]
        lib.orf_reflectance.restype = C.c_double
        lib.orf_trace_srm.argtypes = [C.POINTER(_Mesh), C.POINTER(_Surfaces), C.POINTER(Srm), _f64p, _f64p,
                                      C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _f64p, _u32p, _u32p, _u32p,
                                      _u64p]
        lib.orf_evaluate_srm.argtypes = [C.POINTER(_Mesh), C.POINTER(_Surfaces), C.POINTER(Srm), _f64p, _f64p,
                                         C.c_uint32, C.c_double, C.c_double, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.c_double, C.c_uint32, _f32p, _f64p, _f64p, _f64p, _u32p, _u32p, _u32p,
                                         _u64p]
        lib.orf_evaluate_srm.restype = C.c_int
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _vec(v):
    a = np.ascontiguousarray(v, dtype=np.float64).reshape(3)
    return a, a.ctypes.data_as(_f64p)


class ForwardOracle:
    """A prepared Tet4 mesh (face planes, neighbours, volumes, beta*V prefix,
    lumped material vertex volumes) and the oracle's operations on it."""

    def __init__(self, mesh):
        lib = _load()
        self.mesh = mesh
        n, npts = mesh.n_cells, mesh.n_points
        self._px = np.ascontiguousarray(mesh.points[:, 0])
        self._py = np.ascontiguousarray(mesh.points[:, 1])
        self._pz = np.ascontiguousarray(mesh.points[:, 2])
        self._cells = np.ascontiguousarray(mesh.cells.reshape(-1), dtype=np.uint32)
        self._types = None if mesh.cell_types is None else np.ascontiguousarray(mesh.cell_types, dtype=np.uint32)
        self.beta = np.ascontiguousarray(mesh.beta, dtype=np.float64).copy()
        self.planes = np.zeros(16 * n)
        self.nbr = np.zeros(4 * n, np.int32)
        self.nbr_face = np.zeros(4 * n, np.int32)
        self.volume = np.zeros(n)
        self.bv_prefix = np.zeros(n)
        self.lumped = np.zeros(2 * npts)
        self._m = _Mesh(npts, n, _p(self._px, _f64p), _p(self._py, _f64p), _p(self._pz, _f64p),
                        _p(self._cells, _u32p), _p(self._types, _u32p), int(mesh.cladding_number),
                        float(mesh.cladding_absorption), _p(self.beta, _f64p), float(mesh.n_tot), float(mesh.tau),
                        _p(self.planes, _f64p), _p(self.nbr, _i32p), _p(self.nbr_face, _i32p),
                        _p(self.volume, _f64p), _p(self.bv_prefix, _f64p), _p(self.lumped, _f64p))
        rc = lib.orf_prepare(C.byref(self._m))
        if rc != 0:
            raise ValueError({-1: "degenerate cell", -2: "a face is shared by more than two cells",
                              -3: "allocation failed"}.get(rc, f"prepare failed ({rc})"))

    def set_beta(self, beta):
        self.beta[:] = np.asarray(beta, dtype=np.float64).reshape(-1)
        _load().orf_set_beta(C.byref(self._m), _p(self.beta, _f64p))

    def next_face(self, cell, p, u, forbidden=-1):
        (pa, pp), (ua, up) = _vec(p), _vec(u)
        return _load().orf_next_face(C.byref(self._m), cell, pp, up, forbidden)

    def cross(self, cell, hit, p, u):
        (pa, pp), (ua, up) = _vec(p), _vec(u)
        return _load().orf_cross(C.byref(self._m), cell, hit, pp, up)

    def ray_start(self, seed, launch, rays, ray, n_lambda):
        r = Ray()
        _load().orf_ray_start(C.byref(self._m), seed, launch, rays, ray, n_lambda, C.byref(r))
        return r

    def walk(self, cell, p, u, sigma_a=0.0, sigma_e=0.0):
        """(scores [2 Np], visits, dropped, segments, status) of one ray."""
        (pa, pp), (ua, up) = _vec(p), _vec(u)
        scores = np.zeros(2 * self.mesh.n_points)
        visits = np.zeros(self.mesh.n_cells, np.uint32)
        dropped = np.zeros(self.mesh.n_cells, np.uint32)
        seg = C.c_uint64(0)
        st = _load().orf_walk(C.byref(self._m), sigma_a, sigma_e, cell, pp, up, _p(scores, _f64p),
                              _p(visits, _u32p), _p(dropped, _u32p), C.byref(seg))
        return scores, visits, dropped, int(seg.value), int(st)

    def trace(self, sigma_a, sigma_e, seed, launch, rays, threads=0):
        """Raw accumulators of one launch: scores (8, 2 Np), visits, dropped,
        batch ray counts, segments."""
        sa = np.ascontiguousarray(sigma_a, dtype=np.float64)
        se = np.ascontiguousarray(sigma_e, dtype=np.float64)
        scores = np.zeros(BATCHES * 2 * self.mesh.n_points)
        visits = np.zeros(self.mesh.n_cells, np.uint32)
        dropped = np.zeros(self.mesh.n_cells, np.uint32)
        counts = np.zeros(BATCHES, np.uint32)
        seg = C.c_uint64(0)
        _load().orf_trace(C.byref(self._m), _p(sa, _f64p), _p(se, _f64p), sa.size, seed, launch, rays,
                          _p(scores, _f64p), _p(visits, _u32p), _p(dropped, _u32p), _p(counts, _u32p),
                          C.byref(seg), threads)
        return scores.reshape(BATCHES, -1), visits, dropped, counts, int(seg.value)

    def finalize(self, scores, counts, rays, dropped, sigma_a, sigma_e):
        n = self.mesh.n_cells
        sc = np.ascontiguousarray(scores, dtype=np.float64).reshape(-1)
        ct = np.ascontiguousarray(counts, dtype=np.uint32)
        dr = np.ascontiguousarray(dropped, dtype=np.uint32)
        phi, se, rse, dndt = np.zeros(n, np.float32), np.zeros(n), np.zeros(n), np.zeros(n)
        _load().orf_finalize(C.byref(self._m), _p(sc, _f64p), _p(ct, _u32p), rays, _p(dr, _u32p), sigma_a, sigma_e,
                             _p(phi, _f32p), _p(se, _f64p), _p(rse, _f64p), _p(dndt, _f64p))
        return dict(phiAse=phi, standardError=se, relativeStandardError=rse, dndtAse=dndt)

    def evaluate(self, sigma_a, sigma_e, *, max_sigma_a=None, max_sigma_e=None, min_rays, max_rays, steps=4,
                 threshold=0.005, seed=5489, threads=0):
        sa = np.ascontiguousarray(sigma_a, dtype=np.float64).reshape(-1)
        se = np.ascontiguousarray(sigma_e, dtype=np.float64).reshape(-1)
        msa = float(sa.max() if max_sigma_a is None else max_sigma_a)
        mse = float(se.max() if max_sigma_e is None else max_sigma_e)
        n = self.mesh.n_cells
        phi, err, rse, dndt = np.zeros(n, np.float32), np.zeros(n), np.zeros(n), np.zeros(n)
        visits, dropped, conv = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        total = C.c_uint64(0)
        launches = _load().orf_evaluate(C.byref(self._m), _p(sa, _f64p), _p(se, _f64p), sa.size, msa, mse, min_rays,
                                        max_rays, steps, threshold, seed, _p(phi, _f32p), _p(err, _f64p),
                                        _p(rse, _f64p), _p(dndt, _f64p), _p(visits, _u32p), _p(dropped, _u32p),
                                        _p(conv, _u32p), C.byref(total), threads)
        return dict(phiAse=phi, standardError=err, relativeStandardError=rse, dndtAse=dndt, rayVisits=visits,
                    droppedRays=dropped, convergenceRays=conv, rays=int(total.value), launches=launches)


class _Reflecting:
    """Surfaces and reservoir controls of the oracle's surface reservoir method."""

    def __init__(self, oracle, face_surface, reflectivity, n_in=None, n_out=None, *, slots=32, max_passes=40,
                 divergence_streak=3, tolerance=1e-4, order_stride=0):
        self._fs = np.ascontiguousarray(face_surface, dtype=np.int32).reshape(-1)
        self._refl = np.ascontiguousarray(reflectivity, dtype=np.float32)
        self._nin = None if n_in is None else np.ascontiguousarray(n_in, dtype=np.float32)
        self._nout = None if n_out is None else np.ascontiguousarray(n_out, dtype=np.float32)
        self.surfaces = _Surfaces()
        rc = _load().orf_surfaces_build(C.byref(oracle._m), _p(self._fs, _i32p), self._refl.size,
                                        _p(self._refl, C.POINTER(C.c_float)), _p(self._nin, C.POINTER(C.c_float)),
                                        _p(self._nout, C.POINTER(C.c_float)), C.byref(self.surfaces))
        if rc != 0:
            raise MemoryError("surface table")
        self.srm = Srm()
        self.srm.slots, self.srm.max_passes = int(slots), int(max_passes)
        self.srm.divergence_streak, self.srm.tolerance = int(divergence_streak), float(tolerance)
        self.srm.order_stride = int(order_stride)

    @property
    def n_faces(self):
        return int(self.surfaces.n_faces)

    def status(self):
        return (list(self.srm.status), list(self.srm.passes), list(self.srm.remaining))

    def __del__(self):
        try:
            _load().orf_surfaces_free(C.byref(self.surfaces))
        except Exception:
            pass


def reflecting(oracle, face_surface, reflectivity, n_in=None, n_out=None, **controls):
    return _Reflecting(oracle, face_surface, reflectivity, n_in, n_out, **controls)


def trace_srm(oracle, refl, sigma_a, sigma_e, seed, launch, rays):
    """Raw accumulators of one launch with reflections (single-threaded)."""
    sa = np.ascontiguousarray(sigma_a, dtype=np.float64)
    se = np.ascontiguousarray(sigma_e, dtype=np.float64)
    mesh = oracle.mesh
    scores = np.zeros(BATCHES * 2 * mesh.n_points)
    visits = np.zeros(mesh.n_cells, np.uint32)
    dropped = np.zeros(mesh.n_cells, np.uint32)
    counts = np.zeros(BATCHES, np.uint32)
    seg = C.c_uint64(0)
    _load().orf_trace_srm(C.byref(oracle._m), C.byref(refl.surfaces), C.byref(refl.srm), _p(sa, _f64p),
                          _p(se, _f64p), sa.size, seed, launch, rays, _p(scores, _f64p), _p(visits, _u32p),
                          _p(dropped, _u32p), _p(counts, _u32p), C.byref(seg))
    return scores.reshape(BATCHES, -1), visits, dropped, counts, int(seg.value)


def evaluate_srm(oracle, refl, sigma_a, sigma_e, *, max_sigma_a=None, max_sigma_e=None, min_rays, max_rays, steps=4,
                 threshold=0.005, seed=5489):
    sa = np.ascontiguousarray(sigma_a, dtype=np.float64).reshape(-1)
    se = np.ascontiguousarray(sigma_e, dtype=np.float64).reshape(-1)
    msa = float(sa.max() if max_sigma_a is None else max_sigma_a)
    mse = float(se.max() if max_sigma_e is None else max_sigma_e)
    n = oracle.mesh.n_cells
    phi, err, rse, dndt = np.zeros(n, np.float32), np.zeros(n), np.zeros(n), np.zeros(n)
    visits, dropped, conv = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    total = C.c_uint64(0)
    launches = _load().orf_evaluate_srm(C.byref(oracle._m), C.byref(refl.surfaces), C.byref(refl.srm),
                                        _p(sa, _f64p), _p(se, _f64p), sa.size, msa, mse, min_rays, max_rays, steps,
                                        threshold, seed, _p(phi, _f32p), _p(err, _f64p), _p(rse, _f64p),
                                        _p(dndt, _f64p), _p(visits, _u32p), _p(dropped, _u32p), _p(conv, _u32p),
                                        C.byref(total))
    return dict(phiAse=phi, standardError=err, relativeStandardError=rse, dndtAse=dndt, rayVisits=visits,
                droppedRays=dropped, convergenceRays=conv, rays=int(total.value), launches=launches)


def point_in_tet(u3):
    u = np.ascontiguousarray(u3, dtype=np.float64)
    out = np.zeros(4)
    _load().orf_point_in_tet(u.ctypes.data_as(_f64p), out.ctypes.data_as(_f64p))
    return out


def direction(seed, launch, ray):
    out = np.zeros(3)
    calls = _load().orf_direction(seed, launch, ray, out.ctypes.data_as(_f64p))
    return out, calls


def batch_range(rays, batch):
    first, count = C.c_uint32(), C.c_uint32()
    _load().orf_batch_range(rays, batch, C.byref(first), C.byref(count))
    return first.value, count.value

"""The device prism walk (haseongpu_amd/csrc/prism_walk.h) compiled for the
host and checked against the FP64 oracle walk on identical rays (CPU).

The same source the GPU kernels use is built by hipcc for the host
(tests/walk_host/walk_host.cpp) and walked over meshes packed here into the
device's two record layouts (48-byte global records, 32-byte LDS records),
with the device's conventions: heights in layer units, end points on level
planes exactly, beta zero in cladding prisms.  Criterion as on the GPU: the
single-ray gain exp(E_walk - E_oracle) within 1e-5 of 1, the reflectance
product within 1e-6, the same status, the path length within 2e-5, and the
prism count equal on 99% of the rays and within 3 on the others (a line
through a mesh vertex, common on the structured true-cladding mesh, may
count a zero-length prism on one side and not on the other)."""
import ctypes as C
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from haseongpu_amd.geometry import GainMedium, MeshTopology
from haseongpu_amd.laser import CrossSectionData
from haseongpu_amd.workloads import Workload, plain_small, production_disk
from oracle import oracle as O
from tests.helpers import load_true_cladding

HERE = Path(__file__).resolve().parent / "walk_host"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
pytestmark = pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc is needed to build the host walk")


@pytest.fixture(scope="module")
def walk_lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("walk") / "walk_host.so"
    subprocess.run([HIPCC, "-O2", "-fPIC", "-shared", "--cuda-host-only", "-x", "hip", str(HERE / "walk_host.cpp"),
                    "-o", str(out)], check=True, capture_output=True)
    lib = C.CDLL(str(out))
    u32p, f32p, i32p = C.POINTER(C.c_uint32), C.POINTER(C.c_float), C.POINTER(C.c_int32)
    lib.walk_rays_host.argtypes = [C.c_uint32, C.c_uint32, C.c_float, u32p, u32p, C.c_int, f32p, f32p, f32p,
                                   C.c_uint32, f32p, u32p, f32p, u32p, u32p, f32p, u32p, i32p]
    lib.importance_host.argtypes = [C.c_uint32, C.c_uint32, C.c_float, u32p, u32p, C.c_int, f32p, f32p, f32p, f32p,
                                    C.c_float, f32p, f32p, C.c_uint32, C.c_float, C.c_float, C.c_int, C.c_uint32,
                                    C.c_int, f32p, C.POINTER(C.c_ulonglong)]
    return lib


def _f32_exact(w):
    t = w.topology
    topo = MeshTopology(t.points.astype(np.float32).astype(np.float64), t.trianglePointIndices, levels=t.levels,
                        thickness=float(np.float32(t.thickness)))
    return Workload(w.name, topo, GainMedium(topo, dict(w.medium.physical)), w.spectra, w.settings)


LAYOUTS = {"global": 0, "lds": 1, "indexed": 2}


def _records(w):
    """(48-byte records as uint32 [12T], {layout: staged records as uint32},
    beta FP32) in the caller's triangle order, packed as the library packs
    them: "lds" the 32-byte compact records [8T], "indexed" the 16-byte
    indexed records [4T] followed by the FP32 vertex table (two points per
    16 bytes)."""
    topo, med = w.topology, w.medium
    T, L = topo.numberOfTriangles, int(topo.levels)
    d = topo.topology()
    tri = topo.trianglePointIndices.astype(np.int64)
    nb = d["triangleNeighbors"].reshape(3, T).T
    fe = d["forbiddenEdge"].reshape(3, T).T
    clad = (med.value("claddingCellTypes") == int(med.value("claddingNumber"))).astype(np.uint32)
    xy = topo.points.astype(np.float32)
    fb = lambda a: np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)   # noqa: E731
    nw = np.where(nb >= 0, (nb.astype(np.uint64) << 2) | fe.astype(np.uint64), 0x7FFFFFFF).astype(np.uint64)
    nw |= clad[:, None].astype(np.uint64) << 31
    rec = np.zeros((T, 12), np.uint32)
    for k in range(3):
        rec[:, 2 * k] = fb(xy[tri[:, k], 0])
        rec[:, 2 * k + 1] = fb(xy[tri[:, k], 1])
    rec[:, 6], rec[:, 7], rec[:, 8] = nw[:, 0], nw[:, 1], nw[:, 2]
    rec[:, 9] = tri[:, 0].astype(np.uint32) | (clad << 31)
    rec[:, 10], rec[:, 11] = tri[:, 1], tri[:, 2]
    lds = np.zeros((T, 8), np.uint32)
    lds[:, :6] = rec[:, :6]
    W = np.zeros(T, np.uint64)
    for k in range(3):
        f = np.where(nb[:, k] >= 0, (nb[:, k].astype(np.uint64) << 2) | fe[:, k].astype(np.uint64), 0x1FFFFF)
        W |= f.astype(np.uint64) << np.uint64(21 * k)
    W |= clad.astype(np.uint64) << np.uint64(63)
    lds[:, 6] = (W & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    lds[:, 7] = (W >> np.uint64(32)).astype(np.uint32)
    Np = topo.numberOfPoints
    assert Np <= 0x10000
    idx = np.zeros((T + (Np + 1) // 2, 4), np.uint32)
    idx[:T, 0] = tri[:, 0].astype(np.uint32) | (tri[:, 1].astype(np.uint32) << np.uint32(16))
    idx[:T, 1] = tri[:, 2].astype(np.uint32) | (clad << np.uint32(31))
    idx[:T, 2], idx[:T, 3] = lds[:, 6], lds[:, 7]
    vtx = np.zeros(2 * ((Np + 1) // 2) * 2, np.float32)
    vtx[:2 * Np] = xy.reshape(-1)
    idx[T:] = vtx.view(np.uint32).reshape(-1, 4)
    beta = med.value("betaVolume").astype(np.float32).reshape(L - 1, T)
    beta[:, clad == 1] = 0.0
    return rec.reshape(-1), {"lds": lds.reshape(-1), "indexed": idx.reshape(-1)}, beta.reshape(-1)


def _rays(w, om, n, seed, reflections):
    rng = np.random.default_rng(seed)
    topo = w.topology
    T, L, Np = topo.numberOfTriangles, int(topo.levels), topo.numberOfPoints
    prisms = rng.integers(0, T * (L - 1), n)
    starts = np.array([O.ray_start(om, seed, 0, 0, i, int(j)) for i, j in enumerate(prisms)])
    samples = rng.integers(0, Np * L, n)
    sp = topo.samplePoints()[samples]
    nr = rng.integers(0, reflections + 1, n) if reflections else np.zeros(n, np.int64)
    fp = rng.integers(0, 2, n)
    ends = np.array([O.image_point(om, sp[i], int(nr[i]), int(fp[i])) for i in range(n)])
    return starts.astype(np.float32).astype(np.float64), prisms, ends.astype(np.float32).astype(np.float64), nr, fp


def _walk(lib, w, rec, lds, beta, starts, prisms, ends, nr, fp, staged):
    topo, med = w.topology, w.medium
    T, L = topo.numberOfTriangles, int(topo.levels)
    h = float(topo.thickness)
    hs = starts.copy()
    hs[:, 2] = starts[:, 2] / h
    he = ends.copy()
    ze = ends[:, 2] / h
    zr = np.rint(ze)
    he[:, 2] = np.where(np.abs(ze - zr) < 1e-5, zr, ze)
    hs, he = hs.astype(np.float32).reshape(-1), he.astype(np.float32).reshape(-1)
    n = len(prisms)
    out = np.zeros(5 * n, np.float32)
    steps = np.zeros(n, np.uint32)
    status = np.zeros(n, np.int32)
    refl = np.ascontiguousarray(med.value("reflectivities"), np.float32)
    nidx = np.ascontiguousarray(med.value("refractiveIndices"), np.float32)
    P = lambda a, t: np.ascontiguousarray(a, dtype=t).ctypes.data_as(C.POINTER(np.ctypeslib.as_ctypes_type(t)))  # noqa
    staged_rec = lds.get(staged) if staged != "global" else None
    keep = [np.ascontiguousarray(a) if a is not None else None for a in (rec, staged_rec, beta, refl, nidx)]
    lib.walk_rays_host(T, L - 1, np.float32(h), keep[0].ctypes.data_as(C.POINTER(C.c_uint32)),
                       keep[1].ctypes.data_as(C.POINTER(C.c_uint32)) if keep[1] is not None else None,
                       LAYOUTS[staged], keep[2].ctypes.data_as(C.POINTER(C.c_float)),
                       keep[3].ctypes.data_as(C.POINTER(C.c_float)),
                       keep[4].ctypes.data_as(C.POINTER(C.c_float)), n,
                       hs.ctypes.data_as(C.POINTER(C.c_float)),
                       np.ascontiguousarray(prisms, np.uint32).ctypes.data_as(C.POINTER(C.c_uint32)),
                       he.ctypes.data_as(C.POINTER(C.c_float)),
                       np.ascontiguousarray(nr, np.uint32).ctypes.data_as(C.POINTER(C.c_uint32)),
                       np.ascontiguousarray(fp, np.uint32).ctypes.data_as(C.POINTER(C.c_uint32)),
                       out.ctypes.data_as(C.POINTER(C.c_float)), steps.ctypes.data_as(C.POINTER(C.c_uint32)),
                       status.ctypes.data_as(C.POINTER(C.c_int32)))
    return out.reshape(n, 5), steps, status


def _case(name):
    if name == "plain_small":
        return plain_small(), 0
    if name == "true_cladding":
        topo, med, sp, ref = load_true_cladding(1)
        cs = CrossSectionData.monochromatic(wavelength=1030e-9, crossSectionAbsorption=sp["sigmaA"],
                                            crossSectionEmission=sp["sigmaE"])
        return Workload("tc", topo, med, cs, {"monochromatic": True}), 0
    w = production_disk(levels=6, n_points=400, wavelengths=4)
    w.medium.set("refractiveIndices", [1.83, 1.0, 1.83, 1.0])
    return w, 3


@pytest.mark.parametrize("staged", ["global", "lds", "indexed"])
@pytest.mark.parametrize("name", ["plain_small", "true_cladding", "disk_reflect"])
def test_host_walk_matches_the_oracle(walk_lib, name, staged):
    w, refl = _case(name)
    w = _f32_exact(w)
    sa, se, _, _ = w.spectra.spectralPairs(True)
    om = O.OracleMesh(w.topology, w.medium)
    rec, lds, beta = _records(w)
    starts, prisms, ends, nr, fp = _rays(w, om, 3000, seed=21, reflections=refl)
    out, steps, status = _walk(walk_lib, w, rec, lds, beta, starts, prisms, ends, nr, fp, staged)
    nt = float(w.medium.value("nTot"))
    alpha = float(w.medium.value("claddingAbsorption"))
    rel, same_steps, worst = [], 0, 0
    for i in range(len(prisms)):
        o = O.walk_ray(om, starts[i], int(prisms[i]), ends[i], int(nr[i]), int(fp[i]))
        assert o["status"] == 0
        assert status[i] == 0, f"ray {i} failed with status {status[i]}"
        bl, gl, cl, rf, ln = (float(v) for v in out[i])
        Eg = nt * (se[0] + sa[0]) * bl - nt * sa[0] * gl - alpha * cl
        Eo = nt * (se[0] + sa[0]) * o["beta_len"] - nt * sa[0] * o["gain_len"] - alpha * o["clad_len"]
        rel.append(abs(np.exp(Eg - Eo) - 1.0))
        assert rf == pytest.approx(o["refl"], rel=1e-6, abs=1e-7)
        assert gl + cl == pytest.approx(o["length"], rel=2e-5)
        same_steps += int(steps[i] == o["steps"])
        worst = max(worst, abs(int(steps[i]) - int(o["steps"])))
    rel = np.array(rel)
    assert rel.max() <= 1e-5, f"max single-ray gain error {rel.max():.3e}"
    assert same_steps >= 0.99 * len(prisms) and worst <= 3


@pytest.mark.parametrize("name", ["true_cladding", "disk_reflect"])
def test_staged_layouts_walk_bit_for_bit_alike(walk_lib, name):
    """The same rays over the three record layouts (global 48-byte, compact
    and indexed LDS records) give the same FP32 results and step counts bit
    for bit: the layouts hold the same FP32 corners and neighbour links."""
    w, refl = _case(name)
    w = _f32_exact(w)
    om = O.OracleMesh(w.topology, w.medium)
    rec, lds, beta = _records(w)
    starts, prisms, ends, nr, fp = _rays(w, om, 2000, seed=5, reflections=refl)
    outs = {k: _walk(walk_lib, w, rec, lds, beta, starts, prisms, ends, nr, fp, k) for k in LAYOUTS}
    for k in ("lds", "indexed"):
        assert np.array_equal(outs[k][0], outs["global"][0]), k
        assert np.array_equal(outs[k][1], outs["global"][1]) and np.array_equal(outs[k][2], outs["global"][2]), k


def test_grazing_reflections_with_equal_indices_never_reflect_totally(walk_lib):
    """Refractive indices 1 on both sides of both planes (the production
    disk's setting): a ray reflected at a plane it runs nearly parallel to
    takes the triangle's reflectivity, never total reflection.  Starts just
    above the bottom plane, ends at bottom-level sample points through one
    reflection at the bottom (the image lies on the sample's own plane), so
    the rays graze the plane; sin^2 of the incidence angle computed as
    |xy|^2 / D^2 rounds above 1 in FP32 for such rays."""
    w = _f32_exact(production_disk(levels=6, n_points=400, wavelengths=4))
    assert list(w.medium.value("refractiveIndices")) == [1.0, 1.0, 1.0, 1.0]
    om = O.OracleMesh(w.topology, w.medium)
    rec, lds, beta = _records(w)
    topo = w.topology
    T, Np = topo.numberOfTriangles, topo.numberOfPoints
    rng = np.random.default_rng(4)
    n = 4000
    prisms = rng.integers(0, T, n)                      # bottom layer
    starts = np.array([O.ray_start(om, 9, 0, 0, i, int(j)) for i, j in enumerate(prisms)])
    starts[:, 2] = rng.uniform(1e-6, 1e-3, n) * topo.thickness
    starts = starts.astype(np.float32).astype(np.float64)
    sp = topo.samplePoints()[rng.integers(0, Np, n)]    # level 0
    nr, fp = np.ones(n, np.int64), np.zeros(n, np.int64)
    ends = np.array([O.image_point(om, sp[i], 1, 0) for i in range(n)]).astype(np.float32).astype(np.float64)
    for staged in ("global", "lds", "indexed"):
        out, _, status = _walk(walk_lib, w, rec, lds, beta, starts, prisms, ends, nr, fp, staged)
        for i in range(n):
            o = O.walk_ray(om, starts[i], int(prisms[i]), ends[i], 1, 0)
            if o["status"] != 0:
                continue
            assert status[i] == 0
            assert out[i, 3] == pytest.approx(o["refl"], rel=1e-6, abs=1e-7), (i, out[i, 3], o["refl"])
            assert out[i, 3] < 1.0


@pytest.mark.parametrize("staged", ["global", "lds", "indexed"])
@pytest.mark.parametrize("name", ["plain_small", "disk_reflect_2"])
def test_host_importance_matches_the_oracle(walk_lib, name, staged):
    """The importance kernel's column walks (one xy side-test walk per
    triangle column, each level's ray split at its own plane crossings)
    against the oracle's per-prism centroid walks, FP32 against FP64 on an
    FP32-exact mesh: rtol 1e-4 per entry, entries below 1e-6 of the largest
    by absolute tolerance (the GPU test's criterion)."""
    if name == "plain_small":
        w, R, reflect = plain_small(), 1, False
    else:
        w, R, reflect = production_disk(levels=5, n_points=300, wavelengths=8), 5, True
        w.medium.set("refractiveIndices", [1.83, 1.0, 1.83, 1.0])
    w = _f32_exact(w)
    topo, med = w.topology, w.medium
    T, L, Np = topo.numberOfTriangles, int(topo.levels), topo.numberOfPoints
    sa, se, msa, mse_ = w.spectra.spectralPairs(bool(w.settings.get("monochromatic", False)))
    om = O.OracleMesh(topo, med)
    op = O.OracleParams(sigma_a=sa, sigma_e=se, max_sigma_a=msa, max_sigma_e=mse_, n_tot=float(med.value("nTot")),
                        tau=float(med.value("crystalTFluo")), beta_cells=med.value("betaCells"), n_samples=w.samples,
                        min_rays=1, use_reflections=reflect, max_reflections=2)
    rec, lds, beta = _records(w)
    h = float(topo.thickness)
    tri = topo.trianglePointIndices
    ea = topo.points[tri[:, 1]] - topo.points[tri[:, 0]]
    eb = topo.points[tri[:, 2]] - topo.points[tri[:, 0]]
    area = 0.5 * np.abs(ea[:, 0] * eb[:, 1] - ea[:, 1] * eb[:, 0])
    bvol = med.value("betaVolume").astype(np.float64).reshape(L - 1, T)
    bv = np.where(bvol > 0, bvol * area[None, :] * h, 0.0).astype(np.float32).reshape(-1)
    nt = float(med.value("nTot"))
    log2e = 1.4426950408889634
    c1 = (nt * (np.asarray(se) + np.asarray(sa)) * log2e).astype(np.float32)
    c2 = (nt * np.asarray(sa) * log2e).astype(np.float32)
    refl = np.ascontiguousarray(med.value("reflectivities"), np.float32) if reflect else None
    nidx = np.ascontiguousarray(med.value("refractiveIndices"), np.float32) if reflect else None
    pts = topo.points.astype(np.float32)
    staged_rec = lds.get(staged) if staged != "global" else np.zeros(4, np.uint32)
    keep = [np.ascontiguousarray(a) for a in (rec, staged_rec, beta, bv, c1, c2)]
    ptr = lambda a, t: a.ctypes.data_as(C.POINTER(t)) if a is not None else None   # noqa: E731
    rng = np.random.default_rng(3)
    for s in rng.integers(0, w.samples, 5):
        p, lv = int(s) % Np, int(s) // Np
        out = np.zeros(R * T * (L - 1), np.float32)
        steps = C.c_ulonglong(0)
        walk_lib.importance_host(T, L - 1, np.float32(h), ptr(keep[0], C.c_uint32),
                                 ptr(keep[1], C.c_uint32) if staged != "global" else None, LAYOUTS[staged],
                                 ptr(keep[2], C.c_float),
                                 ptr(keep[3], C.c_float), ptr(refl, C.c_float), ptr(nidx, C.c_float),
                                 np.float32(med.value("claddingAbsorption")), ptr(keep[4], C.c_float),
                                 ptr(keep[5], C.c_float), len(c1), pts[p, 0], pts[p, 1], lv, R, 0,
                                 out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(steps))
        ref = O.importance(om, op, int(s))
        np.testing.assert_allclose(out.astype(np.float64), ref, rtol=1e-4, atol=1e-6 * ref.max(),
                                   err_msg=f"sample {s}")
        assert steps.value > T * (L - 1)

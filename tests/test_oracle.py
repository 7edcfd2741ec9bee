"""CPU checks that pin the FP64 oracle before it is trusted as the checker.

Each function of the oracle is checked against something it does not share
code with: published known-answer vectors (Philox), brute force (segment /
prism overlap lengths, reflection counting), closed forms (angular
quadrature of the point flux), the reference's own edge cases (relative
standard error, adaptive schedule) and the legacy golden fixtures.
"""
import numpy as np
import pytest

from haseongpu_amd.geometry import GainMedium, Grid, MeshTopology
from haseongpu_amd.laser import CrossSectionData
from haseongpu_amd.workloads import plain_small
from oracle import oracle as O
from tests.helpers import load_compare_serial, load_true_cladding
from tests.quadrature import box_point_flux


# ---------------------------------------------------------------- Philox
def test_philox_known_answer_vectors():
    """Random123 kat_vectors for philox4x32 with 10 rounds (Salmon et al., SC'11)."""
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, expected in kat:
        assert tuple(int(v) for v in O.philox4x32_10(ctr, key)) == expected


def test_unit_conversion_is_open_and_exact():
    for x in (0, 1, 511, 512, 0x7FFFFFFF, 0xFFFFFFFF):
        u = O.u32_to_unit(x)
        assert 0.0 < u < 1.0
        assert float(np.float32(u)) == u      # exactly representable in FP32
    assert O.u32_to_unit(0) == 2.0 ** -24
    assert O.u32_to_unit(0xFFFFFFFF) == 1.0 - 2.0 ** -24
    # uniformity of the stream used for start points
    vals = np.array([O.u32_to_unit(int(O.philox4x32_10((i, 3, 0, 0), (5489, 0x48415345))[0]))
                     for i in range(20000)])
    assert abs(vals.mean() - 0.5) < 0.01 and abs(vals.var() - 1 / 12) < 0.005


# ------------------------------------------------------- walk brute force
def _small_mesh(seed=0, levels=4, clad=True, refl=False):
    topo = MeshTopology.fromGrid(Grid(x=0.6, y=0.5, z=0.3, tileSizeX=0.2, tileSizeY=0.25, tileSizeZ=0.1))
    rng = np.random.default_rng(seed)
    med = GainMedium(topo).withPhysicalProperties(
        betaVolume=rng.uniform(0, 0.4, topo.numberOfPrisms),
        claddingCellTypes=(rng.uniform(size=topo.numberOfTriangles) < 0.3).astype(np.uint32) if clad
        else np.zeros(topo.numberOfTriangles, np.uint32),
        claddingNumber=1, claddingAbsorption=3.0, nTot=1e20, crystalTFluo=1e-3,
        reflectivities=rng.uniform(0.1, 0.9, 2 * topo.numberOfTriangles) if refl else
        np.zeros(2 * topo.numberOfTriangles),
        refractiveIndices=[1.5, 1.0, 1.5, 1.0] if refl else [1.0, 1.0, 1.0, 1.0])
    return topo, med


def _locate(topo, x, y):
    """Triangle containing (x, y) by barycentric sign tests (brute force)."""
    P, tri = topo.points, topo.trianglePointIndices.astype(np.int64)
    a, b, c = P[tri[:, 0]], P[tri[:, 1]], P[tri[:, 2]]

    def side(p, q):
        return (q[:, 0] - p[:, 0])[:, None] * (y[None, :] - p[:, 1][:, None]) - \
               (q[:, 1] - p[:, 1])[:, None] * (x[None, :] - p[:, 0][:, None])

    inside = (side(a, b) >= -1e-15) & (side(b, c) >= -1e-15) & (side(c, a) >= -1e-15)
    return np.argmax(inside, axis=0)


def _brute_force(topo, med, a, e, n=400000):
    """Length sums of a straight unfolded segment by dense sampling; heights
    are folded back into the slab, mirror copies alternate orientation."""
    T, layers, h = topo.numberOfTriangles, int(topo.levels) - 1, float(topo.thickness)
    s = (np.arange(n) + 0.5) / n
    pts = a[None, :] + s[:, None] * (e - a)[None, :]
    D = np.linalg.norm(e - a)
    tri = _locate(topo, pts[:, 0], pts[:, 1])
    u = np.floor(pts[:, 2] / h).astype(np.int64)
    copy = np.floor_divide(u, layers)
    r = u - copy * layers
    level = np.where(copy % 2 == 1, layers - 1 - r, r)
    beta = med.value("betaVolume")[tri + level * T]
    clad = med.value("claddingCellTypes")[tri] == med.value("claddingNumber")
    ds = D / n
    return {"beta_len": float(np.sum(np.where(clad, 0.0, beta)) * ds),
            "gain_len": float(np.sum(~clad) * ds), "clad_len": float(np.sum(clad) * ds),
            "copies": copy, "tri": tri}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_walk_lengths_match_brute_force(seed):
    topo, med = _small_mesh(seed)
    om = O.OracleMesh(topo, med)
    rng = np.random.default_rng(100 + seed)
    sp = topo.samplePoints()
    for _ in range(12):
        j = int(rng.integers(topo.numberOfPrisms))
        a = O.ray_start(om, 7, 0, 0, int(rng.integers(1 << 20)), j)
        e = sp[int(rng.integers(len(sp)))]
        w = O.walk_ray(om, a, j, e)
        assert w["status"] == 0
        bf = _brute_force(topo, med, a, e)
        D = np.linalg.norm(e - a)
        for key in ("beta_len", "gain_len", "clad_len"):
            assert w[key] == pytest.approx(bf[key], abs=2e-4 * D)
        assert w["gain_len"] + w["clad_len"] == pytest.approx(D, rel=1e-12)


@pytest.mark.parametrize("first", [0, 1])
@pytest.mark.parametrize("reflections", [1, 2, 3])
def test_reflected_walk_matches_brute_force(reflections, first):
    topo, med = _small_mesh(4, refl=True)
    om = O.OracleMesh(topo, med)
    rng = np.random.default_rng(reflections * 10 + first)
    sp = topo.samplePoints()
    refl = med.value("reflectivities").astype(np.float64)
    nidx = med.value("refractiveIndices").astype(np.float64)
    for _ in range(8):
        j = int(rng.integers(topo.numberOfPrisms))
        a = O.ray_start(om, 9, 0, 0, int(rng.integers(1 << 20)), j)
        p = sp[int(rng.integers(len(sp)))]
        e = O.image_point(om, p, reflections, first)
        w = O.walk_ray(om, a, j, e, reflections, first)
        assert w["status"] == 0
        bf = _brute_force(topo, med, a, e)
        D = np.linalg.norm(e - a)
        for key in ("beta_len", "gain_len", "clad_len"):
            assert w[key] == pytest.approx(bf[key], abs=2e-4 * D)
        # surface factors: one per copy boundary crossed, at the triangle there
        copies = bf["copies"]
        change = np.flatnonzero(np.diff(copies) != 0)
        crossings = [(int(copies[i]), int(copies[i + 1]), int(bf["tri"][i])) for i in change]
        if len(crossings) < reflections:   # end point on the last plane
            crossings.append((int(copies[-1]), int(copies[-1]) + (1 if e[2] > a[2] else -1), int(bf["tri"][-1])))
        assert len(crossings) == reflections
        vz = (e[2] - a[2]) / D
        sin2 = 1.0 - vz * vz
        expect = 1.0
        for c0, c1, t in crossings:
            plane = (0 if c0 % 2 else 1) if c1 > c0 else (1 if c0 % 2 else 0)
            tir = (nidx[2 * plane] / nidx[2 * plane + 1]) ** 2 * sin2 > 1.0
            expect *= 1.0 if tir else refl[2 * t + plane]
        assert w["refl"] == pytest.approx(expect, rel=1e-9)


def test_image_points():
    topo, med = _small_mesh()
    om = O.OracleMesh(topo, med)
    Z = (int(topo.levels) - 1) * topo.thickness
    p = np.array([0.1, 0.2, 0.1])
    assert O.image_point(om, p, 0, 0)[2] == pytest.approx(0.1)
    assert O.image_point(om, p, 1, 1)[2] == pytest.approx(2 * Z - 0.1)   # mirror in the top plane
    assert O.image_point(om, p, 1, 0)[2] == pytest.approx(-0.1)          # mirror in the bottom plane
    assert O.image_point(om, p, 2, 1)[2] == pytest.approx(2 * Z + 0.1)
    assert O.image_point(om, p, 2, 0)[2] == pytest.approx(-2 * Z + 0.1)


# ------------------------------------------------- closed-form estimator
def _box_case(beta, sigma_a, sigma_e, *, reflect=False, refl=0.0, nin=1.0, tile=0.125):
    topo = MeshTopology.fromGrid(Grid(x=1.0, y=0.75, z=0.5, tileSizeX=tile, tileSizeZ=0.125))
    med = GainMedium(topo).withPhysicalProperties(
        betaVolume=np.full(topo.numberOfPrisms, beta), betaCells=np.full(topo.numberOfSamples, beta),
        nTot=2e20, crystalTFluo=1e-3, reflectivities=np.full(2 * topo.numberOfTriangles, refl),
        refractiveIndices=[nin, 1.0, nin, 1.0])
    return topo, med


@pytest.mark.parametrize("sigma_e,label", [(0.5e-21 / 0.5 - 1e-21, "zero gain"), (2.0e-20, "gain"),
                                           (0.0, "absorption")])
def test_estimator_matches_angular_quadrature(sigma_e, label):
    beta, sa = 0.5, 1e-21
    topo, med = _box_case(beta, sa, sigma_e)
    g = 2e20 * (beta * (sigma_e + sa) - sa)
    om = O.OracleMesh(topo, med)
    samples = [0 + 2 * topo.numberOfPoints + 12, 40, topo.numberOfSamples - 1, 3 * topo.numberOfPoints + 30]
    sp = topo.samplePoints()
    for s in samples:
        op = O.OracleParams(sigma_a=sa, sigma_e=sigma_e, n_tot=2e20, tau=1e-3, beta_cells=med.value("betaCells"),
                            min_rays=300000, sample_min=s, sample_max=s, seed=11)
        r = O.calc_phi_ase(om, op)
        exact = box_point_flux(sp[s], (0.0, 1.0, 0.0, 0.75), 0.5, beta, g, 2e20 / 1e-3)
        rse = np.sqrt(r["mse"][s])
        assert r["phiAse"][s] == pytest.approx(exact, rel=max(5 * rse, 2e-3)), (label, s)


@pytest.mark.parametrize("nin", [1.0, 1.8])
def test_estimator_reflections_match_image_quadrature(nin):
    beta, sa, se = 0.4, 1e-21, 1.2e-20
    topo, med = _box_case(beta, sa, se, reflect=True, refl=0.3, nin=nin)
    g = 2e20 * (beta * (se + sa) - sa)
    om = O.OracleMesh(topo, med)
    sp = topo.samplePoints()
    for s in [5 + 2 * topo.numberOfPoints, topo.numberOfPoints + 20, 4 * topo.numberOfPoints + 33]:
        op = O.OracleParams(sigma_a=sa, sigma_e=se, n_tot=2e20, tau=1e-3, beta_cells=med.value("betaCells"),
                            min_rays=300000, sample_min=s, sample_max=s, seed=13, use_reflections=True,
                            max_reflections=1)
        r = O.calc_phi_ase(om, op)
        exact = box_point_flux(sp[s], (0.0, 1.0, 0.0, 0.75), 0.5, beta, g, 2e20 / 1e-3, reflect=True,
                               refl_bottom=0.3, refl_top=0.3, n_in=nin, n_out=1.0)
        direct = box_point_flux(sp[s], (0.0, 1.0, 0.0, 0.75), 0.5, beta, g, 2e20 / 1e-3)
        assert exact > direct * 1.05          # the reflected part is not negligible
        rse = np.sqrt(r["mse"][s])
        assert r["phiAse"][s] == pytest.approx(exact, rel=max(5 * rse, 3e-3))


def test_estimator_is_unbiased_for_any_defensive_fraction():
    topo, med = _box_case(0.3, 1e-21, 2e-20)
    om = O.OracleMesh(topo, med)
    s = 2 * topo.numberOfPoints + 17
    vals = []
    for eps in (0.0, 0.05, 0.5, 0.95):
        op = O.OracleParams(sigma_a=1e-21, sigma_e=2e-20, n_tot=2e20, tau=1e-3, beta_cells=med.value("betaCells"),
                            min_rays=200000, sample_min=s, sample_max=s, seed=17, defensive=eps)
        r = O.calc_phi_ase(om, op)
        vals.append((r["phiAse"][s], np.sqrt(r["mse"][s]) * r["phiAse"][s]))
    ref = np.mean([v for v, _ in vals])
    for v, se in vals:
        assert abs(v - ref) < 5 * se + 2e-3 * ref


# ------------------------------------------------ statistics and schedule
def test_relative_standard_error_edge_cases():
    """tests/forwardMse.cpp:104-139."""
    assert O.relative_standard_error(4.0, 10.0, 4.0) == pytest.approx(np.sqrt((4 * 10 / 16 - 1) / 4))
    assert O.relative_standard_error(1.0, 1.0, 1.0) == np.finfo(np.float64).max
    assert np.isnan(O.relative_standard_error(0.0, 0.0, 2.0))
    assert O.relative_standard_error(np.inf, 1.0, 2.0) == np.finfo(np.float64).max


def test_adaptive_ray_target():
    """include/core/types.hpp:333-354: geometric growth from min to max."""
    assert [O.adaptive_target(1000, 1000000, 3, k) for k in range(5)] == [1000, 10000, 100000, 1000000, 1000000]
    assert O.adaptive_target(1000, 1000, 4, 2) == 1000
    assert O.adaptive_target(1000, 5000, 0, 3) == 1000
    assert O.adaptive_target(100, 1000, 2, 1) == 317   # ceil(100 * sqrt(10))


def test_adaptive_loop_pools_launches():
    topo, med = _box_case(0.3, 1e-21, 2e-20)
    om = O.OracleMesh(topo, med)
    op = O.OracleParams(sigma_a=1e-21, sigma_e=2e-20, n_tot=2e20, tau=1e-3, beta_cells=med.value("betaCells"),
                        min_rays=1000, max_rays=16000, adaptive_steps=4, max_repetitions=3, mse_threshold=1e-4,
                        seed=3, sample_max=60)
    r = O.calc_phi_ase(om, op)
    rays = r["totalRays"][:61]
    assert set(np.unique(rays)) <= {1000, 2000, 4000, 8000, 16000, 32000, 48000}
    assert np.all((r["mse"][:61] <= 1e-4) | (rays == 48000))
    # a sample that stopped early met the threshold
    assert np.all(r["mse"][:61][rays < 48000] <= 1e-4)


# -------------------------------------------------------------- fixtures
@pytest.mark.parametrize("name", ["cuboid", "cylindrical"])
def test_oracle_reproduces_legacy_serial_solver(name):
    """compareSerial: legacy serial solver, 1e6 rays per sample."""
    topo, med, sp, ref = load_compare_serial(name)
    om = O.OracleMesh(topo, med)
    idx = np.arange(0, topo.numberOfSamples, 13)
    op = O.OracleParams(sigma_a=sp["sigmaA"], sigma_e=sp["sigmaE"], n_tot=float(med.value("nTot")),
                        tau=float(med.value("crystalTFluo")), beta_cells=med.value("betaCells"), min_rays=20000,
                        seed=5489)
    # subset of samples through the sample range, one call per contiguous block
    phi = []
    for s in idx:
        op.s.sample_min = int(s)
        op.s.sample_max = int(s)
        phi.append(O.calc_phi_ase(om, op)["phiAse"][s])
    ratio = np.asarray(phi) / ref["phiASE"][idx]
    assert abs(ratio.mean() - 1.0) < 0.01
    assert abs(np.median(ratio) - 1.0) < 0.01
    # dndtAse = (beta (sigma_e + sigma_a) - sigma_a) * phiASE holds exactly in the fixture
    gain = med.value("betaCells") * (sp["sigmaE"] + sp["sigmaA"]) - sp["sigmaA"]
    np.testing.assert_allclose(ref["dndtAse"], gain * ref["phiASE"], rtol=1e-12)


def test_oracle_matches_true_cladding_fixture_scale():
    """Legacy importance-sampled phiASE at 2000 rays/sample, cladding
    absorption 5.5 cm^-1: the 2000-ray oracle sits on the same level."""
    topo, med, sp, ref = load_true_cladding(1)
    om = O.OracleMesh(topo, med)
    op = O.OracleParams(sigma_a=sp["sigmaA"], sigma_e=sp["sigmaE"], n_tot=float(med.value("nTot")),
                        tau=float(med.value("crystalTFluo")), beta_cells=med.value("betaCells"),
                        min_rays=ref["raysPerSample"], seed=ref["rngSeed"], mse_threshold=1.0)
    r = O.calc_phi_ase(om, op)
    assert np.all(r["totalRays"] == ref["totalRays"])
    assert abs(np.median(r["phiAse"] / ref["phiASE"]) - 1.0) < 0.01
    assert abs(r["phiAse"].mean() / ref["phiASE"].mean() - 1.0) < 0.02


@pytest.mark.parametrize("index", [0, 1, 2])
def test_legacy_mse_is_the_squared_relative_standard_error(index):
    """The legacy mse (F2: importance-sampled, 2000 rays per sample on 4210
    sample points; its formula is not in the reference) against this
    estimator's RSE^2 on the same mesh and ray count.

    The estimator has an unbounded variance (a ray may start next to the
    sample point, 1/r^2), so single-run RSEs scatter over a decade and the
    tails of the two distributions follow the two samplers' different ray
    allocations.  The interquartile range carries the scale: the geometric
    mean over q in [0.25, 0.75] of Q_legacy(q) / Q(q) is 1.45-1.60 for RSE^2
    across absorptions and seeds.  The window [1.2, 2.0] admits RSE^2 and
    rejects the other candidate definitions, whose ratios are a fixed
    multiple of it: mse = RSE (~0.05), 2 RSE^2 (~0.76), RSE^2 / 2 (~3.0).  The
    factor ~1.5 is the legacy sampler's larger variance at equal rays, not a
    property of the formula; it stays unpinned.  The reference's threshold
    translations (mse 0.1 ~ RSE 0.14 / 0.10, tests/data/compareSerial/
    README.md:83; 0.01087 ~ 0.05, fixed_legacy_reflection_pump3_wedge_reference/
    README.md:76) relate the legacy mse to the forward Tet4 estimator's RSE,
    a different estimator at a different effort, so they fix no scale.  Their
    consistency still points the same way (test_threshold_translations_favour
    _the_square below)."""
    topo, med, sp, ref = load_true_cladding(index)
    om = O.OracleMesh(topo, med)
    op = O.OracleParams(sigma_a=sp["sigmaA"], sigma_e=sp["sigmaE"], n_tot=float(med.value("nTot")),
                        tau=float(med.value("crystalTFluo")), beta_cells=med.value("betaCells"),
                        min_rays=ref["raysPerSample"], seed=ref["rngSeed"], mse_threshold=1.0)
    rse2 = O.calc_phi_ase(om, op)["mse"]
    qs = np.linspace(0.25, 0.75, 5)

    def scale(model):
        return float(np.exp(np.mean(np.log(np.quantile(ref["mse"], qs) / np.quantile(model, qs)))))

    assert 1.2 <= scale(rse2) <= 2.0, scale(rse2)
    # the alternatives the window excludes
    for wrong in (np.sqrt(rse2), 2.0 * rse2, 0.5 * rse2):
        assert not (1.2 <= scale(wrong) <= 2.0)
    # the legacy values are finite, positive and capped as the fixture states
    assert np.all(ref["mse"] > 0) and ref["mse"].max() < 0.13


def test_threshold_translations_favour_the_square():
    """The reference translates legacy mse thresholds into forward-estimator
    RSE thresholds at matched effort: 0.1 -> 0.14 (cuboid) and 0.10
    (cylindrical), tests/data/compareSerial/README.md:83; 0.01087 -> 0.05,
    fixed_legacy_reflection_pump3_wedge_reference/README.md:76.  If
    mse = RSE_legacy^p and RSE_legacy is a fixed multiple of the forward RSE
    (the two estimators' effort ratio), mse^(1/p) / RSE is the same for all
    three pairs.  Under p = 2 the three ratios spread by 1.5x (2.1-3.2); under
    p = 1 by 4.6x (0.22-1.0).  The translations therefore reject mse = RSE;
    they cannot tell RSE^2 from a multiple of it (the F2 interquartile scale
    above does that)."""
    pairs = np.array([[0.1, 0.14], [0.1, 0.10], [0.01087, 0.05]])

    def spread(p):
        r = pairs[:, 0] ** (1.0 / p) / pairs[:, 1]
        return r.max() / r.min()

    assert spread(2.0) < 1.6
    assert spread(1.0) > 4.0
    # a constant factor on mse (mse = 2 RSE^2) leaves the spread as it is
    two = pairs.copy()
    two[:, 0] /= 2.0
    assert (two[:, 0] ** 0.5 / two[:, 1]).max() / (two[:, 0] ** 0.5 / two[:, 1]).min() == pytest.approx(spread(2.0))


def test_cladding_absorption_lowers_flux_monotonically():
    means = []
    for idx in range(3):
        topo, med, sp, ref = load_true_cladding(idx)
        means.append(ref["phiASE"].mean())
    assert means[0] > means[1] > means[2]


def test_external_importance_reproduces_the_walked_run():
    """The oracle's external-proposal input, fed its own walked map, gives the
    walked run bit for bit (the learned-proposal parity tests feed it the
    network's map through the same door)."""
    w = plain_small(rays=2000)
    om = O.OracleMesh(w.topology, w.medium)
    kw = dict(sigma_a=[1.16e-21], sigma_e=[2.48e-20], n_tot=2.76e20, tau=9.5e-4,
              beta_cells=w.medium.value("betaCells"), min_rays=2000, sample_min=100, sample_max=139, seed=3,
              n_samples=w.samples)
    base = O.calc_phi_ase(om, O.OracleParams(**kw))
    imp = np.stack([O.importance(om, O.OracleParams(**kw), s) for s in range(100, 140)])
    ext = O.calc_phi_ase(om, O.OracleParams(**kw, importance=imp))
    np.testing.assert_array_equal(ext["phiAse"], base["phiAse"])
    np.testing.assert_array_equal(ext["mse"], base["mse"])
    # a flat (beta*V only) external proposal changes the estimate but not its mean
    flat = O.calc_phi_ase(om, O.OracleParams(**kw, importance=np.zeros_like(imp)))
    sl = slice(100, 140)
    assert not np.array_equal(flat["phiAse"][sl], base["phiAse"][sl])
    z = (flat["phiAse"][sl] - base["phiAse"][sl]) / (base["phiAse"][sl] * np.sqrt(flat["mse"][sl] + base["mse"][sl]))
    assert abs(np.mean(z)) < 1.0 and np.all(np.abs(z) < 5)


def test_flat_schedule_terminates():
    """adaptive_steps = 0 with max_rays > min_rays: the target never grows
    (include/core/types.hpp:338-341), so a sample that misses the threshold
    stops at min_rays instead of looping."""
    w = plain_small(rays=500)
    om = O.OracleMesh(w.topology, w.medium)
    op = O.OracleParams(sigma_a=[1.16e-21], sigma_e=[2.48e-20], n_tot=2.76e20, tau=9.5e-4,
                        beta_cells=w.medium.value("betaCells"), min_rays=500, max_rays=5000, adaptive_steps=0,
                        mse_threshold=1e-12, sample_min=0, sample_max=20, n_samples=w.samples)
    r = O.calc_phi_ase(om, op)
    assert np.all(r["totalRays"][:21] == 500)
    assert O.adaptive_target(500, 5000, 0, 7) == 500


def test_high_budget_fixture_matches_the_oracle_at_a_small_budget():
    """tests/golden/production_disk_1e9.npz (the FP64 oracle at 1e9 rays per
    sample on six production-disk samples) against the oracle rerun here at
    1e6 rays with another seed: within 4 combined standard errors, and its
    own relative standard errors below 1e-3."""
    from pathlib import Path

    from haseongpu_amd.workloads import production_disk

    hb = np.load(Path(__file__).resolve().parent / "golden" / "production_disk_1e9.npz")
    assert np.all(hb["totalRays"] == 1_000_000_000) and np.all(np.sqrt(hb["mse"]) < 1e-3)
    w = production_disk()
    s = w.settings
    sa, se, msa, mse_ = w.spectra.spectralPairs(False)
    om = O.OracleMesh(w.topology, w.medium)
    op = O.OracleParams(sigma_a=sa, sigma_e=se, max_sigma_a=msa, max_sigma_e=mse_,
                        n_tot=float(w.medium.value("nTot")), tau=float(w.medium.value("crystalTFluo")),
                        beta_cells=w.medium.value("betaCells"), min_rays=1_000_000, max_rays=1_000_000,
                        mse_threshold=0.0, use_reflections=True, max_reflections=s["maxReflections"],
                        adaptive_steps=0, seed=4321, n_samples=w.samples, samples=hb["samples"])
    ref = O.calc_phi_ase(om, op)
    smp = hb["samples"]
    z = (ref["phiAse"][smp] - hb["phiAse"]) / np.sqrt(ref["mse"][smp] * ref["phiAse"][smp] ** 2
                                                     + hb["mse"] * hb["phiAse"] ** 2)
    assert np.all(np.abs(z) < 4), z


def test_fp32_geometry_build_stays_within_a_tenth_of_a_standard_error():
    """The CPU baseline's build (oracle/build/libhase_oracle_f32.so: FP32 ray
    geometry, FP64 sums) against the FP64 oracle on identical streams, eight
    production-disk samples at 20,000 rays with reflections and the cladding
    ring: the builds differ (the FP32 walk is really loaded), and no estimate
    moves by more than a tenth of its standard error."""
    from haseongpu_amd.workloads import production_disk

    w = production_disk()
    s = w.settings
    sa, se, msa, mse_ = w.spectra.spectralPairs(False)
    om = O.OracleMesh(w.topology, w.medium)
    picks = np.sort(np.random.default_rng(1).choice(w.samples, 8, replace=False)).astype(np.int64)
    op = O.OracleParams(sigma_a=sa, sigma_e=se, max_sigma_a=msa, max_sigma_e=mse_,
                        n_tot=float(w.medium.value("nTot")), tau=float(w.medium.value("crystalTFluo")),
                        beta_cells=w.medium.value("betaCells"), min_rays=20000, max_rays=20000,
                        mse_threshold=0.0, use_reflections=True, max_reflections=s["maxReflections"],
                        adaptive_steps=0, seed=99, n_samples=w.samples, samples=picks)
    r64 = O.calc_phi_ase(om, op)
    r32 = O.calc_phi_ase(om, op, geometry="f32")
    a, b = r64["phiAse"][picks], r32["phiAse"][picks]
    assert np.all(r32["droppedRays"][picks] == 0) and np.all(r32["totalRays"][picks] == 20000)
    assert np.any(a != b)
    assert np.all(np.abs(b / a - 1) < 0.1 * np.sqrt(r64["mse"][picks])), (b / a - 1, np.sqrt(r64["mse"][picks]))

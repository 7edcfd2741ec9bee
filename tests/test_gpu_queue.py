"""GPU tests of the persistent work-queue ray scheduling: it traces the same
rays as one launch per adaptive round (identical ray counts, phiAse equal up
to the FP64 summation order, i.e. a few FP32 ulps), with and without the
captured hipGraph, through escalation, repetitions, flat schedules and the
learned proposal's deferral; the autotuner returns a valid setting that
leaves the results unchanged."""
import numpy as np
import pytest

from haseongpu_amd import _native
from haseongpu_amd.workloads import plain_small, production_disk
from oracle import oracle as O
from tests.test_gpu_parity import _context, _oracle
from tests.test_gpu_proposal import _proposal_for, _random_net

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert np.array_equal(a.totalRays, b.totalRays)
    assert np.array_equal(a.droppedRays, b.droppedRays)
    np.testing.assert_allclose(a.phiAse, b.phiAse, rtol=3e-7, atol=0)
    ok = np.isfinite(a.mse) & (a.mse < 1e300)
    np.testing.assert_allclose(a.mse[ok], b.mse[ok], rtol=1e-9)


def _modes(ctx):
    out = {}
    for name, kw in (("rounds", dict(mode=_native.TRACE_ROUNDS)),
                     ("queue", dict(mode=_native.TRACE_QUEUE, use_graph=0, stage_lds=0, block_size=512)),
                     ("graph", dict(mode=_native.TRACE_QUEUE, use_graph=1, stage_lds=0, block_size=512)),
                     ("lds", dict(mode=_native.TRACE_QUEUE, use_graph=1, stage_lds=1, block_size=1024))):
        ctx.set_tuning(**kw)
        out[name] = ctx.evaluate()
    return out


@pytest.mark.parametrize("case", ["plain", "reflect"])
def test_queue_traces_the_same_rays_as_rounds(case):
    if case == "plain":
        w = plain_small()
        ctx, _ = _context(w, min_rays=2000, max_rays=64000, mse_threshold=2e-4, adaptive_steps=5,
                          max_repetitions=2, rng_seed=9)
    else:
        w = production_disk(levels=6, n_points=400, wavelengths=8)
        ctx, _ = _context(w, min_rays=3000, max_rays=30000, mse_threshold=3e-4, adaptive_steps=3,
                          use_reflections=True, rng_seed=4)
    with ctx:
        r = _modes(ctx)
        st = ctx.stats()
    assert len(np.unique(r["rounds"].totalRays)) > 2       # the schedule escalated
    _same(r["rounds"], r["queue"])
    _same(r["rounds"], r["graph"])
    # triangle records staged in LDS: the same walks, bit for bit
    np.testing.assert_array_equal(r["lds"].phiAse, r["graph"].phiAse)
    np.testing.assert_array_equal(r["lds"].totalRays, r["graph"].totalRays)
    assert st["rays"] == int(r["graph"].totalRays.astype(np.int64).sum())


@pytest.mark.parametrize("shape", [dict(block_size=64, rays_per_item=64, queue_chunk=1, refill_lanes=1, step_budget=1),
                                   dict(block_size=512, rays_per_item=4096, queue_chunk=7, refill_lanes=64,
                                        step_budget=10001),
                                   dict(block_size=64, rays_per_item=128, queue_chunk=3, stage_lds=1),
                                   dict(block_size=1024, rays_per_item=4096, queue_chunk=1, stage_lds=1)])
def test_queue_extreme_launch_shapes(shape):
    w = plain_small()
    ctx, _ = _context(w, min_rays=3000, max_rays=12000, mse_threshold=1e-4, adaptive_steps=2, rng_seed=3)
    with ctx:
        ctx.set_tuning(mode=_native.TRACE_ROUNDS)
        ref = ctx.evaluate()
        ctx.set_tuning(mode=_native.TRACE_QUEUE, use_graph=1, **shape)
        got = ctx.evaluate()
    _same(ref, got)


def test_flat_schedule_stops_after_the_minimum():
    """adaptive_steps = 0 with max > min: the target never grows, so every
    sample stops at min_rays (types.hpp:338-341), in both paths and the oracle."""
    w = plain_small()
    ctx, spectra = _context(w, min_rays=1500, max_rays=9000, mse_threshold=1e-9, adaptive_steps=0, rng_seed=2,
                            max_sample_range=299)
    with ctx:
        r = _modes(ctx)
    for res in r.values():
        assert np.all(res.totalRays[:300] == 1500) and np.all(res.totalRays[300:] == 0)
    om, op = _oracle(w, spectra, min_rays=1500, max_rays=9000, mse_threshold=1e-9, adaptive_steps=0, seed=2,
                     sample_max=299)
    ref = O.calc_phi_ase(om, op)
    assert np.all(ref["totalRays"][:300] == 1500)
    rel = np.abs(r["graph"].phiAse[:300] / ref["phiAse"][:300] - 1.0)
    assert np.median(rel) <= 1e-5 and np.quantile(rel, 0.99) <= 1e-4


def test_learned_deferral_in_the_queue_matches_rounds():
    w = production_disk(levels=6, n_points=400, wavelengths=8)
    ctx, _ = _context(w, min_rays=4096, max_rays=16384, adaptive_steps=2, use_reflections=True,
                      mse_threshold=2e-3, rng_seed=8)
    with ctx:
        ctx.set_proposal(_proposal_for(w, _random_net(8, 0.5), 4, 2, fallback=True))
        ctx.set_tuning(mode=_native.TRACE_ROUNDS)
        a = ctx.evaluate()
        fa = ctx.proposal_stats()["fallback_samples"]
        ctx.set_tuning(mode=_native.TRACE_QUEUE, use_graph=1)
        b = ctx.evaluate()
        fb = ctx.proposal_stats()["fallback_samples"]
    assert fa == fb and fa > 0
    _same(a, b)


def test_autotune_picks_a_valid_setting_and_keeps_results():
    w = production_disk(levels=6, n_points=400, wavelengths=8)
    ctx, _ = _context(w, min_rays=4096, use_reflections=True, rng_seed=1)
    with ctx:
        ref = ctx.evaluate()
        best = ctx.autotune(256)
        assert best["mode"] == _native.TRACE_QUEUE and best["probe_ms"] > 0
        assert best["block_size"] in (128, 256, 512, 1024) and best["rays_per_item"] in (256, 512, 1024, 2048)
        assert best["stage_lds"] in (0, 1)
        assert ctx.tuning() == best
        got = ctx.evaluate()
    _same(ref, got)


@pytest.mark.parametrize("learned", [False, True])
def test_results_do_not_depend_on_the_batch_size(learned):
    # sample_batch 0 sizes the batch from the free device memory (every sample point of this
    # disk in one batch); 97 splits the same evaluation into 25 batches
    w = production_disk(levels=6, n_points=400, wavelengths=8)
    out = []
    for sb in (0, 97):
        ctx, _ = _context(w, min_rays=4096, max_rays=16384, adaptive_steps=2, use_reflections=True,
                          mse_threshold=2e-3, rng_seed=8, sample_batch=sb)
        with ctx:
            if learned:
                ctx.set_proposal(_proposal_for(w, _random_net(8, 0.5), 4, 2, fallback=True))
            out.append(ctx.evaluate())
            if learned:
                assert ctx.proposal_stats()["fallback_samples"] > 0
    assert w.samples > 97 * 20
    _same(*out)


def test_indexed_lds_records_of_a_larger_mesh():
    """A 6720-triangle disk (beyond the 4800 triangles of the compact LDS
    records, within the indexed ones: 16 B per triangle plus the vertex
    table): the staged trace walks the same rays as the global-record trace,
    bit for bit, and the staged importance map matches the FP64 oracle."""
    from haseongpu_amd.workloads import batch_problems
    from tests.test_gpu_parity import _f32_exact

    w = _f32_exact(batch_problems(1, seed=5, triangles=(6000, 7000), levels=(4, 4), wavelengths=(2, 2),
                                  min_rays=1000, max_rays=4000)[0])
    T, Np = w.topology.numberOfTriangles, w.topology.numberOfPoints
    assert 4800 < T and T + (Np + 1) // 2 <= 10176
    ctx, spectra = _context(w, min_rays=2048, max_rays=8192, mse_threshold=1e-3, adaptive_steps=2, rng_seed=6,
                            min_sample_range=0, max_sample_range=2999)
    om, op = _oracle(w, spectra, min_rays=1)
    with ctx:
        ctx.set_tuning(mode=_native.TRACE_QUEUE, use_graph=1, stage_lds=0, block_size=512)
        glob = ctx.evaluate()
        ctx.set_tuning(mode=_native.TRACE_QUEUE, use_graph=1, stage_lds=1, block_size=1024)
        lds = ctx.evaluate()
        np.testing.assert_array_equal(lds.totalRays, glob.totalRays)
        np.testing.assert_array_equal(lds.phiAse, glob.phiAse)
        for s in np.random.default_rng(8).integers(0, 3000, 3):
            g = ctx.importance(int(s)).astype(np.float64)
            c = O.importance(om, op, int(s))
            np.testing.assert_allclose(g, c, rtol=1e-4, atol=1e-6 * c.max())


def test_compact_records_in_global_memory_of_a_mesh_beyond_lds():
    """A 10,027-triangle disk (beyond both LDS formats): its walks read the
    32-byte compact records from global memory (stage_lds 1) or the 48-byte
    records (stage_lds 0); the same rays bit for bit, and the importance map
    from the compact records matches the FP64 oracle."""
    from haseongpu_amd.workloads import batch_problems
    from tests.test_gpu_parity import _f32_exact

    w = _f32_exact(batch_problems(1, seed=7, triangles=(9500, 10500), levels=(3, 3), wavelengths=(2, 2),
                                  min_rays=1000, max_rays=4000)[0])
    T, Np = w.topology.numberOfTriangles, w.topology.numberOfPoints
    assert T + (Np + 1) // 2 > 10176
    ctx, spectra = _context(w, min_rays=2048, max_rays=8192, mse_threshold=1e-3, adaptive_steps=2, rng_seed=8,
                            min_sample_range=0, max_sample_range=2999)
    om, op = _oracle(w, spectra, min_rays=1)
    with ctx:
        ctx.set_tuning(mode=_native.TRACE_QUEUE, use_graph=1, stage_lds=0, block_size=512)
        wide = ctx.evaluate()
        ctx.set_tuning(mode=_native.TRACE_QUEUE, use_graph=1, stage_lds=1, block_size=512)
        compact = ctx.evaluate()
        np.testing.assert_array_equal(compact.totalRays, wide.totalRays)
        np.testing.assert_array_equal(compact.phiAse, wide.phiAse)
        for s in np.random.default_rng(9).integers(0, 3000, 3):
            g = ctx.importance(int(s)).astype(np.float64)
            c = O.importance(om, op, int(s))
            np.testing.assert_allclose(g, c, rtol=1e-4, atol=1e-6 * c.max())

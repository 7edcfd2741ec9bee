"""GPU tests of batched evaluation: independent problems (different meshes,
levels, spectra, schedules, proposals) evaluated together in one work queue
per round give each problem the results it gets alone (same rays; phiAse
within the FP64 summation order)."""
import numpy as np
import pytest

from haseongpu_amd import _native
from haseongpu_amd.simulation import native_problem
from haseongpu_amd.workloads import batch_problems, production_disk
from tests.test_gpu_proposal import _proposal_for, _random_net
from tests.test_gpu_queue import _same

pytestmark = pytest.mark.gpu


def _contexts(workloads):
    out = []
    for w in workloads:
        mesh, exp, comp, _ = native_problem(w)
        ctx = _native.Context(mesh, exp, comp)
        ctx.set_tuning(mode=_native.TRACE_QUEUE, use_graph=1)
        out.append(ctx)
    return out


def test_batch_equals_each_problem_alone():
    ws = batch_problems(5, seed=1, triangles=(200, 900), levels=(4, 9), wavelengths=(1, 6), min_rays=2000,
                        max_rays=20000, mse_threshold=2e-3, adaptive_steps=2)
    disk = production_disk(levels=6, n_points=400, wavelengths=8, min_rays=4096, max_rays=16384,
                           mse_threshold=2e-3)
    ctxs = _contexts(ws + [disk])
    ctxs[-1].set_proposal(_proposal_for(disk, _random_net(8, 0.5), 4, 2, fallback=True))
    alone = [c.evaluate() for c in ctxs]
    fallback_alone = ctxs[-1].proposal_stats()["fallback_samples"]
    with _native.Batch(ctxs) as batch:
        together = batch.evaluate()
        st = batch.stats()
    for a, b in zip(alone, together):
        _same(a, b)
    assert ctxs[-1].proposal_stats()["fallback_samples"] == fallback_alone > 0
    # traced rays include the first launch of the samples that restarted on the walked map
    assert st["rays"] == sum(int(r.totalRays.astype(np.int64).sum()) for r in together) + fallback_alone * 4096
    assert any(len(np.unique(r.totalRays)) > 1 for r in together)   # uneven budgets
    for c in ctxs:
        c.close()


def test_batch_with_an_empty_problem_and_sample_ranges():
    ws = batch_problems(3, seed=2, triangles=(200, 400), levels=(4, 6), wavelengths=(1, 3), min_rays=1000,
                        max_rays=4000)
    ctxs = _contexts(ws)
    ctxs[1].update_beta(np.zeros(ws[1].topology.numberOfPrisms))
    alone = [c.evaluate() for c in ctxs]
    with _native.Batch(ctxs) as batch:
        together = batch.evaluate()
    for a, b in zip(alone, together):
        _same(a, b)
    assert np.all(together[1].phiAse == 0) and np.all(together[1].totalRays == 0)
    for c in ctxs:
        c.close()


def test_batch_runs_large_lds_problems_on_their_own_queue():
    """A problem of at least 8192 sample points whose records fit in LDS runs
    on its own queue inside the batch; results and ray counts are still
    those of each problem alone, and the batch statistics count its rays."""
    ws = batch_problems(3, seed=3, triangles=(200, 500), levels=(4, 8), wavelengths=(1, 4), min_rays=2000,
                        max_rays=16000, mse_threshold=2e-3, adaptive_steps=2)
    disk = production_disk(levels=4, wavelengths=4, min_rays=2048, max_rays=8192, mse_threshold=2e-3)
    assert disk.samples >= 8192 and disk.topology.numberOfTriangles <= 4800
    ctxs = _contexts([ws[0], disk] + ws[1:])
    ctxs[1].set_tuning(stage_lds=1)
    alone = [c.evaluate() for c in ctxs]
    rays_alone = sum(int(c.stats()["rays"]) for c in ctxs)
    with _native.Batch(ctxs) as batch:
        together = batch.evaluate()
        st = batch.stats()
    for a, b in zip(alone, together):
        _same(a, b)
    assert st["rays"] == rays_alone
    for c in ctxs:
        c.close()

"""hase_calc_phi_ase_noalloc against hase_calc_phi_ase on the same arguments.

The allocation-free entry takes every device buffer from the library's fixed
device scratch (HASE_NOALLOC_SCRATCH_BYTES per GPU) and runs the same kernels
on the same random streams, so it must trace the same rays: identical ray
counts and dropped rays, phi / mse / dndtAse equal up to the FP64 summation
order (relative 1e-6; the estimator sums are pooled in a fixed order, so
they are expected bit for bit).

Calls in a row share the scratch, which keeps the previous call's contents:
a production-disk range that needs many sample batches in the scratch, then
a smaller problem, then a problem too large for the scratch (rejected on the
host with the sizes in the message), then the smaller problem again."""
import numpy as np
import pytest

from haseongpu_amd import _native
from haseongpu_amd.workloads import plain_small, production_disk

pytestmark = pytest.mark.gpu


def _same(a, b, label):
    np.testing.assert_array_equal(a.totalRays, b.totalRays, err_msg=label)
    np.testing.assert_array_equal(a.droppedRays, b.droppedRays, err_msg=label)
    for name in ("phiAse", "mse", "dndtAse", "standardError"):
        x = getattr(a, name).astype(np.float64)
        y = getattr(b, name).astype(np.float64)
        rel = np.abs(x - y) / np.maximum(np.abs(x), 1e-300)
        print(f"{label}: {name} max relative difference {rel.max():.2e}", flush=True)
        assert rel.max() <= 1e-6, (label, name, np.sort(rel)[-5:])


def _production(first, last, max_reflections=1, min_rays=10_000, max_rays=100_000):
    w = production_disk()
    sa, se, msa, mse_ = w.spectra.spectralPairs(False)
    mesh = _native.NativeMesh(w.topology, w.medium)
    exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=min_rays, max_rays=max_rays, max_sigma_a=msa,
                                   max_sigma_e=mse_, mse_threshold=0.005, use_reflections=True,
                                   max_reflections=max_reflections, defensive_fraction=0.05)
    comp = _native.make_compute(n_samples=w.samples, adaptive_steps=4, min_sample_range=first,
                                max_sample_range=last, rng_seed=17)
    return mesh, exp, comp


def _small():
    w = plain_small()
    sa, se, msa, mse_ = w.spectra.spectralPairs(True)
    mesh = _native.NativeMesh(w.topology, w.medium)
    exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=4096, max_rays=65536, max_sigma_a=msa,
                                   max_sigma_e=mse_, mse_threshold=2e-4)
    comp = _native.make_compute(n_samples=w.samples, adaptive_steps=4, rng_seed=91)
    return mesh, exp, comp


def test_allocation_free_entry_matches_the_one_shot_entry():
    # 1) production disk, 600 samples over levels 9-10: 232,104 proposal
    #    entries per sample (2.8 MB of tables), so the 256 MiB scratch takes
    #    ~90 samples per batch and the range runs in several batches
    args = _production(9 * 2112 + 500, 9 * 2112 + 1099)
    print("no-alloc: production disk range, one-shot entry", flush=True)
    ref = _native.calc_phi_ase(*args)
    print("no-alloc: production disk range, allocation-free entry", flush=True)
    got = _native.calc_phi_ase_noalloc(*args)
    _same(ref, got, "production disk range")
    assert np.all(got.totalRays[9 * 2112 + 500: 9 * 2112 + 1100] >= 10_000)
    assert np.all(got.totalRays[: 9 * 2112 + 500] == 0)

    # 2) a smaller problem in the same scratch (plain slab, every sample)
    small = _small()
    print("no-alloc: plain slab, one-shot entry", flush=True)
    ref_s = _native.calc_phi_ase(*small)
    print("no-alloc: plain slab, allocation-free entry", flush=True)
    got_s = _native.calc_phi_ase_noalloc(*small)
    _same(ref_s, got_s, "plain slab")
    assert len(np.unique(got_s.totalRays)) > 1            # the schedule escalated

    # 3) too large: 301 reflection slices x 77,368 prisms x 12 B > 256 MiB
    big = _production(0, 10, max_reflections=150)
    print("no-alloc: oversized problem (expect a host-side rejection)", flush=True)
    with pytest.raises(_native.HaseError) as err:
        _native.calc_phi_ase_noalloc(*big)
    assert err.value.code == -3
    assert "no-allocation scratch" in str(err.value), str(err.value)

    # 4) the smaller problem again, after the rejection
    print("no-alloc: plain slab again", flush=True)
    again = _native.calc_phi_ase_noalloc(*small)
    _same(got_s, again, "plain slab, second call")

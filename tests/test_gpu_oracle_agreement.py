"""Independent-stream agreement with the FP64 oracle: on a reflecting,
cladded disk with 16 wavelengths, the GPU estimate of every checked sample,
in the plain form and in the learned form, lies within 3 combined standard
errors of the FP64 CPU oracle run with a different seed, sqrt(SE_gpu^2 +
SE_oracle^2) (the acceptance rule of the build).  With ~150 samples a few
beyond 3 sigma are expected from chance alone: at most 2% may lie beyond it
and none beyond 5 sigma."""
import numpy as np
import pytest

from haseongpu_amd import _native
from haseongpu_amd import proposal as prop
from haseongpu_amd.workloads import production_disk
from oracle import oracle as O
from tests.test_gpu_parity import _context, _oracle

pytestmark = pytest.mark.gpu


def _z(phi_a, se_a, phi_b, se_b):
    return (phi_a - phi_b) / np.sqrt(se_a ** 2 + se_b ** 2)


def test_plain_and_learned_forms_within_three_sigma_of_the_oracle():
    w = production_disk(levels=8, n_points=700, wavelengths=16)
    S = w.samples
    # a contiguous window of 150 samples (one level's points) keeps the CPU oracle short
    first = int(np.random.default_rng(17).integers(0, S - 150))
    last = first + 149
    picks = np.arange(first, last + 1)
    kw = dict(use_reflections=True, max_reflections=1)
    ctx, spectra = _context(w, min_rays=20000, max_rays=200000, mse_threshold=0.005, adaptive_steps=2,
                            rng_seed=11, **kw)
    net = prop.LearnedProposal(prop.init_params(2), np.zeros(10, np.float32),
                               np.array([0.3, 0.3, 2.0, 0.3, 0.3, 2.0, 0.2, 3.0, 0.3, 0.5], np.float32), 8, 2, True)
    with ctx:
        plain = ctx.evaluate()
        ctx.set_proposal(net)
        learned = ctx.evaluate()
    om, op = _oracle(w, spectra, min_rays=20000, max_rays=200000, mse_threshold=0.005, adaptive_steps=2,
                     seed=987654, sample_min=first, sample_max=last, **kw)
    ref = O.calc_phi_ase(om, op)
    phi_r = ref["phiAse"][picks]
    se_r = np.sqrt(ref["mse"][picks]) * phi_r
    assert np.all(ref["droppedRays"][picks] == 0)
    for res in (plain, learned):
        assert np.all(res.droppedRays[picks] == 0)
        z = _z(res.phiAse[picks].astype(np.float64), res.standardError[picks], phi_r, se_r)
        assert np.mean(np.abs(z) > 3) <= 0.02 and np.max(np.abs(z)) < 5, np.sort(np.abs(z))[-5:]
        assert abs(np.mean(z)) < 0.3

"""The FP64 network restatement (oracle/mlp.py), pinned before it judges the
device kernels: a closed-form network fixes the parameter layout and the
SiLU, central finite differences fix the backward pass, known bit patterns
fix the BF16 rounding, and the C oracle's mirror images fix the feature
geometry of the reflection slices."""
import math

import numpy as np
import pytest

from haseongpu_amd import proposal as prop
from oracle import mlp as M


def _silu(z):
    return z / (1.0 + math.exp(-z))


def test_parameter_count_matches_the_layout():
    assert M.COUNT == 128 * 10 + 128 + 3 * (128 * 128 + 128) + 128 + 1
    parts = M.unpack(np.arange(M.COUNT, dtype=np.float64))
    assert parts["W1"][0, 1] == 1 and parts["W1"][1, 0] == 10          # row-major [out][in]
    assert parts["b1"][0] == 1280 and parts["W2"][0, 0] == 1408
    assert parts["w5"][0] == M.COUNT - 129 and parts["b5"][0] == M.COUNT - 1


def test_closed_form_network():
    parts = {n: np.zeros(s) for n, s in zip(M.NAMES, M.SIZES)}
    parts["W1"][5, 3] = 2.0
    parts["b1"][5] = 0.25
    parts["W2"][7, 5] = -1.5
    parts["W3"][0, 7] = 1.0
    parts["b3"][0] = 0.1
    parts["W4"][127, 0] = 3.0
    parts["w5"][127] = 0.5
    parts["b5"][0] = -0.2
    p = M.pack(parts)
    x = np.random.default_rng(1).normal(size=(6, 10))
    got = M.forward(p, x)
    for i in range(6):
        a1 = _silu(2.0 * x[i, 3] + 0.25)
        a2 = _silu(-1.5 * a1)
        a3 = _silu(a2 + 0.1)
        a4 = _silu(3.0 * a3)
        assert got[i] == pytest.approx(0.5 * a4 - 0.2, rel=1e-13, abs=1e-14)


def test_backward_matches_central_differences():
    rng = np.random.default_rng(7)
    p = M.init_params(3) * 0.7
    p[M.COUNT - 129 - 128:M.COUNT - 129] = rng.normal(0, 0.1, 128)   # nonzero b4
    x = rng.normal(size=(5, 10))
    t = rng.normal(size=5)
    w = rng.uniform(0.5, 2.0, size=5)
    loss, g = M.loss_gradient(p, x, t, w)
    y = M.forward(p, x)
    assert loss == pytest.approx(float((w * (y - t) ** 2).sum() / w.sum()), rel=1e-14)
    idx = np.unique(np.concatenate([rng.integers(0, M.COUNT, 120),
                                    [0, 1279, 1280, 1407, 1408, 17792, 50816, 50943, 50944, 51071, 51072]]))
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (M.loss_gradient(pp, x, t, w)[0] - M.loss_gradient(pm, x, t, w)[0]) / (2 * eps)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def test_bf16_rounding_is_nearest_even():
    one = 1.0
    assert M.to_bf16(one + 2.0 ** -8) == one                      # tie -> even (mantissa ...0)
    assert M.to_bf16(one + 3 * 2.0 ** -8) == one + 2.0 ** -6      # tie -> even (mantissa ...10)
    assert M.to_bf16(one + 2.0 ** -8 + 2.0 ** -20) == one + 2.0 ** -7
    assert M.to_bf16(-3.0) == -3.0
    assert M.to_bf16(1e-3) == pytest.approx(1e-3, rel=2 ** -8)


def test_bf16_forward_stays_close_to_fp64():
    p = M.init_params(5)
    x = np.random.default_rng(2).normal(size=(400, 10))
    y, yb = M.forward(p, x), M.forward_bf16(p, x)
    assert np.max(np.abs(yb - y) / (1.0 + np.abs(y))) < 0.05
    assert np.median(np.abs(yb - y) / (1.0 + np.abs(y))) < 0.01


def test_proposal_inputs_geometry():
    f = prop.proposal_inputs(1.0, 2.0, 0.3, 0.5, 1.0, 2, 0.25, 0.1, 3.0)
    assert f.shape == (10,)
    cz = 2.5 * 0.1
    assert f[5] == pytest.approx(cz, rel=1e-6)
    assert f[8] == pytest.approx(math.sqrt(0.25 + 1.0 + (0.3 - cz) ** 2), rel=1e-6)
    assert list(f[[0, 1, 2, 3, 4, 6, 7, 9]]) == pytest.approx([1.0, 2.0, 0.3, 0.5, 1.0, 2.0, 0.25, 3.0])
    np.testing.assert_array_equal(f, M.proposal_inputs(1.0, 2.0, 0.3, 0.5, 1.0, 2, 0.25, 0.1, 3.0))


@pytest.mark.parametrize("reflections,first", [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (3, 1)])
def test_image_planes_match_the_c_oracle(reflections, first):
    from haseongpu_amd.workloads import plain_small
    from oracle import oracle as O

    w = plain_small(rays=10)
    om = O.OracleMesh(w.topology, w.medium)
    layers = int(w.topology.levels) - 1
    h = float(w.topology.thickness)
    for lv in (0, 3, layers):
        z = O.image_point(om, np.array([0.3, 0.4, lv * h]), reflections, first)[2]
        assert prop.image_plane(lv, layers, reflections, first) * h == pytest.approx(z, abs=1e-12)


def test_normalisation_handles_constant_inputs():
    X = np.random.default_rng(0).normal(size=(1000, 10)).astype(np.float32)
    X[:, 9] = 3.0
    shift, scale = prop.fit_normalisation(X)
    Xn = (X - shift) * scale
    assert np.allclose(Xn[:, :9].std(axis=0), 1.0, atol=1e-3)
    assert np.all(Xn[:, 9] == 0.0) and scale[9] == 1.0


def test_init_params_layout():
    p = prop.init_params(0)
    parts = M.unpack(p.astype(np.float64))
    for b in ("b1", "b2", "b3", "b4", "b5"):
        assert np.all(parts[b] == 0)
    assert abs(parts["W2"].std() - math.sqrt(2.0 / 128)) < 0.01
    assert abs(parts["W1"].std() - math.sqrt(2.0 / 10)) < 0.05

"""Pins of the time-stepped loop's oracle (oracle/time_stepping.py), on the CPU.

Each integrator is checked against what the mathematics fixes for a linear
population equation  d beta/dt = a - k beta  (pump a - b beta, constant flux,
fluorescence 1/tau; k = b + (sigma_e + sigma_a) phi + 1/tau): one step is the
exact solution's decay factor replaced by the method's stability polynomial,
the global error falls with the method's order, the fixed-point implicit
Euler converges to the backward-Euler solution with the iteration error
(-h k)^m, and the exponential Euler step is exact when the only
beta-dependent term is the decay.  The prism flux (six-vertex mean) is
pinned by the wedge integrals recorded in the legacy time-series archive.
"""
import math

import numpy as np
import pytest

from oracle import time_stepping as TS
from tests.helpers import GOLDEN

SIGMA_A, SIGMA_E, TAU = 1.16e-21, 2.48e-20, 9.5e-4


def _two_prisms():
    """One triangle, two layers (two prisms), three points per level."""
    tri = np.array([[0, 1, 2]])
    return tri, 3, 2


def _loop(integrator, *, phi=0.0, a=(300.0, 900.0), b=(200.0, 50.0), beta=(0.1, 0.7), h=1e-4, cladding=None,
          ase_steps=100, **kw):
    tri, n_points, layers = _two_prisms()
    flux = np.full(n_points * (layers + 1), float(phi))
    return TS.TimeLoop(tri=tri, n_points=n_points, layers=layers, beta=np.array(beta), pump_rate=np.array(a),
                       pump_saturation=np.array(b), cladding=cladding, sigma_e=SIGMA_E, sigma_a=SIGMA_A, tau=TAU,
                       time_step=h, integrator=integrator, ase=lambda _b: flux, ase_steps=ase_steps,
                       pump_steps=100, **kw)


def _linear(phi=0.0, a=(300.0, 900.0), b=(200.0, 50.0)):
    a, b = np.array(a), np.array(b)
    k = b + (SIGMA_E + SIGMA_A) * phi + 1.0 / TAU
    src = a + SIGMA_A * phi
    return src, k


def _exact(beta0, t, phi=0.0, **kw):
    src, k = _linear(phi, **kw)
    inf = src / k
    return inf + (np.array(beta0) - inf) * np.exp(-k * t)


POLY = {
    "explicit-euler": lambda z: 1 + z,
    "heun": lambda z: 1 + z + z ** 2 / 2,
    "midpoint": lambda z: 1 + z + z ** 2 / 2,
    "runge-kutta-4": lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24,
    "frozen-phi-ase-runge-kutta-4": lambda z: 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24,
}


@pytest.mark.parametrize("integrator", sorted(POLY))
@pytest.mark.parametrize("phi", [0.0, 4.0e21])
def test_one_step_is_the_stability_polynomial(integrator, phi):
    h, beta0 = 1e-4, np.array([0.1, 0.7])
    loop = _loop(integrator, phi=phi, h=h).step(1)
    src, k = _linear(phi)
    inf = src / k
    expect = inf + (beta0 - inf) * POLY[integrator](-k * h)
    np.testing.assert_allclose(loop.beta, expect, rtol=1e-13)


@pytest.mark.parametrize("integrator,order", [("explicit-euler", 1), ("heun", 2), ("midpoint", 2),
                                              ("runge-kutta-4", 4)])
def test_global_error_falls_with_the_order(integrator, order):
    t_end, beta0 = 2e-3, (0.1, 0.7)
    errs = []
    for n in (8, 16, 32):
        loop = _loop(integrator, phi=2e21, h=t_end / n).step(n)
        errs.append(np.abs(loop.beta - _exact(beta0, t_end, phi=2e21)).max())
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert all(abs(r - order) < 0.25 for r in rates), (errs, rates)


@pytest.mark.parametrize("iterations", [1, 2, 3, 8])
def test_implicit_euler_fixed_point_iterations(iterations):
    h, beta0 = 1e-4, np.array([0.1, 0.7])
    loop = _loop("implicit-euler", phi=1e21, h=h, implicit_iterations=iterations).step(1)
    src, k = _linear(1e21)
    backward = (beta0 + h * src) / (1.0 + h * k)          # the implicit Euler solution
    expect = backward + (-h * k) ** iterations * (beta0 - backward)
    np.testing.assert_allclose(loop.beta, expect, rtol=1e-12)


def test_exponential_euler_is_exact_for_pure_decay_with_constant_pump():
    beta0 = (0.1, 0.7)
    loop = _loop("exponential-euler", b=(0.0, 0.0), h=3e-4).step(5)
    np.testing.assert_allclose(loop.beta, _exact(beta0, 1.5e-3, b=(0.0, 0.0)), rtol=1e-12)


def test_frozen_rk4_equals_rk4_for_a_constant_flux_and_evaluates_once():
    a = _loop("runge-kutta-4", phi=3e21).step(4)
    f = _loop("frozen-phi-ase-runge-kutta-4", phi=3e21).step(4)
    np.testing.assert_array_equal(a.beta, f.beta)
    assert a.ase_evaluations == 16 and f.ase_evaluations == 4


def test_frozen_rk4_differs_when_the_flux_follows_beta():
    def loop(integrator):
        tri, n_points, layers = _two_prisms()
        return TS.TimeLoop(tri=tri, n_points=n_points, layers=layers, beta=np.array([0.1, 0.7]),
                           pump_rate=np.zeros(2), sigma_e=SIGMA_E, sigma_a=SIGMA_A, tau=TAU, time_step=2e-4,
                           integrator=integrator, ase=lambda b: np.full(9, 1e23 * float(np.mean(b))),
                           ase_steps=10)
    a, f = loop("runge-kutta-4").step(1), loop("frozen-phi-ase-runge-kutta-4").step(1)
    assert np.abs(a.beta - f.beta).max() > 1e-6


def test_ase_evaluations_per_step():
    counts = {"explicit-euler": 1, "heun": 2, "midpoint": 2, "runge-kutta-4": 4,
              "frozen-phi-ase-runge-kutta-4": 1, "exponential-euler": 1}
    for m, c in counts.items():
        assert _loop(m, phi=1e21).step(3).ase_evaluations == 3 * c, m
    assert _loop("implicit-euler", phi=1e21, implicit_iterations=3).step(2).ase_evaluations == 6


def test_activity_windows_and_pre_pump():
    # pre_pump: no ASE in step 0; ASE in steps 1 and 2; the pump in steps 0 and 1
    loop = _loop("explicit-euler", phi=1e21, ase_steps=3, pre_pump=True)
    loop.pump_steps = 2
    loop.step(1)
    assert loop.ase_evaluations == 0 and np.all(loop.dndt_ase == 0.0) and np.any(loop.dndt_pump != 0.0)
    loop.step(2)
    assert loop.ase_evaluations == 2 and np.all(loop.dndt_pump == 0.0)
    loop.step(1)
    assert loop.ase_evaluations == 2 and np.all(loop.phi == 0.0)
    # first_simulation_step shifts the windows
    late = _loop("explicit-euler", phi=1e21, ase_steps=3, first_simulation_step=3)
    late.step(2)
    assert late.ase_evaluations == 0


def test_cladding_prisms_hold_and_beta_is_clipped():
    loop = _loop("explicit-euler", phi=1e21, a=(1e6, 1e6), b=(0.0, 0.0), cladding=np.array([False, True]), h=1e-3)
    loop.step(1)
    assert loop.beta[0] == 1.0          # clipped from above
    assert loop.beta[1] == 0.7          # cladding: derivative zero
    neg = _loop("explicit-euler", a=(-1e6, 0.0), b=(0.0, 0.0), h=1e-3).step(1)
    assert neg.beta[0] == 0.0


def test_ase_term_sign_and_emission_peak_cross_section():
    # sigma_a is taken where sigma_e peaks, not at the sigma_a maximum
    assert TS.absorption_at_emission_peak([5.0, 1.0, 3.0], [1.0, 9.0, 2.0], 5.0) == 1.0
    assert TS.absorption_at_emission_peak([], [], 7.0) == 7.0
    # inverted medium (gain > 0): ASE depletes; beta below transparency: ASE pumps
    loop = _loop("explicit-euler", phi=1e22, a=(0.0, 0.0), b=(0.0, 0.0), beta=(0.9, 0.0)).step(1)
    assert loop.dndt_ase[0] > 0 and loop.dndt_ase[1] < 0
    np.testing.assert_allclose(loop.dndt_ase[1], -SIGMA_A * 1e22, rtol=1e-15)


def test_prism_flux_reproduces_the_legacy_wedge_integrals():
    d = np.load(GOLDEN / "legacyPump3Wedge.npz")
    tri, levels, h = d["triangles"], int(d["levels"]), float(d["thickness"])
    vol = TS.wedge_volumes(d["points"], tri, h, levels - 1)
    np.testing.assert_allclose(vol.sum(), float(d["totalWedgeVolume"]), rtol=1e-12)
    integrals = [float(np.sum(TS.prism_flux(phi, tri, d["points"].shape[0], levels - 1) * vol))
                 for phi in d["phiASE"]]
    np.testing.assert_allclose(integrals, d["integrals"], rtol=1e-12, atol=0.0)
    np.testing.assert_allclose(d["phiASE"].mean(axis=1), d["meanPhi"], rtol=0, atol=0)


def test_prism_flux_of_a_linear_field_is_its_centroid_value():
    # a field linear in (x, y, z) averages over the six vertices to its value
    # at the prism centroid
    rng = np.random.default_rng(3)
    pts = rng.random((5, 2))
    tri = np.array([[0, 1, 2], [1, 3, 2], [2, 3, 4]])
    L, h = 4, 0.2
    z = np.repeat(np.arange(L) * h, 5)
    xy = np.tile(pts, (L, 1))
    f = lambda x, y, zz: 1.5 + 2.0 * x - 3.0 * y + 0.7 * zz   # noqa: E731
    phi = f(xy[:, 0], xy[:, 1], z)
    got = TS.prism_flux(phi, tri, 5, L - 1)
    cen = pts[tri].mean(axis=1)
    expect = np.concatenate([f(cen[:, 0], cen[:, 1], (l + 0.5) * h) for l in range(L - 1)])
    np.testing.assert_allclose(got, expect, rtol=1e-13)


def test_rejects_bad_controls():
    with pytest.raises(ValueError):
        _loop("leapfrog")
    with pytest.raises(ValueError):
        _loop("heun", h=0.0)
    with pytest.raises(ValueError):
        _loop("heun").step(0)

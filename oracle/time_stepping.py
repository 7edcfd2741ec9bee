"""CPU restatement of the time-stepped beta loop (FP64, numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP loop
(hase_simulation_*), never as the thing measured or shipped.

Follows the reference's CompiledSimulationRunner
(include/core/timeSteppedSimulation.hpp:132-615) on the prism mesh:

* state: beta per prism j = t + l * T, FP64;
* stage (the evaluateStage lambda, timeSteppedSimulation.hpp:255-314): the pump
  term at the stage beta, optionally a fresh ASE flux (zero when ASE is off
  in this step), then ComposeDerivative (include/kernels/derivativeComposition.hpp:34-59);
* prism flux: the six-vertex mean of the sample flux, the legacy wedge mean of
  tests/data/laserPumpCladding/fixed_legacy_no_reflection_pump3_wedge_reference/generate_reference.py:138-141;
* integrators (timeSteppedSimulation.hpp:319-415) with the update kernels of
  include/kernels/timeIntegrationUpdateKernels.hpp:21-123 and the final clip
  of beta_next (:364);
* activity windows (timeSteppedSimulation.hpp:194-197).

The ASE flux comes from a caller-supplied function ``ase(beta) -> phi``
(phi per sample point); ``oracle_ase`` wraps the FP64 C oracle for it.

Every array operation is written in the order of the reference's
expressions, so that the device kernels (built without floating-point
contraction) round identically on identical inputs.
"""
from __future__ import annotations

import math

import numpy as np

INTEGRATORS = ("explicit-euler", "heun", "midpoint", "runge-kutta-4", "frozen-phi-ase-runge-kutta-4",
               "implicit-euler", "exponential-euler")


def absorption_at_emission_peak(sigma_a, sigma_e, max_sigma_a):
    """sigma_a at the (first) maximum of sigma_e, max_sigma_a without spectra
    (timeSteppedSimulation.hpp:45-54)."""
    sigma_a = np.atleast_1d(np.asarray(sigma_a, dtype=np.float64))
    sigma_e = np.atleast_1d(np.asarray(sigma_e, dtype=np.float64))
    if sigma_a.size == 0 or sigma_e.size == 0:
        return float(max_sigma_a)
    return float(sigma_a[min(int(np.argmax(sigma_e)), sigma_a.size - 1)])


def prism_flux(phi_samples, tri, n_points, layers):
    """Six-vertex mean of the sample flux per prism j = t + l * T.

    ``tri`` is (T, 3) vertex indices; sample s = p + l * n_points.  The sum
    runs lower vertices 0, 1, 2 then upper vertices 0, 1, 2."""
    phi = np.asarray(phi_samples).astype(np.float64)
    tri = np.asarray(tri, dtype=np.int64)
    T = tri.shape[0]
    out = np.empty(T * layers)
    for l in range(layers):
        lo = phi[l * n_points + tri]
        hi = phi[(l + 1) * n_points + tri]
        s = lo[:, 0] + lo[:, 1] + lo[:, 2] + hi[:, 0] + hi[:, 1] + hi[:, 2]
        out[l * T:(l + 1) * T] = s / 6.0
    return out


def wedge_volumes(points_xy, tri, thickness, layers):
    """Prism volumes V_j = area(t) * thickness, j = t + l * T."""
    p = np.asarray(points_xy, dtype=np.float64)
    tri = np.asarray(tri, dtype=np.int64)
    a, b, c = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    area = 0.5 * np.abs((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]))
    return np.tile(area * thickness, layers)


def compose(beta, phi, pump_rate, pump_saturation, cladding, sigma_sum, sigma_a, tau, pump_on):
    """(dndt_pump, dndt_ase, derivative) of derivativeComposition.hpp:34-59."""
    if pump_on:
        pump = pump_rate - pump_saturation * beta if pump_saturation is not None else pump_rate.copy()
    else:
        pump = np.zeros_like(beta)
    gain = beta * sigma_sum - sigma_a
    ase = gain * phi
    deriv = pump - ase - beta / tau
    for arr in (pump, ase, deriv):
        arr[cladding] = 0.0
    return pump, ase, deriv


def add_scaled(base, slope, scale):
    """AddScaled (timeIntegrationUpdateKernels.hpp:21-40)."""
    return base + scale * slope


def combine_heun(base, first, second, h):
    """CombineHeun (timeIntegrationUpdateKernels.hpp:42-62)."""
    return base + 0.5 * h * (first + second)


def combine_rk4(base, k1, k2, k3, k4, h):
    """CombineRungeKutta4 (timeIntegrationUpdateKernels.hpp:64-86)."""
    return base + (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


def exponential_euler(beta, dndt_pump, dndt_ase, h, tau):
    """ExponentialEulerUpdate (timeIntegrationUpdateKernels.hpp:88-111)."""
    decay = math.exp(-h / tau)
    source = dndt_pump - dndt_ase
    return tau * source * (1.0 - decay) + beta * decay


def clip(beta):
    """ClipBeta (timeIntegrationUpdateKernels.hpp:113-123)."""
    return np.minimum(1.0, np.maximum(0.0, beta))


class TimeLoop:
    """The loop of timeSteppedSimulation.hpp:177-241 on prism arrays.

    ``ase(beta) -> phi_samples`` evaluates the ASE flux of a beta state;
    ``cladding`` is a boolean mask over prisms."""

    def __init__(self, *, tri, n_points, layers, beta, pump_rate, pump_saturation=None, cladding=None,
                 sigma_e, sigma_a, tau, time_step, integrator, ase, implicit_iterations=8,
                 first_simulation_step=0, ase_steps=0, pump_steps=0, pre_pump=False):
        if integrator not in INTEGRATORS:
            raise ValueError(f"unsupported time integrator '{integrator}'")
        if not time_step > 0.0:
            raise ValueError("time_step must be positive")
        self.tri = np.asarray(tri, dtype=np.int64)
        self.np_ = int(n_points)
        self.layers = int(layers)
        P = self.tri.shape[0] * self.layers
        self.beta = np.array(beta, dtype=np.float64).reshape(P)
        self.pump_rate = np.asarray(pump_rate, dtype=np.float64).reshape(P)
        self.pump_sat = None if pump_saturation is None else np.asarray(pump_saturation, np.float64).reshape(P)
        self.cladding = np.zeros(P, bool) if cladding is None else np.asarray(cladding, bool).reshape(P)
        self.sigma_sum = float(sigma_e) + float(sigma_a)
        self.sigma_a = float(sigma_a)
        self.tau = max(float(tau), np.finfo(np.float64).tiny)
        self.h = float(time_step)
        self.integrator = integrator
        self.ase = ase
        self.implicit_iterations = int(implicit_iterations)
        self.first = int(first_simulation_step)
        self.ase_steps = int(ase_steps)
        self.pump_steps = int(pump_steps)
        self.pre_pump = bool(pre_pump)
        self.completed = 0
        self.ase_evaluations = 0
        self.phi = np.zeros(P)
        self.dndt_pump = np.zeros(P)
        self.dndt_ase = np.zeros(P)
        self.deriv = np.zeros(P)
        self.phi_samples = np.zeros(self.np_ * (self.layers + 1))

    def _stage(self, beta, pump_on, ase_on, refresh):
        if refresh:
            if ase_on:
                self.phi_samples = np.asarray(self.ase(beta))
                self.ase_evaluations += 1
                self.phi = prism_flux(self.phi_samples, self.tri, self.np_, self.layers)
            else:
                self.phi_samples = np.zeros(self.np_ * (self.layers + 1))
                self.phi = np.zeros_like(beta)
        self.dndt_pump, self.dndt_ase, self.deriv = compose(
            beta, self.phi, self.pump_rate, self.pump_sat, self.cladding, self.sigma_sum, self.sigma_a, self.tau,
            pump_on)
        return self.deriv

    def _advance(self, pump_on, ase_on):
        h, beta = self.h, self.beta
        f = lambda b, refresh=True: self._stage(b, pump_on, ase_on, refresh)   # noqa: E731
        m = self.integrator
        if m == "explicit-euler":
            nxt = add_scaled(beta, f(beta), h)
        elif m == "heun":
            k1 = f(beta).copy()
            stage = add_scaled(beta, k1, h)
            nxt = combine_heun(beta, k1, f(stage), h)
        elif m == "midpoint":
            stage = add_scaled(beta, f(beta), 0.5 * h)
            nxt = add_scaled(beta, f(stage), h)
        elif m in ("runge-kutta-4", "frozen-phi-ase-runge-kutta-4"):
            refresh = m == "runge-kutta-4"
            k1 = f(beta).copy()
            k2 = f(add_scaled(beta, k1, 0.5 * h), refresh).copy()
            k3 = f(add_scaled(beta, k2, 0.5 * h), refresh).copy()
            k4 = f(add_scaled(beta, k3, h), refresh).copy()
            nxt = combine_rk4(beta, k1, k2, k3, k4, h)
        elif m == "implicit-euler":
            stage = beta.copy()
            for _ in range(max(1, self.implicit_iterations)):
                nxt = add_scaled(beta, f(stage), h)
                stage = nxt.copy()
        else:
            f(beta)
            nxt = exponential_euler(beta, self.dndt_pump, self.dndt_ase, h, self.tau)
        self.beta = clip(nxt)

    def step(self, n=1):
        if n <= 0:
            raise ValueError("steps must be positive")
        for _ in range(n):
            s = self.first + self.completed
            pump_on = s < self.pump_steps
            ase_on = s < self.ase_steps and not (self.pre_pump and s == 0)
            self._advance(pump_on, ase_on)
            self.completed += 1
        return self


def oracle_ase(mesh, params):
    """``ase(beta)`` from the FP64 C oracle: phi per sample point for the
    prism beta of each stage (same settings and seed at every call)."""
    from . import oracle as O

    def ase(beta):
        mesh.set_beta(np.ascontiguousarray(beta, dtype=np.float64))
        return O.calc_phi_ase(mesh, params)["phiAse"]

    return ase

"""Time-integrator descriptors and the local pump of the time-stepped loop.

The descriptors mirror pyInclude/timeIntegration.py:15-75 (same class names,
the same ``.name`` strings as include/core/simulationRunControl.hpp:65-74);
the library's loop (hase_simulation_*) executes the named integrator.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["TimeIntegrationSolver", "ExplicitEuler", "Heun", "Midpoint", "RungeKutta4",
           "FrozenPhiAseRungeKutta4", "ImplicitEuler", "ExponentialEuler", "LocalPump"]


@dataclass
class TimeIntegrationSolver:
    """Names a beta time integrator of the library's loop."""

    name: str


class ExplicitEuler(TimeIntegrationSolver):
    """First-order explicit Euler."""

    def __init__(self):
        super().__init__("explicit-euler")


class Heun(TimeIntegrationSolver):
    """Second-order predictor-corrector."""

    def __init__(self):
        super().__init__("heun")


class Midpoint(TimeIntegrationSolver):
    """Second-order midpoint rule."""

    def __init__(self):
        super().__init__("midpoint")


class RungeKutta4(TimeIntegrationSolver):
    """Classical fourth-order Runge-Kutta (four ASE evaluations per step)."""

    def __init__(self):
        super().__init__("runge-kutta-4")


class FrozenPhiAseRungeKutta4(TimeIntegrationSolver):
    """RK4 that reuses the first stage's ASE flux in the later stages."""

    def __init__(self):
        super().__init__("frozen-phi-ase-runge-kutta-4")


class ImplicitEuler(TimeIntegrationSolver):
    """Implicit Euler by a fixed number of fixed-point iterations."""

    def __init__(self, iterations=8, tolerance=1e-10):
        super().__init__("implicit-euler")
        self.iterations = int(iterations)
        self.tolerance = float(tolerance)


class ExponentialEuler(TimeIntegrationSolver):
    """Exponential Euler: exact fluorescence decay, pump and ASE frozen per step."""

    def __init__(self):
        super().__init__("exponential-euler")


@dataclass
class LocalPump:
    """Pump acting locally on every prism, linear in beta:

        d beta/dt |pump = rate - saturation * beta     [1/s]

    which is the rate equation of a two-level absorber under a pump of fixed
    intensity I (rate = sigma_a,p I / (h nu), saturation = (sigma_a,p +
    sigma_e,p) I / (h nu)).  ``rate`` and ``saturation`` have one value per
    prism (shape (T, L-1) or flat j = t + l T); ``pump_steps`` is its
    activity window: the pump acts in steps < pump_steps, and None counts as
    0, as the reference's pump_steps does (pyInclude/simulation.py:707).
    It stands in for the reference's ray-traced surface pump
    (include/kernels/generalPump.hpp), which this library does not include."""

    rate: np.ndarray
    saturation: np.ndarray | None = None
    pump_steps: int | None = None
    name: str = "pump"

    def arrays(self, prisms):
        rate = _prism_flat(self.rate, prisms, "rate")
        sat = None if self.saturation is None else _prism_flat(self.saturation, prisms, "saturation")
        return rate, sat

    def active(self, step):
        return step < (0 if self.pump_steps is None else int(self.pump_steps))


def _prism_flat(values, prisms, name):
    arr = np.asarray(values, dtype=np.float64)
    flat = arr.reshape(-1, order="F") if arr.ndim == 2 else arr.reshape(-1)
    if flat.size != prisms:
        raise ValueError(f"pump {name} needs {prisms} values (one per prism), got {flat.size}")
    if not np.all(np.isfinite(flat)):
        raise ValueError(f"pump {name} must be finite")
    return np.ascontiguousarray(flat)

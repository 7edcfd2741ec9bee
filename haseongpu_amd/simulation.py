"""``PhiASE``: the Python front end of the importance-sampled ASE estimator.

Mirrors the reference's runner ``PhiASE(...).run(gainMedium, crossSections)``
/ ``getResults()`` (pyInclude/simulation.py:111-491), with the legacy run
controls the backward estimator used (minRaysPerSample, maxRaysPerSample,
mseThreshold, repetitions, adaptiveSteps, useReflections, monochromatic,
rngSeed - the keys of
tests/data/laserPumpCladding/fixed_legacy_no_reflection_pump3_wedge_reference/phiASE-no-reflections.yaml).
Instead of serialising to openPMD and spawning the calcPhiASE binary
(pyInclude/openpmd/transport.py:1672-1717) it calls the HIP library through
its C ABI in-process.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _native
from .geometry import GainMedium
from .laser import CrossSectionData

__all__ = ["PhiASE", "Result", "native_problem"]


@dataclass
class Result:
    """Per-sample outputs (pyInclude/structures.py:8-18 plus the legacy mse)."""

    phiAse: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    mse: np.ndarray = field(default_factory=lambda: np.zeros(0))
    totalRays: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    dndtAse: np.ndarray = field(default_factory=lambda: np.zeros(0))
    standardError: np.ndarray = field(default_factory=lambda: np.zeros(0))
    relativeStandardError: np.ndarray = field(default_factory=lambda: np.zeros(0))
    droppedRays: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint32))
    stats: dict = field(default_factory=dict)


_CONFIG_KEYS = {
    "minRaysPerSample": "minRaysPerSample", "minRays": "minRaysPerSample", "min_rays": "minRaysPerSample",
    "maxRaysPerSample": "maxRaysPerSample", "maxRays": "maxRaysPerSample", "max_rays": "maxRaysPerSample",
    "mseThreshold": "mseThreshold", "mse_threshold": "mseThreshold",
    "repetitions": "repetitions", "adaptiveSteps": "adaptiveSteps", "adaptive_steps": "adaptiveSteps",
    "useReflections": "useReflections", "use_reflections": "useReflections",
    "maxReflections": "maxReflections", "monochromatic": "monochromatic",
    "rngSeed": "rngSeed", "rng_seed": "rngSeed", "minSampleRange": "minSampleRange",
    "maxSampleRange": "maxSampleRange", "device": "device", "defensiveFraction": "defensiveFraction",
    "raysPerThread": "raysPerThread", "sampleBatch": "sampleBatch",
}


@dataclass
class PhiASE:
    config: object | None = None
    crossSections: CrossSectionData | None = None
    spectralProperties: CrossSectionData | None = None
    gainMedium: GainMedium | None = None
    minRaysPerSample: int = 100000
    maxRaysPerSample: int = 100000
    mseThreshold: float = 0.005
    repetitions: int = 1
    adaptiveSteps: int = 4
    useReflections: bool = False
    maxReflections: int = 1
    monochromatic: bool = False
    rngSeed: int | None = None
    minSampleRange: int | None = None
    maxSampleRange: int | None = None
    device: int = 0
    defensiveFraction: float = 0.05
    raysPerThread: int = 0
    sampleBatch: int = 0
    proposal: object | None = None   # a proposal.LearnedProposal; None = walked importance map

    _result: Result | None = field(default=None, init=False, repr=False)

    def __post_init__(self):
        if isinstance(self.config, (str, Path)):
            self._applyConfig(self._loadConfig(self.config))
        elif isinstance(self.config, dict):
            self._applyConfig(self.config)
        if self.crossSections is None and self.spectralProperties is not None:
            self.crossSections = self.spectralProperties
        if self.spectralProperties is None:
            self.spectralProperties = self.crossSections

    @staticmethod
    def _loadConfig(filename):
        import yaml

        with open(filename, "r", encoding="utf-8") as handle:
            return yaml.safe_load(handle) or {}

    def _applyConfig(self, cfg):
        flat = {}
        for key, value in cfg.items():
            if isinstance(value, dict):
                flat.update(value)
            else:
                flat[key] = value
        for key, value in flat.items():
            if key in _CONFIG_KEYS:
                setattr(self, _CONFIG_KEYS[key], value)

    @classmethod
    def fromYaml(cls, filename, **overrides):
        obj = cls(filename)
        for name, value in overrides.items():
            setattr(obj, name, value)
        return obj

    def _validate(self, medium, spectra):
        if not isinstance(medium, GainMedium):
            raise TypeError("run() needs a GainMedium")
        if spectra is None:
            raise ValueError("PhiASE needs crossSections")
        if int(self.minRaysPerSample) <= 0:
            raise ValueError("minRaysPerSample must be positive")
        if int(self.maxRaysPerSample) < int(self.minRaysPerSample):
            raise ValueError("maxRaysPerSample must be >= minRaysPerSample")
        if float(self.mseThreshold) < 0.0:
            raise ValueError("mseThreshold must be non-negative")
        if int(self.repetitions) < 1:
            raise ValueError("repetitions must be >= 1")
        if int(self.adaptiveSteps) < 0:
            raise ValueError("adaptiveSteps must be non-negative")
        for name in ("nTot", "crystalTFluo"):
            if not float(medium.value(name)) > 0.0:
                raise ValueError(f"GainMedium.{name} must be positive")

    def nativeArguments(self, gainMedium=None, crossSections=None):
        """(mesh, experiment, compute) C-ABI structures for a medium and spectra."""
        medium = gainMedium if gainMedium is not None else self.gainMedium
        spectra = crossSections if crossSections is not None else self.crossSections
        self._validate(medium, spectra)
        topo = medium.topology
        S = topo.numberOfSamples
        sa, se, msa, mse_ = spectra.spectralPairs(bool(self.monochromatic))
        mesh = _native.NativeMesh(topo, medium)
        exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=int(self.minRaysPerSample),
                                       max_rays=int(self.maxRaysPerSample), max_sigma_a=msa, max_sigma_e=mse_,
                                       mse_threshold=float(self.mseThreshold),
                                       use_reflections=bool(self.useReflections),
                                       max_reflections=int(self.maxReflections),
                                       defensive_fraction=float(self.defensiveFraction))
        seed = 5489 if self.rngSeed is None else int(self.rngSeed)
        comp = _native.make_compute(n_samples=S, max_repetitions=int(self.repetitions),
                                    adaptive_steps=int(self.adaptiveSteps),
                                    min_sample_range=0 if self.minSampleRange is None else int(self.minSampleRange),
                                    max_sample_range=None if self.maxSampleRange is None else int(self.maxSampleRange),
                                    rng_seed=seed, device=int(self.device), rays_per_thread=int(self.raysPerThread),
                                    sample_batch=int(self.sampleBatch))
        return mesh, exp, comp

    def run(self, gainMedium=None, crossSections=None):
        mesh, exp, comp = self.nativeArguments(gainMedium, crossSections)
        with _native.Context(mesh, exp, comp) as ctx:
            if self.proposal is not None:
                ctx.set_proposal(self.proposal)
            res = ctx.evaluate()
            stats = ctx.stats()
            if self.proposal is not None:
                stats["proposal"] = ctx.proposal_stats()
        rse = np.sqrt(np.where(np.isfinite(res.mse) & (res.mse >= 0), res.mse, np.nan))
        self._result = Result(phiAse=res.phiAse, mse=res.mse, totalRays=res.totalRays, dndtAse=res.dndtAse,
                              standardError=res.standardError, relativeStandardError=rse,
                              droppedRays=res.droppedRays, stats=stats)
        return self._result

    def getResults(self):
        if self._result is None:
            raise RuntimeError("PhiASE.run() has not been called")
        return self._result


def native_problem(workload, device=0, **overrides):
    """C-ABI structures of a seeded workload plus log(max_sigma_e / max_sigma_a)."""
    phi = PhiASE(crossSections=workload.spectra, **{**workload.settings, "device": device, **overrides})
    mesh, exp, comp = phi.nativeArguments(workload.medium, workload.spectra)
    log_sigma = float(np.log(exp.struct.max_sigma_e / exp.struct.max_sigma_a))
    return mesh, exp, comp, log_sigma

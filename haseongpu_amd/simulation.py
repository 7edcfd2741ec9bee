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

import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _native
from .geometry import GainMedium
from .laser import CrossSectionData

__all__ = ["PhiASE", "Result", "native_problem", "Simulation", "TimeSteppedSimulation", "TimeStepState",
           "SIMULATION_OUTPUT_FIELDS", "SIMULATION_CONTROL_FIELDS"]


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
    # legacy backward-estimator controls (per sample point)
    "minRaysPerSample": "minRaysPerSample", "min_rays_per_sample": "minRaysPerSample",
    "maxRaysPerSample": "maxRaysPerSample", "max_rays_per_sample": "maxRaysPerSample",
    "mseThreshold": "mseThreshold", "mse_threshold": "mseThreshold",
    # the 2.2.0 names (pyInclude/simulation.py:130-175, 296-339)
    "minRays": "minRays", "min_rays": "minRays", "maxRays": "maxRays", "max_rays": "maxRays",
    "relativeStandardErrorThreshold": "relativeStandardErrorThreshold",
    "relative_standard_error_threshold": "relativeStandardErrorThreshold",
    "forwardRayCount": "forwardRayCount", "forward_ray_count": "forwardRayCount",
    "propagationMode": "propagationMode", "propagation_mode": "propagationMode",
    "reflectionMaxIterations": "reflectionMaxIterations", "reflection_max_iterations": "reflectionMaxIterations",
    "reflectionTolerance": "reflectionTolerance", "reflection_tolerance": "reflectionTolerance",
    "surfaceReservoirSize": "surfaceReservoirSize", "surface_reservoir_size": "surfaceReservoirSize",
    "backend": "backend", "openpmdBackend": "openpmdBackend", "openpmd_backend": "openpmdBackend",
    "parallelMode": "parallelMode", "parallel_mode": "parallelMode",
    "numDevices": "numDevices", "num_devices": "numDevices", "max_gpus": "numDevices",
    "nPerNode": "nPerNode", "n_per_node": "nPerNode", "devices": "devices",
    # shared
    "repetitions": "repetitions", "adaptiveSteps": "adaptiveSteps", "adaptive_steps": "adaptiveSteps",
    "useReflections": "useReflections", "use_reflections": "useReflections",
    "maxReflections": "maxReflections", "monochromatic": "monochromatic",
    "rngSeed": "rngSeed", "rng_seed": "rngSeed", "minSampleRange": "minSampleRange",
    "min_sample_range": "minSampleRange", "maxSampleRange": "maxSampleRange",
    "max_sample_range": "maxSampleRange", "device": "device", "defensiveFraction": "defensiveFraction",
    "raysPerThread": "raysPerThread", "sampleBatch": "sampleBatch",
    "writeVtk": "writeVtk", "write_vtk": "writeVtk", "outputPath": "outputPath", "output_path": "outputPath",
}

_DEFAULT_RAYS = 100000
_DEFAULT_MSE = 0.005


@dataclass
class RunControls:
    """The controls one evaluation runs with, after reconciling the legacy
    per-sample names with the 2.2.0 names of ``PhiASE``."""

    min_rays: int
    max_rays: int
    mse_threshold: float
    adaptive_steps: int
    device: int


@dataclass
class PhiASE:
    config: object | None = None
    crossSections: CrossSectionData | None = None
    spectralProperties: CrossSectionData | None = None
    gainMedium: GainMedium | None = None
    # legacy per-sample controls; None: take the 2.2.0 name below, else the default
    minRaysPerSample: int | None = None
    maxRaysPerSample: int | None = None
    mseThreshold: float | None = None
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
    ase_steps: int | None = None     # time-stepped runs: ASE in steps < ase_steps (None counts as 0)
    writeVtk: bool = False           # run(): per-sample results as VTK files under <outputPath>/vtk
    outputPath: object = "."
    # 2.2.0 keyword surface (pyInclude/simulation.py:130-175).  minRays / maxRays
    # are the ray counts of each sample point here (the reference's migration
    # maps minRaysPerSample onto minRays, simulation.py:318-323);
    # relativeStandardErrorThreshold r stops a sample at mse = RSE^2 <= r^2;
    # forwardRayCount fixes the rays of every sample (no escalation).
    minRays: int | None = None
    maxRays: int | None = None
    relativeStandardErrorThreshold: float | None = None
    forwardRayCount: int | None = None
    propagationMode: str | None = None   # None / "backward": this estimator; "forward" is ForwardContext
    reflectionMaxIterations: int = 40    # surface-reservoir controls of the forward estimator; the
    reflectionTolerance: float = 1e-4    # importance-sampled path reflects by mirror images up to
    surfaceReservoirSize: int = 32       # maxReflections and does not read them
    laserProperties: object | None = None
    backend: str | None = None           # None, or a HIP / AMD GPU backend name
    openpmdBackend: str | None = "auto"  # no openPMD transport: the library is called in-process
    parallelMode: str = "single"
    numDevices: int = 1
    nPerNode: int = 1
    devices: list = field(default_factory=list)

    _result: Result | None = field(default=None, init=False, repr=False)

    def __post_init__(self):
        if isinstance(self.config, (str, Path)):
            self._applyConfig(self._loadConfig(self.config))
        elif isinstance(self.config, dict):
            self._applyConfig(self.config)
        if self.crossSections is None and self.spectralProperties is not None:
            self.crossSections = self.spectralProperties
        if self.crossSections is None and self.laserProperties is not None:
            lp = self.laserProperties.toDict()
            self.crossSections = CrossSectionData(lp["l_abs"], lp["s_abs"], lp["l_ems"], lp["s_ems"],
                                                  resolution=lp["l_res"])
        if self.spectralProperties is None:
            self.spectralProperties = self.crossSections
        self.controls()   # reject contradictory or unsupported settings at construction
        if self.ase_steps is not None:
            if isinstance(self.ase_steps, bool) or not isinstance(self.ase_steps, (int, np.integer)):
                raise TypeError("PhiASE.ase_steps must be an integer or None")
            if self.ase_steps < 0:
                raise ValueError("PhiASE.ase_steps must be non-negative")

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

    def controls(self):
        """Effective ray counts, threshold, schedule and device.

        Each quantity may be given under its legacy name or its 2.2.0 name;
        giving both with different meanings is an error, as are the 2.2.0
        settings this one-GPU, importance-sampled path cannot honour (the
        forward propagation mode, MPI, several devices, a CPU backend)."""
        def pick(legacy_name, legacy, modern_name, modern, default, same=lambda a, b: a == b):
            if legacy is not None and modern is not None and not same(legacy, modern):
                raise ValueError(f"PhiASE: {legacy_name}={legacy} contradicts {modern_name}={modern}")
            return legacy if legacy is not None else modern if modern is not None else default

        mode = None if self.propagationMode is None else str(self.propagationMode).strip().lower()
        if mode not in (None, "backward", "importance"):
            raise ValueError(f"PhiASE runs the importance-sampled (backward) estimator; propagationMode="
                             f"{self.propagationMode!r} is the Tet4 forward method, available as "
                             f"haseongpu_amd.ForwardContext on a TetMesh")
        if str(self.parallelMode).strip().lower() != "single":
            raise ValueError(f"PhiASE schedules every sample point of a run on one GPU; parallelMode="
                             f"{self.parallelMode!r} is not available (split the sample range instead)")
        if int(self.numDevices) != 1 or len(self.devices) > 1:
            raise ValueError("PhiASE runs on one GPU: numDevices must be 1 and devices may name one device")
        if self.backend is not None and not any(k in str(self.backend).lower() for k in ("hip", "amd", "gpu")):
            raise ValueError(f"backend {self.backend!r} is not available: this library runs on the HIP device")
        lo = pick("minRaysPerSample", self.minRaysPerSample, "minRays", self.minRays, _DEFAULT_RAYS)
        hi = pick("maxRaysPerSample", self.maxRaysPerSample, "maxRays", self.maxRays, max(int(lo), _DEFAULT_RAYS))
        steps = int(self.adaptiveSteps)
        if self.forwardRayCount is not None and int(self.forwardRayCount) > 0:
            lo = hi = int(self.forwardRayCount)
        rse = self.relativeStandardErrorThreshold
        mse = pick("mseThreshold", self.mseThreshold, "relativeStandardErrorThreshold^2",
                   None if rse is None else float(rse) ** 2, _DEFAULT_MSE,
                   same=lambda a, b: abs(float(a) - float(b)) <= 1e-12 * max(abs(float(a)), abs(float(b))))
        if rse is not None and float(rse) < 0.0:
            raise ValueError("relativeStandardErrorThreshold must be non-negative")
        device = int(self.devices[0]) if len(self.devices) == 1 else int(self.device)
        if int(lo) <= 0:
            raise ValueError("minRays (minRaysPerSample) must be positive")
        if int(hi) < int(lo):
            raise ValueError("maxRays (maxRaysPerSample) must be >= minRays (minRaysPerSample)")
        if float(mse) < 0.0:
            raise ValueError("mseThreshold must be non-negative")
        return RunControls(min_rays=int(lo), max_rays=int(hi), mse_threshold=float(mse), adaptive_steps=steps,
                           device=device)

    def _validate(self, medium, spectra):
        if not isinstance(medium, GainMedium):
            raise TypeError("run() needs a GainMedium")
        if spectra is None:
            raise ValueError("PhiASE needs crossSections")
        self.controls()
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
        ctl = self.controls()
        mesh = _native.NativeMesh(topo, medium)
        exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=ctl.min_rays,
                                       max_rays=ctl.max_rays, max_sigma_a=msa, max_sigma_e=mse_,
                                       mse_threshold=ctl.mse_threshold,
                                       use_reflections=bool(self.useReflections),
                                       max_reflections=int(self.maxReflections),
                                       defensive_fraction=float(self.defensiveFraction))
        seed = 5489 if self.rngSeed is None else int(self.rngSeed)
        comp = _native.make_compute(n_samples=S, max_repetitions=int(self.repetitions),
                                    adaptive_steps=ctl.adaptive_steps,
                                    min_sample_range=0 if self.minSampleRange is None else int(self.minSampleRange),
                                    max_sample_range=None if self.maxSampleRange is None else int(self.maxSampleRange),
                                    rng_seed=seed, device=ctl.device, rays_per_thread=int(self.raysPerThread),
                                    sample_batch=int(self.sampleBatch))
        return mesh, exp, comp

    def run(self, gainMedium=None, crossSections=None):
        mesh, exp, comp = self.nativeArguments(gainMedium, crossSections)
        with _native.Context(mesh, exp, comp) as ctx:
            if self.proposal is not None:
                ctx.set_proposal(self.proposal)
            t0 = time.perf_counter()
            res = ctx.evaluate()
            runtime = time.perf_counter() - t0
            stats = ctx.stats()
            if self.proposal is not None:
                stats["proposal"] = ctx.proposal_stats()
        rse = np.sqrt(np.where(np.isfinite(res.mse) & (res.mse >= 0), res.mse, np.nan))
        self._result = Result(phiAse=res.phiAse, mse=res.mse, totalRays=res.totalRays, dndtAse=res.dndtAse,
                              standardError=res.standardError, relativeStandardError=rse,
                              droppedRays=res.droppedRays, stats=stats)
        if self.writeVtk:
            self._writeVtk(gainMedium if gainMedium is not None else self.gainMedium, int(stats["rays"]), runtime)
        return self._result

    def _writeVtk(self, medium, rays, runtime):
        """Per-sample results as POINT_DATA files <outputPath>/vtk/<name>_<unix time>.vtk
        (include/core/simulation.hpp:161-224)."""
        from .vtk_wedge import writePointsToVtk

        r = self._result
        ctl = self.controls()
        folder = Path(self.outputPath) / "vtk"
        stamp = str(int(time.time()))
        for name, values in (("dndt", r.dndtAse), ("phiase", r.phiAse), ("standard_error", r.standardError),
                             ("total_rays", r.totalRays), ("relative_standard_error", r.relativeStandardError)):
            writePointsToVtk(medium, values, folder / f"{name}_{stamp}.vtk", rays, ctl.max_rays,
                             ctl.mse_threshold, bool(self.useReflections), runtime)

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


# ---------------------------------------------------------------------------
# Time-stepped simulation (pyInclude/simulation.py:494-1004)

SIMULATION_OUTPUT_FIELDS = ("beta_volume", "phi_ase", "standard_error", "relative_standard_error", "total_rays",
                            "dndt_ase", "dndt_pump")
SIMULATION_CONTROL_FIELDS = ("beta_volume",)


@dataclass
class TimeStepState:
    """Snapshot after a completed step (pyInclude/simulation.py:494-542).

    Prism arrays have the shape of ``GainMedium`` betaVolume, (T, L-1);
    ``standardError``, ``relativeStandardError`` and ``totalRays`` belong to
    the sample points of the last ASE evaluation, shape (Np, L); fields that
    were not selected are None."""

    step: int
    time: float
    betaVolume: np.ndarray | None
    phiAse: np.ndarray | None
    dndtAse: np.ndarray | None
    dndtPump: np.ndarray | None
    aseResult: object | None
    standardError: np.ndarray | None = None
    relativeStandardError: np.ndarray | None = None
    totalRays: np.ndarray | None = None
    topology: object | None = None

    @property
    def beta_volume(self):
        return self.betaVolume

    @property
    def phi_ase(self):
        return self.phiAse

    @property
    def dndt_ase(self):
        return self.dndtAse

    @property
    def dndt_pump(self):
        return self.dndtPump

    @property
    def ase_result(self):
        return self.aseResult


class Simulation:
    """Time-stepped beta simulation on the library's device loop.

    Mirrors the reference's ``Simulation`` (pyInclude/simulation.py:545-1002):
    the same constructor keywords, ``add_pump``, ``on_init`` / ``on_step`` /
    ``before_step`` callbacks, ``step`` / ``run_until``, output steps and
    fields, synchronized-debug control of ``beta_volume``.  The loop runs in
    the HIP library (hase_simulation_*) with beta resident on the device;
    Python is contacted only at output steps and at the steps where the set of
    active pumps changes.  Pumps are ``time_integration.LocalPump`` objects
    (linear in beta per prism).  The device loop persists across ``step``
    calls; ``close()`` releases it.
    """

    def __init__(self, *, gain_medium, phi_ase, time_integrator, time_step_size, cross_sections=None,
                 simulation_steps=None, max_time=None, pre_pump=False, report_timings=False,
                 execution_mode="autonomous", output_steps=None, output_fields=None, control_fields=()):
        self.gainMedium = gain_medium
        self.phiASE = phi_ase
        self.timeIntegrationSolver = time_integrator
        self.timeStep = float(time_step_size)
        self.crossSections = cross_sections
        self.endTime = max_time
        self.simulationSteps = None if simulation_steps is None else int(simulation_steps)
        self.prePump = bool(pre_pump)
        self.reportTimings = bool(report_timings)
        self.executionMode = str(execution_mode)
        self.outputSteps = None if output_steps is None else tuple(int(s) for s in output_steps)
        self.outputFields = tuple(SIMULATION_OUTPUT_FIELDS if output_fields is None else
                                  (str(f) for f in output_fields))
        self.controlFields = tuple(str(f) for f in control_fields)
        self._pumps = []
        self._time = 0.0
        self._step = 0
        self._initialized = False
        self._init_callbacks = []
        self._before_step_callbacks = []
        self._step_callbacks = []
        self._lastState = None
        self._ctx = None
        self._loop = None
        self._pump_key = None
        self._validate()

    # -- configuration (validation of pyInclude/simulation.py:630-673) --------
    def _validate(self):
        solver = self.timeIntegrationSolver
        if solver is None:
            raise ValueError("Simulation requires a time_integrator")
        name = solver if isinstance(solver, str) else getattr(solver, "name", None)
        if name not in _native.INTEGRATORS:
            raise ValueError(f"unsupported time integrator '{name}'")
        if not self.timeStep > 0.0:
            raise ValueError("time_step_size must be positive")
        if self.simulationSteps is not None and self.simulationSteps <= 0:
            raise ValueError("simulation_steps must be positive")
        if self.simulationSteps is not None and self.endTime is not None:
            raise ValueError("configure either simulation_steps or max_time, not both")
        if self.executionMode not in ("autonomous", "synchronized-debug"):
            raise ValueError("execution_mode must be 'autonomous' or 'synchronized-debug'")
        if self.outputSteps is not None:
            if not self.outputSteps:
                raise ValueError("output_steps must be omitted or contain at least one completed-step index")
            if any(s <= 0 for s in self.outputSteps):
                raise ValueError("output_steps must contain positive completed-step indices")
            if tuple(sorted(set(self.outputSteps))) != self.outputSteps:
                raise ValueError("output_steps must be strictly increasing and unique")
        if not self.outputFields:
            raise ValueError("output_fields must contain at least one field")
        unknown = sorted(set(self.outputFields) - set(SIMULATION_OUTPUT_FIELDS))
        if unknown:
            raise ValueError(f"unsupported output_fields: {unknown}")
        if len(set(self.outputFields)) != len(self.outputFields):
            raise ValueError("output_fields must be unique")
        if self.executionMode == "autonomous" and self.controlFields:
            raise ValueError("control_fields require execution_mode='synchronized-debug'")
        if self.executionMode == "synchronized-debug" and self.outputSteps is not None:
            raise ValueError("synchronized-debug emits every completed step; output_steps must be omitted")
        unknown = sorted(set(self.controlFields) - set(SIMULATION_CONTROL_FIELDS))
        if unknown:
            raise ValueError(f"unsupported control_fields: {unknown}")
        if len(set(self.controlFields)) != len(self.controlFields):
            raise ValueError("control_fields must be unique")
        if self.crossSections is None:
            self.crossSections = self.phiASE.crossSections
        if self.phiASE.crossSections is None:
            self.phiASE.crossSections = self.crossSections

    @property
    def integratorName(self):
        solver = self.timeIntegrationSolver
        return solver if isinstance(solver, str) else solver.name

    def add_pump(self, pump, injection_method=None):
        """Register a pump (a ``LocalPump``); returns ``self``."""
        from .time_integration import LocalPump

        if self._initialized:
            raise RuntimeError("pumps must be added before the simulation is initialized")
        if not isinstance(pump, LocalPump):
            raise TypeError("pump must be a LocalPump")
        self._pumps.append(pump)
        return self

    def on_step(self, callback, *args, **kwargs):
        self._step_callbacks.append((callback, args, kwargs))
        return self

    def on_init(self, callback, *args, **kwargs):
        self._init_callbacks.append((callback, args, kwargs))
        return self

    def before_step(self, callback, *args, **kwargs):
        self._before_step_callbacks.append((callback, args, kwargs))
        return self

    # -- running ---------------------------------------------------------------
    def _setup(self):
        if self._loop is not None:
            return
        if self.crossSections is None:
            raise ValueError("Simulation requires cross sections (cross_sections= or PhiASE.crossSections)")
        mesh, exp, comp = self.phiASE.nativeArguments(self.gainMedium, self.crossSections)
        if comp.min_sample_range != 0 or comp.max_sample_range + 1 != self.gainMedium.topology.numberOfSamples:
            raise ValueError("a time-stepped simulation evaluates every sample point (no sample range)")
        self._ctx = _native.Context(mesh, exp, comp)
        if self.phiASE.proposal is not None:
            self._ctx.set_proposal(self.phiASE.proposal)
        iters = getattr(self.timeIntegrationSolver, "iterations", 8)
        ase_steps = 0 if self.phiASE.ase_steps is None else int(self.phiASE.ase_steps)
        P = self._ctx.prisms
        self._loop = _native.NativeSimulation(
            self._ctx, time_step=self.timeStep, integrator=self.integratorName, implicit_iterations=iters,
            first_simulation_step=self._step, ase_steps=ase_steps, pump_steps=0xFFFFFFFF,
            pre_pump=self.prePump, beta_volume=self.gainMedium.value("betaVolume"), pump_rate=np.zeros(P))

    def _apply_pumps(self, step):
        """Sum of the pumps active in simulation step ``step``."""
        active = tuple(i for i, p in enumerate(self._pumps) if p.active(step))
        if active == self._pump_key:
            return
        P = self._ctx.prisms
        rate, sat = np.zeros(P), None
        for i in active:
            r, s = self._pumps[i].arrays(P)
            rate = rate + r
            if s is not None:
                sat = s.copy() if sat is None else sat + s
        self._loop.set_pump(rate, sat)
        self._pump_key = active

    def _segment_end(self, step, last):
        """Last step (exclusive) before the active pump set may change."""
        end = last
        for p in self._pumps:
            n = 0 if p.pump_steps is None else int(p.pump_steps)
            if step < n < end:
                end = n
        return end

    def _emit(self, completed_in_run, steps_in_run):
        st = self._loop.state()
        T, L = self.gainMedium.topology.numberOfTriangles, int(self.gainMedium.topology.levels)
        Np = self.gainMedium.topology.numberOfPoints
        want = set(self.outputFields)
        prism = lambda a: a.reshape(T, L - 1, order="F").copy()      # noqa: E731
        sample = lambda a: a.reshape(Np, L, order="F").copy()        # noqa: E731
        rse = np.sqrt(np.where(st["mse"] >= 0, st["mse"], np.nan))
        ase_result = Result(phiAse=st["phi_samples"], mse=st["mse"], totalRays=st["total_rays"],
                            standardError=st["standard_error"], relativeStandardError=rse)
        state = TimeStepState(
            step=self._step, time=self._time,
            betaVolume=prism(st["beta_volume"]) if "beta_volume" in want else None,
            phiAse=prism(st["phi_ase"]) if "phi_ase" in want else None,
            dndtAse=prism(st["dndt_ase"]) if "dndt_ase" in want else None,
            dndtPump=prism(st["dndt_pump"]) if "dndt_pump" in want else None,
            aseResult=ase_result,
            standardError=sample(st["standard_error"]) if "standard_error" in want else None,
            relativeStandardError=sample(rse) if "relative_standard_error" in want else None,
            totalRays=sample(st["total_rays"]) if "total_rays" in want else None,
            topology=self.gainMedium.topology)
        if state.betaVolume is not None:
            self.gainMedium.set("betaVolume", state.betaVolume)
        self._lastState = state
        for cb, args, kwargs in self._step_callbacks:
            cb(state, *args, **kwargs)
        if self.executionMode == "synchronized-debug" and completed_in_run < steps_in_run:
            for cb, args, kwargs in self._before_step_callbacks:
                cb(self, *args, **kwargs)
            if "beta_volume" in self.controlFields:
                self._loop.set_beta(self.gainMedium.value("betaVolume"))

    def _should_output(self, step):
        if self.executionMode == "synchronized-debug" or self.outputSteps is None:
            return True
        return step in self.outputSteps

    def runSteps(self, steps):
        """Run exactly ``steps`` time steps on the device and return ``self``."""
        steps = int(steps)
        if steps <= 0:
            raise ValueError("steps must be positive")
        if self._before_step_callbacks and self.executionMode != "synchronized-debug":
            raise ValueError("before_step callbacks require execution_mode='synchronized-debug'")
        if not self._initialized:
            self._initialized = True
            for cb, args, kwargs in self._init_callbacks:
                cb(self, *args, **kwargs)
        self._setup()
        done = 0
        while done < steps:
            self._apply_pumps(self._step)
            # run up to the next output step or pump change in one library call
            end = self._step + (steps - done)
            end = self._segment_end(self._step, end)
            nxt = self._step + 1
            while nxt < end and not self._should_output(nxt):
                nxt += 1
            n = min(nxt, end) - self._step
            self._loop.step(n)
            done += n
            self._step += n
            self._time += n * self.timeStep
            if self._should_output(self._step):
                self._emit(done, steps)
        return self

    def _derived_simulation_steps(self):
        activity = [0 if self.phiASE.ase_steps is None else int(self.phiASE.ase_steps)]
        activity += [0 if p.pump_steps is None else int(p.pump_steps) for p in self._pumps]
        derived = max(activity, default=0)
        if derived <= 0:
            raise ValueError("simulation_steps is required when all pump_steps and ase_steps are disabled")
        return derived

    def step(self, nsteps=None):
        """Advance ``nsteps`` time steps (PICMI call pattern)."""
        if nsteps is None:
            if self.simulationSteps is not None:
                nsteps = self.simulationSteps
            elif self.endTime is not None:
                return self.run_until()
            else:
                nsteps = self._derived_simulation_steps()
        if int(nsteps) <= 0:
            raise ValueError("nsteps must be positive")
        return self.runSteps(int(nsteps))

    def run_until(self, max_time=None):
        target = self.endTime if max_time is None else max_time
        if target is None:
            raise ValueError("run_until requires max_time or a configured max_time")
        steps = 0
        while self._time + steps * self.timeStep < float(target) - 0.5 * self.timeStep:
            steps += 1
        if steps:
            self.step(steps)
        return self

    def times(self):
        """Milliseconds in ASE evaluations and in the integrator kernels, rays traced."""
        return {"ase_ms": 0.0, "integrator_ms": 0.0, "rays": 0} if self._loop is None else self._loop.times()

    def close(self):
        if self._loop is not None:
            self._loop.close()
            self._loop = None
        if self._ctx is not None:
            self._ctx.close()
            self._ctx = None

    def get_last_state(self):
        return self.getLastState()

    def getLastState(self):
        if self._lastState is None:
            raise RuntimeError("simulation has not completed a time step yet")
        return self._lastState

    getResults = getLastState

    @property
    def lastState(self):
        return self.getLastState()

    @property
    def current_step(self):
        return self._step

    @property
    def current_time(self):
        return self._time

    time = current_time
    stepIndex = current_step

    @property
    def gain_medium(self):
        return self.gainMedium

    @property
    def phi_ase(self):
        return self.phiASE

    @property
    def time_integrator(self):
        return self.timeIntegrationSolver

    @property
    def time_step_size(self):
        return self.timeStep

    @property
    def simulation_steps(self):
        return self.simulationSteps

    @property
    def pumps(self):
        return tuple(self._pumps)


TimeSteppedSimulation = Simulation

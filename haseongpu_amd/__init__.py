"""MI355X-native importance-sampled ASE ray tracer for extruded prism meshes.

Public names follow the reference's front end (HASEonGPU.py,
pyInclude/__init__.py:1-75): ``PhiASE``, ``GainMedium``, ``MeshTopology``,
``Grid``, ``CrossSectionData`` / ``SpectralDecomposition`` and ``Result``,
plus the time-stepped ``Simulation`` with the integrator descriptors of
pyInclude/timeIntegration.py and the legacy VTK writers (``vtkWedge``,
``writePointsToVtk``, ``writePrismToVtk``, ``compareVtk``).
"""
from .geometry import GainMedium, Grid, MeshTopology
from .laser import CrossSectionData, SpectralDecomposition
from .simulation import PhiASE, Result, Simulation, TimeSteppedSimulation, TimeStepState
from .time_integration import (ExplicitEuler, ExponentialEuler, FrozenPhiAseRungeKutta4, Heun, ImplicitEuler,
                               LocalPump, Midpoint, RungeKutta4, TimeIntegrationSolver)
from .vtk_wedge import compareVtk, vtkWedge, writePointsToVtk, writePrismToVtk

__all__ = ["PhiASE", "Result", "GainMedium", "MeshTopology", "Grid", "CrossSectionData",
           "SpectralDecomposition", "Simulation", "TimeSteppedSimulation", "TimeStepState",
           "TimeIntegrationSolver", "ExplicitEuler", "Heun", "Midpoint", "RungeKutta4", "FrozenPhiAseRungeKutta4",
           "ImplicitEuler", "ExponentialEuler", "LocalPump", "vtkWedge", "writePointsToVtk", "writePrismToVtk",
           "compareVtk"]
__version__ = "0.1.0"

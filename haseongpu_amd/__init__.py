"""MI355X-native importance-sampled ASE ray tracer for extruded prism meshes.

Public names follow the reference's front end (HASEonGPU.py,
pyInclude/__init__.py:1-75): ``PhiASE``, ``GainMedium``, ``MeshTopology``,
``Grid``, ``CrossSectionData`` / ``SpectralDecomposition`` and ``Result``.
"""
from .geometry import GainMedium, Grid, MeshTopology
from .laser import CrossSectionData, SpectralDecomposition
from .simulation import PhiASE, Result

__all__ = ["PhiASE", "Result", "GainMedium", "MeshTopology", "Grid", "CrossSectionData",
           "SpectralDecomposition"]
__version__ = "0.1.0"

"""Absorption / emission spectra (``CrossSectionData``).

Mirrors pyInclude/laser.py:183-323 (``CrossSectionData``, ``monochromatic``)
and the maximum cross sections of pyInclude/laser.py:786-797.  Spectra are
resampled to equidistant wavelengths as src/utils/interpolation.cpp:66-118
does (linear pieces, the last sample pinned to the table's last value).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

__all__ = ["CrossSectionData", "SpectralDecomposition", "interpolateLinear"]


def interpolateLinear(y, x, nInterpolations):
    """Resample y(x) at ``nInterpolations`` equidistant points spanning x."""
    y = np.asarray(y, dtype=np.float64).reshape(-1)
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    n = int(nInterpolations)
    if x.size == 0 or x.size != y.size:
        raise ValueError("x and y must be non-empty and of equal length")
    if n < y.size:
        raise ValueError("nInterpolations must be at least the table length")
    if y.size == 1:
        return np.full(n, y[0])
    if np.any(np.diff(x) <= 0.0):
        raise ValueError("x must be strictly increasing")
    xi = x[0] + np.arange(n, dtype=np.float64) * (x[-1] - x[0]) / (n - 1)
    out = np.interp(xi, x, y)
    out[-1] = y[-1]
    return out


@dataclass
class CrossSectionData:
    wavelengthsAbsorption: object
    crossSectionAbsorption: object
    wavelengthsEmission: object
    crossSectionEmission: object
    resolution: int = 1

    def __post_init__(self):
        self.wavelengthsAbsorption = np.asarray(self.wavelengthsAbsorption, dtype=np.float64).reshape(-1)
        self.crossSectionAbsorption = np.asarray(self.crossSectionAbsorption, dtype=np.float64).reshape(-1)
        self.wavelengthsEmission = np.asarray(self.wavelengthsEmission, dtype=np.float64).reshape(-1)
        self.crossSectionEmission = np.asarray(self.crossSectionEmission, dtype=np.float64).reshape(-1)
        self.resolution = int(self.resolution)
        if self.wavelengthsAbsorption.size != self.crossSectionAbsorption.size:
            raise ValueError("wavelengthsAbsorption and crossSectionAbsorption must have the same length")
        if self.wavelengthsEmission.size != self.crossSectionEmission.size:
            raise ValueError("wavelengthsEmission and crossSectionEmission must have the same length")
        if self.crossSectionAbsorption.size == 0 or self.crossSectionEmission.size == 0:
            raise ValueError("spectra must not be empty")
        if self.resolution < 1:
            raise ValueError("resolution must be positive")
        if np.any(self.crossSectionAbsorption < 0) or np.any(self.crossSectionEmission < 0):
            raise ValueError("cross sections must be non-negative")

    @classmethod
    def monochromatic(cls, *, wavelength, crossSectionAbsorption, crossSectionEmission):
        return cls([wavelength], [crossSectionAbsorption], [wavelength], [crossSectionEmission], resolution=1)

    @property
    def maxSigmaA(self):
        return float(np.max(self.crossSectionAbsorption))

    @property
    def maxSigmaE(self):
        return float(np.max(self.crossSectionEmission))

    def spectralPairs(self, monochromatic=False):
        """(sigma_a[W], sigma_e[W], max_sigma_a, max_sigma_e) handed to the
        estimator.  Monochromatic runs use the first pair only
        (tests/data/compareSerial/README.md: the legacy parser reduced each
        spectrum to its first pair); otherwise both spectra are resampled to
        max(resolution, table length) equidistant samples."""
        if monochromatic:
            sa = self.crossSectionAbsorption[:1].copy()
            se = self.crossSectionEmission[:1].copy()
            return sa, se, float(sa[0]), float(se[0])
        same_grid = (self.wavelengthsAbsorption.size == self.wavelengthsEmission.size
                     and np.array_equal(self.wavelengthsAbsorption, self.wavelengthsEmission))
        n = max(self.resolution, self.crossSectionAbsorption.size, self.crossSectionEmission.size)
        if same_grid and n == self.crossSectionAbsorption.size:
            sa, se = self.crossSectionAbsorption.copy(), self.crossSectionEmission.copy()
        else:
            sa = interpolateLinear(self.crossSectionAbsorption, self.wavelengthsAbsorption, n)
            se = interpolateLinear(self.crossSectionEmission, self.wavelengthsEmission, n)
        return sa, se, self.maxSigmaA, self.maxSigmaE


SpectralDecomposition = CrossSectionData

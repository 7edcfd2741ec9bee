"""Seeded synthetic workloads: the small plain slab and the production disk.

* ``plain_small`` - square slab 1 x 1 cm, 10 x 10 grid split into 200
  triangles / 121 points, 10 levels over 0.7 cm (1800 prisms), per-prism
  beta ~ U(0, 0.3) from numpy's default_rng(seed), one wavelength
  (sigma_a 1.16e-21, sigma_e 2.48e-20 cm^2), nTot 2.76e20 cm^-3,
  tau 9.5e-4 s, no reflections, no cladding, fixed rays per sample.
* ``batch_problems`` - independent seeded Delaunay disks for batched runs:
  triangles log-uniform in a range (200-8000 by default), 5-40 levels, beta
  from 1-4 Gaussian pump spots (widths 0.3-1.5 cm), 1-16 Gaussian-bump
  wavelengths, adaptive rays 1e4 -> 1e7 with mse threshold 0.005.
* ``production_disk`` - disk of radius 3 cm, ~2k seeded points (jittered
  hexagonal lattice plus a boundary ring) Delaunay-triangulated into ~4k
  triangles, 20 levels over 0.7 cm, beta = 0.3 exp(-r^2 / (2 * 1.2^2)) +
  N(0, 0.01) clipped to [0, 1] at prism centroids, 64 wavelengths of
  Gaussian-bump spectra (peaks ~1e-21 / 2.5e-20 cm^2), per-triangle top and
  bottom reflectivity ~U(0, 0.1), the outer ring of triangles (those with an
  exterior edge) as cladding with absorption 5.5 cm^-1 and beta 0, adaptive
  rays 1e5 -> 1e7 per sample with mse threshold 0.005.
* ``combined_disk`` - the same construction at scale: radius 5 cm, ~10.5k
  points, 20,682 triangles, 30 levels over 1.0 cm (599,778 prisms, 315,300
  sample points), 128 wavelengths, reflections, cladding ring, adaptive
  rays 1e5 -> 1e8; ``depletion_field`` gives its seeded smooth
  pump-depletion factors, one beta update per time step.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .geometry import GainMedium, Grid, MeshTopology
from .laser import CrossSectionData


@dataclass
class Workload:
    name: str
    topology: MeshTopology
    medium: GainMedium
    spectra: CrossSectionData
    settings: dict = field(default_factory=dict)

    @property
    def samples(self):
        return self.topology.numberOfSamples


def plain_small(seed=5489, rays=100_000):
    grid = Grid(x=1.0, y=1.0, z=0.7, tileSizeX=0.1, tileSizeZ=0.7 / 9)
    topo = MeshTopology.fromGrid(grid)
    rng = np.random.default_rng(seed)
    beta = rng.uniform(0.0, 0.3, size=topo.numberOfPrisms)
    medium = GainMedium(topo).withPhysicalProperties(
        betaVolume=beta,
        betaCells=np.full(topo.numberOfSamples, 0.15),
        claddingCellTypes=np.zeros(topo.numberOfTriangles, np.uint32),
        claddingNumber=1,
        claddingAbsorption=0.0,
        nTot=2.76e20,
        crystalTFluo=9.5e-4,
        refractiveIndices=[1.83, 1.0, 1.83, 1.0],
    )
    spectra = CrossSectionData.monochromatic(wavelength=1030e-9, crossSectionAbsorption=1.16e-21,
                                             crossSectionEmission=2.48e-20)
    settings = dict(minRaysPerSample=rays, maxRaysPerSample=rays, mseThreshold=0.005, repetitions=1,
                    adaptiveSteps=1, useReflections=False, monochromatic=True, rngSeed=seed)
    return Workload("plain_small", topo, medium, spectra, settings)


def disk_points(radius=3.0, n_target=2060, seed=2, ring=None):
    """Jittered hexagonal lattice inside a disk plus a uniform boundary ring."""
    rng = np.random.default_rng(seed)
    area = np.pi * radius * radius
    a = np.sqrt(2.0 * area / (np.sqrt(3.0) * n_target))
    ring = int(round(2.0 * np.pi * radius / a)) if ring is None else int(ring)
    nx = int(np.ceil(2 * radius / a)) + 2
    ny = int(np.ceil(2 * radius / (a * np.sqrt(3) / 2))) + 2
    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    x = (ii - nx / 2) * a + (jj % 2) * a / 2
    y = (jj - ny / 2) * a * np.sqrt(3) / 2
    pts = np.column_stack([x.ravel(), y.ravel()])
    pts = pts + rng.uniform(-0.2 * a, 0.2 * a, size=pts.shape)
    pts = pts[np.hypot(pts[:, 0], pts[:, 1]) < radius - 0.6 * a]
    phi = (np.arange(ring) + rng.uniform(0, 1)) * 2 * np.pi / ring
    boundary = np.column_stack([radius * np.cos(phi), radius * np.sin(phi)])
    return np.vstack([pts, boundary])


def gaussian_spectra(n=64):
    lam = np.linspace(900.0, 1100.0, n)
    sa = 1.0e-21 * np.exp(-0.5 * ((lam - 975.0) / 18.0) ** 2) + 4.0e-23
    se = 2.5e-20 * np.exp(-0.5 * ((lam - 1030.0) / 14.0) ** 2) + 1.0e-22
    return CrossSectionData(lam, sa, lam, se, resolution=n)


def production_disk(seed=2, levels=20, n_points=2060, wavelengths=64, min_rays=100_000, max_rays=10_000_000,
                    mse_threshold=0.005, radius=3.0, thickness_total=0.7, max_reflections=1):
    pts = disk_points(radius, n_points, seed)
    topo = MeshTopology.fromPoints(pts, numberOfLevels=levels, thickness=thickness_total / (levels - 1))
    rng = np.random.default_rng(seed + 1000)
    cent = topo.prismCentroids()
    r = np.hypot(cent[:, 0], cent[:, 1])
    beta = 0.3 * np.exp(-r * r / (2 * 1.2 ** 2)) + rng.normal(0.0, 0.01, size=r.size)
    beta = np.clip(beta, 0.0, 1.0)
    clad_tri = topo.boundaryTriangles()
    beta[np.tile(clad_tri, levels - 1)] = 0.0
    pr = np.hypot(topo.points[:, 0], topo.points[:, 1])
    beta_cells = np.tile(0.3 * np.exp(-pr * pr / (2 * 1.2 ** 2)), levels)
    refl = rng.uniform(0.0, 0.1, size=2 * topo.numberOfTriangles)
    medium = GainMedium(topo).withPhysicalProperties(
        betaVolume=beta,
        betaCells=beta_cells,
        claddingCellTypes=clad_tri.astype(np.uint32),
        claddingNumber=1,
        claddingAbsorption=5.5,
        nTot=2.76e20,
        crystalTFluo=9.5e-4,
        reflectivities=refl,
        refractiveIndices=[1.0, 1.0, 1.0, 1.0],
    )
    settings = dict(minRaysPerSample=min_rays, maxRaysPerSample=max_rays, mseThreshold=mse_threshold,
                    repetitions=1, adaptiveSteps=4, useReflections=True, maxReflections=max_reflections,
                    monochromatic=False, rngSeed=seed)
    return Workload("production_disk", topo, medium, gaussian_spectra(wavelengths), settings)


def combined_disk(seed=5, levels=30, n_points=10_400, wavelengths=128, min_rays=100_000, max_rays=100_000_000,
                  mse_threshold=0.005, radius=5.0, thickness_total=1.0, max_reflections=1):
    """Combined case at scale (production-disk construction, radius 5 cm,
    ~20.7k triangles, 30 levels over 1 cm, 128 wavelengths, 1e5 -> 1e8 rays)."""
    w = production_disk(seed=seed, levels=levels, n_points=n_points, wavelengths=wavelengths, min_rays=min_rays,
                        max_rays=max_rays, mse_threshold=mse_threshold, radius=radius,
                        thickness_total=thickness_total, max_reflections=max_reflections)
    w.name = "combined_disk"
    return w


def depletion_field(topology, *, seed=5, spots=3, rate=(0.004, 0.012), widths=(1.0, 3.0), absorption=0.8):
    """Seeded smooth pump-depletion factors per prism, applied once per time
    step (beta_{k+1} = beta_k * factor): 1 - sum of Gaussian spots, each with
    a depletion rate per step drawn from `rate`, a width from `widths` (cm)
    and a centre within 60% of the disk radius, decaying along z with
    coefficient `absorption` (cm^-1)."""
    rng = np.random.default_rng(seed)
    cent = topology.prismCentroids()
    r_max = float(np.hypot(topology.points[:, 0], topology.points[:, 1]).max())
    loss = np.zeros(cent.shape[0])
    for _ in range(int(spots)):
        rr = 0.6 * r_max * np.sqrt(rng.uniform())
        ang = rng.uniform(0.0, 2.0 * np.pi)
        w = rng.uniform(*widths)
        d2 = (cent[:, 0] - rr * np.cos(ang)) ** 2 + (cent[:, 1] - rr * np.sin(ang)) ** 2
        loss += rng.uniform(*rate) * np.exp(-d2 / (2.0 * w * w))
    return np.clip(1.0 - loss * np.exp(-absorption * cent[:, 2]), 0.0, 1.0)


def batch_problems(n=64, seed=0, triangles=(200, 8000), levels=(5, 40), wavelengths=(1, 16), min_rays=10_000,
                   max_rays=10_000_000, mse_threshold=0.005, adaptive_steps=4):
    """`n` independent seeded problems of the batching workload."""
    from .proposal import random_pump_beta

    out = []
    for i in range(n):
        rng = np.random.default_rng([seed, i])
        T = float(np.exp(rng.uniform(np.log(triangles[0]), np.log(triangles[1]))))
        L = int(rng.integers(levels[0], levels[1] + 1))
        W = int(rng.integers(wavelengths[0], wavelengths[1] + 1))
        radius = float(rng.uniform(1.5, 3.0))
        pts = disk_points(radius, max(16, int(round(T / 2 + 1))), seed=int(rng.integers(1 << 30)))
        topo = MeshTopology.fromPoints(pts, numberOfLevels=L, thickness=float(rng.uniform(0.5, 1.0)) / (L - 1))
        beta = random_pump_beta(topo, rng, widths=(0.3, 1.5))
        pr = np.hypot(topo.points[:, 0], topo.points[:, 1])
        medium = GainMedium(topo).withPhysicalProperties(
            betaVolume=beta,
            betaCells=np.tile(np.interp(pr, [0, radius], [float(beta.max()), 0.0]), L),
            claddingCellTypes=np.zeros(topo.numberOfTriangles, np.uint32),
            claddingNumber=1,
            claddingAbsorption=0.0,
            nTot=2.76e20,
            crystalTFluo=9.5e-4,
            refractiveIndices=[1.83, 1.0, 1.83, 1.0],
        )
        settings = dict(minRaysPerSample=min_rays, maxRaysPerSample=max_rays, mseThreshold=mse_threshold,
                        repetitions=1, adaptiveSteps=adaptive_steps, useReflections=False, monochromatic=False,
                        rngSeed=seed * 1000 + i)
        out.append(Workload(f"batch_problem_{i}", topo, medium, gaussian_spectra(W), settings))
    return out


def local_pump(topology, *, peak_rate=1500.0, width=1.2, absorption=1.5, ratio=1.21, pump_steps=None,
               center=(0.0, 0.0)):
    """Deterministic smooth pump for the time-stepped loop: a Gaussian spot of
    1/e^2-type width ``width`` (cm) entering at z = 0 and absorbed along z
    with coefficient ``absorption`` (cm^-1); rate = peak_rate * profile and
    saturation = ratio * rate (ratio = 1 + sigma_e,p / sigma_a,p; 1.21 for
    Yb:YAG at 940 nm)."""
    from .time_integration import LocalPump

    cent = topology.prismCentroids()
    r2 = (cent[:, 0] - center[0]) ** 2 + (cent[:, 1] - center[1]) ** 2
    rate = peak_rate * np.exp(-r2 / (2.0 * width * width)) * np.exp(-absorption * cent[:, 2])
    return LocalPump(rate=rate, saturation=ratio * rate, pump_steps=pump_steps)

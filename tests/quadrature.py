"""Deterministic angular quadrature of the point flux in a box-shaped slab.

For uniform beta and a uniform gain coefficient g in the box
[x0,x1] x [y0,y1] x [0,Z], the point flux at p is (nTot/tau) times
    (beta / 4 pi) * integral dOmega  integral_{r in medium} exp(g r) dr,
the track-length form of integral beta G / (4 pi r^2) dV (the same identity
gives the sphere closed form of tests/python/phiAse/test_analyticalSphere.py:19-27).
First-order reflections at z = 0 and z = Z are added by the image method: a
ray that reaches a face while still inside the lateral walls continues
through the mirrored box, weighted by the surface factor of that face
(reflectivity, or 1 under total internal reflection for
(n_in/n_out)^2 sin^2(theta) > 1, include/kernels/reflection.hpp:33-45).

``hat_smoothed_cell_flux`` turns such a point flux into the expectation of
the forward estimator's cell value (vertex hat-function averages).
"""
from __future__ import annotations

import numpy as np


def _lateral_exit(px, py, ux, uy, box):
    x0, x1, y0, y1 = box
    with np.errstate(divide="ignore", invalid="ignore"):
        tx = np.where(ux > 0, (x1 - px) / ux, np.where(ux < 0, (x0 - px) / ux, np.inf))
        ty = np.where(uy > 0, (y1 - py) / uy, np.where(uy < 0, (y0 - py) / uy, np.inf))
    return np.minimum(tx, ty)


def _track(g, a, b):
    """integral_a^b exp(g r) dr (0 where b <= a)."""
    b = np.maximum(a, b)
    if abs(g) < 1e-14:
        return b - a
    return (np.exp(g * b) - np.exp(g * a)) / g


def box_point_flux(p, box, Z, beta, g, scale, *, reflect=False, refl_bottom=0.0, refl_top=0.0,
                   n_in=1.0, n_out=1.0, n_mu=1200, n_phi=2400):
    """Point flux at p (already multiplied by ``scale`` = nTot/tau)."""
    px, py, pz = p
    mu = -1.0 + (np.arange(n_mu) + 0.5) * (2.0 / n_mu)
    ph = (np.arange(n_phi) + 0.5) * (2.0 * np.pi / n_phi)
    MU, PH = np.meshgrid(mu, ph, indexing="ij")
    sin_t = np.sqrt(1.0 - MU * MU)
    ux, uy, uz = sin_t * np.cos(PH), sin_t * np.sin(PH), MU
    lat = _lateral_exit(px, py, ux, uy, box)
    with np.errstate(divide="ignore", invalid="ignore"):
        d_top = np.where(uz > 0, (Z - pz) / uz, np.inf)
        d_bot = np.where(uz < 0, -pz / uz, np.inf)
    d_face = np.minimum(d_top, d_bot)
    direct = _track(g, 0.0, np.minimum(lat, d_face))
    total = direct
    if reflect:
        sin2 = 1.0 - uz * uz
        tir = (n_in / n_out) ** 2 * sin2 > 1.0
        f_top = np.where(tir, 1.0, refl_top)
        f_bot = np.where(tir, 1.0, refl_bottom)
        with np.errstate(divide="ignore", invalid="ignore"):
            d_top2 = np.where(uz > 0, (2 * Z - pz) / uz, np.inf)
            d_bot2 = np.where(uz < 0, -(pz + Z) / uz, np.inf)
        up = uz > 0
        refl = np.where(up, f_top * _track(g, d_top, np.minimum(lat, d_top2)),
                        f_bot * _track(g, d_bot, np.minimum(lat, d_bot2)))
        refl = np.where(np.isfinite(d_face) & (lat > d_face), refl, 0.0)
        total = total + refl
    w = (2.0 / n_mu) * (2.0 * np.pi / n_phi)
    return scale * beta * float(np.sum(total)) * w / (4.0 * np.pi)


# Stroud's degree-3 rule on the tetrahedron (barycentric points, weights
# summing to 1): exact for the product of a quadratic and a linear function.
_TET3_POINTS = np.array([[0.25, 0.25, 0.25, 0.25], [0.5, 1 / 6, 1 / 6, 1 / 6], [1 / 6, 0.5, 1 / 6, 1 / 6],
                         [1 / 6, 1 / 6, 0.5, 1 / 6], [1 / 6, 1 / 6, 1 / 6, 0.5]])
_TET3_WEIGHTS = np.array([-0.8, 0.45, 0.45, 0.45, 0.45])


def hat_smoothed_cell_flux(points, cells, volumes, targets, flux):
    """Expected value of the forward estimator's cell flux for each target
    cell of a gain-only Tet4 mesh: the mean over the cell's four vertices of
    the hat-function average of the point flux,
        D_v = integral phi(x) lambda_v(x) dx / integral lambda_v(x) dx,
    over the cells around v (the vertex deposit weights the track by the
    barycentric coordinate of its midpoint, which is exact in expectation for
    a linear lambda and zero gain).  ``flux(p)`` gives the point flux."""
    patch = {}
    for c, cv in enumerate(cells):
        for k, v in enumerate(cv):
            patch.setdefault(int(v), []).append((c, k))
    cache = {}

    def tet_values(c):
        if c not in cache:
            x = _TET3_POINTS @ points[cells[c]]
            cache[c] = np.array([flux(p) for p in x])
        return cache[c]

    out = []
    for c in targets:
        dens = []
        for v in cells[c]:
            num = den = 0.0
            for t, k in patch[int(v)]:
                lam = _TET3_POINTS[:, k]
                num += volumes[t] * np.sum(_TET3_WEIGHTS * lam * tet_values(t))
                den += volumes[t] * 0.25
            dens.append(num / den)
        out.append(np.mean(dens))
    return np.array(out)

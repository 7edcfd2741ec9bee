"""Learned proposal: a small network that stands in for the walked importance map.

The walked map (one centroid-to-sample prism walk per prism, slice and
sample) is the costly half of the estimator.  The learned proposal predicts
the walk's transfer factor (spectral gain times reflectance) from geometry
instead, then forms the same defensive mixture with the exact per-prism
factor beta V / (4 pi d^2).  The estimator stays unbiased whatever the
network predicts (any proposal with full support is), and samples whose
estimate misses the threshold after the first launch restart on the walked
map (``fallback``), so a poor network costs time, never accuracy.

Pipeline:
  1. ``random_pump_beta``: seeded random pump fields (Gaussian spots);
  2. ``training_rows``: the device's own transfer maps (``Context.transfer_map``)
     on those fields become (input, log transfer) rows;
  3. ``fit``: normalisation and AdamW training on the GPU (``_native.mlp_train``);
  4. ``Context.set_proposal`` / ``PhiASE(proposal=...)`` use it, evaluated per
     prism cluster by the BF16-MFMA kernel.

The reference has no learned proposal; its importance map is the retired
centroid-propagation pass (include/kernels/importanceSampling.hpp:10-18).
"""
from __future__ import annotations

import json
from dataclasses import dataclass, field

import numpy as np

from . import _native

__all__ = ["LearnedProposal", "random_pump_beta", "image_plane", "proposal_inputs", "training_rows",
           "fit_normalisation", "init_params", "fit", "train_on_workload"]

INPUT_NAMES = ("sample_x", "sample_y", "sample_z", "prism_x", "prism_y", "prism_z", "level", "beta", "distance",
               "log_sigma_ratio")


@dataclass
class LearnedProposal:
    params: np.ndarray
    shift: np.ndarray
    scale: np.ndarray
    cluster_triangles: int = 4
    cluster_levels: int = 2
    fallback: bool = True
    meta: dict = field(default_factory=dict)

    def save(self, path):
        np.savez(path, params=np.asarray(self.params, np.float32), shift=np.asarray(self.shift, np.float32),
                 scale=np.asarray(self.scale, np.float32),
                 layout=np.array([self.cluster_triangles, self.cluster_levels, int(self.fallback)], np.int64),
                 meta=np.array(json.dumps(self.meta)))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            ct, cl, fb = (int(v) for v in z["layout"])
            return cls(params=z["params"].copy(), shift=z["shift"].copy(), scale=z["scale"].copy(),
                       cluster_triangles=ct, cluster_levels=cl, fallback=bool(fb),
                       meta=json.loads(str(z["meta"])))

    def with_clusters(self, triangles, levels, fallback=None):
        return LearnedProposal(self.params, self.shift, self.scale, int(triangles), int(levels),
                               self.fallback if fallback is None else bool(fallback), dict(self.meta))


def image_plane(level, layers, reflections, first):
    """Unfolded plane (in layers) of a sample's mirror image after `reflections`
    bounces, the first on the bottom (first=0) or top (first=1) surface."""
    level = np.asarray(level, dtype=np.int64)
    m = int(reflections)
    if m == 0:
        return level
    if first == 1:
        return m * layers + (layers - level) if m & 1 else m * layers + level
    return -(m - 1) * layers - level if m & 1 else -m * layers + level


def proposal_inputs(sx, sy, sz, cx, cy, level, beta, h, log_sigma):
    """Unnormalised network inputs in float32 arithmetic (same order as the
    device: sample x, y, z; prism x, y, z; level; beta; distance; log ratio)."""
    f32 = np.float32
    sx, sy, sz, cx, cy, level, beta = np.broadcast_arrays(*(np.asarray(v, dtype=f32)
                                                           for v in (sx, sy, sz, cx, cy, level, beta)))
    cz = (level + f32(0.5)) * f32(h)
    dx, dy, dz = sx - cx, sy - cy, sz - cz
    dist = np.sqrt(dx * dx + dy * dy + dz * dz)
    cols = [sx, sy, sz, cx, cy, cz, level, beta, dist, np.full(sx.shape, f32(log_sigma))]
    return np.stack(cols, axis=-1).astype(f32)


def triangle_centroids_f32(topology):
    """Triangle centroids in the device's float32 arithmetic."""
    p = topology.points.astype(np.float32)
    tri = topology.trianglePointIndices
    x = (p[tri[:, 0], 0] + p[tri[:, 1], 0] + p[tri[:, 2], 0]) * np.float32(1.0 / 3.0)
    y = (p[tri[:, 0], 1] + p[tri[:, 1], 1] + p[tri[:, 2], 1]) * np.float32(1.0 / 3.0)
    return x.astype(np.float32), y.astype(np.float32)


def random_pump_beta(topology, rng, *, spots=(1, 4), widths=(0.3, 1.5), amplitude=(0.1, 0.45), noise=0.01,
                     cladding=None, center=(0.0, 0.0)):
    """Seeded random pump field: 1-4 Gaussian spots of random width and height
    (plus a random absorption profile through the thickness and N(0, noise)),
    clipped to [0, 1]; cladding triangles get beta = 0.  Spot centres fall
    within 0.77 of the mesh's extent around `center`."""
    cent = topology.prismCentroids()
    pr = np.hypot(topology.points[:, 0] - center[0], topology.points[:, 1] - center[1])
    extent = float(pr.max())
    beta = np.zeros(cent.shape[0])
    H = float(topology.thickness) * (int(topology.levels) - 1)
    for _ in range(int(rng.integers(spots[0], spots[1] + 1))):
        r0 = extent * np.sqrt(rng.uniform(0.0, 0.6))
        a0 = rng.uniform(0.0, 2.0 * np.pi)
        cx, cy = center[0] + r0 * np.cos(a0), center[1] + r0 * np.sin(a0)
        w = rng.uniform(*widths)
        amp = rng.uniform(*amplitude)
        alpha = rng.uniform(0.0, 2.0) / max(H, 1e-9)
        zfac = np.exp(-alpha * cent[:, 2]) if rng.uniform() < 0.5 else np.exp(-alpha * (H - cent[:, 2]))
        beta += amp * zfac * np.exp(-((cent[:, 0] - cx) ** 2 + (cent[:, 1] - cy) ** 2) / (2.0 * w * w))
    beta += rng.normal(0.0, noise, size=beta.size)
    beta = np.clip(beta, 0.0, 1.0)
    if cladding is not None and np.any(cladding):
        T = topology.numberOfTriangles
        beta[np.tile(np.asarray(cladding, bool), beta.size // T)] = 0.0
    return beta


def training_rows(ctx, topology, beta, *, log_sigma, rng, samples=64, block=16, rows_per_sample=2048,
                  importance_share=0.5):
    """(inputs, log transfer) rows from the context's transfer maps for the
    current beta: `samples` sample points in random runs of `block`
    consecutive samples; per sample, rows drawn half uniformly over walkable
    (slice, prism) pairs and half in proportion to their walked importance."""
    T = topology.numberOfTriangles
    layers = int(topology.levels) - 1
    h = float(topology.thickness)
    S = topology.numberOfSamples
    np_ = topology.numberOfPoints
    R = ctx.slices
    P = T * layers
    tx, ty = triangle_centroids_f32(topology)
    pts = topology.points.astype(np.float32)
    vol = topology.prismVolumes()
    cz = (np.arange(layers) + 0.5) * h
    X, Y = [], []
    starts = rng.integers(0, max(1, S - block), size=max(1, samples // block))
    for s0 in starts:
        count = int(min(block, S - s0))
        tm = ctx.transfer_map(int(s0), count)
        for i in range(count):
            s = int(s0) + i
            p, lv = s % np_, s // np_
            planes = np.array([image_plane(lv, layers, (k + 1) // 2, k & 1) for k in range(R)])
            tr = tm[i].reshape(-1)
            ok = np.flatnonzero(tr > 0.0)
            if ok.size == 0:
                continue
            k_ok, j_ok = ok // P, ok % P
            t_ok, l_ok = j_ok % T, j_ok // T
            d2 = ((tx[t_ok] - pts[p, 0]) ** 2 + (ty[t_ok] - pts[p, 1]) ** 2
                  + (cz[l_ok] - planes[k_ok] * h) ** 2)
            imp = beta[j_ok] * vol[j_ok] * tr[ok] / np.maximum(d2, 1e-12)
            n_imp = int(rows_per_sample * importance_share) if imp.sum() > 0 else 0
            pick = [rng.integers(0, ok.size, size=rows_per_sample - n_imp)]
            if n_imp:
                pick.append(rng.choice(ok.size, size=n_imp, p=imp / imp.sum()))
            sel = np.concatenate(pick)
            k, j = k_ok[sel], j_ok[sel]
            t, l = j % T, j // T
            X.append(proposal_inputs(pts[p, 0], pts[p, 1], planes[k] * h, tx[t], ty[t], l, beta[j], h, log_sigma))
            Y.append(np.log(tr[ok][sel]).astype(np.float32))
    return np.concatenate(X), np.concatenate(Y)


def fit_normalisation(X):
    """Per-input shift (mean) and scale (1 / std; 1 for constant inputs)."""
    X = np.asarray(X, dtype=np.float64)
    mu = X.mean(axis=0)
    sd = X.std(axis=0)
    scale = np.where(sd > 1e-6 * np.maximum(1.0, np.abs(mu)), 1.0 / np.where(sd > 0, sd, 1.0), 1.0)
    return mu.astype(np.float32), scale.astype(np.float32)


def init_params(seed=0):
    """He-style uniform initialisation (variance 2 / fan_in), zero biases, in
    the ABI order W1 b1 W2 b2 W3 b3 W4 b4 w5 b5."""
    rng = np.random.default_rng(seed)
    H, I = _native.MLP_HIDDEN, _native.MLP_INPUTS
    shapes = [(H, I), (H,), (H, H), (H,), (H, H), (H,), (H, H), (H,), (H,), (1,)]
    parts = []
    for i, shape in enumerate(shapes):
        if i % 2 == 0:   # W1 W2 W3 W4 w5
            lim = np.sqrt(6.0 / shape[-1] if len(shape) == 2 else 6.0 / H)
            parts.append(rng.uniform(-lim, lim, size=shape).reshape(-1))
        else:
            parts.append(np.zeros(int(np.prod(shape))))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == _native.mlp_param_count()
    return p


def fit(X, Y, *, epochs=12, batch=4096, learning_rate=2e-3, weight_decay=1e-5, seed=0, device=0,
        cluster_triangles=4, cluster_levels=2, fallback=True, decay_epochs=None):
    """Normalise the inputs and train the network on the GPU.  The learning
    rate drops tenfold for the last third of the epochs."""
    shift, scale = fit_normalisation(X)
    Xn = ((np.asarray(X, np.float32) - shift) * scale).astype(np.float32)
    params = init_params(seed)
    epochs = int(epochs)
    first = max(1, (2 * epochs) // 3) if decay_epochs is None else max(1, epochs - int(decay_epochs))
    params, h1 = _native.mlp_train(params, Xn, Y, epochs=first, batch=batch, learning_rate=learning_rate,
                                   weight_decay=weight_decay, seed=seed, device=device)
    hist = list(h1)
    if epochs > first:
        params, h2 = _native.mlp_train(params, Xn, Y, epochs=epochs - first, batch=batch,
                                       learning_rate=learning_rate * 0.1, weight_decay=weight_decay,
                                       seed=seed + 1, device=device)
        hist += list(h2)
    var = float(np.var(Y))
    meta = {"rows": int(X.shape[0]), "epochs": epochs, "batch": int(batch), "learning_rate": learning_rate,
            "loss_history": [float(v) for v in hist], "target_variance": var}
    return LearnedProposal(params, shift, scale, int(cluster_triangles), int(cluster_levels), bool(fallback), meta)


def train_on_workload(workload, *, fields=32, samples_per_field=64, rows_per_sample=1024, seed=0, device=0,
                      epochs=12, batch=4096, learning_rate=2e-3, cluster_triangles=4, cluster_levels=2,
                      log=None, center=None):
    """Training set from seeded random pump fields on the workload's mesh and
    a network fitted to it (pump spots around `center`, default the origin).
    Returns (proposal, (X, Y))."""
    from .simulation import native_problem

    rng = np.random.default_rng(seed)
    topo = workload.topology
    center = (0.0, 0.0) if center is None else center
    clad = np.asarray(workload.medium.value("claddingCellTypes")) == int(workload.medium.value("claddingNumber"))
    mesh, exp, comp, log_sigma = native_problem(workload, device=device)
    Xs, Ys = [], []
    with _native.Context(mesh, exp, comp) as ctx:
        for f in range(int(fields)):
            beta = random_pump_beta(topo, rng, cladding=clad, center=center)
            ctx.update_beta(beta)
            X, Y = training_rows(ctx, topo, beta, log_sigma=log_sigma, rng=rng, samples=samples_per_field,
                                 rows_per_sample=rows_per_sample)
            Xs.append(X)
            Ys.append(Y)
            if log:
                log(f"field {f + 1}/{fields}: {X.shape[0]} rows")
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    prop = fit(X, Y, epochs=epochs, batch=batch, learning_rate=learning_rate, seed=seed, device=device,
               cluster_triangles=cluster_triangles, cluster_levels=cluster_levels)
    prop.meta.update({"workload": workload.name, "fields": int(fields), "seed": int(seed)})
    return prop, (X, Y)

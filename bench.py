"""Benchmark: adaptive importance-sampled ASE on the production disk, 1 GPU.

One step = one full evaluation of the hot path (importance map of every
sample point, adaptive rays until every sample's mse <= 0.005 or 1e7 rays,
FP64 statistics) on the seeded production disk: radius 3 cm, ~2.1k points,
~4.1k triangles, 20 levels over 0.7 cm, 64 wavelengths, first-order top /
bottom reflections with per-triangle reflectivity ~U(0, 0.1), cladding ring
(5.5 cm^-1), adaptive 1e5 -> 1e7 rays per sample.  The mesh is resident on
the device before timing starts (persistent context).

Prints one JSON line.  ``value`` is rays/s of the whole job;
``wall_time_to_mse_s`` is the median warm time of a step.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "wall-time to max MSE<=0.005 + rays/s: learned-proposal vs plain IS, 1 MI355X"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# algorithmic cost of one prism step of the walk (rays and importance walks)
BYTES_PER_STEP = 56
FLOPS_PER_STEP = 40
FLOPS_PER_RAY = 40


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_context(w):
    from haseongpu_amd import _native

    s = w.settings
    sa, se, msa, mse_ = w.spectra.spectralPairs(bool(s["monochromatic"]))
    mesh = _native.NativeMesh(w.topology, w.medium)
    exp = _native.NativeExperiment(sigma_a=sa, sigma_e=se, min_rays=s["minRaysPerSample"],
                                   max_rays=s["maxRaysPerSample"], max_sigma_a=msa, max_sigma_e=mse_,
                                   mse_threshold=s["mseThreshold"], use_reflections=s["useReflections"],
                                   max_reflections=s.get("maxReflections", 1), defensive_fraction=0.05)
    comp = _native.make_compute(n_samples=w.samples, max_repetitions=s["repetitions"],
                                adaptive_steps=s["adaptiveSteps"], rng_seed=s["rngSeed"])
    return _native.Context(mesh, exp, comp), (sa, se, msa, mse_)


def cpu_baseline(w, spectra, seconds=15.0):
    """The FP64 oracle (a port of the estimator, OpenMP) on this host's cores,
    over a contiguous block of mid-slab sample points with the same settings."""
    from oracle import oracle as O

    s = w.settings
    sa, se, msa, mse_ = spectra
    threads = len(os.sched_getaffinity(0))
    threads = min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))
    om = O.OracleMesh(w.topology, w.medium)
    Np = w.topology.numberOfPoints
    first = (int(w.topology.levels) // 2) * Np
    block = max(threads, 8)
    done_rays, done_samples, elapsed = 0, 0, 0.0
    while elapsed < seconds and done_samples < Np:
        op = O.OracleParams(sigma_a=sa, sigma_e=se, max_sigma_a=msa, max_sigma_e=mse_,
                            n_tot=float(w.medium.value("nTot")), tau=float(w.medium.value("crystalTFluo")),
                            beta_cells=w.medium.value("betaCells"), min_rays=s["minRaysPerSample"],
                            max_rays=s["maxRaysPerSample"], mse_threshold=s["mseThreshold"],
                            use_reflections=s["useReflections"], max_reflections=s.get("maxReflections", 1),
                            adaptive_steps=s["adaptiveSteps"], seed=s["rngSeed"],
                            sample_min=first + done_samples, sample_max=first + done_samples + block - 1,
                            threads=threads, n_samples=w.samples)
        t0 = time.perf_counter()
        r = O.calc_phi_ase(om, op)
        elapsed += time.perf_counter() - t0
        done_rays += int(r["totalRays"].astype(np.int64).sum())
        done_samples += block
        log(f"cpu baseline: {done_samples} samples, {done_rays} rays, {elapsed:.1f}s")
    return {"value": done_rays / elapsed, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"{done_samples} sample points (contiguous, mid-slab level) of the same workload, "
                      f"full importance map + adaptive rays, {elapsed:.1f}s on {threads} threads",
            "samples_per_s": done_samples / elapsed}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--levels", type=int, default=20, help="override for quick local runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every sample point of a run is scheduled on a single device "
              f"(requested --gpus {args.gpus}); run it with --gpus 1", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))

    from haseongpu_amd.workloads import production_disk

    w = production_disk(levels=args.levels)
    topo = w.topology
    log(f"workload: {topo.numberOfPoints} points, {topo.numberOfTriangles} triangles, {topo.levels} levels, "
        f"{topo.numberOfPrisms} prisms, {w.samples} samples")
    t0 = time.perf_counter()
    ctx, spectra = build_context(w)
    upload_s = time.perf_counter() - t0
    log(f"context created (mesh resident) in {upload_s:.2f}s")

    t0 = time.perf_counter()
    res = ctx.evaluate()
    cold_s = time.perf_counter() - t0
    log(f"cold step {cold_s:.2f}s, rays {int(res.totalRays.astype(np.int64).sum())}")
    for i in range(max(0, args.warmup - 1)):
        ctx.evaluate()
        log(f"warmup {i + 2} done")

    times, rays, stats = [], 0, []
    for i in range(args.steps):
        t0 = time.perf_counter()
        res = ctx.evaluate()
        dt = time.perf_counter() - t0
        st = ctx.stats()
        times.append(dt)
        rays += int(res.totalRays.astype(np.int64).sum())
        stats.append(st)
        log(f"step {i}: {dt:.3f}s rays {st['rays']} max mse {np.nanmax(res.mse):.3e} "
            f"importance {st['importance_ms']:.0f}ms trace {st['trace_ms']:.0f}ms")
    total = float(np.sum(times))
    converged = bool(np.all(res.mse <= w.settings["mseThreshold"]))

    trace_ms = sum(s["trace_ms"] for s in stats)
    imp_ms = sum(s["importance_ms"] for s in stats)
    launches = sum(s["trace_launches"] for s in stats)
    imp_launches = sum(s["importance_launches"] for s in stats)
    ray_steps = sum(s["ray_steps"] for s in stats)
    imp_steps = sum(s["importance_steps"] for s in stats)
    total_rays = sum(s["rays"] for s in stats)
    kernels = {
        "hase_trace_kernel": {"launches": launches, "avg_ms": trace_ms / max(launches, 1),
                              "bytes_per_launch": ray_steps * BYTES_PER_STEP / max(launches, 1)},
        "hase_importance_kernel": {"launches": imp_launches, "avg_ms": imp_ms / max(imp_launches, 1),
                                   "bytes_per_launch": imp_steps * BYTES_PER_STEP / max(imp_launches, 1)},
    }
    dominant = max(kernels, key=lambda k: kernels[k]["avg_ms"] * kernels[k]["launches"])
    kd = kernels[dominant]
    achieved = kd["bytes_per_launch"] / (kd["avg_ms"] * 1e-3) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get(dominant)
    valu_tflops = ((ray_steps + imp_steps) * FLOPS_PER_STEP + total_rays * FLOPS_PER_RAY) / ((trace_ms + imp_ms) * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": rays / total,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "accumulate": "f64",
        "data": "synthetic (seeded production disk; no external data)",
        "config": {
            "workload": "production_disk",
            "radius_cm": 3.0, "points": topo.numberOfPoints, "triangles": topo.numberOfTriangles,
            "levels": int(topo.levels), "prisms": topo.numberOfPrisms, "sample_points": w.samples,
            "wavelengths": int(spectra[0].size), "reflection_order": w.settings.get("maxReflections", 1),
            "rays_per_sample": [w.settings["minRaysPerSample"], w.settings["maxRaysPerSample"]],
            "mse_threshold": w.settings["mseThreshold"], "proposal": "plain importance map",
        },
        "wall_time_to_mse_s": float(np.median(times)),
        "cold_first_call_s": cold_s,
        "all_samples_converged": converged,
        "prism_steps_per_s": (ray_steps + imp_steps) / total,
        "rays_per_step": rays / args.steps,
        "learned_proposal": None,
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_step": BYTES_PER_STEP, "fp32_valu_tflops": valu_tflops,
                     "fp32_valu_frac": valu_tflops / FP32_PEAK_TFLOPS},
        "kernels": kernels,
    }
    ctx.close()
    if rank == 0 and not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, spectra, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

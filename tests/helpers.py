"""Shared loaders for the golden fixtures and synthetic meshes used by tests."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from haseongpu_amd.geometry import GainMedium, MeshTopology

GOLDEN = Path(__file__).resolve().parent / "golden"


def load_compare_serial(name):
    """Legacy serial solver output (1e6 rays/sample) and its prism input.
    Returns (topology, medium, spectra, reference)."""
    d = np.load(GOLDEN / f"compareSerial_{name}.npz")
    topo = MeshTopology(d["points"], d["triangles"], levels=int(d["levels"]),
                        thickness=float(d["thickness"]))
    medium = GainMedium(topo).withPhysicalProperties(
        betaVolume=d["betaVolume"],
        betaCells=d["betaCells"],
        claddingCellTypes=d["claddingCellTypes"],
        claddingNumber=int(d["claddingNumber"]),
        claddingAbsorption=float(d["claddingAbsorption"]),
        nTot=float(d["nTot"]),
        crystalTFluo=float(d["crystalTFluo"]),
        refractiveIndices=d["refractiveIndices"],
    )
    spectra = {"sigmaA": float(d["sigmaA"]), "sigmaE": float(d["sigmaE"])}
    ref = {"phiASE": d["phiASE"], "dndtAse": d["dndtAse"], "raysPerSample": int(d["raysPerSample"])}
    return topo, medium, spectra, ref


def load_true_cladding(index):
    """Legacy importance-sampled phiASE/mse/totalRays at 2000 rays/sample for
    cladding absorption index 0 (0), 1 (5.5) or 2 (55 cm^-1)."""
    d = np.load(GOLDEN / "trueCladding.npz")
    topo = MeshTopology(d["points"], d["triangles"], levels=int(d["levels"]),
                        thickness=float(d["thickness"]))
    S = topo.numberOfSamples
    medium = GainMedium(topo).withPhysicalProperties(
        betaVolume=d["betaVolume"],
        betaCells=np.full(S, float(d["betaCells"])),
        claddingCellTypes=d["claddingCellTypes"],
        claddingNumber=int(d["claddingNumber"]),
        claddingAbsorption=float(d["claddingAbsorptions"][index]),
        nTot=float(d["nTot"]),
        crystalTFluo=float(d["crystalTFluo"]),
    )
    spectra = {"sigmaA": float(d["sigmaA"]), "sigmaE": float(d["sigmaE"])}
    ref = {"phiASE": d["phiASE"][index], "mse": d["mse"][index], "totalRays": d["totalRays"][index],
           "raysPerSample": int(d["raysPerSample"]), "rngSeed": int(d["rngSeed"])}
    return topo, medium, spectra, ref

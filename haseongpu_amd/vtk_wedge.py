"""Legacy ASCII VTK output on the extruded wedge mesh.

Two writers, both producing an UNSTRUCTURED_GRID of VTK_WEDGE cells (type 13)
over the mesh's Np x L points (point p of level l is VTK point p + l * Np, at
z = l * thickness) and T x (L-1) prisms (prism j = t + l * T is VTK cell j,
its six points the triangle's corners on levels l and l + 1):

* ``vtkWedge`` -- the front end's writer (pyInclude/vtkWedge.py:203-265):
  one file with any number of named point- or prism-shaped scalar arrays,
  ``{step}`` / ``{time}`` / ``{field}`` placeholders in the file name, and a
  callback-factory form for ``Simulation.on_step``.  Values are written with
  17 significant digits, so a file read back gives the FP64 values bit for bit.
* ``writePointsToVtk`` / ``writePrismToVtk`` -- the run-level writers of the
  C++ entry (include/utils/writeToVtk.hpp:27-97, src/utils/writeToVtk.cpp:
  24-124): one array per file, a header line carrying the ray count, the
  adaptive budget, the threshold, the reflection request and the runtime,
  values in fixed notation with six decimals.  ``PhiASE(writeVtk=True)``
  writes its per-sample results through them, as include/core/simulation.hpp:
  161-224 does (dndt_, phiase_, standard_error_, total_rays_ and
  relative_standard_error_<unix time>.vtk under <outputPath>/vtk).

``compareVtk`` reads the scalars of such a file back and returns the
per-entry relative difference in percent (src/utils/writeToVtk.cpp:
126-220); ``readVtkScalars`` returns every named scalar array of a file.
"""
from __future__ import annotations

from collections.abc import Mapping, Sequence
from pathlib import Path

import numpy as np

__all__ = ["vtkWedge", "writePointsToVtk", "writePrismToVtk", "compareVtk", "readVtkScalars"]

VTK_WEDGE = 13


def _topology_of(geometry):
    """MeshTopology of a MeshTopology or of a GainMedium that owns one."""
    from .geometry import GainMedium, MeshTopology

    if isinstance(geometry, GainMedium):
        return geometry.topology
    if isinstance(geometry, MeshTopology):
        return geometry
    raise TypeError("geometry must be a GainMedium or MeshTopology")


def _is_geometry(obj):
    from .geometry import GainMedium, MeshTopology

    return isinstance(obj, (GainMedium, MeshTopology))


def _is_list(obj):
    return isinstance(obj, Sequence) and not isinstance(obj, (str, bytes))


def _label(field):
    if isinstance(field, Mapping):
        return "_".join(str(k) for k in field)
    if _is_list(field):
        return "_".join(str(k) for k in field)
    return str(field)


def _file_path(file_name, state, field):
    text = str(file_name)
    if state is not None:
        text = text.format(step=getattr(state, "step", 0), time=getattr(state, "time", 0.0), field=_label(field))
    if not text.endswith(".vtk"):
        text += ".vtk"
    return Path(text)


def _lookup(data, key):
    """data[key] for a mapping, data.key for an object; None when absent."""
    if isinstance(data, Mapping):
        return data.get(key)
    return getattr(data, key, None)


def _has(data, key):
    return (key in data) if isinstance(data, Mapping) else hasattr(data, key)


def _named_arrays(data, field, scalar_name, fields):
    """[(name, float64 array)] selected from ``data`` by ``field`` / ``fields``."""
    out = []
    if fields is not None:
        if not isinstance(fields, Mapping):
            raise TypeError("fields must be a mapping of scalar names to arrays or attribute names")
        for name, source in fields.items():
            values = _lookup(data, source) if isinstance(source, str) and _has(data, source) else source
            if values is None:
                raise ValueError(f"no data available for VTK field '{name}'")
            out.append((str(name), np.asarray(values, dtype=np.float64)))
        return out

    def alias(key, index):
        if scalar_name is None:
            return None
        if isinstance(scalar_name, Mapping):
            return scalar_name.get(key)
        if _is_list(scalar_name):
            return scalar_name[index]
        return scalar_name

    def one(key, name):
        if _has(data, key):
            values, label = _lookup(data, key), name or key
        else:
            values, label = data, name or "scalars"
        if values is None:
            raise ValueError(f"no data available for VTK field '{key}'")
        return str(label), np.asarray(values, dtype=np.float64)

    if isinstance(field, Mapping):
        for i, (name, source) in enumerate(field.items()):
            label = alias(name, i) or name
            if isinstance(source, str):
                out.append((str(label), one(source, label)[1]))
            else:
                out.append((str(label), np.asarray(source, dtype=np.float64)))
        return out
    if _is_list(field):
        return [one(key, alias(key, i)) for i, key in enumerate(field)]
    return [one(field, alias(field, 0))]


def _classify(values, topology):
    """("POINT_DATA" | "CELL_DATA", values flattened in VTK order)."""
    Np, T, L = topology.numberOfPoints, topology.numberOfTriangles, int(topology.levels)
    point_shape, cell_shape = (Np, L), (T, L - 1)
    if values.shape == point_shape or (values.ndim == 1 and values.size == Np * L):
        return "POINT_DATA", values.reshape(-1, order="F")
    if values.shape == cell_shape or (values.ndim == 1 and values.size == T * (L - 1)):
        return "CELL_DATA", values.reshape(-1, order="F")
    raise ValueError(f"VTK wedge data must match point data shape {point_shape} or cell data shape "
                     f"{cell_shape}, got {values.shape}")


def _grid_lines(topology, coord_fmt):
    """POINTS / CELLS / CELL_TYPES sections of the wedge grid, as text."""
    topology._require_levels()
    topology._require_thickness()
    Np, L = topology.numberOfPoints, int(topology.levels)
    tri = np.asarray(topology.trianglePointIndices, dtype=np.int64).reshape(-1, 3)
    T = tri.shape[0]
    xy = np.asarray(topology.points, dtype=np.float64).reshape(-1, 2)
    z = topology.levelCoordinates()
    pts = np.column_stack([np.tile(xy, (L, 1)), np.repeat(z, Np)])
    lower = tri[None, :, :] + (np.arange(L - 1) * Np)[:, None, None]
    cells = np.concatenate([lower, lower + Np], axis=2).reshape(-1, 6)
    P = T * (L - 1)
    parts = [f"POINTS {Np * L} float\n"]
    parts.append("".join(f"{coord_fmt(a)} {coord_fmt(b)} {coord_fmt(c)}\n" for a, b, c in pts.tolist()))
    parts.append(f"CELLS {P} {P * 7}\n")
    parts.append("".join("6 %d %d %d %d %d %d\n" % tuple(c) for c in cells.tolist()))
    parts.append(f"CELL_TYPES {P}\n")
    parts.append(f"{VTK_WEDGE}\n" * P)
    return parts


def _g17(v):
    return f"{v:.17g}"


def _write_wedge(path, arrays, topology):
    sections = {"POINT_DATA": [], "CELL_DATA": []}
    for name, values in arrays:
        kind, flat = _classify(values, topology)
        sections[kind].append((name, flat))
    parts = ["# vtk DataFile Version 2.0\n", "HASEonGPU wedge output\n", "ASCII\n", "DATASET UNSTRUCTURED_GRID\n"]
    parts += _grid_lines(topology, _g17)
    for kind in ("POINT_DATA", "CELL_DATA"):
        if not sections[kind]:
            continue
        parts.append(f"{kind} {sections[kind][0][1].size}\n")
        for name, flat in sections[kind]:
            parts.append(f"SCALARS {name} float 1\nLOOKUP_TABLE default\n")
            parts.append("".join(f"{v:.17g}\n" for v in flat.tolist()))
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text("".join(parts), encoding="utf-8")
    return path


def vtkWedge(file_name, data=None, geometry=None, field="phiAse", scalar_name=None, every=1, fields=None):
    """Write scalar arrays on the wedge mesh to a legacy ASCII VTK file
    (pyInclude/vtkWedge.py:203-265).

    ``vtkWedge(name, state, fields=...)`` writes a ``TimeStepState`` (its
    ``topology`` gives the grid); ``vtkWedge(name, values, geometry=...)``
    writes plain arrays on a ``MeshTopology`` / ``GainMedium``; and
    ``vtkWedge(name, geometry, ...)`` returns a callback for
    ``Simulation.on_step`` that writes every ``every``-th step.

    ``field`` names attributes or keys of the data: a string writes one array,
    a sequence several under their own names, a mapping {alias: source}
    several under aliases.  ``fields`` maps scalar names to arrays (or to
    attribute names) and takes precedence.  Arrays of shape (Np, L) are
    POINT_DATA, (T, L-1) CELL_DATA (flat arrays of those sizes too, in the
    column-major order of the mesh); both kinds may share a file.  The file
    name takes ``{step}``, ``{time}`` and ``{field}`` placeholders and gets a
    ``.vtk`` suffix when it has none.  Returns the written path (the callback
    returns None on skipped steps)."""
    if _is_geometry(data) and geometry is None:
        topology = _topology_of(data)

        def on_step(state):
            if int(every) <= 0:
                raise ValueError("every must be a positive integer")
            if getattr(state, "step", 0) % int(every) != 0:
                return None
            return _write_wedge(_file_path(file_name, state, field),
                                _named_arrays(state, field, scalar_name, fields), topology)

        return on_step

    if geometry is None:
        topology = getattr(data, "topology", None)
        if topology is None:
            raise TypeError("vtkWedge needs a TimeStepState with a topology or an explicit GainMedium/MeshTopology")
    else:
        topology = _topology_of(geometry)
    return _write_wedge(_file_path(file_name, data, field), _named_arrays(data, field, scalar_name, fields), topology)


def _run_header(rays, max_rays, threshold, use_reflections, runtime):
    # the C++ writer streams the two floats with the default (%g, 6 digits) format
    return (f"RAYS={int(rays)} MAXRAYS={int(max_rays)} REFLECTIONS=0 EXPECTATION={float(threshold):g} "
            f"RUNTIME={float(runtime):g} REQUESTED_REFLECTIONS={int(bool(use_reflections))}\n")


def _write_run(kind, geometry, values, filename, rays, max_rays, threshold, use_reflections, runtime):
    topology = _topology_of(geometry)
    values = np.asarray(values, dtype=np.float64)
    got_kind, flat = _classify(values, topology)
    if got_kind != kind:
        raise ValueError(f"{kind} needs {'(Np, L)' if kind == 'POINT_DATA' else '(T, L-1)'} values, "
                         f"got {values.shape}")
    fixed6 = lambda v: f"{v:.6f}"   # noqa: E731
    parts = ["# vtk DataFile Version 2.0\n", _run_header(rays, max_rays, threshold, use_reflections, runtime),
             "ASCII\n", "DATASET UNSTRUCTURED_GRID\n"]
    parts += _grid_lines(topology, fixed6)
    parts.append(f"{kind} {flat.size}\nSCALARS scalars float 1\nLOOKUP_TABLE default\n")
    parts.append("".join(f"{v:.6f}\n" for v in flat.tolist()))
    path = Path(filename)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text("".join(parts), encoding="utf-8")
    return 0


def writePointsToVtk(geometry, ase, filename, raysLaunched, maxRays, relativeStandardErrorThreshold,
                     useReflections, runtime):
    """One per-sample array (Np x L, sample s = p + l * Np) as POINT_DATA
    (include/utils/writeToVtk.hpp:27-53).  Returns 0."""
    return _write_run("POINT_DATA", geometry, ase, filename, raysLaunched, maxRays, relativeStandardErrorThreshold,
                      useReflections, runtime)


def writePrismToVtk(geometry, prismData, filename, raysLaunched, maxRays, relativeStandardErrorThreshold,
                    useReflections, runtime):
    """One per-prism array (T x (L-1), prism j = t + l * T) as CELL_DATA
    (include/utils/writeToVtk.hpp:55-81).  Returns 0."""
    return _write_run("CELL_DATA", geometry, prismData, filename, raysLaunched, maxRays,
                      relativeStandardErrorThreshold, useReflections, runtime)


def readVtkScalars(filename):
    """{scalar name: (data kind, float64 array)} of a legacy ASCII VTK file."""
    tokens = Path(filename).read_text(encoding="utf-8").split()
    out, kind, count, i = {}, None, 0, 0
    while i < len(tokens):
        tok = tokens[i].upper()
        if tok in ("POINT_DATA", "CELL_DATA"):
            kind, count = tok, int(tokens[i + 1])
            i += 2
        elif tok == "SCALARS" and kind is not None:
            name, i = tokens[i + 1], i + 3
            comps = 1
            if tokens[i].upper() != "LOOKUP_TABLE":
                comps, i = int(tokens[i]), i + 1
            if tokens[i].upper() == "LOOKUP_TABLE":
                i += 2
            n = count * comps
            out[name] = (kind, np.array(tokens[i:i + n], dtype=np.float64))
            i += n
        else:
            i += 1
    return out


def compareVtk(compare, filename):
    """Relative difference in percent of each entry of ``compare`` against the
    first scalar array of ``filename`` (src/utils/writeToVtk.cpp:126-220):
    |larger / smaller| - 1 over the two magnitudes, times 100.  An empty
    file name gives an empty array; an unreadable file raises."""
    if str(filename) == "":
        return np.zeros(0)
    path = Path(filename)
    if not path.is_file():
        raise FileNotFoundError(f"can't open {path} for comparison")
    scalars = readVtkScalars(path)
    if not scalars:
        raise ValueError(f"{path} holds no scalar data")
    ref = next(iter(scalars.values()))[1]
    ours = np.asarray(compare, dtype=np.float64).reshape(-1)
    n = min(ref.size, ours.size)
    a, b = np.abs(ref[:n]), np.abs(ours[:n])
    with np.errstate(divide="ignore", invalid="ignore"):
        diff = np.where(a > b, a / b, b / a) - 1.0
    return diff * 100.0

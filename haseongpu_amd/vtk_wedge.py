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
_ABSENT = object()


def _mesh(obj):
    """MeshTopology behind ``obj`` (itself, or a GainMedium's), None otherwise."""
    from .geometry import GainMedium, MeshTopology

    if isinstance(obj, MeshTopology):
        return obj
    if isinstance(obj, GainMedium):
        return obj.topology
    return None


def _several(obj):
    return isinstance(obj, Sequence) and not isinstance(obj, (str, bytes))


def _stored(data, key):
    """``data``'s entry (mappings) or attribute named ``key``; _ABSENT if none."""
    if isinstance(data, Mapping) and key in data:
        return data[key]
    return getattr(data, key, _ABSENT)


class _Request:
    """One array to write: its VTK name, and either the key it is stored
    under in the data or the array itself.  ``else_data`` lets a missing key
    stand for the data object itself (a bare array passed as ``data``)."""

    __slots__ = ("name", "key", "array", "else_data")

    def __init__(self, name, key=None, array=None, else_data=False):
        self.name, self.key, self.array, self.else_data = name, key, array, else_data

    def resolve(self, data):
        if self.key is None:
            values = self.array
        else:
            values = _stored(data, self.key)
            if values is _ABSENT:
                if not self.else_data:
                    raise ValueError(f"VTK array '{self.name}': the data has no entry or attribute '{self.key}'")
                values = data
        if values is None:
            raise ValueError(f"VTK array '{self.name}' is None; there is nothing to write")
        return str(self.name), np.asarray(values, dtype=np.float64)


def _requests(field, scalar_name, fields):
    """Normalise the three ways of choosing arrays into a list of _Request.

    ``fields`` (a mapping {VTK name: array or key}) wins over ``field``.
    ``field`` is a key, a list of keys or a mapping {VTK name: key or array};
    ``scalar_name`` renames a key's array: one string, a list by position or
    a mapping by key.  A key the data lacks makes the data itself the array,
    written as ``scalars`` unless renamed."""
    if fields is not None:
        if not isinstance(fields, Mapping):
            raise TypeError("fields= expects a mapping {VTK scalar name: array or key in the data}")
        out = []
        for name, source in fields.items():
            if isinstance(source, str):
                out.append(_Request(name, key=source))
            else:
                out.append(_Request(name, array=source))
        return out

    def renamed(key, position):
        if isinstance(scalar_name, Mapping):
            return scalar_name.get(key)
        if _several(scalar_name):
            return scalar_name[position]
        return scalar_name

    if isinstance(field, Mapping):
        out = []
        for position, (label, source) in enumerate(field.items()):
            name = renamed(label, position) or label
            if isinstance(source, str):
                out.append(_Request(name, key=source, else_data=True))
            else:
                out.append(_Request(name, array=source))
        return out
    keys = list(field) if _several(field) else [field]
    return [_KeyRequest(key, renamed(key, position)) for position, key in enumerate(keys)]


class _KeyRequest(_Request):
    """A key whose array is written under its own name (or ``rename``); when
    the data lacks the key the data is the array, named ``rename`` or
    ``scalars``."""

    __slots__ = ("rename",)

    def __init__(self, key, rename):
        super().__init__(rename or key, key=key, else_data=True)
        self.rename = rename

    def resolve(self, data):
        if _stored(data, self.key) is _ABSENT:
            return _Request(self.rename or "scalars", array=data).resolve(data)
        return super().resolve(data)


def _label(field):
    """File-name label of a field choice: its keys joined by underscores."""
    names = list(field) if isinstance(field, Mapping) or _several(field) else [field]
    return "_".join(str(n) for n in names)


def _file_path(template, state, field):
    text = str(template)
    if state is not None:
        text = text.format(step=getattr(state, "step", 0), time=getattr(state, "time", 0.0), field=_label(field))
    return Path(text if text.endswith(".vtk") else text + ".vtk")


class _WedgeSizes:
    """Point and prism counts of the extruded mesh, and where an array goes."""

    def __init__(self, topology):
        L = int(topology.levels)
        self.points = (topology.numberOfPoints, L)
        self.prisms = (topology.numberOfTriangles, L - 1)

    def place(self, values):
        """("POINT_DATA" | "CELL_DATA", values in VTK order).  Matrix input is
        (Np, L) or (T, L-1) in the mesh's column-major order; flat input of
        either size is taken as already ordered (sample points first)."""
        for kind, shape in (("POINT_DATA", self.points), ("CELL_DATA", self.prisms)):
            if values.shape == shape:
                return kind, values.reshape(-1, order="F")
        if values.ndim == 1:
            for kind, shape in (("POINT_DATA", self.points), ("CELL_DATA", self.prisms)):
                if values.size == shape[0] * shape[1]:
                    return kind, values
        raise ValueError(f"an array of shape {values.shape} is neither per sample point {self.points} "
                         f"nor per prism {self.prisms} of this wedge mesh")


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
    sizes = _WedgeSizes(topology)
    by_kind = {"POINT_DATA": [], "CELL_DATA": []}
    for name, values in arrays:
        kind, flat = sizes.place(values)
        by_kind[kind].append((name, flat))
    text = ["# vtk DataFile Version 2.0\n", "HASEonGPU wedge output\n", "ASCII\n", "DATASET UNSTRUCTURED_GRID\n"]
    text += _grid_lines(topology, _g17)
    for kind, entries in by_kind.items():
        if entries:
            text.append(f"{kind} {entries[0][1].size}\n")
        for name, flat in entries:
            text.append(f"SCALARS {name} float 1\nLOOKUP_TABLE default\n")
            text.append("".join(f"{_g17(v)}\n" for v in flat.tolist()))
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    path.write_text("".join(text), encoding="utf-8")
    return path


class _StepWriter:
    """``Simulation.on_step`` callback: writes the chosen arrays of every
    ``every``-th step's state on a fixed mesh; other steps return None."""

    def __init__(self, template, topology, field, scalar_name, fields, every):
        self.template, self.topology, self.every = template, topology, every
        self.field = field
        self.requests = _requests(field, scalar_name, fields)

    def __call__(self, state):
        period = int(self.every)
        if period < 1:
            raise ValueError(f"every must be a positive step count, got {self.every}")
        if getattr(state, "step", 0) % period:
            return None
        return _write_wedge(_file_path(self.template, state, self.field),
                            [r.resolve(state) for r in self.requests], self.topology)


def vtkWedge(file_name, data=None, geometry=None, field="phiAse", scalar_name=None, every=1, fields=None):
    """Write scalar arrays on the wedge mesh to a legacy ASCII VTK file
    (the front end's writer, pyInclude/vtkWedge.py:203-265).

    Three call forms: ``vtkWedge(name, state, fields=...)`` writes a
    ``TimeStepState`` on the mesh the state carries; ``vtkWedge(name, values,
    geometry=mesh)`` writes arrays on a ``MeshTopology`` / ``GainMedium``; and
    ``vtkWedge(name, mesh, ...)`` returns a callback for ``Simulation.on_step``
    that writes every ``every``-th step.

    Arrays are chosen by ``field`` (a key, a list of keys, or {VTK name:
    key or array}), renamed by ``scalar_name``, or given outright by
    ``fields`` ({VTK name: array or key}), which wins.  (Np, L) arrays are
    written as POINT_DATA and (T, L-1) arrays as CELL_DATA, flat arrays of
    those sizes too; one file may hold both.  ``{step}``, ``{time}`` and
    ``{field}`` in the name are filled from the state, and ``.vtk`` is
    appended when missing.  Returns the path written (None from the callback
    on skipped steps)."""
    if geometry is None and _mesh(data) is not None:
        return _StepWriter(file_name, _mesh(data), field, scalar_name, fields, every)
    if geometry is not None:
        topology = _mesh(geometry)
        if topology is None:
            raise TypeError(f"geometry must be a MeshTopology or GainMedium, not {type(geometry).__name__}")
    else:
        topology = getattr(data, "topology", None)
        if topology is None:
            raise TypeError("vtkWedge cannot tell the mesh: pass geometry=, or a state that carries its topology")
    arrays = [r.resolve(data) for r in _requests(field, scalar_name, fields)]
    return _write_wedge(_file_path(file_name, data, field), arrays, topology)


def _run_header(rays, max_rays, threshold, use_reflections, runtime):
    # the C++ writer streams the two floats with the default (%g, 6 digits) format
    return (f"RAYS={int(rays)} MAXRAYS={int(max_rays)} REFLECTIONS=0 EXPECTATION={float(threshold):g} "
            f"RUNTIME={float(runtime):g} REQUESTED_REFLECTIONS={int(bool(use_reflections))}\n")


def _write_run(kind, geometry, values, filename, rays, max_rays, threshold, use_reflections, runtime):
    topology = _mesh(geometry)
    if topology is None:
        raise TypeError(f"geometry must be a MeshTopology or GainMedium, not {type(geometry).__name__}")
    values = np.asarray(values, dtype=np.float64)
    got_kind, flat = _WedgeSizes(topology).place(values)
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

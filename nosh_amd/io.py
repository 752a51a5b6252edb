"""Mesh and state I/O and the continuation observers (host side).

Mirrors the reference's output side:

* ``CsvWriter`` -- ``nosh::csv_writer`` (src/csv_writer.cpp:1-141): a header
  line that starts with "#", columns joined by the delimiter plus two spaces,
  doubles in scientific notation with 15 digits, left-aligned in 22
  characters, integers in max(5, len(label)) characters, rows indented by
  len("#") + 1 spaces.
* ``ContinuationObserver`` -- ``nosh::observer`` (src/observer.cpp:120-160):
  one row per accepted continuation step with "(0) step", "(1) <param>",
  "(2) Gibbs energy" and "(2) ||x||_2 scaled" = sqrt(inner_product(psi,
  psi)). The reference's ``nls::gibbs_energy`` and ``nls::inner_product``
  return 0.0 (src/model_evaluator_nls.cpp:697-770); the values written here
  are the ones those functions compute before returning.
* ``EigenDataSaver`` -- ``nosh::save_eigen_data`` (src/save_eigen_data.cpp:
  40-150): per step "##step", "#0unstable", "#1null", "#2stable" (an
  eigenvalue above 1e-5 counts as unstable, within +-1e-5 as null, below as
  stable: LOCA's psi_t = F sign convention) and "NN-0Re()" per eigenvalue.
* ``ContinuationDataSaver`` -- ``nosh::continuation_data_saver``
  (src/continuation_data_saver.hpp:24-50): one state file per accepted step,
  "out%04d" + extension, holding the mesh, psi and the parameter.

The reference writes MOAB .h5m files (HDF5); HDF5 is not available here, so
state files are VTK XML unstructured grids (.vtu, raw little-endian appended
data, readable by ParaView) with point data "psi" (Re, Im) and field data
with the parameters, and checkpoints are NumPy .npz archives (no pickles).
Meshes are read from .vtu, legacy ASCII .vtk and Exodus II (netCDF classic)
files; cells keep their file order, so ``nosh_amd.Mesh`` numbers edges as the reference does for the same
cell order (first appearance, src/mesh.cpp:629-691).
"""
from __future__ import annotations

import os
import re
import struct
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# VTK cell types.
VTK_TRIANGLE, VTK_TETRA = 5, 10


# --------------------------------------------------------------------- CSV
class CsvWriter:
    """Column writer with the reference csv_writer's layout."""

    header_start = "#"
    double_prec = 15
    double_width = double_prec + 7
    int_width = 5

    def __init__(self, file_name: str, delimiter: str = " "):
        self.delimiter = delimiter
        self._f = open(file_name, "w") if file_name else None

    def close(self):
        if self._f:
            self._f.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _width(self, label, value) -> int:
        if isinstance(value, (bool, np.bool_)):
            raise TypeError(f'Invalid data type for item "{label}".')
        if isinstance(value, (int, np.integer)):
            return max(self.int_width, len(label))
        if isinstance(value, (float, np.floating)):
            return self.double_width
        if isinstance(value, str):
            return len(value)
        raise TypeError(f'Invalid data type for item "{label}".')

    def _fmt(self, value) -> str:
        if isinstance(value, (float, np.floating)):
            return f"{float(value):.{self.double_prec}e}"
        return str(value)

    def header_line(self, items: Sequence[Tuple[str, object]]) -> str:
        sep = self.delimiter + "  "
        return self.header_start + " " + sep.join(
            label.ljust(self._width(label, v)) for label, v in items)

    def row_line(self, items: Sequence[Tuple[str, object]]) -> str:
        sep = self.delimiter + "  "
        return " " * (len(self.header_start) + 1) + sep.join(
            self._fmt(v).ljust(self._width(label, v)) for label, v in items)

    def write_header(self, items):
        if self._f:
            self._f.write(self.header_line(items) + "\n")
            self._f.flush()

    def write_row(self, items):
        if self._f:
            self._f.write(self.row_line(items) + "\n")
            self._f.flush()


def read_csv(file_name: str) -> Tuple[List[str], np.ndarray]:
    """(labels, rows) of a file written by CsvWriter (labels from the header;
    columns split on runs of three or more blanks)."""
    with open(file_name) as f:
        lines = [ln.rstrip("\n") for ln in f if ln.strip()]
    labels = [s.strip() for s in re.split(r" {3,}", lines[0][2:].strip())]
    rows = [[float(v) for v in ln.split()] for ln in lines[1:]]
    return labels, np.array(rows)


class ContinuationObserver:
    """One CSV row per accepted step, as nosh::observer writes.

    With ``model`` the Gibbs energy and the scaled norm are computed by the
    library; without one, pass them to ``observe``.
    """

    def __init__(self, csv_file: str, param_name: str = "mu", model=None):
        self.writer = CsvWriter(csv_file, " ")
        self.param_name = param_name
        self.model = model
        self.index = -1

    def observe(self, psi=None, param: float = 0.0, gibbs: Optional[float] = None,
                norm: Optional[float] = None):
        self.index += 1
        if gibbs is None or norm is None:
            import nosh_amd as nosh
            if self.model is None or psi is None:
                raise ValueError("observe needs psi and a model, or gibbs and norm")
            if gibbs is None:
                gibbs = nosh.gibbs_energy(self.model, psi)
            if norm is None:
                norm = float(np.sqrt(nosh.inner_product(self.model, psi, psi)))
        items = [("(0) step", int(self.index)), (f"(1) {self.param_name}", float(param)),
                 ("(2) Gibbs energy", float(gibbs)), ("(2) ||x||_2 scaled", float(norm))]
        if self.index == 0:
            self.writer.write_header(items)
        self.writer.write_row(items)

    def close(self):
        self.writer.close()


class EigenDataSaver:
    """Eigenvalue statistics per step, as nosh::save_eigen_data writes."""

    tol = 1.0e-5

    def __init__(self, csv_file: str):
        self.writer = CsvWriter(csv_file, " ")
        self.step = 0

    def classify(self, evals) -> Tuple[int, int, int]:
        ev = np.asarray(evals, dtype=np.float64)
        unstable = int(np.sum(ev > self.tol))
        stable = int(np.sum(ev < -self.tol))
        return unstable, len(ev) - unstable - stable, stable

    def save(self, evals, step: Optional[int] = None):
        if step is not None:
            self.step = step
        u, z, s = self.classify(evals)
        items = [("##step", int(self.step)), ("#0unstable", u), ("#1null", z), ("#2stable", s)]
        items += [(f"{k:02d}-0Re()", float(v)) for k, v in enumerate(evals)]
        if self.step == 0:
            self.writer.write_header(items)
        self.writer.write_row(items)
        self.step += 1

    def close(self):
        self.writer.close()


# --------------------------------------------------------------------- VTU
_VTK_TYPES = {np.dtype("float64"): "Float64", np.dtype("float32"): "Float32",
              np.dtype("int32"): "Int32", np.dtype("int64"): "Int64",
              np.dtype("uint8"): "UInt8"}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


def write_vtu(file_name: str, coords, cells, point_data: Optional[Dict[str, np.ndarray]] = None,
              field_data: Optional[Dict[str, float]] = None) -> None:
    """VTK XML unstructured grid, appended raw data (little endian, UInt64
    block headers). coords: n x 2 or n x 3; cells: triangles or tetrahedra."""
    coords = np.asarray(coords, dtype=np.float64)
    if coords.shape[1] == 2:
        coords = np.hstack([coords, np.zeros((coords.shape[0], 1))])
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    nv = cells.shape[1]
    if nv not in (3, 4):
        raise ValueError("cells must be triangles or tetrahedra")
    ctype = VTK_TRIANGLE if nv == 3 else VTK_TETRA
    n, T = coords.shape[0], cells.shape[0]
    arrays = []  # (xml attributes, bytes)

    def add(name, a, ncomp=None, tag=""):
        a = np.ascontiguousarray(a)
        if a.dtype.byteorder == ">":
            a = a.astype(a.dtype.newbyteorder("<"))
        nc = ncomp if ncomp is not None else (1 if a.ndim == 1 else a.shape[1])
        attrs = f'type="{_VTK_TYPES[a.dtype]}" Name="{name}" NumberOfComponents="{nc}"'
        if tag:
            attrs += f' NumberOfTuples="{a.shape[0] if a.ndim else 1}"'
        arrays.append((attrs, a.tobytes()))
        return len(arrays) - 1

    i_fd = [add(k, np.array([float(v)]), 1, tag="fd") for k, v in (field_data or {}).items()]
    i_pts = add("Points", coords, 3)
    i_conn = add("connectivity", cells.ravel(), 1)
    i_off = add("offsets", np.arange(1, T + 1, dtype=np.int64) * nv, 1)
    i_typ = add("types", np.full(T, ctype, dtype=np.uint8), 1)
    i_pd = [add(k, np.asarray(v, dtype=np.float64)) for k, v in (point_data or {}).items()]
    offsets, pos = [], 0
    for _, b in arrays:
        offsets.append(pos)
        pos += 8 + len(b)

    def da(i):
        return (f'<DataArray {arrays[i][0]} format="appended" '
                f'offset="{offsets[i]}"/>')

    xml = ['<?xml version="1.0"?>',
           '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
           'header_type="UInt64">', "<UnstructuredGrid>"]
    if i_fd:
        xml.append("<FieldData>")
        xml += [da(i) for i in i_fd]
        xml.append("</FieldData>")
    xml.append(f'<Piece NumberOfPoints="{n}" NumberOfCells="{T}">')
    xml += ["<Points>", da(i_pts), "</Points>", "<Cells>", da(i_conn), da(i_off), da(i_typ),
            "</Cells>"]
    if i_pd:
        xml.append("<PointData>")
        xml += [da(i) for i in i_pd]
        xml.append("</PointData>")
    xml += ["</Piece>", "</UnstructuredGrid>", '<AppendedData encoding="raw">']
    with open(file_name, "wb") as f:
        f.write(("\n".join(xml) + "\n_").encode())
        for _, b in arrays:
            f.write(struct.pack("<Q", len(b)))
            f.write(b)
        f.write(b"\n</AppendedData>\n</VTKFile>\n")


def _attrs(tag: str) -> Dict[str, str]:
    return dict(re.findall(r'(\w+)="([^"]*)"', tag))


def read_vtu(file_name: str):
    """(coords n x 3, cells, point_data, field_data) of a .vtu with ASCII or
    raw appended (uncompressed) data arrays of triangles or tetrahedra."""
    with open(file_name, "rb") as f:
        raw = f.read()
    head_end = raw.find(b"<AppendedData")
    head = raw[:head_end if head_end >= 0 else len(raw)].decode("utf-8", "replace")
    blob = b""
    if head_end >= 0:
        start = raw.index(b"_", head_end) + 1
        blob = raw[start:]
    vtk = _attrs(re.search(r"<VTKFile[^>]*>", head).group(0))
    if vtk.get("byte_order", "LittleEndian") != "LittleEndian":
        raise ValueError("only little-endian .vtu files are read")
    if "compressor" in vtk:
        raise ValueError("compressed .vtu files are not read")
    hdr = "<Q" if vtk.get("header_type", "UInt32") == "UInt64" else "<I"
    hsize = struct.calcsize(hdr)

    def array(m):
        a = _attrs(m.group(1))
        dt = np.dtype(_NP_TYPES[a["type"]]).newbyteorder("<")
        nc = int(a.get("NumberOfComponents", "1"))
        fmt = a.get("format", "ascii")
        if fmt == "appended":
            off = int(a["offset"])
            nbytes = struct.unpack(hdr, blob[off:off + hsize])[0]
            v = np.frombuffer(blob[off + hsize:off + hsize + nbytes], dtype=dt).copy()
        elif fmt == "ascii":
            v = np.array(m.group(2).split(), dtype=dt)
        else:
            raise ValueError(f"DataArray format {fmt!r} is not read")
        return a["Name"] if "Name" in a else "", (v.reshape(-1, nc) if nc > 1 else v)

    pat = re.compile(r"<DataArray([^>]*?)(?:/>|>(.*?)</DataArray>)", re.S)

    def section(name):
        m = re.search(rf"<{name}>(.*?)</{name}>", head, re.S)
        return dict(array(x) for x in pat.finditer(m.group(1))) if m else {}

    fd = {k: float(v.ravel()[0]) for k, v in section("FieldData").items()}
    pts = next(iter(section("Points").values()))
    cells_sec = section("Cells")
    conn, offs, types = cells_sec["connectivity"], cells_sec["offsets"], cells_sec["types"]
    if np.all(types == VTK_TRIANGLE):
        nv = 3
    elif np.all(types == VTK_TETRA):
        nv = 4
    else:
        raise ValueError("only all-triangle or all-tetrahedron meshes are read")
    if not np.array_equal(offs, np.arange(1, len(types) + 1) * nv):
        raise ValueError("inconsistent cell offsets")
    cells = conn.astype(np.int32).reshape(-1, nv)
    return pts.astype(np.float64), cells, section("PointData"), fd


def read_vtk_legacy(file_name: str):
    """(coords n x 3, cells, point_data) of a legacy ASCII .vtk
    UNSTRUCTURED_GRID of triangles or tetrahedra; POINT_DATA SCALARS (any
    component count) and VECTORS arrays are returned by name."""
    with open(file_name) as f:
        tok = f.read().split()
    i = 0

    def find(word):
        nonlocal i
        while tok[i].upper() != word:
            i += 1
        i += 1

    if "ASCII" not in (t.upper() for t in tok[:40]):
        raise ValueError("only ASCII legacy .vtk files are read")
    find("POINTS")
    n = int(tok[i])
    i += 2
    pts = np.array(tok[i:i + 3 * n], dtype=np.float64).reshape(n, 3)
    i += 3 * n
    find("CELLS")
    T, size = int(tok[i]), int(tok[i + 1])
    i += 2
    vals = np.array(tok[i:i + size], dtype=np.int64)
    i += size
    find("CELL_TYPES")
    types = np.array(tok[i + 1:i + 1 + T], dtype=np.int64)
    i += 1 + T
    nv = 3 if np.all(types == VTK_TRIANGLE) else 4 if np.all(types == VTK_TETRA) else 0
    if nv == 0:
        raise ValueError("only all-triangle or all-tetrahedron meshes are read")
    rows = vals.reshape(T, nv + 1)
    if not np.all(rows[:, 0] == nv):
        raise ValueError("inconsistent cell sizes")
    cells = rows[:, 1:].astype(np.int32)
    pd = {}
    while i < len(tok):
        w = tok[i].upper()
        if w == "SCALARS":
            # SCALARS name type [components] [LOOKUP_TABLE table] values
            name = tok[i + 1]
            j, nc = i + 3, 1
            if tok[j].isdigit():
                nc = int(tok[j])
                j += 1
            if tok[j].upper() == "LOOKUP_TABLE":
                j += 2
            v = np.array(tok[j:j + n * nc], dtype=np.float64)
            pd[name] = v.reshape(n, nc) if nc > 1 else v
            i = j + n * nc
        elif w == "VECTORS":
            name = tok[i + 1]
            i += 3
            pd[name] = np.array(tok[i:i + 3 * n], dtype=np.float64).reshape(n, 3)
            i += 3 * n
        else:
            i += 1
    return pts, cells, pd


# ------------------------------------------------------------- Exodus II
# Exodus II meshes (the reference's test meshes are .e files, read through
# MOAB, src/mesh_reader.cpp:19-162) in the netCDF classic formats (CDF-1 and
# the 64-bit-offset CDF-2), through scipy.io.netcdf_file. netCDF-4 (HDF5)
# Exodus files need HDF5, which is not available. Vertices keep their file
# order and cells the order of their element blocks, block by block, the
# order MOAB gives handles in, so nosh_amd.Mesh numbers the edges as the
# reference does (first appearance over the cells in that order).
_EXO_TRI = ("TRI", "TRI3", "TRIANGLE", "TRIANGLE3")
_EXO_TET = ("TETRA", "TETRA4", "TET", "TET4")
_EXO_COMPONENTS = ("_X", "_Y", "_Z")


def _exo_text(a) -> str:
    """A netCDF char array (or attribute bytes) as a stripped str."""
    if isinstance(a, (bytes, bytearray)):
        b = bytes(a)
    else:
        b = b"".join(np.asarray(a).ravel().tolist())
    return b.split(b"\0", 1)[0].decode("ascii", "replace").strip()


def read_exodus(file_name: str, step: int = -1):
    """(coords n x 3, cells T x 3 or T x 4, point_data) of an Exodus II file.

    point_data holds the nodal variables at time step `step` (default: the
    last); variables named <base>_X, <base>_Y[, <base>_Z] become one n x 2 or
    n x 3 array <base> (so "A_X", "A_Y", "A_Z" give the magnetic potential
    "A"), <base>_real / <base>_imag one n x 2 array. Element blocks must all
    be linear triangles or all linear tetrahedra; connectivity is 1-based in
    the file and 0-based here."""
    from scipy.io import netcdf_file
    with netcdf_file(file_name, "r", mmap=False) as f:
        dims, var = f.dimensions, f.variables
        nn = int(dims["num_nodes"])
        nd = int(dims["num_dim"])
        coords = np.zeros((nn, 3))
        if "coord" in var:
            coords[:, :nd] = np.array(var["coord"][:], dtype=np.float64).T
        else:
            for d, nm in enumerate(("coordx", "coordy", "coordz")[:nd]):
                coords[:, d] = np.array(var[nm][:], dtype=np.float64)
        nblk = int(dims.get("num_el_blk", 0) or 0)
        status = np.array(var["eb_status"][:]) if "eb_status" in var else np.ones(nblk, int)
        blocks, kind = [], None
        for b in range(1, nblk + 1):
            if status[b - 1] == 0 or f"connect{b}" not in var:
                continue
            v = var[f"connect{b}"]
            et = _exo_text(getattr(v, "elem_type", b"")).upper()
            k = "tri" if et in _EXO_TRI else "tet" if et in _EXO_TET else None
            if k is None:
                raise ValueError(f"{file_name}: element block {b} has type {et!r}; only "
                                 f"linear triangles (TRI3) and tetrahedra (TETRA4) are supported")
            if kind is not None and k != kind:
                raise ValueError(f"{file_name}: mixed triangle and tetrahedron blocks")
            kind = k
            blocks.append(np.array(v[:], dtype=np.int64) - 1)
        if not blocks:
            raise ValueError(f"{file_name}: no element blocks")
        cells = np.ascontiguousarray(np.concatenate(blocks), dtype=np.int32)
        if cells.min() < 0 or cells.max() >= nn:
            raise ValueError(f"{file_name}: connectivity outside the {nn} nodes")
        names = []
        if "name_nod_var" in var:
            names = [_exo_text(row) for row in np.asarray(var["name_nod_var"][:])]
        raw = {}
        for i, nm in enumerate(names):
            if f"vals_nod_var{i + 1}" in var:
                raw[nm] = np.array(var[f"vals_nod_var{i + 1}"][step], dtype=np.float64)
            elif "vals_nod_var" in var:
                raw[nm] = np.array(var["vals_nod_var"][step, i], dtype=np.float64)
    pd: Dict[str, np.ndarray] = {}
    used = set()
    for nm in names:
        up = nm.upper()
        for suf in (_EXO_COMPONENTS[0], "_REAL"):
            if up.endswith(suf):
                base = nm[:-len(suf)]
                parts = ([base + c for c in ("_X", "_Y", "_Z")] if suf == "_X" else
                         [base + "_real", base + "_imag"])
                found = [next((m for m in names if m.upper() == q.upper()), None) for q in parts]
                if suf == "_X" and found[2] is None:
                    found = found[:2]
                if all(m is not None for m in found) and len(found) >= 2:
                    pd[base] = np.stack([raw[m] for m in found], axis=1)
                    used.update(found)
    for nm in names:
        if nm not in used:
            pd[nm] = raw[nm]
    return coords, cells, pd


def write_exodus(file_name: str, coords, cells, point_data: Optional[Dict[str, np.ndarray]] = None,
                 time: float = 0.0, title: str = "nosh_amd") -> None:
    """Writes a one-block, one-time-step Exodus II file (netCDF CDF-2). Point
    data of shape n become one nodal variable, n x 2 or n x 3 arrays the
    variables <name>_X, <name>_Y[, <name>_Z] (read back by read_exodus as one
    array). 2D meshes (coords n x 2, or n x 3 with z = 0, and triangles) are
    written with num_dim 2."""
    from scipy.io import netcdf_file
    coords = np.asarray(coords, dtype=np.float64)
    cells = np.asarray(cells, dtype=np.int64)
    if cells.ndim != 2 or cells.shape[1] not in (3, 4):
        raise ValueError("cells must be T x 3 (triangles) or T x 4 (tetrahedra)")
    nn = coords.shape[0]
    nd = 2 if cells.shape[1] == 3 and (coords.shape[1] == 2 or not np.any(coords[:, 2])) else 3
    names, vals = [], []
    for k, a in (point_data or {}).items():
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 1:
            names.append(k)
            vals.append(a)
        else:
            for j in range(a.shape[1]):
                names.append(k + _EXO_COMPONENTS[j])
                vals.append(a[:, j])
    with netcdf_file(file_name, "w", version=2) as f:
        f.api_version = np.float32(5.0)
        f.version = np.float32(5.0)
        f.floating_point_word_size = np.int32(8)
        f.file_size = np.int32(1)
        f.title = title
        f.createDimension("time_step", None)  # the record dimension comes first
        f.createDimension("len_string", 33)
        f.createDimension("len_line", 81)
        f.createDimension("four", 4)
        f.createDimension("len_name", 33)
        f.createDimension("num_dim", nd)
        f.createDimension("num_nodes", nn)
        f.createDimension("num_elem", cells.shape[0])
        f.createDimension("num_el_blk", 1)
        f.createDimension("num_el_in_blk1", cells.shape[0])
        f.createDimension("num_nod_per_el1", cells.shape[1])
        t = f.createVariable("time_whole", "d", ("time_step",))
        t[0] = time
        f.createVariable("eb_status", "i", ("num_el_blk",))[:] = 1
        prop = f.createVariable("eb_prop1", "i", ("num_el_blk",))
        prop[:] = 1
        prop.name = "ID"
        for d, nm in enumerate(("coordx", "coordy", "coordz")[:nd]):
            f.createVariable(nm, "d", ("num_nodes",))[:] = coords[:, d]
        cn = f.createVariable("coor_names", "c", ("num_dim", "len_name"))
        for d, nm in enumerate("xyz"[:nd]):
            cn[d, :] = np.array(list(nm.ljust(33, "\0")), dtype="S1")
        conn = f.createVariable("connect1", "i", ("num_el_in_blk1", "num_nod_per_el1"))
        conn[:] = (cells + 1).astype(np.int32)
        conn.elem_type = "TETRA4" if cells.shape[1] == 4 else "TRI3"
        if names:
            f.createDimension("num_nod_var", len(names))
            nv = f.createVariable("name_nod_var", "c", ("num_nod_var", "len_name"))
            for i, nm in enumerate(names):
                if len(nm) > 32:
                    raise ValueError(f"variable name {nm!r} longer than 32 characters")
                nv[i, :] = np.array(list(nm.ljust(33, "\0")), dtype="S1")
            for i, a in enumerate(vals):
                v = f.createVariable(f"vals_nod_var{i + 1}", "d", ("time_step", "num_nodes"))
                v[0, :] = a


def read_mesh(file_name: str):
    """(coords, cells, point_data) from a .vtu, legacy .vtk or Exodus II
    (.e, .exo, .ex2, .g; netCDF classic) file."""
    if file_name.endswith(".vtu"):
        c, t, pd, _ = read_vtu(file_name)
        return c, t, pd
    if file_name.endswith(".vtk"):
        return read_vtk_legacy(file_name)
    if file_name.endswith((".e", ".exo", ".ex2", ".g")):
        return read_exodus(file_name)
    raise ValueError(f"unknown mesh format: {file_name} (.vtu, .vtk and Exodus II .e are read; "
                     f"MOAB .h5m needs HDF5, which is not available)")


# ------------------------------------------------------------------ states
def psi_point_data(psi) -> np.ndarray:
    """Interleaved R^{2n} state -> n x 2 (Re, Im) point data."""
    psi = np.asarray(psi, dtype=np.float64)
    return psi.reshape(-1, 2)


class ContinuationDataSaver:
    """One state file per accepted step: <directory>/out%04d.vtu with point
    data "psi" (Re, Im) and field data with the parameter."""

    def __init__(self, coords, cells, directory: str = ".", param_name: str = "mu",
                 prefix: str = "out"):
        self.coords, self.cells = np.asarray(coords), np.asarray(cells)
        self.directory, self.param_name, self.prefix = directory, param_name, prefix
        self.index = 0

    def save_solution(self, psi, param: float, extra: Optional[Dict[str, float]] = None) -> str:
        path = os.path.join(self.directory, f"{self.prefix}{self.index:04d}.vtu")
        fd = {self.param_name: float(param)}
        fd.update(extra or {})
        write_vtu(path, self.coords, self.cells, {"psi": psi_point_data(psi)}, fd)
        self.index += 1
        return path


def read_state(file_name: str):
    """(interleaved psi, field data) of a state file written by the saver."""
    _, _, pd, fd = read_vtu(file_name)
    return np.ascontiguousarray(pd["psi"]).reshape(-1), fd


def save_checkpoint(file_name: str, coords, cells, psi, params: Dict[str, float],
                    step: int = 0, **arrays) -> None:
    """Raw checkpoint (NumPy .npz, no pickled objects)."""
    names = sorted(params)
    np.savez(file_name, coords=np.asarray(coords, dtype=np.float64),
             cells=np.asarray(cells, dtype=np.int32), psi=np.asarray(psi, dtype=np.float64),
             param_names=np.array(names, dtype="U64"),
             param_values=np.array([params[k] for k in names], dtype=np.float64),
             step=np.int64(step), **arrays)


def load_checkpoint(file_name: str):
    """(coords, cells, psi, params, step, extra arrays) of a checkpoint."""
    with np.load(file_name, allow_pickle=False) as z:
        params = {str(k): float(v) for k, v in zip(z["param_names"], z["param_values"])}
        extra = {k: z[k] for k in z.files
                 if k not in ("coords", "cells", "psi", "param_names", "param_values", "step")}
        return z["coords"], z["cells"], z["psi"], params, int(z["step"]), extra

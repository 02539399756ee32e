"""VTK ImageData (.vti) snapshots and a ParaView .pvd time collection.

Output files keep the layout of the reference's writer (glide/io.py:23-389)
so that existing ParaView state files and readers work unchanged; the layout
itself is the VTK XML ImageData format:

* one ``<ImageData>`` piece whose point data are ``Float32`` arrays in
  ``appended`` raw encoding; the appended section starts with ``_`` and holds,
  per array, a little-endian uint32 byte count followed by the values;
* points run x fastest, then y, then z, so a (ny, nx, nz) enthalpy array is
  written as its (nz, ny, nx) transpose;
* image row 0 is the bottom: the grid's row 0 (the top) is written last
  unless ``flip_y=False``;
* an optional ``TimeValue`` in the piece's field data.

Arrays may be torch tensors (on the device or the host) or numpy arrays; a
device tensor is copied to the host once.  The reference's Zarr/xarray writer
has no counterpart: neither library is available here.
"""
from __future__ import annotations

import struct
import xml.etree.ElementTree as ET
from pathlib import Path
from typing import Any, Mapping

import numpy as np

__all__ = ["write_vti", "read_vti", "VTIWriter"]

_XML_DECL = "<?xml version='1.0' encoding='utf-8'?>\n"
_ESCAPES = (("&", "&amp;"), ("<", "&lt;"), (">", "&gt;"), ('"', "&quot;"), ("\n", "&#10;"))


def _q(text) -> str:
    """A quoted XML attribute value."""
    s = str(text)
    for a, b in _ESCAPES:
        s = s.replace(a, b)
    return f'"{s}"'


def _to_host_f32(value) -> np.ndarray:
    if hasattr(value, "detach"):
        value = value.detach().to("cpu").numpy()
    return np.asarray(value, dtype=np.float32)


def _point_order(a: np.ndarray, flip_y: bool) -> np.ndarray:
    """A grid array in VTK point order (x fastest), bottom row first if flipped."""
    if flip_y:
        a = a[::-1]
    return np.ascontiguousarray(np.moveaxis(a, 2, 0) if a.ndim == 3 else a)


class _Arrays:
    """The point-data arrays of one file: name, component count, raw bytes."""

    def __init__(self, data: Mapping[str, Any], flip_y: bool):
        self.scalars, self.vectors = [], []
        for name, value in data.items():
            if isinstance(value, (list, tuple)):
                comps = [_point_order(_to_host_f32(c), flip_y) for c in value]
                self.vectors.append((name, len(comps), np.stack(comps, axis=-1)))
            else:
                self.scalars.append((name, 1, _point_order(_to_host_f32(value), flip_y)))
        if not self.scalars and not self.vectors:
            raise ValueError("write_vti needs at least one field")
        lead = (self.scalars or self.vectors)[0][2]
        self.ndim = lead.ndim - (0 if self.scalars else 1)
        if self.ndim not in (2, 3):
            raise ValueError(f"fields must be (ny, nx) or (ny, nx, nz) arrays, not {lead.shape}")
        if self.ndim == 3 and self.vectors:
            raise NotImplementedError("vector fields are written for 2-D grids only")
        for name, _, a in self.scalars:
            if a.ndim != self.ndim:
                raise ValueError(f"field {name!r} is {a.ndim}-D but the file's first field is {self.ndim}-D")
        # point counts along x, y, z (the arrays are already in point order)
        if self.ndim == 2:
            ny, nx = lead.shape[:2]
            self.npts = (nx, ny, 1)
        else:
            nz, ny, nx = lead.shape
            self.npts = (nx, ny, nz)

    def ordered(self):
        return self.scalars + self.vectors


def write_vti(filename, data, dx, dy=None, dz=None, origin=(0.0, 0.0), time_value=None, flip_y=True):
    """Write ``data`` (name -> (ny, nx) or (ny, nx, nz) array, or for 2-D grids
    name -> list of component arrays) to one .vti file; 3-D output needs dz.
    Returns the path."""
    arrays = _Arrays(data, flip_y)
    if arrays.ndim == 3 and dz is None:
        raise ValueError("3-D fields need the vertical spacing dz")
    dy = dx if dy is None else dy
    ox, oy = float(origin[0]), float(origin[1])
    oz = float(origin[2]) if len(origin) > 2 else 0.0
    nx, ny, nz = arrays.npts
    extent = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    spacing = f"{dx} {dy} {dz if arrays.ndim == 3 else 1.0}"

    parts = [_XML_DECL, '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian">',
             f"<ImageData WholeExtent={_q(extent)} Origin={_q(f'{ox} {oy} {oz}')} Spacing={_q(spacing)}>",
             f"<Piece Extent={_q(extent)}>"]
    if time_value is not None:
        parts.append('<FieldData><DataArray type="Float32" Name="TimeValue" NumberOfComponents="1" '
                     f'format="ascii">\n{float(time_value)}\n</DataArray></FieldData>')
    active = ""
    if arrays.scalars:
        active += f" Scalars={_q(arrays.scalars[0][0])}"
    three = [name for name, nc, _ in arrays.vectors if nc == 3]
    if three:
        active += f" Vectors={_q(three[0])}"
    parts.append(f"<PointData{active}>")
    blobs, offset = [], 0
    for name, ncomp, a in arrays.ordered():
        raw = a.tobytes()
        parts.append(f'<DataArray type="Float32" Name={_q(name)} NumberOfComponents="{ncomp}" format="appended" '
                     f'offset="{offset}" />')
        blobs.append(struct.pack("<I", len(raw)) + raw)
        offset += 4 + len(raw)
    parts.append("</PointData></Piece></ImageData>")

    path = Path(filename)
    path.parent.mkdir(parents=True, exist_ok=True)
    with open(path, "wb") as f:
        f.write("".join(parts).encode("utf-8"))
        f.write(b'  <AppendedData encoding="raw">\n   _')
        for b in blobs:
            f.write(b)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")
    return path


def read_vti(filename) -> dict[str, Any]:
    """Read a file in the layout above: ``{"extent", "origin", "spacing",
    "time", "fields"}``, each field in point order ((ny, nx) or (nz, ny, nx),
    with a trailing component axis for vectors)."""
    blob = Path(filename).read_bytes()
    tag = b'<AppendedData encoding="raw">'
    at = blob.index(tag)
    data0 = blob.index(b"_", at + len(tag)) + 1
    root = ET.fromstring(blob[:at] + b"</VTKFile>")
    image = root.find("ImageData")
    e = [int(t) for t in image.get("WholeExtent").split()]
    nx, ny, nz = e[1] - e[0] + 1, e[3] - e[2] + 1, e[5] - e[4] + 1
    piece = image.find("Piece")
    time = None
    for arr in piece.iterfind("FieldData/DataArray"):
        if arr.get("Name") == "TimeValue":
            time = float(arr.text)
    shape = (ny, nx) if nz == 1 else (nz, ny, nx)
    fields = {}
    for arr in piece.find("PointData").iterfind("DataArray"):
        start = data0 + int(arr.get("offset"))
        (n,) = struct.unpack_from("<I", blob, start)
        vals = np.frombuffer(blob, dtype="<f4", count=n // 4, offset=start + 4)
        nc = int(arr.get("NumberOfComponents"))
        fields[arr.get("Name")] = vals.reshape(shape + ((nc,) if nc > 1 else ()))
    return {"extent": e, "origin": [float(t) for t in image.get("Origin").split()],
            "spacing": [float(t) for t in image.get("Spacing").split()], "time": time, "fields": fields}


def _field_values(value: Any) -> Any:
    """What a configured output entry writes: callables are evaluated now,
    grid fields are collocated to cell centres, arrays pass through."""
    if callable(value) and not hasattr(value, "data"):
        value = value()
    if hasattr(value, "to_cell") and not hasattr(value, "detach"):
        return value.to_cell().data
    if hasattr(value, "data") and not isinstance(value, np.ndarray) and not hasattr(value, "detach"):
        return value.data
    return value


class VTIWriter:
    """Snapshot series of a simulation (interface of glide/io.py:188-389):
    ``static_fields`` go once to ``{base}_static.vti``; every :meth:`append`
    writes ``dynamic_fields`` to ``{base}_{step:04d}.vti``; :meth:`write_pvd`
    writes the ``{base}.pvd`` collection of the steps.  Entries are grid
    fields, arrays, callables returning either, or lists of them (vectors)."""

    def __init__(self, out_dir, static_fields: Mapping[str, Any] | None = None,
                 dynamic_fields: Mapping[str, Any] | None = None, base: str = "output", dx: float = 1.0,
                 dy: float | None = None, dz: float | None = None, origin: tuple = (0.0, 0.0), flip_y: bool = True):
        self.out_dir = Path(out_dir)
        self.out_dir.mkdir(parents=True, exist_ok=True)
        self.static_fields = dict(static_fields or {})
        self.dynamic_fields = dict(dynamic_fields or {})
        self.base = base
        self.dx = dx
        self.dy = dx if dy is None else dy
        self.dz = dz
        self.origin = origin
        self.flip_y = flip_y
        self.records: list[tuple[float, str]] = []
        self._next_step = 0
        self._static_done = False

    def _write(self, path: Path, entries: Mapping[str, Any], time_value) -> Path:
        payload = {k: ([_field_values(c) for c in v] if isinstance(v, (list, tuple)) else _field_values(v))
                   for k, v in entries.items()}
        return write_vti(path, payload, self.dx, dy=self.dy, dz=self.dz, origin=self.origin, time_value=time_value,
                         flip_y=self.flip_y)

    def initialize(self, grid: Any, *, overwrite: bool = False, write_dynamic_initial: bool = False,
                   time: Any | None = None) -> None:
        """Start a series: with ``overwrite`` remove this base's earlier files
        and records; write the static fields; optionally a first snapshot."""
        if overwrite:
            for old in list(self.out_dir.glob(f"{self.base}_*.vti")) + [self.out_dir / f"{self.base}.pvd"]:
                old.unlink(missing_ok=True)
            self.records = []
            self._next_step = 0
            self._static_done = False
        self.write_static(grid, overwrite=overwrite)
        if write_dynamic_initial:
            if time is None:
                raise ValueError("an initial snapshot needs its time")
            self.append(grid, time=time)

    def write_static(self, grid: Any, *, overwrite: bool = False) -> Path | None:
        if not self.static_fields:
            return None
        path = self.out_dir / f"{self.base}_static.vti"
        if overwrite or not self._static_done:
            self._write(path, self.static_fields, None)
            self._static_done = True
        return path

    def append(self, grid: Any, *, time: Any) -> Path:
        if not self.dynamic_fields:
            raise ValueError("this writer has no dynamic fields to append")
        return self.append_dynamic(grid, time=time)

    def append_dynamic(self, grid: Any, *, time: Any) -> Path:
        step = self._next_step
        self._next_step += 1
        payload = {k: ([_field_values(c) for c in v] if isinstance(v, (list, tuple)) else _field_values(v))
                   for k, v in self.dynamic_fields.items()}
        return self.write_step(step, time, payload)

    def write_step(self, step_idx: int, time_value: Any, data: Mapping[str, Any]) -> Path:
        name = f"{self.base}_{step_idx:04d}.vti"
        path = write_vti(self.out_dir / name, data, self.dx, dy=self.dy, dz=self.dz, origin=self.origin,
                         time_value=time_value, flip_y=self.flip_y)
        self.records.append((float(time_value), name))
        return path

    def write_pvd(self, pvd_name: str | None = None) -> Path:
        lines = ['<?xml version="1.0" ?>', '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">',
                 "  <Collection>"]
        lines += [f'    <DataSet timestep={_q(t)} group="" part="0" file={_q(name)}/>' for t, name in self.records]
        lines += ["  </Collection>", "</VTKFile>", ""]
        path = self.out_dir / (pvd_name or f"{self.base}.pvd")
        path.write_text("\n".join(lines))
        return path

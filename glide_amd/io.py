"""VTK ImageData (.vti) output with a PVD time manifest (glide/io.py:23-389).

Files are byte-compatible with the reference's writer: an XML header whose
point-data arrays are ``Float32`` in ``appended`` raw encoding, each block
preceded by its byte count as a little-endian uint32, values ordered x fastest
(then y, then z), rows flipped so that row 0 is the top of the image unless
``flip_y=False``, and an optional ``TimeValue`` field.  Fields are torch
tensors (device or host) or NumPy arrays; device tensors are copied to the host
once per field.  The reference's ``ZarrWriter`` (xarray/zarr) has no
counterpart here: neither library is available in this environment.
"""
from __future__ import annotations

import struct
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Mapping
from xml.dom import minidom

import numpy as np

__all__ = ["write_vti", "read_vti", "VTIWriter"]


def _host_f32(value) -> np.ndarray:
    if hasattr(value, "detach"):
        value = value.detach().to("cpu").numpy()
    return np.asarray(value, dtype=np.float32)


def write_vti(filename, data, dx, dy=None, dz=None, origin=(0.0, 0.0), time_value=None, flip_y=True):
    """Fields to one .vti file (glide/io.py:23-185).

    ``data`` maps names to 2-D ``(ny, nx)`` or 3-D ``(ny, nx, nz)`` arrays, or
    (2-D only) to a list of component arrays for a vector field.  All fields of
    one file share the first field's dimensionality; 3-D output needs ``dz``.
    """
    dy = dx if dy is None else dy
    path = Path(filename)
    path.parent.mkdir(parents=True, exist_ok=True)
    scalars, vectors = {}, {}
    for name, value in data.items():
        if isinstance(value, (list, tuple)):
            comps = [_host_f32(c) for c in value]
            vectors[name] = [np.flip(c, axis=0) if flip_y else c for c in comps]
        else:
            a = _host_f32(value)
            scalars[name] = np.flip(a, axis=0) if flip_y else a
    first = next(iter(scalars.values())) if scalars else next(iter(vectors.values()))[0]
    if first.ndim == 2:
        three_d = False
        ny, nx = first.shape
        nz = 1
    elif first.ndim == 3:
        three_d = True
        ny, nx, nz = first.shape
        if dz is None:
            raise ValueError(f"dz must be provided for 3D output (got 3D field with shape {first.shape})")
        if vectors:
            raise NotImplementedError("Vector fields are only supported in 2D mode.")
    else:
        raise ValueError(f"Field arrays must be 2D or 3D; got shape {first.shape}")
    for name, a in scalars.items():
        if a.ndim != first.ndim:
            raise ValueError(f"Field '{name}' has shape {a.shape}; expected {'3D' if three_d else '2D'} like the "
                             f"first field.")
    o = (float(origin[0]), float(origin[1]), float(origin[2]) if len(origin) > 2 else 0.0)
    extent = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    root = ET.Element("VTKFile", type="ImageData", version="1.0", byte_order="LittleEndian")
    img = ET.SubElement(root, "ImageData", WholeExtent=extent, Origin=f"{o[0]} {o[1]} {o[2]}",
                        Spacing=f"{dx} {dy} {dz if three_d else 1.0}")
    piece = ET.SubElement(img, "Piece", Extent=extent)
    if time_value is not None:
        fd = ET.SubElement(piece, "FieldData")
        tv = ET.SubElement(fd, "DataArray", type="Float32", Name="TimeValue", NumberOfComponents="1",
                           format="ascii")
        tv.text = f"\n{float(time_value)}\n"
    attrs = {}
    if scalars:
        attrs["Scalars"] = next(iter(scalars))
    for name, comps in vectors.items():
        if len(comps) == 3:
            attrs["Vectors"] = name
            break
    pd = ET.SubElement(piece, "PointData", **attrs)
    blocks = []
    offset = 0
    for name, a in scalars.items():
        # ImageData points run x fastest, then y, then z: (ny, nx, nz) -> (nz, ny, nx)
        raw = np.ascontiguousarray(a.transpose(2, 0, 1) if a.ndim == 3 else a).tobytes()
        ET.SubElement(pd, "DataArray", type="Float32", Name=name, NumberOfComponents="1", format="appended",
                      offset=str(offset))
        blocks.append(raw)
        offset += len(raw) + 4
    for name, comps in vectors.items():
        raw = np.ascontiguousarray(np.stack(comps, axis=-1), dtype=np.float32).tobytes()
        ET.SubElement(pd, "DataArray", type="Float32", Name=name, NumberOfComponents=str(len(comps)),
                      format="appended", offset=str(offset))
        blocks.append(raw)
        offset += len(raw) + 4
    xml = ET.tostring(root, encoding="utf-8", xml_declaration=True).decode("utf-8")
    head = xml[: xml.rindex("</VTKFile>")]
    with open(path, "wb") as f:
        f.write(head.encode("utf-8"))
        f.write(b'  <AppendedData encoding="raw">\n   _')
        for raw in blocks:
            f.write(struct.pack("<I", len(raw)))
            f.write(raw)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")
    return path


def read_vti(filename) -> dict[str, Any]:
    """Parse a file written by :func:`write_vti` (or the reference's writer):
    returns ``{"extent", "origin", "spacing", "time", "fields"}`` with each
    field as stored (x fastest; scalars reshaped to (nz, ny, nx) or (ny, nx),
    vectors to (..., ncomp))."""
    blob = Path(filename).read_bytes()
    marker = b'<AppendedData encoding="raw">'
    start = blob.index(marker)
    us = blob.index(b"_", start + len(marker)) + 1
    header = blob[:start].decode("utf-8") + "</VTKFile>"
    root = ET.fromstring(header)
    img = root.find("ImageData")
    ext = [int(v) for v in img.get("WholeExtent").split()]
    nx, ny, nz = ext[1] - ext[0] + 1, ext[3] - ext[2] + 1, ext[5] - ext[4] + 1
    out = {"extent": ext, "origin": [float(v) for v in img.get("Origin").split()],
           "spacing": [float(v) for v in img.get("Spacing").split()], "time": None, "fields": {}}
    piece = img.find("Piece")
    fd = piece.find("FieldData")
    if fd is not None:
        for da in fd.findall("DataArray"):
            if da.get("Name") == "TimeValue":
                out["time"] = float(da.text.strip())
    for da in piece.find("PointData").findall("DataArray"):
        off = us + int(da.get("offset"))
        (nbytes,) = struct.unpack("<I", blob[off:off + 4])
        vals = np.frombuffer(blob[off + 4:off + 4 + nbytes], dtype="<f4")
        ncomp = int(da.get("NumberOfComponents"))
        shape = (ny, nx) if nz == 1 else (nz, ny, nx)
        out["fields"][da.get("Name")] = vals.reshape(shape + ((ncomp,) if ncomp > 1 else ()))
    return out


def _pretty_xml(element) -> str:
    return minidom.parseString(ET.tostring(element)).toprettyxml(indent="  ")


@dataclass
class VTIWriter:
    """Static and time-dependent .vti output with a .pvd manifest
    (glide/io.py:188-389): configure ``static_fields`` / ``dynamic_fields``
    (fields, constants, arrays, callables evaluated at write time, or lists of
    them for vectors, collocated to cells through ``to_cell()``), call
    :meth:`initialize` once and :meth:`append` per snapshot."""

    out_dir: str | Path
    static_fields: Mapping[str, Any] = field(default_factory=dict)
    dynamic_fields: Mapping[str, Any] = field(default_factory=dict)
    base: str = "output"
    dx: float = 1.0
    dy: float | None = None
    dz: float | None = None
    origin: tuple = (0.0, 0.0)
    flip_y: bool = True

    _step_idx: int = field(default=0, init=False, repr=False)
    _static_written: bool = field(default=False, init=False, repr=False)
    records: list = field(default_factory=list, init=False, repr=False)

    def __post_init__(self) -> None:
        self.out_dir = Path(self.out_dir)
        self.out_dir.mkdir(parents=True, exist_ok=True)
        if self.dy is None:
            self.dy = self.dx

    def initialize(self, grid: Any, *, overwrite: bool = False, write_dynamic_initial: bool = False,
                   time: Any | None = None) -> None:
        """Prepare the directory; write the static fields and, on request, a
        first dynamic snapshot."""
        if overwrite:
            for p in self.out_dir.glob(f"{self.base}_*.vti"):
                p.unlink()
            pvd = self.out_dir / f"{self.base}.pvd"
            if pvd.exists():
                pvd.unlink()
            self.records.clear()
            self._step_idx = 0
            self._static_written = False
        if self.static_fields:
            self.write_static(grid, overwrite=overwrite)
        if write_dynamic_initial:
            if time is None:
                raise ValueError("Argument `time` must be provided when write_dynamic_initial=True.")
            self.append(grid, time=time)

    def append(self, grid: Any, *, time: Any) -> Path:
        """One dynamic snapshot at ``time``."""
        if not self.dynamic_fields:
            raise ValueError("No dynamic fields were configured for this writer.")
        return self.append_dynamic(grid, time=time)

    def write_static(self, grid: Any, *, overwrite: bool = False) -> Path | None:
        """The static fields, once, to ``{base}_static.vti``."""
        if not self.static_fields:
            return None
        path = self.out_dir / f"{self.base}_static.vti"
        if self._static_written and not overwrite:
            return path
        write_vti(path, self._payload(self.static_fields), self.dx, dy=self.dy, dz=self.dz, origin=self.origin,
                  time_value=None, flip_y=self.flip_y)
        self._static_written = True
        return path

    def append_dynamic(self, grid: Any, *, time: Any) -> Path:
        path = self.write_step(self._step_idx, time, self._payload(self.dynamic_fields))
        self._step_idx += 1
        return path

    def write_step(self, step_idx: int, time_value: Any, data: Mapping[str, Any]) -> Path:
        """``data`` to ``{base}_{step_idx:04d}.vti``, recorded for the manifest."""
        name = f"{self.base}_{step_idx:04d}.vti"
        write_vti(self.out_dir / name, data, self.dx, dy=self.dy, dz=self.dz, origin=self.origin,
                  time_value=time_value, flip_y=self.flip_y)
        self.records.append((float(time_value), name))
        return self.out_dir / name

    def write_pvd(self, pvd_name: str | None = None) -> Path:
        """The .pvd collection of the written steps."""
        pvd_name = pvd_name or f"{self.base}.pvd"
        root = ET.Element("VTKFile", type="Collection", version="0.1", byte_order="LittleEndian")
        coll = ET.SubElement(root, "Collection")
        for t, name in self.records:
            ET.SubElement(coll, "DataSet", timestep=str(t), group="", part="0", file=str(name))
        path = self.out_dir / pvd_name
        path.write_text(_pretty_xml(root))
        return path

    def _payload(self, fields: Mapping[str, Any]) -> dict[str, Any]:
        return {name: ([self._array(c) for c in v] if isinstance(v, (list, tuple)) else self._array(v))
                for name, v in fields.items()}

    @staticmethod
    def _array(value: Any) -> Any:
        if callable(value) and not hasattr(value, "data"):
            value = value()
        if isinstance(value, np.ndarray) or hasattr(value, "detach"):
            return value
        if hasattr(value, "to_cell"):
            return value.to_cell().data
        if hasattr(value, "data"):
            return value.data
        return value

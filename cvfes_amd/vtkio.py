"""Minimal VTK XML PolyData (``.vtp``) reader and writer.

CVFES reads its wall mesh with ``vtkXMLPolyDataReader`` (``mesh.py:130-173``) and writes per-sample
displacement arrays ``u_000 ...`` (``SolidMesh.SaveDisplacement``, ``mesh.py:709-726``) and stresses
``xx_000 ...`` (``SolidMesh.Save``, ``mesh.py:683-707``) with ``vtkXMLPolyDataWriter``.  VTK is not
installed here, so this module implements the subset those calls need: points, triangle polys,
point and cell data arrays, in the ``ascii`` or uncompressed ``binary`` (base64, UInt64 header)
encodings.  Files written here open in ParaView/VTK.
"""
from __future__ import annotations

import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np

__all__ = ["PolyData", "read_vtp", "write_vtp", "save_displacement", "save_solution"]

_VTK_TYPES = {
    np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
    np.dtype(np.int64): "Int64", np.dtype(np.int32): "Int32", np.dtype(np.uint8): "UInt8",
}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}


class PolyData:
    """Points ``[N,3]``, triangles ``[E,3]`` and named point / cell arrays."""

    def __init__(self, points, polys, point_data=None, cell_data=None):
        self.points = np.asarray(points, np.float64)
        self.polys = np.asarray(polys, np.int64)
        self.point_data = dict(point_data or {})
        self.cell_data = dict(cell_data or {})


def _encode(arr, fmt):
    arr = np.ascontiguousarray(arr)
    if fmt == "ascii":
        if arr.dtype.kind == "f":
            return " ".join(repr(float(x)) for x in arr.ravel())
        return " ".join(str(int(x)) for x in arr.ravel())
    raw = arr.tobytes()
    return base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")


def _data_array(parent, name, arr, fmt, ncomp=None):
    arr = np.asarray(arr)
    dt = arr.dtype if arr.dtype in _VTK_TYPES else np.dtype(np.float64)
    arr = arr.astype(dt, copy=False)
    attrs = {"type": _VTK_TYPES[dt], "format": fmt}
    if name is not None:
        attrs["Name"] = name
    if ncomp is None:
        ncomp = 1 if arr.ndim == 1 else int(np.prod(arr.shape[1:]))
    attrs["NumberOfComponents"] = str(ncomp)
    el = ET.SubElement(parent, "DataArray", attrs)
    el.text = _encode(arr, fmt)
    return el


def write_vtp(path, pd: PolyData, fmt="binary"):
    """Write ``pd`` to ``path`` (``fmt`` = ``binary`` or ``ascii``)."""
    if fmt not in ("binary", "ascii"):
        raise ValueError("fmt must be 'binary' or 'ascii'")
    root = ET.Element("VTKFile", {"type": "PolyData", "version": "1.0", "byte_order": "LittleEndian",
                                  "header_type": "UInt64"})
    poly = ET.SubElement(root, "PolyData")
    n, e = pd.points.shape[0], pd.polys.shape[0]
    piece = ET.SubElement(poly, "Piece", {"NumberOfPoints": str(n), "NumberOfVerts": "0",
                                          "NumberOfLines": "0", "NumberOfStrips": "0",
                                          "NumberOfPolys": str(e)})
    ptd = ET.SubElement(piece, "PointData")
    for name, arr in pd.point_data.items():
        _data_array(ptd, name, arr, fmt)
    cd = ET.SubElement(piece, "CellData")
    for name, arr in pd.cell_data.items():
        _data_array(cd, name, arr, fmt)
    pts = ET.SubElement(piece, "Points")
    _data_array(pts, "Points", pd.points, fmt, 3)
    polys = ET.SubElement(piece, "Polys")
    _data_array(polys, "connectivity", pd.polys.ravel(), fmt, 1)
    _data_array(polys, "offsets", np.arange(1, e + 1, dtype=np.int64) * pd.polys.shape[1], fmt, 1)
    ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")


def _decode(el, header64):
    dt = _NP_TYPES[el.get("type")]
    fmt = el.get("format", "ascii")
    text = (el.text or "").strip()
    ncomp = int(el.get("NumberOfComponents", "1"))
    if fmt == "ascii":
        arr = np.array(text.split(), dtype=dt) if text else np.zeros(0, dt)
    elif fmt == "binary":
        raw = base64.b64decode(text)
        hs = 8 if header64 else 4
        nbytes = struct.unpack("<Q" if header64 else "<I", raw[:hs])[0]
        arr = np.frombuffer(raw[hs:hs + nbytes], dtype=dt).copy()
    else:
        raise ValueError(f"unsupported DataArray format {fmt!r} (appended data is not supported)")
    return arr.reshape(-1, ncomp) if ncomp > 1 else arr


def read_vtp(path) -> PolyData:
    """Read a ``.vtp`` written in ``ascii`` or uncompressed ``binary`` encoding (triangles only)."""
    tree = ET.parse(path)
    root = tree.getroot()
    if root.get("compressor"):
        raise ValueError("compressed VTK files are not supported")
    header64 = root.get("header_type", "UInt32") == "UInt64"
    piece = root.find("PolyData/Piece")
    points = _decode(piece.find("Points/DataArray"), header64)
    conn = offs = None
    for da in piece.findall("Polys/DataArray"):
        if da.get("Name") == "connectivity":
            conn = _decode(da, header64)
        elif da.get("Name") == "offsets":
            offs = _decode(da, header64)
    if conn is None or offs is None:
        polys = np.zeros((0, 3), np.int64)
    else:
        sizes = np.diff(np.concatenate([[0], offs]))
        if np.any(sizes != 3):
            raise ValueError("only triangle polys are supported")
        polys = conn.reshape(-1, 3).astype(np.int64)
    pdata = {da.get("Name"): _decode(da, header64) for da in piece.findall("PointData/DataArray")}
    cdata = {da.get("Name"): _decode(da, header64) for da in piece.findall("CellData/DataArray")}
    return PolyData(points, polys, pdata, cdata)


def _numbered(filename, counter):
    import os
    base, ext = os.path.splitext(filename)
    return f"{base}{counter}{ext}"


def save_displacement(filename, counter, nodes, elements, u, name="u"):
    """``SaveDisplacement``: ``u [nSmp, nNodes, 3]`` as arrays ``<name>_000 ...`` in ``<file><counter>.vtp``."""
    pd = PolyData(nodes, elements)
    for s in range(u.shape[0]):
        pd.point_data[f"{name}_{s:03d}"] = u[s]
    path = _numbered(filename, counter)
    write_vtp(path, pd)
    return path


def save_solution(filename, counter, nodes, elements, u, stress, uname="displacement"):
    """``SolidMesh.Save``: stresses ``[nElm, nSmp, 5]`` as ``xx_000 ...`` plus displacements."""
    pd = PolyData(nodes, elements)
    for i, comp in enumerate(("xx", "yy", "xy", "xz", "yz")):
        for s in range(stress.shape[1]):
            pd.cell_data[f"{comp}_{s:03d}"] = stress[:, s, i]
    for s in range(u.shape[0]):
        pd.point_data[f"{uname}_{s:03d}"] = u[s]
    path = _numbered(filename, counter)
    write_vtp(path, pd)
    return path

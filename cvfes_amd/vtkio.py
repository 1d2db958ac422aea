"""Minimal VTK XML PolyData (``.vtp``) / UnstructuredGrid (``.vtu``) reader and ``.vtp`` writer.

CVFES reads its wall mesh, face files and initial conditions with ``vtkXMLPolyDataReader`` or, for
names not ending in ``vtp``, ``vtkXMLUnstructuredGridReader`` (``mesh.py:130-131,59-63,145-152``), and
writes per-sample displacement arrays ``u_000 ...`` (``SolidMesh.SaveDisplacement``,
``mesh.py:709-726``) and stresses ``xx_000 ...`` (``SolidMesh.Save``, ``mesh.py:683-707``) with
``vtkXMLPolyDataWriter``.  VTK is not installed here, so this module implements what those calls
need from the VTK XML file format: points, triangle cells, point and cell data arrays, stored as

* ``ascii`` text, ``binary`` (base64 inside the DataArray) or ``appended`` (one ``AppendedData``
  section after the XML, ``raw`` bytes or ``base64`` text, each array at its ``offset``);
* uncompressed (a byte-count header before the data) or compressed block by block
  (``compressor="vtkZLibDataCompressor"`` or ``"vtkLZMADataCompressor"``: a header
  ``[n_blocks, block_size, last_block_size, compressed sizes...]`` then the compressed blocks);
* ``UInt32`` or ``UInt64`` headers, little- or big-endian data.

VTK's own writers default to appended, base64-encoded, zlib-compressed data, which is what real CVFES
case files hold.  The writer here produces inline ``binary`` (default), ``ascii``, or ``appended``
raw data, optionally zlib-compressed; its files open in ParaView/VTK.
"""
from __future__ import annotations

import base64
import lzma
import os
import re
import struct
import xml.etree.ElementTree as ET
import zlib

import numpy as np

__all__ = ["PolyData", "read_vtp", "read_vtu", "read_vtk", "write_vtp", "save_displacement",
           "save_solution"]

_VTK_TYPES = {
    np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
    np.dtype(np.int64): "Int64", np.dtype(np.int32): "Int32", np.dtype(np.int16): "Int16",
    np.dtype(np.int8): "Int8", np.dtype(np.uint64): "UInt64", np.dtype(np.uint32): "UInt32",
    np.dtype(np.uint16): "UInt16", np.dtype(np.uint8): "UInt8",
}
_NP_TYPES = {v: k for k, v in _VTK_TYPES.items()}
_NP_TYPES["Float"] = np.dtype(np.float32)     # names of legacy (version 0.1) files
_NP_TYPES["Double"] = np.dtype(np.float64)
_NP_TYPES["Int"] = np.dtype(np.int32)

VTK_TRIANGLE = 5
_COMPRESSORS = ("vtkZLibDataCompressor", "vtkLZMADataCompressor")


class PolyData:
    """Points ``[N,3]``, triangles ``[E,3]`` and named point / cell arrays."""

    def __init__(self, points, polys, point_data=None, cell_data=None):
        self.points = np.asarray(points, np.float64)
        self.polys = np.asarray(polys, np.int64)
        self.point_data = dict(point_data or {})
        self.cell_data = dict(cell_data or {})


# ------------------------------------------------------------------------------------------------
# writing
# ------------------------------------------------------------------------------------------------
def _b64(b: bytes) -> str:
    return base64.b64encode(b).decode("ascii")


def _binary_block(raw: bytes, compress: bool) -> tuple[bytes, bytes]:
    """(header, payload) of one array: the byte count, or the zlib block table and the blocks."""
    if not compress:
        return struct.pack("<Q", len(raw)), raw
    bs = 1 << 15
    blocks = [zlib.compress(raw[i:i + bs]) for i in range(0, len(raw), bs)]
    last = len(raw) - bs * (len(blocks) - 1) if blocks else 0
    hdr = struct.pack(f"<{3 + len(blocks)}Q", len(blocks), bs, 0 if last == bs else last,
                      *[len(b) for b in blocks])
    return hdr, b"".join(blocks)


class _Appended:
    """Raw bytes of the AppendedData section being written."""

    def __init__(self):
        self.parts = []
        self.size = 0

    def add(self, b: bytes) -> int:
        off = self.size
        self.parts.append(b)
        self.size += len(b)
        return off


def _data_array(parent, name, arr, fmt, ncomp=None, compress=False, appended=None):
    arr = np.asarray(arr)
    dt = arr.dtype if arr.dtype in _VTK_TYPES else np.dtype(np.float64)
    arr = np.ascontiguousarray(arr.astype(dt.newbyteorder("<"), copy=False))
    attrs = {"type": _VTK_TYPES[dt], "format": fmt}
    if name is not None:
        attrs["Name"] = name
    if ncomp is None:
        ncomp = 1 if arr.ndim == 1 else int(np.prod(arr.shape[1:]))
    attrs["NumberOfComponents"] = str(ncomp)
    if fmt == "ascii":
        el = ET.SubElement(parent, "DataArray", attrs)
        el.text = (" ".join(repr(float(x)) for x in arr.ravel()) if arr.dtype.kind == "f"
                   else " ".join(str(int(x)) for x in arr.ravel()))
        return el
    hdr, payload = _binary_block(arr.tobytes(), compress)
    if fmt == "appended":
        attrs["offset"] = str(appended.add(hdr + payload))
        return ET.SubElement(parent, "DataArray", attrs)
    el = ET.SubElement(parent, "DataArray", attrs)
    # VTK's inline layout: the header and the data are base64-encoded separately when compressed,
    # together when not
    el.text = _b64(hdr) + _b64(payload) if compress else _b64(hdr + payload)
    return el


def write_vtp(path, pd: PolyData, fmt="binary", compress=False):
    """Write ``pd`` to ``path``.

    ``fmt``: ``binary`` (base64 inside each DataArray), ``appended`` (raw bytes after the XML) or
    ``ascii``; ``compress`` zlib-compresses binary and appended data (VTK's default compressor).
    """
    if fmt not in ("binary", "ascii", "appended"):
        raise ValueError("fmt must be 'binary', 'appended' or 'ascii'")
    if compress and fmt == "ascii":
        raise ValueError("ascii data cannot be compressed")
    attrs = {"type": "PolyData", "version": "1.0", "byte_order": "LittleEndian",
             "header_type": "UInt64"}
    if compress:
        attrs["compressor"] = "vtkZLibDataCompressor"
    root = ET.Element("VTKFile", attrs)
    poly = ET.SubElement(root, "PolyData")
    n, e = pd.points.shape[0], pd.polys.shape[0]
    piece = ET.SubElement(poly, "Piece", {"NumberOfPoints": str(n), "NumberOfVerts": "0",
                                          "NumberOfLines": "0", "NumberOfStrips": "0",
                                          "NumberOfPolys": str(e)})
    app = _Appended() if fmt == "appended" else None
    kw = dict(compress=compress, appended=app)
    ptd = ET.SubElement(piece, "PointData")
    for name, arr in pd.point_data.items():
        _data_array(ptd, name, arr, fmt, **kw)
    cd = ET.SubElement(piece, "CellData")
    for name, arr in pd.cell_data.items():
        _data_array(cd, name, arr, fmt, **kw)
    pts = ET.SubElement(piece, "Points")
    _data_array(pts, "Points", pd.points, fmt, 3, **kw)
    polys = ET.SubElement(piece, "Polys")
    _data_array(polys, "connectivity", pd.polys.ravel(), fmt, 1, **kw)
    _data_array(polys, "offsets", np.arange(1, e + 1, dtype=np.int64) * pd.polys.shape[1], fmt, 1,
                **kw)
    if app is None:
        ET.ElementTree(root).write(path, xml_declaration=True, encoding="utf-8")
        return
    ET.SubElement(root, "AppendedData", {"encoding": "raw"}).text = "@@APPENDED@@"
    text = ET.tostring(root, encoding="unicode")
    head, tail = text.split("@@APPENDED@@")
    with open(path, "wb") as f:
        f.write(b'<?xml version="1.0" encoding="utf-8"?>\n')
        f.write(head.encode("utf-8") + b"_")
        for part in app.parts:
            f.write(part)
        f.write(tail.encode("utf-8"))


# ------------------------------------------------------------------------------------------------
# reading
# ------------------------------------------------------------------------------------------------
class _FileInfo:
    """Encoding attributes of one VTK XML file and its AppendedData section."""

    def __init__(self, root, appended: bytes | None, app_encoding: str | None):
        self.header = np.dtype("<u8" if root.get("header_type", "UInt32") == "UInt64" else "<u4")
        self.big = root.get("byte_order", "LittleEndian") == "BigEndian"
        if self.big:
            self.header = self.header.newbyteorder(">")
        self.compressor = root.get("compressor") or None
        if self.compressor is not None and self.compressor not in _COMPRESSORS:
            raise ValueError(f"unsupported VTK compressor {self.compressor!r} "
                             f"(supported: {', '.join(_COMPRESSORS)})")
        self.appended = appended
        self.app_encoding = app_encoding

    def decompress(self, block: bytes) -> bytes:
        if self.compressor == "vtkZLibDataCompressor":
            return zlib.decompress(block)
        return lzma.decompress(block)


def _b64_chars(nbytes: int) -> int:
    return -(-nbytes // 3) * 4


def _payload_from_raw(buf: bytes, pos: int, info: _FileInfo) -> bytes:
    """Array bytes stored at ``buf[pos:]`` as header + (compressed) data."""
    hs = info.header.itemsize
    if info.compressor is None:
        nbytes = int(np.frombuffer(buf, info.header, 1, pos)[0])
        return buf[pos + hs:pos + hs + nbytes]
    nb = int(np.frombuffer(buf, info.header, 1, pos)[0])
    hdr = np.frombuffer(buf, info.header, 3 + nb, pos).astype(np.int64)
    out, q = [], pos + hs * (3 + nb)
    for c in hdr[3:]:
        out.append(info.decompress(buf[q:q + int(c)]))
        q += int(c)
    return b"".join(out)


def _payload_from_b64(text: str, info: _FileInfo) -> bytes:
    """Array bytes stored as base64 ``text`` (inline binary, or appended base64 from the offset)."""
    hs = info.header.itemsize
    if info.compressor is None:
        hc = _b64_chars(hs)
        if len(text) >= hc and text[hc - 1] == "=":
            # header encoded on its own, then the data
            nbytes = int(np.frombuffer(base64.b64decode(text[:hc]), info.header, 1)[0])
            return base64.b64decode(text[hc:hc + _b64_chars(nbytes)])[:nbytes]
        first = base64.b64decode(text[:_b64_chars(hs)])
        nbytes = int(np.frombuffer(first[:hs], info.header, 1)[0])
        return base64.b64decode(text[:_b64_chars(hs + nbytes)])[hs:hs + nbytes]
    nb = int(np.frombuffer(base64.b64decode(text[:_b64_chars(hs)])[:hs], info.header, 1)[0])
    hc = _b64_chars(hs * (3 + nb))
    hdr = np.frombuffer(base64.b64decode(text[:hc])[:hs * (3 + nb)], info.header).astype(np.int64)
    total = int(hdr[3:].sum())
    data = base64.b64decode(text[hc:hc + _b64_chars(total)])[:total]
    out, q = [], 0
    for c in hdr[3:]:
        out.append(info.decompress(data[q:q + int(c)]))
        q += int(c)
    return b"".join(out)


def _decode(el, info: _FileInfo):
    tname = el.get("type")
    if tname not in _NP_TYPES:
        raise ValueError(f"unsupported DataArray type {tname!r}")
    dt = _NP_TYPES[tname].newbyteorder(">" if info.big else "<")
    fmt = el.get("format", "ascii")
    ncomp = int(el.get("NumberOfComponents", "1"))
    if fmt == "ascii":
        text = (el.text or "").strip()
        arr = np.array(text.split(), dtype=_NP_TYPES[tname]) if text else np.zeros(0, dt)
    else:
        if fmt == "binary":
            raw = _payload_from_b64("".join((el.text or "").split()), info)
        elif fmt == "appended":
            if info.appended is None:
                raise ValueError("appended DataArray without an AppendedData section")
            off = int(el.get("offset", "0"))
            if info.app_encoding == "base64":
                raw = _payload_from_b64(info.appended[off:].decode("ascii"), info)
            else:
                raw = _payload_from_raw(info.appended, off, info)
        else:
            raise ValueError(f"unsupported DataArray format {fmt!r}")
        if len(raw) % dt.itemsize:
            raise ValueError(f"DataArray {el.get('Name')!r}: {len(raw)} bytes is not a whole "
                             f"number of {tname} values")
        arr = np.frombuffer(raw, dtype=dt).astype(dt.newbyteorder("="))
    return arr.reshape(-1, ncomp) if ncomp > 1 else arr


def _parse(path):
    """XML tree of a VTK file and its encoding information (the AppendedData bytes kept apart,
    since raw appended data is not XML)."""
    with open(path, "rb") as f:
        buf = f.read()
    i = buf.find(b"<AppendedData")
    appended = enc = None
    if i >= 0:
        j = buf.index(b">", i)
        m = re.search(rb'encoding\s*=\s*"([^"]*)"', buf[i:j])
        enc = m.group(1).decode("ascii") if m else "raw"
        if enc not in ("raw", "base64"):
            raise ValueError(f"unsupported AppendedData encoding {enc!r}")
        k = buf.index(b"_", j)
        end = buf.rfind(b"</AppendedData>")
        if end < k:
            raise ValueError("unterminated AppendedData section")
        appended = buf[k + 1:end]
        if enc == "base64":
            appended = b"".join(appended.split())
        tail = buf[end + len(b"</AppendedData>"):]
        buf = buf[:i] + tail
    root = ET.fromstring(buf)
    if root.tag != "VTKFile":
        raise ValueError(f"{path}: not a VTK XML file")
    return root, _FileInfo(root, appended, enc)


def _arrays(piece, tag, info):
    return {da.get("Name"): _decode(da, info) for da in piece.findall(f"{tag}/DataArray")}


def _triangles(conn, offs, what):
    if conn is None or offs is None or offs.size == 0:
        return np.zeros((0, 3), np.int64)
    sizes = np.diff(np.concatenate([[0], offs.astype(np.int64)]))
    if np.any(sizes != 3):
        raise ValueError(f"only triangle {what} are supported (membrane wall meshes)")
    return conn.reshape(-1, 3).astype(np.int64)


def read_vtp(path) -> PolyData:
    """Read a PolyData ``.vtp`` file (triangle polys) in any of the encodings of the module."""
    root, info = _parse(path)
    if root.get("type") != "PolyData":
        raise ValueError(f"{path}: VTK file of type {root.get('type')!r}, expected PolyData")
    piece = root.find("PolyData/Piece")
    points = _decode(piece.find("Points/DataArray"), info)
    polys = _arrays(piece, "Polys", info)
    tris = _triangles(polys.get("connectivity"), polys.get("offsets"), "polys")
    return PolyData(points, tris, _arrays(piece, "PointData", info), _arrays(piece, "CellData", info))


def read_vtu(path) -> PolyData:
    """Read an UnstructuredGrid ``.vtu`` file whose cells are all triangles."""
    root, info = _parse(path)
    if root.get("type") != "UnstructuredGrid":
        raise ValueError(f"{path}: VTK file of type {root.get('type')!r}, expected UnstructuredGrid")
    piece = root.find("UnstructuredGrid/Piece")
    points = _decode(piece.find("Points/DataArray"), info)
    cells = _arrays(piece, "Cells", info)
    types = cells.get("types")
    if types is not None and types.size and np.any(types.astype(np.int64) != VTK_TRIANGLE):
        raise ValueError("only triangle cells (VTK type 5) are supported (membrane wall meshes)")
    tris = _triangles(cells.get("connectivity"), cells.get("offsets"), "cells")
    return PolyData(points, tris, _arrays(piece, "PointData", info), _arrays(piece, "CellData", info))


def read_vtk(path) -> PolyData:
    """``.vtp`` files through :func:`read_vtp`, anything else through :func:`read_vtu`, as the
    reference picks ``vtkXMLPolyDataReader`` or ``vtkXMLUnstructuredGridReader`` by the name
    (``mesh.py:130``)."""
    return read_vtp(path) if str(path).endswith("vtp") else read_vtu(path)


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

"""VTK XML reading in every encoding VTK writes (CPU), and a CVFES case read from files.

The files are built byte by byte by ``tests/vtkfiles.py`` from the file-format description, not by
the writer under test.  The reference reads meshes, faces and initial conditions with
``vtkXMLPolyDataReader`` / ``vtkXMLUnstructuredGridReader`` (``mesh.py:59-63,130-131,145-152``) and
the property fields from ``.npy`` files with the rho suffix (``mesh.py:620-655``).
"""
import numpy as np
import pytest

from cvfes_amd import vtkio
from cvfes_amd.mesh import tube_geometry

import vtkfiles as vf

ENCODINGS = [
    dict(mode="ascii"),
    dict(mode="binary", header="UInt64", separate=False),
    dict(mode="binary", header="UInt32", separate=True),
    dict(mode="binary", header="UInt64", compressor="vtkZLibDataCompressor"),
    dict(mode="appended-raw", header="UInt64"),
    dict(mode="appended-raw", header="UInt32", compressor="vtkZLibDataCompressor"),
    dict(mode="appended-base64", header="UInt32", compressor="vtkZLibDataCompressor"),
    dict(mode="appended-base64", header="UInt64", separate=True),
    dict(mode="appended-base64", header="UInt64", separate=False),
    dict(mode="appended-raw", header="UInt64", compressor="vtkLZMADataCompressor"),
    dict(mode="binary", header="UInt32", big=True),
    dict(mode="appended-raw", header="UInt64", big=True, compressor="vtkZLibDataCompressor"),
]


def _mesh_data(seed=0):
    nodes, el, _ = tube_geometry(9, 5, 0.3, 1.0)
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((nodes.shape[0], 3))
    gid = (100 + rng.permutation(nodes.shape[0])).astype(np.int32)
    cd = np.arange(len(el), dtype=np.float32) * 0.5
    return nodes, el, u, gid, cd


@pytest.mark.parametrize("enc", ENCODINGS, ids=lambda e: "-".join(str(v) for v in e.values()))
def test_read_vtp_every_encoding(tmp_path, enc):
    nodes, el, u, gid, cd = _mesh_data()
    # small blocks so that compressed arrays span several blocks with a partial last one
    path = vf.polydata_file(str(tmp_path / "m.vtp"), nodes, el,
                            point_data=[("u_000", u, 3), ("GlobalNodeID", gid, 1)],
                            cell_data=[("xx_000", cd, 1)], block=200, **enc)
    back = vtkio.read_vtp(path)
    assert np.array_equal(back.points, nodes)
    assert np.array_equal(back.polys, el)
    assert np.array_equal(back.point_data["u_000"], u)
    assert np.array_equal(back.point_data["GlobalNodeID"], gid)
    assert back.point_data["GlobalNodeID"].dtype == np.int32
    assert np.array_equal(back.cell_data["xx_000"], cd)


@pytest.mark.parametrize("enc", ENCODINGS[:7], ids=lambda e: "-".join(str(v) for v in e.values()))
def test_read_vtu_triangles(tmp_path, enc):
    nodes, el, u, gid, _ = _mesh_data(1)
    path = vf.unstructured_file(str(tmp_path / "m.vtu"), nodes, el,
                                point_data=[("velocity", u, 3), ("GlobalNodeID", gid, 1)], **enc)
    back = vtkio.read_vtk(path)
    assert np.array_equal(back.points, nodes) and np.array_equal(back.polys, el)
    assert np.array_equal(back.point_data["velocity"], u)
    assert np.array_equal(back.point_data["GlobalNodeID"], gid)


def test_non_triangle_cells_are_refused(tmp_path):
    nodes, el, _, _, _ = _mesh_data()
    quads = np.stack([el[0::2, 0], el[0::2, 1], el[0::2, 2], el[1::2, 2]], axis=1)
    path = vf.unstructured_file(str(tmp_path / "q.vtu"), nodes, quads, cell_type=9)
    with pytest.raises(ValueError, match="triangle"):
        vtkio.read_vtu(path)
    with pytest.raises(ValueError, match="expected PolyData"):
        vtkio.read_vtp(path)


def test_unknown_compressor_is_refused(tmp_path):
    nodes, el, _, _, _ = _mesh_data()
    path = vf.polydata_file(str(tmp_path / "m.vtp"), nodes, el, mode="binary")
    text = open(path).read().replace('header_type="UInt64"',
                                     'header_type="UInt64" compressor="vtkLZ4DataCompressor"')
    open(path, "w").write(text)
    with pytest.raises(ValueError, match="compressor"):
        vtkio.read_vtp(path)


@pytest.mark.parametrize("fmt,compress", [("binary", False), ("binary", True), ("appended", False),
                                          ("appended", True), ("ascii", False)])
def test_writer_round_trip(tmp_path, fmt, compress):
    nodes, el, u, gid, cd = _mesh_data(2)
    pd = vtkio.PolyData(nodes, el, {"u_000": u, "GlobalNodeID": gid}, {"xx_000": cd})
    path = str(tmp_path / "w.vtp")
    vtkio.write_vtp(path, pd, fmt, compress=compress)
    back = vtkio.read_vtp(path)
    assert np.array_equal(back.points, nodes) and np.array_equal(back.polys, el)
    assert np.array_equal(back.point_data["u_000"], u)
    assert np.array_equal(back.point_data["GlobalNodeID"], gid)
    assert np.array_equal(back.cell_data["xx_000"], cd)


def test_cvfes_case_from_files(tmp_path):
    """The file route of the driver: wall, faces by GlobalNodeID, property files, initial state."""
    from cvfes_amd.cvfes import CVFES
    cfg, want = vf.write_cvfes_case(str(tmp_path))
    cv = CVFES().ReadInputFile(cfg)
    mesh = cv.meshes["wall"]
    assert np.array_equal(mesh.nodes, want["nodes"])
    assert np.array_equal(mesh.elementNodeIds, want["elements"])
    assert np.array_equal(np.sort(mesh.boundary), want["boundary"])
    assert np.array_equal(mesh.vthickness, want["thickness"])
    assert np.array_equal(mesh.vE, want["E"])
    assert mesh.elmE is None
    assert np.array_equal(mesh.iniU, want["u0"]) and np.array_equal(mesh.iniDu, want["du0"])
    assert mesh.nSmp == 4 and mesh.ndof == 3 * want["nodes"].shape[0]


def test_missing_property_file_is_reported(tmp_path):
    import os

    from cvfes_amd.cvfes import CVFES
    cfg, _ = vf.write_cvfes_case(str(tmp_path))
    os.remove(str(tmp_path / "YoungsModulus7.2.npy"))
    with pytest.raises(FileNotFoundError, match="YoungsModulus7.2.npy"):
        CVFES().ReadInputFile(cfg)

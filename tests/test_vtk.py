"""VTK output (hgdg_vtk_write / HpgemAPIDevice.writeVTK) in the layout of
hpGEM's VTKSpecificTimeWriter at order 1 (kernel/Output/
VTKSpecificTimeWriter_Impl.h, VTKStandardElements.cpp), CPU only.

Pins: the files parse as VTK XML with base64 "binary" arrays (uint32 byte
count, then data); cells, offsets and types are VTK's quad / hexahedron /
tetra; the points are each element's vertices in VTK's vertex order; the
point values are the DG solution there, checked against the oracle's basis
evaluation on random states and, in closed form, for states made only of the
(nodal) vertex functions, whose values at the vertices are the coefficients.
"""
import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import _oracle as orc
from hpgem_amd import Mesh, write_vtk
from hpgem_amd._lib import HgdgError

# VTK vertex k is element vertex VTK_ORDER[shape][k] (hpGEM numbering)
VTK_ORDER = {orc.SQUARE: [0, 1, 3, 2], orc.CUBE: [0, 1, 3, 2, 4, 5, 7, 6], orc.TETRAHEDRON: [0, 1, 2, 3]}
VTK_TYPE = {orc.SQUARE: 9, orc.CUBE: 12, orc.TETRAHEDRON: 10}


def decode(node, dtype):
    text = node.text.strip()
    count = struct.unpack("<I", base64.b64decode(text[:8]))[0]
    data = base64.b64decode(text[8:])
    assert len(data) == count
    return np.frombuffer(data, dtype=dtype)


def read_piece(base):
    master = ET.parse(str(base) + ".pvtu").getroot()
    assert master.get("type") == "PUnstructuredGrid"
    source = master.find("PUnstructuredGrid/Piece").get("Source")
    names = [a.get("Name") for a in master.findall("PUnstructuredGrid/PPointData/PDataArray")]
    root = ET.parse(str(base.parent / source)).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    pts = decode(piece.find("Points/DataArray"), "<f8").reshape(-1, 3)
    cells = {a.get("Name"): a for a in piece.findall("Cells/DataArray")}
    conn = decode(cells["connectivity"], "<u4")
    offs = decode(cells["offsets"], "<u4")
    types = decode(cells["types"], "u1")
    data = {a.get("Name"): decode(a, "<f8") for a in piece.findall("PointData/DataArray")}
    assert int(piece.get("NumberOfPoints")) == len(pts) and int(piece.get("NumberOfCells")) == len(types)
    return names, pts, conn, offs, types, data


def reference_vertex(shape, v):
    if shape == orc.TETRAHEDRON:
        x = np.zeros(3)
        if v:
            x[v - 1] = 1.0
        return x
    dim = orc.DIM[shape]
    return np.array([1.0 if (v >> d) & 1 else -1.0 for d in range(dim)])


CASES = [("square", [3, 2], False, 2), ("cube", [2, 2, 3], False, 3), ("cube", [3, 3, 3], True, 1),
         ("tet", [2, 1, 2], False, None)]


@pytest.mark.parametrize("kind,n,periodic,p", CASES)
def test_vtk_points_cells_and_values(kind, n, periodic, p, tmp_path):
    mesh = Mesh.structured(n, periodic=periodic, simplices=(kind == "tet"))
    shape = mesh.shape
    rng = np.random.default_rng(7)
    orders = rng.integers(1, 5, size=mesh.n_elements) if kind == "tet" else np.full(mesh.n_elements, p)
    n_fields = 2
    nbs = [orc.load().orc_basis_size(shape, int(q)) for q in orders]
    state = rng.standard_normal(n_fields * sum(nbs))
    base = tmp_path / "out"
    write_vtk(base, mesh, orders, state, ["a", "b"])
    names, pts, conn, offs, types, data = read_piece(base)
    nv = len(VTK_ORDER[shape])
    assert names == ["a", "b"] and list(data) == ["a", "b"]
    np.testing.assert_array_equal(types, VTK_TYPE[shape])
    np.testing.assert_array_equal(conn, np.arange(mesh.n_elements * nv))
    np.testing.assert_array_equal(offs, nv * np.arange(1, mesh.n_elements + 1))
    off = 0
    for e in range(mesh.n_elements):
        nb = nbs[e]
        for k, v in enumerate(VTK_ORDER[shape]):
            pt = e * nv + k
            np.testing.assert_array_equal(pts[pt, :mesh.dim], mesh.coords[e, v])
            phi, _ = orc.basis_eval(shape, int(orders[e]), reference_vertex(shape, v))
            for f, name in enumerate(names):
                c = state[off + f * nb: off + (f + 1) * nb]
                assert abs(data[name][pt] - c @ phi) <= 1e-13 * max(1.0, np.abs(c).sum())
        off += n_fields * nb


@pytest.mark.parametrize("kind,n", [("cube", [2, 3, 2]), ("tet", [2, 2, 1])])
def test_vertex_functions_give_their_coefficients(kind, n, tmp_path):
    """hpGEM's vertex functions are nodal (one at their vertex, zero at the
    others) and every other function vanishes at the vertices, so a state of
    vertex coefficients only shows exactly those coefficients."""
    mesh = Mesh.structured(n, periodic=False, simplices=(kind == "tet"))
    shape, p = mesh.shape, 3
    nb = orc.load().orc_basis_size(shape, p)
    nv = len(VTK_ORDER[shape])
    vals = np.array([orc.basis_eval(shape, p, reference_vertex(shape, v))[0] for v in range(nv)])  # [vertex][i]
    vertex_fn = [int(np.flatnonzero(np.abs(vals[v] - 1) < 1e-14)[0]) for v in range(nv)]
    rng = np.random.default_rng(3)
    state = np.zeros(mesh.n_elements * nb)
    coef = rng.standard_normal((mesh.n_elements, nv))
    for e in range(mesh.n_elements):
        state[e * nb + np.array(vertex_fn)] = coef[e]
    base = tmp_path / "v"
    write_vtk(base, mesh, p, state, ["u"])
    _, _, _, _, _, data = read_piece(base)
    expected = coef[:, VTK_ORDER[shape]].ravel()
    np.testing.assert_array_equal(data["u"], expected)


def test_vtk_rejects_bad_input(tmp_path):
    mesh = Mesh.structured([2, 2])
    with pytest.raises(HgdgError):
        write_vtk(tmp_path / "x", mesh, 2, np.zeros(mesh.n_elements * 9), ["a", "b"])  # 1 field, 2 names
    with pytest.raises(HgdgError):
        write_vtk(tmp_path / "missing_dir" / "x", mesh, 2, np.zeros(mesh.n_elements * 9), ["a"])


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["hex_advection", "tet_acoustic"])
def test_write_vtk_from_a_device_context(kind, tmp_path):
    from hpgem_amd import OP_ACOUSTIC_CENTRAL, HpgemAPIDevice
    if kind == "hex_advection":
        mesh = Mesh.structured([3, 3, 3])
        dev = HpgemAPIDevice(mesh, 2, velocity=[1.0, 0.5, 0.25])
        dev.setInitialSolutionModes([1.0], [[1.0, 1.0, 0.0]])
        orders, names = np.full(mesh.n_elements, 2), ["u"]
    else:
        mesh = Mesh.structured([2, 2, 2], periodic=[0, 0, 0], simplices=True)
        orders = np.random.default_rng(2).integers(1, 5, size=mesh.n_elements).astype(np.int32)
        dev = HpgemAPIDevice(mesh, orders, kind=OP_ACOUSTIC_CENTRAL, c_inv=np.ones(mesh.n_elements))
        dev.setInitialSolutionGaussian([0.5, 0.5, 0.5], 0.3, 1.0, field=0)
        names = ["v", "s_x", "s_y", "s_z"]
    base = tmp_path / "dev"
    dev.writeVTK(base)
    state = dev.getTimeIntegrationVector(0)
    got_names, pts, _, _, _, data = read_piece(base)
    assert got_names == names
    nv = len(VTK_ORDER[mesh.shape])
    off = 0
    for e in range(mesh.n_elements):
        nb = orc.load().orc_basis_size(mesh.shape, int(orders[e]))
        for k, v in enumerate(VTK_ORDER[mesh.shape]):
            phi, _ = orc.basis_eval(mesh.shape, int(orders[e]), reference_vertex(mesh.shape, v))
            for f, name in enumerate(names):
                c = state[off + f * nb: off + (f + 1) * nb]
                assert abs(data[name][e * nv + k] - c @ phi) <= 1e-13 * max(1.0, np.abs(c).sum())
        off += len(names) * nb
    dev.close()

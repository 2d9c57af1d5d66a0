"""The .hpgem reader (hgdg_mesh_read) on mesh files the reference's own tests
use (tests/files/*.hpgem, written by its Preprocessor from the structured
descriptions in tests/files/sources/), CPU only.

The files pin three things at once: the reader, the structured generator's
element/node/coordinate order (hgdg_structured_mesh and the oracle's, both
restated from applications/Preprocessor/lib/hpgem.cpp:100-235), and the face
matching (the file lists every face with its elements and local face numbers).
"""
import os

import numpy as np
import pytest

import _oracle as orc
from hpgem_amd import Mesh, _lib
from hpgem_amd._lib import HgdgError

MESHES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes")

# file, cells, periodic, simplices (tests/files/sources/<name>)
STRUCTURED = [
    ("unitCubeN4Per.hpgem", [4, 4, 4], [1, 1, 1], False),
    ("unitCubeN4.hpgem", [4, 4, 4], [0, 0, 0], False),
    ("unitCubeN4Tet.hpgem", [4, 4, 4], [0, 0, 0], True),
    ("2Drectangular1mesh.hpgem", [2, 3], [0, 0], False),
    ("2Drectangular2mesh.hpgem", [3, 2], [0, 0], False),
    ("3Drectangular2mesh.hpgem", [2, 3, 2], [0, 0, 0], False),
    ("3Dtriangular2mesh.hpgem", [2, 3, 2], [0, 0, 0], True),
]


def read(name):
    return Mesh.read(os.path.join(MESHES, name))


def face_set(elements, local):
    out = set()
    for (l, r), (fl, fr) in zip(elements, local):
        out.add(frozenset([(int(l), int(fl)), (int(r), int(fr))]) if r >= 0 else frozenset([(int(l), int(fl))]))
    return out


@pytest.mark.parametrize("name,n,periodic,simplices", STRUCTURED)
def test_file_equals_structured_generator(name, n, periodic, simplices):
    m = read(name)
    gen = Mesh.structured(n, periodic=periodic, simplices=simplices)
    assert m.shape == gen.shape and m.dim == len(n)
    assert m.n_elements == gen.n_elements
    np.testing.assert_array_equal(m.coords, gen.coords)       # bit-exact coordinates
    np.testing.assert_array_equal(m.node_ids, gen.node_ids)   # Preprocessor node numbering
    shape, X, nodes = orc.structured_mesh(n, periodic=periodic, simplices=simplices)
    np.testing.assert_array_equal(m.coords, X)
    np.testing.assert_array_equal(m.node_ids, nodes)


@pytest.mark.parametrize("name,n,periodic,simplices", STRUCTURED)
def test_file_faces_equal_matched_faces(name, n, periodic, simplices):
    """Every face the file lists is a face the oracle finds from shared node
    ids, with the same elements and local face numbers, and nothing else."""
    m = read(name)
    P = orc.Problem(m.shape, m.coords, m.node_ids, 1)
    f = P.faces()
    assert len(f) == len(m.face_elements)
    assert face_set(f[:, :2], f[:, 2:4]) == face_set(m.face_elements, m.face_local)
    assert int(np.sum(m.face_elements[:, 1] < 0)) == int(np.sum(f[:, 1] < 0))


def test_version1_and_version2_files_agree():
    v1, v2 = read("meshD3N2v1.hpgem"), read("meshD3N2v2.hpgem")
    assert (v1.version, v2.version) == (1, 2)
    assert v1.zone_names == ["Main"] and v2.zone_names == ["TestZone"]
    np.testing.assert_array_equal(v1.coords, v2.coords)
    np.testing.assert_array_equal(v1.node_ids, v2.node_ids)
    np.testing.assert_array_equal(v1.face_elements, v2.face_elements)
    np.testing.assert_array_equal(v1.face_local, v2.face_local)
    assert v1.n_elements == 8 and v1.n_nodes == 27 and v1.n_edges == 54 and len(v1.face_elements) == 36


def test_periodic_node_copies_resolve_to_element_coordinates():
    """unitCubeN4Per: 64 nodes (one per periodic class), each element's
    vertices at the copy it names; every element is an axis-aligned cube of
    side 1/4 inside [0,1]^3."""
    m = read("unitCubeN4Per.hpgem")
    assert m.n_nodes == 64 and m.n_elements == 64
    assert m.node_ids.max() == 63
    lo = m.coords.min(axis=1)
    hi = m.coords.max(axis=1)
    np.testing.assert_array_equal(hi - lo, 0.25)
    assert lo.min() == 0.0 and hi.max() == 1.0
    assert np.all(m.face_elements[:, 1] >= 0)  # no boundary on a periodic box


def test_other_shapes_are_read():
    tri = read("unitSquareTrianglesN2.hpgem")
    assert tri.shape == _lib.SHAPE_TRIANGLE and tri.dim == 2 and tri.n_vertices == 3
    line = read("unitLineN4Per.hpgem")
    assert line.shape == _lib.SHAPE_LINE and line.dim == 1 and line.n_elements == 4


def test_malformed_files_are_rejected(tmp_path):
    with pytest.raises(HgdgError, match="cannot open"):
        Mesh.read(tmp_path / "missing.hpgem")
    bad = tmp_path / "bad.hpgem"
    bad.write_text("notamesh 2\n")
    with pytest.raises(HgdgError, match="not an hpGEM mesh"):
        Mesh.read(bad)
    two = tmp_path / "two.hpgem"
    two.write_text("mesh 2\n1 1 3 6 12\n2 1 0\nzones\n1\nMain\n")
    with pytest.raises(HgdgError, match="partitions"):
        Mesh.read(two)
    src = open(os.path.join(MESHES, "2Drectangular2mesh.hpgem")).read()
    cut = tmp_path / "cut.hpgem"
    cut.write_text(src[: len(src) // 2])
    with pytest.raises(HgdgError):
        Mesh.read(cut)
    wrong = tmp_path / "wrong.hpgem"
    wrong.write_text(src.replace("elements", "elementz", 1))
    with pytest.raises(HgdgError, match="elements"):
        Mesh.read(wrong)

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


# Files the reference's Preprocessor wrote (output_impl.h:91-208): reading one
# and writing it back reproduces it byte for byte. meshD3N2v2.hpgem was
# written by another tool (no trailing blank after the partition size on line
# 3) and meshD3N2v1.hpgem is format version 1; both are compared below the
# header.
PREPROCESSOR_FILES = ["2Drectangular1mesh.hpgem", "2Drectangular2mesh.hpgem", "3Drectangular2mesh.hpgem",
                      "3Dtriangular2mesh.hpgem", "unitCubeN4.hpgem", "unitCubeN4Per.hpgem", "unitCubeN4Tet.hpgem",
                      "unitLineN4Per.hpgem", "unitSquareTrianglesN2.hpgem"]


@pytest.mark.parametrize("name", PREPROCESSOR_FILES)
def test_written_file_equals_the_preprocessor_file(name, tmp_path):
    out = tmp_path / name
    Mesh.read(os.path.join(MESHES, name)).write(out)
    assert out.read_bytes() == open(os.path.join(MESHES, name), "rb").read()


def test_other_writer_file_differs_only_in_the_partition_line(tmp_path):
    out = tmp_path / "m.hpgem"
    Mesh.read(os.path.join(MESHES, "meshD3N2v2.hpgem")).write(out)
    ref = open(os.path.join(MESHES, "meshD3N2v2.hpgem"), "rb").read().split(b"\n")
    got = out.read_bytes().split(b"\n")
    assert len(ref) == len(got)
    assert [i for i in range(len(ref)) if ref[i] != got[i]] == [2]
    assert got[2] == ref[2] + b" "


@pytest.mark.parametrize("name,n,periodic,simplices", [
    ("unitCubeN4Per.hpgem", [4, 4, 4], True, False),
    ("unitCubeN4.hpgem", [4, 4, 4], False, False),
    ("unitCubeN4Tet.hpgem", [4, 4, 4], False, True),
])
def test_structured_mesh_written_equals_the_preprocessor_file(name, n, periodic, simplices, tmp_path):
    """Generator and writer together reproduce what the reference's
    Preprocessor writes for the same box: node numbering, periodic coordinate
    copies, element node lists, and the face and edge numbering."""
    out = tmp_path / name
    Mesh.structured(n, periodic=periodic, simplices=simplices).write(out)
    assert out.read_bytes() == open(os.path.join(MESHES, name), "rb").read()


def test_write_read_round_trip_keeps_coordinates_nodes_and_zones(tmp_path):
    base = Mesh.structured([3, 2, 3], periodic=[1, 0, 1])
    rng = np.random.default_rng(4)
    shear = np.array([[1.0, 0.25, 0.0], [0.0, 1.0, -0.125], [0.3, 0.0, 1.0]])
    zones = rng.integers(0, 3, size=base.n_elements).astype(np.int32)
    mesh = Mesh(base.shape, base.coords @ shear.T, base.node_ids, zones, ["Left", "Middle", "Right"])
    out = tmp_path / "sheared.hpgem"
    mesh.write(out)
    back = Mesh.read(out)
    np.testing.assert_array_equal(back.coords, mesh.coords)
    np.testing.assert_array_equal(back.node_ids, mesh.node_ids)
    np.testing.assert_array_equal(back.zones, zones)
    assert back.zone_names == ["Left", "Middle", "Right"]
    out2 = tmp_path / "again.hpgem"
    back.write(out2)
    assert out2.read_bytes() == out.read_bytes()


def test_writer_rejects_inconsistent_input(tmp_path):
    base = Mesh.structured([3, 3])
    gap = Mesh(base.shape, base.coords, base.node_ids + (base.node_ids > 4))  # node 5 never used
    with pytest.raises(HgdgError):
        gap.write(tmp_path / "gap.hpgem")
    bad_zone = Mesh(base.shape, base.coords, base.node_ids, np.full(base.n_elements, 2, np.int32), ["A"])
    with pytest.raises(HgdgError):
        bad_zone.write(tmp_path / "zone.hpgem")

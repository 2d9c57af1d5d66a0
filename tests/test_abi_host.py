"""CPU checks of the C ABI library: it loads, exports every declared entry
point, and its host-side tables agree with the oracle (no GPU calls)."""
import os
import re

import numpy as np
import pytest

import _oracle as orc
import hpgem_amd
from hpgem_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(hgdg_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_entry_points()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    # and the binding covers them all
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


@pytest.mark.parametrize("shape,dim", [(orc.SQUARE, 2), (orc.CUBE, 3)])
@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_tensor_basis_matches_oracle_order_and_values(shape, dim, p):
    rng = np.random.default_rng(p)
    for _ in range(5):
        x = rng.uniform(-1, 1, dim)
        v_lib = hpgem_amd.basis_eval(shape, p, x)
        v_orc, _ = orc.basis_eval(shape, p, x)
        np.testing.assert_allclose(v_lib, v_orc, rtol=0, atol=1e-14)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_tet_basis_matches_oracle_order_and_values(p):
    rng = np.random.default_rng(10 + p)
    for _ in range(6):
        x = rng.dirichlet([1, 1, 1, 1])[:3]
        v_lib = hpgem_amd.basis_eval(orc.TETRAHEDRON, p, x)
        v_orc, _ = orc.basis_eval(orc.TETRAHEDRON, p, x)
        np.testing.assert_allclose(v_lib, v_orc, rtol=0, atol=1e-14)


@pytest.mark.parametrize("n,periodic,simplices", [
    ([4, 3], [1, 1], False), ([3, 4, 5], [1, 0, 1], False), ([2, 3, 2], [0, 0, 0], True),
    ([3, 3, 3], [1, 1, 1], True), ([16, 16], [1, 1], False)])
def test_structured_mesh_matches_oracle(n, periodic, simplices):
    m = hpgem_amd.Mesh.structured(n, periodic=periodic, simplices=simplices)
    shape, X, nodes = orc.structured_mesh(n, periodic=periodic, simplices=simplices)
    assert m.shape == shape
    np.testing.assert_array_equal(m.node_ids, nodes)
    np.testing.assert_allclose(m.coords, X, rtol=0, atol=1e-15)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libhgdg.so"))
    with pytest.raises(ImportError):
        _lib.load()


@pytest.mark.parametrize("p", [1, 2])
def test_hcurl_basis_matches_reference_fixture(p):
    """The library's Ainsworth-Coyle set against the values and curls the
    reference's own BasisFunctions3DAinsworthCoyle.cpp printed."""
    from hpgem_amd.api import hcurl_basis_eval
    t = [x for x in orc.fixture_tables() if x["kind"] == "hcurl_basis" and x["p"] == p][0]
    n = t["n"]
    for pt, vals, curls in zip(t["points"], t["values"], t["curls"]):
        v, c = hcurl_basis_eval(p, pt)
        assert v.shape == (n, 3)
        np.testing.assert_allclose(v.ravel(), vals, rtol=0, atol=1e-14)
        np.testing.assert_allclose(c.ravel(), curls, rtol=0, atol=1e-13)

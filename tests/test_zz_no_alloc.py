"""hgdg_step_error_noalloc (the step error estimate with its partial sums in
the library's fixed device scratch array) against hgdg_step_error, through
the C ABI: two calls in a row on one GPU, the second on a smaller mesh (the
scratch array keeps the first call's contents), a tetrahedral mesh, and the
size limit."""
import ctypes

import numpy as np
import pytest

from hpgem_amd import DORMAND_PRINCE, OP_ACOUSTIC_CENTRAL, HpgemAPIDevice, Mesh, _lib
from hpgem_amd._lib import dptr

pytestmark = pytest.mark.gpu


def both_estimates(dev, dt):
    lib = _lib.load()
    t = dev.tableau
    aux = (ctypes.c_int * t.stages)(*dev.auxiliaryVectorIds_)
    e = np.ascontiguousarray(t.error, dtype=np.float64)
    out = []
    for fn in (lib.hgdg_step_error, lib.hgdg_step_error_noalloc):
        err, nrm = ctypes.c_double(), ctypes.c_double()
        rc = fn(dev.h, dev.solutionVectorId_, aux, t.stages, dptr(e), dt, ctypes.byref(err), ctypes.byref(nrm))
        assert rc == 0, lib.hgdg_last_error().decode()
        out.append((err.value, nrm.value))
    return out


def stepped(mesh, p, dt, precision="fp64", **kw):
    dev = HpgemAPIDevice(mesh, p, tableau=DORMAND_PRINCE, precision=precision, **kw)
    rng = np.random.default_rng(p)
    dev.setTimeIntegrationVector(0, rng.standard_normal(dev.n_dof))
    dev.computeOneTimeStep(0.0, dt, 1)
    return dev


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_step_error_matches_step_error_twice(precision):
    print(f"[case] step_error_noalloc vs step_error, 16^3 then 6^3 hex, {precision}", flush=True)
    for n, p in ((16, 3), (6, 2)):  # second call: smaller mesh over the same scratch array
        dt = 0.02 / n
        dev = stepped(Mesh.structured([n, n, n]), p, dt, precision, velocity=[1.0, 0.7, 0.3])
        (e0, n0), (e1, n1) = both_estimates(dev, dt)
        print(f"  {n}^3 p={p}: error {e0:.17e} / {e1:.17e}, norm {n0:.17e} / {n1:.17e}", flush=True)
        assert e0 > 0 and n0 > 0
        assert abs(e1 - e0) <= 1e-12 * e0
        assert abs(n1 - n0) <= 1e-12 * n0
        dev.close()


def test_noalloc_step_error_on_tetrahedra():
    print("[case] step_error_noalloc vs step_error, 3^3 x 5 hp tets", flush=True)
    mesh = Mesh.structured([3, 3, 3], periodic=[0, 0, 0], simplices=True)
    rng = np.random.default_rng(2)
    orders = rng.integers(1, 5, size=mesh.n_elements).astype(np.int32)
    cinv = rng.uniform(0.3, 3.0, size=mesh.n_elements)
    dev = HpgemAPIDevice(mesh, orders, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=DORMAND_PRINCE)
    dev.setTimeIntegrationVector(0, rng.standard_normal(dev.n_dof))
    dev.computeOneTimeStep(0.0, 1e-3, 1)
    (e0, n0), (e1, n1) = both_estimates(dev, 1e-3)
    assert abs(e1 - e0) <= 1e-12 * e0 and abs(n1 - n0) <= 1e-12 * n0
    dev.close()


def test_noalloc_step_error_rejects_meshes_beyond_its_scratch():
    print("[case] step_error_noalloc size limit, 65^3 hex p=1", flush=True)
    lib = _lib.load()
    dev = HpgemAPIDevice(Mesh.structured([65, 65, 65]), 1, velocity=[1.0, 0.0, 0.0], tableau=DORMAND_PRINCE)
    aux = (ctypes.c_int * dev.tableau.stages)(*dev.auxiliaryVectorIds_)
    e = np.ascontiguousarray(dev.tableau.error, dtype=np.float64)
    err, nrm = ctypes.c_double(), ctypes.c_double()
    rc = lib.hgdg_step_error_noalloc(dev.h, 0, aux, dev.tableau.stages, dptr(e), 1e-3, ctypes.byref(err),
                                     ctypes.byref(nrm))
    assert rc == 3  # HGDG_ERR_UNSUPPORTED
    assert "262144" in lib.hgdg_last_error().decode()
    dev.close()

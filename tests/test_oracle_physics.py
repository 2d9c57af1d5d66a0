"""Pins the oracle's operator layer with closed forms, invariants and the
reference's own stored error tables (no GPU).

* tests/self/Basic/090Advection_SelfTest.cpp:330-335: 3D periodic hex, p=1,
  a = (0.1, 0.2, 0.3), u0 = sin(2 pi x) sin(2 pi y) sin(2 pi z), T = 0.02 in
  2 Dormand-Prince steps: L2 errors 5.55789642e-02, 1.44021875e-02,
  3.81732285e-03 on 4^3, 8^3, 16^3 meshes, 1 % tolerance
  (tests/self/ConvergenceTest.h:56-86).
* constant states are steady (periodic upwind advection);
* the upwind operator dissipates energy, the central acoustic operator
  conserves it exactly (u^T M k = 0);
* affine elements: M_e = |det J| M_ref.
"""
import numpy as np
import pytest

import _oracle as orc

# Dormand-Prince 5(4), kernel/Base/TimeIntegration/DormandPrince.cpp:45-80
DP_A = [[], [1 / 5], [3 / 40, 9 / 40], [44 / 45, -56 / 15, 32 / 9],
        [19372 / 6561, -25360 / 2187, 64448 / 6561, -212 / 729],
        [9017 / 3168, -355 / 33, 46732 / 5247, 49 / 176, -5103 / 18656],
        [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84]]
DP_B = [35 / 384, 0, 500 / 1113, 125 / 192, -2187 / 6784, 11 / 84, 0]

# sin(2pi x) sin(2pi y) sin(2pi z) as a sum of four sine modes
PROD3_AMP = np.array([0.25, 0.25, 0.25, -0.25])
PROD3_K = np.array([[1, 1, -1], [1, -1, 1], [-1, 1, 1], [1, 1, 1]], dtype=float)


def hex_problem(n, p, periodic=True, fields=1):
    shape, X, nodes = orc.structured_mesh([n, n, n], periodic=[int(periodic)] * 3)
    return orc.Problem(shape, X, nodes, p, fields)


@pytest.mark.parametrize("n,expected", [(4, 5.55789642e-02), (8, 1.44021875e-02)])
def test_reference_error_table_3d_periodic_hex_p1(n, expected):
    P = hex_problem(n, 1)
    P.set_advection([0.1, 0.2, 0.3])
    u = P.project_modes(PROD3_AMP, PROD3_K)
    u = P.rk_steps(u, 0.01, 2, DP_A, DP_B)
    err = P.l2_error_modes(u, PROD3_AMP, PROD3_K, 0.02)
    assert abs(err - expected) < 0.01 * expected, (err, expected)


def test_product_modes_match_product_of_sines():
    pts = np.random.default_rng(0).random((20, 3))
    f = lambda x: sum(a * np.sin(2 * np.pi * (k @ x)) for a, k in zip(PROD3_AMP, PROD3_K))
    for x in pts:
        assert abs(f(x) - np.prod(np.sin(2 * np.pi * x))) < 1e-14


@pytest.mark.parametrize("shape_kind,p", [("square", 2), ("cube", 3), ("cube", 1)])
def test_constant_state_is_steady(shape_kind, p):
    n = [4, 4] if shape_kind == "square" else [3, 3, 3]
    shape, X, nodes = orc.structured_mesh(n)
    P = orc.Problem(shape, X, nodes, p)
    P.set_advection([1.0, 0.5, 0.3][: len(n)])
    u = P.project_modes(np.array([0.0]), np.zeros((1, len(n))))
    # constant function 1.7 projected = vertex functions only
    u_const = np.zeros(P.n_dof)
    nb = P.n_dof // P.n_elem
    nv = 2 ** len(n)
    for e in range(P.n_elem):
        u_const[e * nb: e * nb + nv] = 1.7
    k = P.time_derivative([u_const], [1.0])
    # |a| |u| / h sets the scale of the operator; the rest is rounding through
    # the hierarchical mass inverse
    scale = 1.7 * 1.2 * n[0]
    assert np.max(np.abs(k)) < 1e-11 * scale


@pytest.mark.parametrize("p", [1, 2, 3])
def test_upwind_dissipates_energy(p):
    P = hex_problem(3, p)
    P.set_advection([1.0, 0.7, 0.3])
    u = np.random.default_rng(p).standard_normal(P.n_dof)
    k = P.time_derivative([u], [1.0])
    nb = P.n_dof // P.n_elem
    rate = 0.0
    for e in range(P.n_elem):
        M = P.mass_matrix(e)
        ue = u[e * nb:(e + 1) * nb]
        rate += ue @ M @ k[e * nb:(e + 1) * nb]
    assert rate < -1e-8


def test_affine_mass_is_scaled_reference_mass():
    shape, X, nodes = orc.structured_mesh([3, 5, 4])
    P = orc.Problem(shape, X, nodes, 2)
    P.set_advection([1.0, 0.0, 0.0])
    M0 = P.mass_matrix(0)
    # reference mass: same basis on [-1,1]^3 has |det J| = 1
    h = np.array([1 / 3, 1 / 5, 1 / 4])
    detJ = np.prod(h / 2)
    shape2, X2, nodes2 = orc.structured_mesh([1, 1, 1], lo=[-1, -1, -1], hi=[1, 1, 1], periodic=[0, 0, 0])
    R = orc.Problem(shape2, X2, nodes2, 2)
    R.set_advection([1.0, 0.0, 0.0])
    np.testing.assert_allclose(M0, detJ * R.mass_matrix(0), rtol=0, atol=1e-15)
    assert np.all(np.linalg.eigvalsh(M0) > 0)


def test_structured_hex_mesh_layout():
    shape, X, nodes = orc.structured_mesh([3, 4, 3])
    assert shape == orc.CUBE and X.shape == (36, 8, 3)
    # x fastest, vertex order x fastest within the element
    np.testing.assert_allclose(X[1, 0], [1 / 3, 0, 0])
    np.testing.assert_allclose(X[3, 0], [0, 0.25, 0])
    np.testing.assert_allclose(X[2, 1], [1.0, 0, 0])  # periodic copy coordinate
    assert nodes[2, 1] == 0                           # ... of node 0
    P = orc.Problem(shape, X, nodes, 1)
    assert P.n_faces == 3 * 36 and P.n_internal == 3 * 36


def test_structured_tet_mesh_faces_match():
    shape, X, nodes = orc.structured_mesh([2, 2, 2], periodic=[0, 0, 0], simplices=True)
    assert shape == orc.TETRAHEDRON and X.shape == (40, 4, 3)
    P = orc.Problem(shape, X, nodes, 1)
    # every tet face is either shared or on the boundary of the box
    n_bdry = P.n_faces - P.n_internal
    assert n_bdry == 6 * 2 * 2 * 2
    assert 4 * 40 == 2 * P.n_internal + n_bdry
    # all tets have positive volume under the reference's rotation scheme? (signed
    # volumes are allowed; |det J| is used) -- the sum must be the box volume
    vol = 0.0
    for e in range(40):
        J = (X[e, 1:] - X[e, 0]).T
        vol += abs(np.linalg.det(J)) / 6
    assert abs(vol - 1.0) < 1e-14


def test_acoustic_central_flux_conserves_energy():
    shape, X, nodes = orc.structured_mesh([2, 2, 2], periodic=[0, 0, 0], simplices=True)
    rng = np.random.default_rng(3)
    orders = rng.integers(1, 4, size=len(X))
    P = orc.Problem(shape, X, nodes, orders, n_fields=4)
    cinv = rng.uniform(0.5, 2.0, size=len(X))
    P.set_acoustic(cinv)
    u = rng.standard_normal(P.n_dof)
    k = P.time_derivative([u], [1.0])
    scale = 0.0
    # energy rate through the weighted mass: u^T M k = u^T r
    r_rate = 0.0
    off = 0
    for e in range(P.n_elem):
        nb4 = ((orders[e] + 1) * (orders[e] + 2) * (orders[e] + 3) // 6) * 4
        Mw = _weighted_mass(P, e, nb4)
        ue = u[off:off + nb4]
        r_rate += ue @ Mw @ k[off:off + nb4]
        scale += abs(ue) @ abs(Mw @ k[off:off + nb4])
        off += nb4
    assert abs(r_rate) < 1e-12 * scale


def _weighted_mass(P, e, n):
    out = np.zeros(n * n)
    orc.check(P.lib.orc_mass_matrix(P.h, e, orc._dp(out)))
    return out.reshape(n, n, order="F")

"""The oracle's DG-Max interior-penalty Maxwell assembly (CPU), pinned by the
reference's own Ainsworth-Coyle basis values (reference-built fixture) and by
closed-form energies of fields the discrete space holds exactly.

For a field u in the H(curl) space of every element with continuous
tangential traces, the assembled form reduces to
    u^T K u = int |curl u|^2 - 2 int_{boundary} (n x u).curl u
              + sum_{boundary faces} stab/diam int |n x u|^2,
    u^T M u = int eps |u|^2,
which fixes the curl-curl term, the boundary factor, the sign of n x phi on
each side (interior jumps must vanish), the face diameter and the
penalty scaling (DGMaxDiscretization.cpp:193-260, 344-410).
"""
import numpy as np
import pytest

import _oracle as orc

TET = orc.TETRAHEDRON


@pytest.mark.parametrize("p", [1, 2])
def test_oracle_hcurl_basis_matches_reference_fixture(p):
    t = [x for x in orc.fixture_tables() if x["kind"] == "hcurl_basis" and x["p"] == p][0]
    for pt, vals, curls in zip(t["points"], t["values"], t["curls"]):
        v, c = orc.hcurl_eval(p, pt)
        np.testing.assert_allclose(v.ravel(), vals, rtol=0, atol=1e-14)
        np.testing.assert_allclose(c.ravel(), curls, rtol=0, atol=1e-13)


def assemble_global(P, S, F):
    """Block assembly of K = sum_e S_e + sum_faces F (dense, for tiny meshes)."""
    ne, n = S.shape[0], S.shape[1]
    K = np.zeros((ne * n, ne * n))
    for e in range(ne):
        K[e * n:(e + 1) * n, e * n:(e + 1) * n] += S[e]
    faces = P.faces()
    for fi, (L, R, fL, fR, o) in enumerate(faces):
        sides = [L] if R < 0 else [L, R]
        for t, et in enumerate(sides):
            for u, eu in enumerate(sides):
                K[et * n:(et + 1) * n, eu * n:(eu + 1) * n] += F[fi, t * 2 + u]
    return K


def project(P, X, p, field):
    """Element-wise L2 projection of a vector field into the oracle's H(curl)
    space (covariant transform phi = J^{-T} phi_ref), with M_e from unit eps."""
    pts, w, _ = orc.rule(TET, 9)
    S, M, F = P.maxwell(p, np.ones(len(X)), 1.0)
    n = M.shape[1]
    U = np.zeros(len(X) * n)
    for e, Xe in enumerate(X):
        J = np.stack([Xe[1] - Xe[0], Xe[2] - Xe[0], Xe[3] - Xe[0]], axis=1)
        Jit = np.linalg.inv(J).T
        dJ = abs(np.linalg.det(J))
        b = np.zeros(n)
        for q, xi in enumerate(pts):
            v, _ = orc.hcurl_eval(p, xi)
            phys = v @ Jit.T
            x = Xe[0] + J @ xi
            b += w[q] * dJ * phys @ field(x)
        U[e * n:(e + 1) * n] = np.linalg.solve(M[e], b)
    return U


@pytest.mark.parametrize("p,field,curl2,bdry_ct,bdry_tt,mass", [
    # u = (0, 0, x): curl u = (0, -1, 0)
    (1, lambda x: np.array([0.0, 0.0, x[0]]), 1.0, 1.0, 1 + 2 / 3, 1 / 3),
    (2, lambda x: np.array([0.0, 0.0, x[0]]), 1.0, 1.0, 1 + 2 / 3, 1 / 3),
    # u = (0, 0, x^2): curl u = (0, -2x, 0)
    (2, lambda x: np.array([0.0, 0.0, x[0] ** 2]), 4 / 3, 2.0, 1 + 2 / 5, 1 / 5),
])
def test_energy_of_exactly_represented_fields(p, field, curl2, bdry_ct, bdry_tt, mass):
    n_cells = 2
    h = 1.0 / n_cells
    shape, X, nodes = orc.structured_mesh([n_cells] * 3, periodic=[0, 0, 0], simplices=True)
    P = orc.Problem(shape, X, nodes, p)
    stab = 10.0 * p * p
    eps = 2.25
    S, M, F = P.maxwell(p, np.full(len(X), eps), stab)
    K = assemble_global(P, S, F)
    np.testing.assert_allclose(K, K.T, rtol=0, atol=1e-12 * np.abs(K).max())
    U = project(P, X, p, field)
    # boundary faces of the 5-tet split are right triangles with legs h
    expected = curl2 - 2 * bdry_ct + stab / (h * np.sqrt(2)) * bdry_tt
    energy = U @ K @ U
    assert abs(energy - expected) < 1e-10 * abs(expected), (energy, expected)
    Mg = np.zeros_like(K)
    n = M.shape[1]
    for e in range(len(X)):
        Mg[e * n:(e + 1) * n, e * n:(e + 1) * n] = M[e]
    assert abs(U @ Mg @ U - eps * mass) < 1e-11 * mass


@pytest.mark.parametrize("p", [1, 2])
def test_gradients_have_no_curl_energy(p):
    """The degree-1 edge functions are gradients grad(l_o l_i): their columns
    of every element curl-curl matrix vanish; S and M are symmetric, M is
    positive definite and linear in eps."""
    shape, X, nodes = orc.structured_mesh([1, 1, 1], periodic=[0, 0, 0], simplices=True)
    P = orc.Problem(shape, X, nodes, p)
    rng = np.random.default_rng(3)
    eps = rng.choice([1.0, 2.25, 12.0], size=len(X))
    S, M, F = P.maxwell(p, eps, 10.0 * p * p)
    S1, M1, _ = P.maxwell(p, np.ones(len(X)), 10.0 * p * p)
    for e in range(len(X)):
        assert np.abs(S[e][:, 6:12]).max() < 1e-13 * np.abs(S[e]).max()
        np.testing.assert_allclose(S[e], S[e].T, atol=1e-13 * np.abs(S[e]).max())
        np.testing.assert_allclose(M[e], M[e].T, atol=1e-14 * np.abs(M[e]).max())
        assert np.linalg.eigvalsh(M[e]).min() > 0
        np.testing.assert_allclose(M[e], eps[e] * M1[e], rtol=0, atol=1e-14 * eps[e] * np.abs(M1[e]).max())
        np.testing.assert_array_equal(S[e], S1[e])

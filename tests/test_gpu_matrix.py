"""The matrix-operator path (hgdg_operator_set_matrices): HpgemAPILinear's
precomputed path on the device, taking the matrices an application builds
from its own integrands (createMassMatrices / createStiffnessMatrices,
kernel/Base/HpgemAPILinear_Impl.h:131-139,250-288) and running
multiplyStiffnessMatrices (:360-434) and solveMassMatrixEquations (:143-156)
with everything built on them.

* The oracle's assembled matrices (hpGEM's integrands and quadrature) on the
  device give the oracle's time derivative and RK4 run (FP64 bar of
  tests/_parity.py), for advection and acoustics on squares, cubes (also
  rotated and displaced) and hp tetrahedra.
* An arbitrary linear operator (random element and face blocks of varying
  sizes, symmetric positive definite masses), as a user integrand could
  produce it: the device reproduces numpy's restatement of the reference's
  stiffness product and mass solve to 1e-12, and the right-hand side, mass
  solve, difference norm, step error and RK4 steps built on them.
"""
import numpy as np
import pytest

import _oracle as orc
import hpgem_amd
from _meshes import perturb, rotate_elements
from _parity import assert_parity
from hpgem_amd import HpgemAPIMatrixDevice, Mesh, RK4

pytestmark = pytest.mark.gpu


def oracle_pair(mesh, orders, kind, nf):
    refs = []
    for ext in (False, True):
        P = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, orders, n_fields=nf, extended=ext)
        if kind == "advection":
            P.set_advection([1.0, 0.7, 0.3][: mesh.dim])
        else:
            P.set_acoustic(np.linspace(0.5, 2.0, mesh.n_elements))
        refs.append(P)
    return refs


def device_from_oracle(P, **kw):
    M, S, fe, blocks, sizes = P.matrices()
    n = [int(np.sqrt(m.size)) for m in M]
    return HpgemAPIMatrixDevice([m.reshape(k, k, order="F") for m, k in zip(M, n)],
                                [s.reshape(k, k, order="F") for s, k in zip(S, n)], fe, blocks,
                                n_fields=P.n_fields, **kw)


CASES = [
    # (label, mesh builder, orders, kind)
    ("quads 5x4 p=2 advection", lambda: Mesh.structured([5, 4]), 2, "advection"),
    ("hex 3^3 p=3 advection", lambda: Mesh.structured([3, 3, 3]), 3, "advection"),
    ("hex 3^3 p=2 acoustic, rotated + displaced, walls", "both", 2, "acoustic"),
    ("tets 2^3x5 random orders acoustic", lambda: Mesh.structured([2, 2, 2], periodic=[0, 0, 0], simplices=True),
     "random", "acoustic"),
]


def build(spec):
    if spec == "both":
        m = Mesh.structured([3, 3, 3], periodic=[0, 0, 0])
        X, nodes, _ = rotate_elements(perturb(m.coords), m.node_ids, 5)
        return Mesh(m.shape, X, nodes)
    return spec()


@pytest.mark.parametrize("label,spec,orders,kind", CASES, ids=[c[0] for c in CASES])
def test_oracle_matrices_give_the_oracle_time_derivative(label, spec, orders, kind):
    print(f"[case] matrix operator: {label}", flush=True)
    mesh = build(spec)
    if orders == "random":
        orders = np.random.default_rng(1).integers(1, 5, size=mesh.n_elements).astype(np.int32)
    nf = 1 if kind == "advection" else mesh.dim + 1
    ref, ref_ld = oracle_pair(mesh, orders, kind, nf)
    dev = device_from_oracle(ref, n_extra_vectors=2)
    assert dev.n_dof == ref.n_dof
    rng = np.random.default_rng(3)
    u, k = rng.standard_normal(ref.n_dof), rng.standard_normal(ref.n_dof)
    dev.setTimeIntegrationVector(0, u)
    dev.setTimeIntegrationVector(2, k)
    dev.computeTimeDerivative([0, 2], [1.0, 0.25], 1)
    p_max = int(np.max(orders))
    assert_parity(dev.getTimeIntegrationVector(1), ref.time_derivative([u, k], [1.0, 0.25]),
                  ref_ld.time_derivative([u, k], [1.0, 0.25]), p_max, ref, f"matrix operator {label}")
    # the stiffness product alone and the mass solve
    dev.computeRightHandSide([0, 2], [1.0, 0.25], 1)
    assert_parity(dev.getTimeIntegrationVector(1), ref.right_hand_side([u, k], [1.0, 0.25]),
                  ref_ld.right_hand_side([u, k], [1.0, 0.25]), p_max, ref, f"matrix right-hand side {label}")
    dev.solveMassMatrixEquations(1)
    assert_parity(dev.getTimeIntegrationVector(1), ref.time_derivative([u, k], [1.0, 0.25]),
                  ref_ld.time_derivative([u, k], [1.0, 0.25]), p_max, ref, f"matrix rhs + solve {label}")
    dev.close()


def test_oracle_matrices_run_rk4_like_the_oracle():
    print("[case] matrix operator RK4 x5, hex 3^3 p=2 advection", flush=True)
    mesh = Mesh.structured([3, 3, 3])
    ref, ref_ld = oracle_pair(mesh, 2, "advection", 1)
    dev = device_from_oracle(ref)
    u0 = ref.project_modes([0.6, -0.3], [[1, 2, 1], [2, 1, 3]])
    dev.setTimeIntegrationVector(0, u0)
    dev.computeOneTimeStep(0.0, 0.01, 5)
    assert_parity(dev.getTimeIntegrationVector(0), ref.rk4(u0, 0.01, 5), ref_ld.rk4(u0, 0.01, 5), 2, ref,
                  "matrix operator RK4 x5")
    dev.close()


def random_operator(seed, n_elem=40):
    """Element blocks of sizes 1..60, faces between random element pairs and
    boundary faces, SPD masses: a linear operator no integrand in this
    library produces."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, 61, size=n_elem)
    mass, stiff = [], []
    for n in sizes:
        A = rng.standard_normal((n, n))
        mass.append(A @ A.T / n + np.eye(n))
        stiff.append(rng.standard_normal((n, n)))
    faces, blocks = [], []
    for _ in range(3 * n_elem):
        L, R = (int(x) for x in rng.choice(n_elem, 2, replace=False))
        if L > R:
            L, R = R, L
        nL, nR = sizes[L], sizes[R]
        faces.append((L, R))
        blocks.append(np.concatenate([rng.standard_normal(nL * nL), rng.standard_normal(nL * nR),
                                      rng.standard_normal(nR * nL), rng.standard_normal(nR * nR)]))
    for _ in range(n_elem // 4):
        L = int(rng.integers(n_elem))
        faces.append((L, -1))
        blocks.append(rng.standard_normal(sizes[L] ** 2))
    return sizes, mass, stiff, np.array(faces), blocks


def numpy_operator(sizes, mass, stiff, faces, blocks, w):
    """multiplyStiffnessMatrices restated: element products, then per face the
    four block products; returns r and M^{-1} r."""
    off = np.concatenate([[0], np.cumsum(sizes)])
    r = np.zeros_like(w)
    for e, S in enumerate(stiff):
        r[off[e]:off[e + 1]] = S @ w[off[e]:off[e + 1]]
    for (L, R), b in zip(faces, blocks):
        nL = sizes[L]
        wL = w[off[L]:off[L + 1]]
        LL = b[: nL * nL].reshape(nL, nL, order="F")
        r[off[L]:off[L + 1]] += LL @ wL
        if R >= 0:
            nR = sizes[R]
            wR = w[off[R]:off[R + 1]]
            o = nL * nL
            LR = b[o:o + nL * nR].reshape(nL, nR, order="F")
            o += nL * nR
            RL = b[o:o + nR * nL].reshape(nR, nL, order="F")
            o += nR * nL
            RR = b[o:o + nR * nR].reshape(nR, nR, order="F")
            r[off[L]:off[L + 1]] += LR @ wR
            r[off[R]:off[R + 1]] += RL @ wL + RR @ wR
    k = np.concatenate([np.linalg.solve(M, r[off[e]:off[e + 1]]) for e, M in enumerate(mass)])
    return r, k


def test_arbitrary_linear_operator_matches_numpy():
    print("[case] matrix operator: random operator, 40 elements of 1..60 unknowns", flush=True)
    sizes, mass, stiff, faces, blocks = random_operator(7)
    dev = HpgemAPIMatrixDevice(mass, stiff, faces, blocks, tableau=hpgem_amd.DORMAND_PRINCE, n_extra_vectors=2)
    n = int(np.sum(sizes))
    assert dev.n_dof == n
    rng = np.random.default_rng(8)
    u, k = rng.standard_normal(n), rng.standard_normal(n)
    w = u - 0.4 * k
    r_ref, k_ref = numpy_operator(sizes, mass, stiff, faces, blocks, w)
    dev.setTimeIntegrationVector(0, u)
    dev.setTimeIntegrationVector(8, k)
    dev.computeTimeDerivative([0, 8], [1.0, -0.4], 1)
    got = dev.getTimeIntegrationVector(1)
    err = np.max(np.abs(got - k_ref)) / np.max(np.abs(k_ref))
    dev.computeRightHandSide([0, 8], [1.0, -0.4], 2)
    r = dev.getTimeIntegrationVector(2)
    err_r = np.max(np.abs(r - r_ref)) / np.max(np.abs(r_ref))
    dev.solveMassMatrixEquations(2)
    err_s = np.max(np.abs(dev.getTimeIntegrationVector(2) - k_ref)) / np.max(np.abs(k_ref))
    print(f"  time derivative {err:.2e}, right-hand side {err_r:.2e}, mass solve {err_s:.2e}")
    assert err < 1e-12 and err_r < 1e-12 and err_s < 1e-12
    # the difference norm: sum_e d_e^T M_e d_e
    l2, mx = dev.difference(0, 8)
    off = np.concatenate([[0], np.cumsum(sizes)])
    d = u - k
    want = np.sqrt(sum(d[off[e]:off[e + 1]] @ M @ d[off[e]:off[e + 1]] for e, M in enumerate(mass)))
    assert abs(l2 - want) <= 1e-12 * want and mx == np.max(np.abs(d))
    # one Dormand-Prince step and its error estimate against numpy
    t = hpgem_amd.DORMAND_PRINCE
    dt = 0.01
    dev.setTimeIntegrationVector(0, u)
    dev.computeOneTimeStep(0.0, dt, 1)
    A = t.a_matrix()
    ks = []
    for i in range(t.stages):
        wi = u + sum(dt * A[i][j] * ks[j] for j in range(i))
        ks.append(numpy_operator(sizes, mass, stiff, faces, blocks, wi)[1])
    un = u + sum(dt * b * kk for b, kk in zip(t.b, ks))
    got = dev.getTimeIntegrationVector(0)
    assert np.max(np.abs(got - un)) <= 1e-12 * np.max(np.abs(un))
    est, nrm = dev.computeErrorAndNormOfUpdate(dt)
    est_ref = np.linalg.norm(sum(dt * e * kk for e, kk in zip(t.error, ks)))
    nrm_ref = sum(np.linalg.norm(un[off[e]:off[e + 1]]) for e in range(len(sizes)))
    print(f"  DP step: estimate {est:.6e} / {est_ref:.6e}, norm {nrm:.6e} / {nrm_ref:.6e}")
    assert abs(est - est_ref) <= 1e-9 * est_ref and abs(nrm - nrm_ref) <= 1e-12 * nrm_ref
    dev.close()


def test_matrix_operator_refuses_projection_and_mixed_storage_keeps_1e6():
    sizes, mass, stiff, faces, blocks = random_operator(9, n_elem=12)
    dev = HpgemAPIMatrixDevice(mass, stiff, faces, blocks, precision="mixed", n_extra_vectors=1)
    with pytest.raises(hpgem_amd.HgdgError):
        dev.setInitialSolutionModes([1.0], [[1, 0, 0]])
    n = int(np.sum(sizes))
    u = np.random.default_rng(2).standard_normal(n)
    dev.setTimeIntegrationVector(0, u)
    for v in dev.auxiliaryVectorIds_:
        dev.setStorage(v, "mixed")
    dev.computeTimeDerivative([0], [1.0], 1)
    _, k_ref = numpy_operator(sizes, mass, stiff, faces, blocks, u)
    got = dev.getTimeIntegrationVector(1)
    assert np.linalg.norm(got - k_ref) <= 1e-6 * np.linalg.norm(k_ref)
    dev.close()


@pytest.mark.parametrize("precision,tab,steps", [("fp64", "RK4", 6), ("mixed", "RK4", 5),
                                                 ("fp64", "DORMAND_PRINCE", 4)])
def test_matrix_operator_fused_update_steps_bitwise_equal(precision, tab, steps):
    """Pairs of RK steps with the update in each last-stage kernel: bitwise
    the same as the separate update kernel, for an arbitrary operator."""
    print(f"[case] matrix operator fused update {precision} {tab} steps={steps}", flush=True)
    tableau = getattr(hpgem_amd.tableaux, tab)
    sizes, mass, stiff, faces, blocks = random_operator(11, n_elem=24)
    u0 = np.random.default_rng(12).standard_normal(int(np.sum(sizes)))
    out = {}
    for fused in (False, True):
        dev = HpgemAPIMatrixDevice(mass, stiff, faces, blocks, precision=precision, tableau=tableau,
                                   n_extra_vectors=2)
        dev.setFusedTimeStepping(fused)
        dev.setTimeIntegrationVector(0, u0)
        dev.computeOneTimeStep(0.0, 0.002, steps)
        dev.computeOneTimeStep(0.0, 0.002, 3)
        out[fused] = [dev.getTimeIntegrationVector(v) for v in range(tableau.stages + 1)]
        dev.close()
    assert np.all(np.isfinite(out[False][0]))
    assert not np.array_equal(out[False][0], u0)
    for v in range(tableau.stages + 1):
        np.testing.assert_array_equal(out[True][v], out[False][v])

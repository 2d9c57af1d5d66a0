"""GPU parity of the advection RHS (volume + upwind faces + inverse mass, with
the RK stage combination) and of RK time stepping against the oracle.

FP64 bar (tests/_parity.py): <= 1e-12 relative (max abs coefficient
difference / max abs value) against the oracle run in long double for every
case; <= 1e-12 against the oracle in double (hpGEM's own rounding) for every
case of order <= 2; at p >= 3, where hpGEM's per-call LU of a cube mass
matrix of condition ~1.9e5 (p = 3) / ~3e6 (p = 4) leaves the double oracle
itself up to 1.35e-12 / 2e-11 from the exact operator, <= 1e-12 relative in
the mass-weighted L2 norm (every distance is printed and kept). Mixed FP32
path: within 1e-6 relative L2 of the FP64 solution.
"""
import numpy as np
import pytest

import _oracle as orc
import hpgem_amd
from _parity import assert_parity as _assert_parity
from _parity import rel
from hpgem_amd import HpgemAPIDevice, Mesh, RK4

pytestmark = pytest.mark.gpu


def assert_parity(got, ref64, ref_ld, p, oracle, label=""):
    _assert_parity(got, ref64, ref_ld, p, oracle, label)


def make_pair(n, p, a, periodic=True, shear=None, precision="fp64", tableau=RK4, extended=False):
    mesh = Mesh.structured(n, periodic=[int(periodic)] * len(n))
    if shear is not None:
        # affine shear x -> S x keeps every element a parallelogram/parallelepiped
        mesh = Mesh(mesh.shape, mesh.coords @ np.asarray(shear).T, mesh.node_ids)
    dev = HpgemAPIDevice(mesh, p, velocity=a, precision=precision, tableau=tableau, n_extra_vectors=2)
    ref = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p)
    ref.set_advection(a)
    if not extended:
        return dev, ref
    ref_ld = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, extended=True)
    ref_ld.set_advection(a)
    return dev, ref, ref_ld


def seeded_state(n_dof, seed):
    return np.random.default_rng(seed).standard_normal(n_dof)


CASES = [
    # (cells, p, velocity, periodic, shear)
    ([16, 16], 2, [1.0, 0.5], True, None),              # the 2D quad p=2 workload's mesh
    ([5, 4], 1, [1.0, -0.5], True, None),
    ([4, 5], 4, [-0.3, 0.8], True, None),
    ([6, 5], 3, [1.0, 0.0], True, None),                # a.n = 0 on y faces: averaged flux
    ([4, 4], 2, [1.0, 0.5], False, None),               # boundary faces: outflow / inflow / none
    ([4, 3, 3], 1, [1.0, 0.7, 0.3], True, None),
    ([3, 4, 3], 2, [-1.0, 0.7, 0.3], True, None),
    ([4, 4, 4], 3, [1.0, 0.7, 0.3], True, None),         # the 3D hex p=3 workload's element
    ([3, 3, 3], 4, [0.2, -0.7, 0.3], True, None),
    ([3, 3, 4], 3, [1.0, 0.7, 0.3], False, None),
    ([3, 3, 3], 2, [1.0, 0.0, 0.3], False, None),
    ([4, 3], 2, [1.0, 0.5], True, [[1.0, 0.3], [0.0, 1.0]]),            # sheared quads
    ([3, 3, 3], 2, [1.0, 0.7, 0.3], True, [[1, 0.2, 0], [0, 1, 0.1], [0.1, 0, 1]]),
]


@pytest.mark.parametrize("n,p,a,periodic,shear", CASES)
def test_time_derivative_matches_oracle(n, p, a, periodic, shear):
    print(f"[case] rhs n={n} p={p} a={a} periodic={periodic} shear={shear is not None}", flush=True)
    dev, ref, ref_ld = make_pair(n, p, a, periodic, shear, extended=True)
    u = seeded_state(dev.n_dof, 1)
    k = seeded_state(dev.n_dof, 2)
    dev.setTimeIntegrationVector(0, u)
    dev.setTimeIntegrationVector(1, k)
    # single input
    dev.computeTimeDerivative([0], [1.0], 2)
    got = dev.getTimeIntegrationVector(2)
    label = f"advection rhs n={n} p={p} periodic={periodic} shear={shear is not None}"
    assert_parity(got, ref.time_derivative([u], [1.0]), ref_ld.time_derivative([u], [1.0]), p, ref, label)
    # RK-style stage combination u + 0.37 k
    dev.computeTimeDerivative([0, 1], [1.0, 0.37], 3)
    got = dev.getTimeIntegrationVector(3)
    assert_parity(got, ref.time_derivative([u, k], [1.0, 0.37]), ref_ld.time_derivative([u, k], [1.0, 0.37]), p, ref,
                  label + " u+0.37k")
    dev.close()


def test_face_and_dof_counts_match_oracle():
    for n, periodic in [([4, 4, 4], True), ([3, 4, 2], False), ([5, 6], True)]:
        dev, ref = make_pair(n, 2, [1.0, 0.5, 0.2][: len(n)], periodic)
        assert dev.n_dof == ref.n_dof
        assert dev.n_faces == ref.n_faces
        assert dev.n_internal_faces == ref.n_internal
        dev.close()


def test_rk4_steps_match_oracle_2d_quads_p2():
    """16x16 periodic quads, p=2, a=(1, 0.5), RK4 dt = 0.1 h/|a|, 100 steps,
    four seeded sine modes: the 2D workload end to end."""
    rng = np.random.default_rng(7)
    amp = rng.uniform(-1, 1, 4)
    kvec = rng.integers(1, 5, size=(4, 2)).astype(float)
    a = [1.0, 0.5]
    dev, ref, ref_ld = make_pair([16, 16], 2, a, extended=True)
    dt = 0.1 * (1 / 16) / np.hypot(*a)
    u0 = ref.project_modes(amp, kvec)
    dev.setInitialSolutionModes(amp, kvec)
    assert_parity(dev.getTimeIntegrationVector(0), u0, ref_ld.project_modes(amp, kvec), 2, ref, "quad16 p=2 projection")
    dev.computeOneTimeStep(0.0, dt, 100)
    got = dev.getTimeIntegrationVector(0)
    assert_parity(got, ref.rk4(u0, dt, 100), ref_ld.rk4(u0, dt, 100), 2, ref, "quad16 p=2 RK4 x100")
    dev.close()


def test_rk4_steps_match_oracle_3d_hex_p3():
    rng = np.random.default_rng(11)
    amp = rng.uniform(-1, 1, 8)
    kvec = rng.integers(1, 7, size=(8, 3)).astype(float)
    a = [1.0, 0.7, 0.3]
    dev, ref, ref_ld = make_pair([4, 4, 4], 3, a, extended=True)
    dt = 0.1 * (1 / 4) / np.linalg.norm(a)
    u0 = ref.project_modes(amp, kvec)
    dev.setInitialSolutionModes(amp, kvec)
    assert_parity(dev.getTimeIntegrationVector(0), u0, ref_ld.project_modes(amp, kvec), 3, ref, "hex4 p=3 projection")
    dev.computeOneTimeStep(0.0, dt, 5)
    assert_parity(dev.getTimeIntegrationVector(0), ref.rk4(u0, dt, 5), ref_ld.rk4(u0, dt, 5), 3, ref,
                  "hex4 p=3 RK4 x5")
    dev.close()


def test_dormand_prince_reference_error_table_on_gpu():
    """tests/self/Basic/090Advection_SelfTest.cpp:330-335 run through the GPU
    path: 3D periodic hex p=1, 2 Dormand-Prince steps to T = 0.02."""
    from test_oracle_physics import PROD3_AMP, PROD3_K
    for n, expected in [(4, 5.55789642e-02), (8, 1.44021875e-02), (16, 3.81732285e-03)]:
        dev, ref = make_pair([n, n, n], 1, [0.1, 0.2, 0.3], tableau=hpgem_amd.DORMAND_PRINCE)
        dev.setInitialSolutionModes(PROD3_AMP, PROD3_K)
        dev.computeOneTimeStep(0.0, 0.01, 2)
        u = dev.getTimeIntegrationVector(0)
        err = ref.l2_error_modes(u, PROD3_AMP, PROD3_K, 0.02)
        assert abs(err - expected) < 0.01 * expected, (n, err, expected)
        # the library's own error norm (computeTotalError on the device)
        l2, _ = dev.computeErrorModes(PROD3_AMP, PROD3_K, shift=[0.1 * 0.02, 0.2 * 0.02, 0.3 * 0.02])
        assert abs(l2 - err) < 1e-12 * err, (n, l2, err)
        dev.close()


@pytest.mark.parametrize("n,p", [([3, 4], 2), ([3, 3, 3], 2), ([3, 4, 3], 3)])
def test_max_error_matches_pointwise_evaluation(n, p):
    """computeMaxError: the largest |f - u_h| at the element rule's (p+1)^dim
    Gauss points, checked against the oracle's basis evaluated at those
    points (kernel/Base/HpgemAPISimplified_Impl.h:403-530)."""
    a = [1.0, 0.5, 0.25][: len(n)]
    dev, ref = make_pair(n, p, a)
    amp, kv = [0.8, -0.3], [[1, 2, 1][: len(n)], [2, 1, 1][: len(n)]]
    u = seeded_state(dev.n_dof, 5) * 0.05
    dev.setTimeIntegrationVector(0, u)
    shift = [0.01, 0.02, 0.03][: len(n)]
    l2, mx = dev.computeErrorModes(amp, kv, shift=shift)
    mesh = Mesh.structured(n)
    xg, _ = np.polynomial.legendre.leggauss(p + 1)
    dim = len(n)
    grids = np.stack(np.meshgrid(*([xg] * dim), indexing="ij"), -1).reshape(-1, dim)
    nb = (p + 1) ** dim
    best = 0.0
    for e in range(mesh.n_elements):
        x0 = mesh.coords[e, 0]
        edges = [mesh.coords[e, 1 << d] - x0 for d in range(dim)]
        for xi in grids:
            phi, _ = orc.basis_eval(mesh.shape, p, xi)
            val = phi @ u[e * nb:(e + 1) * nb]
            x = x0 + sum((xi[d] + 1) / 2 * edges[d] for d in range(dim))
            ex = sum(A * np.sin(2 * np.pi * np.dot(k, x - np.asarray(shift))) for A, k in zip(amp, kv))
            best = max(best, abs(ex - val))
    print(f"max error {mx:.6e} vs pointwise {best:.6e}; L2 {l2:.4e}")
    assert abs(mx - best) < 1e-12 * max(best, 1.0)
    assert l2 > 0
    dev.close()


def test_scale_and_add_matches_numpy():
    dev, _ = make_pair([4, 4], 2, [1.0, 0.5])
    x = seeded_state(dev.n_dof, 3)
    y = seeded_state(dev.n_dof, 4)
    dev.setTimeIntegrationVector(0, x)
    dev.setTimeIntegrationVector(1, y)
    dev.scaleAndAddVector(0, 1, -0.25)
    np.testing.assert_array_equal(dev.getTimeIntegrationVector(0), x + (-0.25) * y)
    dev.close()


@pytest.mark.parametrize("n,p,a", [([16, 16], 2, [1.0, 0.5]), ([6, 6, 6], 3, [1.0, 0.7, 0.3]),
                                   ([4, 4, 4], 3, [1.0, 0.0, -0.3])])
def test_mixed_rhs_single_eval(n, p, a):
    """One mixed-precision RHS (FP64 u, FP32 k input, FP32 output) against the
    FP64 kernel on the same inputs: FP32 rounding only (~1e-7)."""
    dev64, _ = make_pair(n, p, a)
    dev32, _ = make_pair(n, p, a, precision="mixed")
    rng = np.random.default_rng(3)
    u = rng.standard_normal(dev64.n_dof)
    k = rng.standard_normal(dev64.n_dof)
    dev32.setStorage(1, "mixed")
    dev32.setStorage(2, "mixed")
    k32 = k.astype(np.float32).astype(np.float64)
    for d, kk in ((dev64, k32), (dev32, k)):
        d.setTimeIntegrationVector(0, u)
        d.setTimeIntegrationVector(1, kk)
        d.computeTimeDerivative([0, 1], [1.0, 0.01], 2)
    r64 = dev64.getTimeIntegrationVector(2)
    r32 = dev32.getTimeIntegrationVector(2)
    err = rel(r32, r64)
    print(f"mixed single RHS rel err {err:.3e}")
    assert err < 2e-6
    dev64.close()
    dev32.close()


@pytest.mark.parametrize("n,p,a", [([16, 16], 2, [1.0, 0.5]), ([8, 8, 8], 3, [1.0, 0.7, 0.3])])
def test_mixed_precision_within_1e6_of_fp64(n, p, a):
    rng = np.random.default_rng(5)
    dim = len(n)
    amp = rng.uniform(-1, 1, 8)
    kvec = rng.integers(1, 4, size=(8, dim)).astype(float)
    # 0.05 h/|a| lies inside RK4's stability region for p <= 3 on these meshes
    dt = 0.05 * (1 / n[0]) / np.linalg.norm(a)
    dev64, _ = make_pair(n, p, a)
    dev32, _ = make_pair(n, p, a, precision="mixed")
    for d in (dev64, dev32):
        d.setInitialSolutionModes(amp, kvec)
        d.computeOneTimeStep(0.0, dt, 20)
    u64 = dev64.getTimeIntegrationVector(0)
    u32 = dev32.getTimeIntegrationVector(0)
    # L2 (mass-weighted) difference via the device norm on one context
    dev64.setTimeIntegrationVector(dev64.n_vectors - 1, u32)
    l2_diff, max_diff = dev64.difference(0, dev64.n_vectors - 1)
    dev64.setTimeIntegrationVector(dev64.n_vectors - 2, np.zeros_like(u64))
    l2_ref, _ = dev64.difference(0, dev64.n_vectors - 2)
    print(f"mixed vs fp64: rel L2 {l2_diff / l2_ref:.3e}, max {max_diff:.3e}")
    assert l2_diff / l2_ref < 1e-6
    dev64.close()
    dev32.close()


@pytest.mark.parametrize("n,p,precision,periodic", [
    ([32, 32, 32], 3, "fp64", True),     # the headline workload
    ([32, 32, 32], 3, "mixed", True),
    ([16, 16], 2, "fp64", True),
    ([5, 4, 3], 2, "fp64", False),       # 60 elements: a ragged last tile, boundary faces
    ([7, 5], 4, "mixed", False),
    ([5, 4, 3], 2, "fp64", True),        # periodic, ragged last tile
    ([3, 3, 3], 3, "mixed", True),       # every neighbour inside one or two blocks
    ([24, 16, 8], 3, "fp64", True),      # non-cubic box
])
def test_tile_counts_give_bitwise_identical_rhs(n, p, precision, periodic):
    """Every launch shape of the advection kernel (one tile per workgroup, or
    2/4/8 tiles with the next tile prefetched) performs the same arithmetic
    per element, so the RHS is bitwise the same; the timed choice is one of
    the candidates."""
    a = [1.0, 0.7, 0.3][:len(n)]
    dev, _ = make_pair(n, p, a, periodic=periodic, precision=precision)
    tiles, ms = dev.getAdvectionTiles()
    print(f"tuned tiles {tiles}, ms {ms}")
    assert tiles in (1, 2, 4, 8) and np.all(ms > 0)
    rng = np.random.default_rng(11)
    if precision == "mixed":
        dev.setStorage(1, "mixed")
        dev.setStorage(2, "mixed")
    dev.setTimeIntegrationVector(0, rng.standard_normal(dev.n_dof))
    dev.setTimeIntegrationVector(1, rng.standard_normal(dev.n_dof))
    outs = {}
    shapes = (1, 2, 4, 8)
    for t in shapes:
        dev.setAdvectionTiles(t)
        assert dev.getAdvectionTiles()[0] == t
        dev.setTimeIntegrationVector(2, np.full(dev.n_dof, np.nan))
        dev.computeTimeDerivative([0, 1], [1.0, 0.25], 2)
        outs[t] = dev.getTimeIntegrationVector(2)
    assert np.all(np.isfinite(outs[1]))
    for t in shapes[1:]:
        np.testing.assert_array_equal(outs[t], outs[1])
    dev.setAdvectionTiles(0)
    assert dev.getAdvectionTiles()[0] in shapes
    dev.close()


def test_full_size_hex_p3_properties():
    """32^3 periodic hex p=3 (the headline workload) checked through
    size-independent properties: constant states are steady, the operator is
    linear, and on the uniform periodic mesh a one-cell shift of the input
    shifts the output exactly."""
    a = [1.0, 0.7, 0.3]
    mesh = Mesh.structured([32, 32, 32])
    dev = HpgemAPIDevice(mesh, 3, velocity=a, n_extra_vectors=3)
    nb = 64
    ne = mesh.n_elements
    # constant state: the vertex functions all carry the constant
    c = np.zeros((ne, nb))
    c[:, :8] = 2.5
    dev.setTimeIntegrationVector(0, c.ravel())
    dev.computeTimeDerivative([0], [1.0], 1)
    k = dev.getTimeIntegrationVector(1)
    assert np.max(np.abs(k)) < 1e-10 * 2.5 * 32
    # shift invariance by one cell in x
    u = seeded_state(dev.n_dof, 9).reshape(32, 32, 32, nb)
    dev.setTimeIntegrationVector(0, u.ravel())
    dev.computeTimeDerivative([0], [1.0], 1)
    k0 = dev.getTimeIntegrationVector(1).reshape(32, 32, 32, nb)
    dev.setTimeIntegrationVector(0, np.roll(u, 1, axis=2).ravel())
    dev.computeTimeDerivative([0], [1.0], 1)
    k1 = dev.getTimeIntegrationVector(1).reshape(32, 32, 32, nb)
    assert np.array_equal(k1, np.roll(k0, 1, axis=2))
    # linearity
    v = seeded_state(dev.n_dof, 10)
    dev.setTimeIntegrationVector(0, u.ravel())
    dev.setTimeIntegrationVector(2, v)
    dev.computeTimeDerivative([0, 2], [2.0, -0.5], 3)
    k_comb = dev.getTimeIntegrationVector(3)
    dev.computeTimeDerivative([2], [1.0], 4)
    k_v = dev.getTimeIntegrationVector(4)
    lin = 2.0 * k0.ravel() - 0.5 * k_v
    assert rel(k_comb, lin) < 1e-13
    dev.close()


@pytest.mark.parametrize("name,p,a", [("unitCubeN4Per.hpgem", 3, [1.0, 0.7, 0.3]),
                                      ("unitCubeN4.hpgem", 2, [-0.4, 1.0, 0.6]),
                                      ("2Drectangular2mesh.hpgem", 4, [1.0, 0.5])])
def test_rhs_on_reference_mesh_file(name, p, a):
    """Meshes read from the reference's own .hpgem files (periodic copies,
    boundary faces) through hgdg_mesh_read, then the operator on the GPU."""
    import os
    print(f"[case] mesh file {name} p={p}", flush=True)
    mesh = Mesh.read(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes", name))
    dev = HpgemAPIDevice(mesh, p, velocity=a, tableau=RK4)
    ref = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p)
    ref.set_advection(a)
    ref_ld = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, extended=True)
    ref_ld.set_advection(a)
    assert dev.n_faces == len(mesh.face_elements)
    u = seeded_state(dev.n_dof, 7)
    dev.setTimeIntegrationVector(0, u)
    dev.computeTimeDerivative([0], [1.0], 1)
    assert_parity(dev.getTimeIntegrationVector(1), ref.time_derivative([u], [1.0]), ref_ld.time_derivative([u], [1.0]),
                  p, ref, f"mesh file {name} p={p}")
    dev.close()


def test_step_error_and_norm_match_their_definition():
    """computeErrorAndNormOfUpdate: ||sum_j dt e_j k_j||_2 and the sum of the
    elements' coefficient norms, against numpy on the downloaded vectors."""
    dev, _ = make_pair([4, 4, 4], 2, [1.0, 0.7, 0.3], tableau=hpgem_amd.DORMAND_PRINCE)
    dev.setInitialSolutionModes([0.6, -0.3], [[1, 2, 1], [2, 1, 3]])
    dt = 0.004
    dev.computeOneTimeStep(0.0, dt, 1)
    err, nrm = dev.computeErrorAndNormOfUpdate(dt)
    t = hpgem_amd.DORMAND_PRINCE
    ks = [dev.getTimeIntegrationVector(i) for i in dev.auxiliaryVectorIds_]
    dev_sum = sum(dt * e * k for e, k in zip(t.error, ks))
    u = dev.getTimeIntegrationVector(0)
    nb = u.size // dev.n_elements
    ref_err = np.linalg.norm(dev_sum)
    ref_nrm = np.linalg.norm(u.reshape(dev.n_elements, nb), axis=1).sum()
    assert abs(err - ref_err) < 1e-12 * ref_err and abs(nrm - ref_nrm) < 1e-12 * ref_nrm
    dev.close()


def test_adaptive_step_accepts_and_estimates_like_the_reference_loop():
    """A loose tolerance accepts dt_initial at once: the state equals one fixed
    Dormand-Prince step and the next estimate is 0.9 dt (tol norm/err)^(1/5)."""
    a = [1.0, 0.7, 0.3]
    d1, _ = make_pair([4, 4, 4], 2, a, tableau=hpgem_amd.DORMAND_PRINCE)
    d2, _ = make_pair([4, 4, 4], 2, a, tableau=hpgem_amd.DORMAND_PRINCE)
    for d in (d1, d2):
        d.setInitialSolutionModes([0.6, -0.3], [[1, 2, 1], [2, 1, 3]])
    t1, dt1, err, nrm = d1.computeOneTimeStepAdaptive(0.0, 1.0, 1.0, 0.01)
    d2.computeOneTimeStep(0.0, 0.01, 1)
    assert t1 == 0.01 and dt1 == 0.01 and err <= 1.0 * nrm
    np.testing.assert_array_equal(d1.getTimeIntegrationVector(0), d2.getTimeIntegrationVector(0))
    # the next step starts from the estimate (capped by dt_max)
    t2, dt2, err2, nrm2 = d1.computeOneTimeStepAdaptive(t1, 1.0, 0.015, 0.01)
    assert dt2 == min(0.015, 0.9 * 0.01 * (1.0 / err * nrm) ** 0.2)
    assert abs(t2 - (t1 + dt2)) < 1e-15
    d1.close()
    d2.close()


def test_adaptive_step_rejects_until_the_error_bound_holds():
    """A tight tolerance forces rejections: the accepted step satisfies
    error <= tol * norm with a smaller dt, and the run stays accurate."""
    a = [1.0, 0.7, 0.3]
    amp, kv = [0.6, -0.3], [[1, 2, 1], [2, 1, 3]]
    dev, ref = make_pair([6, 6, 6], 3, a, tableau=hpgem_amd.DORMAND_PRINCE)
    dev.setInitialSolutionModes(amp, kv)
    tol = 1e-9
    t, dts = 0.0, []
    while t < 0.02 - 1e-15:
        t, dt, err, nrm = dev.computeOneTimeStepAdaptive(t, tol, 0.02 - t, 0.05)
        assert err <= tol * nrm
        dts.append(dt)
    assert dts[0] < 0.05 and len(dts) >= 2
    l2, _ = dev.computeErrorModes(amp, kv, shift=[x * t for x in a])
    u0 = ref.project_modes(amp, kv)
    l2_0 = ref.l2_error_modes(u0, amp, kv, 0.0)
    print(f"adaptive DP: {len(dts)} steps, dt {dts}, L2 error {l2:.3e} (projection error {l2_0:.3e})")
    assert l2 < 1.5 * l2_0  # time error negligible next to the spatial one
    dev.close()


FUSED_CASES = [
    # (cells, p, velocity, periodic, precision, tableau, steps)
    ([16, 16], 2, [1.0, 0.5], True, "fp64", "RK4", 20),   # the 2D quad p=2 workload
    ([16, 16], 2, [1.0, 0.5], True, "mixed", "RK4", 20),
    ([4, 4], 2, [1.0, 0.5], False, "fp64", "RK4", 7),     # walls: inflow / outflow / no-flux faces
    ([5, 3], 3, [-0.4, 1.0], True, "fp64", "RK3_TVD", 9),
    ([3, 3], 1, [1.0, 0.3], True, "fp64", "DORMAND_PRINCE", 6),  # seven stages, zero tableau entries
    ([3, 3], 1, [1.0, 0.3], True, "fp64", "RK4", 1100),  # more steps than one launch takes
    ([3, 3, 3], 2, [1.0, 0.7, 0.3], True, "fp64", "RK4", 5),
    ([3, 3, 3], 2, [1.0, -0.7, 0.3], False, "mixed", "RK4", 5),
]


@pytest.mark.parametrize("n,p,a,periodic,precision,tab,steps", FUSED_CASES)
def test_fused_rk_steps_bitwise_equal_graph_replay(n, p, a, periodic, precision, tab, steps):
    """RK steps of a mesh whose state fits in LDS run in one single-workgroup
    launch that repeats the graph path's arithmetic operation for operation:
    the solution and every stage vector come out bitwise the same as from the
    replayed step graph."""
    print(f"[case] fused RK n={n} p={p} {precision} {tab} steps={steps}", flush=True)
    tableau = getattr(hpgem_amd.tableaux, tab)
    mesh = Mesh.structured(n, periodic=[int(periodic)] * len(n))
    u0 = None
    out = {}
    for fused in (False, True):
        dev = HpgemAPIDevice(mesh, p, velocity=a, precision=precision, tableau=tableau, n_extra_vectors=2)
        dev.setFusedTimeStepping(fused)
        if u0 is None:
            u0 = seeded_state(dev.n_dof, 5)
        dev.setTimeIntegrationVector(0, u0)
        h = 1.0 / max(n)
        dev.computeOneTimeStep(0.0, 0.05 * h / np.linalg.norm(a), steps)
        enabled, launches = dev.getFusedTimeStepping()
        assert enabled == fused and (launches > 0) == fused
        out[fused] = [dev.getTimeIntegrationVector(v) for v in range(tableau.stages + 1)]
        dev.close()
    assert np.all(np.isfinite(out[True][0]))
    for g, f in zip(out[False], out[True]):
        np.testing.assert_array_equal(f, g)


@pytest.mark.parametrize("n,p,periodic", [([4, 4], 2, True), ([3, 3, 3], 3, False)])
def test_right_hand_side_and_mass_solve_match_oracle(n, p, periodic):
    """hgdg_right_hand_side and hgdg_solve_mass (the two halves of
    computeTimeDerivative in HpgemAPILinear) against the oracle."""
    print(f"[case] advection right-hand side and mass solve n={n} p={p}", flush=True)
    a = [1.0, 0.7, 0.3][: len(n)]
    dev, ref, ref_ld = make_pair(n, p, a, periodic, extended=True)
    rng = np.random.default_rng(31)
    u = rng.standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(0, u)
    dev.computeRightHandSide([0], [1.0], 1)
    assert_parity(dev.getTimeIntegrationVector(1), ref.right_hand_side([u], [1.0]), ref_ld.right_hand_side([u], [1.0]),
                  p, ref, f"advection right-hand side n={n} p={p}")
    dev.setTimeIntegrationVector(2, u)
    dev.solveMassMatrixEquations(2)
    assert_parity(dev.getTimeIntegrationVector(2), ref.solve_mass(u), ref_ld.solve_mass(u), p, ref,
                  f"advection mass solve n={n} p={p}")
    dev.close()


PAIR_CASES = [
    # (cells, p, velocity, periodic, precision, tableau, steps): meshes too large for the single-workgroup path
    ([8, 8, 8], 3, [1.0, 0.7, 0.3], True, "fp64", "RK4", 6),
    ([8, 8, 8], 3, [1.0, 0.7, 0.3], True, "mixed", "RK4", 7),    # three pairs and one separate step
    ([10, 6, 4], 2, [-0.5, 1.0, 0.2], False, "fp64", "RK3_TVD", 5),
    ([12, 8, 6], 1, [0.3, -0.2, 1.0], True, "fp64", "DORMAND_PRINCE", 4),  # last stage has b = 0
    ([40, 30], 3, [1.0, 0.5], True, "mixed", "RK4", 2),
]


@pytest.mark.parametrize("n,p,a,periodic,precision,tab,steps", PAIR_CASES)
def test_fused_update_steps_bitwise_equal_graph_replay(n, p, a, periodic, precision, tab, steps):
    """Pairs of RK steps with the update fused into each step's last stage
    (the solution ping-pongs through a spare vector) perform the FP64
    operations of the separate update kernel: the solution and every stage
    vector come out bitwise the same as from the replayed one-step graph."""
    print(f"[case] fused update n={n} p={p} {precision} {tab} steps={steps}", flush=True)
    tableau = getattr(hpgem_amd.tableaux, tab)
    mesh = Mesh.structured(n, periodic=[int(periodic)] * len(n))
    u0 = None
    out = {}
    for fused in (False, True):
        dev = HpgemAPIDevice(mesh, p, velocity=a, precision=precision, tableau=tableau, n_extra_vectors=2)
        dev.setFusedTimeStepping(fused)
        if u0 is None:
            u0 = seeded_state(dev.n_dof, 6)
        dev.setTimeIntegrationVector(0, u0)
        dt = 0.05 / max(n) / np.linalg.norm(a)
        dev.computeOneTimeStep(0.0, dt, steps)
        dev.computeOneTimeStep(0.0, dt, 3)  # the same graphs again, odd count
        enabled, launches = dev.getFusedTimeStepping()
        assert enabled == fused and launches == 0
        out[fused] = [dev.getTimeIntegrationVector(v) for v in range(tableau.stages + 1)]
        dev.close()
    assert np.all(np.isfinite(out[False][0]))
    for v in range(tableau.stages + 1):
        np.testing.assert_array_equal(out[True][v], out[False][v])

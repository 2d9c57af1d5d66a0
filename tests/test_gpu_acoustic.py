"""GPU parity of the acoustic-wave RHS (AcousticWaveLinear: central flux,
reflecting walls, per-element c_inv) on tetrahedra of per-element order 1..4
and on squares/cubes, against the oracle, with the FP64 bar of
tests/_parity.py (tetrahedra: <= 1e-12 against the double and the
long-double oracle at every order)."""
import numpy as np
import pytest

import _oracle as orc
from _parity import assert_parity, rel
from hpgem_amd import OP_ACOUSTIC_CENTRAL, RK4, HpgemAPIDevice, Mesh

pytestmark = pytest.mark.gpu


def make(n, orders, seed, periodic=False, extended=True):
    mesh = Mesh.structured(n, periodic=[int(periodic)] * 3, simplices=True)
    rng = np.random.default_rng(seed)
    if orders == "random":
        p = rng.integers(1, 5, size=mesh.n_elements).astype(np.int32)
    else:
        p = np.full(mesh.n_elements, orders, dtype=np.int32)
    c = rng.uniform(0.5, 2.0, size=mesh.n_elements)
    cinv = 1.0 / c ** 2
    dev = HpgemAPIDevice(mesh, p, precision="fp64", kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4,
                         n_extra_vectors=2)
    ref = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, n_fields=4)
    ref.set_acoustic(cinv)
    refs = [ref]
    if extended:
        ref_ld = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, n_fields=4, extended=True)
        ref_ld.set_acoustic(cinv)
        refs.append(ref_ld)
    return dev, refs


CASES = [([2, 2, 2], 1, False), ([2, 2, 2], 2, False), ([2, 2, 2], 3, False), ([2, 2, 2], 4, False),
         ([2, 2, 2], "random", False), ([3, 2, 2], "random", False), ([3, 3, 3], "random", True)]


@pytest.mark.parametrize("n,orders,periodic", CASES)
def test_acoustic_rhs_matches_oracle(n, orders, periodic):
    print(f"[case] tet acoustic n={n} p={orders} periodic={periodic}", flush=True)
    dev, (ref, ref_ld) = make(n, orders, seed=len(str(orders)) + n[0], periodic=periodic)
    assert dev.n_dof == ref.n_dof
    assert dev.n_faces == ref.n_faces and dev.n_internal_faces == ref.n_internal
    rng = np.random.default_rng(4)
    u = rng.standard_normal(dev.n_dof)
    k = rng.standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(0, u)
    dev.setTimeIntegrationVector(1, k)
    dev.computeTimeDerivative([0], [1.0], 2)
    pm = int(ref.orders.max())
    label = f"tet acoustic n={n} p={orders} periodic={periodic}"
    assert_parity(dev.getTimeIntegrationVector(2), ref.time_derivative([u], [1.0]),
                  ref_ld.time_derivative([u], [1.0]), pm, ref, label)
    dev.computeTimeDerivative([0, 1], [1.0, -0.2], 3)
    assert_parity(dev.getTimeIntegrationVector(3), ref.time_derivative([u, k], [1.0, -0.2]),
                  ref_ld.time_derivative([u, k], [1.0, -0.2]), pm, ref, label + " u-0.2k")
    dev.close()


def test_acoustic_rk4_matches_oracle_hp():
    dev, (ref, ref_ld) = make([2, 2, 2], "random", seed=21)
    rng = np.random.default_rng(8)
    u0 = rng.standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(0, u0)
    dt = 0.01
    dev.computeOneTimeStep(0.0, dt, 4)
    assert_parity(dev.getTimeIntegrationVector(0), ref.rk4(u0, dt, 4), ref_ld.rk4(u0, dt, 4), int(ref.orders.max()),
                  ref, "tet acoustic hp RK4 x4")
    dev.close()


def test_acoustic_energy_is_conserved_on_gpu():
    """Central fluxes and reflecting walls conserve u^T M_w k = 0; checked
    through the device output and the oracle's weighted mass matrices."""
    dev, (ref,) = make([2, 2, 2], "random", seed=5, extended=False)
    rng = np.random.default_rng(1)
    u = rng.standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(0, u)
    dev.computeTimeDerivative([0], [1.0], 1)
    k = dev.getTimeIntegrationVector(1)
    off = 0
    rate = scale = 0.0
    for e in range(dev.n_elements):
        nb4 = 4 * ((ref.orders[e] + 1) * (ref.orders[e] + 2) * (ref.orders[e] + 3) // 6)
        M = np.zeros(nb4 * nb4)
        orc.check(ref.lib.orc_mass_matrix(ref.h, e, orc._dp(M)))
        M = M.reshape(nb4, nb4, order="F")
        rate += u[off:off + nb4] @ M @ k[off:off + nb4]
        scale += np.abs(u[off:off + nb4]) @ np.abs(M @ k[off:off + nb4])
        off += nb4
    assert abs(rate) < 1e-12 * scale
    dev.close()


TENSOR_CASES = [([4, 4], 2, True), ([4, 3], 3, False), ([3, 3, 3], 1, True), ([3, 3, 3], 2, False),
                ([3, 4, 3], 3, True), ([3, 3, 3], 4, False),
                # grids whose rows are whole blocks and whose 8 rows split over the XCDs: banded block order
                ([4, 8, 3], 3, True), ([2, 8, 2], 4, False)]


@pytest.mark.parametrize("n,p,periodic", TENSOR_CASES)
def test_tensor_acoustic_rhs_matches_oracle(n, p, periodic):
    """The same acoustic operator on squares and cubes (walls on the
    non-periodic boxes)."""
    print(f"[case] tensor acoustic n={n} p={p} periodic={periodic}", flush=True)
    mesh = Mesh.structured(n, periodic=[int(periodic)] * len(n))
    rng = np.random.default_rng(p + len(n))
    cinv = rng.uniform(0.25, 4.0, size=mesh.n_elements)
    dev = HpgemAPIDevice(mesh, p, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4, n_extra_vectors=2)
    nf = len(n) + 1
    ref = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, n_fields=nf)
    ref.set_acoustic(cinv)
    ref_ld = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, n_fields=nf, extended=True)
    ref_ld.set_acoustic(cinv)
    assert dev.n_dof == ref.n_dof
    u = rng.standard_normal(dev.n_dof)
    k = rng.standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(0, u)
    dev.setTimeIntegrationVector(1, k)
    dev.computeTimeDerivative([0, 1], [1.0, 0.3], 2)
    assert_parity(dev.getTimeIntegrationVector(2), ref.time_derivative([u, k], [1.0, 0.3]),
                  ref_ld.time_derivative([u, k], [1.0, 0.3]), p, ref, f"tensor acoustic n={n} p={p} periodic={periodic}")
    dev.close()


# every instance's mixed path: the 3D kernels at p = 2, 3, 4 (p = 4 in FP32
# gathers the neighbour traces one input vector at a time) and the 2D
# four-lane kernel
MIXED_CASES = [([6, 6, 6], 2), ([6, 6, 6], 3), ([5, 5, 5], 4), ([12, 12], 3)]


@pytest.mark.parametrize("n,p", MIXED_CASES)
def test_tensor_acoustic_mixed_within_1e6(n, p):
    print(f"[case] tensor acoustic mixed n={n} p={p}", flush=True)
    mesh = Mesh.structured(n)
    rng = np.random.default_rng(2)
    cinv = rng.uniform(0.25, 4.0, size=mesh.n_elements)
    modes = [[1, 2, 1], [2, 1, 3]] if len(n) == 3 else [[1, 2], [2, 1]]
    runs = {}
    for prec in ("fp64", "mixed"):
        dev = HpgemAPIDevice(mesh, p, precision=prec, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4,
                             n_extra_vectors=2)
        dev.setInitialSolutionModes([0.7, -0.4], modes, field=0)
        dev.computeOneTimeStep(0.0, 2e-3, 20)
        runs[prec] = (dev, dev.getTimeIntegrationVector(0))
    dev64, u64 = runs["fp64"]
    dev64.setTimeIntegrationVector(5, runs["mixed"][1])
    l2d, _ = dev64.difference(0, 5)
    dev64.setTimeIntegrationVector(6, np.zeros_like(u64))
    l2, _ = dev64.difference(0, 6)
    print(f"acoustic mixed vs fp64 n={n} p={p} rel L2 {l2d / l2:.3e}")
    assert l2d / l2 < 1e-6
    for d, _ in runs.values():
        d.close()


@pytest.mark.parametrize("name", ["3Dtriangular2mesh.hpgem", "unitCubeN4Tet.hpgem"])
def test_tet_acoustic_on_reference_mesh_file(name):
    """Tetrahedral meshes read from the reference's .hpgem files, random
    per-element orders 1..4, walls on every boundary face."""
    import os
    print(f"[case] tet acoustic mesh file {name}", flush=True)
    mesh = Mesh.read(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "meshes", name))
    rng = np.random.default_rng(11)
    p = rng.integers(1, 5, size=mesh.n_elements).astype(np.int32)
    cinv = 1.0 / rng.uniform(0.5, 2.0, size=mesh.n_elements) ** 2
    dev = HpgemAPIDevice(mesh, p, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4)
    ref = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, n_fields=4)
    ref.set_acoustic(cinv)
    ref_ld = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, n_fields=4, extended=True)
    ref_ld.set_acoustic(cinv)
    assert dev.n_faces == len(mesh.face_elements)
    assert dev.n_faces - dev.n_internal_faces == int(np.sum(mesh.face_elements[:, 1] < 0))
    u = rng.standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(0, u)
    dev.computeTimeDerivative([0], [1.0], 1)
    assert_parity(dev.getTimeIntegrationVector(1), ref.time_derivative([u], [1.0]), ref_ld.time_derivative([u], [1.0]),
                  int(p.max()), ref, f"tet acoustic mesh file {name}")
    dev.close()


def test_tet_mixed_rhs_single_eval():
    """One mixed-precision tet RHS (FP32 volume/lift arithmetic and FP32
    stage derivative, FP64 stage input and jumps) against the FP64 path."""
    mesh = Mesh.structured([3, 3, 3], periodic=[0, 0, 0], simplices=True)
    rng = np.random.default_rng(21)
    p = rng.integers(1, 5, size=mesh.n_elements).astype(np.int32)
    cinv = 1.0 / rng.uniform(0.5, 2.0, size=mesh.n_elements) ** 2
    devs = {prec: HpgemAPIDevice(mesh, p, precision=prec, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4,
                                 n_extra_vectors=2) for prec in ("fp64", "mixed")}
    u = rng.standard_normal(devs["fp64"].n_dof)
    k = rng.standard_normal(devs["fp64"].n_dof)
    out = {}
    for prec, d in devs.items():
        if prec == "mixed":
            d.setStorage(1, "mixed")
            d.setStorage(2, "mixed")
        d.setTimeIntegrationVector(0, u)
        d.setTimeIntegrationVector(1, k)
        d.computeTimeDerivative([0, 1], [1.0, 0.01], 2)
        out[prec] = d.getTimeIntegrationVector(2)
    k32 = np.float32(k).astype(np.float64)  # the mixed path reads k in FP32
    devs["fp64"].setTimeIntegrationVector(1, k32)
    devs["fp64"].computeTimeDerivative([0, 1], [1.0, 0.01], 3)
    ref = devs["fp64"].getTimeIntegrationVector(3)
    err = rel(out["mixed"], ref)
    print(f"tet mixed single RHS rel err {err:.3e}")
    assert err < 2e-6
    for d in devs.values():
        d.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_tet_threads_per_element_give_bitwise_identical_rhs(precision):
    """Every threads-per-element candidate of every order (the timed choice
    and both fixed ones) computes each row and item with the same
    instructions, so the RHS is bitwise the same; the timed choice is one of
    the candidates and every order present was timed."""
    mesh = Mesh.structured([3, 3, 2], periodic=[0, 0, 0], simplices=True)
    rng = np.random.default_rng(31)
    p = rng.integers(1, 5, size=mesh.n_elements).astype(np.int32)
    cinv = 1.0 / rng.uniform(0.5, 2.0, size=mesh.n_elements) ** 2
    dev = HpgemAPIDevice(mesh, p, precision=precision, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4,
                         n_extra_vectors=2)
    cand = np.array([[8, 16], [12, 32], [32, 64], [35, 64]], dtype=np.int32)
    threads, ms = dev.getSimplexThreads()
    print(f"tuned threads {threads}, ms {ms.tolist()}")
    for g in range(4):
        assert threads[g] in cand[g]
        if np.any(p == g + 1):
            assert np.all(ms[g] > 0)
    if precision == "mixed":
        dev.setStorage(1, "mixed")
        dev.setStorage(2, "mixed")
    dev.setTimeIntegrationVector(0, rng.standard_normal(dev.n_dof))
    dev.setTimeIntegrationVector(1, rng.standard_normal(dev.n_dof))
    outs = []
    for combo in range(16):
        t = np.array([cand[g][(combo >> g) & 1] for g in range(4)], dtype=np.int32)
        dev.setSimplexThreads(t)
        np.testing.assert_array_equal(dev.getSimplexThreads()[0], t)
        dev.setTimeIntegrationVector(2, np.full(dev.n_dof, np.nan))
        dev.computeTimeDerivative([0, 1], [1.0, 0.25], 2)
        outs.append(dev.getTimeIntegrationVector(2))
    assert np.all(np.isfinite(outs[0]))
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])
    dev.setSimplexThreads([0, 0, 0, 0])
    assert all(dev.getSimplexThreads()[0][g] in cand[g] for g in range(4))
    with pytest.raises(Exception):
        dev.setSimplexThreads([24, 0, 0, 0])
    dev.close()


def test_tet_alternative_threads_match_oracle():
    """The larger threads-per-element candidate of every order against the
    oracle (the default candidates are covered by the other parity tests)."""
    print("[case] tet acoustic n=[3, 2, 2] p=random threads=[16, 32, 64, 64]", flush=True)
    dev, (ref, ref_ld) = make([3, 2, 2], "random", 41)
    dev.setSimplexThreads([16, 32, 64, 64])
    rng = np.random.default_rng(42)
    u = rng.standard_normal(dev.n_dof)
    k = rng.standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(0, u)
    dev.setTimeIntegrationVector(1, k)
    dev.computeTimeDerivative([0, 1], [1.0, -0.2], 2)
    assert_parity(dev.getTimeIntegrationVector(2), ref.time_derivative([u, k], [1.0, -0.2]),
                  ref_ld.time_derivative([u, k], [1.0, -0.2]), int(ref.orders.max()), ref,
                  "tet acoustic n=[3, 2, 2] threads=[16, 32, 64, 64]")
    dev.close()


def test_tet_mixed_within_1e6_after_rk4():
    """hp tets (p in 1..4), RK4 with FP32 stage derivatives: relative L2
    distance to the FP64 run stays below 1e-6."""
    mesh = Mesh.structured([4, 4, 4], periodic=[0, 0, 0], simplices=True)
    rng = np.random.default_rng(5)
    p = rng.integers(1, 5, size=mesh.n_elements).astype(np.int32)
    c = rng.uniform(0.5, 2.0, size=mesh.n_elements)
    cinv = 1.0 / c ** 2
    h = 0.25
    dt = 0.05 * h / (c.max() * 9)
    finals = {}
    devs = {}
    for prec in ("fp64", "mixed"):
        d = HpgemAPIDevice(mesh, p, precision=prec, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4,
                           n_extra_vectors=2)
        d.setInitialSolutionModes([0.8, -0.5], [[1, 1, 2], [2, 1, 1]], field=0)
        d.computeOneTimeStep(0.0, dt, 40)
        finals[prec] = d.getTimeIntegrationVector(0)
        devs[prec] = d
    d64 = devs["fp64"]
    d64.setTimeIntegrationVector(5, finals["mixed"])
    l2d, mx = d64.difference(0, 5)
    d64.setTimeIntegrationVector(6, np.zeros(d64.n_dof))
    l2, _ = d64.difference(0, 6)
    print(f"tet mixed vs fp64 after 40 RK4 steps: rel L2 {l2d / l2:.3e}, max {mx:.3e}")
    assert l2 > 0 and l2d / l2 < 1e-6
    for d in devs.values():
        d.close()


def test_tet_projection_is_the_l2_projection():
    """Sine modes projected into the pressure-like field of an hp tet mesh
    (on the device, with hpGEM's tetrahedral rule of order 2p+1) equal the
    oracle's projection, and the other fields stay zero."""
    mesh = Mesh.structured([3, 3, 3], periodic=[0, 0, 0], simplices=True)
    rng = np.random.default_rng(8)
    p = rng.integers(1, 5, size=mesh.n_elements).astype(np.int32)
    cinv = np.ones(mesh.n_elements)
    dev = HpgemAPIDevice(mesh, p, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4)
    amp, k = [0.7, -0.4], [[1, 0, 1], [0, 1, 1]]
    dev.setInitialSolutionModes(amp, k, field=0)
    u = dev.getTimeIntegrationVector(0)
    ref = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, n_fields=4)
    ref.set_acoustic(cinv)
    u_ref = ref.project_modes(amp, k, field_scale=[1.0, 0.0, 0.0, 0.0])
    e_dev = ref.l2_error_modes(u, amp, k, 0.0)
    e_ref = ref.l2_error_modes(u_ref, amp, k, 0.0)
    print(f"projection L2 error: library {e_dev:.6e}, oracle {e_ref:.6e}; coefficients {rel(u, u_ref):.2e}")
    assert rel(u, u_ref) < 1e-12
    assert abs(e_dev - e_ref) < 1e-10 * e_ref
    nb = (p + 1) * (p + 2) * (p + 3) // 6
    off = np.concatenate([[0], np.cumsum(4 * nb)])
    for e in range(mesh.n_elements):
        assert np.all(u[off[e] + nb[e]: off[e + 1]] == 0.0)
    dev.close()


@pytest.mark.parametrize("simplices", [False, True])
def test_gaussian_pulse_projection_keeps_its_norm(simplices):
    """||Pi f||^2 of a resolved pulse f = A exp(-|x-c|^2/w^2) inside the box
    equals int f^2 = A^2 (pi w^2 / 2)^(3/2) to the projection error."""
    n = [8, 8, 8] if simplices else [10, 10, 10]
    mesh = Mesh.structured(n, periodic=[0, 0, 0], simplices=simplices)
    cinv = np.ones(mesh.n_elements)
    dev = HpgemAPIDevice(mesh, 3, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4, n_extra_vectors=1)
    A, w, c = 1.7, 0.15, [0.45, 0.52, 0.5]
    dev.setInitialSolutionGaussian(c, w, A, field=0)
    l2, _ = dev.difference(0, dev.n_vectors - 1)  # the extra vector is zero
    exact = A ** 2 * (np.pi * w * w / 2) ** 1.5
    print(f"||Pi f||^2 = {l2 ** 2:.6e}, int f^2 = {exact:.6e}")
    assert abs(l2 ** 2 - exact) < 2e-3 * exact
    dev.close()


def test_tet_error_norm_matches_oracle_to_quadrature_accuracy():
    """computeTotalError on an hp tet mesh (on the device, hpGEM's
    tetrahedral rule of order 2p+2) against the oracle's."""
    mesh = Mesh.structured([3, 3, 3], periodic=[0, 0, 0], simplices=True)
    rng = np.random.default_rng(9)
    p = rng.integers(1, 5, size=mesh.n_elements).astype(np.int32)
    cinv = np.ones(mesh.n_elements)
    dev = HpgemAPIDevice(mesh, p, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4)
    amp, k = [0.7, -0.4], [[1, 0, 1], [0, 1, 1]]
    dev.setInitialSolutionModes(amp, k, field=0)
    u = dev.getTimeIntegrationVector(0)
    ref = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, n_fields=4)
    l2, mx = dev.computeErrorModes(amp, k, field=0)
    e_ref = ref.l2_error_modes(u, amp, k, 0.0)
    print(f"tet L2 error: library {l2:.6e}, oracle {e_ref:.6e}; max {mx:.3e}")
    assert abs(l2 - e_ref) < 1e-12 * e_ref
    assert mx > 0
    dev.close()


@pytest.mark.parametrize("kind", ["tet", "cube"])
def test_right_hand_side_and_mass_solve_match_oracle(kind):
    """hgdg_right_hand_side (r before the inverse mass,
    HpgemAPILinear::computeRightHandSide) and hgdg_solve_mass
    (solveMassMatrixEquations) against the oracle's two halves of
    computeTimeDerivative; their composition equals the fused path."""
    print(f"[case] right-hand side and mass solve, acoustic {kind}", flush=True)
    if kind == "tet":
        dev, (ref, ref_ld) = make([2, 2, 2], "random", seed=17)
    else:
        mesh = Mesh.structured([3, 3, 3], periodic=[0, 0, 0])
        cinv = np.random.default_rng(3).uniform(0.25, 4.0, size=mesh.n_elements)
        dev = HpgemAPIDevice(mesh, 3, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=RK4, n_extra_vectors=2)
        ref = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, 3, n_fields=4)
        ref.set_acoustic(cinv)
        ref_ld = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, 3, n_fields=4, extended=True)
        ref_ld.set_acoustic(cinv)
    pm = int(ref.orders.max())
    rng = np.random.default_rng(23)
    u = rng.standard_normal(dev.n_dof)
    k = rng.standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(0, u)
    dev.setTimeIntegrationVector(1, k)
    dev.computeRightHandSide([0, 1], [1.0, 0.25], 2)
    r = dev.getTimeIntegrationVector(2)
    assert_parity(r, ref.right_hand_side([u, k], [1.0, 0.25]), ref_ld.right_hand_side([u, k], [1.0, 0.25]), pm, ref,
                  f"acoustic {kind} right-hand side")
    # mass solve of an arbitrary vector
    dev.setTimeIntegrationVector(3, k)
    dev.solveMassMatrixEquations(3)
    assert_parity(dev.getTimeIntegrationVector(3), ref.solve_mass(k), ref_ld.solve_mass(k), pm, ref,
                  f"acoustic {kind} mass solve")
    # r then M^{-1} = the fused time derivative
    dev.solveMassMatrixEquations(2)
    dev.computeTimeDerivative([0, 1], [1.0, 0.25], 4)
    assert rel(dev.getTimeIntegrationVector(2), dev.getTimeIntegrationVector(4)) < 1e-12
    dev.close()


FUSED_CASES = [
    # (cells, p, periodic, precision, tableau, steps): 3D four-lane kernel at p = 1..4, 2D four-lane (p >= 2)
    # and one-thread-per-pencil (p = 1) kernels; odd step counts end with one separate step
    ([6, 6, 6], 3, True, "fp64", "RK4", 6),
    ([5, 5, 5], 4, False, "mixed", "RK4", 7),
    ([6, 5, 4], 1, False, "fp64", "RK3_TVD", 5),
    ([6, 6, 6], 2, True, "mixed", "DORMAND_PRINCE", 4),  # last stage has b = 0
    ([20, 16], 1, True, "fp64", "RK4", 4),
    ([16, 12], 3, False, "mixed", "RK4", 2),
]


@pytest.mark.parametrize("n,p,periodic,precision,tab,steps", FUSED_CASES)
def test_acoustic_fused_update_steps_bitwise_equal(n, p, periodic, precision, tab, steps):
    """Pairs of RK steps with the update fused into each step's last stage
    perform the separate update kernel's FP64 operations: the solution and
    every stage vector come out bitwise the same."""
    import hpgem_amd.tableaux
    print(f"[case] acoustic fused update n={n} p={p} {precision} {tab} steps={steps}", flush=True)
    tableau = getattr(hpgem_amd.tableaux, tab)
    mesh = Mesh.structured(n, periodic=[int(periodic)] * len(n))
    rng = np.random.default_rng(p + 7 * len(n))
    cinv = rng.uniform(0.25, 4.0, size=mesh.n_elements)
    u0 = None
    out = {}
    for fused in (False, True):
        dev = HpgemAPIDevice(mesh, p, precision=precision, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=tableau,
                             n_extra_vectors=2)
        dev.setFusedTimeStepping(fused)
        if u0 is None:
            u0 = rng.standard_normal(dev.n_dof)
        dev.setTimeIntegrationVector(0, u0)
        dt = 0.02 / max(n)
        dev.computeOneTimeStep(0.0, dt, steps)
        dev.computeOneTimeStep(0.0, dt, 3)  # the same graphs again, odd count
        out[fused] = [dev.getTimeIntegrationVector(v) for v in range(tableau.stages + 1)]
        dev.close()
    assert np.all(np.isfinite(out[False][0]))
    assert not np.array_equal(out[False][0], u0)
    for v in range(tableau.stages + 1):
        np.testing.assert_array_equal(out[True][v], out[False][v])


@pytest.mark.parametrize("precision,tab,steps", [("fp64", "RK4", 5), ("mixed", "RK4", 4),
                                                 ("fp64", "DORMAND_PRINCE", 4)])
def test_tet_fused_update_steps_bitwise_equal(precision, tab, steps):
    """The tetrahedral kernel's fused RK update (hp orders 1..4, walls):
    bitwise the same as the separate update kernel."""
    import hpgem_amd.tableaux
    print(f"[case] tet fused update {precision} {tab} steps={steps}", flush=True)
    tableau = getattr(hpgem_amd.tableaux, tab)
    mesh = Mesh.structured([4, 3, 3], periodic=[0, 0, 0], simplices=True)
    rng = np.random.default_rng(11)
    orders = rng.integers(1, 5, size=mesh.n_elements).astype(np.int32)
    cinv = rng.uniform(0.25, 4.0, size=mesh.n_elements)
    u0 = None
    out = {}
    for fused in (False, True):
        dev = HpgemAPIDevice(mesh, orders, precision=precision, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv,
                             tableau=tableau, n_extra_vectors=2)
        dev.setFusedTimeStepping(fused)
        if u0 is None:
            u0 = rng.standard_normal(dev.n_dof)
        dev.setTimeIntegrationVector(0, u0)
        dev.computeOneTimeStep(0.0, 0.004, steps)
        dev.computeOneTimeStep(0.0, 0.004, 3)
        out[fused] = [dev.getTimeIntegrationVector(v) for v in range(tableau.stages + 1)]
        dev.close()
    assert np.all(np.isfinite(out[False][0]))
    assert not np.array_equal(out[False][0], u0)
    for v in range(tableau.stages + 1):
        np.testing.assert_array_equal(out[True][v], out[False][v])

"""GPU parity of the Gauss-point kernels (tensor_general.hip) against the
oracle: quadrilateral/hexahedral meshes whose elements are relabelled by
rotations (neighbours meet through other local faces, in every orientation of
the face, including the quarter turns where hpGEM pairs face points by the
opposite quarter turn), whose elements are bilinear/trilinear (a smooth
periodic displacement of the vertices), or both; and structured meshes forced
onto these kernels, which must agree with the sum-factorised kernels.

FP64 bar as in tests/_parity.py: <= 1e-12 relative to the long-double
oracle; to the double oracle <= 1e-12 up to p = 2, and at p >= 3 <= 1e-12 in
the mass-weighted L2 norm (hpGEM's per-call LU of ill-conditioned tensor
mass matrices).
"""
import numpy as np
import pytest

import _oracle as orc
import hpgem_amd
from _meshes import perturb, rotate_elements
from _parity import assert_parity, rel
from _exact import advection_rate_window
from hpgem_amd import OP_ACOUSTIC_CENTRAL, HpgemAPIDevice, Mesh, RK4

pytestmark = pytest.mark.gpu

VEL = [1.0, 0.7, 0.3]


def variant_mesh(n, variant, periodic=True, seed=5, quarter_turns=True):
    mesh = Mesh.structured(n, periodic=[int(periodic)] * len(n))
    X, nodes = mesh.coords, mesh.node_ids
    if variant in ("perturbed", "both"):
        X = perturb(X)
    if variant in ("rotated", "both"):
        X, nodes, _ = rotate_elements(X, nodes, seed, quarter_turns)
    return Mesh(mesh.shape, X, nodes)


def device_and_oracles(mesh, p, kind, cinv=None, precision="fp64", tableau=RK4, force=False):
    dim = mesh.dim
    nf = 1 if kind == "advection" else dim + 1
    if kind == "advection":
        dev = HpgemAPIDevice(mesh, p, velocity=VEL[:dim], precision=precision, tableau=tableau, n_extra_vectors=2,
                             tensor_general=force)
    else:
        dev = HpgemAPIDevice(mesh, p, precision=precision, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=tableau,
                             n_extra_vectors=2, tensor_general=force)
    refs = []
    for ext in (False, True):
        P = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, n_fields=nf, extended=ext)
        if kind == "advection":
            P.set_advection(VEL[:dim])
        else:
            P.set_acoustic(cinv)
        refs.append(P)
    return dev, refs[0], refs[1]


CASES = [
    # (cells, p, kind, variant, periodic)
    ([4, 3], 1, "advection", "both", True),
    ([4, 3], 2, "advection", "rotated", True),
    ([3, 4], 3, "advection", "perturbed", True),
    ([4, 3], 4, "advection", "both", True),
    ([3, 3], 5, "advection", "both", True),
    ([4, 4], 2, "advection", "both", False),     # inflow / outflow boundary faces
    ([4, 3], 2, "acoustic", "both", True),
    ([3, 4], 3, "acoustic", "both", False),      # reflecting walls
    ([3, 3, 3], 1, "advection", "both", True),
    ([3, 3, 3], 2, "advection", "rotated", True),
    ([3, 3, 3], 2, "advection", "perturbed", True),
    ([3, 3, 3], 3, "advection", "both", True),
    ([3, 3, 3], 4, "advection", "both", True),
    ([3, 3, 4], 2, "advection", "both", False),
    ([3, 3, 3], 1, "acoustic", "both", True),
    ([3, 3, 3], 2, "acoustic", "both", True),
    ([3, 3, 3], 3, "acoustic", "both", False),
    ([4, 3], 3, "advection", "structured", True),  # forced onto the Gauss-point kernels
    ([3, 3, 3], 2, "acoustic", "structured", True),
]


@pytest.mark.parametrize("n,p,kind,variant,periodic", CASES)
def test_time_derivative_matches_oracle(n, p, kind, variant, periodic):
    label = f"general {kind} n={n} p={p} {variant} periodic={periodic}"
    print(f"[case] {label}", flush=True)
    mesh = variant_mesh(n, variant, periodic)
    cinv = np.linspace(0.5, 2.0, mesh.n_elements)
    dev, ref, ref_ld = device_and_oracles(mesh, p, kind, cinv, force=(variant == "structured"))
    assert dev.usesTensorGeneral()
    rng = np.random.default_rng(11)
    u, k = rng.standard_normal(dev.n_dof), rng.standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(0, u)
    dev.setTimeIntegrationVector(2, k)
    dev.computeTimeDerivative([0, 2], [1.0, -0.3], 1)
    got = dev.getTimeIntegrationVector(1)
    assert_parity(got, ref.time_derivative([u, k], [1.0, -0.3]), ref_ld.time_derivative([u, k], [1.0, -0.3]), p,
                  ref, label)
    dev.close()


@pytest.mark.parametrize("n,p,kind", [([4, 3], 3, "advection"), ([3, 3, 3], 2, "advection"),
                                      ([3, 3, 3], 3, "acoustic")])
def test_forced_gauss_point_kernels_agree_with_sum_factorised(n, p, kind):
    """The same structured mesh on both kernel families: hpGEM's operator both
    times, to rounding."""
    print(f"[case] forced vs sum-factorised {kind} n={n} p={p}", flush=True)
    mesh = Mesh.structured(n)
    cinv = np.linspace(0.5, 2.0, mesh.n_elements)
    out = []
    for force in (False, True):
        dev, _, _ = device_and_oracles(mesh, p, kind, cinv, force=force)
        assert dev.usesTensorGeneral() == force
        u = np.random.default_rng(3).standard_normal(dev.n_dof)
        dev.setTimeIntegrationVector(0, u)
        dev.computeTimeDerivative([0], [1.0], 1)
        out.append(dev.getTimeIntegrationVector(1))
        dev.close()
    d = rel(out[1], out[0])
    print(f"  max rel difference {d:.2e}")
    assert d < 1e-12


def test_rk4_on_rotated_non_affine_hexes_matches_oracle():
    print("[case] general RK4 x5, 3^3 hex p=2, rotated + perturbed", flush=True)
    mesh = variant_mesh([3, 3, 3], "both")
    dev, ref, ref_ld = device_and_oracles(mesh, 2, "advection")
    amp, kv = [0.6, -0.3], [[1, 2, 1], [2, 1, 3]]
    u0 = ref.project_modes(amp, kv)
    dev.setTimeIntegrationVector(0, u0)
    dt = 0.01
    dev.computeOneTimeStep(0.0, dt, 5)
    assert_parity(dev.getTimeIntegrationVector(0), ref.rk4(u0, dt, 5), ref_ld.rk4(u0, dt, 5), 2, ref,
                  "general RK4 x5 hex p=2 both")
    dev.close()


def test_mixed_storage_stays_within_1e6_of_fp64():
    print("[case] general mixed storage, 3^3 hex p=3, 20 RK4 steps", flush=True)
    mesh = variant_mesh([3, 3, 3], "both")
    finals = []
    for prec in ("fp64", "mixed"):
        dev, _, _ = device_and_oracles(mesh, 3, "advection", precision=prec)
        dev.setInitialSolutionModes([0.6, -0.3], [[1, 2, 1], [2, 1, 3]])
        if prec == "mixed":
            for v in dev.auxiliaryVectorIds_:
                dev.setStorage(v, "mixed")
        dev.computeOneTimeStep(0.0, 0.005, 20)
        finals.append(dev.getTimeIntegrationVector(0))
        dev.close()
    d = np.linalg.norm(finals[1] - finals[0]) / np.linalg.norm(finals[0])
    print(f"  relative L2 (coefficients) {d:.2e}")
    assert d < 1e-6


@pytest.mark.parametrize("n,p", [([4, 3], 3), ([3, 3, 3], 2)])
def test_projection_and_error_norms_match_oracle(n, p):
    print(f"[case] general projection + norms n={n} p={p}", flush=True)
    mesh = variant_mesh(n, "both")
    dev, ref, ref_ld = device_and_oracles(mesh, p, "advection")
    dim = len(n)
    amp = [0.6, -0.3]
    kv = [[1, 2, 1][:dim], [2, 1, 3][:dim]]
    dev.setInitialSolutionModes(amp, kv)
    got = dev.getTimeIntegrationVector(0)
    assert_parity(got, ref.project_modes(amp, kv), ref_ld.project_modes(amp, kv), p, ref, f"general projection n={n}")
    # the error of a perturbed state against the modes, shifted
    v = got + 1e-3 * np.random.default_rng(2).standard_normal(got.size)
    dev.setTimeIntegrationVector(0, v)
    t = 0.013
    l2, _ = dev.computeErrorModes(amp, kv, shift=[a * t for a in VEL[:dim]])
    want = ref_ld.l2_error_modes(v, amp, kv, t)
    print(f"  L2 error {l2:.15e} / oracle {want:.15e}")
    assert abs(l2 - want) <= 1e-12 * want
    dev.close()


def test_right_hand_side_mass_solve_and_difference_norm_match_oracle():
    print("[case] general right-hand side, mass solve, difference norm, 3^3 hex p=3 acoustic", flush=True)
    mesh = variant_mesh([3, 3, 3], "both")
    cinv = np.linspace(0.5, 2.0, mesh.n_elements)
    dev, ref, ref_ld = device_and_oracles(mesh, 3, "acoustic", cinv)
    u = np.random.default_rng(4).standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(0, u)
    dev.computeRightHandSide([0], [1.0], 1)
    assert_parity(dev.getTimeIntegrationVector(1), ref.right_hand_side([u], [1.0]),
                  ref_ld.right_hand_side([u], [1.0]), 3, ref, "general right-hand side")
    dev.setTimeIntegrationVector(2, u)
    dev.solveMassMatrixEquations(2)
    assert_parity(dev.getTimeIntegrationVector(2), ref.solve_mass(u), ref_ld.solve_mass(u), 3, ref,
                  "general mass solve")
    # mass-weighted L2 of a difference: sum_e d_e^T M_e d_e with the plain (unweighted) mass
    w = np.random.default_rng(5).standard_normal(dev.n_dof)
    dev.setTimeIntegrationVector(3, w)
    l2, mx = dev.difference(0, 3)
    P1 = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, 3, n_fields=4)
    P1.set_acoustic(np.ones(mesh.n_elements))  # unit c_inv: the plain mass on every field
    from _parity import element_masses
    d = u - w
    want = np.sqrt(sum(d[o:o + M.shape[0]] @ M @ d[o:o + M.shape[0]] for o, M in element_masses(P1)))
    print(f"  difference norm {l2:.15e} / oracle masses {want:.15e}")
    assert abs(l2 - want) <= 1e-12 * want and mx == np.max(np.abs(d))
    dev.close()


def advection_run_error(mesh, p, t_end=0.1, cfl=0.05):
    """L2 error after RK4 to t_end of the projected modes, on the device."""
    dim, n = mesh.dim, round(mesh.n_elements ** (1 / mesh.dim))
    dev = HpgemAPIDevice(mesh, p, velocity=VEL[:dim], tableau=RK4)
    amp = np.array([1.0, -0.6])
    k = np.array([[1.0, 0.0, 1.0], [0.0, 1.0, -1.0]])[:, :dim]
    dev.setInitialSolutionModes(amp, k)
    steps = int(np.ceil(t_end / (cfl / n / np.linalg.norm(VEL[:dim]))))
    dev.computeOneTimeStep(0.0, t_end / steps, steps)
    l2, _ = dev.computeErrorModes(amp, k, shift=[a * t_end for a in VEL[:dim]])
    dev.close()
    return l2


@pytest.mark.parametrize("dim,ns,p", [(2, (8, 16), 2), (3, (4, 8), 2)])
def test_advection_converges_on_rotated_non_affine_meshes(dim, ns, p):
    """Every rotation of the square; for cubes the half turns (see below)."""
    print(f"[case] general convergence dim {dim} p={p}", flush=True)
    e = [advection_run_error(variant_mesh([n] * dim, "both", quarter_turns=(dim == 2)), p) for n in ns]
    rate = np.log2(e[0] / e[1])
    lo, hi = advection_rate_window(p)
    print(f"  L2 errors {e[0]:.4e} {e[1]:.4e}, rate {rate:.2f} (window {lo:.2f}..{hi:.2f})")
    assert lo <= rate <= hi


def test_quarter_turned_cube_faces_run_like_the_reference():
    """Hexahedra whose shared faces are quarter-turned against each other:
    hpGEM pairs the face points of such faces by the opposite quarter turn, so
    its discretisation there is not consistent (the oracle's L2 error on
    4^3 / 8^3 rotated, displaced meshes stays at 0.25 / 0.18). The device
    keeps hpGEM's pairing: the same run gives the oracle's error."""
    print("[case] general quarter-turned hexes, 4^3 p=2, RK4 run vs oracle", flush=True)
    mesh = variant_mesh([4, 4, 4], "both")
    got = advection_run_error(mesh, 2)
    P = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, 2)
    P.set_advection(VEL)
    amp = np.array([1.0, -0.6])
    k = np.array([[1.0, 0.0, 1.0], [0.0, 1.0, -1.0]])
    u = P.project_modes(amp, k)
    steps = int(np.ceil(0.1 / (0.05 / 4 / np.linalg.norm(VEL))))
    want = P.l2_error_modes(P.rk4(u, 0.1 / steps, steps), amp, k, 0.1)
    print(f"  L2 error {got:.12e} / oracle {want:.12e}")
    assert abs(got - want) <= 1e-9 * want and want > 0.1


GEN_FUSED_CASES = [
    # (cells, p, kind, variant, precision, tableau, steps) on the Gauss-point kernels
    ([3, 3, 3], 2, "advection", "both", "fp64", "RK4", 6),
    ([4, 3], 3, "acoustic", "rotated", "mixed", "RK4", 4),
    ([3, 3, 3], 1, "acoustic", "perturbed", "fp64", "DORMAND_PRINCE", 4),
]


@pytest.mark.parametrize("n,p,kind,variant,precision,tab,steps", GEN_FUSED_CASES)
def test_fused_update_steps_bitwise_equal_on_gauss_point_kernels(n, p, kind, variant, precision, tab, steps):
    """Pairs of RK steps with the update in each last-stage kernel, on
    rotated and non-affine meshes: bitwise the same as the separate update."""
    import hpgem_amd.tableaux
    print(f"[case] general fused update n={n} p={p} {kind} {variant} {precision} {tab}", flush=True)
    tableau = getattr(hpgem_amd.tableaux, tab)
    mesh = variant_mesh(n, variant)
    rng = np.random.default_rng(3)
    cinv = rng.uniform(0.25, 4.0, size=mesh.n_elements)
    u0 = None
    out = {}
    for fused in (False, True):
        if kind == "advection":
            dev = HpgemAPIDevice(mesh, p, velocity=VEL[:len(n)], precision=precision, tableau=tableau,
                                 n_extra_vectors=2)
        else:
            dev = HpgemAPIDevice(mesh, p, precision=precision, kind=OP_ACOUSTIC_CENTRAL, c_inv=cinv, tableau=tableau,
                                 n_extra_vectors=2)
        assert dev.usesTensorGeneral()
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

"""Pins the oracle on quadrilateral/hexahedral meshes that are not structured
boxes (no GPU): elements whose local vertex order is rotated against their
neighbours' (hpGEM's face search, orientation indices and maps,
MeshManipulator_Impl.h:1860-1950, ReferenceSquare.cpp:131-180) and elements
whose bilinear/trilinear map is not affine
(MappingToPhysHypercubeLinear.cpp:222-304).

* Relabelling an element's vertices by a rotation of the reference shape
  changes its basis functions but not the space they span, so the DG time
  derivative, as a function, must not change: checked at random points of
  every element, for all rotations of the square and the half turns of the
  cube. A wrong face map, orientation map or normal sign moves face points
  and fails this by O(1).
* hpGEM's ReferenceSquare picks, for the two quarter-turn orientations of a
  square face, the opposite quarter turn (tests/test_oracle_tables.py pins
  the maps and the index lookup against the reference-built fixture): on
  cubes rotated by quarter turns the function does change. The oracle keeps
  the reference's pairing; the check documents that it does.
* On a sequence of meshes displaced by a smooth periodic map (non-affine
  elements) the L2 error of upwind advection and of the reference's acoustic
  wave still converges at the order of the method: a wrong Jacobian, area
  element or point mapping of the non-affine map leaves an error that does
  not shrink.
"""
import numpy as np
import pytest

import _oracle as orc
from _exact import acoustic_rate_window, acoustic_wave, advection_rate_window
from _meshes import perturb, rotate_elements

AMP = [0.7, -0.4]


def _modes(dim):
    return [[1, 2, 1][:dim] + [0] * (3 - dim), [2, 1, 3][:dim] + [0] * (3 - dim)]


def _rhs_function_distance(dim, n, p, kind, quarter_turns, seed=3):
    """Max over random points of every element of |k_structured - k_rotated| /
    max |k|, with k the oracle's time derivative of the same projected
    state on the structured mesh and on its rotated relabelling."""
    shape, X, nodes = orc.structured_mesh(n)
    X2, N2, Rs = rotate_elements(X, nodes, seed, quarter_turns)
    nf = 1 if kind == "advection" else dim + 1
    res = []
    for XX, NN in ((X, nodes), (X2, N2)):
        P = orc.Problem(shape, XX, NN, p, n_fields=nf)
        if kind == "advection":
            P.set_advection([1.0, 0.7, 0.3][:dim])
            u = P.project_modes(AMP, _modes(dim))
        else:
            P.set_acoustic(np.linspace(0.5, 2.0, P.n_elem))
            u = P.project_modes(AMP, _modes(dim), np.arange(1, nf + 1, dtype=float))
        res.append((P, P.time_derivative([u], [1.0])))
    nb = len(orc.basis_eval(shape, p, np.zeros(dim))[0])

    def value(k, e, xi):
        phi = orc.basis_eval(shape, p, xi)[0]
        return np.array([k[(e * nf + f) * nb:(e * nf + f + 1) * nb] @ phi for f in range(nf)])

    rng = np.random.default_rng(seed + 1)
    err = scale = 0.0
    for e in range(len(X)):
        for xi in rng.uniform(-1, 1, (3, dim)):
            a = value(res[0][1], e, xi)
            b = value(res[1][1], e, Rs[e].T @ xi)  # the same physical point
            err = max(err, float(np.max(np.abs(a - b))))
            scale = max(scale, float(np.max(np.abs(a))))
    return err / scale


@pytest.mark.parametrize("dim,n,p,kind", [(2, [4, 3], 3, "advection"), (2, [3, 4], 2, "acoustic"),
                                          (3, [3, 3, 3], 2, "advection"), (3, [3, 3, 3], 2, "acoustic")])
def test_time_derivative_does_not_depend_on_the_vertex_order(dim, n, p, kind):
    # squares: every rotation (line faces meet reversed or not); cubes: half turns
    d = _rhs_function_distance(dim, n, p, kind, quarter_turns=(dim == 2))
    print(f"dim {dim} {kind} p={p}: {d:.2e}")
    assert d < 1e-12


def test_quarter_turned_cube_faces_keep_the_reference_pairing():
    d = _rhs_function_distance(3, [3, 3, 3], 2, "advection", quarter_turns=True)
    print(f"quarter turns: {d:.2e}")
    assert d > 1e-2


def _perturbed(n, dim, eps=0.04):
    shape, X, nodes = orc.structured_mesh([n] * dim)
    return shape, perturb(X, eps), nodes


def _advection_error_perturbed(dim, n, p, a=(1.0, 0.7, 0.3), t_end=0.1, cfl=0.05):
    shape, X, nodes = _perturbed(n, dim)
    P = orc.Problem(shape, X, nodes, p)
    P.set_advection(list(a[:dim]))
    amp = np.array([1.0, -0.6])
    k = np.array([[1.0, 0.0, 1.0], [0.0, 1.0, -1.0]])[:, :dim]
    u = P.project_modes(amp, k)
    speed = np.linalg.norm(a[:dim])
    steps = int(np.ceil(t_end / (cfl / n / speed)))
    u = P.rk4(u, t_end / steps, steps)
    return P.l2_error_modes(u, amp, k, t_end)


@pytest.mark.parametrize("dim,ns,p", [(2, (8, 16), 2), (2, (8, 16), 3), (3, (4, 8), 2)])
def test_advection_converges_on_non_affine_meshes(dim, ns, p):
    e = [_advection_error_perturbed(dim, n, p) for n in ns]
    rate = np.log2(e[0] / e[1])
    lo, hi = advection_rate_window(p)
    print(f"non-affine advection dim {dim} p={p}: L2 errors {e[0]:.4e} {e[1]:.4e}, rate {rate:.2f}")
    assert lo <= rate <= hi, (e, rate)


def _acoustic_error_perturbed(n, p, cinv, t_frac=0.25, cfl=0.05):
    dim = 2
    shape, X, nodes = _perturbed(n, dim)
    P = orc.Problem(shape, X, nodes, p, n_fields=dim + 1)
    P.set_acoustic(np.full(P.n_elem, cinv))
    amp, k, _ = acoustic_wave(dim, cinv, 0.0)
    u = np.zeros(P.n_dof)
    for v in range(dim + 1):
        fs = np.zeros(dim + 1)
        fs[v] = 1.0
        u += P.project_modes(amp[v], k, fs)
    c = 1 / np.sqrt(cinv)
    T = t_frac / (np.sqrt(dim) * c)
    steps = int(np.ceil(T / (cfl / n / c)))
    u = P.rk4(u, T / steps, steps)
    amp, k, sh = acoustic_wave(dim, cinv, T)
    return float(np.sqrt(P.l2_error_fields(u, amp, k, sh).sum()))


@pytest.mark.parametrize("p,cinv", [(2, 4.0), (3, 0.25)])
def test_acoustic_wave_converges_on_non_affine_quads(p, cinv):
    e = [_acoustic_error_perturbed(n, p, cinv) for n in (8, 16)]
    rate = np.log2(e[0] / e[1])
    lo, hi = acoustic_rate_window(p)
    print(f"non-affine acoustic p={p} cInv={cinv}: L2 errors {e[0]:.4e} {e[1]:.4e}, rate {rate:.2f}")
    assert lo <= rate <= hi, (e, rate)

"""Pins the oracle's operators by convergence to exact solutions (no GPU).

* Acoustic system: the reference's own exact travelling wave
  (applications/ExampleMultipleVariableProblem/AcousticWaveLinear_Impl.h:60-90,
  see _exact.py) at cInv = 0.25, 1, 4, on squares, cubes and tetrahedra:
  project, run RK4 to a fixed time with a small step, take the L2 error of
  every field, and require the error ratio under mesh halving to lie in the
  window of the method's order. A coupling of the wrong sign sends the wave
  the other way, and cInv on the wrong fields changes the amplitude relation
  between v and s at cInv != 1: either leaves an O(1) error that does not
  shrink with h.
* Upwind advection at p = 2, 3 on cubes, in the style of
  tests/self/ConvergenceTest.h:56-86 (successive meshes, error ratios).
"""
import numpy as np
import pytest

import _oracle as orc
from _exact import acoustic_rate_window, acoustic_wave, acoustic_wave_value, advection_rate_window, eval_modes


def test_wave_decomposition_matches_reference_formula():
    rng = np.random.default_rng(5)
    for dim in (2, 3):
        for cinv in (0.25, 1.0, 4.0):
            for t in (0.0, 0.13):
                amp, k, sh = acoustic_wave(dim, cinv, t)
                for x in rng.random((6, dim)):
                    want = acoustic_wave_value(dim, cinv, t, x)
                    got = [eval_modes(amp[v], k, sh, x) for v in range(dim + 1)]
                    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.max(np.abs(amp)))


def test_wave_solves_the_acoustic_system():
    # v_t = div s and cInv s_t = grad v, by central differences of the formula
    dim, cinv, t, h = 3, 4.0, 0.07, 1e-5
    x = np.array([0.31, 0.12, 0.77])
    f = lambda tt, xx: acoustic_wave_value(dim, cinv, tt, xx)
    dt_ = (f(t + h, x) - f(t - h, x)) / (2 * h)
    grad = np.zeros((dim, dim + 1))
    for d in range(dim):
        e = np.zeros(dim)
        e[d] = h
        grad[d] = (f(t, x + e) - f(t, x - e)) / (2 * h)
    div_s = sum(grad[d, 1 + d] for d in range(dim))
    scale = np.max(np.abs(grad))
    assert abs(dt_[0] - div_s) < 1e-6 * scale
    for d in range(dim):
        assert abs(cinv * dt_[1 + d] - grad[d, 0]) < 1e-6 * scale


def _acoustic_error(kind, n, p, cinv, t_frac=0.25, cfl=0.05):
    dim = 2 if kind == "square" else 3
    shape, X, nodes = orc.structured_mesh([n] * dim, simplices=(kind == "tet"))
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


ACOUSTIC_CASES = [("square", p, cinv, (8, 16)) for p in (1, 2, 3) for cinv in (0.25, 1.0, 4.0)]
ACOUSTIC_CASES += [("cube", 1, cinv, (4, 8)) for cinv in (0.25, 4.0)]
ACOUSTIC_CASES += [("tet", 1, cinv, (4, 8)) for cinv in (0.25, 1.0, 4.0)]
ACOUSTIC_CASES += [("tet", 2, 4.0, (4, 8))]


@pytest.mark.parametrize("kind,p,cinv,ns", ACOUSTIC_CASES)
def test_acoustic_oracle_converges_to_reference_wave(kind, p, cinv, ns):
    e = [_acoustic_error(kind, n, p, cinv) for n in ns]
    rate = np.log2(e[0] / e[1])
    lo, hi = acoustic_rate_window(p)
    print(f"{kind} p={p} cInv={cinv}: L2 errors {e[0]:.4e} {e[1]:.4e}, rate {rate:.2f}")
    assert lo <= rate <= hi, (e, rate)


def _advection_error(n, p, a=(1.0, 0.7, 0.3), t_end=0.1, cfl=0.05):
    shape, X, nodes = orc.structured_mesh([n, n, n])
    P = orc.Problem(shape, X, nodes, p)
    P.set_advection(list(a))
    amp = np.array([1.0, -0.6])
    k = np.array([[1.0, 0.0, 1.0], [0.0, 1.0, -1.0]])
    u = P.project_modes(amp, k)
    steps = int(np.ceil(t_end / (cfl / n / np.linalg.norm(a))))
    u = P.rk4(u, t_end / steps, steps)
    return P.l2_error_modes(u, amp, k, t_end)


@pytest.mark.parametrize("p", [2, 3])
def test_advection_oracle_converges_at_optimal_order(p):
    e = [_advection_error(n, p) for n in (4, 8)]
    rate = np.log2(e[0] / e[1])
    lo, hi = advection_rate_window(p)
    print(f"advection p={p}: L2 errors {e[0]:.4e} {e[1]:.4e}, rate {rate:.2f}")
    assert lo <= rate <= hi, (e, rate)

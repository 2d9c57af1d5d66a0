"""The GPU path against exact solutions, at sizes beyond the oracle's reach.

* Acoustic system, the reference's own exact travelling wave
  (applications/ExampleMultipleVariableProblem/AcousticWaveLinear_Impl.h:60-90,
  tests/_exact.py) at cInv = 0.25, 1, 4 on periodic cubes (p = 1..3) and
  periodic tetrahedra (p = 1..3): projected on the device, RK4 on the device,
  error of every field from the library's own computeTotalError. The error
  ratio under mesh halving must lie in the window of the method's order; at
  the coarse level the GPU's error must equal the oracle's (same initial
  projection, same RK4 run) to 1e-9 relative.
* Upwind advection at p = 2, 3, 4 on periodic cubes (tests/self/
  ConvergenceTest.h:56-86 style: successive meshes, error ratios).
"""
import numpy as np
import pytest

import _oracle as orc
from _exact import acoustic_rate_window, acoustic_wave, advection_rate_window
from hpgem_amd import OP_ACOUSTIC_CENTRAL, RK4, HpgemAPIDevice, Mesh

pytestmark = pytest.mark.gpu


def _acoustic_setup(kind, n, p, cinv, t_frac=0.25, cfl=0.05):
    dim = 2 if kind == "square" else 3
    mesh = Mesh.structured([n] * dim, simplices=(kind == "tet"))
    c = 1 / np.sqrt(cinv)
    T = t_frac / (np.sqrt(dim) * c)
    steps = int(np.ceil(T / (cfl / n / c)))
    return mesh, dim, T, steps


def gpu_acoustic_error(kind, n, p, cinv):
    mesh, dim, T, steps = _acoustic_setup(kind, n, p, cinv)
    dev = HpgemAPIDevice(mesh, p, kind=OP_ACOUSTIC_CENTRAL, c_inv=np.full(mesh.n_elements, cinv), tableau=RK4)
    amp, k, _ = acoustic_wave(dim, cinv, 0.0)
    for v in range(dim + 1):
        dev.setInitialSolutionModes(amp[v], k, field=v)
    u0 = dev.getTimeIntegrationVector(0)
    dev.computeOneTimeStep(0.0, T / steps, steps)
    amp, k, sh = acoustic_wave(dim, cinv, T)
    e2 = sum(dev.computeErrorModes(amp[v], k, shift=sh, field=v)[0] ** 2 for v in range(dim + 1))
    u = dev.getTimeIntegrationVector(0)
    dev.close()
    return float(np.sqrt(e2)), mesh, u0, u, T, steps


ACOUSTIC_CASES = [("cube", p, cinv, (8, 16)) for p in (1, 2, 3) for cinv in (0.25, 1.0, 4.0)]
ACOUSTIC_CASES += [("tet", p, cinv, (4, 8)) for p in (1, 2, 3) for cinv in (0.25, 1.0, 4.0)]


@pytest.mark.parametrize("kind,p,cinv,ns", ACOUSTIC_CASES)
def test_acoustic_gpu_converges_to_reference_wave(kind, p, cinv, ns):
    print(f"[case] acoustic exact wave {kind} p={p} cInv={cinv} n={ns}", flush=True)
    runs = [gpu_acoustic_error(kind, n, p, cinv) for n in ns]
    e = [r[0] for r in runs]
    rate = np.log2(e[0] / e[1])
    lo, hi = acoustic_rate_window(p)
    print(f"  L2 errors {e[0]:.4e} {e[1]:.4e}, rate {rate:.2f} (window {lo:.2f}..{hi:.2f})", flush=True)
    assert lo <= rate <= hi, (e, rate)


@pytest.mark.parametrize("kind,n,p,cinv", [("cube", 4, 2, 4.0), ("tet", 4, 2, 0.25), ("square", 8, 3, 0.25)])
def test_acoustic_gpu_error_equals_oracle_error(kind, n, p, cinv):
    """Same projected wave, same RK4 run, same error integral: the GPU's
    result and the oracle's give the same L2 error."""
    print(f"[case] acoustic exact wave vs oracle {kind} n={n} p={p} cInv={cinv}", flush=True)
    e_gpu, mesh, u0, u, T, steps = gpu_acoustic_error(kind, n, p, cinv)
    dim = mesh.dim
    P = orc.Problem(mesh.shape, mesh.coords, mesh.node_ids, p, n_fields=dim + 1)
    P.set_acoustic(np.full(mesh.n_elements, cinv))
    amp, k, sh = acoustic_wave(dim, cinv, T)
    e_of_gpu_state = float(np.sqrt(P.l2_error_fields(u, amp, k, sh).sum()))
    u_orc = P.rk4(u0, T / steps, steps)
    e_orc = float(np.sqrt(P.l2_error_fields(u_orc, amp, k, sh).sum()))
    print(f"  GPU {e_gpu:.12e}, oracle norm of the GPU state {e_of_gpu_state:.12e}, oracle run {e_orc:.12e}")
    assert abs(e_gpu - e_of_gpu_state) <= 1e-9 * e_orc
    assert abs(e_gpu - e_orc) <= 1e-9 * e_orc


def gpu_advection_error(n, p, a=(1.0, 0.7, 0.3), t_end=0.1, cfl=0.05):
    mesh = Mesh.structured([n, n, n])
    dev = HpgemAPIDevice(mesh, p, velocity=list(a), tableau=RK4)
    amp = np.array([1.0, -0.6])
    k = np.array([[1.0, 0.0, 1.0], [0.0, 1.0, -1.0]])
    dev.setInitialSolutionModes(amp, k)
    steps = int(np.ceil(t_end / (cfl / n / np.linalg.norm(a))))
    dev.computeOneTimeStep(0.0, t_end / steps, steps)
    l2, _ = dev.computeErrorModes(amp, k, shift=[ai * t_end for ai in a])
    dev.close()
    return l2


@pytest.mark.parametrize("p", [2, 3, 4])
def test_advection_gpu_converges_at_optimal_order(p):
    print(f"[case] advection convergence p={p}", flush=True)
    ns = (8, 16) if p < 4 else (4, 8)
    e = [gpu_advection_error(n, p) for n in ns]
    rate = np.log2(e[0] / e[1])
    lo, hi = advection_rate_window(p)
    print(f"  L2 errors {e[0]:.4e} {e[1]:.4e}, rate {rate:.2f} (window {lo:.2f}..{hi:.2f})", flush=True)
    assert lo <= rate <= hi, (e, rate)

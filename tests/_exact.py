"""Exact solutions the parity tests measure convergence against.

acoustic_wave: the travelling wave of hpGEM's acoustic example,
applications/ExampleMultipleVariableProblem/AcousticWaveLinear_Impl.h:60-90,
with c = 1/sqrt(cInv) and x1 = sum of the transverse coordinates:

    v   = -sqrt(D) c 2pi cos(2pi x0) cos(2pi (x1 - sqrt(D) c t))
    s_0 = -2pi sin(2pi x0) sin(2pi (x1 - sqrt(D) c t)) / cInv
    s_d =  2pi cos(2pi x0) cos(2pi (x1 - sqrt(D) c t)) / cInv   (d >= 1)

It solves v_t = div s, cInv s_t = grad v (the system the acoustic integrands
discretise) on the periodic unit box. With A = 2pi (k_a . x + w t), B = 2pi
(k_b . x - w t), k_a = (1, -1, .., -1), k_b = (1, .., 1), w = sqrt(D) c, each
field is a combination of cos A and cos B. The tests use the same wave
translated by a quarter period in x0 (x -> x + (1/4, 0, ..)): the equations
and the periodic box are translation invariant, and the translated wave is a
sum of the two sine modes sin(2pi k . (x - shift(t))) with shift(t) = (0, w t
/ (D - 1), ..), which the projection and error routines take directly.
"""
import numpy as np


def acoustic_wave(dim, cinv, t):
    """(amp [D+1][2], k [2][D], shift [D]): field v of the translated wave at
    time t is sum_m amp[v][m] sin(2 pi k[m] . (x - shift))."""
    c = 1.0 / np.sqrt(cinv)
    w = np.sqrt(dim) * c
    k = np.array([[1.0] + [-1.0] * (dim - 1), [1.0] * dim])
    amp = np.zeros((dim + 1, 2))
    # cos(theta + pi/2) = -sin(theta): every cos A, cos B turns into -sin
    amp[0] = [w * np.pi, w * np.pi]                # v   = -w pi (cos A + cos B)
    amp[1] = [np.pi / cinv, -np.pi / cinv]         # s_0 = -(pi / cInv)(cos A - cos B)
    for d in range(2, dim + 1):
        amp[d] = [-np.pi / cinv, -np.pi / cinv]    # s_d = (pi / cInv)(cos A + cos B)
    shift = np.zeros(dim)
    shift[1:] = w * t / (dim - 1)
    return amp, k, shift


def acoustic_wave_value(dim, cinv, t, x):
    """The reference's formula itself, at x translated by the quarter period
    (a direct check of the mode decomposition above)."""
    c = 1.0 / np.sqrt(cinv)
    xs = np.array(x, dtype=float).copy()
    xs[0] += 0.25
    x0, x1 = xs[0], xs[1:].sum()
    ph = 2 * np.pi * (x1 - np.sqrt(dim) * c * t)
    out = np.zeros(dim + 1)
    out[0] = -np.sqrt(dim) * c * 2 * np.pi * np.cos(2 * np.pi * x0) * np.cos(ph)
    out[1] = -2 * np.pi * np.sin(2 * np.pi * x0) * np.sin(ph) / cinv
    for d in range(2, dim + 1):
        out[d] = 2 * np.pi * np.cos(2 * np.pi * x0) * np.cos(ph) / cinv
    return out


def eval_modes(amp, k, shift, x):
    return sum(a * np.sin(2 * np.pi * (np.asarray(kk) @ (np.asarray(x) - shift))) for a, kk in zip(amp, k))


# Orders of convergence of the L2 error under mesh halving, as the tests
# assert them: upwind advection is optimal (p + 1); the central flux of the
# acoustic system gives p for odd p and p + 1 for even p on these uniform
# meshes. Windows are generous on the high side (pre-asymptotic meshes
# superconverge), tight on the low side, where a wrong operator lands.
def advection_rate_window(p):
    return p + 1 - 0.35, p + 1 + 0.8


def acoustic_rate_window(p):
    # p = 3 on squares measures 2.55 from 8^2 to 16^2 and 2.99 from 16^2 to
    # 32^2: the low side allows for that pre-asymptotic range
    return p - 0.5, p + 2.2

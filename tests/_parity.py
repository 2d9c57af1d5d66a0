"""The FP64 parity bar shared by the GPU suites.

got = the GPU's FP64 result, ref64 = the oracle in double (hpGEM's algorithm
with its FP64 rounding), ref_ld = the same algorithm in long double (the exact
operator to ~1e-18). "Relative" = max abs difference / max abs value of the
reference, over all coefficients.

* got vs ref_ld: <= 1e-12 for every case (measured: 1e-16 .. 4e-16).
* got vs ref64: <= 1e-12 for every tetrahedral case and every square/cube
  case whose highest order is <= 2.
* got vs ref64 on squares/cubes at order >= 3: <= 1e-12 relative in the
  mass-weighted L2 norm (the DG function-space norm,
  sqrt(sum_e x_e^T M_e x_e)); the coefficient distance is printed and kept.
  hpGEM solves every element's hierarchical mass matrix by LU on every call,
  and the tensor-product sets are ill-conditioned (kappa(M_e) ~ 1.9e5 at
  p = 3, ~3e6 at p = 4 on cubes), so the double oracle's own coefficients
  sit up to 1.35e-12 (p = 3, the reference's unitCubeN4Per mesh) and 2.0e-11
  (p = 4, 3^3 cubes) from the exact operator, while in the mass-weighted norm
  they sit within 4e-14 and 3e-13: that rounding lives in directions of tiny
  L2 size, which the function-space norm sees at their true weight. A plain
  1e-12 coefficient bar against the double oracle would test the
  reference's own rounding there, not this implementation, which stays
  within 4e-16 of the exact operator. hpGEM's tetrahedral mass matrices are
  far better conditioned: the double oracle stays within 7e-14 of the exact
  operator up to p = 4, so tetrahedra take the plain bar at every order.

Every call prints the three distances and appends them to DISTANCES, which
the suites' session fixture writes to $HGDG_PARITY_DIR/parity_distances.json
when that variable names an existing folder.
"""
import numpy as np

import _oracle as orc

FP64_TOL = 1e-12
DISTANCES = []


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def element_masses(P):
    """Per-element mass matrices of an oracle problem (acoustic: the cInv
    weighting on the s fields), with each element's DOF offset."""
    out = []
    off = 0
    for e in range(P.n_elem):
        n = orc.load().orc_basis_size(P.shape, int(P.orders[e])) * P.n_fields
        M = np.zeros(n * n)
        orc.check(P.lib.orc_mass_matrix(P.h, e, orc._dp(M)))
        out.append((off, M.reshape(n, n, order="F")))
        off += n
    return out


def mass_rel(masses, x, ref):
    num = den = 0.0
    d = x - ref
    for off, M in masses:
        n = M.shape[0]
        num += d[off:off + n] @ M @ d[off:off + n]
        den += ref[off:off + n] @ M @ ref[off:off + n]
    return float(np.sqrt(num / den))


def assert_parity(got, ref64, ref_ld, p_max, oracle, label=""):
    """The bar above. p_max = the highest element order of the case; oracle =
    the double oracle problem (its shape, and its mass matrices at p >= 3)."""
    own = rel(ref64, ref_ld)
    e_ld = rel(got, ref_ld)
    e_64 = rel(got, ref64)
    rec = {"case": label, "p_max": int(p_max), "vs_long_double": e_ld, "vs_double": e_64, "double_vs_long_double": own}
    msg = f"  [{label}] vs long double {e_ld:.2e}, vs double {e_64:.2e} (double vs long double {own:.2e})"
    conditioning_limited = p_max >= 3 and oracle.shape in (orc.SQUARE, orc.CUBE)
    if conditioning_limited:
        masses = element_masses(oracle)
        m64 = mass_rel(masses, got, ref64)
        rec["vs_double_mass_l2"] = m64
        msg += f", vs double in the mass-weighted L2 norm {m64:.2e}"
    print(msg, flush=True)
    DISTANCES.append(rec)
    assert e_ld <= FP64_TOL, rec
    if conditioning_limited:
        assert rec["vs_double_mass_l2"] <= FP64_TOL, rec
    else:
        assert e_64 <= FP64_TOL, rec

"""ctypes binding to the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: the checker the GPU path is compared against. Only the
tests, the smoke check and the CPU-baseline leg of bench.py use it.
"""
import ctypes
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
LIB_PATH_LD = os.path.join(ROOT, "oracle", "liboracle_ld.so")
LIB_PATH_BENCH = os.path.join(ROOT, "oracle", "liboracle_bench.so")
FIXTURE = os.path.join(ROOT, "tests", "golden", "ref_tables.json")

LINE, SQUARE, TRIANGLE, CUBE, TETRAHEDRON = 1, 2, 3, 4, 5
SHAPE_IDS = {"line": LINE, "square": SQUARE, "triangle": TRIANGLE, "cube": CUBE,
             "tetrahedron": TETRAHEDRON}
DIM = {LINE: 1, SQUARE: 2, TRIANGLE: 2, CUBE: 3, TETRAHEDRON: 3}

_libs = {}
_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int)


def _dp(a):
    return a.ctypes.data_as(_D)


def load(extended=False):
    """The double-precision oracle, or (extended=True) the same algorithm in
    long double, or (extended="bench") the double build at -O3 with OpenMP
    that bench.py times as the CPU baseline."""
    if extended in _libs:
        return _libs[extended]
    path = LIB_PATH_BENCH if extended == "bench" else LIB_PATH_LD if extended else LIB_PATH
    if not os.path.exists(path):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)
    lib = ctypes.CDLL(path)
    lib.orc_last_error.restype = ctypes.c_char_p
    lib.orc_problem_create.restype = ctypes.c_void_p
    lib.orc_problem_create.argtypes = [ctypes.c_int, ctypes.c_int64, _D, _I64, _I32, ctypes.c_int]
    lib.orc_problem_destroy.argtypes = [ctypes.c_void_p]
    for name in ["orc_problem_info", "orc_faces", "orc_set_advection", "orc_set_acoustic",
                 "orc_assemble", "orc_time_derivative", "orc_rk4", "orc_project_modes",
                 "orc_l2_error_modes", "orc_mass_matrix"]:
        getattr(lib, name).restype = ctypes.c_int
    lib.orc_problem_info.argtypes = [ctypes.c_void_p, _I64, _I64, _I64]
    lib.orc_faces.argtypes = [ctypes.c_void_p, _I64]
    lib.orc_set_advection.argtypes = [ctypes.c_void_p, _D]
    lib.orc_set_acoustic.argtypes = [ctypes.c_void_p, _D]
    lib.orc_assemble.argtypes = [ctypes.c_void_p]
    lib.orc_time_derivative.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(_D), _D, _D]
    lib.orc_time_derivative_par.restype = ctypes.c_int
    lib.orc_time_derivative_par.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(_D), _D, _D]
    lib.orc_threads.restype = ctypes.c_int
    lib.orc_threads.argtypes = []
    lib.orc_rk4.argtypes = [ctypes.c_void_p, _D, ctypes.c_double, ctypes.c_int]
    lib.orc_project_modes.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, _D, _D, _D]
    lib.orc_l2_error_modes.argtypes = [ctypes.c_void_p, _D, ctypes.c_int, _D, _D, ctypes.c_double, _D]
    lib.orc_mass_matrix.argtypes = [ctypes.c_void_p, ctypes.c_int64, _D]
    lib.orc_hcurl_eval.restype = ctypes.c_int
    lib.orc_hcurl_eval.argtypes = [ctypes.c_int, _D, _D, _D]
    lib.orc_maxwell_assemble.restype = ctypes.c_int
    lib.orc_maxwell_assemble.argtypes = [ctypes.c_void_p, ctypes.c_int, _D, ctypes.c_double, _D, _D, _D]
    _libs[extended] = lib
    _load_simplex_rules(lib)
    return lib


def check(rc, lib=None):
    if rc != 0:
        raise RuntimeError("oracle: " + (lib or load()).orc_last_error().decode())


def fixture_tables():
    with open(FIXTURE) as f:
        return json.load(f)["tables"]


def _load_simplex_rules(lib):
    """Simplex Gauss rules are data taken from the reference-built fixture."""
    seen = set()
    for t in fixture_tables():
        if t["kind"] != "rule" or t["shape"] not in ("triangle", "tetrahedron"):
            continue
        key = (t["shape"], t["name"])
        if key in seen:
            continue
        seen.add(key)
        pts = np.ascontiguousarray(t["points"], dtype=np.float64).ravel()
        w = np.ascontiguousarray(t["weights"], dtype=np.float64)
        check(lib.orc_set_simplex_rule(SHAPE_IDS[t["shape"]], t["order"], len(w), _dp(pts), _dp(w)))


def rule(shape, order):
    lib = load()
    n = ctypes.c_int()
    o = ctypes.c_int()
    check(lib.orc_rule(shape, order, ctypes.byref(n), ctypes.byref(o), None, None))
    pts = np.zeros(n.value * DIM[shape])
    w = np.zeros(n.value)
    check(lib.orc_rule(shape, order, ctypes.byref(n), ctypes.byref(o), _dp(pts), _dp(w)))
    return pts.reshape(n.value, DIM[shape]), w, o.value


def basis_eval(shape, p, pt):
    lib = load()
    n = lib.orc_basis_size(shape, p)
    if n < 0:
        check(1)
    d = DIM[shape]
    vals = np.zeros(n)
    grads = np.zeros(n * d)
    x = np.ascontiguousarray(pt, dtype=np.float64)
    check(lib.orc_basis_eval(shape, p, _dp(x), _dp(vals), _dp(grads)))
    return vals, grads.reshape(n, d)


def face_map(shape, face, pin):
    lib = load()
    d = DIM[shape]
    pout = np.zeros(d)
    sign = ctypes.c_double()
    x = np.ascontiguousarray(pin if len(pin) else [0.0], dtype=np.float64)
    check(lib.orc_face_map(shape, face, _dp(x), _dp(pout), ctypes.byref(sign)))
    return pout, sign.value


def orientation_map(fshape, idx, pin):
    lib = load()
    pout = np.zeros(DIM[fshape])
    x = np.ascontiguousarray(pin, dtype=np.float64)
    check(lib.orc_orientation_map(fshape, idx, _dp(x), _dp(pout)))
    return pout


def orientation_index(fshape, l1, l2):
    lib = load()
    a = np.ascontiguousarray(l1, dtype=np.int64)
    b = np.ascontiguousarray(l2, dtype=np.int64)
    r = lib.orc_orientation_index(fshape, len(a), a.ctypes.data_as(_I64), b.ctypes.data_as(_I64))
    if r < 0:
        check(1)
    return r


def hcurl_eval(p, pt):
    """Values and curls ([n][3]) of the oracle's Ainsworth-Coyle set."""
    lib = load()
    n = 6 * (p + 1) + (12 * (p - 1) if p >= 2 else 0)
    v, c = np.zeros(3 * n), np.zeros(3 * n)
    check(lib.orc_hcurl_eval(p, _dp(np.ascontiguousarray(pt, dtype=np.float64)), _dp(v), _dp(c)))
    return v.reshape(n, 3), c.reshape(n, 3)


def structured_mesh(n, lo=None, hi=None, periodic=None, simplices=False):
    lib = load()
    dim = len(n)
    lo = np.zeros(dim) if lo is None else np.asarray(lo, dtype=np.float64)
    hi = np.ones(dim) if hi is None else np.asarray(hi, dtype=np.float64)
    per = np.ones(dim, dtype=np.int32) if periodic is None else np.asarray(periodic, dtype=np.int32)
    nn = np.asarray(n, dtype=np.int32)
    ne = ctypes.c_int64()
    nv = ctypes.c_int()
    shape = ctypes.c_int()
    args = [dim, nn.ctypes.data_as(_I32), _dp(lo), _dp(hi), per.ctypes.data_as(_I32), int(simplices)]
    check(lib.orc_structured_mesh(*args, ctypes.byref(ne), ctypes.byref(nv), ctypes.byref(shape), None, None))
    X = np.zeros((ne.value, nv.value, dim))
    nodes = np.zeros((ne.value, nv.value), dtype=np.int64)
    check(lib.orc_structured_mesh(*args, ctypes.byref(ne), ctypes.byref(nv), ctypes.byref(shape),
                                  _dp(X), nodes.ctypes.data_as(_I64)))
    return shape.value, X, nodes


class Problem:
    """Mesh + DG space + operator in the oracle, state in reference layout."""

    def __init__(self, shape, X, nodes, orders, n_fields=1, extended=False):
        self.lib = load(extended)
        self.shape = shape
        self.dim = DIM[shape]
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        ne = self.nodes.shape[0]
        self.orders = np.ascontiguousarray(np.broadcast_to(orders, (ne,)), dtype=np.int32)
        self.n_elem = ne
        self.n_fields = n_fields
        h = self.lib.orc_problem_create(shape, ne, _dp(self.X), self.nodes.ctypes.data_as(_I64),
                                        self.orders.ctypes.data_as(_I32), n_fields)
        if not h:
            check(1)
        self.h = ctypes.c_void_p(h)
        nf, ni, nd = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(self.lib.orc_problem_info(self.h, ctypes.byref(nf), ctypes.byref(ni), ctypes.byref(nd)))
        self.n_faces, self.n_internal, self.n_dof = nf.value, ni.value, nd.value

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_problem_destroy(self.h)
            self.h = None

    def faces(self):
        out = np.zeros((self.n_faces, 5), dtype=np.int64)
        check(self.lib.orc_faces(self.h, out.ctypes.data_as(_I64)))
        return out

    def set_advection(self, a):
        av = np.zeros(3)
        av[: len(a)] = a
        check(self.lib.orc_set_advection(self.h, _dp(av)))

    def set_acoustic(self, cinv):
        c = np.ascontiguousarray(cinv, dtype=np.float64)
        check(self.lib.orc_set_acoustic(self.h, _dp(c)))

    def time_derivative(self, vecs, coefs):
        vecs = [np.ascontiguousarray(v, dtype=np.float64) for v in vecs]
        arr = (_D * len(vecs))(*[_dp(v) for v in vecs])
        c = np.ascontiguousarray(coefs, dtype=np.float64)
        out = np.zeros(self.n_dof)
        check(self.lib.orc_time_derivative(self.h, len(vecs), arr, _dp(c), _dp(out)))
        return out

    def time_derivative_par(self, vecs, coefs):
        """The same operator with the element loop on all OpenMP threads (the
        bench build): each element gathers its faces' blocks."""
        vecs = [np.ascontiguousarray(v, dtype=np.float64) for v in vecs]
        arr = (_D * len(vecs))(*[_dp(v) for v in vecs])
        c = np.ascontiguousarray(coefs, dtype=np.float64)
        out = np.zeros(self.n_dof)
        check(self.lib.orc_time_derivative_par(self.h, len(vecs), arr, _dp(c), _dp(out)))
        return out

    def assemble(self):
        check(self.lib.orc_assemble(self.h))

    def right_hand_side(self, vecs, coefs):
        """r = S w before the mass solve (HpgemAPILinear::computeRightHandSide)."""
        vecs = [np.ascontiguousarray(v, dtype=np.float64) for v in vecs]
        arr = (_D * len(vecs))(*[_dp(v) for v in vecs])
        c = np.ascontiguousarray(coefs, dtype=np.float64)
        out = np.zeros(self.n_dof)
        self.lib.orc_right_hand_side.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(_D), _D, _D]
        check(self.lib.orc_right_hand_side(self.h, len(vecs), arr, _dp(c), _dp(out)))
        return out

    def solve_mass(self, v):
        """M^{-1} v per element by LU (solveMassMatrixEquations)."""
        out = np.array(v, dtype=np.float64, copy=True)
        self.lib.orc_solve_mass.argtypes = [ctypes.c_void_p, _D]
        check(self.lib.orc_solve_mass(self.h, _dp(out)))
        return out

    def rk4(self, u, dt, n_steps):
        u = np.array(u, dtype=np.float64, copy=True)
        check(self.lib.orc_rk4(self.h, _dp(u), dt, n_steps))
        return u

    def rk_steps(self, u, dt, n_steps, a, b):
        u = np.array(u, dtype=np.float64, copy=True)
        s = len(b)
        A = np.zeros((s, s))
        for i, row in enumerate(a):
            A[i, : len(row)] = row
        A = np.ascontiguousarray(A)
        B = np.ascontiguousarray(b, dtype=np.float64)
        self.lib.orc_rk_steps.argtypes = [ctypes.c_void_p, _D, ctypes.c_double, ctypes.c_int, ctypes.c_int, _D, _D]
        check(self.lib.orc_rk_steps(self.h, _dp(u), dt, n_steps, s, _dp(A), _dp(B)))
        return u

    def project_modes(self, amp, k, field_scale=None):
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        kk = np.zeros((len(amp), 3))
        kk[:, : np.asarray(k).shape[1]] = k
        fs = np.ones(self.n_fields) if field_scale is None else np.ascontiguousarray(field_scale, dtype=np.float64)
        u = np.zeros(self.n_dof)
        check(self.lib.orc_project_modes(self.h, len(amp), _dp(amp), _dp(kk), _dp(fs), _dp(u)))
        return u

    def l2_error_modes(self, u, amp, k, t):
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        kk = np.zeros((len(amp), 3))
        kk[:, : np.asarray(k).shape[1]] = k
        u = np.ascontiguousarray(u, dtype=np.float64)
        err = ctypes.c_double()
        check(self.lib.orc_l2_error_modes(self.h, _dp(u), len(amp), _dp(amp), _dp(kk), t, ctypes.byref(err)))
        return err.value

    def maxwell(self, p, eps, stab):
        """DG-Max element matrices S (curl-curl), M (permittivity mass) and face
        blocks F[face][(test side, trial side)], each n x n with [row, col] =
        [test, trial]."""
        n = 6 * (p + 1) + (12 * (p - 1) if p >= 2 else 0)
        nf = self.n_faces
        S = np.zeros(self.n_elem * n * n)
        M = np.zeros(self.n_elem * n * n)
        F = np.zeros(nf * 4 * n * n)
        e = np.ascontiguousarray(eps, dtype=np.float64)
        check(self.lib.orc_maxwell_assemble(self.h, p, _dp(e), stab, _dp(S), _dp(M), _dp(F)))
        # column-major blocks -> [row][col]
        S = S.reshape(self.n_elem, n, n).transpose(0, 2, 1)
        M = M.reshape(self.n_elem, n, n).transpose(0, 2, 1)
        F = F.reshape(nf, 4, n, n).transpose(0, 1, 3, 2)
        return S, M, F

    def l2_error_fields(self, u, amp, k, shift):
        """Per-field squared L2 errors against f_v(x - shift), f_v = sum_m
        amp[v][m] sin(2 pi k_m . x) (computeTotalError's integrand, rule 2p+2)."""
        amp = np.ascontiguousarray(np.broadcast_to(amp, (self.n_fields, len(k))), dtype=np.float64)
        kk = np.zeros((len(k), 3))
        kk[:, : np.asarray(k).shape[1]] = k
        sh = np.zeros(3)
        sh[: len(shift)] = shift
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(self.n_fields)
        self.lib.orc_l2_error_fields.argtypes = [ctypes.c_void_p, _D, ctypes.c_int, _D, _D, _D, _D]
        check(self.lib.orc_l2_error_fields(self.h, _dp(u), len(kk), _dp(amp), _dp(kk), _dp(sh), _dp(out)))
        return out

    def matrices(self):
        """(M_e list, S_e list, face rows [L, R], face blocks concatenated LL, LR,
        RL, RR (boundary: LL), all column-major) as HpgemAPILinear holds them."""
        faces = self.faces()
        M, S = [], []
        sizes = [self.lib.orc_basis_size(self.shape, int(p)) * self.n_fields for p in self.orders]
        for e in range(self.n_elem):
            n = sizes[e]
            m, st = np.zeros(n * n), np.zeros(n * n)
            check(self.lib.orc_mass_matrix(self.h, e, _dp(m)))
            check(self.lib.orc_element_stiffness(self.h, e, _dp(st)))
            M.append(m)
            S.append(st)
        blocks = []
        for f in range(self.n_faces):
            L, R = int(faces[f][0]), int(faces[f][1])
            nL = sizes[L]
            nR = sizes[R] if R >= 0 else 0
            out = np.zeros((nL + nR) ** 2)
            check(self.lib.orc_face_blocks(self.h, f, _dp(out)))
            blocks.append(out[: nL * nL] if R < 0 else out)
        return M, S, faces[:, :2].copy(), blocks, sizes

    def mass_matrix(self, e):
        nb = self.n_dof // self.n_elem if np.all(self.orders == self.orders[0]) else None
        n = nb
        out = np.zeros(n * n)
        check(self.lib.orc_mass_matrix(self.h, e, _dp(out)))
        return out.reshape(n, n, order="F")

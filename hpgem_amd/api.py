"""hpGEM-shaped host API over the C ABI (include/hgdg.h).

Method names and argument meaning follow hpGEM's explicit solver API
(kernel/Base/HpgemAPISimplified.h:190-442, HpgemAPILinear.h:168-301):
time-integration vectors are addressed by id, computeTimeDerivative takes
input ids and coefficients, computeOneTimeStep advances (time, dt). The
integrands are chosen declaratively (operator descriptor), since a C++/Python
virtual integrand cannot run inside a GPU kernel. All arithmetic runs in the
HIP library; this module only marshals arguments.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr
from .tableaux import RK4, ButcherTableau

SHAPES = {"square": _lib.SHAPE_SQUARE, "cube": _lib.SHAPE_CUBE, "tetrahedron": _lib.SHAPE_TETRAHEDRON}
PRECISIONS = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED_FP32, "mixed_fp32": _lib.PREC_MIXED_FP32}


class Mesh:
    """Element vertex coordinates + topological node ids (a read .hpgem mesh)."""

    def __init__(self, shape, coords, node_ids, zones=None, zone_names=("Main",), faces=None):
        self.shape = int(shape)
        self.coords = np.ascontiguousarray(coords, dtype=np.float64)
        self.node_ids = np.ascontiguousarray(node_ids, dtype=np.int64)
        self.n_elements, self.n_vertices, self.dim = self.coords.shape
        self.zones = np.zeros(self.n_elements, dtype=np.int32) if zones is None else np.asarray(zones, np.int32)
        self.zone_names = list(zone_names)
        # faces listed by a mesh file: (element, local face) per side, -1 on the boundary side
        self.face_elements, self.face_local = (None, None) if faces is None else faces

    @classmethod
    def read(cls, path):
        """Read a .hpgem file (MeshManipulator::readMesh, kernel/Base/MeshManipulator_Impl.h:1505-1800)."""
        lib = _lib.load()
        info = _lib.MeshFileInfo()
        name = str(path).encode()
        check(lib.hgdg_mesh_read(name, ctypes.byref(info), None, None, None, None, None, None))
        ne, nv, dim, nf = info.n_elements, info.n_vertices, info.dim, info.n_faces
        coords = np.zeros((ne, nv, dim))
        nodes = np.zeros((ne, nv), dtype=np.int64)
        zones = np.zeros(ne, dtype=np.int32)
        fel = np.zeros((nf, 2), dtype=np.int64)
        flo = np.zeros((nf, 2), dtype=np.int32)
        names = ctypes.create_string_buffer(info.zone_name_bytes)
        check(lib.hgdg_mesh_read(name, ctypes.byref(info), dptr(coords), nodes.ctypes.data_as(_lib._I64),
                                 zones.ctypes.data_as(_lib._I32), fel.ctypes.data_as(_lib._I64),
                                 flo.ctypes.data_as(_lib._I32), names))
        mesh = cls(info.shape, coords, nodes, zones, names.value.decode().split("\n"), (fel, flo))
        mesh.version, mesh.n_nodes, mesh.n_edges = info.version, info.n_nodes, info.n_edges
        return mesh

    def write(self, path):
        """Write a .hpgem file, version 2, one partition (the Preprocessor's
        outputMesh, applications/Preprocessor/lib/output_impl.h:91-208)."""
        lib = _lib.load()
        coords = np.ascontiguousarray(self.coords, dtype=np.float64)
        nodes = np.ascontiguousarray(self.node_ids, dtype=np.int64)
        zones = None if self.zones is None else np.ascontiguousarray(self.zones, dtype=np.int32)
        names = "\n".join(self.zone_names).encode() if self.zone_names else None
        check(lib.hgdg_mesh_write(str(path).encode(), int(self.shape), int(self.n_elements), dptr(coords),
                                  nodes.ctypes.data_as(_lib._I64),
                                  None if zones is None else zones.ctypes.data_as(_lib._I32), names))

    @classmethod
    def structured(cls, n, lower=None, upper=None, periodic=True, simplices=False):
        """hpGEM Preprocessor structured mesh (applications/Preprocessor/lib/hpgem.cpp:100-235)."""
        lib = _lib.load()
        dim = len(n)
        nn = np.ascontiguousarray(n, dtype=np.int32)
        lo = np.zeros(dim) if lower is None else np.ascontiguousarray(lower, dtype=np.float64)
        hi = np.ones(dim) if upper is None else np.ascontiguousarray(upper, dtype=np.float64)
        per = np.ascontiguousarray(np.broadcast_to(np.asarray(periodic, dtype=np.int32), (dim,)))
        ne, nv, shape = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        args = [dim, nn.ctypes.data_as(_lib._I32), dptr(lo), dptr(hi), per.ctypes.data_as(_lib._I32),
                int(bool(simplices))]
        check(lib.hgdg_structured_mesh(*args, ctypes.byref(ne), ctypes.byref(nv), ctypes.byref(shape), None, None))
        coords = np.zeros((ne.value, nv.value, dim))
        nodes = np.zeros((ne.value, nv.value), dtype=np.int64)
        check(lib.hgdg_structured_mesh(*args, ctypes.byref(ne), ctypes.byref(nv), ctypes.byref(shape),
                                       dptr(coords), nodes.ctypes.data_as(_lib._I64)))
        return cls(shape.value, coords, nodes)


def write_vtk(base_name, mesh, orders, state, names):
    """hpGEM's VTK output at order 1 (VTKSpecificTimeWriter,
    kernel/Output/VTKSpecificTimeWriter_Impl.h): <base_name>.pvtu and
    <base_name>.0.vtu with one point array per field, each element's values
    at its own vertices. state in hpGEM's layout (getTimeIntegrationVector)."""
    lib = _lib.load()
    ordv = np.ascontiguousarray(np.broadcast_to(np.asarray(orders, dtype=np.int32), (mesh.n_elements,)))
    st = np.ascontiguousarray(state, dtype=np.float64)
    names = list(names)
    check(lib.hgdg_vtk_write(str(base_name).encode(), int(mesh.shape), int(mesh.n_elements), dptr(mesh.coords),
                             ordv.ctypes.data_as(_lib._I32), int(ordv[0]) if len(ordv) else 1, len(names),
                             dptr(st), int(st.size), "\n".join(names).encode()))


def basis_size(shape, p):
    return _lib.load().hgdg_basis_size(int(shape), int(p))


def basis_eval(shape, p, point):
    lib = _lib.load()
    n = basis_size(shape, p)
    out = np.zeros(n)
    x = np.ascontiguousarray(point, dtype=np.float64)
    check(lib.hgdg_basis_eval(int(shape), int(p), dptr(x), dptr(out)))
    return out


def hcurl_basis_eval(p, point):
    """Values and curls ([n][3] each) of hpGEM's Ainsworth-Coyle set of order p."""
    lib = _lib.load()
    n = lib.hgdg_hcurl_basis_size(int(p))
    v, c = np.zeros((n, 3)), np.zeros((n, 3))
    x = np.ascontiguousarray(point, dtype=np.float64)
    check(lib.hgdg_hcurl_basis_eval(int(p), dptr(x), dptr(v), dptr(c)))
    return v, c


def maxwell_assemble(mesh, p, epsilon, stab, device=0):
    """Interior-penalty Maxwell stiffness (block-CSR) and permittivity mass of
    a tetrahedral mesh on the GPU (DG-Max, DGMaxDiscretization.cpp:193-410)."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    check(lib.hgdg_ctx_create(ctypes.byref(h), device))
    try:
        order = np.ones(mesh.n_elements, dtype=np.int32)
        desc = _lib.MeshDesc(mesh.shape, mesh.dim, mesh.n_elements, dptr(mesh.coords),
                             mesh.node_ids.ctypes.data_as(_lib._I64), order.ctypes.data_as(_lib._I32), 1)
        check(lib.hgdg_mesh_upload(h, ctypes.byref(desc)))
        eps = np.ascontiguousarray(epsilon, dtype=np.float64)
        md = _lib.MaxwellDesc(int(p), dptr(eps), float(stab))
        nblk, bs, ms = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_double()
        check(lib.hgdg_maxwell_assemble(h, ctypes.byref(md), ctypes.byref(nblk), ctypes.byref(bs), None, None, None,
                                        None, None))
        n = bs.value
        row_ptr = np.zeros(mesh.n_elements + 1, dtype=np.int64)
        col_idx = np.zeros(nblk.value, dtype=np.int64)
        K = np.zeros((nblk.value, n, n))
        M = np.zeros((mesh.n_elements, n, n))
        check(lib.hgdg_maxwell_assemble(h, ctypes.byref(md), ctypes.byref(nblk), ctypes.byref(bs),
                                        row_ptr.ctypes.data_as(_lib._I64), col_idx.ctypes.data_as(_lib._I64),
                                        dptr(K), dptr(M), ctypes.byref(ms)))
        return {"row_ptr": row_ptr, "col_idx": col_idx, "K": K, "M": M, "block_size": n, "kernel_ms": ms.value}
    finally:
        check(lib.hgdg_ctx_destroy(h))


class HpgemAPIDevice:
    """Explicit DG solver whose RHS, mass solve and RK loop run on one GPU.

    Mirrors HpgemAPISimplified/HpgemAPILinear: vector 0 is the solution
    (solutionVectorId_), vectors 1..stages are the stage derivatives
    (auxiliaryVectorIds_), extra vectors follow.
    """

    def __init__(self, mesh, polynomial_order, velocity=None, precision="fp64", tableau=RK4,
                 n_extra_vectors=0, device=0, kind=_lib.OP_ADVECTION_UPWIND, c_inv=None, tensor_general=False):
        self.lib = _lib.load()
        self.mesh = mesh
        self.p = polynomial_order
        self.tableau = tableau
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        h = ctypes.c_void_p()
        check(self.lib.hgdg_ctx_create(ctypes.byref(h), device))
        self.h = h
        if tensor_general:
            check(self.lib.hgdg_set_tensor_general(self.h, 1))
        self._order = np.ascontiguousarray(np.broadcast_to(np.asarray(polynomial_order, dtype=np.int32),
                                                           (mesh.n_elements,)))
        desc = _lib.MeshDesc(mesh.shape, mesh.dim, mesh.n_elements, dptr(mesh.coords),
                             mesh.node_ids.ctypes.data_as(_lib._I64), self._order.ctypes.data_as(_lib._I32),
                             int(self._order[0]))
        check(self.lib.hgdg_mesh_upload(self.h, ctypes.byref(desc)))
        vel = (ctypes.c_double * 3)(*(list(velocity or []) + [0.0, 0.0, 0.0])[:3])
        self._cinv = None if c_inv is None else np.ascontiguousarray(c_inv, dtype=np.float64)
        op = _lib.OpDesc(kind, self.precision, vel, dptr(self._cinv) if self._cinv is not None else None)
        check(self.lib.hgdg_operator_set(self.h, ctypes.byref(op)))
        self._query_info()
        self.solutionVectorId_ = 0
        self.auxiliaryVectorIds_ = list(range(1, tableau.stages + 1))
        self.n_vectors = 1 + tableau.stages + n_extra_vectors
        check(self.lib.hgdg_state_alloc(self.h, self.n_vectors))

    def _query_info(self):
        # The operator fixes the number of fields (acoustics: dim + 1), so the
        # DOF count is read after hgdg_operator_set.
        ne, nf, ni, nd = (ctypes.c_int64() for _ in range(4))
        check(self.lib.hgdg_mesh_info(self.h, ctypes.byref(ne), ctypes.byref(nf), ctypes.byref(ni), ctypes.byref(nd)))
        self.n_elements, self.n_faces, self.n_internal_faces, self.n_dof = ne.value, nf.value, ni.value, nd.value

    def close(self):
        if getattr(self, "h", None):
            check(self.lib.hgdg_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def usesTensorGeneral(self):
        """Whether the mesh runs on the Gauss-point kernels (non-affine
        quadrilaterals/hexahedra, unaligned neighbours, or forced by
        tensor_general=True) rather than the sum-factorised ones."""
        f, u = ctypes.c_int(), ctypes.c_int()
        check(self.lib.hgdg_get_tensor_general(self.h, ctypes.byref(f), ctypes.byref(u)))
        return bool(u.value)

    # ---- launch shape of the tensor advection kernel ----------------------
    def setAdvectionTiles(self, tiles):
        """Tiles of elements per workgroup (1, 2, 4, 8), or 0 to time every
        choice and keep the fastest (done by default when the operator is set)."""
        check(self.lib.hgdg_set_advection_tiles(self.h, int(tiles)))

    def getAdvectionTiles(self):
        """(tiles in use, timed ms per launch of candidates 1, 2, 4, 8)."""
        t = ctypes.c_int()
        ms = np.zeros(4)
        check(self.lib.hgdg_get_advection_tiles(self.h, ctypes.byref(t), dptr(ms)))
        return t.value, ms

    # ---- fused RK steps ----------------------------------------------------
    def setFusedTimeStepping(self, enable):
        """True (default): computeOneTimeStep may run all its steps in one
        single-workgroup launch when the state fits in LDS (small advection
        meshes), and otherwise pairs of steps with the RK update in each
        last-stage kernel (every operator); False: always replay the captured
        one-step graph. Results are bitwise the same."""
        check(self.lib.hgdg_set_rk_fused(self.h, 1 if enable else 0))

    def getFusedTimeStepping(self):
        """(enabled, single-workgroup RK launches made so far)."""
        e = ctypes.c_int()
        n = ctypes.c_int64()
        check(self.lib.hgdg_get_rk_fused(self.h, ctypes.byref(e), ctypes.byref(n)))
        return bool(e.value), n.value

    # ---- threads per element of the tetrahedral kernel ---------------------
    def setSimplexThreads(self, threads):
        """Threads per element for orders 1..4 (four values): each 0 to time
        both candidates of that order and keep the faster (done by default
        when the operator is set), or one of the candidates (16/32, 32/64,
        32/64, 35/64)."""
        t = np.ascontiguousarray(threads, dtype=np.int32)
        if t.shape != (4,):
            raise ValueError("four values, orders 1..4")
        check(self.lib.hgdg_set_simplex_threads(self.h, t.ctypes.data_as(_lib._I32)))

    def getSimplexThreads(self):
        """(threads per element in use for orders 1..4, timed ms per launch
        [order][candidate], 0 where not timed)."""
        t = np.zeros(4, dtype=np.int32)
        ms = np.zeros(8)
        check(self.lib.hgdg_get_simplex_threads(self.h, t.ctypes.data_as(_lib._I32), dptr(ms)))
        return t, ms.reshape(4, 2)

    # ---- time-integration vectors (reference layout) ----------------------
    def setTimeIntegrationVector(self, vec, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.size != self.n_dof:
            raise ValueError(f"expected {self.n_dof} values, got {v.size}")
        check(self.lib.hgdg_state_set(self.h, vec, dptr(v)))

    def writeVTK(self, base_name, vec=None, names=None):
        """Write time-integration vector `vec` (default: the solution) as VTK
        (VTKSpecificTimeWriter at order 1). Fields are named u (advection) or
        v, s_x, s_y, s_z (acoustics) unless names are given."""
        vec = self.solutionVectorId_ if vec is None else vec
        orders, counts = np.unique(self._order, return_counts=True)
        nb_total = sum(basis_size(self.mesh.shape, int(p)) * int(c) for p, c in zip(orders, counts))
        n_fields = self.n_dof // nb_total
        if names is None:
            names = ["u"] if n_fields == 1 else ["v", "s_x", "s_y", "s_z"][:n_fields]
        write_vtk(base_name, self.mesh, self._order, self.getTimeIntegrationVector(vec), names)

    def getTimeIntegrationVector(self, vec):
        out = np.zeros(self.n_dof)
        check(self.lib.hgdg_state_get(self.h, vec, dptr(out)))
        return out

    def setStorage(self, vec, precision):
        check(self.lib.hgdg_state_set_storage(self.h, vec, PRECISIONS[precision]))

    def setInitialSolutionModes(self, amp, k, vec=None, field=-1):
        """L2 projection of sum_m amp_m sin(2 pi k_m . x) (setInitialSolution),
        into every field (field=-1) or into one field."""
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        kk = np.zeros((len(amp), 3))
        kk[:, : np.asarray(k).shape[1]] = k
        check(self.lib.hgdg_project_field_modes(self.h, self.solutionVectorId_ if vec is None else vec, field,
                                                len(amp), dptr(amp), dptr(kk)))

    def computeErrorModes(self, amp, k, shift=None, vec=None, field=-1):
        """(L2 error, max error) of a vector against sum_m amp_m sin(2 pi k_m . (x - shift))
        (computeTotalError / computeMaxError; for advection shift = a t)."""
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        kk = np.zeros((len(amp), 3))
        kk[:, : np.asarray(k).shape[1]] = k
        sh = np.zeros(3)
        if shift is not None:
            sh[: len(shift)] = shift
        l2, mx = ctypes.c_double(), ctypes.c_double()
        check(self.lib.hgdg_error_modes(self.h, self.solutionVectorId_ if vec is None else vec, field, len(amp),
                                        dptr(amp), dptr(kk), dptr(sh), ctypes.byref(l2), ctypes.byref(mx)))
        return l2.value, mx.value

    def setInitialSolutionGaussian(self, center, width, amplitude=1.0, vec=None, field=-1):
        """L2 projection of amplitude * exp(-|x - center|^2 / width^2)."""
        c = np.zeros(3)
        c[: len(center)] = center
        check(self.lib.hgdg_project_gaussian(self.h, self.solutionVectorId_ if vec is None else vec, field, dptr(c),
                                             float(width), float(amplitude)))

    # ---- the hot path ------------------------------------------------------
    def computeTimeDerivative(self, inputVectorIds, coefficientsInputVectors, resultVectorId, time=0.0):
        ids = (ctypes.c_int * len(inputVectorIds))(*inputVectorIds)
        co = np.ascontiguousarray(coefficientsInputVectors, dtype=np.float64)
        check(self.lib.hgdg_time_derivative(self.h, ids, dptr(co), len(inputVectorIds), resultVectorId, time))

    def computeRightHandSide(self, inputVectorIds, coefficientsInputVectors, resultVectorId, time=0.0):
        """r = S w before the inverse mass (HpgemAPILinear::computeRightHandSide)."""
        ids = (ctypes.c_int * len(inputVectorIds))(*inputVectorIds)
        co = np.ascontiguousarray(coefficientsInputVectors, dtype=np.float64)
        check(self.lib.hgdg_right_hand_side(self.h, ids, dptr(co), len(inputVectorIds), resultVectorId, time))

    def solveMassMatrixEquations(self, timeIntegrationVectorId):
        """vector = M^{-1} vector, element by element (HpgemAPILinear::solveMassMatrixEquations)."""
        check(self.lib.hgdg_solve_mass(self.h, timeIntegrationVectorId))

    def scaleAndAddVector(self, vectorToChangeId, vectorToAddId, scale):
        check(self.lib.hgdg_scale_and_add(self.h, vectorToChangeId, vectorToAddId, scale))

    def computeOneTimeStep(self, time, dt, n_steps=1):
        """Advance n_steps of the tableau; returns the new time."""
        t = self.tableau
        aux = (ctypes.c_int * t.stages)(*self.auxiliaryVectorIds_)
        a = np.ascontiguousarray(t.a_matrix(), dtype=np.float64)
        b = np.ascontiguousarray(t.b, dtype=np.float64)
        check(self.lib.hgdg_rk_steps(self.h, self.solutionVectorId_, aux, t.stages, dptr(a), dptr(b), dt, n_steps))
        return time + n_steps * dt

    def computeErrorAndNormOfUpdate(self, dt):
        """(error estimate, solution size) of the last step of an embedded tableau."""
        t = self.tableau
        if not t.has_error_estimate:
            raise ValueError(f"{t.name} has no embedded error estimate")
        aux = (ctypes.c_int * t.stages)(*self.auxiliaryVectorIds_)
        e = np.ascontiguousarray(t.error, dtype=np.float64)
        err, nrm = ctypes.c_double(), ctypes.c_double()
        check(self.lib.hgdg_step_error(self.h, self.solutionVectorId_, aux, t.stages, dptr(e), dt, ctypes.byref(err),
                                       ctypes.byref(nrm)))
        return err.value, nrm.value

    def computeOneTimeStepAdaptive(self, time, maximum_relative_error, dt_max, dt_initial):
        """One adaptive step (computeOneTimeStep(time, maximumRelativeError, dtMax));
        returns (new time, accepted dt, error estimate, solution size)."""
        t = self.tableau
        if not t.has_error_estimate:
            raise ValueError(f"{t.name} has no embedded error estimate")
        aux = (ctypes.c_int * t.stages)(*self.auxiliaryVectorIds_)
        a = np.ascontiguousarray(t.a_matrix(), dtype=np.float64)
        b = np.ascontiguousarray(t.b, dtype=np.float64)
        e = np.ascontiguousarray(t.error, dtype=np.float64)
        tm, dt, err, nrm = ctypes.c_double(time), ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self.lib.hgdg_rk_adaptive_step(self.h, self.solutionVectorId_, aux, t.stages, dptr(a), dptr(b), dptr(e),
                                             t.order, maximum_relative_error, dt_max, dt_initial, ctypes.byref(tm),
                                             ctypes.byref(dt), ctypes.byref(err), ctypes.byref(nrm)))
        return tm.value, dt.value, err.value, nrm.value

    def solve(self, start_time, end_time, dt):
        n = int(round((end_time - start_time) / dt))
        return self.computeOneTimeStep(start_time, dt, n)

    def copyVector(self, dst, src):
        check(self.lib.hgdg_copy(self.h, dst, src))

    def synchronize(self):
        check(self.lib.hgdg_sync(self.h))

    def stream(self):
        s = ctypes.c_void_p()
        check(self.lib.hgdg_get_stream(self.h, ctypes.byref(s)))
        return s.value

    def difference(self, x, y):
        """(mass-weighted L2 norm, max abs coefficient) of vector x - vector y."""
        l2, mx = ctypes.c_double(), ctypes.c_double()
        check(self.lib.hgdg_state_diff(self.h, x, y, ctypes.byref(l2), ctypes.byref(mx)))
        return l2.value, mx.value


class HpgemAPIMatrixDevice(HpgemAPIDevice):
    """HpgemAPILinear's precomputed path on the device: the element mass and
    stiffness matrices and the face matrices an application builds from its
    own integrands (createMassMatrices / createStiffnessMatrices) are handed
    over once; computeTimeDerivative, computeRightHandSide,
    solveMassMatrixEquations, computeOneTimeStep and the adaptive step then
    run on the device. Vectors are in hpGEM's layout. mass, stiffness: lists
    of n_e x n_e arrays; face_elements: [F][2] (left, right or -1); face_blocks:
    per face the column-major LL, LR, RL, RR blocks concatenated (LL only on
    boundary faces)."""

    def __init__(self, mass, stiffness, face_elements, face_blocks, n_fields=1, precision="fp64", tableau=RK4,
                 n_extra_vectors=0, device=0):
        self.lib = _lib.load()
        self.tableau = tableau
        self.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        h = ctypes.c_void_p()
        check(self.lib.hgdg_ctx_create(ctypes.byref(h), device))
        self.h = h
        sizes = np.ascontiguousarray([np.asarray(m).shape[0] for m in mass], dtype=np.int32)
        self._mass = np.ascontiguousarray(np.concatenate([np.asarray(m, dtype=np.float64).ravel(order="F")
                                                          for m in mass]))
        self._stiff = np.ascontiguousarray(np.concatenate([np.asarray(a, dtype=np.float64).ravel(order="F")
                                                           for a in stiffness]))
        self._fe = np.ascontiguousarray(face_elements, dtype=np.int64).reshape(-1, 2)
        fb = [np.asarray(b, dtype=np.float64).ravel() for b in face_blocks]
        self._fb = np.ascontiguousarray(np.concatenate(fb) if fb else np.zeros(1))
        self._sizes = sizes
        desc = _lib.MatrixDesc(len(sizes), n_fields, sizes.ctypes.data_as(_lib._I32), dptr(self._mass),
                               dptr(self._stiff), len(self._fe), self._fe.ctypes.data_as(_lib._I64), dptr(self._fb))
        check(self.lib.hgdg_operator_set_matrices(self.h, ctypes.byref(desc), self.precision))
        self._query_info()
        self.solutionVectorId_ = 0
        self.auxiliaryVectorIds_ = list(range(1, tableau.stages + 1))
        self.n_vectors = 1 + tableau.stages + n_extra_vectors
        check(self.lib.hgdg_state_alloc(self.h, self.n_vectors))

"""ctypes binding to libhgdg.so, the HIP library behind include/hgdg.h.

There is no CPU fallback: if the library is missing, loading fails loudly.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhgdg.so")

SHAPE_SQUARE, SHAPE_CUBE, SHAPE_TETRAHEDRON = 2, 4, 5
SHAPE_LINE, SHAPE_TRIANGLE, SHAPE_PRISM, SHAPE_PYRAMID = 1, 3, 6, 7
OP_ADVECTION_UPWIND, OP_ACOUSTIC_CENTRAL, OP_MATRIX = 1, 2, 3
PREC_FP64, PREC_MIXED_FP32 = 0, 1

ERRORS = {1: "invalid argument", 2: "HIP error", 3: "unsupported", 4: "call out of order", 5: "memory"}

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)
_INT = ctypes.POINTER(ctypes.c_int)


class MeshDesc(ctypes.Structure):
    _fields_ = [("shape", ctypes.c_int32), ("dim", ctypes.c_int32), ("n_elements", ctypes.c_int64),
                ("coords", _D), ("node_ids", _I64), ("order", _I32), ("uniform_order", ctypes.c_int32)]


class OpDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("precision", ctypes.c_int32), ("velocity", ctypes.c_double * 3),
                ("c_inv", _D)]


class MatrixDesc(ctypes.Structure):
    _fields_ = [("n_elements", ctypes.c_int64), ("n_fields", ctypes.c_int32), ("block_size", _I32),
                ("mass", _D), ("stiffness", _D), ("n_faces", ctypes.c_int64), ("face_elements", _I64),
                ("face_blocks", _D)]


# (name, restype, argtypes) of every entry point declared in include/hgdg.h
SIGNATURES = [
    ("hgdg_last_error", ctypes.c_char_p, []),
    ("hgdg_ctx_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    ("hgdg_ctx_destroy", ctypes.c_int, [ctypes.c_void_p]),
    ("hgdg_mesh_upload", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MeshDesc)]),
    ("hgdg_mesh_info", ctypes.c_int, [ctypes.c_void_p, _I64, _I64, _I64, _I64]),
    ("hgdg_operator_set", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(OpDesc)]),
    ("hgdg_set_advection_tiles", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hgdg_get_advection_tiles", ctypes.c_int, [ctypes.c_void_p, _INT, _D]),
    ("hgdg_set_rk_fused", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hgdg_get_rk_fused", ctypes.c_int, [ctypes.c_void_p, _INT, _I64]),
    ("hgdg_operator_set_matrices", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MatrixDesc), ctypes.c_int]),
    ("hgdg_set_tensor_general", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hgdg_get_tensor_general", ctypes.c_int, [ctypes.c_void_p, _INT, _INT]),
    ("hgdg_set_simplex_threads", ctypes.c_int, [ctypes.c_void_p, _I32]),
    ("hgdg_get_simplex_threads", ctypes.c_int, [ctypes.c_void_p, _I32, _D]),
    ("hgdg_state_alloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hgdg_state_set", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    ("hgdg_state_get", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _D]),
    ("hgdg_state_set_storage", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("hgdg_time_derivative", ctypes.c_int, [ctypes.c_void_p, _INT, _D, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_double]),
    ("hgdg_right_hand_side", ctypes.c_int, [ctypes.c_void_p, _INT, _D, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_double]),
    ("hgdg_solve_mass", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hgdg_scale_and_add", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    ("hgdg_rk_steps", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _INT, ctypes.c_int, _D, _D,
                                     ctypes.c_double, ctypes.c_int]),
    ("hgdg_project_modes", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, _D]),
    ("hgdg_project_field_modes", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D]),
    ("hgdg_step_error_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _INT, ctypes.c_int, _D,
                                               ctypes.c_double, _D, _D]),
    ("hgdg_step_error", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _INT, ctypes.c_int, _D, ctypes.c_double, _D,
                                       _D]),
    ("hgdg_rk_adaptive_step", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _INT, ctypes.c_int, _D, _D, _D,
                                             ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D,
                                             _D, _D, _D]),
    ("hgdg_error_modes", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _D, _D, _D,
                                        _D]),
    ("hgdg_project_gaussian", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, ctypes.c_double,
                                             ctypes.c_double]),
    ("hgdg_state_diff", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _D, _D]),
    ("hgdg_copy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ("hgdg_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("hgdg_get_stream", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    ("hgdg_basis_size", ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    ("hgdg_basis_eval", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _D, _D]),
    ("hgdg_hcurl_basis_size", ctypes.c_int, [ctypes.c_int]),
    ("hgdg_hcurl_basis_eval", ctypes.c_int, [ctypes.c_int, _D, _D, _D]),
    ("hgdg_structured_mesh", ctypes.c_int, [ctypes.c_int, _I32, _D, _D, _I32, ctypes.c_int, _I64, _I32, _I32,
                                            _D, _I64]),
]

_lib = None


class HgdgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hgdg error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


def load():
    """Load libhgdg.so (built by __graft_entry__.build()); raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build the HIP extension first "
                          "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        raise HgdgError(rc, load().hgdg_last_error().decode())


def dptr(a):
    return a.ctypes.data_as(_D)


class SolveDesc(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("kind", ctypes.c_int32), ("order", ctypes.c_int32),
                ("cells", ctypes.c_int32 * 3), ("velocity", ctypes.c_double * 3), ("wave_speed", ctypes.c_double),
                ("n_modes", ctypes.c_int32), ("amp", _D), ("k", _D), ("dt", ctypes.c_double),
                ("n_steps", ctypes.c_int32), ("keep_state", ctypes.c_int32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("active", ctypes.c_int64), ("latency_p50_s", ctypes.c_double), ("latency_p99_s", ctypes.c_double),
                ("latency_mean_s", ctypes.c_double), ("dof_rhs_per_s", ctypes.c_double), ("wall_s", ctypes.c_double),
                ("peak_bytes", ctypes.c_int64), ("ticks", ctypes.c_int64), ("restages", ctypes.c_int64),
                ("restage_s", ctypes.c_double), ("tick_s", ctypes.c_double)]


SIGNATURES += [
    ("hgdg_server_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int64, ctypes.c_int]),
    ("hgdg_server_submit", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SolveDesc), ctypes.POINTER(ctypes.c_int64)]),
    ("hgdg_server_wait", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]),
    ("hgdg_server_result", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), _D, _D, _D]),
    ("hgdg_server_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ServerStats)]),
    ("hgdg_server_set_profiling", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("hgdg_server_family_times", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _I32, _I32, _I32, _D, _I64,
                                                 _INT, _D]),
    ("hgdg_server_destroy", ctypes.c_int, [ctypes.c_void_p]),
]


class MeshFileInfo(ctypes.Structure):
    _fields_ = [("version", ctypes.c_int32), ("dim", ctypes.c_int32), ("shape", ctypes.c_int32),
                ("n_vertices", ctypes.c_int32), ("n_nodes", ctypes.c_int64), ("n_elements", ctypes.c_int64),
                ("n_faces", ctypes.c_int64), ("n_boundary_faces", ctypes.c_int64), ("n_edges", ctypes.c_int64),
                ("n_zones", ctypes.c_int32), ("zone_name_bytes", ctypes.c_int32)]


SIGNATURES += [
    ("hgdg_mesh_read", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(MeshFileInfo), _D, _I64, _I32, _I64, _I32,
                                      ctypes.c_char_p]),
    ("hgdg_mesh_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int64, _D, _I64, _I32,
                                       ctypes.c_char_p]),
    ("hgdg_vtk_write", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int64, _D, _I32, ctypes.c_int,
                                      ctypes.c_int, _D, ctypes.c_int64, ctypes.c_char_p]),
]


class MaxwellDesc(ctypes.Structure):
    _fields_ = [("order", ctypes.c_int32), ("epsilon", _D), ("stab", ctypes.c_double)]


SIGNATURES += [
    ("hgdg_maxwell_assemble", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(MaxwellDesc), _I64, _I32, _I64, _I64, _D,
                                             _D, _D]),
]

"""ctypes binding of libhgdt.so (the C ABI declared in include/hgdt.h).

The library is the only compute path: if it cannot be loaded, or no HIP device
is usable, every call raises -- there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libhgdt.so")

# status codes (include/hgdt.h), mapped to the reference's exception families
OK = 0
_STATUS_NAMES = {
    1: "wrong_input_given",
    2: "shapes_do_not_match",
    3: "operator_error",
    4: "integrand_error",
    5: "space_error",
    6: "mapper_error",
    7: "linear_solver_failed",
    8: "device_error",
    9: "not_supported",
    10: "out_of_memory",
}


class HgdtError(RuntimeError):
    """Base class; `status` holds the hgdt_status code."""

    def __init__(self, status, message):
        super().__init__(f"[{_STATUS_NAMES.get(status, status)}] {message}")
        self.status = status


class wrong_input_given(HgdtError):
    pass


class shapes_do_not_match(HgdtError):
    pass


class operator_error(HgdtError):
    pass


class integrand_error(HgdtError):
    pass


class space_error(HgdtError):
    pass


class mapper_error(HgdtError):
    pass


class linear_solver_failed(HgdtError):
    pass


class device_error(HgdtError):
    pass


class not_supported(HgdtError):
    pass


class out_of_memory(HgdtError):
    pass


_EXC = {
    1: wrong_input_given,
    2: shapes_do_not_match,
    3: operator_error,
    4: integrand_error,
    5: space_error,
    6: mapper_error,
    7: linear_solver_failed,
    8: device_error,
    9: not_supported,
    10: out_of_memory,
}

QUAD, TRI, HEX, TET, MIXED = 0, 1, 2, 3, 4
CONTINUOUS_LAGRANGE, DISCONTINUOUS_LAGRANGE = 0, 1
STENCIL_ELEMENT, STENCIL_INTERSECTION, STENCIL_ELEMENT_AND_INTERSECTION = 0, 1, 2
SOURCE_CONSTANT, SOURCE_PER_ELEMENT, SOURCE_PER_QUADRATURE_POINT = 0, 1, 2
FORMAT_AUTO, FORMAT_SELL, FORMAT_DIA, FORMAT_DIA_NONSYMMETRIC, FORMAT_MATRIX_FREE, FORMAT_BLOCK = 0, 1, 2, 3, 4, 5
BOUNDARY_DIRICHLET_COUPLING, BOUNDARY_PENALTY, BOUNDARY_PRODUCT = 0, 1, 2

KERNEL_ELEMENT, KERNEL_INTERSECTION, KERNEL_GATHER, KERNEL_APPLY, KERNEL_VECTOR = 0, 1, 2, 3, 4


class SolverOptions(ctypes.Structure):
    _fields_ = [
        ("rtol", ctypes.c_double),
        ("max_iter", ctypes.c_int64),
        ("check_every", ctypes.c_int32),
        ("use_graph", ctypes.c_int32),
        ("time_spmv", ctypes.c_int32),
        ("zero_initial_guess", ctypes.c_int32),
        ("true_residual", ctypes.c_int32),
        ("restarts", ctypes.c_int32),
    ]


class SolverStats(ctypes.Structure):
    _fields_ = [
        ("converged", ctypes.c_int32),
        ("breakdown", ctypes.c_int32),
        ("iterations", ctypes.c_int64),
        ("relres", ctypes.c_double),
        ("spmv_avg_ms", ctypes.c_double),
        ("spmv_count", ctypes.c_int64),
        ("solve_ms", ctypes.c_double),
        ("step_bytes", ctypes.c_int64),
        ("compactions", ctypes.c_int64),
        ("true_relres", ctypes.c_double),
        ("column_staged", ctypes.c_int32),
        ("restarts", ctypes.c_int32),
        ("iteration_bytes", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_pvp = ctypes.POINTER(ctypes.c_void_p)
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pdbl = ctypes.POINTER(ctypes.c_double)

# every function declared in include/hgdt.h: name -> (restype, argtypes)
SIGNATURES = {
    "hgdt_last_error": (ctypes.c_char_p, []),
    "hgdt_version": (ctypes.c_int, []),
    "hgdt_device_count": (ctypes.c_int, [_pi32]),
    "hgdt_memcpy": (ctypes.c_int, [_vp, _vp, _i64, _vp]),
    "hgdt_device_alloc": (ctypes.c_int, [_i64, _pvp]),
    "hgdt_device_free": (ctypes.c_int, [_vp]),
    "hgdt_stream_synchronize": (ctypes.c_int, [_vp]),
    "hgdt_grid_create_structured": (ctypes.c_int, [ctypes.c_int, _pi32, _pdbl, _pdbl, _pvp]),
    "hgdt_grid_create": (ctypes.c_int, [ctypes.c_int, _i64, _pdbl, _i64, _pi32, _pvp]),
    "hgdt_grid_create_mixed": (ctypes.c_int, [_i64, _pdbl, _i64, _pi32, _pi64, _pi32, _pvp]),
    "hgdt_grid_create_union": (ctypes.c_int, [_pvp, _i32, _pvp]),
    "hgdt_grid_destroy": (ctypes.c_int, [_vp]),
    "hgdt_grid_info": (ctypes.c_int, [_vp, _pi64]),
    "hgdt_grid_copy_coords": (ctypes.c_int, [_vp, _pdbl]),
    "hgdt_grid_copy_connectivity": (ctypes.c_int, [_vp, _pi32]),
    "hgdt_grid_instance_elements": (ctypes.c_int, [_vp, _pi64]),
    "hgdt_grid_quadrature_points": (ctypes.c_int, [_vp, ctypes.c_int, _pdbl, _pi32]),
    "hgdt_grid_quadrature_offsets": (ctypes.c_int, [_vp, ctypes.c_int, _pi64]),
    "hgdt_grid_boundary_quadrature_points": (ctypes.c_int, [_vp, ctypes.c_int, _pdbl, _pi32]),
    "hgdt_space_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _pvp]),
    "hgdt_space_destroy": (ctypes.c_int, [_vp]),
    "hgdt_space_size": (ctypes.c_int, [_vp, _pi64, _pi32]),
    "hgdt_space_copy_global_indices": (ctypes.c_int, [_vp, _pi32]),
    "hgdt_space_instance_dofs": (ctypes.c_int, [_vp, _pi64]),
    "hgdt_space_error_norms": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp, _vp, _vp, _dbl, _pdbl, _vp]),
    "hgdt_operator_create": (ctypes.c_int, [_vp, ctypes.c_int, _pvp]),
    "hgdt_operator_create_between": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _pvp]),
    "hgdt_operator_create_on_matrix": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _pi64, _pi32, ctypes.c_int64,
                                                      ctypes.c_int64, _vp, _pvp]),
    "hgdt_operator_destroy": (ctypes.c_int, [_vp]),
    "hgdt_operator_shape": (ctypes.c_int, [_vp, _pi64, _pi64, _pi64]),
    "hgdt_operator_copy_pattern": (ctypes.c_int, [_vp, _pi64, _pi32]),
    "hgdt_operator_append_laplace": (ctypes.c_int, [_vp, _vp, _dbl, _i32, _dbl]),
    "hgdt_operator_append_product": (ctypes.c_int, [_vp, _vp, _dbl, _i32, _dbl]),
    "hgdt_operator_append_laplace_points": (ctypes.c_int, [_vp, _vp, _i32, _i32, _dbl]),
    "hgdt_operator_set_boundary_source_tensor": (ctypes.c_int, [_vp, _vp]),
    "hgdt_operator_set_face_tensors": (ctypes.c_int, [_vp, _i32, _vp, _vp]),
    "hgdt_operator_append_laplace_tensor": (ctypes.c_int, [_vp, _vp, _vp, _i32, _dbl]),
    "hgdt_operator_append_product_points": (ctypes.c_int, [_vp, _vp, _i32, _i32, _dbl]),
    "hgdt_operator_append_inner_coupling": (ctypes.c_int, [_vp, _dbl, _vp, _dbl, _vp, _dbl, _i32, _dbl]),
    "hgdt_operator_append_inner_penalty": (ctypes.c_int, [_vp, _dbl, _vp, _dbl, _i32, _dbl]),
    "hgdt_operator_append_dirichlet_coupling": (ctypes.c_int, [_vp, _dbl, _vp, _dbl, _i32, _dbl]),
    "hgdt_operator_append_boundary_penalty": (ctypes.c_int, [_vp, _dbl, _vp, _dbl, _i32, _dbl]),
    "hgdt_operator_append_source": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _dbl, _i32, _i32]),
    "hgdt_operator_append_boundary_source": (ctypes.c_int, [_vp, ctypes.c_int, _dbl, _vp, _dbl, _vp, _dbl, _i32, _i32]),
    "hgdt_operator_set_dirichlet": (ctypes.c_int, [_vp, _i32]),
    "hgdt_operator_set_dirichlet_mode": (ctypes.c_int, [_vp, _i32, _i32]),
    "hgdt_operator_dirichlet_dofs": (ctypes.c_int, [_vp, _pi64, _pi32]),
    "hgdt_operator_assemble": (ctypes.c_int, [_vp, _vp]),
    "hgdt_operator_values_device": (ctypes.c_int, [_vp, _pvp]),
    "hgdt_operator_rhs_device": (ctypes.c_int, [_vp, _pvp]),
    "hgdt_operator_copy_values": (ctypes.c_int, [_vp, _pdbl]),
    "hgdt_operator_copy_rhs": (ctypes.c_int, [_vp, _pdbl]),
    "hgdt_operator_set_format": (ctypes.c_int, [_vp, ctypes.c_int]),
    "hgdt_operator_format_info": (ctypes.c_int, [_vp, _pi32, _pi64, _pi32]),
    "hgdt_operator_apply": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "hgdt_solver_options_default": (ctypes.c_int, [ctypes.POINTER(SolverOptions)]),
    "hgdt_operator_apply_inverse": (
        ctypes.c_int,
        [_vp, _vp, _vp, ctypes.POINTER(SolverOptions), ctypes.POINTER(SolverStats), _vp],
    ),
    "hgdt_operator_instance_stats": (ctypes.c_int, [_vp, _pi64, _pdbl]),
    "hgdt_operator_instance_true_residuals": (ctypes.c_int, [_vp, _pdbl]),
    "hgdt_space_error_norms_noalloc": (ctypes.c_int, [_vp, _vp, _i32, _vp, _vp, _vp, _dbl, _pdbl, _vp]),
    "hgdt_noalloc_scratch_bytes": (ctypes.c_int, [_pi64]),
    "hgdt_operator_assembly_info": (ctypes.c_int, [_vp, _pi32, _pi32]),
    "hgdt_operator_assembly_bytes": (ctypes.c_int, [_vp, _pi64]),
    "hgdt_grid_boundary_intersections": (ctypes.c_int, [_vp, _pdbl, _pdbl]),
    "hgdt_operator_set_boundary_filter": (ctypes.c_int, [_vp, _vp, _i64]),
    "hgdt_operator_set_dirichlet_faces": (ctypes.c_int, [_vp, _vp, _i64]),
    "hgdt_operator_adjoint_log_diffusion_gradient": (ctypes.c_int, [_vp, _i32, _vp, _vp, _vp, _vp]),
    "hgdt_operator_set_launch": (ctypes.c_int, [_vp, _i32, _i32, _i32]),
    "hgdt_operator_get_launch": (ctypes.c_int, [_vp, _i32, _pi32, _pi32]),
    "hgdt_operator_set_column_staging": (ctypes.c_int, [_vp, _i32]),
}

_lib = None


def load():
    """Load libhgdt.so (raises if it is missing: build it with __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libhgdt.so not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != OK:
        msg = load().hgdt_last_error().decode(errors="replace")
        raise _EXC.get(status, HgdtError)(status, msg)


def call(name, *args):
    check(getattr(load(), name)(*args))

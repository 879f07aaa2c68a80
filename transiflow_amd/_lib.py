"""ctypes binding of ``libtransiflow_hip.so`` (C ABI in include/transiflow_hip.h).

The library is built in-tree by ``python -m transiflow_amd.build``.  There is
no fallback: if the library is missing every backend call raises.
"""

import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libtransiflow_hip.so')

_c_int = ctypes.c_int
_c_long = ctypes.c_long
_c_double = ctypes.c_double
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    'tf_last_error': (ctypes.c_char_p, []),
    'tf_device_count': (_c_int, []),
    'tf_prof_enable': (_c_int, [_c_int]),
    'tf_prof_read': (_c_int, [_c_int, ctypes.POINTER(_c_long), _dp]),
    'tf_capture_fallbacks': (_c_int, []),
    'tf_prof_reset': (_c_int, []),
    'tf_prof_work': (_c_int, [_c_int, ctypes.POINTER(ctypes.c_double)]),
    'tf_problem_create': (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_int, _c_int,
                                   _dp, _dp, _dp, _c_int, _c_int, _c_int]),
    'tf_problem_destroy': (_c_int, [_vp]),
    'tf_problem_set_physics': (_c_int, [_vp, _c_double, _c_double, _c_double, _c_double, _c_int, _c_int]),
    'tf_problem_set_bcs': (_c_int, [_vp, _c_int, _ip, _ip, _ip, _ip, _dp, _dp, _dp]),
    'tf_problem_rows': (_c_int, [_vp]),
    'tf_rhs': (_c_int, [_vp, _vp, _vp, _vp]),
    'tf_jacobian_count': (_c_int, [_vp, _vp, _vp, _ip, _vp]),
    'tf_jacobian_fill': (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'tf_mass_count': (_c_int, [_vp, _vp, _ip, _vp]),
    'tf_mass_fill': (_c_int, [_vp, _vp, _vp, _vp, _vp]),
    'tf_axpby_f64': (_c_int, [_c_long, _c_double, _vp, _c_double, _vp, _vp, _vp]),
    'tf_scale_f64': (_c_int, [_c_long, _c_double, _vp, _vp, _vp]),
    'tf_div_f64': (_c_int, [_c_long, _vp, _c_double, _vp, _vp]),
    'tf_f64_to_f32': (_c_int, [_c_long, _vp, _vp, _vp]),
    'tf_f32_to_f64': (_c_int, [_c_long, _vp, _vp, _vp]),
    'tf_reduce_work_size': (_c_int, []),
    'tf_dot_f64': (_c_int, [_c_long, _vp, _vp, _vp, _vp, _dp, _c_int, _vp]),
    'tf_spmv_f64': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _vp]),
    'tf_refine_residual': (_c_int, [_c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    'tf_refine_update': (_c_int, [_c_long, _vp, _vp, _vp, _vp]),
    'tf_read_f64': (_c_int, [_vp, _c_int, _dp, _vp]),
    'tf_spmv_f32': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _vp]),
    'tf_set_entry_f64': (_c_int, [_vp, _c_long, _c_long, _c_double, _vp]),
    'tf_fill_f64': (_c_int, [_c_long, _c_double, _vp, _vp]),
    'tf_theta_residual': (_c_int, [_c_long, _vp, _vp, _vp, _vp, _vp, _c_double, _c_double, _vp, _vp]),
    'tf_csr_diagonal': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    'tf_csr_interleave_count': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _ip, _vp]),
    'tf_csr_interleave_fill': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _vp, _vp, _vp, _vp]),
    'tf_krylov_work_size': (_c_int, []),
    'tf_krylov_max_basis': (_c_int, []),
    'tf_multidot_f64': (_c_int, [_c_int, _c_int, _vp, _c_long, _vp, _vp, _vp, _c_int, _vp]),
    'tf_multidot_f32': (_c_int, [_c_int, _c_int, _vp, _c_long, _vp, _vp, _vp, _c_int, _vp]),
    'tf_multiaxpy_f64': (_c_int, [_c_int, _c_int, _vp, _c_long, _vp, _vp, _c_double, _vp]),
    'tf_multiaxpy_f32': (_c_int, [_c_int, _c_int, _vp, _c_long, _vp, _vp, _c_double, _vp]),
    'tf_normalize_f64': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    'tf_normalize_f32': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp]),
    'tf_cellblock_setup': (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp]),
    'tf_cellblock_apply_f64': (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    'tf_cellblock_apply_f32': (_c_int, [_c_int, _c_int, _vp, _vp, _vp, _vp]),
    'tf_linelu_create': (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int]),
    'tf_linelu_destroy': (_c_int, [_vp]),
    'tf_linelu_factor': (_c_int, [_vp, _vp, _vp, _vp, _c_int, _vp]),
    'tf_linelu_status': (_c_int, [_vp, _ip, _vp]),
    'tf_linelu_solve_f64': (_c_int, [_vp, _vp, _vp, _vp]),
    'tf_linelu_solve_f32': (_c_int, [_vp, _vp, _vp, _vp]),
    'tf_conv_fwd': (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                             _c_int, _vp, _vp, _vp]),
    'tf_conv_wgrad_work_size': (_c_int, [_c_long, _c_int, _c_int]),
    'tf_conv_wgrad': (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                               _vp, _vp, _vp]),
    'tf_mse_grad': (_c_int, [_c_long, _vp, _vp, _vp, _vp, _vp]),
    'tf_relu_bwd': (_c_int, [_c_long, _vp, _vp, _c_int, _vp, _vp, _vp]),
    'tf_conv_pack': (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp]),
    'tf_adam': (_c_int, [_c_long, _vp, _vp, _vp, _vp, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                         ctypes.c_float, _c_int, _vp]),
    'tf_to_bf16': (_c_int, [_c_long, _vp, ctypes.c_float, _vp, _vp]),
    'tf_krylov_create': (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int]),
    'tf_krylov_destroy': (_c_int, [_vp]),
    'tf_krylov_set_graph': (_c_int, [_vp, _c_int]),
    'tf_fgmres': (_c_int, [_vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_int, _c_int, _vp, _vp, _c_double, _c_int,
                           _c_int, _ip, _dp, _vp]),
    'tf_fgmres2': (_c_int, [_vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp,
                            _c_double, _c_double, _c_int, _c_int, _ip, _dp, _vp]),
    'tf_csr_add_count': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _ip, _vp]),
    'tf_csr_add_count_noalloc': (_c_int, [_c_int, _vp, _vp, _vp, _vp, _vp, _ip, _vp]),
    'tf_csr_add_fill': (_c_int, [_c_int, _vp, _vp, _vp, _c_double, _vp, _vp, _vp, _c_double, _vp, _vp, _vp, _vp]),
    'tf_gemm_tuning': (_c_int, [_c_int]),
    'tf_gemm_table': (_c_int, [_ip, _c_int]),
    'tf_bcr_create': (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_int, _c_int, _c_int]),
    'tf_bcr_destroy': (_c_int, [_vp]),
    'tf_bcr_factor': (_c_int, [_vp, _vp, _vp, _vp, _c_int, _vp]),
    'tf_bcr_status': (_c_int, [_vp, _ip, _vp]),
    'tf_bcr_solve_f64': (_c_int, [_vp, _vp, _vp, _vp]),
    'tf_bcr_solve_f32': (_c_int, [_vp, _vp, _vp, _vp]),
    'tf_bcr_solve2_f64': (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'tf_bcr_solve2_f32': (_c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    'tf_bcr_set_graph': (_c_int, [_vp, _c_int]),
    'tf_bcr_set_top': (_c_int, [_vp, _c_int]),
    'tf_bcr_set_small': (_c_int, [_vp, _c_int]),
    'tf_bcr_top_errors': (_c_int, [_vp, _ip, _vp]),
    'tf_bcr_set_half': (_c_int, [_vp, _c_int, _vp]),
    'tf_bcr_set_gemm_split': (_c_int, [_vp, _c_int]),
    'tf_bcr_set_fused': (_c_int, [_vp, _c_int]),
    'tf_bcr_set_dense_top': (_c_int, [_vp, _c_int]),
    'tf_lineilu_create': (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_int, _c_int]),
    'tf_lineilu_destroy': (_c_int, [_vp]),
    'tf_lineilu_factor': (_c_int, [_vp, _vp, _vp, _vp, _c_int, _vp]),
    'tf_lineilu_status': (_c_int, [_vp, _ip, _vp]),
    'tf_lineilu_solve_f64': (_c_int, [_vp, _vp, _vp, _vp]),
    'tf_lineilu_solve_f32': (_c_int, [_vp, _vp, _vp, _vp]),
    'tf_mg_restrict_state': (_c_int, [_c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    'tf_mg_create': (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_int, _c_int, _c_int]),
    'tf_mg_destroy': (_c_int, [_vp]),
    'tf_mg_set': (_c_int, [_vp, _vp, _vp, _vp, _vp, _c_int, _vp, _c_int]),
    'tf_mg_solve_f64': (_c_int, [_vp, _vp, _vp, _vp]),
    'tf_mg_solve_f32': (_c_int, [_vp, _vp, _vp, _vp]),
    'tf_fill_f32': (_c_int, [_c_long, _c_double, _vp, _vp]),
    'tf_axpby_f32': (_c_int, [_c_long, _c_double, _vp, _c_double, _vp, _vp, _vp]),
    'tf_scale_f32': (_c_int, [_c_long, _c_double, _vp, _vp, _vp]),
}

_lib = None


class BackendError(RuntimeError):
    """A HIP backend call failed; the message is the library's last error."""


def load():
    """Load the shared library and declare every exported signature.

    PyTorch is imported first: it carries its own libamdhip64.so.7 (same
    SONAME as /opt/rocm's), and loading it before our library makes the
    dynamic linker bind our library to that one HIP runtime instead of
    bringing a second runtime into the process.
    """
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BackendError('HIP backend library not built (%s); run `python -m transiflow_amd.build`'
                           % LIB_PATH)
    import torch  # noqa: F401  (see docstring)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def call(name, *args):
    """Call ``name`` and raise BackendError on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise BackendError('%s failed (%d): %s' % (name, rc, lib.tf_last_error().decode()))
    return rc


def dptr(arr):
    """ctypes double* of a contiguous numpy float64 array."""
    return arr.ctypes.data_as(_dp)


def iptr(arr):
    return arr.ctypes.data_as(_ip)

"""ctypes binding of the in-tree HIP library libtbem_fmm.so (C ABI: include/tbem_fmm.h).

The library is the only compute path: if it is not built, or no GPU is present,
the operators raise instead of falling back to anything else.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtbem_fmm.so")

_D = ctypes.POINTER(ctypes.c_double)
_L = ctypes.POINTER(ctypes.c_long)
_SZ = ctypes.c_size_t
_SZP = ctypes.POINTER(ctypes.c_size_t)
_VP = ctypes.c_void_p

# every symbol declared in include/tbem_fmm.h
EXPORTED = (
    "tbem_fmm_create", "tbem_fmm_plan_save", "tbem_fmm_create_from_plan", "tbem_fmm_apply", "tbem_fmm_apply_dev", "tbem_fmm_apply_adjoint",
    "tbem_fmm_apply_adjoint_dev", "tbem_fmm_obs_grad", "tbem_fmm_obs_grad_dev",
    "tbem_fmm_forward_token", "tbem_fmm_obs_grad_reuse_dev",
    "tbem_fmm_set_obs_locs", "tbem_fmm_set_obs_locs_dev", "tbem_fmm_obs_version", "tbem_fmm_n_rows", "tbem_fmm_n_cols", "tbem_fmm_clone",
    "tbem_fmm_destroy", "tbem_fmm_info", "tbem_fmm_c2e", "tbem_fmm_pairs", "tbem_fmm_octree",
    "tbem_fmm_stage_times", "tbem_fmm_m2l_ms", "tbem_fmm_set_overlap", "tbem_nbody_create", "tbem_nbody_apply", "tbem_nbody_apply_dev", "tbem_nbody_apply_noalloc",
    "tbem_nbody_apply_adjoint", "tbem_nbody_n_rows", "tbem_nbody_n_cols", "tbem_nbody_destroy",
    "tbem_sparse_create", "tbem_sparse_from_coo", "tbem_sparse_apply", "tbem_sparse_apply_dev",
    "tbem_sparse_csr", "tbem_sparse_clone", "tbem_sparse_right_multiply", "tbem_sparse_n_rows", "tbem_sparse_n_cols", "tbem_sparse_nnz",
    "tbem_sparse_destroy", "tbem_integral_create", "tbem_integral_apply", "tbem_integral_apply_dev",
    "tbem_integral_n_rows", "tbem_integral_n_cols", "tbem_integral_destroy",
    "tbem_constraints_create", "tbem_constraints_size", "tbem_constraints_n_terms", "tbem_constraints_get",
    "tbem_constraints_homogenize", "tbem_distribute_vector", "tbem_distribute_vector_dev", "tbem_condense_vector",
    "tbem_condense_vector_dev", "tbem_condense_sparse", "tbem_constraints_destroy",
    "tbem_kernel_eval", "tbem_fp64_sqrt_div", "tbem_surface", "tbem_pinv", "tbem_last_error", "tbem_device_available",
)
FARFIELD_FMM, FARFIELD_DIRECT = 0, 1

FP64, MIXED = 0, 1


class FmmDesc(ctypes.Structure):
    _fields_ = [
        ("kernel", ctypes.c_int), ("shear_modulus", ctypes.c_double), ("poisson_ratio", ctypes.c_double),
        ("n_obs", _SZ), ("obs_locs", _D), ("obs_normals", _D),
        ("n_src", _SZ), ("src_locs", _D), ("src_normals", _D), ("src_weights", _D),
        ("mac", ctypes.c_double), ("order", _SZ), ("min_pts_per_cell", _SZ), ("d", ctypes.c_double),
        ("account_for_small_cells", ctypes.c_int), ("precision", ctypes.c_int),
        ("up_check_to_equiv", _D), ("n_up_levels", _SZ),
        ("down_check_to_equiv", _D), ("n_down_levels", _SZ), ("tree_boxes", ctypes.c_int),
        ("translations", ctypes.c_int), ("m2l_output_budget", _SZ), ("point_pairs", ctypes.c_int),
    ]


class FmmInfo(ctypes.Structure):
    _fields_ = [
        ("n_e", _SZ), ("n_src_cells", _SZ), ("n_obs_cells", _SZ), ("n_up_levels", _SZ),
        ("n_down_levels", _SZ), ("p2m", _SZ), ("m2m", _SZ), ("p2p", _SZ), ("m2p", _SZ), ("p2l", _SZ),
        ("m2l", _SZ), ("l2l", _SZ), ("l2p", _SZ), ("p2p_pairs", ctypes.c_double),
        ("m2l_classes", ctypes.c_double), ("plan_seconds", ctypes.c_double),
        ("m2l_output_bytes", ctypes.c_double), ("work_bytes", ctypes.c_double),
        ("leaf_wave_items", ctypes.c_int), ("leaf_chunked", ctypes.c_int), ("leaf_block", ctypes.c_int),
        ("leaf_symmetric", ctypes.c_int), ("symmetric_pairs", ctypes.c_double),
    ]


class NbodyDesc(ctypes.Structure):
    _fields_ = [
        ("kernel", ctypes.c_int), ("shear_modulus", ctypes.c_double), ("poisson_ratio", ctypes.c_double),
        ("n_obs", _SZ), ("obs_locs", _D), ("obs_normals", _D),
        ("n_src", _SZ), ("src_locs", _D), ("src_normals", _D), ("src_weights", _D),
        ("precision", ctypes.c_int),
    ]


class TbemError(RuntimeError):
    pass


_lib = None


def load(path=LIB_PATH):
    """Load libtbem_fmm.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TbemError(f"{path} is not built; run `python 3bem_amd/build.py` (or __graft_entry__.build())")
    # One HIP runtime per process: torch wheels bundle their own libamdhip64.so.7.
    # Loading torch first makes this library bind to that same copy (same SONAME),
    # so device pointers and streams from torch are valid here and vice versa.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(path)
    sig = {
        "tbem_fmm_create": (ctypes.c_int, [ctypes.POINTER(FmmDesc), ctypes.POINTER(_VP)]),
        "tbem_fmm_plan_save": (ctypes.c_int, [_VP, _VP, _SZ, ctypes.POINTER(_SZ)]),
        "tbem_fmm_create_from_plan": (ctypes.c_int, [ctypes.POINTER(FmmDesc), _VP, _SZ, ctypes.POINTER(_VP)]),
        "tbem_fmm_apply": (ctypes.c_int, [_VP, _D, _D]),
        "tbem_fmm_apply_dev": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tbem_fmm_apply_adjoint": (ctypes.c_int, [_VP, _D, _D]),
        "tbem_fmm_apply_adjoint_dev": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tbem_fmm_obs_grad": (ctypes.c_int, [_VP, _D, _D, _D]),
        "tbem_fmm_obs_grad_dev": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP]),
        "tbem_fmm_forward_token": (ctypes.c_uint64, [_VP]),
        "tbem_fmm_obs_grad_reuse_dev": (ctypes.c_int, [_VP, ctypes.c_uint64, _VP, _VP, _VP, _VP]),
        "tbem_fmm_set_obs_locs": (ctypes.c_int, [_VP, _D]),
        "tbem_fmm_set_obs_locs_dev": (ctypes.c_int, [_VP, _VP, _VP]),
        "tbem_fmm_obs_version": (ctypes.c_uint64, [_VP]),
        "tbem_fmm_n_rows": (_SZ, [_VP]),
        "tbem_fmm_n_cols": (_SZ, [_VP]),
        "tbem_fmm_clone": (ctypes.c_int, [_VP, ctypes.POINTER(_VP)]),
        "tbem_fmm_destroy": (None, [_VP]),
        "tbem_fmm_info": (ctypes.c_int, [_VP, ctypes.POINTER(FmmInfo)]),
        "tbem_fmm_c2e": (ctypes.c_int, [_VP, ctypes.c_int, _SZ, _D]),
        "tbem_fmm_pairs": (ctypes.c_int, [_VP, ctypes.c_int, _D, ctypes.POINTER(_SZ)]),
        "tbem_fmm_octree": (ctypes.c_int, [_VP, ctypes.c_int, _D, _L, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
        "tbem_fmm_stage_times": (ctypes.c_int, [_VP, _D]),
        "tbem_fmm_m2l_ms": (ctypes.c_int, [_VP, _D]),
        "tbem_fmm_set_overlap": (ctypes.c_int, [_VP, ctypes.c_int]),
        "tbem_nbody_create": (ctypes.c_int, [ctypes.POINTER(NbodyDesc), ctypes.POINTER(_VP)]),
        "tbem_nbody_apply": (ctypes.c_int, [_VP, _D, _D]),
        "tbem_nbody_apply_dev": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tbem_nbody_apply_noalloc": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tbem_nbody_apply_adjoint": (ctypes.c_int, [_VP, _D, _D]),
        "tbem_nbody_n_rows": (_SZ, [_VP]),
        "tbem_nbody_n_cols": (_SZ, [_VP]),
        "tbem_nbody_destroy": (None, [_VP]),
        "tbem_sparse_create": (ctypes.c_int, [_SZ, _SZ, _SZ, _D, _SZP, _SZP, ctypes.POINTER(_VP)]),
        "tbem_sparse_from_coo": (ctypes.c_int, [_SZ, _SZ, _SZ, _SZP, _SZP, _D, ctypes.POINTER(_VP)]),
        "tbem_sparse_apply": (ctypes.c_int, [_VP, _D, _D]),
        "tbem_sparse_apply_dev": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tbem_sparse_csr": (ctypes.c_int, [_VP, _D, _SZP, _SZP]),
        "tbem_sparse_clone": (ctypes.c_int, [_VP, ctypes.POINTER(_VP)]),
        "tbem_sparse_right_multiply": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(_VP)]),
        "tbem_sparse_n_rows": (_SZ, [_VP]),
        "tbem_sparse_n_cols": (_SZ, [_VP]),
        "tbem_sparse_nnz": (_SZ, [_VP]),
        "tbem_sparse_destroy": (None, [_VP]),
        "tbem_integral_create": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int, _VP, _VP, ctypes.POINTER(_VP)]),
        "tbem_integral_apply": (ctypes.c_int, [_VP, _D, _D]),
        "tbem_integral_apply_dev": (ctypes.c_int, [_VP, _VP, _VP, _VP]),
        "tbem_integral_n_rows": (_SZ, [_VP]),
        "tbem_integral_n_cols": (_SZ, [_VP]),
        "tbem_integral_destroy": (None, [_VP]),
        "tbem_constraints_create": (ctypes.c_int, [_SZ, _SZP, _SZP, _D, _D, ctypes.POINTER(_VP)]),
        "tbem_constraints_size": (_SZ, [_VP]),
        "tbem_constraints_n_terms": (_SZ, [_VP]),
        "tbem_constraints_get": (ctypes.c_int, [_VP, _SZP, _SZP, _SZP, _D, _D]),
        "tbem_constraints_homogenize": (ctypes.c_int, [_VP, ctypes.POINTER(_VP)]),
        "tbem_distribute_vector": (ctypes.c_int, [_VP, _D, _SZ, _D]),
        "tbem_distribute_vector_dev": (ctypes.c_int, [_VP, _VP, _SZ, _VP, _VP]),
        "tbem_condense_vector": (ctypes.c_int, [_VP, _D, _SZ, _D]),
        "tbem_condense_vector_dev": (ctypes.c_int, [_VP, _VP, _SZ, _VP, _VP]),
        "tbem_condense_sparse": (ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(_VP)]),
        "tbem_constraints_destroy": (None, [_VP]),
        "tbem_kernel_eval": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, _SZ, _D, _D, _D, _D, _D]),
        "tbem_fp64_sqrt_div": (ctypes.c_int, [_SZ, _D, _D, _D, _D]),
        "tbem_surface": (ctypes.c_long, [_SZ, _D]),
        "tbem_pinv": (ctypes.c_int, [_D, ctypes.c_int, ctypes.c_double, _D]),
        "tbem_last_error": (ctypes.c_char_p, []),
        "tbem_device_available": (ctypes.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != 0:
        msg = _lib.tbem_last_error().decode(errors="replace") if _lib else "?"
        raise TbemError(f"tbem error {code}: {msg}")


def require_device():
    lib = load()
    if not lib.tbem_device_available():
        raise TbemError("no HIP device available: the MI355X operators need a GPU")
    return lib

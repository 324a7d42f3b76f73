"""ctypes binding of the C ABI in ``include/bartels_hip.h`` (``bartels_amd/libbartels_hip.so``).

The product path has no fallback: if the HIP library is missing this module raises, loudly.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbartels_hip.so")

c_int = ctypes.c_int
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
P = ctypes.POINTER

# status codes (bh_status)
BH_OK, BH_BAD_ARG, BH_NONFINITE, BH_SOLVE_FAILED, BH_MAXIT, BH_DEVICE_ERROR, BH_UNSUPPORTED = range(7)
STATUS_NAMES = ["BH_OK", "BH_BAD_ARG", "BH_NONFINITE", "BH_SOLVE_FAILED", "BH_MAXIT", "BH_DEVICE_ERROR",
                "BH_UNSUPPORTED"]

# models (bh_model)
NEOHOOKEAN, STVK, COROTATIONAL, ARAP, NEOHOOKEAN_LEGACY = 0, 1, 2, 3, 4
MODEL_PER_TET = 255


class BhPsd(ctypes.Structure):
    _fields_ = [("enabled", c_int), ("tol", c_double)]


class BhElementOpts(ctypes.Structure):
    _fields_ = [("stretch_reference_layout", c_int)]


class BhTuning(ctypes.Structure):
    """schedule knobs of a mesh: gather lanes and XCD grouping (same per-slot fold order), and the vertex
    colouring of the colour sweeps (0 keep, 1 DSatur, 2 greedy) and the lanes per vertex of the Gauss-Seidel
    sweeps (0 keep, 4, 8)"""
    _fields_ = [("gather_lanes", c_int), ("xcd_groups", c_int), ("coloring", c_int), ("sweep_lanes", c_int),
                ("gather_order", c_int)]


# void (*on_system)(void* user, int iteration, const float* vals32, const double* vals64, const double* b,
#                   const double* x)
ON_SYSTEM = ctypes.CFUNCTYPE(None, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p)
# bh_element_callback: (element_matrix double[144], element, user)
ELEMENT_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), c_int, c_void_p)


class BhNewtonOpts(ctypes.Structure):
    _fields_ = [
        ("model", c_int),
        ("d_model_ids", c_void_p),
        ("psd", BhPsd),
        ("gradient_tol", c_double),
        ("relative_tol", c_int),
        ("max_iterations", c_int),
        ("max_iterations_ls", c_int),
        ("alpha", c_double),
        ("c", c_double),
        ("p", c_double),
        ("armijo_with_alpha", c_int),
        ("reject_nonfinite", c_int),
        ("pcg_tol", c_double),
        ("pcg_max_iterations", c_int),
        ("precond", c_int),
        ("precond_handle", c_void_p),
        ("precond_shift", c_double),
        ("element_precision", c_int),
        ("jacobi_fallback", c_int),
        ("mgn", c_void_p),
        ("on_system", ON_SYSTEM),
        ("user", c_void_p),
        ("d_mass_vals", c_void_p),
        ("d_q_pred", c_void_p),
        ("inertia", c_double),
        ("relaxation", c_double),
        ("coarse_aggregates", c_int),
        ("coarse_correction", c_int),
    ]


class BhNewtonReport(ctypes.Structure):
    _fields_ = [
        ("iterations", c_int),
        ("converged", c_int),
        ("pcg_iterations_total", c_int),
        ("pcg_failures", c_int),
        ("pcg_fallbacks", c_int),
        ("ls_iterations_total", c_int),
        ("energy", c_double),
        ("gradient_norm", c_double),
        ("gradient_norm0", c_double),
    ]


class BartelsError(RuntimeError):
    def __init__(self, status, where):
        name = STATUS_NAMES[status] if 0 <= status < len(STATUS_NAMES) else str(status)
        super().__init__(f"{where} failed with {name}")
        self.status = status


# every symbol include/bartels_hip.h declares: name -> (restype, argtypes)
_VP = c_void_p
SIGNATURES = {
    "bh_linear_tetmesh_dphi_dX": (c_int, [c_int, _VP, c_int, _VP, _VP, _VP]),
    "bh_linear_tri2dmesh_dphi_dX": (c_int, [c_int, _VP, c_int, _VP, _VP]),
    "bh_linear_tri3dmesh_dphi_dX": (c_int, [c_int, _VP, c_int, _VP, _VP]),
    "bh_trimesh_create": (c_int, [c_int, c_int, c_int, _VP, P(c_void_p)]),
    "bh_trimesh_destroy": (c_int, [_VP]),
    "bh_trimesh_nnz": (c_int, [_VP, P(c_int64)]),
    "bh_trimesh_pattern": (c_int, [_VP, _VP, _VP]),
    "bh_linear_tri2dmesh_q": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, P(c_double)]),
    "bh_linear_tri2dmesh_dq": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, _VP]),
    "bh_linear_tri2dmesh_dq2": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, P(BhPsd), _VP]),
    "bh_linear_tri2dmesh_q_device": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, _VP, _VP]),
    "bh_linear_tri2dmesh_dq_device": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, _VP, _VP]),
    "bh_linear_tri2dmesh_dq2_device": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, P(BhPsd), _VP, _VP]),
    "bh_linear_tri2dmesh_mass_matrix": (c_int, [_VP, _VP, _VP, _VP]),
    "bh_linear_tri3dmesh_stvk_dq2": (c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, P(BhPsd), _VP]),
    "bh_linear_tri3dmesh_stvk_dq2_device": (c_int, [_VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, P(BhPsd), _VP, _VP]),
    "bh_linear_tetmesh_dphi_dX_device": (c_int, [_VP, _VP, _VP, _VP, _VP]),
    "bh_mesh_create": (c_int, [c_int, c_int, _VP, _VP, _VP, P(c_void_p)]),
    "bh_mesh_create_device": (c_int, [c_int, c_int, _VP, _VP, _VP, _VP, P(c_void_p)]),
    "bh_mesh_destroy": (c_int, [_VP]),
    "bh_mesh_trim": (c_int, [_VP]),
    "bh_mesh_nnz": (c_int, [_VP, P(c_int64)]),
    "bh_mesh_pattern": (c_int, [_VP, _VP, _VP]),
    "bh_mesh_block_pattern_device": (c_int, [_VP, P(c_void_p), P(c_void_p), P(c_int64)]),
    "bh_mesh_set_element_opts": (c_int, [_VP, P(BhElementOpts)]),
    "bh_mesh_set_rest_positions": (c_int, [_VP, _VP]),
    "bh_mesh_set_gather_lanes": (c_int, [_VP, c_int]),
    "bh_mesh_set_tuning": (c_int, [_VP, P(BhTuning)]),
    "bh_linear_tetmesh_q": (c_int, [_VP, c_int, _VP, _VP, P(c_double)]),
    "bh_linear_tetmesh_dq": (c_int, [_VP, c_int, _VP, _VP, _VP]),
    "bh_linear_tetmesh_dq2": (c_int, [_VP, c_int, _VP, _VP, P(BhPsd), _VP]),
    "bh_linear_tetmesh_dq2_callback": (c_int, [_VP, c_int, _VP, _VP, _VP, ELEMENT_CALLBACK, _VP, _VP]),
    "bh_linear_tetmesh_dq2_callback_device": (c_int, [_VP, c_int, _VP, _VP, _VP, ELEMENT_CALLBACK, _VP, _VP, _VP]),
    "bh_linear_tetmesh_generic_elastic_q": (c_int, [_VP, _VP, _VP, _VP, P(c_double)]),
    "bh_linear_tetmesh_generic_elastic_dq": (c_int, [_VP, _VP, _VP, _VP, _VP]),
    "bh_linear_tetmesh_generic_elastic_dq2": (c_int, [_VP, _VP, _VP, _VP, P(BhPsd), _VP]),
    "bh_linear_tetmesh_q_device": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, _VP]),
    "bh_linear_tetmesh_q_device_noalloc": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, _VP]),
    "bh_linear_tetmesh_dq_device": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, _VP]),
    "bh_linear_tetmesh_dq2_device": (c_int, [_VP, c_int, _VP, _VP, _VP, P(BhPsd), _VP, _VP]),
    "bh_linear_tetmesh_mass_matrix_device": (c_int, [_VP, _VP, _VP, _VP]),
    "bh_linear_tetmesh_const_ext_force_device": (c_int, [_VP, _VP, _VP, _VP]),
    "bh_newton_default_opts": (None, [P(BhNewtonOpts)]),
    "bh_newton_static_device": (c_int, [_VP, P(BhNewtonOpts), _VP, _VP, _VP, _VP, P(BhNewtonReport), _VP, _VP]),
    "bh_spmv_device": (c_int, [_VP, _VP, _VP, _VP, _VP]),
    "bh_linear_tetmesh_q_f32_device": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, _VP]),
    "bh_linear_tetmesh_dq_f32_device": (c_int, [_VP, c_int, _VP, _VP, _VP, _VP, _VP]),
    "bh_linear_tetmesh_dq2_f32_device": (c_int, [_VP, c_int, _VP, _VP, _VP, P(BhPsd), _VP, _VP]),
    "bh_precond_create": (c_int, [_VP, P(c_void_p)]),
    "bh_precond_destroy": (c_int, [_VP]),
    "bh_precond_info": (c_int, [_VP, P(c_int), P(c_int64)]),
    "bh_precond_factor_bic0_device": (c_int, [_VP, _VP, _VP, c_double, P(c_int), _VP]),
    "bh_precond_set_factor_device": (c_int, [_VP, _VP, _VP, _VP]),
    "bh_precond_lists_device": (c_int, [_VP, P(c_void_p), P(c_void_p), P(c_void_p), P(c_void_p), P(c_void_p)]),
    "bh_precond_apply_device": (c_int, [_VP, _VP, _VP, _VP]),
    "bh_mgn_create": (c_int, [ctypes.c_uint64, P(_VP)]),
    "bh_mgn_destroy": (c_int, [_VP]),
    "bh_mgn_set_form": (c_int, [_VP, c_int, c_int, c_int]),
    "bh_mgn_set_inference": (c_int, [_VP, c_int]),
    "bh_mgn_num_params": (c_int, [_VP, P(c_int64)]),
    "bh_mgn_get_params": (c_int, [_VP, _VP, _VP]),
    "bh_mgn_set_params": (c_int, [_VP, _VP]),
    "bh_mgn_loss_device": (c_int, [_VP, _VP, _VP, _VP, _VP, _VP, c_int, P(c_double), _VP]),
    "bh_mgn_adam_step": (c_int, [_VP, c_double, c_double, c_double, c_double, _VP]),
    "bh_mgn_train_step_device": (c_int, [_VP, _VP, _VP, _VP, _VP, c_int, ctypes.c_uint64, c_double, P(c_double), _VP]),
    "bh_mgn_factor_device": (c_int, [_VP, _VP, _VP, _VP, _VP]),
    "bh_mgn_debug_buffers": (c_int, [_VP, P(_VP), P(_VP), P(_VP), P(_VP)]),
    "bh_mgn_debug_latents": (c_int, [_VP, c_int, P(_VP), P(_VP), P(_VP)]),
    "bh_read_tetgen": (c_int, [ctypes.c_char_p, ctypes.c_char_p, P(c_int), P(c_int), _VP, _VP]),
    "bh_read_medit": (c_int, [ctypes.c_char_p, P(c_int), P(c_int), _VP, _VP]),
    "bh_save_market": (c_int, [ctypes.c_char_p, c_int64, c_int64, _VP, _VP, _VP]),
    "bh_load_market": (c_int, [ctypes.c_char_p, P(c_int64), P(c_int64), P(c_int64), _VP, _VP, _VP]),
    "bh_mesh_set_profiling": (c_int, [_VP, c_int]),
    "bh_mesh_psd_count": (c_int, [_VP, P(c_int64)]),
    "bh_mesh_kernel_times": (c_int, [_VP, _VP, _VP]),
}

KERNEL_CLASSES = ["energy", "gradient", "hessian_gather", "psd"]

_lib = None


def load():
    """Loads the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m bartels_amd.build` (there is no CPU fallback)")
    # torch ships its own HIP runtime; importing it first lets the library bind to the runtime torch
    # uses (loaded the other way round, torch finds no GPU in the same process)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, where):
    if status != BH_OK:
        raise BartelsError(status, where)

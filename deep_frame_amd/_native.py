"""ctypes binding of libdfmi.so (include/dfmi.h).

Mirrors the reference's native-FFI convention (``CudaDirectSolver``,
deep_frame/topology_optimization.py:57-78, 142-161): the library is loaded with
``ctypes.CDLL``, every symbol gets ``argtypes``/``restype``, and a nonzero status
becomes ``RuntimeError(f"{name} failed: {message}")``.  There is no fallback: a
missing or unloadable library raises at import of the first user.
"""
import ctypes
from ctypes import POINTER, Structure, c_char_p, c_double, c_int32, c_int64, c_uint8, c_void_p
from pathlib import Path

LIBRARY = Path(__file__).resolve().parent / "libdfmi.so"


class Settings(Structure):
    _fields_ = [("coarsest_dofs", c_int32), ("max_levels", c_int32), ("levels", c_int32),
                ("chebyshev_degree", c_int32), ("chebyshev_ratio", c_double), ("growth", c_double),
                ("power_iterations", c_int32), ("precision", c_int32), ("tolerance", c_double),
                ("max_iterations", c_int32), ("dead_ratio", c_double), ("seed", c_int32), ("refine", c_int32),
                ("inner_tolerance", c_double), ("fine_smoother", c_int32), ("retries", c_int32),
                ("coarse_krylov", c_int32), ("block_krylov", c_int32), ("smoother", c_int32), ("sweeps", c_int32),
                ("damping", c_double), ("cycle", c_int32), ("coarse", c_int32), ("krylov_levels", c_int32),
                ("chebyshev_kind", c_int32)]


class Report(Structure):
    _fields_ = [("iterations", c_int32), ("relative_residual", c_double), ("converged", c_int32),
                ("retried", c_int32)]


SIGNATURES = {
    "dfmi_last_error": ([], c_char_p),
    "dfmi_default_settings": ([POINTER(Settings)], None),
    "dfmi_create": ([POINTER(c_int32), POINTER(c_double), POINTER(c_double), POINTER(Settings), POINTER(c_void_p)],
                    c_int32),
    "dfmi_destroy": ([c_void_p], c_int32),
    "dfmi_set_stream": ([c_void_p, c_void_p], c_int32),
    "dfmi_levels": ([c_void_p, POINTER(c_int32), POINTER(c_int32)], c_int32),
    "dfmi_update": ([c_void_p, c_void_p], c_int32),
    "dfmi_build_hierarchy": ([c_void_p, c_int32, c_void_p, POINTER(c_double)], c_int32),
    "dfmi_solve": ([c_void_p, c_int32, c_void_p, c_void_p, c_int32, POINTER(Report)], c_int32),
    "dfmi_apply": ([c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_int32], c_int32),
    "dfmi_set_mass": ([c_void_p, POINTER(c_double)], c_int32),
    "dfmi_mass_apply": ([c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32], c_int32),
    "dfmi_mask": ([c_void_p, c_int32, c_void_p, c_int32], c_int32),
    "dfmi_build_hierarchy_kernel": ([c_void_p, c_int32, c_void_p, POINTER(c_double), c_int32, POINTER(c_double)],
                                    c_int32),
    "dfmi_kernel_component": ([c_void_p, c_int32, c_void_p, c_int32, POINTER(c_int32), POINTER(c_double)], c_int32),
    "dfmi_modal_sensitivity": ([c_void_p, c_void_p, c_int32, POINTER(c_double), POINTER(c_double), c_void_p, c_void_p,
                                c_double, c_void_p], c_int32),
    "dfmi_element_energy": ([c_void_p, c_void_p, c_int32, c_void_p], c_int32),
    "dfmi_compliance_sensitivity": ([c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_double, c_void_p, c_void_p],
                                    c_int32),
    "dfmi_level_apply": ([c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_int32], c_int32),
    "dfmi_precondition": ([c_void_p, c_int32, c_void_p, c_void_p, c_int32], c_int32),
    "dfmi_transfer": ([c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_int32], c_int32),
    "dfmi_time_apply": ([c_void_p, c_int32, c_int32, c_int32, c_int32, c_int32, POINTER(c_double)], c_int32),
    "dfmi_tune_fine": ([c_void_p, c_int32, c_int32, POINTER(c_int32), POINTER(c_double)], c_int32),
    "dfmi_diagnostics": ([c_void_p, c_int32, POINTER(c_double), POINTER(c_int32)], c_int32),
    "dfmi_history": ([c_void_p, POINTER(c_double), c_int32, POINTER(c_int32)], c_int32),
    "dfmi_kernel_check": ([c_void_p, c_int32, POINTER(c_double), POINTER(c_int32)], c_int32),
    "dfmi_gram": ([c_void_p, c_void_p, c_int64, c_int32, c_int32, c_void_p, c_void_p, c_int32, c_void_p], c_int32),
    "dfmi_block_work_doubles": ([c_int64, c_int32, c_int32, POINTER(c_int64)], c_int32),
    "dfmi_block_gram": ([c_void_p, c_int32, c_void_p, c_int32, c_void_p, c_int32, c_int64, c_void_p, c_void_p,
                         c_void_p], c_int32),
    "dfmi_block_combine": ([c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int64, c_void_p], c_int32),
    "dfmi_block_update": ([c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int64, c_void_p, c_void_p,
                           c_void_p], c_int32),
    "dfmi_block_combine2": ([c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32, c_int64,
                             c_void_p], c_int32),
    "dfmi_block_combine_flush": ([c_void_p, c_void_p, c_void_p, c_int32, c_int64, c_void_p, c_void_p, c_void_p,
                                  c_void_p], c_int32),
    "dfmi_gram_static": ([c_void_p, c_void_p, c_int64, c_int32, c_int32, c_void_p, c_void_p], c_int32),
    "dfmi_combine": ([c_void_p, c_int64, c_int32, c_void_p, c_int32, c_void_p, c_int32, c_void_p], c_int32),
    "dfmi_project": ([c_void_p, c_void_p, c_int64, c_double, POINTER(c_double), c_int32, c_int32, c_void_p, c_void_p,
                      c_void_p], c_int32),
    "dfmi_simp": ([c_void_p, c_void_p, c_int64, c_double, c_double, c_double, c_void_p, c_void_p, c_void_p], c_int32),
    "dfmi_filter_create": ([POINTER(c_int32), POINTER(c_double), c_double, POINTER(c_uint8), POINTER(c_uint8),
                            POINTER(c_void_p)], c_int32),
    "dfmi_filter_destroy": ([c_void_p], c_int32),
    "dfmi_filter_set_stream": ([c_void_p, c_void_p], c_int32),
    "dfmi_filter_forward": ([c_void_p, c_void_p, c_void_p], c_int32),
    "dfmi_filter_pullback": ([c_void_p, c_void_p, c_void_p, c_void_p], c_int32),
    "dfmi_filter_set_corridor": ([c_void_p, c_double, POINTER(c_double)], c_int32),
    "dfmi_filter_volume": ([c_void_p, c_void_p, POINTER(c_double)], c_int32),
    "dfmi_oc_update": ([c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_double, c_void_p], c_int32),
    "dfmi_mma_step": ([c_int64, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                       c_void_p, POINTER(c_double), c_void_p, c_double, c_double, c_void_p], c_int32),
    "dfmi_net_create": ([POINTER(c_void_p)], c_int32),
    "dfmi_net_destroy": ([c_void_p], c_int32),
    "dfmi_net_set_stream": ([c_void_p, c_void_p], c_int32),
    "dfmi_net_param_count": ([c_void_p, POINTER(c_int64)], c_int32),
    "dfmi_net_set_params": ([c_void_p, c_void_p], c_int32),
    "dfmi_net_get": ([c_void_p, c_int32, c_void_p], c_int32),
    "dfmi_net_forward": ([c_void_p, c_void_p, c_int32, POINTER(c_int32), c_void_p], c_int32),
    "dfmi_net_backward": ([c_void_p, c_void_p], c_int32),
    "dfmi_net_adam": ([c_void_p, c_double, c_double, c_double, c_double], c_int32),
    "dfmi_net_train_step": ([c_void_p, c_void_p, c_int32, c_void_p, c_int32, c_double, POINTER(c_double)], c_int32),
    "dfmi_attach_net": ([c_void_p, c_void_p], c_int32),
    "dfmi_harvest": ([c_void_p, c_void_p, c_int64], c_int32),
    "dfmi_harvest_count": ([c_void_p, POINTER(c_int64)], c_int32),
}

_lib = None


def library():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise RuntimeError(f"{LIBRARY} is missing: run deep_frame_amd/build.py (or __graft_entry__.build())")
        lib = ctypes.CDLL(str(LIBRARY))
        missing = [name for name in SIGNATURES if not hasattr(lib, name)]
        if missing:
            raise RuntimeError(f"{LIBRARY} lacks symbols {missing}: rebuild it")
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def call(name, *args):
    status = getattr(library(), name)(*args)
    if status != 0:
        message = library().dfmi_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed: {message}")


def default_settings():
    s = Settings()
    library().dfmi_default_settings(ctypes.byref(s))
    return s


def ptr(tensor):
    """Device pointer of a contiguous torch tensor that lives on the GPU (host memory is refused)."""
    if tensor.device.type != "cuda":
        raise ValueError(f"expected a GPU tensor, got one on {tensor.device}")
    if not tensor.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return c_void_p(tensor.data_ptr())


def stream_handle(torch, device=None):
    return c_void_p(torch.cuda.current_stream(device).cuda_stream)

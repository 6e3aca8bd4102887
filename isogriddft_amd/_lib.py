"""ctypes binding of ``libisogrid_hip.so`` (the C ABI declared in include/isogrid_hip.h).

The library is loaded from this package directory.  There is no fallback: if the
library is missing every entry point raises, so no CPU path can silently stand in
for the HIP kernels.
"""

from __future__ import annotations

import ctypes as C
import functools
import inspect
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libisogrid_hip.so"

P = C.c_void_p
I = C.c_int
L = C.c_longlong
D = C.c_double

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "ig_monitor_stencil_f64": [P, P, I, L, I, I, I, P, P, P, D, I, I, I, I, I, P],
    "ig_monitor_stencil_f32": [P, P, I, L, I, I, I, P, P, P, D, I, I, I, I, I, P],
    "ig_monitor_stencil_dots_work_size": [I, I, I, I],
    "ig_monitor_stencil_dots_f64": [P, P, I, L, I, I, I, P, P, P, D, P, P, L, P, P],
    "ig_monitor_stencil_dots_f32": [P, P, I, L, I, I, I, P, P, P, D, P, P, L, P, P],
    "ig_legacy_stencil_f64": [P, P, I, L, I, I, I, P, P, D, I, I, P],
    "ig_legacy_stencil_f32": [P, P, I, L, I, I, I, P, P, D, I, I, P],
    "ig_reduction_work_size": [L],
    "ig_dots_f64": [I, P, P, P, P, P, P, P, L, P, P, I, I, I, P],
    "ig_dots_f32": [I, P, P, P, P, P, P, P, L, P, P, I, I, I, P],
    "ig_dots_noalloc_f64": [I, P, P, P, P, P, P, P, L, P, I, I, I, P],
    "ig_bicgstab_p_f64": [L, P, P, P, P, P, P, P],
    "ig_bicgstab_p_f32": [L, P, P, P, P, P, P, P],
    "ig_bicgstab_s_f64": [L, P, P, P, P, P, P, P],
    "ig_bicgstab_s_f32": [L, P, P, P, P, P, P, P],
    "ig_bicgstab_x_f64": [L, P, P, P, P, P, P, P, P, P, P],
    "ig_bicgstab_x_f32": [L, P, P, P, P, P, P, P, P, P, P],
    "ig_mg_fine_restrict_f64": [P, P, I, I, I, P, I, I, I, P],
    "ig_mg_fine_restrict_f32": [P, P, I, I, I, P, I, I, I, P],
    "ig_mg_fine_prolong_f64": [P, P, I, I, I, I, I, I, P],
    "ig_mg_fine_prolong_f32": [P, P, I, I, I, I, I, I, P],
    "ig_mg_smooth_f64": [P, P, P, P, L, P],
    "ig_mg_smooth_f32": [P, P, P, P, L, P],
    "ig_mg7_jacobi_f64": [P, P, P, P, D, I, I, I, P],
    "ig_mg7_jacobi_f32": [P, P, P, P, D, I, I, I, P],
    "ig_mg7_restrict_f64": [P, P, P, P, I, I, I, I, I, I, P],
    "ig_mg7_restrict_f32": [P, P, P, P, I, I, I, I, I, I, P],
    "ig_mg7_prolong_f64": [P, P, I, I, I, I, I, I, P],
    "ig_mg7_prolong_f32": [P, P, I, I, I, I, I, I, P],
    "ig_mg_probe_vector_f64": [P, I, I, I, I, P],
    "ig_mg_probe_read_f64": [P, P, I, I, I, I, P],
    "ig_gram_work_size": [I, I, L],
    "ig_gram_f32": [P, L, I, P, L, I, P, L, P, I, P, L, P],
    "ig_gram_f64": [P, L, I, P, L, I, P, L, P, I, P, L, P],
    "ig_rotate_f32": [P, I, P, L, P, L, I, I, L, I, P],
    "ig_rotate_f64": [P, I, P, L, P, L, I, I, L, I, P],
    "ig_gemm_nt_work_size": [I, I, L],
    "ig_gemm_nt_f32": [P, L, I, P, L, I, L, P, L, P, L, P],
    "ig_gemm_nt_f64": [P, L, I, P, L, I, L, P, L, P, L, P],
    "ig_physics_work_size": [L],
    "ig_vloc_gth_f64": [P, L, I, P, P, I, P],
    "ig_vloc_gth_grad_f64": [P, P, L, I, P, P, P, P],
    "ig_lsda_pw92_f64": [P, P, L, P, P, P, P, P],
    "ig_lsda_vwn_f64": [P, P, L, P, P, P, P, P],
    "ig_projector_desc_size": [],
    "ig_projector_fields_f64": [P, I, I, I, P, P, I, P, P, I, P],
    "ig_nonlocal_work_size": [I, I],
    "ig_nonlocal_project_f64": [P, L, I, I, I, I, P, P, I, P, P, P, P, P, P],
    "ig_nonlocal_project_f32": [P, L, I, I, I, I, P, P, I, P, P, P, P, P, P],
    "ig_nonlocal_apply_f64": [P, L, I, I, I, I, P, P, I, P, P],
    "ig_nonlocal_apply_f32": [P, L, I, I, I, I, P, P, I, P, P],
    "ig_density_f64": [P, L, I, L, P, P, I, P],
    "ig_density_f32": [P, L, I, L, P, P, I, P],
    "ig_multipole_moments_f64": [P, P, P, P, L, P, I, P, P, P],
    "ig_boundary_values_f64": [P, I, I, I, P, I, P, P, P],
    "ig_cell_fit_f64": [P, P, I, I, I, P, I, P, P],
    "ig_cell_moments_f64": [P, P, I, I, I, P, I, P, D, P, I, P, P, P],
    "ig_cell_split_moments_f64": [P, P, P, I, I, I, P, I, P, D, P, I, P, P, P],
    "ig_moments_adjoint_f64": [P, P, P, L, P, I, P, P, I, P],
    "ig_cell_moments_adjoint_f64": [P, P, I, I, I, P, I, D, P, I, P, P, P],
    "ig_knn_work_bytes": [I, I],
    "ig_knn_f64": [P, L, P, I, I, P, L, P, P, P],
    "ig_idw_quadrature_f64": [P, P, P, I, I, P, P, P, P, P],
    "ig_patch_main_f64": [P, P, P, L, P, D, P, P, P],
    "ig_patch_field_f64": [P, L, P, D, I, D, P, P],
    "ig_grid_work_size": [],
    "ig_monitor_field_f64": [P, L, P, I, P, P],
    "ig_monitor_smooth_f64": [P, I, I, I, D, P, P],
    "ig_harmonic_sweep_f64": [P, P, P, I, I, I, D, D, D, D, P, P, P],
    "ig_mapping_min_jacobian_f64": [P, I, I, I, D, D, D, P, P, P],
    "ig_mapping_metric_f64": [P, I, I, I, D, D, D, P, P, P, P],
    "ig_row_dots_work_size": [I],
    "ig_row_dots_f64": [P, P, P, I, L, L, P, P, P],
    "ig_row_dots_f32": [P, P, P, I, L, L, P, P, P],
    "ig_cg_update_f64": [P, P, P, P, P, P, P, I, L, L, P, P, P],
    "ig_cg_update_f32": [P, P, P, P, P, P, P, I, L, L, P, P, P],
    "ig_cg_direction_f64": [P, P, P, P, I, L, L, P],
    "ig_cg_direction_f32": [P, P, P, P, I, L, L, P],
    "ig_cg_start_f64": [P, L, P, P, P, P, P, P, I, L, L, P, P, P],
    "ig_cg_start_f32": [P, L, P, P, P, P, P, P, I, L, L, P, P, P],
    "ig_cg_finish_f64": [P, P, P, P, P, L, I, L, L, P],
    "ig_cg_finish_f32": [P, P, P, P, P, L, I, L, L, P],
    "ig_green_identity_work_size": [],
    "ig_green_identity_f64": [P, P, P, P, P, I, I, I, D, D, D, I, P, P, P],
    "ig_stream_blocks": [L],
    "ig_stream_copy_f32": [P, P, L, P],
    "ig_stream_read_f32": [P, L, P, P],
}
_RESTYPE = {"ig_gram_work_size": L, "ig_knn_work_bytes": L, "ig_gemm_nt_work_size": L,
            "ig_monitor_stencil_dots_work_size": L}

_lib = None


class HipLibraryError(RuntimeError):
    pass


def load():
    """Load the library (once) and bind every declared symbol."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise HipLibraryError(
            f"{LIB_PATH} is missing: build it with `python -m isogriddft_amd.build` "
            "(there is no CPU fallback for the HIP kernels)"
        )
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, I)
    _lib = lib
    return lib


def call(name, *args):
    """Invoke an entry point and translate its status code."""
    fn = getattr(load(), name)
    st = fn(*args)
    if st == 0:
        return
    if st in (-1, -2):
        raise ValueError(f"{name}: invalid or unsupported arguments (status {st})")
    raise RuntimeError(f"{name}: HIP launch failed (status {st})")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("HIP kernels take device tensors; got a host tensor")
    return C.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


def on_stream(fn):
    """Decorator for functions with a ``stream`` argument (a torch CUDA stream or None):
    with a stream, the body runs with that stream current, so its tensor operations and
    the library kernels it launches (on the current stream when ``stream`` is None) are
    ordered on that one stream.  Without one, everything runs on the current stream.  The
    caller orders later uses of the results, as with any stream argument."""
    idx = list(inspect.signature(fn).parameters).index("stream")

    @functools.wraps(fn)
    def wrapped(*args, **kw):
        if len(args) > idx:
            stream = args[idx]
            if stream is None:
                return fn(*args, **kw)
            args = args[:idx] + (None,) + args[idx + 1:]
        else:
            stream = kw.get("stream")
            if stream is None:
                return fn(*args, **kw)
            kw["stream"] = None
        with torch.cuda.stream(stream):
            return fn(*args, **kw)

    return wrapped

"""ctypes binding of libdmsc.so (the C ABI declared in include/dmsc.h).

The library is built in-tree by damascene_amd/build.py.  Loading fails
loudly when it is missing: there is no CPU fallback for any stage.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdmsc.so")

c_int, c_float, c_void_p, c_uint64 = ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_uint64
P = c_void_p


class GpbParams(ctypes.Structure):
    _fields_ = [("n_eig", c_int), ("tol", c_float), ("texton_choice", c_int), ("precision", c_int)]


class GpbOutputs(ctypes.Structure):
    _fields_ = [
        ("d_textons", P), ("d_mpb", P), ("d_combined_g", P), ("d_nmax", P), ("d_eigvecs", P),
        ("d_spb", P), ("d_gpb_all", P), ("d_gpb", P), ("pitch", c_int),
        ("h_eigvals", c_float * 32), ("kmeans_iterations", c_int), ("lanczos_iterations", c_int),
        ("stage_ms", c_float * 8),
    ]


class StepShape(ctypes.Structure):
    _fields_ = [("mixed_fused", c_int), ("tile", c_int), ("k1_threads", c_int), ("ms_per_step", c_float)]


class BatchImage(ctypes.Structure):
    _fields_ = [("w", c_int), ("h", c_int), ("d_rgb", P), ("out", ctypes.POINTER(GpbOutputs))]


# name -> (restype, argtypes)
SIGNATURES = {
    "dmsc_version": (ctypes.c_char_p, []),
    "dmsc_last_error_string": (ctypes.c_char_p, []),
    "dmsc_ctx_create": (c_int, [ctypes.POINTER(P), c_int, P]),
    "dmsc_ctx_destroy": (c_int, [P]),
    "dmsc_ctx_synchronize": (c_int, [P]),
    "dmsc_ctx_stream": (P, [P]),
    "dmsc_pitch_in_floats": (c_int, [c_int]),
    "dmsc_generate_image": (c_int, [c_int, c_int, c_int, c_uint64, P, P]),
    "dmsc_boundary_mask": (c_int, [c_int, c_int, P, P]),
    "dmsc_texton_filters": (c_int, [P, P]),
    "dmsc_parabola_filters": (c_int, [c_int, P]),
    "dmsc_spb_filters": (c_int, [P]),
    "dmsc_rotation_tables": (c_int, [c_int, c_int, c_float, P, P, P, P, P, P]),
    "dmsc_halfdisc_offsets": (c_int, [c_int, c_int, c_float, c_int, P, P, P, P, c_int, P]),
    "dmsc_blur_kernel": (c_int, [c_float, P, P]),
    "dmsc_stencil": (c_int, [c_int, c_int, P, P]),
    "dmsc_gamma_lut": (c_int, [c_float, P]),
    "dmsc_load_ppm": (c_int, [ctypes.c_char_p, ctypes.POINTER(c_int), ctypes.POINTER(c_int), P]),
    "dmsc_postprocess": (c_int, [P, c_int, c_int, P, P, P, ctypes.POINTER(c_int)]),
    "dmsc_normalize_gpb_all": (c_int, [P, c_int, c_int, c_int, P]),
    "dmsc_write_pb": (c_int, [ctypes.c_char_p, c_int, c_int, P]),
    "dmsc_write_gpb_all": (c_int, [ctypes.c_char_p, c_int, c_int, c_int, P]),
    "dmsc_write_pgm": (c_int, [ctypes.c_char_p, c_int, c_int, P]),
    "dmsc_lanczos_time_steps": (c_int, [P, c_int, c_int, P, c_int, c_int, c_int, P]),
    "dmsc_chebyshev_time_products": (c_int, [P, c_int, c_int, P, c_int, c_int, P]),
    "dmsc_autotune_chebyshev": (c_int, [P, c_int, c_int, P, c_int, c_int, P, P, P, c_int, P]),
    "dmsc_chebyshev_tile": (c_int, [P, c_int, c_int, P]),
    "dmsc_tune_set": (c_int, [ctypes.c_char_p, c_int]),
    "dmsc_tune_get": (c_int, [ctypes.c_char_p, ctypes.POINTER(c_int)]),
    "dmsc_rgb_to_grey": (c_int, [P, c_int, c_int, P, P]),
    "dmsc_rgb_to_lab": (c_int, [P, c_int, c_int, c_float, P, P, P, P]),
    "dmsc_normalize_lab": (c_int, [P, c_int, c_int, P, P, P]),
    "dmsc_mirror_image_f": (c_int, [P, c_int, c_int, c_int, P, P]),
    "dmsc_mirror_image_i": (c_int, [P, c_int, c_int, c_int, P, P]),
    "dmsc_unmirror_image_f": (c_int, [P, c_int, c_int, c_int, P, P]),
    "dmsc_quantize_image": (c_int, [P, c_int, c_int, c_int, P, P]),
    "dmsc_texton_responses": (c_int, [P, c_int, c_int, P, P]),
    "dmsc_kmeans": (c_int, [P, c_int, c_int, c_int, c_int, P, P, c_int, c_int, c_int, P, P]),
    "dmsc_find_textons": (c_int, [P, c_int, c_int, P, P, c_int, c_int, P]),
    "dmsc_local_cues": (c_int, [P, c_int, c_int, P, P, P, P, P, P, P, P, c_int, c_int]),
    "dmsc_local_cues_raw": (c_int, [P, c_int, c_int, P, P, P, P, P, P, P, P, c_int, P, c_int]),
    "dmsc_combine": (c_int, [P, c_int, c_int, c_int, P, P, P, P, P, P, c_int]),
    "dmsc_nonmax": (c_int, [P, c_int, c_int, P, c_int, P]),
    "dmsc_calc_gpb": (c_int, [P, c_int, c_int, c_int, P, P, P, P, P]),
    "dmsc_intervene": (c_int, [P, c_int, c_int, P, P, c_int, c_float]),
    "dmsc_convert_matrix": (c_int, [P, c_int, c_int, P, c_int, P, P]),
    "dmsc_stencil_spmv": (c_int, [P, c_int, c_int, P, c_int, P, P]),
    "dmsc_generalized_eigensolve": (c_int, [P, c_int, c_int, P, c_int, c_int, c_float, c_int, P, P, P]),
    "dmsc_spectral_pb": (c_int, [P, P, P, c_int, c_int, c_int, P, c_int]),
    "dmsc_gpb_default_params": (None, [P]),
    "dmsc_gpb": (c_int, [P, c_int, c_int, P, P, P, c_int]),
    "dmsc_gpb_batch": (c_int, [P, c_int, P, P, c_int]),
    "dmsc_kmeans_rechecks": (c_int, [P, ctypes.POINTER(c_int)]),
    "dmsc_kmeans_noalloc": (c_int, [P, c_int, c_int, c_int, c_int, P, P, c_int, c_int, c_int, P, P]),
    "dmsc_kmeans_noalloc_max_pixels": (ctypes.c_size_t, []),
    "dmsc_eigensolve_report": (c_int, [P, ctypes.POINTER(c_int), ctypes.POINTER(ctypes.c_double)]),
    "dmsc_autotune_eigensolver": (c_int, [P, c_int, c_int, P, c_int, c_int, c_int, P, P, c_int, P]),
    "dmsc_autotune_spmv": (c_int, [P, c_int, c_int, P, c_int, P, P, c_int, P, P]),
    "dmsc_autotune_clear": (c_int, []),
    "dmsc_combine_weights": (c_int, [c_int, P, P]),
    "dmsc_symmetric_eigen": (c_int, [c_int, P, P, P]),
    "dmsc_spb_combine_forward": (c_int, [P, c_int, c_int, c_int, P, P, P, P, c_int, P, P, P, c_int, P, P]),
    "dmsc_spb_combine_backward": (c_int, [P, c_int, c_int, c_int, P, P, P, P, c_int, P, P, P, P, P, P]),
}

_lib = None


class DmscError(RuntimeError):
    pass


def lib():
    """Load libdmsc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DmscError(f"{LIB_PATH} is missing; build it with `python -m damascene_amd.build`")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().dmsc_last_error_string()
        raise DmscError(f"dmsc error {rc}: {msg.decode() if msg else ''}")
    return rc

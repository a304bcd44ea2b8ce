"""ctypes binding of libatsim.so (include/atsim.h).

Device buffers are passed as raw device pointers; the helpers below accept
torch CUDA tensors (torch is used only for device memory and streams).  There
is no fallback: if libatsim.so is missing or no GPU is present, the calls
raise.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libatsim.so")

_lib = None


class AtsimError(RuntimeError):
    pass


class SimDesc(C.Structure):
    _fields_ = [
        ("N", C.c_long), ("Nout", C.c_long), ("M", C.c_long), ("L", C.c_long), ("W", C.c_long),
        ("D", C.c_long), ("dx", C.c_double), ("dt", C.c_double), ("zangle", C.c_double),
        ("bicubic", C.c_int), ("fresnel", C.c_int), ("amplitude", C.c_int),
        ("slayer_alt", C.POINTER(C.c_double)), ("super_layer_index", C.POINTER(C.c_long)),
        ("lambda_", C.POINTER(C.c_double)), ("scoeff", C.POINTER(C.c_double)),
        ("unwrap_mode", C.POINTER(C.c_int)),
        ("xpos0", C.POINTER(C.c_double)), ("ypos0", C.POINTER(C.c_double)),
        ("vx", C.POINTER(C.c_double)), ("vy", C.POINTER(C.c_double)),
    ]


# name -> (restype, argtypes)
_SIGS = {
    "atsim_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "atsim_ctx_destroy": (None, [C.c_void_p]),
    "atsim_error": (C.c_char_p, [C.c_void_p]),
    "atsim_sync": (C.c_int, [C.c_void_p]),
    "atsim_stream": (C.c_void_p, [C.c_void_p]),
    "atsim_device_ok": (C.c_int, [C.c_void_p]),
    "atsim_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "atsim_malloc": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_size_t]),
    "atsim_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "atsim_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int]),
    "atsim_get_tuning": (C.c_int, [C.c_void_p, C.c_char_p]),
    "atsim_fft2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_int]),
    "atsim_permut": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_long]),
    "atsim_fresnel": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_long,
                                C.c_double, C.c_double, C.c_double]),
    "atsim_fft2d_variant": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_int, C.c_int]),
    "atsim_row_fft_variant": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_int, C.c_int]),
    "atsim_fft2d_nm": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_long, C.c_int]),
    "atsim_fresnel_nm": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_long,
                                   C.c_double, C.c_double, C.c_double]),
    "atsim_fresnel_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_long,
                                        C.c_double, C.c_double, C.c_double]),
    "atsim_fresnel_tf_init": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_double, C.c_double,
                                        C.c_double, C.c_double]),
    "atsim_fresnel_tf_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_long]),
    "atsim_fresnel_cube": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_double,
                                     C.c_double, C.c_double, C.c_long, C.c_double]),
    "atsim_noise_c64": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_uint64, C.c_uint32]),
    "atsim_make_master": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_double, C.c_double, C.c_double,
                                    C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "atsim_strfn": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_long, C.POINTER(C.c_double)]),
    "atsim_r0_measure": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.POINTER(C.c_double)]),
    "atsim_scale": (C.c_int, [C.c_void_p, C.c_void_p, C.c_long, C.c_double]),
    "atsim_superlayers": (C.c_long, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_long, C.c_double,
                                     C.POINTER(C.c_double), C.POINTER(C.c_long)]),
    "atsim_sim_create": (C.c_int, [C.c_void_p, C.POINTER(SimDesc), C.POINTER(C.c_void_p)]),
    "atsim_sim_destroy": (None, [C.c_void_p]),
    "atsim_sim_master": (C.c_void_p, [C.c_void_p, C.c_long]),
    "atsim_sim_set_master": (C.c_int, [C.c_void_p, C.c_long, C.c_void_p]),
    "atsim_sim_frame": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "atsim_sim_positions": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "atsim_sim_set_positions": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "atsim_sim_gaps": (C.c_long, [C.c_void_p]),
    "atsim_sim_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "atsim_sim_synth_masters": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "atsim_sim_normalize_masters": (C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_double), C.c_double,
                                              C.POINTER(C.c_double)]),
    "atsim_sim_profile_read": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_long)]),
    "atsim_sim_layer_screens": (C.c_int, [C.c_void_p, C.c_void_p]),
    "atsim_sims_frame": (C.c_int, [C.POINTER(C.c_void_p), C.c_long, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "atsim_sim_use_masters": (C.c_int, [C.c_void_p, C.c_void_p]),
    "atsim_sim_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_char_p, C.c_long]),
    "atsim_sim_stats_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "atsim_sim_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_long]),
    "atsim_sim_block_sums": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_long]),
    "atsim_block_fixup": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_long, C.POINTER(C.c_double)]),
    "atsim_block_counts": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_long, C.POINTER(C.c_float)]),
    "atsim_host_alloc": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_size_t]),
    "atsim_host_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "atsim_copy_async": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_size_t]),
    "atsim_copy_wait": (C.c_int, [C.c_void_p, C.c_int]),
    "atsim_stream_wait_copy": (C.c_int, [C.c_void_p, C.c_int]),
    "atsim_sim_advance": (C.c_int, [C.c_void_p]),
    "atsim_gemm_bf16": (C.c_int, [C.c_void_p, C.c_long, C.c_long, C.c_long, C.c_long, C.c_float,
                                  C.c_void_p, C.c_long, C.c_long, C.c_long,
                                  C.c_void_p, C.c_long, C.c_long, C.c_long, C.c_float,
                                  C.c_void_p, C.c_long, C.c_long, C.c_long]),
    "atsim_gemm_bf16_ex": (C.c_int, [C.c_void_p, C.c_long, C.c_long, C.c_long, C.c_long, C.c_float,
                                     C.c_void_p, C.c_long, C.c_long, C.c_long,
                                     C.c_void_p, C.c_long, C.c_long, C.c_long, C.c_float,
                                     C.c_void_p, C.c_long, C.c_long, C.c_long, C.c_void_p, C.c_void_p, C.c_int]),
    "atsim_nn_add_bias_act": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_long, C.c_long, C.c_long, C.c_int]),
    "atsim_nn_gelu_bwd": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long]),
    "atsim_nn_bias_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_long, C.c_int]),
    "atsim_nn_mse": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.POINTER(C.c_double)]),
    "atsim_nn_reduce_batches": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_int]),
    "atsim_nn_make_pairs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long,
                                      C.c_long, C.c_long, C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_double,
                                      C.c_double, C.c_double, C.c_double]),
    "atsim_nn_unpack": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_long, C.c_long, C.c_double,
                                  C.c_double, C.c_double, C.c_void_p, C.c_void_p]),
    "atsim_nn_adam": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_float,
                                C.c_float, C.c_float, C.c_float, C.c_long]),
}


def lib():
    """Load libatsim.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AtsimError(f"{LIB_PATH} not built: run `python -m atmosphericturbulencesimul_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)

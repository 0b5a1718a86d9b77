"""ctypes binding of libflf_hip.so (C ABI declared in include/flf_hip.h).

The library is the only compute path: if it is missing or cannot be loaded,
importing the projector fails loudly (there is no CPU fallback).
"""
import ctypes
import os

LIB_NAME = "libflf_hip.so"
LIB_PATH = os.environ.get("FLF_HIP_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

FLF_OK = 0
FLF_EINVAL = -1
FLF_ENOMEM = -2
FLF_EHIP = -3
FLF_ESTATE = -4

# Every symbol include/flf_hip.h declares, with (restype, argtypes).
_c = ctypes
SIGNATURES = {
    "flf_version": (_c.c_int, []),
    "flf_create": (_c.c_int, [_c.c_int, _c.POINTER(_c.c_void_p)]),
    "flf_destroy": (None, [_c.c_void_p]),
    "flf_last_error": (_c.c_char_p, [_c.c_void_p]),
    "flf_plan": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int,
                            _c.POINTER(_c.c_int)]),
    "flf_plane_fshape": (_c.c_int, [_c.c_void_p, _c.c_int, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int)]),
    "flf_device_bytes": (_c.c_longlong, [_c.c_void_p]),
    "flf_set_psf": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_void_p]),
    "flf_forward": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.POINTER(_c.c_int), _c.c_int,
                               _c.c_void_p]),
    "flf_backward": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.POINTER(_c.c_int), _c.c_int,
                                _c.c_void_p]),
    "flf_backward_rl_update": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_void_p,
                                          _c.c_int, _c.POINTER(_c.c_int), _c.c_int, _c.c_void_p]),
    "flf_forward_ratio": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_void_p,
                                     _c.POINTER(_c.c_int), _c.c_int, _c.c_void_p]),
    "flf_datafit": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_int, _c.POINTER(_c.c_double),
                               _c.c_void_p]),
    "flf_fourier_len": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int]),
    "flf_grid_check": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _c.POINTER(_c.c_longlong)]),
    "flf_radix_plan": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int)]),
    "flf_set_frames": (_c.c_int, [_c.c_void_p, _c.c_int]),
    "flf_forward_pair": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_void_p, _c.POINTER(_c.c_int),
                                    _c.c_int, _c.c_void_p]),
    "flf_rl_ratio": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _c.c_longlong, _c.c_int,
                                _c.c_void_p]),
    "flf_set_frame_dims": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_int, _c.c_void_p]),
    "flf_frame_capacity": (_c.c_int, [_c.c_void_p]),
    "flf_set_option": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.c_int]),
    "flf_get_option": (_c.c_int, [_c.c_void_p, _c.c_char_p, _c.POINTER(_c.c_int)]),
    "flf_rl_scale": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _c.c_longlong, _c.c_int,
                                _c.c_void_p]),
    "flf_net_create": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.POINTER(_c.c_void_p)]),
    "flf_net_destroy": (None, [_c.c_void_p]),
    "flf_net_last_error": (_c.c_char_p, [_c.c_void_p]),
    "flf_net_num_params": (_c.c_longlong, [_c.c_void_p]),
    "flf_net_set_params": (_c.c_int, [_c.c_void_p, _c.c_void_p]),
    "flf_net_get_params": (_c.c_int, [_c.c_void_p, _c.c_void_p]),
    "flf_net_get_grads": (_c.c_int, [_c.c_void_p, _c.c_void_p]),
    "flf_net_forward": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _c.c_int, _c.c_void_p,
                                   _c.c_void_p]),
    "flf_net_train_step": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _c.c_int,
                                      _c.c_float, _c.POINTER(_c.c_double), _c.c_void_p]),
    "flf_net_upsample": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_void_p,
                                    _c.c_void_p]),
    "flf_mfma_gemm_nt": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _c.c_int,
                                    _c.c_void_p]),
    "flf_check_device_buffer": (_c.c_int, [_c.c_int, _c.c_void_p, _c.c_longlong]),
    "flf_datafit_noalloc": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _c.c_longlong, _c.c_void_p,
                                       _c.c_void_p]),
    "flf_psf_line": (_c.c_int, [_c.c_void_p, _c.c_double, _c.c_int, _c.c_void_p, _c.c_void_p]),
    "flf_psf_wave": (_c.c_int, [_c.c_void_p, _c.c_double, _c.c_int, _c.c_int, _c.c_void_p, _c.c_void_p]),
    "flf_psf_lightfield": (_c.c_int, [_c.c_void_p, _c.c_void_p, _c.c_int, _c.c_int, _c.c_int, _c.c_void_p,
                                      _c.c_void_p]),
    "flf_psf_normalise": (_c.c_int, [_c.c_void_p, _c.c_int, _c.c_int, _c.c_double, _c.c_int, _c.c_void_p,
                                     _c.c_void_p]),
    "flf_profile_enable": (_c.c_int, [_c.c_void_p, _c.c_int]),
    "flf_profile_read": (_c.c_int, [_c.c_void_p, _c.POINTER(_c.c_double), _c.POINTER(_c.c_longlong),
                                    _c.POINTER(_c.c_double)]),
}

_lib = None


def load():
    """Load the library (once) and declare every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "%s not found: build the HIP library first (python -m flfhip.build, "
            "or make -C fast-light-field_amd/csrc)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FLFError(RuntimeError):
    pass


def check(lib, ctx, rc, what):
    """Map a status code to the exception the reference would raise
    (py_light_field.cpp:1425-1493: TypeError/ValueError; MemoryError)."""
    if rc == FLF_OK:
        return
    msg = lib.flf_last_error(ctx)
    msg = msg.decode() if msg else ""
    text = "%s failed (%d): %s" % (what, rc, msg)
    if rc == FLF_EINVAL:
        raise ValueError(text)
    if rc == FLF_ENOMEM:
        raise MemoryError(text)
    if rc == FLF_ESTATE:
        raise RuntimeError(text)
    raise FLFError(text)

"""ctypes binding of the C-ABI in include/seistorch_amd.h.

The library is loaded after ``torch`` so that it shares torch's HIP runtime
(one runtime per process; streams and device pointers are torch's).  There is
no fallback: if the library is missing or a call fails, a RuntimeError says so.
"""
import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libseistorch_amd.so")

_p = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_f = ctypes.c_float

# name -> (restype, argtypes); kept in the order of include/seistorch_amd.h
SIGNATURES = {
    "seis_last_error": (ctypes.c_char_p, []),
    "seis_abi_version": (_i32, []),
    "seis_profile_enable": (ctypes.c_int, [_i32]),
    "seis_profile_collect": (ctypes.c_int, [_i32, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(_i64)]),
    "seis_stream_copy": (ctypes.c_int, [_p, _p, _i64, _i32, _p]),
    "seis_ring_len": (_i64, [_i64, _i64, _i32, _i32]),
    "seis_snapshot_count": (_i64, [_i64, _i32]),
    "seis_acoustic_step": (ctypes.c_int, [_p, _p, _p, _p, _p, _i64, _i64, _i64, _f, _f, _p]),
    "seis_acoustic_step_vjp": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64,
                                              _f, _f, _p]),
    "seis_acoustic_forward": (ctypes.c_int, [_p, _p, _p, _i64, _p, _p, _p, _i64, _p, _i64, _p,
                                             _p, _p, _i32, _p, _p, _p, _p, _p, _p, _i64, _i64,
                                             _i64, _f, _f, _i32, _i32, _i32, _i32, _i32, _p]),
    "seis_acoustic_backward": (ctypes.c_int, [_p, _p, _p, _i64, _p, _p, _p, _i64, _p, _i64, _p,
                                              _p, _p, _i32, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                              _p, _p, _i64, _i64, _i64, _f, _f, _i32, _i32, _i32,
                                              _i32, _i32, _p]),
    "seis_acoustic_forward_batch": (ctypes.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                                   _p, _i32, _p, _p, _p, _p, _p, _i32, _i32, _i32,
                                                   _i32, _i32, _p]),
    "seis_acoustic_backward_batch": (ctypes.c_int, [_i64, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                                    _p, _i32, _p, _p, _p, _p, _p, _p, _p, _i32,
                                                    _i32, _i32, _i32, _i32, _p]),
    "seis_habc_step": (ctypes.c_int, [_p, _p, _p, _p, _p, _i64, _i64, _i64, _f, _f, _i32, _p]),
    "seis_habc_step_vjp": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _f,
                                          _f, _i32, _p]),
    "seis_habc_forward": (ctypes.c_int, [_p, _p, _p, _i64, _p, _p, _i64, _p, _p, _p, _i32, _p,
                                         _p, _p, _p, _i64, _i64, _i64, _f, _f, _i32, _i32, _i32,
                                         _p]),
    "seis_rho_habc_step": (ctypes.c_int, [_p, _p, _p, _p, _p, _p, _i64, _i64, _i64,
                                          ctypes.c_double, ctypes.c_double, _i32, _i32, _p]),
    "seis_rho_habc_forward": (ctypes.c_int, [_p, _p, _p, _p, _i64, _p, _p, _i64, _p, _p, _p,
                                             _i32, _p, _p, _p, _p, _p, _i64, _i64, _i64,
                                             ctypes.c_double, ctypes.c_double, _i32, _i32, _i32,
                                             _p]),
    "seis_rho_habc_backward": (ctypes.c_int, [_p, _p, _p, _p, _i64, _p, _p, _i64, _p, _p, _p,
                                              _i32, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i64,
                                              _i64, _i64, ctypes.c_double, ctypes.c_double, _i32,
                                              _i32, _i32, _p]),
    "seis_habc_backward": (ctypes.c_int, [_p, _p, _p, _i64, _p, _p, _i64, _p, _p, _p, _i32, _p,
                                          _p, _p, _p, _p, _p, _p, _i64, _i64, _i64, _f, _f, _i32,
                                          _i32, _i32, _p]),
    "seis_elastic_coefficients": (ctypes.c_int, [_p, _p, _p, _p, _p, _i64, _i64, _f, _f, _p]),
    "seis_elastic_stage": (ctypes.c_int, [_i32, _p, _p, _f, _i32, _i32, _i32, _i32, _i64, _i64,
                                          _i64, _f, _f, _i32, _p]),
    "seis_elastic_stage_vjp": (ctypes.c_int, [_i32, _p, _p, _p, _p, _p, _f, _i32, _i32, _i32,
                                              _i32, _i64, _i64, _i64, _f, _f, _i32, _p]),
    "seis_elastic_chain": (ctypes.c_int, [_p, _i32, _p, _p, _p, _p, _p, _p, _i64, _i64, _f, _f,
                                          _p]),
    "seis_elastic_forward": (ctypes.c_int, [_p, _p, _i64, _p, _i32, _p, _i64, _p, _i32, _p, _p,
                                            _p, _i32, _p, _i32, _p, _p, _i64, _i64, _i64, _f, _f,
                                            _i32, _i32, _i32, _p]),
    "seis_elastic_backward": (ctypes.c_int, [_p, _p, _i64, _p, _i32, _p, _i64, _p, _i32, _p, _p,
                                             _p, _i32, _p, _p, _p, _p, _p, _i64, _i64, _i64, _f,
                                             _f, _i32, _i32, _i32, _p]),
    "seis_filtfilt": (ctypes.c_int, [_p, _p, _p, _i64, _i64, _p, _p, _i32, _p]),
    "seis_gaussian_filter": (ctypes.c_int, [_p, _p, _i64, _i64, _i32, _p, _i32, _p]),
    "seis_smooth_gradient": (ctypes.c_int, [_p, _p, _p, _i64, _i64, _p, _i32, _p, _i32, _i32,
                                            _p, _p, _p, _p]),
    "seis_smooth_gradient_noalloc": (ctypes.c_int, [_p, _p, _i64, _i64, _p, _i32, _p, _i32,
                                                    _i32, _p, _p, _p, _p]),
    "seis_unet_conv3x3": (ctypes.c_int, [_p, _i32, _i32, _i32, _i32, _p, _p, _i32, _i32, _p, _p,
                                         _p, _p]),
    "seis_unet_flip_weights": (ctypes.c_int, [_p, _i32, _i32, _p, _p]),
    "seis_unet_conv3x3_wgrad": (ctypes.c_int, [_p, _p, _i32, _i32, _i32, _i32, _i32, _p, _p,
                                               _p]),
    "seis_unet_gelu_bwd": (ctypes.c_int, [_p, _p, _i64, _p, _p]),
    "seis_unet_to_bf16": (ctypes.c_int, [_p, _i64, _p, _p]),
    "seis_unet_avgpool2": (ctypes.c_int, [_p, _i32, _i32, _i32, _i32, _p, _p]),
    "seis_unet_avgpool2_bwd": (ctypes.c_int, [_p, _i32, _i32, _i32, _i32, _p, _i32, _p]),
    "seis_unet_up_concat": (ctypes.c_int, [_p, _p, _i32, _i32, _i32, _i32, _i32, _p, _p]),
    "seis_unet_up_concat_bwd": (ctypes.c_int, [_p, _i32, _i32, _i32, _i32, _i32, _p, _p, _i32,
                                               _p]),
    "seis_unet_pack_input": (ctypes.c_int, [_p, _p, _i32, _i32, _i32, _p, _p, _p]),
    "seis_unet_head": (ctypes.c_int, [_p, _i64, _i32, _p, _p, _p, _p]),
    "seis_unet_head_bwd": (ctypes.c_int, [_p, _p, _i64, _i32, _p, _p, _p, _p, _p]),
    "seis_unet_mse": (ctypes.c_int, [_p, _p, _p, _i64, _f, _p, _p, _p]),
    "seis_unet_adamw": (ctypes.c_int, [_p, _p, _p, _p, _i64, _f, _f, _f, _f, _f, _i32, _p, _p]),
    "seis_unet_fill": (ctypes.c_int, [_p, _i64, _f, _p]),
}

ABI_VERSION = 15
BATCH_COLS = 23  # SEIS_BATCH_COLS: int64 columns per instance of the batch table

_lock = threading.Lock()
_lib = None


def load():
    """Load (once) and return the library; raise RuntimeError if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"seistorch_amd: HIP library not built ({LIB_PATH} missing); "
                    "run `python -m seistorch_amd.build`")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            if lib.seis_abi_version() != ABI_VERSION:
                raise RuntimeError("seistorch_amd: library ABI version mismatch; rebuild it")
            _lib = lib
    return _lib


def call(name, *args):
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.seis_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed (code {rc}): {msg}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def ptr_array(tensors):
    """A C array of device pointers (float* const*), kept alive by the caller."""
    return (ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])


def stream_of(t):
    """The current torch stream on the tensor's device, as a void*."""
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def require_device(*tensors):
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise RuntimeError("seistorch_amd kernels need tensors on a HIP device; "
                               f"got a tensor on {t.device}")
        if t.dtype != torch.float32 and t.dtype not in (torch.int32, torch.int64):
            raise RuntimeError(f"seistorch_amd kernels compute in float32; got {t.dtype}")
        if not t.is_contiguous():
            raise RuntimeError("seistorch_amd kernels need contiguous tensors")

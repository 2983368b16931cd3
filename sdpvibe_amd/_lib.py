"""ctypes binding of the gfx950 library libsdpvibe_hip.so (declared in include/sdpvibe_hip.h).

Torch tensors are only containers here: their device pointers and sizes go through the C ABI,
and every launch is enqueued on torch's current HIP stream. There is no fallback: a missing
library, a CPU tensor or an error code raises.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsdpvibe_hip.so")

_P, _I64, _I32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
SIGNATURES = {
    "sdpv_version": ([], _I32),
    "sdpv_source_hash": ([], ctypes.c_uint64),
    "sdpv_host_alloc": ([_I64], _P),
    "sdpv_host_free": ([_P], _I32),
    "sdpv_shift_round": ([_P, _I64, _I64, _P, _P, _I64, _P, _I32, _P], _I32),
    "sdpv_add_same": ([_P, _P, _P, _I64, _I64, _I32, _P], _I32),
    "sdpv_bitlen": ([_P, _I64, _I64, _P, _P], _I32),
    "sdpv_badd_align": ([_P, _P, _P, _P, _P, _P, _I64, _I64, _I64, _I64, _I32, _P], _I32),
    "sdpv_exact_dot_work": ([_I64, _I64, _I64, _I64, _I64], _I64),
    "sdpv_exact_dot": ([_P, _I64, _I64, _I64, _I64, _P, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _P,
                        _I64, _P, _P, _I64, _P], _I32),
    "sdpv_mpf_top": ([_P, _P, _I64, _I64, _P, _P], _I32),
    "sdpv_mpf_mul": ([_P, _P, _P, _P, _P, _P, _I64, _I64, _I32, _P], _I32),
    "sdpv_mpf_add": ([_P, _P, _P, _P, _P, _P, _I64, _I64, _I32, _P], _I32),
    "sdpv_mpf_norm": ([_P, _P, _I64, _P, _P, _I64, _I64, _P], _I32),
    "sdpv_row_exp": ([_P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _P, _P], _I32),
    "sdpv_rns_split": ([_P, _P, _P, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64, _I64,
                        _P, _P, _I64, _P, _P, _P, _I64, _I64, _P, _P, _P, _P], _I32),
    "sdpv_rns_gemm": ([_P, _P, _P, _P, _I64, _I64, _I64, _I64, _I64, _P, _I64, _P, _P, _P, _I64, _I32, _I32, _P],
                      _I32),
    "sdpv_rns_crt": ([_P, _I64, _I64, _I64, _I64, _I64, _I64, _P, _P, _P, _P, _I64, _I64, _I32, _P, _P,
                      _P, _P, _I64, _I64, _I64, _I64, _I64, _P, _P, _P, _P, _P], _I32),
    "sdpv_tile_chol": ([_P, _I64, _I64, _I64, _P, _P, _P, _P, _P], _I32),
    "sdpv_tile_trsv": ([_P, _P, _P, _I64, _I64, _I64, _I64, _P], _I32),
    "sdpv_mpf_to_double": ([_P, _P, _I64, _I64, _P, _I64, _P, _P], _I32),
    "sdpv_mpf_from_double": ([_P, _I64, _I64, _P, _P, _P], _I32),
    "sdpv_symeig": ([_P, _I64, _I64, _P, _P, _P, _P, _P, _P], _I32),
    "sdpv_symeig_static": ([_P, _I64, _I64, _P, _P, _P, _P, _P], _I32),
    "sdpv_mfma_selftest": ([_P, _P, _P, _P], _I32),
    "sdpv_decode_bigfloats": ([_P, _I64, _I64, _I64, _I32, _I32, _P, _P, _I64], _I64),
    "sdpv_encode_bigfloats": ([_P, _P, _I64, _I32, _I32, _P, _I64, _I64], _I64),
}
ERRORS = {-1: "bad argument", -2: "unsupported size", -3: "kernel launch failed"}
EXPECTED_VERSION = 1

_lib = None


class HipLibraryError(RuntimeError):
    pass


def lib():
    """The loaded library (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HipLibraryError(f"{LIB_PATH} is missing: build it with `python -m sdpvibe_amd.build`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes, fn.restype = args, res
        if handle.sdpv_version() != EXPECTED_VERSION:
            raise HipLibraryError("libsdpvibe_hip.so is stale: rebuild it")
        from .build import source_hash
        if handle.sdpv_source_hash() != source_hash():
            raise HipLibraryError("libsdpvibe_hip.so was built from other sources than the ones shipped "
                                  "with it: rebuild it with `python -m sdpvibe_amd.build`")
        _lib = handle
    return _lib


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise HipLibraryError(f"{name}: {ERRORS.get(rc, rc)}")


def ptr(t):
    """Device pointer of a contiguous GPU tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("the HIP library takes GPU tensors only")
    if not t.is_contiguous():
        raise ValueError("the HIP library takes contiguous tensors only")
    return ctypes.c_void_p(t.data_ptr())


def ptr_strided(t):
    """Device pointer of a GPU tensor view whose strides the caller passes to the entry point
    (a 4-d limb view: non-negative strides, unit stride along the last dimension)"""
    if not t.is_cuda:
        raise ValueError("the HIP library takes GPU tensors only")
    if t.dim() != 4 or t.stride(-1) != 1 or min(t.stride()) < 0:
        raise ValueError("strided operands are 4-d views with unit stride along the last dimension")
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

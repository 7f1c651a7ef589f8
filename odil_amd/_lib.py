"""Loader for the C-ABI library `libodil_amd.so` (declared in
`include/odil_amd.h`).

The library is loaded with ctypes; device buffers are torch tensors whose
`data_ptr()` is passed as a plain pointer, and every call takes the HIP
stream torch is currently using. There is no fallback: if the library is
missing or fails to load, every accelerated entry point raises.
"""

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libodil_amd.so')

F32, F64 = 0, 1
# ABI of include/odil_amd.h these bindings call (odil_abi_version).
ABI_VERSION = 7
_lib = None

c_int = ctypes.c_int
c_double = ctypes.c_double
c_vp = ctypes.c_void_p
c_ip = ctypes.POINTER(ctypes.c_int)


class OdilError(RuntimeError):
    """Failure reported by the native library (launch error, capacity)."""


def lib():
    """Return the loaded library, loading it on first use."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OdilError(
                'odil_amd native library not built ({}); run '
                '`python -m odil_amd.build`'.format(LIB_PATH))
        handle = ctypes.CDLL(LIB_PATH)
        handle.odil_last_error.restype = ctypes.c_char_p
        found = handle.odil_abi_version()
        if found != ABI_VERSION:
            raise OdilError(
                'odil_amd native library {} has ABI {}, these bindings need '
                '{}; rebuild with `python -m odil_amd.build`'.format(
                    LIB_PATH, found, ABI_VERSION))
        _lib = handle
    return _lib


def check(status, what=''):
    """Map a C-ABI status code to the exception types of the reference:
    1 (bad argument) -> ValueError, anything else -> OdilError."""
    if status == 0:
        return
    msg = lib().odil_last_error().decode(errors='replace')
    if status == 1:
        raise ValueError('{}: {}'.format(what, msg))
    raise OdilError('{}: {} (status {})'.format(what, msg, status))


def dtype_tag(dtype):
    if dtype == torch.float64:
        return F64
    if dtype == torch.float32:
        return F32
    raise ValueError('unsupported dtype {}'.format(dtype))


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return c_vp(t.data_ptr())


def ptrs(ts):
    """C array of device pointers (entries may be None)."""
    arr = (c_vp * len(ts))()
    for k, t in enumerate(ts):
        arr[k] = None if t is None else t.data_ptr()
    return arr


def stream():
    """Handle of torch's current HIP stream on the current device."""
    return c_vp(torch.cuda.current_stream().cuda_stream)


def require_device(*ts):
    for t in ts:
        if t is not None and not t.is_cuda:
            raise ValueError('odil_amd kernels need device tensors; got a '
                             'tensor on {}'.format(t.device))
        if t is not None and not t.is_contiguous():
            raise ValueError('odil_amd kernels need contiguous tensors')

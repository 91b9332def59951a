"""rocPRIM flagged select over the packed MSB-first mask -- the vendor
baseline reported beside the engine (the analogue of the reference's
cub::DeviceSelect::Flagged, src/benchmarks.cuh:384-394).  A separate library
(libcsx_baseline.so); nothing in the engine depends on it."""
import ctypes
import os

import torch

from . import PKG, CsxError, _ptr, _require_device, _stream

LIB_PATH = os.path.join(PKG, "libcsx_baseline.so")
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m gpu_compressstore2_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.csxb_rocprim_select_temp_size.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(ctypes.c_size_t)]
        L.csxb_rocprim_select_temp_size.restype = ctypes.c_int
        L.csxb_rocprim_select.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                          ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.csxb_rocprim_select.restype = ctypes.c_int
        L.csxb_rocprim_select_bytes_temp_size.argtypes = L.csxb_rocprim_select_temp_size.argtypes
        L.csxb_rocprim_select_bytes_temp_size.restype = ctypes.c_int
        L.csxb_rocprim_select_bytes.argtypes = L.csxb_rocprim_select.argtypes
        L.csxb_rocprim_select_bytes.restype = ctypes.c_int
        _LIB = L
    return _LIB


def temp_size(elem_size, n):
    b = ctypes.c_size_t(0)
    e = lib().csxb_rocprim_select_temp_size(elem_size, n, ctypes.byref(b))
    if e != 0:
        raise CsxError(3, f"rocprim select size query failed ({e})")
    return b.value


def rocprim_select(inp, mask, n=None, out=None, count=None, temp=None, stream=None):
    """(out, count tensor): rocprim::select with flags extracted from the mask."""
    _require_device(inp, mask, out, count)
    n = inp.numel() if n is None else n
    if out is None:
        out = torch.empty(max(n, 1), dtype=inp.dtype, device=inp.device)
    if count is None:
        count = torch.zeros(1, dtype=torch.int64, device=inp.device)
    if temp is None:
        temp = torch.empty(max(temp_size(inp.element_size(), n), 16), dtype=torch.uint8, device=inp.device)
    e = lib().csxb_rocprim_select(inp.element_size(), _ptr(inp), _ptr(mask), _ptr(out), _ptr(count), n, _ptr(temp), temp.numel(),
                                  _stream(stream))
    if e != 0:
        raise CsxError(3, f"rocprim select failed ({e})")
    return out, count


def temp_size_bytes(elem_size, n):
    b = ctypes.c_size_t(0)
    e = lib().csxb_rocprim_select_bytes_temp_size(elem_size, n, ctypes.byref(b))
    if e != 0:
        raise CsxError(3, f"rocprim select size query failed ({e})")
    return b.value


def rocprim_select_bytes(inp, flags_u8, n=None, out=None, count=None, temp=None, stream=None):
    """rocprim::select with one flag byte per element (flags expanded beforehand)."""
    _require_device(inp, flags_u8, out, count)
    n = inp.numel() if n is None else n
    if out is None:
        out = torch.empty(max(n, 1), dtype=inp.dtype, device=inp.device)
    if count is None:
        count = torch.zeros(1, dtype=torch.int64, device=inp.device)
    if temp is None:
        temp = torch.empty(max(temp_size_bytes(inp.element_size(), n), 16), dtype=torch.uint8, device=inp.device)
    e = lib().csxb_rocprim_select_bytes(inp.element_size(), _ptr(inp), _ptr(flags_u8), _ptr(out), _ptr(count), n, _ptr(temp), temp.numel(),
                                        _stream(stream))
    if e != 0:
        raise CsxError(3, f"rocprim select failed ({e})")
    return out, count

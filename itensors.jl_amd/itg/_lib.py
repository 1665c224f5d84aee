"""ctypes binding of libitg.so (the C ABI in include/itg.h).

This is the Python image of the `ccall` layer the Julia package extension
would use: every function here forwards plain pointers, sizes and labels.
Device memory is owned by the caller (torch tensors on the HIP device).
There is no fallback: if the library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libitg.so")

ITG_OK = 0
ITG_ERR_INVALID = 1
ITG_ERR_UNSUPPORTED = 2
ITG_ERR_HIP = 3
ITG_ERR_NOMEM = 4

ITG_F32 = 0
ITG_F64 = 1
ITG_C64 = 2
ITG_C128 = 3

PREC = {"fp64": 0, "split_bf16": 1, "split_int8": 2}

ITG_FLAG_CONJ_A = 1
ITG_FLAG_CONJ_B = 2
ITG_FLAG_KEEP_SLICES_1 = 4
ITG_FLAG_KEEP_SLICES_2 = 8
ITG_FLAG_REUSE_SLICES_1 = 16
ITG_FLAG_REUSE_SLICES_2 = 32
ITG_BS_ORDER_SEQUENTIAL = 0
ITG_BS_ORDER_THREADED = 1

# every symbol include/itg.h declares
EXPORTS = (
    "itg_init", "itg_destroy", "itg_last_error", "itg_set_stream", "itg_synchronize",
    "itg_version", "itg_permute", "itg_contract_dense", "itg_contract_dense_ex",
    "itg_bs_plan_create", "itg_bs_plan_sizes", "itg_bs_plan_blocks", "itg_bs_plan_pairs",
    "itg_bs_execute", "itg_bs_execute_noalloc", "itg_bs_plan_destroy", "itg_bs_permute", "itg_bs_autotune",
    "itg_batch_plan_create", "itg_batch_execute", "itg_batch_plan_destroy",
    "itg_malloc", "itg_free", "itg_memcpy_htod", "itg_memcpy_dtoh", "itg_memcpy_dtod",
    "itg_timer_enable", "itg_timer_count", "itg_timer_read", "itg_set_split_tile",
    "itg_bs_copy", "itg_bs_chain_create", "itg_bs_chain_execute", "itg_bs_chain_destroy",
    "itg_bs_diag_plan_create", "itg_bs_diag_execute",
    "itg_graph_begin", "itg_graph_end", "itg_graph_launch", "itg_graph_destroy",
    "itg_bs_permute_ex", "itg_convert", "itg_set_gemm_backend", "itg_bs_plan_info",
    "itg_set_blocksparse_order",
)


class ItgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"libitg status {status}: {msg}")
        self.status = status


class ItgUnsupported(ItgError):
    """The backend does not handle this case (ITG_ERR_UNSUPPORTED)."""


_lib = None
_lock = threading.Lock()

P = ctypes.c_void_p
I64P = ctypes.POINTER(ctypes.c_int64)
I32P = ctypes.POINTER(ctypes.c_int32)
DP = ctypes.POINTER(ctypes.c_double)
I = ctypes.c_int
I64 = ctypes.c_int64


def lib():
    """Load libitg.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C itensors.jl_amd/csrc` "
                "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.itg_init.argtypes = [I, ctypes.POINTER(P)]
        L.itg_destroy.argtypes = [P]
        L.itg_last_error.argtypes = [P]
        L.itg_last_error.restype = ctypes.c_char_p
        L.itg_set_stream.argtypes = [P, P]
        L.itg_synchronize.argtypes = [P]
        L.itg_version.restype = ctypes.c_char_p
        L.itg_permute.argtypes = [P, I, I, I64P, I32P, P, P, DP, DP]
        dense = [P, I, I, I64P, I64P, P, I, I64P, I64P, P, I, I64P, I64P, P, DP, DP, I]
        L.itg_contract_dense.argtypes = dense
        L.itg_contract_dense_ex.argtypes = dense + [ctypes.c_uint]
        L.itg_bs_plan_create.argtypes = [P, I, I64, I64P, I64P, I64P, I64P, I64P,
                                         I, I64, I64P, I64P, I64P, I64P, I64P,
                                         I, I64P, ctypes.POINTER(P)]
        L.itg_bs_plan_sizes.argtypes = [P, I64P, I64P, I64P]
        L.itg_bs_plan_blocks.argtypes = [P, I64P, I64P]
        L.itg_bs_plan_pairs.argtypes = [P, I64P]
        L.itg_bs_plan_info.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
        L.itg_bs_execute.argtypes = [P, P, I, P, P, P, DP, DP, I, ctypes.c_uint]
        L.itg_bs_execute_noalloc.argtypes = L.itg_bs_execute.argtypes
        L.itg_bs_plan_destroy.argtypes = [P]
        L.itg_bs_permute.argtypes = [P, I, I, I64, I64P, I64P, I64P, I32P, P, P]
        L.itg_bs_permute_ex.argtypes = [P, I, I, I64, I64P, I64P, I64P, P, I32P, P, P, DP, DP,
                                        I32P]
        L.itg_bs_autotune.argtypes = [P, P, I, P, P, P, I]
        L.itg_bs_copy.argtypes = [P, I, I64, I, I64P, I64P, I64P, I64P, I64P, P, I64, P, I64]
        L.itg_bs_chain_create.argtypes = [P, P, P, ctypes.POINTER(P)]
        L.itg_bs_diag_plan_create.argtypes = L.itg_bs_plan_create.argtypes
        L.itg_bs_diag_execute.argtypes = [P, P, I, P, P, P, DP, DP]
        L.itg_bs_chain_execute.argtypes = [P, P, I, P, P, P, P]
        L.itg_bs_chain_destroy.argtypes = [P]
        PP = ctypes.POINTER(P)
        L.itg_batch_plan_create.argtypes = [P, I, I64, I, I64P, I64P, PP, I, I64P, I64P, PP,
                                            I, I64P, I64P, PP, ctypes.POINTER(P)]
        L.itg_batch_execute.argtypes = [P, P, DP, I]
        L.itg_batch_plan_destroy.argtypes = [P]
        SZ = ctypes.c_size_t
        L.itg_malloc.argtypes = [P, SZ, ctypes.POINTER(P)]
        L.itg_free.argtypes = [P, P]
        for n in ("itg_memcpy_htod", "itg_memcpy_dtoh", "itg_memcpy_dtod"):
            getattr(L, n).argtypes = [P, P, P, SZ]
        L.itg_convert.argtypes = [P, I, P, I, P, I64]
        L.itg_graph_begin.argtypes = [P]
        L.itg_graph_end.argtypes = [P, ctypes.POINTER(P)]
        L.itg_graph_launch.argtypes = [P, P]
        L.itg_graph_destroy.argtypes = [P]
        L.itg_timer_enable.argtypes = [P, I]
        L.itg_set_split_tile.argtypes = [P, I]
        L.itg_set_gemm_backend.argtypes = [P, I]
        L.itg_set_blocksparse_order.argtypes = [P, I]
        L.itg_timer_count.argtypes = [P]
        L.itg_timer_read.argtypes = [P, I64, ctypes.POINTER(ctypes.c_char_p), I64P, DP]
        for name in EXPORTS:
            if name not in ("itg_last_error", "itg_version"):
                getattr(L, name).restype = I
        L.itg_timer_count.restype = I64
        _lib = L
    return _lib


def i64(seq):
    a = np.ascontiguousarray(np.asarray(seq, dtype=np.int64).reshape(-1))
    return a


def ptr64(a):
    if a.size == 0:
        a = np.zeros(1, dtype=np.int64)
    return a.ctypes.data_as(I64P), a


def ptr32(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))
    if a.size == 0:
        a = np.zeros(1, dtype=np.int32)
    return a.ctypes.data_as(I32P), a


def scalar2(x):
    c = complex(x)
    a = (ctypes.c_double * 2)(c.real, c.imag)
    return ctypes.cast(a, DP), a


class Handle:
    """One itg_handle per (device, host thread)."""

    def __init__(self, device):
        L = lib()
        h = P()
        st = L.itg_init(int(device), ctypes.byref(h))
        if st != ITG_OK:
            raise ItgError(st, f"itg_init(device={device}) failed (no HIP device?)")
        self.h = h
        self.device = device

    def check(self, st):
        if st == ITG_OK:
            return
        msg = lib().itg_last_error(self.h).decode()
        if st == ITG_ERR_UNSUPPORTED:
            raise ItgUnsupported(st, msg)
        raise ItgError(st, msg)

    def set_stream(self, stream_ptr):
        self.check(lib().itg_set_stream(self.h, P(stream_ptr)))

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.itg_destroy(self.h)
        except Exception:
            pass


_handles = threading.local()


# itg_set_gemm_backend code applied to every handle (ndtensors.backend_*)
backend = [0]
# itg_set_blocksparse_order code applied to every handle
# (ndtensors.enable_threaded_blocksparse)
bs_order = [ITG_BS_ORDER_SEQUENTIAL]


def handle_for(device_index):
    d = getattr(_handles, "d", None)
    if d is None:
        d = _handles.d = {}
    h = d.get(device_index)
    if h is None:
        h = d[device_index] = Handle(device_index)
        if backend[0]:
            h.check(lib().itg_set_gemm_backend(h.h, backend[0]))
        if bs_order[0]:
            h.check(lib().itg_set_blocksparse_order(h.h, bs_order[0]))
    return h


def open_handles():
    """The handles this thread has opened (one per device)."""
    return list(getattr(_handles, "d", {}).values())

"""NDTensors-level interface of the MI355X backend.

Mirrors the NDTensors functions the reference's `*` / `contract` path goes
through, with the same argument meaning (label tuples in the ncon
convention, column-major storage, alpha/beta with beta == 0 overwriting):

  contract(T1, labels1, T2, labels2[, labelsR])    generic_tensor_operations.jl:87-119
  contract_(R, labelsR, T1, labels1, T2, labels2, alpha, beta)
                                   == contract!   dense/tensoralgebra/contract.jl:160-216
  contract_labels(labels1, labels2)                contraction_logic.jl:5-34
  contract_blockoffsets(...)                       blocksparse/contract_sequential.jl:1-41
  permutedims(T, perm) / permutedims_(R, T, perm, alpha, beta)
                                                   dense/densetensor.jl:212-275

Data lives on the HIP device as flat torch tensors; every contraction runs in
libitg.so. There is no CPU fallback: without the library or a GPU the calls
raise.
"""
from __future__ import annotations

import ctypes
import math
from collections import OrderedDict

import numpy as np

from . import _lib
from ._lib import ItgError, ItgUnsupported  # noqa: F401  (re-export)

try:  # torch provides device memory and streams
    import torch
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

# ------------------------------------------------------------------ flags
# Process-global settings in the style of NDTensors' Refs (using_tblis,
# gemm_backend: NDTensors/src/NDTensors.jl:243-255,
# abstractarray/tensoralgebra/contract.jl:15-24).
_precision = ["fp64"]


def set_precision(mode: str) -> str:
    """Select the arithmetic of GEMM-shaped contractions: "fp64" (FP64 MFMA),
    "split_int8" (FP64 operands as 7 int8 slices under per-row scales, slice
    products on INT8 MFMA with exact INT32 sums, FP64 recombination) or
    "split_bf16" (reserved; raises ItgUnsupported). Returns the previous mode."""
    if mode not in _lib.PREC:
        raise ValueError(f"unknown precision mode {mode!r}")
    old = _precision[0]
    _precision[0] = mode
    return old


def precision() -> str:
    return _precision[0]


# The reference's algorithm switches. gemm_backend selects the kernel family
# of contractions planned afterwards (itg_set_gemm_backend on every device
# handle): "Auto" by shape, "BLAS" the grouped MFMA GEMM whenever a mode is
# contracted, "Generic" the generic strided kernel. The threaded-block-sparse
# switch selects the block-pair order of block-sparse plans made afterwards
# (itg_set_blocksparse_order): the order of contract_blocks(
# ::Algorithm"threaded_threads") (NDTensors/src/blocksparse/
# contract_threaded.jl:2-73), which the reference uses when the switch is on
# and Julia runs more than one thread; it fixes the output's block order and
# the order of the sums into each output block. On the device every plan
# runs all its output blocks in one grouped launch either way. The TBLIS
# switch is accepted and changes nothing: there is no TBLIS.
_flags = {"threaded_blocksparse": False, "tblis": False, "gemm_backend": "Auto"}
_BACKEND_CODE = {"Auto": 0, "BLAS": 1, "Generic": 2}


def using_threaded_blocksparse() -> bool:
    """NDTensors/src/NDTensors.jl:106-184."""
    return _flags["threaded_blocksparse"]


def _set_threaded(on):
    old = _flags["threaded_blocksparse"]
    _flags["threaded_blocksparse"] = on
    code = _lib.ITG_BS_ORDER_THREADED if on else _lib.ITG_BS_ORDER_SEQUENTIAL
    _lib.bs_order[0] = code
    for h in _lib.open_handles():
        h.check(_lib.lib().itg_set_blocksparse_order(h.h, code))
    return old


def enable_threaded_blocksparse(enable=None) -> bool:
    """Block-sparse contractions planned from now on enumerate their block
    pairs in the threaded order (T2's blocks outer when T1 has no more blocks
    than T2; NDTensors/src/blocksparse/contract_threaded.jl:2-73), as the
    reference does with the switch on in a Julia session of more than one
    thread. enable=False disables (ITensors' enable_threaded_blocksparse(::Bool),
    src/global_variables.jl:156-162). Returns the previous state
    (NDTensors.jl:140-161)."""
    return _set_threaded(enable is None or bool(enable))


def disable_threaded_blocksparse() -> bool:
    """Back to the sequential pair order (contract_sequential.jl:26-39).
    Returns the previous state (NDTensors.jl:163-169)."""
    return _set_threaded(False)


def using_tblis() -> bool:
    """NDTensors.jl:243-255."""
    return _flags["tblis"]


def enable_tblis():
    _flags["tblis"] = True


def disable_tblis():
    _flags["tblis"] = False


def gemm_backend() -> str:
    """abstractarray/tensoralgebra/contract.jl:15-24 (:Auto, :BLAS, :Generic)."""
    return _flags["gemm_backend"]


def _set_backend(name):
    _flags["gemm_backend"] = name
    _lib.backend[0] = _BACKEND_CODE[name]
    for h in _lib.open_handles():
        h.check(_lib.lib().itg_set_gemm_backend(h.h, _BACKEND_CODE[name]))


def backend_auto():
    """Plans made from now on pick their kernels by shape."""
    _set_backend("Auto")


def backend_blas():
    """Plans made from now on use the grouped MFMA GEMM (TTGT) whenever a mode
    is contracted (the reference's BLAS.gemm! path)."""
    _set_backend("BLAS")


def backend_generic():
    """Plans made from now on use the generic strided kernel (no permutation,
    no MFMA; the reference's generic matmul path)."""
    _set_backend("Generic")


# ------------------------------------------------------------------ types
def _dtype_code(tdtype):
    if tdtype == torch.float64:
        return _lib.ITG_F64
    if tdtype == torch.complex128:
        return _lib.ITG_C128
    if tdtype == torch.float32:
        return _lib.ITG_F32
    if tdtype == torch.complex64:
        return _lib.ITG_C64
    raise ItgUnsupported(_lib.ITG_ERR_UNSUPPORTED,
                         f"dtype {tdtype} (Float32, Float64, ComplexF32, ComplexF64 only)")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else ctypes.c_void_p(0)


def _handle(dev):
    h = _lib.handle_for(dev.index if dev.index is not None else torch.cuda.current_device())
    h.set_stream(torch.cuda.current_stream(dev).cuda_stream)
    return h


def set_split_tile(tile: int, device=None) -> None:
    """Tile shape of the split-precision GEMM on `device`: 0 = per plan
    (autotuned, else by size), 1 = 64 x 64, 2 = 128 x 64, 3 = 64 x 64 with
    the A operand in registers (real data; complex uses 1). The shapes agree
    bit for bit; this exists for tests and tuning."""
    dev = torch.device(device) if device is not None else torch.device(
        "cuda", torch.cuda.current_device())
    h = _handle(dev)
    h.check(_lib.lib().itg_set_split_tile(h.h, int(tile)))


class timer:
    """Per-kernel device timing, the counterpart of NDTensors' global
    `timer::TimerOutput` (NDTensors/src/NDTensors.jl:96-99): inside the block
    every kernel launched by the library on `device` (outside graph capture)
    is bracketed by HIP events on its stream; on exit `results` maps kernel
    name -> (launches, total ms).

        with itg.ndtensors.timer() as t:
            seq.run_step(0)
        t.results["split_gemm_kernel"]
    """

    def __init__(self, device=None):
        self.dev = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.results = {}

    def __enter__(self):
        h = _handle(self.dev)
        h.check(_lib.lib().itg_timer_enable(h.h, 1))
        return self

    def __exit__(self, *exc):
        import ctypes

        h = _handle(self.dev)
        L = _lib.lib()
        n = L.itg_timer_count(h.h)
        for i in range(n):
            name = ctypes.c_char_p()
            calls = ctypes.c_int64()
            ms = ctypes.c_double()
            h.check(L.itg_timer_read(h.h, i, ctypes.byref(name), ctypes.byref(calls),
                                     ctypes.byref(ms)))
            self.results[name.value.decode()] = (calls.value, ms.value)
        h.check(L.itg_timer_enable(h.h, 0))
        return False


class Graph:
    """A HIP graph of library launches, recorded through the C ABI
    (itg_graph_begin / itg_graph_end) and replayed with launch() on the
    device's current stream. Everything run inside the recording must have
    run once before (plans, factor tables and workspace exist).

        g = Graph.capture(lambda: seq.run("split_int8"))
        g.launch()
    """

    def __init__(self, dev, graph):
        self.dev = dev
        self.graph = graph

    @staticmethod
    def capture(fn, device=None):
        dev = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        h = _handle(dev)
        L = _lib.lib()
        h.check(L.itg_graph_begin(h.h))
        graph = ctypes.c_void_p()
        try:
            fn()
        finally:
            st = L.itg_graph_end(h.h, ctypes.byref(graph))
        h.check(st)
        return Graph(dev, graph)

    def launch(self):
        h = _handle(self.dev)
        h.check(_lib.lib().itg_graph_launch(h.h, self.graph))

    def __del__(self):
        try:
            if getattr(self, "graph", None) and _lib._lib is not None:
                _lib._lib.itg_graph_destroy(self.graph)
        except Exception:
            pass


class DenseTensor:
    """Dense storage + dims (NDTensors/src/dense/dense.jl:5-16): a flat device
    vector holding a column-major array."""

    def __init__(self, data, dims):
        dims = tuple(int(d) for d in dims)
        if data.dim() != 1 or data.numel() != math.prod(dims):
            raise ValueError("data must be a flat vector with prod(dims) elements")
        self.data = data
        self.dims = dims

    @property
    def eltype(self):
        return self.data.dtype

    @property
    def ndims(self):
        return len(self.dims)

    def nnz(self):
        return self.data.numel()

    def to_numpy(self):
        return self.data.detach().cpu().numpy().reshape(self.dims, order="F")

    @staticmethod
    def from_numpy(a, device="cuda"):
        a = np.asarray(a)
        flat = np.asfortranarray(a).reshape(-1, order="F")
        return DenseTensor(torch.from_numpy(np.ascontiguousarray(flat)).to(device), a.shape)


class BlockSparseTensor:
    """BlockSparse storage: flat device data + insertion-ordered
    {Block (1-based tuple): offset} + per-mode sector sizes
    (NDTensors/src/blocksparse/blocksparse.jl:5-13, blockoffsets.jl:7-11,
    blockdims.jl:8-19)."""

    def __init__(self, data, blockoffsets, blockdims, *, structure_only=False):
        self.data = data
        self.blockoffsets = OrderedDict((tuple(int(c) for c in b), int(o)) for b, o in blockoffsets.items())
        self.blockdims = tuple(tuple(int(x) for x in bd) for bd in blockdims)
        if structure_only:  # block structure for planning; holds no values
            return
        self._validate()

    @classmethod
    def _checked(cls, data, blockoffsets, blockdims, end):
        """A tensor whose block structure was already checked: the output
        structure of a plan (built and checked by the library, `end` = its
        nnz), or the structure of an existing tensor with new data of the
        same length. Only the data length is checked."""
        T = cls.__new__(cls)
        T.data = data
        T.blockoffsets = blockoffsets if isinstance(blockoffsets, OrderedDict) \
            else OrderedDict(blockoffsets)
        T.blockdims = tuple(tuple(bd) for bd in blockdims)
        if data is None or data.dim() != 1 or data.numel() < end:
            raise ValueError("BlockSparseTensor data must be a flat vector covering its blocks")
        T._end = end
        return T

    def _validate(self):
        # every block must lie inside the data: the device kernels index the
        # flat vector through these offsets without further checks
        data = self.data
        if data is None or data.dim() != 1:
            raise ValueError("BlockSparseTensor data must be a flat vector")
        end = 0
        for b, o in self.blockoffsets.items():
            if len(b) != len(self.blockdims) or any(
                    not 1 <= c <= len(self.blockdims[i]) for i, c in enumerate(b)):
                raise ValueError(f"block {b} is outside the block dimensions")
            if o < 0:
                raise ValueError(f"negative offset {o} of block {b}")
            end = max(end, o + math.prod(self.blockdims_of(b)))
        if data.numel() < end:
            raise ValueError(f"data holds {data.numel()} elements, the blocks reach {end}")
        self._end = end

    @property
    def eltype(self):
        return self.data.dtype

    @property
    def ndims(self):
        return len(self.blockdims)

    @property
    def dims(self):
        return tuple(sum(bd) for bd in self.blockdims)

    def blockdims_of(self, block):
        return tuple(self.blockdims[i][c - 1] for i, c in enumerate(block))

    def nnzblocks(self):
        return len(self.blockoffsets)

    def nnz(self):
        return sum(math.prod(self.blockdims_of(b)) for b in self.blockoffsets)

    def block_numpy(self, block, host=None):
        host = self.data.detach().cpu().numpy() if host is None else host
        d = self.blockdims_of(block)
        o = self.blockoffsets[block]
        return host[o:o + math.prod(d)].reshape(d, order="F")

    def to_numpy_dense(self):
        """dense(T) (blocksparsetensor.jl:357-368) on the host."""
        host = self.data.detach().cpu().numpy()
        out = np.zeros(self.dims, dtype=host.dtype)
        starts = [np.concatenate([[0], np.cumsum(bd)]) for bd in self.blockdims]
        for b in self.blockoffsets:
            sl = tuple(slice(starts[i][c - 1], starts[i][c]) for i, c in enumerate(b))
            out[sl] = self.block_numpy(b, host)
        return out


def contract_labels(labels1, labels2):
    """Output labels: the positive labels of T1 in order, then those of T2
    (NDTensors/src/tensoroperations/contraction_logic.jl:5-34)."""
    return tuple(x for x in labels1 if x > 0) + tuple(x for x in labels2 if x > 0)


def _out_dims(T1, l1, T2, l2, lR):
    out = []
    for r in lR:
        if r in l1:
            out.append(T1.dims[l1.index(r)])
        elif r in l2:
            out.append(T2.dims[l2.index(r)])
        else:
            raise ValueError(f"label {r} of R is not in T1 or T2")
    return tuple(out)


class DiagBlockSparseTensor:
    """DiagBlockSparse storage (NDTensors/src/blocksparse/diagblocksparse.jl:
    1-60): blocks (k, ..., k) only, each storing its diagonal, min(block
    dims) values, at its offset in the flat device array `data`; or, when
    `uniform` is a number, that one value on every diagonal element and no
    array (UniformDiagBlockSparse, as `delta` builds). A tensor without QNs
    (Diag storage) is the one-block case."""

    def __init__(self, data, blockoffsets, blockdims, uniform=None, dtype=None):
        self.data = data
        self.uniform = uniform
        self._dtype = data.dtype if data is not None else (dtype or torch.float64)
        self.blockoffsets = OrderedDict((tuple(int(c) for c in b), int(o))
                                        for b, o in blockoffsets.items())
        self.blockdims = tuple(tuple(int(x) for x in bd) for bd in blockdims)
        for b in self.blockoffsets:
            if any(c != b[0] for c in b):
                raise ValueError(f"block {b} is off the block diagonal")
            if len(b) != len(self.blockdims) or any(
                    not 1 <= c <= len(self.blockdims[i]) for i, c in enumerate(b)):
                raise ValueError(f"block {b} is outside the block dimensions")
        if data is None:
            if uniform is None:
                raise ValueError("a DiagBlockSparseTensor needs data or a uniform value")
        else:
            if data.dim() != 1:
                raise ValueError("DiagBlockSparseTensor data must be a flat vector")
            end = max([o + self.diaglen_of(b) for b, o in self.blockoffsets.items()] or [0])
            if min(self.blockoffsets.values(), default=0) < 0 or data.numel() < end:
                raise ValueError(f"data holds {data.numel()} elements, the diagonal blocks "
                                 f"reach {end}")

    @property
    def eltype(self):
        return self._dtype

    @property
    def ndims(self):
        return len(self.blockdims)

    @property
    def dims(self):
        return tuple(sum(bd) for bd in self.blockdims)

    def blockdims_of(self, block):
        return tuple(self.blockdims[i][c - 1] for i, c in enumerate(block))

    def diaglen_of(self, block):
        return min(self.blockdims_of(block)) if self.ndims else 1

    def nnzblocks(self):
        return len(self.blockoffsets)

    def nnz(self):
        return sum(self.diaglen_of(b) for b in self.blockoffsets)

    def diag_numpy(self, block):
        n = self.diaglen_of(block)
        if self.uniform is not None:
            return np.full(n, self.uniform, dtype=np.complex128 if self._dtype == torch.complex128
                           else np.float64)
        o = self.blockoffsets[block]
        return self.data.detach().cpu().numpy()[o:o + n]

    def to_numpy_dense(self):
        """dense(T) on the host (for inspection and tests)."""
        out = np.zeros(self.dims, dtype=np.complex128 if self._dtype == torch.complex128
                       else np.float64)
        starts = [np.concatenate([[0], np.cumsum(bd)]).astype(int) for bd in self.blockdims]
        for b in self.blockoffsets:
            d = self.diag_numpy(b)
            for j in range(len(d)):
                out[tuple(starts[i][c - 1] + j for i, c in enumerate(b))] = d[j]
        return out

    def as_complex(self):
        if self._dtype == torch.complex128:
            return self
        if self.uniform is not None:
            return DiagBlockSparseTensor(None, self.blockoffsets, self.blockdims,
                                         uniform=complex(self.uniform), dtype=torch.complex128)
        return DiagBlockSparseTensor(convert(self.data, torch.complex128), self.blockoffsets,
                                     self.blockdims)


def _as_blocksparse(T):
    """A DenseTensor viewed as a one-block BlockSparseTensor (same data)."""
    if isinstance(T, DenseTensor):
        return BlockSparseTensor(T.data, {(1,) * T.ndims: 0}, tuple((d,) for d in T.dims))
    return T


def _conj_storage(T):
    """conj of a storage's values (a copy for complex data; real data and
    real uniform values are returned as they are)."""
    if isinstance(T, DiagBlockSparseTensor):
        if T.uniform is not None:
            u = complex(T.uniform).conjugate() if T.eltype == torch.complex128 else T.uniform
            return DiagBlockSparseTensor(None, T.blockoffsets, T.blockdims, uniform=u,
                                         dtype=T.eltype)
        if not T.data.is_complex():
            return T
        return DiagBlockSparseTensor(T.data.conj().resolve_conj(), T.blockoffsets, T.blockdims)
    if not T.data.is_complex():
        return T
    return _with_data(T, T.data.conj().resolve_conj())


def _contract_diag(T1, labels1, T2, labels2, labelsR, alpha=1.0):
    """contract(BlockSparse, DiagBlockSparse) in either order
    (NDTensors/src/blocksparse/diagblocksparse.jl:596-634); Dense with a
    one-block Diag goes the same way. The output labels default to the
    non-diagonal operand's first (diagblocksparse.jl:626-634)."""
    d1 = isinstance(T1, DiagBlockSparseTensor)
    d2 = isinstance(T2, DiagBlockSparseTensor)
    if d1 and d2:
        raise ItgUnsupported(_lib.ITG_ERR_UNSUPPORTED, "contraction of two diagonal tensors")
    if d1:
        T1, labels1, T2, labels2 = T2, labels2, T1, labels1
    if labelsR is None:
        labelsR = contract_labels(labels1, labels2)
    labelsR = tuple(labelsR)
    dense_out = isinstance(T1, DenseTensor)
    B, D = _as_blocksparse(T1), T2
    if B.eltype != D.eltype:
        if B.eltype == torch.float64:
            B = BlockSparseTensor(convert(B.data, torch.complex128), B.blockoffsets, B.blockdims)
        else:
            D = D.as_complex()
    p = bs_plan(B, labels1, D, labels2, labelsR, diag=True)
    blockdims = [B.blockdims[labels1.index(r)] if r in labels1 else D.blockdims[labels2.index(r)]
                 for r in labelsR]
    data = torch.empty(p.nnzR, dtype=B.eltype, device=B.data.device)
    if p.npairs:
        p.execute_diag(data, B.data, D, alpha=alpha)
    if dense_out:
        dims = tuple(sum(bd) for bd in blockdims)
        if not p.npairs:
            data = torch.zeros(math.prod(dims), dtype=B.eltype, device=B.data.device)
        return DenseTensor(data, dims)
    return BlockSparseTensor._checked(data, OrderedDict(zip(p.blocksR, p.offsR)), blockdims,
                                      p.nnzR)


def _promote(T1, T2):
    """Mixed element types are promoted by allocation to their common type
    (Float32 < Float64, real < complex; promote_type), as
    dense/tensoralgebra/contract.jl:196-211 does."""
    if T1.eltype == T2.eltype:
        return T1, T2
    dt = torch.promote_types(T1.eltype, T2.eltype)
    if T1.eltype != dt:
        T1 = _with_data(T1, convert(T1.data, dt))
    if T2.eltype != dt:
        T2 = _with_data(T2, convert(T2.data, dt))
    return T1, T2


def convert(data, dtype):
    """A widened copy of a flat device vector (itg_convert): Float32 ->
    Float64 / complex, real -> complex with zero imaginary parts."""
    if data.dtype == dtype:
        return data
    out = torch.empty(data.numel(), dtype=dtype, device=data.device)
    h = _handle(data.device)
    h.check(_lib.lib().itg_convert(h.h, _dtype_code(data.dtype), _ptr(data), _dtype_code(dtype),
                                   _ptr(out), data.numel()))
    return out


def _with_data(T, data):
    if isinstance(T, DenseTensor):
        return DenseTensor(data, T.dims)
    end = getattr(T, "_end", None)
    if end is not None and type(T) is BlockSparseTensor:
        return BlockSparseTensor._checked(data, T.blockoffsets, T.blockdims, end)
    return BlockSparseTensor(data, T.blockoffsets, T.blockdims)


def contract_(R, labelsR, T1, labels1, T2, labels2, alpha=1.0, beta=0.0, *, prec=None,
              conj1=False, conj2=False):
    """contract!(R, labelsR, T1, labels1, T2, labels2, alpha, beta) for dense
    tensors: R = alpha*T1*T2 + beta*R; beta == 0 never reads R."""
    if not (isinstance(R, DenseTensor) and isinstance(T1, DenseTensor) and isinstance(T2, DenseTensor)):
        raise TypeError("contract_ takes DenseTensors (block-sparse outputs come from contract)")
    T1, T2 = _promote(T1, T2)
    if R.eltype != T1.eltype:
        raise TypeError(f"output eltype {R.eltype} does not match promoted input eltype {T1.eltype}")
    dev = R.data.device
    h = _handle(dev)
    dt = _dtype_code(R.eltype)
    lR, l1, l2 = _lib.i64(labelsR), _lib.i64(labels1), _lib.i64(labels2)
    dR, d1, d2 = _lib.i64(R.dims), _lib.i64(T1.dims), _lib.i64(T2.dims)
    a, _ka = _lib.scalar2(alpha)
    b, _kb = _lib.scalar2(beta)
    flags = (_lib.ITG_FLAG_CONJ_A if conj1 else 0) | (_lib.ITG_FLAG_CONJ_B if conj2 else 0)
    p = _lib.PREC[prec or _precision[0]]
    args = []
    for arr in (dR, lR, d1, l1, d2, l2):
        ptr, keep = _lib.ptr64(arr)
        args.append((ptr, keep))
    st = _lib.lib().itg_contract_dense_ex(
        h.h, dt,
        len(R.dims), args[0][0], args[1][0], _ptr(R.data),
        len(T1.dims), args[2][0], args[3][0], _ptr(T1.data),
        len(T2.dims), args[4][0], args[5][0], _ptr(T2.data),
        a, b, p, flags)
    h.check(st)
    return R


# ------------------------------------------------------- block-sparse plans
class BSPlan:
    """A libitg block-sparse contraction plan: the output block list (in the
    reference's first-appearance order), its offsets, and the device tables of
    the grouped launches."""

    def __init__(self, T1, labels1, T2, labels2, labelsR, diag=False):
        dev = T1.data.device
        self.h = _handle(dev)
        L = _lib.lib()

        def side(T):
            nb = T.nnzblocks()
            blocks = _lib.i64([c for b in T.blockoffsets for c in b]) if nb else _lib.i64([])
            offs = _lib.i64(list(T.blockoffsets.values()))
            nsec = _lib.i64([len(bd) for bd in T.blockdims])
            secd = _lib.i64([x for bd in T.blockdims for x in bd])
            return nb, blocks, offs, nsec, secd

        n1, b1, o1, ns1, sd1 = side(T1)
        n2, b2, o2, ns2, sd2 = side(T2)
        l1, l2, lR = _lib.i64(labels1), _lib.i64(labels2), _lib.i64(labelsR)
        keep = []

        def pp(a):
            ptr, k = _lib.ptr64(a)
            keep.append(k)
            return ptr

        plan = ctypes.c_void_p()
        create = L.itg_bs_diag_plan_create if diag else L.itg_bs_plan_create
        st = create(self.h.h, T1.ndims, n1, pp(b1), pp(o1), pp(ns1), pp(sd1), pp(l1),
                    T2.ndims, n2, pp(b2), pp(o2), pp(ns2), pp(sd2), pp(l2),
                    len(labelsR), pp(lR), ctypes.byref(plan))
        self.h.check(st)
        self.plan = plan
        nblk, nnz, npairs = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        L.itg_bs_plan_sizes(plan, ctypes.byref(nblk), ctypes.byref(nnz), ctypes.byref(npairs))
        self.nblkR, self.nnzR, self.npairs = nblk.value, nnz.value, npairs.value
        # elements each operand's data must hold (the kernels read through
        # the offsets unchecked; execute() verifies the lengths first)
        self.need1, self.need2 = _extent(T1), _extent(T2)
        NR = len(labelsR)
        blocks = np.zeros(max(NR * self.nblkR, 1), dtype=np.int64)
        offs = np.zeros(max(self.nblkR, 1), dtype=np.int64)
        L.itg_bs_plan_blocks(plan, blocks.ctypes.data_as(_lib.I64P), offs.ctypes.data_as(_lib.I64P))
        self.blocksR = [tuple(int(x) for x in blocks[i * NR:(i + 1) * NR]) for i in range(self.nblkR)]
        self.offsR = [int(x) for x in offs[:self.nblkR]]

    def info(self):
        """{"strategy": "direct" | "fiber" | "gemm" | None, "perm1", "perm2",
        "permR"}: how the plan runs (itg_bs_plan_info); the perm flags say
        whether the FP64 GEMM permutes T1, T2 or its output."""
        arr = (ctypes.c_int32 * 4)()
        _lib.lib().itg_bs_plan_info(self.plan, arr)
        strat = {0: "direct", 1: "fiber", 2: "gemm"}.get(arr[0])
        return {"strategy": strat, "perm1": bool(arr[1]), "perm2": bool(arr[2]),
                "permR": bool(arr[3])}

    def pairs(self):
        L = _lib.lib()
        arr = np.zeros(max(3 * self.npairs, 1), dtype=np.int64)
        L.itg_bs_plan_pairs(self.plan, arr.ctypes.data_as(_lib.I64P))
        return [tuple(int(x) for x in arr[3 * i:3 * i + 3]) for i in range(self.npairs)]

    def _check_lengths(self, R_data, T1_data, T2_data):
        for name, t, need in (("R", R_data, self.nnzR), ("T1", T1_data, self.need1),
                              ("T2", T2_data, self.need2)):
            if t is not None and t.numel() < need:
                raise ValueError(f"{name} holds {t.numel()} elements, the plan needs {need}")

    def execute(self, R_data, T1_data, T2_data, alpha=1.0, prec=None, flags=0, pair_alpha=None):
        self._check_lengths(R_data, T1_data, T2_data)
        h = _handle(R_data.device)
        dt = _dtype_code(R_data.dtype)
        a, _ka = _lib.scalar2(alpha)
        pa = None
        if pair_alpha is not None:
            arr = np.zeros(2 * self.npairs, dtype=np.float64)
            c = np.asarray(pair_alpha, dtype=np.complex128)
            arr[0::2], arr[1::2] = c.real, c.imag
            self._pa_keep = arr
            pa = arr.ctypes.data_as(_lib.DP)
        st = _lib.lib().itg_bs_execute(h.h, self.plan, dt, _ptr(R_data), _ptr(T1_data), _ptr(T2_data),
                                       a, pa, _lib.PREC[prec or _precision[0]], int(flags))
        h.check(st)

    def execute_diag(self, R_data, T1_data, D, alpha=1.0):
        """R = alpha * T1 * D for a diagonal plan; D a DiagBlockSparseTensor."""
        self._check_lengths(R_data, T1_data, D.data if D.uniform is None else None)
        h = _handle(R_data.device)
        a, _ka = _lib.scalar2(alpha)
        u, _ku = (_lib.scalar2(D.uniform) if D.uniform is not None else (None, None))
        st = _lib.lib().itg_bs_diag_execute(h.h, self.plan, _dtype_code(R_data.dtype),
                                            _ptr(R_data), _ptr(T1_data),
                                            _ptr(D.data) if D.data is not None else None, u, a)
        h.check(st)

    def autotune(self, R_data, T1_data, T2_data, prec=None):
        self._check_lengths(R_data, T1_data, T2_data)
        h = _handle(R_data.device)
        st = _lib.lib().itg_bs_autotune(h.h, self.plan, _dtype_code(R_data.dtype), _ptr(R_data),
                                        _ptr(T1_data), _ptr(T2_data), _lib.PREC[prec or _precision[0]])
        h.check(st)

    def __del__(self):
        try:
            if getattr(self, "plan", None) and _lib._lib is not None:
                _lib._lib.itg_bs_plan_destroy(self.plan)
        except Exception:
            pass


class BSChain:
    """Two consecutive block-sparse contractions in one pass: Z = (X*B1)*B2
    from the plans p1 (X*B1 -> Y) and p2 (Y*B2 -> Z), Y never stored.
    Raises ItgUnsupported when the pair has no fused form. Z is
    bit-identical to p1.execute then p2.execute with alpha = 1 (FP64 in
    every precision mode)."""

    def __init__(self, p1, p2):
        self.h = p1.h
        chain = ctypes.c_void_p()
        self.h.check(_lib.lib().itg_bs_chain_create(self.h.h, p1.plan, p2.plan,
                                                    ctypes.byref(chain)))
        self.chain = chain
        self.need = (p2.nnzR, p1.need1, p1.need2, p2.need2)

    def execute(self, Z_data, X_data, B1_data, B2_data):
        for name, t, need in zip(("Z", "X", "B1", "B2"), (Z_data, X_data, B1_data, B2_data),
                                 self.need):
            if t.numel() < need:
                raise ValueError(f"{name} holds {t.numel()} elements, the chain needs {need}")
        h = _handle(Z_data.device)
        h.check(_lib.lib().itg_bs_chain_execute(h.h, self.chain, _dtype_code(Z_data.dtype),
                                                _ptr(Z_data), _ptr(X_data), _ptr(B1_data),
                                                _ptr(B2_data)))

    def __del__(self):
        try:
            if getattr(self, "chain", None) and _lib._lib is not None:
                _lib._lib.itg_bs_chain_destroy(self.chain)
        except Exception:
            pass


def _extent(T):
    """One past the last element any stored block of T occupies."""
    if isinstance(T, DiagBlockSparseTensor):
        if T.uniform is not None:
            return 0
        return max([o + T.diaglen_of(b) for b, o in T.blockoffsets.items()] or [0])
    return max([o + math.prod(T.blockdims_of(b)) for b, o in T.blockoffsets.items()] or [0])


_plan_cache: "OrderedDict[tuple, BSPlan]" = OrderedDict()
_PLAN_CACHE_MAX = 64


def _structure_key(T):
    return (tuple(T.blockoffsets.items()), T.blockdims)


def bs_plan(T1, labels1, T2, labels2, labelsR, diag=False):
    """Plan of a block-sparse contraction, cached by block structure (the
    structure repeats every time a network is applied). diag: T2 is a
    DiagBlockSparseTensor."""
    key = (_structure_key(T1), tuple(labels1), _structure_key(T2), tuple(labels2), tuple(labelsR),
           T1.data.device.index, diag, _flags["gemm_backend"], _flags["threaded_blocksparse"])
    p = _plan_cache.get(key)
    if p is not None:
        _plan_cache.move_to_end(key)
        return p
    p = BSPlan(T1, labels1, T2, labels2, labelsR, diag=diag)
    _plan_cache[key] = p
    while len(_plan_cache) > _PLAN_CACHE_MAX:
        _plan_cache.popitem(last=False)
    return p


def contract_blockoffsets(T1, labels1, T2, labels2, labelsR):
    """(blockoffsetsR, nnzR, plan) with plan = [(block1, block2, blockR)] as
    1-based block tuples in the reference's order: sequential
    (blocksparse/contract_sequential.jl:1-41), or threaded
    (contract_threaded.jl:2-73, contract_generic.jl:3-32) while
    enable_threaded_blocksparse() is in effect."""
    p = bs_plan(T1, labels1, T2, labels2, labelsR)
    offs = OrderedDict(zip(p.blocksR, p.offsR))
    b1 = list(T1.blockoffsets)
    b2 = list(T2.blockoffsets)
    plan = [(b1[i], b2[j], p.blocksR[r]) for i, j, r in p.pairs()]
    return offs, p.nnzR, plan


def contract(T1, labels1, T2, labels2, labelsR=None, *, alpha=1.0, prec=None, conj1=False,
             conj2=False):
    """NDTensors.contract(T1, labels1, T2, labels2[, labelsR]): allocate the
    (uninitialised) output and contract into it."""
    labels1, labels2 = tuple(labels1), tuple(labels2)
    if isinstance(T1, DiagBlockSparseTensor) or isinstance(T2, DiagBlockSparseTensor):
        # the diagonal kernel takes no conjugation flags: conjugate first
        if conj1:
            T1 = _conj_storage(T1)
        if conj2:
            T2 = _conj_storage(T2)
        return _contract_diag(T1, labels1, T2, labels2, labelsR, alpha)
    if labelsR is None:
        labelsR = contract_labels(labels1, labels2)
    labelsR = tuple(labelsR)
    T1, T2 = _promote(T1, T2)
    if isinstance(T1, DenseTensor) and isinstance(T2, DenseTensor):
        dims = _out_dims(T1, labels1, T2, labels2, labelsR)
        R = DenseTensor(torch.empty(math.prod(dims), dtype=T1.eltype, device=T1.data.device), dims)
        return contract_(R, labelsR, T1, labels1, T2, labels2, alpha, 0.0, prec=prec,
                         conj1=conj1, conj2=conj2)
    if isinstance(T1, BlockSparseTensor) and isinstance(T2, BlockSparseTensor):
        p = bs_plan(T1, labels1, T2, labels2, labelsR)
        blockdims = []
        for r in labelsR:
            if r in labels1:
                blockdims.append(T1.blockdims[labels1.index(r)])
            else:
                blockdims.append(T2.blockdims[labels2.index(r)])
        data = torch.empty(p.nnzR, dtype=T1.eltype, device=T1.data.device)
        R = BlockSparseTensor._checked(data, OrderedDict(zip(p.blocksR, p.offsR)), blockdims,
                                       p.nnzR)
        if p.npairs:
            flags = (_lib.ITG_FLAG_CONJ_A if conj1 else 0) | (_lib.ITG_FLAG_CONJ_B if conj2 else 0)
            p.execute(data, T1.data, T2.data, alpha=alpha, prec=prec, flags=flags)
        return R
    raise ItgUnsupported(_lib.ITG_ERR_UNSUPPORTED, "mixed Dense / BlockSparse contraction")


def _bs_permutedims_(R, T, perm, alpha, beta):
    """permutedims!(R::BlockSparseTensor, T::BlockSparseTensor, perm, f) with
    f(r, t) = beta*r + alpha*t (NDTensors/src/blocksparse/blocksparsetensor.jl:
    834-881): over union(R's blocks, T's permuted blocks), R's block (zero
    if absent) becomes f(r, perm(t)) (t zero if T lacks the block); blocks R
    lacks are appended to R's storage, and kept only when nonzero. R is
    updated in place (data and blockoffsets) and returned."""
    perm = tuple(int(x) for x in perm)
    N = len(perm)
    if T.ndims != N or R.ndims != N:
        raise ValueError("permutation length does not match the tensors")
    inv = [0] * N
    for i, p in enumerate(perm):
        inv[p - 1] = i + 1
    if tuple(R.blockdims) != tuple(T.blockdims[p - 1] for p in perm):
        raise ValueError("R's block dimensions are not T's permuted ones")
    if R.eltype != T.eltype:
        raise TypeError("R and T must share one element type")
    dev = T.data.device
    h = _handle(dev)
    L = _lib.lib()
    tb = {tuple(b[p - 1] for p in perm): (b, o) for b, o in T.blockoffsets.items()}
    order = list(R.blockoffsets)
    new = [bp for bp in tb if bp not in R.blockoffsets]
    extent = _extent(R)
    offs_dst = list(R.blockoffsets.values())
    acc = extent
    for bp in new:
        offs_dst.append(acc)
        acc += math.prod(R.blockdims_of(bp))
    order += new
    data = torch.empty(acc, dtype=R.eltype, device=dev)
    if complex(beta) != 0 and extent:
        h.check(L.itg_memcpy_dtod(h.h, _ptr(data), _ptr(R.data), extent * R.data.element_size()))
    dims, offs_src = [], []
    for bp in order:
        bt = tb.get(bp)
        dims.extend(R.blockdims_of(bp)[i - 1] for i in inv)  # T's mode order
        offs_src.append(bt[1] if bt else -1)
    fresh = np.array([0] * len(R.blockoffsets) + [1] * len(new), dtype=np.uint8)
    nz = np.zeros(len(order), dtype=np.int32)
    keep = []

    def pp(a):
        ptr, k = _lib.ptr64(a)
        keep.append(k)
        return ptr

    a, _ka = _lib.scalar2(alpha)
    b, _kb = _lib.scalar2(beta)
    p32, _k32 = _lib.ptr32(list(perm))
    if order:
        h.check(L.itg_bs_permute_ex(
            h.h, _dtype_code(T.eltype), N, len(order), pp(_lib.i64(dims)), pp(_lib.i64(offs_src)),
            pp(_lib.i64(offs_dst)), fresh.ctypes.data_as(ctypes.c_void_p), p32, _ptr(T.data),
            _ptr(data), a, b, nz.ctypes.data_as(_lib.I32P)))
    nR = len(R.blockoffsets)
    kept = [i for i in range(nR, len(order)) if nz[i]]
    if len(kept) < len(new):  # drop the new blocks that came out all zero
        total = extent + sum(math.prod(R.blockdims_of(order[i])) for i in kept)
        final = torch.empty(total, dtype=R.eltype, device=dev)
        es = data.element_size()
        if extent:
            h.check(L.itg_memcpy_dtod(h.h, _ptr(final), _ptr(data), extent * es))
        o = extent
        new_offs = offs_dst[:nR]
        for i in kept:
            n = math.prod(R.blockdims_of(order[i]))
            h.check(L.itg_memcpy_dtod(h.h, ctypes.c_void_p(final.data_ptr() + o * es),
                                      ctypes.c_void_p(data.data_ptr() + offs_dst[i] * es), n * es))
            new_offs.append(o)
            o += n
        data = final
        order = order[:nR] + [order[i] for i in kept]
        offs_dst = new_offs
    R.data = data
    R.blockoffsets = OrderedDict(zip(order, offs_dst))
    return R


def permutedims_(R, T, perm, alpha=1.0, beta=0.0):
    """R = alpha*permutedims(T, perm) + beta*R (perm 1-based, Julia's); for
    block-sparse tensors the union-of-blocks form of permutedims!."""
    if isinstance(T, BlockSparseTensor):
        return _bs_permutedims_(R, T, perm, alpha, beta)
    h = _handle(T.data.device)
    d, _k = _lib.ptr64(_lib.i64(T.dims))
    pp, _kp = _lib.ptr32([int(x) for x in perm])
    a, _ka = _lib.scalar2(alpha)
    b, _kb = _lib.scalar2(beta)
    h.check(_lib.lib().itg_permute(h.h, _dtype_code(T.eltype), T.ndims, d, pp, _ptr(T.data),
                                   _ptr(R.data), a, b))
    return R


def permutedims(T, perm):
    """permutedims(T, perm) for Dense and BlockSparse tensors (perm 1-based).
    Block-sparse: every block is permuted; block coordinates are permuted
    and the blocks keep their order and offsets."""
    perm = tuple(int(x) for x in perm)
    if isinstance(T, DenseTensor):
        dims = tuple(T.dims[p - 1] for p in perm)
        R = DenseTensor(torch.empty_like(T.data), dims)
        return permutedims_(R, T, perm)
    h = _handle(T.data.device)
    blocks = list(T.blockoffsets)
    bd = _lib.i64([d for b in blocks for d in T.blockdims_of(b)])
    offs = _lib.i64(list(T.blockoffsets.values()))
    keep = []

    def pp(a):
        p, k = _lib.ptr64(a)
        keep.append(k)
        return p

    # output offsets accumulate in block order (blockoffsets.jl:103-111)
    dst = []
    acc = 0
    for b in blocks:
        dst.append(acc)
        acc += math.prod(T.blockdims_of(b))
    out = torch.empty(acc, dtype=T.eltype, device=T.data.device)
    p32, _k32 = _lib.ptr32(list(perm))
    h.check(_lib.lib().itg_bs_permute(h.h, _dtype_code(T.eltype), T.ndims, len(blocks), pp(bd),
                                      pp(offs), pp(_lib.i64(dst)), p32, _ptr(T.data), _ptr(out)))
    new_offs = OrderedDict((tuple(b[p - 1] for p in perm), o) for b, o in zip(blocks, dst))
    return BlockSparseTensor(out, new_offs, tuple(T.blockdims[p - 1] for p in perm))

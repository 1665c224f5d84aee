"""ITensors-level interface of the MI355X backend: Index / QN / ITensor and
`*` / `contract`, with the reference's semantics.

  QN arithmetic               src/lib/QuantumNumbers/src/qn.jl:6-30,178-207
  Index identity, dag         src/index.jl:24-33,354-355,362
  contraction labels          src/indexset.jl:672-707 (arrow check :684-688)
  flux, nzblocks              src/indexset.jl:876-883, src/qn/qnindexset.jl:9-18
  ITensor(flux, inds), random_itensor
                              src/qn/qnitensor.jl:158-166,383-460
  A * B, contract(A, B)       src/tensor_operations/tensor_algebra.jl:1-77
  contract(As; sequence)      src/tensor_operations/tensor_algebra.jl:106-159
  dag(T)                      src/itensor.jl:1806-1846
  combiner, combinedind       src/qn/qnitensor.jl:462-467, src/tensor_operations/
                              itensor_combiner.jl:1-34; index fusion src/qn/qnindex.jl:
                              359-388 with NDTensors/src/blocksparse/blockdims.jl:165-174;
                              combineblocks src/qn/qnindex.jl:407-424

Storage is an `ndtensors.DenseTensor` (no QNs) or `ndtensors.BlockSparseTensor`
(QN indices) on the HIP device; every contraction runs in libitg.so.
"""
from __future__ import annotations

import itertools
import math
import struct
import warnings
from collections import OrderedDict

import numpy as np

from . import combiners as CB
from . import ndtensors as ND

Out, In, Neither = 1, -1, 0


# -------------------------------------------------------------------- QN
class QN:
    """Up to four named values, each with an optional modulus; a missing name
    counts as zero (qn.jl:6-30)."""

    __slots__ = ("vals",)
    maxQNs = 4

    def __init__(self, *args):
        vals = {}
        if len(args) == 1 and isinstance(args[0], (int, np.integer)):
            args = (("", int(args[0])),)
        elif len(args) == 2 and all(isinstance(a, (int, np.integer)) for a in args):
            args = (("", int(args[0]), int(args[1])),)
        for a in args:
            name, v = a[0], int(a[1])
            m = int(a[2]) if len(a) > 2 else 1
            vals[name] = (v % m if m > 1 else v, m)
        if len(vals) > self.maxQNs:
            raise ValueError("a QN holds at most 4 named values")
        self.vals = vals

    @classmethod
    def _of(cls, vals):
        q = cls.__new__(cls)
        q.vals = vals
        return q

    def __add__(self, other):
        out = dict(self.vals)
        for n, (v, m) in other.vals.items():
            if n in out:
                v0, m0 = out[n]
                if m0 != m:
                    raise ValueError("QNVals must have same modulus")
                s = v0 + v
                out[n] = (s % m if m > 1 else s, m)
            else:
                if len(out) >= self.maxQNs:
                    raise ValueError("Cannot add QN, maximum number of QNVals reached")
                out[n] = (v, m)
        return QN._of(out)

    def __neg__(self):
        return QN._of({n: ((-v) % m if m > 1 else -v, m) for n, (v, m) in self.vals.items()})

    def __sub__(self, other):
        return self + (-other)

    def __rmul__(self, d):
        """Arrow * QN (qn.jl:156-162)."""
        if d == 1:
            return self
        if d == -1:
            return -self
        return QN._of({n: (0, m) for n, (v, m) in self.vals.items()})

    def _canon(self):
        return tuple(sorted((n, v) for n, (v, m) in self.vals.items() if v != 0))

    def __eq__(self, other):
        return isinstance(other, QN) and self._canon() == other._canon()

    def __hash__(self):
        return hash(self._canon())

    def __repr__(self):
        return "QN(" + ", ".join(f"({n!r}, {v})" for n, (v, m) in self.vals.items()) + ")"


# ------------------------------------------------------------------ TagSet
def _tag_key(t):
    """Order of two tags: each stored as 16 UInt16 characters and compared as
    the big-endian integer of their little-endian bytes
    (src/lib/SmallStrings/src/smallstring.jl:37-45,81-85,103), i.e. bytewise
    over (low byte, high byte) of each character, shorter tags first."""
    return b"".join(struct.pack("<H", ord(c)) for c in t).ljust(32, b"\0")


def tagstring(tags):
    """Canonical string of a tag set: whitespace dropped, tags cut to 16
    characters, empty and repeated tags dropped, at most four kept (in the
    order given), sorted (src/lib/TagSets/src/TagSets.jl:81-95,131-180,
    317-327)."""
    out = []
    for t in str(tags).replace(" ", "").split(","):
        t = t[:16]
        if t and t not in out and len(out) < 4:
            out.append(t)
    return ",".join(sorted(out, key=_tag_key))


# ------------------------------------------------------------------- Index
_next_id = itertools.count(1)


class Index:
    """An index: identity (id, plev, tags), a space (an int, or a list of
    (QN, dim) sectors) and an arrow. QN indices default to Out
    (src/qn/qnindex.jl:128)."""

    __slots__ = ("id", "space", "dir", "tags", "plev")

    def __init__(self, space, tags="", plev=0, dir=None, id=None):
        if isinstance(space, (int, np.integer)):
            self.space = int(space)
        else:
            self.space = tuple((q if isinstance(q, QN) else QN(q), int(d)) for q, d in space)
        self.tags = tagstring(tags)
        self.plev = int(plev)
        self.id = next(_next_id) if id is None else id
        if dir is None:
            dir = Out if self.hasqns else Neither
        self.dir = dir

    @property
    def hasqns(self):
        return not isinstance(self.space, int)

    @property
    def dim(self):
        return self.space if not self.hasqns else sum(d for _, d in self.space)

    def nblocks(self):
        return 1 if not self.hasqns else len(self.space)

    def blockdims(self):
        return (self.space,) if not self.hasqns else tuple(d for _, d in self.space)

    def qn(self, b):
        """QN of sector b (1-based, as Block numbers are)."""
        return self.space[b - 1][0]

    def __eq__(self, other):
        return isinstance(other, Index) and (self.id, self.plev, self.tags) == (other.id, other.plev, other.tags)

    def __hash__(self):
        return hash((self.id, self.plev, self.tags))

    def __repr__(self):
        return f"(dim={self.dim}|id={self.id % 1000}|{self.tags!r}){chr(39) * self.plev}"

    def prime(self, n=1):
        return Index(self.space, self.tags, self.plev + n, self.dir, self.id)

    def sim(self):
        return Index(self.space, self.tags, self.plev, self.dir)

    def dag(self):
        return Index(self.space, self.tags, self.plev, -self.dir, self.id)


def dag_index(i):
    """src/index.jl:362: flip the arrow, keep the identity."""
    return i.dag()


def prime(i, n=1):
    return i.prime(n)


def flux_of_block(inds, block):
    """flux(inds, block) = sum_n dir(i_n)*qn(i_n, block_n) (src/indexset.jl:876-883)."""
    q = QN()
    for i, b in zip(inds, block):
        q = q + i.dir * i.qn(b)
    return q


def eachblock(inds):
    """Block grid in column-major order (blockdims.jl:104-106), 1-based."""
    ranges = [range(1, i.nblocks() + 1) for i in inds]
    for rev in itertools.product(*reversed(ranges)):
        yield tuple(reversed(rev))


def nzblocks(flux, inds):
    """Blocks whose flux equals `flux` (src/qn/qnindexset.jl:9-18)."""
    return [b for b in eachblock(inds) if flux_of_block(inds, b) == flux]


def compute_contraction_labels(Ais, Bis):
    """src/indexset.jl:672-707 -- matched pairs get -1, -2, ... in
    (A position outer, B position inner) order; the rest are numbered from
    ncont+1 over A then B. Matching QN indices must have opposite arrows."""
    have_qns = all(i.hasqns for i in Ais) and all(j.hasqns for j in Bis)
    la = [0] * len(Ais)
    lb = [0] * len(Bis)
    ncont = 0
    for i, a in enumerate(Ais):
        for j, b in enumerate(Bis):
            if a == b:
                if have_qns and a.dir != -b.dir:
                    raise ValueError(
                        f"Attempting to contract IndexSet {Ais} with IndexSet {Bis}: QN indices "
                        f"must have opposite direction to contract, but {a} and {b} do not")
                ncont += 1
                la[i] = lb[j] = -ncont
    u = ncont
    for i in range(len(la)):
        if la[i] == 0:
            u += 1
            la[i] = u
    for j in range(len(lb)):
        if lb[j] == 0:
            u += 1
            lb[j] = u
    return tuple(la), tuple(lb)


# ------------------------------------------------------------------ ITensor
class ITensor:
    """An ITensor: indices + NDTensors storage on the device."""

    def __init__(self, tensor, inds):
        self.tensor = tensor
        self.inds = tuple(inds)
        if len(self.inds) != tensor.ndims:
            raise ValueError("number of indices does not match the storage")

    # ---- introspection
    @property
    def eltype(self):
        return self.tensor.eltype

    def hasqns(self):
        if isinstance(self.tensor, (CB.CombinerTensor, ND.DiagBlockSparseTensor)):
            return bool(self.inds) and all(i.hasqns for i in self.inds)
        return isinstance(self.tensor, ND.BlockSparseTensor)

    def isdiag(self):
        """src/itensor.jl:893."""
        return isinstance(self.tensor, ND.DiagBlockSparseTensor)

    def nnzblocks(self):
        return self.tensor.nnzblocks() if self.hasqns() else 1

    def nnz(self):
        return self.tensor.nnz()

    def blocks(self):
        return list(self.tensor.blockoffsets) if self.hasqns() else [(1,) * len(self.inds)]

    def flux(self):
        """Flux of the first stored block (src/qn/flux.jl:55-60); None when
        there are no blocks or no QNs."""
        if not self.hasqns() or self.nnzblocks() == 0:
            return None
        return flux_of_block(self.inds, next(iter(self.tensor.blockoffsets)))

    def checkflux(self):
        f = self.flux()
        for b in self.tensor.blockoffsets:
            if flux_of_block(self.inds, b) != f:
                raise ValueError("ITensor has blocks of different flux")
        return True

    def array(self, *order):
        """Dense host array (numpy), optionally in a given index order."""
        a = (self.tensor.to_numpy_dense() if self.hasqns() or self.isdiag()
             else self.tensor.to_numpy())
        if order:
            perm = [self.inds.index(i) for i in order]
            a = np.transpose(a, perm)
        return a

    def scalar(self):
        """T[] for a 0-index ITensor."""
        a = self.array()
        return a.reshape(-1)[0] if a.size else 0.0

    def norm(self):
        import torch
        return float(torch.linalg.vector_norm(self.tensor.data).item())

    def __mul__(self, other):
        if isinstance(other, ITensor):
            return contract(self, other)
        # a number: scale the stored values, keeping the storage kind (Diag,
        # uniform Diag, BlockSparse, Dense), as NDTensors' scalar * tensor does
        if not isinstance(other, (int, float, complex, np.number)):
            return NotImplemented
        if iscombiner(self):
            raise TypeError("a combiner holds no values to scale")
        return _scaled(self, other)

    __rmul__ = __mul__

    def __repr__(self):
        kind = "BlockSparse" if self.hasqns() else "Dense"
        return f"ITensor({kind}, inds={self.inds}, eltype={self.eltype})"


def dag(T):
    """dag(T): conjugate the data and flip every arrow (src/itensor.jl:1806-1846).
    Real data is aliased, as the reference does when aliasing is allowed. A
    combiner flips its arrows and its storage's conjugation flag
    (NDTensors/src/combiner/combiner.jl:39-40)."""
    if iscombiner(T):
        return ITensor(T.tensor.conj(), tuple(i.dag() for i in T.inds))
    if T.isdiag():
        D = T.tensor
        if D.uniform is not None:
            S = ND.DiagBlockSparseTensor(None, D.blockoffsets, D.blockdims,
                                         uniform=np.conj(D.uniform).item(), dtype=D.eltype)
        else:
            d = D.data.conj().resolve_conj() if D.data.is_complex() else D.data
            S = ND.DiagBlockSparseTensor(d, D.blockoffsets, D.blockdims)
        return ITensor(S, tuple(i.dag() for i in T.inds))
    data = T.tensor.data
    if data.is_complex():
        data = data.conj().resolve_conj()
    return ITensor(ND._with_data(T.tensor, data), tuple(i.dag() for i in T.inds))


# --------------------------------------------------------------- building
def _torch_dtype(dtype):
    import torch
    if dtype in (float, np.float64, "Float64", torch.float64):
        return torch.float64
    if dtype in (complex, np.complex128, "ComplexF64", torch.complex128):
        return torch.complex128
    raise ValueError(f"unsupported element type {dtype}")


def itensor_from_array(a, *inds, device="cuda"):
    """Dense ITensor holding the (column-major) array `a`."""
    return ITensor(ND.DenseTensor.from_numpy(a, device), inds)


def random_itensor(*args, dtype=np.float64, rng=None, device="cuda"):
    """random_itensor([flux,] inds...): dense N(0,1) entries, or for QN
    indices the blocks allowed by `flux` (default QN()) in eachblock order,
    data drawn in storage order (src/qn/qnitensor.jl:383-460). Complex data
    draws the real and imaginary parts as independent N(0,1)."""
    import torch
    if rng is None:
        rng = np.random.default_rng()
    if args and isinstance(args[0], QN):
        flux, inds = args[0], args[1:]
    else:
        flux, inds = None, args
    if len(inds) == 1 and isinstance(inds[0], (tuple, list)):
        inds = tuple(inds[0])
    tdt = _torch_dtype(dtype)
    cplx = tdt == torch.complex128
    qns = bool(inds) and all(i.hasqns for i in inds)
    if not qns:
        n = math.prod(i.dim for i in inds)
        x = rng.standard_normal(n)
        if cplx:
            x = x + 1j * rng.standard_normal(n)
        data = torch.from_numpy(np.ascontiguousarray(x)).to(device=device, dtype=tdt)
        return ITensor(ND.DenseTensor(data, tuple(i.dim for i in inds)), inds)
    if flux is None:
        flux = QN()
    blocks = nzblocks(flux, inds)
    if not blocks:
        raise ValueError(f"ITensor with flux={flux} resulted in no allowed blocks")
    offs = OrderedDict()
    nnz = 0
    bdims = tuple(i.blockdims() for i in inds)
    for b in blocks:
        offs[b] = nnz
        nnz += math.prod(bdims[m][c - 1] for m, c in enumerate(b))
    x = rng.standard_normal(nnz)
    if cplx:
        x = x + 1j * rng.standard_normal(nnz)
    data = torch.from_numpy(np.ascontiguousarray(x)).to(device=device, dtype=tdt)
    return ITensor(ND.BlockSparseTensor(data, offs, bdims), inds)


def blocksparse_itensor(data, blocks, inds, device="cuda"):
    """ITensor from host data (flat, storage order) and an ordered block list."""
    import torch
    offs = OrderedDict()
    nnz = 0
    bdims = tuple(i.blockdims() for i in inds)
    for b in blocks:
        offs[tuple(b)] = nnz
        nnz += math.prod(bdims[m][c - 1] for m, c in enumerate(b))
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(data))).to(device)
    if t.numel() != nnz:
        raise ValueError("data length does not match the blocks")
    return ITensor(ND.BlockSparseTensor(t, offs, bdims), inds)


# --------------------------------------------------------------- combiner
def _qn_order_key(names):
    """QN ordering (src/lib/QuantumNumbers/src/qn.jl:273-291): values compared
    in name order, a missing name counting as zero."""
    return lambda q: tuple(q.vals.get(n, (0, 1))[0] for n in names)


def combineblocks(space):
    """src/qn/qnindex.jl:407-424: stable sort of the sectors by QN, then
    merge neighbours with equal QN. Returns (merged space, perm, comb),
    perm and comb 1-based (comb[i] = merged sector of sorted sector i)."""
    names = sorted({n for q, _ in space for n in q.vals})
    key = _qn_order_key(names)
    perm = sorted(range(len(space)), key=lambda i: key(space[i][0]))
    merged, comb = [], []
    for i in perm:
        q, d = space[i]
        if merged and merged[-1][0] == q:
            merged[-1] = (q, merged[-1][1] + d)
        else:
            merged.append((q, d))
        comb.append(len(merged))
    return merged, [i + 1 for i in perm], comb


def fused_space(inds, dir):
    """Sectors of ⊗(inds...; dir): dir * sum_k dir(i_k) * qn_k with the first
    index's sectors fastest, sizes multiplied (src/qn/qnindex.jl:359-388;
    the left fold of NDTensors/src/blocksparse/blockdims.jl:165-167 leaves
    the intermediate arrows cancelled)."""
    spaces = [[(i.dir * q, d) for q, d in i.space] for i in inds]
    out = []
    for combo in itertools.product(*reversed(spaces)):
        q, d = QN(), 1
        for qk, dk in reversed(combo):
            q = q + qk
            d *= dk
        out.append((dir * q, d))
    return out


def combiner(*inds, dir=None, tags="CMB,Link"):
    """combiner(is...; dir, tags): an ITensor with Combiner storage whose
    indices are (combined index, dag.(is)...) (src/qn/qnitensor.jl:462-467,
    src/tensor_operations/itensor_combiner.jl:1-12). For QN indices the
    combined index holds the product sectors sorted by QN and merged; its
    arrow defaults to In when every index is In, else Out."""
    if len(inds) == 1 and isinstance(inds[0], (tuple, list)):
        inds = tuple(inds[0])
    if not inds:
        return ITensor(CB.CombinerTensor((), (), ()), ())
    if all(i.hasqns for i in inds):
        if dir is None:
            dir = In if all(i.dir == In for i in inds) else Out
        merged, perm, comb = combineblocks(fused_space(inds, dir))
        c = Index(merged, tags, dir=dir)
    else:
        c = Index(math.prod(i.dim for i in inds), tags)
        perm, comb = (), ()
    C = CB.CombinerTensor(perm, comb, (c.blockdims(),) + tuple(i.blockdims() for i in inds))
    return ITensor(C, (c,) + tuple(i.dag() for i in inds))


def iscombiner(T):
    return isinstance(T.tensor, CB.CombinerTensor)


def combinedind(C):
    """The combined index of a combiner, its first index (itensor_combiner.jl:25-30)."""
    return C.inds[0] if iscombiner(C) and C.inds else None


def _contract_combiner(A, B):
    lA, lB = compute_contraction_labels(A.inds, B.inds)
    first = iscombiner(A)
    C, T, lC, lT = (A, B, lA, lB) if first else (B, A, lB, lA)
    if iscombiner(T):
        raise ValueError("contracting two combiners")
    R, lR = CB.contract(T.tensor, lT, C.tensor, lC, combiner_first=first)
    indsR = tuple(C.inds[lC.index(x)] if x in lC else T.inds[lT.index(x)] for x in lR)
    return ITensor(R, indsR)


# ------------------------------------------------------------- contraction
_warn_order = [14]  # src/global_variables.jl: default_warn_order


def get_warn_order():
    return _warn_order[0]


def set_warn_order(n):
    """Returns the previous threshold; None disables the warning."""
    old = _warn_order[0]
    _warn_order[0] = n
    return old


def reset_warn_order():
    return set_warn_order(14)


def disable_warn_order():
    return set_warn_order(None)


# ------------------------------------------------------ diagonal ITensors
def nzdiagblocks(flux, inds):
    """Diagonal blocks (k, ..., k) whose flux is `flux`
    (src/qn/qnindexset.jl:20-29 over eachdiagblock)."""
    if not inds:
        return []
    n = min(i.nblocks() for i in inds)
    return [b for b in ((k,) * len(inds) for k in range(1, n + 1)) if flux_of_block(inds, b) == flux]


def _diag_storage(inds, flux, values=None, uniform=None, dtype=np.float64, device="cuda"):
    import torch
    tdt = torch.complex128 if np.dtype(dtype) == np.complex128 else torch.float64
    if all(i.hasqns for i in inds) and inds:
        blocks = nzdiagblocks(QN() if flux is None else flux, inds)
        bdims = tuple(i.blockdims() for i in inds)
    elif not any(i.hasqns for i in inds):
        blocks = [(1,) * len(inds)]
        bdims = tuple((i.dim,) for i in inds)
    else:
        raise ValueError("diagonal ITensor over a mix of QN and dense indices")
    offs, n = OrderedDict(), 0
    for b in blocks:
        offs[b] = n
        n += min(bdims[i][c - 1] for i, c in enumerate(b)) if inds else 1
    if uniform is not None:
        return ND.DiagBlockSparseTensor(None, offs, bdims, uniform=uniform, dtype=tdt)
    if values is None:
        data = torch.zeros(n, dtype=tdt, device=device)
    elif np.isscalar(values):
        data = torch.full((n,), values, dtype=tdt, device=device)
    else:
        v = np.asarray(values)
        if v.size != n:
            raise ValueError(f"{v.size} diagonal values for {n} diagonal elements")
        data = torch.from_numpy(np.ascontiguousarray(v, dtype=np.dtype(dtype))).to(device)
    return ND.DiagBlockSparseTensor(data, offs, bdims)


def delta(*inds, flux=None, dtype=np.float64):
    """delta([flux,] is...): ones on the diagonal of the blocks allowed by
    `flux` (UniformDiagBlockSparse; src/qn/qnitensor.jl:520-543, Diag for
    indices without QNs, src/itensor.jl:607-613)."""
    if inds and isinstance(inds[0], QN):
        flux, inds = inds[0], inds[1:]
    if len(inds) == 1 and isinstance(inds[0], (tuple, list)):
        inds = tuple(inds[0])
    one = 1.0 if np.dtype(dtype) != np.complex128 else complex(1.0)
    return ITensor(_diag_storage(inds, flux, uniform=one, dtype=dtype), inds)


def diag_itensor(x, *inds, flux=None, dtype=None, device="cuda"):
    """diag_itensor(x, [flux,] is...): the diagonal set to x, a number or a
    vector of every stored diagonal value in block order (DiagBlockSparse;
    src/qn/qnitensor.jl:484-517, src/itensor.jl:479-600)."""
    if inds and isinstance(inds[0], QN):
        flux, inds = inds[0], inds[1:]
    if len(inds) == 1 and isinstance(inds[0], (tuple, list)):
        inds = tuple(inds[0])
    if dtype is None:
        dtype = np.complex128 if np.iscomplexobj(x) else np.float64
    return ITensor(_diag_storage(inds, flux, values=x, dtype=dtype, device=device), inds)


δ = delta


def _scaled(T, x):
    """x * T, on the device through the library's permute (identity
    permutation, alpha = x) over the flat data; the storage kind is kept."""
    import torch
    S = T.tensor
    cplx = isinstance(x, (complex, np.complexfloating)) and complex(x).imag != 0
    if isinstance(S, ND.DiagBlockSparseTensor) and S.uniform is not None:
        u = S.uniform * x
        dt = torch.complex128 if (cplx or S.eltype == torch.complex128) else torch.float64
        return ITensor(ND.DiagBlockSparseTensor(None, S.blockoffsets, S.blockdims,
                                                uniform=u if dt == torch.complex128 else float(np.real(u)),
                                                dtype=dt), T.inds)
    data = S.data
    if cplx and not data.is_complex():
        data = ND.convert(data, torch.complex128)
    out = torch.empty_like(data)
    if data.numel():
        ND.permutedims_(ND.DenseTensor(out, (data.numel(),)), ND.DenseTensor(data, (data.numel(),)),
                        (1,), alpha=x if data.is_complex() else float(np.real(x)))
    if isinstance(S, ND.DiagBlockSparseTensor):
        return ITensor(ND.DiagBlockSparseTensor(out, S.blockoffsets, S.blockdims), T.inds)
    return ITensor(ND._with_data(S, out), T.inds)


def contract(A, B, *, prec=None, conjA=False, conjB=False):
    """A * B (src/tensor_operations/tensor_algebra.jl:1-77): contract every
    index A and B share; the result holds A's free indices (in A's order)
    then B's. QN indices must meet with opposite arrows."""
    if not iscombiner(A) and not iscombiner(B) and (len(A.inds) == 0 or len(B.inds) == 0):
        # an order-0 operand is a number: A[] * B (tensor_algebra.jl:66-77)
        if conjA:
            A = ITensor(ND._conj_storage(A.tensor), A.inds)
        if conjB:
            B = ITensor(ND._conj_storage(B.tensor), B.inds)
        if len(A.inds) == 0 and len(B.inds) == 0:
            return itensor_from_array(np.asarray(A.scalar() * B.scalar()),
                                      device=A.tensor.data.device)
        return _scaled(B, A.scalar()) if len(A.inds) == 0 else _scaled(A, B.scalar())
    if iscombiner(A) or iscombiner(B):
        # a combiner holds no values: a conjugation flag acts on the other
        # operand's data, conjugated before the data movement
        if conjA and not iscombiner(A):
            A = ITensor(ND._conj_storage(A.tensor), A.inds)
        if conjB and not iscombiner(B):
            B = ITensor(ND._conj_storage(B.tensor), B.inds)
        return _contract_combiner(A, B)
    lA, lB = compute_contraction_labels(A.inds, B.inds)
    lR = ND.contract_labels(lA, lB)
    if A.isdiag() and not B.isdiag():
        # contract(::DiagBlockSparse, l1, ::BlockSparse, l2) puts the other
        # operand's labels first (diagblocksparse.jl:626-634)
        lR = ND.contract_labels(lB, lA)
    indsR = tuple(A.inds[lA.index(x)] if x in lA else B.inds[lB.index(x)] for x in lR)
    if A.hasqns() != B.hasqns():
        raise ND.ItgUnsupported(2, "contracting a QN ITensor with a dense one")
    if _warn_order[0] is not None and len(indsR) >= _warn_order[0]:
        # the order warning of _contract (tensor_algebra.jl:8-22)
        warnings.warn(f"Contraction resulted in ITensor with {len(indsR)} indices, which is "
                      f"greater than or equal to the ITensor order warning threshold "
                      f"{_warn_order[0]}.")
    T = ND.contract(A.tensor, lA, B.tensor, lB, lR, prec=prec, conj1=conjA, conj2=conjB)
    return ITensor(T, indsR)


_sequence_optimization = [False]


def using_contraction_sequence_optimization():
    """src/global_variables.jl (default off)."""
    return _sequence_optimization[0]


def enable_contraction_sequence_optimization():
    _sequence_optimization[0] = True


def disable_contraction_sequence_optimization():
    _sequence_optimization[0] = False


def left_associative_contraction_sequence(n):
    """[[[1, 2], 3], ...] (src/tensor_operations/contraction_cost.jl:1-2)."""
    seq = 1
    for k in range(2, n + 1):
        seq = [seq, k] if k > 2 else [1, 2]
    return seq


def _seq_inds(As, seq, costs):
    if isinstance(seq, int):
        return list(As[seq - 1].inds)
    a = _seq_inds(As, seq[0], costs)
    b = _seq_inds(As, seq[1], costs)
    union = a + [i for i in b if i not in a]
    costs.append(math.prod(i.dim for i in union))
    return [i for i in a if i not in b] + [i for i in b if i not in a]


def contraction_cost(As, sequence=None):
    """Pairwise costs of contracting As along `sequence` (1-based, as the
    reference's): each pair costs dim(union(inds1, inds2)), sparsity
    ignored (src/tensor_operations/contraction_cost.jl:4-42)."""
    costs = []
    _seq_inds(list(As), sequence or left_associative_contraction_sequence(len(As)), costs)
    return costs


def optimal_contraction_sequence(As):
    """A sequence of least total cost (the cost model of contraction_cost)
    by dynamic programming over subsets of the tensors -- the search the
    reference delegates to TensorOperations.optimaltree
    (ext/ITensorsTensorOperationsExt/ITensorsTensorOperationsExt.jl:6-13).
    Exact; meant for the few-tensor networks it is used on (up to ~14)."""
    As = list(As)
    n = len(As)
    if n == 1:
        return 1
    if n > 16:
        raise ValueError("optimal_contraction_sequence: more than 16 tensors")
    full = (1 << n) - 1
    inds = [list(A.inds) for A in As]
    # indices of the tensor a subset contracts to: each pairwise step keeps
    # the symmetric difference, which is associative, so whatever the order
    # they are the indices met an odd number of times in the subset
    open_of, best, split = {}, {}, {}

    def open_inds(S):
        if S not in open_of:
            seen = {}
            for k in range(n):
                if S >> k & 1:
                    for i in inds[k]:
                        seen[i] = seen.get(i, 0) + 1
            open_of[S] = [i for i, c in seen.items() if c % 2 == 1]
        return open_of[S]

    for k in range(n):
        best[1 << k] = 0.0
    for S in range(1, full + 1):
        if S & (S - 1) == 0:
            continue
        low = S & -S
        T = (S - 1) & S
        while T:
            if T & low:  # each unordered split once
                U = S ^ T
                a, b = open_inds(T), open_inds(U)
                union = a + [i for i in b if i not in a]
                c = best[T] + best[U] + math.prod(i.dim for i in union)
                if S not in best or c < best[S]:
                    best[S], split[S] = c, (T, U)
            T = (T - 1) & S

    def build(S):
        if S & (S - 1) == 0:
            return S.bit_length()
        T, U = split[S]
        return [build(T), build(U)]

    return build(full)


def contract_network(As, sequence=None, *, prec=None):
    """contract(As; sequence) (src/tensor_operations/tensor_algebra.jl:121-159):
    "left_associative" (the default, unless sequence optimisation is
    enabled), "right_associative", "automatic" (optimal_contraction_sequence)
    or an explicit tree of 1-based positions, e.g. [[1, 2], 3]."""
    As = list(As)
    if sequence is None:
        sequence = "automatic" if using_contraction_sequence_optimization() else "left_associative"
    if sequence == "left_associative":
        out = As[0]
        for B in As[1:]:
            out = contract(out, B, prec=prec)
        return out
    if sequence == "right_associative":
        out = As[-1]
        for B in reversed(As[:-1]):
            out = contract(B, out, prec=prec)
        return out
    if sequence == "automatic":
        sequence = optimal_contraction_sequence(As)

    def rec(s):
        if isinstance(s, int):
            return As[s - 1]
        a, b = s
        return contract(rec(a), rec(b), prec=prec)

    return rec(sequence)

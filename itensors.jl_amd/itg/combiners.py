"""NDTensors-level combiner contraction on the MI355X backend.

  Combiner storage (blockperm, blockcomb)  NDTensors/src/combiner/combiner.jl:7-20
  combining / uncombining test, validity   NDTensors/src/combiner/combiner.jl:76-140
  dense combine / uncombine / index
  replacement / empty combiner             NDTensors/src/combiner/contract.jl:15-89
  block-sparse combine                     NDTensors/src/blocksparse/combiner.jl:25-81,
                                           blocksparse/blocksparsetensor.jl:487-630
  block-sparse uncombine                   NDTensors/src/blocksparse/combiner.jl:82-140,
                                           blocksparse/blocksparsetensor.jl:632-795

The block bookkeeping runs on the host, as it does in the reference; every
element of tensor data moves on the device, in one grouped strided copy
(itg_bs_copy) or one permutation (itg_permute). Blocks are 1-based tuples.
"""
from __future__ import annotations

import math
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from . import ndtensors as ND


class CombinerTensor:
    """A tensor with Combiner storage: `blockperm` sorts the product blocks
    of the uncombined indices (first index fastest) by QN, `blockcomb` gives
    the combined block of each sorted block (both 1-based; empty for dense
    combiners). `blockdims` are the sector sizes of its indices, combined
    index first. `isconj` is the storage's conjugation flag (flipped by dag;
    it affects nothing here, as in the reference without fermions)."""

    def __init__(self, blockperm, blockcomb, blockdims, isconj=False):
        self.blockperm = tuple(int(x) for x in blockperm)
        self.blockcomb = tuple(int(x) for x in blockcomb)
        self.blockdims = tuple(tuple(int(x) for x in bd) for bd in blockdims)
        self.isconj = bool(isconj)

    @property
    def ndims(self):
        return len(self.blockdims)

    @property
    def dims(self):
        return tuple(sum(bd) for bd in self.blockdims)

    @property
    def eltype(self):
        return None

    def nnz(self):
        return 0

    def conj(self):
        return CombinerTensor(self.blockperm, self.blockcomb, self.blockdims, not self.isconj)


def _invalid(lT, lC):
    raise ValueError(f"invalid combiner contraction: tensor labels {lT}, combiner labels {lC} "
                     "(combining leaves exactly the combined index uncontracted; uncombining "
                     "contracts exactly the combined index)")


def is_combining(lT, lC):
    """combinedind_label(C) not in the tensor's labels (combiner.jl:96-101),
    checked as in combiner.jl:103-130."""
    comb = lC[0] not in lT
    n = sum(1 for x in lC if (x not in lT) == comb)
    if n != 1:
        _invalid(lT, lC)
    return comb


def bs_copy(dst, src, descs):
    """Grouped strided copy on the device (itg_bs_copy). descs: list of
    (dims, src_off, src_strides, dst_off, dst_strides), element units."""
    if not descs:
        return
    nm = max(1, max(len(d[0]) for d in descs))
    dims, so, ss, do, ds = [], [], [], [], []
    for d, o1, s1, o2, s2 in descs:
        pad = nm - len(d)
        dims += list(d) + [1] * pad
        ss += list(s1) + [0] * pad
        ds += list(s2) + [0] * pad
        so.append(o1)
        do.append(o2)
    keep = []

    def p(a):
        ptr, k = _lib.ptr64(_lib.i64(a))
        keep.append(k)
        return ptr

    h = ND._handle(dst.device)
    h.check(_lib.lib().itg_bs_copy(h.h, ND._dtype_code(src.dtype), len(descs), nm, p(dims),
                                   p(so), p(ss), p(do), p(ds), ND._ptr(src), src.numel(),
                                   ND._ptr(dst), dst.numel()))


def _strides(dims):
    s, acc = [], 1
    for d in dims:
        s.append(acc)
        acc *= d
    return s


def _lin(idx, shape):
    """1-based column-major linear index of a 1-based multi-index."""
    lin, acc = 0, 1
    for i, n in zip(idx, shape):
        lin += (i - 1) * acc
        acc *= n
    return lin + 1


def _cart(lin, shape):
    """Inverse of _lin."""
    lin -= 1
    out = []
    for n in shape:
        out.append(lin % n + 1)
        lin //= n
    return tuple(out)


def _product_dims(bdims_uc):
    """Sector sizes of the product of the uncombined indices, blocks in
    column-major order over their block grids (blockdims.jl:165-174)."""
    shape = [len(b) for b in bdims_uc]
    return [math.prod(bdims_uc[k][c - 1] for k, c in enumerate(_cart(l, shape)))
            for l in range(1, math.prod(shape) + 1)]


# ------------------------------------------------------------ block sparse
def _combine_bs(T, lT, C, lC):
    """permutedims_combine (blocksparsetensor.jl:537-630): output labels
    (combined, rest of T in order)."""
    clabel, luc = lC[0], tuple(lC[1:])
    lR = (clabel,) + tuple(x for x in lT if x not in lC)
    ucpos = [lT.index(x) for x in luc]
    rest = [lT.index(x) for x in lR[1:]]
    bd_uc = [T.blockdims[p] for p in ucpos]
    shape = [len(b) for b in bd_uc]
    pdims = _product_dims(bd_uc)
    bp, comb = C.blockperm, C.blockcomb
    if sorted(bp) != list(range(1, len(pdims) + 1)) or len(comb) != len(bp):
        raise ValueError("combiner does not match the blocks of the combined indices")
    ipos = {b: j + 1 for j, b in enumerate(bp)}  # product block -> sorted position
    sdim = [pdims[b - 1] for b in bp]  # sector size at each sorted position
    ent = []
    for b, off in T.blockoffsets.items():
        lin = _lin([b[p] for p in ucpos], shape)
        pos = ipos[lin]
        nb = comb[pos - 1]
        o, j = 0, pos - 1
        while j >= 1 and comb[j - 1] == nb:  # preceding sorted blocks merged into nb
            o += sdim[j - 1]
            j -= 1
        ent.append((b, off, pos, nb, o))
    # output blocks: (sorted position, rest) in column-major order, then
    # merged and de-duplicated keeping first occurrences
    # (blocksparsetensor.jl:560-567, blockoffsets.jl:35-37, combine_blocks)
    keys = sorted(((pos,) + tuple(b[p] for p in rest) for b, _o, pos, _n, _x in ent),
                  key=lambda t: t[::-1])
    blocksR = list(OrderedDict(((comb[k[0] - 1],) + k[1:], None) for k in keys))
    bdR = (C.blockdims[0],) + tuple(T.blockdims[p] for p in rest)
    offsR, acc = OrderedDict(), 0
    for b in blocksR:
        offsR[b] = acc
        acc += math.prod(bdR[m][c - 1] for m, c in enumerate(b))
    data = torch.zeros(acc, dtype=T.data.dtype, device=T.data.device)
    descs = []
    for b, off, pos, nb, o in ent:
        dT = T.blockdims_of(b)
        bR = (nb,) + tuple(b[p] for p in rest)
        dR = tuple(bdR[m][c - 1] for m, c in enumerate(bR))
        sR = _strides(dR)
        dstr = [0] * len(dT)
        acc = 1
        for p in ucpos:  # uncombined modes: column-major inside the combined range
            dstr[p] = acc
            acc *= dT[p]
        for r, p in enumerate(rest):
            dstr[p] = sR[r + 1]
        descs.append((dT, off, _strides(dT), offsR[bR] + o, dstr))
    bs_copy(data, T.data, descs)
    return ND.BlockSparseTensor(data, offsR, bdR), lR


def _uncombine_bs(T, lT, C, lC):
    """uncombine (blocksparse/combiner.jl:82-140, blocksparsetensor.jl:664-795):
    output labels (uncombined in combiner order, rest of T in order)."""
    clabel, luc = lC[0], tuple(lC[1:])
    cpos = lT.index(clabel)
    rest = [p for p in range(len(lT)) if p != cpos]
    lR = luc + tuple(lT[p] for p in rest)
    bd_uc = list(C.blockdims[1:])
    shape = [len(b) for b in bd_uc]
    pdims = _product_dims(bd_uc)
    bp, comb = C.blockperm, C.blockcomb
    if sorted(bp) != list(range(1, len(pdims) + 1)) or len(comb) != len(bp):
        raise ValueError("combiner does not match the blocks of the combined indices")
    bdR = tuple(bd_uc) + tuple(T.blockdims[p] for p in rest)
    offsR, acc = OrderedDict(), 0
    plan = []
    for b, off in T.blockoffsets.items():
        nb = b[cpos]
        dT = T.blockdims_of(b)
        sT = _strides(dT)
        o = 0
        for j in range(1, len(comb) + 1):  # sorted blocks merged into nb, in order
            if comb[j - 1] != nb:
                continue
            lin = bp[j - 1]
            bR = _cart(lin, shape) + tuple(b[p] for p in rest)
            dR = tuple(bdR[m][c - 1] for m, c in enumerate(bR))
            if bR in offsR:
                raise ValueError("uncombining produced a block twice")
            offsR[bR] = acc
            # R block = T block restricted to [o, o + pdims[lin]) along the
            # combined mode; R's uncombined modes are that range, column-major
            dims = (pdims[lin - 1],) + tuple(dT[p] for p in rest)
            src_str = [sT[cpos]] + [sT[p] for p in rest]
            plan.append((dims, off + o * sT[cpos], src_str, acc, _strides(dims)))
            acc += math.prod(dR)
            o += pdims[lin - 1]
        if o != dT[cpos]:
            raise ValueError("combined block size does not match the combiner")
    data = torch.zeros(acc, dtype=T.data.dtype, device=T.data.device)
    bs_copy(data, T.data, plan)
    return ND.BlockSparseTensor(data, offsR, bdR), lR


# ------------------------------------------------------------------- dense
def _copy_data(x):
    """A device copy of a flat tensor (itg_memcpy_dtod)."""
    out = torch.empty_like(x)
    if x.numel():
        h = ND._handle(x.device)
        h.check(_lib.lib().itg_memcpy_dtod(h.h, ND._ptr(out), ND._ptr(x),
                                           x.numel() * x.element_size()))
    return out


def _copy_dense(T):
    return _copy_data(T.data)


def _contract_dense(T, lT, C, lC, combiner_first):
    """contract!! for Dense x Combiner (combiner/contract.jl:15-89). Output
    labels follow contract_labels of the operands in call order."""
    lR = ND.contract_labels(lC, lT) if combiner_first else ND.contract_labels(lT, lC)
    if C.ndims <= 1:  # empty combiner: a permutation to the output labels
        perm = tuple(lT.index(x) + 1 for x in lR)
        return ND.permutedims(T, perm), lR
    if C.ndims == 2 and sum(1 for x in lC if x in lT) == 1:
        # index replacement: same data, the one combiner index T holds
        # becomes the other one
        (cpos2,) = [i for i, x in enumerate(lT) if x in lC]
        new = [x for x in lC if x not in lT][0]
        labels = list(lT)
        labels[cpos2] = new
        dims = list(T.dims)
        dims[cpos2] = C.dims[lC.index(new)]
        return ND.DenseTensor(_copy_dense(T), dims), tuple(labels)
    if is_combining(lT, lC):
        clabel, luc = lC[0], tuple(lC[1:])
        cpos = lR.index(clabel)
        lperm = lR[:cpos] + luc + lR[cpos + 1:]
        perm = tuple(lT.index(x) + 1 for x in lperm)
        P = ND.permutedims(T, perm)
        dims = tuple(C.dims[0] if x == clabel else T.dims[lT.index(x)] for x in lR)
        return ND.DenseTensor(P.data, dims), lR
    # uncombining: same data, the combined index replaced by the uncombined
    # ones at its position (storage copied)
    cpos2 = lT.index(lC[0])
    labels = lT[:cpos2] + tuple(lC[1:]) + lT[cpos2 + 1:]
    dims = T.dims[:cpos2] + tuple(C.dims[1:]) + T.dims[cpos2 + 1:]
    return ND.DenseTensor(_copy_dense(T), dims), labels


def contract(T, lT, C, lC, combiner_first=False):
    """contract(T, labels, C::CombinerTensor, labels) -> (R, labelsR) with
    NDTensors' output label order."""
    lT, lC = tuple(lT), tuple(lC)
    if isinstance(T, ND.BlockSparseTensor):
        if C.ndims == 0:  # no indices combined: a copy (blocksparse/combiner.jl:154-163)
            data = _copy_data(T.data) if T.data.numel() else T.data
            return ND.BlockSparseTensor(data, T.blockoffsets, T.blockdims), lT
        if is_combining(lT, lC):
            return _combine_bs(T, lT, C, lC)
        return _uncombine_bs(T, lT, C, lC)
    return _contract_dense(T, lT, C, lC, combiner_first)

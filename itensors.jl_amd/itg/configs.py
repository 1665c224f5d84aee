"""Workload synthesis: the two-site effective-Hamiltonian network
L[a',w,a] psi[a,s1,s2,b] W1[w,s1,s1',w'] W2[w',s2,s2',w''] R[b',w'',b]
in dense and U(1) block-sparse form.

Conventions (chosen here; the reference fixes none of them):
  * bond sectors Sz = q in [-12, 12], sizes n_q = round(chi*e^{-q^2/32}/Z)
    for q != 0 and n_0 = chi - sum_{q!=0} n_q (so the sizes sum to chi);
  * physical sectors [Sz=+1 => 1, Sz=-1 => 1];
  * MPO link [Sz=0 => 1, Sz=+2 => 1, Sz=-2 => 1, Sz=0 => 2];
  * arrows: psi (a Out, s1 Out, s2 Out, b In); L (a' Out, w In, a In);
    W (w Out, s In, s' Out, w' In); R (b' In, w Out, b Out); every tensor
    has flux 0, i.e. q_b = q_a+s1+s2, q_a' = q_a+q_w, q_w' = q_w+s'-s,
    q_b' = q_b+q_w.
  * random entries: one numpy PCG64(seed) stream, tensors drawn in the order
    L, psi, W1, W2, R, each in storage order (blocks in eachblock order,
    column-major inside a block); complex data draws the real parts of a
    tensor, then its imaginary parts.
"""
from __future__ import annotations

import math

import numpy as np

from .itensors import QN, Index, In, Out, random_itensor

SZ = "Sz"


def sector_sizes(chi, qmax=12, width=32.0):
    qs = np.arange(-qmax, qmax + 1)
    w = np.exp(-qs.astype(np.float64) ** 2 / width)
    w /= w.sum()
    n = np.rint(chi * w).astype(np.int64)
    n[qmax] = chi - (n.sum() - n[qmax])
    return [(int(q), int(x)) for q, x in zip(qs, n)]


def bond_space(chi, qns=True):
    if not qns:
        return chi
    return [(QN((SZ, q)), n) for q, n in sector_sizes(chi) if n > 0]


def phys_space(qns=True):
    return [(QN((SZ, 1)), 1), (QN((SZ, -1)), 1)] if qns else 2


def link_space(qns=True):
    return [(QN((SZ, 0)), 1), (QN((SZ, 2)), 1), (QN((SZ, -2)), 1), (QN((SZ, 0)), 2)] if qns else 5


class HeffIndices:
    """The eleven indices of the network (QN or dense)."""

    def __init__(self, chi, qns=True):
        bd, ph, lk = bond_space(chi, qns), phys_space(qns), link_space(qns)
        d = (lambda x: x) if qns else (lambda x: None)
        self.a = Index(bd, "a", dir=d(Out))
        self.b = Index(bd, "b", dir=d(In))
        self.s1 = Index(ph, "s1", dir=d(Out))
        self.s2 = Index(ph, "s2", dir=d(Out))
        self.ap = Index(bd, "a'", dir=d(Out))
        self.bp = Index(bd, "b'", dir=d(In))
        self.w = Index(lk, "w", dir=d(In))        # L - W1
        self.wp = Index(lk, "w'", dir=d(In))      # W1 - W2
        self.wpp = Index(lk, "w''", dir=d(In))    # W2 - R
        self.s1p = Index(ph, "s1'", dir=d(Out))
        self.s2p = Index(ph, "s2'", dir=d(Out))
        self.qns = qns

    def L(self):
        return (self.ap, self.w, self.a.dag())

    def psi(self):
        return (self.a, self.s1, self.s2, self.b)

    def W1(self):
        return (self.w.dag(), self.s1.dag(), self.s1p, self.wp)

    def W2(self):
        return (self.wp.dag(), self.s2.dag(), self.s2p, self.wpp)

    def R(self):
        return (self.bp, self.wpp.dag(), self.b.dag())

    def out(self):
        """Index order of ((((L*psi)*W1)*W2)*R)."""
        return (self.ap, self.s1p, self.s2p, self.bp)


def heff_network(chi, qns=True, dtype=np.float64, seed=0, device="cuda"):
    """(L, psi, W1, W2, R, indices) with N(0,1) entries."""
    ix = HeffIndices(chi, qns)
    rng = np.random.Generator(np.random.PCG64(seed))
    kw = dict(dtype=dtype, rng=rng, device=device)
    if qns:
        z = QN((SZ, 0))
        ts = [random_itensor(z, *inds, **kw) for inds in (ix.L(), ix.psi(), ix.W1(), ix.W2(), ix.R())]
    else:
        ts = [random_itensor(*inds, **kw) for inds in (ix.L(), ix.psi(), ix.W1(), ix.W2(), ix.R())]
    return ts + [ix]


def heff_flops(L, psi, W1, W2, R):
    """Algorithmic flops of the four pairwise contractions (2*m*n*k summed
    over block pairs, real arithmetic; x4 for complex)."""
    from . import ndtensors as ND
    from .itensors import compute_contraction_labels

    total = []
    cur = L
    for B in (psi, W1, W2, R):
        lA, lB = compute_contraction_labels(cur.inds, B.inds)
        lR = ND.contract_labels(lA, lB)
        if cur.hasqns():
            p = ND.bs_plan(cur.tensor, lA, B.tensor, lB, lR)
            f = 0
            b1 = list(cur.tensor.blockoffsets)
            b2 = list(B.tensor.blockoffsets)
            for i, j, _r in p.pairs():
                d1 = cur.tensor.blockdims_of(b1[i])
                d2 = B.tensor.blockdims_of(b2[j])
                m = math.prod(d for d, l in zip(d1, lA) if l > 0)
                n = math.prod(d for d, l in zip(d2, lB) if l > 0)
                k = math.prod(d for d, l in zip(d1, lA) if l < 0)
                f += 2 * m * n * k
        else:
            d1, d2 = cur.tensor.dims, B.tensor.dims
            m = math.prod(d for d, l in zip(d1, lA) if l > 0)
            n = math.prod(d for d, l in zip(d2, lB) if l > 0)
            k = math.prod(d for d, l in zip(d1, lA) if l < 0)
            f = 2 * m * n * k
        total.append(f)
        # structure of the next left operand, without computing it
        cur = _shape_only_contract(cur, B, lA, lB, lR)
    return total


def _shape_only_contract(A, B, lA, lB, lR):
    """An ITensor with the block structure of A*B and no data."""
    import torch
    from . import ndtensors as ND
    from .itensors import ITensor

    indsR = tuple(A.inds[lA.index(x)] if x in lA else B.inds[lB.index(x)] for x in lR)
    if A.hasqns():
        p = ND.bs_plan(A.tensor, lA, B.tensor, lB, lR)
        bd = tuple(i.blockdims() for i in indsR)
        T = ND.BlockSparseTensor(torch.empty(0, dtype=A.eltype, device=A.tensor.data.device),
                                 dict(zip(p.blocksR, p.offsR)), bd, structure_only=True)
        return ITensor(T, indsR)
    dims = tuple(i.dim for i in indsR)
    return ITensor(_FakeDense(dims), indsR)


class _FakeDense:
    def __init__(self, dims):
        self.dims = dims
        self.ndims = len(dims)

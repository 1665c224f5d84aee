"""Reverse-mode rule of contraction on the device.

rrule(contract, x1, x2) follows ext/ITensorsChainRulesCoreExt/itensor.jl:94-103:
    y = x1 * x2
    pullback(ybar) = (NoTangent, project(x1)(ybar * dag(x2)),
                                 project(x2)(dag(x1) * ybar))
where dag conjugates the data and flips the arrows, and the projection keeps
the real part for real inputs (projection.jl:1-10). The two adjoint
contractions are ordinary device contractions; conj is applied inside the
GEMM / direct kernels (no conjugated copy of the data is made).
"""
from __future__ import annotations

from . import ndtensors as ND
from .itensors import ITensor, contract


class NoTangent:
    def __repr__(self):
        return "NoTangent()"


def _dag_view(x):
    """dag(x) without copying: flipped arrows, data conjugated lazily (the
    contraction takes a conj flag)."""
    return ITensor(x.tensor, tuple(i.dag() for i in x.inds))


def _project(x, dx):
    if not x.eltype.is_complex and dx.eltype.is_complex:
        return ITensor(ND._with_data(dx.tensor, dx.tensor.data.real.contiguous()), dx.inds)
    return dx


def permute_like(x, inds):
    """Reorder an ITensor's storage to the index order `inds` (the same index
    set), as ITensors.permute does."""
    if tuple(x.inds) == tuple(inds):
        return x
    perm = [x.inds.index(i) + 1 for i in inds]
    return ITensor(ND.permutedims(x.tensor, perm), inds)


def rrule_contract(x1, x2, *, prec=None):
    y = contract(x1, x2, prec=prec)
    cplx1 = x1.eltype.is_complex
    cplx2 = x2.eltype.is_complex

    def contract_pullback(ybar):
        x1bar = contract(ybar, _dag_view(x2), prec=prec, conjB=cplx2)
        x2bar = contract(_dag_view(x1), ybar, prec=prec, conjA=cplx1)
        return NoTangent(), _project(x1, x1bar), _project(x2, x2bar)

    return y, contract_pullback

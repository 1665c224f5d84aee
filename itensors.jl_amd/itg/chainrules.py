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


def replaceinds(x, inds):
    """The same storage under other indices (each new index must have the
    space of the one it replaces)."""
    inds = tuple(inds)
    for a, b in zip(x.inds, inds):
        if a.space != b.space:
            raise ValueError(f"index {b} does not have the space of {a}")
    return ITensor(x.tensor, inds)


_LAYOUT_MAPS = {}  # (src structure, dst structure, device) -> map, block-sparse only
_LAYOUT_MAPS_MAX = 8


def _layout_map(src, dst):
    """Element positions in dst's storage of every element of src's storage
    (same index set; src's blocks must be blocks of dst). Returns a device
    int64 tensor, or None when the layouts are identical. Block-sparse maps
    (one small index computation per block) are kept for the structures seen
    last: an eigensolver's gradients repeat them."""
    if tuple(src.inds) == tuple(dst.inds):
        if not src.hasqns() or list(src.tensor.blockoffsets.items()) == list(dst.tensor.blockoffsets.items()):
            return None
    if not src.hasqns():
        return _build_layout_map(src, dst)
    key = (tuple(src.inds), tuple(dst.inds), src.tensor.data.device,
           tuple(src.tensor.blockoffsets.items()), tuple(dst.tensor.blockoffsets.items()))
    m = _LAYOUT_MAPS.get(key)
    if m is None:
        m = _build_layout_map(src, dst)
        if len(_LAYOUT_MAPS) >= _LAYOUT_MAPS_MAX:
            _LAYOUT_MAPS.pop(next(iter(_LAYOUT_MAPS)))
        _LAYOUT_MAPS[key] = m
    return m


def _build_layout_map(src, dst):
    import math

    import torch

    perm = [src.inds.index(i) for i in dst.inds]  # dst mode k = src mode perm[k]
    chunks = []
    if not src.hasqns():
        dims = src.tensor.dims
        # dst position of each src element
        ddims = tuple(dims[p] for p in perm)
        dst_lin = torch.arange(math.prod(ddims), device=src.tensor.data.device).reshape(tuple(reversed(ddims)))
        # dst coordinate k = src coordinate perm[k]; permute dst_lin's axes to src order
        inv = [perm.index(m) for m in range(len(dims))]
        rev = len(dims) - 1
        src_view = dst_lin.permute(*[rev - inv[rev - a] for a in range(len(dims))])
        return src_view.reshape(-1).contiguous()
    dev = src.tensor.data.device
    for b, o in src.tensor.blockoffsets.items():
        db = tuple(b[p] for p in perm)
        if db not in dst.tensor.blockoffsets:
            raise ValueError("gradient block outside the tensor's block structure")
        od = dst.tensor.blockoffsets[db]
        dims = src.tensor.blockdims_of(b)
        n = math.prod(dims)
        ddims = tuple(dims[p] for p in perm)
        dst_lin = torch.arange(n, device=dev).reshape(tuple(reversed(ddims)))
        inv = [perm.index(m) for m in range(len(dims))]
        rev = len(dims) - 1
        src_view = dst_lin.permute(*[rev - inv[rev - a] for a in range(len(dims))])
        chunks.append(src_view.reshape(-1) + od)
    if not chunks:
        return torch.empty(0, dtype=torch.int64, device=dev)
    return torch.cat(chunks).contiguous()


def axpy_(y, a, x):
    """y += a*x for ITensors over the same index set (x may store its modes
    in another order and hold a subset of y's blocks). Returns y."""
    m = _layout_map(x, y)
    if m is None:
        y.tensor.data.add_(x.tensor.data, alpha=a)
    else:
        y.tensor.data.index_add_(0, m, x.tensor.data, alpha=a)
    return y


def inner(x, y):
    """<x|y> = sum conj(x)*y over matching layouts (returns a Python number)."""
    m = _layout_map(y, x)
    yd = y.tensor.data if m is None else y.tensor.data.new_zeros(x.tensor.data.numel()).index_add_(0, m, y.tensor.data)
    v = (x.tensor.data.conj() * yd).sum().item()
    return v


def heff_energy_and_grad(L, psi, W1, W2, R, out_inds, *, prec=None):
    """E(psi) = Re<psi|H psi>/<psi|psi> for H psi = ((((L*psi)*W1)*W2)*R) with
    H psi carrying `out_inds` (psi's spaces), and dE/dpsi by reverse mode
    through rrule_contract -- every backward step is a device contraction
    (the adjoints conjugate inside the kernels). For real psi the result is
    ((H + H^T) psi - 2 E psi) / <psi|psi>. Returns (E, grad ITensor in psi's
    layout)."""
    import torch

    y1, pb1 = rrule_contract(L, psi, prec=prec)
    y2, pb2 = rrule_contract(y1, W1, prec=prec)
    y3, pb3 = rrule_contract(y2, W2, prec=prec)
    y4, pb4 = rrule_contract(y3, R, prec=prec)
    bra = replaceinds(psi, out_inds)
    bra = ITensor(bra.tensor, tuple(i.dag() for i in bra.inds))  # dag, data conj in-kernel
    num_t, pbn = rrule_contract(bra, y4, prec=prec)
    num = complex(num_t.tensor.data.reshape(-1)[0].item()) if num_t.tensor.data.numel() else 0j
    den = float(torch.linalg.vector_norm(psi.tensor.data).item() ** 2)
    E = num.real / den
    # cotangent of num (a 0-index ITensor) for E = Re(num)/den
    seed = ITensor(ND._with_data(num_t.tensor, torch.full_like(num_t.tensor.data, 1.0 / den)), num_t.inds)
    _, bra_bar, y4_bar = pbn(seed)
    _, y3_bar, _ = pb4(y4_bar)
    _, y2_bar, _ = pb3(y3_bar)
    _, y1_bar, _ = pb2(y2_bar)
    _, _, psi_bar = pb1(y1_bar)
    grad = ITensor(ND._with_data(psi.tensor, psi.tensor.data.clone()), psi.inds)
    grad.tensor.data.mul_(-2.0 * E / den)
    axpy_(grad, 1.0, psi_bar)
    # the bra holds conj(psi): its cotangent maps back to psi conjugated
    bb = replaceinds(ITensor(bra_bar.tensor, tuple(i.dag() for i in bra_bar.inds)),
                     tuple(psi.inds[out_inds.index(i)] if i in out_inds else i
                           for i in (j.dag() for j in bra_bar.inds)))
    if bb.eltype.is_complex:
        bb = ITensor(ND._with_data(bb.tensor, bb.tensor.data.conj().resolve_conj()), bb.inds)
    axpy_(grad, 1.0, bb)
    return E, grad


def rrule_contract(x1, x2, *, prec=None):
    y = contract(x1, x2, prec=prec)
    cplx1 = x1.eltype.is_complex
    cplx2 = x2.eltype.is_complex

    def contract_pullback(ybar):
        x1bar = contract(ybar, _dag_view(x2), prec=prec, conjB=cplx2)
        x2bar = contract(_dag_view(x1), ybar, prec=prec, conjA=cplx1)
        return NoTangent(), _project(x1, x1bar), _project(x2, x2bar)

    return y, contract_pullback

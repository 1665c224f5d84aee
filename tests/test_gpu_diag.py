"""DiagBlockSparse x BlockSparse on the device (itg_bs_diag_plan_create /
itg_bs_diag_execute) vs the oracle: block lists and offsets exactly, values
to 1e-12 relative Frobenius (one product per term, so in practice exact).

Cases: delta replacing an index (either operand order, the result's index
order), U * diag(s), order-3 diagonal tensors with one or two contracted
modes, the fully contracted diagonal, complex and mixed real / complex, the
reference's test_diagblocksparse.jl:49-78 structures, dense Diag x Dense,
and the off-block-diagonal rejection of the C ABI.
"""
from collections import OrderedDict

import numpy as np
import pytest

from helpers import FP64_RTOL, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


def oinds(O, inds):
    out = []
    for i in inds:
        if not i.hasqns:
            out.append(O.Index(i.dim, i.tags, i.plev, i.dir, i.id))
            continue
        space = [(O.QN(*[(n, v, m) for n, (v, m) in q.vals.items()]), d) for q, d in i.space]
        out.append(O.Index(space, tags=i.tags, plev=i.plev, dir=i.dir, id=i.id))
    return tuple(out)


def to_oracle(O, T):
    offs = OrderedDict((tuple(c - 1 for c in b), o) for b, o in T.tensor.blockoffsets.items())
    if T.isdiag():
        D = T.tensor
        host = None if D.data is None else D.data.detach().cpu().numpy()
        return O.DiagBlockSparse(host, offs, oinds(O, T.inds), uniform=D.uniform)
    return O.BlockSparse(T.tensor.data.detach().cpu().numpy(), offs, oinds(O, T.inds))


def check(env, A, B):
    """A * B on the device vs the oracle (which takes the block-sparse
    operand first, as the reference's dispatch does)."""
    torch, itg, ND, O = env
    C = A * B
    lA, lB = itg.compute_contraction_labels(A.inds, B.inds)
    if A.isdiag():
        R, _ = O.diagblocksparse_contract(to_oracle(O, B), lB, to_oracle(O, A), lA)
    else:
        R, _ = O.diagblocksparse_contract(to_oracle(O, A), lA, to_oracle(O, B), lB)
    got = [tuple(c - 1 for c in b) for b in C.tensor.blockoffsets]
    assert got == list(R.offsets.keys())
    assert list(C.tensor.blockoffsets.values()) == list(R.offsets.values())
    assert relerr(C.tensor.data.cpu().numpy(), R.data) < FP64_RTOL
    # the index order the reference gives: the non-diagonal operand's first
    other, diag = (B, A) if A.isdiag() else (A, B)
    exp = [i for i in other.inds if i not in diag.inds] + [i for i in diag.inds if i not in other.inds]
    assert list(C.inds) == exp
    return C


def qidx(itg, pairs, tags, **kw):
    return itg.Index([(itg.QN(("Sz", q)), d) for q, d in pairs], tags, **kw)


SP = [(-2, 3), (-1, 5), (0, 7), (1, 5), (2, 3)]


@pytest.mark.parametrize("diag_first", [False, True])
@pytest.mark.parametrize("cplx", [False, True])
def test_delta_replaces_an_index(env, diag_first, cplx):
    torch, itg, ND, O = env
    i, j = qidx(itg, SP, "i"), qidx(itg, SP, "j")
    k = qidx(itg, [(-1, 2), (1, 3)], "k", dir=itg.In)
    A = itg.random_itensor(itg.QN(("Sz", 1)), i, j, k, rng=np.random.default_rng(0),
                           dtype=np.complex128 if cplx else np.float64)
    d = itg.delta(j.dag(), j.prime())
    C = check(env, d, A) if diag_first else check(env, A, d)
    assert np.array_equal(C.array(i, j.prime(), k), A.array(i, j, k))


@pytest.mark.parametrize("cplx_u,cplx_s", [(False, False), (True, False), (False, True)])
def test_columns_scaled(env, cplx_u, cplx_s):
    torch, itg, ND, O = env
    a, b = qidx(itg, SP, "a"), qidx(itg, SP, "b", dir=itg.In)
    U = itg.random_itensor(a, b, rng=np.random.default_rng(1),
                           dtype=np.complex128 if cplx_u else np.float64)
    rng = np.random.default_rng(2)
    s = rng.standard_normal(23) + (1j * rng.standard_normal(23) if cplx_s else 0)
    bp = qidx(itg, SP, "bp", dir=itg.In)
    S = itg.diag_itensor(s, b.dag(), bp)
    C = check(env, U, S)
    assert np.allclose(C.array(a, bp), U.array(a, b) * s[None, :], rtol=0, atol=1e-15)


@pytest.mark.parametrize("ncontract", [1, 2])
def test_order3_diagonal(env, ncontract):
    torch, itg, ND, O = env
    i, j = qidx(itg, SP, "i"), qidx(itg, SP, "j", dir=itg.In)
    A = itg.random_itensor(i, j, rng=np.random.default_rng(3))
    jp = qidx(itg, SP, "jp", dir=itg.In)
    if ncontract == 1:
        D = itg.diag_itensor(0.0, j.dag(), jp, jp.prime())
    else:
        D = itg.diag_itensor(0.0, i.dag(), j.dag(), jp.dag())
    assert D.tensor.nnzblocks() >= 1
    D.tensor.data.copy_(torch.from_numpy(
        np.random.default_rng(4).standard_normal(D.tensor.data.numel())))
    check(env, A, D)
    check(env, D, A)


def test_fully_contracted_diagonal(env):
    torch, itg, ND, O = env
    i, j = qidx(itg, SP, "i"), qidx(itg, SP, "j", dir=itg.In)
    A = itg.random_itensor(i, j, rng=np.random.default_rng(5))
    d = np.random.default_rng(6).standard_normal(23)
    D = itg.diag_itensor(d, i.dag(), j.dag())
    C = check(env, A, D)
    Ad = A.array(i, j)
    assert np.isclose(C.tensor.data.cpu().numpy()[0], sum(d[x] * Ad[x, x] for x in range(23)),
                      rtol=1e-13, atol=0)


@pytest.mark.parametrize("dims", [([2, 2], [2, 2]), ([3, 2, 3], [2, 2])])
def test_reference_structures(env, dims):
    """test_diagblocksparse.jl:49-78: blocks (1,1), (2,2), uniform one."""
    torch, itg, ND, O = env
    QN = itg.QN
    i = itg.Index([(QN(("A", k)), x) for k, x in enumerate(dims[0])], "i")
    j = itg.Index([(QN(("A", k)), x) for k, x in enumerate(dims[1])], "j", dir=itg.In)
    A = itg.random_itensor(i, j, rng=np.random.default_rng(7))
    for D in (itg.delta(j.dag(), j.prime()), itg.delta(i.dag(), j.dag())):
        check(env, A, D)


def test_dense_diag_times_dense(env):
    torch, itg, ND, O = env
    i, j, k = itg.Index(5, "i"), itg.Index(4, "j"), itg.Index(3, "k")
    A = itg.random_itensor(i, j, k, rng=np.random.default_rng(8))
    s = np.random.default_rng(9).standard_normal(4)
    S = itg.diag_itensor(s, j, j.prime())
    C = A * S
    assert C.inds == (i, k, j.prime())
    ref = np.einsum("ijk,j->ikj", A.array(i, j, k), s)
    assert np.allclose(C.array(i, k, j.prime()), ref, rtol=0, atol=1e-15)
    D = itg.delta(k, k.prime())
    assert np.array_equal((D * A).array(i, j, k.prime()), A.array(i, j, k))


def test_off_block_diagonal_rejected_by_the_abi(env):
    """The C ABI refuses a diagonal operand with a block off the block
    diagonal (the reference raises, test_diagblocksparse.jl:31-46)."""
    torch, itg, ND, O = env
    i = qidx(itg, [(0, 1), (1, 1)], "i")
    j = qidx(itg, [(0, 1), (1, 1)], "j", dir=itg.In)
    A = itg.random_itensor(i, j, rng=np.random.default_rng(10))

    class Fake:  # bypasses the host-side check of DiagBlockSparseTensor
        blockoffsets = OrderedDict([((1, 2), 0), ((2, 1), 1)])
        blockdims = ((1, 1), (1, 1))
        ndims = 2
        data = torch.ones(2, dtype=torch.float64, device="cuda")

        def nnzblocks(self):
            return 2

    with pytest.raises(itg.ItgError):
        ND.BSPlan(A.tensor, (1, -1), Fake(), (-1, 2), (1, 2), diag=True)
    with pytest.raises(ValueError):
        ND.DiagBlockSparseTensor(None, Fake.blockoffsets, Fake.blockdims, uniform=1.0)


def test_conjugation_flags_reach_the_diagonal_path(env):
    """contract(...; conjA / conjB) with a diagonal operand (the adjoint of
    a contraction conjugates one side) equals contracting conj'd copies."""
    torch, itg, ND, O = env
    a, b = qidx(itg, SP, "a"), qidx(itg, SP, "b", dir=itg.In)
    U = itg.random_itensor(a, b, rng=np.random.default_rng(11), dtype=np.complex128)
    rng = np.random.default_rng(12)
    s = rng.standard_normal(23) + 1j * rng.standard_normal(23)
    bp = qidx(itg, SP, "bp", dir=itg.In)
    S = itg.diag_itensor(s, b.dag(), bp)
    ref = U.array(a, b).conj() * s.conj()[None, :]
    C = itg.contract(U, S, conjA=True, conjB=True)
    assert np.allclose(C.array(a, bp), ref, rtol=0, atol=1e-15)
    C2 = itg.contract(S, U, conjA=True)
    assert np.allclose(C2.array(a, bp), U.array(a, b) * s.conj()[None, :], rtol=0, atol=1e-15)


@pytest.mark.parametrize("seed", range(6))
def test_random_structures(env, seed):
    """Random U(1) sector layouts and arrows; a stored diagonal over one of
    the block-sparse tensor's indices and an open one, either operand
    order, real or complex: device vs the oracle."""
    torch, itg, ND, O = env
    rng = np.random.default_rng(100 + seed)

    def space():
        qs = sorted(rng.choice(np.arange(-3, 4), size=int(rng.integers(2, 5)), replace=False))
        return [(int(q), int(rng.integers(1, 6))) for q in qs]

    sa, sb = space(), space()
    a = qidx(itg, sa, "a")
    b = qidx(itg, sb, "b", dir=itg.In)
    c = qidx(itg, space(), "c", dir=itg.Out if seed % 2 else itg.In)
    A = itg.random_itensor(a, b, c, rng=rng, dtype=np.complex128 if seed % 3 == 0 else np.float64)
    if A.nnzblocks() == 0:
        pytest.skip("no allowed blocks")
    if seed % 2:
        D = itg.diag_itensor(0.0, b.dag(), qidx(itg, sb, "bp", dir=itg.In))
    else:
        D = itg.diag_itensor(0.0, a.dag(), qidx(itg, sa, "x", dir=itg.Out))
    if D.tensor.nnzblocks() == 0:
        pytest.skip("no diagonal blocks with zero flux")
    D.tensor.data.copy_(torch.from_numpy(rng.standard_normal(D.tensor.data.numel())))
    check(env, A, D)
    check(env, D, A)


@pytest.mark.parametrize("cplx_x", [False, True])
def test_scaled_diagonal_keeps_its_storage(env, cplx_x):
    """x * diag_itensor and x * delta stay diagonal (scaled values, or a
    scaled uniform value) and contract like the diagonal they are; the plan
    checks every operand's data length against its blocks first."""
    torch, itg, ND, O = env
    x = (1.5 - 0.5j) if cplx_x else -2.5
    a, b = qidx(itg, SP, "a"), qidx(itg, SP, "b", dir=itg.In)
    U = itg.random_itensor(a, b, rng=np.random.default_rng(8))
    s = np.random.default_rng(9).standard_normal(23)
    bp = qidx(itg, SP, "bp", dir=itg.In)
    S = itg.diag_itensor(s, b.dag(), bp)
    for xS in (x * S, S * x):
        assert xS.isdiag() and xS.tensor.uniform is None
        assert xS.tensor.data.numel() == S.tensor.data.numel()
        C = check(env, U, xS)
        assert np.allclose(C.array(a, bp), x * U.array(a, b) * s[None, :], rtol=0, atol=1e-14)
    d = itg.delta(b.dag(), bp)
    for xd in (x * d, d * x):
        assert xd.isdiag() and xd.tensor.uniform == (x if cplx_x else x)
        C = check(env, U, xd)
        assert np.allclose(C.array(a, bp), x * U.array(a, b), rtol=0, atol=1e-14)
    # a BlockSparseTensor whose data is shorter than its blocks is refused
    short = U.tensor.data[:-1]
    with pytest.raises(ValueError):
        ND.BlockSparseTensor(short, U.tensor.blockoffsets, U.tensor.blockdims)

"""The threaded block-sparse switch on the device.

With NDTensors.enable_threaded_blocksparse() in a Julia session of more than
one thread, contract_blockoffsets enumerates the block pairs with
contract_blocks(::Algorithm"threaded_threads") (NDTensors/src/blocksparse/
contract.jl:42-55, contract_threaded.jl:2-73): T2's blocks outer when T1 has
no more blocks than T2. That order fixes the output's block order and
offsets (first appearance) and the order of the sums into each output block
(contract_generic.jl:37-129). The device plans follow it while
itg.enable_threaded_blocksparse() is in effect: block lists and offsets
bit-exact against the oracle's threaded plan, values to 1e-12, in FP64 and
split precision; per-pair factors are taken in that enumeration order.
Known-answer cases: test/threading/test_threading.jl:16-78.
"""
import numpy as np
import pytest

from helpers import FP64_RTOL, relerr
from test_gpu_blocksparse import to_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


@pytest.fixture
def threaded(env):
    _, itg, _, _ = env
    was = itg.enable_threaded_blocksparse()
    yield
    itg.enable_threaded_blocksparse(was)


def _oracle_product(O, A, B, threaded_order, pair_alpha=None):
    OA, OB = to_oracle(O, A), to_oracle(O, B)
    lA, lB = O.compute_contraction_labels(OA.inds, OB.inds)
    lR = O.contract_labels(lA, lB)
    return O.blocksparse_contract(OA, lA, OB, lB, lR, threaded=threaded_order,
                                  pair_alpha=pair_alpha)


def _check(O, C, OC):
    got = [tuple(c - 1 for c in b) for b in C.tensor.blockoffsets]
    assert got == list(OC.offsets.keys())
    assert list(C.tensor.blockoffsets.values()) == list(OC.offsets.values())
    want = OC.data.astype(np.complex128 if np.iscomplexobj(OC.data) else np.float64)
    err = relerr(C.tensor.data.cpu().numpy(), want)
    assert err < FP64_RTOL, err


def test_switch_state(env):
    """test_threading.jl:16-28: the switch returns the previous state, with
    and without a Bool argument."""
    _, itg, _, _ = env
    e0 = itg.enable_threaded_blocksparse(False)
    assert not itg.using_threaded_blocksparse()
    assert not itg.enable_threaded_blocksparse(True)
    assert itg.using_threaded_blocksparse()
    assert itg.enable_threaded_blocksparse(False)
    assert not itg.using_threaded_blocksparse()
    assert not itg.enable_threaded_blocksparse()
    assert itg.disable_threaded_blocksparse()
    itg.enable_threaded_blocksparse(e0)
    assert itg.using_threaded_blocksparse() == e0
    from itg import _lib
    h = _lib.handle_for(0)
    assert _lib.lib().itg_set_blocksparse_order(h.h, 7) == _lib.ITG_ERR_INVALID


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_random_structures_threaded_order(env, threaded, cplx):
    torch, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(31)
    dt = np.complex128 if cplx else np.float64
    changed = 0
    for trial in range(6):
        def idx(tag, lo, hi, big):
            n = int(rng.integers(lo, hi))
            return Index([(QN(int(q)), int(rng.integers(1, 40 if big else 6)))
                          for q in range(-(n // 2), n - n // 2)], tag)
        a, b, c, d = idx("a", 2, 4, True), idx("b", 2, 5, True), idx("c", 2, 4, False), \
            idx("d", 3, 6, True)
        A = itg.random_itensor(QN(0), a, b.dag(), c, dtype=dt, rng=rng)
        B = itg.random_itensor(QN(1), b, c.dag(), d, dtype=dt, rng=rng)
        for X, Y in ((A, B), (B, A)):
            OC, plan = _oracle_product(O, X, Y, True)
            _, plan_seq = _oracle_product(O, X, Y, False)
            changed += plan != plan_seq
            for prec in ("fp64", "split_int8"):
                itg.set_precision(prec)
                try:
                    C = X * Y
                finally:
                    itg.set_precision("fp64")
                _check(O, C, OC)
    assert changed > 0  # the T2-outer order differed from the sequential one


def test_plan_order_and_pair_factors(env, threaded):
    """contract_blockoffsets reports the threaded plan; per-pair factors
    (compute_alpha's slot) are applied in that order."""
    torch, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(7)
    a = Index([(QN(0), 5), (QN(1), 7)], "a")
    b = Index([(QN(-1), 3), (QN(0), 6), (QN(1), 4), (QN(2), 2)], "b")
    c = Index([(QN(-1), 4), (QN(0), 5), (QN(1), 3), (QN(2), 6)], "c")
    A = itg.random_itensor(QN(0), a, b.dag(), rng=rng)
    B = itg.random_itensor(QN(0), b, c.dag(), rng=rng)
    lA, lB = itg.compute_contraction_labels(A.inds, B.inds)
    lR = ND.contract_labels(lA, lB)
    offs, nnz, plan = ND.contract_blockoffsets(A.tensor, lA, B.tensor, lB, lR)
    OC, oplan = _oracle_product(O, A, B, True)
    assert [tuple(tuple(x - 1 for x in blk) for blk in pr) for pr in plan] == oplan
    assert len(A.tensor.blockoffsets) <= len(B.tensor.blockoffsets)
    fac = rng.standard_normal(len(plan))
    p = ND.bs_plan(A.tensor, lA, B.tensor, lB, lR)
    OCa, _ = _oracle_product(O, A, B, True, pair_alpha=list(fac))
    for prec in ("fp64", "split_int8"):
        out = torch.full((p.nnzR,), float("nan"), dtype=torch.float64, device="cuda")
        p.execute(out, A.tensor.data, B.tensor.data, prec=prec, pair_alpha=fac)
        torch.cuda.synchronize()
        assert relerr(out.cpu().numpy(), OCa.data) < FP64_RTOL, prec


def test_kat_threaded_equals_sequential(env):
    """test_threading.jl:30-58: A' * A with threaded block sparse off and on
    gives the same tensor (both operands hold two blocks, so the threaded
    plan runs T2's blocks outer; the output blocks come out in the same
    order here)."""
    torch, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(5)
    i = Index([(QN(0), 500), (QN(1), 500)], "i")
    A = itg.random_itensor(QN(0), i.prime(), i.dag(), rng=rng)
    Ap = itg.ITensor(A.tensor, tuple(x.prime() for x in A.inds))
    was = itg.enable_threaded_blocksparse(False)
    try:
        R = Ap * A
        itg.enable_threaded_blocksparse(True)
        Rt = Ap * A
        OC, _ = _oracle_product(O, Ap, A, True)
        _check(O, Rt, OC)
    finally:
        itg.enable_threaded_blocksparse(was)
    dense = lambda T: to_oracle(O, T).dense()  # noqa: E731
    assert np.allclose(dense(R), dense(Rt), rtol=0, atol=1e-12 * np.abs(dense(R)).max())


def test_kat_no_blocks_threaded(env, threaded):
    """test_threading.jl:60-78: no matching pair gives no blocks."""
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i = Index([(QN(0), 1), (QN(1), 1)], "i")
    A = itg.blocksparse_itensor(np.array([11.0]), [(1, 1)], (i.prime(2), i.dag().prime()))
    B = itg.blocksparse_itensor(np.array([22.0]), [(2, 2)], (i.prime(), i.dag()))
    C = A * B
    assert C.nnzblocks() == 0 and C.tensor.data.numel() == 0


def test_network_threaded_fused_and_oracle(env, threaded):
    """The two-site network under the threaded order: the fused MPO pair is
    bit-identical to the two separate steps, and the result (block order,
    offsets, values) matches the oracle's threaded fold."""
    torch, itg, ND, O = env
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    net = heff_network(256, qns=True, seed=9, device="cuda")[:5]
    fused = ContractionSequence(net)
    plain = ContractionSequence(net, fuse=False)
    assert list(fused.chains) == [1]
    a = fused.run("fp64").tensor.data.clone()
    b = plain.run("fp64").tensor.data.clone()
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    cur = to_oracle(O, net[0])
    for T in net[1:]:
        OT = to_oracle(O, T)
        lA, lB = O.compute_contraction_labels(cur.inds, OT.inds)
        cur, _ = O.blocksparse_contract(cur, lA, OT, lB, O.contract_labels(lA, lB), threaded=True)
    res = fused.result
    # the result's index order is the fold's; compare in the oracle's order
    assert tuple(i.id for i in res.inds) == tuple(i.id for i in cur.inds)
    _check(O, res, cur)

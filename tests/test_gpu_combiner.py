"""Combiner contraction on the device vs the oracle (pure data movement:
block lists, offsets and values must match bit for bit).

Cases follow the reference's combiner tests: test/base/test_qncombiner.jl:
4-13 (merging equal-QN sectors of one index, round trip),
test/base/test_qnitensor.jl:586-870 (empty combiner, the arrow keyword,
every subset and order of the indices of order-3 / order-4 tensors, the
two-QN-name examples), test/base/test_combiner.jl:5-264 (dense: combine
and uncombine in both operand orders, index replacement, a contraction
between combining and uncombining).
"""
import itertools

import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


def oracle_inds(O, inds):
    out = []
    for i in inds:
        space = [(O.QN(*[(n, v, m) for n, (v, m) in q.vals.items()]), d) for q, d in i.space]
        out.append(O.Index(space, tags=i.tags, plev=i.plev, dir=i.dir, id=i.id))
    return tuple(out)


def to_oracle(O, T):
    from collections import OrderedDict
    host = T.tensor.data.detach().cpu().numpy()
    offs = OrderedDict((tuple(c - 1 for c in b), o) for b, o in T.tensor.blockoffsets.items())
    return O.BlockSparse(host, offs, oracle_inds(O, T.inds))


def same_storage(T, OR):
    got = [tuple(c - 1 for c in b) for b in T.tensor.blockoffsets]
    assert got == list(OR.offsets.keys())
    assert list(T.tensor.blockoffsets.values()) == list(OR.offsets.values())
    assert np.array_equal(T.tensor.data.detach().cpu().numpy(), OR.data)


def combine_and_check(env, A, picks, combiner_first=False, **kw):
    """A * combiner(A.inds[picks]...) on the device vs the oracle; then the
    round trip through dag(C), also vs the oracle."""
    _, itg, ND, O = env
    C = itg.combiner(*[A.inds[p] for p in picks], **kw)
    AC = C * A if combiner_first else A * C
    c = itg.combinedind(C)
    assert AC.inds[0] == c
    assert AC.nnz() == A.nnz()
    lA, lC = itg.compute_contraction_labels(A.inds, C.inds)
    OA = to_oracle(O, A)
    OR, lR, oc = O.combine_blocksparse(OA, lA, [OA.inds[p] for p in picks], lC,
                                       dir=kw.get("dir"))
    assert [d for _q, d in c.space] == [d for _q, d in oc.space]
    same_storage(AC, OR)
    # uncombine
    back = AC * itg.dag(C)
    assert back.inds == tuple(A.inds[p] for p in picks) + tuple(
        i for k, i in enumerate(A.inds) if k not in picks)
    lAC, lD = itg.compute_contraction_labels(AC.inds, itg.dag(C).inds)
    _s, order, merge, _p, _d = O.combined_index([OA.inds[p] for p in picks], dir=kw.get("dir"))
    OB, _lB = O.uncombine_blocksparse(to_oracle(O, AC), lAC, oc,
                                      [OA.inds[p] for p in picks], lD, order, merge)
    same_storage(back, OB)
    perm = [back.inds.index(i) for i in A.inds]
    assert np.array_equal(np.transpose(back.array(), perm), A.array())
    return AC, C


def test_single_index_merge_round_trip(env):
    """test_qncombiner.jl:4-13."""
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i = Index([(QN(0), 1), (QN(0), 1)], "i")
    A = itg.random_itensor(i, rng=np.random.default_rng(0))
    AC, C = combine_and_check(env, A, [0])
    assert itg.combinedind(C).space == ((QN(0), 2),)


@pytest.mark.parametrize("order4", [False, True])
@pytest.mark.parametrize("cplx", [False, True])
def test_every_subset_and_order(env, order4, cplx):
    """test_qnitensor.jl:666-784 (order 3 and order 4, every subset)."""
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i1 = Index([(QN(("", 0, 2)), 2), (QN(("", 1, 2)), 2)], "i1")
    i2 = Index([(QN(("", 0, 2)), 2), (QN(("", 1, 2)), 2)], "i2")
    inds = [i1, i2, i1.prime().dag()] + ([i2.prime().dag()] if order4 else [])
    A = itg.random_itensor(QN(), *inds, dtype=complex if cplx else float,
                           rng=np.random.default_rng(4))
    for r in range(1, len(inds) + 1):
        for picks in itertools.permutations(range(len(inds)), r):
            combine_and_check(env, A, list(picks), tags="c")


def test_mixed_arrows_uneven_sectors_both_operand_orders(env):
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i = Index([(QN(("Sz", 0)), 2), (QN(("Sz", 1)), 3)], "i")
    j = Index([(QN(("Sz", 0)), 1), (QN(("Sz", -1)), 2), (QN(("Sz", 1)), 1)], "j", dir=itg.In)
    k = Index([(QN(("Sz", 0)), 3), (QN(("Sz", 2)), 2), (QN(("Sz", -1)), 1)], "k")
    A = itg.random_itensor(QN(("Sz", 1)), i, j, k.dag(), rng=np.random.default_rng(9))
    for picks in ([0, 1], [1, 0], [2, 0], [0, 1, 2], [2, 1, 0]):
        X, _ = combine_and_check(env, A, picks)
        Y, _ = combine_and_check(env, A, picks, combiner_first=True)
        assert np.array_equal(X.tensor.data.cpu().numpy(), Y.tensor.data.cpu().numpy())


def test_direction_keyword(env):
    """test_qnitensor.jl:603-625."""
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i1 = Index([(QN(0), 2), (QN(1), 3)], "i1")
    A = itg.random_itensor(i1.prime(), i1.dag(), rng=np.random.default_rng(3))
    AC, C = combine_and_check(env, A, [1], dir=itg.Out)
    assert itg.combinedind(C).dir == itg.Out
    AC.checkflux()


def test_two_name_examples(env):
    """test_qnitensor.jl:786-870 (Examples 1 and 2)."""
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    s1 = Index([(QN(("Sz", 0), ("Nf", 0)), 1), (QN(("Sz", 1), ("Nf", 1)), 1),
                (QN(("Sz", -1), ("Nf", 1)), 1), (QN(("Sz", 0), ("Nf", 2)), 1)], "s1")
    s2 = Index(s1.space, "s2")
    A = itg.random_itensor(QN(), s1, s2, s1.dag().prime(), s2.dag().prime(),
                           rng=np.random.default_rng(11))
    AC, _ = combine_and_check(env, A, [2, 3])
    assert AC.nnzblocks() < A.nnz()
    t1 = Index([(QN(("Nf", 0)), 1), (QN(("Nf", 1)), 1)], "t1")
    t2 = Index(t1.space, "t2")
    B = itg.random_itensor(QN(), t2.dag().prime(), t2, t1.dag().prime(), t1,
                           rng=np.random.default_rng(12))
    combine_and_check(env, B, [0, 2])


def test_empty_combiner_is_a_copy(env):
    """test_qnitensor.jl:586-601."""
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i1 = Index([(QN(("", 0, 2)), 2), (QN(("", 1, 2)), 2)], "i1")
    A = itg.random_itensor(QN(), i1, i1.dag().prime(), rng=np.random.default_rng(2))
    C = itg.combiner()
    assert itg.combinedind(C) is None
    AC = A * C
    assert AC.inds == A.inds and AC.nnzblocks() == A.nnzblocks()
    assert np.array_equal(AC.tensor.data.cpu().numpy(), A.tensor.data.cpu().numpy())
    assert AC.tensor.data.data_ptr() != A.tensor.data.data_ptr()


def test_invalid_combiner_contraction_raises(env):
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i = Index([(QN(0), 2), (QN(1), 2)], "i")
    j = Index([(QN(0), 1), (QN(1), 1)], "j")
    A = itg.random_itensor(QN(), i, j.dag(), rng=np.random.default_rng(1))
    C = itg.combiner(i, Index([(QN(0), 1)], "x"))
    with pytest.raises(ValueError):
        A * C  # one combined index is not in A: two uncontracted combiner indices


# ------------------------------------------------------------------- dense
def test_dense_combine_uncombine_both_orders(env):
    """test_combiner.jl:5-213: combine any subset in both operand orders;
    A * C holds A's other indices then the combined one, C * A the combined
    one first; the combined index is the column-major fusion of the
    combined indices in combiner order; uncombining restores A."""
    _, itg, ND, O = env
    Index = itg.Index
    i, j, k, l = Index(2, "i"), Index(3, "j"), Index(4, "k"), Index(5, "l")
    A = itg.random_itensor(i, j, k, l, rng=np.random.default_rng(6))
    a = A.array()
    for picks in ([0, 1], [2, 0], [3, 1, 0], [1, 3]):
        C = itg.combiner(*[A.inds[p] for p in picks])
        c = itg.combinedind(C)
        rest = [p for p in range(4) if p not in picks]
        for first in (False, True):
            AC = C * A if first else A * C
            order = ([c] + [A.inds[p] for p in rest]) if first else ([A.inds[p] for p in rest] + [c])
            assert AC.inds == tuple(order)
            if first:
                exp = np.transpose(a, picks + rest).reshape((-1,) + tuple(a.shape[p] for p in rest),
                                                            order="F")
            else:
                exp = np.transpose(a, rest + picks).reshape(tuple(a.shape[p] for p in rest) + (-1,),
                                                            order="F")
            assert np.array_equal(AC.array(), exp)
            back = AC * itg.dag(C)
            assert set(back.inds) == set(A.inds)
            assert np.array_equal(back.array(*A.inds), a)


def test_dense_index_replacement_and_contraction_between(env):
    """test_combiner.jl:214-264: a one-index combiner replaces the index;
    (A*C)*B uncombined equals A*B."""
    _, itg, ND, O = env
    Index = itg.Index
    i, j, k, l = Index(2, "i"), Index(3, "j"), Index(4, "k"), Index(5, "l")
    A = itg.random_itensor(i, j, k, l, rng=np.random.default_rng(8))
    C = itg.combiner(l, tags="nl")
    c = itg.combinedind(C)
    B = A * C
    assert B.inds == (i, j, k, c)
    assert np.array_equal(B.array(), A.array())
    C3 = itg.combiner(i, j, k)
    Ac = A * C3
    Bl = itg.random_itensor(l, rng=np.random.default_rng(10))
    AB = Ac * Bl
    ref = A * Bl
    assert relerr((AB * C3).array(*ref.inds), ref.array()) < 1e-14


def test_block_copy_rejects_out_of_range_descriptors(env):
    """itg_bs_copy checks every access against the array lengths on the
    host and launches nothing when one falls outside."""
    torch, itg, ND, O = env
    from itg import combiners as CB
    src = torch.arange(12, dtype=torch.float64, device="cuda")
    dst = torch.zeros(12, dtype=torch.float64, device="cuda")
    CB.bs_copy(dst, src, [((3, 4), 0, (1, 3), 0, (4, 1))])  # a 3x4 transpose
    assert np.array_equal(dst.cpu().numpy().reshape(3, 4), np.arange(12.0).reshape(4, 3).T)
    for bad in ([((3, 5), 0, (1, 3), 0, (5, 1))],   # source past its end
                [((3, 4), 0, (1, 3), 1, (4, 1))],   # destination past its end
                [((2, 2), 0, (1, 2), -1, (1, 2))]):  # negative destination
        with pytest.raises(itg.ItgError):
            CB.bs_copy(dst, src, bad)


def test_conjugation_flag_with_a_combiner(env):
    """contract(A, C; conjA=True) with a combiner C conjugates A's values
    (the adjoint of a combiner contraction does this)."""
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i = Index([(QN(("Sz", 0)), 2), (QN(("Sz", 1)), 3)], "i")
    j = Index([(QN(("Sz", 0)), 1), (QN(("Sz", -1)), 2)], "j", dir=itg.In)
    A = itg.random_itensor(QN(), i, j, rng=np.random.default_rng(13), dtype=complex)
    C = itg.combiner(i, j)
    X = itg.contract(A, C, conjA=True)
    Y = A * C
    assert np.array_equal(X.tensor.data.cpu().numpy(), np.conj(Y.tensor.data.cpu().numpy()))

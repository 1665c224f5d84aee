"""CPU checks of the oracle's combiner restatement (no GPU).

Pins: a hand-derived known answer for the fused / sorted / merged index and
the element position map; the reference's own combiner tests
(test/base/test_qncombiner.jl:4-13, test/base/test_qnitensor.jl:586-870):
single-index merge, the arrow keyword, combine-then-uncombine identity over
every index subset and order of order-3 and order-4 tensors, nnz
preservation, per-block flux, fewer blocks than elements for the
two-QN-name example.
"""
import itertools
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import ndtensors_oracle as O  # noqa: E402


def _qn(q):
    """0 -> QN(0); ("Sz", 1) or ("", 0, 2) -> one named value; a tuple of
    such tuples -> several named values."""
    if isinstance(q, tuple) and q and isinstance(q[0], tuple):
        return O.QN(*q)
    return O.QN(q)


def qidx(pairs, tags="", dir=1, plev=0, id=None):
    return O.Index([(_qn(q), d) for q, d in pairs], tags, plev, dir, id)


def test_combined_index_single_index_merge():
    """test_qncombiner.jl:4-13: [QN(0)=>1, QN(0)=>1] fuses to one sector of 2."""
    i = qidx([(0, 1), (0, 1)])
    sectors, order, merge, _prod, d = O.combined_index([i])
    assert len(sectors) == 1 and sectors[0][1] == 2 and O.qn_eq(sectors[0][0], O.QN(0))
    assert order == [0, 1] and merge == [0, 0] and d == 1


def test_combined_index_hand_derived():
    """i = [Sz0:2, Sz1:3] (Out), j = [Sz0:1, Sz-1:2, Sz1:1] (In): mixed arrows
    give an Out index; product sectors (i fastest) carry Sz = q_i - q_j:
    0/2, 1/3, 1/4, 2/6, -1/2, 0/3; sorted (stable) -1, 0, 0, 1, 1, 2;
    merged -1:2, 0:5, 1:7, 2:6."""
    i = qidx([(("Sz", 0), 2), (("Sz", 1), 3)], "i", 1)
    j = qidx([(("Sz", 0), 1), (("Sz", -1), 2), (("Sz", 1), 1)], "j", -1)
    sectors, order, merge, _prod, d = O.combined_index([i, j])
    assert d == 1
    assert order == [4, 0, 5, 1, 2, 3]
    assert merge == [0, 1, 1, 2, 2, 3]
    assert [s[1] for s in sectors] == [2, 5, 7, 6]
    assert [s[0].get("Sz", (0, 1))[0] for s in sectors] == [-1, 0, 1, 2]
    pos = O.combined_position_map([i, j], order)
    # sorted sector starts 0, 2, 4, 7, 10, 14
    assert pos[(0, 0)] == 2      # product sector (0,0), sorted second
    assert pos[(2, 0)] == 7      # (1,0): fourth, local 0
    assert pos[(2, 2)] == 17     # (1,1): sixth, local (0,1) -> 0 + 3*1
    assert sorted(pos.values()) == list(range(i.dim * j.dim))


def test_combined_index_direction_keyword():
    """test_qnitensor.jl:603-625: combiner(dag(i1); dir=Out) gives an Out index."""
    i1 = qidx([(0, 2), (1, 3)], "i1", 1)
    _s, _o, _m, _p, d = O.combined_index([i1.dag()], dir=1)
    assert d == 1
    _s, _o, _m, _p, d = O.combined_index([i1.dag()])
    assert d == -1


def _round_trip(A, lA, picks):
    cinds = [A.inds[p].dag() for p in picks]  # the combiner holds dag.(is)
    lC = (100,) + tuple(lA[p] for p in picks)
    R, lR, c = O.combine_blocksparse(A, lA, [A.inds[p] for p in picks], lC)
    assert sum(np.prod(O.blockdims(R.inds, b)) for b in R.offsets) == len(R.data)
    assert len(R.data) == len(A.data)
    _s, order, merge, _p, _d = O.combined_index([A.inds[p] for p in picks])
    back, lB = O.uncombine_blocksparse(R, lR, c, [A.inds[p] for p in picks],
                                       (100,) + tuple(lA[p] for p in picks), order, merge)
    perm = [lB.index(x) for x in lA]
    assert np.array_equal(np.transpose(back.dense(), perm), A.dense())
    assert len(back.offsets) == len(A.offsets)
    return R, c, cinds


@pytest.mark.parametrize("order4", [False, True])
def test_combine_uncombine_identity_all_subsets(order4):
    """test_qnitensor.jl:666-784: every subset and order of the indices."""
    rng = np.random.default_rng(5)
    i1 = qidx([(("", 0, 2), 2), (("", 1, 2), 2)], "i1", 1)
    i2 = qidx([(("", 0, 2), 2), (("", 1, 2), 2)], "i2", 1)
    inds = [i1, i2, i1.prime().dag()] + ([i2.prime().dag()] if order4 else [])
    A = O.random_blocksparse(rng, O.QN(), inds)
    lA = tuple(range(1, len(inds) + 1))
    for r in range(1, len(inds) + 1):
        for picks in itertools.permutations(range(len(inds)), r):
            R, c, _ = _round_trip(A, lA, list(picks))
            flux0 = O.flux_of_block(A.inds, next(iter(A.offsets)))
            for b in R.offsets:
                assert O.qn_eq(O.flux_of_block(R.inds, b), O.qn_scale(1, flux0))


def test_two_name_example_has_fewer_blocks_than_elements():
    """test_qnitensor.jl:786-828 ("Order 4, Combine 2, Example 1")."""
    rng = np.random.default_rng(7)
    sp = [((("Sz", 0), ("Nf", 0)), 1), ((("Sz", 1), ("Nf", 1)), 1),
          ((("Sz", -1), ("Nf", 1)), 1), ((("Sz", 0), ("Nf", 2)), 1)]
    s1 = qidx(sp, "s1", 1)
    s2 = qidx(sp, "s2", 1)
    inds = [s1, s2, s1.prime().dag(), s2.prime().dag()]
    A = O.random_blocksparse(rng, O.QN(), inds)
    R, c, _ = _round_trip(A, (1, 2, 3, 4), [2, 3])
    assert len(R.offsets) < len(A.data)
    for b in R.offsets:
        assert O.qn_eq(O.flux_of_block(R.inds, b), {})

"""Contraction sequences on the host (no GPU): the cost model of
contraction_cost (src/tensor_operations/contraction_cost.jl:4-42), the
left-associative sequence, and optimal_contraction_sequence against brute
force over every binary tree (the reference delegates the search to
TensorOperations.optimaltree, ext/ITensorsTensorOperationsExt)."""
import itertools
from types import SimpleNamespace

import numpy as np

import helpers  # noqa: F401


def trees(items):
    """Every binary bracketing of the set `items` (unordered children)."""
    items = list(items)
    if len(items) == 1:
        yield items[0]
        return
    first, rest = items[0], items[1:]
    for r in range(0, len(rest)):
        for left_rest in itertools.combinations(rest, r):
            left = [first] + list(left_rest)
            right = [x for x in rest if x not in left_rest]
            for a in trees(left):
                for b in trees(right):
                    yield [a, b]


def network(itg, rng, n):
    idx = [itg.Index(int(rng.integers(2, 9)), f"x{k}") for k in range(n + 2)]
    As = []
    for k in range(n):
        pick = sorted(set(rng.choice(len(idx), size=3, replace=False).tolist()))
        As.append(SimpleNamespace(inds=tuple(idx[p] for p in pick)))
    return As


def test_reference_example():
    """test/ext/ITensorsTensorOperationsExt/runtests.jl:12-59: A(i', i) with
    A'' and A': the best trees cost 2 d^3 (the reference's answer
    [1, [3, 2]] is one of them)."""
    import itg
    d = 100
    i = itg.Index(d, "i")

    def A(p):  # only the indices enter the cost
        return SimpleNamespace(inds=(i.prime(p + 1), i.prime(p)))

    As = [A(0), A(2), A(1)]
    seq = itg.optimal_contraction_sequence(As)
    assert sum(itg.contraction_cost(As, seq)) == 2 * d ** 3
    assert sum(itg.contraction_cost(As, [1, [3, 2]])) == 2 * d ** 3
    # the outer product first: d^4, then a four-index result against A': d^4
    assert sum(itg.contraction_cost(As, [[1, 2], 3])) == 2 * d ** 4


def test_left_associative_costs():
    import itg
    from itg.itensors import left_associative_contraction_sequence
    assert left_associative_contraction_sequence(4) == [[[1, 2], 3], 4]
    a, b, c = itg.Index(2, "a"), itg.Index(3, "b"), itg.Index(5, "c")
    def T(*ix):
        return SimpleNamespace(inds=ix)

    As = [T(a, b), T(b, c), T(c)]
    # (a b) x (b c): union a, b, c -> 30; result (a, c); x (c): union a, c -> 10
    assert itg.contraction_cost(As) == [30, 10]


def test_optimal_matches_brute_force():
    import itg
    rng = np.random.default_rng(0)
    for n in (3, 4, 5, 6):
        for _ in range(4):
            As = network(itg, rng, n)
            best = min(sum(itg.contraction_cost(As, t)) for t in trees(range(1, n + 1)))
            seq = itg.optimal_contraction_sequence(As)
            assert sum(itg.contraction_cost(As, seq)) == best
            assert sorted(_leaves(seq)) == list(range(1, n + 1))


def _leaves(s):
    return [s] if isinstance(s, int) else _leaves(s[0]) + _leaves(s[1])


def test_sequence_optimization_switch():
    import itg
    assert not itg.using_contraction_sequence_optimization()
    itg.enable_contraction_sequence_optimization()
    assert itg.using_contraction_sequence_optimization()
    itg.disable_contraction_sequence_optimization()
    assert not itg.using_contraction_sequence_optimization()


def test_warn_order_switches():
    """src/global_variables.jl: the threshold can be read, set, reset and
    disabled."""
    import itg.itensors as I
    assert I.get_warn_order() == 14
    assert I.set_warn_order(5) == 14 and I.get_warn_order() == 5
    I.disable_warn_order()
    assert I.get_warn_order() is None
    I.reset_warn_order()
    assert I.get_warn_order() == 14

"""contract(As; sequence) on the device: every sequence form gives the same
tensor (test/ext/ITensorsTensorOperationsExt/runtests.jl:12-55)."""
import numpy as np
import pytest

from helpers import FP64_RTOL, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("qns", [False, True])
def test_sequences_agree(qns):
    import torch
    import itg
    torch.cuda.init()
    if qns:
        i = itg.Index([(itg.QN(("Sz", q)), d) for q, d in ((-1, 20), (0, 40), (1, 20))], "i")
    else:
        i = itg.Index(80, "i")
    A = itg.random_itensor(i.prime(), i.dag(), rng=np.random.default_rng(1))

    def P(T, n):
        return itg.ITensor(T.tensor, tuple(x.prime(n) for x in T.inds))

    net = [P(A, 2), P(A, 1), A]
    ref = itg.contract_network(net)  # left associative
    assert set(ref.inds) == {i.prime(3), i.dag()}
    order = ref.inds
    for seq in ("automatic", "left_associative", "right_associative", [[1, 2], 3], [[2, 3], 1],
                [[1, 3], 2]):
        got = itg.contract_network(net, sequence=seq)
        assert set(got.inds) == set(order)
        assert relerr(got.array(*order), ref.array(*order)) < FP64_RTOL, seq
    itg.enable_contraction_sequence_optimization()
    try:
        got = itg.contract_network([A, P(A, 2), P(A, 1)])
        assert relerr(got.array(*order), ref.array(*order)) < FP64_RTOL
    finally:
        itg.disable_contraction_sequence_optimization()

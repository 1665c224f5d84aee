"""Intermediate mode orders chosen by ContractionSequence: in the two-site
network L * psi * W1 * W2 * R the intermediate Z = (a', b, s1', s2', w'')
meets R(b', w'', b) with its contracted modes b, w'' apart, so the grouped
FP64 GEMM of the last step would permute Z first. With reorder=True the
fused W1 * W2 chain writes Z as (a', s1', s2', w'', b) instead, the last
step reads it in place, and the result (indices and values) is the one of
the plain left-to-right fold (tensor_algebra.jl:121-130).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(t):
    import torch
    return (torch.view_as_real(t) if t.is_complex() else t).contiguous().view(torch.int64)


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_reordered_intermediate_same_result(cplx):
    import torch
    from itg import ndtensors as ND
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    torch.cuda.init()
    dt = np.complex128 if cplx else np.float64
    net = heff_network(512, qns=True, dtype=dt, seed=5, device="cuda")[:5]
    plain = ContractionSequence(net, reorder=False)
    seq = ContractionSequence(net)
    # the plain fold's last GEMM permutes Z; the reordered one does not
    assert plain.steps[-1][1].info()["strategy"] == "gemm"
    assert plain.steps[-1][1].info()["perm1"]
    assert not seq.steps[-1][1].info()["perm1"]
    assert list(seq.chains) == list(plain.chains) == [1]
    assert seq.result.inds == plain.result.inds
    assert seq.steps[2][4].blockdims != plain.steps[2][4].blockdims  # Z's modes moved
    assert seq.result.tensor.blockoffsets == plain.result.tensor.blockoffsets
    for prec in ("fp64", "split_int8"):
        with ND.timer() as t0:
            want = plain.run(prec).tensor.data.clone()
            torch.cuda.synchronize()
        with ND.timer() as t:
            got = seq.run(prec).tensor.data.clone()
            torch.cuda.synchronize()
        if prec == "fp64":
            assert "permute_kernel" in t0.results, t0.results
            assert "permute_kernel" not in t.results, t.results
        assert torch.equal(_bits(got), _bits(want)), prec


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_dense_reordered_intermediate(cplx):
    """The dense network: step 3 (K = 10, strided / fiber kernels) writes Z
    with R's contracted modes last, so step 4's GEMM reads it in place; the
    result matches the plain fold's to rounding (the GEMM may sum in another
    order) and the oracle's restriction to sampled rows."""
    import torch
    from itg import ndtensors as ND
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    torch.cuda.init()
    dt = np.complex128 if cplx else np.float64
    net = heff_network(96, qns=False, dtype=dt, seed=6, device="cuda")[:5]
    plain = ContractionSequence(net, reorder=False)
    seq = ContractionSequence(net)
    assert seq.result.inds == plain.result.inds
    assert seq.steps[2][4].dims != plain.steps[2][4].dims  # Z's modes moved
    for prec in ("fp64", "split_int8"):
        with ND.timer() as t0:
            want = plain.run(prec).tensor.data.clone()
            torch.cuda.synchronize()
        with ND.timer() as t:
            got = seq.run(prec).tensor.data.clone()
            torch.cuda.synchronize()
        if prec == "fp64":
            assert "permute_kernel" in t0.results, t0.results
            assert "permute_kernel" not in t.results, t.results
        err = float(torch.linalg.vector_norm(got - want) / torch.linalg.vector_norm(want))
        assert err < 1e-13, (prec, err)

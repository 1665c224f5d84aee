"""Small-K contractions with one operand much larger than the other (the MPO
steps of an effective-Hamiltonian apply) against the oracle, in both operand
orders (the large operand as T1 and as T2), real and complex with conj flags,
alpha/beta, dense and QN block-sparse.

Bar: relative Frobenius 1e-12 against the extended-precision oracle; block
structure and offsets exact.
"""
import numpy as np
import pytest

from helpers import FP64_RTOL, rand, relerr
from test_gpu_blocksparse import check_against_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


@pytest.mark.parametrize("cplx", [False, True])
def test_dense_mpo_step_both_orders(env, cplx):
    torch, itg, ND, O = env
    rng = np.random.default_rng(21)
    X = rand(rng, (37, 5, 2, 2, 41), cplx)   # (a', w, s1, s2, b)
    W = rand(rng, (5, 2, 2, 5), cplx)        # (w, s1', s1, w')
    lX, lW = (1, -1, -2, 2, 3), (-1, 4, -2, 5)
    R0 = None
    for first_big in (True, False):
        if first_big:
            T1, l1, T2, l2 = X, lX, W, lW
        else:
            T1, l1, T2, l2 = W, lW, X, lX
        lR = ND.contract_labels(l1, l2)
        for c1, c2 in ([(False, False)] if not cplx else [(False, False), (True, False), (False, True)]):
            A1 = ND.DenseTensor.from_numpy(T1)
            A2 = ND.DenseTensor.from_numpy(T2)
            T = ND.contract(A1, l1, A2, l2, lR, conj1=c1, conj2=c2)
            torch.cuda.synchronize()
            ref = O.dense_contract(None, lR, np.conj(T1) if c1 else T1, l1,
                                   np.conj(T2) if c2 else T2, l2, exact=True)
            err = relerr(T.to_numpy(), ref)
            assert err < FP64_RTOL, (first_big, c1, c2, err)
        # alpha / beta on an existing output
        shape = ref.shape
        R0 = rand(rng, shape, cplx)
        T = ND.DenseTensor.from_numpy(R0)
        al, be = (0.75 - 0.5j, 1.5 + 0.25j) if cplx else (0.75, -1.5)
        ND.contract_(T, lR, ND.DenseTensor.from_numpy(T1), l1, ND.DenseTensor.from_numpy(T2), l2,
                     al, be)
        torch.cuda.synchronize()
        ref = O.dense_contract(R0, lR, T1, l1, T2, l2, al, be, exact=True)
        err = relerr(T.to_numpy(), ref)
        assert err < FP64_RTOL, (first_big, "alpha/beta", err)


@pytest.mark.parametrize("cplx", [False, True])
def test_dense_beyond_matrix_form(env, cplx):
    """More distinct A values (K = 20 > 16 slots) or more outputs per fiber
    (40 > 32) than the matrix form holds: the loop form of the kernel."""
    torch, itg, ND, O = env
    rng = np.random.default_rng(22)
    for X, lX, W, lW in [
        (rand(rng, (37, 20, 41), cplx), (1, -1, 2), rand(rng, (20, 3), cplx), (-1, 3)),
        (rand(rng, (37, 4, 41), cplx), (1, -1, 2), rand(rng, (4, 40), cplx), (-1, 3)),
    ]:
        for T1, l1, T2, l2 in ((X, lX, W, lW), (W, lW, X, lX)):
            lR = ND.contract_labels(l1, l2)
            T = ND.contract(ND.DenseTensor.from_numpy(T1), l1, ND.DenseTensor.from_numpy(T2), l2,
                            lR, conj1=cplx)
            torch.cuda.synchronize()
            ref = O.dense_contract(None, lR, np.conj(T1) if cplx else T1, l1, T2, l2, exact=True)
            err = relerr(T.to_numpy(), ref)
            assert err < FP64_RTOL, (T1.shape, T2.shape, err)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_blocksparse_mpo_steps_both_orders(env, dtype):
    torch, itg, ND, O = env
    from itg.configs import heff_network
    L, psi, W1, W2, R, ix = heff_network(128, qns=True, seed=7, dtype=dtype)
    X = itg.contract(L, psi)
    torch.cuda.synchronize()
    for A, B in ((X, W1), (W1, X)):
        C = itg.contract(A, B)
        torch.cuda.synchronize()
        check_against_oracle(env, A, B, C)
    Y = itg.contract(X, W1)
    for A, B in ((Y, W2), (W2, Y)):
        C = itg.contract(A, B)
        torch.cuda.synchronize()
        check_against_oracle(env, A, B, C)

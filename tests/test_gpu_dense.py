"""Dense contraction on the device vs the oracle (FP64 paths).

Mirrors the reference's dense tests: the index-permutation matrix of
test/base/test_contract.jl:7-264, alpha/beta and NaN-overwrite semantics of
NDTensors/test/test_dense.jl:237-287, real/complex promotion
(test_contract.jl:267-324), outer products and full contractions.
Tolerance: relative Frobenius error <= 1e-12 against the oracle.
"""
import itertools

import numpy as np
import pytest

from helpers import FP64_RTOL, labels_for, rand, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


def run_dense(env, A, lA, B, lB, lR=None, alpha=1.0, beta=0.0, R0=None, prec=None):
    torch, itg, ND, O = env
    TA = ND.DenseTensor.from_numpy(A)
    TB = ND.DenseTensor.from_numpy(B)
    if lR is None:
        lR = ND.contract_labels(lA, lB)
    if R0 is None:
        T = ND.contract(TA, lA, TB, lB, lR, alpha=alpha, prec=prec)
    else:
        T = ND.DenseTensor.from_numpy(R0)
        ND.contract_(T, lR, TA, lA, TB, lB, alpha, beta, prec=prec)
    torch.cuda.synchronize()
    return T.to_numpy(), lR


@pytest.mark.parametrize("cplx", [False, True])
def test_matrix_products_all_orientations(env, cplx):
    rng = np.random.default_rng(1)
    _, _, _, O = env
    M, N, K = 37, 53, 70  # not multiples of any tile
    for ta, tb, tc in itertools.product([0, 1], repeat=3):
        A = rand(rng, (K, M) if ta else (M, K), cplx)
        B = rand(rng, (N, K) if tb else (K, N), cplx)
        lA = (-1, 1) if ta else (1, -1)
        lB = (2, -1) if tb else (-1, 2)
        lR = (2, 1) if tc else (1, 2)
        got, _ = run_dense(env, A, lA, B, lB, lR)
        ref = O.dense_contract(None, lR, A, lA, B, lB, exact=True)
        err = relerr(got, ref)
        assert err < FP64_RTOL, (err, (ta, tb, tc))


@pytest.mark.parametrize("cplx", [False, True])
def test_3x3_permutation_matrix(env, cplx):
    """Every placement of one contracted index in two 3-tensors and every
    output order (test_contract.jl:7-264 checks the same matrix)."""
    rng = np.random.default_rng(2)
    _, _, _, O = env
    dims_free = {0: 9, 1: 7, 2: 11, 3: 6}
    kdim = 33  # K >= 24 exercises the GEMM strategy
    for pa in range(3):
        for pb in range(3):
            la, lb = labels_for(3, 3, [(pa, pb)])
            shA = [0] * 3
            shB = [0] * 3
            fi = 0
            for i in range(3):
                if i == pa:
                    shA[i] = kdim
                else:
                    shA[i] = dims_free[fi]
                    fi += 1
            for j in range(3):
                if j == pb:
                    shB[j] = kdim
                else:
                    shB[j] = dims_free[fi]
                    fi += 1
            A = rand(rng, tuple(shA), cplx)
            B = rand(rng, tuple(shB), cplx)
            free = [x for x in la if x > 0] + [x for x in lb if x > 0]
            for lR in itertools.permutations(free):
                got, _ = run_dense(env, A, la, B, lb, lR)
                ref = O.dense_contract(None, lR, A, la, B, lb)
                err = relerr(got, ref)
                assert err < FP64_RTOL, (err, (pa, pb, lR))


def test_4x3_two_contracted(env):
    rng = np.random.default_rng(3)
    _, _, _, O = env
    for pairs in [[(0, 1), (2, 0)], [(1, 2), (3, 0)], [(3, 1), (0, 2)]]:
        la, lb = labels_for(4, 3, pairs)
        dims = {}
        shA = []
        for i in range(4):
            k = [p for p in pairs if p[0] == i]
            shA.append(8 + i if not k else 5 + 3 * i)
        shB = []
        for j in range(3):
            k = [p for p in pairs if p[1] == j]
            shB.append(shA[k[0][0]] if k else 10 + j)
        A = rand(rng, tuple(shA))
        B = rand(rng, tuple(shB))
        free = [x for x in la if x > 0] + [x for x in lb if x > 0]
        for lR in itertools.permutations(free):
            got, _ = run_dense(env, A, la, B, lb, lR)
            ref = O.dense_contract(None, lR, A, la, B, lb)
            err = relerr(got, ref)
            assert err < FP64_RTOL, (err, (pairs, lR))
        del dims


@pytest.mark.parametrize("cplx", [False, True])
def test_alpha_beta_and_nan_overwrite(env, cplx):
    """beta == 0 overwrites NaN; beta != 0 accumulates (test_dense.jl:237-287)."""
    rng = np.random.default_rng(4)
    _, _, _, O = env
    A = rand(rng, (40, 30), cplx)
    B = rand(rng, (30, 50), cplx)
    R0 = np.full((40, 50), np.nan, dtype=np.complex128 if cplx else np.float64)
    got, _ = run_dense(env, A, (1, -1), B, (-1, 2), (1, 2), alpha=2.0, beta=0.0, R0=R0)
    assert np.all(np.isfinite(got))
    err = relerr(got, 2.0 * (A @ B))
    assert err < FP64_RTOL, (err)
    R1 = rand(rng, (40, 50), cplx)
    al, be = (2.5 - 1j, -0.5 + 0.25j) if cplx else (2.5, -0.5)
    got, _ = run_dense(env, A, (1, -1), B, (-1, 2), (1, 2), alpha=al, beta=be, R0=R1)
    ref = O.dense_contract(R1, (1, 2), A, (1, -1), B, (-1, 2), al, be)
    err = relerr(got, ref)
    assert err < FP64_RTOL, (err)
    # small-K (direct) path with beta
    A2 = rand(rng, (40, 3), cplx)
    B2 = rand(rng, (3, 50), cplx)
    got, _ = run_dense(env, A2, (1, -1), B2, (-1, 2), (2, 1), alpha=al, beta=be, R0=R1.T.copy())
    ref = O.dense_contract(R1.T, (2, 1), A2, (1, -1), B2, (-1, 2), al, be)
    err = relerr(got, ref)
    assert err < FP64_RTOL, (err)


def test_outer_and_full_contractions(env):
    rng = np.random.default_rng(5)
    _, _, _, O = env
    a = rand(rng, (6, 5))
    b = rand(rng, (7,))
    got, lR = run_dense(env, a, (1, 2), b, (3,), (2, 3, 1))
    err = relerr(got, O.dense_contract(None, (2, 3, 1), a, (1, 2), b, (3,)))
    assert err < FP64_RTOL, (err)
    # full contraction to a scalar, small and GEMM-sized K
    for shape in [(3, 4), (64, 33)]:
        x = rand(rng, shape)
        y = rand(rng, shape)
        got, _ = run_dense(env, x, (-1, -2), y, (-1, -2), ())
        assert abs(got.reshape(-1)[0] - np.sum(x * y)) < 1e-12 * np.abs(x * y).sum()


def test_mixed_real_complex_promotes(env):
    rng = np.random.default_rng(6)
    _, _, _, O = env
    A = rand(rng, (20, 30))
    B = rand(rng, (30, 25), True)
    got, _ = run_dense(env, A, (1, -1), B, (-1, 2))
    assert got.dtype == np.complex128
    err = relerr(got, A @ B)
    assert err < FP64_RTOL, (err)


def test_config1_cpu_reference_case(env):
    """A[a=128,s=2,b=128] * L[a'=128,w=5,a=128] over a, then * W[w,s,s',w']
    over (w, s); N(0,1) entries from PCG64(0) drawn A, L, W."""
    _, _, _, O = env
    rng = np.random.Generator(np.random.PCG64(0))
    A = rng.standard_normal(128 * 2 * 128).reshape((128, 2, 128), order="F")
    L = rng.standard_normal(128 * 5 * 128).reshape((128, 5, 128), order="F")
    W = rng.standard_normal(5 * 2 * 2 * 5).reshape((5, 2, 2, 5), order="F")
    # ITensor labels: L * A -> (a', w, s, b)
    lL, lA = (1, 2, -1), (-1, 3, 4)
    X, lX = run_dense(env, L, lL, A, lA)
    refX = O.dense_contract(None, lX, L, lL, A, lA, exact=True)
    err = relerr(X, refX)
    assert err < FP64_RTOL, (err)
    lX2, lW = (1, -1, -2, 3), (-1, -2, 4, 5)
    Y, lY = run_dense(env, X, lX2, W, lW)
    refY = O.dense_contract(None, lY, refX.astype(np.float64), lX2, W, lW, exact=True)
    err = relerr(Y, refY)
    assert err < FP64_RTOL, (err)


def test_permutedims(env):
    torch, itg, ND, O = env
    rng = np.random.default_rng(7)
    for shape in [(5, 7, 3, 4), (64, 3, 40), (33, 65)]:
        x = rand(rng, shape)
        T = ND.DenseTensor.from_numpy(x)
        for perm in itertools.permutations(range(len(shape))):
            R = ND.permutedims(T, [p + 1 for p in perm])
            torch.cuda.synchronize()
            assert np.array_equal(R.to_numpy(), np.transpose(x, perm)), (shape, perm)


def test_stale_hip_error_from_the_host_is_not_reported(env):
    """An error left in HIP's per-thread last-error state by an unrelated
    host call (here hipSetDevice with an invalid ordinal, which touches no
    device) must not make the next contraction report a launch failure."""
    import ctypes
    torch, itg, ND, O = env
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipSetDevice.restype = ctypes.c_int
    hip.hipPeekAtLastError.restype = ctypes.c_int
    assert hip.hipSetDevice(9999) != 0
    assert hip.hipPeekAtLastError() != 0
    rng = np.random.default_rng(30)
    A, B = rng.standard_normal((64, 48)), rng.standard_normal((48, 80))
    T = ND.contract(ND.DenseTensor.from_numpy(A), (1, -1), ND.DenseTensor.from_numpy(B), (-1, 2),
                    (1, 2))
    torch.cuda.synchronize()
    assert relerr(T.to_numpy(), A @ B) < FP64_RTOL


def test_order_zero_operands_scale(env):
    """An order-0 ITensor is a number: A[] * B, B * A[], A[] * B[]
    (tensor_algebra.jl:66-77), also against a QN tensor and a diagonal one."""
    import itg
    rng = np.random.default_rng(14)
    i, j = itg.Index(4, "i"), itg.Index(3, "j")
    B = itg.random_itensor(i, j, rng=rng)
    s = itg.itensor_from_array(np.asarray(2.5))
    assert np.array_equal((s * B).array(i, j), 2.5 * B.array(i, j))
    assert np.array_equal((B * s).array(i, j), 2.5 * B.array(i, j))
    assert (s * s).scalar() == 6.25
    z = itg.itensor_from_array(np.asarray(1.0 - 2.0j))
    assert np.allclose((z * B).array(i, j), (1.0 - 2.0j) * B.array(i, j), rtol=0, atol=1e-15)
    q = itg.Index([(itg.QN(0), 2), (itg.QN(1), 3)], "q")
    Q = itg.random_itensor(q, q.dag().prime(), rng=rng)
    SQ = s * Q
    assert SQ.hasqns() and list(SQ.tensor.blockoffsets) == list(Q.tensor.blockoffsets)
    assert np.array_equal(SQ.tensor.data.cpu().numpy(), 2.5 * Q.tensor.data.cpu().numpy())
    D = s * itg.delta(q, q.dag().prime())
    assert D.isdiag() and D.tensor.uniform == 2.5

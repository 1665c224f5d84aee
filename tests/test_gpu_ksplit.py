"""Contractions with a small output and a long contracted extent (inner
products, the cotangents of small tensors such as MPO factors): the plan's
FP64 GEMM runs in split-K form (partial blocks over slices of K, added in
slice order by ksplit_reduce_kernel). Against the oracle at 1e-12 in every
orientation, real and complex, with alpha / beta (beta == 0 never reads R),
for dense and QN block-sparse operands; split precision hands such plans
to the same FP64 form once an output block sums over 2^17 values or more.
"""
import itertools

import numpy as np
import pytest

from helpers import FP64_RTOL, rand, relerr
from test_gpu_dense import run_dense

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_small_output_long_k_all_orientations(env, cplx):
    torch, itg, ND, O = env
    rng = np.random.default_rng(11)
    M, N, K = 10, 7, 150001  # one 64 x 64 tile, K not a multiple of any slice
    for ta, tb, tc in itertools.product([0, 1], repeat=3):
        A = rand(rng, (K, M) if ta else (M, K), cplx)
        B = rand(rng, (N, K) if tb else (K, N), cplx)
        lA = (-1, 1) if ta else (1, -1)
        lB = (2, -1) if tb else (-1, 2)
        lR = (2, 1) if tc else (1, 2)
        with ND.timer() as t:
            got, _ = run_dense(env, A, lA, B, lB, lR, alpha=0.75)
        assert "ksplit_reduce_kernel" in t.results, t.results
        ref = O.dense_contract(None, lR, A, lA, B, lB, 0.75, exact=True)
        assert relerr(got, ref) < FP64_RTOL, (ta, tb, tc)


def test_inner_product_and_beta(env):
    torch, itg, ND, O = env
    rng = np.random.default_rng(12)
    A = rng.standard_normal((64, 3, 2048))
    B = rng.standard_normal((64, 3, 2048))
    with ND.timer() as t:
        got, _ = run_dense(env, A, (-1, -2, -3), B, (-1, -2, -3))
    ref = float(np.sum(A * B))
    assert abs(got.reshape(-1)[0] - ref) < 1e-12 * np.sqrt(np.sum(A * A) * np.sum(B * B))
    if "ksplit_reduce_kernel" in t.results:  # a split plan: the streaming form
        assert "ksplit_dot_kernel" in t.results and "gemm_f64_kernel" not in t.results
    # beta != 0 adds beta * R; beta == 0 overwrites a NaN-filled R
    A2 = rng.standard_normal((20, 30000))
    B2 = rng.standard_normal((30000, 9))
    R0 = rng.standard_normal((20, 9))
    got, _ = run_dense(env, A2, (1, -1), B2, (-1, 2), (1, 2), alpha=2.0, beta=-0.5, R0=R0)
    assert relerr(got, 2.0 * A2 @ B2 - 0.5 * R0) < FP64_RTOL
    got, _ = run_dense(env, A2, (1, -1), B2, (-1, 2), (1, 2), alpha=1.0, beta=0.0,
                       R0=np.full((20, 9), np.nan))
    assert relerr(got, A2 @ B2) < FP64_RTOL


def test_split_precision_hands_long_reductions_to_fp64(env):
    torch, itg, ND, O = env
    rng = np.random.default_rng(13)
    A = rng.standard_normal((12, 1 << 18))
    B = rng.standard_normal((1 << 18, 5))
    with ND.timer() as t:
        got, _ = run_dense(env, A, (1, -1), B, (-1, 2), prec="split_int8")
    assert "ksplit_reduce_kernel" in t.results and "split_gemm_kernel" not in t.results
    assert relerr(got, A @ B) < FP64_RTOL
    # a shorter K keeps the split form
    A = rng.standard_normal((12, 20000))
    B = rng.standard_normal((20000, 5))
    with ND.timer() as t:
        got, _ = run_dense(env, A, (1, -1), B, (-1, 2), prec="split_int8")
    assert "split_gemm_kernel" in t.results
    assert relerr(got, A @ B) < 1e-12


def test_blocksparse_small_output(env):
    """A(i, j, k) * B(dag i, dag j, l) with small k and l: few output blocks,
    each summing many pairs over long K."""
    torch, itg, ND, O = env
    from test_gpu_blocksparse import check_against_oracle
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(14)
    i = Index([(QN(q), 300) for q in range(-2, 3)], "i")
    j = Index([(QN(q), 200) for q in range(-2, 3)], "j")
    k = Index([(QN(0), 3), (QN(1), 2)], "k")
    l = Index([(QN(0), 2), (QN(-1), 4)], "l")
    for cplx in (False, True):
        dt = np.complex128 if cplx else np.float64
        A = itg.random_itensor(QN(0), i, j.dag(), k, dtype=dt, rng=rng)
        B = itg.random_itensor(QN(0), i.dag(), j, l, dtype=dt, rng=rng)
        with ND.timer() as t:
            C = A * B
            torch.cuda.synchronize()
        assert "ksplit_reduce_kernel" in t.results, t.results
        check_against_oracle(env, A, B, C)


@pytest.mark.parametrize("shape", [(8, 6), (2, 3), (3000, 2)], ids=["mfma", "dot", "skinny"])
def test_graph_replay(env, shape):
    """Each split-K form (MFMA partial tiles, the streaming dot kernel, the
    skinny kernel) replays from a graph captured through the C ABI with the
    eager result (no memset in the path)."""
    torch, itg, ND, O = env
    M, N = shape
    kname = {(8, 6): "gemm_f64_kernel", (2, 3): "ksplit_dot_kernel",
             (3000, 2): "ksplit_skinny_kernel"}[shape]
    rng = np.random.default_rng(15)
    K = 100000 if M * N < 1000 else 20000
    TA = ND.DenseTensor.from_numpy(rng.standard_normal((M, K)))
    TB = ND.DenseTensor.from_numpy(rng.standard_normal((K, N)))
    R = ND.DenseTensor(torch.empty(M * N, dtype=torch.float64, device="cuda"), (M, N))
    run = lambda: ND.contract_(R, (1, 2), TA, (1, -1), TB, (-1, 2))  # noqa: E731
    with ND.timer() as t:
        run()
        torch.cuda.synchronize()
    assert "ksplit_reduce_kernel" in t.results and kname in t.results, t.results
    eager = R.data.clone()
    g = ND.Graph.capture(run)
    for rep in range(3):
        R.data.fill_(float("nan"))
        _ = torch.ones(4, device="cuda").sum().item()  # eager work between replays
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(R.data, eager), rep


@pytest.mark.parametrize("qns", [False, True], ids=["dense", "u1"])
def test_norm_squared_complex(env, qns):
    """A * dag(A) = ||A||^2 (test/base/test_qnitensor.jl:937-947 at scale):
    conj inside the streaming dot kernel, one output element."""
    torch, itg, ND, O = env
    rng = np.random.default_rng(16)
    if qns:
        QN, Index = itg.QN, itg.Index
        i = Index([(QN(q), 700) for q in range(-2, 3)], "i")
        j = Index([(QN(q), 300) for q in range(-2, 3)], "j")
        A = itg.random_itensor(QN(0), i, j.dag(), dtype=np.complex128, rng=rng)
    else:
        A = itg.random_itensor(itg.Index(900, "i"), itg.Index(700, "j"), dtype=np.complex128,
                               rng=rng)
    with ND.timer() as t:
        C = A * itg.dag(A)
        torch.cuda.synchronize()
    assert "ksplit_dot_kernel" in t.results, t.results
    got = complex(C.tensor.data.reshape(-1)[0].item())
    ref = float(torch.sum(torch.abs(A.tensor.data) ** 2).item())
    assert abs(got - ref) < 1e-12 * ref and abs(got.imag) < 1e-12 * ref


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_skinny_matrix_times_vectors(env, cplx):
    """Matrix times one to four vectors over a long K (every output block has
    a side of at most 4): the skinny split-K kernel, every orientation, the
    long side on either operand, alpha and beta."""
    torch, itg, ND, O = env
    rng = np.random.default_rng(17)
    K = 20000
    for M, N in ((3000, 1), (3, 2500)):
        for ta, tb, tc in itertools.product([0, 1], repeat=3):
            A = rand(rng, (K, M) if ta else (M, K), cplx)
            B = rand(rng, (N, K) if tb else (K, N), cplx)
            lA = (-1, 1) if ta else (1, -1)
            lB = (2, -1) if tb else (-1, 2)
            lR = (2, 1) if tc else (1, 2)
            R0 = rand(rng, (N, M) if tc else (M, N), cplx)
            with ND.timer() as t:
                got, _ = run_dense(env, A, lA, B, lB, lR, alpha=0.5, beta=2.0, R0=R0)
            assert "ksplit_skinny_kernel" in t.results, (M, N, t.results)
            ref = O.dense_contract(R0, lR, A, lA, B, lB, 0.5, 2.0)
            assert relerr(got, ref) < FP64_RTOL, (M, N, ta, tb, tc)


def test_skinny_blocksparse_and_conj(env):
    """Block-sparse A(i, j, k) * dag(B)(dag i, dag j, l) with a short l:
    several output blocks of one or two columns, conj(B) inside the kernel,
    against the oracle."""
    torch, itg, ND, O = env
    from test_gpu_blocksparse import check_against_oracle
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(18)
    i = Index([(QN(q), 120) for q in range(-2, 3)], "i")
    j = Index([(QN(q), 90) for q in range(-2, 3)], "j")
    k = Index([(QN(0), 300), (QN(1), 200)], "k")
    l = Index([(QN(0), 1), (QN(1), 2)], "l")
    A = itg.random_itensor(QN(0), i, j.dag(), k, dtype=np.complex128, rng=rng)
    B = itg.random_itensor(QN(0), i, j.dag(), l, dtype=np.complex128, rng=rng)
    with ND.timer() as t:
        C = A * itg.dag(B)
        torch.cuda.synchronize()
    assert "ksplit_skinny_kernel" in t.results, t.results
    Bd = itg.ITensor(ND._with_data(B.tensor, B.tensor.data.conj().resolve_conj()),
                     tuple(x.dag() for x in B.inds))
    check_against_oracle(env, A, Bd, C)

"""Split-precision contraction (FP64 operands as int8 slices on the INT8
matrix cores, FP64 recombination) vs the oracle.

Bar: relative Frobenius error <= 1e-12 against the extended-precision oracle,
for dense products in every operand orientation, QN block-sparse steps and
the full effective-Hamiltonian network. Rows with very different magnitudes
are exercised because the scale is per row / per column.
"""
import itertools

import numpy as np
import pytest

from helpers import labels_for, rand, relerr

pytestmark = pytest.mark.gpu
SPLIT_RTOL = 1e-12
PREC = "split_int8"


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


def run(env, A, lA, B, lB, lR, alpha=1.0, beta=0.0, R0=None):
    torch, itg, ND, O = env
    TA, TB = ND.DenseTensor.from_numpy(A), ND.DenseTensor.from_numpy(B)
    if R0 is None:
        T = ND.contract(TA, lA, TB, lB, lR, alpha=alpha, prec=PREC)
    else:
        T = ND.DenseTensor.from_numpy(R0)
        ND.contract_(T, lR, TA, lA, TB, lB, alpha, beta, prec=PREC)
    torch.cuda.synchronize()
    return T.to_numpy()


def test_split_matrix_orientations(env):
    _, _, _, O = env
    rng = np.random.default_rng(11)
    for (M, N, K) in [(37, 53, 70), (130, 65, 300), (1, 40, 64), (64, 64, 1000)]:
        for ta, tb, tc in itertools.product([0, 1], repeat=3):
            A = rand(rng, (K, M) if ta else (M, K))
            B = rand(rng, (N, K) if tb else (K, N))
            lA = (-1, 1) if ta else (1, -1)
            lB = (2, -1) if tb else (-1, 2)
            lR = (2, 1) if tc else (1, 2)
            got = run(env, A, lA, B, lB, lR)
            ref = O.dense_contract(None, lR, A, lA, B, lB, exact=True)
            err = relerr(got, ref)
            assert err < SPLIT_RTOL, (err, M, N, K, ta, tb, tc)


def test_split_permutation_matrix(env):
    _, _, _, O = env
    rng = np.random.default_rng(12)
    for pa, pb in itertools.product(range(3), range(3)):
        la, lb = labels_for(3, 3, [(pa, pb)])
        shA = [9, 7, 11]
        shB = [6, 8, 5]
        shA[pa] = shB[pb] = 40
        A = rand(rng, tuple(shA))
        B = rand(rng, tuple(shB))
        free = [x for x in la if x > 0] + [x for x in lb if x > 0]
        for lR in itertools.permutations(free):
            got = run(env, A, la, B, lb, lR)
            ref = O.dense_contract(None, lR, A, la, B, lb, exact=True)
            err = relerr(got, ref)
            assert err < SPLIT_RTOL, (err, pa, pb, lR)


def test_split_row_scales_zero_rows_alpha_beta(env):
    _, _, _, O = env
    rng = np.random.default_rng(13)
    A = rand(rng, (50, 96)) * np.exp2(rng.integers(-40, 40, size=(50, 1)))
    B = rand(rng, (96, 30)) * np.exp2(rng.integers(-40, 40, size=(1, 30)))
    A[7, :] = 0.0
    B[:, 3] = 0.0
    R1 = rand(rng, (50, 30))
    got = run(env, A, (1, -1), B, (-1, 2), (1, 2), alpha=-1.5, beta=0.25, R0=R1)
    ref = O.dense_contract(R1, (1, 2), A, (1, -1), B, (-1, 2), -1.5, 0.25, exact=True)
    # elementwise relative to the row/column scale product
    scale = np.abs(A).max(1, keepdims=True) * np.abs(B).max(0, keepdims=True) * 96 + np.abs(R1)
    assert np.all(np.abs(got - ref) <= 1e-12 * scale + 1e-300)
    assert np.all(got[7, :] == 0.25 * R1[7, :])
    err = relerr(got, ref)
    assert err < SPLIT_RTOL, err


def test_split_complex_orientations_conj_and_beta(env):
    """ComplexF64: Re/Im slices share a per-row scale; both output components
    come from K-stacked real slice products (conj folded into the slice-set
    choice)."""
    torch, itg, ND, O = env
    rng = np.random.default_rng(15)
    M, N, K = 45, 70, 130
    for ta, tb in itertools.product([0, 1], repeat=2):
        A = rand(rng, (K, M) if ta else (M, K), True)
        B = rand(rng, (N, K) if tb else (K, N), True)
        lA = (-1, 1) if ta else (1, -1)
        lB = (2, -1) if tb else (-1, 2)
        got = run(env, A, lA, B, lB, (1, 2))
        ref = O.dense_contract(None, (1, 2), A, lA, B, lB, exact=True)
        err = relerr(got, ref)
        assert err < SPLIT_RTOL, (err, ta, tb)
        for c1, c2 in [(True, False), (False, True), (True, True)]:
            TA, TB = ND.DenseTensor.from_numpy(A), ND.DenseTensor.from_numpy(B)
            T = ND.contract(TA, lA, TB, lB, (1, 2), prec=PREC, conj1=c1, conj2=c2)
            torch.cuda.synchronize()
            ref = O.dense_contract(None, (1, 2), np.conj(A) if c1 else A, lA,
                                   np.conj(B) if c2 else B, lB, exact=True)
            err = relerr(T.to_numpy(), ref)
            assert err < SPLIT_RTOL, (err, ta, tb, c1, c2)
    A = rand(rng, (40, 96), True)
    B = rand(rng, (96, 33), True)
    R1 = rand(rng, (40, 33), True)
    for be in (0.5, 0.25 - 0.75j):
        got = run(env, A, (1, -1), B, (-1, 2), (1, 2), alpha=2.0, beta=be, R0=R1)
        ref = O.dense_contract(R1, (1, 2), A, (1, -1), B, (-1, 2), 2.0, be, exact=True)
        err = relerr(got, ref)
        assert err < SPLIT_RTOL, (err, be)


def test_split_complex_heff_network(env):
    torch, itg, ND, O = env
    from itg.configs import heff_network
    for qns, chi in [(False, 96), (True, 256)]:
        L, psi, W1, W2, R, ix = heff_network(chi, qns=qns, seed=2, dtype=np.complex128)
        exact = L
        for T in (psi, W1, W2, R):
            exact = itg.contract(exact, T, prec="fp64")
        cur = L
        for T in (psi, W1, W2, R):
            cur = itg.contract(cur, T, prec=PREC)
        torch.cuda.synchronize()
        err = relerr(cur.tensor.data.cpu().numpy(), exact.tensor.data.cpu().numpy())
        assert err < SPLIT_RTOL, (qns, err)


@pytest.mark.parametrize("cplx", [False, True])
def test_split_large_dense_tiles(env, cplx):
    """A 2048 x 2048 product has enough 128 x 64 tiles for the three-stage
    128-row split kernel; K = 300 gives a partial last K step. FP64 numpy
    products are the reference (their own error is ~1e-15)."""
    torch, itg, ND, O = env
    rng = np.random.default_rng(16)
    M, N, K = 2048, 2048, 300
    A = rand(rng, (M, K), cplx)
    B = rand(rng, (K, N), cplx)
    got = run(env, A, (1, -1), B, (-1, 2), (1, 2))
    err = relerr(got, A @ B)
    assert err < SPLIT_RTOL, err
    if cplx:
        TA, TB = ND.DenseTensor.from_numpy(A), ND.DenseTensor.from_numpy(B)
        T = ND.contract(TA, (1, -1), TB, (-1, 2), (1, 2), prec=PREC, conj1=True, conj2=True)
        torch.cuda.synchronize()
        err = relerr(T.to_numpy(), np.conj(A) @ np.conj(B))
        assert err < SPLIT_RTOL, err


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
@pytest.mark.parametrize("K", [8192, 8200, 20480])
def test_split_long_k_chunks(env, cplx, K):
    """Contracted extents above one K chunk (8192: the INT32 level sums of a
    chunk are exact, chunks are summed in FP64): one chunk exactly, a chunk
    plus a partial one, and the 20480 of the dense chi = 4096 network's last
    step. Complex operands keep three slice sets per block, so a chunk's
    offset inside a set's planes must not be scaled by the set count. Also
    the conjugated complex product. FP64 numpy products are the reference."""
    torch, itg, ND, O = env
    rng = np.random.default_rng(K + cplx)
    M, N = 96, 72
    A = rand(rng, (M, K), cplx)
    B = rand(rng, (K, N), cplx)
    got = run(env, A, (1, -1), B, (-1, 2), (1, 2))
    err = relerr(got, A @ B)
    assert err < SPLIT_RTOL, err
    if cplx:
        TA, TB = ND.DenseTensor.from_numpy(A), ND.DenseTensor.from_numpy(B)
        T = ND.contract(TA, (1, -1), TB, (-1, 2), (1, 2), prec=PREC, conj1=True)
        torch.cuda.synchronize()
        err = relerr(T.to_numpy(), np.conj(A) @ B)
        assert err < SPLIT_RTOL, err


@pytest.mark.parametrize("shape", [(200, 150, 300), (70, 64, 20), (64, 130, 2000)])
def test_split_tile_shapes_agree(env, shape):
    """The split tile shapes (64 x 64, 128 x 64, 64 x 64 with A in registers)
    form the same exact INT32 level sums and recombine them in the same
    order, so their results agree bit for bit. K = 20 is a single K step,
    K = 300 ends in a partial one, K = 2000 runs the stage ring through all
    its phases; M, N leave partial tiles."""
    torch, itg, ND, O = env
    rng = np.random.default_rng(21)
    M, N, K = shape
    A = rand(rng, (M, K))
    B = rand(rng, (K, N))
    outs = []
    try:
        for tile in (1, 2, 3):
            ND.set_split_tile(tile)
            outs.append(run(env, A, (1, -1), B, (-1, 2), (1, 2)))
    finally:
        ND.set_split_tile(0)
    assert relerr(outs[0], A @ B) < SPLIT_RTOL
    for tile, o in zip((2, 3), outs[1:]):
        assert np.array_equal(o, outs[0]), tile


def test_split_tile_shapes_agree_blocksparse(env):
    """Grouped block-sparse products (several block pairs per output block,
    ragged block sizes): every tile shape gives the same bits."""
    torch, itg, ND, O = env
    from itg.configs import heff_network
    L, psi, W1, W2, R, ix = heff_network(128, qns=True, seed=3)
    outs = []
    try:
        for tile in (1, 2, 3):
            ND.set_split_tile(tile)
            cur = L
            for T in (psi, W1, W2, R):
                cur = itg.contract(cur, T, prec=PREC)
            torch.cuda.synchronize()
            outs.append(cur.tensor.data.cpu().numpy())
    finally:
        ND.set_split_tile(0)
    for tile, o in zip((2, 3), outs[1:]):
        assert np.array_equal(o, outs[0]), tile


def test_split_repeated_calls_one_plan(env):
    """The same (cached) plan run again on the same data gives the same
    bits, and on new data the new product: nothing of the previous call's
    partial row maxima, exponents or slices leaks into the next. K = 1000
    gives 8 K parts per row tile; a zero row and a non-finite row are
    included."""
    torch, itg, ND, O = env
    rng = np.random.default_rng(22)
    M, N, K = 300, 200, 1000
    A = rand(rng, (M, K))
    B = rand(rng, (K, N))
    A[7, :] = 0.0
    first = run(env, A, (1, -1), B, (-1, 2), (1, 2))
    again = run(env, A, (1, -1), B, (-1, 2), (1, 2))
    assert np.array_equal(first, again)
    assert relerr(first, A @ B) < SPLIT_RTOL
    assert np.all(first[7, :] == 0.0)
    A2 = rand(rng, (M, K))
    A2[11, 500] = np.inf
    third = run(env, A2, (1, -1), B, (-1, 2), (1, 2))
    assert np.all(np.isnan(third[11, :]))
    keep = np.delete(np.arange(M), 11)
    assert relerr(third[keep], (A2 @ B)[keep]) < SPLIT_RTOL


def test_split_nonfinite_propagates(env):
    rng = np.random.default_rng(14)
    A = rand(rng, (20, 64))
    B = rand(rng, (64, 20))
    A[3, 5] = np.nan
    got = run(env, A, (1, -1), B, (-1, 2), (1, 2))
    assert np.all(np.isnan(got[3, :]))
    assert np.all(np.isfinite(np.delete(got, 3, axis=0)))


def test_split_blocksparse_network_chi512(env):
    torch, itg, ND, O = env
    from itg.configs import heff_network
    L, psi, W1, W2, R, ix = heff_network(512, qns=True, seed=0)
    exact = L
    for T in (psi, W1, W2, R):
        exact = itg.contract(exact, T, prec="fp64")
    cur = L
    for T in (psi, W1, W2, R):
        cur = itg.contract(cur, T, prec=PREC)
    torch.cuda.synchronize()
    assert list(cur.tensor.blockoffsets) == list(exact.tensor.blockoffsets)
    err = relerr(cur.tensor.data.cpu().numpy(), exact.tensor.data.cpu().numpy())
    assert err < SPLIT_RTOL, err


def test_split_heff_chi2048_vs_oracle(env):
    """The config-3 network (chi = 2048) in split precision against the numpy
    oracle, overall and per output block."""
    torch, itg, ND, O = env
    from collections import OrderedDict
    from itg.configs import heff_network
    L, psi, W1, W2, R, ix = heff_network(2048, qns=True, seed=0)
    cur = L
    for T in (psi, W1, W2, R):
        cur = itg.contract(cur, T, prec=PREC)
    torch.cuda.synchronize()

    def to_oracle(T):
        host = T.tensor.data.detach().cpu().numpy()
        offs = OrderedDict((tuple(c - 1 for c in b), o) for b, o in T.tensor.blockoffsets.items())
        inds = []
        for i in T.inds:
            space = [(O.QN(*[(n, v, m) for n, (v, m) in q.vals.items()]), d) for q, d in i.space]
            inds.append(O.Index(space, tags=i.tags, plev=i.plev, dir=i.dir, id=i.id))
        return O.BlockSparse(host, offs, tuple(inds))

    oc = to_oracle(L)
    for T in (psi, W1, W2, R):
        OT = to_oracle(T)
        lA, lB = O.compute_contraction_labels(oc.inds, OT.inds)
        oc, _ = O.blocksparse_contract(oc, lA, OT, lB, O.contract_labels(lA, lB))
    x = cur.tensor.data.cpu().numpy()
    err = relerr(x, oc.data)
    assert err < SPLIT_RTOL, err
    worst = 0.0
    for b, o in oc.offsets.items():
        n = oc.block(b).size
        worst = max(worst, relerr(x[o:o + n], oc.data[o:o + n]))
    assert worst < 1e-11, worst

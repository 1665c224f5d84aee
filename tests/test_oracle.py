"""Pins the CPU oracle (oracle/ndtensors_oracle.py) before it is trusted.

Each check is independent of the formula it tests: known-answer tests taken
from the reference's own test-suite, brute force over all block pairs,
invariants (flux conservation), densify-and-compare against numpy's einsum,
explicit loops, and central finite differences. A planted mistake in the
oracle (a dropped pair, a wrong label sign, a transposed operand) fails at
least one of them.
"""
import itertools
from collections import OrderedDict

import numpy as np
import pytest

from helpers import relerr
from oracle import ndtensors_oracle as O


def qidx(pairs, tag="", dir=1):
    return O.Index([(O.QN(q), d) for q, d in pairs], tags=tag, dir=dir)


# ----------------------------------------------------------------- QN / labels
def test_qn_arithmetic():
    assert O.qn_eq(O.qn_add(O.QN(0), O.QN(1)), O.QN(1))
    assert O.qn_eq(O.QN(), O.QN(("Sz", 0)))  # missing names are zero
    assert O.qn_eq(O.qn_add(O.QN(("Sz", 1)), O.QN(("N", 2))), O.QN(("N", 2), ("Sz", 1)))
    assert O.qn_eq(O.qn_add(O.QN(("P", 1, 2)), O.QN(("P", 1, 2))), O.QN(("P", 0, 2)))
    assert not O.qn_eq(O.QN(1), O.QN(-1))
    assert O.qn_eq(O.qn_scale(-1, O.QN(("Sz", 3))), O.QN(("Sz", -3)))


def test_contraction_labels_closed_form():
    i, j, k, l = (O.Index(2, t) for t in "ijkl")
    la, lb = O.compute_contraction_labels((i, j, k), (k, l, i))
    # matches are numbered A-position-major: (i: A0~B2) -> -1, (k: A2~B0) -> -2
    assert la == (-1, 3, -2) and lb == (-2, 4, -1)
    assert O.contract_labels(la, lb) == (3, 4)
    la, lb = O.compute_contraction_labels((i, j), (k, l))
    assert la == (1, 2) and lb == (3, 4)  # outer product
    la, lb = O.compute_contraction_labels((i, j), (j, i))
    assert la == (-1, -2) and lb == (-2, -1)
    assert O.contract_labels(la, lb) == ()
    assert O.contract_inds((i, j, k), (-1, 3, -2), (k, l, i), (-2, 4, -1), (4, 3)) == (l, j)


def test_kat_arrow_mismatch():
    """test/base/test_qnitensor.jl:1799-1805: matching QN indices must have
    opposite arrows."""
    i = qidx([(0, 2), (1, 2)], "i")
    with pytest.raises(ValueError):
        O.compute_contraction_labels((i, i.dag().prime()), (i, i.dag().prime(2)))
    O.compute_contraction_labels((i, i.dag().prime()), (i.dag(), i.prime(3)))


# ----------------------------------------------------------- block structure
def test_kat_flux_and_nnzblocks():
    """test/base/test_qnitensor.jl:564-583."""
    rng = np.random.default_rng(0)
    i = qidx([(0, 1), (1, 2)], "i")
    j = qidx([(0, 3), (1, 4)], "j")
    A = O.random_blocksparse(rng, O.QN(0), (i, j.dag()))
    assert len(A.offsets) == 2
    B = O.random_blocksparse(rng, O.QN(1), (j, i.dag().prime()))
    assert len(B.offsets) == 1
    la, lb = O.compute_contraction_labels(A.inds, B.inds)
    C, plan = O.blocksparse_contract(A, la, B, lb)
    assert set(C.inds) == {i, i.prime()}
    assert len(C.offsets) == 1
    for b in C.offsets:
        assert O.qn_eq(O.flux_of_block(C.inds, b), O.QN(1))


def test_kat_contract_to_scalar_is_norm2():
    """test/base/test_qnitensor.jl:937-947: A * dag(A) has one block equal to
    norm(A)^2."""
    rng = np.random.default_rng(1)
    i = qidx([(0, 1), (1, 1)], "i")
    A = O.random_blocksparse(rng, O.QN(0), (i, i.dag().prime()))
    Ad = O.BlockSparse(np.conj(A.data), A.offsets, tuple(x.dag() for x in A.inds))
    la, lb = O.compute_contraction_labels(A.inds, Ad.inds)
    C, _ = O.blocksparse_contract(A, la, Ad, lb)
    assert len(C.offsets) == 1 and C.data.size == 1
    assert abs(C.data[0] - np.sum(A.data ** 2)) < 1e-14 * np.sum(A.data ** 2)


def test_zero_block_contraction():
    """test/threading/test_threading.jl:60-78: no pair matches -> 0 blocks."""
    i = qidx([(0, 2), (1, 3)], "i")
    j = qidx([(0, 2), (1, 3)], "j")
    offA, _ = O.blockoffsets([(0, 0)], (i, j.dag()))
    offB, _ = O.blockoffsets([(1, 1)], (j, i.dag().prime()))
    A = O.BlockSparse(np.arange(4.0), offA, (i, j.dag()))
    B = O.BlockSparse(np.arange(9.0), offB, (j, i.dag().prime()))
    la, lb = O.compute_contraction_labels(A.inds, B.inds)
    C, plan = O.blocksparse_contract(A, la, B, lb)
    assert plan == [] and len(C.offsets) == 0 and C.data.size == 0


def _random_case(rng, nA=3, nB=3, ncont=1):
    idx = []
    for t in range(nA + nB - ncont):
        n = int(rng.integers(1, 4))
        idx.append(qidx([(q, int(rng.integers(1, 4))) for q in range(-(n // 2), n - n // 2)], f"t{t}"))
    a_inds = idx[:nA]
    b_inds = [x.dag() for x in idx[:ncont]] + idx[nA:]
    rng.shuffle(b_inds)
    A = O.random_blocksparse(rng, O.QN(int(rng.integers(-1, 2))), tuple(a_inds))
    B = O.random_blocksparse(rng, O.QN(int(rng.integers(-1, 2))), tuple(b_inds))
    return A, B


def test_plan_brute_force_and_flux_invariant():
    rng = np.random.default_rng(2)
    done = 0
    while done < 25:
        try:
            A, B = _random_case(rng, 3, 3, int(rng.integers(1, 3)))
        except ValueError:
            continue
        la, lb = O.compute_contraction_labels(A.inds, B.inds)
        lR = O.contract_labels(la, lb)
        offs, nnz, plan, indsR = O.contract_blockoffsets_sequential(A, la, B, lb, lR)
        # brute force: every (b1, b2) whose contracted coordinates agree, in
        # (b1 outer, b2 inner) order; R blocks in first-appearance order
        brute = []
        for b1 in A.offsets:
            for b2 in B.offsets:
                if all(b1[i] == b2[lb.index(x)] for i, x in enumerate(la) if x < 0):
                    bR = tuple(b1[la.index(x)] if x in la else b2[lb.index(x)] for x in lR)
                    brute.append((b1, b2, bR))
        assert plan == brute
        seen = []
        for _, _, bR in brute:
            if bR not in seen:
                seen.append(bR)
        assert list(offs) == seen
        fA = O.flux_of_block(A.inds, next(iter(A.offsets)))
        fB = O.flux_of_block(B.inds, next(iter(B.offsets)))
        for bR in offs:
            assert O.qn_eq(O.flux_of_block(indsR, bR), O.qn_add(fA, fB))
        done += 1


def test_threaded_plan_against_sequential():
    """contract_blockoffsets_threaded (contract_threaded.jl:2-73) holds the
    sequential plan's pairs; it is the sequential plan itself when T1 has
    more blocks than T2, and otherwise the sequential plan of the swapped
    product T2 * T1 with every pair swapped back (T2's blocks outer). Output
    offsets follow first appearance over its own plan (contract_generic.jl:
    23-30), and the contraction's values do not depend on the order."""
    rng = np.random.default_rng(21)
    kinds = set()
    done = 0
    while done < 30:
        try:
            A, B = _random_case(rng, 3, int(rng.integers(2, 4)), int(rng.integers(1, 3)))
        except ValueError:
            continue
        la, lb = O.compute_contraction_labels(A.inds, B.inds)
        lR = O.contract_labels(la, lb)
        offs_s, nnz_s, plan_s, _ = O.contract_blockoffsets_sequential(A, la, B, lb, lR)
        offs_t, nnz_t, plan_t, _ = O.contract_blockoffsets_threaded(A, la, B, lb, lR)
        assert sorted(plan_t) == sorted(plan_s) and nnz_t == nnz_s
        if len(A.offsets) > len(B.offsets):
            assert plan_t == plan_s and offs_t == offs_s
        else:
            _, _, plan_sw, _ = O.contract_blockoffsets_sequential(B, lb, A, la, lR)
            assert plan_t == [(b1, b2, bR) for b2, b1, bR in plan_sw]
        first = []
        for _, _, bR in plan_t:
            if bR not in first:
                first.append(bR)
        assert list(offs_t) == first
        n = 0
        for bR, o in offs_t.items():
            assert o == n
            n += int(np.prod(O.blockdims(O.contract_inds(A.inds, la, B.inds, lb, lR), bR)))
        Cs, _ = O.blocksparse_contract(A, la, B, lb, lR)
        Ct, _ = O.blocksparse_contract(A, la, B, lb, lR, threaded=True)
        ref = Cs.dense()
        assert np.allclose(Ct.dense(), ref, rtol=0, atol=1e-13 * max(1.0, np.abs(ref).max()))
        kinds.add((len(A.offsets) > len(B.offsets), plan_t == plan_s))
        done += 1
    # both branches ran, and the T2-outer one changed the order somewhere
    assert (True, True) in kinds and (False, False) in kinds


def test_kat_threaded_contraction():
    """test/threading/test_threading.jl:30-58 (A' * A with and without
    threaded block sparse: equal tensors) and :60-78 (no matching pair: no
    blocks either way), on the oracle."""
    rng = np.random.default_rng(4)
    i = qidx([(0, 50), (1, 50)], "i")
    A = O.random_blocksparse(rng, O.QN(0), (i.prime(), i.dag()))
    Ap = O.BlockSparse(A.data, A.offsets, tuple(x.prime() for x in A.inds))
    la, lb = O.compute_contraction_labels(Ap.inds, A.inds)
    R, plan = O.blocksparse_contract(Ap, la, A, lb)
    Rt, plan_t = O.blocksparse_contract(Ap, la, A, lb, threaded=True)
    assert sorted(plan) == sorted(plan_t)
    assert np.allclose(R.dense(), Rt.dense(), rtol=0, atol=1e-12 * np.abs(R.dense()).max())
    j = qidx([(0, 1), (1, 1)], "j")
    offA, _ = O.blockoffsets([(0, 0)], (j.prime(), j.dag()))
    offB, _ = O.blockoffsets([(1, 1)], (j.prime(), j.dag()))
    A1 = O.BlockSparse(np.array([11.0]), offA, (j.prime(), j.dag()))
    B1 = O.BlockSparse(np.array([22.0]), offB, (j.prime(), j.dag()))
    A1p = O.BlockSparse(A1.data, A1.offsets, tuple(x.prime() for x in A1.inds))
    la, lb = O.compute_contraction_labels(A1p.inds, B1.inds)
    for threaded in (False, True):
        C, plan = O.blocksparse_contract(A1p, la, B1, lb, threaded=threaded)
        assert plan == [] and len(C.offsets) == 0 and C.data.size == 0


@pytest.mark.parametrize("cplx", [False, True])
def test_blocksparse_values_vs_dense(cplx):
    rng = np.random.default_rng(3)
    done = 0
    while done < 20:
        try:
            A, B = _random_case(rng, 3, 3, int(rng.integers(1, 3)))
        except ValueError:
            continue
        if cplx:
            A.data = A.data + 1j * rng.standard_normal(A.data.size)
        la, lb = O.compute_contraction_labels(A.inds, B.inds)
        C, plan = O.blocksparse_contract(A, la, B, lb)
        ref = np.einsum(O._einsum_spec(la, lb, O.contract_labels(la, lb)), A.dense(), B.dense())
        assert np.allclose(C.dense(), ref, rtol=0, atol=1e-12 * max(1.0, np.abs(ref).max()))
        done += 1


def _only_block(T, b):
    """T with every stored block but b zeroed (same structure)."""
    data = np.zeros_like(T.data)
    o = T.offsets[b]
    n = T.block(b).size
    data[o:o + n] = T.data[o:o + n]
    return O.BlockSparse(data, T.offsets, T.inds)


def test_blocksparse_pair_alpha():
    """Per-pair factors (compute_alpha's slot, contract_generic.jl:91): unit
    factors change nothing; a one-hot factor on pair n gives exactly that
    pair's product (densified operands holding only its two blocks, einsum);
    complex factors c_n give sum_n c_n * (pair n's product)."""
    rng = np.random.default_rng(12)
    done = 0
    while done < 8:
        try:
            A, B = _random_case(rng, 3, 3, int(rng.integers(1, 3)))
        except ValueError:
            continue
        la, lb = O.compute_contraction_labels(A.inds, B.inds)
        lr = O.contract_labels(la, lb)
        full, plan = O.blocksparse_contract(A, la, B, lb)
        if len(plan) < 2:
            continue
        ones, _ = O.blocksparse_contract(A, la, B, lb, pair_alpha=[1.0] * len(plan))
        assert np.array_equal(ones.data, full.data)
        spec = O._einsum_spec(la, lb, lr)
        c = rng.standard_normal(len(plan)) + 1j * rng.standard_normal(len(plan))
        want = np.zeros(full.dense().shape, dtype=np.complex128)
        total = np.zeros_like(full.dense())
        for n, (b1, b2, _bR) in enumerate(plan):
            hot = [0.0] * len(plan)
            hot[n] = 1.0
            got, _ = O.blocksparse_contract(A, la, B, lb, pair_alpha=hot)
            ref = np.einsum(spec, _only_block(A, b1).dense(), _only_block(B, b2).dense())
            assert np.allclose(got.dense(), ref, rtol=0, atol=1e-13 * max(1.0, np.abs(ref).max()))
            total += ref
            want += c[n] * ref
        assert np.allclose(full.dense(), total, rtol=0, atol=1e-12 * max(1.0, np.abs(total).max()))
        got, _ = O.blocksparse_contract(A, la, B, lb, alpha=0.5, pair_alpha=list(c))
        assert np.allclose(got.dense(), 0.5 * want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()))
        with pytest.raises(ValueError):
            O.blocksparse_contract(A, la, B, lb, pair_alpha=[1.0])
        done += 1


# ------------------------------------------------------------ dense contract
def test_dense_contract_vs_loops():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((3, 4, 5))
    B = rng.standard_normal((5, 2, 3))
    # contract A's mode 0 with B's mode 2 and A's mode 2 with B's mode 0
    la, lb = (-1, 1, -2), (-2, 2, -1)
    R = O.dense_contract(None, (2, 1), A, la, B, lb, exact=True)
    loop = np.zeros((2, 4))
    for a, b, c, d in itertools.product(range(3), range(4), range(5), range(2)):
        loop[d, b] += A[a, b, c] * B[c, d, a]
    assert np.allclose(R.astype(np.float64), loop, atol=1e-13)


def test_dense_alpha_beta_nan():
    """beta == 0 overwrites NaN; otherwise R = alpha*AB + beta*R
    (NDTensors/test/test_dense.jl:237-287)."""
    rng = np.random.default_rng(5)
    A, B = rng.standard_normal((4, 3)), rng.standard_normal((3, 5))
    R = np.full((4, 5), np.nan)
    out = O.dense_contract(R, (1, 2), A, (1, -1), B, (-1, 2), 2.0, 0.0)
    assert np.allclose(out, 2 * A @ B)
    R = rng.standard_normal((4, 5))
    out = O.dense_contract(R, (1, 2), A, (1, -1), B, (-1, 2), -1.0, 0.5)
    assert np.allclose(out, -A @ B + 0.5 * R)


def test_exact_mode_is_tighter_than_float64():
    rng = np.random.default_rng(6)
    A = rng.standard_normal((60, 500))
    B = rng.standard_normal((500, 40))
    ex = O.dense_contract(None, (1, 2), A, (1, -1), B, (-1, 2), exact=True)
    fp = O.dense_contract(None, (1, 2), A, (1, -1), B, (-1, 2))
    # both agree to float64 rounding, and the exact one is a longdouble result
    assert ex.dtype == np.longdouble
    assert relerr(fp, ex.astype(np.float64)) < 1e-14


# ------------------------------------------------------------------ adjoint
def test_rrule_against_finite_differences():
    """ext/ITensorsChainRulesCoreExt/itensor.jl:94-103 checked with central
    differences of f(x1, x2) = Re<ybar, x1*x2>."""
    rng = np.random.default_rng(7)
    x1 = rng.standard_normal((3, 4)) + 1j * rng.standard_normal((3, 4))
    x2 = rng.standard_normal((4, 5))
    l1, l2, lR = (1, -1), (-1, 2), (1, 2)
    ybar = rng.standard_normal((3, 5)) + 1j * rng.standard_normal((3, 5))

    def f(a, b):
        y = O.dense_contract(None, lR, a, l1, b, l2)
        return np.real(np.sum(np.conj(ybar) * y))

    g1, g2 = O.rrule_contract(x1, l1, x2, l2, ybar, lR)
    h = 1e-6
    for idx in [(0, 0), (2, 3), (1, 2)]:
        e = np.zeros_like(x1)
        e[idx] = 1
        d_re = (f(x1 + h * e, x2) - f(x1 - h * e, x2)) / (2 * h)
        d_im = (f(x1 + 1j * h * e, x2) - f(x1 - 1j * h * e, x2)) / (2 * h)
        # ChainRules convention: the cotangent is d_re + i d_im
        assert abs(g1[idx] - (d_re + 1j * d_im)) < 1e-7
    for idx in [(0, 0), (3, 4)]:
        e = np.zeros_like(x2)
        e[idx] = 1
        d = (f(x1, x2 + h * e) - f(x1, x2 - h * e)) / (2 * h)
        assert np.isreal(g2).all() and abs(g2[idx] - d) < 1e-7


# -------------------------------------------------------- workload structure
def test_heff_sector_sizes_and_counts():
    """The U(1) chi = 2048 configuration: sector sizes, stored blocks of L,
    psi, W and R, and block pairs per contraction step."""
    qs = np.arange(-12, 13)
    w = np.exp(-qs ** 2 / 32.0)
    n = np.rint(2048 * w / w.sum()).astype(int)
    n[12] = 2048 - (n.sum() - n[12])
    assert list(n) == [2, 5, 9, 16, 28, 44, 66, 94, 124, 154, 181, 198, 206, 198, 181, 154,
                       124, 94, 66, 44, 28, 16, 9, 5, 2]
    bond = [(O.QN(("Sz", int(q))), int(x)) for q, x in zip(qs, n)]
    phys = [(O.QN(("Sz", 1)), 1), (O.QN(("Sz", -1)), 1)]
    link = [(O.QN(("Sz", 0)), 1), (O.QN(("Sz", 2)), 1), (O.QN(("Sz", -2)), 1), (O.QN(("Sz", 0)), 2)]
    a, b = O.Index(bond, "a", dir=1), O.Index(bond, "b", dir=-1)
    s1, s2 = O.Index(phys, "s1", dir=1), O.Index(phys, "s2", dir=1)
    ap, bp = O.Index(bond, "a'", dir=1), O.Index(bond, "b'", dir=-1)
    w0, w1, w2 = (O.Index(link, t, dir=-1) for t in ("w", "w'", "w''"))
    s1p, s2p = O.Index(phys, "s1'", dir=1), O.Index(phys, "s2'", dir=1)
    z = O.QN(("Sz", 0))
    rng = np.random.default_rng(0)
    L = O.random_blocksparse(rng, z, (ap, w0, a.dag()))
    psi = O.random_blocksparse(rng, z, (a, s1, s2, b))
    W1 = O.random_blocksparse(rng, z, (w0.dag(), s1.dag(), s1p, w1))
    W2 = O.random_blocksparse(rng, z, (w1.dag(), s2.dag(), s2p, w2))
    R = O.random_blocksparse(rng, z, (bp, w2.dag(), b.dag()))
    assert [len(t.offsets) for t in (L, psi, W1, W2, R)] == [96, 96, 20, 20, 96]
    cur, pairs = L, []
    for T in (psi, W1, W2, R):
        la, lb = O.compute_contraction_labels(cur.inds, T.inds)
        offs, nnz, plan, indsR = O.contract_blockoffsets_sequential(cur, la, T, lb, O.contract_labels(la, lb))
        pairs.append(len(plan))
        # stand-in with the right structure for the next step
        cur = O.BlockSparse(np.zeros(nnz), offs, indsR)
    assert pairs == [372, 940, 952, 372]


@pytest.mark.parametrize("cplx", [False, True])
def test_heff_apply_dense_vs_brute_force(cplx):
    """heff_apply_dense (the left-to-right fold of the two-site network) against
    the network's defining sum written as seven nested loops on tiny
    dimensions: O[x,k,l,y] = sum L[x,w,a] psi[a,i,j,b] W1[w,i,k,u]
    W2[u,j,l,v] R[y,v,b]. A transposed operand, a wrong contracted pair or a
    dropped step changes the result."""
    rng = np.random.default_rng(11)
    chi, d, w = 3, 2, 2

    def draw(*shape):
        x = rng.standard_normal(shape)
        return x + 1j * rng.standard_normal(shape) if cplx else x

    L, psi = draw(chi, w, chi), draw(chi, d, d, chi)
    W1, W2, R = draw(w, d, d, w), draw(w, d, d, w), draw(chi, w, chi)
    want = np.zeros((chi, d, d, chi), dtype=psi.dtype)
    rs, ds, ws = range(chi), range(d), range(w)
    for x, k, l, y in itertools.product(rs, ds, ds, rs):
        acc = 0.0
        for wl, a, i, j, b, u, v in itertools.product(ws, rs, ds, ds, rs, ws, ws):
            acc += L[x, wl, a] * psi[a, i, j, b] * W1[wl, i, k, u] * W2[u, j, l, v] * R[y, v, b]
        want[x, k, l, y] = acc
    for exact in (False, True):
        got = O.heff_apply_dense(L, psi, W1, W2, R, exact=exact)
        assert got.shape == want.shape
        assert relerr(got, want) < 1e-13
    # a row subset of L gives the same rows of the output (the sampled-row
    # oracle of the large dense configurations relies on it)
    rows = [2, 0]
    assert relerr(O.heff_apply_dense(L[rows], psi, W1, W2, R), want[rows]) < 1e-13


def test_blocksparse_permutedims_union():
    """blocksparse_permutedims_ (blocksparsetensor.jl:834-881) against its
    dense meaning: dense(R') = beta*dense(R) + alpha*permutedims(dense(T));
    R' holds R's blocks in R's order, then the permuted blocks of T that R
    lacks in T's order, except new blocks that come out all zero."""
    rng = np.random.default_rng(13)
    done = 0
    while done < 10:
        try:
            A, _B = _random_case(rng, 3, 3, 1)
        except ValueError:
            continue
        perm = tuple(rng.permutation(3))
        pinds = tuple(A.inds[p] for p in perm)
        blocks = [tuple(b[p] for p in perm) for b in A.offsets]
        if len(blocks) < 3:
            continue
        # R: some of T's permuted blocks missing, R's own block order shuffled
        keep = [blocks[i] for i in rng.permutation(len(blocks))[:len(blocks) - 1]]
        # and a block T does not have (another flux), so that r is kept as beta*r
        others = [b for b in itertools.product(*[range(i.nblocks()) for i in pinds])
                  if b not in blocks]
        if not others:
            continue
        keep.insert(int(rng.integers(0, len(keep) + 1)), others[int(rng.integers(0, len(others)))])
        offs, n = OrderedDict(), 0
        for b in keep:
            offs[b] = n
            n += int(np.prod(O.blockdims(pinds, b)))
        R = O.BlockSparse(rng.standard_normal(n), offs, pinds)
        T = A
        if rng.random() < 0.5:  # an all-zero block of T that R lacks is not inserted
            miss = [b for b in blocks if b not in R.offsets][0]
            inv = tuple(np.argsort(perm))
            bt = tuple(miss[i] for i in inv)
            o = T.offsets[bt]
            T = O.BlockSparse(T.data.copy(), T.offsets, T.inds)
            T.data[o:o + T.block(bt).size] = 0.0
        for alpha, beta in ((1.0, 0.0), (0.5, -2.0), (0.0, 1.5)):
            out = O.blocksparse_permutedims_(R, T, perm, alpha, beta)
            want = beta * R.dense() + alpha * np.transpose(T.dense(), perm)
            assert np.allclose(out.dense(), want, rtol=0, atol=1e-14)
            assert list(out.offsets)[:len(R.offsets)] == list(R.offsets)
            new = list(out.offsets)[len(R.offsets):]
            tnew = [b for b in blocks if b not in R.offsets]
            assert new == [b for b in tnew if alpha != 0 and
                           np.any(np.transpose(T.block(tuple(b[i] for i in np.argsort(perm))),
                                               perm) != 0)]
        # beta == 0 does not read R
        Rn = O.BlockSparse(np.full(R.data.size, np.nan), R.offsets, R.inds)
        out = O.blocksparse_permutedims_(Rn, T, perm, 1.0, 0.0)
        assert not np.isnan(out.data).any()
        done += 1

"""CPU checks of the oracle's DiagBlockSparse restatement (no GPU).

Pins diagblocksparse_contract (densify the diagonal blocks, then the
block-sparse contraction) against what the mathematics fixes, so a wrong
diagonal offset, a transposed block or a dropped pair fails:
  * delta(dag(j), j') replaces j by j' (values unchanged);
  * U * diag(s) scales U's columns by s;
  * a fully contracted diagonal gives sum_j d_j A[j, j] (explicit loop);
  * the reference's cases (test_diagblocksparse.jl:31-78): square and
    non-square blocks against the dense contraction, the off-diagonal error;
  * random QN structures: densify(result) == einsum(densify(A), densify(D)).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import ndtensors_oracle as O  # noqa: E402


def qidx(pairs, tags="", dir=1, plev=0, id=None):
    return O.Index([(O.QN(("Sz", q)), d) for q, d in pairs], tags, plev, dir, id)


def diag_of(inds, values=None, uniform=None, flux=None):
    blocks = O.nzdiagblocks(O.QN() if flux is None else flux, inds)
    offs, n = {}, 0
    for b in blocks:
        offs[b] = n
        n += min(O.blockdims(inds, b))
    return O.DiagBlockSparse(None if values is None else np.asarray(values)[:n], offs, inds,
                             uniform=uniform), n


def test_delta_replaces_an_index():
    rng = np.random.default_rng(0)
    i = qidx([(-1, 2), (0, 3), (1, 2)], "i")
    j = qidx([(-2, 1), (0, 2), (2, 3)], "j")
    k = qidx([(-1, 1), (1, 2)], "k", dir=-1)
    A = O.random_blocksparse(rng, O.QN(("Sz", 1)), [i, j, k])
    jp = qidx([(-2, 1), (0, 2), (2, 3)], "j", plev=1, id=j.id)
    D, _ = diag_of([j.dag(), jp], uniform=1.0)
    R, _plan = O.diagblocksparse_contract(A, (1, -1, 2), D, (-1, 3))
    assert [x.plev for x in R.inds] == [0, 0, 1]
    assert np.array_equal(np.transpose(R.dense(), (0, 2, 1)), A.dense())


def test_columns_scaled_by_the_diagonal():
    rng = np.random.default_rng(1)
    a = qidx([(-1, 3), (0, 2), (1, 4)], "a")
    b = qidx([(-1, 3), (0, 2), (1, 4)], "b", dir=-1)
    U = O.random_blocksparse(rng, O.QN(), [a, b])
    s = rng.standard_normal(9)
    bp = qidx([(-1, 3), (0, 2), (1, 4)], "b", plev=1, dir=-1, id=b.id + 999)
    S, n = diag_of([b.dag(), bp], values=s)
    assert n == 9
    R, _plan = O.diagblocksparse_contract(U, (1, -1), S, (-1, 2))
    assert np.allclose(R.dense(), U.dense() * s[None, :], rtol=0, atol=1e-15)


def test_fully_contracted_diagonal_sums_the_diagonal():
    rng = np.random.default_rng(2)
    i = qidx([(-1, 2), (0, 3), (1, 2)], "i")
    j = qidx([(-1, 2), (0, 3), (1, 2)], "j", dir=-1)
    A = O.random_blocksparse(rng, O.QN(), [i, j])
    d = rng.standard_normal(7)
    D, _ = diag_of([i.dag(), j.dag()], values=d)
    R, _plan = O.diagblocksparse_contract(A, (-1, -2), D, (-1, -2))
    Ad = A.dense()
    assert np.isclose(R.data[0], sum(d[x] * Ad[x, x] for x in range(7)), rtol=1e-14, atol=0)


@pytest.mark.parametrize("dims", [([2, 2], [2, 2]), ([3, 2, 3], [2, 2])])
def test_reference_square_and_nonsquare_blocks(dims):
    """test_diagblocksparse.jl:49-78 (blocks (1,1), (2,2); uniform one): the
    three label patterns agree with the dense contraction against the dense
    diagonal tensor."""
    rng = np.random.default_rng(3)
    i = O.Index([(O.QN(("A", k)), x) for k, x in enumerate(dims[0])], "i", 0, 1)
    j = O.Index([(O.QN(("A", k)), x) for k, x in enumerate(dims[1])], "j", 0, -1)
    offs, nnz = O.blockoffsets([(0, 0), (1, 1)], [i, j])
    A = O.BlockSparse(rng.standard_normal(nnz), offs, [i, j])
    doffs, n = {}, 0
    for b in [(0, 0), (1, 1)]:
        doffs[b] = n
        n += min(O.blockdims([i, j], b))
    t = O.DiagBlockSparse(None, doffs, [i, j], uniform=1.0)
    Ad, td = A.dense(), t.dense()
    for la, lt in (((1, -2), (3, -2)), ((-2, 1), (-2, 3))):
        R, _ = O.diagblocksparse_contract(A, la, t, lt)
        ref = np.einsum("ab,cb->ac" if la == (1, -2) else "ba,bc->ac", Ad, td)
        assert np.allclose(R.dense(), ref, rtol=0, atol=1e-15)
    R, _ = O.diagblocksparse_contract(A, (-1, -2), t, (-1, -2))
    assert np.isclose(R.data[0], np.sum(Ad * td), rtol=1e-14, atol=0)


def test_off_diagonal_blocks_are_an_error():
    """test_diagblocksparse.jl:31-46."""
    i = O.Index([(O.QN(("A", 0)), 1), (O.QN(("A", 1)), 1)], "i", 0, 1)
    j = O.Index([(O.QN(("A", 0)), 1), (O.QN(("A", 1)), 1)], "j", 0, -1)
    offs, nnz = O.blockoffsets([(0, 1), (1, 0)], [i, j])
    a1 = O.BlockSparse(np.random.default_rng(4).standard_normal(nnz), offs, [i, j])
    a2 = O.DiagBlockSparse(None, {(0, 1): 0, (1, 0): 1}, [i.dag(), j.dag()], uniform=1.0)
    for l1, l2 in (((1, -1), (-1, 2)), ((-1, -2), (-1, -2))):
        with pytest.raises(ValueError):
            O.diagblocksparse_contract(a1, l1, a2, l2)


@pytest.mark.parametrize("seed", range(4))
def test_random_structures_match_dense_einsum(seed):
    rng = np.random.default_rng(10 + seed)
    sp = [(-2, 1), (-1, 2), (0, 3), (1, 2), (2, 1)]
    i, j, k = qidx(sp, "i"), qidx(sp, "j", dir=-1), qidx([(0, 2), (1, 1)], "k")
    A = O.random_blocksparse(rng, O.QN(), [i, j, k], dtype=np.complex128 if seed % 2 else np.float64)
    # order-3 diagonal tensor: contract j (and for odd seeds also k)
    jp = qidx(sp, "j", plev=1, id=j.id + 77)
    D, n = diag_of([j.dag(), jp, jp.prime()], values=rng.standard_normal(64))
    R, _ = O.diagblocksparse_contract(A, (1, -1, 2), D, (-1, 3, 4))
    ref = np.einsum("ajk,jbc->akbc", A.dense(), D.dense())
    assert np.allclose(R.dense(), ref, rtol=0, atol=1e-13)


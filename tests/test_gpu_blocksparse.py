"""QN block-sparse contraction on the device vs the oracle.

Structure (block list, order, offsets, flux) must match exactly; values to
relative Frobenius 1e-12. Known-answer cases from the reference tests:
test/base/test_qnitensor.jl:564-583 (flux and nnzblocks of a product),
:937-947 (A*dag(A) == norm(A)^2 in one block), :1799-1805 (arrow mismatch),
test/threading/test_threading.jl:60-78 (a contraction with zero blocks).
"""
import numpy as np
import pytest

from helpers import FP64_RTOL, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


def oracle_inds(O, inds):
    """Oracle Index objects with the same sectors, arrows and identity."""
    out = []
    for i in inds:
        space = [(O.QN(*[(n, v, m) for n, (v, m) in q.vals.items()]), d) for q, d in i.space]
        out.append(O.Index(space, tags=i.tags, plev=i.plev, dir=i.dir, id=i.id))
    return tuple(out)


def to_oracle(O, T):
    host = T.tensor.data.detach().cpu().numpy()
    offs = {tuple(c - 1 for c in b): o for b, o in T.tensor.blockoffsets.items()}
    from collections import OrderedDict
    return O.BlockSparse(host, OrderedDict(offs), oracle_inds(O, T.inds))


def check_against_oracle(env, A, B, C, exact=False):
    """C = A*B from the device; compare structure and values with the oracle."""
    _, itg, ND, O = env
    OA, OB = to_oracle(O, A), to_oracle(O, B)
    lA, lB = O.compute_contraction_labels(OA.inds, OB.inds)
    lR = O.contract_labels(lA, lB)
    OC, plan = O.blocksparse_contract(OA, lA, OB, lB, lR, exact=exact)
    got_blocks = [tuple(c - 1 for c in b) for b in C.tensor.blockoffsets]
    assert got_blocks == list(OC.offsets.keys())
    assert list(C.tensor.blockoffsets.values()) == list(OC.offsets.values())
    err = relerr(C.tensor.data.cpu().numpy(), OC.data.astype(np.complex128 if np.iscomplexobj(OC.data) else np.float64))
    assert err < FP64_RTOL, (err)
    return plan


def test_kat_flux_and_nnzblocks(env):
    _, itg, ND, O = env
    QN, Index, dag = itg.QN, itg.Index, itg.dag
    i = Index([(QN(0), 1), (QN(1), 2)], "i")
    j = Index([(QN(0), 3), (QN(1), 4)], "j")
    rng = np.random.default_rng(0)
    A = itg.random_itensor(QN(0), i, j.dag(), rng=rng)
    assert A.flux() == QN(0)
    assert A.nnzblocks() == 2
    B = itg.random_itensor(QN(1), j, i.dag().prime(), rng=rng)
    assert B.flux() == QN(1)
    assert B.nnzblocks() == 1
    C = A * B
    assert set(C.inds) == {i, i.prime()}
    assert C.flux() == QN(1)
    assert C.nnzblocks() == 1
    check_against_oracle(env, A, B, C)
    del dag


def test_kat_contract_to_scalar(env):
    torch, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i = Index([(QN(0), 1), (QN(1), 1)], "i")
    A = itg.random_itensor(QN(0), i, i.dag().prime(), rng=np.random.default_rng(1))
    c = A * itg.dag(A)
    assert c.nnz() == 1 and c.nnzblocks() == 1
    assert abs(c.scalar() - A.norm() ** 2) < 1e-12 * A.norm() ** 2


def test_kat_arrow_mismatch_raises(env):
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i = Index([(QN(0), 2), (QN(1), 2)], "i")
    A = itg.random_itensor(QN(0), i, i.dag().prime(), rng=np.random.default_rng(2))
    B = itg.random_itensor(QN(0), i, i.dag().prime(2), rng=np.random.default_rng(3))
    with pytest.raises(ValueError):
        A * B  # i meets i with the same arrow


def test_zero_block_contraction(env):
    """Both inputs hold blocks but no pair matches: the product has no blocks
    (test/threading/test_threading.jl:60-78)."""
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i = Index([(QN(0), 2), (QN(1), 3)], "i")
    j = Index([(QN(0), 2), (QN(1), 3)], "j")
    A = itg.blocksparse_itensor(np.arange(4.0), [(1, 1)], (i, j.dag()))
    B = itg.blocksparse_itensor(np.arange(9.0), [(2, 2)], (j, i.dag().prime()))
    C = A * B
    assert C.nnzblocks() == 0
    assert C.tensor.data.numel() == 0


@pytest.mark.parametrize("cplx", [False, True])
def test_random_qn_structures(env, cplx):
    _, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(10)
    dt = np.complex128 if cplx else np.float64
    for trial in range(6):
        def idx(tag, big):
            n = rng.integers(2, 5)
            return Index([(QN(int(q)), int(rng.integers(1, 40 if big else 6))) for q in range(-(n // 2), n - n // 2)], tag)
        a, b, c, d = idx("a", True), idx("b", True), idx("c", False), idx("d", True)
        A = itg.random_itensor(QN(0), a, b.dag(), c, dtype=dt, rng=rng)
        B = itg.random_itensor(QN(1), b, c.dag(), d, dtype=dt, rng=rng)
        C = A * B
        check_against_oracle(env, A, B, C)
        # product with reversed operand order
        C2 = B * A
        check_against_oracle(env, B, A, C2)


def test_heff_small_chi_end_to_end(env):
    torch, itg, ND, O = env
    from itg.configs import heff_network
    L, psi, W1, W2, R, ix = heff_network(64, qns=True, seed=0)
    cur = L
    for T in (psi, W1, W2, R):
        nxt = cur * T
        check_against_oracle(env, cur, T, nxt)
        cur = nxt
    assert cur.inds == ix.out()
    assert cur.flux() == itg.QN(("Sz", 0))


def test_heff_chi2048_structure_and_values(env):
    """Block-sparse U(1) network at chi=2048: block counts 96/96/20/96 and
    pairs per step 372/940/952/372; every step's values vs the oracle."""
    torch, itg, ND, O = env
    from itg.configs import heff_network
    L, psi, W1, W2, R, ix = heff_network(2048, qns=True, seed=0)
    assert [t.nnzblocks() for t in (L, psi, W1, W2, R)] == [96, 96, 20, 20, 96]
    cur = L
    npairs = []
    for T in (psi, W1, W2, R):
        nxt = cur * T
        plan = check_against_oracle(env, cur, T, nxt)
        npairs.append(len(plan))
        cur = nxt
    assert npairs == [372, 940, 952, 372]


def test_plan_shared_by_real_and_complex_after_autotune(env):
    """Plans are cached by block structure, so a real and a complex tensor of
    one structure share a plan. Autotuning the real contraction caches every
    tile choice, including the mixed one (large tiles for the big output
    blocks, 64 x 64 for the rest), whose shapes differ for complex data; the
    complex autotune and run that follow must use their own."""
    torch, itg, ND, O = env
    QN, Index = itg.QN, itg.Index
    i = Index([(QN(0), 160), (QN(1), 40)], "i")
    j = Index([(QN(0), 150), (QN(1), 30)], "j")
    k = Index([(QN(0), 140), (QN(1), 50)], "k")
    for dtype in (float, complex):
        A = itg.random_itensor(QN(0), i, j.dag(), dtype=dtype, rng=np.random.default_rng(40))
        B = itg.random_itensor(QN(0), j, k.dag(), dtype=dtype, rng=np.random.default_rng(41))
        lA, lB = itg.compute_contraction_labels(A.inds, B.inds)
        lR = ND.contract_labels(lA, lB)
        p = ND.bs_plan(A.tensor, lA, B.tensor, lB, lR)
        R = torch.empty(p.nnzR, dtype=A.eltype, device="cuda")
        p.autotune(R, A.tensor.data, B.tensor.data, prec="fp64")
        C = A * B
        check_against_oracle(env, A, B, C)

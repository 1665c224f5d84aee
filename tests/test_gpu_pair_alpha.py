"""Per-pair factors (the slot compute_alpha fills, contract_utilities.jl:2-15,
called once per block pair at contract_generic.jl:91) and complex scalars on
the device, against the oracle with the same factors in plan order.

Covers the GEMM strategy (FP64 MFMA and split precision), the small-K
strategies, real and complex data, real (fermion-sign-like +-1) and complex
factors, and a complex global alpha in split precision; plus that a plan
keeps the tables of earlier factor sets (a graph captured with them stays
valid) and refuses new factors while the stream is being captured.
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


def _oracle(O, T):
    from collections import OrderedDict
    inds = []
    for i in T.inds:
        space = [(O.QN(*[(n, v, m) for n, (v, m) in q.vals.items()]), d) for q, d in i.space]
        inds.append(O.Index(space, tags=i.tags, plev=i.plev, dir=i.dir, id=i.id))
    offs = OrderedDict((tuple(c - 1 for c in b), o) for b, o in T.tensor.blockoffsets.items())
    return O.BlockSparse(T.tensor.data.detach().cpu().numpy(), offs, tuple(inds))


def _case(itg, big, cplx, seed):
    """A(i, j, k) * B(dag j, dag k, l): every output block (i, l) sums the
    pairs of several (j, k) sector combinations. big: sectors of 40-72 (GEMM
    strategy); otherwise 1-3 (small contracted extent)."""
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(seed)
    if big:
        sz = [(0, 64), (1, 48), (-1, 40), (2, 24)]
        sk = [(0, 72), (1, 56), (-1, 48)]
    else:
        sz = [(0, 3), (1, 2), (-1, 2), (2, 1)]
        sk = [(0, 2), (1, 1), (-1, 3)]
    sp = lambda s: [(QN(("Sz", q)), n) for q, n in s]
    i = Index(sp(sz), "i", dir=itg.Out)
    j = Index(sp(sk), "j", dir=itg.In)
    k = Index(sp(sk), "k", dir=itg.Out)
    l = Index(sp(sz), "l", dir=itg.In)
    dt = np.complex128 if cplx else np.float64
    A = itg.random_itensor(QN(("Sz", 0)), i, j, k, dtype=dt, rng=rng)
    B = itg.random_itensor(QN(("Sz", 0)), j.dag(), k.dag(), l, dtype=dt, rng=rng)
    return A, B


def _run(env, A, B, prec, pair_alpha, alpha=1.0):
    torch, itg, ND, O = env
    lA, lB = itg.compute_contraction_labels(A.inds, B.inds)
    lR = ND.contract_labels(lA, lB)
    p = ND.bs_plan(A.tensor, lA, B.tensor, lB, lR)
    out = torch.full((p.nnzR,), float("nan"), dtype=A.tensor.data.dtype, device=A.tensor.data.device)
    p.execute(out, A.tensor.data, B.tensor.data, alpha=alpha, prec=prec, pair_alpha=pair_alpha)
    torch.cuda.synchronize()
    OA, OB = _oracle(O, A), _oracle(O, B)
    la, lb = O.compute_contraction_labels(OA.inds, OB.inds)
    ref, plan = O.blocksparse_contract(OA, la, OB, lb, O.contract_labels(la, lb), alpha=alpha,
                                       pair_alpha=pair_alpha)
    assert [tuple(c - 1 for c in b) for b in p.blocksR] == list(ref.offsets)
    return p, out.cpu().numpy(), ref


@pytest.mark.parametrize("big", [True, False], ids=["gemm", "smallK"])
@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
@pytest.mark.parametrize("prec", ["fp64", "split_int8"])
def test_pair_alpha_signs(env, big, cplx, prec):
    """Fermion-sign-like factors (+-1, compute_alpha's values) and a scale."""
    _, itg, ND, _ = env
    A, B = _case(itg, big, cplx, seed=3)
    lA, lB = itg.compute_contraction_labels(A.inds, B.inds)
    npairs = ND.bs_plan(A.tensor, lA, B.tensor, lB, ND.contract_labels(lA, lB)).npairs
    assert npairs >= 8
    signs = np.random.default_rng(5).choice([-1.0, 1.0], npairs)
    for alpha in (1.0, -0.75):
        p, got, ref = _run(env, A, B, prec, list(signs), alpha)
        assert relerr(got, ref.data) < FP64_RTOL
    if big:
        assert p.pairs() and max(n for n in np.bincount([r for _, _, r in p.pairs()])) >= 2


@pytest.mark.parametrize("big", [True, False], ids=["gemm", "smallK"])
@pytest.mark.parametrize("prec", ["fp64", "split_int8"])
def test_pair_alpha_complex_factors(env, big, prec):
    """Complex per-pair factors and a complex global alpha on complex data:
    split precision forms the real-factor and imaginary-factor sums
    separately and combines them in FP64."""
    _, itg, ND, _ = env
    A, B = _case(itg, big, True, seed=7)
    lA, lB = itg.compute_contraction_labels(A.inds, B.inds)
    npairs = ND.bs_plan(A.tensor, lA, B.tensor, lB, ND.contract_labels(lA, lB)).npairs
    rng = np.random.default_rng(9)
    fac = rng.standard_normal(npairs) + 1j * rng.standard_normal(npairs)
    for alpha in (1.0, 0.5 - 1.25j):
        _p, got, ref = _run(env, A, B, prec, list(fac), alpha)
        assert relerr(got, ref.data) < FP64_RTOL
    # a complex alpha without per-pair factors
    _p, got, ref = _run(env, A, B, prec, None, 0.3 + 0.9j)
    assert relerr(got, ref.data) < FP64_RTOL


def test_pair_alpha_plans_kept_and_capture_refusal(env):
    """Switching between factor sets reuses each set's tables (no rebuild,
    nothing freed); a factor set first seen under stream capture is refused
    instead of uploading tables into the graph."""
    torch, itg, ND, O = env
    A, B = _case(itg, True, False, seed=11)
    lA, lB = itg.compute_contraction_labels(A.inds, B.inds)
    lR = ND.contract_labels(lA, lB)
    p = ND.bs_plan(A.tensor, lA, B.tensor, lB, lR)
    f1 = [1.0] * p.npairs
    f2 = [(-1.0) ** n for n in range(p.npairs)]
    outs = []
    for f in (f1, f2, f1, f2):
        out = torch.empty(p.nnzR, dtype=torch.float64, device="cuda")
        p.execute(out, A.tensor.data, B.tensor.data, prec="fp64", pair_alpha=f)
        outs.append(out)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[2]) and torch.equal(outs[1], outs[3])
    assert not torch.equal(outs[0], outs[1])
    # a graph captured with f2's tables replays correctly after other sets ran
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    out = torch.empty(p.nnzR, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        p.execute(out, A.tensor.data, B.tensor.data, prec="fp64", pair_alpha=f2)
    p.execute(outs[0], A.tensor.data, B.tensor.data, prec="fp64",
              pair_alpha=[0.5] * p.npairs)
    for _ in range(3):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, outs[1])
    f3 = [2.0] * p.npairs
    g2 = torch.cuda.CUDAGraph()
    with pytest.raises(itg.ItgError):
        with torch.cuda.graph(g2, stream=s):
            p.execute(out, A.tensor.data, B.tensor.data, prec="fp64", pair_alpha=f3)
    torch.cuda.synchronize()


@pytest.mark.parametrize("perm_out", [False, True], ids=["R_in_place", "R_permuted"])
def test_dense_split_complex_alpha_beta(env, perm_out):
    """Dense contract! in split precision with complex alpha and beta (and a
    NaN-free R read through beta), against the oracle's dense_contract; with
    R's modes interleaving the operands' free modes the combined sum goes
    through the output permutation."""
    torch, itg, ND, O = env
    rng = np.random.default_rng(21)
    a = rng.standard_normal((40, 36, 30)) + 1j * rng.standard_normal((40, 36, 30))
    b = rng.standard_normal((36, 50)) + 1j * rng.standard_normal((36, 50))
    la, lb = (1, -1, 2), (-1, 3)
    lr = (1, 3, 2) if perm_out else (1, 2, 3)
    dims = {1: 40, 2: 30, 3: 50}
    r0 = rng.standard_normal([dims[x] for x in lr]) + 1j * rng.standard_normal([dims[x] for x in lr])
    alpha, beta = 0.7 - 0.4j, -0.3 + 1.1j
    want = O.dense_contract(r0, lr, a, la, b, lb, alpha=alpha, beta=beta)
    A = ND.DenseTensor.from_numpy(a)
    Bt = ND.DenseTensor.from_numpy(b)
    R = ND.DenseTensor.from_numpy(r0)
    ND.contract_(R, lr, A, la, Bt, lb, alpha=alpha, beta=beta, prec="split_int8")
    torch.cuda.synchronize()
    assert relerr(R.to_numpy(), want) < FP64_RTOL

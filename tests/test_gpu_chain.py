"""Fused pair of block-sparse contractions (itg_bs_chain_*): the two MPO
steps of the two-site effective Hamiltonian, ((L*psi)*W1)*W2, run as one
kernel that never stores (L*psi)*W1.

The fused result must be bit-identical to the two separate contractions
(same operations in the same order), and the whole apply must match the
oracle to 1e-12 relative Frobenius. Pairs without a fused form (physical
sectors wider than one, mismatched plans) must be refused, and the sequence
must then fall back to the separate steps.
"""
from collections import OrderedDict

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


def to_oracle(O, T):
    host = T.tensor.data.detach().cpu().numpy()
    offs = OrderedDict((tuple(c - 1 for c in b), o) for b, o in T.tensor.blockoffsets.items())
    inds = []
    for i in T.inds:
        space = [(O.QN(*[(n, v, m) for n, (v, m) in q.vals.items()]), d) for q, d in i.space]
        inds.append(O.Index(space, tags=i.tags, plev=i.plev, dir=i.dir, id=i.id))
    return O.BlockSparse(host, offs, tuple(inds))


def oracle_apply(O, net):
    cur = to_oracle(O, net[0])
    for T in net[1:]:
        OT = to_oracle(O, T)
        lA, lB = O.compute_contraction_labels(cur.inds, OT.inds)
        cur, _ = O.blocksparse_contract(cur, lA, OT, lB, O.contract_labels(lA, lB))
    return cur


def same_blocks(T, OR):
    got = [tuple(c - 1 for c in b) for b in T.tensor.blockoffsets]
    assert got == list(OR.offsets.keys())
    assert list(T.tensor.blockoffsets.values()) == list(OR.offsets.values())


@pytest.mark.parametrize("chi,cplx", [(128, False), (300, False), (256, True), (1024, False)])
def test_fused_mpo_steps_bit_identical_and_match_oracle(env, chi, cplx):
    torch, itg, ND, O = env
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    net = heff_network(chi, qns=True, seed=chi, dtype=np.complex128 if cplx else np.float64)[:5]
    fused = ContractionSequence(net)
    plain = ContractionSequence(net, fuse=False)
    assert list(fused.chains) == [1], "the two MPO steps should fuse"
    assert not plain.chains
    for prec in ("fp64", "split_int8"):
        zf = fused.run(prec).tensor.data.clone()
        zp = plain.run(prec).tensor.data.clone()
        torch.cuda.synchronize()
        assert torch.equal(zf, zp), prec
    # the fused pair alone against the separate pair, on the same input
    _k, p1, X, B1, Y, _l = plain.steps[1]
    _k, p2, _Y, B2, Z, _l = plain.steps[2]
    p1.execute(Y.data, X.data, B1.data)
    p2.execute(Z.data, Y.data, B2.data)
    zsep = Z.data.clone()
    Z.data.fill_(float("nan"))
    ND.BSChain(p1, p2).execute(Z.data, X.data, B1.data, B2.data)
    torch.cuda.synchronize()
    assert torch.equal(Z.data, zsep)
    # the whole apply against the oracle
    if chi <= 300:
        fused.run("fp64")
        ref = oracle_apply(O, net)
        same_blocks(fused.result, ref)
        assert relerr(fused.result.tensor.data.cpu().numpy(), ref.data) < FP64_RTOL


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_fused_pair_non_finite_values(env, cplx):
    """A wave whose x (or y) values are not all finite forms its sums over
    the coefficient masks only (Inf or NaN times a missing coefficient must
    not reach an output); every other wave adds all terms, which is the same
    sum for finite values. Both must equal the separate steps bit for bit,
    NaN positions included."""
    torch, itg, ND, O = env
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    net = heff_network(128, qns=True, seed=7, dtype=np.complex128 if cplx else np.float64)[:5]
    plain = ContractionSequence(net, fuse=False)
    plain.run("fp64")
    _k, p1, X, B1, Y, _l = plain.steps[1]
    _k, p2, _Y, B2, Z, _l = plain.steps[2]
    xs = X.data.clone()
    flat = X.data.view(torch.float64) if not cplx else torch.view_as_real(X.data).reshape(-1)
    n = flat.numel()
    flat[n // 7] = float("inf")
    flat[n // 3] = float("-inf")
    flat[n // 2] = float("nan")
    p1.execute(Y.data, X.data, B1.data)
    p2.execute(Z.data, Y.data, B2.data)
    zsep = Z.data.clone()
    Z.data.fill_(0.0)
    ND.BSChain(p1, p2).execute(Z.data, X.data, B1.data, B2.data)
    torch.cuda.synchronize()
    a = torch.view_as_real(Z.data) if cplx else Z.data
    b = torch.view_as_real(zsep) if cplx else zsep
    assert torch.equal(torch.isnan(a), torch.isnan(b))
    assert int(torch.isnan(b).sum()) > 0  # the NaN reached some outputs
    fin = ~torch.isnan(b)
    assert torch.equal(a[fin], b[fin])
    assert int(torch.isinf(b).sum()) + int(torch.isnan(b).sum()) < b.numel() // 4  # not everywhere
    X.data.copy_(xs)


@pytest.mark.parametrize("cplx", [False, True])
def test_fused_pair_with_wide_link_sectors(env, cplx):
    """A 3-wide MPO link sector gives more than eight x / y values per
    super-fiber (the 16-wide form of the kernel): still bit-identical to
    the separate steps and equal to the oracle."""
    torch, itg, ND, O = env
    from itg.configs import HeffIndices, SZ
    from itg.network import ContractionSequence
    QN, Index = itg.QN, itg.Index
    ix = HeffIndices(256, True)
    lk = [(QN((SZ, 0)), 1), (QN((SZ, 2)), 1), (QN((SZ, -2)), 1), (QN((SZ, 0)), 3)]
    ix.w, ix.wp, ix.wpp = (Index(lk, t, dir=itg.In) for t in ("w", "w'", "w''"))
    rng = np.random.default_rng(21)
    z = QN((SZ, 0))
    dt = np.complex128 if cplx else np.float64
    net = [itg.random_itensor(z, *inds, rng=rng, dtype=dt)
           for inds in (ix.L(), ix.psi(), ix.W1(), ix.W2(), ix.R())]
    fused = ContractionSequence(net)
    plain = ContractionSequence(net, fuse=False)
    assert list(fused.chains) == [1]
    zf = fused.run("fp64").tensor.data.clone()
    zp = plain.run("fp64").tensor.data.clone()
    torch.cuda.synchronize()
    assert torch.equal(zf, zp)
    ref = oracle_apply(O, net)
    same_blocks(fused.result, ref)
    assert relerr(zf.cpu().numpy(), ref.data) < FP64_RTOL


def test_fused_sequence_graph_replays(env):
    """Captured once, replayed several times with eager work in between:
    every replay equals the eager result."""
    torch, itg, ND, O = env
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    net = heff_network(512, qns=True, seed=3)[:5]
    seq = ContractionSequence(net)
    assert seq.chains
    eager = seq.run("fp64").tensor.data.clone()
    g = seq.capture("fp64")
    for _ in range(3):
        seq.result.tensor.data.fill_(0.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(seq.result.tensor.data, eager)
        _ = (net[1].tensor.data * 2.0).sum().item()  # unrelated eager work


def _wide_physical_network(itg, chi, seed):
    from itg.configs import HeffIndices, SZ
    QN, Index = itg.QN, itg.Index
    ix = HeffIndices(chi, True)
    ph2 = [(QN((SZ, 1)), 2), (QN((SZ, -1)), 2)]
    ix.s1 = Index(ph2, "s1", dir=itg.Out)
    ix.s2 = Index(ph2, "s2", dir=itg.Out)
    ix.s1p = Index(ph2, "s1'", dir=itg.Out)
    ix.s2p = Index(ph2, "s2'", dir=itg.Out)
    rng = np.random.default_rng(seed)
    z = QN((SZ, 0))
    return [itg.random_itensor(z, *inds, rng=rng, device="cuda")
            for inds in (ix.L(), ix.psi(), ix.W1(), ix.W2(), ix.R())]


def test_pairs_without_fused_form_fall_back(env):
    """Physical sectors of size 2: the step-1-only fiber mode is not unit,
    so the pair has no fused form; the sequence runs the steps apart."""
    torch, itg, ND, O = env
    from itg.network import ContractionSequence
    net = _wide_physical_network(itg, 128, 1)
    seq = ContractionSequence(net)
    assert not seq.chains
    _k, p1, *_r = seq.steps[1]
    _k, p2, *_r = seq.steps[2]
    with pytest.raises(itg.ItgUnsupported):
        ND.BSChain(p1, p2)
    seq.run("fp64")
    ref = oracle_apply(O, net)
    same_blocks(seq.result, ref)
    assert relerr(seq.result.tensor.data.cpu().numpy(), ref.data) < FP64_RTOL


def test_mismatched_plans_rejected(env):
    """The first plan's output must be the second plan's first operand."""
    torch, itg, ND, O = env
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    seq = ContractionSequence(heff_network(128, qns=True, seed=2)[:5], fuse=False)
    p0, p2 = seq.steps[0][1], seq.steps[2][1]
    with pytest.raises(itg.ItgError) as e:
        ND.BSChain(p0, p2)
    assert not isinstance(e.value, itg.ItgUnsupported)

"""Differentiation and batching on the device.

* rrule(contract) (ext/ITensorsChainRulesCoreExt/itensor.jl:94-103): the
  pullback of one dense and one QN block-sparse contraction against the
  oracle's closed form, and central finite differences (h = 1e-6).
* E(psi) = Re<psi|H psi>/<psi|psi> over the two-site network: the reverse-mode
  gradient against ((H + H^T) psi - 2 E psi)/<psi|psi> computed by the oracle
  (dense einsum), for dense and U(1) block-sparse psi, plus directional
  central differences at chi = 1024.
* 256 independent dense networks with bond dimensions uniform in [16, 96]:
  one grouped launch per kernel and step, captured in a HIP graph; every
  replay must equal the eager result and the per-network contraction.
"""
import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu

HPSI = "pwa,asub,wsxv,vuyz,qzb->pxyq"
HTPHI = "pwa,pxyq,wsxv,vuyz,qzb->asub"


def h_apply(L, psi, W1, W2, R):
    """H psi (the HPSI network) as a left-to-right chain of pairwise
    contractions (one 5-operand einsum picks a path that takes minutes)."""
    t = np.einsum("pwa,asub->pwsub", L, psi, optimize=True)
    t = np.einsum("pwsub,wsxv->pubxv", t, W1, optimize=True)
    t = np.einsum("pubxv,vuyz->pbxyz", t, W2, optimize=True)
    return np.einsum("pbxyz,qzb->pxyq", t, R, optimize=True)


def ht_apply(L, phi, W1, W2, R):
    """H^T phi (the HTPHI network), pairwise."""
    t = np.einsum("pwa,pxyq->waxyq", L, phi, optimize=True)
    t = np.einsum("waxyq,wsxv->ayqsv", t, W1, optimize=True)
    t = np.einsum("ayqsv,vuyz->aqsuz", t, W2, optimize=True)
    return np.einsum("aqsuz,qzb->asub", t, R, optimize=True)


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import chainrules as CR
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, CR, O


def test_rrule_dense_and_blocksparse_vs_oracle(env):
    torch, itg, CR, O = env
    rng = np.random.default_rng(21)
    i, j, k = itg.Index(7, "i"), itg.Index(40, "j"), itg.Index(9, "k")
    x1 = itg.random_itensor(i, j, rng=rng, dtype=np.complex128)
    x2 = itg.random_itensor(j, k, rng=rng)
    y, pb = CR.rrule_contract(x1, x2)
    ybar = itg.random_itensor(*y.inds, rng=rng, dtype=np.complex128)
    _, g1, g2 = pb(ybar)
    r1, r2 = O.rrule_contract(x1.array(), (1, -1), x2.array(), (-1, 2), ybar.array(), (1, 2))
    assert relerr(g1.array(*x1.inds), r1) < 1e-12
    assert not g2.eltype.is_complex and relerr(g2.array(*x2.inds), r2) < 1e-12
    # block-sparse: densified gradients vs the dense closed form
    QN = itg.QN
    a = itg.Index([(QN(0), 30), (QN(1), 20)], "a")
    b = itg.Index([(QN(0), 25), (QN(1), 35)], "b")
    c = itg.Index([(QN(0), 4), (QN(1), 6)], "c")
    x1 = itg.random_itensor(QN(0), a, b.dag(), rng=rng)
    x2 = itg.random_itensor(QN(1), b, c, rng=rng)
    y, pb = CR.rrule_contract(x1, x2)
    ybar = itg.ITensor(itg.ndtensors._with_data(y.tensor, torch.randn_like(y.tensor.data)), y.inds)
    _, g1, g2 = pb(ybar)
    r1, r2 = O.rrule_contract(x1.array(), (1, -1), x2.array(), (-1, 2), ybar.array(), (1, 2))
    mask1 = x1.array() != 0
    assert relerr(g1.array(*x1.inds)[mask1], r1[mask1]) < 1e-12
    mask2 = x2.array() != 0
    assert relerr(g2.array(*x2.inds)[mask2], r2[mask2]) < 1e-12


def closed_form_grad(L, psi, W1, W2, R):
    hp = h_apply(L, psi, W1, W2, R)
    htp = ht_apply(L, psi, W1, W2, R)
    den = np.sum(psi * psi)
    E = np.sum(psi * hp) / den
    return E, (hp + htp - 2 * E * psi) / den


@pytest.mark.parametrize("qns,chi", [(False, 12), (True, 64)])
def test_energy_gradient_closed_form(env, qns, chi):
    torch, itg, CR, O = env
    from itg.configs import heff_network
    L, psi, W1, W2, R, ix = heff_network(chi, qns=qns, seed=3)
    for prec in ("fp64", "split_int8"):
        E, g = CR.heff_energy_and_grad(L, psi, W1, W2, R, ix.out(), prec=prec)
        Er, gr = closed_form_grad(L.array(), psi.array(), W1.array(), W2.array(), R.array())
        assert abs(E - Er) < 1e-12 * abs(Er) + 1e-14
        got = g.array(*psi.inds)
        assert relerr(got, gr) < 1e-11, prec


@pytest.mark.parametrize("qns", [False, True])
def test_energy_gradient_fd_chi1024(env, qns):
    """Directional central differences (h = 1e-6) of E at chi = 1024."""
    torch, itg, CR, O = env
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    L, psi, W1, W2, R, ix = heff_network(1024, qns=qns, seed=4)
    E, g = CR.heff_energy_and_grad(L, psi, W1, W2, R, ix.out())
    rng = np.random.default_rng(5)
    x0 = psi.tensor.data.clone()

    def energy():
        seq = ContractionSequence([L, psi, W1, W2, R])
        out = seq.run()
        bra = CR.replaceinds(psi, ix.out())
        num = CR.inner(bra, out)
        return float(np.real(num)) / float(torch.sum(psi.tensor.data ** 2).item())

    for _ in range(3):
        v = torch.from_numpy(rng.standard_normal(x0.numel())).to(x0.device)
        h = 1e-6
        psi.tensor.data.copy_(x0 + h * v)
        ep = energy()
        psi.tensor.data.copy_(x0 - h * v)
        em = energy()
        psi.tensor.data.copy_(x0)
        fd = (ep - em) / (2 * h)
        an = float(torch.dot(g.tensor.data, v).item())
        assert abs(fd - an) < 1e-5 * max(1.0, abs(an)), (fd, an)


def test_batch_of_256_networks_one_graph(env):
    torch, itg, CR, O = env
    from itg.configs import heff_network
    from itg.network import BatchedSequence, ContractionSequence
    rng = np.random.default_rng(6)
    chis = rng.integers(16, 97, size=256)
    nets = []
    for n, chi in enumerate(chis):
        L, psi, W1, W2, R, ix = heff_network(int(chi), qns=False, seed=1 + n)
        nets.append([L, psi, W1, W2, R])
    batch = BatchedSequence(nets)
    from itg import ndtensors as ND
    with ND.timer() as t:
        eager = [r.tensor.data.clone() for r in batch.run()]
        torch.cuda.synchronize()
    # the intermediate before the product with R is written in the order
    # that product reads, so no step permutes
    assert "permute_kernel" not in t.results, t.results
    for n in (0, 17, 255):
        ref = ContractionSequence(nets[n]).run().tensor.data
        assert relerr(eager[n].cpu().numpy(), ref.cpu().numpy()) < 1e-13
    L, psi, W1, W2, R = (t.array() for t in nets[3])
    ref = h_apply(L, psi, W1, W2, R)
    assert relerr(batch.results[3].array(), ref) < 1e-12
    g = batch.capture()
    for rep in range(3):
        for r in batch.results:
            r.tensor.data.fill_(float("nan"))
        g.replay()
        torch.cuda.synchronize()
        for n in range(256):
            assert torch.equal(batch.results[n].tensor.data, eager[n]), (rep, n)
        _ = torch.ones(10, device="cuda").sum().item()  # eager work between replays

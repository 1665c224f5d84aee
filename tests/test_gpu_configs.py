"""The metric's configurations at their full sizes on the device, against the
oracle.

* Headline: the U(1) block-sparse two-site H_eff apply at chi = 4096,
  left to right with the MPO pair fused (ContractionSequence), FP64 MFMA and
  split precision, graph-replayed as the benchmark runs it. The whole output
  is compared with the oracle's block-sparse contraction chain
  (blocksparse/contract.jl:3-76, contract_generic.jl:78-129): block list and
  offsets exactly, relative Frobenius <= 1e-12 overall and <= 1e-11 in every
  output block.
* C2: the dense Float64 network at chi = 1024, both forms.
* C4: the dense ComplexF64 network at chi = 4096, both forms.
  The dense outputs are checked on sampled rows of L's free index a' (a'
  is never contracted, so those rows of the output are the network applied
  to the same rows of L: the oracle's heff_apply_dense on L[rows]), and at
  chi = 96 on the whole output. Tolerance: relative Frobenius <= 1e-12.
"""
import numpy as np
import pytest

from helpers import FP64_RTOL, relerr

pytestmark = pytest.mark.gpu

BLOCK_RTOL = 1e-11


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


def _oracle_bs(O, T):
    from collections import OrderedDict
    inds = []
    for i in T.inds:
        space = [(O.QN(*[(n, v, m) for n, (v, m) in q.vals.items()]), d) for q, d in i.space]
        inds.append(O.Index(space, tags=i.tags, plev=i.plev, dir=i.dir, id=i.id))
    offs = OrderedDict((tuple(c - 1 for c in b), o) for b, o in T.tensor.blockoffsets.items())
    return O.BlockSparse(T.tensor.data.detach().cpu().numpy(), offs, tuple(inds))


def _oracle_chain(O, net):
    ts = [_oracle_bs(O, t) for t in net]
    cur = ts[0]
    for B in ts[1:]:
        lA, lB = O.compute_contraction_labels(cur.inds, B.inds)
        cur, _plan = O.blocksparse_contract(cur, lA, B, lB, O.contract_labels(lA, lB))
    return cur


def _replay(torch, seq, mode):
    """Autotune, capture one apply in a graph and replay it twice (the
    benchmark's timed form); returns the host copy of the output."""
    seq.autotune(mode)
    g = seq.capture(mode)
    seq.result.tensor.data.fill_(float("nan"))
    for _ in range(2):
        g.replay()
    torch.cuda.synchronize()
    return seq.result.tensor.data.detach().cpu().numpy().copy()


@pytest.fixture(scope="module")
def headline(env):
    torch, itg, ND, O = env
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    net = heff_network(4096, qns=True, seed=0, device="cuda")[:5]
    seq = ContractionSequence(net)
    print("[headline] chi=4096 U(1) network built; oracle chain on the host", flush=True)
    ref = _oracle_chain(O, net)
    yield seq, ref
    del seq, net
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("mode", ["fp64", "split_int8"])
def test_headline_chi4096_fused_vs_oracle(env, headline, mode):
    torch, itg, ND, O = env
    seq, ref = headline
    assert seq.chains, "the MPO pair must run fused"
    got = _replay(torch, seq, mode)
    blocks = [tuple(c - 1 for c in b) for b in seq.result.tensor.blockoffsets]
    assert blocks == list(ref.offsets)
    assert list(seq.result.tensor.blockoffsets.values()) == list(ref.offsets.values())
    assert got.shape == ref.data.shape
    err = relerr(got, ref.data)
    worst = 0.0
    for b, o in ref.offsets.items():
        n = ref.block(b).size
        worst = max(worst, relerr(got[o:o + n], ref.data[o:o + n]))
    print(f"[headline] {mode}: rel err {err:.3e}, worst block {worst:.3e}", flush=True)
    assert err < FP64_RTOL
    assert worst < BLOCK_RTOL


def _dense_rows_check(env, chi, dtype, rows, modes, full=False):
    torch, itg, ND, O = env
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    net = heff_network(chi, qns=False, dtype=dtype, seed=0, device="cuda")[:5]
    L, psi, W1, W2, R = (t.tensor.to_numpy() for t in net)
    sel = slice(None) if full else rows
    print(f"[dense chi={chi} {np.dtype(dtype).name}] oracle on {'all' if full else len(rows)} "
          f"rows of L", flush=True)
    want = O.heff_apply_dense(L[sel], psi, W1, W2, R)
    del L, psi, R
    seq = ContractionSequence(net)
    shape = tuple(i.dim for i in seq.result.inds)
    for mode in modes:
        got = _replay(torch, seq, mode).reshape(shape, order="F")[sel]
        err = relerr(got, want)
        print(f"[dense chi={chi}] {mode}: rel err {err:.3e}", flush=True)
        assert err < FP64_RTOL, (mode, err)
    del seq, net
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_dense_network_chi96_full_vs_oracle(env, cplx):
    _dense_rows_check(env, 96, np.complex128 if cplx else np.float64, None,
                      ("fp64", "split_int8"), full=True)


def test_c2_dense_f64_chi1024_rows_vs_oracle(env):
    rows = np.sort(np.random.default_rng(1).choice(1024, 8, replace=False))
    _dense_rows_check(env, 1024, np.float64, rows, ("fp64", "split_int8"))


def test_c4_dense_c128_chi4096_rows_vs_oracle(env):
    rows = np.sort(np.random.default_rng(1).choice(4096, 8, replace=False))
    _dense_rows_check(env, 4096, np.complex128, rows, ("fp64", "split_int8"))

"""Fixed operands in split precision (ITG_FLAG_KEEP_SLICES_* /
ITG_FLAG_REUSE_SLICES_*): an effective Hamiltonian applied many times with
the same environments L and R need not re-slice them.

* The two-site network at chi = 512 (both GEMM steps run the split GEMM)
  with L and R fixed: the first run keeps their digit planes, later runs
  (eager and graph-replayed) reuse them; every output is bit-identical to
  the plain sequence's.
* Reuse really skips the slicing: after a KEEP call, changing T1's data and
  calling with REUSE gives the result of the old data (the documented
  contract: the caller keeps the data fixed); KEEP again picks up the new
  data.
* Refusals: REUSE with no kept slices, or for another data pointer, is
  ITG_ERR_INVALID; the allocation-free execute refuses KEEP and accepts
  REUSE. FP64 ignores the flags.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import _lib
    from itg import ndtensors as ND
    torch.cuda.init()
    return torch, itg, _lib, ND


def _bits(t):
    import torch
    return (torch.view_as_real(t) if t.is_complex() else t).contiguous().view(torch.int64)


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_fixed_environments_bit_identical(env, cplx):
    torch, itg, _lib, ND = env
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    dt = np.complex128 if cplx else np.float64
    net = heff_network(512, qns=True, dtype=dt, seed=3, device="cuda")[:5]
    plain = ContractionSequence(net)
    want = plain.run("split_int8").tensor.data.clone()
    fixed = ContractionSequence(net, fixed=(0, 4))
    with ND.timer() as t1:
        got1 = fixed.run("split_int8").tensor.data.clone()  # keeps L's and R's planes
        torch.cuda.synchronize()
    with ND.timer() as t2:
        got2 = fixed.run("split_int8").tensor.data.clone()  # reuses them
        torch.cuda.synchronize()
    assert torch.equal(_bits(got1), _bits(want))
    assert torch.equal(_bits(got2), _bits(want))
    # the reusing run slices only psi and Z: fewer slicing launches' worth of time
    k = "slice_max_kernel+slice_kernel"
    print(f"[fixed] slicing ms keep {t1.results[k][1]:.4f} reuse {t2.results[k][1]:.4f}", flush=True)
    g = fixed.capture("split_int8")
    fixed.result.tensor.data.fill_(float("nan"))
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(_bits(fixed.result.tensor.data), _bits(want))


def _pair(itg, seed):
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(seed)
    sz = [(0, 72), (1, 48), (-1, 40)]
    sk = [(0, 64), (1, 40), (-1, 48)]
    sp = lambda s: [(QN(("Sz", q)), n) for q, n in s]
    i = Index(sp(sz), "i", dir=itg.Out)
    j = Index(sp(sk), "j", dir=itg.In)
    l = Index(sp(sz), "l", dir=itg.In)
    A = itg.random_itensor(QN(("Sz", 0)), i, j, rng=rng)
    B = itg.random_itensor(QN(("Sz", 0)), j.dag(), l, rng=rng)
    return A, B


def _plan(itg, ND, A, B):
    lA, lB = itg.compute_contraction_labels(A.inds, B.inds)
    return ND.bs_plan(A.tensor, lA, B.tensor, lB, ND.contract_labels(lA, lB))


def test_reuse_skips_slicing_and_refusals(env):
    torch, itg, _lib, ND = env
    A, B = _pair(itg, 11)
    p = _plan(itg, ND, A, B)
    a, b = A.tensor.data, B.tensor.data
    plain = torch.empty(p.nnzR, dtype=torch.float64, device="cuda")
    p.execute(plain, a, b, prec="split_int8")
    # REUSE before any KEEP
    out = torch.empty_like(plain)
    with pytest.raises(_lib.ItgError):
        p.execute(out, a, b, prec="split_int8", flags=_lib.ITG_FLAG_REUSE_SLICES_1)
    # KEEP T1, then REUSE with the same pointer: identical
    p.execute(out, a, b, prec="split_int8", flags=_lib.ITG_FLAG_KEEP_SLICES_1)
    assert torch.equal(out, plain)
    p.execute(out, a, b, prec="split_int8", flags=_lib.ITG_FLAG_REUSE_SLICES_1)
    assert torch.equal(out, plain)
    # change T1 in place: REUSE still gives the old result (no slicing)
    saved = a.clone()
    a.mul_(2.0)
    p.execute(out, a, b, prec="split_int8", flags=_lib.ITG_FLAG_REUSE_SLICES_1)
    torch.cuda.synchronize()
    assert torch.equal(out, plain)
    # KEEP again picks up the new data (doubling scales exactly)
    p.execute(out, a, b, prec="split_int8", flags=_lib.ITG_FLAG_KEEP_SLICES_1)
    torch.cuda.synchronize()
    assert torch.equal(out, 2.0 * plain)
    a.copy_(saved)
    # REUSE for another data pointer holding the same values
    a2 = a.clone()
    with pytest.raises(_lib.ItgError):
        p.execute(out, a2, b, prec="split_int8", flags=_lib.ITG_FLAG_REUSE_SLICES_1)
    # both operands kept, then the allocation-free execute reuses them
    p.execute(out, a, b, prec="split_int8",
              flags=_lib.ITG_FLAG_KEEP_SLICES_1 | _lib.ITG_FLAG_KEEP_SLICES_2)
    assert torch.equal(out, plain)
    import ctypes
    h = ND._handle(out.device)
    al, _k = _lib.scalar2(1.0)
    na = torch.full_like(out, float("nan"))
    st = _lib.lib().itg_bs_execute_noalloc(
        h.h, p.plan, _lib.ITG_F64, ctypes.c_void_p(na.data_ptr()), ctypes.c_void_p(a.data_ptr()),
        ctypes.c_void_p(b.data_ptr()), al, None, _lib.PREC["split_int8"],
        _lib.ITG_FLAG_REUSE_SLICES_1 | _lib.ITG_FLAG_REUSE_SLICES_2)
    torch.cuda.synchronize()
    assert st == 0, _lib.lib().itg_last_error(h.h)
    assert torch.equal(na, plain)
    st = _lib.lib().itg_bs_execute_noalloc(
        h.h, p.plan, _lib.ITG_F64, ctypes.c_void_p(na.data_ptr()), ctypes.c_void_p(a.data_ptr()),
        ctypes.c_void_p(b.data_ptr()), al, None, _lib.PREC["split_int8"],
        _lib.ITG_FLAG_KEEP_SLICES_1)
    assert st == 1  # ITG_ERR_INVALID: keeping allocates
    # FP64 ignores the flags
    f64 = torch.empty_like(plain)
    p.execute(f64, a, b, prec="fp64")
    out.fill_(float("nan"))
    p.execute(out, a, b, prec="fp64", flags=_lib.ITG_FLAG_REUSE_SLICES_1 | _lib.ITG_FLAG_REUSE_SLICES_2)
    assert torch.equal(out, f64)

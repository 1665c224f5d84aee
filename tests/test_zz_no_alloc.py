"""itg_bs_execute_noalloc (the block-sparse execute that allocates no device
memory: its workspace is the library's fixed scratch array) against
itg_bs_execute on the same plan and inputs, through the C ABI.

Two calls in a row on one GPU, the second with a smaller contraction, in
FP64 MFMA and split precision, real and complex: every output must be bit
for bit the one itg_bs_execute gives (same kernels, same tables; the scratch
contents left by the first call must not leak into the second), including
an inner product (split-K partial sums in the scratch array). Also: a
plan whose launch tables do not exist yet, and a contraction whose
workspace exceeds the scratch array, are refused with ITG_ERR_INVALID.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ITG_ERR_INVALID = 1


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import _lib
    from itg import ndtensors as ND
    torch.cuda.init()
    return torch, itg, _lib, ND


def _pair(itg, sectors, cplx, seed):
    """A(i, j, k) * B(dag j, dag k, l) with QN sectors scaled from `sectors`:
    several (j, k) pairs per output block (GEMM strategy)."""
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(seed)
    sz = [(0, 4 * sectors), (1, 3 * sectors), (-1, 2 * sectors + 5), (2, sectors + 3)]
    sk = [(0, 4 * sectors + 8), (1, 3 * sectors + 8), (-1, 3 * sectors)]
    sp = lambda s: [(QN(("Sz", q)), n) for q, n in s]
    i = Index(sp(sz), "i", dir=itg.Out)
    j = Index(sp(sk), "j", dir=itg.In)
    k = Index(sp(sk), "k", dir=itg.Out)
    l = Index(sp(sz), "l", dir=itg.In)
    dt = np.complex128 if cplx else np.float64
    A = itg.random_itensor(QN(("Sz", 0)), i, j, k, dtype=dt, rng=rng)
    B = itg.random_itensor(QN(("Sz", 0)), j.dag(), k.dag(), l, dtype=dt, rng=rng)
    return A, B


def _plan(itg, ND, A, B):
    lA, lB = itg.compute_contraction_labels(A.inds, B.inds)
    return ND.bs_plan(A.tensor, lA, B.tensor, lB, ND.contract_labels(lA, lB))


def _noalloc(env, p, out, A, B, prec, alpha=1.0, pair_alpha=None):
    torch, itg, _lib, ND = env
    a, _ka = _lib.scalar2(alpha)
    pa, keep = None, None
    if pair_alpha is not None:
        keep = np.zeros(2 * p.npairs, dtype=np.float64)
        c = np.asarray(pair_alpha, dtype=np.complex128)
        keep[0::2], keep[1::2] = c.real, c.imag
        pa = keep.ctypes.data_as(_lib.DP)
    h = ND._handle(out.device)
    st = _lib.lib().itg_bs_execute_noalloc(
        h.h, p.plan, ND._dtype_code(out.dtype), ctypes.c_void_p(out.data_ptr()),
        ctypes.c_void_p(A.tensor.data.data_ptr()), ctypes.c_void_p(B.tensor.data.data_ptr()),
        a, pa, _lib.PREC[prec], 0)
    torch.cuda.synchronize()
    msg = _lib.lib().itg_last_error(h.h)
    return st, (msg.decode() if isinstance(msg, bytes) else str(msg))


def _bits(t):
    import torch
    return (torch.view_as_real(t) if t.is_complex() else t).contiguous().view(torch.int64)


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
@pytest.mark.parametrize("prec", ["fp64", "split_int8"])
def test_noalloc_matches_execute_two_calls(env, prec, cplx):
    torch = env[0]
    itg, ND = env[1], env[3]
    for call, (sectors, seed) in enumerate(((8, 1), (3, 2))):  # second call smaller
        A, B = _pair(itg, sectors, cplx, seed)
        p = _plan(itg, ND, A, B)
        dt = A.tensor.data.dtype
        want = torch.full((p.nnzR,), float("nan"), dtype=dt, device="cuda")
        p.execute(want, A.tensor.data, B.tensor.data, alpha=0.75, prec=prec)
        got = torch.full((p.nnzR,), float("nan"), dtype=dt, device="cuda")
        print(f"[noalloc] call {call}: {prec} {'c128' if cplx else 'f64'} sectors={sectors} "
              f"nnzR={p.nnzR} pairs={p.npairs}", flush=True)
        st, msg = _noalloc(env, p, got, A, B, prec, alpha=0.75)
        assert st == 0, msg
        assert bool(torch.isfinite(want).all())
        assert torch.equal(_bits(got), _bits(want)), "outputs differ"


@pytest.mark.parametrize("prec", ["fp64", "split_int8"])
def test_noalloc_pair_factors(env, prec):
    torch, itg, _lib, ND = env
    A, B = _pair(itg, 6, False, 3)
    p = _plan(itg, ND, A, B)
    fac = np.where(np.arange(p.npairs) % 3 == 1, -1.0, 1.0) * (1.0 + 0.01 * np.arange(p.npairs))
    want = torch.empty((p.nnzR,), dtype=torch.float64, device="cuda")
    p.execute(want, A.tensor.data, B.tensor.data, prec=prec, pair_alpha=fac)
    got = torch.full((p.nnzR,), float("nan"), dtype=torch.float64, device="cuda")
    st, msg = _noalloc(env, p, got, A, B, prec, pair_alpha=fac)
    assert st == 0, msg
    assert torch.equal(_bits(got), _bits(want))
    # a set of factors without tables is refused, not built
    st, msg = _noalloc(env, p, got, A, B, prec, pair_alpha=-fac)
    assert st == ITG_ERR_INVALID and "itg_bs_execute first" in msg, msg


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_noalloc_inner_product_two_calls(env, cplx):
    """A(i, dag j) * B(dag i, j) -> a number: the split-K form with the
    streaming dot kernel, its partial sums in the scratch array; two calls,
    the second smaller, bit for bit as itg_bs_execute."""
    torch, itg, _lib, ND = env
    QN, Index = itg.QN, itg.Index
    dt = np.complex128 if cplx else np.float64
    for call, n in enumerate((400, 150)):
        rng = np.random.default_rng(10 + call)
        i = Index([(QN(("Sz", q)), n) for q in range(-2, 3)], "i", dir=itg.Out)
        j = Index([(QN(("Sz", q)), n + 7) for q in range(-2, 3)], "j", dir=itg.Out)
        A = itg.random_itensor(QN(("Sz", 0)), i, j.dag(), dtype=dt, rng=rng)
        B = itg.random_itensor(QN(("Sz", 0)), i.dag(), j, dtype=dt, rng=rng)
        p = _plan(itg, ND, A, B)
        want = torch.full((p.nnzR,), float("nan"), dtype=A.tensor.data.dtype, device="cuda")
        with ND.timer() as t:
            p.execute(want, A.tensor.data, B.tensor.data, prec="fp64")
            torch.cuda.synchronize()
        assert "ksplit_dot_kernel" in t.results, t.results
        got = torch.full_like(want, float("nan"))
        print(f"[noalloc] inner call {call}: n={n} pairs={p.npairs}", flush=True)
        st, msg = _noalloc(env, p, got, A, B, "fp64")
        assert st == 0, msg
        assert torch.equal(_bits(got), _bits(want))


def test_noalloc_refuses_unbuilt_tables(env):
    torch, itg, _lib, ND = env
    A, B = _pair(itg, 6, False, 4)
    p = _plan(itg, ND, A, B)  # never executed: no launch tables yet
    got = torch.zeros((p.nnzR,), dtype=torch.float64, device="cuda")
    st, msg = _noalloc(env, p, got, A, B, "fp64")
    assert st == ITG_ERR_INVALID and "itg_bs_execute" in msg, msg
    assert bool((got == 0).all())  # nothing ran


def test_noalloc_refuses_workspace_above_scratch(env):
    torch, itg, _lib, ND = env
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(5)
    # one 3000 x 3000 block per operand: split precision needs
    # (9e6 + 9e6) * 6 bytes of digit planes, above the 64 MiB scratch
    i = Index([(QN(("Sz", 0)), 3000)], "i", dir=itg.Out)
    j = Index([(QN(("Sz", 0)), 3000)], "j", dir=itg.In)
    l = Index([(QN(("Sz", 0)), 3000)], "l", dir=itg.In)
    A = itg.random_itensor(QN(("Sz", 0)), i, j, rng=rng)
    B = itg.random_itensor(QN(("Sz", 0)), j.dag(), l, rng=rng)
    p = _plan(itg, ND, A, B)
    want = torch.empty((p.nnzR,), dtype=torch.float64, device="cuda")
    p.execute(want, A.tensor.data, B.tensor.data, prec="split_int8")
    got = torch.zeros_like(want)
    st, msg = _noalloc(env, p, got, A, B, "split_int8")
    assert st == ITG_ERR_INVALID and "scratch" in msg, msg
    assert bool((got == 0).all())
    # the FP64 form of the same contraction needs no workspace (both
    # operands are read in place) and runs once its tables exist
    ref = torch.empty_like(want)
    p.execute(ref, A.tensor.data, B.tensor.data, prec="fp64")
    st, msg = _noalloc(env, p, got, A, B, "fp64")
    assert st == 0, msg
    assert torch.equal(_bits(got), _bits(ref))

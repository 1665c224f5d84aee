"""The GEMM-backend switch (NDTensors' gemm_backend, abstractarray/
tensoralgebra/contract.jl:15-24: :Auto, :BLAS, :Generic) selects the kernel
family on the device: Auto by shape, BLAS the grouped MFMA GEMM whenever a
mode is contracted, Generic the strided kernel for everything. Each
backend's block-sparse and dense results are checked against the oracle
(1e-12 relative Frobenius), and the kernel that ran is read from the
library's per-kernel timer."""
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
    yield torch, itg, ND, O
    ND.backend_auto()


def _oracle(O, T):
    from collections import OrderedDict
    inds = []
    for i in T.inds:
        space = [(O.QN(*[(n, v, m) for n, (v, m) in q.vals.items()]), d) for q, d in i.space]
        inds.append(O.Index(space, tags=i.tags, plev=i.plev, dir=i.dir, id=i.id))
    offs = OrderedDict((tuple(c - 1 for c in b), o) for b, o in T.tensor.blockoffsets.items())
    return O.BlockSparse(T.tensor.data.detach().cpu().numpy(), offs, tuple(inds))


def _qn_pair(itg, big, seed):
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(seed)
    if big:
        sz, sk = [(0, 48), (1, 40), (-1, 32)], [(0, 40), (1, 24), (-1, 32)]
    else:
        sz, sk = [(0, 5), (1, 3), (-1, 4)], [(0, 2), (1, 1), (-1, 3)]
    sp = lambda s: [(QN(("Sz", q)), n) for q, n in s]
    i = Index(sp(sz), "i", dir=itg.Out)
    j = Index(sp(sk), "j", dir=itg.In)
    k = Index(sp(sk), "k", dir=itg.Out)
    l = Index(sp(sz), "l", dir=itg.In)
    A = itg.random_itensor(QN(("Sz", 0)), i, j, k, rng=rng)
    B = itg.random_itensor(QN(("Sz", 0)), k.dag(), j.dag(), l, rng=rng)
    return A, B


# (backend, operand shape) -> the kernel the plan must use
EXPECT = {("Auto", True): "gemm_f64_kernel", ("Auto", False): None,
          ("BLAS", True): "gemm_f64_kernel", ("BLAS", False): "gemm_f64_kernel",
          ("Generic", True): "direct_kernel", ("Generic", False): "direct_kernel"}


@pytest.mark.parametrize("big", [True, False], ids=["largeK", "smallK"])
@pytest.mark.parametrize("backend", ["Auto", "BLAS", "Generic"])
def test_backend_blocksparse_vs_oracle(env, backend, big):
    torch, itg, ND, O = env
    {"Auto": ND.backend_auto, "BLAS": ND.backend_blas, "Generic": ND.backend_generic}[backend]()
    assert ND.gemm_backend() == backend
    A, B = _qn_pair(itg, big, 7 if big else 8)
    with ND.timer() as t:
        C = itg.contract(A, B, prec="fp64")
        torch.cuda.synchronize()
    OA, OB = _oracle(O, A), _oracle(O, B)
    la, lb = O.compute_contraction_labels(OA.inds, OB.inds)
    ref, _p = O.blocksparse_contract(OA, la, OB, lb, O.contract_labels(la, lb))
    assert [tuple(c - 1 for c in b) for b in C.tensor.blockoffsets] == list(ref.offsets)
    err = relerr(C.tensor.data.cpu().numpy(), ref.data)
    print(f"[backend {backend} {'largeK' if big else 'smallK'}] kernels {sorted(t.results)} "
          f"rel err {err:.2e}", flush=True)
    assert err < FP64_RTOL
    want = EXPECT[(backend, big)]
    if want is not None:
        assert want in t.results, (backend, big, sorted(t.results))
    if backend == "Generic":
        assert "gemm_f64_kernel" not in t.results and "permute_kernel" not in t.results
    ND.backend_auto()


@pytest.mark.parametrize("backend", ["Auto", "BLAS", "Generic"])
def test_backend_dense_vs_oracle(env, backend):
    torch, itg, ND, O = env
    {"Auto": ND.backend_auto, "BLAS": ND.backend_blas, "Generic": ND.backend_generic}[backend]()
    rng = np.random.default_rng(3)
    a, b, c, d = (itg.Index(n, t) for n, t in ((60, "a"), (50, "b"), (40, "c"), (30, "d")))
    A = itg.random_itensor(a, b, c, rng=rng)
    B = itg.random_itensor(c, d, b, rng=rng)
    with ND.timer() as t:
        C = itg.contract(A, B, prec="fp64")
        torch.cuda.synchronize()
    ref = np.einsum("abc,cdb->ad", A.array(), B.array())
    got = C.array(a, d)
    err = relerr(got, ref)
    print(f"[backend {backend} dense] kernels {sorted(t.results)} rel err {err:.2e}", flush=True)
    assert err < FP64_RTOL
    assert ("direct_kernel" if backend == "Generic" else "gemm_f64_kernel") in t.results
    ND.backend_auto()

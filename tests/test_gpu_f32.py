"""Float32 / ComplexF32 at the boundary (ITG_F32 / ITG_C64).

Contractions run in FP64 on widened copies and round once when R is
written, so against the oracle evaluated in FP64 on the same (Float32)
inputs each output element is within one Float32 rounding: the tests
assert relative Frobenius <= 1e-6 (2^-24 = 6e-8 per element). Permutations
move the values unchanged (exact). Mixed element types promote to the
common type (promote_type; dense/tensoralgebra/contract.jl:196-211).
"""
import numpy as np
import pytest

from helpers import FP64_RTOL, relerr

pytestmark = pytest.mark.gpu

F32_RTOL = 1e-6


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


def _arr(rng, shape, cplx):
    x = rng.standard_normal(shape)
    if cplx:
        x = x + 1j * rng.standard_normal(shape)
    return x.astype(np.complex64 if cplx else np.float32)


@pytest.mark.parametrize("cplx", [False, True], ids=["f32", "c64"])
@pytest.mark.parametrize("prec", ["fp64", "split_int8"])
def test_dense_contract_f32(env, cplx, prec):
    torch, itg, ND, O = env
    rng = np.random.default_rng(50)
    for la, lb, lr, sa, sb in (((1, -1), (-1, 2), (1, 2), (70, 90), (90, 60)),
                               ((-1, 1, -2), (2, -2, -1), (2, 1), (30, 40, 20), (50, 20, 30)),
                               ((1, 2), (3,), (3, 1, 2), (8, 9), (7,))):
        a, b = _arr(rng, sa, cplx), _arr(rng, sb, cplx)
        dims = dict(zip(la, sa))
        dims.update(zip(lb, sb))
        r0 = _arr(rng, [dims[x] for x in lr], cplx)
        alpha, beta = (0.75 - 0.5j, 0.5 + 0.25j) if cplx else (0.75, -0.5)
        want = O.dense_contract(r0.astype(np.complex128 if cplx else np.float64), lr,
                                a.astype(np.complex128), la, b.astype(np.complex128), lb,
                                alpha=alpha, beta=beta)
        A, B = ND.DenseTensor.from_numpy(a), ND.DenseTensor.from_numpy(b)
        R = ND.DenseTensor.from_numpy(r0)
        ND.contract_(R, lr, A, la, B, lb, alpha=alpha, beta=beta, prec=prec)
        torch.cuda.synchronize()
        got = R.to_numpy()
        assert got.dtype == r0.dtype
        assert relerr(got, want) < F32_RTOL
        # beta = 0 overwrites a NaN-filled R
        R = ND.DenseTensor.from_numpy(np.full(r0.shape, np.nan, dtype=r0.dtype))
        ND.contract_(R, lr, A, la, B, lb, alpha=alpha, beta=0.0, prec=prec)
        torch.cuda.synchronize()
        want0 = O.dense_contract(None, lr, a.astype(np.complex128), la, b.astype(np.complex128),
                                 lb, alpha=alpha)
        assert relerr(R.to_numpy(), want0) < F32_RTOL


@pytest.mark.parametrize("cplx", [False, True], ids=["f32", "c64"])
def test_blocksparse_contract_f32(env, cplx):
    torch, itg, ND, O = env
    from collections import OrderedDict
    QN, Index = itg.QN, itg.Index
    rng = np.random.default_rng(51)
    sp = lambda s: [(QN(("Sz", q)), n) for q, n in s]
    i = Index(sp([(0, 30), (1, 20), (-1, 25)]), "i", dir=itg.Out)
    j = Index(sp([(0, 40), (1, 35), (-1, 30)]), "j", dir=itg.In)
    k = Index(sp([(0, 10), (2, 5), (-2, 7)]), "k", dir=itg.Out)
    dt = np.complex128 if cplx else np.float64
    A = itg.random_itensor(QN(("Sz", 0)), i, j, k, dtype=dt, rng=rng)
    B = itg.random_itensor(QN(("Sz", 0)), j.dag(), k.dag(), i.prime().dag(), dtype=dt, rng=rng)
    t32 = torch.complex64 if cplx else torch.float32
    A32 = itg.ITensor(ND._with_data(A.tensor, A.tensor.data.to(t32)), A.inds)
    B32 = itg.ITensor(ND._with_data(B.tensor, B.tensor.data.to(t32)), B.inds)
    C = A32 * B32
    torch.cuda.synchronize()
    assert C.tensor.data.dtype == t32

    def to_o(T):
        offs = OrderedDict((tuple(c - 1 for c in b), o) for b, o in T.tensor.blockoffsets.items())
        inds = tuple(O.Index([(O.QN(*[(n, v, m) for n, (v, m) in q.vals.items()]), d)
                              for q, d in x.space], tags=x.tags, plev=x.plev, dir=x.dir, id=x.id)
                     for x in T.inds)
        return O.BlockSparse(T.tensor.data.detach().cpu().numpy().astype(dt), offs, inds)

    OA, OB = to_o(A32), to_o(B32)
    la, lb = O.compute_contraction_labels(OA.inds, OB.inds)
    ref, _ = O.blocksparse_contract(OA, la, OB, lb, O.contract_labels(la, lb))
    assert [tuple(c - 1 for c in b) for b in C.tensor.blockoffsets] == list(ref.offsets)
    assert relerr(C.tensor.data.cpu().numpy(), ref.data) < F32_RTOL
    # mixed Float32 x Float64 promotes to Float64 (FP64 result)
    Cm = A32 * B
    torch.cuda.synchronize()
    assert Cm.tensor.data.dtype == (torch.complex128 if cplx else torch.float64)
    ref2, _ = O.blocksparse_contract(OA, la, to_o(B), lb, O.contract_labels(la, lb))
    assert relerr(Cm.tensor.data.cpu().numpy(), ref2.data) < FP64_RTOL


@pytest.mark.parametrize("cplx", [False, True], ids=["f32", "c64"])
def test_permute_f32_exact(env, cplx):
    torch, itg, ND, O = env
    rng = np.random.default_rng(52)
    a = _arr(rng, (7, 33, 5, 40), cplx)
    T = ND.DenseTensor.from_numpy(a)
    for perm in ((2, 1, 3, 4), (4, 3, 2, 1), (1, 3, 4, 2), (3, 1, 2, 4)):
        P = ND.permutedims(T, perm)
        torch.cuda.synchronize()
        assert np.array_equal(P.to_numpy(), np.transpose(a, [p - 1 for p in perm]))

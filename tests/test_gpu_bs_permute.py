"""Block-sparse permutedims on the device vs the oracle.

* permutedims(T, perm): every block permuted, block list kept (the common
  case), for every permutation of order-3 and order-4 QN tensors.
* permutedims!(R, T, perm, f) with f(r, t) = beta*r + alpha*t over the union
  of R's blocks and T's permuted blocks
  (NDTensors/src/blocksparse/blocksparsetensor.jl:834-881): R missing some
  of T's blocks and holding blocks T lacks, new all-zero blocks dropped,
  beta == 0 never reading R (NaN-filled R), real and complex.
Structure bit-exact, values exact (a permutation and one multiply-add per
element in FP64: relative error 1e-15 at most).
"""
import itertools
from collections import OrderedDict

import numpy as np
import pytest

from helpers import relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    from oracle import ndtensors_oracle as O
    torch.cuda.init()
    return torch, itg, ND, O


def _qn_tensor(itg, rng, n, cplx):
    while True:
        try:
            return _qn_tensor_try(itg, rng, n, cplx)
        except ValueError:  # no block has flux 0
            continue


def _qn_tensor_try(itg, rng, n, cplx):
    QN, Index = itg.QN, itg.Index
    inds = []
    for t in range(n):
        k = int(rng.integers(2, 4))
        space = [(QN(("Sz", q)), int(rng.integers(1, 9))) for q in range(-(k // 2), k - k // 2)]
        inds.append(Index(space, f"t{t}", dir=itg.Out if rng.random() < 0.5 else itg.In))
    return itg.random_itensor(QN(("Sz", 0)), *inds, dtype=np.complex128 if cplx else np.float64,
                              rng=rng)


def _oracle_bs(O, tensor, sectors):
    offs = OrderedDict((tuple(c - 1 for c in b), o) for b, o in tensor.blockoffsets.items())
    inds = tuple(O.Index([(O.QN(0), d) for d in bd]) for bd in sectors)
    return O.BlockSparse(tensor.data.detach().cpu().numpy(), offs, inds)


@pytest.mark.parametrize("n", [3, 4])
def test_bs_permutedims_every_perm(env, n):
    torch, itg, ND, O = env
    rng = np.random.default_rng(30 + n)
    A = _qn_tensor(itg, rng, n, cplx=False)
    T = A.tensor
    for perm in itertools.permutations(range(n)):
        P = ND.permutedims(T, [p + 1 for p in perm])
        torch.cuda.synchronize()
        want = np.transpose(T.to_numpy_dense(), perm)
        assert np.array_equal(P.to_numpy_dense(), want)
        assert list(P.blockoffsets) == [tuple(b[p] for p in perm) for b in T.blockoffsets]


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c128"])
def test_bs_permutedims_union_alpha_beta(env, cplx):
    torch, itg, ND, O = env
    rng = np.random.default_rng(41 if cplx else 40)
    done = 0
    while done < 6:
        n = int(rng.integers(3, 5))
        A = _qn_tensor(itg, rng, n, cplx)
        T = A.tensor
        perm = tuple(int(x) for x in rng.permutation(n))
        sectors = tuple(T.blockdims[p] for p in perm)
        blocks = [tuple(b[p] for p in perm) for b in T.blockoffsets]
        others = [b for b in itertools.product(*[range(1, len(s) + 1) for s in sectors])
                  if b not in blocks]
        if len(blocks) < 3 or not others:
            continue
        keep = [blocks[i] for i in rng.permutation(len(blocks))[:len(blocks) - 2]]
        keep.insert(int(rng.integers(0, len(keep) + 1)), others[int(rng.integers(0, len(others)))])
        if rng.random() < 0.5:
            # an all-zero block of T that R lacks: not inserted
            miss = [b for b in blocks if b not in keep][0]
            inv = np.argsort(perm)
            bt = tuple(miss[i] for i in inv)
            o = T.blockoffsets[bt]
            T.data[o:o + int(np.prod(T.blockdims_of(bt)))] = 0
        for alpha, beta in ((1.0, 0.0), (0.5, -2.0), (0.0, 1.5), (-1.25, 0.75)):
            if cplx:
                alpha, beta = alpha * (1 - 0.5j), beta * (0.25 + 1j)
            offs, nn = OrderedDict(), 0
            for b in keep:
                offs[b] = nn
                nn += int(np.prod([sectors[m][c - 1] for m, c in enumerate(b)]))
            r0 = rng.standard_normal(nn) + (1j * rng.standard_normal(nn) if cplx else 0)
            if beta == 0:
                r0 = np.full(nn, np.nan, dtype=r0.dtype)
            dt = torch.complex128 if cplx else torch.float64
            R = ND.BlockSparseTensor(torch.from_numpy(r0).to("cuda", dt), offs, sectors)
            OR = _oracle_bs(O, R, sectors)
            OT = _oracle_bs(O, T, T.blockdims)
            want = O.blocksparse_permutedims_(OR, OT, perm, alpha, beta)
            out = ND.permutedims_(R, T, [p + 1 for p in perm], alpha=alpha, beta=beta)
            torch.cuda.synchronize()
            assert out is R
            assert [tuple(c - 1 for c in b) for b in R.blockoffsets] == list(want.offsets)
            assert list(R.blockoffsets.values()) == list(want.offsets.values())
            got = R.data.cpu().numpy()
            assert got.shape == want.data.shape
            assert relerr(got, want.data) < 1e-15
        done += 1

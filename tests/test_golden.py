"""Golden fixtures (tests/golden/*.npz, written by make_golden.py from the
oracle in extended precision): the oracle must keep reproducing them (CPU)
and the device path must match them (GPU)."""
import hashlib
import os

import numpy as np
import pytest

from helpers import ROOT, relerr

GOLD = os.path.join(ROOT, "tests", "golden")


def c1_inputs():
    rng = np.random.Generator(np.random.PCG64(0))
    A = rng.standard_normal(128 * 2 * 128).reshape((128, 2, 128), order="F")
    L = rng.standard_normal(128 * 5 * 128).reshape((128, 5, 128), order="F")
    W = rng.standard_normal(5 * 2 * 2 * 5).reshape((5, 2, 2, 5), order="F")
    return A, L, W


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_c1_inputs_reproducible():
    g = np.load(os.path.join(GOLD, "golden_c1.npz"))
    A, L, W = c1_inputs()
    assert sha(A) == str(g["sha_A"]) and sha(L) == str(g["sha_L"]) and sha(W) == str(g["sha_W"])


def test_oracle_reproduces_c1():
    from oracle import ndtensors_oracle as O
    g = np.load(os.path.join(GOLD, "golden_c1.npz"))
    A, L, W = c1_inputs()
    X = O.dense_contract(None, (1, 2, 3, 4), L, (1, 2, -1), A, (-1, 3, 4))
    Y = O.dense_contract(None, (1, 4, 5, 6), X, (1, -1, -2, 4), W, (-1, -2, 5, 6))
    assert tuple(X.shape) == tuple(g["X_shape"]) and tuple(Y.shape) == tuple(g["Y_shape"])
    assert np.allclose(X.reshape(-1, order="F")[g["X_idx"]], g["X_val"], rtol=1e-13, atol=1e-12)
    assert np.allclose(Y.reshape(-1, order="F")[g["Y_idx"]], g["Y_val"], rtol=1e-13, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "split_int8"])
def test_device_matches_c1_golden(prec):
    import torch
    import itg
    from itg import ndtensors as ND
    g = np.load(os.path.join(GOLD, "golden_c1.npz"))
    A, L, W = c1_inputs()
    TA, TL, TW = (ND.DenseTensor.from_numpy(x) for x in (A, L, W))
    X = ND.contract(TL, (1, 2, -1), TA, (-1, 3, 4), prec=prec)
    Y = ND.contract(X, (1, -1, -2, 4), TW, (-1, -2, 5, 6), prec=prec)
    torch.cuda.synchronize()
    x = X.to_numpy().reshape(-1, order="F")
    y = Y.to_numpy().reshape(-1, order="F")
    assert relerr(x[g["X_idx"]], g["X_val"]) < 1e-12
    assert relerr(y[g["Y_idx"]], g["Y_val"]) < 1e-12
    assert abs(np.linalg.norm(y) - float(g["Y_norm"])) < 1e-12 * float(g["Y_norm"])


def heff32_tensors(itg, g):
    QN, Index = itg.QN, itg.Index
    bond = [(QN(("Sz", int(q))), int(d)) for q, d in zip(g["bond_q"], g["bond_d"])]
    phys = [(QN(("Sz", 1)), 1), (QN(("Sz", -1)), 1)]
    link = [(QN(("Sz", 0)), 1), (QN(("Sz", 2)), 1), (QN(("Sz", -2)), 1), (QN(("Sz", 0)), 2)]
    a, b = Index(bond, "a", dir=1), Index(bond, "b", dir=-1)
    s1, s2 = Index(phys, "s1", dir=1), Index(phys, "s2", dir=1)
    ap, bp = Index(bond, "a'", dir=1), Index(bond, "b'", dir=-1)
    w0, w1, w2 = (Index(link, t, dir=-1) for t in ("w", "w'", "w''"))
    s1p, s2p = Index(phys, "s1'", dir=1), Index(phys, "s2'", dir=1)
    inds = ((ap, w0, a.dag()), (a, s1, s2, b), (w0.dag(), s1.dag(), s1p, w1),
            (w1.dag(), s2.dag(), s2p, w2), (bp, w2.dag(), b.dag()))
    out = []
    for name, ix in zip(("L", "psi", "W1", "W2", "R"), inds):
        blocks = [tuple(int(c) + 1 for c in blk) for blk in g[name + "_blocks"]]
        out.append(itg.blocksparse_itensor(g[name + "_data"], blocks, ix))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "split_int8"])
def test_device_matches_heff_u1_chi32_golden(prec):
    import torch
    import itg
    g = np.load(os.path.join(GOLD, "golden_heff_u1_chi32.npz"))
    ts = heff32_tensors(itg, g)
    cur = ts[0]
    for T in ts[1:]:
        cur = itg.contract(cur, T, prec=prec)
    torch.cuda.synchronize()
    got_blocks = [tuple(c - 1 for c in b) for b in cur.tensor.blockoffsets]
    assert got_blocks == [tuple(b) for b in g["out_blocks"]]
    assert list(cur.tensor.blockoffsets.values()) == list(g["out_offsets"])
    assert relerr(cur.tensor.data.cpu().numpy(), g["out_data"]) < 1e-12

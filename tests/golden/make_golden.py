#!/usr/bin/env python3
"""Writes the golden fixtures of this directory from the CPU oracle.

  python tests/golden/make_golden.py

golden_c1.npz   dense case: A[a=128,s=2,b=128] * L[a'=128,w=5,a=128] (over a),
                then * W[w=5,s=2,s'=2,w'=5] (over w, s); N(0,1) entries from
                numpy PCG64(0) drawn A, L, W in column-major order. Stores the
                SHA-256 of the inputs, the output norms and 512 sampled output
                entries computed in extended precision.
golden_heff_u1_chi32.npz
                the U(1) block-sparse two-site network at chi = 32 (Sz
                sectors -12..12 with sizes n_q ~ exp(-q^2/32), physical
                +-1, MPO link [0,+2,-2,0(x2)]), inputs and the full output
                (blocks, offsets, data) in extended precision.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import ndtensors_oracle as O  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def c1():
    rng = np.random.Generator(np.random.PCG64(0))
    A = rng.standard_normal(128 * 2 * 128).reshape((128, 2, 128), order="F")
    L = rng.standard_normal(128 * 5 * 128).reshape((128, 5, 128), order="F")
    W = rng.standard_normal(5 * 2 * 2 * 5).reshape((5, 2, 2, 5), order="F")
    X = O.dense_contract(None, (1, 2, 3, 4), L, (1, 2, -1), A, (-1, 3, 4), exact=True)
    Y = O.dense_contract(None, (1, 4, 5, 6), X, (1, -1, -2, 4), W, (-1, -2, 5, 6), exact=True)
    sel = np.random.Generator(np.random.PCG64(123))
    ix = sel.integers(0, X.size, 512)
    iy = sel.integers(0, Y.size, 512)
    np.savez(os.path.join(HERE, "golden_c1.npz"),
             sha_A=sha(A), sha_L=sha(L), sha_W=sha(W),
             X_shape=X.shape, Y_shape=Y.shape,
             X_norm=float(np.linalg.norm(X.astype(np.float64))),
             Y_norm=float(np.linalg.norm(Y.astype(np.float64))),
             X_idx=ix, X_val=X.reshape(-1, order="F")[ix].astype(np.float64),
             Y_idx=iy, Y_val=Y.reshape(-1, order="F")[iy].astype(np.float64))


def sectors(chi):
    qs = np.arange(-12, 13)
    w = np.exp(-qs ** 2 / 32.0)
    n = np.rint(chi * w / w.sum()).astype(int)
    n[12] = chi - (n.sum() - n[12])
    return [(int(q), int(x)) for q, x in zip(qs, n) if x > 0]


def heff32():
    chi = 32
    bond = [(O.QN(("Sz", q)), d) for q, d in sectors(chi)]
    phys = [(O.QN(("Sz", 1)), 1), (O.QN(("Sz", -1)), 1)]
    link = [(O.QN(("Sz", 0)), 1), (O.QN(("Sz", 2)), 1), (O.QN(("Sz", -2)), 1), (O.QN(("Sz", 0)), 2)]
    a, b = O.Index(bond, "a", dir=1), O.Index(bond, "b", dir=-1)
    s1, s2 = O.Index(phys, "s1", dir=1), O.Index(phys, "s2", dir=1)
    ap, bp = O.Index(bond, "a'", dir=1), O.Index(bond, "b'", dir=-1)
    w0, w1, w2 = (O.Index(link, t, dir=-1) for t in ("w", "w'", "w''"))
    s1p, s2p = O.Index(phys, "s1'", dir=1), O.Index(phys, "s2'", dir=1)
    z = O.QN(("Sz", 0))
    rng = np.random.Generator(np.random.PCG64(0))
    ts = [O.random_blocksparse(rng, z, inds) for inds in
          ((ap, w0, a.dag()), (a, s1, s2, b), (w0.dag(), s1.dag(), s1p, w1),
           (w1.dag(), s2.dag(), s2p, w2), (bp, w2.dag(), b.dag()))]
    cur = ts[0]
    for T in ts[1:]:
        la, lb = O.compute_contraction_labels(cur.inds, T.inds)
        cur, _ = O.blocksparse_contract(cur, la, T, lb, exact=True)
    out = {"bond_q": [q for q, _ in sectors(chi)], "bond_d": [d for _, d in sectors(chi)]}
    for name, T in zip(("L", "psi", "W1", "W2", "R"), ts):
        out[name + "_data"] = T.data
        out[name + "_blocks"] = np.array(list(T.offsets.keys()), dtype=np.int64)
    out["out_data"] = cur.data.astype(np.float64)
    out["out_blocks"] = np.array(list(cur.offsets.keys()), dtype=np.int64)
    out["out_offsets"] = np.array(list(cur.offsets.values()), dtype=np.int64)
    np.savez(os.path.join(HERE, "golden_heff_u1_chi32.npz"), **out)


if __name__ == "__main__":
    c1()
    heff32()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))

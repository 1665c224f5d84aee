"""Shared helpers of the test-suite (host side only)."""
import itertools
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "itensors.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

# FP64 paths: products of K terms in a different order than numpy/BLAS; the
# relative Frobenius difference stays ~1e-15. 1e-12 leaves ample margin and is
# the bound the split-precision forms must also meet.
FP64_RTOL = 1e-12


def relerr(x, ref):
    x = np.asarray(x)
    ref = np.asarray(ref)
    n = np.linalg.norm(ref.astype(np.complex128).reshape(-1))
    d = np.linalg.norm((x.astype(np.complex128) - ref.astype(np.complex128)).reshape(-1))
    return float(d / n) if n > 0 else float(d)


def rand(rng, shape, cplx=False):
    x = rng.standard_normal(shape)
    if cplx:
        x = x + 1j * rng.standard_normal(shape)
    return x


def labels_for(ndim_a, ndim_b, contracted_pairs):
    """Labels (ncon style) for A (ndim_a modes) and B with the given
    (posA, posB) contracted pairs: contracted -1, -2, ...; free positive."""
    la = [0] * ndim_a
    lb = [0] * ndim_b
    for k, (i, j) in enumerate(contracted_pairs):
        la[i] = lb[j] = -(k + 1)
    u = len(contracted_pairs)
    for i in range(ndim_a):
        if la[i] == 0:
            u += 1
            la[i] = u
    for j in range(ndim_b):
        if lb[j] == 0:
            u += 1
            lb[j] = u
    return tuple(la), tuple(lb)


def all_perms(n):
    return list(itertools.permutations(range(n)))

"""A fixed contraction sequence with its plans and intermediates kept on the
device, replayable and capturable in a HIP graph.

`contract(As; sequence)` (src/tensor_operations/tensor_algebra.jl:121-159)
recomputes labels, plans and output allocations on every call. Applying the
same network many times (an eigensolver applying H_eff) repeats identical
work, so this object does it once: the first call lowers the left-associative
fold to a list of (plan, operand, output buffer) steps, and `run()` then only
enqueues the kernels. `capture()` records those launches in a graph
(torch.cuda.CUDAGraph over the handle's stream), removing launch overhead.
"""
from __future__ import annotations

import math

from . import ndtensors as ND
from .itensors import ITensor, compute_contraction_labels


class ContractionSequence:
    def __init__(self, tensors, *, prec=None):
        import torch

        self.tensors = list(tensors)
        self.prec = prec
        self.steps = []
        cur = self.tensors[0]
        for B in self.tensors[1:]:
            lA, lB = compute_contraction_labels(cur.inds, B.inds)
            lR = ND.contract_labels(lA, lB)
            indsR = tuple(cur.inds[lA.index(x)] if x in lA else B.inds[lB.index(x)] for x in lR)
            if cur.hasqns():
                p = ND.bs_plan(cur.tensor, lA, B.tensor, lB, lR)
                bd = []
                for r in lR:
                    bd.append(cur.tensor.blockdims[lA.index(r)] if r in lA else B.tensor.blockdims[lB.index(r)])
                data = torch.empty(p.nnzR, dtype=cur.eltype, device=cur.tensor.data.device)
                R = ND.BlockSparseTensor(data, dict(zip(p.blocksR, p.offsR)), bd)
                self.steps.append(("bs", p, cur.tensor, B.tensor, R, (lA, lB, lR)))
            else:
                dims = tuple(i.dim for i in indsR)
                R = ND.DenseTensor(torch.empty(math.prod(dims), dtype=cur.eltype,
                                               device=cur.tensor.data.device), dims)
                self.steps.append(("dense", None, cur.tensor, B.tensor, R, (lA, lB, lR)))
            cur = ITensor(R, indsR)
        self.result = cur

    def run_step(self, i, prec=None):
        kind, p, A, B, R, (lA, lB, lR) = self.steps[i]
        prec = prec or self.prec
        if kind == "bs":
            if p.npairs:
                p.execute(R.data, A.data, B.data, prec=prec)
        else:
            ND.contract_(R, lR, A, lA, B, lB, 1.0, 0.0, prec=prec)

    def run(self, prec=None):
        for i in range(len(self.steps)):
            self.run_step(i, prec)
        return self.result

    def autotune(self, prec=None):
        for kind, p, A, B, R, _l in self.steps:
            if kind == "bs" and p.npairs:
                p.autotune(R.data, A.data, B.data, prec=prec or self.prec)

    def capture(self, prec=None):
        """Record one run() in a graph; replay with graph.replay()."""
        import torch

        self.run(prec)  # warm: workspace sized, device tables uploaded
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run(prec)
        return g

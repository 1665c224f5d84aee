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

from . import _lib
from . import ndtensors as ND
from ._lib import ItgUnsupported
from .itensors import ITensor, compute_contraction_labels


class ContractionSequence:
    """fuse: run consecutive block-sparse steps that admit it (ND.BSChain:
    two fiber-path contractions whose intermediate is only read by the
    next) as one kernel that never stores the intermediate. The result is
    bit-identical either way; the skipped intermediates are left unwritten.

    fixed: positions in `tensors` whose data stays the same from one run to
    the next (the environments L and R of an effective Hamiltonian, (0, 4) for
    the two-site network). In split precision their digit planes are made
    by the first run and kept by the plans (ITG_FLAG_KEEP_SLICES_*), and later
    runs reuse them (ITG_FLAG_REUSE_SLICES_*) instead of slicing again; call
    invalidate_fixed() after changing such a tensor's data. FP64 runs are
    unaffected.

    reorder: let an intermediate produced by a fiber or strided step hold the
    indices its consumer contracts last (in the consumer's order) when the
    consumer is a grouped GEMM that would otherwise permute it first (the Z
    of the two-site network before its product with R). Block-sparse steps
    ask their plans (itg_bs_plan_info); a dense step with a short contracted
    extent is reordered whenever a next step exists. Only the order of
    intermediates changes: the result's index order and values are the ones
    of the left-to-right fold of tensor_algebra.jl:121-130 (dense values to
    rounding: a GEMM that reads Z in place may sum in another order than
    one that reads a permuted copy)."""

    def __init__(self, tensors, *, prec=None, fuse=True, fixed=(), reorder=True):
        self.tensors = list(tensors)
        self.prec = prec
        self.steps = []
        self.chains = {}  # first step index -> BSChain over it and the next
        # step -> slice-keeping flag bits for its fixed operands: position 0
        # is step 0's T1, position j >= 1 is step j-1's T2
        self._fixed_bits = {}
        for j in fixed:
            step, op = (0, 1) if j == 0 else (j - 1, 2)
            self._fixed_bits[step] = self._fixed_bits.get(step, 0) | op
        self._kept = set()  # precision modes whose fixed slices are kept
        cur = self.tensors[0]
        nxt = self.tensors[2:] + [None]
        for k, B in enumerate(self.tensors[1:]):
            a_it = cur
            step, cur = self._build_step(a_it, B)
            # An intermediate's mode order is internal (only the result's is
            # the reference's): when the next step is a grouped GEMM that
            # would permute this operand, and this step is not a GEMM (whose
            # output order could then need a permutation of its own), put the
            # modes the next step contracts last, in the order the next
            # operand holds them. The modes that stay free keep their order,
            # so every later output, the result included, is unchanged.
            C = nxt[k]
            if reorder and C is not None and step[0] == "bs" and step[1].npairs and \
                    step[1].info()["strategy"] != "gemm":
                probe = self._probe(cur, C)
                if probe is not None and probe["strategy"] == "gemm" and probe["perm1"]:
                    order = self._order_for_next(cur.inds, C.inds)
                    step2, cur2 = self._build_step(a_it, B, out_inds=order)
                    probe2 = self._probe(cur2, C)
                    # kept only if this step still runs the same way
                    if probe2 is not None and not probe2["perm1"] and \
                            step2[1].info()["strategy"] == step[1].info()["strategy"]:
                        step, cur = step2, cur2
            elif reorder and C is not None and step[0] == "dense" and \
                    ND.gemm_backend() == "Auto" and \
                    _contracted_extent(a_it, step[5][0]) < BatchedSequence._GEMM_MIN_K:
                # dense steps have no plan to ask: a step with a short
                # contracted extent runs on the strided or fiber kernels
                # (any output order costs the same), so its output takes the
                # order a GEMM-shaped next step reads in place
                order = self._order_for_next(cur.inds, C.inds)
                if order != tuple(cur.inds):
                    step, cur = self._build_step(a_it, B, out_inds=order)
            self.steps.append(step)
        self.result = cur
        i = 0
        while fuse and i + 1 < len(self.steps):
            s1, s2 = self.steps[i], self.steps[i + 1]
            if s1[0] == "bs" and s2[0] == "bs" and s1[1].npairs and s2[1].npairs:
                try:
                    self.chains[i] = ND.BSChain(s1[1], s2[1])
                    i += 2
                    continue
                except ItgUnsupported:
                    pass
            i += 1

    @staticmethod
    def _build_step(A, B, out_inds=None):
        """The step A * B: (step tuple, output ITensor). out_inds: the output's
        index order (default: the reference's, A's free indices then B's)."""
        import torch

        lA, lB = compute_contraction_labels(A.inds, B.inds)
        lR = ND.contract_labels(lA, lB)
        if out_inds is not None:
            pos = {}
            for x in lR:
                pos[A.inds[lA.index(x)] if x in lA else B.inds[lB.index(x)]] = x
            lR = tuple(pos[i] for i in out_inds)
        indsR = tuple(A.inds[lA.index(x)] if x in lA else B.inds[lB.index(x)] for x in lR)
        if A.hasqns():
            p = ND.bs_plan(A.tensor, lA, B.tensor, lB, lR)
            bd = [A.tensor.blockdims[lA.index(r)] if r in lA else B.tensor.blockdims[lB.index(r)]
                  for r in lR]
            data = torch.empty(p.nnzR, dtype=A.eltype, device=A.tensor.data.device)
            R = ND.BlockSparseTensor._checked(data, dict(zip(p.blocksR, p.offsR)), bd, p.nnzR)
            return ("bs", p, A.tensor, B.tensor, R, (lA, lB, lR)), ITensor(R, indsR)
        dims = tuple(i.dim for i in indsR)
        R = ND.DenseTensor(torch.empty(math.prod(dims), dtype=A.eltype,
                                       device=A.tensor.data.device), dims)
        return ("dense", None, A.tensor, B.tensor, R, (lA, lB, lR)), ITensor(R, indsR)

    @staticmethod
    def _probe(A, B):
        """How A * B would run (the plan's info), or None without QNs."""
        if not A.hasqns():
            return None
        lA, lB = compute_contraction_labels(A.inds, B.inds)
        p = ND.bs_plan(A.tensor, lA, B.tensor, lB, ND.contract_labels(lA, lB))
        return p.info() if p.npairs else None

    @staticmethod
    def _order_for_next(inds, next_inds):
        """inds with the indices `next_inds` shares moved last, in next_inds'
        order; the others keep their order."""
        shared = [i for i in next_inds if i in inds]
        return tuple(i for i in inds if i not in shared) + tuple(shared)

    def run_chain(self, i):
        """Steps i and i + 1 as one fused launch."""
        _k, _p, X, B1, _Y, _l = self.steps[i]
        _k, _p, _Y2, B2, Z, _l = self.steps[i + 1]
        self.chains[i].execute(Z.data, X.data, B1.data, B2.data)

    def _flags(self, i, prec):
        ops = self._fixed_bits.get(i, 0)
        if not ops:
            return 0
        reuse = (prec or ND.precision()) in self._kept
        f = 0
        if ops & 1:
            f |= _lib.ITG_FLAG_REUSE_SLICES_1 if reuse else _lib.ITG_FLAG_KEEP_SLICES_1
        if ops & 2:
            f |= _lib.ITG_FLAG_REUSE_SLICES_2 if reuse else _lib.ITG_FLAG_KEEP_SLICES_2
        return f

    def invalidate_fixed(self):
        """The fixed tensors' data changed: the next run slices them again."""
        self._kept.clear()

    def run_step(self, i, prec=None):
        kind, p, A, B, R, (lA, lB, lR) = self.steps[i]
        prec = prec or self.prec
        if kind == "bs":
            if p.npairs:
                p.execute(R.data, A.data, B.data, prec=prec, flags=self._flags(i, prec))
        else:
            ND.contract_(R, lR, A, lA, B, lB, 1.0, 0.0, prec=prec)

    def run(self, prec=None):
        i = 0
        while i < len(self.steps):
            if i in self.chains:
                self.run_chain(i)
                i += 2
            else:
                self.run_step(i, prec)
                i += 1
        if self._fixed_bits:
            self._kept.add(prec or self.prec or ND.precision())
        return self.result

    def autotune(self, prec=None):
        for kind, p, A, B, R, _l in self.steps:
            if kind == "bs" and p.npairs:
                p.autotune(R.data, A.data, B.data, prec=prec or self.prec)

    def capture(self, prec=None):
        """Record one run() in a graph; replay with graph.replay()."""
        return _capture(self.run, prec)


def _capture(run, prec):
    import torch

    run(prec)  # warm: workspace sized, device tables uploaded
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run(prec)
    return g


def _contracted_extent(A, lA):
    """Product of the dimensions of A's modes that a step contracts (labels
    lA of A's modes, contracted ones negative)."""
    return math.prod(A.inds[i].dim for i, x in enumerate(lA) if x < 0)


class BatchPlan:
    """libitg batch plan: nbatch dense contractions with one label pattern,
    bound to their device buffers, executed as grouped launches."""

    def __init__(self, A_list, lA, B_list, lB, R_list, lR):
        import ctypes
        n = len(A_list)
        dev = R_list[0].data.device
        self.h = ND._handle(dev)
        keep = []

        def arr64(v):
            a = _lib.i64(v)
            keep.append(a)
            return _lib.ptr64(a)[0]

        def ptrs(ts):
            a = (ctypes.c_void_p * n)(*[t.data.data_ptr() for t in ts])
            keep.append(a)
            return a

        plan = ctypes.c_void_p()
        st = _lib.lib().itg_batch_plan_create(
            self.h.h, ND._dtype_code(R_list[0].eltype), n,
            len(lR), arr64([d for t in R_list for d in t.dims]), arr64(lR), ptrs(R_list),
            len(lA), arr64([d for t in A_list for d in t.dims]), arr64(lA), ptrs(A_list),
            len(lB), arr64([d for t in B_list for d in t.dims]), arr64(lB), ptrs(B_list),
            ctypes.byref(plan))
        self.h.check(st)
        self.plan = plan
        self.dev = dev
        # the plan holds raw device addresses: the tensors must outlive it
        # (a graph captured from execute() replays against these addresses)
        self.tensors = (list(A_list), list(B_list), list(R_list))

    def execute(self, prec=None, alpha=1.0):
        h = ND._handle(self.dev)
        a, _k = _lib.scalar2(alpha)
        h.check(_lib.lib().itg_batch_execute(h.h, self.plan, a, _lib.PREC[prec or ND.precision()]))

    def __del__(self):
        # no import here: at interpreter exit the import system is gone
        try:
            if getattr(self, "plan", None) and _lib._lib is not None:
                _lib._lib.itg_batch_plan_destroy(self.plan)
                self.plan = None
        except Exception:
            pass


class BatchedSequence:
    """Many independent dense networks with the same index pattern (each with
    its own dimensions) contracted left to right, one grouped launch per
    kernel of each step for the whole batch. The intermediates of a step
    share one allocation.

    reorder: an intermediate made by a step with a short contracted extent
    (the strided or fiber kernels, which write any mode order at the same
    cost) holds the indices the next step contracts last, in the order the
    next operand holds them, so that a GEMM-shaped next step reads it in
    place instead of permuting it (the Z of the two-site network before its
    product with R). The result's index order is the fold's either way."""

    # contracted extent below which the planner's automatic choice is not the
    # grouped GEMM (the library's kBackendAuto rule, flop-weighted mean
    # contracted extent >= 24)
    _GEMM_MIN_K = 24

    def __init__(self, networks, *, prec=None, reorder=True):
        import torch

        self.prec = prec
        self.plans = []
        self.arenas = []  # one per step; every plan points into these
        cur = [net[0] for net in networks]
        nsteps = len(networks[0]) - 1
        for s in range(nsteps):
            Bs = [net[s + 1] for net in networks]
            lA, lB = compute_contraction_labels(cur[0].inds, Bs[0].inds)
            lR = ND.contract_labels(lA, lB)
            if reorder and s + 1 < nsteps and ND.gemm_backend() == "Auto":
                if _contracted_extent(cur[0], lA) < self._GEMM_MIN_K:
                    out0 = tuple(cur[0].inds[lA.index(x)] if x in lA else Bs[0].inds[lB.index(x)]
                                 for x in lR)
                    order = ContractionSequence._order_for_next(out0, networks[0][s + 2].inds)
                    lab = dict(zip(out0, lR))
                    lR = tuple(lab[i] for i in order)
            outs, sizes, dims_all, inds_all = [], [], [], []
            for A, B in zip(cur, Bs):
                la, lb = compute_contraction_labels(A.inds, B.inds)
                if (la, lb) != (lA, lB):
                    raise ValueError("networks of a batch must share one index pattern")
                inds = tuple(A.inds[la.index(x)] if x in la else B.inds[lb.index(x)] for x in lR)
                dims = tuple(i.dim for i in inds)
                dims_all.append(dims)
                inds_all.append(inds)
                sizes.append(math.prod(dims))
            arena = torch.empty(sum(sizes), dtype=cur[0].eltype, device=cur[0].tensor.data.device)
            self.arenas.append(arena)
            o = 0
            for dims, n in zip(dims_all, sizes):
                outs.append(ND.DenseTensor(arena[o:o + n], dims))
                o += n
            self.plans.append(BatchPlan([c.tensor for c in cur], lA, [b.tensor for b in Bs], lB, outs, lR))
            cur = [ITensor(R, inds) for R, inds in zip(outs, inds_all)]
        self.results = cur

    def run(self, prec=None):
        for p in self.plans:
            p.execute(prec or self.prec)
        return self.results

    def capture(self, prec=None):
        return _capture(self.run, prec)

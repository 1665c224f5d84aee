"""CPU oracle for the NDTensors contraction path -- TEST INFRASTRUCTURE ONLY.

This module is a from-scratch numpy restatement of what ITensors.jl /
NDTensors.jl compute on the `*` / `contract` path. It exists to CHECK the
MI355X backend: only the test-suite, the smoke check and the CPU-baseline leg
of the benchmark may import it. The product (the `itg` package and libitg.so)
never imports, calls or links anything in this directory.

Each function cites the reference file:line it restates (paths relative to the
ITensors.jl source tree, v0.9.27 / NDTensors v0.4.27).

Pinning status (what checks each function; see tests/test_oracle.py):
  qn_add / qn_eq / flux / nzblocks   -- known-answer tests from the reference's
                                        test_qnitensor.jl:564-583, :937-947,
                                        brute-force flux invariants
  compute_contraction_labels         -- closed-form rules + arrow-error KAT
                                        (test_qnitensor.jl:1799-1805)
  contract_labels / contract_inds    -- closed-form rules, brute force vs einsum
  contract_blockoffsets_sequential   -- brute force over all block pairs, flux
                                        invariant flux(C)=flux(A)+flux(B), the
                                        zero-block case (test_threading.jl:60-78),
                                        densify-and-compare
  contract_blockoffsets_threaded     -- the same pairs as the sequential plan;
                                        equal to it when T1 has more blocks;
                                        otherwise the sequential plan of the
                                        swapped product T2*T1 with each pair
                                        swapped back; offsets by first
                                        appearance; contraction values equal
                                        the sequential ones and the
                                        reference's threading KATs
                                        (test_threading.jl:30-78)
  dense_contract                     -- textbook reductions: matrix product vs
                                        explicit loops, outer product, scalar
                                        results, alpha/beta + NaN overwrite
                                        (test_dense.jl:237-287)
  blocksparse_contract               -- densify(result) == dense_contract(
                                        densify(A), densify(B)) on random
                                        structures; per-pair factors: unit
                                        factors change nothing, one-hot
                                        factors give one pair's product
                                        (densified, einsum) and their sum
                                        is the whole contraction
  rrule_contract                     -- central finite differences
  blocksparse_permutedims_           -- densified: dense(R') == beta*dense(R) +
                                        alpha*transpose(dense(T)) on random QN
                                        structures (R missing some of T's
                                        blocks and holding blocks T lacks);
                                        block order = R's then T's new ones;
                                        all-zero new blocks dropped
  combined_index /
  combined_position_map              -- hand-derived known answer (mixed arrows,
                                        uneven sectors: order, merge, sizes,
                                        element positions) and the reference's
                                        one-index merge (test_qncombiner.jl:4-13),
                                        arrow keyword (test_qnitensor.jl:603-625)
  combine_blocksparse /
  uncombine_blocksparse              -- combine-then-uncombine identity over every
                                        subset and order of order-3/4 tensors
                                        (test_qnitensor.jl:666-784), nnz kept,
                                        per-block flux kept, the two-QN-name
                                        example (test_qnitensor.jl:786-828)
  DiagBlockSparse / nzdiagblocks /
  diagblocksparse_contract           -- closed forms: delta replaces an index,
                                        U * diag(s) scales columns, the fully
                                        contracted case is sum_j d_j A[j, j];
                                        the reference's cases
                                        (test_diagblocksparse.jl:31-78: square
                                        and non-square blocks, the off-diagonal
                                        error); densify-and-einsum on random QN
                                        structures (tests/test_oracle_diag.py)
"""
from __future__ import annotations

import itertools
from collections import OrderedDict

import numpy as np

# ---------------------------------------------------------------- QN algebra
# A QN is a dict {name: (value, modulus)}; a missing name means value 0.
# Restates src/lib/QuantumNumbers/src/qn.jl:6-30 (up to 4 named values, an
# optional modulus) and :178-207 (+ adds matching names, missing = 0).


def QN(*vals):
    """QN(v) (unnamed), QN(("Sz", v)), QN(("P", v, 2), ("Sz", v)) ...

    qn.jl:23-75: a bare integer is an unnamed value; tuples are
    (name, value[, modulus]).
    """
    out = {}
    if len(vals) == 1 and isinstance(vals[0], (int, np.integer)):
        vals = (("", int(vals[0])),)
    if len(vals) == 2 and isinstance(vals[0], (int, np.integer)) and isinstance(vals[1], (int, np.integer)):
        vals = (("", int(vals[0]), int(vals[1])),)
    for v in vals:
        name, val = v[0], int(v[1])
        mod = int(v[2]) if len(v) > 2 else 1
        if mod > 1:
            val %= mod
        out[name] = (val, mod)
    return out


def qn_add(a, b):
    """qn.jl:178-207 -- elementwise addition of matching names (mod m)."""
    out = dict(a)
    for name, (v, m) in b.items():
        if name in out:
            va, ma = out[name]
            if ma != m:
                raise ValueError("QNVals must have same modulus")
            s = va + v
            out[name] = (s % m if m > 1 else s, m)
        else:
            out[name] = (v, m)
    return out


def qn_scale(d, a):
    """Arrow * QN (qn.jl:156-162): multiply every value by the arrow (+1/-1)."""
    return {n: ((d * v) % m if m > 1 else d * v, m) for n, (v, m) in a.items()}


def qn_eq(a, b):
    """Equality with missing names treated as zero (qn.jl:248-280)."""
    for name in set(a) | set(b):
        va = a.get(name, (0, 1))[0]
        vb = b.get(name, (0, 1))[0]
        if va != vb:
            return False
    return True


# ------------------------------------------------------------------- Index
class Index:
    """Index identity (id, plev, tags) + space + arrow (src/index.jl:24-33).

    `space` is an int (dense) or a list of (QN, dim) sectors (QNIndex,
    src/qn/qnindex.jl:6-8). dir: +1 = Out, -1 = In, 0 = Neither; QN indices
    default to Out (qnindex.jl:128).
    """

    _next_id = [1]

    def __init__(self, space, tags="", plev=0, dir=None, id=None):
        self.space = space
        self.tags = tags
        self.plev = plev
        if id is None:
            id = Index._next_id[0]
            Index._next_id[0] += 1
        self.id = id
        if dir is None:
            dir = 1 if self.hasqns else 0
        self.dir = dir

    @property
    def hasqns(self):
        return not isinstance(self.space, (int, np.integer))

    @property
    def dim(self):
        return int(self.space) if not self.hasqns else sum(d for _, d in self.space)

    def nblocks(self):
        return 1 if not self.hasqns else len(self.space)

    def blockdim(self, b):
        return self.dim if not self.hasqns else self.space[b][1]

    def qn(self, b):
        return self.space[b][0]

    def __eq__(self, other):
        # src/index.jl:354-355: id, plev and tags; the arrow is not compared
        return (self.id, self.plev, self.tags) == (other.id, other.plev, other.tags)

    def __hash__(self):
        return hash((self.id, self.plev, self.tags))

    def dag(self):
        """src/index.jl:362 -- flip the arrow, keep identity."""
        return Index(self.space, self.tags, self.plev, -self.dir, self.id)

    def prime(self, n=1):
        return Index(self.space, self.tags, self.plev + n, self.dir, self.id)

    def __repr__(self):
        return f"Index(id={self.id}, dim={self.dim}, dir={self.dir}, plev={self.plev})"


def compute_contraction_labels(Ais, Bis):
    """src/indexset.jl:672-707.

    Every (i, j) with Ais[i] == Bis[j] gets the next negative label
    -1, -2, ... in (i outer, j inner) order; QN indices that match must have
    opposite arrows. Free labels are then numbered ncont+1, ... over A, then B.
    """
    have_qns = all(i.hasqns for i in Ais) and all(j.hasqns for j in Bis)
    la = [0] * len(Ais)
    lb = [0] * len(Bis)
    ncont = 0
    for i, a in enumerate(Ais):
        for j, b in enumerate(Bis):
            if a == b:
                if have_qns and a.dir != -b.dir:
                    raise ValueError("QN indices must have opposite direction to contract")
                ncont += 1
                la[i] = lb[j] = -ncont
    u = ncont
    for i in range(len(la)):
        if la[i] == 0:
            u += 1
            la[i] = u
    for j in range(len(lb)):
        if lb[j] == 0:
            u += 1
            lb[j] = u
    return tuple(la), tuple(lb)


def contract_labels(l1, l2):
    """NDTensors/src/tensoroperations/contraction_logic.jl:5-34: output labels
    are the positive labels of T1 in order, then those of T2."""
    return tuple(x for x in l1 if x > 0) + tuple(x for x in l2 if x > 0)


def contract_inds(is1, l1, is2, l2, lR):
    """contraction_logic.jl:36-62: each output label takes the index of the
    tensor (T1 first) that carries it."""
    out = []
    for r in lR:
        if r in l1:
            out.append(is1[l1.index(r)])
        else:
            out.append(is2[l2.index(r)])
    return tuple(out)


# --------------------------------------------------------- block structure
def eachblock(inds):
    """NDTensors/src/blocksparse/blockdims.jl:104-106: CartesianIndices over
    the block grid, i.e. column-major (first mode fastest). 0-based here."""
    shape = [i.nblocks() for i in inds]
    for rev in itertools.product(*[range(n) for n in reversed(shape)]):
        yield tuple(reversed(rev))


def flux_of_block(inds, block):
    """src/indexset.jl:876-883 with flux(i, b) = dir(i) * qn(i, b)
    (src/qn/qnindex.jl:242)."""
    tot = {}
    for i, b in zip(inds, block):
        tot = qn_add(tot, qn_scale(i.dir, i.qn(b)))
    return tot


def nzblocks(flux, inds):
    """src/qn/qnindexset.jl:9-18: blocks (eachblock order) whose flux equals
    the requested one."""
    return [b for b in eachblock(inds) if qn_eq(flux_of_block(inds, b), flux)]


def blockdims(inds, block):
    return tuple(i.blockdim(b) for i, b in zip(inds, block))


def blockoffsets(blocks, inds):
    """NDTensors/src/blocksparse/blockoffsets.jl:70-79: offsets accumulate in
    the given block order."""
    offs = OrderedDict()
    nnz = 0
    for b in blocks:
        offs[b] = nnz
        nnz += int(np.prod(blockdims(inds, b)))
    return offs, nnz


class BlockSparse:
    """A BlockSparseTensor: flat data + ordered {block: offset} + inds
    (NDTensors/src/blocksparse/blocksparse.jl:5-13). Blocks are 0-based
    tuples here (the reference's Block{N} is the same tuple plus one)."""

    def __init__(self, data, offsets, inds):
        self.data = data
        self.offsets = offsets
        self.inds = tuple(inds)

    def block(self, b):
        """blockview (blocksparsetensor.jl:327-352): a column-major view."""
        dims = blockdims(self.inds, b)
        o = self.offsets[b]
        n = int(np.prod(dims))
        return self.data[o:o + n].reshape(dims, order="F")

    def dense(self):
        """dense(T) (blocksparsetensor.jl:357-368): scatter blocks into a
        zero array."""
        shape = tuple(i.dim for i in self.inds)
        out = np.zeros(shape, dtype=self.data.dtype)
        starts = [np.concatenate([[0], np.cumsum([i.blockdim(b) for b in range(i.nblocks())])]) for i in self.inds]
        for b in self.offsets:
            sl = tuple(slice(starts[m][bm], starts[m][bm + 1]) for m, bm in enumerate(b))
            out[sl] = self.block(b)
        return out


def random_blocksparse(rng, flux, inds, dtype=np.float64):
    """ITensor(ElT, flux, inds) + randn! (src/qn/qnitensor.jl:158-166,
    :387-393): allowed blocks in eachblock order, data filled in flat
    storage order from the generator (N(0,1) real; for complex the real and
    imaginary parts are each N(0,1))."""
    blocks = nzblocks(flux, inds)
    if not blocks:
        raise ValueError("ITensor with flux resulted in no allowed blocks")
    offs, nnz = blockoffsets(blocks, inds)
    data = rng.standard_normal(nnz)
    if np.dtype(dtype).kind == "c":
        data = data + 1j * rng.standard_normal(nnz)
    return BlockSparse(data.astype(dtype), offs, inds)


# ------------------------------------------------------- dense contraction
def _einsum_spec(l1, l2, lR):
    letters = {}

    def let(x):
        if x not in letters:
            letters[x] = chr(ord("a") + len(letters)) if len(letters) < 26 else chr(ord("A") + len(letters) - 26)
        return letters[x]

    return "".join(let(x) for x in l1) + "," + "".join(let(x) for x in l2) + "->" + "".join(let(x) for x in lR)


def dense_contract(R, lR, A, lA, B, lB, alpha=1.0, beta=0.0, exact=False):
    """R = alpha * A*B + beta * R in the label convention
    (NDTensors/src/dense/tensoralgebra/contract.jl:160-216).

    beta == 0 overwrites R without reading it (NaN-safe,
    NDTensors/test/test_dense.jl:237-259). `R` may be None when beta == 0.
    `exact=True` evaluates in extended precision (numpy longdouble, 63-bit
    mantissa) for use as a tight value oracle; otherwise float64 BLAS.
    Returns the new R (a new array).
    """
    spec = _einsum_spec(tuple(lA), tuple(lB), tuple(lR))
    cplx = np.iscomplexobj(A) or np.iscomplexobj(B) or np.iscomplexobj(alpha) or (R is not None and np.iscomplexobj(R))
    if exact:
        dt = np.clongdouble if cplx else np.longdouble
        AB = np.einsum(spec, np.asarray(A, dtype=dt), np.asarray(B, dtype=dt), optimize=False)
    else:
        AB = np.einsum(spec, A, B, optimize=True)
    out = alpha * AB
    if beta != 0:
        out = out + beta * R
    return out


def contract_dense(A, lA, B, lB, lR=None, **kw):
    """NDTensors.contract(T1, l1, T2, l2[, lR]) for Dense data
    (generic_tensor_operations.jl:87-119): allocate R and contract into it."""
    if lR is None:
        lR = contract_labels(lA, lB)
    return dense_contract(None, lR, A, lA, B, lB, **kw), lR


# ---------------------------------------------- block-sparse contraction
def _find_matching_positions(t1, t2):
    """contract_utilities.jl:45-55 (0 = absent; 1-based positions)."""
    out = []
    for x in t1:
        pos = 0
        for j, y in enumerate(t2):
            if x == y:
                pos = j + 1
        out.append(pos)
    return out


def contract_blockoffsets_sequential(T1, l1, T2, l2, lR):
    """NDTensors/src/blocksparse/contract_sequential.jl:1-41 with the block
    rules of contract_utilities.jl:57-91.

    Iterates block1 over T1's blocks (outer), block2 over T2's (inner); a pair
    contracts when every contracted mode has the same block number. The
    output block takes the free block numbers into their R positions. Output
    offsets are assigned on first appearance. Returns (offsetsR, nnzR, plan)
    with plan = [(block1, block2, blockR), ...].
    """
    indsR = contract_inds(T1.inds, l1, T2.inds, l2, lR)
    m12 = _find_matching_positions(l1, l2)
    m1R = _find_matching_positions(l1, lR)
    m2R = _find_matching_positions(l2, lR)
    offsR = OrderedDict()
    nnzR = 0
    plan = []
    for b1 in T1.offsets:
        for b2 in T2.offsets:
            ok = all(p == 0 or b1[i] == b2[p - 1] for i, p in enumerate(m12))
            if not ok:
                continue
            bR = [0] * len(lR)
            for i, p in enumerate(m1R):
                if p > 0:
                    bR[p - 1] = b1[i]
            for i, p in enumerate(m2R):
                if p > 0:
                    bR[p - 1] = b2[i]
            bR = tuple(bR)
            plan.append((b1, b2, bR))
            if bR not in offsR:
                offsR[bR] = nnzR
                nnzR += int(np.prod(blockdims(indsR, bR)))
    return offsR, nnzR, plan, indsR


def contract_blockoffsets_threaded(T1, l1, T2, l2, lR):
    """contract_blockoffsets with Algorithm"threaded_threads"
    (NDTensors/src/blocksparse/contract_generic.jl:3-32 over contract_blocks,
    contract_threaded.jl:2-73), the algorithm the reference picks when
    using_threaded_blocksparse() and Julia runs more than one thread
    (contract.jl:42-55).

    contract_blocks partitions T1's blocks over the tasks when T1 has more
    blocks than T2 (T1 outer, T2 inner, as the sequential loop), else T2's
    blocks (T2 outer, T1 inner); each task keeps its loop order and the task
    results are concatenated in partition order, so the plan is the plain
    nested loop in that outer/inner order. Offsets are then assigned by
    first appearance over the plan. Same return value as the sequential
    version."""
    indsR = contract_inds(T1.inds, l1, T2.inds, l2, lR)
    m12 = _find_matching_positions(l1, l2)
    m1R = _find_matching_positions(l1, lR)
    m2R = _find_matching_positions(l2, lR)
    blocks1, blocks2 = list(T1.offsets), list(T2.offsets)
    if len(blocks1) > len(blocks2):
        order = [(b1, b2) for b1 in blocks1 for b2 in blocks2]
    else:
        order = [(b1, b2) for b2 in blocks2 for b1 in blocks1]
    plan = []
    for b1, b2 in order:  # maybe_contract_blocks (contract_utilities.jl:17-26)
        if not all(p == 0 or b1[i] == b2[p - 1] for i, p in enumerate(m12)):
            continue
        bR = [0] * len(lR)
        for i, p in enumerate(m1R):
            if p > 0:
                bR[p - 1] = b1[i]
        for i, p in enumerate(m2R):
            if p > 0:
                bR[p - 1] = b2[i]
        plan.append((b1, b2, tuple(bR)))
    offsR = OrderedDict()
    nnzR = 0
    for _, _, bR in plan:
        if bR not in offsR:
            offsR[bR] = nnzR
            nnzR += int(np.prod(blockdims(indsR, bR)))
    return offsR, nnzR, plan, indsR


def blocksparse_contract(T1, l1, T2, l2, lR=None, alpha=1.0, exact=False, conj1=False, conj2=False,
                         pair_alpha=None, threaded=False):
    """NDTensors.contract(::BlockSparseTensor, ...) (blocksparse/contract.jl:3-76,
    contract_generic.jl:37-129): plan, allocate R, then for each output block
    the pairs in plan order with beta = 0 for the first and 1 afterwards.
    pair_alpha: None, or one factor per pair of the plan (enumeration order)
    multiplying alpha -- the value compute_alpha returns for that pair
    (contract_utilities.jl:2-15, called per pair at contract_generic.jl:91).
    threaded: the plan of contract_blockoffsets_threaded (the reference with
    threaded block sparse on); the sums into each output block then follow
    that plan's order.
    Returns (R, plan)."""
    if lR is None:
        lR = contract_labels(l1, l2)
    plan_of = contract_blockoffsets_threaded if threaded else contract_blockoffsets_sequential
    offsR, nnzR, plan, indsR = plan_of(T1, l1, T2, l2, lR)
    if pair_alpha is not None and len(pair_alpha) != len(plan):
        raise ValueError(f"pair_alpha has {len(pair_alpha)} factors for {len(plan)} pairs")
    cplx = (np.iscomplexobj(T1.data) or np.iscomplexobj(T2.data) or np.iscomplexobj(alpha)
            or (pair_alpha is not None and np.iscomplexobj(np.asarray(pair_alpha))))
    dt = np.complex128 if cplx else np.float64
    if exact:
        dt = np.clongdouble if cplx else np.longdouble
    data = np.full(nnzR, np.nan, dtype=dt)  # uninitialised output (similar.jl:5-8)
    R = BlockSparse(data, offsR, indsR)
    written = set()
    for n, (b1, b2, bR) in enumerate(plan):
        A = T1.block(b1)
        B = T2.block(b2)
        if conj1:
            A = np.conj(A)
        if conj2:
            B = np.conj(B)
        view = R.block(bR)
        beta = 1.0 if bR in written else 0.0
        a = alpha if pair_alpha is None else alpha * pair_alpha[n]
        view[...] = dense_contract(view, lR, A, l1, B, l2, a, beta, exact=exact)
        written.add(bR)
    return R, plan


# ------------------------------------------------ block-sparse permutedims!
def blocksparse_permutedims_(R, T, perm, alpha=1.0, beta=0.0):
    """permutedims!(R, T, perm, f) with f(r, t) = beta*r + alpha*t
    (NDTensors/src/blocksparse/blocksparsetensor.jl:834-881): over
    union(R's blocks, T's permuted blocks) in that order, R's block (zeros if
    absent) becomes f.(r, permutedims(T's block or zeros, perm)); a block R
    did not have is inserted (appended, blockoffsets.jl insertion order) only
    when the result is nonzero. beta == 0 does not read R (NaN-safe).
    perm is 0-based here. Returns the new R (new data and offsets)."""
    perm = tuple(perm)
    inv = tuple(np.argsort(perm))
    order = list(R.offsets)
    for b in T.offsets:
        pb = tuple(b[p] for p in perm)
        if pb not in R.offsets and pb not in order:
            order.append(pb)
    dt = np.result_type(R.data, T.data, np.asarray(alpha), np.asarray(beta))
    blocks, offs, n = [], OrderedDict(), 0
    for bR in order:
        bT = tuple(bR[i] for i in inv)
        shape = blockdims(R.inds, bR)
        t = (np.transpose(T.block(bT), perm) if bT in T.offsets
             else np.zeros(shape, dtype=dt))
        if bR in R.offsets:
            r = R.block(bR)
            v = alpha * t + (beta * r if beta != 0 else 0)
        else:
            v = alpha * t
            if not np.any(v != 0):
                continue  # a new all-zero block is not inserted
        v = np.asarray(v, dtype=dt)
        offs[bR] = n
        n += v.size
        blocks.append(v.reshape(-1, order="F"))
    data = np.concatenate(blocks) if blocks else np.zeros(0, dtype=dt)
    return BlockSparse(data, offs, R.inds)


# ------------------------------------------------------------ DiagBlockSparse
def nzdiagblocks(flux, inds):
    """src/qn/qnindexset.jl:20-29: the diagonal blocks (k, ..., k) (0-based
    here) over k < min(nblocks), kept when their flux is `flux`."""
    n = min(i.nblocks() for i in inds)
    return [b for b in ((k,) * len(inds) for k in range(n)) if qn_eq(flux_of_block(inds, b), flux)]


class DiagBlockSparse:
    """DiagBlockSparse (NDTensors/src/blocksparse/diagblocksparse.jl:1-60):
    blocks on the block diagonal only, each storing its diagonal (min of its
    dims values) from its offset in `data`; `uniform` = one value on every
    diagonal element (UniformDiagBlockSparse, what delta builds)."""

    def __init__(self, data, offsets, inds, uniform=None):
        self.data = data
        self.offsets = offsets
        self.inds = tuple(inds)
        self.uniform = uniform

    def diag(self, b):
        n = min(blockdims(self.inds, b))
        if self.uniform is not None:
            return np.full(n, self.uniform)
        o = self.offsets[b]
        return self.data[o:o + n]

    def dense_block(self, b):
        """dense of one diagonal block (diag/diagtensor.jl: zeros with the
        diagonal set)."""
        d = self.diag(b)
        out = np.zeros(blockdims(self.inds, b), dtype=np.result_type(d, np.float64))
        for j, v in enumerate(d):
            out[(j,) * len(self.inds)] = v
        return out

    def as_blocksparse(self):
        """The same block list with every block made dense: the tensor the
        reference contracts against (convert_to_dense)."""
        offs, nnz = blockoffsets(list(self.offsets), self.inds)
        dt = np.complex128 if (np.iscomplexobj(self.data) or np.iscomplexobj(self.uniform)) else np.float64
        data = np.zeros(nnz, dtype=dt)
        for b, o in offs.items():
            blk = self.dense_block(b)
            data[o:o + blk.size] = blk.reshape(-1, order="F")
        return BlockSparse(data, offs, self.inds)

    def dense(self):
        return self.as_blocksparse().dense()


def diagblocksparse_contract(T, lT, D, lD, lR=None, alpha=1.0):
    """contract(::BlockSparseTensor, ::DiagBlockSparseTensor)
    (NDTensors/src/blocksparse/diagblocksparse.jl:596-675): the output
    structure from contract_blockoffsets, then per pair in plan order the
    dense contract! of the T block with the dense form of the diagonal block
    (diag/tensoralgebra/contract.jl:103-160 with convert_to_dense, :208-216
    for the argument order), beta 0 for an output block's first pair and 1
    after. A block off the block diagonal is an error
    (diagblocksparse.jl:645-649). The output labels default to T's first
    (contract_labels(lT, lD); diagonal-first calls swap, :626-634)."""
    for b in D.offsets:
        if any(c != b[0] for c in b):
            raise ValueError("When contracting a BlockSparse tensor with a DiagBlockSparse "
                             "tensor, the DiagBlockSparse tensor must be block diagonal")
    return blocksparse_contract(T, lT, D.as_blocksparse(), lD, lR, alpha)


# ------------------------------------------------------------------- adjoint
def rrule_contract(x1, l1, x2, l2, ybar, lR):
    """ext/ITensorsChainRulesCoreExt/itensor.jl:94-103 on dense data:
    x1bar = ybar * dag(x2), x2bar = dag(x1) * ybar, projected to each input's
    element type (projection.jl:1-10). dag conjugates the data."""
    x1bar = np.einsum(_einsum_spec(tuple(lR), tuple(l2), tuple(l1)), ybar, np.conj(x2))
    x2bar = np.einsum(_einsum_spec(tuple(l1), tuple(lR), tuple(l2)), np.conj(x1), ybar)
    if not np.iscomplexobj(x1):
        x1bar = np.real(x1bar)
    if not np.iscomplexobj(x2):
        x2bar = np.real(x2bar)
    return x1bar, x2bar


# ------------------------------------------------ H_eff network (oracle)
def heff_apply_dense(L, psi, W1, W2, R, exact=False):
    """((((L*psi)*W1)*W2)*R) with the left-associative fold of
    src/tensor_operations/tensor_algebra.jl:121-134, on dense arrays with
    mode orders L[a',w,a], psi[a,s1,s2,b], W[wl,s,s',wr], R[b',w,b]."""
    X, lX = contract_dense(L, (1, 2, -1), psi, (-1, 3, 4, 5), exact=exact)      # X[a',w,s1,s2,b]
    Y, lY = contract_dense(X, (1, -1, -2, 4, 5), W1, (-1, -2, 6, 7), exact=exact)  # Y[a',s2,b,s1',w']
    Z, lZ = contract_dense(Y, (1, -1, 3, 4, -2), W2, (-2, -1, 6, 7), exact=exact)  # Z[a',b,s1',s2',w'']
    O, lO = contract_dense(Z, (1, -1, 3, 4, -2), R, (5, -2, -1), exact=exact)     # O[a',s1',s2',b']
    return O


# ------------------------------------------------------------------ combiner
# A combiner fuses indices i_1..i_n (first fastest) into one index whose
# sectors are the product sectors sorted by QN and merged when equal
# (src/qn/qnitensor.jl:462-467, src/qn/qnindex.jl:359-424). Restated here
# element by element: every multi-index (x_1..x_n) of the fused indices has
# one position in the combined index, and the combined tensor is the dense
# tensor re-indexed through that map, cut into the reference's blocks.


def qn_order_key(names, q):
    """QN ordering (src/lib/QuantumNumbers/src/qn.jl:273-291): values in
    name order, missing names counting as zero."""
    return tuple(q.get(n, (0, 1))[0] for n in names)


def combined_index(inds, dir=None):
    """(sectors of the combined index, sorted product sector order, merge map).

    dir defaults to In (-1) when every index is In, else Out (+1): the left
    fold of outer (NDTensors/src/blocksparse/blockdims.jl:165-167) with
    src/qn/qnindex.jl:369-379 at each step. Product sectors: first index
    fastest, QN = dir * sum_k dir_k * q_k, size = product. Sorting is stable
    (Julia's sortperm); merged neighbours have equal QN (qnindex.jl:407-424).
    Returned order / merge are 0-based lists over sorted positions."""
    if dir is None:
        dir = -1 if all(i.dir == -1 for i in inds) else 1
    grids = [range(i.nblocks()) for i in inds]
    prod = []
    for rev in itertools.product(*reversed(grids)):
        bs = tuple(reversed(rev))
        q = {}
        d = 1
        for i, b in zip(inds, bs):
            q = qn_add(q, qn_scale(i.dir, i.qn(b)))
            d *= i.blockdim(b)
        prod.append((bs, qn_scale(dir, q), d))
    names = sorted({n for _b, q, _d in prod for n in q})
    order = sorted(range(len(prod)), key=lambda k: qn_order_key(names, prod[k][1]))
    sectors, merge = [], []
    for k in order:
        _bs, q, d = prod[k]
        if sectors and qn_eq(sectors[-1][0], q):
            sectors[-1] = (sectors[-1][0], sectors[-1][1] + d)
        else:
            sectors.append((q, d))
        merge.append(len(sectors) - 1)
    return sectors, order, merge, prod, dir


def combined_position_map(inds, order):
    """For every multi-index x of the fused indices (dense, 0-based), its
    position in the combined index: sorted product sectors laid end to end,
    each sector's elements column-major over its block-local indices."""
    grids = [range(i.nblocks()) for i in inds]
    prods = []
    for rev in itertools.product(*reversed(grids)):
        prods.append(tuple(reversed(rev)))
    starts = [np.concatenate([[0], np.cumsum([i.blockdim(b) for b in range(i.nblocks())])]) for i in inds]
    pos = {}
    at = 0
    for k in order:
        bs = prods[k]
        dims = [inds[m].blockdim(b) for m, b in enumerate(bs)]
        for rev in itertools.product(*[range(d) for d in reversed(dims)]):
            loc = tuple(reversed(rev))
            x = tuple(int(starts[m][b]) + l for m, (b, l) in enumerate(zip(bs, loc)))
            pos[x] = at
            at += 1
    return pos


def _sector_of(starts, x):
    return int(np.searchsorted(starts, x, side="right") - 1)


def combine_blocksparse(T, lT, cinds, lC, dir=None):
    """A * combiner(cinds...) for a block-sparse A (NDTensors/src/blocksparse/
    combiner.jl:25-81, blocksparsetensor.jl:537-630). lC = (combined label,
    labels of cinds in T). Returns (BlockSparse R, labels (combined, rest of
    T in order), combined Index).

    Output blocks: each T block's (sorted position of its product sector,
    rest) in column-major order, first coordinate then replaced by its
    merged sector, first occurrences kept (blocksparsetensor.jl:560-567,
    blockoffsets.jl:35-37)."""
    clabel, luc = lC[0], tuple(lC[1:])
    ucpos = [lT.index(x) for x in luc]
    rest = [p for p in range(len(lT)) if lT[p] not in luc]
    lR = (clabel,) + tuple(lT[p] for p in rest)
    sectors, order, merge, prod, cdir = combined_index(cinds, dir)
    cind = Index(sectors, "CMB,Link", dir=cdir)
    where = {prod[k][0]: s for s, k in enumerate(order)}  # product sector -> sorted position
    keys = []
    for b in T.offsets:
        s = where[tuple(b[p] for p in ucpos)]
        keys.append((s,) + tuple(b[p] for p in rest))
    keys.sort(key=lambda t: t[::-1])
    blocks = []
    for k in keys:
        nb = (merge[k[0]],) + k[1:]
        if nb not in blocks:
            blocks.append(nb)
    indsR = (cind,) + tuple(T.inds[p] for p in rest)
    offs, nnz = blockoffsets(blocks, indsR)
    # dense re-indexing through the position map, then cut into blocks
    pos = combined_position_map([T.inds[p] for p in ucpos], order)
    Td = T.dense()
    Rd = np.zeros(tuple(i.dim for i in indsR), dtype=Td.dtype)
    it = np.nditer(Td, flags=["multi_index"])
    for v in it:
        x = it.multi_index
        Rd[(pos[tuple(x[p] for p in ucpos)],) + tuple(x[p] for p in rest)] = v
    data = np.zeros(nnz, dtype=Td.dtype)
    R = BlockSparse(data, offs, indsR)
    starts = [np.concatenate([[0], np.cumsum([i.blockdim(b) for b in range(i.nblocks())])]) for i in indsR]
    for b, o in offs.items():
        sl = tuple(slice(starts[m][c], starts[m][c + 1]) for m, c in enumerate(b))
        blk = Rd[sl]
        data[o:o + blk.size] = blk.reshape(-1, order="F")
    return R, lR, cind


def uncombine_blocksparse(T, lT, cind, cinds, lC, order, merge):
    """A * dag(combiner) for a block-sparse A holding the combined index
    (NDTensors/src/blocksparse/combiner.jl:82-140, blocksparsetensor.jl:
    664-795). lC = (combined label, labels the uncombined indices get).
    Returns (BlockSparse R, labels (uncombined, rest of T in order)).
    Output blocks: for each T block in order, the sorted sectors merged into
    its combined sector in sorted order, as product sectors."""
    clabel, luc = lC[0], tuple(lC[1:])
    cpos = lT.index(clabel)
    rest = [p for p in range(len(lT)) if p != cpos]
    lR = luc + tuple(lT[p] for p in rest)
    grids = [range(i.nblocks()) for i in cinds]
    prods = [tuple(reversed(rev)) for rev in itertools.product(*reversed(grids))]
    blocks = []
    for b in T.offsets:
        for s in range(len(order)):
            if merge[s] == b[cpos]:
                blocks.append(prods[order[s]] + tuple(b[p] for p in rest))
    indsR = tuple(cinds) + tuple(T.inds[p] for p in rest)
    offs, nnz = blockoffsets(blocks, indsR)
    pos = combined_position_map(list(cinds), order)
    Td = T.dense()
    Rd = np.zeros(tuple(i.dim for i in indsR), dtype=Td.dtype)
    n = len(cinds)
    it = np.nditer(Rd, flags=["multi_index"])
    for _v in it:
        x = it.multi_index
        src = [0] * len(lT)
        src[cpos] = pos[x[:n]]
        for r, p in enumerate(rest):
            src[p] = x[n + r]
        Rd[x] = Td[tuple(src)]
    data = np.zeros(nnz, dtype=Td.dtype)
    starts = [np.concatenate([[0], np.cumsum([i.blockdim(b) for b in range(i.nblocks())])]) for i in indsR]
    for b, o in offs.items():
        sl = tuple(slice(starts[m][c], starts[m][c + 1]) for m, c in enumerate(b))
        blk = Rd[sl]
        data[o:o + blk.size] = blk.reshape(-1, order="F")
    return BlockSparse(data, offs, indsR), lR

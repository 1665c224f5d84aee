"""CPU checks of the boundary and of the host-side logic (no GPU calls).

* libitg.so loads and exports every function include/itg.h declares, with
  the status / dtype / precision constants the Python binding assumes.
* The product's host logic (itg.itensors: QN, labels, nzblocks, eachblock;
  itg.configs: sector sizes) agrees with the oracle on the same inputs.
"""
import os
import re

import numpy as np
import pytest

from helpers import ROOT

import itg
from itg import _lib
from itg.configs import sector_sizes
from oracle import ndtensors_oracle as O

HEADER = os.path.join(ROOT, "include", "itg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(itg_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    decl = declared_functions()
    assert len(decl) >= 15
    for name in decl:
        assert hasattr(L, name), name
    assert set(decl) == set(_lib.EXPORTS)


def test_header_constants_match_binding():
    text = open(HEADER).read()

    def val(name):
        return int(re.search(name + r"\s*=\s*(\d+)", text).group(1))

    assert val("ITG_OK") == _lib.ITG_OK
    assert val("ITG_ERR_INVALID") == _lib.ITG_ERR_INVALID
    assert val("ITG_ERR_UNSUPPORTED") == _lib.ITG_ERR_UNSUPPORTED
    assert val("ITG_ERR_HIP") == _lib.ITG_ERR_HIP
    assert val("ITG_F64") == _lib.ITG_F64 and val("ITG_C128") == _lib.ITG_C128
    assert val("ITG_F32") == _lib.ITG_F32 and val("ITG_C64") == _lib.ITG_C64
    assert val("ITG_PREC_FP64") == _lib.PREC["fp64"]
    assert val("ITG_PREC_SPLIT_BF16") == _lib.PREC["split_bf16"]
    assert val("ITG_PREC_SPLIT_INT8") == _lib.PREC["split_int8"]
    assert val("ITG_BS_ORDER_SEQUENTIAL") == _lib.ITG_BS_ORDER_SEQUENTIAL
    assert val("ITG_BS_ORDER_THREADED") == _lib.ITG_BS_ORDER_THREADED


def test_version_and_init_without_device():
    L = _lib.lib()
    assert L.itg_version().decode().count(".") == 2
    import ctypes
    h = ctypes.c_void_p()
    st = L.itg_init(0, ctypes.byref(h))
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        assert st == _lib.ITG_ERR_HIP  # no device: a clean error, no crash
    else:
        assert st == _lib.ITG_OK
        L.itg_destroy(h)


def test_null_handle_is_rejected():
    L = _lib.lib()
    assert L.itg_synchronize(None) == _lib.ITG_ERR_INVALID
    assert L.itg_bs_plan_destroy(None) == _lib.ITG_OK
    assert L.itg_timer_enable(None, 1) == _lib.ITG_ERR_INVALID
    assert L.itg_timer_count(None) == 0
    assert L.itg_set_split_tile(None, 1) == _lib.ITG_ERR_INVALID
    assert L.itg_set_gemm_backend(None, 0) == _lib.ITG_ERR_INVALID
    assert L.itg_set_blocksparse_order(None, _lib.ITG_BS_ORDER_THREADED) == _lib.ITG_ERR_INVALID
    assert L.itg_bs_copy(None, _lib.ITG_F64, 1, 1, None, None, None, None, None, None, 0,
                         None, 0) == _lib.ITG_ERR_INVALID


# ---------------------------------------------------------------- host logic
def _pair(space, tag, dir=1):
    p = itg.Index([(itg.QN(q), d) for q, d in space], tag, dir=dir)
    o = O.Index([(O.QN(q), d) for q, d in space], tags=tag, dir=dir, id=p.id)
    return p, o


def test_labels_match_oracle():
    rng = np.random.default_rng(0)
    for _ in range(50):
        n = int(rng.integers(2, 7))
        pool = [itg.Index(2, f"t{i}") for i in range(n)]
        opool = [O.Index(2, tags=f"t{i}", id=pool[i].id) for i in range(n)]
        ia = list(rng.choice(n, size=int(rng.integers(1, n + 1)), replace=False))
        ib = list(rng.choice(n, size=int(rng.integers(1, n + 1)), replace=False))
        got = itg.compute_contraction_labels([pool[i] for i in ia], [pool[i] for i in ib])
        ref = O.compute_contraction_labels([opool[i] for i in ia], [opool[i] for i in ib])
        assert got == ref
        assert itg.ndtensors.contract_labels(*got) == O.contract_labels(*ref)


def test_nzblocks_and_flux_match_oracle():
    rng = np.random.default_rng(1)
    for _ in range(30):
        inds, oinds = [], []
        for t in range(int(rng.integers(1, 5))):
            sp = [(int(q), int(rng.integers(1, 4))) for q in range(-1, int(rng.integers(1, 3)))]
            p, o = _pair(sp, f"i{t}", dir=int(rng.choice([-1, 1])))
            inds.append(p)
            oinds.append(o)
        f = int(rng.integers(-2, 3))
        got = itg.nzblocks(itg.QN(f), inds)
        ref = O.nzblocks(O.QN(f), oinds)
        assert [tuple(c - 1 for c in b) for b in got] == ref


def test_qn_semantics():
    QN = itg.QN
    assert QN(0) + QN(1) == QN(1)
    assert QN() == QN(("Sz", 0))
    assert QN(("Sz", 1)) + QN(("N", 1)) == QN(("N", 1), ("Sz", 1))
    assert QN(("P", 1, 2)) + QN(("P", 1, 2)) == QN(("P", 0, 2))
    assert -1 * QN(("Sz", 2)) == QN(("Sz", -2))


def test_arrow_check_raises_on_host():
    p, _ = _pair([(0, 2), (1, 2)], "i")
    with pytest.raises(ValueError):
        itg.compute_contraction_labels((p, p.dag().prime()), (p, p.prime(2)))


def test_sector_sizes_sum_and_symmetry():
    for chi in (64, 1024, 2048, 4096):
        s = sector_sizes(chi)
        n = [x for _, x in s]
        assert sum(n) == chi and n == n[::-1]
    assert [x for _, x in sector_sizes(2048)] == [2, 5, 9, 16, 28, 44, 66, 94, 124, 154, 181, 198,
                                                 206, 198, 181, 154, 124, 94, 66, 44, 28, 16, 9, 5, 2]


def test_precision_flag():
    old = itg.set_precision("split_int8")
    assert itg.precision() == "split_int8"
    itg.set_precision(old)
    with pytest.raises(ValueError):
        itg.set_precision("fp32")


def test_reference_algorithm_switches_are_accepted():
    """NDTensors' threading / TBLIS / GEMM-backend switches exist and report
    their state (NDTensors.jl:106-184,243-255; tensoralgebra/contract.jl:15-24)."""
    from itg import ndtensors as ND
    assert ND.enable_threaded_blocksparse() is False
    assert ND.using_threaded_blocksparse() is True
    assert ND.disable_threaded_blocksparse() is True
    ND.enable_tblis()
    assert ND.using_tblis()
    ND.disable_tblis()
    ND.backend_blas()
    assert ND.gemm_backend() == "BLAS"
    ND.backend_auto()
    assert ND.gemm_backend() == "Auto"

"""itg -- MI355X (gfx950) backend for the tensor contraction path of
ITensors.jl / NDTensors.jl.

Layers (host side mirrors the reference's interface, device side is libitg.so):
  itg.itensors    Index / QN / ITensor, `*` and `contract`
  itg.ndtensors   NDTensors.contract / contract! / permutedims on device storage
  itg.chainrules  rrule(contract) with a device adjoint
  itg.configs     effective-Hamiltonian workloads (dense and U(1) block-sparse)
  itg.network     fixed contraction sequences (fused steps, graph capture)
  itg.hdf5        HDF5 read / write in the ITensors.jl layout (host, libhdf5)
  itg._lib        the ctypes image of the C ABI (include/itg.h)
"""
from ._lib import LIB_PATH, ItgError, ItgUnsupported, lib  # noqa: F401
from . import ndtensors  # noqa: F401
from .itensors import (  # noqa: F401
    QN, Index, ITensor, In, Neither, Out, compute_contraction_labels, contract,
    contract_network, dag, eachblock, flux_of_block, itensor_from_array, nzblocks, prime,
    random_itensor, blocksparse_itensor, combiner, combinedind, combineblocks, iscombiner,
    delta, diag_itensor, nzdiagblocks, contraction_cost, optimal_contraction_sequence,
    enable_contraction_sequence_optimization, disable_contraction_sequence_optimization,
    using_contraction_sequence_optimization,
)
from .ndtensors import precision, set_precision  # noqa: F401
from .ndtensors import (  # noqa: F401
    disable_threaded_blocksparse, enable_threaded_blocksparse, using_threaded_blocksparse)
from .itensors import tagstring  # noqa: F401
from . import hdf5  # noqa: F401

__version__ = "0.1.0"

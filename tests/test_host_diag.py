"""Host mirror of the diagonal ITensors (no GPU): dag of stored and
uniform diagonals."""
import numpy as np

import helpers  # noqa: F401



def test_dag_of_diagonal_itensors():
    """dag conjugates the stored diagonal (or the uniform value) and flips
    every arrow, keeping the diagonal storage (src/itensor.jl:1806-1846)."""
    import itg
    i = itg.Index([(itg.QN(("Sz", 0)), 2), (itg.QN(("Sz", 1)), 3)], "i")
    v = np.arange(5) + 1j * np.arange(5)
    D = itg.diag_itensor(v, i.dag(), i.prime(), device="cpu")
    Dd = itg.dag(D)
    assert Dd.isdiag() and [x.dir for x in Dd.inds] == [-x.dir for x in D.inds]
    assert np.array_equal(Dd.tensor.data.numpy(), np.conj(v))
    d = itg.delta(i.dag(), i.prime(), dtype=np.complex128)
    dd = itg.dag(d)
    assert dd.isdiag() and dd.tensor.uniform == 1.0 and dd.tensor.data is None

"""HDF5 round trip of device tensors: block-sparse tensors written from the
device and read back onto it contract to the same bits as the originals."""
import numpy as np
import pytest

import helpers  # noqa: F401

pytestmark = pytest.mark.gpu


def test_device_round_trip_then_contract(tmp_path):
    import torch
    import itg
    from itg import hdf5 as H5
    from itg.configs import heff_network
    try:
        H5._h5()
    except H5.H5Error as e:
        pytest.skip(f"libhdf5 not available: {e}")
    L, psi, W1, W2, R, ix = heff_network(256, qns=True, seed=7, device="cuda:0")
    p = str(tmp_path / "net.h5")
    with H5.h5open(p, "w") as f:
        for name, T in (("L", L), ("psi", psi)):
            H5.write(f, name, T)
    with H5.h5open(p) as f:
        rL = H5.read(f, "L", itg.ITensor, device="cuda:0")
        rpsi = H5.read(f, "psi", itg.ITensor, device="cuda:0")
    assert rL.tensor.data.is_cuda and rpsi.tensor.data.is_cuda
    assert rL.inds == L.inds and rpsi.inds == psi.inds
    a = (L * psi).tensor.data
    b = (rL * rpsi).tensor.data
    torch.cuda.synchronize()
    assert torch.equal(a, b)

"""HIP graphs through the C ABI (itg_graph_begin / _end / _launch): a whole
block-sparse H_eff apply and a dense contraction recorded once and replayed;
every replay (with eager work and synchronisations between replays) must
equal the eager result bit for bit. Misuse returns ITG_ERR_INVALID."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import itg
    from itg import ndtensors as ND
    torch.cuda.init()
    return torch, itg, ND


@pytest.mark.parametrize("mode", ["fp64", "split_int8"])
def test_graph_replays_heff_apply(env, mode):
    torch, itg, ND = env
    from itg.configs import heff_network
    from itg.network import ContractionSequence
    net = heff_network(512, qns=True, seed=2, device="cuda")[:5]
    seq = ContractionSequence(net)
    seq.run(mode)  # plans, tables and workspace built outside the capture
    torch.cuda.synchronize()
    eager = seq.result.tensor.data.clone()
    g = ND.Graph.capture(lambda: seq.run(mode))
    scratch = torch.randn(1 << 20, dtype=torch.float64, device="cuda")
    for _ in range(3):
        seq.result.tensor.data.fill_(float("nan"))
        scratch.mul_(1.5)  # eager work between replays
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(seq.result.tensor.data, eager)


def test_graph_replays_dense_contract(env):
    torch, itg, ND = env
    rng = np.random.default_rng(4)
    a = ND.DenseTensor.from_numpy(rng.standard_normal((64, 40, 30)))
    b = ND.DenseTensor.from_numpy(rng.standard_normal((30, 64, 50)))
    r = ND.DenseTensor(torch.empty(40 * 50, dtype=torch.float64, device="cuda"), (40, 50))
    run = lambda: ND.contract_(r, (2, 3), a, (-1, 2, -2), b, (-2, -1, 3))
    run()
    torch.cuda.synchronize()
    eager = r.data.clone()
    g = ND.Graph.capture(run)
    for _ in range(3):
        r.data.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(r.data, eager)


def test_graph_misuse_is_refused(env):
    torch, itg, ND = env
    from itg import _lib
    L = _lib.lib()
    h = ND._handle(torch.device("cuda:0"))
    g = ctypes.c_void_p()
    assert L.itg_graph_end(h.h, ctypes.byref(g)) == _lib.ITG_ERR_INVALID
    assert L.itg_graph_begin(h.h) == _lib.ITG_OK
    assert L.itg_graph_begin(h.h) == _lib.ITG_ERR_INVALID
    assert L.itg_graph_launch(h.h, None) == _lib.ITG_ERR_INVALID
    assert L.itg_graph_end(h.h, ctypes.byref(g)) == _lib.ITG_OK  # an empty graph
    assert L.itg_graph_destroy(g) == _lib.ITG_OK

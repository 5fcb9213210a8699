"""Graph replay of the preconditioner: applyPrecondDev on the same device
buffers is captured once and replayed.  Every replay, with other device work
and synchronisations in between, must equal the direct launches exactly
(same kernels in the same order)."""
import math

import numpy as np
import pytest

from tests._common import make_pair

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,nrhs", [(0, 1), (1, 1), (1, 3)])
def test_replays_equal_direct_launches(prec, nrhs):
    import torch
    cb, dd = make_pair((3, 3), (19, 19), 2 * math.pi * 3.3, precision=prec)
    dd.factorize()
    dd.set_graphs(True)
    rng = np.random.default_rng(21)
    X = rng.standard_normal((nrhs, dd.N)) + 1j * rng.standard_normal((nrhs, dd.N))
    ref = dd.apply_precond_block(X) if nrhs > 1 else dd.apply_precond(X[0])[None, :]  # direct launches
    x = torch.from_numpy(X.reshape(-1)).to("cuda:0")
    y = torch.zeros_like(x)
    for rep in range(4):
        y.fill_(float("nan"))
        if nrhs > 1:
            dd.apply_precond_block_dev(nrhs, x.data_ptr(), y.data_ptr())
        else:
            dd.apply_precond_dev(x.data_ptr(), y.data_ptr())
        dd.sync()
        got = y.cpu().numpy().reshape(nrhs, dd.N)
        assert np.array_equal(got, ref), f"replay {rep}"
        # unrelated eager work between replays
        z = torch.randn(1 << 20, device="cuda:0").sum()
        torch.cuda.synchronize()
        _ = dd.apply_a(X[0])
        assert torch.isfinite(z)

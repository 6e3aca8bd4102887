"""Argument validation of the row-wise CG wrappers, on the host: every malformed call is
refused with ValueError before any kernel could be launched (no GPU needed)."""

import pytest
import torch

from isogriddft_amd import reductions as red

K, N = 3, 40


def _blk(dt=torch.float64):
    return torch.zeros((K, N), dtype=dt)


def _vec(n=N, dt=torch.float64):
    return torch.ones(n, dtype=dt)


def _alpha(k=K):
    return torch.zeros(k, dtype=torch.float64)


@pytest.mark.parametrize("call", [
    lambda: red.row_dots(_blk(), _blk(torch.float32)),                      # dtype mismatch
    lambda: red.row_dots(_blk(), _blk()[:, :N - 1]),                        # shape mismatch
    lambda: red.row_dots(_blk(), _blk(), _vec(N - 1)),                      # short weights
    lambda: red.row_dots(_blk().t(), _blk().t()),                           # column stride != 1
    lambda: red.cg_update(_blk(), _blk(), _blk(), _blk(), _vec(N + 1), None, _alpha()),
    lambda: red.cg_update(_blk(), _blk(), _blk(), _blk(), _vec(), _vec(dt=torch.float32), _alpha()),
    lambda: red.cg_update(_blk(), _blk(), _blk(), _blk(), _vec(), None, _alpha(K + 1)),
    lambda: red.cg_direction(_blk(), _blk(), _vec(), _alpha().float()),     # FP32 beta
    lambda: red.cg_start(_blk(), _vec(), _blk(), _vec(), _vec(), _blk(), _blk(torch.float32)),
    lambda: red.cg_start(_blk()[:, :N - 1], _vec(), _blk(), _vec(), _vec(), _blk(), _blk()),
    lambda: red.cg_start(_blk(), None, _blk(), _vec(), _vec(), _blk(), _blk()),   # isw required
    lambda: red.cg_finish(_blk(), _blk(), _vec(N - 2), _alpha(), _blk()),
    lambda: red.cg_finish(_blk(), _blk(), _vec(), _alpha(), _blk(torch.float32)),
    lambda: red.cg_finish(_blk(), _blk(), _vec(), _alpha(2), _blk()),
    lambda: red.gram(_blk(), _blk(torch.float32), None),                    # dtype mismatch
    lambda: red.gram(_blk(), torch.zeros((2, N + 1), dtype=torch.float64), None),
    lambda: red.gram(_blk(), _blk(), _vec(N - 1)),
    lambda: red.gram(_blk(), _blk(), None, out=torch.zeros((K, K - 1), dtype=torch.float64)),
    lambda: red.gram(_blk(), _blk(), None, out=torch.zeros((K, K), dtype=torch.float32)),
])
def test_malformed_calls_are_refused(call):
    with pytest.raises(ValueError) as e:
        call()
    assert "device tensors" not in str(e.value)  # refused by the argument checks themselves


def test_well_formed_host_tensors_are_refused_as_host():
    """A well-formed call on host tensors reaches the pointer check, which refuses host
    memory: there is no CPU path."""
    if not red._lib.LIB_PATH.exists():
        pytest.skip("library not built")
    with pytest.raises(ValueError, match="device tensors"):
        red.cg_finish(_blk(), _blk(), _vec(), _alpha(), _blk())

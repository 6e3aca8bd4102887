"""The allocation-free weighted-dot entry point (ig_dots_noalloc_f64) against the
workspace entry point (ig_dots_f64), through the C ABI: the same kernels and the same
block partition, so the results are bit-identical.  Two calls in a row, the second with a
smaller input (the scratch array keeps the first call's partials; nothing may leak into the
second); one, two and three products, weighted and not."""

import pytest
import torch

from isogriddft_amd import _lib

pytestmark = pytest.mark.gpu


def _both(nd, n, weighted, seed):
    g = torch.Generator().manual_seed(seed)
    vecs = [torch.randn(n, dtype=torch.float64, generator=g).cuda() for _ in range(6)]
    w = torch.rand(n, dtype=torch.float64, generator=g).cuda() if weighted else None
    P = _lib.ptr
    args = [P(vecs[i]) if i < 2 * nd else None for i in range(6)]
    h = _lib.stream_handle()
    work = torch.empty(_lib.load().ig_reduction_work_size(n), dtype=torch.float64, device="cuda")
    ref = torch.full((4,), -1.0, dtype=torch.float64, device="cuda")
    out = torch.full((4,), -2.0, dtype=torch.float64, device="cuda")
    _lib.call("ig_dots_f64", nd, *args, P(w), n, P(work), P(ref), 0, 2, 3, h)
    _lib.call("ig_dots_noalloc_f64", nd, *args, P(w), n, P(out), 0, 2, 3, h)
    torch.cuda.synchronize()
    exact = [float((vecs[2 * i] * vecs[2 * i + 1] * (w if weighted else 1.0)).sum()) for i in range(nd)]
    return ref.cpu(), out.cpu(), exact


@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("weighted", [False, True])
def test_noalloc_dots_match_workspace_dots(nd, weighted):
    slots = (0, 2, 3)
    for n, seed in ((3_000_000, 1), (777, 2)):  # large (1024 blocks) then small (1 block)
        ref, out, exact = _both(nd, n, weighted, seed + 10 * nd)
        for i in range(nd):
            assert out[slots[i]].item() == ref[slots[i]].item(), (n, i)
            assert abs(out[slots[i]].item() - exact[i]) <= 1e-10 * max(1.0, abs(exact[i])), (n, i)
        # untouched output slots keep their values
        assert out[1].item() == -2.0


def test_noalloc_rejects_invalid_arguments():
    with pytest.raises(ValueError):
        _lib.call("ig_dots_noalloc_f64", 4, None, None, None, None, None, None, None, 10, None, 0, 0, 0,
                  _lib.stream_handle())

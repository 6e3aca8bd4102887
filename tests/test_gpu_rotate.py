"""GPU parity of the MFMA basis-rotation kernel Y = Q X (csrc/rotate.hip), the product of
the block iteration's subspace update (ks/eigensolver_jax.py:184-201).

Reference: the FP64 product of the same operands (torch on the GPU).  Bars (max-abs error
over max |reference|): FP64 1e-14, FP32 2e-6 (K <= 300 products of unit-scale operands
accumulated in FP32).  Cases: every row-tile count of both precisions, ragged M, K and N,
padded row strides, the in-place form (Y = the first M rows of X), a row-shifted overlap,
and the argument checks (misaligned overlap, more than 256 rows).
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import _lib
from isogriddft_amd.eigensolver import _mm, _rotate_kernel

pytestmark = pytest.mark.gpu

BARS = {torch.float64: 1e-14, torch.float32: 2e-6}


def _call(Q, X, Y, M, K, N, ct=0):
    sfx = "f32" if X.dtype == torch.float32 else "f64"
    _lib.call(f"ig_rotate_{sfx}", _lib.ptr(Q), Q.stride(0), _lib.ptr(X), X.stride(0), _lib.ptr(Y), Y.stride(0), M, K,
              N, ct, _lib.stream_handle())


def _err(got, ref):
    return float((got.double() - ref).abs().max() / ref.abs().max())


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_rotation_matches_fp64_product(dt):
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    shapes = [(1, 1, 1), (5, 7, 1000), (32, 64, 4097), (33, 65, 3001), (100, 3, 777), (192, 240, 10007),
              (256, 300, 5003), (17, 129, 64), (64, 16, 130)]
    for M, K, N in shapes:
        for ct in (0, 1, 2):
            print(f"case {dt} M={M} K={K} N={N} ct={ct}", flush=True)
            Q = torch.randn((M, K + 3), dtype=torch.float64, device="cuda", generator=g)[:, :K]
            X = torch.randn((K, N + 5), dtype=torch.float64, device="cuda", generator=g)[:, :N]
            ref = Q @ X
            Y = torch.full((M, N + 2), float("nan"), dtype=dt, device="cuda")[:, :N]
            Qd, Xd = Q.to(dt), X.to(dt)
            _call(Qd, Xd, Y, M, K, N, ct)
            assert _err(Y, ref) <= BARS[dt], (dt, M, K, N, ct, _err(Y, ref))
            # the padding columns of Y are left alone
            assert bool(torch.isnan(Y.as_strided((M, 2), (N + 2, 1), Y.storage_offset() + N)).all())


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_rotation_in_place_and_shifted(dt):
    g = torch.Generator(device="cuda")
    g.manual_seed(12)
    for M, K, N in ((48, 60, 5000), (192, 240, 20000), (200, 200, 3000)):
        print(f"case in place {dt} M={M} K={K} N={N}", flush=True)
        Q = torch.randn((M, K), dtype=dt, device="cuda", generator=g)
        B = torch.randn((K + 4, N), dtype=dt, device="cuda", generator=g)
        ref = Q.double() @ B[:K].double()
        tail = B[K:].clone()
        _call(Q, B, B, M, K, N)
        assert _err(B[:M], ref) <= BARS[dt]
        assert torch.equal(B[K:], tail)
        # Y = rows [4, 4 + M) over X = rows [0, K): the S Z rotation of the block buffer
        B = torch.randn((max(K, 4 + M), N), dtype=dt, device="cuda", generator=g)
        ref = Q.double() @ B[:K].double()
        _call(Q, B, B[4:4 + M], M, K, N)
        assert _err(B[4:4 + M], ref) <= BARS[dt]


def test_rotation_argument_checks():
    X = torch.zeros((8, 100), dtype=torch.float64, device="cuda")
    Q = torch.zeros((4, 8), dtype=torch.float64, device="cuda")
    flat = X.reshape(-1)
    Y = flat[3:3 + 400].view(4, 100)  # overlaps X off by 3 columns
    with pytest.raises(ValueError):
        _call(Q, X, Y, 4, 8, 100)
    Qb = torch.zeros((257, 8), dtype=torch.float64, device="cuda")
    Yb = torch.zeros((257, 100), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        _call(Qb, X, Yb, 257, 8, 100)
    assert not _rotate_kernel(Qb, X, Yb)
    # _mm falls back to the vendor product for more rows
    Qr = torch.randn((300, 8), dtype=torch.float64, device="cuda")
    Xr = torch.randn((8, 100), dtype=torch.float64, device="cuda")
    assert _err(_mm(Qr, Xr), Qr @ Xr) <= 1e-14
    # ragged ldx < N is rejected
    with pytest.raises(ValueError):
        _lib.call("ig_rotate_f64", _lib.ptr(Q), 8, _lib.ptr(X), 50, _lib.ptr(Yb), 100, 4, 8, 100, 0,
                  _lib.stream_handle())


def test_mm_uses_rotation_kernel_with_host_coefficients():
    """The eigensolver passes host FP64 coefficient matrices; they are moved and cast."""
    X = torch.randn((40, 3000), dtype=torch.float32, device="cuda")
    Qh = torch.from_numpy(np.random.default_rng(3).standard_normal((24, 40)))
    got = _mm(Qh, X)
    ref = Qh.cuda() @ X.double()
    assert got.dtype == torch.float32 and _err(got, ref) <= BARS[torch.float32]

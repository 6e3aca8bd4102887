"""GPU parity of the split-K MFMA product C = A B^T (ig_gemm_nt_f32 / ig_gemm_nt_f64), the
Gram blocks of the block iteration.

Reference: the FP64 product of the same operands on the host (NumPy).  Bars: FP64 within
1e-13 of the entries' scale sqrt(sum_k A_ik^2 sum_k B_jk^2); FP32 (FP32 accumulation inside
each K panel, FP64 across panels) within 5e-6 of that scale.  Covered: every output-block
layout (4 wavefronts of one row tile, 8 of one, 8 of two, several row and column blocks), ragged
m / n / K (partial tiles, a partial last K chunk), padded row strides, rows that alias the
same storage, and the refusal of misaligned rows.
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import _lib, eigensolver

pytestmark = pytest.mark.gpu


def _run(A, B):
    a, K = A.shape
    b = B.shape[0]
    lib = _lib.load()
    wsize = lib.ig_gemm_nt_work_size(a, b, K)
    assert wsize >= a * b
    work = torch.empty(wsize, dtype=torch.float64, device="cuda")
    G = torch.full((a, b), float("nan"), dtype=torch.float64, device="cuda")
    sfx = "f32" if A.dtype == torch.float32 else "f64"
    _lib.call(f"ig_gemm_nt_{sfx}", _lib.ptr(A), A.stride(0), a, _lib.ptr(B), B.stride(0), b, K, _lib.ptr(work),
              wsize, _lib.ptr(G), b, _lib.stream_handle())
    torch.cuda.synchronize()
    return G.cpu().numpy()


def _check(A, B, tol):
    G = _run(A, B)
    An = A.double().cpu().numpy()
    Bn = B.double().cpu().numpy()
    ref = An @ Bn.T
    scale = np.sqrt(np.outer((An * An).sum(1), (Bn * Bn).sum(1)))
    err = np.max(np.abs(G - ref) / scale)
    assert np.all(np.isfinite(G))
    assert err <= tol, err
    return err


SHAPES = [(1, 1, 77), (17, 5, 1000), (48, 48, 262144 + 13), (64, 96, 200003), (100, 70, 300000),
          (240, 96, 1 << 20), (256, 96, 500009), (300, 130, 100001)]


@pytest.mark.parametrize("m,n,K", SHAPES)
def test_gemm_nt_matches_fp64_product(m, n, K):
    print(f"case m={m} n={n} K={K}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(m * 1000 + n)
    ld = -(-K // 4) * 4  # rows padded to 16 bytes (the kernel's 128-bit loads)
    for dt, tol in ((torch.float64, 1e-13), (torch.float32, 5e-6)):
        A = torch.randn((m, ld), generator=g, device="cuda", dtype=dt)[:, :K]
        B = torch.randn((n, ld), generator=g, device="cuda", dtype=dt)[:, :K]
        _check(A, B, tol)


def test_gemm_nt_padded_strides_and_aliased_rows():
    print("case padded strides", flush=True)
    g = torch.Generator(device="cuda").manual_seed(5)
    for dt, tol in ((torch.float64, 1e-13), (torch.float32, 5e-6)):
        K = 150001
        ld = K + 7 if dt == torch.float64 else K + 11
        ld = -(-ld // 4) * 4  # 16-byte aligned rows in both precisions
        S = torch.randn((80, ld), generator=g, device="cuda", dtype=dt)
        A = S[:, :K]
        _check(A, A, tol)              # Z Z^T
        _check(A[10:70], A[:40], tol)  # overlapping row ranges of one block


def test_gemm_nt_refuses_misaligned_rows():
    A = torch.randn((8, 1001), device="cuda", dtype=torch.float32)
    lib = _lib.load()
    wsize = lib.ig_gemm_nt_work_size(8, 8, 1001)
    work = torch.empty(wsize, dtype=torch.float64, device="cuda")
    G = torch.empty((8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        _lib.call("ig_gemm_nt_f32", _lib.ptr(A), A.stride(0), 8, _lib.ptr(A), A.stride(0), 8, 1001, _lib.ptr(work),
                  wsize, _lib.ptr(G), 8, _lib.stream_handle())
    # the eigensolver's product falls back to the library GEMMs for such operands
    ref = A.double() @ A.double().T
    out = eigensolver._gemm_nt_split(A, A)
    assert torch.allclose(out, ref, rtol=1e-5, atol=1e-3)


def test_gemm_nt_too_small_workspace_is_refused():
    A = torch.randn((32, 4096), device="cuda", dtype=torch.float64)
    work = torch.empty(10, dtype=torch.float64, device="cuda")
    G = torch.empty((32, 32), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        _lib.call("ig_gemm_nt_f64", _lib.ptr(A), A.stride(0), 32, _lib.ptr(A), A.stride(0), 32, 4096, _lib.ptr(work),
                  10, _lib.ptr(G), 32, _lib.stream_handle())

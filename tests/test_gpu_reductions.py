"""GPU parity of the weighted dots and Gram products (FP64 bar 1e-13, FP32 bar 2e-6)."""

import numpy as np
import pytest
import torch

from isogriddft_amd import reductions

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,N", [(40, 70, 5003), (16, 16, 884736 // 8), (100, 3, 777), (65, 130, 20000)])
def test_gram_f64(m, n, N):
    rng = np.random.default_rng(m + n)
    X = rng.normal(size=(m, N))
    Y = rng.normal(size=(n, N))
    w = rng.uniform(0.1, 2.0, size=N)
    G = reductions.gram(torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda(), torch.from_numpy(w).cuda())
    ref = X @ (w[:, None] * Y.T)
    err = np.max(np.abs(G.cpu().numpy() - ref)) / np.max(np.abs(ref))
    assert err < 1e-13, err


@pytest.mark.parametrize("m,n,N", [(40, 70, 5003), (48, 48, 200000), (7, 130, 999)])
def test_gram_f32_mfma(m, n, N):
    rng = np.random.default_rng(m * n)
    X = rng.normal(size=(m, N)).astype(np.float32)
    Y = rng.normal(size=(n, N)).astype(np.float32)
    w = rng.uniform(0.1, 2.0, size=N).astype(np.float32)
    G = reductions.gram(torch.from_numpy(X).cuda(), torch.from_numpy(Y).cuda(), torch.from_numpy(w).cuda())
    Xw = (X * w).astype(np.float64)  # the kernel rounds X*w to FP32 before the product
    Xw = (X * w).astype(np.float32).astype(np.float64)
    ref = Xw @ Y.astype(np.float64).T
    scale = np.sqrt(np.outer(np.sum(Xw * Xw, 1), np.sum(Y.astype(np.float64) ** 2, 1)))
    err = np.max(np.abs(G.cpu().numpy() - ref) / scale)
    assert err < 2e-6, err


def test_dots_multi():
    rng = np.random.default_rng(5)
    n = 123457
    a, b, c, d, e, f, w = (rng.normal(size=n) for _ in range(7))
    t = [torch.from_numpy(v).cuda() for v in (a, b, c, d, e, f, w)]
    out = torch.zeros(8, dtype=torch.float64, device="cuda")
    reductions.dots_into(out, [(t[0], t[1]), (t[2], t[3]), (t[4], t[5])], t[6], slots=(2, 5, 7))
    o = out.cpu().numpy()
    np.testing.assert_allclose(o[[2, 5, 7]], [np.sum(a * b * w), np.sum(c * d * w), np.sum(e * f * w)],
                               rtol=1e-12)
    out32 = torch.zeros(4, dtype=torch.float64, device="cuda")
    reductions.dots_into(out32, [(t[0].float(), t[1].float())], None)
    ref = np.sum(a.astype(np.float32).astype(np.float64) * b.astype(np.float32).astype(np.float64))
    assert abs(out32[0].item() - ref) < 1e-9 * np.sum(np.abs(a * b))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,N,weighted", [(216, 32, 221184, False), (16, 200, 40000, True), (256, 64, 1001, True),
                                             (64, 256, 2048, False), (1, 1, 3, True), (200, 16, 132651, False)])
def test_gram_narrow_shapes(dtype, m, n, N, weighted):
    """The split-K narrow kernel (one side <= 64 rows, the other <= 256): both operand
    orientations, paired and scalar loads (odd N), with and without weights, and the
    accumulate flag.  FP32 inputs accumulate in FP64 (bar 1e-12 on the FP32 products)."""
    rng = np.random.default_rng(m * 7 + n)
    X = rng.normal(size=(m, N)).astype(dtype)
    Y = rng.normal(size=(n, N)).astype(dtype)
    w = rng.uniform(0.1, 2.0, size=N).astype(dtype) if weighted else None
    # the weights are applied, in the input precision, to the operand the kernel stages as
    # the wide side: X unless X has over 256 rows or Y over 64 (then the roles swap)
    swapped = not (n <= 64 and m <= 256)
    Xw = (X * w).astype(dtype) if weighted and not swapped else X
    Yw = (Y * w).astype(dtype) if weighted and swapped else Y
    ref = Xw.astype(np.float64) @ Yw.astype(np.float64).T
    t = lambda a: None if a is None else torch.from_numpy(a).cuda()  # noqa: E731
    G = reductions.gram(t(X), t(Y), t(w))
    scale = np.sqrt(np.outer(np.sum(Xw.astype(np.float64) ** 2, 1), np.sum(Yw.astype(np.float64) ** 2, 1)))
    assert np.max(np.abs(G.cpu().numpy() - ref) / scale) < 1e-12
    G2 = reductions.gram(t(X), t(Y), t(w), out=G.clone(), accumulate=True)
    assert np.max(np.abs(G2.cpu().numpy() - 2 * ref) / scale) < 1e-12


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_split_k_nt_products(dt):
    """The eigensolver's split-K A B^T (strided-batched panels plus a tail) on row views of
    one buffer, against an FP64 product."""
    from isogriddft_amd.eigensolver import _gemm_nt_split

    N = 20000  # four 4096-wide panels and a 3616-wide tail
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    buf = torch.randn((70, N), dtype=torch.float64, device="cuda", generator=g).to(dt)
    for (a0, a1, b0, b1) in ((0, 40, 40, 70), (5, 65, 0, 12), (0, 70, 0, 70)):
        A, B = buf[a0:a1], buf[b0:b1]
        got = _gemm_nt_split(A, B)
        want = A.double() @ B.double().T
        err = float((got - want).abs().max() / want.abs().max())
        assert got.dtype == torch.float64 and got.shape == (a1 - a0, b1 - b0)
        assert err <= (1e-12 if dt == torch.float64 else 2e-6), err


@pytest.mark.parametrize("N", [70001, 70000])  # scalar path, 16-byte vector path
@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_row_wise_cg_kernels(dt, N):
    """Row dots, the fused CG update and the direction update against torch on a ragged
    (row stride > N) block; FP64 sums (the FP32 case compares with FP64 sums of the same
    FP32 values)."""
    from isogriddft_amd.reductions import cg_direction, cg_update, row_dots

    k, ld = 5, 70016
    g = torch.Generator(device="cuda")
    g.manual_seed(9)

    def block():
        return torch.randn((k, ld), dtype=torch.float64, device="cuda", generator=g).to(dt)[:, :N]

    x, r, p, q = block(), block(), block(), block()
    w = torch.rand(N, dtype=torch.float64, device="cuda", generator=g).to(dt) + 0.5
    dinv = torch.rand(N, dtype=torch.float64, device="cuda", generator=g).to(dt) + 0.1
    tol = 1e-13 if dt == torch.float64 else 1e-6
    d = row_dots(p, q, w)
    want = (p.double() * q.double() * w.double()).sum(1)
    assert torch.allclose(d, want, rtol=tol, atol=tol * want.abs().max().item())
    assert torch.equal(row_dots(p, q, w), d)  # deterministic
    d0 = row_dots(p, q, None)
    assert torch.allclose(d0, (p.double() * q.double()).sum(1), rtol=tol, atol=tol * N)
    alpha = torch.linspace(-0.7, 0.9, k, dtype=torch.float64, device="cuda")
    x0, r0 = x.clone(), r.clone()
    rz = cg_update(x, r, p, q, dinv, w, alpha)
    a = alpha.to(dt)[:, None]
    xw = x0 + a * p
    rw = r0 - a * q
    assert torch.allclose(x, xw, rtol=0, atol=4 * torch.finfo(dt).eps * xw.abs().max().item())
    assert torch.allclose(r, rw, rtol=0, atol=4 * torch.finfo(dt).eps * rw.abs().max().item())
    rzw = (r.double() ** 2 * dinv.double() * w.double()).sum(1)
    assert torch.allclose(rz, rzw, rtol=tol, atol=0)
    beta = torch.linspace(0.1, 1.3, k, dtype=torch.float64, device="cuda")
    p0 = p.clone()
    cg_direction(p, r, dinv, beta)
    pw = dinv * r + beta.to(dt)[:, None] * p0
    assert torch.allclose(p, pw, rtol=0, atol=4 * torch.finfo(dt).eps * pw.abs().max().item())


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("N", [70000, 70001])
def test_cg_start_and_finish(dt, N):
    """The fused first residual (r = isw b - q, p = dinv r, sum w r p) and the fused last
    update (y = sw (x + alpha p)) against torch, on ragged blocks whose row strides differ
    (b and y: stride 70008; q, r, p, x: 70016); N = 70001 takes the scalar kernels."""
    from isogriddft_amd.reductions import cg_finish, cg_start

    k = 7
    g = torch.Generator(device="cuda")
    g.manual_seed(11)

    def block(ld):
        return torch.randn((k, ld), dtype=torch.float64, device="cuda", generator=g).to(dt)[:, :N]

    b, q, x, p = block(70008), block(70016), block(70016), block(70016)
    r, pv = block(70016), block(70016)
    w = torch.rand(N, dtype=torch.float64, device="cuda", generator=g).to(dt) + 0.5
    dinv = torch.rand(N, dtype=torch.float64, device="cuda", generator=g).to(dt) + 0.1
    isw = torch.rand(N, dtype=torch.float64, device="cuda", generator=g).to(dt) + 0.2
    eps = torch.finfo(dt).eps
    rz = cg_start(b, isw, q, dinv, w, r, pv)
    rw = isw * b - q
    assert torch.allclose(r, rw, rtol=0, atol=4 * eps * rw.abs().max().item())
    pw = dinv * r
    assert torch.allclose(pv, pw, rtol=0, atol=2 * eps * pw.abs().max().item())
    rzw = (r.double() * pv.double() * w.double()).sum(1)
    assert torch.allclose(rz, rzw, rtol=1e-13 if dt == torch.float64 else 1e-6, atol=0)
    assert torch.equal(cg_start(b, isw, q, dinv, w, r, pv), rz)  # deterministic
    rz0 = cg_start(b, isw, q, dinv, None, r, pv)
    assert torch.allclose(rz0, (r.double() * pv.double()).sum(1), rtol=1e-13 if dt == torch.float64 else 1e-6)

    alpha = torch.linspace(-0.6, 1.1, k, dtype=torch.float64, device="cuda")
    y = block(70008)
    x0 = x.clone()
    cg_finish(x, p, isw, alpha, y)
    yw = isw * (x0 + alpha.to(dt)[:, None] * p)
    assert torch.equal(x, x0)
    assert torch.allclose(y, yw, rtol=0, atol=4 * eps * yw.abs().max().item())
    with pytest.raises(ValueError):
        cg_finish(x, p, isw[:-1], alpha, y)
    with pytest.raises(ValueError):
        cg_start(b, isw, q, dinv.to(torch.float64 if dt == torch.float32 else torch.float32), w, r, pv)


def test_fp32_block_products_at_large_n():
    """(M, K) x (K, N) FP32 products at N = 224^3 through the eigensolver's panelled GEMM:
    (192, 248) is one of the shapes for which a single library FP32 GEMM call was measured
    to return wrong values there; the panelled product must match FP64 to FP32 rounding."""
    from isogriddft_amd.eigensolver import _mm

    N = 224 ** 3
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    for M, K in ((192, 248), (48, 200)):
        B = torch.randn((K, N), dtype=torch.float32, device="cuda", generator=g)
        A = torch.randn((M, K), dtype=torch.float32, device="cuda", generator=g)
        got = _mm(A, B)
        ref = A.double() @ B.double()
        err = float((got.double() - ref).abs().max() / ref.abs().max())
        print(M, K, err)
        assert err < 1e-5
        out = torch.empty((M + 8, N), dtype=torch.float32, device="cuda")
        _mm(A, B, out=out[:M])
        assert torch.equal(out[:M], got)
        del B, A, got, ref, out
        torch.cuda.empty_cache()


def test_fp32_split_k_products_at_large_n():
    """The eigensolver's split-K Gram products in FP32 at N = 224^3 (one strided-batched
    library GEMM over 2^8 K panels) against FP64, for the shapes of a k = 48, m = 192
    iteration."""
    from isogriddft_amd.eigensolver import _gemm_nt_split

    N = 224 ** 3
    g = torch.Generator(device="cuda")
    g.manual_seed(12)
    buf = torch.randn((288, N), dtype=torch.float32, device="cuda", generator=g)
    for (a0, a1, b0, b1) in ((0, 240, 192, 288), (192, 240, 0, 192), (192, 240, 192, 240)):
        got = _gemm_nt_split(buf[a0:a1], buf[b0:b1])
        ref = buf[a0:a1].double() @ buf[b0:b1].double().T
        err = float((got - ref).abs().max() / ref.abs().max())
        print((a0, a1, b0, b1), err)
        assert err < 1e-5

"""Weighted reductions and Gram products on the GPU.

``wdot``      sum a*b*w with FP64 accumulation     (ops/kinetic.py:61-74)
``gram``      X W Y^T for orbital-major blocks       (ks/eigensolver.py:246-272,
              ops/reductions_jax.py:89-104); FP32 blocks run on MFMA.
Workspaces come from the torch caching allocator and are reused per device.
"""

from __future__ import annotations

import torch

from . import _lib

_WORK: dict = {}


def workspace(device, n_doubles: int, slot: str = "red") -> torch.Tensor:
    key = (str(device), slot)
    t = _WORK.get(key)
    if t is None or t.numel() < n_doubles:
        t = torch.empty(max(n_doubles, 1024), dtype=torch.float64, device=device)
        _WORK[key] = t
    return t


@_lib.on_stream
def dots_into(out: torch.Tensor, pairs, w=None, slots=(0, 1, 2), stream=None):
    """out[slots[i]] = sum(a_i * b_i [* w]) for up to three pairs (device-resident)."""
    nd = len(pairs)
    a0, b0 = pairs[0]
    dt = a0.dtype
    n = a0.numel()
    lib = _lib.load()
    work = workspace(a0.device, lib.ig_reduction_work_size(n))
    ptrs = []
    for i in range(3):
        if i < nd:
            ptrs += [_lib.ptr(pairs[i][0]), _lib.ptr(pairs[i][1])]
        else:
            ptrs += [None, None]
    if w is not None and w.dtype != dt:
        w = w.to(dt)
    sfx = "f64" if dt == torch.float64 else "f32"
    _lib.call(f"ig_dots_{sfx}", nd, *ptrs, _lib.ptr(w), n, _lib.ptr(work), _lib.ptr(out),
              slots[0], slots[1] if nd > 1 else 0, slots[2] if nd > 2 else 0,
              _lib.stream_handle(stream))
    return out


def wdot(dg, a: torch.Tensor, b: torch.Tensor | None, weighted: bool = True) -> float:
    """Host scalar sum(a * b * w) (b None -> sum(a * w))."""
    a = a.reshape(-1).contiguous()
    if b is None:
        b = torch.ones_like(a)
    b = b.reshape(-1).contiguous()
    out = torch.zeros(4, dtype=torch.float64, device=a.device)
    dots_into(out, [(a, b)], dg.weights(a.dtype) if weighted else None)
    return float(out[0].item())


def gram(X: torch.Tensor, Y: torch.Tensor, w: torch.Tensor, out: torch.Tensor | None = None,
         accumulate: bool = False, stream=None) -> torch.Tensor:
    """G = X W Y^T (FP64 result) for X (m, N), Y (n, N) of the same dtype (w None: W = I)."""
    if X.dtype != Y.dtype:
        raise ValueError("X and Y must share a dtype")
    X2 = X.reshape(X.shape[0], -1)
    Y2 = Y.reshape(Y.shape[0], -1)
    if X2.stride(1) != 1 or Y2.stride(1) != 1:
        raise ValueError("Gram operands must be row-contiguous")
    m, N = X2.shape
    n = Y2.shape[0]
    if Y2.shape[1] != N or (w is not None and w.numel() != N):
        raise ValueError("weights must match the orbital grid size.")
    if out is None:
        out = torch.zeros((m, n), dtype=torch.float64, device=X.device)
    elif out.dtype != torch.float64 or tuple(out.shape) != (m, n) or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous ({m}, {n}) float64 tensor; received "
                         f"{tuple(out.shape)} {out.dtype}.")
    lib = _lib.load()
    need = lib.ig_gram_work_size(m, n, N)
    work = workspace(X.device, need, "gram")
    sfx = "f64" if X.dtype == torch.float64 else "f32"
    ww = None
    if w is not None:
        ww = w.reshape(-1)
        if ww.dtype != X.dtype:
            ww = ww.to(X.dtype)
        ww = ww.contiguous()
    _lib.call(f"ig_gram_{sfx}", _lib.ptr(X2), X2.stride(0), m, _lib.ptr(Y2), Y2.stride(0), n,
              _lib.ptr(ww), N, _lib.ptr(out), int(accumulate), _lib.ptr(work), work.numel(),
              _lib.stream_handle(stream))
    return out


# ------------------------------------------------------------------ row-wise block CG


def _rows(t):
    if t.dim() != 2 or t.stride(1) != 1:
        raise ValueError("row-wise kernels take (k, N) blocks with unit column stride")
    return t


def _field(f, like, name, optional=False):
    """A per-point vector (N values, contiguous, the blocks' dtype) shared by every row."""
    if f is None:
        if optional:
            return None
        raise ValueError(f"{name} is required")
    if f.dim() != 1 or f.numel() != like.shape[1] or not f.is_contiguous() or f.dtype != like.dtype:
        raise ValueError(f"{name} must be a contiguous vector of N = {like.shape[1]} values of dtype {like.dtype}")
    return f


def _alpha(a, k, name):
    if a.dtype != torch.float64 or a.numel() != k or not a.is_contiguous():
        raise ValueError(f"{name} must be a contiguous FP64 device vector of length k = {k}")
    return a


@_lib.on_stream
def row_dots(a: torch.Tensor, b: torch.Tensor, w: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    """FP64 device vector out[i] = sum_n w[n] a[i, n] b[i, n] (w None: 1)."""
    a, b = _rows(a), _rows(b)
    k, N = a.shape
    if b.shape != a.shape or b.stride(0) != a.stride(0) or b.dtype != a.dtype:
        raise ValueError("row_dots operands must share shape, row stride and dtype")
    _field(w, a, "w", optional=True)
    lib = _lib.load()
    work = workspace(a.device, lib.ig_row_dots_work_size(k), "rows")
    out = torch.empty(k, dtype=torch.float64, device=a.device)
    sfx = "f64" if a.dtype == torch.float64 else "f32"
    _lib.call(f"ig_row_dots_{sfx}", _lib.ptr(a), _lib.ptr(b), _lib.ptr(w), k, N, a.stride(0), _lib.ptr(work),
              _lib.ptr(out), _lib.stream_handle(stream))
    return out


@_lib.on_stream
def cg_update(x, r, p, q, dinv, w, alpha, stream=None) -> torch.Tensor:
    """Row i: x += alpha_i p, r -= alpha_i q; returns the FP64 vector sum w dinv r_i^2."""
    for t in (x, r, p, q):
        _rows(t)
        if t.shape != x.shape or t.stride(0) != x.stride(0) or t.dtype != x.dtype:
            raise ValueError("cg_update blocks must share shape, row stride and dtype")
    k, N = x.shape
    _field(dinv, x, "dinv"), _field(w, x, "w", optional=True), _alpha(alpha, k, "alpha")
    lib = _lib.load()
    work = workspace(x.device, lib.ig_row_dots_work_size(k), "rows")
    out = torch.empty(k, dtype=torch.float64, device=x.device)
    sfx = "f64" if x.dtype == torch.float64 else "f32"
    _lib.call(f"ig_cg_update_{sfx}", _lib.ptr(x), _lib.ptr(r), _lib.ptr(p), _lib.ptr(q), _lib.ptr(dinv),
              _lib.ptr(w), _lib.ptr(alpha), k, N, x.stride(0), _lib.ptr(work), _lib.ptr(out),
              _lib.stream_handle(stream))
    return out


@_lib.on_stream
def cg_direction(p, r, dinv, beta, stream=None) -> None:
    """Row i: p = dinv r + beta_i p."""
    _rows(p), _rows(r)
    k, N = p.shape
    if r.shape != p.shape or r.stride(0) != p.stride(0) or r.dtype != p.dtype:
        raise ValueError("cg_direction blocks must share shape, row stride and dtype")
    _field(dinv, p, "dinv"), _alpha(beta, k, "beta")
    sfx = "f64" if p.dtype == torch.float64 else "f32"
    _lib.call(f"ig_cg_direction_{sfx}", _lib.ptr(p), _lib.ptr(r), _lib.ptr(dinv), _lib.ptr(beta), k, N,
              p.stride(0), _lib.stream_handle(stream))


@_lib.on_stream
def cg_start(b, isw, q, dinv, w, r, p, stream=None) -> torch.Tensor:
    """First residual of the row-wise CG from x0 = dinv isw b (q = A x0 given): row i:
    r = isw b - q, p = dinv r; returns the FP64 vector sum w r_i p_i."""
    for t in (q, r, p):
        _rows(t)
        if t.shape != q.shape or t.stride(0) != q.stride(0) or t.dtype != q.dtype:
            raise ValueError("cg_start blocks q, r, p must share shape, row stride and dtype")
    _rows(b)
    if b.shape != q.shape or b.dtype != q.dtype:
        raise ValueError("cg_start: b must match q in shape and dtype")
    k, N = q.shape
    _field(isw, q, "isw"), _field(dinv, q, "dinv"), _field(w, q, "w", optional=True)
    lib = _lib.load()
    work = workspace(q.device, lib.ig_row_dots_work_size(k), "rows")
    out = torch.empty(k, dtype=torch.float64, device=q.device)
    sfx = "f64" if q.dtype == torch.float64 else "f32"
    _lib.call(f"ig_cg_start_{sfx}", _lib.ptr(b), b.stride(0), _lib.ptr(isw), _lib.ptr(q), _lib.ptr(dinv),
              _lib.ptr(w), _lib.ptr(r), _lib.ptr(p), k, N, q.stride(0), _lib.ptr(work), _lib.ptr(out),
              _lib.stream_handle(stream))
    return out


@_lib.on_stream
def cg_finish(x, p, sw, alpha, y, stream=None) -> None:
    """Last CG update mapped back: row i: y = sw (x + alpha_i p); x is not written."""
    _rows(x), _rows(p), _rows(y)
    if p.shape != x.shape or p.stride(0) != x.stride(0) or p.dtype != x.dtype:
        raise ValueError("cg_finish blocks x, p must share shape, row stride and dtype")
    if y.shape != x.shape or y.dtype != x.dtype:
        raise ValueError("cg_finish: y must match x in shape and dtype")
    k, N = x.shape
    _field(sw, x, "sw"), _alpha(alpha, k, "alpha")
    sfx = "f64" if x.dtype == torch.float64 else "f32"
    _lib.call(f"ig_cg_finish_{sfx}", _lib.ptr(x), _lib.ptr(p), _lib.ptr(sw), _lib.ptr(alpha), _lib.ptr(y),
              y.stride(0), k, N, x.stride(0), _lib.stream_handle(stream))

"""Kinetic and Laplacian operators on the GPU (drop-in for the reference's ops/kinetic.py).

Call surfaces (same names and argument meaning as the reference):
  apply_legacy_laplacian_operator / apply_legacy_kinetic_operator      ops/kinetic.py:117-153
  apply_monitor_grid_laplacian_operator[_trial_boundary_fix]           ops/kinetic.py:279-337
  apply_monitor_grid_kinetic_operator[_trial_boundary_fix]             ops/kinetic.py:340-358
  apply_laplacian_operator / apply_kinetic_operator                    ops/kinetic.py:361-374
  integrate_field / weighted_l2_norm                                   ops/kinetic.py:61-74
  green_identity_terms        audit/h2_monitor_grid_kinetic_green_identity_audit.py:143-218, :295-330
  green_identity_defect       <u, T v>_W - <T u, v>_W (the operator/adjoint mismatch)
NumPy in -> NumPy out (FP64), as the reference; a CUDA tensor in -> a CUDA tensor
out in the tensor's dtype (float32 or float64), which is the device-resident path
used by the solvers.  ``apply_operator`` is the block-level entry point.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import DeviceGrid, device_grid

CLOSURES = {"trial_fix": 0, "production": 1}


def _closure_id(closure: str) -> int:
    try:
        return CLOSURES[closure]
    except KeyError:
        raise ValueError("closure must be `production` or `trial_fix`; received " f"`{closure}`.") from None


@_lib.on_stream
def apply_operator(
    dg: DeviceGrid,
    x: torch.Tensor,
    out: torch.Tensor | None = None,
    *,
    scale: float = 1.0,
    closure: str = "production",
    veff: torch.Tensor | None = None,
    transpose: bool = False,
    interior_only: bool = False,
    xchunk: int = 0,
    tile: int = 0,
    invj: torch.Tensor | None = None,
    stream=None,
) -> torch.Tensor:
    """out = scale * Lap(x) + veff * x for a block x of shape (k, *shape) or (*shape).

    x, out and veff are CUDA tensors of one dtype (float32 / float64); the geometry
    copy of that dtype is used.  ``closure`` is ignored on legacy grids.  ``invj``
    (monitor grids) replaces the 1/J factor of the divergence: with invj = s / J and
    veff = s V the output comes out scaled point-wise by s at no extra pass.
    """
    if not x.is_cuda:
        raise ValueError("apply_operator takes a CUDA tensor")
    if x.dtype not in (torch.float32, torch.float64):
        raise ValueError("only float32 / float64 blocks are supported")
    shape = dg.shape
    if tuple(x.shape[-3:]) != shape:
        raise ValueError(f"field must match the grid shape {shape}; received {tuple(x.shape[-3:])}.")
    xb = x.reshape(-1, *shape)
    if not xb.is_contiguous():
        xb = xb.contiguous()
    nblk = xb.shape[0]
    if out is None:
        out = torch.empty_like(xb)
    elif (not out.is_cuda or out.dtype != x.dtype or out.numel() != xb.numel() or not out.is_contiguous()):
        # the kernel writes nblk * N values through a raw pointer: a shorter, strided or
        # differently typed `out` would be overrun or silently left unwritten
        raise ValueError(f"out must be a contiguous CUDA tensor of {xb.numel()} {x.dtype} values "
                         f"(the shape of the input block); received {tuple(out.shape)} {out.dtype}.")
    ob = out.reshape(-1, *shape)
    if ob.data_ptr() == xb.data_ptr():
        raise ValueError("out must not alias the input block")
    if veff is not None:
        veff = veff.reshape(-1)
        if veff.numel() != dg.N:
            raise ValueError(f"veff must have {dg.N} points; received {veff.numel()}.")
        if veff.dtype != x.dtype:
            veff = veff.to(x.dtype)
        veff = veff.contiguous()
    dt = x.dtype
    sfx = "f64" if dt == torch.float64 else "f32"
    st = _lib.stream_handle(stream)
    nx, ny, nz = shape
    if invj is not None:
        if dg.kind != "monitor":
            raise ValueError("an invj override applies to monitor grids only")
        if invj.dtype != dt or invj.numel() != dg.N:
            raise ValueError("invj must be a field of the block dtype")
        invj = invj.reshape(-1).contiguous()
    if dg.kind == "monitor":
        cl = _closure_id(closure)
        ij = dg.invj(dt) if invj is None else invj
        _lib.call(f"ig_monitor_stencil_{sfx}", _lib.ptr(xb), _lib.ptr(ob), nblk, dg.N, nx, ny, nz,
                  _lib.ptr(dg.get("scoef", dt)), _lib.ptr(ij), _lib.ptr(veff), float(scale), cl,
                  int(transpose), int(interior_only), int(xchunk), int(tile), st)
    else:
        _lib.call(f"ig_legacy_stencil_{sfx}", _lib.ptr(xb), _lib.ptr(ob), nblk, dg.N, nx, ny, nz,
                  _lib.ptr(dg.coef(dt)), _lib.ptr(veff), float(scale), int(transpose),
                  int(interior_only), st)
    return out.reshape(x.shape)



@_lib.on_stream
def apply_operator_dots(dg: DeviceGrid, x: torch.Tensor, out: torch.Tensor, w: torch.Tensor, *, scale: float = 1.0,
                        closure: str = "trial_fix", veff: torch.Tensor | None = None,
                        stream=None) -> torch.Tensor | None:
    """out = scale * Lap(x) + veff * x (trial_fix, forward, monitor grid) for a block x
    (k, *shape) and, from the same pass, the row dots sum_n w[n] x[m][n] out[m][n] (FP64,
    a fixed summation order): the (p, A p) products of a CG step.  Returns the dots (FP64
    device vector of length k), or None, with nothing launched, where the fused kernel
    does not take the block (another closure, a legacy grid, z rows wider than one
    wavefront); the caller then applies the operator and takes the dots separately."""
    if dg.kind != "monitor" or _closure_id(closure) != 0:
        return None
    if not x.is_cuda or x.dtype not in (torch.float32, torch.float64):
        raise ValueError("apply_operator_dots takes a float32 / float64 CUDA block")
    shape = dg.shape
    if tuple(x.shape[-3:]) != shape:
        raise ValueError(f"field must match the grid shape {shape}; received {tuple(x.shape[-3:])}.")
    xb = x.reshape(-1, *shape)
    if not xb.is_contiguous():
        raise ValueError("apply_operator_dots takes a contiguous block")
    nblk = xb.shape[0]
    if not out.is_cuda or out.dtype != x.dtype or out.numel() != xb.numel() or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous CUDA tensor of {xb.numel()} {x.dtype} values.")
    ob = out.reshape(-1, *shape)
    if ob.data_ptr() == xb.data_ptr():
        raise ValueError("out must not alias the input block")
    w = w.reshape(-1)
    if w.dtype != x.dtype or w.numel() != dg.N or not w.is_contiguous() or not w.is_cuda:
        raise ValueError(f"w must be a contiguous CUDA field of {dg.N} {x.dtype} values.")
    if veff is not None:
        veff = veff.reshape(-1)
        if veff.numel() != dg.N:
            raise ValueError(f"veff must have {dg.N} points; received {veff.numel()}.")
        veff = veff.to(x.dtype).contiguous()
    from .reductions import workspace

    nx, ny, nz = shape
    lib = _lib.load()
    nwork = int(lib.ig_monitor_stencil_dots_work_size(nblk, nx, ny, nz))
    work = workspace(x.device, nwork, "stencil_dots")
    dots = torch.empty(nblk, dtype=torch.float64, device=x.device)
    dt = x.dtype
    sfx = "f64" if dt == torch.float64 else "f32"
    st = getattr(lib, f"ig_monitor_stencil_dots_{sfx}")(
        _lib.ptr(xb), _lib.ptr(ob), nblk, dg.N, nx, ny, nz, _lib.ptr(dg.get("scoef", dt)), _lib.ptr(dg.invj(dt)),
        _lib.ptr(veff), float(scale), _lib.ptr(w), _lib.ptr(work), work.numel(), _lib.ptr(dots),
        _lib.stream_handle())
    if st == -2:
        return None
    if st != 0:
        if st == -1:
            raise ValueError(f"ig_monitor_stencil_dots_{sfx}: invalid arguments (status {st})")
        raise RuntimeError(f"ig_monitor_stencil_dots_{sfx}: HIP launch failed (status {st})")
    return dots


def _run(psi, grid_geometry, scale, closure):
    dg = device_grid(grid_geometry)
    if isinstance(psi, torch.Tensor):
        if tuple(psi.shape) != dg.shape:
            raise ValueError(f"psi must match the grid shape {dg.shape}; received {tuple(psi.shape)}.")
        return apply_operator(dg, psi, scale=scale, closure=closure)
    field = np.asarray(psi)
    if field.ndim != 3:
        raise ValueError(f"psi must be a 3D field; received ndim={field.ndim}.")
    if field.shape != dg.shape:
        raise ValueError(f"psi must match the grid shape {dg.shape}; received {field.shape}.")
    x = torch.from_numpy(np.ascontiguousarray(field, dtype=np.float64)).to(dg.device)
    return apply_operator(dg, x, scale=scale, closure=closure).cpu().numpy()


def apply_monitor_grid_laplacian_operator(psi, grid_geometry):
    return _run(psi, grid_geometry, 1.0, "production")


def apply_monitor_grid_laplacian_operator_trial_boundary_fix(psi, grid_geometry):
    return _run(psi, grid_geometry, 1.0, "trial_fix")


def apply_monitor_grid_kinetic_operator(psi, grid_geometry):
    return _run(psi, grid_geometry, -0.5, "production")


def apply_monitor_grid_kinetic_operator_trial_boundary_fix(psi, grid_geometry):
    return _run(psi, grid_geometry, -0.5, "trial_fix")


def apply_legacy_laplacian_operator(psi, grid_geometry):
    return _run(psi, grid_geometry, 1.0, "production")


def apply_legacy_kinetic_operator(psi, grid_geometry):
    return _run(psi, grid_geometry, -0.5, "production")


def apply_laplacian_operator(psi, grid_geometry):
    return _run(psi, grid_geometry, 1.0, "production")


def apply_kinetic_operator(psi, grid_geometry):
    return _run(psi, grid_geometry, -0.5, "production")


def integrate_field(field, grid_geometry):
    """sum(field * cell_volumes) with FP64 accumulation on the device."""
    dg = device_grid(grid_geometry)
    t = field if isinstance(field, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(field, dtype=np.float64)).to(dg.device)
    if tuple(t.shape) != dg.shape:
        raise ValueError(f"field must match the grid shape {dg.shape}; received {tuple(t.shape)}.")
    from .reductions import wdot
    return wdot(dg, t.reshape(-1), None)


def weighted_l2_norm(field, grid_geometry):
    dg = device_grid(grid_geometry)
    t = field if isinstance(field, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(field, dtype=np.float64)).to(dg.device)
    from .reductions import wdot
    v = t.reshape(-1)
    return float(np.sqrt(max(wdot(dg, v, v), 0.0)))


# ---------------------------------------------------------------- Green identity


@dataclass(frozen=True)
class GreenIdentityTerms:
    """Discrete Green-identity energies of one real field, all normalised by <f, f>_W
    (the scalar fields of audit/h2_monitor_grid_kinetic_green_identity_audit.py:88-111)."""

    closure: str
    weighted_norm: float
    operator_kinetic_ha: float
    gradient_kinetic_ha: float
    delta_kinetic_mha: float
    boundary_term_ha: float
    closure_mismatch_mha: float


def _field_tensor(field, dg):
    t = field if isinstance(field, torch.Tensor) else torch.from_numpy(
        np.ascontiguousarray(field, dtype=np.float64)).to(dg.device)
    if tuple(t.shape) != dg.shape:
        raise ValueError(f"field must match the grid shape {dg.shape}; received {tuple(t.shape)}.")
    if t.dtype != torch.float64:
        t = t.to(torch.float64)
    return t.contiguous()


@_lib.on_stream
def green_identity_terms(field, grid_geometry, closure: str = "trial_fix", *, stream=None) -> GreenIdentityTerms:
    """Operator energy <f, T f>_W with T of ``closure``, the gradient-form energy
    1/2 sum w d_a f g^{ab} d_b f (np.gradient edge_order=2), the explicit boundary-flux
    term with F^a = J g^{ab} d_b f of ``closure``, and the closure mismatch
    K_op - K_grad - K_bdry, each divided by <f, f>_W
    (audit/h2_monitor_grid_kinetic_green_identity_audit.py:143-218, :295-330; the
    production-closure numbers follow the earlier audit recorded in
    audit/baselines.py:1473-1545).  Everything runs on the device in FP64."""
    dg = device_grid(grid_geometry)
    if dg.kind != "monitor":
        raise ValueError("the Green-identity terms are defined on monitor grids")
    cl = _closure_id(closure)
    f = _field_tensor(field, dg)
    tf = apply_operator(dg, f, scale=-0.5, closure=closure, stream=stream)
    from .reductions import workspace
    work = workspace(dg.device, int(_lib.load().ig_green_identity_work_size()), slot="green")
    out = torch.empty(4, dtype=torch.float64, device=dg.device)
    nx, ny, nz = dg.shape
    hx, hy, hz = dg.h
    _lib.call("ig_green_identity_f64", _lib.ptr(f), _lib.ptr(tf), _lib.ptr(dg.coef()), _lib.ptr(dg.invj()),
              _lib.ptr(dg.weights()), nx, ny, nz, float(hx), float(hy), float(hz), cl, _lib.ptr(work),
              _lib.ptr(out), _lib.stream_handle(stream))
    k_op, k_grad, k_bdry, den = (float(v) for v in out.cpu())
    if not den > 0.0:
        raise ValueError("field has zero weighted norm")
    return GreenIdentityTerms(
        closure=closure,
        weighted_norm=float(np.sqrt(den)),
        operator_kinetic_ha=k_op / den,
        gradient_kinetic_ha=k_grad / den,
        delta_kinetic_mha=1000.0 * (k_op - k_grad) / den,
        boundary_term_ha=k_bdry / den,
        closure_mismatch_mha=1000.0 * (k_op - k_grad - k_bdry) / den,
    )


@_lib.on_stream
def green_identity_defect(u, v, grid_geometry, closure: str = "production", *, stream=None) -> float:
    """<u, T v>_W - <T u, v>_W: the mismatch between the kinetic operator and its adjoint
    under the quadrature weights (zero to round-off for the zero-ghost closure, whose
    transpose is W T W^-1; nonzero for the one-sided production ends).  Legacy grids use
    the product weights under which the 7-point operator is self-adjoint."""
    dg = device_grid(grid_geometry)
    a = _field_tensor(u, dg)
    b = _field_tensor(v, dg)
    ta = apply_operator(dg, a, scale=-0.5, closure=closure, stream=stream)
    tb = apply_operator(dg, b, scale=-0.5, closure=closure, stream=stream)
    w = dg.operator_weights()
    return float(torch.sum(a.reshape(-1) * tb.reshape(-1) * w) - torch.sum(ta.reshape(-1) * b.reshape(-1) * w))

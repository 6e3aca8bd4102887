"""Kinetic and Laplacian operators on the GPU (drop-in for the reference's ops/kinetic.py).

Call surfaces (same names and argument meaning as the reference):
  apply_legacy_laplacian_operator / apply_legacy_kinetic_operator      ops/kinetic.py:117-153
  apply_monitor_grid_laplacian_operator[_trial_boundary_fix]           ops/kinetic.py:279-337
  apply_monitor_grid_kinetic_operator[_trial_boundary_fix]             ops/kinetic.py:340-358
  apply_laplacian_operator / apply_kinetic_operator                    ops/kinetic.py:361-374
  integrate_field / weighted_l2_norm                                   ops/kinetic.py:61-74
NumPy in -> NumPy out (FP64), as the reference; a CUDA tensor in -> a CUDA tensor
out in the tensor's dtype (float32 or float64), which is the device-resident path
used by the solvers.  ``apply_operator`` is the block-level entry point.
"""

from __future__ import annotations

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
    stream=None,
) -> torch.Tensor:
    """out = scale * Lap(x) + veff * x for a block x of shape (k, *shape) or (*shape).

    x, out and veff are CUDA tensors of one dtype (float32 / float64); the geometry
    copy of that dtype is used.  ``closure`` is ignored on legacy grids.
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
    ob = out.reshape(-1, *shape)
    if ob.data_ptr() == xb.data_ptr():
        raise ValueError("out must not alias the input block")
    if veff is not None:
        veff = veff.reshape(-1)
        if veff.dtype != x.dtype:
            veff = veff.to(x.dtype)
        veff = veff.contiguous()
    dt = x.dtype
    sfx = "f64" if dt == torch.float64 else "f32"
    st = _lib.stream_handle(stream)
    nx, ny, nz = shape
    if dg.kind == "monitor":
        cl = _closure_id(closure)
        _lib.call(f"ig_monitor_stencil_{sfx}", _lib.ptr(xb), _lib.ptr(ob), nblk, dg.N, nx, ny, nz,
                  _lib.ptr(dg.get("scoef", dt)), _lib.ptr(dg.invj(dt)), _lib.ptr(veff), float(scale), cl,
                  int(transpose), int(interior_only), int(xchunk), int(tile), st)
    else:
        _lib.call(f"ig_legacy_stencil_{sfx}", _lib.ptr(xb), _lib.ptr(ob), nblk, dg.N, nx, ny, nz,
                  _lib.ptr(dg.coef(dt)), _lib.ptr(veff), float(scale), int(transpose),
                  int(interior_only), st)
    return out.reshape(x.shape) if out.numel() == x.numel() else out


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

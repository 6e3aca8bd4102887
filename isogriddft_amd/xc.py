"""Spin-polarised LSDA on the GPU (drop-in for the reference's xc/lsda.py call surface).

Two functionals, both Slater exchange plus
  * `lda,vwn` (the reference's default, which it evaluates through PySCF/libxc,
    xc/lsda.py:22, :95-100): VWN5 correlation (Vosko, Wilk, Nusair, Can. J. Phys. 58,
    1200 (1980)) with the VWN spin interpolation -- ig_lsda_vwn_f64;
  * `lda,pw92`: Perdew-Wang 1992 correlation -- ig_lsda_pw92_f64.
Density validation follows xc/lsda.py:57-67.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import device_grid

SUPPORTED_LSDA_FUNCTIONAL = "lda,vwn"
_CANONICAL = {
    "lda,vwn": "lda,vwn", "lda,vwn5": "lda,vwn", "slater,vwn": "lda,vwn", "lda_x,lda_c_vwn": "lda,vwn",
    "lda,pw92": "lda,pw92", "lda,pw": "lda,pw92", "slater,pw92": "lda,pw92", "lda_x,lda_c_pw": "lda,pw92",
}
_ENTRY = {"lda,vwn": "ig_lsda_vwn_f64", "lda,pw92": "ig_lsda_pw92_f64"}
_NEGATIVE_DENSITY_TOLERANCE = 1.0e-14


@dataclass(frozen=True)
class LSDAEvaluation:
    """Field names follow xc/lsda.py:28-39."""

    functional: str
    backend: str
    rho_up: np.ndarray
    rho_down: np.ndarray
    rho_total: np.ndarray
    eps_xc: np.ndarray
    energy_density: np.ndarray
    v_xc_up: np.ndarray
    v_xc_down: np.ndarray


def normalize_functional(functional) -> str:
    name = SUPPORTED_LSDA_FUNCTIONAL if functional is None else str(functional).strip().lower()
    if name not in _CANONICAL:
        raise ValueError(f"The HIP LSDA kernels evaluate `lda,vwn` and `lda,pw92`; received `{functional}`.")
    return _CANONICAL[name]


def validate_density_field(rho, name="rho"):
    values = np.asarray(rho, dtype=np.float64)
    if values.ndim != 3:
        raise ValueError(f"{name} must be a 3D field; received ndim={values.ndim}.")
    if not np.all(np.isfinite(values)):
        raise ValueError(f"{name} must contain only finite values.")
    if np.any(values < -_NEGATIVE_DENSITY_TOLERANCE):
        raise ValueError(f"{name} must be non-negative up to roundoff.")
    return np.maximum(values, 0.0)


@_lib.on_stream
def lsda_device(rho_up: torch.Tensor, rho_down: torch.Tensor, stream=None, functional=None):
    """(eps_xc, v_up, v_down) FP64 on the device; raises ValueError on invalid densities."""
    entry = _ENTRY[normalize_functional(functional)]
    if rho_up.shape != rho_down.shape:
        raise ValueError("rho_up and rho_down must have the same shape.")
    ru = rho_up.to(torch.float64).contiguous().reshape(-1)
    rd = rho_down.to(torch.float64).contiguous().reshape(-1)
    n = ru.numel()
    eps = torch.empty_like(ru)
    vu = torch.empty_like(ru)
    vd = torch.empty_like(ru)
    flag = torch.zeros(1, dtype=torch.int32, device=ru.device)
    _lib.call(entry, _lib.ptr(ru), _lib.ptr(rd), n, _lib.ptr(eps), _lib.ptr(vu), _lib.ptr(vd),
              _lib.ptr(flag), _lib.stream_handle(stream))
    if int(flag.item()) != 0:
        raise ValueError("densities must be finite and non-negative up to roundoff.")
    shp = rho_up.shape
    return eps.view(shp), vu.view(shp), vd.view(shp)


def evaluate_lsda_terms(rho_up, rho_down, functional=SUPPORTED_LSDA_FUNCTIONAL) -> LSDAEvaluation:
    """xc/lsda.py:70-129 on the GPU."""
    name = normalize_functional(functional)
    up = validate_density_field(rho_up, "rho_up")
    dn = validate_density_field(rho_down, "rho_down")
    if up.shape != dn.shape:
        raise ValueError(f"rho_up and rho_down must have the same 3D shape; received {up.shape} and {dn.shape}.")
    dev = torch.device("cuda")
    eps, vu, vd = lsda_device(torch.from_numpy(up).to(dev), torch.from_numpy(dn).to(dev), functional=name)
    tot = up + dn
    eps = eps.cpu().numpy()
    return LSDAEvaluation(name, "hip", up, dn, tot, eps, tot * eps, vu.cpu().numpy(), vd.cpu().numpy())


def evaluate_lsda_energy_density(rho_up, rho_down, functional=SUPPORTED_LSDA_FUNCTIONAL):
    return evaluate_lsda_terms(rho_up, rho_down, functional).eps_xc


def evaluate_lsda_potential(rho_up, rho_down, functional=SUPPORTED_LSDA_FUNCTIONAL):
    ev = evaluate_lsda_terms(rho_up, rho_down, functional)
    return ev.v_xc_up, ev.v_xc_down


def evaluate_lsda_energy(rho_up, rho_down, grid_geometry, functional=SUPPORTED_LSDA_FUNCTIONAL) -> float:
    """E_xc = int rho eps_xc (xc/lsda.py:161-174), FP64 reduction on the GPU."""
    from .reductions import wdot

    name = normalize_functional(functional)
    dg = device_grid(grid_geometry)
    up = torch.from_numpy(validate_density_field(rho_up, "rho_up")).to(dg.device)
    dn = torch.from_numpy(validate_density_field(rho_down, "rho_down")).to(dg.device)
    eps, _, _ = lsda_device(up, dn, functional=name)
    return wdot(dg, up + dn, eps)

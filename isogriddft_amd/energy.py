"""Energy components and orbital densities on the GPU (drop-in for the reference's
scf/driver.py:630-690 and :884-1015).

  T_s   = sum_occ <psi|T psi>_W            kinetic block apply + Gram diagonal
  E_loc = int rho V_loc                    ig_vloc_gth_f64 + FP64 weighted dot
  E_nl  = sum_occ <psi|V_nl psi>_W         projections on the support boxes
  E_H   = 1/2 int rho v_H                  open-boundary Poisson solve
  E_xc  = int rho eps_xc                   ig_lsda_vwn_f64 (default) / ig_lsda_pw92_f64
  E_II  = sum_{i<j} Z_i Z_j / R_ij         host (a few atoms)
Densities sum_m occ_m psi_m^2 come from ig_density_*.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, pseudo
from .device import DeviceGrid, device_grid
from .hamiltonian import H2_BENCHMARK_CASE
from .kinetic import apply_operator
from .reductions import gram, wdot


@dataclass(frozen=True)
class SpinOccupations:
    """Fields of the reference's SpinOccupations (scf/driver.py)."""

    n_alpha: float
    n_beta: float
    occupations_up: tuple
    occupations_down: tuple


@dataclass(frozen=True)
class SinglePointEnergyComponents:
    kinetic: float
    local_ionic: float
    nonlocal_ionic: float
    hartree: float
    xc: float
    ion_ion_repulsion: float
    total: float


def evaluate_ion_ion_repulsion(case=H2_BENCHMARK_CASE) -> float:
    """scf/driver.py:884-902 with the GTH valence charges."""
    atoms = pseudo.case_atoms(case)
    e = 0.0
    for i, (ei, pi) in enumerate(atoms):
        for ej, pj in atoms[i + 1:]:
            d = float(np.linalg.norm(np.subtract(pi, pj)))
            if d <= 0.0:
                raise ValueError("Distinct ions must not occupy the same position.")
            e += pseudo.GTH_PADE[ei].ionic_charge * pseudo.GTH_PADE[ej].ionic_charge / d
    return e


def density_device(dg: DeviceGrid, orbitals: torch.Tensor, occupations, out=None, accumulate=False, stream=None):
    """rho = sum_m occ_m psi_m^2 (FP64, flat) for an orbital-major block."""
    blk = orbitals.reshape(-1, dg.N)
    occ = torch.as_tensor(np.asarray(occupations, dtype=np.float64)).to(dg.device)
    n = occ.numel()
    if out is None:
        out = torch.zeros(dg.N, dtype=torch.float64, device=dg.device)
        accumulate = False
    if n == 0:
        if not accumulate:
            out.zero_()
        return out
    if blk.shape[0] < n:
        raise ValueError("Not enough orbitals were provided for the requested occupations.")
    if blk.stride(1) != 1:
        blk = blk.contiguous()
    sfx = "f64" if blk.dtype == torch.float64 else "f32"
    _lib.call(f"ig_density_{sfx}", _lib.ptr(blk), blk.stride(0), n, dg.N, _lib.ptr(occ), _lib.ptr(out),
              int(accumulate), _lib.stream_handle(stream))
    return out


def renormalize_density_device(dg: DeviceGrid, rho: torch.Tensor, target: float) -> torch.Tensor:
    """scf/driver.py:679-690."""
    if target == 0.0:
        return torch.zeros_like(rho)
    integral = wdot(dg, rho, None)
    if integral <= 0.0:
        raise ValueError("A positive density integral is required for renormalization.")
    return rho * (target / integral)


def expectation_values(dg, X, AX):
    """<x_m | A x_m>_W for every orbital of a block (FP64, host)."""
    k = X.reshape(-1, dg.N).shape[0]
    G = gram(X.reshape(k, -1), AX.reshape(k, -1), dg.weights())
    return torch.diagonal(G).cpu().numpy()


def kinetic_energy_terms(dg, X, closure="production"):
    """<psi_m | T psi_m>_W per orbital."""
    TX = apply_operator(dg, X, scale=-0.5, closure=closure)
    return expectation_values(dg, X, TX)


def energy_components_device(dg, orbitals_up, orbitals_down, occupations, case, rho_up=None, rho_down=None, *,
                             poisson_tolerance=1e-8, poisson_precision="fp64", closure="production",
                             functional=None):
    """All components from device blocks; densities from the orbitals unless given."""
    from . import poisson, xc

    atoms = pseudo.case_atoms(case)
    occ_u = tuple(occupations.occupations_up)
    occ_d = tuple(occupations.occupations_down)
    if rho_up is None:
        rho_up = density_device(dg, orbitals_up, occ_u)
    if rho_down is None:
        rho_down = density_device(dg, orbitals_down, occ_d)
    ru = renormalize_density_device(dg, rho_up.reshape(-1).to(torch.float64), occupations.n_alpha)
    rd = renormalize_density_device(dg, rho_down.reshape(-1).to(torch.float64), occupations.n_beta)
    rho = ru + rd
    plan = pseudo.nonlocal_plan(dg, case)
    kin = 0.0
    enl = 0.0
    for blk, occ in ((orbitals_up, occ_u), (orbitals_down, occ_d)):
        if not occ:
            continue
        X = blk.reshape(-1, dg.N)[:len(occ)].to(torch.float64).contiguous()
        kin += float(np.dot(occ, kinetic_energy_terms(dg, X.view((-1,) + dg.shape), closure)))
        if not plan.empty:
            enl += float(np.dot(occ, plan.energy_terms(X)))
    vloc = pseudo.local_potential_device(dg, atoms)
    eloc = wdot(dg, rho, vloc)
    sol = poisson.solve_poisson_device(dg, rho.clamp_min(0.0), tolerance=poisson_tolerance,
                                       precision=poisson_precision)
    eh = 0.5 * wdot(dg, rho, sol["potential"].reshape(-1))
    eps, _, _ = xc.lsda_device(ru.view(dg.shape), rd.view(dg.shape), functional=functional)
    exc = wdot(dg, rho, eps.reshape(-1))
    eii = evaluate_ion_ion_repulsion(case)
    total = kin + eloc + enl + eh + exc + eii
    return SinglePointEnergyComponents(float(kin), float(eloc), float(enl), float(eh), float(exc), float(eii),
                                       float(total))


def evaluate_single_point_energy(rho_up, rho_down, orbitals_up, orbitals_down, occupations, grid_geometry,
                                 case=H2_BENCHMARK_CASE) -> SinglePointEnergyComponents:
    """scf/driver.py:905-1015 (NumPy in, components out; every term on the GPU)."""
    dg = device_grid(grid_geometry)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64))).to(dg.device)  # noqa: E731
    ou = t(orbitals_up).reshape(-1, dg.N) if len(occupations.occupations_up) else \
        torch.zeros((0, dg.N), dtype=torch.float64, device=dg.device)
    od = t(orbitals_down).reshape(-1, dg.N) if len(occupations.occupations_down) else \
        torch.zeros((0, dg.N), dtype=torch.float64, device=dg.device)
    return energy_components_device(dg, ou, od, occupations, case, t(rho_up), t(rho_down))

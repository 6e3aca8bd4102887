"""Energy components and orbital densities on the GPU (drop-in for the reference's
scf/driver.py:630-690 and :884-1015).

  T_s   = sum_occ <psi|T psi>_W            kinetic block apply + Gram diagonal
  E_loc = int rho V_loc                    ig_vloc_gth_f64 + FP64 weighted dot
  E_nl  = sum_occ <psi|V_nl psi>_W         projections on the support boxes
  E_H   = 1/2 int rho v_H                  open-boundary Poisson solve
  E_xc  = int rho eps_xc                   ig_lsda_vwn_f64 (default) / ig_lsda_pw92_f64
  E_II  = sum_{i<j} Z_i Z_j / R_ij         host (a few atoms)
Densities sum_m occ_m psi_m^2 come from ig_density_*.
The local-only variants (scf/driver.py:2654-2937) drop E_nl and, on monitor grids with
the near-core patch, take E_loc from the patch-corrected quadrature.
"""

from __future__ import annotations

import time
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


@_lib.on_stream
def density_device(dg: DeviceGrid, orbitals: torch.Tensor, occupations, out=None, accumulate=False, stream=None):
    """rho = sum_m occ_m psi_m^2 (FP64, flat) for an orbital-major block."""
    if orbitals.dtype not in (torch.float32, torch.float64):
        raise ValueError(f"orbitals must be float32 or float64; received {orbitals.dtype}.")
    blk = orbitals.reshape(-1, dg.N)
    occ = torch.as_tensor(np.asarray(occupations, dtype=np.float64)).to(dg.device)
    n = occ.numel()
    if out is None:
        out = torch.zeros(dg.N, dtype=torch.float64, device=dg.device)
        accumulate = False
    elif out.dtype != torch.float64 or out.numel() != dg.N or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float64 field of {dg.N} points; received "
                         f"{tuple(out.shape)} {out.dtype}.")
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


# ------------------------------------------------------------------ local-only energies


@dataclass(frozen=True)
class StaticLocalEnergyEvaluationProfile:
    """Fields of scf/driver.py:282-289."""

    kinetic_wall_time_seconds: float
    local_ionic_wall_time_seconds: float
    hartree_energy_wall_time_seconds: float
    xc_energy_wall_time_seconds: float
    ion_ion_wall_time_seconds: float


def default_monitor_patch_parameters():
    """scf/driver.py:1018-1024."""
    return pseudo.LocalPotentialPatchParameters(patch_radius_scale=0.75, patch_grid_shape=(25, 25, 25),
                                                correction_strength=1.30, interpolation_neighbors=8)


def _static_local_closure(dg, kinetic_version):
    """scf/driver.py:2654-2665: trial_fix only on monitor grids."""
    return "trial_fix" if (kinetic_version == "trial_fix" and dg.kind == "monitor") else "production"


def _occupied_kinetic(dg, blocks, closure):
    kin = 0.0
    for blk, occ in blocks:
        if not occ:
            continue
        X = blk.reshape(-1, dg.N)[:len(occ)].to(torch.float64).contiguous()
        kin += float(np.dot(occ, kinetic_energy_terms(dg, X.view((-1,) + dg.shape), closure)))
    return kin


def evaluate_static_local_single_point_energy(rho_up, rho_down, orbitals_up, orbitals_down, occupations,
                                              grid_geometry, case=H2_BENCHMARK_CASE, *, use_monitor_patch=False,
                                              patch_parameters=None, kinetic_version="production",
                                              hartree_backend="python", use_jax_hartree_cached_operator=False,
                                              jax_hartree_cg_impl="baseline", jax_hartree_cg_preconditioner="none",
                                              jax_hartree_line_preconditioner_impl="baseline",
                                              monitor_boundary_construction_mode="corrected_moments"):
    """scf/driver.py:2668-2810 on the GPU: T_s + E_loc + E_H + E_xc + E_II (E_nl = 0).
    ``monitor_boundary_construction_mode`` (an addition) selects the Hartree faces as in
    the Poisson surfaces."""
    from . import poisson, xc
    from .reductions import wdot

    dg = device_grid(grid_geometry)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64))).to(dg.device)  # noqa: E731
    ru = renormalize_density_device(dg, t(rho_up).reshape(-1), occupations.n_alpha)
    rd = renormalize_density_device(dg, t(rho_down).reshape(-1), occupations.n_beta)
    rho = ru + rd
    closure = _static_local_closure(dg, kinetic_version)
    blocks = []
    if occupations.n_alpha:
        blocks.append((t(orbitals_up), tuple(occupations.occupations_up)))
    if occupations.n_beta:
        blocks.append((t(orbitals_down), tuple(occupations.occupations_down)))
    kin = _occupied_kinetic(dg, blocks, closure)
    atoms = pseudo.case_atoms(case)
    eloc = wdot(dg, rho, pseudo.local_potential_device(dg, atoms))
    if dg.kind == "monitor" and use_monitor_patch:
        params = default_monitor_patch_parameters() if patch_parameters is None else patch_parameters
        plan = pseudo.patch_plan(dg, case, grid_geometry, params)
        eloc += float(sum(c[3] for c in plan.corrections(dg, rho)))
    sol = poisson.solve_poisson_device(dg, rho.clamp_min(0.0), tolerance=1e-8, mode=monitor_boundary_construction_mode)
    eh = 0.5 * wdot(dg, rho, sol["potential"].reshape(-1))
    eps, _, _ = xc.lsda_device(ru.view(dg.shape), rd.view(dg.shape))
    exc = wdot(dg, rho, eps.reshape(-1))
    eii = evaluate_ion_ion_repulsion(case)
    total = kin + eloc + eh + exc + eii
    return SinglePointEnergyComponents(float(kin), float(eloc), 0.0, float(eh), float(exc), float(eii), float(total))


def evaluate_static_local_single_point_energy_from_context(operator_context, orbitals_up, orbitals_down,
                                                           occupations):
    """scf/driver.py:2813-2937: reuses the frozen context's density, local slice, Hartree
    result and LSDA evaluation."""
    from . import poisson
    from .reductions import wdot

    ctx = operator_context
    geom = ctx.grid_geometry
    dg = device_grid(geom)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64))).to(dg.device)  # noqa: E731
    t0 = time.perf_counter()
    blocks = []
    if occupations.n_alpha:
        blocks.append((t(orbitals_up), tuple(occupations.occupations_up)))
    if occupations.n_beta:
        blocks.append((t(orbitals_down), tuple(occupations.occupations_down)))
    kin = _occupied_kinetic(dg, blocks, _static_local_closure(dg, ctx.kinetic_version))
    t1 = time.perf_counter()
    rho = t(ctx.rho_total).reshape(-1)
    if ctx.frozen_patch_local_embedding is not None:
        eloc = float(ctx.frozen_patch_local_embedding.patch_evaluation.corrected_local_energy)
    elif ctx.local_ionic_evaluation is not None:
        eloc = wdot(dg, rho, t(ctx.local_ionic_evaluation.total_local_potential).reshape(-1))
    else:
        eloc = wdot(dg, rho, t(ctx.local_ionic_potential).reshape(-1))
    t2 = time.perf_counter()
    hp = ctx.hartree_poisson_result if ctx.hartree_poisson_result is not None else ctx.hartree_potential
    eh = poisson.evaluate_hartree_energy(ctx.rho_total, geom, hp)
    t3 = time.perf_counter()
    lsda = ctx.lsda_evaluation
    if lsda is None:
        from . import xc

        lsda = xc.evaluate_lsda_terms(ctx.rho_up, ctx.rho_down)
    exc = wdot(dg, t(lsda.energy_density).reshape(-1), None)
    t4 = time.perf_counter()
    eii = evaluate_ion_ion_repulsion(ctx.case)
    t5 = time.perf_counter()
    total = kin + eloc + eh + exc + eii
    comps = SinglePointEnergyComponents(float(kin), float(eloc), 0.0, float(eh), float(exc), float(eii),
                                        float(total))
    return comps, StaticLocalEnergyEvaluationProfile(t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4)

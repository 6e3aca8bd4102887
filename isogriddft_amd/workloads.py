"""Molecular fixed-potential workloads on atom-centred monitor grids, end to end on the GPU.

``python -m isogriddft_amd.workloads benzene [--n 224] [--k 48]`` runs (``h2_legacy`` runs the
reference's CPU-runnable H2 case on its legacy grid instead, see ``run_h2_legacy``):

  grid       the harmonic-map monitor generator for the molecule (grid/monitor_builder.py
             defaults for H/C/N/O), n^3 points over +-half_extent Bohr
  density    a superposition of normalised atomic Gaussians carrying the GTH valence
             charges (a frozen start density; the reference starts its SCF from a
             superposition as well), spin-unpolarised
  potential  V_loc (GTH-PADE) + v_H (open-boundary Poisson, multipole order 4, multigrid-
             preconditioned BiCGSTAB in FP64 and FP32 + FP64 defect correction; one
             unpreconditioned FP64 solve for comparison) + v_xc (LSDA, lda,vwn) and the
             separable GTH projectors (``--poisson-only`` stops here)
  eigensolve the lowest k states of that fixed potential, in FP64 and in FP32 + FP64
             refinement (kinetic preconditioner, working subspace `block`)
  forces     Hellmann-Feynman forces -dE_ion/dR at the FP64 orbitals (ionic-energy adjoint)

and prints one JSON line with the timings, iteration counts and the mixed-vs-FP64
eigenvalue difference.  Benzene: D6h, C-C 2.637 Bohr, C-H 2.048 Bohr, 30 valence
electrons, k = 48 (15 occupied + 33 buffer states), +-14 Bohr.
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import autograd, configs, eigensolver, monitor_grid, poisson, pseudo, xc
from .device import DeviceGrid
from .hamiltonian import BlockHamiltonian
from .reductions import wdot

# Gaussian exponents (Bohr^-2) of the atomic valence densities of the start density
START_EXPONENTS = {"H": 1.0, "C": 0.55, "N": 0.65, "O": 0.75}

MOLECULES = {"benzene": (configs.BENZENE, 14.0, 48, 15), "h2o": (configs.H2O, 10.0, 8, 4)}


def _log(msg):
    print(f"[workload] {msg}", file=sys.stderr, flush=True)


def start_density(dg, atoms):
    """sum_a Z_a (alpha/pi)^(3/2) exp(-alpha r_a^2), renormalised on the grid to sum Z_a."""
    xyz = dg.xyz()
    rho = torch.zeros(dg.N, dtype=torch.float64, device=dg.device)
    ztot = 0.0
    for el, pos in atoms:
        z = pseudo.GTH_PADE[el].ionic_charge
        a = START_EXPONENTS[el]
        r2 = (xyz[0] - pos[0]) ** 2 + (xyz[1] - pos[1]) ** 2 + (xyz[2] - pos[2]) ** 2
        rho += z * (a / np.pi) ** 1.5 * torch.exp(-a * r2)
        ztot += z
    return rho * (ztot / wdot(dg, rho, None))


EDGE_GAP_HA = 1e-3


def occupation_edge(eigenvalues, nocc, min_gap=EDGE_GAP_HA):
    """(n, gap): the occupied-state count nearest to nocc (ties: the lower) whose edge
    eps[n] - eps[n - 1] is at least min_gap, among 1 <= n < len(eigenvalues); (len, None) if
    no computed edge is that wide."""
    ev = np.asarray(eigenvalues, dtype=np.float64)
    cands = [n for n in range(1, ev.size) if ev[n] - ev[n - 1] >= min_gap]
    if not cands:
        return int(ev.size), None
    n = min(cands, key=lambda c: (abs(c - nocc), c))
    return int(n), float(ev[n] - ev[n - 1])


def energy_components(dg, mol, orbitals, nocc):
    """T_s, E_loc, E_nl, E_H, E_xc (lda,vwn), E_II and the total (Ha) of the closed-shell
    occupation (2 electrons in each of the lowest nocc states) of an orbital block, as the
    reference's evaluate_single_point_energy (scf/driver.py:905-1015) forms them; the
    trial_fix kinetic operator of the eigensolve."""
    from .energy import SpinOccupations, energy_components_device

    orbs = orbitals.reshape(-1, dg.N)[:nocc].to(torch.float64).contiguous()
    occ = SpinOccupations(float(nocc), float(nocc), (1.0,) * nocc, (1.0,) * nocc)
    e = energy_components_device(dg, orbs, orbs, occ, mol, closure="trial_fix", functional="lda,vwn")
    return {"T_s": e.kinetic, "E_loc": e.local_ionic, "E_nl": e.nonlocal_ionic, "E_H": e.hartree, "E_xc": e.xc,
            "E_II": e.ion_ion_repulsion, "E_total": e.total}


def run(name="benzene", n=224, k=None, block=None, tol=3e-4, multipole_order=4, grid_device="cuda",
        poisson_only=False):
    mol, half, k_default, nocc = MOLECULES[name]
    k = k or k_default
    atoms = pseudo.case_atoms(mol)
    out = {"workload": f"{mol.name} fixed-potential eigensolve on a {n}^3 monitor grid (+-{half} Bohr)",
           "k": k, "occupied": nocc}

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    geom = monitor_grid.build_monitor_grid_for_case(
        mol, shape=(n, n, n), box_half_extents=(half, half, half),
        element_parameters=monitor_grid.build_default_near_core_element_parameters(mol), device=grid_device)
    torch.cuda.synchronize()
    out["grid_s"] = round(time.perf_counter() - t0, 3)
    out["grid_min_jacobian"] = float(np.min(geom.jacobian))
    _log(f"grid {n}^3 in {out['grid_s']} s")
    dg = DeviceGrid(geom)

    t0 = time.perf_counter()
    rho = start_density(dg, atoms)
    vloc = pseudo.local_potential_device(dg, atoms)
    half_rho = 0.5 * rho
    _, vxc, _ = xc.lsda_device(half_rho, half_rho)
    torch.cuda.synchronize()
    out["potential_setup_s"] = round(time.perf_counter() - t0, 3)
    pots = {}
    # per-grid set-up: the geometry data of the open-boundary solve (cells of the moment
    # fit), then the multigrid hierarchy of its interior operator
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    poisson.poisson_data(dg)
    torch.cuda.synchronize()
    out["poisson_grid_setup_s"] = round(time.perf_counter() - t0, 4)
    t0 = time.perf_counter()
    poisson.poisson_multigrid(dg)
    torch.cuda.synchronize()
    out["poisson_mg_setup_s"] = round(time.perf_counter() - t0, 4)
    for prec in ("fp64", "mixed"):
        kw = dict(multipole_order=multipole_order, tolerance=1e-8, precision=prec, max_iterations=4000)
        poisson.solve_poisson_device(dg, rho, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol = poisson.solve_poisson_device(dg, rho, **kw)
        torch.cuda.synchronize()
        out[f"poisson_{prec}_s"] = round(time.perf_counter() - t0, 4)
        out[f"poisson_{prec}_iterations"] = sol["iterations"]
        pots[prec] = sol["potential"].reshape(-1)
    out["poisson_mixed_vs_fp64_max_rel"] = float((pots["mixed"] - pots["fp64"]).abs().max()
                                                 / pots["fp64"].abs().max())
    # the same FP64 solve without a preconditioner (the reference's iteration), for comparison
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol = poisson.solve_poisson_device(dg, rho, multipole_order=multipole_order, tolerance=1e-8, max_iterations=4000,
                                       preconditioner="none")
    torch.cuda.synchronize()
    out["poisson_fp64_unpreconditioned_s"] = round(time.perf_counter() - t0, 4)
    out["poisson_fp64_unpreconditioned_iterations"] = sol["iterations"]
    _log(f"poisson fp64 {out['poisson_fp64_s']} s, mixed {out['poisson_mixed_s']} s "
         f"(unpreconditioned fp64 {out['poisson_fp64_unpreconditioned_s']} s)")
    if poisson_only:
        return out
    veff = vloc + pots["fp64"] + vxc.reshape(-1)
    t0 = time.perf_counter()
    plan = pseudo.NonlocalPlan(dg, atoms)
    out["projector_setup_s"] = round(time.perf_counter() - t0, 3)
    H = BlockHamiltonian(dg, veff, closure="trial_fix", nonlocal_plan=plan)
    m = block or 4 * k
    X0 = eigensolver.default_guess_block(geom, atoms, m, dg)
    _log(f"starting block of {m} orbitals built")
    ev = {}
    res64 = None
    for prec in ("fp64", "mixed"):
        eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, tol, 3, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, tol, 600, precision=prec, profile=True)
        torch.cuda.synchronize()
        out[f"eig_{prec}_s"] = round(time.perf_counter() - t0, 3)
        out[f"eig_{prec}_iterations"] = res.iterations
        out[f"eig_{prec}_converged"] = bool(res.converged)
        out[f"eig_{prec}_phase_s"] = {key: round(v, 3) for key, v in res.timings.items()}
        # mean share of the k residual rows that entered the expansion (soft locking)
        out[f"eig_{prec}_active_fraction"] = round(float(np.mean(res.active_rows)) / k, 3) if res.active_rows else 1.0
        ev[prec] = np.asarray(res.eigenvalues)
        if prec == "fp64":
            res64 = res
        else:
            resm = res
            out["eig_fp32_iterations"] = res.fp32_iterations
        _log(f"eigensolve {prec}: {out[f'eig_{prec}_s']} s, {res.iterations} iterations")
    out["eigenvalues_fp64_Ha"] = [round(float(e), 10) for e in ev["fp64"]]
    out["eig_mixed_vs_fp64_abs_Ha"] = [float(abs(a - b)) for a, b in zip(ev["mixed"], ev["fp64"])]
    out["eig_mixed_vs_fp64_max_abs_Ha"] = float(np.max(np.abs(ev["mixed"] - ev["fp64"])))
    # Energy components of a closed-shell occupation, from the orbitals of each precision.
    # They are functions of the occupied subspace, which is only defined where the occupation
    # edge sits in a gap: the lowest states of these fixed potentials come in clusters split by
    # 1e-10 Ha or less (modes of the discrete operator localised in the GTH wells, one set per
    # heavy atom), and an edge inside a cluster makes the occupied density arbitrary.  So the
    # edge is moved to the nearest gap of at least EDGE_GAP_HA; the count used is reported.
    out["occupation_edge_gap_Ha"] = float(ev["fp64"][nocc] - ev["fp64"][nocc - 1]) if k > nocc else None
    n_e, gap = occupation_edge(ev["fp64"], nocc)
    out["energy_occupied_states"] = n_e
    out["energy_edge_gap_Ha"] = gap
    comps = {prec: energy_components(dg, mol, r.orbitals, n_e) for prec, r in (("fp64", res64), ("mixed", resm))}
    out["energy_components_fp64_Ha"] = comps["fp64"]
    out["energy_mixed_vs_fp64_abs_Ha"] = {key: abs(comps["mixed"][key] - comps["fp64"][key]) for key in comps["fp64"]}
    out["eigensolve_plus_poisson_fp64_s"] = round(out["eig_fp64_s"] + out["poisson_fp64_s"], 3)
    out["eigensolve_plus_poisson_mixed_s"] = round(out["eig_mixed_s"] + out["poisson_mixed_s"], 3)

    occupations = [2.0] * nocc
    orbs = res64.orbitals.reshape(k, -1)[:nocc].to(torch.float64)
    rho_orb = (2.0 * orbs * orbs).sum(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    F = autograd.atomic_forces(dg, [el for el, _ in atoms], [p for _, p in atoms], rho_orb, orbs, occupations)
    torch.cuda.synchronize()
    out["forces_s"] = round(time.perf_counter() - t0, 3)
    out["forces_Ha_per_Bohr"] = np.round(F, 8).tolist()
    out["forces_net_Ha_per_Bohr"] = np.round(F.sum(0), 10).tolist()
    _log(f"forces in {out['forces_s']} s")
    return out


def run_h2_legacy(shape=(49, 49, 61), k=2, tol=1e-6):
    """The reference's CPU-runnable case on the GPU: H2 (R = 1.4 Bohr on z), GTH-PADE H,
    LSDA (Slater + PW92), legacy sinh grid 49 x 49 x 61 over +-8 Bohr; the fixed-potential
    eigensolve of the lowest k = 2 states on the potential of a superposition of atomic
    densities, then one open-boundary Poisson solve (order 2, rtol 1e-10) on the singlet
    density of the lowest orbital."""
    mol = configs.H2
    atoms = pseudo.case_atoms(mol)
    out = {"workload": f"H2 legacy {shape[0]}x{shape[1]}x{shape[2]} (+-8 Bohr), lda,pw92, k={k}"}
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    geom = configs.legacy_h2_geometry(shape=shape)
    dg = DeviceGrid(geom)
    rho = start_density(dg, atoms)
    vloc = pseudo.local_potential_device(dg, atoms)
    half_rho = 0.5 * rho
    _, vxc, _ = xc.lsda_device(half_rho, half_rho, functional="lda,pw92")
    vh = poisson.solve_poisson_device(dg, rho, multipole_order=2, tolerance=1e-10, max_iterations=4000)
    H = BlockHamiltonian(dg, vloc + vh["potential"].reshape(-1) + vxc.reshape(-1), closure="production")
    m = eigensolver.reference_block_size(k)
    X0 = eigensolver.default_guess_block(geom, atoms, m, dg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = eigensolver.block_subspace_iteration(H, dg.operator_weights(), X0, k, tol, 600)
    torch.cuda.synchronize()
    out["eig_s"] = round(time.perf_counter() - t0, 4)
    out["eig_iterations"] = res.iterations
    out["eig_converged"] = bool(res.converged)
    out["eigenvalues_Ha"] = [float(e) for e in res.eigenvalues]
    psi = res.orbitals.reshape(k, -1)[0].to(torch.float64)
    psi = psi / float(np.sqrt(wdot(dg, psi, psi)))  # cell-volume normalisation, as the reference returns it
    rho1 = 2.0 * psi * psi
    t0 = time.perf_counter()
    sol = poisson.solve_poisson_device(dg, rho1, multipole_order=2, tolerance=1e-10, max_iterations=4000)
    torch.cuda.synchronize()
    out["poisson_s"] = round(time.perf_counter() - t0, 4)
    out["poisson_iterations"] = sol["iterations"]
    out["hartree_energy_Ha"] = 0.5 * wdot(dg, rho1, sol["potential"].reshape(-1))
    out["total_s"] = round(time.perf_counter() - t_all, 3)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("molecule", choices=sorted(MOLECULES) + ["h2_legacy"])
    ap.add_argument("--n", type=int, default=224, help="grid points per axis")
    ap.add_argument("--k", type=int, default=None, help="states (default: the molecule's k)")
    ap.add_argument("--block", type=int, default=None, help="working subspace (default 4k)")
    ap.add_argument("--tol", type=float, default=3e-4, help="eigensolver residual tolerance")
    ap.add_argument("--poisson-only", action="store_true",
                    help="stop after the Poisson solves (grid, density, set-up, FP64 / mixed / unpreconditioned)")
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("the workloads run on the GPU only", file=sys.stderr)
        sys.exit(2)
    if args.molecule == "h2_legacy":
        print(json.dumps(run_h2_legacy()), flush=True)
        return
    print(json.dumps(run(args.molecule, args.n, args.k, args.block, args.tol, poisson_only=args.poisson_only)),
          flush=True)


if __name__ == "__main__":
    main()

"""A minimal SCF closed loop driven only through the drop-in surfaces, pinned to the
reference's recorded H2 result.

The loop restates the reference's minimal SCF (scf/driver.py:2953-3110) in test code:
the SCF controller itself is outside the path; what is checked is that the path's
surfaces (solve_fixed_potential_eigenproblem, evaluate_single_point_energy) close
the loop on the reference's fixed point.  Reference: H2 at R = 1.4 Bohr, singlet,
lda,vwn, legacy 51^3 grid over (8, 8, 10) Bohr.  Its recorded run
(audit/baselines.py:1924-1945) stopped after 5 iterations of mixing 0.6, at a density
residual of 1.75e-3 and an energy change of -1.7e-6 Ha, with E_total =
-1.1461203815144159 Ha and a lowest eigenvalue of -0.3804835585677118 Ha.  Here the loop
runs to a density residual of 1e-7.  Measured on one MI355X: iteration 5 reproduces the
reference's residual to 1.4e-4 relative and its energy to 3.4e-7 Ha; the converged
energy is -1.14612087 Ha, 4.9e-7 Ha below the reference's stopping point.  The bars
(1e-4 Ha total, 2e-3 Ha eigenvalue at convergence; 1e-5 residual, 2e-6 Ha energy at
iteration 5) cover the reference's looser eigensolver.
"""

import numpy as np
import pytest

from isogriddft_amd import eigensolver, energy, kinetic
from oracle import grid as ogrid

pytestmark = pytest.mark.gpu

E_TOTAL_REF = -1.1461203815144159
EPS1_REF = -0.3804835585677118


def _bonding(g):
    f = 0.0
    for a in ogrid.H2_CASE.atoms:
        dx, dy, dz = g.x_points - a.position[0], g.y_points - a.position[1], g.z_points - a.position[2]
        f = f + np.exp(-0.8 * (dx * dx + dy * dy + dz * dz))
    return f / np.sqrt(np.sum(f * f * g.cell_volumes))


def _renorm(rho, n, g):
    return rho * (n / float(np.sum(rho * g.cell_volumes)))


def test_h2_legacy_singlet_scf_matches_reference():
    g = ogrid.build_legacy_geometry()
    occ = energy.SpinOccupations(1.0, 1.0, (1.0,), (1.0,))
    psi = _bonding(g)
    rho = _renorm(psi * psi, 1.0, g)
    guess = psi[None]
    history = []
    for it in range(1, 61):
        sol = eigensolver.solve_fixed_potential_eigenproblem(g, rho, rho, "up", 1, initial_guess_orbitals=guess,
                                                             tolerance=1e-8)
        assert sol.converged
        psi = np.asarray(sol.orbitals)[0]
        guess = psi[None]
        rho_out = _renorm(psi * psi, 1.0, g)
        residual = float(np.sqrt(2.0) * kinetic.weighted_l2_norm(rho_out - rho, g))
        e = energy.evaluate_single_point_energy(rho_out, rho_out, psi[None], psi[None], occ, g)
        history.append((it, residual, e.total, float(sol.eigenvalues[0])))
        rho = _renorm(0.4 * rho + 0.6 * rho_out, 1.0, g)
        if residual < 1e-7:
            break
    for row in history:
        print(row)
    it, residual, e_total, eps1 = history[-1]
    print("E_total", e_total, "vs", E_TOTAL_REF, "eps1", eps1, "vs", EPS1_REF, "components", e)
    assert residual < 1e-7
    assert abs(e_total - E_TOTAL_REF) < 1e-4
    assert abs(eps1 - EPS1_REF) < 2e-3
    # the energy is stationary at the fixed point: the last changes are tiny
    assert abs(history[-1][2] - history[-2][2]) < 1e-8
    # the reference stopped after its 5th iteration; the same loop retraces it: density
    # residual 1.750856831863596e-3 and E_total -1.1461203815144159 there (its eigensolver
    # ran at tolerance 5e-3, this one at 1e-8)
    _, r5, e5, _ = history[4]
    assert abs(r5 - 1.750856831863596e-3) < 1e-5
    assert abs(e5 - E_TOTAL_REF) < 2e-6


def _antibonding(g):
    a, b = ogrid.H2_CASE.atoms
    fa = np.exp(-0.8 * ((g.x_points - a.position[0]) ** 2 + (g.y_points - a.position[1]) ** 2
                        + (g.z_points - a.position[2]) ** 2))
    fb = np.exp(-0.8 * ((g.x_points - b.position[0]) ** 2 + (g.y_points - b.position[1]) ** 2
                        + (g.z_points - b.position[2]) ** 2))
    f = fa - fb
    return f / np.sqrt(np.sum(f * f * g.cell_volumes))


def test_h2_legacy_triplet_scf_matches_reference():
    """The triplet route of the same recorded run (audit/baselines.py:1966-1977): both
    electrons spin-up in the two lowest states, converged by the reference after 6
    iterations (residual 2.248e-3) at E_total = -0.7660687190478731 Ha, lowest eigenvalue
    -0.6255731151927522 Ha.  Measured: iteration 6 at -0.76606849 Ha, converged
    -0.76607043 Ha (19 iterations, 1.7e-6 Ha below the reference's stopping point)."""
    g = ogrid.build_legacy_geometry()
    occ = energy.SpinOccupations(2.0, 0.0, (1.0, 1.0), ())
    guess = np.stack([_bonding(g), _antibonding(g)])
    rho = _renorm(guess[0] ** 2 + guess[1] ** 2, 2.0, g)
    zero = np.zeros_like(rho)
    empty = np.zeros((0,) + rho.shape)
    history = []
    for it in range(1, 81):
        sol = eigensolver.solve_fixed_potential_eigenproblem(g, rho, zero, "up", 2, initial_guess_orbitals=guess,
                                                             tolerance=1e-8)
        assert sol.converged
        orbs = np.asarray(sol.orbitals)
        guess = orbs
        rho_out = _renorm(orbs[0] ** 2 + orbs[1] ** 2, 2.0, g)
        residual = float(kinetic.weighted_l2_norm(rho_out - rho, g))
        e = energy.evaluate_single_point_energy(rho_out, zero, orbs, empty, occ, g)
        history.append((it, residual, e.total, float(sol.eigenvalues[0])))
        rho = _renorm(0.4 * rho + 0.6 * rho_out, 2.0, g)
        if residual < 1e-7:
            break
    for row in history:
        print(row)
    _, residual, e_total, eps1 = history[-1]
    print("triplet E_total", e_total, "eps1", eps1, e)
    assert residual < 1e-7
    assert abs(e_total - (-0.7660687190478731)) < 1e-5
    assert abs(eps1 - (-0.6255731151927522)) < 2e-3
    # at the reference's last iteration (6) the energy agrees to 2.3e-7 Ha; the residual
    # (2.54e-3 against 2.25e-3) follows the two-state subspace, which the reference's
    # eigensolver resolved only to its tolerance 5e-3
    _, r6, e6, _ = history[5]
    assert abs(r6 - 2.2480823759300898e-3) < 0.2 * 2.2480823759300898e-3
    assert abs(e6 - (-0.7660687190478731)) < 1e-6


def test_h2_monitor_patch_triplet_scf_matches_reference():
    """The monitor-grid route of the same recorded run (audit/baselines.py:1990-2001 via
    audit/h2_monitor_grid_scf_dry_run_audit.py:28-33): the 67 x 67 x 81 development grid
    (here from the GPU generator), near-core patch (0.75, 25^3, 1.30, 8 neighbours),
    trial_fix kinetic, static-local chain, triplet, linear mixing 0.20.  The reference
    stopped after 18 iterations at residual 4.552787297010315e-3, E_total =
    -1.2214418066604806 Ha and lowest eigenvalue -0.4168423341628571 Ha.

    Those numbers were made with the nodal multipole faces of earlier reference versions.
    With the corrected-moment faces of the current reference the same loop retraces the
    residuals but sits 0.6 mHa lower in energy and eigenvalue: a near-constant offset of
    v_H.  So the loop runs with monitor_boundary_construction_mode="nodal" (faces from the
    nodal moments alone).  Measured: at iteration 18 the residual is 4.5476e-3, E_total
    -1.22144894 Ha (7e-6 from the record) and eps_1 -0.41683993 Ha (2.4e-6)."""
    from isogriddft_amd import monitor_grid, pseudo

    g = monitor_grid.build_monitor_grid_for_case(
        ogrid.H2_CASE, shape=(67, 67, 81), box_half_extents=(8.0, 8.0, 10.0),
        element_parameters=monitor_grid.build_h2_local_patch_development_element_parameters())
    params = pseudo.LocalPotentialPatchParameters(patch_radius_scale=0.75, patch_grid_shape=(25, 25, 25),
                                                  correction_strength=1.30, interpolation_neighbors=8)
    occ = energy.SpinOccupations(2.0, 0.0, (1.0, 1.0), ())
    guess = np.stack([_bonding(g), _antibonding(g)])
    rho = _renorm(guess[0] ** 2 + guess[1] ** 2, 2.0, g)
    zero = np.zeros_like(rho)
    empty = np.zeros((0,) + rho.shape)
    history = []
    for it in range(1, 19):
        sol = eigensolver.solve_fixed_potential_static_local_eigenproblem(
            g, rho, zero, "up", 2, initial_guess_orbitals=guess, tolerance=1e-8, use_monitor_patch=True,
            patch_parameters=params, kinetic_version="trial_fix", monitor_boundary_construction_mode="nodal")
        assert sol.converged
        orbs = np.asarray(sol.orbitals)
        guess = orbs
        rho_out = _renorm(orbs[0] ** 2 + orbs[1] ** 2, 2.0, g)
        residual = float(kinetic.weighted_l2_norm(rho_out - rho, g))
        e = energy.evaluate_static_local_single_point_energy(
            rho_out, zero, orbs, empty, occ, g, use_monitor_patch=True, patch_parameters=params,
            kinetic_version="trial_fix", monitor_boundary_construction_mode="nodal")
        history.append((it, residual, e.total, float(sol.eigenvalues[0])))
        rho = _renorm(0.8 * rho + 0.2 * rho_out, 2.0, g)
    for row in history:
        print(row)
    _, r18, e18, eps18 = history[-1]
    print("monitor triplet at iteration 18", r18, e18, eps18, e)
    assert abs(r18 - 4.552787297010315e-3) < 0.01 * 4.552787297010315e-3
    assert abs(e18 - (-1.2214418066604806)) < 2e-5
    assert abs(eps18 - (-0.4168423341628571)) < 1e-5

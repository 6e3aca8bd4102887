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

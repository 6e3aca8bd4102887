"""The GPU path reproduces the reference-held scalars of the default legacy H2 grid.

Reference values (src/isogrid/audit/baselines.py of the reference, bonding trial orbital
psi = normalised sum_A exp(-0.8 r_A^2) on the 51^3 legacy grid, (8, 8, 10) Bohr, stretch
(2.5, 2.5, 2.8)):  <T> = 0.997532314836 (:1066), <psi|V_loc|psi> = -2.173678262801
(:1067), E_H = 1.748670071699 Ha (:1068, rho = 2 psi^2, BiCGSTAB rtol 1e-8),
E_xc = -0.869107874980 (:822) and <psi|v_xc|psi> = -0.568636849315 (:1069) for
rho_up = rho_down = psi^2 with libxc 'lda,vwn'.
Bars: the recorded digits (1e-11 Ha for <T>, <V_loc>; 1e-8 Ha for E_H, whose recorded value
carries the reference's own 1e-8 solver tolerance).
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import kinetic, poisson, pseudo
from oracle import grid as ogrid

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def legacy51():
    return ogrid.build_legacy_geometry()


def _trial(g):
    f = 0.0
    for a in ogrid.H2_CASE.atoms:
        dx, dy, dz = g.x_points - a.position[0], g.y_points - a.position[1], g.z_points - a.position[2]
        f = f + np.exp(-0.8 * (dx * dx + dy * dy + dz * dz))
    return f / np.sqrt(np.sum(f * f * g.cell_volumes))


def test_kinetic_and_local_expectations(legacy51):
    g = legacy51
    psi = _trial(g)
    tpsi = kinetic.apply_kinetic_operator(psi, g)
    T = kinetic.integrate_field(psi * tpsi, g)
    assert abs(T - 0.997532314836) < 1e-11, T
    ev = pseudo.evaluate_local_ionic_potential(ogrid.H2_CASE, g)
    vloc = kinetic.integrate_field(psi * psi * ev.total_local_potential, g)
    assert abs(vloc - (-2.173678262801)) < 1e-11, vloc
    assert abs(pseudo.evaluate_local_ionic_energy(ev, 2 * psi * psi) - (-4.347356525602)) < 1e-11


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hartree_energy(legacy51, precision):
    g = legacy51
    rho = 2.0 * _trial(g) ** 2
    res = poisson.solve_hartree_potential(g, rho, tolerance=1e-10, max_iterations=2000, precision=precision)
    assert res.converged
    eh = poisson.evaluate_hartree_energy(rho, g, res)
    assert abs(eh - 1.748670071699) < 1e-8, (precision, eh)


def test_xc_energy_and_potential_expectation(legacy51):
    from isogriddft_amd import xc

    g = legacy51
    psi = _trial(g)
    r = psi * psi
    exc = xc.evaluate_lsda_energy(r, r, g)
    assert abs(exc - (-0.869107874980)) < 1e-11, exc
    ev = xc.evaluate_lsda_terms(r, r)
    vexp = kinetic.integrate_field(psi * psi * ev.v_xc_up, g)
    assert abs(vexp - (-0.568636849315)) < 1e-11, vexp

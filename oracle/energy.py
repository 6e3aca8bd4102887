"""Oracle: density build and single-point energy components.

Restates scf/driver.py:630-702 (weighted expectations, occupied density,
renormalisation), :884-902 (E_II) and the component sums of :905-1015 / :2668-2937.
Pinning: elementwise sums; the trial-orbital components against the reference-held
<T>, <V_loc>, E_H scalars on the default H2 grids (tests/test_oracle_pins.py).
"""

from __future__ import annotations

import numpy as np

from .pseudo import GTH_PADE


def expectation(psi, hpsi, w):
    """<psi|H psi>_W for one orbital, scf/driver.py:630-640."""
    return float(np.sum(psi * hpsi * w))


def density(orbitals, occupations):
    """sum_i f_i |psi_i|^2, scf/driver.py:663-676."""
    rho = np.zeros(orbitals.shape[1:])
    for i, f in enumerate(occupations):
        rho += f * orbitals[i] ** 2
    return rho


def renormalize(rho, target, w):
    """scf/driver.py:679-690."""
    if target == 0.0:
        return np.zeros_like(rho)
    total = float(np.sum(rho * w))
    if total <= 0.0:
        raise ValueError("positive density integral required")
    return rho * (target / total)


def ion_ion(case):
    """scf/driver.py:884-902."""
    e = 0.0
    atoms = case.atoms
    for i, a in enumerate(atoms):
        for b in atoms[i + 1:]:
            d = np.linalg.norm(np.subtract(a.position, b.position))
            e += GTH_PADE[a.element].zion * GTH_PADE[b.element].zion / d
    return e

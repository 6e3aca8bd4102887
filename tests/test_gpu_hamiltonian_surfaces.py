"""GPU parity of the Hamiltonian drop-in surfaces (ks/local_hamiltonian.py:69-112,
ks/static_hamiltonian.py:238-343 of the reference) against the FP64 oracle.

Cases: the reference-generated 19x19x23 monitor grid and a 25x27x29 legacy sinh grid, both
with an H2O-like molecule (the O atom carries the s projector).  Every field of
LocalHamiltonianTerms / StaticKSHamiltonianTerms is compared with its oracle restatement:
kinetic (production closure on monitor grids, the 7-point operator on legacy grids),
V_loc, the nonlocal action, v_xc of the chosen spin channel (lda,vwn at a spin-polarised
density, and lda,pw92), and the Hartree term.

Bars: 1e-12 (max-abs over max|ref|) for every term built from pinned operators, and for
the Hartree action when the oracle's v_H is passed in; the Hartree potential the surface
solves for itself (BiCGSTAB, tolerance 1e-8) is held to 1e-6 of the oracle's direct solve.
"""

import numpy as np
import pytest

from isogriddft_amd import hamiltonian
from oracle import grid as ogrid
from oracle import kinetic as okin
from oracle import poisson as opoi
from oracle import pseudo as opseudo
from oracle import xc as oxc

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _rel(a, b):
    b = np.asarray(b)
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


def _case():
    half = 0.5 * np.radians(104.5)
    hx, hz = 1.809 * np.sin(half), 1.809 * np.cos(half)
    return ogrid.Case("H2O-like", (ogrid.Atom("O", (0.13, -0.07, 0.21)), ogrid.Atom("H", (-hx, 0.1, hz + 0.2)),
                                   ogrid.Atom("H", (hx, -0.05, hz + 0.2))))


def _grids(small_monitor):
    yield "monitor19x19x23", small_monitor
    yield "legacy25x27x29", ogrid.build_legacy_geometry(shape=(25, 27, 29), half_extents=(6.0, 6.0, 7.0),
                                                        stretch=(2.5, 2.5, 2.8))


def _fields(g):
    X, Y, Z = g.x_points, g.y_points, g.z_points
    r2 = X ** 2 + Y ** 2 + (Z - 0.3) ** 2
    psi = np.exp(-0.6 * r2) * (1.0 + 0.3 * X - 0.2 * Z) + 0.05 * np.cos(X) * np.exp(-0.1 * r2)
    up = 0.7 * np.exp(-0.9 * r2)
    down = 0.4 * np.exp(-1.1 * (X ** 2 + (Y - 0.2) ** 2 + (Z - 0.1) ** 2))
    return psi, up, down


def test_local_hamiltonian_surfaces_match_oracle(small_monitor):
    case = _case()
    for name, g in _grids(small_monitor):
        print(f"case {name}", flush=True)
        psi, _, _ = _fields(g)
        extra = 0.1 * np.sin(g.x_points) * np.exp(-0.05 * g.z_points ** 2)
        vloc = opseudo.local_ionic_potential(case, g)
        kin = okin.kinetic_operator(psi, g, "production")
        for ex in (None, extra):
            v = vloc if ex is None else vloc + ex
            ref = kin + v * psi
            got = hamiltonian.apply_local_hamiltonian(psi, g, case=case, extra_local_potential=ex)
            assert _rel(got, ref) <= TOL, (name, _rel(got, ref))
            t = hamiltonian.evaluate_local_hamiltonian_terms(psi, g, case=case, extra_local_potential=ex)
            assert _rel(t.local_potential, v) <= TOL
            assert _rel(t.kinetic_action, kin) <= TOL
            assert _rel(t.local_potential_action, v * psi) <= TOL
            assert _rel(t.total_action, ref) <= TOL
            assert np.array_equal(t.psi, psi)
        # a caller-supplied local potential replaces the case's
        got = hamiltonian.apply_local_hamiltonian(psi, g, case=case, local_ionic_potential=extra)
        assert _rel(got, kin + extra * psi) <= TOL


@pytest.mark.parametrize("spin", ["up", "down"])
def test_static_ks_surfaces_match_oracle(small_monitor, spin):
    case = _case()
    for name, g in _grids(small_monitor):
        print(f"case {name} spin={spin}", flush=True)
        psi, up, down = _fields(g)
        tot = up + down
        kin = okin.kinetic_operator(psi, g, "production")
        vloc = opseudo.local_ionic_potential(case, g)
        vnl = opseudo.nonlocal_action(psi, case, g)
        assert np.max(np.abs(vnl)) > 1e-6 * np.max(np.abs(kin))  # the O projector is active
        vh, _, _, _ = opoi.solve_poisson(g, tot, order=2, method="direct")
        eh = opoi.hartree_energy(tot, g, vh)
        _, vu, vd = oxc.lsda_vwn(up, down)
        vxc = vu if spin == "up" else vd
        ref = kin + vloc * psi + vnl + vh * psi + vxc * psi

        # 1) the oracle's v_H passed in: every term to 1e-12
        t = hamiltonian.evaluate_static_ks_terms(psi, g, up, down, spin, case=case, hartree_potential=vh)
        assert t.spin_channel == spin
        for got, want, what in ((t.kinetic_action, kin, "kinetic"), (t.local_ionic_potential, vloc, "vloc"),
                                (t.local_ionic_action, vloc * psi, "vloc psi"), (t.nonlocal_ionic_action, vnl, "vnl"),
                                (t.hartree_action, vh * psi, "vh psi"), (t.xc_potential, vxc, "vxc"),
                                (t.xc_action, vxc * psi, "vxc psi"), (t.total_action, ref, "total"),
                                (t.rho_total, tot, "rho")):
            assert _rel(got, want) <= TOL, (name, what, _rel(got, want))
        assert abs(t.hartree_energy - eh) <= TOL * abs(eh)
        got = hamiltonian.apply_static_ks_hamiltonian(psi, g, up, down, spin, case=case, hartree_potential=vh)
        assert _rel(got, ref) <= TOL

        # 2) precomputed nonlocal action and v_xc, lda,pw92 through xc_functional
        _, pu, pd = oxc.lsda_pw92(up, down)
        vxc_pw = pu if spin == "up" else pd
        t2 = hamiltonian.evaluate_static_ks_terms(psi, g, up, down, spin, case=case, hartree_potential=vh,
                                                  xc_functional="lda,pw92")
        assert _rel(t2.xc_potential, vxc_pw) <= TOL
        got = hamiltonian.apply_static_ks_hamiltonian(psi, g, up, down, spin, case=case, nonlocal_ionic_action=vnl,
                                                      hartree_potential=vh, xc_potential=vxc)
        assert _rel(got, ref) <= TOL

        # 3) v_H solved by the surface itself (BiCGSTAB to 1e-8) against the direct solve
        t3 = hamiltonian.evaluate_static_ks_terms(psi, g, up, down, spin, case=case)
        assert _rel(t3.hartree_potential, vh) <= 1e-6, (name, _rel(t3.hartree_potential, vh))
        assert abs(t3.hartree_energy - eh) <= 1e-6 * abs(eh)
        assert _rel(t3.total_action, ref) <= 1e-6

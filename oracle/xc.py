"""Oracle: spin-polarised LSDA = Slater exchange + Perdew-Wang 1992 correlation.

The reference evaluates ``lda,vwn`` through PySCF/libxc (xc/lsda.py:22, :42-50,
:95-100), which is absent here; the north-star functional is Slater + PW92 instead.
This module therefore restates published closed forms, not the reference:
  * Slater exchange with the spin-scaling relation E_x[a,b] = (E_x[2a] + E_x[2b]) / 2.
  * PW92 correlation, J. P. Perdew and Y. Wang, Phys. Rev. B 45, 13244 (1992), Table I
    parameters with p = 1, interpolation via f(zeta) and f''(0) = 1.709921.
The reference's functional, lda,vwn (Slater + VWN5), is restated below as well.

Pinning (tests/test_oracle_pins.py, tests/test_oracle_math.py):
  * VWN at zeta = 0 against the reference's recorded libxc values: E_xc = -0.869107874980 Ha
    and <psi|v_xc|psi> = -0.568636849315 Ha (audit/baselines.py:822, :1069), to 1e-11.
  * Both functionals: the paramagnetic and ferromagnetic limits in closed form, the
    potentials v_sigma = d(rho eps)/d rho_sigma by central differences, eps_c even in zeta,
    the spin stiffness d^2 eps_c / d zeta^2 following the exact high-density law
    -(1/(6 pi^2)) ln rs, and VWN5 and PW92 agreeing within 1 mHa at intermediate zeta.
  * Parity unpinned: VWN5 at 0 < |zeta| < 1 against libxc itself.  The reference holds no
    polarised libxc value and PySCF/libxc is not available here; the checks above fix the
    interpolation's structure and its leading stiffness, not the fitted subleading
    stiffness parameters.
The density validation follows xc/lsda.py:57-67.
"""

from __future__ import annotations

import numpy as np

NEGATIVE_DENSITY_TOLERANCE = 1.0e-14
DENSITY_FLOOR = 1.0e-30

# PW92 Table I: (A, alpha1, beta1, beta2, beta3, beta4) for ec(rs,0), ec(rs,1), -alpha_c(rs)
PW92_PARAMS = (
    (0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294),
    (0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517),
    (0.016887, 0.11125, 10.357, 3.6231, 0.88026, 0.49671),
)
FZ20 = 1.709921
FZ_DEN = 2.0 ** (4.0 / 3.0) - 2.0


def validate_density(rho, name="rho"):
    """xc/lsda.py:57-67."""
    v = np.asarray(rho, dtype=np.float64)
    if not np.all(np.isfinite(v)):
        raise ValueError(f"{name} must contain only finite values.")
    if np.any(v < -NEGATIVE_DENSITY_TOLERANCE):
        raise ValueError(f"{name} must be non-negative up to roundoff.")
    return np.maximum(v, 0.0)


def _pw_g(rs, p):
    A, a1, b1, b2, b3, b4 = p
    srs = np.sqrt(rs)
    q0 = -2.0 * A * (1.0 + a1 * rs)
    q1 = 2.0 * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs)
    q1p = A * (b1 / srs + 2.0 * b2 + 3.0 * b3 * srs + 4.0 * b4 * rs)
    lg = np.log1p(1.0 / q1)
    g = q0 * lg
    dg = -2.0 * A * a1 * lg - q0 * q1p / (q1 * q1 + q1)
    return g, dg


def lsda_pw92(rho_up, rho_down):
    """Return (eps_xc, v_up, v_down) per point; eps_xc is energy per electron."""
    a = validate_density(rho_up, "rho_up")
    b = validate_density(rho_down, "rho_down")
    rho = a + b
    ok = rho > DENSITY_FLOOR
    rs_ = np.where(ok, (3.0 / (4.0 * np.pi * np.where(ok, rho, 1.0))) ** (1.0 / 3.0), 1.0)
    zeta = np.where(ok, (a - b) / np.where(ok, rho, 1.0), 0.0)
    zeta = np.clip(zeta, -1.0, 1.0)
    # exchange
    cx = (6.0 / np.pi) ** (1.0 / 3.0)
    ex_vol = -0.75 * cx * (a ** (4.0 / 3.0) + b ** (4.0 / 3.0))
    vx_up = -cx * np.cbrt(a)
    vx_dn = -cx * np.cbrt(b)
    # correlation
    g0, dg0 = _pw_g(rs_, PW92_PARAMS[0])
    g1, dg1 = _pw_g(rs_, PW92_PARAMS[1])
    g2, dg2 = _pw_g(rs_, PW92_PARAMS[2])
    ac, dac = -g2, -dg2
    opz = np.cbrt(1.0 + zeta)
    omz = np.cbrt(1.0 - zeta)
    fz = ((1.0 + zeta) * opz + (1.0 - zeta) * omz - 2.0) / FZ_DEN
    dfz = (4.0 / 3.0) * (opz - omz) / FZ_DEN
    z3 = zeta ** 3
    z4 = zeta * z3
    ec = g0 + ac * fz / FZ20 * (1.0 - z4) + (g1 - g0) * fz * z4
    decdrs = dg0 + dac * fz / FZ20 * (1.0 - z4) + (dg1 - dg0) * fz * z4
    decdz = ac / FZ20 * (dfz * (1.0 - z4) - 4.0 * z3 * fz) + (g1 - g0) * (dfz * z4 + 4.0 * z3 * fz)
    vc_common = ec - rs_ / 3.0 * decdrs
    vc_up = vc_common - (zeta - 1.0) * decdz
    vc_dn = vc_common - (zeta + 1.0) * decdz
    eps = np.where(ok, ex_vol / np.where(ok, rho, 1.0) + ec, 0.0)
    v_up = np.where(ok, vx_up + vc_up, 0.0)
    v_dn = np.where(ok, vx_dn + vc_dn, 0.0)
    return eps, v_up, v_dn


def lsda_energy_density(rho_up, rho_down):
    eps, _, _ = lsda_pw92(rho_up, rho_down)
    return (validate_density(rho_up) + validate_density(rho_down)) * eps


# ---------------------------------------------------------------- VWN (libxc 'lda,vwn')
# Slater exchange + Vosko-Wilk-Nusair correlation, parameterisation "VWN5"
# (S. H. Vosko, L. Wilk, M. Nusair, Can. J. Phys. 58, 1200 (1980), fit to the Ceperley-
# Alder data; the functional behind PySCF's 'lda,vwn', xc/lsda.py:22, :95-100 of the
# reference).  Parameter sets (A, b, c, x0) for the paramagnetic and ferromagnetic
# correlation energies and the spin stiffness; x = sqrt(rs).
VWN_P = (0.0310907, 3.72744, 12.9352, -0.10498)
VWN_F = (0.01554535, 7.06042, 18.0578, -0.32500)
VWN_A = (-1.0 / (6.0 * np.pi ** 2), 1.13107, 13.0045, -0.0047584)
VWN_FPP = 4.0 / (9.0 * (2.0 ** (1.0 / 3.0) - 1.0))


def _vwn_g(x, p):
    """VWN interpolation function and its derivative with respect to x."""
    A, b, c, x0 = p
    X = x * x + b * x + c
    X0 = x0 * x0 + b * x0 + c
    Q = np.sqrt(4.0 * c - b * b)
    at = np.arctan(Q / (2.0 * x + b))
    g = A * (np.log(x * x / X) + 2.0 * b / Q * at
             - b * x0 / X0 * (np.log((x - x0) ** 2 / X) + 2.0 * (b + 2.0 * x0) / Q * at))
    den = (2.0 * x + b) ** 2 + Q * Q
    dg = A * (2.0 / x - (2.0 * x + b) / X - 4.0 * b / den
              - b * x0 / X0 * (2.0 / (x - x0) - (2.0 * x + b) / X - 4.0 * (b + 2.0 * x0) / den))
    return g, dg


def lsda_vwn(rho_up, rho_down):
    """(eps_xc, v_up, v_down) of Slater + VWN5 with the VWN spin interpolation
    eps_c = eps_P + alpha f(z)/f''(0) (1 - z^4) + (eps_F - eps_P) f(z) z^4."""
    a = validate_density(rho_up, "rho_up")
    b = validate_density(rho_down, "rho_down")
    rho = a + b
    ok = rho > DENSITY_FLOOR
    rs = np.where(ok, (3.0 / (4.0 * np.pi * np.where(ok, rho, 1.0))) ** (1.0 / 3.0), 1.0)
    z = np.clip(np.where(ok, (a - b) / np.where(ok, rho, 1.0), 0.0), -1.0, 1.0)
    cx = (6.0 / np.pi) ** (1.0 / 3.0)
    ex_vol = -0.75 * cx * (a ** (4.0 / 3.0) + b ** (4.0 / 3.0))
    vx_up = -cx * np.cbrt(a)
    vx_dn = -cx * np.cbrt(b)
    x = np.sqrt(rs)
    gp, dp = _vwn_g(x, VWN_P)
    gf, df = _vwn_g(x, VWN_F)
    ga, da = _vwn_g(x, VWN_A)
    opz, omz = np.cbrt(1.0 + z), np.cbrt(1.0 - z)
    fz = ((1.0 + z) * opz + (1.0 - z) * omz - 2.0) / FZ_DEN
    dfz = (4.0 / 3.0) * (opz - omz) / FZ_DEN
    z3 = z ** 3
    z4 = z * z3
    ec = gp + ga * fz / VWN_FPP * (1.0 - z4) + (gf - gp) * fz * z4
    # d/drs = (d/dx) / (2 x)
    decdx = dp + da * fz / VWN_FPP * (1.0 - z4) + (df - dp) * fz * z4
    decdrs = decdx / (2.0 * x)
    decdz = ga / VWN_FPP * (dfz * (1.0 - z4) - 4.0 * z3 * fz) + (gf - gp) * (dfz * z4 + 4.0 * z3 * fz)
    vc_common = ec - rs / 3.0 * decdrs
    eps = np.where(ok, ex_vol / np.where(ok, rho, 1.0) + ec, 0.0)
    v_up = np.where(ok, vx_up + vc_common - (z - 1.0) * decdz, 0.0)
    v_dn = np.where(ok, vx_dn + vc_common - (z + 1.0) * decdz, 0.0)
    return eps, v_up, v_dn

"""Synthetic SCF-like spin-density histories for the density-mixing fixtures.

Pure NumPy, deterministic: used by make_mixing_golden.py (which feeds them to the
reference's mixers) and by the tests (which feed the same histories to the oracle and to
the GPU mixers).  The histories follow a linear fixed-point map
    F(rho) = rho* + 0.6 s(x) (rho - rho*),  residual = F(rho) - rho,
with a smooth contraction profile s in (0.3, 0.9) and perturbations that shrink by half
per step, i.e. the iterates of a damped SCF approaching rho*.
"""

from __future__ import annotations

import numpy as np


def _axes(shape):
    return np.meshgrid(*[np.linspace(-1.0, 1.0, n) for n in shape], indexing="ij")


def fixed_point(shape, w, n_alpha, n_beta):
    i, j, k = _axes(shape)
    up = np.exp(-3.0 * (i * i + j * j + (k - 0.2) ** 2)) + 1e-3
    dn = np.exp(-2.5 * ((i + 0.1) ** 2 + j * j + (k + 0.3) ** 2)) + 1e-3
    up = up * (n_alpha / np.sum(up * w)) if n_alpha else np.zeros(shape)
    dn = dn * (n_beta / np.sum(dn * w)) if n_beta else np.zeros(shape)
    s = 0.3 + 0.6 * (0.5 + 0.5 * np.cos(2.0 * i) * np.cos(1.5 * k))
    return up, dn, s


def history(shape, w, length, n_alpha=1, n_beta=1, collinear=False, seed=0):
    """[(rho_up, rho_down, residual_up, residual_down)] of `length` steps; with collinear
    the last residual difference repeats the direction of the previous one."""
    i, j, k = _axes(shape)
    star_up, star_dn, s = fixed_point(shape, w, n_alpha, n_beta)
    scale_up = float(np.max(star_up)) if n_alpha else 0.0
    scale_dn = float(np.max(star_dn)) if n_beta else 0.0
    env = np.exp(-(i * i + j * j + k * k))
    out = []
    for t in range(length):
        amp = 0.25 * 0.5 ** t
        ph = 0.7 * t + 0.3 * seed
        d_up = amp * scale_up * env * np.sin((t + 1) * 1.3 * i + ph) * np.cos((t + 2) * 0.9 * k - ph)
        d_dn = amp * scale_dn * env * np.cos((t + 1) * 1.1 * j - ph) * np.sin((t + 1) * 0.7 * k + ph)
        if collinear and t == length - 1:
            # rho_t - rho_{t-1} = 0.5 (rho_{t-1} - rho_{t-2}): the last two secants are parallel
            p_up, p_dn = out[-1][0] - star_up, out[-1][1] - star_dn
            q_up, q_dn = out[-2][0] - star_up, out[-2][1] - star_dn
            d_up = p_up + 0.5 * (p_up - q_up)
            d_dn = p_dn + 0.5 * (p_dn - q_dn)
        m_up, m_dn = star_up + d_up, star_dn + d_dn
        out.append((m_up, m_dn, (0.6 * s - 1.0) * d_up, (0.6 * s - 1.0) * d_dn))
    return out


# (name, kind, length, n_alpha, n_beta, collinear, parameters)
CASES = (
    ("diis_a", "diis", 4, 1, 1, False, {}),
    ("diis_b", "diis", 3, 1, 0, False, {}),
    ("anderson_a", "anderson", 5, 1, 1, False,
     dict(regularization=1e-10, damping=0.7, step_clip_factor=2.0, adaptive_damping_enabled=True,
          min_damping=0.3, max_damping=1.0, acceptance_residual_ratio_threshold=1.0,
          collinearity_cosine_threshold=0.999)),
    ("anderson_b", "anderson", 4, 1, 1, True,
     dict(regularization=1e-8, damping=1.0, step_clip_factor=None, adaptive_damping_enabled=False,
          min_damping=0.2, max_damping=1.0, acceptance_residual_ratio_threshold=0.95,
          collinearity_cosine_threshold=0.99)),
    ("anderson_c", "anderson", 4, 1, 1, False,
     dict(regularization=1e-10, damping=0.8, step_clip_factor=None, adaptive_damping_enabled=True,
          min_damping=0.3, max_damping=0.9, acceptance_residual_ratio_threshold=1e-6,
          collinearity_cosine_threshold=0.999)),
    ("anderson_d", "anderson", 4, 1, 0, False,
     dict(regularization=1e-10, damping=0.9, step_clip_factor=0.05, adaptive_damping_enabled=True,
          min_damping=0.3, max_damping=1.0, acceptance_residual_ratio_threshold=1.0,
          collinearity_cosine_threshold=0.999)),
    ("broyden_a", "broyden", 4, 1, 1, False, dict(mixing=0.3, regularization=1e-12, damping=0.8)),
    ("broyden_b", "broyden", 3, 1, 0, False, dict(mixing=0.5, regularization=1e-12, damping=1.0)),
)

"""CPU restatement of the reference's SCF density mixers (test infrastructure only: the
product path is isogriddft_amd.mixing on the GPU; this module is the checker).

A history entry is a tuple (rho_up, rho_down, residual_up, residual_down) of FP64
fields; w are the cell volumes.  The spin-density inner product is
<a, b> = sum w (a_up b_up + a_down b_down)            (scf/driver.py:1533-1559).

  diis      scf/driver.py:2234-2294   Pulay extrapolation on the residual history
  anderson  scf/driver.py:2297-2545   secant (type-II) Anderson with collinearity filter
            (:2125-2193), adaptive damping (:2196-2231), acceptance test and step clip
  broyden   scf/driver.py:2548-2650   good-Broyden-like inverse-Jacobian update

Pinned against the reference's own outputs on SCF-like histories
(tests/golden/mixing_vectors.npz, written by tests/golden/make_mixing_golden.py) and by
the closed-form checks in tests/test_oracle_mixing.py.
"""

from __future__ import annotations

import numpy as np

from .energy import renormalize


def dot(a_up, a_dn, b_up, b_dn, w):
    return float(np.sum((a_up * b_up + a_dn * b_dn) * w))


def norm(a_up, a_dn, w):
    return float(np.sqrt(max(dot(a_up, a_dn, a_up, a_dn, w), 0.0)))


def _finish(up, dn, n_alpha, n_beta, w):
    """Clamp, check the integrals, renormalise to the electron counts (None on failure)."""
    if n_alpha > 0:
        up = np.maximum(up, 0.0)
        if float(np.sum(up * w)) <= 0.0:
            return None
    else:
        up = np.zeros_like(up)
    if n_beta > 0:
        dn = np.maximum(dn, 0.0)
        if float(np.sum(dn * w)) <= 0.0:
            return None
    else:
        dn = np.zeros_like(dn)
    return renormalize(up, n_alpha, w), renormalize(dn, n_beta, w)


def diis(history, w, n_alpha, n_beta):
    h = len(history)
    if h < 2:
        return None
    B = np.full((h + 1, h + 1), -1.0)
    B[h, h] = 0.0
    for i, a in enumerate(history):
        for j, b in enumerate(history):
            B[i, j] = dot(a[2], a[3], b[2], b[3], w)
    B[:h, :h] += 1e-12 * max(float(np.max(np.abs(B[:h, :h]))), 1.0) * np.eye(h)
    rhs = np.zeros(h + 1)
    rhs[h] = -1.0
    try:
        c = np.linalg.solve(B, rhs)[:h]
    except np.linalg.LinAlgError:
        return None
    if not np.all(np.isfinite(c)) or float(np.max(np.abs(c))) > 20.0:
        return None
    up = sum(ci * e[0] for ci, e in zip(c, history))
    dn = sum(ci * e[1] for ci, e in zip(c, history))
    return _finish(up, dn, n_alpha, n_beta, w)


def anderson_secants(history, w, cosine_threshold):
    """Consecutive differences (d rho_up, d rho_dn, d r_up, d r_dn), dropping tiny residual
    differences and those nearly parallel to one already kept; returns (kept, len(kept)+1)."""
    if len(history) < 2:
        return [], len(history)
    kept = []
    for prev, nxt in zip(history[:-1], history[1:]):
        dr_up, dr_dn = nxt[2] - prev[2], nxt[3] - prev[3]
        n = norm(dr_up, dr_dn, w)
        if not np.isfinite(n) or n <= 1e-14:
            continue
        parallel = False
        for k in kept:
            kn = norm(k[2], k[3], w)
            if not np.isfinite(kn) or kn <= 1e-14:
                continue
            if abs(dot(dr_up, dr_dn, k[2], k[3], w) / (n * kn)) >= cosine_threshold:
                parallel = True
                break
        if not parallel:
            kept.append((nxt[0] - prev[0], nxt[1] - prev[1], dr_up, dr_dn))
    return kept, len(kept) + 1


def anderson_damping(history, w, base, adaptive, lo, hi):
    b = float(np.clip(base, lo, hi))
    if not adaptive or len(history) < 2:
        return b
    cur = norm(history[-1][2], history[-1][3], w)
    prev = norm(history[-2][2], history[-2][3], w)
    if not np.isfinite(cur) or not np.isfinite(prev) or prev <= 1e-16:
        return b
    ratio = cur / prev
    if ratio > 1.05:
        return float(lo)
    if ratio > 0.99:
        return float(np.clip(0.5 * (b + lo), lo, hi))
    if ratio < 0.90:
        return float(hi)
    return b


def anderson(history, w, n_alpha, n_beta, *, regularization, damping, step_clip_factor,
             adaptive_damping_enabled, min_damping, max_damping, acceptance_residual_ratio_threshold,
             collinearity_cosine_threshold):
    """Returns (result or None, accepted, effective damping, filtered length, ratio or None)."""
    h = len(history)
    if h < 2 or not (0.0 < damping <= 1.0):
        return None, False, float(damping), h, None
    cur = history[-1]
    sec, flen = anderson_secants(history, w, collinearity_cosine_threshold)
    d = anderson_damping(history, w, damping, adaptive_damping_enabled, min_damping, max_damping)
    if not sec:
        return None, False, d, flen, None
    n = len(sec)
    G = np.array([[dot(a[2], a[3], b[2], b[3], w) for b in sec] for a in sec])
    rhs = np.array([dot(a[2], a[3], cur[2], cur[3], w) for a in sec])
    G = G + max(float(regularization), 0.0) * max(float(np.max(np.abs(G))), 1.0) * np.eye(n)
    try:
        c = np.linalg.solve(G, rhs)
    except np.linalg.LinAlgError:
        return None, False, d, flen, None
    if not np.all(np.isfinite(c)) or float(np.max(np.abs(c))) > 50.0:
        return None, False, d, flen, None
    up = cur[0] - sum(ci * s[0] for ci, s in zip(c, sec))
    dn = cur[1] - sum(ci * s[1] for ci, s in zip(c, sec))
    pr_up = cur[2] - sum(ci * s[2] for ci, s in zip(c, sec))
    pr_dn = cur[3] - sum(ci * s[3] for ci, s in zip(c, sec))
    up = (1.0 - d) * cur[0] + d * up
    dn = (1.0 - d) * cur[1] + d * dn
    pr_up = (1.0 - d) * cur[2] + d * pr_up
    pr_dn = (1.0 - d) * cur[3] + d * pr_dn
    rn = norm(cur[2], cur[3], w)
    ratio = 0.0 if rn <= 1e-16 else float(norm(pr_up, pr_dn, w) / rn)
    if not np.isfinite(ratio) or ratio > acceptance_residual_ratio_threshold:
        return None, False, d, flen, ratio
    if step_clip_factor is not None:
        cf = float(step_clip_factor)
        if not np.isfinite(cf) or cf <= 0.0:
            return None, False, d, flen, ratio
        st_up, st_dn = up - cur[0], dn - cur[1]
        sn = norm(st_up, st_dn, w)
        cap = cf * max(rn, 1e-16)
        if sn > cap:
            f = cap / max(sn, 1e-16)
            up, dn = cur[0] + f * st_up, cur[1] + f * st_dn
    res = _finish(up, dn, n_alpha, n_beta, w)
    if res is None:
        return None, False, d, flen, ratio
    return res, True, d, flen, ratio


def broyden(history, w, n_alpha, n_beta, *, mixing, regularization, damping):
    if len(history) < 2 or not (0.0 < mixing <= 1.0) or not (0.0 < damping <= 1.0):
        return None
    upd, sres, den = [], [], []

    def inv_map(v_up, v_dn):
        m_up, m_dn = mixing * v_up, mixing * v_dn
        for (u_up, u_dn), (s_up, s_dn), q in zip(upd, sres, den):
            c = dot(s_up, s_dn, v_up, v_dn, w) / q
            m_up = m_up + c * u_up
            m_dn = m_dn + c * u_dn
        return m_up, m_dn

    for prev, nxt in zip(history[:-1], history[1:]):
        dr_up, dr_dn = nxt[0] - prev[0], nxt[1] - prev[1]
        ds_up, ds_dn = nxt[2] - prev[2], nxt[3] - prev[3]
        q = dot(ds_up, ds_dn, ds_up, ds_dn, w)
        if q <= max(float(regularization), 1e-16):
            continue
        m_up, m_dn = inv_map(ds_up, ds_dn)
        upd.append((dr_up - m_up, dr_dn - m_dn))
        sres.append((ds_up, ds_dn))
        den.append(q)
    if not den:
        return None
    c_up, c_dn = inv_map(history[-1][2], history[-1][3])
    return _finish(history[-1][0] + damping * c_up, history[-1][1] + damping * c_dn, n_alpha, n_beta, w)

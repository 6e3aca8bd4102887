"""SCF density mixers on the GPU (the reference's scf/driver.py:2125-2650: DIIS, Anderson,
Broyden-like), the step after the density build that keeps an SCF iteration in HBM.

A spin density (rho_up, rho_down) is held as one stacked vector of 2N values, so the
reference's spin inner product <a, b> = sum w (a_up b_up + a_down b_down)
(scf/driver.py:1533-1559) is a weighted product over 2N with weights [w; w].  Every
inner product a mixer needs comes from one pass of the split-K Gram kernel over the
stacked history (secant differences are formed on the device first, not from
differences of Gram entries, which would cancel late in an SCF); the secant filter, the
Broyden recursion and the small linear solves run on those Gram entries on the host; the
mixed density is one device combination of the history rows, clamped, checked and
renormalised on the device.

Inputs: history entries with NumPy or device-tensor fields (field names of the
reference's dataclasses); outputs: NumPy when the history holds NumPy arrays, device
tensors otherwise.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .device import device_grid
from .reductions import gram, wdot


@dataclass(frozen=True)
class MonitorGridDiisHistoryEntry:
    """scf/driver.py:239-246."""

    mixed_rho_up: object
    mixed_rho_down: object
    residual_up: object
    residual_down: object


@dataclass(frozen=True)
class MonitorGridAndersonHistoryEntry:
    """scf/driver.py:249-256."""

    mixed_rho_up: object
    mixed_rho_down: object
    residual_up: object
    residual_down: object


@dataclass(frozen=True)
class MonitorGridBroydenHistoryEntry:
    """scf/driver.py:271-278."""

    rho_up: object
    rho_down: object
    residual_up: object
    residual_down: object


@dataclass(frozen=True)
class MonitorGridAndersonApplyResult:
    """scf/driver.py:259-268."""

    mixed_rho_up: object
    mixed_rho_down: object
    accepted: bool
    effective_damping: float
    filtered_history_length: int
    projected_residual_ratio: float | None


class _Stack:
    """The history on the device: rho (h, 2N) and residuals (h, 2N), FP64."""

    def __init__(self, history, grid_geometry, rho_names):
        self.dg = device_grid(grid_geometry)
        dev = self.dg.device
        self.numpy = not isinstance(getattr(history[0], rho_names[0]), torch.Tensor)
        N = self.dg.N

        def flat(a):
            t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64))
            t = t.to(device=dev, dtype=torch.float64).reshape(-1)
            if t.numel() != N:
                raise ValueError(f"history fields must have {N} points; received {t.numel()}.")
            return t

        h = len(history)
        self.rho = torch.empty((h, 2 * N), dtype=torch.float64, device=dev)
        self.res = torch.empty_like(self.rho)
        for i, e in enumerate(history):
            self.rho[i, :N] = flat(getattr(e, rho_names[0]))
            self.rho[i, N:] = flat(getattr(e, rho_names[1]))
            self.res[i, :N] = flat(e.residual_up)
            self.res[i, N:] = flat(e.residual_down)
        w = self.dg.weights()
        self.w2 = torch.cat([w, w])
        self.N = N

    def gram(self, A, B):
        return gram(A, B, self.w2).cpu().numpy()

    def combine(self, coef, rows):
        """sum_i coef_i rows_i (one device GEMV)."""
        c = torch.as_tensor(np.asarray(coef, dtype=np.float64), device=rows.device)
        return torch.matmul(c, rows)

    def _wdot2(self, a, b):
        return wdot(self.dg, a[:self.N], b[:self.N]) + wdot(self.dg, a[self.N:], b[self.N:])

    def finish(self, v, n_alpha, n_beta):
        """Clamp each spin channel at zero, require a positive integral, renormalise to the
        electron counts (scf/driver.py:2271-2294); None when a channel has no charge."""
        out = []
        for part, n in ((v[:self.N], n_alpha), (v[self.N:], n_beta)):
            if n > 0:
                part = part.clamp_min(0.0)
                total = wdot(self.dg, part, None)
                if total <= 0.0:
                    return None
                part = part * (n / total)
            else:
                part = torch.zeros_like(part)
            out.append(part)
        return self.emit(out[0], out[1])

    def emit(self, up, dn):
        shape = self.dg.shape
        if self.numpy:
            return up.reshape(shape).cpu().numpy(), dn.reshape(shape).cpu().numpy()
        return up.reshape(shape), dn.reshape(shape)


def apply_monitor_grid_density_diis(history, *, grid_geometry, n_alpha, n_beta):
    """scf/driver.py:2234-2294: Pulay extrapolation on the residual history."""
    h = len(history)
    if h < 2:
        return None
    st = _Stack(history, grid_geometry, ("mixed_rho_up", "mixed_rho_down"))
    B = np.full((h + 1, h + 1), -1.0)
    B[h, h] = 0.0
    B[:h, :h] = st.gram(st.res, st.res)
    B[:h, :h] += 1e-12 * max(float(np.max(np.abs(B[:h, :h]))), 1.0) * np.eye(h)
    rhs = np.zeros(h + 1)
    rhs[h] = -1.0
    try:
        c = np.linalg.solve(B, rhs)[:h]
    except np.linalg.LinAlgError:
        return None
    if not np.all(np.isfinite(c)) or float(np.max(np.abs(c))) > 20.0:
        return None
    return st.finish(st.combine(c, st.rho), n_alpha, n_beta)


def filter_monitor_grid_anderson_secants(G, collinearity_cosine_threshold):
    """scf/driver.py:2125-2193 on the Gram matrix G of the residual differences: indices
    of the secants kept (tiny or nearly parallel to an earlier kept one are dropped)."""
    kept = []
    n = np.sqrt(np.clip(np.diag(G), 0.0, None))
    for i in range(G.shape[0]):
        if not np.isfinite(n[i]) or n[i] <= 1e-14:
            continue
        parallel = False
        for j in kept:
            if not np.isfinite(n[j]) or n[j] <= 1e-14:
                continue
            if abs(G[i, j] / (n[i] * n[j])) >= collinearity_cosine_threshold:
                parallel = True
                break
        if not parallel:
            kept.append(i)
    return kept


def select_monitor_grid_anderson_damping(current_norm, previous_norm, *, base_damping, adaptive_damping_enabled,
                                         min_damping, max_damping):
    """scf/driver.py:2196-2231 from the last two residual norms (previous_norm None: one
    entry only)."""
    b = float(np.clip(base_damping, min_damping, max_damping))
    if not adaptive_damping_enabled or previous_norm is None:
        return b
    if not np.isfinite(current_norm) or not np.isfinite(previous_norm) or previous_norm <= 1e-16:
        return b
    ratio = current_norm / previous_norm
    if ratio > 1.05:
        return float(min_damping)
    if ratio > 0.99:
        return float(np.clip(0.5 * (b + min_damping), min_damping, max_damping))
    if ratio < 0.90:
        return float(max_damping)
    return b


def apply_monitor_grid_density_anderson(history, *, grid_geometry, n_alpha, n_beta, regularization, damping,
                                        step_clip_factor, adaptive_damping_enabled, min_damping, max_damping,
                                        acceptance_residual_ratio_threshold, collinearity_cosine_threshold):
    """scf/driver.py:2297-2545: secant Anderson with the collinearity filter, adaptive
    damping, the projected-residual acceptance test and the step clip."""
    h = len(history)

    def reject(d, flen, ratio=None):
        return MonitorGridAndersonApplyResult(None, None, False, float(d), int(flen), ratio)

    if h < 2 or not (0.0 < damping <= 1.0):
        return reject(damping, h)
    st = _Stack(history, grid_geometry, ("mixed_rho_up", "mixed_rho_down"))
    dres = st.res[1:] - st.res[:-1]
    drho = st.rho[1:] - st.rho[:-1]
    # one pass: [dR; R_cur; R_prev] against [dR; R_cur; R_prev]
    rows = torch.cat([dres, st.res[-1:], st.res[-2:-1]], 0)
    G = st.gram(rows, rows)
    ns = h - 1
    kept = filter_monitor_grid_anderson_secants(G[:ns, :ns], collinearity_cosine_threshold)
    flen = len(kept) + 1
    cur_norm = float(np.sqrt(max(G[ns, ns], 0.0)))
    prev_norm = float(np.sqrt(max(G[ns + 1, ns + 1], 0.0)))
    d = select_monitor_grid_anderson_damping(cur_norm, prev_norm, base_damping=damping,
                                             adaptive_damping_enabled=adaptive_damping_enabled,
                                             min_damping=min_damping, max_damping=max_damping)
    if not kept:
        return reject(d, flen)
    K = np.asarray(kept)
    Gk = G[np.ix_(K, K)]
    rhs = G[K, ns]
    Gk = Gk + max(float(regularization), 0.0) * max(float(np.max(np.abs(Gk))), 1.0) * np.eye(len(K))
    try:
        c = np.linalg.solve(Gk, rhs)
    except np.linalg.LinAlgError:
        return reject(d, flen)
    if not np.all(np.isfinite(c)) or float(np.max(np.abs(c))) > 50.0:
        return reject(d, flen)
    kt = torch.as_tensor(K, device=st.rho.device)
    step = -d * st.combine(c, drho.index_select(0, kt))          # rho - rho_cur
    pres = st.res[-1] - d * st.combine(c, dres.index_select(0, kt))
    rn = float(np.sqrt(max(G[ns, ns], 0.0)))
    ratio = 0.0 if rn <= 1e-16 else float(np.sqrt(max(st._wdot2(pres, pres), 0.0)) / rn)
    if not np.isfinite(ratio) or ratio > acceptance_residual_ratio_threshold:
        return reject(d, flen, ratio)
    if step_clip_factor is not None:
        cf = float(step_clip_factor)
        if not np.isfinite(cf) or cf <= 0.0:
            return reject(d, flen, ratio)
        sn = float(np.sqrt(max(st._wdot2(step, step), 0.0)))
        cap = cf * max(rn, 1e-16)
        if sn > cap:
            step = step * (cap / max(sn, 1e-16))
    out = st.finish(st.rho[-1] + step, n_alpha, n_beta)
    if out is None:
        return reject(d, flen, ratio)
    return MonitorGridAndersonApplyResult(out[0], out[1], True, float(d), int(flen), ratio)


def apply_monitor_grid_density_broyden_like(history, *, grid_geometry, n_alpha, n_beta, mixing, regularization,
                                            damping):
    """scf/driver.py:2548-2650: inverse map mixing * I plus rank-one secant updates.

    The update vectors u_i = d rho_i - G_{i-1} d r_i are kept as coefficient rows over the
    stacked basis [d rho; d r; r_cur] (host), so the correction G r_cur is one device
    combination; every inner product comes from one Gram of [d r; r_cur]."""
    h = len(history)
    if h < 2 or not (0.0 < mixing <= 1.0) or not (0.0 < damping <= 1.0):
        return None
    st = _Stack(history, grid_geometry, ("rho_up", "rho_down"))
    drho = st.rho[1:] - st.rho[:-1]
    dres = st.res[1:] - st.res[:-1]
    ns = h - 1
    rows = torch.cat([dres, st.res[-1:]], 0)
    G = st.gram(rows, rows)       # <d r_i, d r_j>, <d r_i, r_cur>
    nb = 2 * ns + 1               # basis: d rho_0.., d r_0.., r_cur

    def inv_map(v_coef, v_dots):
        """Coefficients of G v over the basis for a vector v given by its coefficients and
        its inner products with the kept secant residuals."""
        out = mixing * v_coef
        for (u, q, i) in updates:
            out = out + (v_dots[i] / q) * u
        return out

    updates = []
    for i in range(ns):
        q = G[i, i]
        if q <= max(float(regularization), 1e-16):
            continue
        e = np.zeros(nb)
        e[ns + i] = 1.0
        mapped = inv_map(e, G[i, :ns])
        u = -mapped
        u[i] += 1.0
        updates.append((u, float(q), i))
    if not updates:
        return None
    e = np.zeros(nb)
    e[2 * ns] = 1.0
    corr = inv_map(e, G[ns, :ns])
    basis = torch.cat([drho, dres, st.res[-1:]], 0)
    return st.finish(st.rho[-1] + damping * st.combine(corr, basis), n_alpha, n_beta)

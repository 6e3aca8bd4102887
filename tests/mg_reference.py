"""Reference V-cycle of the Poisson multigrid preconditioner, written with SciPy sparse
matrices (aggregation P, 7-point level operators, dense coarsest blocks) so that the
kernels' index arithmetic is checked against an independent formulation.  Test helper,
imported by the host and GPU multigrid tests."""

import numpy as np
import scipy.sparse as sp


def interior_mask(shape):
    m = np.zeros(shape, dtype=bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m.reshape(-1)


def fine_aggregation(shape):
    """P1 (n x 8 c^3): interior node i -> aggregate (parity p, (i - p) >> 2)."""
    c = [(n + 3) // 4 for n in shape]
    I = np.indices(shape).reshape(3, -1)
    p = I & 1
    S = (I - p) >> 2
    col = (((4 * p[0] + 2 * p[1] + p[2]) * c[0] + S[0]) * c[1] + S[1]) * c[2] + S[2]
    rows = np.flatnonzero(interior_mask(shape))
    return sp.csr_matrix((np.ones(rows.size), (rows, col[rows])), shape=(I.shape[1], 8 * c[0] * c[1] * c[2]))


def level_aggregation(c):
    """P (8 c^3 x 8 d^3) between coarse levels: (p, s) -> (p, s >> 1)."""
    d = [(n + 1) // 2 for n in c]
    m = 8 * int(np.prod(c))
    J = np.indices((8, *c)).reshape(4, -1)
    col = ((J[0] * d[0] + (J[1] >> 1)) * d[1] + (J[2] >> 1)) * d[2] + (J[3] >> 1)
    return sp.csr_matrix((np.ones(m), (np.arange(m), col)), shape=(m, 8 * int(np.prod(d))))


def seven_point_matrix(coef):
    """Sparse matrix of the level operators coef (7, 8, cx, cy, cz)."""
    c = coef.shape[2:]
    m = 8 * int(np.prod(c))
    idx = np.arange(m).reshape(8, *c)
    rows, cols, vals = [idx.reshape(-1)], [idx.reshape(-1)], [coef[0].reshape(-1)]
    k = 1
    for axis in range(3):
        for step in (-1, 1):
            src = [slice(None)] * 4
            dst = [slice(None)] * 4
            if step < 0:
                src[axis + 1], dst[axis + 1] = slice(1, None), slice(None, -1)
            else:
                src[axis + 1], dst[axis + 1] = slice(None, -1), slice(1, None)
            rows.append(idx[tuple(src)].reshape(-1))
            cols.append(idx[tuple(dst)].reshape(-1))
            vals.append(coef[k][tuple(src)].reshape(-1))
            k += 1
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(m, m))


class ReferenceVCycle:
    """x = M b for the full-grid operator A (zero face rows / columns), the level
    operators ``levels`` (list of (7, 8, c...) arrays), the coarsest inverses ``inv``
    (8, m, m) and the fine / coarse damping factors."""

    def __init__(self, A, shape, levels, inv, omega, coarse_omega):
        self.A = A
        d = A.diagonal()
        self.dinv = np.where(d != 0.0, 1.0 / np.where(d != 0.0, d, 1.0), 0.0)
        self.P1 = fine_aggregation(shape)
        self.As = [seven_point_matrix(c) for c in levels]
        self.Ds = [c[0].reshape(-1) for c in levels]
        self.Ps = [level_aggregation(c.shape[2:]) for c in levels[:-1]]
        self.inv = inv
        self.omega = omega
        self.cw = coarse_omega

    def _coarse(self, lev, rhs):
        if lev == len(self.As) - 1:
            r = rhs.reshape(8, -1)
            return np.einsum("pij,pj->pi", self.inv, r).reshape(-1)
        A, D, P = self.As[lev], self.Ds[lev], self.Ps[lev]
        xa = self.cw * rhs / D
        e = self._coarse(lev + 1, P.T @ (rhs - A @ xa))
        xa = xa + P @ e
        return xa + self.cw * (rhs - A @ xa) / D

    def __call__(self, b):
        w = self.omega * self.dinv
        x = w * b
        e = self._coarse(0, self.P1.T @ (b - self.A @ x))
        x = x + self.P1 @ e
        return x + w * (b - self.A @ x)


def bicgstab(A, b, M=None, rtol=1e-10, maxiter=500):
    """Right-preconditioned BiCGSTAB (the recurrences of the device solver); returns
    (x, iterations, relative residual)."""
    x = np.zeros_like(b)
    r = b.copy()
    rhat = b.copy()
    p = np.zeros_like(b)
    v = np.zeros_like(b)
    rho = alpha = omega = 1.0
    bn = np.linalg.norm(b)
    rel = 1.0
    for it in range(1, maxiter + 1):
        rho_new = rhat @ r
        beta = (rho_new / rho) * (alpha / omega)
        rho = rho_new
        p = r + beta * (p - omega * v)
        ph = M(p) if M else p
        v = A @ ph
        alpha = rho / (rhat @ v)
        s = r - alpha * v
        sh = M(s) if M else s
        t = A @ sh
        omega = (t @ s) / (t @ t)
        x = x + alpha * ph + omega * sh
        r = s - omega * t
        rel = np.linalg.norm(r) / bn
        if rel <= rtol:
            return x, it, rel
    return x, maxiter, rel


SHIFTS = (0, 6, 1, 5, 2, 4, 3)  # colour shift of the 7-point offsets (o_x + 2 o_y + 3 o_z) / 2 mod 7


def probe_entries(A, shape):
    """The 28-probe reading of the same-parity entries, restated with NumPy masks: probe
    q = 7 pair + colour holds the interior nodes of parity pair (class p or 7 - p) and
    sub-lattice colour ((ix >> 1) + 2 (iy >> 1) + 3 (iz >> 1)) mod 7; row i reads its
    offset-k entry from the probe of its own pair and colour + SHIFTS[k]."""
    I = np.indices(shape).reshape(3, -1)
    cls = 4 * (I[0] & 1) + 2 * (I[1] & 1) + (I[2] & 1)
    pair = np.where(cls >= 4, 7 - cls, cls)
    col = ((I[0] >> 1) + 2 * (I[1] >> 1) + 3 * (I[2] >> 1)) % 7
    inside = interior_mask(shape)
    ent = np.zeros((7, I.shape[1]))
    for q in range(28):
        e = (inside & (7 * pair + col == q)).astype(np.float64)
        out = A @ e
        for k in range(7):
            sel = inside & (pair == q // 7) & ((col + SHIFTS[k]) % 7 == q % 7)
            ent[k, sel] = out[sel]
    return ent.reshape((7,) + tuple(shape))

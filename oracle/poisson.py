"""Oracle: open-boundary Hartree Poisson solve on legacy and monitor grids.

Restates poisson/open_boundary.py and poisson/hartree.py:
  * nodal multipole moments (orders 0-2)          open_boundary.py:223-264, :816-861
  * corrected moments (top-10 % Jacobian cells,
    4x4x4 quadratic least squares, 2x2x2 samples)  open_boundary.py:182-203, :267-464
  * boundary values from moments                   open_boundary.py:467-495, :884-986
  * monitor split v = u + b, rhs = 4 pi rho + L(b) open_boundary.py:1108-1197
  * legacy interior operator + rhs                 open_boundary.py:990-1014, :1200-1226
  * E_H = 1/2 int rho v_H                          hartree.py:117-129
Extension (no reference counterpart, D3): raw Cartesian moments and the boundary
expansion to orders 3 and 4 from the Taylor series of 1/|r - r'|.

Pinning: discrete solutions by a sparse direct solve of the assembled interior
matrix (scipy.sparse.linalg.spsolve); multipole formulas by point charges and
analytic Gaussian moments; the corrected-moment construction by a reference
fixture and the reference invariants (tests/test_hartree_poisson.py:234-530 of the
reference); E_H by the reference-held 1.748670071699 / 1.765755769546 Ha.
"""

from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import kinetic
from .grid import _boundary_mask

FOUR_PI = 4.0 * np.pi
_QUANTILE = 0.9


def raw_moments(weighted, dx, dy, dz, order):
    """Raw Cartesian moments M[l][(a1..al)] = sum q x_a1 ... x_al for l <= order."""
    comps = (dx, dy, dz)
    out = []
    for l in range(order + 1):
        t = np.zeros((3,) * l) if l else np.zeros(())
        for idx in itertools.product(range(3), repeat=l):
            prod = weighted
            for a in idx:
                prod = prod * comps[a]
            t[idx] = np.sum(prod)
        out.append(t)
    return out


def traceless_quadrupole(M2):
    """Q_ab = 3 M_ab - tr(M) delta_ab (open_boundary.py:244-263)."""
    return 3.0 * M2 - np.trace(M2) * np.eye(3)


def boundary_potential(moments, dx, dy, dz, order):
    """Free-space expansion sum_l (-1)^l / l! M_{a..} d_{a..}(1/r).

    For l <= 2 this equals q/r + p.r/r^3 + 1/2 Q_ab r_a r_b / r^5 of
    open_boundary.py:467-495 with Q traceless.
    """
    r2 = dx * dx + dy * dy + dz * dz
    r = np.sqrt(r2)
    rv = (dx, dy, dz)
    phi = moments[0] / r
    if order >= 1:
        M1 = moments[1]
        phi = phi + (M1[0] * dx + M1[1] * dy + M1[2] * dz) / r ** 3
    if order >= 2:
        M2 = moments[2]
        quad = sum(M2[a, b] * rv[a] * rv[b] for a in range(3) for b in range(3))
        phi = phi + 0.5 * (3.0 * quad - np.trace(M2) * r2) / r ** 5
    if order >= 3:
        M3 = moments[3]
        c3 = sum(M3[a, b, c] * rv[a] * rv[b] * rv[c] for a in range(3) for b in range(3) for c in range(3))
        tr3 = sum(sum(M3[a, a, c] for a in range(3)) * rv[c] for c in range(3))
        phi = phi + (15.0 * c3 / r ** 7 - 9.0 * tr3 / r ** 5) / 6.0
    if order >= 4:
        M4 = moments[4]
        c4 = 0.0
        for a, b, c, d in itertools.product(range(3), repeat=4):
            c4 = c4 + M4[a, b, c, d] * rv[a] * rv[b] * rv[c] * rv[d]
        t2 = 0.0
        for c, d in itertools.product(range(3), repeat=2):
            t2 = t2 + sum(M4[a, a, c, d] for a in range(3)) * rv[c] * rv[d]
        t0 = sum(M4[a, a, c, c] for a in range(3) for c in range(3))
        phi = phi + (105.0 * c4 / r ** 9 - 90.0 * t2 / r ** 7 + 9.0 * t0 / r ** 5) / 24.0
    return phi


# ---------------------------------------------------------------- corrected moments


def _cell_average(f):
    return (
        f[:-1, :-1, :-1] + f[1:, :-1, :-1] + f[:-1, 1:, :-1] + f[1:, 1:, :-1]
        + f[:-1, :-1, 1:] + f[1:, :-1, 1:] + f[:-1, 1:, 1:] + f[1:, 1:, 1:]
    ) / 8.0


def selected_cells(geom):
    """open_boundary.py:182-203."""
    cj = _cell_average(geom.jacobian)
    bm = np.zeros(cj.shape, dtype=bool)
    bm[0], bm[-1] = True, True
    bm[:, 0], bm[:, -1] = True, True
    bm[:, :, 0], bm[:, :, -1] = True, True
    inner = ~bm
    thr = float(np.quantile(cj[inner], _QUANTILE))
    mask = inner & (cj >= thr)
    return mask, np.argwhere(mask)


def _incident_counts(cell_mask):
    m = cell_mask.astype(np.float64)
    c = np.zeros(tuple(s + 1 for s in m.shape))
    for ox, oy, oz in itertools.product((0, 1), repeat=3):
        c[ox:ox + m.shape[0], oy:oy + m.shape[1], oz:oz + m.shape[2]] += m
    return c


def node_selected_fraction(cell_mask):
    """open_boundary.py:211-220, :267-296."""
    sel = _incident_counts(cell_mask)
    tot = _incident_counts(np.ones_like(cell_mask, dtype=bool))
    return np.divide(sel, tot, out=np.zeros_like(sel), where=tot > 0.0)


def _trilinear(c, u, v, w):
    um, vm, wm = 1.0 - u, 1.0 - v, 1.0 - w
    return (
        um * vm * wm * c[0, 0, 0] + u * vm * wm * c[1, 0, 0] + um * v * wm * c[0, 1, 0]
        + u * v * wm * c[1, 1, 0] + um * vm * w * c[0, 0, 1] + u * vm * w * c[1, 0, 1]
        + um * v * w * c[0, 1, 1] + u * v * w * c[1, 1, 1]
    )


def _design(x, y, z):
    return np.column_stack([np.ones_like(x), x, y, z, x * x, y * y, z * z, x * y, x * z, y * z])


def _subcell_samples(geom, rho, cells):
    """Per selected cell: least-squares quadratic over the clipped 4x4x4 stencil, sampled at
    the 2x2x2 sub-cell midpoints (open_boundary.py:397-457, :584-626).  Returns the sample
    positions, the trilinear Jacobian, the quadratic and the trilinear density values and
    the sub-cell logical volume."""
    lx, ly, lz = geom.logical_x, geom.logical_y, geom.logical_z
    subvol = (lx[1] - lx[0]) * (ly[1] - ly[0]) * (lz[1] - lz[0]) / 8.0
    X, Y, Z, J = geom.x_points, geom.y_points, geom.z_points, geom.jacobian
    nx, ny, nz = geom.jacobian.shape
    out = []
    for i, j, k in cells:
        sl = (slice(max(i - 1, 0), min(i + 3, nx)), slice(max(j - 1, 0), min(j + 3, ny)),
              slice(max(k - 1, 0), min(k + 3, nz)))
        D = _design(X[sl].ravel(), Y[sl].ravel(), Z[sl].ravel())
        coef, *_ = np.linalg.lstsq(D, rho[sl].ravel(), rcond=None)
        c = (slice(i, i + 2), slice(j, j + 2), slice(k, k + 2))
        for u, v, ww in itertools.product((0.25, 0.75), repeat=3):
            xv, yv, zv = _trilinear(X[c], u, v, ww), _trilinear(Y[c], u, v, ww), _trilinear(Z[c], u, v, ww)
            jv = _trilinear(J[c], u, v, ww)
            rq = float(_design(np.array([xv]), np.array([yv]), np.array([zv]))[0] @ coef)
            out.append((xv, yv, zv, jv, rq, _trilinear(rho[c], u, v, ww)))
    xs, ys, zs, js, rq, rt = (np.asarray(col, dtype=np.float64) for col in zip(*out))
    return xs, ys, zs, js, rq, rt, subvol


def corrected_moment_delta(geom, rho, center, order):
    """Quadratic sub-cell moments minus nodal-region moments, open_boundary.py:370-464."""
    mask, cells = selected_cells(geom)
    zero = [np.zeros((3,) * l) if l else 0.0 for l in range(order + 1)]
    if cells.size == 0:
        return zero
    frac = node_selected_fraction(mask)
    w = rho * geom.cell_volumes * frac
    nodal = raw_moments(w, geom.x_points - center[0], geom.y_points - center[1], geom.z_points - center[2], order)
    xs, ys, zs, js, rq, _, subvol = _subcell_samples(geom, rho, cells)
    quad = raw_moments(rq * js * subvol, xs - center[0], ys - center[1], zs - center[2], order)
    return [q - n for q, n in zip(quad, nodal)]


def split_boundary_correction(geom, rho, order, bx, by, bz):
    """Boundary-value correction of the `legacy_split` mode, open_boundary.py:559-719:
    d = (rho_quadratic - rho_trilinear) J subvol at the sub-cell samples, expanded to the
    same multipole order about the |d|-weighted centre of the samples.  Returns the
    correction at the points (bx, by, bz) and the centre (None when there is none)."""
    _, cells = selected_cells(geom)
    if cells.size == 0:
        return np.zeros_like(bx), None
    xs, ys, zs, js, rq, rt, subvol = _subcell_samples(geom, rho, cells)
    d = (rq - rt) * js * subvol
    a = np.abs(d)
    if a.sum() <= 0.0:
        return np.zeros_like(bx), None
    cc = np.array([np.sum(a * xs), np.sum(a * ys), np.sum(a * zs)]) / a.sum()
    mom = raw_moments(d, xs - cc[0], ys - cc[1], zs - cc[2], order)
    return boundary_potential(mom, bx - cc[0], by - cc[1], bz - cc[2], order), cc


# ---------------------------------------------------------------- boundary condition


def multipole_boundary(geom, rho, order=2, mode="corrected_moments", center=None):
    """Boundary values (zero inside) and the moment list used, open_boundary.py:794-987."""
    if center is None:
        center = geom.reference_center
    dx = geom.x_points - center[0]
    dy = geom.y_points - center[1]
    dz = geom.z_points - center[2]
    # "nodal" (nodal moments only, the construction the reference's recorded 67x67x81
    # Hartree energy used) exists for pinning; the reference's modes are the other two.
    if mode not in ("corrected_moments", "legacy_split", "nodal"):
        raise ValueError(f"unknown monitor boundary construction mode {mode!r}")
    baseline = raw_moments(rho * geom.cell_volumes, dx, dy, dz, order)
    moments = baseline
    monitor = getattr(geom, "kind", "") == "monitor" and mode != "nodal"
    if monitor:
        delta = corrected_moment_delta(geom, rho, center, order)
        moments = [m + d for m, d in zip(baseline, delta)]
    bmask = _boundary_mask(geom.cell_volumes.shape)
    b = np.zeros(geom.cell_volumes.shape)
    if monitor and mode == "legacy_split":
        # open_boundary.py:923-934: nodal-moment boundary plus the split correction; the
        # reported moments stay the corrected ones.
        corr, _ = split_boundary_correction(geom, rho, order, geom.x_points[bmask], geom.y_points[bmask],
                                            geom.z_points[bmask])
        b[bmask] = boundary_potential(baseline, dx[bmask], dy[bmask], dz[bmask], order) + corr
    else:
        b[bmask] = boundary_potential(moments, dx[bmask], dy[bmask], dz[bmask], order)
    return b, moments


# ---------------------------------------------------------------- solves


def _interior_system(geom, closure="production"):
    if getattr(geom, "kind", "") == "legacy":
        L = kinetic.legacy_laplacian_matrix(geom)
    else:
        L = kinetic.monitor_laplacian_matrix(geom, closure)
    interior = ~_boundary_mask(geom.cell_volumes.shape).reshape(-1)
    return L, interior


def laplacian(field, geom):
    if getattr(geom, "kind", "") == "legacy":
        return kinetic.legacy_laplacian(field, geom)
    return kinetic.monitor_laplacian(field, geom, "production")


def solve_poisson(geom, rho, order=2, mode="corrected_moments", method="direct", tol=1e-13, maxiter=20000):
    """Full open-boundary solve. Returns (potential, boundary field, moments, info).

    The discrete interior problem is -L u = 4 pi rho + L b on interior nodes with u = 0
    on the boundary (monitor: open_boundary.py:1124-1146; legacy: the same system,
    open_boundary.py:1200-1226, since its rhs folds a+- b into the faces).
    """
    rho = np.maximum(np.asarray(rho, dtype=np.float64), 0.0)
    b, moments = multipole_boundary(geom, rho, order, mode)
    rhs = FOUR_PI * rho + laplacian(b, geom)
    L, interior = _interior_system(geom)
    A = (-L)[interior][:, interior]
    r = rhs.reshape(-1)[interior]
    info = {}
    if method == "direct":
        u = spla.spsolve(A.tocsc(), r)
        info["iterations"] = 0
    else:
        dinv = 1.0 / A.diagonal()
        M = spla.LinearOperator(A.shape, matvec=lambda v: dinv * v)
        it = [0]

        def cb(_):
            it[0] += 1

        u, code = spla.bicgstab(A, r, rtol=tol, atol=0.0, maxiter=maxiter, M=M, callback=cb)
        info["iterations"] = it[0]
        info["code"] = code
    pot = b.copy().reshape(-1)
    pot[interior] = u
    pot = pot.reshape(rho.shape)
    info["residual_max"] = float(np.max(np.abs((A @ u) - r)))
    return pot, b, moments, info


def hartree_energy(rho, geom, potential):
    """hartree.py:117-129."""
    return 0.5 * float(np.sum(np.maximum(rho, 0.0) * potential * geom.cell_volumes))

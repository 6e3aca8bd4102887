"""Oracle: kinetic / Laplacian operators on the legacy and monitor grids.

Restates ops/kinetic.py:
  * legacy separable flux form, zero ghosts        ops/kinetic.py:87-153
  * monitor "trial_fix" (zero-ghost centered)      ops/kinetic.py:165-187, :212-276, :310-358
  * monitor "production" (np.gradient edge_order=2) ops/kinetic.py:190-209, :273-307, :340-346
  * JAX twins with identical closures               ops/kinetic_jax.py:39-127
plus the exact Euclidean transposes (new; the reference has no adjoints) and a
sparse-matrix assembly used by the Poisson direct solve.

Pinning (tests/test_oracle_kinetic.py): skew-symmetry of the zero-ghost derivative
(W-self-adjointness to round-off), the closed-form spectrum of the zero-ghost
operator on an affine map, exactness of the production closure on quadratics,
brute-force dense matrices for the transposes, and the reference-held <T> scalars.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def _spacing(axis):
    return float(axis[1] - axis[0])


# ---------------------------------------------------------------- legacy (7-point)


def legacy_axis_coefficients(point_jacobian, logical):
    """Per-axis (a_minus, a_plus) so that lap_axis(i) = a+ (f[i+1]-f[i]) - a- (f[i]-f[i-1]).

    ops/kinetic.py:87-114: face scale = mean of neighbouring point scales; the two
    end fluxes use the end point scale with a zero ghost.
    """
    h = _spacing(logical)
    s = np.asarray(point_jacobian, dtype=np.float64)
    face = 0.5 * (s[:-1] + s[1:])
    am = np.empty_like(s)
    ap = np.empty_like(s)
    am[0] = 1.0 / (h * s[0]) / (h * s[0])
    am[1:] = 1.0 / (h * face) / (h * s[1:])
    ap[:-1] = 1.0 / (h * face) / (h * s[:-1])
    ap[-1] = 1.0 / (h * s[-1]) / (h * s[-1])
    return am, ap


def _legacy_axis_laplacian(field, scale, h, axis):
    """Literal restatement of ops/kinetic.py:87-114."""
    f = np.moveaxis(field, axis, 0).astype(np.float64)
    face = 0.5 * (scale[:-1] + scale[1:])
    sh = (slice(None),) + (None,) * (f.ndim - 1)
    fm = np.empty_like(f)
    fp = np.empty_like(f)
    fm[0] = f[0] / (h * scale[0])
    fm[1:] = (f[1:] - f[:-1]) / (h * face[sh])
    fp[:-1] = (f[1:] - f[:-1]) / (h * face[sh])
    fp[-1] = -f[-1] / (h * scale[-1])
    return np.moveaxis((fp - fm) / (h * scale[sh]), 0, axis)


def legacy_laplacian(psi, geom):
    out = _legacy_axis_laplacian(psi, geom.x_point_jacobian, _spacing(geom.x_logical), 0)
    out = out + _legacy_axis_laplacian(psi, geom.y_point_jacobian, _spacing(geom.y_logical), 1)
    out = out + _legacy_axis_laplacian(psi, geom.z_point_jacobian, _spacing(geom.z_logical), 2)
    return out


def legacy_kinetic(psi, geom):
    """ops/kinetic.py:147-153."""
    return -0.5 * legacy_laplacian(psi, geom)


def legacy_laplacian_transpose(y, geom):
    """Euclidean transpose of ``legacy_laplacian`` (brute-force pinned)."""
    out = np.zeros_like(y, dtype=np.float64)
    for axis, (s, lg) in enumerate(
        [
            (geom.x_point_jacobian, geom.x_logical),
            (geom.y_point_jacobian, geom.y_logical),
            (geom.z_point_jacobian, geom.z_logical),
        ]
    ):
        am, ap = legacy_axis_coefficients(s, lg)
        f = np.moveaxis(y, axis, 0)
        sh = (slice(None),) + (None,) * (f.ndim - 1)
        t = -(am + ap)[sh] * f
        t[1:] += ap[:-1][sh] * f[:-1]
        t[:-1] += am[1:][sh] * f[1:]
        out += np.moveaxis(t, 0, axis)
    return out


# ---------------------------------------------------------------- monitor (19-point)


def _grad_zero_ghost(f, h, axis):
    """ops/kinetic.py:165-187."""
    m = np.moveaxis(f, axis, 0)
    g = np.empty_like(m, dtype=np.float64)
    g[1:-1] = (m[2:] - m[:-2]) / (2.0 * h)
    g[0] = (m[1] - 0.0) / (2.0 * h)
    g[-1] = (0.0 - m[-2]) / (2.0 * h)
    return np.moveaxis(g, 0, axis)


def _grad_production(f, h, axis):
    """np.gradient(edge_order=2) on a uniform axis, ops/kinetic.py:202-208 / kinetic_jax.py:58-60."""
    m = np.moveaxis(f, axis, 0)
    g = np.empty_like(m, dtype=np.float64)
    g[1:-1] = (m[2:] - m[:-2]) / (2.0 * h)
    g[0] = (-3.0 * m[0] + 4.0 * m[1] - m[2]) / (2.0 * h)
    g[-1] = (3.0 * m[-1] - 4.0 * m[-2] + m[-3]) / (2.0 * h)
    return np.moveaxis(g, 0, axis)


def _grad_transpose(f, h, axis, closure):
    """Exact transpose of the 1D derivative matrix for the given closure."""
    m = np.moveaxis(f, axis, 0)
    n = m.shape[0]
    g = np.zeros_like(m, dtype=np.float64)
    # interior rows j: D[j, j-1] = -1/2h, D[j, j+1] = +1/2h
    lo = 1 if closure == "production" else 0
    hi = n - 1 if closure == "production" else n
    for j in range(lo, hi):
        if j - 1 >= 0:
            g[j - 1] += -m[j] / (2.0 * h)
        if j + 1 < n:
            g[j + 1] += m[j] / (2.0 * h)
    if closure == "production":
        g[0] += -3.0 * m[0] / (2.0 * h)
        g[1] += 4.0 * m[0] / (2.0 * h)
        g[2] += -1.0 * m[0] / (2.0 * h)
        g[n - 1] += 3.0 * m[n - 1] / (2.0 * h)
        g[n - 2] += -4.0 * m[n - 1] / (2.0 * h)
        g[n - 3] += 1.0 * m[n - 1] / (2.0 * h)
    return np.moveaxis(g, 0, axis)


def _grad(f, h, axis, closure):
    if closure == "trial_fix":
        return _grad_zero_ghost(f, h, axis)
    if closure == "production":
        return _grad_production(f, h, axis)
    raise ValueError(f"unknown closure {closure!r}")


def monitor_laplacian(psi, geom, closure="production"):
    """(1/J) d_a (J g^{ab} d_b psi), ops/kinetic.py:212-337."""
    hs = (_spacing(geom.logical_x), _spacing(geom.logical_y), _spacing(geom.logical_z))
    J = geom.jacobian
    gi = geom.inverse_metric_tensor
    d = [_grad(psi, hs[a], a, closure) for a in range(3)]
    div = 0.0
    for a in range(3):
        flux = J * (gi[..., a, 0] * d[0] + gi[..., a, 1] * d[1] + gi[..., a, 2] * d[2])
        div = div + _grad(flux, hs[a], a, closure)
    return div / J


def monitor_kinetic(psi, geom, closure="production"):
    """ops/kinetic.py:340-358."""
    return -0.5 * monitor_laplacian(psi, geom, closure)


def monitor_laplacian_transpose(y, geom, closure="production"):
    """Euclidean transpose of ``monitor_laplacian``: D_b^T C^{ab} D_a^T (J^-1 y)."""
    hs = (_spacing(geom.logical_x), _spacing(geom.logical_y), _spacing(geom.logical_z))
    J = geom.jacobian
    gi = geom.inverse_metric_tensor
    u = y / J
    d = [_grad_transpose(u, hs[a], a, closure) for a in range(3)]
    out = 0.0
    for b in range(3):
        flux = J * (gi[..., 0, b] * d[0] + gi[..., 1, b] * d[1] + gi[..., 2, b] * d[2])
        out = out + _grad_transpose(flux, hs[b], b, closure)
    return out


def kinetic_operator(psi, geom, closure="production"):
    """Dispatcher, ops/kinetic.py:369-374 (legacy grids ignore ``closure``)."""
    if getattr(geom, "kind", "") == "legacy":
        return legacy_kinetic(psi, geom)
    return monitor_kinetic(psi, geom, closure)


# ---------------------------------------------------------------- sparse assembly


def _deriv_matrix(n, h, closure):
    rows, cols, vals = [], [], []
    for j in range(n):
        if closure == "production" and j == 0:
            entries = [(0, -3.0), (1, 4.0), (2, -1.0)]
        elif closure == "production" and j == n - 1:
            entries = [(n - 3, 1.0), (n - 2, -4.0), (n - 1, 3.0)]
        else:
            entries = [(j - 1, -1.0), (j + 1, 1.0)]
        for c, v in entries:
            if 0 <= c < n:
                rows.append(j)
                cols.append(c)
                vals.append(v / (2.0 * h))
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def monitor_laplacian_matrix(geom, closure="production"):
    """Sparse (N x N) matrix of ``monitor_laplacian`` (C order)."""
    nx, ny, nz = geom.jacobian.shape
    hs = (_spacing(geom.logical_x), _spacing(geom.logical_y), _spacing(geom.logical_z))
    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    Dx = sp.kron(sp.kron(_deriv_matrix(nx, hs[0], closure), Iy), Iz).tocsr()
    Dy = sp.kron(sp.kron(Ix, _deriv_matrix(ny, hs[1], closure)), Iz).tocsr()
    Dz = sp.kron(sp.kron(Ix, Iy), _deriv_matrix(nz, hs[2], closure)).tocsr()
    D = [Dx, Dy, Dz]
    J = geom.jacobian.reshape(-1)
    gi = geom.inverse_metric_tensor.reshape(-1, 3, 3)
    L = None
    for a in range(3):
        flux = None
        for b in range(3):
            term = sp.diags(J * gi[:, a, b]) @ D[b]
            flux = term if flux is None else flux + term
        t = D[a] @ flux
        L = t if L is None else L + t
    return (sp.diags(1.0 / J) @ L).tocsr()


def legacy_laplacian_matrix(geom):
    nx, ny, nz = geom.cell_volumes.shape
    mats = []
    for n, s, lg in [
        (nx, geom.x_point_jacobian, geom.x_logical),
        (ny, geom.y_point_jacobian, geom.y_logical),
        (nz, geom.z_point_jacobian, geom.z_logical),
    ]:
        am, ap = legacy_axis_coefficients(s, lg)
        main = -(am + ap)
        mats.append(sp.diags([am[1:], main, ap[:-1]], [-1, 0, 1], shape=(n, n)))
    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    return (
        sp.kron(sp.kron(mats[0], Iy), Iz)
        + sp.kron(sp.kron(Ix, mats[1]), Iz)
        + sp.kron(sp.kron(Ix, Iy), mats[2])
    ).tocsr()


# ---------------------------------------------------------------- Green identity


def green_identity_terms(f, geom, closure="trial_fix"):
    """(K_op, K_grad, K_bdry, <f,f>_W), unnormalised, for a real field f on a monitor grid
    (audit/h2_monitor_grid_kinetic_green_identity_audit.py:143-218, :295-330):
      K_op   = sum w f T f                      T of ``closure``
      K_grad = 1/2 sum w d_a f g^{ab} d_b f     np.gradient edge_order=2 (:156-176)
      K_bdry = sum_faces s 1/2 f F^a dS         face nodes off every edge, s = +1 at the
               minimum face, -1 at the maximum face, F^a = J g^{ab} d_b f of ``closure``
               (:179-218)."""
    hs = (_spacing(geom.logical_x), _spacing(geom.logical_y), _spacing(geom.logical_z))
    w = np.asarray(geom.cell_volumes, dtype=np.float64)
    J = np.asarray(geom.jacobian, dtype=np.float64)
    gi = np.asarray(geom.inverse_metric_tensor, dtype=np.float64)
    f = np.asarray(f, dtype=np.float64)
    k_op = float(np.sum(w * f * monitor_kinetic(f, geom, closure)))
    gp = [_grad_production(f, hs[a], a) for a in range(3)]
    dens = np.zeros_like(f)
    for a in range(3):
        for b in range(3):
            dens += gp[a] * gi[..., a, b] * gp[b]
    k_grad = float(np.sum(0.5 * dens * w))
    gc = [_grad(f, hs[a], a, closure) for a in range(3)]
    flux = [J * (gi[..., a, 0] * gc[0] + gi[..., a, 1] * gc[1] + gi[..., a, 2] * gc[2]) for a in range(3)]
    inner = slice(1, -1)
    k_bdry = 0.0
    for a in range(3):
        area = hs[(a + 1) % 3] * hs[(a + 2) % 3]
        for idx, sgn in ((0, 0.5), (-1, -0.5)):
            sl = [inner, inner, inner]
            sl[a] = idx
            sl = tuple(sl)
            k_bdry += float(np.sum(sgn * f[sl] * flux[a][sl]) * area)
    den = float(np.sum(w * f * f))
    return k_op, k_grad, k_bdry, den

"""Oracle: structured-grid geometry (legacy sinh grid and monitor harmonic-map grid).

Restates:
  * legacy separable sinh mapping          grid/mapping.py:22-135, grid/geometry.py:38-104,
                                            grid/__init__.py:46-104
  * monitor field, smoothing, harmonic map  grid/monitor_geometry.py:51-136, :268-321
  * covariant basis / metric / Jacobian     grid/monitor_geometry.py:324-373
  * backtracking + generator                grid/monitor_geometry.py:376-396, :481-577
  * element parameters                      grid/monitor_builder.py:25-245

Pinning: the legacy grid is pinned by the reference-held kinetic / local-ionic
scalars on the default 51^3 H2 grid (tests/test_oracle_pins.py); the monitor
generator is pinned by a reference-generated 19x19x23 fixture (tests/golden) and by
the reference-held <T> = 0.9671226925733339 Ha on the 67x67x81 H2 grid.
"""

from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

# ----------------------------------------------------------------------------------
# Molecules (config/defaults.py:14-131)
# ----------------------------------------------------------------------------------


@dataclass(frozen=True)
class Atom:
    element: str
    position: tuple


@dataclass(frozen=True)
class Case:
    name: str
    atoms: tuple


H2_CASE = Case("H2", (Atom("H", (0.0, 0.0, -0.7)), Atom("H", (0.0, 0.0, 0.7))))


def h2o_case() -> Case:
    """H2O geometry, config/defaults.py:17-18, :109-131."""
    half = 0.5 * np.radians(104.5)
    hx = 1.809 * np.sin(half)
    hz = 1.809 * np.cos(half)
    return Case(
        "H2O",
        (
            Atom("O", (0.0, 0.0, 0.0)),
            Atom("H", (-hx, 0.0, hz)),
            Atom("H", (hx, 0.0, hz)),
        ),
    )


# ----------------------------------------------------------------------------------
# Legacy sinh grid
# ----------------------------------------------------------------------------------


@dataclass
class LegacyGeometry:
    shape: tuple
    x_logical: np.ndarray
    y_logical: np.ndarray
    z_logical: np.ndarray
    x_coordinates: np.ndarray
    y_coordinates: np.ndarray
    z_coordinates: np.ndarray
    x_points: np.ndarray
    y_points: np.ndarray
    z_points: np.ndarray
    x_point_jacobian: np.ndarray
    y_point_jacobian: np.ndarray
    z_point_jacobian: np.ndarray
    cell_volumes: np.ndarray
    reference_center: tuple
    kind: str = "legacy"


def _sinh_axis(n: int, lower: float, upper: float, stretch: float, center: float):
    """grid/mapping.py:22-88: piecewise sinh stretch about ``center``."""
    u = np.linspace(-1.0, 1.0, n, dtype=np.float64)
    a = np.abs(u)
    if stretch == 0.0:
        frac = a
        dfrac = np.ones_like(a)
    else:
        frac = np.sinh(stretch * a) / np.sinh(stretch)
        dfrac = stretch * np.cosh(stretch * a) / np.sinh(stretch)
    neg = u < 0.0
    x = np.empty_like(u)
    jac = np.empty_like(u)
    x[neg] = center - (-lower) * frac[neg]
    x[~neg] = center + upper * frac[~neg]
    jac[neg] = (-lower) * dfrac[neg]
    jac[~neg] = upper * dfrac[~neg]
    return u, x, jac


def _cell_widths(points: np.ndarray) -> np.ndarray:
    """grid/geometry.py:38-57."""
    edges = np.empty(points.size + 1)
    edges[1:-1] = 0.5 * (points[1:] + points[:-1])
    edges[0] = points[0] - 0.5 * (points[1] - points[0])
    edges[-1] = points[-1] + 0.5 * (points[-1] - points[-2])
    return np.diff(edges)


def build_legacy_geometry(
    shape=(51, 51, 51),
    half_extents=(8.0, 8.0, 10.0),
    stretch=(2.5, 2.5, 2.8),
    center=None,
    case: Case = H2_CASE,
) -> LegacyGeometry:
    """grid/__init__.py:46-104 + grid/geometry.py:60-104 (defaults = the H2 default grid)."""
    if center is None:
        pos = np.array([a.position for a in case.atoms], dtype=np.float64)
        center = tuple(float(v) for v in pos.mean(axis=0))
    axes = [
        _sinh_axis(n, -h, h, s, c)
        for n, h, s, c in zip(shape, half_extents, stretch, center)
    ]
    (ux, xc, jx), (uy, yc, jy), (uz, zc, jz) = axes
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    wx, wy, wz = _cell_widths(xc), _cell_widths(yc), _cell_widths(zc)
    vol = wx[:, None, None] * wy[None, :, None] * wz[None, None, :]
    return LegacyGeometry(
        tuple(shape), ux, uy, uz, xc, yc, zc, X, Y, Z, jx, jy, jz, vol, tuple(center)
    )


# ----------------------------------------------------------------------------------
# Monitor grid
# ----------------------------------------------------------------------------------


@dataclass(frozen=True)
class ElementParams:
    """grid/monitor_model.py:14-43."""

    element: str
    near_core_radius: float
    local_radius: float
    projector_radius: float
    patch_radius: float
    kinetic_weight: float
    local_weight: float
    projector_weight: float
    kinetic_exponent: float
    local_exponent: float
    projector_exponent: float
    monitor_cap: float


_SEEDS = {
    # grid/monitor_builder.py:25-50
    "H": (1, 0.2, (-4.1802368, 0.72507482), ()),
    "C": (4, 0.34883045, (-8.5137711, 1.22843203), (0.30455321,)),
    "N": (5, 0.28917923, (-12.23481988, 1.76640728), (0.25660487,)),
    "O": (6, 0.24762086, (-16.58031797, 2.39570092), (0.22178614,)),
}


def default_element_parameters(case: Case) -> dict:
    """grid/monitor_builder.py:63-102."""
    out = {}
    for atom in case.atoms:
        e = atom.element
        if e in out:
            continue
        z, rloc, coeffs, radii = _SEEDS[e]
        max_nl = max(radii, default=1.5 * rloc)
        cscale = sum(abs(v) for v in coeffs)
        out[e] = ElementParams(
            element=e,
            near_core_radius=max(0.35, 1.8 * rloc + 0.15),
            local_radius=max(0.45, 2.5 * rloc + 0.15),
            projector_radius=max(0.55, 2.4 * max_nl),
            patch_radius=max(1.0, 4.0 * max(rloc, max_nl) + 0.25),
            kinetic_weight=1.5 + 0.9 * z,
            local_weight=2.0 + 1.1 * z + 0.05 * cscale,
            projector_weight=0.45 * z if radii else 0.0,
            kinetic_exponent=2.0,
            local_exponent=4.0,
            projector_exponent=2.0,
            monitor_cap=4.0 + 0.8 * z,
        )
    return out


def h2_patch_development_parameters() -> dict:
    """grid/monitor_builder.py:204-216 (weight x4, radius x0.70)."""
    base = default_element_parameters(H2_CASE)
    return {
        e: replace(
            p,
            near_core_radius=p.near_core_radius * 0.70,
            local_radius=p.local_radius * 0.70,
            kinetic_weight=p.kinetic_weight * 4.0,
            local_weight=p.local_weight * 4.0,
        )
        for e, p in base.items()
    }


@dataclass
class MonitorGeometry:
    shape: tuple
    box_bounds: tuple
    logical_x: np.ndarray
    logical_y: np.ndarray
    logical_z: np.ndarray
    x_points: np.ndarray
    y_points: np.ndarray
    z_points: np.ndarray
    covariant_basis: np.ndarray
    jacobian: np.ndarray
    metric_tensor: np.ndarray
    inverse_metric_tensor: np.ndarray
    cell_volumes: np.ndarray
    patch_radii: tuple = ()
    kind: str = "monitor"

    @property
    def reference_center(self):
        b = self.box_bounds
        return tuple(0.5 * (lo + hi) for lo, hi in b)


def _monitor_field(case, params, X, Y, Z):
    """grid/monitor_geometry.py:51-112 (only the capped total is used by the generator)."""
    values = np.full(X.shape, 1.0)
    for atom in case.atoms:
        p = params[atom.element]
        dx, dy, dz = X - atom.position[0], Y - atom.position[1], Z - atom.position[2]
        r = np.sqrt(dx * dx + dy * dy + dz * dz, dtype=np.float64)
        kin = p.kinetic_weight * np.exp(-np.power(r / p.near_core_radius, p.kinetic_exponent))
        loc = p.local_weight / (1.0 + np.power(r / p.local_radius, p.local_exponent))
        if p.projector_weight > 0.0:
            prj = p.projector_weight * np.exp(-np.power(r / p.projector_radius, p.projector_exponent))
        else:
            prj = np.zeros(X.shape)
        raw = kin + loc + prj
        values += p.monitor_cap * (1.0 - np.exp(-raw / p.monitor_cap))
    return values


def _smooth(values, s):
    """grid/monitor_geometry.py:115-136."""
    if s <= 0.0:
        return values
    m = values
    out = m.copy()
    avg = (
        m[1:-1, 1:-1, 1:-1] + m[:-2, 1:-1, 1:-1] + m[2:, 1:-1, 1:-1]
        + m[1:-1, :-2, 1:-1] + m[1:-1, 2:, 1:-1]
        + m[1:-1, 1:-1, :-2] + m[1:-1, 1:-1, 2:]
    ) / 7.0
    out[1:-1, 1:-1, 1:-1] = (1.0 - s) * m[1:-1, 1:-1, 1:-1] + s * avg
    return np.maximum(out, 1.0)


def _boundary_mask(shape):
    mask = np.zeros(shape, dtype=bool)
    mask[0], mask[-1] = True, True
    mask[:, 0], mask[:, -1] = True, True
    mask[:, :, 0], mask[:, :, -1] = True, True
    return mask


def _harmonic(mon, lx, ly, lz, bcoords, coords, inner, tol, relax):
    """grid/monitor_geometry.py:268-321 (damped Jacobi on div(M grad X) = 0)."""
    coords = coords.copy()
    bmask = _boundary_mask(mon.shape)
    coords[bmask] = bcoords[bmask]
    dx, dy, dz = lx[1] - lx[0], ly[1] - ly[0], lz[1] - lz[0]
    c = mon[1:-1, 1:-1, 1:-1]
    axm = 0.5 * (c + mon[:-2, 1:-1, 1:-1]) / (dx * dx)
    axp = 0.5 * (c + mon[2:, 1:-1, 1:-1]) / (dx * dx)
    aym = 0.5 * (c + mon[1:-1, :-2, 1:-1]) / (dy * dy)
    ayp = 0.5 * (c + mon[1:-1, 2:, 1:-1]) / (dy * dy)
    azm = 0.5 * (c + mon[1:-1, 1:-1, :-2]) / (dz * dz)
    azp = 0.5 * (c + mon[1:-1, 1:-1, 2:]) / (dz * dz)
    diag = axm + axp + aym + ayp + azm + azp
    for _ in range(inner):
        cand = (
            axm[..., None] * coords[:-2, 1:-1, 1:-1, :]
            + axp[..., None] * coords[2:, 1:-1, 1:-1, :]
            + aym[..., None] * coords[1:-1, :-2, 1:-1, :]
            + ayp[..., None] * coords[1:-1, 2:, 1:-1, :]
            + azm[..., None] * coords[1:-1, 1:-1, :-2, :]
            + azp[..., None] * coords[1:-1, 1:-1, 2:, :]
        ) / diag[..., None]
        upd = coords.copy()
        upd[1:-1, 1:-1, 1:-1, :] = (1.0 - relax) * coords[1:-1, 1:-1, 1:-1, :] + relax * cand
        upd[bmask] = bcoords[bmask]
        change = float(np.max(np.abs(upd[1:-1, 1:-1, 1:-1, :] - coords[1:-1, 1:-1, 1:-1, :])))
        coords = upd
        if change < tol:
            break
    return coords


def metric_from_coordinates(lx, ly, lz, coords):
    """grid/monitor_geometry.py:324-373: B = d x / d xi via np.gradient(edge_order=2)."""
    basis = np.zeros(coords.shape[:-1] + (3, 3))
    for c in range(3):
        basis[..., 0, c] = np.gradient(coords[..., c], lx, axis=0, edge_order=2)
        basis[..., 1, c] = np.gradient(coords[..., c], ly, axis=1, edge_order=2)
        basis[..., 2, c] = np.gradient(coords[..., c], lz, axis=2, edge_order=2)
    g = np.einsum("...ai,...bi->...ab", basis, basis)
    ginv = np.linalg.inv(g)
    jac = np.linalg.det(basis)
    vol = jac * (lx[1] - lx[0]) * (ly[1] - ly[0]) * (lz[1] - lz[0])
    return basis, jac, g, ginv, vol


def generate_monitor_geometry(
    case: Case,
    shape,
    box_half_extents=None,
    element_params=None,
    outer=6,
    inner=180,
    tol=5.0e-5,
    relax=0.55,
    inner_relax=0.82,
    smoothing=0.18,
) -> MonitorGeometry:
    """grid/monitor_builder.py:105-201 + grid/monitor_geometry.py:481-577."""
    if element_params is None:
        element_params = default_element_parameters(case)
    pos = np.array([a.position for a in case.atoms], dtype=np.float64)
    center = pos.mean(axis=0)
    if box_half_extents is None:
        support = max(
            max(p.patch_radius, p.local_radius, p.projector_radius)
            for p in element_params.values()
        )
        pad = max(6.0, 4.0 * support)
        mad = np.max(np.abs(pos - center), axis=0)
        box_half_extents = tuple(float(m + pad) for m in mad)
    bounds = tuple(
        (float(c - h), float(c + h)) for c, h in zip(center, box_half_extents)
    )
    nx, ny, nz = shape
    lx = np.linspace(0.0, 1.0, nx)
    ly = np.linspace(0.0, 1.0, ny)
    lz = np.linspace(0.0, 1.0, nz)
    px = np.linspace(bounds[0][0], bounds[0][1], nx)
    py = np.linspace(bounds[1][0], bounds[1][1], ny)
    pz = np.linspace(bounds[2][0], bounds[2][1], nz)
    X0, Y0, Z0 = np.meshgrid(px, py, pz, indexing="ij")
    bcoords = np.stack([X0, Y0, Z0], axis=-1)
    coords = bcoords.copy()
    for _ in range(outer):
        mon = _monitor_field(case, element_params, coords[..., 0], coords[..., 1], coords[..., 2])
        mon = _smooth(mon, smoothing)
        solved = _harmonic(mon, lx, ly, lz, bcoords, coords, inner, tol, inner_relax)
        # backtracking, grid/monitor_geometry.py:376-396
        alpha = relax
        updated = coords
        while alpha >= 1.0e-3:
            trial = coords + alpha * (solved - coords)
            _, jac, _, _, _ = metric_from_coordinates(lx, ly, lz, trial)
            if float(np.min(jac)) > 1.0e-12:
                updated = trial
                break
            alpha *= 0.5
        disp = float(np.max(np.abs(updated - coords)))
        coords = updated
        if disp < tol:
            break
    basis, jac, g, ginv, vol = metric_from_coordinates(lx, ly, lz, coords)
    return MonitorGeometry(
        tuple(shape), bounds, lx, ly, lz,
        coords[..., 0].copy(), coords[..., 1].copy(), coords[..., 2].copy(),
        basis, jac, g, ginv, vol,
        tuple(element_params[a.element].patch_radius for a in case.atoms),
    )


def monitor_geometry_from_coordinates(coords, box_bounds, patch_radii=()) -> MonitorGeometry:
    """Geometry container for arbitrary mapped coordinates (metric by np.gradient)."""
    nx, ny, nz = coords.shape[:3]
    lx, ly, lz = np.linspace(0, 1, nx), np.linspace(0, 1, ny), np.linspace(0, 1, nz)
    basis, jac, g, ginv, vol = metric_from_coordinates(lx, ly, lz, coords)
    return MonitorGeometry(
        (nx, ny, nz), tuple(box_bounds), lx, ly, lz,
        coords[..., 0].copy(), coords[..., 1].copy(), coords[..., 2].copy(),
        basis, jac, g, ginv, vol, tuple(patch_radii),
    )

"""Host-side grid geometry: containers and builders (NumPy, FP64, one-off setup).

The kernels consume only the arrays named in the reference's geometry contract
(grid/geometry.py:15-35, grid/monitor_model.py:153-174): logical axes, physical
points, Jacobian, inverse metric tensor and cell volumes.  Any object exposing
those attributes (including the reference's own geometry dataclasses) is accepted
by the device layer; this module only builds the geometries that the benchmark
configurations need:

* ``build_legacy_geometry``       legacy separable sinh grid, grid/mapping.py:22-135,
                                  grid/geometry.py:38-104
* ``build_mapped_monitor_geometry`` monitor-style curvilinear grid from an analytic
                                  multi-centre sinh mapping with its exact Jacobian
                                  (the synthetic operator-level configuration; the
                                  reference's harmonic-map generator is not on the path)
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass(eq=False)
class LegacyGridGeometry:
    shape: tuple
    x_logical: np.ndarray
    y_logical: np.ndarray
    z_logical: np.ndarray
    x_points: np.ndarray
    y_points: np.ndarray
    z_points: np.ndarray
    x_point_jacobian: np.ndarray
    y_point_jacobian: np.ndarray
    z_point_jacobian: np.ndarray
    cell_volumes: np.ndarray
    reference_center: tuple
    kind: str = "legacy"


@dataclass(eq=False)
class MappedMonitorGeometry:
    shape: tuple
    box_bounds: tuple
    logical_x: np.ndarray
    logical_y: np.ndarray
    logical_z: np.ndarray
    x_points: np.ndarray
    y_points: np.ndarray
    z_points: np.ndarray
    jacobian: np.ndarray
    inverse_metric_tensor: np.ndarray
    cell_volumes: np.ndarray
    patch_radii: tuple = ()
    kind: str = "monitor"
    description: str = ""

    @property
    def reference_center(self):
        return tuple(0.5 * (lo + hi) for lo, hi in self.box_bounds)


def _sinh_axis(n, half, stretch, center):
    u = np.linspace(-1.0, 1.0, n)
    a = np.abs(u)
    if stretch == 0.0:
        frac, dfrac = a, np.ones_like(a)
    else:
        frac = np.sinh(stretch * a) / np.sinh(stretch)
        dfrac = stretch * np.cosh(stretch * a) / np.sinh(stretch)
    sign = np.where(u < 0.0, -1.0, 1.0)
    return u, center + sign * half * frac, half * dfrac


def _widths(p):
    e = np.empty(p.size + 1)
    e[1:-1] = 0.5 * (p[1:] + p[:-1])
    e[0] = p[0] - 0.5 * (p[1] - p[0])
    e[-1] = p[-1] + 0.5 * (p[-1] - p[-2])
    return np.diff(e)


def build_legacy_geometry(shape, half_extents, stretch, center) -> LegacyGridGeometry:
    """Separable sinh grid about ``center`` (symmetric half extents)."""
    (ux, xc, jx), (uy, yc, jy), (uz, zc, jz) = [
        _sinh_axis(n, h, s, c) for n, h, s, c in zip(shape, half_extents, stretch, center)
    ]
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    vol = _widths(xc)[:, None, None] * _widths(yc)[None, :, None] * _widths(zc)[None, None, :]
    return LegacyGridGeometry(tuple(shape), ux, uy, uz, X, Y, Z, jx, jy, jz, vol, tuple(center))


def build_mapped_monitor_geometry(
    shape,
    half_extents,
    centers,
    amplitudes,
    b=1.5,
    sigma=1.0,
    box_center=(0.0, 0.0, 0.0),
    patch_radii=(),
    description="",
) -> MappedMonitorGeometry:
    """Analytic curvilinear map x = xi + sum_c a_c w_c(xi) sinh(b (xi - C_c)).

    xi spans the box linearly in the logical coordinate u in [0, 1]; w_c is a
    Gaussian of width ``sigma`` about centre C_c.  The covariant basis is the exact
    Jacobian dx/du (no finite differences); g = B B^T, J = det B and the cell
    volumes J * h_x h_y h_z follow the reference definitions
    (grid/monitor_geometry.py:353-373).
    """
    nx, ny, nz = shape
    lx, ly, lz = (np.linspace(0.0, 1.0, n) for n in shape)
    lo = np.asarray(box_center) - np.asarray(half_extents)
    span = 2.0 * np.asarray(half_extents, dtype=np.float64)
    XI = np.stack(np.meshgrid(lo[0] + span[0] * lx, lo[1] + span[1] * ly, lo[2] + span[2] * lz,
                              indexing="ij"), axis=-1)
    X = XI.copy()
    A = np.zeros(XI.shape[:3] + (3, 3))
    A[..., 0, 0] = A[..., 1, 1] = A[..., 2, 2] = 1.0
    for C, a in zip(centers, amplitudes):
        d = XI - np.asarray(C, dtype=np.float64)
        wgt = np.exp(-np.sum(d * d, axis=-1) / (2.0 * sigma * sigma))
        sh = np.sinh(b * d)
        ch = np.cosh(b * d)
        X += a * wgt[..., None] * sh
        dw = -wgt[..., None] * d / (sigma * sigma)  # d w / d xi_j
        # A[i][j] = d x_i / d xi_j
        A += a * sh[..., :, None] * dw[..., None, :]
        for i in range(3):
            A[..., i, i] += a * wgt * b * ch[..., i]
    B = np.empty_like(A)  # B[a][i] = d x_i / d u_a = A[i][a] * span_a
    for a_ in range(3):
        B[..., a_, :] = A[..., :, a_] * span[a_]
    g = np.einsum("...ai,...bi->...ab", B, B)
    ginv = np.linalg.inv(g)
    J = np.linalg.det(B)
    if not np.all(J > 0.0):
        raise ValueError("mapped monitor geometry folds over (J <= 0)")
    h = (1.0 / (nx - 1)) * (1.0 / (ny - 1)) * (1.0 / (nz - 1))
    bounds = tuple((float(l), float(l + s)) for l, s in zip(lo, span))
    return MappedMonitorGeometry(
        tuple(shape), bounds, lx, ly, lz,
        np.ascontiguousarray(X[..., 0]), np.ascontiguousarray(X[..., 1]), np.ascontiguousarray(X[..., 2]),
        J, ginv, J * h, tuple(patch_radii), "monitor", description,
    )


def grid_kind(geom) -> str:
    if hasattr(geom, "inverse_metric_tensor"):
        return "monitor"
    if hasattr(geom, "x_point_jacobian"):
        return "legacy"
    raise TypeError("grid_geometry must be a legacy or monitor geometry")


def grid_shape(geom) -> tuple:
    return tuple(int(v) for v in np.shape(geom.cell_volumes))


def box_bounds(geom):
    if hasattr(geom, "box_bounds"):
        return geom.box_bounds
    return geom.spec.box_bounds


def reference_center(geom):
    """Monitor: box midpoint (poisson/open_boundary.py:768-778); legacy: spec centre."""
    if grid_kind(geom) == "monitor":
        return tuple(0.5 * (lo + hi) for lo, hi in box_bounds(geom))
    if hasattr(geom, "reference_center"):
        return tuple(geom.reference_center)
    return tuple(geom.spec.reference_center)


def logical_axes(geom):
    if grid_kind(geom) == "monitor":
        return geom.logical_x, geom.logical_y, geom.logical_z
    return geom.x_logical, geom.y_logical, geom.z_logical


def uniform_spacing(axis, label="axis") -> float:
    a = np.asarray(axis, dtype=np.float64)
    d = np.diff(a)
    if not np.allclose(d, d[0]):
        raise ValueError(f"The kernels expect a uniform logical {label}-axis.")
    return float(d[0])

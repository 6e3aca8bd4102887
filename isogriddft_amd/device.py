"""Device-resident geometry: the HBM layout every kernel streams.

Monitor grids (SoA, C order, z fastest):
    coef  [6, N]  c^{ab} = J g^{ab} in the order xx, yy, zz, xy, xz, yz
    invj  [N]     1/J
    w     [N]     cell volumes (the quadrature weights W)
    xyz   [3, N]  physical coordinates (projectors, multipoles, V_loc)
Legacy grids: per-axis flux coefficients (a-, a+) of the 7-point operator
(ops/kinetic.py:87-114) plus w and xyz.  FP64 copies are built once per geometry;
FP32 copies are made on first use.  Geometry validation follows
ops/kinetic.py:156-162 (finite, strictly positive Jacobian).
"""

from __future__ import annotations

import weakref

import numpy as np
import torch

from . import grid as _grid

_CACHE: dict = {}


def _legacy_axis_coefficients(point_jacobian, logical):
    h = _grid.uniform_spacing(logical)
    s = np.asarray(point_jacobian, dtype=np.float64)
    face = 0.5 * (s[:-1] + s[1:])
    am = np.empty_like(s)
    ap = np.empty_like(s)
    am[0] = 1.0 / (h * s[0]) / (h * s[0])
    am[1:] = 1.0 / (h * face) / (h * s[1:])
    ap[:-1] = 1.0 / (h * face) / (h * s[:-1])
    ap[-1] = 1.0 / (h * s[-1]) / (h * s[-1])
    return am, ap


class DeviceGrid:
    """Geometry arrays uploaded to one GPU, in FP64 and (lazily) FP32."""

    def __init__(self, geom, device="cuda"):
        self.kind = _grid.grid_kind(geom)
        self.shape = _grid.grid_shape(geom)
        self.N = int(np.prod(self.shape))
        self.device = torch.device(device)
        dev = self.device
        f64 = torch.float64
        w = np.asarray(geom.cell_volumes, dtype=np.float64).reshape(-1)
        self._t = {}
        self._t[("w", f64)] = torch.from_numpy(np.ascontiguousarray(w)).to(dev)
        xyz = np.stack([np.asarray(geom.x_points, dtype=np.float64).reshape(-1),
                        np.asarray(geom.y_points, dtype=np.float64).reshape(-1),
                        np.asarray(geom.z_points, dtype=np.float64).reshape(-1)])
        self._t[("xyz", f64)] = torch.from_numpy(np.ascontiguousarray(xyz)).to(dev)
        self.reference_center = _grid.reference_center(geom)
        if self.kind == "monitor":
            J = np.asarray(geom.jacobian, dtype=np.float64)
            gi = np.asarray(geom.inverse_metric_tensor, dtype=np.float64)
            if not np.all(np.isfinite(J)):
                raise ValueError("Monitor-grid Jacobian must be finite.")
            if np.any(J <= 0.0):
                raise ValueError("Monitor-grid Jacobian must stay strictly positive.")
            if not np.all(np.isfinite(gi)):
                raise ValueError("Monitor-grid inverse metric tensor must be finite.")
            Jf = J.reshape(-1)
            gf = gi.reshape(-1, 3, 3)
            # c^{ab} = J g^{ab}; the off-diagonals take the mean of (a,b) and (b,a) so the
            # stored tensor is exactly symmetric.
            coef = np.stack([
                Jf * gf[:, 0, 0], Jf * gf[:, 1, 1], Jf * gf[:, 2, 2],
                Jf * 0.5 * (gf[:, 0, 1] + gf[:, 1, 0]),
                Jf * 0.5 * (gf[:, 0, 2] + gf[:, 2, 0]),
                Jf * 0.5 * (gf[:, 1, 2] + gf[:, 2, 1]),
            ])
            self._t[("coef", f64)] = torch.from_numpy(np.ascontiguousarray(coef)).to(dev)
            self._t[("invj", f64)] = torch.from_numpy(1.0 / Jf).to(dev)
            self._t[("jac", f64)] = torch.from_numpy(np.ascontiguousarray(Jf)).to(dev)
            lx, ly, lz = _grid.logical_axes(geom)
            self.h = (_grid.uniform_spacing(lx, "x"), _grid.uniform_spacing(ly, "y"),
                      _grid.uniform_spacing(lz, "z"))
            # stencil-scaled coefficients c^{ab} / (4 h_a h_b), the kernels' operand
            hx, hy, hz = self.h
            fac = np.array([1 / (4 * hx * hx), 1 / (4 * hy * hy), 1 / (4 * hz * hz),
                            1 / (4 * hx * hy), 1 / (4 * hx * hz), 1 / (4 * hy * hz)])
            self._t[("scoef", f64)] = torch.from_numpy(np.ascontiguousarray(coef * fac[:, None])).to(dev)
            self.box_bounds = tuple(tuple(b) for b in _grid.box_bounds(geom))
        else:
            parts = []
            for s, lg in ((geom.x_point_jacobian, geom.x_logical),
                          (geom.y_point_jacobian, geom.y_logical),
                          (geom.z_point_jacobian, geom.z_logical)):
                am, ap = _legacy_axis_coefficients(s, lg)
                parts += [am, ap]
            self._t[("lcoef", f64)] = torch.from_numpy(np.concatenate(parts)).to(dev)
            self.h = tuple(_grid.uniform_spacing(a) for a in _grid.logical_axes(geom))
            # The 7-point flux operator is self-adjoint under the product weights
            # (h s_x)(h s_y)(h s_z), not under the midpoint cell volumes used for integrals.
            sx, sy, sz = (h * np.asarray(s, dtype=np.float64) for h, s in
                          zip(self.h, (geom.x_point_jacobian, geom.y_point_jacobian, geom.z_point_jacobian)))
            wsa = sx[:, None, None] * sy[None, :, None] * sz[None, None, :]
            self._t[("wsa", f64)] = torch.from_numpy(np.ascontiguousarray(wsa.reshape(-1))).to(dev)
            self.box_bounds = None
        self.patch_radii = tuple(getattr(geom, "patch_radii", ()) or ())

    def get(self, name, dtype=torch.float64):
        key = (name, dtype)
        t = self._t.get(key)
        if t is None:
            t = self._t[(name, torch.float64)].to(dtype).contiguous()
            self._t[key] = t
        return t

    def coef(self, dtype=torch.float64):
        return self.get("coef" if self.kind == "monitor" else "lcoef", dtype)

    def invj(self, dtype=torch.float64):
        return self.get("invj", dtype)

    def weights(self, dtype=torch.float64):
        return self.get("w", dtype)

    def kinetic_diagonal(self):
        """Diagonal of the discrete Laplacian (FP64, flat), cached.  Legacy grids: exact,
        -(a-_a + a+_a) summed over axes.  Monitor grids: the centred-difference form
        J^-1 sum_a d_a(c^aa d_a) gives -J^-1 sum_a (c^aa_{i+e_a} + c^aa_{i-e_a}) / (4 h_a^2)
        from the in-domain neighbours (the mixed terms have no diagonal entry); the
        one-sided boundary rows of the production closure are not represented."""
        key = ("ldiag", torch.float64)
        if key in self._t:
            return self._t[key]
        shape = self.shape
        if self.kind == "monitor":
            sc = self.get("scoef").view(6, *shape)
            d = torch.zeros(shape, dtype=torch.float64, device=self.device)
            for a in range(3):
                c = sc[a]
                n = shape[a]
                lo = [slice(None)] * 3
                hi = [slice(None)] * 3
                lo[a] = slice(0, n - 1)
                hi[a] = slice(1, n)
                d[tuple(lo)] += c[tuple(hi)]   # c_{i+1} for i < n-1
                d[tuple(hi)] += c[tuple(lo)]   # c_{i-1} for i > 0
            diag = -(d * self.get("invj").view(shape)).reshape(-1)
        else:
            lc = self.get("lcoef")
            nx, ny, nz = shape
            off = 0
            diag = torch.zeros(shape, dtype=torch.float64, device=self.device)
            for a, n in enumerate((nx, ny, nz)):
                am, ap = lc[off:off + n], lc[off + n:off + 2 * n]
                off += 2 * n
                v = -(am + ap)
                view = [1, 1, 1]
                view[a] = n
                diag += v.view(view)
            diag = diag.reshape(-1)
        self._t[key] = diag.contiguous()
        return self._t[key]

    def operator_weights(self, dtype=torch.float64):
        """Weights under which the trial_fix / legacy kinetic operator is self-adjoint:
        the cell volumes on monitor grids, (h s_x)(h s_y)(h s_z) on legacy grids."""
        return self.get("w" if self.kind == "monitor" else "wsa", dtype)

    def xyz(self):
        return self.get("xyz")

    def nbytes(self):
        return sum(t.numel() * t.element_size() for t in self._t.values())


def device_grid(geom, device="cuda") -> DeviceGrid:
    """Cached DeviceGrid for a geometry object (keyed by identity, dropped with it)."""
    if isinstance(geom, DeviceGrid):
        return geom
    key = (id(geom), str(device))
    hit = _CACHE.get(key)
    if hit is not None:
        ref, dg = hit
        if ref() is geom:
            return dg
    dg = DeviceGrid(geom, device)
    try:
        ref = weakref.ref(geom)
        weakref.finalize(geom, _CACHE.pop, key, None)
    except TypeError:
        ref = lambda g=geom: g  # noqa: E731  (objects without weakref support stay cached)
    _CACHE[key] = (ref, dg)
    return dg

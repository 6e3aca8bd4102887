"""Open-boundary Hartree Poisson solve on the GPU (drop-in for the reference's
poisson/open_boundary.py and poisson/hartree.py).

The discrete problem is the reference's: L v = -4 pi rho with Dirichlet data from a
free-space multipole expansion, split as v = u + b (b = boundary values, zero inside)
and solved for the interior unknown u from

    -L u = 4 pi rho + L b      on interior nodes, u = 0 on the six faces

(monitor: poisson/open_boundary.py:1108-1197 with the production-closure Laplacian;
legacy: :1200-1276 with the separable 7-point operator, whose boundary-folded rhs of
:1200-1213 equals 4 pi rho + L b on interior nodes).

Device pieces, all HIP kernels of libisogrid_hip.so:
  moments      ig_multipole_moments_f64         (open_boundary.py:223-264, :816-861)
  corrected    ig_cell_fit_f64 + ig_cell_moments_f64 minus the nodal-region moments
               (open_boundary.py:370-464; the selected cells and node fractions of
               :182-220 are geometry-only and are built once per grid on the host)
  boundary     ig_boundary_values_f64           (open_boundary.py:467-495)
  matvec       ig_monitor_stencil_* / ig_legacy_stencil_* with interior_only
  Krylov       ig_bicgstab_{p,s,x}_* + ig_dots_*, scalars resident on the device
               (the reference's SciPy BiCGSTAB, :1017-1064 unpreconditioned on monitor
               grids, :1229-1276 Jacobi-preconditioned on legacy grids)
Precision modes: "fp64" (everything FP64) and "mixed" (FP32 BiCGSTAB inner solves
with FP64 residual / defect correction, stopping on the FP64 residual).

Stopping rule: ||r||_2 <= tolerance * ||rhs||_2 on interior nodes, as SciPy's rtol;
the residual is inspected every ``check_every`` iterations.  Multipole orders 0-4
(orders 3 and 4 extend the reference, which stops at 2).
"""

from __future__ import annotations

import itertools
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, multigrid
from .device import DeviceGrid, device_grid
from .kinetic import apply_operator
from .reductions import dots_into, workspace

FOUR_PI = 4.0 * math.pi
_NEGATIVE_DENSITY_TOLERANCE = 1.0e-14
_QUANTILE = 0.9  # open_boundary.py:36
# "nodal" (faces from the nodal raw moments alone) is the monitor-grid construction of
# earlier reference versions, with which the reference's recorded SCF audits were made
_MODES = ("corrected_moments", "legacy_split", "nodal")
_SOLVERS = ("auto", "scipy_bicgstab", "jacobi", "bicgstab")
_JACOBI_DAMPING = 0.8  # open_boundary.py:34
_JACOBI_CHECK_INTERVAL = 25  # open_boundary.py:35


# ------------------------------------------------------------------ result types
# Field names and meaning follow poisson/open_boundary.py:44-102.


@dataclass(frozen=True)
class OpenBoundaryBoundaryConstructionDiagnostics:
    baseline_total_charge: float
    baseline_dipole_moment: np.ndarray
    baseline_quadrupole_tensor: np.ndarray
    correction_total_charge: float
    correction_dipole_moment: np.ndarray
    correction_quadrupole_tensor: np.ndarray
    boundary_value_correction_rms: float
    boundary_value_correction_max_abs: float
    corrected_moment_boundary_rms_mismatch: float
    corrected_moment_boundary_max_abs_mismatch: float


@dataclass(frozen=True)
class OpenBoundaryMultipoleBoundary:
    reference_center: tuple
    multipole_order: int
    total_charge: float
    dipole_moment: np.ndarray
    quadrupole_tensor: np.ndarray
    boundary_values: np.ndarray
    description: str
    diagnostics: OpenBoundaryBoundaryConstructionDiagnostics | None = None
    raw_moments: tuple = ()


@dataclass(frozen=True)
class OpenBoundaryPoissonResult:
    rho: np.ndarray
    potential: np.ndarray
    boundary_condition: OpenBoundaryMultipoleBoundary
    solver_method: str
    solver_iterations: int
    residual_max: float
    description: str
    response_diagnostics: object | None = None
    converged: bool = True
    residual_relative_l2: float = 0.0


# ------------------------------------------------------------------ monomial algebra


# smallest relative residual asked of one FP32 inner solve
MIXED_INNER_FLOOR = 1e-6

def monomial_exponents(order: int):
    """Exponent triples of the raw-moment table: degree by degree, (px, py, pz)
    descending lexicographically (the layout the moment kernels write)."""
    out = []
    for d in range(order + 1):
        for px in range(d, -1, -1):
            for py in range(d - px, -1, -1):
                out.append((px, py, d - px - py))
    return out


def nmono(order: int) -> int:
    return (order + 1) * (order + 2) * (order + 3) // 6


def moment_tensors(flat, order: int):
    """Symmetric tensors M_l[a1..al] = sum q x_a1 ... x_al from the flat table."""
    index = {e: i for i, e in enumerate(monomial_exponents(order))}
    out = []
    for l in range(order + 1):
        if l == 0:
            out.append(float(flat[0]))
            continue
        t = np.zeros((3,) * l)
        for idx in itertools.product(range(3), repeat=l):
            e = [0, 0, 0]
            for a in idx:
                e[a] += 1
            t[idx] = flat[index[tuple(e)]]
        out.append(t)
    return out


def _poly_add(p, e, c):
    p[e] = p.get(e, 0.0) + c


def _poly_mul_r2(p, times=1):
    for _ in range(times):
        q = {}
        for (a, b, c), v in p.items():
            _poly_add(q, (a + 2, b, c), v)
            _poly_add(q, (a, b + 2, c), v)
            _poly_add(q, (a, b, c + 2), v)
        p = q
    return p


def _contract(M, l):
    """sum M[a1..al] x_a1 ... x_al as an exponent -> coefficient map."""
    p = {}
    for idx in itertools.product(range(3), repeat=l):
        e = [0, 0, 0]
        for a in idx:
            e[a] += 1
        _poly_add(p, tuple(e), float(M[idx]))
    return p


def boundary_coefficients(flat, order: int) -> np.ndarray:
    """Coefficients of the homogeneous Q_l with phi = sum_l Q_l(r) / r^(2l+1).

    The free-space expansion phi = sum_l (-1)^l / l! M_{a..} d_{a..}(1/r); for l <= 2
    it is q/r + p.r/r^3 + 1/2 Q_ab r_a r_b / r^5 with Q = 3M - tr(M) 1, the formula
    of open_boundary.py:467-495.
    """
    M = moment_tensors(flat, order)
    polys = [{(0, 0, 0): M[0]}]
    if order >= 1:
        polys.append(_contract(M[1], 1))
    if order >= 2:
        p = {e: 1.5 * v for e, v in _contract(M[2], 2).items()}
        for e, v in _poly_mul_r2({(0, 0, 0): -0.5 * float(np.trace(M[2]))}).items():
            _poly_add(p, e, v)
        polys.append(p)
    if order >= 3:
        p = {e: 15.0 / 6.0 * v for e, v in _contract(M[3], 3).items()}
        tr = {}
        for c in range(3):
            e = [0, 0, 0]
            e[c] = 1
            _poly_add(tr, tuple(e), -9.0 / 6.0 * sum(M[3][a, a, c] for a in range(3)))
        for e, v in _poly_mul_r2(tr).items():
            _poly_add(p, e, v)
        polys.append(p)
    if order >= 4:
        p = {e: 105.0 / 24.0 * v for e, v in _contract(M[4], 4).items()}
        t2 = {}
        for c, d in itertools.product(range(3), repeat=2):
            e = [0, 0, 0]
            e[c] += 1
            e[d] += 1
            _poly_add(t2, tuple(e), -90.0 / 24.0 * sum(M[4][a, a, c, d] for a in range(3)))
        for e, v in _poly_mul_r2(t2).items():
            _poly_add(p, e, v)
        t0 = sum(M[4][a, a, c, c] for a in range(3) for c in range(3))
        for e, v in _poly_mul_r2({(0, 0, 0): 9.0 / 24.0 * t0}, 2).items():
            _poly_add(p, e, v)
        polys.append(p)
    table = monomial_exponents(order)
    out = np.zeros(len(table))
    for i, e in enumerate(table):
        d = sum(e)
        out[i] = polys[d].get(e, 0.0)
    return out


def quadrupole_from_flat(flat) -> np.ndarray:
    M2 = moment_tensors(flat, 2)[2]
    return 3.0 * M2 - np.trace(M2) * np.eye(3)


# ------------------------------------------------------------------ per-grid setup


class PoissonGridData:
    """Geometry-only data of the open-boundary solve, built once per grid."""

    def interior_mask(self, dtype):
        """1 on interior nodes, 0 on the faces, in the given dtype (flat)."""
        m = self._masks.get(dtype)
        if m is None:
            m = self.interior.to(dtype)
            self._masks[dtype] = m
        return m

    def __init__(self, dg: DeviceGrid):
        self.dg = dg
        nx, ny, nz = dg.shape
        dev = dg.device
        interior = np.zeros(dg.shape, dtype=bool)
        interior[1:-1, 1:-1, 1:-1] = True
        self.interior = torch.from_numpy(interior.reshape(-1)).to(dev)
        self._masks = {}
        self.n_interior = int(interior.sum())
        self.dinv = None
        self.cells = None
        self.frac = None
        self.ncells = 0
        self.subvol = 0.0
        if dg.kind == "legacy":
            # Jacobi preconditioner: inverse diagonal of -L (open_boundary.py:990-1014)
            lc = dg.get("lcoef").cpu().numpy()
            ax = lc[0:nx] + lc[nx:2 * nx]
            ay = lc[2 * nx:2 * nx + ny] + lc[2 * nx + ny:2 * nx + 2 * ny]
            az = lc[2 * nx + 2 * ny:2 * nx + 2 * ny + nz] + lc[2 * nx + 2 * ny + nz:]
            diag = ax[:, None, None] + ay[None, :, None] + az[None, None, :]
            dinv = np.where(interior, 1.0 / diag, 0.0)
            self.dinv = torch.from_numpy(dinv.reshape(-1)).to(dev)
        else:
            J = dg.get("jac").cpu().numpy().reshape(dg.shape)
            cj = (J[:-1, :-1, :-1] + J[1:, :-1, :-1] + J[:-1, 1:, :-1] + J[1:, 1:, :-1]
                  + J[:-1, :-1, 1:] + J[1:, :-1, 1:] + J[:-1, 1:, 1:] + J[1:, 1:, 1:]) / 8.0
            inner = np.zeros(cj.shape, dtype=bool)
            inner[1:-1, 1:-1, 1:-1] = True
            if inner.any():
                thr = float(np.quantile(cj[inner], _QUANTILE))
                mask = inner & (cj >= thr)
            else:
                mask = inner
            cells = np.argwhere(mask).astype(np.int32)
            self.ncells = int(cells.shape[0])
            if self.ncells:
                self.cells = torch.from_numpy(np.ascontiguousarray(cells)).to(dev)
                m = mask.astype(np.float64)
                sel = np.zeros(dg.shape)
                tot = np.zeros(dg.shape)
                one = np.ones_like(m)
                for ox, oy, oz in itertools.product((0, 1), repeat=3):
                    sel[ox:ox + m.shape[0], oy:oy + m.shape[1], oz:oz + m.shape[2]] += m
                    tot[ox:ox + m.shape[0], oy:oy + m.shape[1], oz:oz + m.shape[2]] += one
                frac = np.divide(sel, tot, out=np.zeros_like(sel), where=tot > 0.0)
                self.frac = torch.from_numpy(frac.reshape(-1)).to(dev)
            hx, hy, hz = dg.h
            self.subvol = hx * hy * hz / 8.0


def poisson_data(dg: DeviceGrid) -> PoissonGridData:
    pd = getattr(dg, "_poisson", None)
    if pd is None:
        pd = PoissonGridData(dg)
        dg._poisson = pd
    return pd


# ------------------------------------------------------------------ multipole boundary


def _center_arg(center):
    import ctypes

    arr = (ctypes.c_double * 3)(*[float(c) for c in center])
    return ctypes.cast(arr, ctypes.c_void_p), arr


def _cell_fit(dg, pdata, rho, stream=None):
    """Per selected cell: the quadratic least-squares fit over its 4x4x4 stencil."""
    nx, ny, nz = dg.shape
    coef = workspace(dg.device, 14 * pdata.ncells, "cellfit")
    _lib.call("ig_cell_fit_f64", _lib.ptr(rho), _lib.ptr(dg.xyz()), nx, ny, nz, _lib.ptr(pdata.cells),
              pdata.ncells, _lib.ptr(coef), _lib.stream_handle(stream))
    return coef


def _corrected_delta(dg, pdata, rho, order, stream=None):
    """Quadratic sub-cell moments minus nodal-region moments (open_boundary.py:370-464)."""
    if pdata.ncells == 0:
        return torch.zeros(nmono(order), dtype=torch.float64, device=dg.device)
    lib = _lib.load()
    nx, ny, nz = dg.shape
    coef = _cell_fit(dg, pdata, rho, stream)
    work = workspace(dg.device, lib.ig_physics_work_size(dg.N), "phys")
    quad = torch.zeros(nmono(order), dtype=torch.float64, device=dg.device)
    cptr, keep = _center_arg(dg.reference_center)
    _lib.call("ig_cell_moments_f64", _lib.ptr(dg.xyz()), _lib.ptr(dg.get("jac")), nx, ny, nz,
              _lib.ptr(pdata.cells), pdata.ncells, _lib.ptr(coef), pdata.subvol, cptr, order,
              _lib.ptr(work), _lib.ptr(quad), _lib.stream_handle(stream))
    nodal = _moments_with_frac(dg, rho, order, pdata.frac, stream)
    del keep
    return quad - nodal


def _split_correction(dg, pdata, rho, order, stream=None):
    """Boundary-value correction of the `legacy_split` mode (open_boundary.py:559-719):
    d = (rho_quadratic - rho_trilinear) J subvol at the sub-cell samples, expanded to the
    same order about the |d|-weighted sample centre and evaluated on the faces.
    Returns (field or None, centre or None)."""
    if pdata.ncells == 0:
        return None, None
    lib = _lib.load()
    nx, ny, nz = dg.shape
    coef = _cell_fit(dg, pdata, rho, stream)
    work = workspace(dg.device, lib.ig_physics_work_size(dg.N), "phys")
    h = _lib.stream_handle(stream)
    acc = torch.zeros(4, dtype=torch.float64, device=dg.device)
    _lib.call("ig_cell_split_moments_f64", _lib.ptr(rho), _lib.ptr(dg.xyz()), _lib.ptr(dg.get("jac")), nx, ny, nz,
              _lib.ptr(pdata.cells), pdata.ncells, _lib.ptr(coef), pdata.subvol, None, order, _lib.ptr(work),
              _lib.ptr(acc), h)
    a = acc.cpu().numpy()
    if not a[0] > 0.0:
        return None, None
    center = a[1:] / a[0]
    cptr, keep = _center_arg(center)
    mom = torch.zeros(nmono(order), dtype=torch.float64, device=dg.device)
    _lib.call("ig_cell_split_moments_f64", _lib.ptr(rho), _lib.ptr(dg.xyz()), _lib.ptr(dg.get("jac")), nx, ny, nz,
              _lib.ptr(pdata.cells), pdata.ncells, _lib.ptr(coef), pdata.subvol, cptr, order, _lib.ptr(work),
              _lib.ptr(mom), h)
    del keep
    return boundary_field(dg, mom.cpu().numpy(), order, center, stream), center


@_lib.on_stream
def boundary_field(dg, flat, order, center=None, stream=None):
    """Multipole potential of the flat raw moments `flat` (about `center`, default the grid
    reference centre) on the six faces, zero inside (open_boundary.py:467-495)."""
    nx, ny, nz = dg.shape
    qcoef = torch.from_numpy(boundary_coefficients(np.asarray(flat, dtype=np.float64), order)).to(dg.device)
    field = torch.empty(dg.N, dtype=torch.float64, device=dg.device)
    cptr, keep = _center_arg(dg.reference_center if center is None else center)
    _lib.call("ig_boundary_values_f64", _lib.ptr(dg.xyz()), nx, ny, nz, cptr, order, _lib.ptr(qcoef),
              _lib.ptr(field), _lib.stream_handle(stream))
    del keep
    return field


def _moments_with_frac(dg, rho, order, frac, stream=None):
    lib = _lib.load()
    work = workspace(dg.device, lib.ig_physics_work_size(dg.N), "phys")
    out = torch.zeros(nmono(order), dtype=torch.float64, device=dg.device)
    cptr, keep = _center_arg(dg.reference_center)
    _lib.call("ig_multipole_moments_f64", _lib.ptr(rho), _lib.ptr(dg.weights()), _lib.ptr(frac),
              _lib.ptr(dg.xyz()), dg.N, cptr, order, _lib.ptr(work), _lib.ptr(out),
              _lib.stream_handle(stream))
    del keep
    return out


def _device_rho(dg, rho):
    """rho as a flat contiguous FP64 CUDA field of the grid, or ValueError."""
    if not isinstance(rho, torch.Tensor) or not rho.is_cuda or rho.dtype != torch.float64 or rho.numel() != dg.N:
        raise ValueError(f"rho must be a float64 CUDA field of {dg.N} points.")
    return rho.reshape(-1).contiguous()


@_lib.on_stream
def multipole_boundary_device(dg, rho, order=2, mode="corrected_moments", stream=None):
    """Boundary field b (zero inside) and the flat raw moments used, on the device.

    rho: FP64 CUDA tensor of the grid shape (already validated and clamped).
    Returns (b, moments_flat (device), baseline_flat (device)).
    """
    if order not in (0, 1, 2, 3, 4):
        raise ValueError(f"multipole_order must be 0-4; received {order}.")
    if mode not in _MODES:
        raise ValueError(f"monitor_boundary_construction_mode must be one of {_MODES}; received `{mode}`.")
    rho = _device_rho(dg, rho)
    base = _moments_with_frac(dg, rho, order, None, stream)
    moments = base
    split = mode == "legacy_split" and dg.kind == "monitor"
    if dg.kind == "monitor" and mode != "nodal":
        moments = base + _corrected_delta(dg, poisson_data(dg), rho, order, stream)
    # legacy_split (open_boundary.py:923-934): faces from the nodal moments plus the split
    # correction; the reported moments stay the corrected ones.
    b = boundary_field(dg, (base if split else moments).cpu().numpy(), order, None, stream)
    if split:
        corr, _ = _split_correction(dg, poisson_data(dg), rho, order, stream)
        if corr is not None:
            b += corr
    return b, moments, base


# ------------------------------------------------------------------ Krylov


class _KrylovSpace:
    def __init__(self, n, dtype, device):
        self.n = n
        self.dtype = dtype
        z = lambda: torch.zeros(n, dtype=dtype, device=device)  # noqa: E731
        self.r, self.rhat, self.p, self.v = z(), z(), z(), z()
        self.s, self.t, self.phat, self.shat = z(), z(), z(), z()
        self.sc = torch.zeros(16, dtype=torch.float64, device=device)
        lib = _lib.load()
        self.work = torch.zeros(max(lib.ig_reduction_work_size(n), 64), dtype=torch.float64, device=device)


_SPACES: dict = {}


def _space(n, dtype, device):
    key = (n, dtype, str(device))
    sp = _SPACES.get(key)
    if sp is None:
        sp = _KrylovSpace(n, dtype, device)
        _SPACES[key] = sp
    return sp


def _matvec(dg, x, out, stream=None, transpose=False):
    """out = -L x on interior nodes, 0 on the faces (production closure on monitor grids);
    with transpose, the transposed interior operator S (-L^T) E for x with zero faces."""
    if not transpose:
        apply_operator(dg, x.view(dg.shape), out.view(dg.shape), scale=-1.0, closure="production",
                       interior_only=True, stream=stream)
        return
    apply_operator(dg, x.view(dg.shape), out.view(dg.shape), scale=-1.0, closure="production", transpose=True,
                   stream=stream)
    out.mul_(poisson_data(dg).interior_mask(out.dtype))


@_lib.on_stream
def bicgstab(dg, b, x, *, dinv=None, rtol=1e-8, maxiter=400, check_every=8, stream=None, transpose=False,
             mg=None):
    """Solve (-L) x = b on interior nodes (transpose: (-L)^T x = b); x (flat CUDA tensor)
    is overwritten.

    b and x share a dtype (float32 / float64); Krylov scalars stay FP64 on the device.
    Preconditioner (right): the diagonal ``dinv``, or one multigrid V-cycle ``mg``
    (``poisson_multigrid`` of the same operator, forward or transposed): the fused p / s updates then
    write its first damped-Jacobi sweep and ``mg.finish`` completes the cycle in place.
    Returns (iterations, converged, relative residual ||r|| / ||b||).
    """
    if mg is not None and (dinv is not None or getattr(mg, "transpose", False) != bool(transpose)):
        raise ValueError("the multigrid preconditioner must be built for the operator solved "
                         "(forward or transposed) and replaces dinv")
    dt = b.dtype
    n = b.numel()
    sp = _space(n, dt, b.device)
    sfx = "f64" if dt == torch.float64 else "f32"
    h = _lib.stream_handle(stream)
    x.zero_()
    sp.r.copy_(b)
    sp.rhat.copy_(b)
    sp.p.zero_()
    sp.v.zero_()
    sc = sp.sc
    sc.zero_()
    dots_into(sc, [(b, b)], None, slots=(0, 0, 0), stream=stream)
    bb = float(sc[0].item())
    if not (bb > 0.0):
        return 0, True, 0.0
    bnorm = math.sqrt(bb)
    sc[1] = 1.0
    sc[2] = 1.0
    sc[3] = 1.0
    sc[7] = bb
    if mg is not None:
        dinv = mg.scaled_inverse_diagonal(dt)
        check_every = min(check_every, 2)  # a few tens of iterations: test often
    dv = _lib.ptr(dinv.to(dt) if dinv is not None and dinv.dtype != dt else dinv)
    it = 0
    rel = 1.0
    converged = False
    while it < maxiter:
        it += 1
        _lib.call(f"ig_bicgstab_p_{sfx}", n, _lib.ptr(sp.r), _lib.ptr(sp.p), _lib.ptr(sp.v), dv,
                  _lib.ptr(sp.phat), _lib.ptr(sc), h)
        if mg is not None:
            mg.finish(sp.p, sp.phat, stream)
        _matvec(dg, sp.phat, sp.v, stream, transpose)
        dots_into(sc, [(sp.rhat, sp.v)], None, slots=(4, 0, 0), stream=stream)
        _lib.call(f"ig_bicgstab_s_{sfx}", n, _lib.ptr(sp.r), _lib.ptr(sp.v), dv, _lib.ptr(sp.s),
                  _lib.ptr(sp.shat), _lib.ptr(sc), h)
        if mg is not None:
            mg.finish(sp.s, sp.shat, stream)
        _matvec(dg, sp.shat, sp.t, stream, transpose)
        dots_into(sc, [(sp.t, sp.s), (sp.t, sp.t)], None, slots=(5, 6, 0), stream=stream)
        _lib.call(f"ig_bicgstab_x_{sfx}", n, _lib.ptr(x), _lib.ptr(sp.phat), _lib.ptr(sp.shat), _lib.ptr(sp.s),
                  _lib.ptr(sp.t), _lib.ptr(sp.r), _lib.ptr(sp.rhat), _lib.ptr(sc), _lib.ptr(sp.work), h)
        if it % check_every == 0 or it == maxiter:
            host = sc[:9].cpu().numpy()
            rel = math.sqrt(max(host[7], 0.0)) / bnorm
            if rel <= rtol:
                converged = True
                break
            if host[8] != 0.0:
                break
    return it, converged, rel


def _jacobi(dg, b, x, dinv_diag, tol, maxiter, stream=None):
    """Damped Jacobi on interior nodes (the reference fallback, open_boundary.py:1080-1105,
    :1279-1306); stops on the max-abs residual as the reference."""
    r = torch.empty_like(b)
    x.zero_()
    it = 0
    res = float("inf")
    for it in range(1, maxiter + 1):
        _matvec(dg, x, r, stream)
        r.neg_().add_(b)
        x.add_(r * dinv_diag, alpha=_JACOBI_DAMPING)
        if it % _JACOBI_CHECK_INTERVAL == 0 or it == maxiter:
            _matvec(dg, x, r, stream)
            res = float((r - b).abs().max().item())
            if res < tol:
                return it, True, res
    return it, False, res


def _diag_inverse(dg, pdata):
    """Inverse diagonal of -L on interior nodes (zero on faces)."""
    if pdata.dinv is not None:
        return pdata.dinv
    if getattr(pdata, "dinv_monitor", None) is not None:
        return pdata.dinv_monitor
    # monitor: the diagonal of -L is read off by applying the operator to unit impulses
    # at stride 5 in every axis (the 19-point stencil does not reach 5 nodes away)
    dev = dg.device
    diag = torch.zeros(dg.N, dtype=torch.float64, device=dev)
    e = torch.zeros(dg.shape, dtype=torch.float64, device=dev)
    out = torch.empty(dg.N, dtype=torch.float64, device=dev)
    for a, b_, c in itertools.product(range(5), repeat=3):
        e.zero_()
        e[a::5, b_::5, c::5] = 1.0
        _matvec(dg, e.view(-1), out, None)
        diag += out * e.view(-1)
    inv = torch.where(pdata.interior, 1.0 / torch.where(diag != 0, diag, torch.ones_like(diag)),
                      torch.zeros_like(diag))
    pdata.dinv_monitor = inv
    return inv


def poisson_multigrid(dg, transpose=False):
    """The multigrid preconditioner of this grid's interior operator A (transpose: of A^T,
    probed through the transposed stencil; the same parity structure), built on first use,
    then cached with the grid's Poisson data; monitor grids only."""
    pdata = poisson_data(dg)
    key = "multigrid_t" if transpose else "multigrid"
    mg = getattr(pdata, key, None)
    if mg is None:
        if dg.kind != "monitor":
            raise ValueError("the multigrid preconditioner is built for monitor grids")
        mg = multigrid.PoissonMultigrid(dg.shape, lambda x, out: _matvec(dg, x, out, transpose=transpose),
                                        dg.device)
        mg.transpose = bool(transpose)
        setattr(pdata, key, mg)
    return mg


@_lib.on_stream
def solve_poisson_device(dg, rho, *, multipole_order=2, tolerance=1e-8, max_iterations=400, solver="auto",
                         mode="corrected_moments", precision="fp64", check_every=8, inner_tolerance=None,
                         stream=None, preconditioner="auto"):
    """Device-resident open-boundary solve.  rho: FP64 CUDA tensor (validated, >= 0).

    preconditioner: "auto" (multigrid V-cycle on monitor grids, the inverse diagonal on
    legacy grids), "multigrid" (monitor grids), "jacobi" (the inverse diagonal) or "none".
    Returns a dict: potential (FP64, grid shape), boundary (b), moments (flat, host),
    baseline (flat, host), iterations, converged, residual (relative 2-norm), method.
    """
    if preconditioner not in ("auto", "multigrid", "jacobi", "none"):
        raise ValueError("preconditioner must be `auto`, `multigrid`, `jacobi` or `none`; "
                         f"received `{preconditioner}`.")
    if preconditioner == "multigrid" and dg.kind != "monitor":
        raise ValueError("the multigrid preconditioner is built for monitor grids")
    if solver not in _SOLVERS:
        raise ValueError(
            "The open-boundary Poisson solve supports solver=`auto`, `scipy_bicgstab`, `bicgstab` "
            f"or `jacobi`; received `{solver}`.")
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be `fp64` or `mixed`; received `{precision}`.")
    pdata = poisson_data(dg)
    rho = _device_rho(dg, rho)
    b, moments, base = multipole_boundary_device(dg, rho, multipole_order, mode, stream)
    # rhs = 4 pi rho + L b on interior nodes (L b = -(-L b))
    rhs = torch.empty_like(rho)
    _matvec(dg, b, rhs, stream)
    rhs.neg_().add_(rho * pdata.interior, alpha=FOUR_PI)
    u = torch.zeros_like(rho)
    mg = None
    dinv = None
    pre = ("multigrid" if dg.kind == "monitor" else "jacobi") if preconditioner == "auto" else preconditioner
    if solver != "jacobi" and pre == "multigrid":
        mg = poisson_multigrid(dg)
    elif pre == "jacobi":
        dinv = pdata.dinv if pdata.dinv is not None else _diag_inverse(dg, pdata)
    tag = "_mg" if mg is not None else ""
    if solver == "jacobi":
        it, conv, res = _jacobi(dg, rhs, u, _diag_inverse(dg, pdata), tolerance, max_iterations, stream)
        method = "hip_jacobi" if conv else "hip_jacobi_maxiter"
        rel = res
    elif precision == "fp64":
        it, conv, rel = bicgstab(dg, rhs, u, dinv=dinv, rtol=tolerance, maxiter=max_iterations,
                                 check_every=check_every, stream=stream, mg=mg)
        method = f"hip_bicgstab{tag}_fp64" + ("_monitor" if dg.kind == "monitor" else "")
    else:
        # FP32 BiCGSTAB corrections of the FP64 residual (defect correction).  Each inner
        # solve restarts its Krylov space, so it is asked for the whole remaining
        # reduction (at most what FP32 attains, MIXED_INNER_FLOOR): the first one goes to
        # ~1e-6, the next only closes the gap to the tolerance.
        r64 = rhs.clone()
        tmp = torch.empty_like(rhs)
        d32 = torch.zeros(rho.numel(), dtype=torch.float32, device=rho.device)
        rn = float(torch.linalg.vector_norm(rhs).item())
        it = 0
        conv = rn == 0.0
        rel = 0.0
        dinv32 = dinv.float() if dinv is not None else None
        outer = 0
        while not conv and it < max_iterations:
            outer += 1
            scale = float(torch.linalg.vector_norm(r64).item())
            r32 = (r64 / scale).float()
            inner = inner_tolerance if inner_tolerance is not None else \
                max(0.5 * tolerance * rn / scale, MIXED_INNER_FLOOR)
            k, _, _ = bicgstab(dg, r32, d32, dinv=dinv32, rtol=inner, maxiter=max_iterations - it,
                               check_every=check_every, stream=stream, mg=mg)
            it += k
            u.add_(d32.double(), alpha=scale)
            _matvec(dg, u, tmp, stream)
            torch.sub(rhs, tmp, out=r64)
            prev = rel if outer > 1 else math.inf
            rel = float(torch.linalg.vector_norm(r64).item()) / rn
            conv = rel <= tolerance
            # a correction that does not reduce the FP64 residual (an inner breakdown or
            # stagnation at the FP32 floor) ends the loop; so does a non-finite residual
            if outer > 64 or not math.isfinite(rel) or (not conv and rel >= prev):
                break
        method = f"hip_bicgstab{tag}_mixed_fp32_fp64" + ("_monitor" if dg.kind == "monitor" else "")
    potential = b + u
    # interior residual max |(-L) u - rhs|
    res = torch.empty_like(rho)
    _matvec(dg, u, res, stream)
    res.sub_(rhs)
    residual_max = float((res * pdata.interior).abs().max().item())
    return {
        "potential": potential.view(dg.shape),
        "boundary": b.view(dg.shape),
        "rhs": rhs.view(dg.shape),
        "moments": moments.cpu().numpy(),
        "baseline": base.cpu().numpy(),
        "iterations": int(it),
        "converged": bool(conv),
        "residual": float(rel),
        "residual_max": residual_max,
        "method": method,
    }


# ------------------------------------------------------------------ adjoint


@_lib.on_stream
def moments_adjoint_device(dg, q, order, mode="corrected_moments", out=None, stream=None):
    """sum_j q_j d moment_j / d rho (flat FP64 field) for the flat raw moments that
    multipole_boundary_device reports (nodal; on monitor grids plus the corrected-moment
    delta, whose sub-cell quadratic-fit part is scattered by ig_cell_moments_adjoint_f64)."""
    q = np.ascontiguousarray(np.asarray(q, dtype=np.float64))
    if q.size != nmono(order):
        raise ValueError(f"q must hold {nmono(order)} moment weights for order {order}.")
    pdata = poisson_data(dg)
    if out is None:
        out = torch.empty(dg.N, dtype=torch.float64, device=dg.device)
    import ctypes

    qp = q.ctypes.data_as(ctypes.c_void_p)
    cptr, keep = _center_arg(dg.reference_center)
    corrected = dg.kind == "monitor" and pdata.ncells > 0 and mode != "nodal"
    _lib.call("ig_moments_adjoint_f64", _lib.ptr(dg.weights()), _lib.ptr(pdata.frac if corrected else None),
              _lib.ptr(dg.xyz()), dg.N, cptr, order, qp, _lib.ptr(out), 0, _lib.stream_handle(stream))
    if corrected:
        nx, ny, nz = dg.shape
        _lib.call("ig_cell_moments_adjoint_f64", _lib.ptr(dg.xyz()), _lib.ptr(dg.get("jac")), nx, ny, nz,
                  _lib.ptr(pdata.cells), pdata.ncells, pdata.subvol, cptr, order, qp, _lib.ptr(out),
                  _lib.stream_handle(stream))
    del keep
    return out


@_lib.on_stream
def poisson_adjoint_device(dg, g, multipole_order=2, mode="corrected_moments", tolerance=1e-10,
                           max_iterations=4000, stream=None, preconditioner="auto"):
    """P^T g for the open-boundary Poisson map v = P rho of solve_poisson_device: the
    gradient of <g, v(rho)> with respect to rho.

    With b = B rho the multipole faces, S / E interior restriction / extension and
    A = S (-L) E, the forward map is v = B rho + E A^-1 S (4 pi rho + L B rho), so
        P^T g = 4 pi E y + B^T (g + L^T E y),   A^T y = S g,
    solved by BiCGSTAB on the transposed stencil (no stored forward iterates; preconditioner
    as in solve_poisson_device, the multigrid one built for A^T), and
    B^T h = M^T F^T h with F the faces of each unit moment and M^T the moment adjoint.
    The legacy_split construction is not linear in rho (its correction centre is
    |d|-weighted) and has no adjoint here.  Returns (gradient, info dict)."""
    if dg.kind == "monitor" and mode == "legacy_split":
        raise NotImplementedError("the legacy_split boundary is not linear in rho; its adjoint is not built.")
    if mode not in _MODES:
        raise ValueError(f"monitor_boundary_construction_mode must be one of {_MODES}; received `{mode}`.")
    if preconditioner not in ("auto", "multigrid", "jacobi", "none"):
        raise ValueError("preconditioner must be `auto`, `multigrid`, `jacobi` or `none`; "
                         f"received `{preconditioner}`.")
    if preconditioner == "multigrid" and dg.kind != "monitor":
        raise ValueError("the multigrid preconditioner is built for monitor grids")
    pdata = poisson_data(dg)
    g = g.reshape(-1).to(torch.float64).contiguous()
    b = g * pdata.interior_mask(torch.float64)
    y = torch.empty_like(b)
    pre = ("multigrid" if dg.kind == "monitor" else "jacobi") if preconditioner == "auto" else preconditioner
    mg = poisson_multigrid(dg, transpose=True) if pre == "multigrid" else None
    dinv = (pdata.dinv if pdata.dinv is not None else _diag_inverse(dg, pdata)) if pre == "jacobi" else None
    it, conv, rel = bicgstab(dg, b, y, dinv=dinv, rtol=tolerance, maxiter=max_iterations, stream=stream,
                             transpose=True, mg=mg)
    lty = apply_operator(dg, y.view(dg.shape), scale=1.0, closure="production", transpose=True,
                         stream=stream).reshape(-1)
    h = g + lty
    from .reductions import wdot

    nm = nmono(multipole_order)
    q = np.empty(nm)
    for j in range(nm):
        e = np.zeros(nm)
        e[j] = 1.0
        q[j] = wdot(dg, boundary_field(dg, e, multipole_order, None, stream), h, weighted=False)
    grad = moments_adjoint_device(dg, q, multipole_order, mode, stream=stream)
    grad.add_(y, alpha=FOUR_PI)
    return grad, {"iterations": int(it), "converged": bool(conv), "residual": float(rel)}


# ------------------------------------------------------------------ drop-in call surfaces


def validate_density_field(rho, grid_geometry, name="rho"):
    """poisson/hartree.py:32-44: finite, >= -1e-14, clamped at zero (NumPy in/out)."""
    from .grid import grid_shape

    values = np.asarray(rho, dtype=np.float64)
    shape = grid_shape(grid_geometry)
    if values.shape != shape:
        raise ValueError(f"{name} must have shape {shape}; received {values.shape}.")
    if not np.all(np.isfinite(values)):
        raise ValueError(f"{name} must contain only finite values.")
    if np.any(values < -_NEGATIVE_DENSITY_TOLERANCE):
        raise ValueError(f"{name} must be non-negative up to roundoff.")
    return np.maximum(values, 0.0)


def _to_device_density(rho, dg, name="rho"):
    if isinstance(rho, torch.Tensor):
        t = rho.to(device=dg.device, dtype=torch.float64)
        if tuple(t.shape) != dg.shape:
            raise ValueError(f"{name} must have shape {dg.shape}; received {tuple(t.shape)}.")
        bad = (~torch.isfinite(t)).any() | (t < -_NEGATIVE_DENSITY_TOLERANCE).any()
        if bool(bad.item()):
            raise ValueError(f"{name} must be finite and non-negative up to roundoff.")
        return t.clamp_min(0.0).contiguous()
    return torch.from_numpy(np.ascontiguousarray(rho)).to(dg.device)


def solve_open_boundary_poisson(grid_geometry, rho, multipole_order=2, tolerance=1.0e-8, max_iterations=400,
                                solver="auto", monitor_boundary_construction_mode="corrected_moments", *,
                                precision="fp64", device="cuda"):
    """poisson/open_boundary.py:1379-1413 on the GPU."""
    if solver not in ("auto", "scipy_bicgstab", "jacobi", "bicgstab"):
        raise ValueError(
            "The current open-boundary Poisson slice supports solver=`auto`, `scipy_bicgstab`, "
            f"or `jacobi`; received `{solver}`.")
    dg = device_grid(grid_geometry, device)
    density = rho if isinstance(rho, torch.Tensor) else validate_density_field(rho, grid_geometry)
    rt = _to_device_density(density, dg)
    sol = solve_poisson_device(dg, rt, multipole_order=multipole_order, tolerance=tolerance,
                               max_iterations=max_iterations, solver=solver,
                               mode=monitor_boundary_construction_mode, precision=precision)
    flat = sol["moments"]
    base = sol["baseline"]
    order = multipole_order
    dip = flat[1:4].copy() if order >= 1 else np.zeros(3)
    quad = quadrupole_from_flat(flat) if order >= 2 else np.zeros((3, 3))
    diag = None
    if dg.kind == "monitor":
        bdip = base[1:4].copy() if order >= 1 else np.zeros(3)
        bquad = quadrupole_from_flat(base) if order >= 2 else np.zeros((3, 3))
        # open_boundary.py:498-556: face values against the nodal-moment and the
        # corrected-moment expansions.
        faces = ~poisson_data(dg).interior.reshape(-1)
        actual = sol["boundary"].reshape(-1)[faces]
        corr = actual - boundary_field(dg, base, order).reshape(-1)[faces]
        mis = actual - boundary_field(dg, flat, order).reshape(-1)[faces]
        rms = lambda t: float(torch.sqrt(torch.mean(t * t)).item()) if t.numel() else 0.0  # noqa: E731
        amax = lambda t: float(t.abs().max().item()) if t.numel() else 0.0  # noqa: E731
        diag = OpenBoundaryBoundaryConstructionDiagnostics(
            baseline_total_charge=float(base[0]), baseline_dipole_moment=bdip,
            baseline_quadrupole_tensor=bquad, correction_total_charge=float(flat[0] - base[0]),
            correction_dipole_moment=dip - bdip, correction_quadrupole_tensor=quad - bquad,
            boundary_value_correction_rms=rms(corr), boundary_value_correction_max_abs=amax(corr),
            corrected_moment_boundary_rms_mismatch=rms(mis), corrected_moment_boundary_max_abs_mismatch=amax(mis))
    bvals = sol["boundary"].cpu().numpy()
    bc = OpenBoundaryMultipoleBoundary(
        reference_center=tuple(float(c) for c in dg.reference_center), multipole_order=order,
        total_charge=float(flat[0]), dipole_moment=dip, quadrupole_tensor=quad, boundary_values=bvals,
        description=(f"Free-space multipole Dirichlet data of order {order} about the grid reference "
                     f"centre (HIP kernels; mode {monitor_boundary_construction_mode})."),
        diagnostics=diag, raw_moments=tuple(float(v) for v in flat))
    return OpenBoundaryPoissonResult(
        rho=np.asarray(rt.cpu().numpy()), potential=sol["potential"].cpu().numpy(), boundary_condition=bc,
        solver_method=sol["method"], solver_iterations=sol["iterations"], residual_max=sol["residual_max"],
        description="Open-boundary Poisson solve on the GPU: v = u + b, -L u = 4 pi rho + L b inside.",
        converged=sol["converged"], residual_relative_l2=sol["residual"])


def solve_hartree_potential(grid_geometry, rho, multipole_order=2, tolerance=1.0e-8, max_iterations=400,
                            solver="auto", backend="python", use_jax_cached_operator=False, cg_impl="baseline",
                            cg_preconditioner="none", line_preconditioner_impl="baseline",
                            monitor_boundary_construction_mode="corrected_moments", *, precision="fp64"):
    """poisson/hartree.py:47-98.  Every backend value runs the HIP path; the reference's
    validation of `backend` is kept."""
    nb = str(backend).strip().lower()
    if nb not in ("python", "jax", "hip"):
        raise ValueError(f"backend must be `python`, `jax` or `hip`; received `{backend}`.")
    density = validate_density_field(rho, grid_geometry) if not isinstance(rho, torch.Tensor) else rho
    return solve_open_boundary_poisson(grid_geometry, density, multipole_order=multipole_order,
                                       tolerance=tolerance, max_iterations=max_iterations,
                                       solver="auto" if solver == "auto" else solver,
                                       monitor_boundary_construction_mode=monitor_boundary_construction_mode,
                                       precision=precision)


def _potential_array(grid_geometry, hartree_potential):
    if isinstance(hartree_potential, OpenBoundaryPoissonResult):
        return np.asarray(hartree_potential.potential, dtype=np.float64)
    return np.asarray(hartree_potential, dtype=np.float64)


def evaluate_hartree_energy(rho, grid_geometry, hartree_potential) -> float:
    """E_H = 1/2 int rho v_H (poisson/hartree.py:117-129), FP64 reduction on the GPU."""
    from .reductions import wdot

    density = validate_density_field(rho, grid_geometry)
    pot = _potential_array(grid_geometry, hartree_potential)
    dg = device_grid(grid_geometry)
    a = torch.from_numpy(np.ascontiguousarray(density)).to(dg.device)
    b = torch.from_numpy(np.ascontiguousarray(pot)).to(dg.device)
    return 0.5 * wdot(dg, a, b)


def build_hartree_action(psi, grid_geometry, hartree_potential):
    """v_H * psi (poisson/hartree.py:132-144), formed on the device."""
    pot = _potential_array(grid_geometry, hartree_potential)
    dg = device_grid(grid_geometry)
    a = torch.from_numpy(np.ascontiguousarray(pot)).to(dg.device)
    b = torch.from_numpy(np.ascontiguousarray(np.asarray(psi, dtype=np.float64))).to(dg.device)
    return (a * b).cpu().numpy()

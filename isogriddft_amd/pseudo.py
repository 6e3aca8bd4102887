"""GTH pseudopotential terms on the GPU (drop-in for the reference's pseudo/local.py
and pseudo/nonlocal.py).

Local part: ig_vloc_gth_f64 evaluates, per grid point and atom,
    V_loc(r) = -Z erf(r / (sqrt2 r_loc)) / r + exp(-r^2 / 2 r_loc^2) sum_i C_i (r/r_loc)^(2i-2)
(pseudo/local.py:138-180; r = 0 limit -Z sqrt(2/pi) / r_loc).
Nonlocal part: separable projectors p_i^l(r) Y_lm on a per-atom support box
(pseudo/nonlocal.py:99-321).  The reference evaluates every projector on the whole
grid; here each projector lives on the index box of the nodes within 10 r_l of the
atom (outside it |p| < e^-50 of its peak), is evaluated once per (grid, atoms) and
kept in HBM, and the block apply is project -> h-matrix -> scatter on that box.
The near-core patch correction field (pseudo/local.py:379-703) is not built here;
callers pass it (or any extra local field) as an explicit potential.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import DeviceGrid, device_grid
from .reductions import workspace


@dataclass(frozen=True)
class GTHChannel:
    angular_momentum: int
    radius: float
    h_matrix: tuple

    @property
    def projector_count(self):
        return len(self.h_matrix)


@dataclass(frozen=True)
class GTHPseudoData:
    element: str
    ionic_charge: int
    rloc: float
    local_coefficients: tuple
    nonlocal_channels: tuple


# gth-pade parameters of H, C, N, O, the elements the reference ships
# (pseudo/gth_data.py:28-117).
GTH_PADE = {
    "H": GTHPseudoData("H", 1, 0.2, (-4.1802368, 0.72507482), ()),
    "C": GTHPseudoData("C", 4, 0.34883045, (-8.5137711, 1.22843203),
                       (GTHChannel(0, 0.30455321, ((9.52284179,),)), GTHChannel(1, 0.2326773, ()))),
    "N": GTHPseudoData("N", 5, 0.28917923, (-12.23481988, 1.76640728),
                       (GTHChannel(0, 0.25660487, ((13.55224272,),)), GTHChannel(1, 0.27013369, ()))),
    "O": GTHPseudoData("O", 6, 0.24762086, (-16.58031797, 2.39570092),
                       (GTHChannel(0, 0.22178614, ((18.26691718,),)), GTHChannel(1, 0.2568289, ()))),
}
SUPPORTED_GTH_ELEMENTS = tuple(GTH_PADE)
_SUPPORT_RADII = 10.0


def case_atoms(case):
    """[(element, (x, y, z))] of a benchmark case, a molecule or a list of atoms."""
    atoms = getattr(getattr(case, "geometry", None), "atoms", None)
    if atoms is None:
        atoms = getattr(case, "atoms", case)
    out = []
    for a in atoms:
        el = getattr(a, "element", None)
        pos = getattr(a, "position", None)
        if el is None:
            el, pos = a
        if el not in GTH_PADE:
            raise ValueError(f"No GTH-PADE data for element `{el}`; supported: {SUPPORTED_GTH_ELEMENTS}.")
        out.append((el, tuple(float(c) for c in pos)))
    return out


def _atom_rows(atoms):
    rows = np.zeros((len(atoms), 9))
    for n, (el, pos) in enumerate(atoms):
        d = GTH_PADE[el]
        rows[n, :3] = pos
        rows[n, 3] = d.ionic_charge
        rows[n, 4] = d.rloc
        for i, c in enumerate(d.local_coefficients[:4]):
            rows[n, 5 + i] = c
    return rows


def local_potential_device(dg: DeviceGrid, atoms, out=None, accumulate=False, stream=None):
    """FP64 V_loc summed over atoms on the device (flat, length N)."""
    if out is None:
        out = torch.empty(dg.N, dtype=torch.float64, device=dg.device)
        accumulate = False
    if not atoms:
        if not accumulate:
            out.zero_()
        return out
    rows = torch.from_numpy(_atom_rows(atoms)).to(dg.device)
    _lib.call("ig_vloc_gth_f64", _lib.ptr(dg.xyz()), dg.N, len(atoms), _lib.ptr(rows), _lib.ptr(out),
              int(accumulate), _lib.stream_handle(stream))
    return out


def local_potential_gradient_device(dg: DeviceGrid, atoms, rho_w, stream=None):
    """dE/dR_A (natoms x 3, host) of E = sum rho w V_loc; rho_w = rho * w on the device."""
    if len(atoms) > 64:
        raise ValueError("at most 64 atoms per call")
    rows = torch.from_numpy(_atom_rows(atoms)).to(dg.device)
    lib = _lib.load()
    work = workspace(dg.device, lib.ig_physics_work_size(dg.N), "phys")
    out = torch.zeros(3 * len(atoms), dtype=torch.float64, device=dg.device)
    _lib.call("ig_vloc_gth_grad_f64", _lib.ptr(dg.xyz()), _lib.ptr(rho_w.reshape(-1)), dg.N, len(atoms),
              _lib.ptr(rows), _lib.ptr(work), _lib.ptr(out), _lib.stream_handle(stream))
    return out.cpu().numpy().reshape(len(atoms), 3)


# ------------------------------------------------------------------ local drop-ins
# Result types mirror pseudo/local.py:30-52.


@dataclass(frozen=True)
class AtomicLocalPotentialContribution:
    atom_index: int
    element: str
    position: tuple
    pseudo_data: GTHPseudoData
    local_potential: np.ndarray


@dataclass(frozen=True)
class LocalIonicPotentialEvaluation:
    pseudo_family: str
    grid_geometry: object
    atom_contributions: tuple
    total_local_potential: np.ndarray


def evaluate_local_ionic_potential(case, grid_geometry) -> LocalIonicPotentialEvaluation:
    """pseudo/local.py:261-357 (legacy and monitor grids alike)."""
    dg = device_grid(grid_geometry)
    atoms = case_atoms(case)
    contribs = []
    total = torch.zeros(dg.N, dtype=torch.float64, device=dg.device)
    for n, (el, pos) in enumerate(atoms):
        f = local_potential_device(dg, [(el, pos)])
        total += f
        contribs.append(AtomicLocalPotentialContribution(n, el, pos, GTH_PADE[el],
                                                         f.cpu().numpy().reshape(dg.shape)))
    return LocalIonicPotentialEvaluation("gth-pade", grid_geometry, tuple(contribs),
                                         total.cpu().numpy().reshape(dg.shape))


evaluate_monitor_grid_local_ionic_potential = evaluate_local_ionic_potential
evaluate_legacy_local_ionic_potential = evaluate_local_ionic_potential


def evaluate_local_ionic_energy(evaluation: LocalIonicPotentialEvaluation, density_field) -> float:
    """E_loc = int rho V_loc (pseudo/local.py:360-376), FP64 reduction on the GPU."""
    from .reductions import wdot

    dg = device_grid(evaluation.grid_geometry)
    rho = np.asarray(density_field, dtype=np.float64)
    if rho.shape != dg.shape:
        raise ValueError(f"density_field must have shape {dg.shape}; received {rho.shape}.")
    a = torch.from_numpy(np.ascontiguousarray(rho)).to(dg.device)
    b = torch.from_numpy(np.ascontiguousarray(evaluation.total_local_potential)).to(dg.device)
    return wdot(dg, a, b)


# ------------------------------------------------------------------ nonlocal projectors

_DESC = np.dtype([("l", "<i4"), ("m", "<i4"), ("i", "<i4"), ("pad", "<i4"), ("rl", "<f8"), ("pref", "<f8")])


def _magnetic_numbers(l):
    """pseudo/nonlocal.py:132-140: m = 0 for s; (-1, 0, 1) for p (real harmonics y, z, x)."""
    if l == 0:
        return (0,)
    if l == 1:
        return (-1, 0, 1)
    raise ValueError("only s and p projector channels are supported (pseudo/nonlocal.py:99-129)")


def _radial_prefactor(l, i, rl):
    """sqrt2 / (rl^(l + (4i+3)/2) sqrt(Gamma(l + (4i+3)/2))), zero-based i (nonlocal.py:143-155)."""
    g = l + 0.5 * (4 * (i + 1) - 1)
    return math.sqrt(2.0) / (rl ** g * math.sqrt(math.gamma(g)))


class _AtomProjectors:
    def __init__(self, dg, atom_index, element, pos, xyz_host):
        d = GTH_PADE[element]
        self.atom_index = atom_index
        self.element = element
        self.position = pos
        descs = []
        meta = []
        blocks = []
        for ch in d.nonlocal_channels:
            if ch.projector_count == 0:
                continue
            for m in _magnetic_numbers(ch.angular_momentum):
                for i in range(ch.projector_count):
                    descs.append((ch.angular_momentum, m, i, 0, ch.radius,
                                  _radial_prefactor(ch.angular_momentum, i, ch.radius)))
                    meta.append((ch.angular_momentum, m, i, ch.radius))
                blocks.append(np.asarray(ch.h_matrix, dtype=np.float64))
        self.nproj = len(descs)
        self.meta = tuple(meta)
        if self.nproj == 0:
            return
        H = np.zeros((self.nproj, self.nproj))
        o = 0
        for b in blocks:
            k = b.shape[0]
            H[o:o + k, o:o + k] = b
            o += k
        self.H_host = H
        rcut = _SUPPORT_RADII * max(m_[3] for m_ in meta)
        r2 = ((xyz_host[0] - pos[0]) ** 2 + (xyz_host[1] - pos[1]) ** 2 + (xyz_host[2] - pos[2]) ** 2)
        idx = np.argwhere(r2.reshape(dg.shape) <= rcut * rcut)
        if idx.size == 0:
            # the atom's support misses the grid: the nearest node carries the (tiny) field
            near = np.unravel_index(int(np.argmin(r2)), dg.shape)
            idx = np.array([near])
        lo = idx.min(axis=0)
        hi = idx.max(axis=0) + 1
        self.box = (ctypes.c_int * 6)(int(lo[0]), int(lo[1]), int(lo[2]),
                                      int(hi[0] - lo[0]), int(hi[1] - lo[1]), int(hi[2] - lo[2]))
        self.nbox = int(np.prod(hi - lo))
        self.pos_c = (ctypes.c_double * 3)(*pos)
        arr = np.array(descs, dtype=_DESC)
        if arr.itemsize != _lib.load().ig_projector_desc_size():
            raise RuntimeError("projector descriptor layout mismatch with the HIP library")
        self.descs = torch.from_numpy(arr.view(np.uint8).copy()).to(dg.device)
        self.H = torch.from_numpy(H).to(dg.device)
        self.P = torch.empty((self.nproj, self.nbox), dtype=torch.float64, device=dg.device)
        nx, ny, nz = dg.shape
        _lib.call("ig_projector_fields_f64", _lib.ptr(dg.xyz()), nx, ny, nz, self.box, self.pos_c, self.nproj,
                  _lib.ptr(self.descs), _lib.ptr(self.P), 0, None)

    def field_derivative(self, dg, axis):
        """d P / d R_axis on the box (the atom moves)."""
        out = torch.empty_like(self.P)
        nx, ny, nz = dg.shape
        _lib.call("ig_projector_fields_f64", _lib.ptr(dg.xyz()), nx, ny, nz, self.box, self.pos_c, self.nproj,
                  _lib.ptr(self.descs), _lib.ptr(out), axis + 1, None)
        return out


class NonlocalPlan:
    """Projectors of every atom of a case on one grid, resident in HBM."""

    def __init__(self, dg: DeviceGrid, case):
        self.dg = dg
        self.atoms = case_atoms(case)
        xyz = dg.xyz().cpu().numpy()
        self.per_atom = []
        for n, (el, pos) in enumerate(self.atoms):
            ap = _AtomProjectors(dg, n, el, pos, xyz)
            if ap.nproj:
                self.per_atom.append(ap)

    @property
    def empty(self):
        return not self.per_atom

    def project(self, psi, stream=None):
        """[(atom, coef (k, nproj), d = coef H^T (k, nproj))] for a block psi (k, *shape)."""
        dg = self.dg
        blk = psi.reshape(-1, dg.N)
        k = blk.shape[0]
        sfx = "f64" if blk.dtype == torch.float64 else "f32"
        nx, ny, nz = dg.shape
        lib = _lib.load()
        out = []
        for ap in self.per_atom:
            work = workspace(dg.device, lib.ig_nonlocal_work_size(k, ap.nproj), "nl")
            coef = torch.empty((k, ap.nproj), dtype=torch.float64, device=dg.device)
            d = torch.empty((k, ap.nproj), dtype=torch.float64, device=dg.device)
            _lib.call(f"ig_nonlocal_project_{sfx}", _lib.ptr(blk), blk.stride(0), k, nx, ny, nz, ap.box,
                      _lib.ptr(ap.P), ap.nproj, _lib.ptr(dg.weights()), _lib.ptr(ap.H), _lib.ptr(work),
                      _lib.ptr(coef), _lib.ptr(d), _lib.stream_handle(stream))
            out.append((ap, coef, d))
        return out

    def apply(self, psi, out, stream=None, projections=None):
        """out += V_nl psi for a block (out, psi of one dtype and shape).  Returns the projections."""
        dg = self.dg
        if projections is None:
            projections = self.project(psi, stream)
        ob = out.reshape(-1, dg.N)
        k = ob.shape[0]
        sfx = "f64" if ob.dtype == torch.float64 else "f32"
        nx, ny, nz = dg.shape
        for ap, _, d in projections:
            _lib.call(f"ig_nonlocal_apply_{sfx}", _lib.ptr(ob), ob.stride(0), k, nx, ny, nz, ap.box, _lib.ptr(ap.P),
                      ap.nproj, _lib.ptr(d), _lib.stream_handle(stream))
        return projections

    def energy_terms(self, psi, stream=None):
        """<psi_m | V_nl | psi_m>_W per orbital (host, FP64)."""
        dg = self.dg
        k = psi.reshape(-1, dg.N).shape[0]
        e = np.zeros(k)
        for _, coef, d in self.project(psi, stream):
            e += (coef * d).sum(dim=1).cpu().numpy()
        return e


def nonlocal_plan(dg: DeviceGrid, case) -> NonlocalPlan:
    key = tuple(case_atoms(case))
    cache = getattr(dg, "_nonlocal", None)
    if cache is None:
        cache = {}
        dg._nonlocal = cache
    plan = cache.get(key)
    if plan is None:
        plan = NonlocalPlan(dg, case)
        cache[key] = plan
    return plan


# Result types mirror pseudo/nonlocal.py:55-89.


@dataclass(frozen=True)
class ProjectorFieldEvaluation:
    atom_index: int
    element: str
    position: tuple
    angular_momentum: int
    magnetic_number: int
    projector_index: int
    radius: float
    projector_values: np.ndarray


@dataclass(frozen=True)
class AtomicNonlocalActionContribution:
    atom_index: int
    element: str
    position: tuple
    pseudo_data: GTHPseudoData
    projector_fields: tuple
    projection_coefficients: tuple
    resolved_coefficients: tuple
    nonlocal_action: np.ndarray


@dataclass(frozen=True)
class NonlocalIonicActionEvaluation:
    pseudo_family: str
    grid_geometry: object
    atom_contributions: tuple
    total_nonlocal_action: np.ndarray


def _box_to_full(dg, ap, values):
    full = np.zeros(dg.shape)
    b = list(ap.box)
    full[b[0]:b[0] + b[3], b[1]:b[1] + b[4], b[2]:b[2] + b[5]] = values.reshape(b[3], b[4], b[5])
    return full


def evaluate_nonlocal_ionic_action(case, grid_geometry, psi) -> NonlocalIonicActionEvaluation:
    """pseudo/nonlocal.py:286-321 for one orbital field (NumPy in, NumPy out)."""
    dg = device_grid(grid_geometry)
    field = np.asarray(psi, dtype=np.float64)
    if field.shape != dg.shape:
        raise ValueError(f"psi must have shape {dg.shape}; received {field.shape}.")
    plan = nonlocal_plan(dg, case)
    x = torch.from_numpy(np.ascontiguousarray(field)).to(dg.device).reshape(1, -1)
    total = torch.zeros_like(x)
    contribs = []
    projections = plan.apply(x, total)
    for ap, coef, d in projections:
        one = torch.zeros_like(x)
        plan.apply(x, one, projections=[(ap, coef, d)])
        fields = tuple(
            ProjectorFieldEvaluation(ap.atom_index, ap.element, ap.position, l, m, i, rl,
                                     _box_to_full(dg, ap, ap.P[p].cpu().numpy()))
            for p, (l, m, i, rl) in enumerate(ap.meta))
        contribs.append(AtomicNonlocalActionContribution(
            ap.atom_index, ap.element, ap.position, GTH_PADE[ap.element], fields,
            tuple(float(v) for v in coef[0].cpu().numpy()), tuple(float(v) for v in d[0].cpu().numpy()),
            one.cpu().numpy().reshape(dg.shape)))
    return NonlocalIonicActionEvaluation("gth-pade", grid_geometry, tuple(contribs),
                                         total.cpu().numpy().reshape(dg.shape))

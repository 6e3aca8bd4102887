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
Near-core patch (pseudo/local.py:379-703): per atom a fine Cartesian lattice inside the
patch sphere samples the density by inverse-distance-squared interpolation from its k
nearest main-grid nodes (found once per grid on the GPU, ig_knn_f64); the patch
quadrature minus the main-grid quadrature is embedded as a frozen correction field.
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


@_lib.on_stream
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


@_lib.on_stream
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


# ------------------------------------------------------------------ near-core patch
# pseudo/local.py:379-703.  The patch lattice, its sphere mask and the k nearest main-grid
# nodes of every patch point depend on the geometry only: they are built once per
# (grid, atoms, patch parameters) and kept in HBM; each density is then one gather +
# reduction per atom (ig_idw_quadrature_f64), one masked main-grid reduction
# (ig_patch_main_f64) and, for the frozen embedding, one masked field update.


@dataclass(frozen=True)
class LocalPotentialPatchParameters:
    """pseudo/local.py:55-76 (defaults and validation)."""

    patch_radius_scale: float = 1.0
    patch_grid_shape: tuple = (25, 25, 25)
    correction_strength: float = 1.0
    interpolation_neighbors: int = 8

    def __post_init__(self) -> None:
        if self.patch_radius_scale <= 0.0:
            raise ValueError("patch_radius_scale must be positive.")
        if self.correction_strength < 0.0:
            raise ValueError("correction_strength must be non-negative.")
        if self.interpolation_neighbors <= 0:
            raise ValueError("interpolation_neighbors must be positive.")
        if len(self.patch_grid_shape) != 3:
            raise ValueError("patch_grid_shape must be a 3-tuple.")
        for axis_points in self.patch_grid_shape:
            if axis_points < 3 or axis_points % 2 == 0:
                raise ValueError("Each patch-grid direction must use an odd number of points >= 3.")


@dataclass(frozen=True)
class AtomicLocalPotentialPatchCorrection:
    """pseudo/local.py:79-92."""

    atom_index: int
    element: str
    position: tuple
    patch_radius_bohr: float
    patch_grid_shape: tuple
    correction_strength: float
    main_grid_patch_energy: float
    patch_integrated_energy: float
    correction_energy: float
    patch_point_count: int


@dataclass(frozen=True)
class LocalIonicPotentialPatchEvaluation:
    """pseudo/local.py:95-113."""

    base_evaluation: LocalIonicPotentialEvaluation
    density_field: np.ndarray
    patch_parameters: LocalPotentialPatchParameters
    atomic_patch_corrections: tuple
    uncorrected_local_energy: float
    corrected_local_energy: float
    total_patch_correction: float


@dataclass(frozen=True)
class FrozenPatchLocalPotentialEmbedding:
    """pseudo/local.py:116-135."""

    base_evaluation: LocalIonicPotentialEvaluation
    patch_evaluation: LocalIonicPotentialPatchEvaluation
    corrected_total_local_potential: np.ndarray
    correction_field: np.ndarray
    atomic_scaling_factors: tuple
    embedded_patch_correction_energy: float
    embedding_energy_mismatch: float


def geometry_patch_radii(grid_geometry) -> tuple:
    """Per-atom patch radii of a monitor grid: the reference's `patch_interfaces`
    (which must implement `local_gth`, pseudo/local.py:467-471) or `patch_radii`."""
    interfaces = getattr(grid_geometry, "patch_interfaces", None)
    if interfaces is not None:
        radii = []
        for itf in interfaces:
            if "local_gth" not in tuple(getattr(itf, "implemented_purposes", ("local_gth",))):
                raise ValueError("This monitor-grid patch interface does not implement local-GTH correction.")
            radii.append(float(itf.patch_radius))
        return tuple(radii)
    radii = tuple(float(r) for r in (getattr(grid_geometry, "patch_radii", ()) or ()))
    if not radii:
        raise ValueError("The near-core patch correction needs a monitor grid with per-atom patch radii.")
    return radii


def _host_atom(row):
    arr = (ctypes.c_double * 9)(*[float(v) for v in row])
    return ctypes.cast(arr, ctypes.c_void_p), arr


class _AtomPatch:
    """Geometry-only data of one atom's patch: lattice points inside the sphere (device
    SoA), V_A there, and the k nearest main-grid nodes of each point."""

    def __init__(self, dg: DeviceGrid, atom_index, element, position, radius, shape, neighbors):
        self.atom_index = atom_index
        self.element = element
        self.position = tuple(float(c) for c in position)
        self.radius = float(radius)
        self.shape = tuple(int(n) for n in shape)
        self.k = int(neighbors)
        R = self.radius
        # pseudo/local.py:379-413 (np.linspace lattice, cell volume from the first steps)
        axes = [np.linspace(c - R, c + R, n, dtype=np.float64) for c, n in zip(self.position, self.shape)]
        self.cell_volume = float(axes[0][1] - axes[0][0]) * float(axes[1][1] - axes[1][0]) * \
            float(axes[2][1] - axes[2][0])
        PX, PY, PZ = np.meshgrid(*axes, indexing="ij")
        dx, dy, dz = PX - self.position[0], PY - self.position[1], PZ - self.position[2]
        mask = np.sqrt(dx * dx + dy * dy + dz * dz, dtype=np.float64) <= R
        pts = np.ascontiguousarray(np.stack([PX[mask], PY[mask], PZ[mask]]))
        self.nq = int(pts.shape[1])
        dev = dg.device
        self.row = _atom_rows([(element, self.position)])[0]
        self.q = torch.from_numpy(pts).to(dev)
        self.vq = torch.empty(self.nq, dtype=torch.float64, device=dev)
        self.idx = torch.empty((self.nq, self.k), dtype=torch.int32, device=dev)
        self.dist = torch.empty((self.nq, self.k), dtype=torch.float64, device=dev)
        if self.nq == 0:
            return
        rows = torch.from_numpy(self.row[None, :].copy()).to(dev)
        _lib.call("ig_vloc_gth_f64", _lib.ptr(self.q), self.nq, 1, _lib.ptr(rows), _lib.ptr(self.vq), 0,
                  _lib.stream_handle())
        if self.k > dg.N:
            raise ValueError("interpolation_neighbors exceeds the number of grid nodes.")
        lib = _lib.load()
        nbytes = int(lib.ig_knn_work_bytes(self.nq, self.k))
        if nbytes <= 0:
            raise ValueError("interpolation_neighbors must be between 1 and 32 on the HIP path.")
        work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        _lib.call("ig_knn_f64", _lib.ptr(dg.xyz()), dg.N, _lib.ptr(self.q), self.nq, self.k, _lib.ptr(work),
                  nbytes, _lib.ptr(self.idx), _lib.ptr(self.dist), _lib.stream_handle())
        del work

    @_lib.on_stream
    def quadrature(self, dg, rho, values=None, stream=None):
        """(sum over patch points of rho_interp V_A) * patch cell volume, on the device."""
        out = torch.zeros(1, dtype=torch.float64, device=dg.device)
        if self.nq == 0:
            return out
        lib = _lib.load()
        work = workspace(dg.device, lib.ig_physics_work_size(dg.N), "phys")
        _lib.call("ig_idw_quadrature_f64", _lib.ptr(rho), _lib.ptr(self.idx), _lib.ptr(self.dist), self.nq, self.k,
                  _lib.ptr(self.vq), _lib.ptr(values), _lib.ptr(work), _lib.ptr(out), _lib.stream_handle(stream))
        return out * self.cell_volume

    @_lib.on_stream
    def main_grid(self, dg, rho, stream=None):
        """[sum rho V_A w, sum rho w, node count] over the main-grid nodes inside the sphere."""
        lib = _lib.load()
        work = workspace(dg.device, lib.ig_physics_work_size(dg.N), "phys")
        out = torch.zeros(3, dtype=torch.float64, device=dg.device)
        aptr, keep = _host_atom(self.row)
        _lib.call("ig_patch_main_f64", _lib.ptr(dg.xyz()), _lib.ptr(rho), _lib.ptr(dg.weights()), dg.N, aptr,
                  self.radius, _lib.ptr(work), _lib.ptr(out), _lib.stream_handle(stream))
        del keep
        return out

    @_lib.on_stream
    def add_field(self, dg, out, scale, constant, stream=None):
        aptr, keep = _host_atom(self.row)
        _lib.call("ig_patch_field_f64", _lib.ptr(dg.xyz()), dg.N, aptr, self.radius, int(constant), float(scale),
                  _lib.ptr(out), _lib.stream_handle(stream))
        del keep


class PatchPlan:
    """All atoms' patches for one grid and parameter set."""

    def __init__(self, dg: DeviceGrid, atoms, radii, params: LocalPotentialPatchParameters):
        if len(radii) != len(atoms):
            raise ValueError("The grid's patch interfaces do not match the atoms of the case.")
        self.params = params
        self.patches = [
            _AtomPatch(dg, n, el, pos, radii[n] * params.patch_radius_scale, params.patch_grid_shape,
                       params.interpolation_neighbors)
            for n, (el, pos) in enumerate(atoms)
        ]

    @_lib.on_stream
    def corrections(self, dg, rho, stream=None):
        """Per atom (pseudo/local.py:461-517): main-grid and patch energies, correction
        energy, density weight and node count inside the sphere (host floats)."""
        res = []
        for p in self.patches:
            pe = p.quadrature(dg, rho, stream=stream)
            mg = p.main_grid(dg, rho, stream)
            vals = torch.cat([pe, mg]).cpu().numpy()
            patch_e, main_e, dw, cnt = (float(v) for v in vals)
            ce = self.params.correction_strength * (patch_e - main_e)
            res.append((p, main_e, patch_e, ce, dw, int(round(cnt))))
        return res

    @_lib.on_stream
    def embed(self, dg, corrections, out=None, stream=None):
        """Frozen correction field (pseudo/local.py:546-582) accumulated into `out`;
        returns (field, per-atom scaling factors)."""
        if out is None:
            out = torch.zeros(dg.N, dtype=torch.float64, device=dg.device)
        scalings = []
        for p, main_e, _, ce, dw, cnt in corrections:
            if cnt == 0 or abs(ce) <= 1.0e-16:
                scalings.append(0.0)
                continue
            if abs(main_e) > 1.0e-14:
                scale = ce / main_e
                p.add_field(dg, out, scale, False, stream)
            elif dw > 1.0e-16:
                scale = ce / dw
                p.add_field(dg, out, scale, True, stream)
            else:
                scale = 0.0
            scalings.append(float(scale))
        return out, tuple(scalings)


def patch_plan(dg: DeviceGrid, case, grid_geometry, params: LocalPotentialPatchParameters) -> PatchPlan:
    """The grid's patch plan for (atoms, patch radii, parameters), built once and cached on
    the DeviceGrid itself: the kNN indices index that grid's nodes, so the plan lives and
    dies with it (a cache keyed by object identity could hand a recycled id a stale plan)."""
    atoms = case_atoms(case)
    radii = geometry_patch_radii(grid_geometry)
    key = (tuple(atoms), radii, params)
    cache = getattr(dg, "_patch_plans", None)
    if cache is None:
        cache = {}
        dg._patch_plans = cache
    plan = cache.get(key)
    if plan is None:
        plan = PatchPlan(dg, atoms, radii, params)
        cache[key] = plan
    return plan


@_lib.on_stream
def frozen_patch_field_device(dg: DeviceGrid, case, grid_geometry, rho, params=None, stream=None):
    """Device-level frozen patch embedding for a flat FP64 density: (correction field,
    per-atom correction records, scaling factors)."""
    params = LocalPotentialPatchParameters() if params is None else params
    plan = patch_plan(dg, case, grid_geometry, params)
    corr = plan.corrections(dg, rho.reshape(-1), stream)
    field, scalings = plan.embed(dg, corr, stream=stream)
    return field, corr, scalings


def _patch_density(density_field, dg):
    rho = np.asarray(density_field)
    if rho.ndim != 3:
        raise ValueError(f"density_field must be a 3D field; received ndim={rho.ndim}.")
    if rho.shape != dg.shape:
        raise ValueError(f"density_field must match the grid shape {dg.shape}; received {rho.shape}.")
    return np.asarray(rho, dtype=np.float64)


def evaluate_monitor_grid_local_ionic_potential_with_patch(case, grid_geometry, density_field, *,
                                                           patch_parameters=None, base_evaluation=None):
    """pseudo/local.py:520-582 on the GPU."""
    dg = device_grid(grid_geometry)
    density = _patch_density(density_field, dg)
    params = LocalPotentialPatchParameters() if patch_parameters is None else patch_parameters
    if base_evaluation is None:
        base_evaluation = evaluate_local_ionic_potential(case, grid_geometry)
    uncorrected = evaluate_local_ionic_energy(base_evaluation, density)
    plan = patch_plan(dg, case, grid_geometry, params)
    rho = torch.from_numpy(np.ascontiguousarray(density)).to(dg.device).reshape(-1)
    corr = plan.corrections(dg, rho)
    records = tuple(
        AtomicLocalPotentialPatchCorrection(p.atom_index, p.element, p.position, p.radius, p.shape,
                                            float(params.correction_strength), main_e, patch_e, float(ce), p.nq)
        for p, main_e, patch_e, ce, _, _ in corr)
    total = float(sum(r.correction_energy for r in records))
    return LocalIonicPotentialPatchEvaluation(base_evaluation, density, params, records, float(uncorrected),
                                              float(uncorrected + total), total)


def evaluate_monitor_grid_local_ionic_potential_with_frozen_patch_field(case, grid_geometry, density_field, *,
                                                                        patch_parameters=None,
                                                                        patch_evaluation=None,
                                                                        base_evaluation=None):
    """pseudo/local.py:585-703 on the GPU."""
    from .reductions import wdot

    dg = device_grid(grid_geometry)
    density = _patch_density(density_field, dg)
    params = LocalPotentialPatchParameters() if patch_parameters is None else patch_parameters
    if patch_evaluation is None:
        patch_evaluation = evaluate_monitor_grid_local_ionic_potential_with_patch(
            case, grid_geometry, density, patch_parameters=params, base_evaluation=base_evaluation)
    base = patch_evaluation.base_evaluation
    plan = patch_plan(dg, case, grid_geometry, patch_evaluation.patch_parameters)
    rho = torch.from_numpy(np.ascontiguousarray(density)).to(dg.device).reshape(-1)
    corr = []
    for p, rec in zip(plan.patches, patch_evaluation.atomic_patch_corrections):
        mg = p.main_grid(dg, rho).cpu().numpy()
        corr.append((p, rec.main_grid_patch_energy, rec.patch_integrated_energy, rec.correction_energy,
                     float(mg[1]), int(round(float(mg[2])))))
    field, scalings = plan.embed(dg, corr)
    embedded = wdot(dg, rho, field)
    f = field.cpu().numpy().reshape(dg.shape)
    return FrozenPatchLocalPotentialEmbedding(
        base, patch_evaluation, np.asarray(base.total_local_potential + f, dtype=np.float64), f, scalings,
        float(embedded), float(embedded - patch_evaluation.total_patch_correction))


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

    @_lib.on_stream
    def project(self, psi, stream=None):
        """[(atom, coef (k, nproj), d = coef H^T (k, nproj))] for a block psi (k, *shape)."""
        dg = self.dg
        if psi.dtype not in (torch.float32, torch.float64) or psi.numel() % dg.N != 0:
            raise ValueError(f"psi must be a float32 / float64 block of fields of {dg.N} points.")
        blk = psi.reshape(-1, dg.N)
        if blk.stride(1) != 1:
            blk = blk.contiguous()
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

    @_lib.on_stream
    def apply(self, psi, out, stream=None, projections=None):
        """out += V_nl psi for a block (out, psi of one dtype and shape).  Returns the projections."""
        dg = self.dg
        if projections is None:
            projections = self.project(psi, stream)
        if out.dtype not in (torch.float32, torch.float64) or out.numel() % dg.N != 0:
            raise ValueError(f"out must be a float32 / float64 block of fields of {dg.N} points.")
        ob = out.reshape(-1, dg.N)
        if ob.stride(1) != 1 or ob.data_ptr() != out.data_ptr():
            raise ValueError("out is updated in place and must have unit stride along the grid")
        k = ob.shape[0]
        if any(d.shape[0] != k for _, _, d in projections):
            raise ValueError("the projections do not belong to a block of the size of out")
        sfx = "f64" if ob.dtype == torch.float64 else "f32"
        nx, ny, nz = dg.shape
        for ap, _, d in projections:
            _lib.call(f"ig_nonlocal_apply_{sfx}", _lib.ptr(ob), ob.stride(0), k, nx, ny, nz, ap.box, _lib.ptr(ap.P),
                      ap.nproj, _lib.ptr(d), _lib.stream_handle(stream))
        return projections

    @_lib.on_stream
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

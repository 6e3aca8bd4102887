"""Monitor-driven harmonic structured grids generated on the GPU (drop-in for the
reference's grid/monitor_builder.py:63-216 and grid/monitor_geometry.py:481-577).

Per outer iteration (the reference's algorithm):
  M      = 1 + sum_A cap_A (1 - exp(-raw_A / cap_A))     atom-centred monitor on the
                                                          current coordinates
  M     <- max((1 - s) M + s avg7(M), 1) inside          smoothing
  X'     = damped Jacobi sweeps of div(M grad X) = 0 with the faces fixed, until the
           largest change is below the tolerance
  X     <- X + alpha (X' - X), alpha halved from the relaxation until det(dX/dxi) > 1e-12
until the outer displacement is below the tolerance.  The monitor, every sweep, the
Jacobian check and the final metric (J, g^-1, cell volumes) run in HIP kernels on
device-resident SoA coordinates; the host only reads the per-sweep convergence scalar.
"""

from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np
import torch

from . import _lib
from .grid import MappedMonitorGeometry
from .pseudo import case_atoms

_GTH_MONITOR_SEEDS = {
    # element: (ionic charge, r_loc, local coefficients, projector radii)
    # grid/monitor_builder.py:25-50
    "H": (1, 0.2, (-4.1802368, 0.72507482), ()),
    "C": (4, 0.34883045, (-8.5137711, 1.22843203), (0.30455321,)),
    "N": (5, 0.28917923, (-12.23481988, 1.76640728), (0.25660487,)),
    "O": (6, 0.24762086, (-16.58031797, 2.39570092), (0.22178614,)),
}
H2_MONITOR_LOCAL_PATCH_BASELINE_WEIGHT_SCALE = 4.0
H2_MONITOR_LOCAL_PATCH_BASELINE_RADIUS_SCALE = 0.70
_JACOBIAN_FLOOR = 1.0e-12


@dataclass(frozen=True)
class NearCoreElementParameters:
    """grid/monitor_model.py:14-43 (fields and validation)."""

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

    def __post_init__(self) -> None:
        if min(self.near_core_radius, self.local_radius, self.projector_radius, self.patch_radius) <= 0.0:
            raise ValueError("near-core, local, projector and patch radii must be positive.")
        if min(self.kinetic_exponent, self.local_exponent, self.projector_exponent) <= 0.0:
            raise ValueError("All monitor exponents must be positive.")
        if self.monitor_cap <= 1.0:
            raise ValueError("monitor_cap must be larger than 1.")


@dataclass
class MonitorGridSpec:
    """grid/monitor_model.py:98-135."""

    name: str
    description: str
    nx: int
    ny: int
    nz: int
    unit: str
    box_bounds: tuple
    logical_bounds: tuple
    element_parameters: dict
    harmonic_outer_iterations: int
    harmonic_inner_iterations: int
    harmonic_tolerance: float
    harmonic_relaxation: float
    inner_relaxation: float
    monitor_smoothing: float

    def __post_init__(self) -> None:
        if self.nx < 3 or self.ny < 3 or self.nz < 3:
            raise ValueError("Monitor grids require at least three points per direction.")
        if self.harmonic_outer_iterations <= 0 or self.harmonic_inner_iterations <= 0:
            raise ValueError("Harmonic iteration counts must be positive.")
        if self.harmonic_tolerance <= 0.0:
            raise ValueError("harmonic_tolerance must be positive.")
        if not (0.0 < self.harmonic_relaxation <= 1.0):
            raise ValueError("harmonic_relaxation must satisfy 0 < value <= 1.")
        if not (0.0 < self.inner_relaxation <= 1.0):
            raise ValueError("inner_relaxation must satisfy 0 < value <= 1.")
        if not (0.0 <= self.monitor_smoothing < 1.0):
            raise ValueError("monitor_smoothing must satisfy 0 <= value < 1.")
        if not self.element_parameters:
            raise ValueError("element_parameters must not be empty.")

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)


def build_default_near_core_element_parameters(case) -> dict:
    """grid/monitor_builder.py:63-102."""
    out = {}
    for el, _ in case_atoms(case):
        if el in out:
            continue
        if el not in _GTH_MONITOR_SEEDS:
            raise ValueError("The current monitor-grid core supports only GTH-inspired H/C/N/O seeds; "
                             f"received `{el}`.")
        z, rloc, coeffs, radii = _GTH_MONITOR_SEEDS[el]
        max_nl = max(radii, default=1.5 * rloc)
        cscale = sum(abs(v) for v in coeffs)
        out[el] = NearCoreElementParameters(
            element=el,
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


def build_h2_local_patch_development_element_parameters(case=None) -> dict:
    """grid/monitor_builder.py:204-216: H2 parameters with weights x4 and radii x0.70."""
    from .hamiltonian import H2_BENCHMARK_CASE

    base = build_default_near_core_element_parameters(H2_BENCHMARK_CASE if case is None else case)
    r, wgt = H2_MONITOR_LOCAL_PATCH_BASELINE_RADIUS_SCALE, H2_MONITOR_LOCAL_PATCH_BASELINE_WEIGHT_SCALE
    return {e: replace(p, near_core_radius=p.near_core_radius * r, local_radius=p.local_radius * r,
                       kinetic_weight=p.kinetic_weight * wgt, local_weight=p.local_weight * wgt)
            for e, p in base.items()}


def _center(atoms):
    n = float(len(atoms))
    return tuple(sum(p[a] for _, p in atoms) / n for a in range(3))


def build_monitor_grid_spec_for_case(case, *, shape=None, box_half_extents=None, element_parameters=None,
                                     harmonic_outer_iterations=6, harmonic_inner_iterations=180,
                                     harmonic_tolerance=5.0e-5, harmonic_relaxation=0.55, inner_relaxation=0.82,
                                     monitor_smoothing=0.18) -> MonitorGridSpec:
    """grid/monitor_builder.py:105-183."""
    atoms = case_atoms(case)
    if element_parameters is None:
        element_parameters = build_default_near_core_element_parameters(case)
    if shape is None:
        shape = (29, 29, 29) if len(atoms) <= 2 else (31, 31, 31)
    center = _center(atoms)
    if box_half_extents is None:
        support = max(max(p.patch_radius, p.local_radius, p.projector_radius) for p in element_parameters.values())
        mad = [max(abs(p[a] - center[a]) for _, p in atoms) for a in range(3)]
        pad = max(6.0, 4.0 * support)
        box_half_extents = tuple(m + pad for m in mad)
    bounds = tuple((center[a] - box_half_extents[a], center[a] + box_half_extents[a]) for a in range(3))
    name = getattr(getattr(case, "geometry", None), "name", "molecule")
    return MonitorGridSpec(
        name=f"{str(name).lower()}_monitor_grid",
        description="Atom-centred monitor-function grid from a weighted harmonic map (GPU generator).",
        nx=int(shape[0]), ny=int(shape[1]), nz=int(shape[2]),
        unit=getattr(getattr(case, "geometry", None), "unit", "bohr"),
        box_bounds=bounds, logical_bounds=((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)),
        element_parameters=element_parameters, harmonic_outer_iterations=harmonic_outer_iterations,
        harmonic_inner_iterations=harmonic_inner_iterations, harmonic_tolerance=harmonic_tolerance,
        harmonic_relaxation=harmonic_relaxation, inner_relaxation=inner_relaxation,
        monitor_smoothing=monitor_smoothing)


def _atom_table(atoms, params):
    rows = []
    for el, pos in atoms:
        p = params[el]
        rows.append([pos[0], pos[1], pos[2], p.kinetic_weight, p.near_core_radius, p.kinetic_exponent,
                     p.local_weight, p.local_radius, p.local_exponent, p.projector_weight, p.projector_radius,
                     p.projector_exponent, p.monitor_cap])
    return np.asarray(rows, dtype=np.float64).reshape(len(atoms), 13)


@dataclass
class HarmonicMapStatistics:
    outer_iterations: int
    inner_iterations: tuple
    relaxations: tuple
    max_displacements: tuple


def generate_monitor_grid_coordinates(case, spec: MonitorGridSpec, device="cuda"):
    """The harmonic-map iteration on the device: returns (coordinates (3, N) FP64 device
    tensor, logical axes, statistics)."""
    nx, ny, nz = spec.shape
    N = nx * ny * nz
    dev = torch.device(device)
    lx, ly, lz = (np.linspace(lo, hi, n, dtype=np.float64) for (lo, hi), n in zip(spec.logical_bounds, spec.shape))
    axes = [np.linspace(lo, hi, n, dtype=np.float64) for (lo, hi), n in zip(spec.box_bounds, spec.shape)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    coords = torch.from_numpy(np.stack([X.ravel(), Y.ravel(), Z.ravel()])).to(dev).contiguous()
    hx, hy, hz = float(lx[1] - lx[0]), float(ly[1] - ly[0]), float(lz[1] - lz[0])
    for ax, h in ((lx, hx), (ly, hy), (lz, hz)):
        if not np.allclose(np.diff(ax), h):
            raise ValueError("The current monitor-grid generator expects uniform logical coordinates.")
    atoms = torch.from_numpy(_atom_table(case_atoms(case), spec.element_parameters)).to(dev)
    lib = _lib.load()
    work = torch.empty(lib.ig_grid_work_size(), dtype=torch.float64, device=dev)
    scal = torch.empty(1, dtype=torch.float64, device=dev)
    mon = torch.empty(N, dtype=torch.float64, device=dev)
    smon = torch.empty(N, dtype=torch.float64, device=dev)
    a = torch.empty_like(coords)
    b = torch.empty_like(coords)
    h = _lib.stream_handle()
    stats = HarmonicMapStatistics(0, (), (), ())
    inner_counts, relaxations, displacements = [], [], []
    for outer in range(spec.harmonic_outer_iterations):
        _lib.call("ig_monitor_field_f64", _lib.ptr(coords), N, _lib.ptr(atoms), atoms.shape[0], _lib.ptr(mon), h)
        _lib.call("ig_monitor_smooth_f64", _lib.ptr(mon), nx, ny, nz, float(spec.monitor_smoothing), _lib.ptr(smon), h)
        a.copy_(coords)
        n_in = 0
        for _ in range(spec.harmonic_inner_iterations):
            _lib.call("ig_harmonic_sweep_f64", _lib.ptr(smon), _lib.ptr(a), _lib.ptr(b), nx, ny, nz, hx, hy, hz,
                      float(spec.inner_relaxation), _lib.ptr(work), _lib.ptr(scal), h)
            a, b = b, a
            n_in += 1
            if float(scal.item()) < spec.harmonic_tolerance:
                break
        # backtracking: the largest fraction of the step that keeps det(dX/dxi) > 1e-12
        alpha = spec.harmonic_relaxation
        updated = coords
        while alpha >= 1.0e-3:
            trial = coords + alpha * (a - coords)
            _lib.call("ig_mapping_min_jacobian_f64", _lib.ptr(trial), nx, ny, nz, hx, hy, hz, _lib.ptr(work),
                      _lib.ptr(scal), h)
            if float(scal.item()) > _JACOBIAN_FLOOR:
                updated = trial
                break
            alpha *= 0.5
        disp = float((updated - coords).abs().max().item())
        coords = updated.contiguous()
        inner_counts.append(n_in)
        relaxations.append(alpha if alpha >= 1.0e-3 else 0.0)
        displacements.append(disp)
        stats = HarmonicMapStatistics(outer + 1, tuple(inner_counts), tuple(relaxations), tuple(displacements))
        if disp < spec.harmonic_tolerance:
            break
    return coords, (lx, ly, lz), stats


def generate_monitor_grid_geometry(case, spec: MonitorGridSpec, device="cuda") -> MappedMonitorGeometry:
    """grid/monitor_geometry.py:481-577 on the GPU: the mapped grid with its Jacobian,
    inverse metric tensor and cell volumes (the arrays the kernels consume)."""
    coords, (lx, ly, lz), stats = generate_monitor_grid_coordinates(case, spec, device)
    nx, ny, nz = spec.shape
    N = nx * ny * nz
    dev = coords.device
    jac = torch.empty(N, dtype=torch.float64, device=dev)
    ginv = torch.empty((N, 9), dtype=torch.float64, device=dev)
    vol = torch.empty(N, dtype=torch.float64, device=dev)
    _lib.call("ig_mapping_metric_f64", _lib.ptr(coords), nx, ny, nz, float(lx[1] - lx[0]), float(ly[1] - ly[0]),
              float(lz[1] - lz[0]), _lib.ptr(jac), _lib.ptr(ginv), _lib.ptr(vol), _lib.stream_handle())
    c = coords.cpu().numpy().reshape(3, nx, ny, nz)
    geom = MappedMonitorGeometry(
        (nx, ny, nz), tuple(tuple(float(v) for v in b) for b in spec.box_bounds), lx, ly, lz,
        c[0].copy(), c[1].copy(), c[2].copy(), jac.cpu().numpy().reshape(nx, ny, nz),
        ginv.cpu().numpy().reshape(nx, ny, nz, 3, 3), vol.cpu().numpy().reshape(nx, ny, nz),
        tuple(spec.element_parameters[el].patch_radius for el, _ in case_atoms(case)), "monitor",
        f"{spec.description} Harmonic map: {stats.outer_iterations} outer iterations, "
        f"inner sweeps {stats.inner_iterations}.")
    geom.harmonic_statistics = stats
    return geom


def build_monitor_grid_for_case(case, *, shape=None, box_half_extents=None, element_parameters=None,
                                device="cuda") -> MappedMonitorGeometry:
    """grid/monitor_builder.py:186-201."""
    spec = build_monitor_grid_spec_for_case(case, shape=shape, box_half_extents=box_half_extents,
                                            element_parameters=element_parameters)
    return generate_monitor_grid_geometry(case, spec, device)

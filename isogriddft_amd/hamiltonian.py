"""Kohn-Sham Hamiltonian applies on the GPU (drop-in for the reference's
ks/local_hamiltonian.py and ks/static_hamiltonian.py).

``BlockHamiltonian`` is the device-resident operator every solver uses:

    H psi = -1/2 Lap psi + V_eff psi [+ V_nl psi]

with the kinetic stencil and the V_eff product fused in one HIP kernel
(ig_monitor_stencil_* / ig_legacy_stencil_*) and the nonlocal projectors applied on
their support boxes afterwards (ig_nonlocal_*).  Blocks are orbital-major
(k, nx, ny, nz) CUDA tensors in float32 or float64.

The NumPy call surfaces mirror the reference one-to-one:
  evaluate_local_hamiltonian_terms / apply_local_hamiltonian  ks/local_hamiltonian.py:69-112
  evaluate_static_ks_terms / apply_static_ks_hamiltonian      ks/static_hamiltonian.py:238-343
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib, configs
from .device import DeviceGrid, device_grid
from .kinetic import apply_operator, apply_operator_dots

H2_BENCHMARK_CASE = configs.H2
_SPIN = ("up", "down")
KINETIC_VERSIONS = ("production", "trial_fix")


def normalize_spin_channel(spin_channel: str) -> str:
    s = str(spin_channel).strip().lower()
    if s not in _SPIN:
        raise ValueError(f"spin_channel must be `up` or `down`; received `{spin_channel}`.")
    return s


def normalize_kinetic_version(kinetic_version: str) -> str:
    s = str(kinetic_version).strip().lower()
    if s not in KINETIC_VERSIONS:
        raise ValueError(f"kinetic_version must be `production` or `trial_fix`; received `{kinetic_version}`.")
    return s


class BlockHamiltonian:
    """H = -1/2 Lap + V_eff (+ V_nl) frozen on one grid, resident in HBM."""

    def __init__(self, dg: DeviceGrid, veff, *, closure="production", nonlocal_plan=None):
        self.dg = dg
        self.closure = normalize_kinetic_version(closure)
        v = torch.as_tensor(veff, dtype=torch.float64).reshape(-1)
        if v.numel() != dg.N:
            raise ValueError(f"effective potential must have {dg.N} points; received {v.numel()}.")
        self.veff = {torch.float64: v.to(dg.device).contiguous()}
        self.veff[torch.float32] = self.veff[torch.float64].float()
        self.nonlocal_plan = nonlocal_plan if (nonlocal_plan is not None and not nonlocal_plan.empty) else None

    @_lib.on_stream
    def apply(self, X: torch.Tensor, out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
        """out = H X for a block X (k, *shape) or one field (*shape)."""
        v = self.veff[X.dtype].view(self.dg.shape)
        out = apply_operator(self.dg, X, out, scale=-0.5, closure=self.closure, veff=v, stream=stream)
        if self.nonlocal_plan is not None:
            self.nonlocal_plan.apply(X, out, stream=stream)
        return out

    __call__ = apply

    def apply_dots(self, X: torch.Tensor, out: torch.Tensor, w: torch.Tensor) -> torch.Tensor | None:
        """out = H X and, from the same pass, the FP64 row dots sum_n w[n] X[m][n] out[m][n];
        None (nothing launched) where the fused kernel does not take the block or the
        Hamiltonian carries projectors (then apply and take the dots separately)."""
        if self.nonlocal_plan is not None:
            return None
        v = self.veff[X.dtype].view(self.dg.shape)
        return apply_operator_dots(self.dg, X, out, w, scale=-0.5, closure=self.closure, veff=v)

    def diagonal(self) -> torch.Tensor:
        """Diagonal of -1/2 Lap + V_eff (FP64, flat; the nonlocal term is left out): the
        Davidson preconditioner of the block eigensolver."""
        return -0.5 * self.dg.kinetic_diagonal() + self.veff[torch.float64]


# ------------------------------------------------------------------ NumPy call surfaces


def _field(x, dg, name="psi"):
    a = np.asarray(x)
    if a.shape != dg.shape:
        raise ValueError(f"{name} must have shape {dg.shape}; received {a.shape}.")
    if not np.all(np.isfinite(a)):
        raise ValueError(f"{name} must contain only finite values.")
    return np.asarray(a, dtype=np.float64)


def _resolve_local(grid_geometry, dg, case, local_ionic_potential):
    from . import pseudo

    if local_ionic_potential is None:
        ev = pseudo.evaluate_local_ionic_potential(case, grid_geometry)
        return ev.total_local_potential, ev
    if isinstance(local_ionic_potential, pseudo.LocalIonicPotentialEvaluation):
        return np.asarray(local_ionic_potential.total_local_potential, dtype=np.float64), local_ionic_potential
    return _field(local_ionic_potential, dg, "local_ionic_potential"), None


def _to_dev(a, dg):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dg.device)


@dataclass(frozen=True)
class LocalHamiltonianTerms:
    """Fields of ks/local_hamiltonian.py:29-36."""

    psi: np.ndarray
    local_potential: np.ndarray
    kinetic_action: np.ndarray
    local_potential_action: np.ndarray
    total_action: np.ndarray


def evaluate_local_hamiltonian_terms(psi, grid_geometry, case=H2_BENCHMARK_CASE, local_ionic_potential=None,
                                     extra_local_potential=None) -> LocalHamiltonianTerms:
    """T psi, V_local psi and their sum (ks/local_hamiltonian.py:69-94), production closure."""
    dg = device_grid(grid_geometry)
    field = _field(psi, dg)
    local, _ = _resolve_local(grid_geometry, dg, case, local_ionic_potential)
    if extra_local_potential is not None:
        local = local + _field(extra_local_potential, dg, "extra_local_potential")
    x = _to_dev(field, dg)
    kin = apply_operator(dg, x, scale=-0.5, closure="production")
    vt = _to_dev(local, dg)
    vpsi = vt * x
    total = kin + vpsi
    return LocalHamiltonianTerms(field, np.asarray(local), kin.cpu().numpy(), vpsi.cpu().numpy(), total.cpu().numpy())


def apply_local_hamiltonian(psi, grid_geometry, case=H2_BENCHMARK_CASE, local_ionic_potential=None,
                            extra_local_potential=None):
    """ks/local_hamiltonian.py:97-112: T psi + V_local psi in one fused kernel."""
    dg = device_grid(grid_geometry)
    field = _field(psi, dg)
    local, _ = _resolve_local(grid_geometry, dg, case, local_ionic_potential)
    if extra_local_potential is not None:
        local = local + _field(extra_local_potential, dg, "extra_local_potential")
    H = BlockHamiltonian(dg, _to_dev(local, dg), closure="production")
    return H.apply(_to_dev(field, dg)).cpu().numpy()


@dataclass(frozen=True)
class StaticKSHamiltonianTerms:
    """Fields of ks/static_hamiltonian.py:44-65."""

    psi: np.ndarray
    rho_up: np.ndarray
    rho_down: np.ndarray
    rho_total: np.ndarray
    spin_channel: str
    kinetic_action: np.ndarray
    local_ionic_potential: np.ndarray
    local_ionic_action: np.ndarray
    nonlocal_ionic_action: np.ndarray
    hartree_potential: np.ndarray
    hartree_action: np.ndarray
    hartree_energy: float
    xc_potential: np.ndarray
    xc_action: np.ndarray
    total_action: np.ndarray
    local_ionic_evaluation: object
    nonlocal_ionic_evaluation: object
    hartree_poisson_result: object
    lsda_evaluation: object


def resolve_hartree(grid_geometry, dg, rho_total, hartree_potential, **poisson_kwargs):
    """(v_H, E_H, result) as ks/static_hamiltonian.py:123-168."""
    from . import poisson

    if hartree_potential is None:
        res = poisson.solve_hartree_potential(grid_geometry, rho_total, **poisson_kwargs)
        pot = res.potential
    elif isinstance(hartree_potential, poisson.OpenBoundaryPoissonResult):
        res = hartree_potential
        pot = np.asarray(res.potential, dtype=np.float64)
    else:
        res = None
        pot = _field(hartree_potential, dg, "hartree_potential")
    eh = poisson.evaluate_hartree_energy(rho_total, grid_geometry, pot)
    return pot, eh, res


def resolve_xc(dg, rho_up, rho_down, spin, xc_potential, xc_functional):
    """(v_xc of the channel, LSDA evaluation) as ks/static_hamiltonian.py:171-196."""
    from . import xc

    if xc_potential is not None:
        return _field(xc_potential, dg, "xc_potential"), None
    ev = xc.evaluate_lsda_terms(rho_up, rho_down, functional=xc_functional)
    return (ev.v_xc_up if spin == "up" else ev.v_xc_down), ev


def evaluate_static_ks_terms(psi, grid_geometry, rho_up, rho_down, spin_channel, case=H2_BENCHMARK_CASE,
                             local_ionic_potential=None, nonlocal_ionic_action=None, hartree_potential=None,
                             xc_potential=None, xc_functional=None) -> StaticKSHamiltonianTerms:
    """ks/static_hamiltonian.py:238-313 with every piece formed on the GPU."""
    from . import pseudo

    dg = device_grid(grid_geometry)
    field = _field(psi, dg)
    up = _field(rho_up, dg, "rho_up")
    dn = _field(rho_down, dg, "rho_down")
    tot = up + dn
    spin = normalize_spin_channel(spin_channel)
    local, local_ev = _resolve_local(grid_geometry, dg, case, local_ionic_potential)
    if nonlocal_ionic_action is None:
        nl_ev = pseudo.evaluate_nonlocal_ionic_action(case, grid_geometry, field)
        nl = nl_ev.total_nonlocal_action
    elif isinstance(nonlocal_ionic_action, pseudo.NonlocalIonicActionEvaluation):
        nl_ev = nonlocal_ionic_action
        nl = np.asarray(nl_ev.total_nonlocal_action, dtype=np.float64)
    else:
        nl_ev = None
        nl = _field(nonlocal_ionic_action, dg, "nonlocal_ionic_action")
    vh, eh, hres = resolve_hartree(grid_geometry, dg, tot, hartree_potential)
    vxc, lsda = resolve_xc(dg, up, dn, spin, xc_potential, xc_functional)
    x = _to_dev(field, dg)
    kin = apply_operator(dg, x, scale=-0.5, closure="production")
    loc_a = _to_dev(local, dg) * x
    h_a = _to_dev(vh, dg) * x
    x_a = _to_dev(vxc, dg) * x
    total = kin + loc_a + _to_dev(nl, dg) + h_a + x_a
    c = lambda t: t.cpu().numpy()  # noqa: E731
    return StaticKSHamiltonianTerms(field, up, dn, tot, spin, c(kin), np.asarray(local), c(loc_a), np.asarray(nl),
                                    np.asarray(vh), c(h_a), float(eh), np.asarray(vxc), c(x_a), c(total),
                                    local_ev, nl_ev, hres, lsda)


def apply_static_ks_hamiltonian(psi, grid_geometry, rho_up, rho_down, spin_channel, case=H2_BENCHMARK_CASE,
                                local_ionic_potential=None, nonlocal_ionic_action=None, hartree_potential=None,
                                xc_potential=None, xc_functional=None):
    """ks/static_hamiltonian.py:316-343."""
    return evaluate_static_ks_terms(psi, grid_geometry, rho_up, rho_down, spin_channel, case, local_ionic_potential,
                                    nonlocal_ionic_action, hartree_potential, xc_potential,
                                    xc_functional).total_action

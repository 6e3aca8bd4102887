"""Workload definitions: molecules and the synthetic operator-level configuration.

Molecules follow the reference's benchmark cases (config/defaults.py:14-131): H2 at
R = 1.4 Bohr on z, H2O (O-H 1.809 Bohr, 104.5 deg).  Benzene (D6h, C-C 2.637 Bohr,
C-H 2.048 Bohr) has no reference counterpart.

The synthetic operator-level workload ("c2"): a 96^3 monitor-style grid from the
analytic two-centre sinh map (centres 1.4 Bohr apart on z, a = 0.3 with a seeded
+-5 % jitter, b = 1.5, Gaussian weights of width 1 Bohr, box +-8 Bohr), a block of
16 orbitals each a sum of 8 seeded Gaussians (exponents log-uniform in
[0.3, 3] Bohr^-2) and a Poisson right-hand side of 6 seeded normalised Gaussian
charges.  All parameters and seeds are here; nothing is read from disk.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from .grid import build_legacy_geometry, build_mapped_monitor_geometry


@dataclass(frozen=True)
class Atom:
    element: str
    position: tuple


@dataclass(frozen=True)
class Molecule:
    name: str
    atoms: tuple

    @property
    def positions(self):
        return np.array([a.position for a in self.atoms], dtype=np.float64)


H2 = Molecule("H2", (Atom("H", (0.0, 0.0, -0.7)), Atom("H", (0.0, 0.0, 0.7))))


def _h2o():
    half = 0.5 * math.radians(104.5)
    hx, hz = 1.809 * math.sin(half), 1.809 * math.cos(half)
    return Molecule("H2O", (Atom("O", (0.0, 0.0, 0.0)), Atom("H", (-hx, 0.0, hz)), Atom("H", (hx, 0.0, hz))))


def _benzene():
    atoms = []
    for i in range(6):
        t = math.pi / 3.0 * i
        atoms.append(Atom("C", (2.637 * math.cos(t), 2.637 * math.sin(t), 0.0)))
    for i in range(6):
        t = math.pi / 3.0 * i
        r = 2.637 + 2.048
        atoms.append(Atom("H", (r * math.cos(t), r * math.sin(t), 0.0)))
    return Molecule("C6H6", tuple(atoms))


H2O = _h2o()
BENZENE = _benzene()


@dataclass(frozen=True)
class SyntheticConfig:
    shape: tuple = (96, 96, 96)
    half_extents: tuple = (8.0, 8.0, 8.0)
    centers: tuple = ((0.0, 0.0, -0.7), (0.0, 0.0, 0.7))
    a: float = 0.3
    jitter: float = 0.05
    b: float = 1.5
    sigma: float = 1.0
    k: int = 16
    gaussians_per_orbital: int = 8
    exponent_range: tuple = (0.3, 3.0)
    center_box: float = 2.5
    n_charges: int = 6
    seed: int = 20251203


C2 = SyntheticConfig()
# The same analytic map and orbital recipe at the benzene-size block (224^3, k = 48,
# box +-14 Bohr): the operator-level stand-in for the largest configuration.
C5_OPERATOR = SyntheticConfig(shape=(224, 224, 224), half_extents=(14.0, 14.0, 14.0), k=48)


def synthetic_amplitudes(cfg: SyntheticConfig = C2):
    rng = np.random.default_rng(cfg.seed)
    return tuple(float(cfg.a * (1.0 + cfg.jitter * rng.uniform(-1.0, 1.0))) for _ in cfg.centers)


def build_synthetic_geometry(cfg: SyntheticConfig = C2, shape=None):
    return build_mapped_monitor_geometry(
        shape or cfg.shape, cfg.half_extents, cfg.centers, synthetic_amplitudes(cfg),
        b=cfg.b, sigma=cfg.sigma,
        description="analytic two-centre sinh map (synthetic operator-level workload)",
    )


def _gaussian_sum(xyz, centers, exps, coefs, norm=False):
    """sum_g c_g exp(-alpha_g |r - R_g|^2) on points xyz (3, N) (torch, any device)."""
    out = torch.zeros(xyz.shape[1], dtype=xyz.dtype, device=xyz.device)
    for R, a, c in zip(centers, exps, coefs):
        d2 = (xyz[0] - R[0]) ** 2 + (xyz[1] - R[1]) ** 2 + (xyz[2] - R[2]) ** 2
        pref = (a / math.pi) ** 1.5 if norm else 1.0
        out += c * pref * torch.exp(-a * d2)
    return out


def synthetic_orbital_params(cfg: SyntheticConfig = C2, k=None):
    rng = np.random.default_rng(cfg.seed + 1)
    k = cfg.k if k is None else k
    lo, hi = np.log(cfg.exponent_range[0]), np.log(cfg.exponent_range[1])
    params = []
    for _ in range(k):
        g = cfg.gaussians_per_orbital
        centers = rng.uniform(-cfg.center_box, cfg.center_box, size=(g, 3))
        exps = np.exp(rng.uniform(lo, hi, size=g))
        coefs = rng.normal(size=g)
        params.append((centers, exps, coefs))
    return params


def build_synthetic_orbitals(geom, cfg: SyntheticConfig = C2, k=None, device="cpu", dtype=torch.float64):
    """(k, nx, ny, nz) block of seeded Gaussian sums."""
    xyz = torch.stack([torch.as_tensor(np.asarray(getattr(geom, n), dtype=np.float64).reshape(-1))
                       for n in ("x_points", "y_points", "z_points")]).to(device)
    params = synthetic_orbital_params(cfg, k)
    shape = tuple(np.shape(geom.cell_volumes))
    out = torch.empty((len(params),) + shape, dtype=dtype, device=device)
    for m, (c, e, w) in enumerate(params):
        out[m] = _gaussian_sum(xyz, c, e, w).reshape(shape).to(dtype)
    return out


def build_synthetic_density(geom, cfg: SyntheticConfig = C2, device="cpu"):
    """Sum of ``n_charges`` normalised Gaussian charges (total charge n_charges)."""
    rng = np.random.default_rng(cfg.seed + 2)
    centers = rng.uniform(-2.0, 2.0, size=(cfg.n_charges, 3))
    exps = rng.uniform(0.5, 2.0, size=cfg.n_charges)
    xyz = torch.stack([torch.as_tensor(np.asarray(getattr(geom, n), dtype=np.float64).reshape(-1))
                       for n in ("x_points", "y_points", "z_points")]).to(device)
    rho = _gaussian_sum(xyz, centers, exps, np.ones(cfg.n_charges), norm=True)
    return rho.reshape(tuple(np.shape(geom.cell_volumes)))


def legacy_h2_geometry(shape=(49, 49, 61), half_extents=(8.0, 8.0, 8.0), stretch=(2.5, 2.5, 2.8)):
    """Legacy sinh grid about the H2 centre (the reference's default is 51^3, (8,8,10))."""
    return build_legacy_geometry(shape, half_extents, stretch, (0.0, 0.0, 0.0))


def atom_centered_geometry(mol: Molecule, shape, half_extent, a=-0.25, b=1.2, sigma=0.9):
    """Atom-centred monitor-style grid: one refining sinh centre per atom (a < 0)."""
    return build_mapped_monitor_geometry(
        shape, (half_extent,) * 3, [at.position for at in mol.atoms], [a] * len(mol.atoms),
        b=b, sigma=sigma, description=f"atom-centred analytic map for {mol.name}",
    )

"""Oracle: GTH-PADE pseudopotential terms.

Restates:
  * gth-pade H/C/N/O table            pseudo/gth_data.py:28-117
  * local GTH field (erf + Gaussian)   pseudo/local.py:138-180, :303-357
  * near-core patch + frozen embedding pseudo/local.py:379-703
  * separable nonlocal projectors      pseudo/nonlocal.py:99-321

Pinning: V_loc by the closed form with ``math.erf`` (independent scalar loop) and the
reference-held <V_loc> = -2.173678262801 Ha on the default legacy 51^3 grid; the
nonlocal term (tests/test_oracle_nonlocal.py) by the closed-form projector normalisation
int r^2 p^2 dr = 1, the analytic overlaps of the s and p projectors with atom-centred
Gaussians, and the invariants of the dense nonlocal matrix built column by column on tiny
grids (W-self-adjoint, rank = number of projectors, nonzero eigenvalue h <p, p>_W with
eigenvector p, additivity over atoms); the patch correction by a reference-generated
fixture (cKDTree neighbour order is only pinned by import).
"""

from __future__ import annotations

from dataclasses import dataclass
from math import erf, gamma, pi, sqrt

import numpy as np


@dataclass(frozen=True)
class GTHChannel:
    l: int
    radius: float
    h: tuple  # h_ij matrix rows


@dataclass(frozen=True)
class GTHData:
    element: str
    zion: int
    rloc: float
    coefficients: tuple
    channels: tuple


GTH_PADE = {
    "H": GTHData("H", 1, 0.2, (-4.1802368, 0.72507482), ()),
    "C": GTHData(
        "C", 4, 0.34883045, (-8.5137711, 1.22843203),
        (GTHChannel(0, 0.30455321, ((9.52284179,),)), GTHChannel(1, 0.2326773, ())),
    ),
    "N": GTHData(
        "N", 5, 0.28917923, (-12.23481988, 1.76640728),
        (GTHChannel(0, 0.25660487, ((13.55224272,),)), GTHChannel(1, 0.27013369, ())),
    ),
    "O": GTHData(
        "O", 6, 0.24762086, (-16.58031797, 2.39570092),
        (GTHChannel(0, 0.22178614, ((18.26691718,),)), GTHChannel(1, 0.2568289, ())),
    ),
}


def local_potential_radial(r, data: GTHData):
    """pseudo/local.py:138-180 (r = 0 limit -Z sqrt(2/pi)/rloc)."""
    r = np.asarray(r, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        scaled = r / (sqrt(2.0) * data.rloc)
        erfv = np.vectorize(erf, otypes=[np.float64])(scaled)
        coul = -data.zion * erfv / r
    coul = np.where(r == 0.0, -data.zion * sqrt(2.0 / pi) / data.rloc, coul)
    x = r / data.rloc
    poly = np.zeros_like(r)
    for i, c in enumerate(data.coefficients):
        poly = poly + c * x ** (2 * i)
    return coul + np.exp(-0.5 * x * x) * poly


def local_ionic_potential(case, geom):
    """pseudo/local.py:303-357: sum of atomic local GTH potentials on the grid points."""
    total = np.zeros(geom.x_points.shape)
    for atom in case.atoms:
        d = GTH_PADE[atom.element]
        dx = geom.x_points - atom.position[0]
        dy = geom.y_points - atom.position[1]
        dz = geom.z_points - atom.position[2]
        total += local_potential_radial(np.sqrt(dx * dx + dy * dy + dz * dz), d)
    return total


def local_potential_radial_derivative(r, data: GTHData):
    """d V_loc / d r (closed form; pinned against central differences in the tests)."""
    r = np.asarray(r, dtype=np.float64)
    a = 1.0 / (sqrt(2.0) * data.rloc)
    erfv = np.vectorize(erf, otypes=[np.float64])(a * r)
    with np.errstate(divide="ignore", invalid="ignore"):
        dcoul = data.zion * (erfv / (r * r) - 2.0 * a / sqrt(pi) * np.exp(-(a * r) ** 2) / r)
    dcoul = np.where(r == 0.0, 0.0, dcoul)
    x = r / data.rloc
    poly = np.zeros_like(r)
    dpoly = np.zeros_like(r)
    for i, c in enumerate(data.coefficients):
        poly = poly + c * x ** (2 * i)
        if i > 0:
            dpoly = dpoly + c * 2 * i * x ** (2 * i - 1)
    g = np.exp(-0.5 * x * x)
    return dcoul + g * (dpoly - x * poly) / data.rloc


# ---------------------------------------------------------------- nonlocal

_Y00 = 1.0 / sqrt(4.0 * pi)
_Y1 = sqrt(3.0 / (4.0 * pi))


def radial_projector(r, l, i, rl):
    """pseudo/nonlocal.py:143-155 (zero-based projector index i)."""
    order = i + 1
    power = l + 2 * i
    garg = l + 0.5 * (4 * order - 1)
    pref = sqrt(2.0) / (rl ** garg * sqrt(gamma(garg)))
    return pref * r ** power * np.exp(-0.5 * (r / rl) ** 2)


def real_harmonic(l, m, dx, dy, dz, r):
    """pseudo/nonlocal.py:99-129."""
    if l == 0:
        return np.full(dx.shape, _Y00)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = {-1: _Y1 * dy / r, 0: _Y1 * dz / r, 1: _Y1 * dx / r}[m]
    v = np.asarray(v, dtype=np.float64)
    v[r == 0.0] = 0.0
    return v


def projector_fields(case, geom):
    """All (atom, l, m, i) projector fields with their h-matrix blocks.

    Returns a list of (atom_index, fields[(nproj, *shape)], h[(nproj, nproj)]) following
    the channel/m/i ordering of pseudo/nonlocal.py:209-283.
    """
    out = []
    for ia, atom in enumerate(case.atoms):
        d = GTH_PADE[atom.element]
        dx = geom.x_points - atom.position[0]
        dy = geom.y_points - atom.position[1]
        dz = geom.z_points - atom.position[2]
        r = np.sqrt(dx * dx + dy * dy + dz * dz)
        fields, blocks = [], []
        for ch in d.channels:
            npj = len(ch.h)
            if npj == 0:
                continue
            ms = (0,) if ch.l == 0 else (-1, 0, 1)
            for m in ms:
                for i in range(npj):
                    fields.append(radial_projector(r, ch.l, i, ch.radius) * real_harmonic(ch.l, m, dx, dy, dz, r))
                blocks.append(np.asarray(ch.h, dtype=np.float64))
        if fields:
            nproj = len(fields)
            H = np.zeros((nproj, nproj))
            o = 0
            for b in blocks:
                k = b.shape[0]
                H[o:o + k, o:o + k] = b
                o += k
            out.append((ia, np.asarray(fields), H))
    return out


def nonlocal_action(psi, case, geom):
    """pseudo/nonlocal.py:209-321: sum_A sum_ij |p_i> h_ij <p_j|psi>_W."""
    total = np.zeros(psi.shape)
    w = geom.cell_volumes
    for _, fields, H in projector_fields(case, geom):
        coeffs = np.array([np.sum(f * psi * w) for f in fields])
        res = H @ coeffs
        for f, c in zip(fields, res):
            total += c * f
    return total


# ---------------------------------------------------------------- near-core patch


def patch_local_correction_field(
    case, geom, rho, *, patch_radius_scale=1.0, patch_shape=(25, 25, 25),
    correction_strength=1.0, neighbors=8,
):
    """Frozen-density patch embedding, pseudo/local.py:379-703.

    Returns (corrected_total_local_potential, correction_field, total_patch_correction).
    """
    from scipy.spatial import cKDTree

    base_parts = []
    for atom in case.atoms:
        d = GTH_PADE[atom.element]
        dx = geom.x_points - atom.position[0]
        dy = geom.y_points - atom.position[1]
        dz = geom.z_points - atom.position[2]
        base_parts.append(local_potential_radial(np.sqrt(dx * dx + dy * dy + dz * dz), d))
    base_total = np.sum(base_parts, axis=0)
    cloud = np.column_stack((geom.x_points.reshape(-1), geom.y_points.reshape(-1), geom.z_points.reshape(-1)))
    tree = cKDTree(cloud)
    values = rho.reshape(-1)
    corr_field = np.zeros(rho.shape)
    total_corr = 0.0
    for ia, atom in enumerate(case.atoms):
        d = GTH_PADE[atom.element]
        R = geom.patch_radii[ia] * patch_radius_scale
        axes = [np.linspace(c - R, c + R, n) for c, n in zip(atom.position, patch_shape)]
        cellv = float(axes[0][1] - axes[0][0]) * float(axes[1][1] - axes[1][0]) * float(axes[2][1] - axes[2][0])
        PX, PY, PZ = np.meshgrid(*axes, indexing="ij")
        pr = np.sqrt((PX - atom.position[0]) ** 2 + (PY - atom.position[1]) ** 2 + (PZ - atom.position[2]) ** 2)
        pmask = pr <= R
        pts = np.column_stack((PX[pmask], PY[pmask], PZ[pmask]))
        dist, idx = tree.query(pts, k=neighbors)
        if neighbors == 1:
            dist, idx = dist[:, None], idx[:, None]
        interp = np.empty(pts.shape[0])
        zero = dist[:, 0] <= 1.0e-12
        interp[zero] = values[idx[zero, 0]]
        nz = ~zero
        wts = 1.0 / np.maximum(dist[nz], 1.0e-12) ** 2
        wts /= np.sum(wts, axis=1, keepdims=True)
        interp[nz] = np.sum(values[idx[nz]] * wts, axis=1)
        patch_energy = float(np.sum(interp * local_potential_radial(pr[pmask], d)) * cellv)
        rm = np.sqrt((geom.x_points - atom.position[0]) ** 2 + (geom.y_points - atom.position[1]) ** 2
                     + (geom.z_points - atom.position[2]) ** 2)
        mmask = rm <= R
        main_energy = float(np.sum(rho[mmask] * base_parts[ia][mmask] * geom.cell_volumes[mmask]))
        ce = correction_strength * (patch_energy - main_energy)
        total_corr += ce
        if not np.any(mmask) or abs(ce) <= 1.0e-16:
            continue
        if abs(main_energy) > 1.0e-14:
            corr_field[mmask] += (ce / main_energy) * base_parts[ia][mmask]
        else:
            dw = float(np.sum(rho[mmask] * geom.cell_volumes[mmask]))
            if dw > 1.0e-16:
                corr_field[mmask] += ce / dw
    return base_total + corr_field, corr_field, total_corr

"""Adjoints behind torch.autograd.Function, checked against the forward path.

Bars: the H psi adjoint is the exact Euclidean transpose (<g, H psi> = <H^T g, psi> to
1e-12 relative); Ritz-value gradients (dE/dv_ext) and Hellmann-Feynman forces agree with
central finite differences of the FP64 forward path (potential step 1e-4, nuclear step
1e-3 Bohr) to 1e-5 relative.
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import autograd, configs, eigensolver
from isogriddft_amd.device import DeviceGrid
from oracle import grid as ogrid

pytestmark = pytest.mark.gpu


def _smooth(g, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, 3)
    return np.exp(-0.4 * ((g.x_points - c[0]) ** 2 + (g.y_points - c[1]) ** 2 + (g.z_points - c[2]) ** 2)) \
        * (1 + 0.2 * np.sin(g.x_points + 2 * g.z_points))


@pytest.mark.parametrize("kind,closure", [("monitor", "production"), ("monitor", "trial_fix"),
                                          ("legacy", "production")])
def test_hpsi_adjoint_is_the_transpose(kind, closure):
    if kind == "monitor":
        g = configs.build_synthetic_geometry(shape=(20, 22, 24))
    else:
        g = ogrid.build_legacy_geometry(shape=(19, 21, 23), half_extents=(6.0, 6.0, 7.0), stretch=(2.5, 2.5, 2.8))
    dg = DeviceGrid(g)
    rng = np.random.default_rng(2)
    psi = torch.from_numpy(np.stack([_smooth(g, 1), _smooth(g, 2)])).cuda().requires_grad_(True)
    v = torch.from_numpy(0.3 * np.cos(g.y_points)).cuda().requires_grad_(True)
    gw = torch.from_numpy(rng.standard_normal((2,) + dg.shape)).cuda()
    L = (gw * autograd.hpsi(psi, v, dg, closure)).sum()
    dpsi, dv = torch.autograd.grad(L, (psi, v))
    # L is linear in psi and in v: <dL/dpsi, psi> = L, <dL/dv, v> = <g, v psi>
    assert abs((dpsi * psi).sum().item() - L.item()) <= 1e-12 * abs(L.item()) + 1e-12 * gw.abs().sum().item()
    Lv = (gw * v * psi).sum().item()
    assert abs((dv * v).sum().item() - Lv) <= 1e-12 * abs(Lv) + 1e-14


def test_ritz_value_gradient_matches_finite_differences():
    g = configs.build_synthetic_geometry(shape=(14, 15, 16))
    dg = DeviceGrid(g)
    # no symmetry: distinct, non-degenerate lowest eigenvalues (degenerate levels have no
    # individual derivative)
    v0 = -2.0 * np.exp(-0.5 * ((g.x_points - 0.3) ** 2 + 1.3 * g.y_points ** 2 + 0.8 * (g.z_points + 0.2) ** 2)) \
        - 1.2 * np.exp(-0.7 * ((g.x_points + 0.9) ** 2 + (g.y_points - 0.6) ** 2 + (g.z_points - 0.8) ** 2))
    dv = 0.5 * np.exp(-0.3 * ((g.x_points - 0.5) ** 2 + g.y_points ** 2 + (g.z_points + 0.4) ** 2))
    k = 3
    X0 = eigensolver.default_guess_block(g, [("H", (0.0, 0.0, -0.7)), ("H", (0.0, 0.0, 0.7))],
                                         eigensolver.reference_block_size(k), dg)
    v = torch.from_numpy(v0.reshape(-1)).cuda().requires_grad_(True)
    eps = autograd.ritz_values(v, dg, X0, k, tol=1e-10)
    wgt = torch.tensor([1.0, 0.5, 0.25], dtype=torch.float64, device="cuda")
    (grad,) = torch.autograd.grad((wgt * eps).sum(), v)
    analytic = float((grad.cpu().numpy() * dv.reshape(-1)).sum())
    h = 1e-4
    ep = autograd.ritz_values(torch.from_numpy((v0 + h * dv).reshape(-1)).cuda(), dg, X0, k, tol=1e-10)
    em = autograd.ritz_values(torch.from_numpy((v0 - h * dv).reshape(-1)).cuda(), dg, X0, k, tol=1e-10)
    fd = float((wgt * (ep - em)).sum()) / (2 * h)
    print("dE/dv . dv", analytic, fd, "eigenvalues", eps.detach().cpu().numpy())
    gaps = np.diff(eps.detach().cpu().numpy())
    assert gaps.min() > 1e-2  # the check needs separated levels
    assert abs(analytic - fd) <= 1e-5 * abs(fd)


def test_forces_match_finite_differences():
    g = configs.build_synthetic_geometry(shape=(26, 28, 30))
    dg = DeviceGrid(g)
    elements = ["O", "H", "H"]
    R0 = np.array([[0.1, -0.05, 0.2], [-1.2, 0.1, 0.9], [1.3, -0.2, 0.8]])
    orbs = np.stack([_smooth(g, 5), _smooth(g, 6)])
    orbs /= np.sqrt((orbs * orbs * g.cell_volumes).reshape(2, -1).sum(1))[:, None, None, None]
    occ = (2.0, 2.0)
    rho = torch.from_numpy((2.0 * orbs[0] ** 2 + 2.0 * orbs[1] ** 2).reshape(-1)).cuda()
    X = torch.from_numpy(orbs).cuda()
    F = autograd.atomic_forces(dg, elements, R0, rho, X, occ)
    h = 1e-3
    for a in range(3):
        for ax in range(3):
            Rp, Rm = R0.copy(), R0.copy()
            Rp[a, ax] += h
            Rm[a, ax] -= h
            ep = autograd.ionic_energy(torch.from_numpy(Rp), dg, elements, rho, X, occ).item()
            em = autograd.ionic_energy(torch.from_numpy(Rm), dg, elements, rho, X, occ).item()
            fd = -(ep - em) / (2 * h)
            print(a, ax, F[a, ax], fd)
            assert abs(F[a, ax] - fd) <= 1e-5 * max(abs(fd), 1.0), (a, ax, F[a, ax], fd)


@pytest.mark.parametrize("occupations", [(2.0, 2.0, 0.0, 0.0), (2.0, 1.0, 1.0, 0.0)])
def test_density_response_matches_finite_differences(occupations):
    """d/dv of L = sum_p g_p rho_p[v] (rho = sum_i f_i psi_i^2 of the lowest states of
    T + v) from the Sternheimer adjoint, against central differences of re-solved
    eigenproblems; the second case has unequal occupations (occupied-pair terms)."""
    g = configs.build_synthetic_geometry(shape=(14, 15, 16))
    dg = DeviceGrid(g)
    v0 = -2.0 * np.exp(-0.5 * ((g.x_points - 0.3) ** 2 + 1.3 * g.y_points ** 2 + 0.8 * (g.z_points + 0.2) ** 2)) \
        - 1.2 * np.exp(-0.7 * ((g.x_points + 0.9) ** 2 + (g.y_points - 0.6) ** 2 + (g.z_points - 0.8) ** 2))
    dv = 0.5 * np.exp(-0.3 * ((g.x_points - 0.5) ** 2 + g.y_points ** 2 + (g.z_points + 0.4) ** 2))
    gl = np.cos(0.4 * g.x_points) * np.exp(-0.1 * (g.y_points ** 2 + g.z_points ** 2)) + 0.3 * g.z_points
    k = 4
    X0 = eigensolver.default_guess_block(g, [("H", (0.0, 0.0, -0.7)), ("H", (0.0, 0.0, 0.7))],
                                         eigensolver.reference_block_size(k), dg)
    gt = torch.from_numpy(gl.reshape(-1)).cuda()
    v = torch.from_numpy(v0.reshape(-1)).cuda().requires_grad_(True)
    rho = autograd.orbital_density(v, dg, X0, k, occupations, tol=1e-10)
    (grad,) = torch.autograd.grad((gt * rho).sum(), v)
    analytic = float((grad.cpu().numpy() * dv.reshape(-1)).sum())
    h = 1e-4

    def loss(vv):
        r = autograd.orbital_density(torch.from_numpy(vv.reshape(-1)).cuda(), dg, X0, k, occupations, tol=1e-10)
        return float((gt * r).sum())

    fd = (loss(v0 + h * dv) - loss(v0 - h * dv)) / (2 * h)
    print(occupations, "dL/dv . dv", analytic, fd)
    assert abs(analytic - fd) <= 1e-5 * abs(fd)
    # the charge is conserved to first order: d(sum w rho)/dv = 0
    w = dg.operator_weights()
    (gq,) = torch.autograd.grad((w * autograd.orbital_density(v, dg, X0, k, occupations, tol=1e-10)).sum(), v)
    assert float(gq.abs().max()) <= 1e-8 * float(grad.abs().max())


def test_composite_energy_gradient_16_random_directions():
    """L(v) = sum_i f_i eps_i[v] + E_H[rho[v]] through three adjoints chained by autograd
    (Ritz values, the Sternheimer density response, the adjoint Poisson solve), against
    central differences along 16 seeded random smooth directions."""
    g = configs.build_synthetic_geometry(shape=(14, 15, 16))
    dg = DeviceGrid(g)
    v0 = -2.0 * np.exp(-0.5 * ((g.x_points - 0.3) ** 2 + 1.3 * g.y_points ** 2 + 0.8 * (g.z_points + 0.2) ** 2)) \
        - 1.2 * np.exp(-0.7 * ((g.x_points + 0.9) ** 2 + (g.y_points - 0.6) ** 2 + (g.z_points - 0.8) ** 2))
    k, occ = 4, (2.0, 2.0, 0.0, 0.0)
    f = torch.tensor(occ, dtype=torch.float64, device="cuda")
    X0 = eigensolver.default_guess_block(g, [("H", (0.0, 0.0, -0.7)), ("H", (0.0, 0.0, 0.7))],
                                         eigensolver.reference_block_size(k), dg)

    def L(v):
        eps = autograd.ritz_values(v, dg, X0, k, tol=1e-10)
        rho = autograd.orbital_density(v, dg, X0, k, occ, tol=1e-10)
        return (f * eps).sum() + autograd.hartree_energy(rho, dg, tol=1e-11)

    v = torch.from_numpy(v0.reshape(-1)).cuda().requires_grad_(True)
    (grad,) = torch.autograd.grad(L(v), v)
    grad = grad.cpu().numpy()
    rng = np.random.default_rng(2024)
    h = 1e-4
    worst = 0.0
    for j in range(16):
        dv = np.zeros(g.x_points.shape)
        for _ in range(3):
            c = rng.uniform(-2.0, 2.0, 3)
            a = rng.uniform(0.2, 1.0)
            dv += rng.normal() * np.exp(-a * ((g.x_points - c[0]) ** 2 + (g.y_points - c[1]) ** 2
                                               + (g.z_points - c[2]) ** 2))
        analytic = float((grad * dv.reshape(-1)).sum())
        lp = L(torch.from_numpy((v0 + h * dv).reshape(-1)).cuda()).item()
        lm = L(torch.from_numpy((v0 - h * dv).reshape(-1)).cuda()).item()
        fd = (lp - lm) / (2 * h)
        rel = abs(analytic - fd) / max(abs(fd), 1e-8)
        worst = max(worst, rel)
        print(j, analytic, fd, rel)
        assert abs(analytic - fd) <= 1e-5 * max(abs(fd), 1e-3), (j, analytic, fd)
    print("worst relative difference over 16 directions", worst)

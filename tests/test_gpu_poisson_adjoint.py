"""Adjoint of the open-boundary Poisson map (rho -> v_H) on the GPU.

Checks: the moment adjoint is the exact transpose of the HIP moment reductions
(<M^T q, rho> = q . M rho, 1e-12); the Poisson adjoint passes the dot-product test
<P^T g, d> = <g, P d> to the solver tolerance (1e-9 relative); dE_H / d rho from
torch.autograd matches central finite differences of E_H (exactly quadratic: 1e-7).
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import autograd, configs, poisson
from isogriddft_amd.device import DeviceGrid
from oracle import grid as ogrid

pytestmark = pytest.mark.gpu


def _grids(small_monitor):
    yield "monitor19x19x23", small_monitor
    yield "synthetic16x18x20", configs.build_synthetic_geometry(shape=(16, 18, 20))
    yield "legacy15x17x19", ogrid.build_legacy_geometry(shape=(15, 17, 19), half_extents=(6.0, 6.0, 7.0),
                                                        stretch=(2.5, 2.5, 2.8))


def _smooth(g, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, size=3)
    r2 = (g.x_points - c[0]) ** 2 + (g.y_points - c[1]) ** 2 + (g.z_points - c[2]) ** 2
    return np.exp(-0.6 * r2) * (1.0 + 0.3 * rng.standard_normal(g.x_points.shape))


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4])
def test_moment_adjoint_is_the_transpose(small_monitor, order):
    for name, g in _grids(small_monitor):
        dg = DeviceGrid(g)
        rho = torch.from_numpy(_smooth(g, order).reshape(-1)).cuda()
        q = np.random.default_rng(7 + order).standard_normal(poisson.nmono(order))
        _, moms, _ = poisson.multipole_boundary_device(dg, rho, order)
        lhs = float((poisson.moments_adjoint_device(dg, q, order) * rho).sum())
        rhs = float(np.dot(q, moms.cpu().numpy()))
        print(name, order, lhs, rhs)
        assert abs(lhs - rhs) <= 1e-12 * (abs(rhs) + float(np.abs(q) @ np.abs(moms.cpu().numpy()))), name


def test_poisson_adjoint_dot_product(small_monitor):
    for name, g in _grids(small_monitor):
        dg = DeviceGrid(g)
        d = torch.from_numpy(_smooth(g, 1).reshape(-1)).cuda()
        gv = torch.from_numpy(_smooth(g, 2).reshape(-1)).cuda()
        sol = poisson.solve_poisson_device(dg, d, tolerance=1e-12, max_iterations=6000)
        pt, info = poisson.poisson_adjoint_device(dg, gv, tolerance=1e-12, max_iterations=6000)
        assert sol["converged"] and info["converged"], (name, info)
        a = float((gv * sol["potential"].reshape(-1)).sum())
        b = float((pt * d).sum())
        scale = float((gv.abs() * sol["potential"].reshape(-1).abs()).sum())
        print(name, a, b, info)
        assert abs(a - b) <= 1e-9 * scale, (name, a, b)


def test_hartree_energy_gradient_matches_finite_differences(small_monitor):
    for name, g in list(_grids(small_monitor))[:2]:
        dg = DeviceGrid(g)
        rho = torch.from_numpy(np.abs(_smooth(g, 3)).reshape(-1)).cuda().requires_grad_(True)
        e = autograd.hartree_energy(rho, dg, tol=1e-12)
        (grad,) = torch.autograd.grad(e, rho)
        d = torch.from_numpy(_smooth(g, 4).reshape(-1)).cuda()
        eps = 1e-3
        with torch.no_grad():
            ep = float(autograd.hartree_energy(rho + eps * d, dg, tol=1e-12))
            em = float(autograd.hartree_energy(rho - eps * d, dg, tol=1e-12))
        fd = (ep - em) / (2 * eps)
        an = float((grad * d).sum())
        print(name, fd, an)
        assert abs(fd - an) <= 1e-7 * abs(an), (name, fd, an)


def test_legacy_split_has_no_adjoint(small_monitor):
    dg = DeviceGrid(small_monitor)
    with pytest.raises(NotImplementedError):
        poisson.poisson_adjoint_device(dg, torch.ones(dg.N, dtype=torch.float64, device="cuda"),
                                       mode="legacy_split")

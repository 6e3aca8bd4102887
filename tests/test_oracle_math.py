"""Pin the oracle by mathematics: closed forms, invariants and brute force.

Each check is chosen so that a plausible slip (dropped term, wrong sign or index,
transposed operand) fails it; none re-types the oracle's own formula.
"""

import itertools

import numpy as np
import pytest

from oracle import eigen, grid, kinetic, poisson, xc


def _affine_geometry(shape, A, b=(0.0, 0.0, 0.0)):
    nx, ny, nz = shape
    lx, ly, lz = np.linspace(0, 1, nx), np.linspace(0, 1, ny), np.linspace(0, 1, nz)
    U = np.stack(np.meshgrid(lx, ly, lz, indexing="ij"), -1)
    coords = U @ np.asarray(A, dtype=np.float64).T + np.asarray(b)
    return grid.monitor_geometry_from_coordinates(coords, ((-1, 1), (-1, 1), (-1, 1)))


def _dense(op, shape):
    n = int(np.prod(shape))
    M = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        M[:, i] = op(e.reshape(shape)).reshape(-1)
    return M


def test_trial_fix_is_weighted_self_adjoint(small_monitor):
    g = small_monitor
    rng = np.random.default_rng(0)
    u, v = rng.normal(size=(2,) + g.jacobian.shape)
    w = g.cell_volumes
    a = np.sum(u * kinetic.monitor_kinetic(v, g, "trial_fix") * w)
    b = np.sum(kinetic.monitor_kinetic(u, g, "trial_fix") * v * w)
    assert abs(a - b) <= 1e-12 * (abs(a) + abs(b))
    # the production closure is not (the discrete Green-identity defect)
    a = np.sum(u * kinetic.monitor_kinetic(v, g, "production") * w)
    b = np.sum(kinetic.monitor_kinetic(u, g, "production") * v * w)
    assert abs(a - b) > 1e-6 * (abs(a) + abs(b))


def test_trial_fix_closed_form_spectrum():
    shape = (5, 6, 7)
    L = np.diag([3.0, 4.0, 5.0])
    g = _affine_geometry(shape, L)
    M = _dense(lambda f: kinetic.monitor_kinetic(f, g, "trial_fix"), shape)
    ev = np.sort(np.linalg.eigvals(M).real)
    expect = []
    for j in itertools.product(*[range(1, n + 1) for n in shape]):
        s = 0.0
        for a, (ja, n) in enumerate(zip(j, shape)):
            h = 1.0 / (n - 1)
            s += np.cos(ja * np.pi / (n + 1)) ** 2 / (h * h * L[a, a] ** 2)
        expect.append(0.5 * s)
    np.testing.assert_allclose(ev, np.sort(expect), rtol=1e-10, atol=1e-10)


def test_production_exact_on_quadratics_sheared_map():
    A = np.array([[4.0, 0.7, -0.3], [0.2, 5.0, 0.4], [-0.5, 0.3, 6.0]])
    g = _affine_geometry((7, 8, 9), A, b=(0.3, -0.2, 0.1))
    X, Y, Z = g.x_points, g.y_points, g.z_points
    f = 1.3 * X * X - 0.7 * Y * Y + 0.4 * Z * Z + 0.9 * X * Y - 0.6 * Y * Z + 0.25 * X * Z + X - 2 * Z
    lap = kinetic.monitor_laplacian(f, g, "production")
    np.testing.assert_allclose(lap, 2 * 1.3 - 2 * 0.7 + 2 * 0.4, rtol=0, atol=1e-10)


@pytest.mark.parametrize("closure", ["production", "trial_fix"])
def test_monitor_transpose_is_exact(small_monitor, closure):
    rng = np.random.default_rng(1)
    g = grid.monitor_geometry_from_coordinates(
        np.stack([small_monitor.x_points, small_monitor.y_points, small_monitor.z_points], -1)[2:8, 3:10, 4:11],
        small_monitor.box_bounds)
    shape = g.jacobian.shape
    M = _dense(lambda f: kinetic.monitor_laplacian(f, g, closure), shape)
    y = rng.normal(size=shape)
    np.testing.assert_allclose(kinetic.monitor_laplacian_transpose(y, g, closure).reshape(-1),
                               M.T @ y.reshape(-1), rtol=1e-11, atol=1e-9)
    Ms = kinetic.monitor_laplacian_matrix(g, closure).toarray()
    np.testing.assert_allclose(Ms, M, rtol=1e-12, atol=1e-9)


def test_legacy_transpose_and_constant_field():
    g = grid.build_legacy_geometry(shape=(6, 7, 8))
    shape = g.cell_volumes.shape
    M = _dense(lambda f: kinetic.legacy_laplacian(f, g), shape)
    y = np.random.default_rng(2).normal(size=shape)
    np.testing.assert_allclose(kinetic.legacy_laplacian_transpose(y, g).reshape(-1), M.T @ y.reshape(-1),
                               rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(kinetic.legacy_laplacian_matrix(g).toarray(), M, rtol=1e-12, atol=1e-10)
    # constant field: zero interior action (reference tests/test_local_hamiltonian.py:59-64)
    act = kinetic.legacy_kinetic(np.ones(shape), g)
    assert np.max(np.abs(act[1:-1, 1:-1, 1:-1])) < 1e-12


def test_lsda_potentials_are_energy_derivatives():
    rng = np.random.default_rng(3)
    a = rng.uniform(1e-4, 2.0, 200)
    b = rng.uniform(1e-4, 2.0, 200)
    _, vu, vd = xc.lsda_pw92(a, b)
    h = 1e-6
    e = lambda x, y: xc.lsda_energy_density(x, y)  # noqa: E731
    du = (e(a + h * a, b) - e(a - h * a, b)) / (2 * h * a)
    dd = (e(a, b + h * b) - e(a, b - h * b)) / (2 * h * b)
    np.testing.assert_allclose(vu, du, rtol=2e-7, atol=1e-9)
    np.testing.assert_allclose(vd, dd, rtol=2e-7, atol=1e-9)


def test_lsda_closed_forms():
    # unpolarised exchange: eps_x = -(3/4)(3/pi)^(1/3) rho^(1/3)
    rho = np.array([0.01, 0.3, 2.0])
    eps, vu, vd = xc.lsda_pw92(rho / 2, rho / 2)
    rs = (3 / (4 * np.pi * rho)) ** (1 / 3)
    # PW92 unpolarised correlation at rs (paper form, independent evaluation)
    A, a1, b1, b2, b3, b4 = 0.031091, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294
    ec = -2 * A * (1 + a1 * rs) * np.log(1 + 1 / (2 * A * (b1 * rs ** 0.5 + b2 * rs + b3 * rs ** 1.5 + b4 * rs ** 2)))
    ex = -0.75 * (3 / np.pi) ** (1 / 3) * rho ** (1 / 3)
    np.testing.assert_allclose(eps, ex + ec, rtol=1e-12)
    np.testing.assert_allclose(vu, vd, rtol=1e-13)
    # fully polarised: exchange is 2^(1/3) times the unpolarised value at the same rho
    eps_p, _, vdn = xc.lsda_pw92(rho, 0 * rho)
    A, a1, b1, b2, b3, b4 = 0.015545, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517
    ec1 = -2 * A * (1 + a1 * rs) * np.log(1 + 1 / (2 * A * (b1 * rs ** 0.5 + b2 * rs + b3 * rs ** 1.5 + b4 * rs ** 2)))
    np.testing.assert_allclose(eps_p, 2 ** (1 / 3) * ex + ec1, rtol=1e-12)
    with pytest.raises(ValueError):
        xc.lsda_pw92(np.array([-1e-12]), np.array([0.0]))


def test_multipole_expansion_converges_to_point_charge():
    r0 = np.array([0.3, -0.2, 0.45])
    q = np.array([1.7])
    pts = np.random.default_rng(4).normal(size=(3, 50))
    pts = 6.0 * pts / np.linalg.norm(pts, axis=0)
    exact = q[0] / np.linalg.norm(pts - r0[:, None], axis=0)
    errs = []
    for order in range(5):
        m = poisson.raw_moments(q, np.array([r0[0]]), np.array([r0[1]]), np.array([r0[2]]), order)
        phi = poisson.boundary_potential(m, pts[0], pts[1], pts[2], order)
        errs.append(np.max(np.abs(phi - exact)))
    ratio = np.linalg.norm(r0) / 6.0
    for order, e in enumerate(errs):
        assert e < 3.0 * q[0] / 6.0 * ratio ** (order + 1)
    assert all(errs[i + 1] < errs[i] for i in range(4))


def test_direct_and_iterative_poisson_agree(small_monitor, golden):
    rho = golden["rho_shifted"]
    pd, _, _, _ = poisson.solve_poisson(small_monitor, rho, method="direct")
    pi, _, _, info = poisson.solve_poisson(small_monitor, rho, method="iter", tol=1e-13)
    assert np.max(np.abs(pd - pi)) < 1e-10 * np.max(np.abs(pd))
    # far field tracks the total charge; centre potential above the corner (reference test :172-181)
    assert pd[9, 9, 11] > pd[0, 0, 0] > 0


def test_block_iteration_matches_dense_eigh():
    # trial_fix is exactly W-self-adjoint, so the W-similar operator is symmetric
    A = np.array([[8.0, 0.5, 0.0], [0.3, 9.0, 0.2], [0.0, -0.4, 10.0]])
    g = _affine_geometry((9, 9, 11), A, b=(-4.0, -4.5, -5.0))
    w = g.cell_volumes.reshape(-1)
    V = -2.0 * np.exp(-0.5 * (g.x_points ** 2 + g.y_points ** 2 + g.z_points ** 2)).reshape(-1)
    shape = g.cell_volumes.shape

    def apply(block):
        return np.stack([kinetic.monitor_kinetic(b.reshape(shape), g, "trial_fix").reshape(-1) + V * b
                         for b in block])

    H = apply(np.eye(w.size)).T
    sw = np.sqrt(w)
    S = (sw[:, None] * H) / sw[None, :]
    assert np.max(np.abs(S - S.T)) < 1e-9 * np.max(np.abs(S))
    ev_dense = np.linalg.eigvalsh(0.5 * (S + S.T))[:3]
    X0 = eigen.default_guesses(g, [(0, 0, -0.7), (0, 0, 0.7)], 12, (0, 0, 0)).reshape(12, -1)
    ev, Psi, rn, it, ok = eigen.block_subspace_iteration(apply, w, X0, 3, 1e-9, 400)
    assert ok
    np.testing.assert_allclose(ev, ev_dense, rtol=0, atol=1e-10)
    np.testing.assert_allclose(eigen.gram(Psi, Psi, w), np.eye(3), atol=1e-10)


def test_vwn_potentials_are_energy_derivatives_and_limits():
    """Slater + VWN5: v_sigma = d(rho eps)/d rho_sigma (central differences, polarised
    densities), eps at z = 1 is the ferromagnetic fit, eps at z = 0 the paramagnetic one."""
    rng = np.random.default_rng(4)
    a = rng.uniform(1e-4, 2.0, 200)
    b = rng.uniform(1e-4, 2.0, 200)
    _, vu, vd = xc.lsda_vwn(a, b)
    h = 1e-6
    e = lambda x, y: (x + y) * xc.lsda_vwn(x, y)[0]  # noqa: E731
    du = (e(a + h * a, b) - e(a - h * a, b)) / (2 * h * a)
    dd = (e(a, b + h * b) - e(a, b - h * b)) / (2 * h * b)
    np.testing.assert_allclose(vu, du, rtol=2e-7, atol=1e-9)
    np.testing.assert_allclose(vd, dd, rtol=2e-7, atol=1e-9)
    rho = np.array([0.01, 0.3, 2.0])
    rs = (3.0 / (4.0 * np.pi * rho)) ** (1.0 / 3.0)
    x = np.sqrt(rs)
    ex_p = -0.75 * (3.0 / np.pi) ** (1.0 / 3.0) * np.cbrt(rho)
    ex_f = 2.0 ** (1.0 / 3.0) * ex_p
    eps0, _, _ = xc.lsda_vwn(rho / 2, rho / 2)
    np.testing.assert_allclose(eps0, ex_p + xc._vwn_g(x, xc.VWN_P)[0], rtol=1e-13)
    eps1, _, _ = xc.lsda_vwn(rho, 0 * rho)
    np.testing.assert_allclose(eps1, ex_f + xc._vwn_g(x, xc.VWN_F)[0], rtol=1e-13)
    # the paramagnetic fit reproduces the VWN paper's ec(rs = 1) ~ -0.0600 Ha (Ceperley-Alder)
    assert abs(xc._vwn_g(np.array([1.0]), xc.VWN_P)[0][0] - (-0.0600)) < 5e-4


def _eps_c(fn, rs, z):
    """Correlation energy per electron of an LSDA oracle at (rs, zeta): the total minus the
    exact Slater exchange of the two spin densities."""
    rho = 3.0 / (4.0 * np.pi * rs ** 3)
    a, b = rho * (1.0 + z) / 2.0, rho * (1.0 - z) / 2.0
    eps = fn(np.atleast_1d(a), np.atleast_1d(b))[0]
    ex = -0.75 * (6.0 / np.pi) ** (1.0 / 3.0) * (a ** (4.0 / 3.0) + b ** (4.0 / 3.0)) / rho
    return eps - ex


@pytest.mark.parametrize("fn", [xc.lsda_pw92, xc.lsda_vwn])
def test_spin_stiffness_high_density_limit_and_spin_symmetry(fn):
    """The spin-interpolation term of both functionals: the spin stiffness
    alpha_c(rs) = d^2 eps_c / d zeta^2 at zeta = 0 (central differences in zeta) follows the
    exact high-density (RPA) law alpha_c ~ -(1 / (6 pi^2)) ln rs + const, so its slope in
    ln rs tends to -1/(6 pi^2); a dropped or sign-flipped stiffness term, or a wrong f''(0),
    breaks it.  eps_c is even in zeta."""
    d = 1e-3
    rss = np.array([1e-5, 1e-4])
    alpha = [float((_eps_c(fn, rs, d) - 2.0 * _eps_c(fn, rs, 0.0) + _eps_c(fn, rs, -d))[0] / d ** 2) for rs in rss]
    assert alpha[0] > alpha[1] > 0.0
    slope = (alpha[1] - alpha[0]) / np.log(10.0)
    assert abs(slope + 1.0 / (6.0 * np.pi ** 2)) <= 5e-4 / (6.0 * np.pi ** 2), slope
    for rs in (0.5, 2.0, 7.0):
        for z in (0.2, 0.55, 0.9):
            assert abs(_eps_c(fn, rs, z)[0] - _eps_c(fn, rs, -z)[0]) <= 1e-14


def test_vwn_and_pw92_spin_interpolations_agree():
    """VWN5 and PW92 are independent fits of the same uniform-gas correlation energy
    (Ceperley-Alder and its spin dependence); at intermediate polarisations they agree to
    better than 1 mHa for 0.5 <= rs <= 10, while the spin-stiffness term itself is several
    mHa there, so a slip in either interpolation shows."""
    worst = 0.0
    for rs in np.linspace(0.5, 10.0, 20):
        for z in (0.3, 0.6, 0.9):
            worst = max(worst, abs(_eps_c(xc.lsda_vwn, rs, z)[0] - _eps_c(xc.lsda_pw92, rs, z)[0]))
    assert worst <= 1e-3, worst

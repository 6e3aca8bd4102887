"""Host-side algebra of the open-boundary solve (no GPU): the monomial table and the
homogeneous boundary coefficients against the oracle's free-space expansion."""

import numpy as np

from isogriddft_amd import poisson
from oracle import poisson as opoisson


def _flat_from_points(q, pts, order):
    table = poisson.monomial_exponents(order)
    return np.array([np.sum(q * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c) for a, b, c in table])


def test_monomial_table_layout():
    t = poisson.monomial_exponents(4)
    assert len(t) == poisson.nmono(4) == 35
    assert t[:10] == [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (1, 1, 0), (1, 0, 1),
                      (0, 2, 0), (0, 1, 1), (0, 0, 2)]
    assert t[20] == (4, 0, 0) and t[23] == (2, 2, 0) and t[34] == (0, 0, 4)


def test_boundary_coefficients_match_oracle_expansion():
    rng = np.random.default_rng(3)
    pts = rng.normal(scale=0.6, size=(7, 3))
    q = rng.uniform(0.2, 1.0, size=7)
    targets = rng.normal(size=(50, 3))
    targets *= 9.0 / np.linalg.norm(targets, axis=1, keepdims=True)
    errs = []
    for order in range(5):
        flat = _flat_from_points(q, pts, order)
        coef = poisson.boundary_coefficients(flat, order)
        table = poisson.monomial_exponents(order)
        r = np.linalg.norm(targets, axis=1)
        phi = np.zeros(len(targets))
        for c, (a, b, d) in zip(coef, table):
            phi += c * targets[:, 0] ** a * targets[:, 1] ** b * targets[:, 2] ** d / r ** (2 * (a + b + d) + 1)
        moms = opoisson.raw_moments(q, pts[:, 0], pts[:, 1], pts[:, 2], order)
        ref = np.array([opoisson.boundary_potential(moms, *t, order) for t in targets])
        assert np.max(np.abs(phi - ref)) <= 1e-13 * np.max(np.abs(ref))
        # and the expansion converges to the exact point-charge potential with the order
        exact = np.array([np.sum(q / np.linalg.norm(t - pts, axis=1)) for t in targets])
        errs.append(np.max(np.abs(phi - exact)) / np.max(np.abs(exact)))
    assert all(e1 < e0 for e0, e1 in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-4, errs


def test_quadrupole_is_traceless_three_m_minus_trace():
    rng = np.random.default_rng(5)
    pts = rng.normal(size=(5, 3))
    q = rng.uniform(size=5)
    flat = _flat_from_points(q, pts, 2)
    Q = poisson.quadrupole_from_flat(flat)
    M2 = np.einsum("p,pa,pb->ab", q, pts, pts)
    assert np.allclose(Q, opoisson.traceless_quadrupole(M2), rtol=0, atol=1e-13)
    assert abs(np.trace(Q)) < 1e-13


def test_kinetic_diagonal_matches_assembled_operators():
    """The eigensolver's Davidson preconditioner uses diag(Lap): exact for trial_fix and
    legacy grids, and for the interior rows of the production closure."""
    from isogriddft_amd import configs
    from isogriddft_amd.device import DeviceGrid
    from oracle import grid as ogrid
    from oracle import kinetic as okin

    g = configs.build_synthetic_geometry(shape=(10, 11, 12))
    d = DeviceGrid(g, "cpu").kinetic_diagonal().numpy()
    ref = okin.monitor_laplacian_matrix(g, "trial_fix").diagonal()
    assert np.max(np.abs(d - ref)) <= 1e-14 * np.max(np.abs(ref))
    inner = np.zeros(g.x_points.shape, bool)
    inner[2:-2, 2:-2, 2:-2] = True
    refp = okin.monitor_laplacian_matrix(g, "production").diagonal()
    assert np.max(np.abs((d - refp)[inner.ravel()])) <= 1e-14 * np.max(np.abs(refp))
    lg = ogrid.build_legacy_geometry(shape=(9, 10, 11))
    dl = DeviceGrid(lg, "cpu").kinetic_diagonal().numpy()
    assert np.array_equal(dl, okin.legacy_laplacian_matrix(lg).diagonal())

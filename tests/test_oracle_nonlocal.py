"""Pin the oracle's separable GTH nonlocal term (pseudo/nonlocal.py:99-321 of the reference)
by mathematics that does not re-type its formula.

* Normalisation.  The GTH projectors are normalised by construction:
  int_0^inf r^2 p_i^l(r)^2 dr = 1 for every (l, i, r_l).  Evaluated by adaptive quadrature
  (scipy.integrate.quad) of the oracle's radial_projector; a wrong Gamma argument, radial
  power or prefactor breaks it.
* Overlap with an atom-centred Gaussian.  <p_i^0 Y_00 | exp(-a r^2)> has the closed form
  4 pi Y_00 * sqrt2 / (r_l^g sqrt(Gamma(g))) * Gamma(i + 3/2) / (2 beta^(i + 3/2)),
  beta = a + 1 / (2 r_l^2), g = (4i + 3) / 2.  The oracle's projector field on a fine uniform
  grid around an off-centre atom, integrated with the cell volumes, converges to it; a wrong
  atom offset, harmonic or radial factor fails.  The p harmonics are checked the same way:
  <p_0^1 Y_1m | x_n exp(-a r^2)> is the closed form above with one more power of r, times
  delta(m -> n) with m = -1, 0, 1 -> y, z, x.
* Dense brute force on a tiny grid.  The matrix M of nonlocal_action, built column by column
  from unit vectors, is W-self-adjoint (W M symmetric), has rank equal to the number of
  projectors, and for one s-projector atom its only nonzero eigenvalue is h <p, p>_W with
  eigenvector p.  A missing weight, a transposed or squared h, or a dropped projector fails.
"""

from math import gamma, pi, sqrt
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import grid, pseudo


def _uniform_geometry(n, half):
    ax = np.linspace(-half, half, n)
    h = ax[1] - ax[0]
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    return SimpleNamespace(x_points=X, y_points=Y, z_points=Z, cell_volumes=np.full(X.shape, h ** 3))


@pytest.mark.parametrize("l", [0, 1, 2])
@pytest.mark.parametrize("i", [0, 1, 2])
@pytest.mark.parametrize("rl", [0.22178614, 0.30455321, 0.7])
def test_radial_projector_is_normalised(l, i, rl):
    from scipy.integrate import quad

    val, err = quad(lambda r: (r * pseudo.radial_projector(np.array(r), l, i, rl)) ** 2, 0.0, np.inf,
                    epsabs=1e-14, epsrel=1e-13, limit=200)
    assert abs(val - 1.0) < 1e-10, (l, i, rl, val)


def test_table_channels_are_the_gth_pade_values():
    # the active projectors of the configurations: C, N, O s channels with one projector each
    # (pseudo/gth_data.py:54-56, :80-82, :106-108), no p projectors, none for H (:38)
    for el, (rl, h) in {"C": (0.30455321, 9.52284179), "N": (0.25660487, 13.55224272),
                        "O": (0.22178614, 18.26691718)}.items():
        ch = [c for c in pseudo.GTH_PADE[el].channels if len(c.h)]
        assert len(ch) == 1 and ch[0].l == 0 and ch[0].radius == rl and ch[0].h == ((h,),)
    assert pseudo.GTH_PADE["H"].channels == ()


def _gauss_overlap(l, i, rl, a):
    """int r^2 p_i^l(r) r^l exp(-a r^2) dr (closed form)."""
    g = l + (4 * i + 3) / 2.0
    beta = a + 1.0 / (2.0 * rl * rl)
    n = 2 * l + 2 * i + 2  # power of r in the integrand r^2 * r^(l+2i) * r^l
    # int_0^inf r^n exp(-beta r^2) dr = Gamma((n+1)/2) / (2 beta^((n+1)/2))
    return sqrt(2.0) / (rl ** g * sqrt(gamma(g))) * gamma((n + 1) / 2.0) / (2.0 * beta ** ((n + 1) / 2.0))


def test_s_projector_overlap_with_gaussian_converges_to_closed_form():
    rl, a = 0.30455321, 1.3
    centre = (0.11, -0.07, 0.05)
    case = grid.Case("X", (grid.Atom("C", centre),))
    expected = 4.0 * pi * (1.0 / sqrt(4.0 * pi)) * _gauss_overlap(0, 0, rl, a)
    errs = []
    for n in (41, 61, 81):
        g = _uniform_geometry(n, 2.2)
        (_, fields, H), = pseudo.projector_fields(case, g)
        assert fields.shape[0] == 1 and H.shape == (1, 1) and H[0, 0] == 9.52284179
        r2 = (g.x_points - centre[0]) ** 2 + (g.y_points - centre[1]) ** 2 + (g.z_points - centre[2]) ** 2
        val = float(np.sum(fields[0] * np.exp(-a * r2) * g.cell_volumes))
        errs.append(abs(val - expected) / expected)
    assert errs[-1] < 1e-9, errs
    # the oracle's nonlocal action applied to the Gaussian: h <p|g> p, point by point
    g = _uniform_geometry(61, 2.2)
    r2 = (g.x_points - centre[0]) ** 2 + (g.y_points - centre[1]) ** 2 + (g.z_points - centre[2]) ** 2
    act = pseudo.nonlocal_action(np.exp(-a * r2), case, g)
    r = np.sqrt(r2)
    p = pseudo.radial_projector(r, 0, 0, rl) / sqrt(4.0 * pi)
    np.testing.assert_allclose(act, 9.52284179 * expected * p, rtol=1e-8, atol=1e-12)


def test_p_harmonics_match_axes_and_closed_form():
    rl, a = 0.2568289, 0.9
    centre = (-0.05, 0.08, 0.02)
    g = _uniform_geometry(81, 2.0)
    dx, dy, dz = g.x_points - centre[0], g.y_points - centre[1], g.z_points - centre[2]
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    radial = pseudo.radial_projector(r, 1, 0, rl)
    expected = (4.0 * pi / 3.0) * sqrt(3.0 / (4.0 * pi)) * _gauss_overlap(1, 0, rl, a)
    e = np.exp(-a * r * r)
    comps = {"x": dx * e, "y": dy * e, "z": dz * e}
    axis_of = {-1: "y", 0: "z", 1: "x"}
    for m in (-1, 0, 1):
        f = radial * pseudo.real_harmonic(1, m, dx, dy, dz, r)
        for name, c in comps.items():
            val = float(np.sum(f * c * g.cell_volumes))
            if name == axis_of[m]:
                assert abs(val - expected) / expected < 1e-8, (m, name, val, expected)
            else:
                assert abs(val) < 1e-10 * expected, (m, name, val)


def test_dense_nonlocal_matrix_invariants():
    # tiny non-uniform grid (legacy sinh map) with an O atom (one s projector) and an H atom
    # (none) off the grid nodes
    geom = grid.build_legacy_geometry((7, 8, 9), (1.2, 1.1, 1.3), (1.5, 1.4, 1.6),
                                      (0.0, 0.0, 0.0))
    shape = geom.cell_volumes.shape
    n = int(np.prod(shape))
    case = grid.Case("OH", (grid.Atom("O", (0.05, -0.03, 0.1)), grid.Atom("H", (0.2, 0.1, 0.9))))
    M = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        M[:, j] = pseudo.nonlocal_action(e.reshape(shape), case, geom).reshape(-1)
    w = geom.cell_volumes.reshape(-1)
    WM = w[:, None] * M
    np.testing.assert_allclose(WM, WM.T, rtol=0, atol=1e-12 * np.abs(WM).max())
    s = np.linalg.svd(M, compute_uv=False)
    assert s[0] > 0 and s[1] < 1e-12 * s[0]  # rank 1: one projector in total
    # the nonzero eigenvalue is h <p, p>_W, the eigenvector is p
    dx = geom.x_points - 0.05
    dy = geom.y_points + 0.03
    dz = geom.z_points - 0.1
    p = (pseudo.radial_projector(np.sqrt(dx * dx + dy * dy + dz * dz), 0, 0, 0.22178614)
         / sqrt(4.0 * pi)).reshape(-1)
    lam = 18.26691718 * float(np.sum(p * p * w))
    ev = np.linalg.eigvals(M)
    top = ev[np.argmax(np.abs(ev))]
    assert abs(top.imag) < 1e-12 * abs(top) and abs(top.real - lam) < 1e-12 * abs(lam)
    np.testing.assert_allclose(M @ p, lam * p, rtol=0, atol=1e-12 * np.abs(lam * p).max())


def test_two_projector_atoms_add_and_rank_counts_projectors():
    geom = grid.build_legacy_geometry((6, 7, 8), (2.0, 2.0, 2.0), (1.3, 1.3, 1.3),
                                      (0.0, 0.0, 0.0))
    shape = geom.cell_volumes.shape
    n = int(np.prod(shape))
    atoms = (grid.Atom("C", (-0.6, 0.1, 0.0)), grid.Atom("N", (0.7, -0.2, 0.3)), grid.Atom("O", (0.0, 0.5, -0.8)))

    def dense(case):
        M = np.zeros((n, n))
        for j in range(n):
            e = np.zeros(n)
            e[j] = 1.0
            M[:, j] = pseudo.nonlocal_action(e.reshape(shape), case, geom).reshape(-1)
        return M

    Ms = [dense(grid.Case("x", (a,))) for a in atoms]
    Mall = dense(grid.Case("all", atoms))
    np.testing.assert_allclose(Mall, sum(Ms), rtol=0, atol=1e-12 * np.abs(Mall).max())
    s = np.linalg.svd(Mall, compute_uv=False)
    assert s[2] > 1e-8 * s[0] and s[3] < 1e-12 * s[0]  # rank 3

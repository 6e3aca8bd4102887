"""GPU parity of the near-core local-GTH patch (pseudo/local.py:379-703).

Bars: nearest-neighbour index sets bit-exact against a NumPy brute force ordered by
(distance^2, node index), distances 1e-15 relative; patch / main-grid energies and the
frozen field 1e-12 against the oracle (cKDTree neighbours) and the reference-generated
vectors; eigenvalues of the patched operator 1e-10 Ha against the same solve with the
oracle's patched potential passed explicitly.
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import eigensolver, pseudo
from isogriddft_amd.device import DeviceGrid
from oracle import grid as ogrid
from oracle import pseudo as opseudo

pytestmark = pytest.mark.gpu


def _knn_numpy(xyz, q, k):
    d2 = ((xyz[None, :, 0] - q[:, None, 0]) * (xyz[None, :, 0] - q[:, None, 0])
          + (xyz[None, :, 1] - q[:, None, 1]) * (xyz[None, :, 1] - q[:, None, 1])
          + (xyz[None, :, 2] - q[:, None, 2]) * (xyz[None, :, 2] - q[:, None, 2]))
    idx = np.empty((q.shape[0], k), dtype=np.int64)
    for t in range(q.shape[0]):
        order = np.lexsort((np.arange(xyz.shape[0]), d2[t]))
        idx[t] = order[:k]
    return idx, np.sqrt(np.take_along_axis(d2, idx, axis=1))


def _device_knn(dg, q, k):
    from isogriddft_amd import _lib

    lib = _lib.load()
    qd = torch.from_numpy(np.ascontiguousarray(q.T)).cuda()
    nbytes = int(lib.ig_knn_work_bytes(q.shape[0], k))
    work = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    idx = torch.empty((q.shape[0], k), dtype=torch.int32, device="cuda")
    dist = torch.empty((q.shape[0], k), dtype=torch.float64, device="cuda")
    _lib.call("ig_knn_f64", _lib.ptr(dg.xyz()), dg.N, _lib.ptr(qd), q.shape[0], k, _lib.ptr(work), nbytes,
              _lib.ptr(idx), _lib.ptr(dist), _lib.stream_handle())
    return idx.cpu().numpy(), dist.cpu().numpy()


@pytest.mark.parametrize("k", [1, 8, 13, 32])
def test_knn_matches_brute_force_with_ties(k):
    """Uniform grid with binary-fraction spacing: queries on nodes and on midpoints of node
    pairs give exactly tied distances."""
    axes = [0.5 * np.arange(n) - 0.25 * (n - 1) for n in (12, 13, 14)]
    coords = np.stack(np.meshgrid(*axes, indexing="ij"), -1)
    geom = ogrid.monitor_geometry_from_coordinates(coords, tuple((a[0], a[-1]) for a in axes))
    dg = DeviceGrid(geom)
    xyz = np.stack([geom.x_points.ravel(), geom.y_points.ravel(), geom.z_points.ravel()], 1)
    rng = np.random.default_rng(3)
    q = np.concatenate([xyz[rng.integers(0, xyz.shape[0], 40)],
                        0.5 * (xyz[rng.integers(0, xyz.shape[0], 40)] + xyz[rng.integers(0, xyz.shape[0], 40)]),
                        rng.uniform(xyz.min(0) - 1.0, xyz.max(0) + 1.0, size=(300, 3))])
    idx, dist = _device_knn(dg, q, k)
    ridx, rdist = _knn_numpy(xyz, q, k)
    assert np.array_equal(idx, ridx)
    np.testing.assert_allclose(dist, rdist, rtol=1e-15, atol=0)


def test_knn_on_monitor_grid(small_monitor):
    g = small_monitor
    dg = DeviceGrid(g)
    xyz = np.stack([g.x_points.ravel(), g.y_points.ravel(), g.z_points.ravel()], 1)
    rng = np.random.default_rng(5)
    q = rng.uniform(xyz.min(0), xyz.max(0), size=(700, 3))
    idx, dist = _device_knn(dg, q, 8)
    ridx, rdist = _knn_numpy(xyz, q, 8)
    assert np.array_equal(idx, ridx)
    np.testing.assert_allclose(dist, rdist, rtol=1e-15, atol=0)


def test_patch_embedding_matches_reference_vectors(golden, small_monitor):
    params = pseudo.LocalPotentialPatchParameters(patch_radius_scale=0.75, correction_strength=1.3)
    emb = pseudo.evaluate_monitor_grid_local_ionic_potential_with_frozen_patch_field(
        ogrid.H2_CASE, small_monitor, golden["patch_rho"], patch_parameters=params)
    np.testing.assert_allclose(emb.corrected_total_local_potential, golden["patch_corrected_local_potential"],
                               rtol=1e-12, atol=1e-12)
    assert abs(emb.patch_evaluation.total_patch_correction - float(golden["patch_total_correction"])) < 1e-13
    assert abs(emb.embedding_energy_mismatch) < 1e-12


def _axial_density(g):
    """Unequal weight on the two atoms but invariant under x -> -x, y -> -y and x <-> y,
    the mirror symmetries of the H2 monitor grid: patch points on those mirror planes
    have exactly equidistant neighbour pairs, and with this density the choice between
    them does not change the interpolated value."""
    X, Y, Z = g.x_points, g.y_points, g.z_points
    s2 = X * X + Y * Y
    return np.exp(-0.9 * (s2 + (Z - 0.6) ** 2)) + 0.4 * np.exp(-1.7 * (s2 + (Z + 0.8) ** 2))


def _asym_density(g):
    X, Y, Z = g.x_points, g.y_points, g.z_points
    rho = np.exp(-0.9 * ((X - 0.2) ** 2 + (Y + 0.1) ** 2 + (Z - 0.6) ** 2))
    rho += 0.4 * np.exp(-1.7 * ((X + 0.3) ** 2 + Y ** 2 + (Z + 0.8) ** 2))
    return rho


def _patch_energy_brute_force(g, rho, atom, radius, shape, k):
    """Patch quadrature with the neighbours of a NumPy brute force ordered by
    (distance^2, node index) -- the tie rule of the HIP kNN."""
    from oracle import pseudo as op

    axes = [np.linspace(c - radius, c + radius, n) for c, n in zip(atom.position, shape)]
    cellv = float(axes[0][1] - axes[0][0]) * float(axes[1][1] - axes[1][0]) * float(axes[2][1] - axes[2][0])
    PX, PY, PZ = np.meshgrid(*axes, indexing="ij")
    pr = np.sqrt((PX - atom.position[0]) ** 2 + (PY - atom.position[1]) ** 2 + (PZ - atom.position[2]) ** 2)
    m = pr <= radius
    pts = np.stack([PX[m], PY[m], PZ[m]], 1)
    xyz = np.stack([g.x_points.ravel(), g.y_points.ravel(), g.z_points.ravel()], 1)
    vals = rho.ravel()
    interp = np.empty(pts.shape[0])
    for s in range(0, pts.shape[0], 400):
        idx, dist = _knn_numpy(xyz, pts[s:s + 400], k)
        for t in range(idx.shape[0]):
            if dist[t, 0] <= 1e-12:
                interp[s + t] = vals[idx[t, 0]]
            else:
                w = 1.0 / np.maximum(dist[t], 1e-12) ** 2
                interp[s + t] = np.sum(vals[idx[t]] * (w / w.sum()))
    return float(np.sum(interp * op.local_potential_radial(pr[m], op.GTH_PADE[atom.element])) * cellv)


@pytest.mark.parametrize("neighbors", [1, 8, 12])
def test_patch_matches_oracle(small_monitor, neighbors):
    g = small_monitor
    rho = _axial_density(g)
    params = pseudo.LocalPotentialPatchParameters(patch_radius_scale=0.8, patch_grid_shape=(17, 19, 21),
                                                  correction_strength=1.1, interpolation_neighbors=neighbors)
    ev = pseudo.evaluate_monitor_grid_local_ionic_potential_with_patch(ogrid.H2_CASE, g, rho,
                                                                       patch_parameters=params)
    emb = pseudo.evaluate_monitor_grid_local_ionic_potential_with_frozen_patch_field(
        ogrid.H2_CASE, g, rho, patch_parameters=params, patch_evaluation=ev)
    V, field, tot = opseudo.patch_local_correction_field(
        ogrid.H2_CASE, g, rho, patch_radius_scale=0.8, patch_shape=(17, 19, 21), correction_strength=1.1,
        neighbors=neighbors)
    print(neighbors, ev.total_patch_correction, tot)
    assert abs(ev.total_patch_correction - tot) <= 1e-12 * max(abs(tot), 1e-3)
    np.testing.assert_allclose(emb.correction_field, field, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(emb.corrected_total_local_potential, V, rtol=1e-12, atol=1e-12)
    assert abs(ev.corrected_local_energy - ev.uncorrected_local_energy - tot) <= 1e-12
    for rec in ev.atomic_patch_corrections:
        assert rec.patch_point_count > 0 and rec.patch_grid_shape == (17, 19, 21)


@pytest.mark.parametrize("neighbors", [3, 8])
def test_patch_energy_asymmetric_density_tie_rule(small_monitor, neighbors):
    """An asymmetric density, where the choice among equidistant neighbours matters: the
    GPU patch energies equal a brute-force quadrature with the (distance, index) rule."""
    g = small_monitor
    rho = _asym_density(g)
    params = pseudo.LocalPotentialPatchParameters(patch_radius_scale=0.8, patch_grid_shape=(17, 19, 21),
                                                  interpolation_neighbors=neighbors)
    ev = pseudo.evaluate_monitor_grid_local_ionic_potential_with_patch(ogrid.H2_CASE, g, rho,
                                                                       patch_parameters=params)
    for rec, atom in zip(ev.atomic_patch_corrections, ogrid.H2_CASE.atoms):
        ref = _patch_energy_brute_force(g, rho, atom, rec.patch_radius_bohr, (17, 19, 21), neighbors)
        assert abs(rec.patch_integrated_energy - ref) <= 1e-12 * abs(ref), (rec.patch_integrated_energy, ref)


def test_patched_eigensolver_route(small_monitor):
    """use_monitor_patch=True equals passing the oracle's patched potential explicitly."""
    g = small_monitor
    n = _axial_density(g)
    rho = n / np.sum(n * g.cell_volumes)
    params = pseudo.LocalPotentialPatchParameters(patch_radius_scale=0.75, correction_strength=1.3)
    V, _, _ = opseudo.patch_local_correction_field(ogrid.H2_CASE, g, 2 * rho, patch_radius_scale=0.75,
                                                   correction_strength=1.3)
    kw = dict(k=2, tolerance=1e-7, max_iterations=800, kinetic_version="trial_fix")
    a = eigensolver.solve_fixed_potential_static_local_eigenproblem(
        g, rho, rho, "up", case=ogrid.H2_CASE, use_monitor_patch=True, patch_parameters=params, **kw)
    b = eigensolver.solve_fixed_potential_static_local_eigenproblem(
        g, rho, rho, "up", case=ogrid.H2_CASE, local_ionic_potential=V, **kw)
    assert a.converged and b.converged
    assert a.operator_context.frozen_patch_local_embedding is not None
    print(a.eigenvalues, b.eigenvalues, a.iteration_count)
    np.testing.assert_allclose(a.eigenvalues, b.eigenvalues, rtol=0, atol=1e-10)

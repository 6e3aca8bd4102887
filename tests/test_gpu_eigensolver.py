"""GPU block eigensolver and orthonormalisation against dense FP64 references.

Bars: eigenvalues within 1e-8 Ha of a dense eigh of the assembled W-similar
operator (trial_fix closure, exactly W-self-adjoint); Ritz vectors W-orthonormal to
1e-12; the mixed FP32 + FP64-refinement solve within 1e-6 Ha of the FP64 one;
orthonormalisation spans the same subspace as the oracle's SVD route to 1e-12.
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import configs, eigensolver
from isogriddft_amd.device import DeviceGrid
from isogriddft_amd.hamiltonian import BlockHamiltonian
from oracle import eigen as oeigen
from oracle import kinetic as okin

pytestmark = pytest.mark.gpu


def _problem(shape=(11, 12, 13)):
    g = configs.build_synthetic_geometry(shape=shape)
    V = -2.0 * np.exp(-0.5 * (g.x_points ** 2 + g.y_points ** 2 + (g.z_points - 0.3) ** 2)) \
        - 1.5 * np.exp(-0.8 * (g.x_points ** 2 + (g.y_points + 0.4) ** 2 + (g.z_points + 0.6) ** 2))
    return g, V


def _dense_eigs(g, V, n):
    L = okin.monitor_laplacian_matrix(g, "trial_fix").toarray()
    H = -0.5 * L + np.diag(V.reshape(-1))
    sw = np.sqrt(g.cell_volumes.reshape(-1))
    S = (sw[:, None] * H) / sw[None, :]
    assert np.max(np.abs(S - S.T)) < 1e-9 * np.max(np.abs(S))
    return np.linalg.eigvalsh(0.5 * (S + S.T))[:n]


def _start(dg, g, m):
    atoms = [("H", (0.0, 0.0, -0.7)), ("H", (0.0, 0.0, 0.7))]
    return eigensolver.default_guess_block(g, atoms, m, dg)


@pytest.mark.parametrize("precondition", ["diagonal", "kinetic"])
@pytest.mark.parametrize("k", [1, 3])
def test_block_iteration_matches_dense_eigh(k, precondition):
    g, V = _problem()
    dg = DeviceGrid(g)
    H = BlockHamiltonian(dg, torch.from_numpy(V).cuda(), closure="trial_fix")
    m = eigensolver.reference_block_size(k)
    X0 = _start(dg, g, m)
    res = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-7, 400, precondition=precondition)
    ref = _dense_eigs(g, V, k)
    print("k", k, precondition, "iters", res.iterations, res.eigenvalues - ref)
    assert res.converged
    assert np.max(np.abs(res.eigenvalues - ref)) <= 1e-8
    psi = res.orbitals.reshape(k, -1)
    S = (psi * dg.weights()[None, :]) @ psi.T
    assert torch.max(torch.abs(S - torch.eye(k, device=S.device, dtype=S.dtype))).item() <= 1e-12


@pytest.mark.parametrize("steps", [0, 1, 2])
def test_kinetic_preconditioner_step_counts(steps, monkeypatch):
    """The kinetic preconditioner with no CG step (x0 = D^-1 r only), one step (the fused
    start and finish back to back) and two steps converges to the dense eigenvalues."""
    monkeypatch.setattr(eigensolver, "KINETIC_CG_STEPS", steps)
    g, V = _problem()
    dg = DeviceGrid(g)
    H = BlockHamiltonian(dg, torch.from_numpy(V).cuda(), closure="trial_fix")
    k = 3
    X0 = _start(dg, g, eigensolver.reference_block_size(k))
    res = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-7, 400)
    print("steps", steps, "iters", res.iterations)
    assert res.converged
    assert np.max(np.abs(res.eigenvalues - _dense_eigs(g, V, k))) <= 1e-8


@pytest.mark.parametrize("precondition", ["diagonal", "kinetic"])
def test_mixed_precision_refinement_matches_fp64(precondition):
    g, V = _problem((14, 15, 16))
    dg = DeviceGrid(g)
    H = BlockHamiltonian(dg, torch.from_numpy(V).cuda(), closure="trial_fix")
    k = 3
    X0 = _start(dg, g, eigensolver.reference_block_size(k))
    r64 = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 400, precondition=precondition)
    rmx = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 400, precision="mixed",
                                               precondition=precondition)
    print("fp64 iters", r64.iterations, "mixed", rmx.iterations, "fp32 part", rmx.fp32_iterations,
          rmx.eigenvalues - r64.eigenvalues)
    assert r64.converged and rmx.converged and rmx.fp32_iterations > 0
    assert np.max(np.abs(rmx.eigenvalues - r64.eigenvalues)) <= 1e-6
    assert rmx.orbitals.dtype == torch.float64


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_inplace_rotation_matches_two_buffers(monkeypatch, precision):
    """The single-buffer form (rotated in place by the MFMA kernel, or in column panels of
    vendor products when the kernel does not take the shape) gives the two-buffer
    iteration's eigenvalues; a small panel width forces many panels and a ragged last one."""
    g, V = _problem((14, 15, 16))
    dg = DeviceGrid(g)
    H = BlockHamiltonian(dg, torch.from_numpy(V).cuda(), closure="trial_fix")
    k = 3
    X0 = _start(dg, g, eigensolver.reference_block_size(k))
    ref = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 400, precision=precision,
                                               inplace_rotation=False)
    monkeypatch.setattr(eigensolver, "ROTATION_PANEL", 1000)
    monkeypatch.setattr(eigensolver, "APPLY_ROWS", 7)
    got = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 400, precision=precision,
                                               inplace_rotation=True)
    print(precision, ref.iterations, got.iterations, got.eigenvalues - ref.eigenvalues)
    assert ref.converged and got.converged and not got.two_buffers
    assert np.max(np.abs(got.eigenvalues - ref.eigenvalues)) <= 1e-9
    assert got.orbitals.dtype == torch.float64
    # the column-panel form (vendor products) that wider coefficient matrices take
    monkeypatch.setattr(eigensolver, "ROTATE_MAX_ROWS", 0)
    pan = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 400, precision=precision,
                                               inplace_rotation=True)
    assert pan.converged and np.max(np.abs(pan.eigenvalues - ref.eigenvalues)) <= 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_hz_reapply_matches_rotation(precision):
    """S Z re-applied after each rotation gives the eigenvalues of the rotated S Z."""
    g, V = _problem((14, 15, 16))
    dg = DeviceGrid(g)
    H = BlockHamiltonian(dg, torch.from_numpy(V).cuda(), closure="trial_fix")
    k = 3
    X0 = _start(dg, g, eigensolver.reference_block_size(k))
    ref = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 400, precision=precision,
                                               hz_update="rotate")
    got = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 400, precision=precision,
                                               hz_update="apply")
    print(precision, ref.iterations, got.iterations, got.eigenvalues - ref.eigenvalues)
    assert ref.converged and got.converged
    assert np.max(np.abs(got.eigenvalues - ref.eigenvalues)) <= 1e-8
    with pytest.raises(ValueError):
        eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 4, hz_update="bogus")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_small_working_subspace(precision):
    """The smallest working subspace the reference allows (2k + 2) still converges to the
    dense eigenvalues with the kinetic preconditioner."""
    g, V = _problem((14, 15, 16))
    dg = DeviceGrid(g)
    H = BlockHamiltonian(dg, torch.from_numpy(V).cuda(), closure="trial_fix")
    k = 4
    X0 = _start(dg, g, 2 * k + 2)
    res = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 600, precision=precision)
    ref = _dense_eigs(g, V, k)
    print(precision, "iters", res.iterations, res.eigenvalues - ref)
    assert res.converged
    assert np.max(np.abs(res.eigenvalues - ref)) <= 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("variant", ["two_buffers", "inplace", "reapply", "diagonal"])
def test_soft_locking_matches_unlocked(precision, variant, monkeypatch):
    """Soft locking (converged rows add no residual direction) reaches the dense
    eigenvalues and those of the unlocked iteration, with the active S R rows moved in
    front of S Z (two buffers and in-place panels), S Z re-applied, and the diagonal
    preconditioner taking the active rows' Ritz values."""
    g, V = _problem((14, 15, 16))
    dg = DeviceGrid(g)
    H = BlockHamiltonian(dg, torch.from_numpy(V).cuda(), closure="trial_fix")
    k = 6
    X0 = _start(dg, g, eigensolver.reference_block_size(k))
    kw = {}
    if variant == "inplace":
        monkeypatch.setattr(eigensolver, "ROTATION_PANEL", 1000)
        kw["inplace_rotation"] = True
    elif variant == "reapply":
        kw["hz_update"] = "apply"
    elif variant == "diagonal":
        kw["precondition"] = "diagonal"
    monkeypatch.setattr(eigensolver, "LOCK_FRACTION", 0.5)  # lock early, so the tail runs with few rows
    print("case", precision, variant, flush=True)
    ref = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-7, 600, precision=precision, lock=False,
                                               **kw)
    got = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-7, 600, precision=precision, **kw)
    dense = _dense_eigs(g, V, k)
    print(precision, variant, ref.iterations, got.iterations, got.active_rows, got.eigenvalues - dense)
    assert ref.converged and got.converged
    assert min(got.active_rows) < k  # some rows were locked
    assert max(ref.active_rows) == min(ref.active_rows) == k
    assert np.max(np.abs(got.eigenvalues - dense)) <= 1e-8
    assert np.max(np.abs(got.eigenvalues - ref.eigenvalues)) <= 1e-8
    assert np.max(got.residual_norms) <= 1e-7
    psi = got.orbitals.reshape(k, -1)
    S = (psi * dg.weights()[None, :]) @ psi.T
    assert torch.max(torch.abs(S - torch.eye(k, device=S.device, dtype=S.dtype))).item() <= 1e-12


def test_orthonormalize_spans_oracle_subspace():
    g, _ = _problem()
    dg = DeviceGrid(g)
    rng = np.random.default_rng(4)
    blk = rng.normal(size=(6,) + dg.shape)
    blk[5] = blk[0] + 2 * blk[1]  # rank 5
    w = g.cell_volumes.reshape(-1)
    got = eigensolver.weighted_orthonormalize_orbitals(blk, g, require_full_rank=False)
    ref = oeigen.orthonormalize(blk, w, require_full_rank=False)
    assert got.shape[0] == ref.shape[0] == 5
    M = oeigen.gram(got.reshape(5, -1), ref.reshape(5, -1), w)
    assert np.max(np.abs(M @ M.T - np.eye(5))) <= 1e-12
    with pytest.raises(ValueError):
        eigensolver.weighted_orthonormalize_orbitals(blk, g)
    S = eigensolver.weighted_overlap_matrix(got, g)
    assert np.max(np.abs(S - np.eye(5))) <= 1e-12


def test_static_local_drop_in(small_monitor):
    g = small_monitor
    psi = np.exp(-0.8 * (g.x_points ** 2 + g.y_points ** 2 + (g.z_points - 0.7) ** 2)) + \
        np.exp(-0.8 * (g.x_points ** 2 + g.y_points ** 2 + (g.z_points + 0.7) ** 2))
    psi /= np.sqrt(np.sum(psi * psi * g.cell_volumes))
    rho = psi * psi
    vxc = -0.5 * np.cbrt(rho)  # explicit xc potential (the reference's route when xc_potential is given)
    res = eigensolver.solve_fixed_potential_static_local_eigenproblem(
        g, rho, rho, "up", 2, xc_potential=vxc, kinetic_version="trial_fix", tolerance=1e-6)
    assert res.converged and res.eigenvalues.shape == (2,)
    assert res.orbitals.shape == (2,) + g.cell_volumes.shape
    assert res.max_orthogonality_error <= 1e-10
    ctx = res.operator_context
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    L = okin.monitor_laplacian_matrix(g, "trial_fix")
    Hs = (-0.5 * L + sp.diags(ctx.effective_local_potential.reshape(-1))).tocsr()
    sw = np.sqrt(g.cell_volumes.reshape(-1))
    S = sp.diags(sw) @ Hs @ sp.diags(1.0 / sw)
    S = 0.5 * (S + S.T)
    ref = np.sort(spla.eigsh(S, k=2, which="SA", tol=1e-14)[0])
    print("drop-in", res.eigenvalues, ref, res.iteration_count)
    assert np.max(np.abs(res.eigenvalues - ref)) <= 1e-8
    # the block apply surface agrees with the context's operator
    act = eigensolver.apply_fixed_potential_static_local_block(res.orbitals, ctx)
    dense = (Hs @ res.orbitals.reshape(2, -1).T).T.reshape(act.shape)
    assert np.max(np.abs(act - dense)) <= 1e-12 * np.max(np.abs(dense))

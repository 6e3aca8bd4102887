"""Host checks of the Poisson multigrid hierarchy (isogriddft_amd.multigrid) on the CPU.

The coarse operators are built with tensor reshapes; here they are compared with the
generic sparse Galerkin product P^T A P of the assembled interior monitor operator (the
oracle's matrix), restricted to same-parity couplings.  The probing that reads the fine
entries off a matrix-vector product is compared with the matrix entries themselves.
Shapes: odd and even extents, extents that are not multiples of 4 (padded, partly empty
aggregates) and grids small enough that level 1 is already the coarsest.
"""

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from isogriddft_amd import configs, multigrid
from oracle import poisson as opoi

SHAPES = [(9, 10, 11), (16, 16, 16), (13, 22, 18), (34, 20, 25)]


def _system(shape):
    g = configs.build_synthetic_geometry(shape=shape)
    L, interior = opoi._interior_system(g)
    n = int(np.prod(shape))
    keep = np.flatnonzero(interior)
    # interior operator embedded in the full grid: zero rows and columns on the faces
    E = sp.csr_matrix((np.ones(keep.size), (keep, keep)), shape=(n, n))
    A = (E @ (-L) @ E).tocsr()
    return A, interior


def _entries(A, shape):
    nx, ny, nz = shape
    idx = np.arange(A.shape[0]).reshape(shape)
    ent = np.zeros((7, nx, ny, nz))
    for k, (ox, oy, oz) in enumerate(multigrid.OFFSETS):
        src = [slice(max(0, -o), n - max(0, o)) for o, n in zip((ox, oy, oz), shape)]
        dst = [slice(max(0, o), n - max(0, -o)) for o, n in zip((ox, oy, oz), shape)]
        rows = idx[tuple(src)].reshape(-1)
        cols = idx[tuple(dst)].reshape(-1)
        ent[k][tuple(src)] = np.asarray(A[rows, cols]).reshape(ent[k][tuple(src)].shape)
    return ent


def _aggregation(shape, interior):
    """P (n x 8 c^3) of level 1: interior node i -> (parity, (i - p) >> 2)."""
    c = [(n + 3) // 4 for n in shape]
    I = np.indices(shape).reshape(3, -1)
    p = I & 1
    S = (I - p) >> 2
    col = (((4 * p[0] + 2 * p[1] + p[2]) * c[0] + S[0]) * c[1] + S[1]) * c[2] + S[2]
    rows = np.flatnonzero(interior)
    return sp.csr_matrix((np.ones(rows.size), (rows, col[rows])), shape=(I.shape[1], 8 * c[0] * c[1] * c[2])), c


def _same_parity_seven_point(Ac, c):
    """(7, 8, cx, cy, cz) of the same-parity part of a coarse operator in [p][S] order."""
    m = 8 * c[0] * c[1] * c[2]
    idx = np.arange(m).reshape(8, *c)
    out = np.zeros((7, 8, *c))
    Ad = Ac.toarray()
    out[0] = Ad[idx.reshape(-1), idx.reshape(-1)].reshape(8, *c)
    k = 1
    for axis in range(3):
        for step in (-1, 1):
            src = [slice(None)] * 4
            dst = [slice(None)] * 4
            if step < 0:
                src[axis + 1], dst[axis + 1] = slice(1, None), slice(None, -1)
            else:
                src[axis + 1], dst[axis + 1] = slice(None, -1), slice(1, None)
            out[k][tuple(src)] = Ad[idx[tuple(src)].reshape(-1), idx[tuple(dst)].reshape(-1)].reshape(
                out[k][tuple(src)].shape)
            k += 1
    return out


@pytest.mark.parametrize("shape", SHAPES)
def test_probing_reads_the_operator_entries(shape):
    """The 28-probe rule (the device kernels' key arithmetic, restated with NumPy masks)
    recovers every same-parity entry exactly: no probe node other than the intended one
    lies among a row's nonzero columns."""
    import mg_reference as mref

    A, _ = _system(shape)
    ref = _entries(A, shape)
    got = mref.probe_entries(A, shape)
    assert np.array_equal(got, ref)
    # the probed set holds the whole same-parity part of A: nothing else couples i with
    # a node of its own parity
    I = np.indices(shape).reshape(3, -1)
    Acoo = A.tocoo()
    same = np.all((I[:, Acoo.row] - I[:, Acoo.col]) % 2 == 0, axis=0)
    assert np.isclose(np.abs(Acoo.data[same]).sum(), np.abs(ref).sum(), rtol=1e-13)


@pytest.mark.parametrize("shape", SHAPES)
def test_level_one_is_the_same_parity_galerkin_product(shape):
    A, interior = _system(shape)
    P, c = _aggregation(shape, interior)
    ref = _same_parity_seven_point((P.T @ A @ P).tocsr(), c)
    got = multigrid.level_one(torch.from_numpy(_entries(A, shape))).numpy()
    scale = np.abs(ref).max()
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref)) <= 1e-13 * scale


@pytest.mark.parametrize("shape", SHAPES)
def test_deeper_levels_are_galerkin_products_of_the_level_above(shape):
    A, interior = _system(shape)
    coef = multigrid.level_one(torch.from_numpy(_entries(A, shape)))
    c = list(coef.shape[2:])
    # the level-1 7-point operators as a sparse matrix, then 2 x 2 x 2 aggregation
    m = 8 * int(np.prod(c))
    idx = np.arange(m).reshape(8, *c)
    rows, cols, vals = [idx.reshape(-1)], [idx.reshape(-1)], [coef[0].numpy().reshape(-1)]
    k = 1
    for axis in range(3):
        for step in (-1, 1):
            src = [slice(None)] * 4
            dst = [slice(None)] * 4
            if step < 0:
                src[axis + 1], dst[axis + 1] = slice(1, None), slice(None, -1)
            else:
                src[axis + 1], dst[axis + 1] = slice(None, -1), slice(1, None)
            rows.append(idx[tuple(src)].reshape(-1))
            cols.append(idx[tuple(dst)].reshape(-1))
            vals.append(coef[k].numpy()[tuple(src)].reshape(-1))
            k += 1
    A1 = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(m, m))
    d = [(n + 1) // 2 for n in c]
    J = np.indices((8, *c)).reshape(4, -1)
    col = (((J[0] * d[0] + (J[1] >> 1)) * d[1] + (J[2] >> 1)) * d[2] + (J[3] >> 1))
    P2 = sp.csr_matrix((np.ones(m), (np.arange(m), col)), shape=(m, 8 * int(np.prod(d))))
    ref = _same_parity_seven_point((P2.T @ A1 @ P2).tocsr(), d)
    got = multigrid.next_level(coef).numpy()
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.abs(ref).max()


def test_coarsest_blocks_and_empty_aggregates():
    shape = (13, 22, 18)
    A, _ = _system(shape)
    levels, inv = multigrid.build_levels(torch.from_numpy(_entries(A, shape)))
    assert max(levels[-1].shape[2:]) <= multigrid.COARSEST
    for lev in levels:
        assert torch.all(lev[0] != 0.0)  # empty aggregates carry diagonal 1
    dense = multigrid.dense_blocks(levels[-1])
    eye = np.einsum("pij,pjk->pik", dense, inv.numpy())
    assert np.allclose(eye, np.eye(dense.shape[1])[None], atol=1e-10)


def _host_multigrid(shape):
    A, interior = _system(shape)

    def matvec(x, out):
        out.copy_(torch.from_numpy(A @ x.numpy()))

    mg = multigrid.PoissonMultigrid(shape, matvec, "cpu", entries=torch.from_numpy(_entries(A, shape)))
    return A, interior, mg


def test_device_probing_refuses_host_tensors():
    with pytest.raises(ValueError):
        multigrid.probe_fine_entries((8, 8, 8), lambda x, out: None, "cpu")


def test_fine_damping_from_the_spectral_radius():
    """omega = 1.5 / lambda with lambda the power-iteration estimate of the spectral radius
    of D^-1 A; the estimate is within 3 % of the dominant eigenvalue (below it)."""
    import scipy.sparse.linalg as spla

    shape = (20, 22, 24)
    A, interior, mg = _host_multigrid(shape)
    keep = np.flatnonzero(interior)
    Ai = A[keep][:, keep]
    DA = sp.diags(1.0 / Ai.diagonal()) @ Ai
    lam = abs(spla.eigs(DA, k=1, which="LM", return_eigenvectors=False, tol=1e-6)[0])
    assert 0.97 * lam <= mg.lam <= 1.0001 * lam
    assert mg.omega == pytest.approx(multigrid.FINE_OMEGA_SCALE / mg.lam)


@pytest.mark.parametrize("shape", [(24, 24, 24), (30, 26, 35)])
def test_vcycle_preconditioned_bicgstab_converges_in_few_iterations(shape):
    """The solve itself (reference V-cycle with the hierarchy of PoissonMultigrid): the
    preconditioned BiCGSTAB reaches 1e-10 in at most a third of the unpreconditioned
    iterations, and both reach the same solution (the preconditioner changes the
    iterates, not the system)."""
    import mg_reference as ref

    A, interior, mg = _host_multigrid(shape)
    M = ref.ReferenceVCycle(A, shape, [c.numpy() for c in mg.levels64], mg.inv64.numpy(), mg.omega,
                            multigrid.COARSE_OMEGA)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(A.shape[0]) * interior
    x0, it0, rel0 = ref.bicgstab(A, b, None, rtol=1e-10, maxiter=2000)
    x1, it1, rel1 = ref.bicgstab(A, b, M, rtol=1e-10, maxiter=2000)
    assert rel0 <= 1e-10 and rel1 <= 1e-10
    assert it1 * 3 <= it0, (it0, it1)
    assert np.linalg.norm(x1 - x0) <= 1e-8 * np.linalg.norm(x0)
    assert np.all(x1[~interior] == 0.0)

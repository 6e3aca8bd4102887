"""Host-side pieces of the block eigensolver (no GPU): the generalized Rayleigh-Ritz solve
and the split-K Gram panelling."""

import numpy as np
import pytest
import torch

from isogriddft_amd.eigensolver import _gemm_nt_split, _host_geigh


def _spd_pair(n, rank, seed):
    rng = np.random.default_rng(seed)
    A = rng.normal(size=(n, n))
    A = 0.5 * (A + A.T)
    X = rng.normal(size=(n, rank))
    B = X @ X.T if rank < n else X @ X.T + n * np.eye(n)
    return torch.from_numpy(A), torch.from_numpy(B)


def test_geigh_full_rank_matches_scipy_reduction():
    import scipy.linalg as sl

    A, B = _spd_pair(12, 12, 0)
    ev, c = _host_geigh(A, B)
    want = sl.eigh(A.numpy(), B.numpy(), eigvals_only=True)
    assert np.allclose(ev.numpy(), want, rtol=1e-12, atol=1e-12)
    assert torch.allclose(c.T @ B @ c, torch.eye(12, dtype=c.dtype), atol=1e-10)
    assert torch.allclose(c.T @ A @ c, torch.diag(ev), atol=1e-9)


def test_geigh_singular_metric_solves_on_its_range():
    """B of rank 9 in 12 dimensions: 9 pairs, B-orthonormal, and A c = theta B c holds on
    the range of B (the Ritz pairs of the projected pencil)."""
    A, B = _spd_pair(12, 9, 1)
    ev, c = _host_geigh(A, B)
    assert c.shape == (12, 9)
    assert torch.allclose(c.T @ B @ c, torch.eye(9, dtype=c.dtype), atol=1e-9)
    assert torch.allclose(c.T @ A @ c, torch.diag(ev), atol=1e-8)
    # the same pairs from an explicit basis of the range
    lb, U = torch.linalg.eigh(B)
    Ur = U[:, lb > 1e-9 * lb.max()]
    Br, Ar = Ur.T @ B @ Ur, Ur.T @ A @ Ur
    Lr = torch.linalg.cholesky(Br)
    Li = torch.linalg.inv(Lr)
    w = torch.linalg.eigvalsh(Li @ Ar @ Li.T)
    assert torch.allclose(ev, w, atol=1e-9)


@pytest.mark.parametrize("N", [20000, 8191, 300])
def test_split_k_products_on_host_tensors(N):
    g = torch.Generator().manual_seed(4)
    buf = torch.randn((9, N + 7), dtype=torch.float64, generator=g)[:, :N]
    got = _gemm_nt_split(buf[:5], buf[3:])
    assert torch.allclose(got, buf[:5] @ buf[3:].T, rtol=1e-13, atol=1e-11)


def test_occupation_edge_moves_to_the_nearest_gap():
    from isogriddft_amd.workloads import occupation_edge

    ev = [-3.7, -3.7 + 1e-10, -3.7 + 2e-10, -3.2, -3.2 + 1e-11, -0.8, -0.5]
    assert occupation_edge(ev, 4) == (3, ev[3] - ev[2])        # edge 4 splits a cluster; 3 and 5 tie -> 3
    assert occupation_edge(ev, 5)[0] == 5
    assert occupation_edge(ev, 6)[0] == 6
    assert occupation_edge([-1.0, -1.0 + 1e-12], 1) == (2, None)

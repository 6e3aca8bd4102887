"""GPU parity of the H psi apply with fused row dots (kinetic.apply_operator_dots, C ABI
ig_monitor_stencil_dots_*): the full-z trial_fix kernel that also returns
dots[m] = sum_n w[n] x[m][n] out[m][n] from the same pass (the (p, A p) products of the
eigensolver's kinetic CG).

* out is bit-identical to the general kernel (the same expression tree).
* dots equal the FP64 NumPy sum of w x out over the returned out within 1e-12 of
  sum |w x out| (only the summation order differs), FP32 and FP64 blocks.
* Cases: partial orbital groups (k not a multiple of 3), several x chunks, a clamped last
  y tile (ny not a multiple of 8: its recomputed rows must count once), two full bands of
  eight y-tile pairs, z rows of 256 FP32 points, no potential.
* Blocks the fused kernel does not take return None and launch nothing: z rows wider
  than one wavefront, production closure, legacy grids.
* The block iteration with the fused dots gives the eigenvalues of the unfused one.
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import configs, eigensolver, kinetic
from isogriddft_amd.device import DeviceGrid
from isogriddft_amd.hamiltonian import BlockHamiltonian

pytestmark = pytest.mark.gpu

GENERAL = 23208  # the general kernel: 2 orbitals, 8 x 32 tiles


def _case(shape, k, dt, seed):
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.randn((k, *shape), generator=gen, device="cuda", dtype=dt)
    w = (torch.rand(dg.N, generator=gen, device="cuda", dtype=torch.float64) + 0.5).to(dt)
    v = (0.3 * torch.randn(dg.N, generator=gen, device="cuda", dtype=torch.float64)).to(dt)
    return dg, x, w, v


CASES = [
    ((20, 24, 32), 5, torch.float32), ((20, 24, 32), 5, torch.float64),
    ((33, 37, 72), 3, torch.float32), ((33, 37, 72), 4, torch.float64),
    ((130, 24, 40), 7, torch.float32), ((12, 61, 100), 4, torch.float32), ((12, 61, 100), 2, torch.float64),
    ((20, 256, 32), 3, torch.float32), ((12, 30, 256), 6, torch.float32),
]


@pytest.mark.parametrize("shape,k,dt", CASES)
def test_fused_dots_match_apply_and_numpy(shape, k, dt):
    print(f"case shape={shape} k={k} {dt}", flush=True)
    dg, x, w, v = _case(shape, k, dt, 3)
    for veff in (v, None):
        ref = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=veff, tile=GENERAL)
        out = torch.full_like(x, float("nan"))
        dots = kinetic.apply_operator_dots(dg, x, out, w, scale=-0.5, closure="trial_fix", veff=veff)
        assert dots is not None and dots.dtype == torch.float64 and dots.shape == (k,)
        assert torch.equal(out, ref), (shape, k, dt)
        xs = x.reshape(k, -1).double().cpu().numpy()
        os_ = ref.reshape(k, -1).double().cpu().numpy()
        terms = w.double().cpu().numpy()[None, :] * xs * os_
        want = terms.sum(axis=1)
        got = dots.cpu().numpy()
        assert np.all(np.abs(got - want) <= 1e-12 * np.abs(terms).sum(axis=1)), (shape, k, dt, got - want)


def test_fused_dots_decline_other_blocks():
    for shape, dt in (((6, 8, 260), torch.float32), ((6, 8, 130), torch.float64)):
        dg, x, w, v = _case(shape, 3, dt, 1)
        out = torch.zeros_like(x)
        assert kinetic.apply_operator_dots(dg, x, out, w, scale=-0.5, veff=v) is None
        assert torch.count_nonzero(out).item() == 0
    dg, x, w, v = _case((20, 24, 32), 3, torch.float32, 1)
    out = torch.zeros_like(x)
    assert kinetic.apply_operator_dots(dg, x, out, w, scale=-0.5, closure="production", veff=v) is None
    dgl = DeviceGrid(configs.legacy_h2_geometry(shape=(11, 11, 13)))
    xl = torch.ones((3, 11, 11, 13), device="cuda", dtype=torch.float32)
    wl = torch.ones(dgl.N, device="cuda", dtype=torch.float32)
    assert kinetic.apply_operator_dots(dgl, xl, torch.zeros_like(xl), wl, scale=-0.5) is None
    with pytest.raises(ValueError):
        kinetic.apply_operator_dots(dg, x, out, w.double(), scale=-0.5)


def test_block_iteration_with_fused_dots_matches_unfused(monkeypatch):
    g = configs.build_synthetic_geometry(shape=(24, 26, 28))
    V = -2.0 * np.exp(-0.5 * (g.x_points ** 2 + g.y_points ** 2 + (g.z_points - 0.3) ** 2))
    dg = DeviceGrid(g)
    H = BlockHamiltonian(dg, torch.from_numpy(V).cuda(), closure="trial_fix")
    k = 4
    atoms = [("H", (0.0, 0.0, -0.7)), ("H", (0.0, 0.0, 0.7))]
    X0 = eigensolver.default_guess_block(g, atoms, eigensolver.reference_block_size(k), dg)
    res = {}
    for fuse in (False, True):
        monkeypatch.setattr(eigensolver, "FUSE_CG_DOTS", fuse)
        monkeypatch.setattr(eigensolver, "FUSE_CG_MIN_GEOMETRY_BYTES", 0)
        res[fuse] = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 400, precision="mixed")
    print("iterations unfused", res[False].iterations, "fused", res[True].iterations, flush=True)
    assert res[False].converged and res[True].converged
    assert np.max(np.abs(res[True].eigenvalues - res[False].eigenvalues)) <= 1e-8

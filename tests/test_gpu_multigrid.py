"""GPU checks of the Poisson multigrid preconditioner (isogriddft_amd.multigrid,
csrc/multigrid.hip).

* The fine entries probed through the HIP stencil (28 probe vectors, ig_mg_probe_*) equal
  the oracle's assembled interior matrix entries, and the hierarchy built from them the
  one built from the oracle's matrix (1e-12 relative: the operators differ by rounding).
* One V-cycle on the device (ig_mg_* / ig_mg7_* kernels and the stencil) equals the
  SciPy-sparse reference V-cycle of tests/mg_reference.py with the same hierarchy:
  FP64 within 1e-11 relative, FP32 within 2e-5 of the FP64 reference.  Shapes: a grid
  whose level 1 is the coarsest, odd extents, extents that are not multiples of 4, and
  three levels.
* The preconditioned solve reaches the same discrete solution as the unpreconditioned
  one (FP64, 1e-9 relative at 1e-11 residual) on the 96^3 operator-level grid, in at most
  a quarter of the iterations; the mixed FP32 + FP64 solve with the FP32 V-cycle agrees
  with FP64 to 1e-8 and needs no more FP32 iterations than FP64 ones.
* The transposed operator (the adjoint solve): its probed entries equal the oracle
  matrix's transpose, and the adjoint gradient with the transposed hierarchy equals the
  unpreconditioned one (1e-9 relative at 1e-12 residual) in a third of the iterations.
* Given a stream, the V-cycle and the solves enqueue all their work on it: the results
  equal those on the current stream bit for bit (every kernel and reduction order is
  fixed), the adjoint to 1e-13 relative (its corrected-moment scatter adds atomically).
"""

import numpy as np
import pytest
import torch

import mg_reference as ref
from isogriddft_amd import configs, multigrid, poisson
from isogriddft_amd.device import DeviceGrid
from test_multigrid_host import _entries, _system

pytestmark = pytest.mark.gpu

SHAPES = [(9, 10, 11), (30, 26, 35), (40, 36, 44)]


def _mg(shape):
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    return dg, poisson.poisson_multigrid(dg)


@pytest.mark.parametrize("shape", SHAPES)
def test_hierarchy_from_device_probes_matches_oracle_matrix(shape):
    print(f"case shape={shape}", flush=True)
    dg, mg = _mg(shape)
    A, _ = _system(shape)
    want = _entries(A, shape)
    got = multigrid.probe_fine_entries(shape, lambda x, out: poisson._matvec(dg, x, out), "cuda").cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-12 * np.abs(want).max()
    assert np.array_equal(got == 0.0, want == 0.0)
    levels, inv = multigrid.build_levels(torch.from_numpy(_entries(A, shape)))
    assert len(levels) == len(mg.levels64)
    for a, b in zip(mg.levels64, levels):
        b = b.numpy()
        assert np.max(np.abs(a.cpu().numpy() - b)) <= 1e-12 * np.abs(b).max()
    assert torch.allclose(mg.inv64.cpu(), inv, rtol=1e-9, atol=1e-12 * float(inv.abs().max()))
    assert 1.0 < mg.lam < 4.0 and mg.omega == pytest.approx(multigrid.FINE_OMEGA_SCALE / mg.lam)


@pytest.mark.parametrize("shape", SHAPES)
def test_vcycle_kernels_match_sparse_reference(shape):
    print(f"case shape={shape}", flush=True)
    dg, mg = _mg(shape)
    A, interior = _system(shape)
    M = ref.ReferenceVCycle(A, shape, [c.cpu().numpy() for c in mg.levels64], mg.inv64.cpu().numpy(), mg.omega,
                            multigrid.COARSE_OMEGA)
    rng = np.random.default_rng(11)
    for trial in range(2):
        b = rng.standard_normal(A.shape[0]) * interior
        want = M(b)
        scale = np.abs(want).max()
        for dt, tol in ((torch.float64, 1e-11), (torch.float32, 2e-5)):
            bt = torch.from_numpy(b).to("cuda", dt)
            out = torch.full_like(bt, float("nan"))
            mg.apply(bt, out)
            got = out.double().cpu().numpy()
            err = np.max(np.abs(got - want)) / scale
            assert err <= tol, (shape, dt, trial, err)
            assert np.all(got[~interior] == 0.0)


def test_preconditioned_solve_matches_unpreconditioned_at_96():
    """96^3 operator-level grid (two Gaussian-weighted centres), six Gaussian charges."""
    cfg = configs.C2
    g = configs.build_synthetic_geometry(cfg)
    dg = DeviceGrid(g)
    rho = configs.build_synthetic_density(g, cfg, device="cuda").to(torch.float64).reshape(-1)
    print("case 96^3 none / multigrid / mixed", flush=True)
    r0 = poisson.solve_poisson_device(dg, rho, tolerance=1e-11, max_iterations=6000, preconditioner="none")
    r1 = poisson.solve_poisson_device(dg, rho, tolerance=1e-11, max_iterations=6000, preconditioner="multigrid")
    rm = poisson.solve_poisson_device(dg, rho, tolerance=1e-10, max_iterations=6000, precision="mixed",
                                      preconditioner="multigrid")
    print("iterations none", r0["iterations"], "multigrid", r1["iterations"], "mixed", rm["iterations"], flush=True)
    assert r0["converged"] and r1["converged"] and rm["converged"]
    assert r1["method"].startswith("hip_bicgstab_mg_fp64")
    p0 = r0["potential"].cpu().numpy()
    p1 = r1["potential"].cpu().numpy()
    pm = rm["potential"].cpu().numpy()
    assert np.max(np.abs(p1 - p0)) <= 1e-9 * np.abs(p0).max()
    assert np.max(np.abs(pm - p1)) <= 1e-8 * np.abs(p1).max()
    assert 4 * r1["iterations"] <= r0["iterations"]
    assert rm["iterations"] <= r1["iterations"] + 4


@pytest.mark.parametrize("shape", [(30, 26, 35), (40, 36, 44)])
def test_transposed_hierarchy_matches_oracle_transpose(shape):
    print(f"case transposed shape={shape}", flush=True)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    A, _ = _system(shape)
    At = A.T.tocsr()
    want = _entries(At, shape)
    got = multigrid.probe_fine_entries(shape, lambda x, out: poisson._matvec(dg, x, out, transpose=True),
                                       "cuda").cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-12 * np.abs(want).max()
    mg = poisson.poisson_multigrid(dg, transpose=True)
    assert mg.transpose and mg is poisson.poisson_multigrid(dg, transpose=True)
    levels, _ = multigrid.build_levels(torch.from_numpy(want))
    for a, b in zip(mg.levels64, levels):
        b = b.numpy()
        assert np.max(np.abs(a.cpu().numpy() - b)) <= 1e-12 * np.abs(b).max()


def test_adjoint_solve_with_transposed_multigrid():
    shape = (40, 36, 44)
    print(f"case adjoint shape={shape}", flush=True)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    gen = torch.Generator(device="cuda").manual_seed(7)
    w = torch.randn(dg.N, generator=gen, dtype=torch.float64, device="cuda")
    g0, i0 = poisson.poisson_adjoint_device(dg, w, tolerance=1e-12, max_iterations=6000, preconditioner="none")
    g1, i1 = poisson.poisson_adjoint_device(dg, w, tolerance=1e-12, max_iterations=6000)
    print("adjoint iterations none", i0["iterations"], "multigrid", i1["iterations"], flush=True)
    assert i0["converged"] and i1["converged"]
    assert torch.allclose(g1, g0, rtol=0, atol=1e-9 * float(g0.abs().max()))
    assert 3 * i1["iterations"] <= i0["iterations"]


def test_multigrid_refuses_legacy_grids_and_transpose():
    dg = DeviceGrid(configs.legacy_h2_geometry(shape=(11, 11, 13)))
    rho = torch.ones(dg.N, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        poisson.solve_poisson_device(dg, rho, preconditioner="multigrid")
    dgm, mg = _mg((12, 12, 12))
    x = torch.zeros(dgm.N, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        poisson.bicgstab(dgm, x.clone(), x, transpose=True, mg=mg)


def test_solves_on_a_side_stream_equal_the_current_stream():
    shape = (40, 36, 44)
    print(f"case side stream shape={shape}", flush=True)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    gen = torch.Generator(device="cuda").manual_seed(5)
    rho = torch.rand(dg.N, generator=gen, dtype=torch.float64, device="cuda")
    side = torch.cuda.Stream()
    for precision in ("fp64", "mixed"):
        r0 = poisson.solve_poisson_device(dg, rho, tolerance=1e-10, max_iterations=2000, precision=precision)
        r1 = poisson.solve_poisson_device(dg, rho, tolerance=1e-10, max_iterations=2000, precision=precision,
                                          stream=side)
        side.synchronize()
        assert r0["converged"] and r1["converged"] and r0["iterations"] == r1["iterations"], precision
        assert torch.equal(r0["potential"], r1["potential"]), precision
    mg = poisson.poisson_multigrid(dg)
    b = rho * poisson.poisson_data(dg).interior_mask(torch.float64)
    for dt in (torch.float64, torch.float32):
        bt = b.to(dt)
        o0 = torch.empty_like(bt)
        o1 = torch.full_like(bt, float("nan"))
        mg.apply(bt, o0)
        torch.cuda.current_stream().synchronize()
        side.wait_stream(torch.cuda.current_stream())
        mg.apply(bt, o1, stream=side)
        side.synchronize()
        assert torch.equal(o0, o1), dt
    w = torch.randn(dg.N, generator=gen, dtype=torch.float64, device="cuda")
    g0, i0 = poisson.poisson_adjoint_device(dg, w, tolerance=1e-10, max_iterations=2000)
    g1, i1 = poisson.poisson_adjoint_device(dg, w, tolerance=1e-10, max_iterations=2000, stream=side)
    side.synchronize()
    assert i0["iterations"] == i1["iterations"]
    assert float((g1 - g0).abs().max()) <= 1e-13 * float(g0.abs().max())

"""GPU parity of the z-vectorised trial_fix stencil (128-bit lanes, one barrier per step).

Bars: FP64 <= 1e-12 relative against the oracle; FP32 <= 2e-5; the z-vectorised and the
general kernel evaluate the same expression tree, so their results must agree bit for
bit; x chunking and orbital grouping must not change a bit either.
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import configs, kinetic
from isogriddft_amd.device import DeviceGrid
from oracle import kinetic as okin

pytestmark = pytest.mark.gpu

ZV = 7000000  # tile code forcing the z-vectorised kernel (+ orbitals per workgroup)
GENERAL = 23208  # tile code of the general kernel: 2 orbitals, 8 x 32 tiles


def _fields(shape, k, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape], indexing="ij")
    out = []
    for _ in range(k):
        a, b, c = rng.uniform(0.5, 2.0, 3)
        out.append(np.exp(-a * X * X - b * Y * Y - c * Z * Z) * (1 + 0.2 * rng.standard_normal(shape)))
    return np.array(out)


SHAPES = [(20, 24, 32), (11, 30, 36), (33, 37, 96), (9, 40, 128), (40, 96, 96)]


@pytest.mark.parametrize("shape", SHAPES)
def test_zv_matches_oracle_fp64_and_fp32(shape):
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    psi = _fields(shape, 3, 5)
    v = 0.3 * np.sin(g.x_points) * np.cos(g.z_points)
    ref = np.array([okin.monitor_kinetic(p, g, "trial_fix") + v * p for p in psi])
    scale = np.max(np.abs(ref))
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
        x = torch.from_numpy(psi).to("cuda", dt)
        vt = torch.from_numpy(v).to("cuda", dt)
        for nb in (1, 2):
            out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=ZV + nb)
            err = np.max(np.abs(out.double().cpu().numpy() - ref)) / scale
            assert err <= tol, (shape, dt, nb, err)


@pytest.mark.parametrize("shape", [(33, 37, 96), (40, 96, 96)])
def test_zv_bitwise_equal_to_general_kernel(shape):
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    for dt in (torch.float64, torch.float32):
        x = torch.from_numpy(_fields(shape, 5, 9)).to("cuda", dt)
        vt = torch.from_numpy(0.2 * np.cos(g.y_points)).to("cuda", dt)
        gen = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=GENERAL)
        for nb in (1, 2):
            for xc in (0, 5, 16):
                zv = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=ZV + nb, xchunk=xc)
                assert torch.equal(zv, gen), (shape, dt, nb, xc, (zv - gen).abs().max().item())


def test_zv_interior_only_and_no_potential():
    shape = (18, 26, 64)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    x = torch.from_numpy(_fields(shape, 2, 3)).cuda()
    zv = kinetic.apply_operator(dg, x, scale=-1.0, closure="trial_fix", interior_only=True, tile=ZV + 2)
    gen = kinetic.apply_operator(dg, x, scale=-1.0, closure="trial_fix", interior_only=True, tile=GENERAL)
    assert torch.equal(zv, gen)
    assert zv[:, 0].abs().max().item() == 0.0 and zv[:, :, :, -1].abs().max().item() == 0.0


def test_zv_rejects_unsupported_shapes():
    g = configs.build_synthetic_geometry(shape=(10, 12, 66))  # ny < 16 and nz % 4 != 0
    dg = DeviceGrid(g)
    x = torch.from_numpy(_fields((10, 12, 66), 1, 1)).cuda()
    with pytest.raises(ValueError):
        kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", tile=ZV + 1)
    # the automatic choice falls back to the general kernel
    out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix")
    ref = okin.monitor_kinetic(x[0].cpu().numpy(), g, "trial_fix")
    assert np.max(np.abs(out[0].cpu().numpy() - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_large_fp32_default_dispatch_is_bitwise_equal():
    """FP32 blocks whose geometry outgrows the Infinity Cache (here 168^3 > 4.2 M points)
    go by default to the general kernel with four orbitals per workgroup; the result must
    equal the z-vectorised kernel's bit for bit."""
    shape = (168, 168, 168)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.standard_normal((4,) + shape)).float().cuda()
    vt = torch.from_numpy(0.3 * np.sin(g.x_points) * np.cos(g.z_points)).float().cuda().reshape(-1)
    default = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt)
    zv = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=ZV + 2)
    assert torch.equal(default, zv)

"""GPU parity of the 128-bit-lane register-queue trial_fix stencils: the z-vectorised
kernel (32-point z tiles, z-halo points on spare lanes; NB = 1-4 orbitals per workgroup)
and the pipelined full-z kernel (one wavefront per z row, lane-permute z neighbours,
one barrier per plane; the default for large FP32 blocks).

Bars: FP64 <= 1e-12 relative against the oracle; FP32 <= 2e-5.  Both kernels evaluate the
general kernel's expression tree, so their results agree bit for bit, for every orbital
count (partial last group), x chunking, clamped tiles, grids smaller than one tile and the
folded eigensolver weights.  Shapes outside a kernel's range (z extent not a multiple of
the lane width, or wider than a wavefront for the full-z kernel) are refused with a
ValueError and not run.
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import configs, kinetic
from isogriddft_amd.device import DeviceGrid
from oracle import kinetic as okin

pytestmark = pytest.mark.gpu

ZQ = 9000000  # tile code forcing the z-vectorised register-queue kernel (+ orbitals per workgroup)
ZP = 9200000  # tile code forcing the pipelined full-z kernel (+ rows x 10 + orbitals, + 100: deeper prefetch)
ZP_VARIANTS = (81, 82, 83, 182, 122, 43, 64, 1083, 2083)  # + 1000 / 2000: y-sweep block orders
GENERAL = 23208  # the general kernel: 2 orbitals, 8 x 32 tiles


def _fields(shape, k, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape], indexing="ij")
    out = []
    for _ in range(k):
        a, b, c = rng.uniform(0.5, 2.0, 3)
        out.append(np.exp(-a * X * X - b * Y * Y - c * Z * Z) * (1 + 0.2 * rng.standard_normal(shape)))
    return np.array(out)


# y extents below, at and above the FP32 (28) and FP64 (12) tile heights; z extents below,
# at and above the 32-point tile width, multiples of 4 (the FP32 lane width)
SHAPES = [(5, 6, 8), (20, 24, 32), (11, 30, 36), (33, 37, 72), (40, 96, 96)]


@pytest.mark.parametrize("shape", SHAPES)
def test_zq_matches_oracle_fp64_and_fp32(shape):
    print(f"case shape={shape}", flush=True)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    psi = _fields(shape, 5, 5)
    v = 0.3 * np.sin(g.x_points) * np.cos(g.z_points)
    ref = np.array([okin.monitor_kinetic(p, g, "trial_fix") + v * p for p in psi])
    scale = np.max(np.abs(ref))
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
        x = torch.from_numpy(psi).to("cuda", dt)
        vt = torch.from_numpy(v).to("cuda", dt)
        for nb in (1, 2, 3, 4):
            out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=ZQ + nb)
            err = np.max(np.abs(out.double().cpu().numpy() - ref)) / scale
            assert err <= tol, (shape, dt, nb, err)


@pytest.mark.parametrize("shape", [(33, 37, 72), (40, 96, 96), (130, 24, 40), (9, 61, 100)])
def test_zq_bitwise_equal_to_general_kernel(shape):
    print(f"case shape={shape}", flush=True)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    for dt in (torch.float64, torch.float32):
        for k in (3, 13):
            x = torch.from_numpy(_fields(shape, k, 9)).to("cuda", dt)
            vt = torch.from_numpy(0.2 * np.cos(g.y_points)).to("cuda", dt)
            gen = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=GENERAL)
            # + 100 g: items of g consecutive units dealt round robin to the XCDs
            for nb in (1, 2, 3, 4, 103, 203, 303, 802):
                for xchunk in (0, 9, shape[0]):
                    out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=ZQ + nb,
                                                 xchunk=xchunk)
                    assert torch.equal(out, gen), (shape, dt, k, nb, xchunk)
            # without a potential, and with the folded weights of the eigensolver
            gen = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", tile=GENERAL)
            out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", tile=ZQ + 2)
            assert torch.equal(out, gen), (shape, dt, k, "no potential")
            s = torch.from_numpy(np.sqrt(g.cell_volumes).reshape(-1)).to("cuda", dt)
            ij = dg.invj(dt) * s
            gen = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt.reshape(-1) * s, invj=ij,
                                         tile=GENERAL)
            out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt.reshape(-1) * s, invj=ij,
                                         tile=ZQ + 3)
            assert torch.equal(out, gen), (shape, dt, k, "folded")


def test_zq_refuses_unaligned_z():
    g = configs.build_synthetic_geometry(shape=(8, 8, 10))
    dg = DeviceGrid(g)
    x = torch.ones((2, 8, 8, 10), device="cuda", dtype=torch.float32)
    with pytest.raises(ValueError):
        kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", tile=ZQ + 2)


# pipelined full-z kernel: z extents up to 256 (FP32) / 128 (FP64) points; y extents
# below, at and above its 4-, 8- and 12-row tiles
ZP_SHAPES = [(5, 6, 8), (20, 24, 32), (11, 30, 36), (33, 37, 72), (40, 13, 96)]


@pytest.mark.parametrize("shape", ZP_SHAPES)
def test_zp_matches_oracle_fp64_and_fp32(shape):
    print(f"case shape={shape}", flush=True)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    psi = _fields(shape, 5, 11)
    v = 0.3 * np.sin(g.x_points) * np.cos(g.z_points)
    ref = np.array([okin.monitor_kinetic(p, g, "trial_fix") + v * p for p in psi])
    scale = np.max(np.abs(ref))
    for dt, tol in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
        x = torch.from_numpy(psi).to("cuda", dt)
        vt = torch.from_numpy(v).to("cuda", dt)
        for var in ZP_VARIANTS:
            out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=ZP + var)
            err = np.max(np.abs(out.double().cpu().numpy() - ref)) / scale
            assert err <= tol, (shape, dt, var, err)


# (20, 256, 32): 32 tiles of 8 rows, two full bands of eight tile pairs in the y-sweep orders
@pytest.mark.parametrize("shape", [(33, 37, 72), (130, 24, 40), (9, 61, 100), (12, 30, 256), (20, 256, 32)])
def test_zp_bitwise_equal_to_general_kernel(shape):
    print(f"case shape={shape}", flush=True)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    dts = (torch.float64, torch.float32) if shape[2] <= 128 else (torch.float32,)
    for dt in dts:
        for k in (3, 13):
            x = torch.from_numpy(_fields(shape, k, 9)).to("cuda", dt)
            vt = torch.from_numpy(0.2 * np.cos(g.y_points)).to("cuda", dt)
            gen = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=GENERAL)
            for var in ZP_VARIANTS:
                for xchunk in (0, 9, shape[0]):
                    out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=ZP + var,
                                                 xchunk=xchunk)
                    assert torch.equal(out, gen), (shape, dt, k, var, xchunk)
            s = torch.from_numpy(np.sqrt(g.cell_volumes).reshape(-1)).to("cuda", dt)
            ij = dg.invj(dt) * s
            gen = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt.reshape(-1) * s, invj=ij,
                                         tile=GENERAL)
            out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt.reshape(-1) * s, invj=ij,
                                         tile=ZP + 83)
            assert torch.equal(out, gen), (shape, dt, k, "folded")


def test_zp_refuses_wide_z():
    for shape, dt in (((6, 8, 260), torch.float32), ((6, 8, 130), torch.float64)):
        dg = DeviceGrid(configs.build_synthetic_geometry(shape=shape))
        x = torch.ones((2, *shape), device="cuda", dtype=dt)
        with pytest.raises(ValueError):
            kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", tile=ZP + 83)


def test_default_large_block_bitwise_equal_to_general_kernel():
    """The launcher's own choice for a large block (geometry beyond the Infinity Cache):
    128 x 168 x 224 with 32 orbitals (a partial last group of the 3-orbital workgroups) in
    FP32 (the full-z kernel) and FP64 (the z-vectorised kernel), against the general
    kernel."""
    shape, k = (128, 168, 224), 32
    print(f"case default shape={shape} k={k}", flush=True)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    gen_rng = torch.Generator(device="cuda").manual_seed(17)
    for dt in (torch.float32, torch.float64):
        x = torch.randn((k, *shape), generator=gen_rng, device="cuda", dtype=dt)
        vt = torch.from_numpy(0.2 * np.cos(g.y_points) * np.sin(g.x_points)).to("cuda", dt)
        gen = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=GENERAL)
        out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt)
        assert torch.equal(out, gen), dt
        del x, gen, out

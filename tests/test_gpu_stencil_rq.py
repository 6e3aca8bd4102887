"""GPU parity of the register-queue trial_fix stencil (x neighbours from per-lane register
queues, one plane of LDS, NB = 4 / 8 / 12 orbitals per workgroup).

Bars: FP64 <= 1e-12 relative against the oracle; FP32 <= 2e-5; the register-queue and the
general kernel evaluate the same expression tree, so their results agree bit for bit,
for every orbital count (partial last group), x chunking, clamped tiles and tiny grids.
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import configs, kinetic
from isogriddft_amd.device import DeviceGrid
from oracle import kinetic as okin

pytestmark = pytest.mark.gpu

RQ = 8000000  # tile code forcing the register-queue kernel (+ orbitals per workgroup)
GENERAL = 23208  # the general kernel: 2 orbitals, 8 x 32 tiles


def _fields(shape, k, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape], indexing="ij")
    out = []
    for _ in range(k):
        a, b, c = rng.uniform(0.5, 2.0, 3)
        out.append(np.exp(-a * X * X - b * Y * Y - c * Z * Z) * (1 + 0.2 * rng.standard_normal(shape)))
    return np.array(out)


SHAPES = [(5, 6, 7), (20, 24, 32), (11, 30, 36), (33, 37, 70), (40, 96, 96)]


@pytest.mark.parametrize("shape", SHAPES)
def test_rq_matches_oracle_fp64_and_fp32(shape):
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
        for nb in (4, 8, 12):
            out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=RQ + nb)
            err = np.max(np.abs(out.double().cpu().numpy() - ref)) / scale
            assert err <= tol, (shape, dt, nb, err)


@pytest.mark.parametrize("shape", [(33, 37, 70), (40, 96, 96), (130, 24, 40)])
def test_rq_bitwise_equal_to_general_kernel(shape):
    print(f"case shape={shape}", flush=True)
    g = configs.build_synthetic_geometry(shape=shape)
    dg = DeviceGrid(g)
    for dt in (torch.float64, torch.float32):
        for k in (3, 13):
            x = torch.from_numpy(_fields(shape, k, 9)).to("cuda", dt)
            vt = torch.from_numpy(0.2 * np.cos(g.y_points)).to("cuda", dt)
            gen = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=GENERAL)
            for nb in (4, 8, 12):
                for xchunk in (0, 9, shape[0]):
                    out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt, tile=RQ + nb,
                                                 xchunk=xchunk)
                    assert torch.equal(out, gen), (shape, dt, k, nb, xchunk)
            # the folded weights of the eigensolver (invj = s / J, veff = s V) as well
            s = torch.from_numpy(np.sqrt(g.cell_volumes).reshape(-1)).to("cuda", dt)
            ij = dg.invj(dt) * s
            gen = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt.reshape(-1) * s, invj=ij,
                                         tile=GENERAL)
            out = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=vt.reshape(-1) * s, invj=ij,
                                         tile=RQ + 8)
            assert torch.equal(out, gen), (shape, dt, k, "folded")

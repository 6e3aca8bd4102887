"""GPU discrete Green-identity terms (kinetic.green_identity_terms, ig_green_identity_f64)
and the operator/adjoint defect (kinetic.green_identity_defect).

Bars: the four unnormalised sums against the oracle at 1e-12 relative on the reference's
19x19x23 grid; the reference-held 67x67x81 scalars (audit/baselines.py:1483-1545,
:1572-1633) at 1e-10 Ha on the GPU-generated grid (the generator reproduces the
reference coordinates to 1e-12 Bohr); defect of the zero-ghost closure below 1e-13 of
sum |u T v w|, the production defect against the oracle at 1e-12 relative.
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import kinetic, monitor_grid
from oracle import grid as ogrid
from oracle import kinetic as okin

from test_oracle_green import REFERENCE_GREEN, reference_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("closure", ["trial_fix", "production"])
def test_terms_match_oracle(small_monitor, closure):
    g = small_monitor
    rng = np.random.default_rng(11)
    f = np.exp(-0.05 * (g.x_points ** 2 + g.y_points ** 2 + g.z_points ** 2)) + 0.2 * rng.normal(size=g.x_points.shape)
    t = kinetic.green_identity_terms(f, g, closure)
    k_op, k_grad, k_bdry, den = okin.green_identity_terms(f, g, closure)
    assert abs(t.weighted_norm ** 2 - den) <= 1e-12 * den
    for got, want in ((t.operator_kinetic_ha, k_op / den), (t.gradient_kinetic_ha, k_grad / den),
                      (t.boundary_term_ha, k_bdry / den)):
        assert abs(got - want) <= 1e-12 * max(abs(want), abs(k_grad / den)), (closure, got, want)
    mm = 1000 * (k_op - k_grad - k_bdry) / den
    assert abs(t.closure_mismatch_mha - mm) <= 1e-9 * max(abs(mm), 1000 * abs(k_grad / den))


def test_reference_scalars_on_generated_grid():
    g = monitor_grid.build_monitor_grid_for_case(
        ogrid.H2_CASE, shape=(67, 67, 81), box_half_extents=(8.0, 8.0, 10.0),
        element_parameters=monitor_grid.build_h2_local_patch_development_element_parameters())
    for (label, closure), (op, gr, dmha) in REFERENCE_GREEN.items():
        t = kinetic.green_identity_terms(reference_field(label, g), g, closure)
        print(label, closure, t)
        assert abs(t.operator_kinetic_ha - op) < 1e-10, (label, closure)
        assert abs(t.gradient_kinetic_ha - gr) < 1e-10, (label, closure)
        assert abs(t.delta_kinetic_mha - dmha) < 1e-7, (label, closure)
        assert abs(t.boundary_term_ha) < 1e-14
        assert abs(t.closure_mismatch_mha - t.delta_kinetic_mha) < 1e-9


def test_defect_zero_ghost_vanishes_production_matches_oracle(small_monitor):
    g = small_monitor
    rng = np.random.default_rng(5)
    u = rng.normal(size=g.x_points.shape)
    v = rng.normal(size=g.x_points.shape)
    w = g.cell_volumes
    for closure in ("trial_fix", "production"):
        d = kinetic.green_identity_defect(u, v, g, closure)
        tu = okin.monitor_kinetic(u, g, closure)
        tv = okin.monitor_kinetic(v, g, closure)
        scale = float(np.sum(np.abs(u * tv * w)))
        want = float(np.sum(u * tv * w) - np.sum(tu * v * w))
        print(closure, "defect", d, "oracle", want, "scale", scale)
        if closure == "trial_fix":
            assert abs(d) < 1e-13 * scale
        else:
            assert abs(want) > 1e-3 * scale
            assert abs(d - want) < 1e-12 * scale


def test_defect_legacy_grid_vanishes():
    g = ogrid.build_legacy_geometry(shape=(17, 19, 21))
    rng = np.random.default_rng(2)
    u = rng.normal(size=g.x_points.shape)
    v = rng.normal(size=g.x_points.shape)
    d = kinetic.green_identity_defect(u, v, g)
    tv = okin.legacy_kinetic(v, g)
    from isogriddft_amd.device import device_grid
    wsa = device_grid(g).operator_weights().cpu().numpy().reshape(g.x_points.shape)
    assert abs(d) < 1e-13 * float(np.sum(np.abs(u * tv * wsa)))


def test_tensor_input_and_errors(small_monitor):
    g = small_monitor
    f = torch.from_numpy(np.cos(0.2 * g.x_points) * np.exp(-0.1 * g.z_points ** 2)).cuda()
    a = kinetic.green_identity_terms(f, g, "trial_fix")
    b = kinetic.green_identity_terms(f.cpu().numpy(), g, "trial_fix")
    assert a == b
    with pytest.raises(ValueError):
        kinetic.green_identity_terms(f, g, "bogus")
    with pytest.raises(ValueError):
        kinetic.green_identity_terms(f[:-1], g)

"""Host logic of the GPU monitor-grid generator: element parameters and grid specs
against the oracle's restatement of grid/monitor_builder.py:63-216."""

import numpy as np
import pytest

from isogriddft_amd import monitor_grid
from oracle import grid as ogrid


@pytest.mark.parametrize("case", [ogrid.H2_CASE, ogrid.h2o_case()], ids=["H2", "H2O"])
def test_default_element_parameters(case):
    got = monitor_grid.build_default_near_core_element_parameters(case)
    ref = ogrid.default_element_parameters(case)
    assert set(got) == set(ref)
    for el, p in got.items():
        r = ref[el]
        for f in ("near_core_radius", "local_radius", "projector_radius", "patch_radius", "kinetic_weight",
                  "local_weight", "projector_weight", "kinetic_exponent", "local_exponent", "projector_exponent",
                  "monitor_cap"):
            assert getattr(p, f) == getattr(r, f), (el, f)


def test_h2_development_parameters_and_spec(golden):
    got = monitor_grid.build_h2_local_patch_development_element_parameters()
    ref = ogrid.h2_patch_development_parameters()
    assert got["H"].kinetic_weight == ref["H"].kinetic_weight and got["H"].local_radius == ref["H"].local_radius
    spec = monitor_grid.build_monitor_grid_spec_for_case(ogrid.H2_CASE, shape=(19, 19, 23),
                                                         box_half_extents=(8.0, 8.0, 10.0), element_parameters=got)
    np.testing.assert_allclose(np.array(spec.box_bounds), golden["mon_box"], rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        monitor_grid.build_monitor_grid_spec_for_case(ogrid.H2_CASE, shape=(2, 19, 23))
    with pytest.raises(ValueError):
        monitor_grid.NearCoreElementParameters("H", 0.5, 0.5, 0.5, 1.0, 1, 1, 0, 2, 4, 2, 1.0)

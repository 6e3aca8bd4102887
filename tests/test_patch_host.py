"""Host logic of the near-core patch: parameter validation (pseudo/local.py:55-76) and
patch-radius lookup on reference-style and oracle geometries."""

from types import SimpleNamespace

import pytest

from isogriddft_amd import pseudo


def test_parameter_validation():
    p = pseudo.LocalPotentialPatchParameters()
    assert p.patch_grid_shape == (25, 25, 25) and p.interpolation_neighbors == 8
    for bad in (dict(patch_radius_scale=0.0), dict(correction_strength=-1.0), dict(interpolation_neighbors=0),
                dict(patch_grid_shape=(25, 25)), dict(patch_grid_shape=(24, 25, 25)),
                dict(patch_grid_shape=(1, 25, 25))):
        with pytest.raises(ValueError):
            pseudo.LocalPotentialPatchParameters(**bad)


def test_patch_radii_lookup():
    itf = SimpleNamespace(patch_radius=1.25, implemented_purposes=("local_gth", "nonlocal"))
    assert pseudo.geometry_patch_radii(SimpleNamespace(patch_interfaces=(itf, itf))) == (1.25, 1.25)
    bad = SimpleNamespace(patch_radius=1.0, implemented_purposes=("nonlocal",))
    with pytest.raises(ValueError):
        pseudo.geometry_patch_radii(SimpleNamespace(patch_interfaces=(bad,)))
    assert pseudo.geometry_patch_radii(SimpleNamespace(patch_radii=(0.5, 0.75))) == (0.5, 0.75)
    with pytest.raises(ValueError):
        pseudo.geometry_patch_radii(SimpleNamespace(patch_radii=()))


def test_patch_plan_lives_on_its_grid(monkeypatch):
    """The kNN plan indexes one grid's nodes: it is cached on that DeviceGrid, never
    handed to another grid object (not even one that reuses a freed grid's id)."""
    built = []

    class StubPlan:
        def __init__(self, dg, atoms, radii, params):
            built.append(dg)
            self.dg = dg

    monkeypatch.setattr(pseudo, "PatchPlan", StubPlan)
    geom = SimpleNamespace(patch_radii=(0.5, 0.5))
    case = [("H", (0.0, 0.0, -0.7)), ("H", (0.0, 0.0, 0.7))]
    params = pseudo.LocalPotentialPatchParameters()
    dg1 = SimpleNamespace()
    p1 = pseudo.patch_plan(dg1, case, geom, params)
    assert pseudo.patch_plan(dg1, case, geom, params) is p1 and len(built) == 1
    del dg1, p1
    for _ in range(4):  # fresh grids with the same atoms, radii and parameters
        dg = SimpleNamespace()
        assert pseudo.patch_plan(dg, case, geom, params).dg is dg
        del dg
    assert len(built) == 5

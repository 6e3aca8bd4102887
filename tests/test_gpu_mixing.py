"""GPU density mixers against the reference's outputs and the oracle.

Bars: mixed densities 1e-12 relative against the reference vectors
(tests/golden/mixing_vectors.npz) and against the oracle on further histories; Anderson
decisions exactly and its projected-residual ratio to 1e-12.
"""

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from isogriddft_amd import configs, mixing
from oracle import mixing as omixing

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
from mixing_cases import CASES, history  # noqa: E402

pytestmark = pytest.mark.gpu
MIX = Path(__file__).resolve().parent / "golden" / "mixing_vectors.npz"


def _run(kind, h, geom, na, nb, par, as_torch=False):
    if as_torch:
        h = [tuple(torch.from_numpy(a).cuda() for a in e) for e in h]
    if kind == "diis":
        ent = [mixing.MonitorGridDiisHistoryEntry(*e) for e in h]
        return mixing.apply_monitor_grid_density_diis(ent, grid_geometry=geom, n_alpha=na, n_beta=nb), None
    if kind == "anderson":
        ent = [mixing.MonitorGridAndersonHistoryEntry(*e) for e in h]
        r = mixing.apply_monitor_grid_density_anderson(ent, grid_geometry=geom, n_alpha=na, n_beta=nb, **par)
        res = None if r.mixed_rho_up is None else (r.mixed_rho_up, r.mixed_rho_down)
        return res, (r.accepted, r.effective_damping, r.filtered_history_length, r.projected_residual_ratio)
    ent = [mixing.MonitorGridBroydenHistoryEntry(*e) for e in h]
    return mixing.apply_monitor_grid_density_broyden_like(ent, grid_geometry=geom, n_alpha=na, n_beta=nb,
                                                          **par), None


def _close(got, ref):
    got = got.cpu().numpy() if isinstance(got, torch.Tensor) else got
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * max(np.max(np.abs(ref)), 1e-300))


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gpu_mixers_match_reference(golden, small_monitor, case):
    mg = np.load(MIX)
    name, kind, length, na, nb, col, par = case
    w = golden["mon_cell_volumes"]
    h = history(w.shape, w, length, na, nb, col)
    for as_torch in (False, True):
        res, meta = _run(kind, h, small_monitor, na, nb, par, as_torch)
        assert (res is None) == bool(mg[f"{name}_none"])
        if res is not None:
            _close(res[0], mg[f"{name}_up"])
            _close(res[1], mg[f"{name}_down"])
            assert isinstance(res[0], torch.Tensor) == as_torch
        if meta is not None:
            acc, d, flen, ratio = meta
            assert acc == bool(mg[f"{name}_accepted"])
            assert d == float(mg[f"{name}_effective_damping"])
            assert flen == int(mg[f"{name}_filtered_history_length"])
            assert abs(ratio - float(mg[f"{name}_projected_residual_ratio"])) <= 1e-12


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_gpu_mixers_match_oracle_on_other_histories(case):
    name, kind, length, na, nb, col, par = case
    geom = configs.build_synthetic_geometry(shape=(14, 16, 18))
    w = geom.cell_volumes
    h = history(w.shape, w, length + 1, na, nb, col, seed=3)
    res, meta = _run(kind, h, geom, na, nb, par)
    if kind == "diis":
        ref, rmeta = omixing.diis(h, w, na, nb), None
    elif kind == "anderson":
        ref, *rmeta = omixing.anderson(h, w, na, nb, **par)
    else:
        ref, rmeta = omixing.broyden(h, w, na, nb, **par), None
    assert (res is None) == (ref is None)
    if ref is not None:
        _close(res[0], ref[0])
        _close(res[1], ref[1])
    if meta is not None:
        assert meta[:3] == tuple(rmeta[:3])
        assert (meta[3] is None) == (rmeta[3] is None)
        if meta[3] is not None:
            assert abs(meta[3] - rmeta[3]) <= 1e-12

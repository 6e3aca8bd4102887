"""Pin the oracle's density mixers: the reference's own outputs on SCF-like histories
(tests/golden/mixing_vectors.npz from tests/golden/make_mixing_golden.py) and closed forms.

Bar: 1e-12 relative on the mixed densities; Anderson decisions (accepted, effective
damping, kept secants) exactly and its projected-residual ratio to 1e-12.
"""

import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import mixing

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
from mixing_cases import CASES, history  # noqa: E402

MIX = Path(__file__).resolve().parent / "golden" / "mixing_vectors.npz"


@pytest.fixture(scope="module")
def mix_golden():
    return np.load(MIX)


def run_oracle(kind, h, w, na, nb, par):
    if kind == "diis":
        return mixing.diis(h, w, na, nb), None
    if kind == "anderson":
        res, acc, d, flen, ratio = mixing.anderson(h, w, na, nb, **par)
        return res, (acc, d, flen, ratio)
    return mixing.broyden(h, w, na, nb, **par), None


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_oracle_mixers_match_reference(golden, mix_golden, case):
    name, kind, length, na, nb, col, par = case
    w = golden["mon_cell_volumes"]
    h = history(w.shape, w, length, na, nb, col)
    res, meta = run_oracle(kind, h, w, na, nb, par)
    assert (res is None) == bool(mix_golden[f"{name}_none"])
    if res is not None:
        for got, key in zip(res, ("up", "down")):
            ref = mix_golden[f"{name}_{key}"]
            np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * max(np.max(np.abs(ref)), 1e-300))
    if meta is not None:
        acc, d, flen, ratio = meta
        assert acc == bool(mix_golden[f"{name}_accepted"])
        assert d == float(mix_golden[f"{name}_effective_damping"])
        assert flen == int(mix_golden[f"{name}_filtered_history_length"])
        assert abs(ratio - float(mix_golden[f"{name}_projected_residual_ratio"])) <= 1e-12


def _linear_problem(shape=(6, 7, 8)):
    rng = np.random.default_rng(4)
    w = rng.uniform(0.5, 1.5, size=shape)
    star_up = 1.0 + rng.uniform(0, 1, size=shape)
    star_up /= np.sum(star_up * w)
    star_dn = 1.0 + rng.uniform(0, 1, size=shape)
    star_dn /= np.sum(star_dn * w)
    e_up = rng.normal(size=shape) * 0.05 * star_up
    e_up -= np.sum(e_up * w) / np.sum(w) * np.ones(shape)  # zero integral: renormalisation is inert
    e_dn = rng.normal(size=shape) * 0.05 * star_dn
    e_dn -= np.sum(e_dn * w) / np.sum(w) * np.ones(shape)
    return w, star_up, star_dn, e_up, e_dn


def test_diis_two_opposite_residuals_average():
    """Residuals r and -r: the Pulay coefficients are (1/2, 1/2)."""
    w, su, sd, eu, ed = _linear_problem()
    h = [(su + eu, sd + ed, eu, ed), (su - eu, sd - ed, -eu, -ed)]
    up, dn = mixing.diis(h, w, 1, 1)
    np.testing.assert_allclose(up, su, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(dn, sd, rtol=1e-12, atol=1e-15)


def test_anderson_secant_solves_a_uniform_linear_map():
    """r = alpha (rho - rho*) and rho_1 - rho* parallel to the secant: one Anderson step
    with full damping lands on rho*."""
    w, su, sd, eu, ed = _linear_problem()
    a = -0.4
    h = [(su + 2 * eu, sd + 2 * ed, a * 2 * eu, a * 2 * ed), (su + eu, sd + ed, a * eu, a * ed)]
    res, acc, d, flen, ratio = mixing.anderson(
        h, w, 1, 1, regularization=0.0, damping=1.0, step_clip_factor=None, adaptive_damping_enabled=False,
        min_damping=0.1, max_damping=1.0, acceptance_residual_ratio_threshold=1.0,
        collinearity_cosine_threshold=0.999)
    assert acc and flen == 2 and d == 1.0 and ratio < 1e-12
    np.testing.assert_allclose(res[0], su, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(res[1], sd, rtol=1e-12, atol=1e-15)

"""End-to-end molecular workload (grid generator, GTH local + nonlocal, LSDA, Poisson,
both eigensolver precisions, forces) on a small H2O monitor grid."""

import numpy as np
import pytest

from isogriddft_amd import workloads

pytestmark = pytest.mark.gpu


def test_h2o_workload_small():
    out = workloads.run("h2o", n=48, tol=1e-4, multipole_order=2)
    print({k: v for k, v in out.items() if not k.startswith("forces_Ha")})
    assert out["eig_fp64_converged"] and out["eig_mixed_converged"]
    assert out["eig_mixed_vs_fp64_max_abs_Ha"] <= 1e-6
    assert out["poisson_mixed_vs_fp64_max_rel"] <= 1e-5
    F = np.asarray(out["forces_Ha_per_Bohr"])
    assert F.shape == (3, 3) and np.all(np.isfinite(F))
    # the molecule is mirror-symmetric in x and planar in y = 0: the hydrogens' x forces
    # are opposite, and no atom feels a y force (to the eigensolver tolerance)
    assert abs(F[1, 0] + F[2, 0]) <= 1e-4 * max(abs(F[1, 0]), 1e-3)
    assert np.max(np.abs(F[:, 1])) <= 1e-4 * np.max(np.abs(F))


def test_h2_legacy_workload():
    """The reference's CPU-runnable case (legacy 49 x 49 x 61, lda,pw92, k = 2): converges,
    a bound bonding state, a physical Hartree energy of the singlet density."""
    out = workloads.run_h2_legacy()
    print(out)
    assert out["eig_converged"]
    e = out["eigenvalues_Ha"]
    assert -1.0 < e[0] < 0.0 and e[0] < e[1]
    assert 1.0 < out["hartree_energy_Ha"] < 2.5

"""Write golden vectors for the reference functions that only an import can pin.

Run in the build container (where the reference source is mounted) with
    python tests/golden/make_golden.py
It imports the reference's NumPy path from /root/reference/src and writes small
fixtures (inputs and outputs only) to tests/golden/reference_vectors.npz.  Nothing
here runs on the GPU box; the fixtures travel, the reference does not.

Covered (functions no closed form, invariant or recorded scalar pins):
  * monitor harmonic-map generator, 19x19x23 H2 development grid
        grid/monitor_builder.py:186-216, grid/monitor_geometry.py:481-577
  * corrected-moment multipole boundary (top-10 % Jacobian cells, quadratic fits)
        poisson/open_boundary.py:794-987 on a centred and a shifted Gaussian density
  * legacy_split boundary construction    poisson/open_boundary.py:559-719
  * near-core patch frozen embedding       pseudo/local.py:637-703
  * default eigensolver guess block        ks/eigensolver.py:818-913
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/src")
OUT = Path(__file__).resolve().parent / "reference_vectors.npz"


def main():
    sys.path.insert(0, str(REF))
    from isogrid.config import H2_BENCHMARK_CASE
    from isogrid.grid import build_h2_local_patch_development_element_parameters, build_monitor_grid_for_case
    from isogrid.grid import build_grid_geometry, build_default_h2_grid_spec
    from isogrid.poisson.open_boundary import _compute_multipole_boundary_condition
    from isogrid.pseudo import LocalPotentialPatchParameters
    from isogrid.pseudo import evaluate_monitor_grid_local_ionic_potential_with_frozen_patch_field
    from isogrid.ks.eigensolver import _build_default_guess_orbitals
    from dataclasses import replace

    g = build_monitor_grid_for_case(
        H2_BENCHMARK_CASE, shape=(19, 19, 23), box_half_extents=(8.0, 8.0, 10.0),
        element_parameters=build_h2_local_patch_development_element_parameters(),
    )
    out = {
        "mon_x": g.x_points, "mon_y": g.y_points, "mon_z": g.z_points,
        "mon_jacobian": g.jacobian, "mon_cell_volumes": g.cell_volumes,
        "mon_box": np.array(g.spec.box_bounds, dtype=np.float64),
        "mon_patch_radii": np.array([p.patch_radius for p in g.patch_interfaces]),
    }
    r2 = g.x_points ** 2 + g.y_points ** 2 + g.z_points ** 2
    rho_c = np.exp(-0.5 * r2)
    rho_c = 2.0 * rho_c / np.sum(rho_c * g.cell_volumes)
    s2 = g.x_points ** 2 + (g.y_points - 0.4) ** 2 + (g.z_points - 1.1) ** 2
    rho_s = np.exp(-0.9 * s2)
    rho_s = 2.0 * rho_s / np.sum(rho_s * g.cell_volumes)
    for tag, rho in (("centered", rho_c), ("shifted", rho_s)):
        out[f"rho_{tag}"] = rho
        for mode in ("corrected_moments", "legacy_split"):
            bc = _compute_multipole_boundary_condition(g, rho, multipole_order=2,
                                                       monitor_boundary_construction_mode=mode)
            key = f"{tag}_{mode}"
            out[f"{key}_charge"] = np.array(bc.total_charge)
            out[f"{key}_dipole"] = bc.dipole_moment
            out[f"{key}_quadrupole"] = bc.quadrupole_tensor
            out[f"{key}_boundary"] = bc.boundary_values
    # frozen patch embedding on the H2 bonding-trial density
    f = 0.0
    for atom in H2_BENCHMARK_CASE.geometry.atoms:
        dx, dy, dz = g.x_points - atom.position[0], g.y_points - atom.position[1], g.z_points - atom.position[2]
        f = f + np.exp(-0.8 * (dx * dx + dy * dy + dz * dz))
    psi = f / np.sqrt(np.sum(f * f * g.cell_volumes))
    rho_t = 2.0 * psi * psi
    pp = LocalPotentialPatchParameters(patch_radius_scale=0.75, patch_grid_shape=(25, 25, 25),
                                       correction_strength=1.30, interpolation_neighbors=8)
    emb = evaluate_monitor_grid_local_ionic_potential_with_frozen_patch_field(
        H2_BENCHMARK_CASE, g, rho_t, patch_parameters=pp)
    out["patch_rho"] = rho_t
    out["patch_corrected_local_potential"] = emb.corrected_total_local_potential
    out["patch_total_correction"] = np.array(emb.patch_evaluation.total_patch_correction)
    # default guess block on a small legacy grid
    spec = replace(build_default_h2_grid_spec(), nx=15, ny=15, nz=17)
    lg = build_grid_geometry(spec)

    class _Ctx:
        grid_geometry = lg
        case = H2_BENCHMARK_CASE

    out["guess_legacy_shape"] = np.array(spec.shape)
    out["guess_block"] = _build_default_guess_orbitals(_Ctx(), 8)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()

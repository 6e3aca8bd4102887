"""Write golden vectors for the reference's SCF density mixers (DIIS, Anderson,
Broyden-like: scf/driver.py:2125-2650), which no closed form or recorded value pins.

Run in the build container (where the reference source is mounted) with
    python tests/golden/make_mixing_golden.py
It imports the reference's NumPy mixers from /root/reference/src, feeds them the
histories of mixing_cases.py on the 19x19x23 H2 development monitor grid (its cell
volumes are the `mon_cell_volumes` fixture of reference_vectors.npz) and writes the
outputs only to tests/golden/mixing_vectors.npz; the tests regenerate the inputs.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from mixing_cases import CASES, history  # noqa: E402

REF = Path("/root/reference/src")
OUT = HERE / "mixing_vectors.npz"


def main():
    sys.path.insert(0, str(REF))
    from isogrid.config import H2_BENCHMARK_CASE
    from isogrid.grid import build_h2_local_patch_development_element_parameters, build_monitor_grid_for_case
    from isogrid.scf import driver as d

    g = build_monitor_grid_for_case(
        H2_BENCHMARK_CASE, shape=(19, 19, 23), box_half_extents=(8.0, 8.0, 10.0),
        element_parameters=build_h2_local_patch_development_element_parameters(),
    )
    vol = np.load(HERE / "reference_vectors.npz")["mon_cell_volumes"]
    assert np.array_equal(vol, g.cell_volumes)
    out = {}
    for name, kind, length, na, nb, col, par in CASES:
        h = history(g.spec.shape, g.cell_volumes, length, na, nb, col)
        if kind == "diis":
            ent = tuple(d.MonitorGridDiisHistoryEntry(*e) for e in h)
            res = d._apply_monitor_grid_density_diis(ent, grid_geometry=g, n_alpha=na, n_beta=nb)
            up_dn = res
        elif kind == "anderson":
            ent = tuple(d.MonitorGridAndersonHistoryEntry(*e) for e in h)
            r = d._apply_monitor_grid_density_anderson(ent, grid_geometry=g, n_alpha=na, n_beta=nb, **par)
            out[f"{name}_accepted"] = np.array(r.accepted)
            out[f"{name}_effective_damping"] = np.array(r.effective_damping)
            out[f"{name}_filtered_history_length"] = np.array(r.filtered_history_length)
            out[f"{name}_projected_residual_ratio"] = np.array(
                np.nan if r.projected_residual_ratio is None else r.projected_residual_ratio)
            up_dn = None if r.mixed_rho_up is None else (r.mixed_rho_up, r.mixed_rho_down)
        else:
            ent = tuple(d.MonitorGridBroydenHistoryEntry(*e) for e in h)
            up_dn = d._apply_monitor_grid_density_broyden_like(ent, grid_geometry=g, n_alpha=na, n_beta=nb, **par)
        out[f"{name}_none"] = np.array(up_dn is None)
        if up_dn is not None:
            out[f"{name}_up"], out[f"{name}_down"] = up_dn
        print(name, "none" if up_dn is None else "ok")
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()

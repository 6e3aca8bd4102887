"""Host-side checks of the benchmark script: the one-GPU contract and the algorithmic
byte count the reported bandwidth divides by."""

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def test_bench_rejects_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_per_block_apply():
    import bench

    # 96^3 FP32 block of 16 orbitals: psi read and H psi written once per orbital, the six
    # flux coefficients and 1/J once, V once: N (2 * 4 * 16 + 7 * 4 + 4) bytes
    assert bench.algorithmic_bytes(96 ** 3, 16, 4, 4, 4) == 884736 * 160 == 141557760
    # 224^3 FP64 block of 48 orbitals with FP64 geometry
    assert bench.algorithmic_bytes(224 ** 3, 48, 8, 8, 8) == 224 ** 3 * (2 * 8 * 48 + 7 * 8 + 8)


def test_cpu_baseline_leg_on_a_small_grid():
    """The CPU-baseline leg (the FP64 oracle block apply on the host) reports the fields
    of the bench contract, with the bandwidth computed from the same byte count."""
    import numpy as np

    import bench
    from isogriddft_amd import configs

    geom = configs.build_synthetic_geometry(shape=(12, 13, 14))
    rng = np.random.default_rng(1)
    psi = rng.standard_normal((2, 12, 13, 14))
    veff = rng.standard_normal((12, 13, 14))
    r = bench.cpu_baseline(geom, psi, veff, 0.05)
    assert set(r) == {"value", "unit", "cores", "kind", "sample"}
    assert r["unit"] == "GB/s" and r["cores"] == 1 and r["kind"] == "port" and r["value"] > 0
    assert "12x13x14 grid" in r["sample"]


def test_cpu_hpsi_threads_give_the_same_applies():
    """The threaded CPU leg applies the same oracle operator; its byte count per orbital
    apply is the one-orbital block count."""
    import numpy as np

    import bench
    from isogriddft_amd import configs

    geom = configs.build_synthetic_geometry(shape=(10, 11, 12))
    rng = np.random.default_rng(2)
    psi = rng.standard_normal((4, 10, 11, 12))
    veff = rng.standard_normal((10, 11, 12))
    r1 = bench.cpu_hpsi(geom, psi, veff, 0.02, threads=1)
    r2 = bench.cpu_hpsi(geom, psi, veff, 0.02, threads=2)
    for r in (r1, r2):
        assert r["orbital_applies"] % 4 == 0 and r["orbital_applies"] >= 4
        assert abs(r["gbs"] - bench.algorithmic_bytes(1320, 1, 8, 8, 8) * r["orbital_applies_per_s"] / 1e9) < 1e-9
    threads, nproc = bench.host_threads()
    assert 1 <= threads <= min(nproc, bench.MAX_HOST_THREADS)

"""Benchmark: block H.psi on the 224^3, 48-state workload (one MI355X), plus the metric's
eigensolve + Poisson half and the host-CPU baseline.

A step is one fused block apply  H psi = -1/2 Lap psi + V_eff psi  of the 48-orbital FP32
block on the 224^3 analytic two-centre monitor-style grid (box +-14 Bohr, ``trial_fix``
closure), inputs and outputs resident in HBM.  The timed loop cycles over enough distinct
input/output blocks that the working set exceeds twice the 256 MB Infinity Cache.

Prints one JSON line:
  value          whole-job H.psi algorithmic bandwidth (GB/s) of that block apply
  roofline       achieved vs 8 TB/s for the stencil kernel, timed with HIP events on the
                 launch stream; `traffic` = HBM bytes per launch from the committed
                 rocprofv3 PMC summary of the same workload (null if there is none)
  stream_ceiling the bandwidth of this library's 128-bit streaming copy and read kernels
                 on the same GPU (the practical ceiling next to the 8 TB/s datasheet peak)
  hpsi_c2        the same apply on the 96^3, 16-orbital synthetic block
  eigensolve_poisson            Poisson solve and lowest-16 eigensolve on the 96^3 grid,
                                FP64 against FP32 + FP64 refinement
  eigensolve_poisson_benzene    the metric's second part on its own configuration: benzene
                                on the 224^3 monitor grid, k = 48, FP64 against FP32 + FP64
                                refinement, with per-eigenvalue and per-energy-component errors
                                and the adjoint forces
  eigensolve_poisson_h2o        the same for H2O on the 128^3 monitor grid, k = 8
  products_mfma  TFLOP/s and MFMA fraction of the eigensolver's rotation and Gram products
  cpu_baseline   the FP64 NumPy oracle on this host: `value` = H.psi GB/s of a bounded sample
                 of the headline block with all host threads; `runs` = H.psi applies/s,
                 Poisson solves/s and eigensolve wall time on the c1 and c2 configurations
                 with all threads and with one
Algorithmic bytes per block apply: B = N (2 s_psi k + 7 s_g + s_V).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "H·ψ GB/s; eigensolve+Poisson s, FP32+refine vs FP64 GPU at ≤1e-6 Ha"
HBM_PEAK_GBS = 8000.0
LLC_BYTES = 256 * 2 ** 20
MAX_HOST_THREADS = 16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dtype", choices=("f32", "f64"), default="f32")
    ap.add_argument("--closure", choices=("trial_fix", "production"), default="trial_fix")
    ap.add_argument("--tile", type=int, default=0, help="stencil variant code (0 automatic)")
    ap.add_argument("--xchunk", type=int, default=0, help="x planes per workgroup (0 auto)")
    ap.add_argument("--workload", choices=("c2", "c5"), default="c5",
                    help="c5: 224^3, k=48 (default, the metric's configuration); c2: 96^3, k=16")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-seconds", type=float, default=4.0, help="H.psi sample length per CPU leg")
    ap.add_argument("--cpu-quick", action="store_true", help="CPU baseline without the 96^3 Poisson solve")
    ap.add_argument("--no-solvers", action="store_true",
                    help="skip the eigensolve + Poisson timing (FP64 vs FP32 + FP64 refinement)")
    ap.add_argument("--eig-config", choices=("c2", "c5"), default="c2",
                    help="grid of the synthetic eigensolve + Poisson leg: c2 96^3, k=16 (default, order-2 "
                         "Poisson); c5 the 224^3, k=48 block of the headline (order-4 Poisson)")
    ap.add_argument("--eig-tol", type=float, default=1e-3, help="residual tolerance of the synthetic eigensolve")
    ap.add_argument("--eig-block", type=int, default=None,
                    help="working subspace of the 96^3 eigensolve (default: the reference's 12k + 8)")
    ap.add_argument("--molecules", default="benzene,h2o",
                    help="molecular legs, comma separated (benzene: 224^3, k=48; h2o: 128^3, k=8; none)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (the scale for PMC byte counters)")
    return ap.parse_args()


def algorithmic_bytes(N, k, s_psi, s_g=4, s_v=4):
    return N * (2 * s_psi * k + 7 * s_g + s_v)


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, MAX_HOST_THREADS)), n


# ------------------------------------------------------------------ CPU baseline (the oracle)


def _oracle_apply(geom, closure):
    from oracle import kinetic as okin

    if getattr(geom, "kind", "") == "legacy":
        return lambda f: okin.legacy_kinetic(f, geom)
    return lambda f: okin.monitor_kinetic(f, geom, closure)


def cpu_hpsi(geom, psi_np, veff_np, seconds, threads=1, closure="trial_fix"):
    """The oracle's block apply (ks/eigensolver.py:780-783: one orbital at a time,
    T psi + V psi) repeated for about ``seconds``; with threads > 1 the orbitals of the
    block are spread over that many Python threads (NumPy releases the GIL in its array
    operations) and BLAS pools are limited to the same count."""
    from concurrent.futures import ThreadPoolExecutor

    from threadpoolctl import threadpool_limits

    T = _oracle_apply(geom, closure)
    k = psi_np.shape[0]
    N = psi_np[0].size

    def one(f):
        return T(f) + veff_np * f

    done = 0
    with threadpool_limits(limits=threads), ThreadPoolExecutor(max_workers=threads) as ex:
        t0 = time.perf_counter()
        while True:
            if threads == 1:
                for f in psi_np:
                    one(f)
            else:
                list(ex.map(one, psi_np))
            done += k
            if time.perf_counter() - t0 >= seconds:
                break
        dt = time.perf_counter() - t0
    return {"orbital_applies_per_s": done / dt, "gbs": algorithmic_bytes(N, 1, 8, 8, 8) * done / dt / 1e9,
            "orbital_applies": done, "seconds": dt}


def cpu_baseline(geom, psi_np, veff_np, seconds, threads=1):
    """Headline CPU figure: the oracle H.psi bandwidth on a sample of the benchmark block."""
    r = cpu_hpsi(geom, psi_np, veff_np, seconds, threads)
    shape = "x".join(map(str, psi_np.shape[1:]))
    return {
        "value": round(r["gbs"], 4),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": f"FP64 NumPy oracle (ops/kinetic.py restatement, T psi + V psi per orbital), "
                  f"{r['orbital_applies']} orbital applies of {psi_np.shape[0]} sampled orbitals on the {shape} grid "
                  f"in {r['seconds']:.1f} s, {threads} thread(s); bandwidth from the block byte count",
    }


def cpu_runs(threads_all, seconds, quick):
    """The host-CPU reference figures: H.psi applies/s, Poisson solves/s and eigensolve wall time on
    c1 (H2 legacy 49x49x61) and c2 (96^3 synthetic), with all host threads and with one."""
    from threadpoolctl import threadpool_limits

    from isogriddft_amd import configs
    from oracle import eigen as oeig
    from oracle import grid as ogrid
    from oracle import poisson as opoi
    from oracle import pseudo as opseudo

    g1 = configs.legacy_h2_geometry()
    g2 = configs.build_synthetic_geometry(configs.C2)
    r1 = np.exp(-1.6 * (g1.x_points ** 2 + g1.y_points ** 2 + (np.abs(g1.z_points) - 0.7) ** 2))
    v1 = opseudo.local_ionic_potential(ogrid.H2_CASE, g1)
    psi1 = np.stack([r1, g1.z_points * r1])
    psi2 = configs.build_synthetic_orbitals(g2, configs.C2).numpy()
    v2 = -1.0 / np.sqrt(0.3 + g2.x_points ** 2 + g2.y_points ** 2 + g2.z_points ** 2)
    rho2 = configs.build_synthetic_density(g2, configs.C2).numpy()
    w1 = g1.cell_volumes.reshape(-1)
    apply1 = _oracle_apply(g1, "trial_fix")

    def block1(X):
        return np.stack([(apply1(x.reshape(g1.cell_volumes.shape)) + v1 * x.reshape(v1.shape)).reshape(-1)
                         for x in np.asarray(X)])

    out = {}
    for threads in sorted({threads_all, 1}, reverse=True):
        key = f"threads_{threads}"
        res = {}
        h1 = cpu_hpsi(g1, psi1, v1, seconds / 4, threads)
        h2 = cpu_hpsi(g2, psi2, v2, seconds, threads)
        res["hpsi_c1_block_applies_per_s"] = round(h1["orbital_applies_per_s"] / 2, 3)
        res["hpsi_c2_block_applies_per_s"] = round(h2["orbital_applies_per_s"] / 16, 4)
        res["hpsi_c2_gbs"] = round(h2["gbs"], 4)
        with threadpool_limits(limits=threads):
            t0 = time.perf_counter()
            _, _, _, info = opoi.solve_poisson(g1, r1, order=2, method="bicgstab", tol=1e-10)
            dt = time.perf_counter() - t0
            res["poisson_c1_s"] = round(dt, 3)
            res["poisson_c1_solves_per_s"] = round(1.0 / dt, 3)
            res["poisson_c1_iterations"] = info["iterations"]
            t0 = time.perf_counter()
            vals, _ = oeig.eigsh_lowest(block1, w1, 2, psi1[0].reshape(-1), tol=1e-6)
            res["eigensolve_c1_s"] = round(time.perf_counter() - t0, 3)
            res["eigensolve_c1_eigenvalues_Ha"] = [round(float(v), 8) for v in vals]
        out[key] = res
        print(f"[bench] cpu {key}: {res}", file=sys.stderr, flush=True)
    if not quick:
        # the 96^3 solve once: its corrected-moment boundary and the SciPy BiCGSTAB run in
        # single-threaded Python / SciPy sparse code, so the thread count changes nothing
        t0 = time.perf_counter()
        b, _ = opoi.multipole_boundary(g2, rho2, 2, "corrected_moments")
        tb = time.perf_counter() - t0
        _, _, _, info = opoi.solve_poisson(g2, rho2, order=2, method="bicgstab", tol=1e-8)
        dt = time.perf_counter() - t0 - tb
        out["poisson_c2"] = {"s": round(dt, 3), "solves_per_s": round(1.0 / dt, 4), "iterations": info["iterations"],
                             "boundary_s": round(tb, 3), "threads": 1}
        print(f"[bench] cpu poisson c2: {out['poisson_c2']}", file=sys.stderr, flush=True)
    out["workloads"] = ("c1: H2 legacy sinh grid 49x49x61, psi = bonding Gaussian and its z-odd partner, "
                        "V = V_loc(GTH H); Poisson order 2, rtol 1e-10, Jacobi BiCGSTAB (open_boundary.py:1229-1276); "
                        "eigensolve = ARPACK eigsh of the lowest 2 states at tol 1e-6 (ks/eigensolver.py:993-1133). "
                        "c2: the 96^3 synthetic block (k = 16), V = -1/sqrt(0.3 + r^2); Poisson on its 6 Gaussian "
                        "charges, order 2, rtol 1e-8")
    return out


# ------------------------------------------------------------------ GPU legs


def solver_leg(geom, dg, cfg, eig_tol, order=2, block=None):
    """Poisson (multipoles of the given order, rtol 1e-8) on the synthetic charges, then the lowest-k
    eigenproblem of T + V + v_H (trial_fix closure, working subspace `block`, default the
    reference's 12k + 8, residual tolerance eig_tol), each in FP64 and in FP32 with FP64
    refinement."""
    import torch

    from isogriddft_amd import configs, eigensolver, poisson
    from isogriddft_amd.hamiltonian import BlockHamiltonian

    rho = configs.build_synthetic_density(geom, cfg, device="cuda").to(torch.float64)
    out = {}
    pots = {}
    # per-grid set-up: the geometry data of the open-boundary solve (cells of the moment
    # fit), then the multigrid hierarchy of its interior operator
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    poisson.poisson_data(dg)
    torch.cuda.synchronize()
    out["poisson_grid_setup_s"] = round(time.perf_counter() - t0, 4)
    t0 = time.perf_counter()
    poisson.poisson_multigrid(dg)
    torch.cuda.synchronize()
    out["poisson_mg_setup_s"] = round(time.perf_counter() - t0, 4)
    for prec in ("fp64", "mixed"):
        kw = dict(multipole_order=order, tolerance=1e-8, precision=prec, max_iterations=4000)
        poisson.solve_poisson_device(dg, rho, **kw)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sol = poisson.solve_poisson_device(dg, rho, **kw)
        torch.cuda.synchronize()
        out[f"poisson_{prec}_s"] = round(time.perf_counter() - t0, 4)
        out[f"poisson_{prec}_iterations"] = sol["iterations"]
        out[f"poisson_{prec}_solves_per_s"] = round(1.0 / max(out[f"poisson_{prec}_s"], 1e-9), 2)
        pots[prec] = sol["potential"]
        print(f"[bench] poisson {prec}: {out[f'poisson_{prec}_s']} s", file=sys.stderr, flush=True)
    out["poisson_mixed_vs_fp64_max_rel"] = float((pots["mixed"] - pots["fp64"]).abs().max()
                                                 / pots["fp64"].abs().max())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol = poisson.solve_poisson_device(dg, rho, multipole_order=order, tolerance=1e-8, max_iterations=4000,
                                       preconditioner="none")
    torch.cuda.synchronize()
    out["poisson_fp64_unpreconditioned_s"] = round(time.perf_counter() - t0, 4)
    out["poisson_fp64_unpreconditioned_iterations"] = sol["iterations"]
    # a neutral molecule-like potential: two Z = 3 soft-Coulomb centres plus the Hartree
    # potential of the 6-electron synthetic density (trial_fix: exactly W-self-adjoint)
    xyz = dg.xyz()
    vion = sum(-3.0 / torch.sqrt(0.3 + (xyz[0] - c[0]) ** 2 + (xyz[1] - c[1]) ** 2 + (xyz[2] - c[2]) ** 2)
               for c in cfg.centers)
    H = BlockHamiltonian(dg, vion + pots["fp64"].reshape(-1), closure="trial_fix")
    k = cfg.k
    m = block or eigensolver.reference_block_size(k)
    atoms = [("H", c) for c in cfg.centers]
    X0 = eigensolver.default_guess_block(geom, atoms, m, dg)
    ev = {}
    for prec in ("fp64", "mixed"):
        # warm the library GEMM selections of this precision (a 3-iteration solve)
        eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, eig_tol, 3, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, eig_tol, 400, precision=prec,
                                                   profile=True)
        torch.cuda.synchronize()
        out[f"eig_{prec}_s"] = round(time.perf_counter() - t0, 4)
        out[f"eig_{prec}_iterations"] = res.iterations
        out[f"eig_{prec}_converged"] = bool(res.converged)
        ev[prec] = np.asarray(res.eigenvalues)
        out[f"eig_{prec}_phase_s"] = {key: round(v, 4) for key, v in res.timings.items()}
        out[f"eig_{prec}_active_fraction"] = round(float(np.mean(res.active_rows)) / k, 3) if res.active_rows else 1.0
        print(f"[bench] eigensolve {prec}: {out[f'eig_{prec}_s']} s, {res.iterations} iterations",
              file=sys.stderr, flush=True)
    out["eig_fp32_iterations"] = res.fp32_iterations
    out["eig_mixed_vs_fp64_abs_Ha"] = [float(abs(a - b)) for a, b in zip(ev["mixed"], ev["fp64"])]
    out["eig_mixed_vs_fp64_max_abs_Ha"] = float(np.max(np.abs(ev["mixed"] - ev["fp64"])))
    out["eigensolve_plus_poisson_fp64_s"] = round(out["eig_fp64_s"] + out["poisson_fp64_s"], 4)
    out["eigensolve_plus_poisson_mixed_s"] = round(out["eig_mixed_s"] + out["poisson_mixed_s"], 4)
    out["mixed_speedup"] = round(out["eigensolve_plus_poisson_fp64_s"] / out["eigensolve_plus_poisson_mixed_s"], 3)
    out["workload"] = (f"{dg.shape[0]}^3 grid, Poisson: {cfg.n_charges} Gaussian charges, order-{order} multipoles, "
                       f"rtol 1e-8, multigrid-preconditioned BiCGSTAB; eigensolve: k={k}, working subspace {m}, residual tol {eig_tol}, "
                       "trial_fix closure, V = two Z=3 soft-Coulomb centres + v_H")
    return out


MFMA_PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}


def products_leg(n=224, m=192, k=48, reps=3):
    """MFMA utilisation of the eigensolver's two product shapes at the benzene block size
    (N = n^3, working subspace m, k states), timed with HIP events on the launch stream:
    the basis rotation Z' = Q [Z; R] (2 m (m + k) N flop) on the library's own MFMA kernel
    (ig_rotate_*) and the Rayleigh-Ritz Gram block [Z; R] [R; HR]^T (2 (m + k) 2k N flop) on
    the library's split-K MFMA kernel (ig_gemm_nt_*), each also on the vendor GEMMs for
    comparison (rotation: FP32 in 2^20-column panels; Gram: split-K strided-batched), in
    FP32 and FP64, on seeded random blocks."""
    import torch

    from isogriddft_amd import eigensolver as es
    from isogriddft_amd.eigensolver import MM_PANEL, _gemm_nt_kernel, _gemm_nt_split, _rotate_kernel

    N = n ** 3
    out = {"shapes": f"N = {n}^3, rotation ({m} x {m + k}) x ({m + k} x N), Gram ({m + k} x N) x (N x {2 * k})"}
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        g = torch.Generator(device="cuda")
        g.manual_seed(5)
        B = torch.randn((m + k, N), dtype=dt, device="cuda", generator=g)
        Q = torch.randn((m, m + k), dtype=dt, device="cuda", generator=g)
        C = torch.empty((m, N), dtype=dt, device="cuda")
        def vendor_rotation():
            for c0 in range(0, N, MM_PANEL if dt == torch.float32 else N):
                c1 = min(c0 + MM_PANEL, N) if dt == torch.float32 else N
                torch.matmul(Q, B[:, c0:c1], out=C[:, c0:c1])

        def own_rotation():
            if not _rotate_kernel(Q, B, C):
                raise RuntimeError("the MFMA rotation kernel declined the benchmark shape")

        def own_gram():
            if _gemm_nt_kernel(B, B[:2 * k]) is None:
                raise RuntimeError("the split-K MFMA product declined the benchmark shape")

        def vendor_gram():
            es.GEMM_NT_KERNEL = False
            try:
                _gemm_nt_split(B, B[:2 * k])
            finally:
                es.GEMM_NT_KERNEL = True

        legs = (("rotation", own_rotation, 2.0 * m * (m + k) * N),
                ("rotation_vendor", vendor_rotation, 2.0 * m * (m + k) * N),
                ("gram", own_gram, 2.0 * (m + k) * 2 * k * N),
                ("gram_vendor", vendor_gram, 2.0 * (m + k) * 2 * k * N))
        for what, fn, flop in legs:
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            tf = flop / (ms * 1e-3) / 1e12
            out[f"{what}_{name}"] = {"ms": round(ms, 3), "tflops": round(tf, 2), "peak": MFMA_PEAK_TFLOPS[name],
                                     "frac": round(tf / MFMA_PEAK_TFLOPS[name], 4)}
        del B, Q, C
        torch.cuda.empty_cache()
    print(f"[bench] products: {out}", file=sys.stderr, flush=True)
    return out


MOLECULE_LEGS = {"benzene": dict(n=224, k=48, block=192, tol=3e-4), "h2o": dict(n=128, k=8, block=None, tol=3e-4)}


def molecule_leg(name):
    """The metric's second part on a molecular configuration: the fixed-potential
    eigensolve plus open-boundary Poisson solve (multipole order 4) on the molecule's
    atom-centred monitor grid, FP64 against FP32 + FP64 refinement, the error of every
    eigenvalue and energy component, then the adjoint forces (``workloads.run``)."""
    from isogriddft_amd import workloads

    kw = MOLECULE_LEGS[name]
    r = workloads.run(name, kw["n"], kw["k"], kw["block"], kw["tol"])
    keys = ("workload", "k", "occupied", "grid_s", "poisson_grid_setup_s", "poisson_mg_setup_s", "poisson_fp64_s", "poisson_mixed_s",
            "poisson_fp64_iterations", "poisson_mixed_iterations", "poisson_mixed_vs_fp64_max_rel",
            "poisson_fp64_unpreconditioned_s", "poisson_fp64_unpreconditioned_iterations", "eig_fp64_s", "eig_mixed_s",
            "eig_fp64_iterations", "eig_mixed_iterations", "eig_fp32_iterations", "eig_fp64_converged",
            "eig_mixed_converged", "eig_fp64_phase_s", "eig_mixed_phase_s", "eig_fp64_active_fraction",
            "eig_mixed_active_fraction", "eig_mixed_vs_fp64_max_abs_Ha",
            "eig_mixed_vs_fp64_abs_Ha", "eigenvalues_fp64_Ha", "occupation_edge_gap_Ha", "energy_occupied_states",
            "energy_edge_gap_Ha", "energy_components_fp64_Ha",
            "energy_mixed_vs_fp64_abs_Ha", "eigensolve_plus_poisson_fp64_s", "eigensolve_plus_poisson_mixed_s",
            "forces_s", "forces_net_Ha_per_Bohr")
    out = {key: r[key] for key in keys if key in r}
    out["working_subspace"] = kw["block"] or 4 * kw["k"]
    out["residual_tol"] = kw["tol"]
    out["mixed_speedup"] = round(r["eigensolve_plus_poisson_fp64_s"] / r["eigensolve_plus_poisson_mixed_s"], 3)
    return out


def stream_ceiling(reps=10):
    """Bandwidth of the library's 128-bit streaming copy and read kernels over 1 GiB."""
    import torch

    from isogriddft_amd import _lib

    n = 256 * 2 ** 20
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    nb = _lib.load().ig_stream_blocks(n)
    part = torch.empty(nb, dtype=torch.float64, device="cuda")
    h = _lib.stream_handle()
    out = {}
    for what, fn, byt in (("copy", lambda: _lib.call("ig_stream_copy_f32", _lib.ptr(src), _lib.ptr(dst), n, h), 8 * n),
                          ("read", lambda: _lib.call("ig_stream_read_f32", _lib.ptr(src), n, _lib.ptr(part), h),
                           4 * n)):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        out[f"{what}_GBs"] = round(byt / (e0.elapsed_time(e1) / reps) / 1e6, 1)
    del src, dst, part
    torch.cuda.empty_cache()
    return out


def hpsi_timing(cfg, dtype, closure, steps, warmup, tile=0, xchunk=0):
    """Time `steps` block applies of the synthetic block of `cfg` with HIP events on the
    launch stream, cycling over enough blocks to defeat the Infinity Cache.  Returns the
    figures plus the geometry and one input block (for the CPU sample)."""
    import torch

    from isogriddft_amd import configs, kinetic
    from isogriddft_amd.device import DeviceGrid

    t0 = time.perf_counter()
    geom = configs.build_synthetic_geometry(cfg)
    dg = DeviceGrid(geom)
    print(f"[bench] {cfg.shape[0]}^3 geometry in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    N, k = dg.N, cfg.k
    dt = torch.float32 if dtype == "f32" else torch.float64
    s_psi = 4 if dt == torch.float32 else 8
    psi0 = configs.build_synthetic_orbitals(geom, cfg, device="cuda", dtype=torch.float64)
    xyz = dg.xyz()
    veff64 = -1.0 / torch.sqrt(0.3 + xyz[0] ** 2 + xyz[1] ** 2 + (xyz[2] - 0.7) ** 2) \
        - 1.0 / torch.sqrt(0.3 + xyz[0] ** 2 + xyz[1] ** 2 + (xyz[2] + 0.7) ** 2)
    veff = veff64.to(dt).contiguous()
    block_bytes = psi0.numel() * s_psi
    R = max(2, math.ceil(2 * LLC_BYTES / (2 * block_bytes)) + 1)
    ins = [(psi0 * (1.0 + 0.01 * r)).to(dt).contiguous() for r in range(R)]
    outs = [torch.empty_like(ins[0]) for _ in range(R)]
    stream = torch.cuda.current_stream()

    def step(i):
        kinetic.apply_operator(dg, ins[i % R], outs[i % R], scale=-0.5, closure=closure, veff=veff, tile=tile,
                               xchunk=xchunk)

    for i in range(warmup):
        step(i)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(steps):
        step(i)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = ev0.elapsed_time(ev1) / steps
    B = algorithmic_bytes(N, k, s_psi, s_psi, s_psi)
    achieved = B / (ms_per_step * 1e-3) / 1e9
    sample = psi0[: min(k, 16)].cpu().numpy()
    veff_np = veff64.cpu().numpy().reshape(dg.shape)
    del ins, outs, psi0
    torch.cuda.empty_cache()
    return {"geom": geom, "dg": dg, "N": N, "k": k, "R": R, "ms_per_step": ms_per_step, "bytes": B,
            "achieved": achieved, "wall": wall, "sample": sample, "veff_np": veff_np}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch

    from isogriddft_amd import configs

    if args.copy_calibration:
        src = torch.ones(64 * 2 ** 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    cfg = configs.C2 if args.workload == "c2" else configs.C5_OPERATOR
    head = hpsi_timing(cfg, args.dtype, args.closure, args.steps, args.warmup, args.tile, args.xchunk)
    dg = head["dg"]
    achieved = head["achieved"]
    traffic = None
    tf = ROOT / "profiles" / f"hpsi_{args.workload}_{args.dtype}_{args.closure}_traffic.json"
    if tf.exists() and args.tile == 0 and args.xchunk == 0:
        try:
            traffic = json.loads(tf.read_text()).get("bytes_per_launch")
        except (ValueError, OSError):
            traffic = None
    line = {
        "metric": METRIC,
        "value": round(achieved, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload} operator-level synthetic: block H.psi, "
                        f"{dg.shape[0]}x{dg.shape[1]}x{dg.shape[2]} analytic two-centre monitor-style grid "
                        f"(+-{cfg.half_extents[0]:g} Bohr), k={head['k']} orbitals, {args.closure} closure",
            "grid": list(dg.shape),
            "k": head["k"],
            "buffers_cycled": head["R"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "detail": {
            "algorithmic_bytes_per_apply": head["bytes"],
            "block_applies_per_s": round(1e3 / head["ms_per_step"], 2),
            "wall_s_timed_region": round(head["wall"], 4),
        },
    }
    line["stream_ceiling"] = stream_ceiling()
    line["detail"]["frac_of_stream_copy"] = round(achieved / line["stream_ceiling"]["copy_GBs"], 4)
    print(f"[bench] headline {achieved:.1f} GB/s, ceiling {line['stream_ceiling']}", file=sys.stderr, flush=True)
    cpu_geom, cpu_sample, cpu_veff = head["geom"], head["sample"], head["veff_np"]
    del head
    if args.workload != "c2":
        c2 = hpsi_timing(configs.C2, args.dtype, args.closure, args.steps, args.warmup)
        line["hpsi_c2"] = {"gbs": round(c2["achieved"], 2), "frac": round(c2["achieved"] / HBM_PEAK_GBS, 4),
                           "ms_per_step": round(c2["ms_per_step"], 5), "block": "96^3, k=16"}
        geom2, dg2 = c2["geom"], c2["dg"]
        del c2
    else:
        geom2, dg2 = cpu_geom, dg
    torch.cuda.empty_cache()
    if not args.no_solvers:
        if args.eig_config == "c5":
            line["eigensolve_poisson_c5"] = solver_leg(cpu_geom, dg, configs.C5_OPERATOR, args.eig_tol, order=4,
                                                       block=args.eig_block)
        else:
            line["eigensolve_poisson"] = solver_leg(geom2, dg2, configs.C2, args.eig_tol, order=2,
                                                    block=args.eig_block)
        mols = [m for m in args.molecules.split(",") if m and m != "none"]
        for mol in mols:
            torch.cuda.empty_cache()
            line[f"eigensolve_poisson_{mol}"] = molecule_leg(mol)
        torch.cuda.empty_cache()
        line["products_mfma"] = products_leg()
    if not args.no_cpu:
        threads, nproc = host_threads()
        cb = cpu_baseline(cpu_geom, cpu_sample, cpu_veff, args.cpu_seconds, threads)
        cb["nproc"] = nproc
        cb["runs"] = cpu_runs(threads, args.cpu_seconds, args.cpu_quick)
        line["cpu_baseline"] = cb
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

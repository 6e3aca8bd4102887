"""Benchmark: block H.psi on the synthetic 96^3 monitor-style workload (one MI355X).

A step is one fused block apply  H psi = -1/2 Lap psi + V_eff psi  of the 16-orbital
FP32 block on the 96^3 analytic two-centre grid (``trial_fix`` closure, inputs and
outputs resident in HBM).  To keep the number an HBM number, the timed loop cycles
over enough distinct input/output blocks that the working set is larger than twice
the 256 MB last-level (Infinity) cache.

Prints one JSON line:
  value          whole-job H.psi algorithmic bandwidth (GB/s), the metric's first part
  roofline       achieved vs 8 TB/s for the stencil kernel, timed with HIP events on
                 the launch stream; `traffic` from the committed rocprofv3 PMC summary
  cpu_baseline   the FP64 NumPy oracle applying the same block on this host
  eigensolve_poisson          the Poisson solve and lowest-16 eigensolve on the same grid,
                              FP64 against FP32 + FP64 refinement
  eigensolve_poisson_benzene  the metric's second part on its own configuration: benzene
                              on the 224^3 monitor grid, k = 48, eigensolve + Poisson in
                              FP64 and in FP32 + FP64 refinement, and the adjoint forces
  products_mfma               TFLOP/s and MFMA fraction of the eigensolver's rotation and
                              Gram products at that size, FP32 and FP64
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dtype", choices=("f32", "f64"), default="f32")
    ap.add_argument("--closure", choices=("trial_fix", "production"), default="trial_fix")
    ap.add_argument("--tile", type=int, default=0, help="z tile width (0 auto, 32, 64)")
    ap.add_argument("--xchunk", type=int, default=0, help="x planes per workgroup (0 auto)")
    ap.add_argument("--workload", choices=("c2", "c5"), default="c2",
                    help="c2: 96^3, k=16 (default); c5: 224^3, k=48 (benzene-size block)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-solvers", action="store_true",
                    help="skip the eigensolve + Poisson timing (FP64 vs FP32 + FP64 refinement)")
    ap.add_argument("--eig-tol", type=float, default=None,
                    help="eigensolver residual tolerance (default: 1e-3 for c2, 3e-4 for c5)")
    ap.add_argument("--eig-block", type=int, default=None,
                    help="working-subspace size (default: the reference's 12k + 8 for c2, 4k for c5)")
    ap.add_argument("--molecule", choices=("benzene", "h2o", "none"), default="benzene",
                    help="molecular leg: the fixed-potential eigensolve + Poisson of the molecule on its "
                         "monitor grid, FP64 against FP32 + FP64 refinement (benzene: 224^3, k=48)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device copy (the scale for PMC byte counters)")
    return ap.parse_args()


def algorithmic_bytes(N, k, s_psi, s_g=4, s_v=4):
    return N * (2 * s_psi * k + 7 * s_g + s_v)


def cpu_baseline(geom, psi_np, veff_np, seconds):
    """Oracle FP64 block apply, repeated on the host for about ``seconds``."""
    from oracle import kinetic as okin

    N = psi_np[0].size
    k = psi_np.shape[0]
    t0 = time.perf_counter()
    done = 0
    while True:
        for f in psi_np:
            okin.monitor_kinetic(f, geom, "trial_fix") + veff_np * f
            done += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    blocks = done / k
    gbs = algorithmic_bytes(N, k, 8, 8, 8) * blocks / dt / 1e9
    return {
        "value": round(gbs, 4),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": f"FP64 NumPy oracle (ops/kinetic.py restatement), {done} orbital applies "
                  f"({blocks:.2f} blocks of {k}) on the {'x'.join(map(str, psi_np.shape[1:]))} grid in "
                  f"{dt:.1f} s, single thread",
    }


def solver_leg(geom, dg, cfg, veff64, eig_tol, order=2, block=None):
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
    print(f"[bench] starting block of {m} orbitals built", file=sys.stderr, flush=True)
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
        out[f"eig_{prec}_reprojected_dropped_truncated"] = [res.reprojected_rows, res.dropped_rows,
                                                          res.truncated_steps]
        print(f"[bench] eigensolve {prec}: {out[f'eig_{prec}_s']} s, {res.iterations} iterations",
              file=sys.stderr, flush=True)
    out["eig_fp32_iterations"] = res.fp32_iterations
    out["eig_mixed_vs_fp64_max_abs_Ha"] = float(np.max(np.abs(ev["mixed"] - ev["fp64"])))
    out["eigensolve_plus_poisson_fp64_s"] = round(out["eig_fp64_s"] + out["poisson_fp64_s"], 4)
    out["eigensolve_plus_poisson_mixed_s"] = round(out["eig_mixed_s"] + out["poisson_mixed_s"], 4)
    out["workload"] = (f"{dg.shape[0]}^3 grid, Poisson: {cfg.n_charges} Gaussian charges, order-{order} multipoles, "
                       f"rtol 1e-8; eigensolve: k={k}, working subspace {m}, residual tol {eig_tol}, "
                       "trial_fix closure, V = two Z=3 soft-Coulomb centres + v_H")
    return out


MFMA_PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}


def products_leg(n=224, m=192, k=48, reps=3):
    """MFMA utilisation of the eigensolver's two product shapes at the benzene block size
    (N = n^3, working subspace m, k states), timed with HIP events on the launch stream:
    the basis rotation Z' = Q [Z; R] (2 m (m + k) N flop, the panelled library GEMM) and
    the Rayleigh-Ritz Gram block [Z; R] [R; HR]^T (2 (m + k) 2k N flop, split-K batched
    GEMM with FP64 partial sums), in FP32 and FP64, on seeded random blocks."""
    import torch

    from isogriddft_amd.eigensolver import _gemm_nt_split, _mm

    N = n ** 3
    out = {"shapes": f"N = {n}^3, rotation ({m} x {m + k}) x ({m + k} x N), Gram ({m + k} x N) x (N x {2 * k})"}
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        g = torch.Generator(device="cuda")
        g.manual_seed(5)
        B = torch.randn((m + k, N), dtype=dt, device="cuda", generator=g)
        Q = torch.randn((m, m + k), dtype=dt, device="cuda", generator=g)
        C = torch.empty((m, N), dtype=dt, device="cuda")
        legs = (("rotation", lambda: _mm(Q, B, out=C), 2.0 * m * (m + k) * N),
                ("gram", lambda: _gemm_nt_split(B, B[:2 * k]), 2.0 * (m + k) * 2 * k * N))
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
    """The metric's second part on its own configuration: the molecule's fixed-potential
    eigensolve plus open-boundary Poisson solve (multipole order 4) on its atom-centred
    monitor grid, FP64 against FP32 + FP64 refinement, then the adjoint forces at the
    FP64 orbitals (``isogriddft_amd.workloads.run``)."""
    from isogriddft_amd import workloads

    kw = MOLECULE_LEGS[name]
    r = workloads.run(name, kw["n"], kw["k"], kw["block"], kw["tol"])
    keys = ("workload", "k", "occupied", "grid_s", "poisson_fp64_s", "poisson_mixed_s", "poisson_fp64_iterations",
            "poisson_mixed_iterations", "poisson_mixed_vs_fp64_max_rel", "eig_fp64_s", "eig_mixed_s",
            "eig_fp64_iterations", "eig_mixed_iterations", "eig_fp32_iterations", "eig_fp64_converged",
            "eig_mixed_converged", "eig_fp64_phase_s", "eig_mixed_phase_s", "eig_mixed_vs_fp64_max_abs_Ha",
            "eigensolve_plus_poisson_fp64_s", "eigensolve_plus_poisson_mixed_s", "forces_s",
            "forces_net_Ha_per_Bohr")
    out = {key: r[key] for key in keys if key in r}
    out["working_subspace"] = kw["block"] or 4 * kw["k"]
    out["residual_tol"] = kw["tol"]
    out["mixed_speedup"] = round(r["eigensolve_plus_poisson_fp64_s"] / r["eigensolve_plus_poisson_mixed_s"], 3)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch

    from isogriddft_amd import configs, kinetic
    from isogriddft_amd.device import DeviceGrid

    cfg = configs.C2 if args.workload == "c2" else configs.C5_OPERATOR
    geom = configs.build_synthetic_geometry(cfg)
    dg = DeviceGrid(geom)
    N, k = dg.N, cfg.k
    dt = torch.float32 if args.dtype == "f32" else torch.float64
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
        kinetic.apply_operator(dg, ins[i % R], outs[i % R], scale=-0.5, closure=args.closure, veff=veff,
                               tile=args.tile, xchunk=args.xchunk)

    if args.copy_calibration:
        src = torch.ones(64 * 2 ** 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        step(i)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / args.steps
    B = algorithmic_bytes(N, k, s_psi, s_psi, s_psi)
    achieved = B / (ms_per_step * 1e-3) / 1e9

    # correctness spot check of one orbital against the FP64 oracle (small, cheap)
    traffic = None
    tf = ROOT / "profiles" / f"hpsi_c2_{args.dtype}_{args.closure}_traffic.json"
    if tf.exists():
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
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload} operator-level synthetic: block H.psi, "
                        f"{dg.shape[0]}x{dg.shape[1]}x{dg.shape[2]} analytic two-centre "
                        f"monitor-style grid, k={k} orbitals, {args.closure} closure",
            "grid": list(dg.shape),
            "k": k,
            "buffers_cycled": R,
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
            "algorithmic_bytes_per_apply": B,
            "block_applies_per_s": round(1e3 / ms_per_step, 2),
            "wall_s_timed_region": round(wall, 4),
        },
    }
    psi_np = None if args.no_cpu else psi0[: k].cpu().numpy()
    del ins, outs, psi0
    torch.cuda.empty_cache()
    if not args.no_solvers:
        c2 = args.workload == "c2"
        tol = args.eig_tol or (1e-3 if c2 else 3e-4)
        block = args.eig_block or (None if c2 else 4 * k)
        line["eigensolve_poisson"] = solver_leg(geom, dg, cfg, veff64, tol, order=2 if c2 else 4, block=block)
        if args.molecule != "none":
            torch.cuda.empty_cache()
            line[f"eigensolve_poisson_{args.molecule}"] = molecule_leg(args.molecule)
            torch.cuda.empty_cache()
            line["products_mfma"] = products_leg()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(geom, psi_np, veff64.cpu().numpy().reshape(dg.shape), args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

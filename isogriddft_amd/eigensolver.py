"""Fixed-potential block eigensolver on the GPU (drop-in for the reference's
ks/eigensolver.py entry points and the block iteration of ks/eigensolver_jax.py).

Algorithm (ks/eigensolver_jax.py:417-524, restated on device-resident blocks and in the
W^(1/2)-similar form Z = W^(1/2) X, where W-inner products are plain ones and
S = W^(1/2) H W^(-1/2) is the operator):
  Z <- orthonormal(W^(1/2) X0);  HZ <- S Z
  repeat:
    Ritz pairs of span(Z): (theta, C) from P = Z HZ^T and G = Z Z^T, measured every 16
      iterations; in between the basis already is the Ritz basis (C = I, theta from the
      previous step)
    Psi = C^T Z, R = C^T HZ - theta Psi;  stop when max ||R|| <= tol
    R <- M R,  M ~ (T + c)^-1: a few CG steps per row (kinetic preconditioner)
    R <- R - (R Z^T) Z, whitened by the Cholesky factor of R R^T
                                                 (ks/eigensolver_jax.py:119-146)
    HR = S R
    generalized Rayleigh-Ritz on [Z; R]: T c = theta G_B c with T = [[P, Z HR^T],
      [R HZ^T, R HR^T]], G_B the Gram matrix of [Z; R]; lowest m Ritz vectors c
    Z <- c^T [Z; R], HZ <- c^T [HZ; HR]
Device work per iteration: one H apply of k orbitals (the fused HIP stencil), the kinetic
preconditioner (k-row CG: stencil applies and the row-wise CG kernels), the Gram products
as split-K strided-batched library GEMMs, the basis rotation as a library GEMM, and S Z
either rotated along (small m) or re-applied (large m); the (m + k)-sized generalized
eigenproblems run on the host (LAPACK).

The preconditioner is an addition to the reference's block iteration (which has none):
it changes the iterates, not the Ritz pairs converged to.  The kinetic one cuts the c2
workload (96^3, k = 16) from 161 iterations with the diagonal preconditioner to 15, and
makes the 224^3, k = 48 block converge (the diagonal one had not converged after 400
iterations).  W is the inner product that
makes H self-adjoint: the cell volumes on monitor grids; on legacy grids the product
weights (h s_x)(h s_y)(h s_z) of the 7-point flux form, with the returned orbitals then
re-orthonormalised under the cell volumes as the reference returns them.

Precision: "fp64" runs every step in FP64.  "mixed" iterates with FP32 blocks (FP32
stencil, MFMA Gram) until the residuals reach max(tol, fp32 floor), then promotes the
basis to FP64, re-applies H in FP64 and finishes with FP64 iterations (refinement).
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import configs
from .device import DeviceGrid, device_grid
from .hamiltonian import (BlockHamiltonian, H2_BENCHMARK_CASE, normalize_kinetic_version, normalize_spin_channel,
                          resolve_hartree, resolve_xc)
from . import _lib
from .kinetic import apply_operator
from .reductions import cg_direction, cg_finish, cg_start, cg_update, gram, row_dots

FP32_RESIDUAL_FLOOR = 5e-5
# an FP32 plateau promotes early only within this factor of the FP32 floor
FP32_STALL_CEILING = 100.0


# ------------------------------------------------------------------ weighted block algebra


# widest column block handed to one FP32 library GEMM
MM_PANEL = 1 << 20
# Gram blocks of the iteration on the library's split-K MFMA kernel (False: vendor GEMMs
# only).  The kernel streams the grid axis in 128-byte slices with one barrier each, so it
# pays for the large blocks, where it keeps the matrix cores busy (N = 224^3, 240 x 96:
# FP64 10.2 ms against 15.0 ms for the split-K batched vendor GEMM, FP32 5.1 against 5.1),
# and not for short ones (48 x 48: FP64 2.9 against 1.7 ms; 3 x 3: 2.2 against 1.3 ms):
# rows of A from which it is used, per precision.
GEMM_NT_KERNEL = True
GEMM_NT_MIN_ROWS = {torch.float64: (96, 1), torch.float32: (192, 64)}
# rows of Q the MFMA rotation kernel accumulates in registers
ROTATE_MAX_ROWS = 256


def _rotate_kernel(A, B, out):
    """out = A @ B on the library's MFMA rotation kernel (ig_rotate_*).  B and out are
    row-major device blocks with unit column stride; out may overlap B row-aligned (the
    same row stride, offset by whole rows), e.g. out = B[:M] for an in-place rotation.
    Returns False (nothing launched) when the kernel does not take the shape (more than
    ROTATE_MAX_ROWS rows of A, a non-unit column stride, a misaligned overlap)."""
    M, K = A.shape
    N = B.shape[1]
    if (B.dtype not in (torch.float32, torch.float64) or out.dtype != B.dtype or M > ROTATE_MAX_ROWS
            or B.stride(1) != 1 or out.stride(1) != 1 or out.shape != (M, N) or B.shape[0] != K):
        return False
    Ad = A.to(device=B.device, dtype=B.dtype).contiguous()
    sfx = "f32" if B.dtype == torch.float32 else "f64"
    st = getattr(_lib.load(), f"ig_rotate_{sfx}")(_lib.ptr(Ad), Ad.stride(0), _lib.ptr(B), B.stride(0), _lib.ptr(out),
                                                   out.stride(0), M, K, N, 0, _lib.stream_handle())
    if st == -3:
        raise RuntimeError(f"ig_rotate_{sfx}: HIP launch failed")
    return st == 0


def _mm(A, B, out=None):
    """A @ B for a small A (M, K) and a row-major block B (K, N).  Device blocks go to the
    MFMA rotation kernel (up to ROTATE_MAX_ROWS rows of A).  Otherwise FP32 products wider
    than MM_PANEL columns run in column panels: the library's FP32 GEMM (hipBLASLt through
    torch) was measured to return wrong values (NaN or ~1e17) for several (M, K), e.g.
    (192, 248) and (104, 152), at N = 224^3 in one call, and correct ones per panel of
    2^20 columns (FP64 and the rocBLAS path were correct at every shape tried)."""
    N = B.shape[1]
    if B.is_cuda and A.shape[0] <= ROTATE_MAX_ROWS:
        res = out if out is not None else torch.empty((A.shape[0], N), dtype=B.dtype, device=B.device)
        if _rotate_kernel(A, B, res):
            return res
    if B.dtype != torch.float32 or N <= MM_PANEL:
        return torch.matmul(A, B, out=out) if out is not None else torch.matmul(A, B)
    if out is None:
        out = torch.empty((A.shape[0], N), dtype=B.dtype, device=B.device)
    for c0 in range(0, N, MM_PANEL):
        c1 = min(c0 + MM_PANEL, N)
        out[:, c0:c1].copy_(torch.matmul(A, B[:, c0:c1]))
    return out


def _sym(P):
    return 0.5 * (P + P.T)


def _whiten(C, w, floor=1e-10):
    """(T^T C, T) with T = (C^T W C)^(-1/2) (eigenvalue floor 1e-10), else a weighted
    Cholesky-QR transform (ks/eigensolver_jax.py:119-146)."""
    G = _sym(gram(C, C, w))
    ev, V = torch.linalg.eigh(G)
    if bool(torch.isfinite(ev).all()) and float(ev.min()) > floor:
        T = (V * (1.0 / torch.sqrt(ev.clamp_min(1e-14)))[None, :]) @ V.T
    else:
        L = torch.linalg.cholesky(G + 1e-14 * float(ev.abs().max()) * torch.eye(G.shape[0], dtype=G.dtype,
                                                                                  device=G.device))
        T = torch.linalg.inv(L).T  # C^T W C = L L^T  ->  (L^-1 C) is W-orthonormal
    out = _mm(T.T.to(C.dtype), C.reshape(C.shape[0], -1))
    return out, T


def weighted_qr(X, w):
    """Rows of X made W-orthonormal by R^(-1) of the weighted QR (Cholesky-QR on the
    FP64 Gram, ks/eigensolver_jax.py:108-117)."""
    G = _sym(gram(X, X, w))
    L = torch.linalg.cholesky(G)
    Linv = torch.linalg.inv(L)
    return _mm(Linv.to(X.dtype), X.reshape(X.shape[0], -1))


def orthonormalize_device(X, w, rank_tolerance=1e-12, require_full_rank=True):
    """SVD-equivalent weighted orthonormalisation (ks/eigensolver.py:292-341) from the
    FP64 Gram X^T W X = V S^2 V^T: columns X V S^-1 ordered by decreasing S."""
    m = X.shape[0]
    N = X.reshape(m, -1).shape[1]
    G = _sym(gram(X, X, w))
    ev, V = torch.linalg.eigh(G)
    ev = torch.flip(ev, [0])
    V = torch.flip(V, [1])
    s = torch.sqrt(ev.clamp_min(0.0))
    if m == 0 or float(s[0]) == 0.0:
        if require_full_rank:
            raise ValueError("No nonzero orbital directions were provided for orthonormalization.")
        return X.new_zeros((0,) + tuple(X.shape[1:]))
    cutoff = rank_tolerance * max(N, m) * float(s[0])
    rank = int((s > cutoff).sum())
    if require_full_rank and rank < m:
        raise ValueError("The orbital block is rank-deficient under weighted orthonormalization: "
                         f"rank={rank}, requested={m}.")
    T = V[:, :rank] / s[None, :rank]
    out = _mm(T.T.to(X.dtype), X.reshape(m, -1))
    return out.reshape((rank,) + tuple(X.shape[1:]))


@dataclass
class BlockIterationResult:
    eigenvalues: np.ndarray
    orbitals: torch.Tensor
    residual_norms: np.ndarray
    iterations: int
    converged: bool
    residual_history: list = field(default_factory=list)
    ritz_value_history: list = field(default_factory=list)
    subspace_dimensions: list = field(default_factory=list)
    ritz_matrix: np.ndarray | None = None
    basis: torch.Tensor | None = None
    fp32_iterations: int = 0
    timings: dict = field(default_factory=dict)
    # residual rows projected twice / dropped as dependent, and Rayleigh-Ritz steps solved
    # on the range of a singular metric
    reprojected_rows: int = 0
    dropped_rows: int = 0
    truncated_steps: int = 0
    # residual rows that entered the expansion at each iteration (soft locking)
    active_rows: list = field(default_factory=list)
    # wall time of each iteration (profiling runs only) and whether the rotations
    # alternated two buffers (False: column panels rotated in place)
    iteration_seconds: list = field(default_factory=list)
    two_buffers: bool = True


def _host_geigh(A, B):
    """Generalized symmetric eigenproblem A c = theta B c (B SPD) on the host: ascending
    theta and c with c^T B c = I, by Cholesky reduction (torch's LAPACK: a second BLAS
    thread pool in the same process, e.g. SciPy's, was measured to slow these small
    solves down by 50x).  If B is numerically singular (a direction of the expansion
    lies in the span of the others to rounding), the problem is solved on the range of
    B instead: B = U diag(l) U^T, directions with l <= 1e-12 max(l) dropped, so fewer
    than B.shape[0] pairs come back."""
    Ah = A.detach().to("cpu", torch.float64)
    Bh = 0.5 * (B + B.T).detach().to("cpu", torch.float64)
    L, info = torch.linalg.cholesky_ex(Bh)
    if int(info) == 0:
        Linv = torch.linalg.solve_triangular(L, torch.eye(L.shape[0], dtype=L.dtype), upper=False)
        At = Linv @ Ah @ Linv.T
        ev, Y = torch.linalg.eigh(0.5 * (At + At.T))
        return ev, Linv.T @ Y
    lb, U = torch.linalg.eigh(Bh)
    keep = lb > 1e-12 * float(lb.max())
    X = U[:, keep] / torch.sqrt(lb[keep])[None, :]
    At = X.T @ Ah @ X
    ev, Y = torch.linalg.eigh(0.5 * (At + At.T))
    return ev, X @ Y


def _chol_inv(G):
    """L^-1 with G = L L^T for a small SPD FP64 matrix (host), with a tiny diagonal shift
    if the factorisation fails (near rank loss)."""
    Gh = 0.5 * (G + G.T).detach().to("cpu", torch.float64)
    try:
        L = torch.linalg.cholesky(Gh)
    except RuntimeError:
        Gh = Gh + 1e-13 * float(torch.diagonal(Gh).abs().max()) * torch.eye(Gh.shape[0], dtype=Gh.dtype)
        L = torch.linalg.cholesky(Gh)
    return torch.linalg.solve_triangular(L, torch.eye(L.shape[0], dtype=L.dtype), upper=False)


PRECONDITIONER_FLOOR = 1e-2
# kinetic preconditioner: CG steps per residual block and the SPD shift c (Ha)
KINETIC_CG_STEPS = 8
# kinetic CG: take the (p, A p) row dots from the stencil pass itself (FP32 blocks whose
# geometry outgrows the Infinity Cache, where the fused full-z kernel is the stencil the
# launcher would pick anyway)
FUSE_CG_DOTS = True
FUSE_CG_MIN_GEOMETRY_BYTES = 128 << 20
KINETIC_SHIFT = 1.0
# residual rows whose (twice) projection out of the basis keeps less than DEPENDENT_REL *
# eps of their norm are treated as lying in the basis span
DEPENDENT_REL = 10.0
# soft locking: a residual row below LOCK_FRACTION of the phase's target adds no direction
# (0.1 keeps the locked states an order of magnitude inside the tolerance, so quantities
# linear in the orbitals, such as densities and forces, keep most of the unlocked accuracy)
LOCK_FRACTION = 0.1
# K panel of the split-K Gram products: about NT_PANELS panels, none shorter than NT_MIN_PANEL
NT_PANELS = 256
NT_MIN_PANEL = 4096


def _gemm_nt_kernel(A, B):
    """A B^T (FP64) on the library's split-K MFMA kernel (ig_gemm_nt_*), or None (nothing
    launched) when the kernel does not take the operands (16-byte aligned rows)."""
    a, N = A.shape
    b = B.shape[0]
    lib = _lib.load()
    wsize = lib.ig_gemm_nt_work_size(a, b, N)
    if wsize <= 0:
        return None
    work = torch.empty(wsize, dtype=torch.float64, device=A.device)
    G = torch.empty((a, b), dtype=torch.float64, device=A.device)
    sfx = "f32" if A.dtype == torch.float32 else "f64"
    st = getattr(lib, f"ig_gemm_nt_{sfx}")(_lib.ptr(A), A.stride(0), a, _lib.ptr(B), B.stride(0), b, N,
                                          _lib.ptr(work), wsize, _lib.ptr(G), b, _lib.stream_handle())
    if st == -3:
        raise RuntimeError(f"ig_gemm_nt_{sfx}: HIP launch failed")
    return G if st == 0 else None


def _gemm_nt_split(A, B):
    """A B^T (FP64 result) for A (a, N), B (b, N) with unit column stride.  Large device
    blocks (GEMM_NT_MIN_ROWS) go to the library's split-K MFMA kernel (ig_gemm_nt_*: every
    panel of the grid axis read once, all a x b partials of a panel in MFMA accumulators,
    FP64 sum of the partials in a fixed order).  Otherwise the grid axis is cut into P equal panels that run as one
    strided-batched library GEMM (a single GEMM with K = N this long keeps only a few
    workgroups busy), the remainder as one more product, and the partials are added in
    FP64 in a fixed order.  In FP32 each panel accumulates at most N / P products before
    the FP64 sum."""
    a, N = A.shape
    b = B.shape[0]
    if A.stride(1) != 1 or B.stride(1) != 1:
        raise ValueError("Gram operands need unit column stride")
    if (A.is_cuda and A.dtype == B.dtype and A.dtype in GEMM_NT_MIN_ROWS and GEMM_NT_KERNEL
            and a >= GEMM_NT_MIN_ROWS[A.dtype][0] and b >= GEMM_NT_MIN_ROWS[A.dtype][1]):
        G = _gemm_nt_kernel(A, B)
        if G is not None:
            return G
    kp = max(NT_MIN_PANEL, -(-N // NT_PANELS))
    kp = -(-kp // 64) * 64
    P = N // kp
    if P < 2:
        return torch.matmul(A, B.T).to(torch.float64)
    n0 = P * kp
    A3 = A.as_strided((P, a, kp), (kp, A.stride(0), 1))
    B3 = B.as_strided((P, kp, b), (kp, 1, B.stride(0)))
    G = torch.bmm(A3, B3).to(torch.float64).sum(0)
    if n0 < N:
        G += torch.matmul(A[:, n0:], B[:, n0:].T).to(torch.float64)
    return G


# rows per H apply panel and columns per in-place rotation panel
APPLY_ROWS = 64
ROTATION_PANEL = 1 << 20
# working-subspace size m + k from which S Z is re-applied instead of rotated: with
# projectors (the apply is a stencil pass, the projector passes and two scaling passes)
# and on monitor grids without them (the output scaling folds into the stencil).  224^3,
# k = 48, m = 192: without projectors re-applying is faster (eigensolve FP64 23.3 -> 21.5 s,
# mixed 11.8 -> 10.9 s), with the benzene projectors slower (FP64 5.57 -> 5.80 s, mixed
# 3.00 -> 3.19 s).
HZ_APPLY_MIN = 400
HZ_APPLY_MIN_FOLDED = 200
# block buffers above this size release the allocator's cached segments first
BIG_BUFFER_BYTES = 8 << 30


def block_subspace_iteration(H, w, X0, k, tol, max_iter, *, precision="fp64", fp32_floor=FP32_RESIDUAL_FLOOR,
                             keep_basis=False, profile=False, precondition="kinetic", inplace_rotation=None,
                             hz_update=None, lock=True):
    """Lowest-k eigenpairs of the W-self-adjoint H from the starting block X0 (m, N).

    The iteration runs on Z = W^(1/2) X, where the W-inner products become plain ones
    and S = W^(1/2) H W^(-1/2) is the operator (the similarity of the reference's
    weighted-Euclidean route, ks/eigensolver.py:947-990), so every large product is a
    plain tall-skinny GEMM (library, MFMA) with no weighted temporaries.  Rayleigh-Ritz
    results do not depend on the basis, so the whitening of the new basis uses the
    Cholesky factor of its Gram matrix (formed from the small blocks already computed)
    instead of the symmetric inverse square root.

    precondition acts on each residual before it enters the basis.  "kinetic" (default,
    also True) takes KINETIC_CG_STEPS of Jacobi-preconditioned conjugate gradients on
    (T + c) x = r per residual row (c = KINETIC_SHIFT > 0 so the system is SPD): a
    polynomial approximation of the inverse kinetic energy that damps the smooth error as
    well as the rough, at one kinetic block apply of k rows per step.  "diagonal" scales by
    (diag(H) - theta_i)^-1 (Davidson; diag(S) = diag(H) since S is a diagonal similarity
    of H).  False / None: no preconditioner (the reference's iteration).  Every choice
    converges to the same Ritz pairs; only the iteration count changes.

    inplace_rotation: rotate the basis in column panels inside one buffer instead of
    alternating two buffers (half the memory, one extra panel copy per rotation).  None
    (default) takes the two-buffer form when the second buffer fits in 45 % of the free
    device memory.  Mixed precision always alternates the two FP32 halves of one FP64-sized
    storage.

    hz_update: how S Z follows the new basis: "rotate" (HZ' = Q [HR; HZ], 2 m (m + k) N
    flop) or "apply" (HZ' = S Z', a bandwidth-bound block apply).  None (default) applies
    once m + k reaches HZ_APPLY_MIN (HZ_APPLY_MIN_FOLDED on monitor grids without
    projectors), where the rotation's flops outgrow the apply's bytes.

    lock: soft locking.  A state whose residual norm is at most LOCK_FRACTION of the
    current phase's target contributes no residual direction to the expansion: its Ritz
    vector stays in the basis and is still refined by every Rayleigh-Ritz step, its
    residual is still measured each iteration (a state that drifts back above the
    threshold is active again), and convergence is still judged on all k residuals.  The
    preconditioner, the S R apply, the Gram products and the Rayleigh-Ritz size then scale
    with the active rows only.  False expands with all k residuals every iteration.
    """
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be `fp64` or `mixed`; received `{precision}`.")
    if hz_update not in (None, "rotate", "apply"):
        raise ValueError(f"hz_update must be `rotate` or `apply`; received `{hz_update}`.")
    m = X0.shape[0]
    # On monitor grids without projectors the W^(1/2) on the output folds into the
    # stencil's 1/J and V factors (one pass fewer per apply), and re-applying S Z pays from
    # a smaller subspace on
    fold = H.dg.kind == "monitor" and H.nonlocal_plan is None
    reapply = hz_update == "apply" or (
        hz_update is None and m + k >= (HZ_APPLY_MIN_FOLDED if fold else HZ_APPLY_MIN))
    # S is symmetric for the zero-ghost closure and for the legacy operator under its
    # product weights (the W the callers pass); the one-sided production ends are not
    symmetric = H.dg.kind == "legacy" or H.closure == "trial_fix"
    if precondition not in (True, False, None, "diagonal", "kinetic"):
        raise ValueError(f"precondition must be `diagonal`, `kinetic` or False; received `{precondition}`.")
    pmode = "kinetic" if precondition is True else (precondition or None)
    if k > m:
        raise ValueError("the starting block must hold at least k orbitals")
    shape = H.dg.shape
    N = H.dg.N
    dev = X0.device
    w64 = w.reshape(-1).to(torch.float64)
    sw64 = torch.sqrt(w64)
    dt = torch.float32 if precision == "mixed" else torch.float64
    tm = {"hamiltonian_apply": 0.0, "gram": 0.0, "eigh": 0.0, "rotation": 0.0, "precondition": 0.0}
    clock = [time.perf_counter()]

    def mark(key):
        # phase timer (synchronising) when profiling, a no-op otherwise
        if profile:
            torch.cuda.synchronize()
            t = time.perf_counter()
            tm[key] += t - clock[0]
            clock[0] = t

    def setup(dtype):
        sw = sw64.to(dtype)
        return sw, 1.0 / sw

    sw, isw = setup(dt)
    hdiag = H.diagonal() if pmode == "diagonal" else None
    if pmode == "kinetic":
        kop = BlockHamiltonian(H.dg, torch.full((N,), KINETIC_SHIFT, dtype=torch.float64, device=dev),
                               closure=H.closure)
        kdinv64 = 1.0 / (-0.5 * H.dg.kinetic_diagonal() + KINETIC_SHIFT)

    folded = {}

    def apply_s(Zb, out):
        """out = S Zb = W^(1/2) H W^(-1/2) Zb for a block of rows (in panels of at most
        APPLY_ROWS rows, so the unwhitened temporary stays small)."""
        mark("rotation")
        dtb = Zb.dtype
        if fold and dtb not in folded:
            s_ = sw64.to(dtb)
            folded.clear()
            folded[dtb] = (H.dg.invj(dtb) * s_, H.veff[dtb] * s_)
        for r0 in range(0, Zb.shape[0], APPLY_ROWS):
            r1 = min(r0 + APPLY_ROWS, Zb.shape[0])
            xb = (Zb[r0:r1] * isw).view((r1 - r0,) + shape)
            ob = out[r0:r1].view((r1 - r0,) + shape)
            if fold:
                ij, vs = folded[dtb]
                apply_operator(H.dg, xb, ob, scale=-0.5, closure=H.closure, veff=vs, invj=ij)
            else:
                H.apply(xb, ob)
                out[r0:r1].mul_(sw)
        mark("hamiltonian_apply")

    cg_bufs = {}

    def kinetic_cg(Rb):
        """Rb <- ~ (S_T + c)^-1 Rb row by row (Rb: at most k rows).  The CG runs on the X-space rows
        x = W^(-1/2) r with the W-self-adjoint T + c and W-inner products (no similarity
        scaling around each stencil apply): KINETIC_CG_STEPS of Jacobi-preconditioned CG
        from x0 = D^-1 r, D = diag(T) + c, then the result is mapped back by W^(1/2)."""
        dtb = Rb.dtype
        if dtb not in cg_bufs:
            cg_bufs.clear()
            dinv_t = kdinv64.to(dtb)
            isw_t = (1.0 / sw64).to(dtb)
            cg_bufs[dtb] = (torch.empty((4, k, N), dtype=dtb, device=dev), dinv_t, w64.to(dtb), isw_t,
                            isw_t * dinv_t, sw64.to(dtb))
        buf, dinv, wd, iswd, isd, swd = cg_bufs[dtb]
        fuse_dots = FUSE_CG_DOTS and dtb == torch.float32 and N * 4 * 8 > FUSE_CG_MIN_GEOMETRY_BYTES
        kb = Rb.shape[0]
        x, r, pv, q = buf[0, :kb], buf[1, :kb], buf[2, :kb], buf[3, :kb]
        mark("rotation")
        torch.mul(Rb, isd, out=x)                       # x0 = D^-1 W^(-1/2) R
        kop.apply(x.view((kb,) + shape), q.view((kb,) + shape))
        rz = cg_start(Rb, iswd, q, dinv, wd, r, pv)     # r = W^(-1/2) R - A x0, p = D^-1 r
        for step in range(KINETIC_CG_STEPS):
            # q = A p with the (p, q)_W row dots from the same stencil pass where the fused
            # kernel takes the block
            pq = kop.apply_dots(pv.view((kb,) + shape), q.view((kb,) + shape), wd) if fuse_dots else None
            if pq is None:
                kop.apply(pv.view((kb,) + shape), q.view((kb,) + shape))
                pq = row_dots(pv, q, wd)
            alpha = torch.where(pq > 0, rz / torch.where(pq > 0, pq, 1.0), 0.0)
            if step == KINETIC_CG_STEPS - 1:
                cg_finish(x, pv, swd, alpha, Rb)        # R = W^(1/2) (x + alpha p)
                break
            rz_new = cg_update(x, r, pv, q, dinv, wd, alpha)
            beta = torch.where(rz > 0, rz_new / torch.where(rz > 0, rz, 1.0), 0.0)
            cg_direction(pv, r, dinv, beta)
            rz = rz_new
        else:
            torch.mul(x, swd, out=Rb)                   # no CG step: R = W^(1/2) x0
        mark("precondition")

    def gemm_nt(A, B):
        """A B^T as FP64 (see _gemm_nt_split)."""
        return _gemm_nt_split(A, B)

    def orthonormalize_rows(Z):
        """Z <- L^-1 Z with Z Z^T = L L^T, in place (column panels)."""
        G = gemm_nt(Z, Z)
        Linv = _chol_inv(G).to(dev)
        rotate(Linv.to(Z.dtype), Z, Z, None)

    # One buffer holds the rows [Z; R; HR; HZ] (basis, k residual rows, k rows for S R,
    # S Z).  The ka active residuals are compacted to the first rows of R and their S R
    # written right behind them, so [Z; R_a] and [R_a; HR_a] are contiguous: the basis
    # rotation is one GEMM and one pass of the Gram kernel gives Z R_a^T, Z HR_a^T,
    # R_a R_a^T and R_a HR_a^T together.  Before the S Z rotation HR_a moves to the last
    # ka rows of the HR block, so [HR_a; HZ] is contiguous too (no move when ka = k).  Two
    # buffers alternate between iterations, or one is rotated in place when two do not fit.
    def views(B):
        return B[:m], B[m:m + k], B[m + k:m + 2 * k], B[m + 2 * k:]

    rows = 2 * m + 2 * k

    def second_buffer(dtype):
        """A second buffer for the out-of-place rotations, or None for in-place panels."""
        if inplace_rotation is True:
            return None
        nbytes = rows * N * torch.finfo(dtype).bits // 8
        if inplace_rotation is None and dev.type == "cuda":
            free, _ = torch.cuda.mem_get_info(dev)
            free += torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev)  # cached blocks
            if nbytes > 0.45 * free:
                return None
        return torch.empty((rows, N), dtype=dtype, device=dev)

    def rotate(Qm, src, dst, spare):
        """dst = Qm src, out of place into the spare buffer's rows, or in place: dst
        overlaps src row-aligned, which the MFMA rotation kernel takes directly (each of
        its wavefronts reads all of src in its columns before writing them); wider Q
        falls back to column panels."""
        if spare is not None:
            _mm(Qm, src, out=dst)
            return
        if _rotate_kernel(Qm, src, dst):
            return
        for c0 in range(0, N, ROTATION_PANEL):
            c1 = min(c0 + ROTATION_PANEL, N)
            dst[:, c0:c1].copy_(torch.matmul(Qm, src[:, c0:c1]))

    if rows * N * 8 > BIG_BUFFER_BYTES and dev.type == "cuda":
        torch.cuda.empty_cache()  # whole free segments back, so one large buffer fits
    store = None
    if dt == torch.float32:
        # one FP64-sized storage: its two halves are the alternating FP32 buffers, and
        # after the promotion the FP64 buffer (no allocation at the promotion)
        store = torch.empty(rows * N, dtype=torch.float64, device=dev)
        halves = store.view(torch.float32)
        BUF = halves[rows * N:].view(rows, N)
        BUF_next = None if inplace_rotation is True else halves[:rows * N].view(rows, N)
        BUF[:m].copy_(X0.reshape(m, N) * sw64)
    else:
        BUF = torch.empty((rows, N), dtype=dt, device=dev)
        torch.mul(X0.reshape(m, N), sw64, out=BUF[:m])
        BUF_next = second_buffer(dt)
    orthonormalize_rows(BUF[:m])
    apply_s(BUF[:m], BUF[m + 2 * k:])

    res = BlockIterationResult(np.zeros(k), BUF[:k], np.full(k, np.inf), 0, False)
    phase32 = dt == torch.float32
    Psi = None
    it = 0
    # After each Rayleigh-Ritz step the basis is the Ritz basis of the (m + k) space:
    # Z HZ^T = diag(theta_all), Z Z^T = I.  Iterations in between re-measurements take
    # the Ritz vectors and values from there; every 16 iterations (and after the FP32 ->
    # FP64 promotion) both matrices are measured and the small eigenproblem solved.
    theta_all = None
    refresh = True
    for it in range(1, max_iter + 1):
        Z, R, HR, HZ = views(BUF)
        mark("rotation")
        if profile:
            res.iteration_seconds.append(clock[0])
        if refresh or it % 16 == 1:
            Gzz = _sym(gemm_nt(Z, Z))
            P = _sym(gemm_nt(Z, HZ))
            mark("gram")
            theta, V = _host_geigh(P, Gzz)
            mark("eigh")
            Ct = V[:, :k].T.to(dev, dt).contiguous()
            Psi = _mm(Ct, Z)
            _mm(Ct, HZ, out=R)
            th = theta[:k].numpy()
            R.addcmul_(Psi, torch.from_numpy(-th).to(dev, dt)[:, None])
            refresh = False
        else:
            Gzz = None
            P = None
            th = theta_all[:k]
            Psi = Z[:k]
            # R = HZ - theta Z in one pass
            torch.addcmul(HZ[:k], Psi, torch.from_numpy(-th).to(dev, dt)[:, None], out=R)
        rn = np.sqrt(np.clip(np.diag(gemm_nt(R, R).cpu().numpy()), 0.0, None))
        res.residual_history.append(rn)
        res.ritz_value_history.append(th)
        res.subspace_dimensions.append(m)
        ritz_k = P[:k, :k].cpu().numpy() if P is not None else np.diag(th)
        target = max(tol, fp32_floor) if phase32 else tol
        # FP32 stops helping once the residual stagnates near its rounding floor: the best
        # of the last 60 iterations is not 5 % below the best before them while the
        # residual is within FP32_STALL_CEILING of the floor.  A plateau far above the
        # floor (the slowest of many states, in exact arithmetic too) only promotes after
        # 150 iterations without progress.
        hist = res.residual_history

        def plateau(n):
            return len(hist) > n and min(r.max() for r in hist[-n:]) >= 0.95 * min(r.max() for r in hist[:-n])

        stalled = phase32 and ((rn.max() <= FP32_STALL_CEILING * fp32_floor and plateau(60)) or plateau(150))
        if rn.max() <= target or stalled:
            if phase32:
                # refinement: promote the basis to FP64 and rebuild S Z in FP64
                res.fp32_iterations = it
                phase32 = False
                dt = torch.float64
                sw, isw = setup(dt)
                # The FP32 basis goes to the storage's second half (whose rows [0, m) lie
                # past the bytes of the FP64 rows [0, m) since 8 m N < 4 (2m + 2k) N), is
                # widened into the FP64 buffer over the same storage, then re-orthonormalised.
                res.orbitals = Psi.to(torch.float64)
                half2 = halves[rows * N:].view(rows, N)
                if BUF.data_ptr() != half2.data_ptr():
                    half2[:m].copy_(Z)
                del Z, R, HR, HZ, Psi
                BUF = BUF_next = None
                BUF = store.view(rows, N)
                BUF[:m].copy_(half2[:m])
                del half2
                orthonormalize_rows(BUF[:m])
                BUF_next = second_buffer(dt)
                apply_s(BUF[:m], BUF[m + 2 * k:])
                refresh = True
                continue
            res.eigenvalues, res.orbitals, res.residual_norms = th, Psi, rn
            res.iterations, res.converged = it, True
            res.ritz_matrix = ritz_k
            break
        # soft locking: the active residual rows, compacted to the front of R in order
        # (row act[j] >= j, so no row is overwritten before it is read)
        act = np.nonzero(rn > LOCK_FRACTION * target)[0] if lock else np.arange(k)
        if len(act) == 0:
            act = np.arange(k)
        ka = len(act)
        res.active_rows.append(ka)
        for j, i in enumerate(act):
            if i != j:
                R[j].copy_(R[i])
        Ra, HRa = BUF[m:m + ka], BUF[m + ka:m + 2 * ka]
        mark("rotation")
        if hdiag is not None:
            den = hdiag[None, :] - torch.from_numpy(th[act]).to(dev)[:, None]
            Ra.div_(den.clamp_min_(PRECONDITIONER_FLOOR).to(dt))
        elif pmode == "kinetic":
            kinetic_cg(Ra)
        # project the residuals out of the basis and whiten them
        GRZ = gemm_nt(Ra, Z)
        mark("gram")
        # R_a <- R_a - (R_a Z^T) Z as one rotation [-R_a Z^T, I] [Z; R_a] into the rows of R_a
        # (the kernel reads all of its source columns before writing them)
        Qp = torch.cat([-GRZ, torch.eye(ka, dtype=GRZ.dtype, device=GRZ.device)], 1).to(dt)
        if not (Ra.is_cuda and _rotate_kernel(Qp, BUF[:m + ka], Ra)):
            Ra.sub_(_mm(GRZ.to(dt), Z))
        del Qp
        GR = gemm_nt(Ra, Ra).cpu()
        mark("gram")
        # |R_pre|^2 = |R|^2 + |R Z^T|^2 (Z orthonormal).  Rows that lost more than 99 % of
        # their norm to the projection carry its rounding error: they are projected once
        # more ("twice is enough").  A row still below DEPENDENT_REL * eps of its norm lies
        # in span(Z) to rounding and is dropped (zero row, unit pivot) rather than whitened
        # into noise.
        d = torch.diagonal(GR).clone()
        d_pre = d + (GRZ.cpu() ** 2).sum(1)
        again = d < 1e-4 * d_pre
        if bool(again.any()):
            res.reprojected_rows += int(again.sum())
            idx = again.nonzero().flatten().to(dev)
            Rf = Ra.index_select(0, idx)
            Rf.sub_(_mm(gemm_nt(Rf, Z).to(dt), Z))
            Ra.index_copy_(0, idx, Rf)
            del Rf
            GR = gemm_nt(Ra, Ra).cpu()
            d = torch.diagonal(GR).clone()
            mark("gram")
        dead = d <= (DEPENDENT_REL * torch.finfo(dt).eps) ** 2 * d_pre
        if bool(dead.any()):
            res.dropped_rows += int(dead.sum())
            Ra[dead.to(dev)] = 0
            GR[dead, :] = 0
            GR[:, dead] = 0
            GR[dead, dead] = 1.0
        Linv_r = _chol_inv(GR).to(dev)
        mark("eigh")
        if not (Ra.is_cuda and _rotate_kernel(Linv_r.to(dt), Ra, Ra)):  # in place
            Ra.copy_(_mm(Linv_r.to(dt), Ra))
        apply_s(Ra, HRa)
        G23 = gemm_nt(BUF[:m + ka], BUF[m:m + 2 * ka]).cpu()  # [Z; R_a] [R_a; HR_a]^T: (m + ka, 2ka)
        # R_a HZ^T: (ka, m); equal to (Z HR_a^T)^T up to rounding when S is symmetric
        G4 = None if symmetric else gemm_nt(Ra, HZ).cpu()
        mark("gram")
        h0 = m + 2 * k - ka  # first row of [HR_a; HZ] for the S Z rotation
        if not reapply and ka < k:
            # HR_a moves d = 2 (k - ka) rows down; overlapping ranges go back to front in
            # blocks of at most d rows (each block's source is read before it is covered)
            d = h0 - (m + ka)
            for e in range(ka, 0, -d):
                b0 = max(e - d, 0)
                BUF[h0 + b0:h0 + e].copy_(BUF[m + ka + b0:m + ka + e])
        ZR, ZHR = G23[:m, :ka], G23[:m, ka:]
        RR, RHR = G23[m:, :ka], G23[m:, ka:]
        Pz = P.cpu() if P is not None else torch.diag(torch.from_numpy(theta_all))
        Bz = Gzz.cpu() if Gzz is not None else torch.eye(m, dtype=torch.float64)
        T = _sym(torch.cat([torch.cat([Pz, ZHR], 1), torch.cat([ZHR.T if G4 is None else G4, RHR], 1)], 0))
        GB = _sym(torch.cat([torch.cat([Bz, ZR], 1), torch.cat([ZR.T, RR], 1)], 0))
        # generalized Rayleigh-Ritz on [Z; R_a]: c^T GB c = I, c^T T c = diag(theta)
        thT, c = _host_geigh(T, GB)
        if c.shape[1] < m + ka:
            res.truncated_steps += 1
        if c.shape[1] < m:
            raise RuntimeError("the Rayleigh-Ritz space lost rank below the working-subspace size")
        theta_all = thT[:m].numpy().copy()
        Q = c[:, :m].T.contiguous().to(dev)  # new basis = Q [Z; R_a]
        mark("eigh")
        Qd = Q.to(dt)
        out = BUF_next if BUF_next is not None else BUF
        # Z' = Q [Z; R_a] and HZ' = Q [HR_a; HZ] (columns of Q reordered).  In place, each
        # column panel of the sources is read whole before that panel is written, and the
        # two rotations touch disjoint rows ([0, m + ka) -> [0, m), [h0, end) -> [m + 2k, end)).
        rotate(Qd, BUF[:m + ka], out[:m], BUF_next)
        if not reapply:
            Qh = torch.cat([Qd[:, m:], Qd[:, :m]], 1)
            rotate(Qh, BUF[h0:], out[m + 2 * k:], BUF_next)
        if BUF_next is not None:
            BUF, BUF_next = BUF_next, BUF
        if reapply:
            apply_s(BUF[:m], BUF[m + 2 * k:])
        res.eigenvalues, res.orbitals, res.residual_norms = th, Psi, rn
        res.iterations = it
        res.two_buffers = res.two_buffers and BUF_next is not None
        res.ritz_matrix = ritz_k
    mark("rotation")
    if profile:
        t_it = res.iteration_seconds + [clock[0]]
        res.iteration_seconds = [round(b - a, 5) for a, b in zip(t_it[:-1], t_it[1:])]
    res.timings = dict(tm)
    sw, isw = setup(dt)
    if keep_basis:
        res.basis = BUF[:m] * isw
    res.orbitals = (res.orbitals * isw).reshape((k,) + shape)
    return res


# ------------------------------------------------------------------ default guesses


def default_guess_block(grid_geometry, atoms, count, dg=None):
    """Deterministic + rng(12345) guesses, W-orthonormalised (ks/eigensolver.py:818-879)."""
    from .grid import grid_kind, reference_center

    dg = dg or device_grid(grid_geometry)
    if grid_kind(grid_geometry) == "monitor":
        bb = dg.box_bounds
        center = tuple(0.5 * (b[0] + b[1]) for b in bb)
    else:
        center = tuple(reference_center(grid_geometry))
    xyz = dg.xyz()
    x, y, z = (xyz[i] - center[i] for i in range(3))
    r2 = x * x + y * y + z * z
    guesses = []
    gs = []
    for _, pos in atoms:
        d2 = (xyz[0] - pos[0]) ** 2 + (xyz[1] - pos[1]) ** 2 + (xyz[2] - pos[2]) ** 2
        gs.append(torch.exp(-0.85 * d2))
    if gs:
        guesses.append(sum(gs))
        if len(gs) >= 2:
            guesses.append(sum(g if i % 2 == 0 else -g for i, g in enumerate(gs)))
    base = torch.exp(-0.65 * r2)
    guesses += [base, z * torch.exp(-0.55 * r2), x * torch.exp(-0.55 * r2), y * torch.exp(-0.55 * r2),
                torch.exp(-1.10 * r2), (3.0 * z * z - r2) * torch.exp(-0.45 * r2)]
    total = max(count, 8)
    rng = np.random.default_rng(12345)
    while len(guesses) < total:
        f = torch.from_numpy(rng.normal(size=dg.shape).reshape(-1)).to(dg.device)
        guesses.append(base * f)
    block = torch.stack(guesses[:total])
    ortho = orthonormalize_device(block, dg.weights(), require_full_rank=False)
    if ortho.shape[0] < count:
        raise ValueError("Unable to build enough independent default guess orbitals for the fixed-potential solve.")
    return ortho[:count]


def initial_guess_block(grid_geometry, atoms, k, basis_size, initial_guess_orbitals=None, dg=None):
    """ks/eigensolver.py:882-913."""
    dg = dg or device_grid(grid_geometry)
    target = max(k, basis_size)
    default = default_guess_block(grid_geometry, atoms, target, dg)
    if initial_guess_orbitals is None:
        return default[:target]
    user = torch.as_tensor(np.asarray(initial_guess_orbitals, dtype=np.float64)).to(dg.device).reshape(-1, dg.N)
    comb = torch.cat([user, default], 0)
    ortho = orthonormalize_device(comb, dg.weights(), require_full_rank=False)
    if ortho.shape[0] < k:
        raise ValueError("The provided initial guesses do not span enough weighted-independent directions.")
    return ortho[:target]


def reference_block_size(k, initial_subspace_size=None, ncv=None):
    """Working-subspace size of ks/eigensolver_jax.py:81-87."""
    return max(2 * k + 2, 4, 12 * k + 8, 0 if initial_subspace_size is None else int(initial_subspace_size),
               0 if ncv is None else int(ncv))


# ------------------------------------------------------------------ operator contexts


@dataclass(frozen=True)
class FixedPotentialStaticLocalOperatorContext:
    """Fields of ks/eigensolver.py:80-106."""

    grid_geometry: object
    case: object
    spin_channel: str
    rho_up: np.ndarray
    rho_down: np.ndarray
    rho_total: np.ndarray
    local_ionic_potential: np.ndarray
    hartree_potential: np.ndarray
    xc_potential: np.ndarray
    effective_local_potential: np.ndarray
    local_ionic_evaluation: object
    frozen_patch_local_embedding: object
    hartree_poisson_result: object
    lsda_evaluation: object
    kinetic_version: str = "production"


@dataclass(frozen=True)
class FixedPotentialOperatorContext:
    """Fields of ks/eigensolver.py:60-77 (static KS: local + nonlocal + Hartree + XC)."""

    grid_geometry: object
    case: object
    spin_channel: str
    rho_up: np.ndarray
    rho_down: np.ndarray
    rho_total: np.ndarray
    local_ionic_potential: np.ndarray
    hartree_potential: np.ndarray
    xc_potential: np.ndarray
    effective_local_potential: np.ndarray
    local_ionic_evaluation: object
    hartree_poisson_result: object
    lsda_evaluation: object


@dataclass(frozen=True)
class FixedPotentialStaticLocalPreparationProfile:
    density_build_wall_time_seconds: float
    local_ionic_resolve_wall_time_seconds: float
    hartree_resolve_wall_time_seconds: float
    xc_resolve_wall_time_seconds: float
    total_wall_time_seconds: float


def _resolve_local(grid_geometry, dg, case, local_ionic_potential):
    from .hamiltonian import _resolve_local as rl

    return rl(grid_geometry, dg, case, local_ionic_potential)


def _density_pair(grid_geometry, dg, rho_up, rho_down):
    from .hamiltonian import _field

    up = _field(rho_up, dg, "rho_up")
    dn = _field(rho_down, dg, "rho_down")
    return up, dn, up + dn


def prepare_fixed_potential_static_local_operator_profiled(
        grid_geometry, rho_up, rho_down, spin_channel, case=H2_BENCHMARK_CASE, local_ionic_potential=None,
        hartree_potential=None, xc_potential=None, xc_functional=None, *, use_monitor_patch=False,
        patch_parameters=None, kinetic_version="production", base_local_ionic_evaluation=None,
        hartree_backend="python", use_jax_hartree_cached_operator=False, jax_hartree_cg_impl="baseline",
        jax_hartree_cg_preconditioner="none", jax_hartree_line_preconditioner_impl="baseline",
        monitor_boundary_construction_mode="corrected_moments", hartree_precision="fp64"):
    """ks/eigensolver.py:571-671: freeze T + V_loc + V_H + V_xc on one grid."""
    t0 = time.perf_counter()
    dg = device_grid(grid_geometry)
    spin = normalize_spin_channel(spin_channel)
    kv = normalize_kinetic_version(kinetic_version)
    up, dn, tot = _density_pair(grid_geometry, dg, rho_up, rho_down)
    t1 = time.perf_counter()
    embedding = None
    if local_ionic_potential is None and use_monitor_patch and dg.kind == "monitor":
        # ks/eigensolver.py:428-474: V_loc plus the frozen patch field of rho_total
        from . import pseudo

        embedding = pseudo.evaluate_monitor_grid_local_ionic_potential_with_frozen_patch_field(
            case, grid_geometry, tot, patch_parameters=patch_parameters,
            base_evaluation=base_local_ionic_evaluation)
        local, local_ev = embedding.corrected_total_local_potential, embedding.base_evaluation
    else:
        local, local_ev = _resolve_local(grid_geometry, dg, case, local_ionic_potential)
    t2 = time.perf_counter()
    vh, _, hres = resolve_hartree(grid_geometry, dg, tot, hartree_potential, backend=hartree_backend,
                                  monitor_boundary_construction_mode=monitor_boundary_construction_mode,
                                  precision=hartree_precision)
    t3 = time.perf_counter()
    vxc, lsda = resolve_xc(dg, up, dn, spin, xc_potential, xc_functional)
    t4 = time.perf_counter()
    ctx = FixedPotentialStaticLocalOperatorContext(
        grid_geometry, case, spin, up, dn, tot, np.asarray(local), np.asarray(vh), np.asarray(vxc),
        np.asarray(local + vh + vxc, dtype=np.float64), local_ev, embedding, hres, lsda, kv)
    prof = FixedPotentialStaticLocalPreparationProfile(t1 - t0, t2 - t1, t3 - t2, t4 - t3, time.perf_counter() - t0)
    return ctx, prof


def prepare_fixed_potential_static_local_operator(grid_geometry, rho_up, rho_down, spin_channel,
                                                  case=H2_BENCHMARK_CASE, local_ionic_potential=None,
                                                  hartree_potential=None, xc_potential=None, xc_functional=None,
                                                  **kwargs):
    """ks/eigensolver.py:674-726."""
    return prepare_fixed_potential_static_local_operator_profiled(
        grid_geometry, rho_up, rho_down, spin_channel, case, local_ionic_potential, hartree_potential,
        xc_potential, xc_functional, **kwargs)[0]


def prepare_fixed_potential_static_ks_operator(grid_geometry, rho_up, rho_down, spin_channel,
                                               case=H2_BENCHMARK_CASE, local_ionic_potential=None,
                                               hartree_potential=None, xc_potential=None, xc_functional=None):
    """ks/eigensolver.py:477-531 (the nonlocal term is applied per block by the operator)."""
    dg = device_grid(grid_geometry)
    spin = normalize_spin_channel(spin_channel)
    up, dn, tot = _density_pair(grid_geometry, dg, rho_up, rho_down)
    local, local_ev = _resolve_local(grid_geometry, dg, case, local_ionic_potential)
    vh, _, hres = resolve_hartree(grid_geometry, dg, tot, hartree_potential)
    vxc, lsda = resolve_xc(dg, up, dn, spin, xc_potential, xc_functional)
    return FixedPotentialOperatorContext(grid_geometry, case, spin, up, dn, tot, np.asarray(local), np.asarray(vh),
                                         np.asarray(vxc), np.asarray(local + vh + vxc, dtype=np.float64), local_ev,
                                         hres, lsda)


def block_hamiltonian(operator_context) -> BlockHamiltonian:
    """The device operator of a context (static-local: T + V_eff; static KS: + V_nl)."""
    from . import pseudo

    dg = device_grid(operator_context.grid_geometry)
    cache = getattr(dg, "_hops", None)
    if cache is None:
        cache = {}
        dg._hops = cache
    key = id(operator_context)
    hit = cache.get(key)
    if hit is not None and hit[0] is operator_context:
        return hit[1]
    veff = torch.from_numpy(np.ascontiguousarray(operator_context.effective_local_potential)).to(dg.device)
    if isinstance(operator_context, FixedPotentialOperatorContext):
        op = BlockHamiltonian(dg, veff, closure="production",
                              nonlocal_plan=pseudo.nonlocal_plan(dg, operator_context.case))
    else:
        op = BlockHamiltonian(dg, veff, closure=operator_context.kinetic_version)
    cache.clear()
    cache[key] = (operator_context, op)
    return op


def _apply_block_numpy(orbitals, operator_context):
    dg = device_grid(operator_context.grid_geometry)
    blk = np.asarray(orbitals, dtype=np.float64)
    single = blk.shape == dg.shape
    x = torch.from_numpy(np.ascontiguousarray(blk.reshape((-1,) + dg.shape))).to(dg.device)
    out = block_hamiltonian(operator_context).apply(x).cpu().numpy()
    return out[0] if single else out


def apply_fixed_potential_static_local_operator(psi, operator_context):
    """ks/eigensolver.py:729-756."""
    return _apply_block_numpy(psi, operator_context)


def apply_fixed_potential_static_local_block(orbitals, operator_context):
    """ks/eigensolver.py:773-784 (one fused block launch instead of a per-orbital loop)."""
    return _apply_block_numpy(orbitals, operator_context)


def apply_fixed_potential_static_ks_operator(psi, operator_context):
    """ks/eigensolver.py:534-556."""
    return _apply_block_numpy(psi, operator_context)


def apply_fixed_potential_static_ks_block(orbitals, operator_context):
    """ks/eigensolver.py:804-815."""
    return _apply_block_numpy(orbitals, operator_context)


# ------------------------------------------------------------------ results and entry points


@dataclass(frozen=True)
class FixedPotentialEigensolverResult:
    """Fields of ks/eigensolver.py:121-154."""

    target_orbitals: int
    solver_backend: str
    use_scipy_fallback: bool
    solver_method: str
    solver_note: str
    converged: bool
    iteration_count: int
    tolerance: float
    eigenvalues: np.ndarray
    orbitals: np.ndarray
    weighted_overlap: np.ndarray
    max_orthogonality_error: float
    residual_norms: np.ndarray
    residual_history: np.ndarray
    ritz_value_history: np.ndarray
    subspace_dimensions: tuple
    ritz_matrix: np.ndarray
    initial_guess_orbitals: np.ndarray
    final_basis_orbitals: np.ndarray
    operator_context: object
    static_local_preparation_profile: object = None
    use_jax_block_kernels: bool = False
    use_jax_cached_kernels: bool = False
    wall_time_seconds: float | None = None
    jax_internal_profile: object = None
    precision: str = "fp64"
    fp32_iterations: int = 0

    @property
    def internal_profile(self):
        return self.jax_internal_profile


_BACKENDS = ("auto", "jax", "scipy_fallback", "hip")


def _cell_volume_orthonormal(H, dg, X, eigenvalues):
    """Legacy grids: the iteration runs in the inner product that makes the operator
    self-adjoint (exact Ritz pairs); the returned orbitals are made orthonormal under the
    cell volumes as the reference's are (Gram-Schmidt in eigenvalue order, i.e. the
    Cholesky factor of the Gram matrix), with residual norms recomputed for them."""
    w = dg.weights()
    G = 0.5 * (gram(X, X, w) + gram(X, X, w).T)
    Linv = torch.linalg.inv(torch.linalg.cholesky(G))
    Xo = _mm(Linv.to(X.dtype), X.reshape(X.shape[0], -1)).reshape(X.shape)
    HX = torch.empty_like(Xo)
    H.apply(Xo.view((-1,) + dg.shape), HX.view((-1,) + dg.shape))
    lam = torch.as_tensor(np.asarray(eigenvalues, dtype=np.float64), device=X.device)
    R = HX.reshape(X.shape[0], -1) - lam[:, None] * Xo.reshape(X.shape[0], -1)
    res = torch.sqrt(torch.clamp((R * R * w[None, :]).sum(1), min=0.0)).cpu().numpy()
    return Xo, res


def _solve(operator_context, k, initial_guess_orbitals, max_iterations, tolerance, initial_subspace_size, ncv,
           backend, precision, profile, keep_basis=True):
    from . import pseudo

    if k <= 0:
        raise ValueError(f"k must be positive; received {k}.")
    if max_iterations <= 0:
        raise ValueError(f"max_iterations must be positive; received {max_iterations}.")
    if tolerance <= 0.0:
        raise ValueError(f"tolerance must be positive; received {tolerance}.")
    nb = str(backend).strip().lower()
    if nb not in _BACKENDS:
        raise ValueError(f"solver_backend must be one of {_BACKENDS}; received `{backend}`.")
    t0 = time.perf_counter()
    geom = operator_context.grid_geometry
    dg = device_grid(geom)
    m = reference_block_size(k, initial_subspace_size, ncv)
    X0 = initial_guess_block(geom, pseudo.case_atoms(operator_context.case), k, m, initial_guess_orbitals, dg)
    m = min(m, X0.shape[0])
    X0 = X0[:m]
    H = block_hamiltonian(operator_context)
    wop = dg.operator_weights()
    res = block_subspace_iteration(H, wop, X0, k, tolerance, max_iterations, precision=precision,
                                   keep_basis=keep_basis)
    orbs = res.orbitals.double()
    if dg.kind != "monitor":
        orbs, res.residual_norms = _cell_volume_orthonormal(H, dg, orbs, res.eigenvalues)
    S = gram(orbs, orbs, dg.weights()).cpu().numpy()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    c = lambda t: None if t is None else t.double().cpu().numpy().reshape((-1,) + dg.shape)  # noqa: E731
    return FixedPotentialEigensolverResult(
        target_orbitals=k, solver_backend="hip" if nb == "auto" else nb, use_scipy_fallback=False,
        solver_method=f"hip_block_subspace_iteration_{precision}",
        solver_note=("block subspace iteration with residual expansion and Rayleigh-Ritz restart on the GPU "
                     f"(working subspace {m}, precision {precision})"),
        converged=res.converged, iteration_count=res.iterations, tolerance=float(tolerance),
        eigenvalues=np.asarray(res.eigenvalues, dtype=np.float64), orbitals=c(orbs), weighted_overlap=S,
        max_orthogonality_error=float(np.max(np.abs(S - np.eye(k)))), residual_norms=np.asarray(res.residual_norms),
        residual_history=np.asarray(res.residual_history), ritz_value_history=np.asarray(res.ritz_value_history),
        subspace_dimensions=tuple(res.subspace_dimensions), ritz_matrix=res.ritz_matrix,
        initial_guess_orbitals=c(X0), final_basis_orbitals=c(res.basis), operator_context=operator_context,
        static_local_preparation_profile=profile, wall_time_seconds=wall, precision=precision,
        fp32_iterations=res.fp32_iterations)


def solve_fixed_potential_static_local_eigenproblem(
        grid_geometry, rho_up, rho_down, spin_channel, k, case=H2_BENCHMARK_CASE, initial_guess_orbitals=None,
        local_ionic_potential=None, hartree_potential=None, xc_potential=None, xc_functional=None,
        operator_context=None, max_iterations=400, tolerance=1.0e-3, initial_subspace_size=None, ncv=None, *,
        use_monitor_patch=False, patch_parameters=None, kinetic_version="production", use_jax_block_kernels=False,
        operator_preparation_profile=None, base_local_ionic_evaluation=None, hartree_backend="python",
        use_jax_hartree_cached_operator=False, jax_hartree_cg_impl="baseline", jax_hartree_cg_preconditioner="none",
        jax_hartree_line_preconditioner_impl="baseline", monitor_boundary_construction_mode="corrected_moments",
        solver_backend="auto", profile_jax_internals=False, precision="fp64"):
    """ks/eigensolver.py:1183-1320 on the GPU (every backend value runs the block iteration)."""
    if operator_context is None:
        operator_context, operator_preparation_profile = prepare_fixed_potential_static_local_operator_profiled(
            grid_geometry, rho_up, rho_down, spin_channel, case, local_ionic_potential, hartree_potential,
            xc_potential, xc_functional, use_monitor_patch=use_monitor_patch, patch_parameters=patch_parameters,
            kinetic_version=kinetic_version, base_local_ionic_evaluation=base_local_ionic_evaluation,
            hartree_backend=hartree_backend, monitor_boundary_construction_mode=monitor_boundary_construction_mode)
    return _solve(operator_context, k, initial_guess_orbitals, max_iterations, tolerance, initial_subspace_size,
                  ncv, solver_backend, precision, operator_preparation_profile)


def solve_fixed_potential_eigenproblem(grid_geometry, rho_up, rho_down, spin_channel, k, case=H2_BENCHMARK_CASE,
                                       initial_guess_orbitals=None, local_ionic_potential=None,
                                       hartree_potential=None, xc_potential=None, xc_functional=None,
                                       max_iterations=400, tolerance=1.0e-3, initial_subspace_size=None, ncv=None,
                                       *, precision="fp64"):
    """ks/eigensolver.py:1136-1180: the static-KS operator (with nonlocal projectors)."""
    ctx = prepare_fixed_potential_static_ks_operator(grid_geometry, rho_up, rho_down, spin_channel, case,
                                                     local_ionic_potential, hartree_potential, xc_potential,
                                                     xc_functional)
    return _solve(ctx, k, initial_guess_orbitals, max_iterations, tolerance, initial_subspace_size, ncv, "auto",
                  precision, None)


# ------------------------------------------------------------------ overlaps (NumPy surfaces)


def weighted_overlap_matrix(orbitals, grid_geometry, other=None, *, use_jax_block_kernels=False):
    """ks/eigensolver.py:246-272 (X^T W Y with FP64 accumulation on the GPU)."""
    dg = device_grid(grid_geometry)
    a = torch.from_numpy(np.ascontiguousarray(np.asarray(orbitals, dtype=np.float64))).to(dg.device)
    a = a.reshape(-1, dg.N)
    b = a if other is None else torch.from_numpy(
        np.ascontiguousarray(np.asarray(other, dtype=np.float64))).to(dg.device).reshape(-1, dg.N)
    return gram(a, b, dg.weights()).cpu().numpy()


def weighted_orbital_norms(orbitals, grid_geometry, *, use_jax_block_kernels=False):
    S = weighted_overlap_matrix(orbitals, grid_geometry)
    return np.sqrt(np.clip(np.diag(S), 0.0, None))


def weighted_orthonormalize_orbitals(orbitals, grid_geometry, *, require_full_rank=True, rank_tolerance=1.0e-12,
                                     use_jax_block_kernels=False):
    """ks/eigensolver.py:292-341."""
    dg = device_grid(grid_geometry)
    blk = np.asarray(orbitals, dtype=np.float64)
    if blk.shape[0] == 0:
        return blk.copy()
    x = torch.from_numpy(np.ascontiguousarray(blk.reshape(blk.shape[0], -1))).to(dg.device)
    out = orthonormalize_device(x, dg.weights(), rank_tolerance, require_full_rank)
    return out.cpu().numpy().reshape((-1,) + dg.shape)

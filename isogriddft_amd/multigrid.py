"""Multigrid preconditioner of the open-boundary Poisson solve on monitor grids.

The reference solves the interior system -L u = rhs with SciPy BiCGSTAB and no
preconditioner on monitor grids (poisson/open_boundary.py:1017-1064), or with CG and a line
preconditioner on its JAX path (poisson/poisson_jax.py:411-579).  The interior monitor
Laplacian takes 2h-wide centred differences in flux form: its diagonal terms couple node i
with i +- 2 e_a only, so the grid splits into eight parity sub-lattices (i mod 2 per axis)
joined only by the mixed-derivative terms.  This preconditioner is one V-cycle whose
coarse spaces respect that split:

  level 1       the interior nodes of one parity inside a 4 x 4 x 4 block of the grid
                form one aggregate (piecewise-constant prolongation P, restriction P^T)
  level l + 1   2 x 2 x 2 points of one parity of level l
  operators     the same-parity part of the Galerkin product P^T A P: eight independent
                7-point operators per level (the mixed terms are left to the fine smoother)
  coarsest      a dense inverse per parity once every extent is at most COARSEST
  smoothing     one damped-Jacobi sweep before and after the coarse correction on every
                level; the fine damping is 1.5 / lambda with lambda the spectral radius of
                D^-1 A from power iterations, the coarse one COARSE_OMEGA

The fine same-parity entries (diagonal and +-2 e_a couplings, closure included) are read
off the operator kernel itself by applying it to 28 probe vectors
(``probe_fine_entries``).  The hierarchy depends on the geometry only and is built once per
grid.  With it the benzene 224^3 monitor system needs a few tens of BiCGSTAB iterations
instead of ~280, and FP32 converges as fast as FP64 down to its rounding floor.  The
preconditioner changes the iterates, not the discrete solution the solve converges to.

Device kernels: ``csrc/multigrid.hip`` (ig_mg_* / ig_mg7_*); the fine operator applies are
the stencil kernel.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib

# stop coarsening once every extent of a level is at most this (dense inverse per parity)
COARSEST = 8
# damped-Jacobi factor on the coarse 7-point levels (their D^-1 A has spectral radius <= 2)
COARSE_OMEGA = 0.6
# fine damping = FINE_OMEGA_SCALE / (spectral radius of D^-1 A)
FINE_OMEGA_SCALE = 1.5
POWER_ITERATIONS = 20

# order of the 7-point coefficients: diagonal, -x, +x, -y, +y, -z, +z
OFFSETS = ((0, 0, 0), (-2, 0, 0), (2, 0, 0), (0, -2, 0), (0, 2, 0), (0, 0, -2), (0, 0, 2))


PROBES = 28


def probe_fine_entries(shape, matvec, device):
    """A[i, i], A[i, i -+ 2 e_a] of the interior operator (7, nx, ny, nz) FP64, read off
    ``matvec(x, out)`` (out = A x on interior nodes, 0 on the faces) with 28 probe vectors.

    A node i of parity p = i & 1 sits on sub-lattice point s = i >> 1.  Within one parity
    class the operator couples s with s +- e_a only (the 2h-wide differences), and the
    colour (sx + 2 sy + 3 sz) mod 7 differs between any two points of such a 7-point
    neighbourhood.  No entry couples parities that differ in all three axes (the mixed
    terms change two parities, the one-sided face closure one), so classes p and
    p ^ (1, 1, 1) share a probe: probe q = 7 pair + colour holds the interior nodes of that
    pair and colour, and row i of A probe_q sees one probe node among its nonzero columns,
    i + o, whenever its pair matches and colour(s_i + o / 2) = colour (kernels
    ig_mg_probe_vector_f64 / ig_mg_probe_read_f64)."""
    if torch.device(device).type != "cuda":
        raise ValueError("probing runs the device kernels: device must be a GPU")
    nx, ny, nz = shape
    ent = torch.zeros((7,) + tuple(shape), dtype=torch.float64, device=device)
    e = torch.empty(shape, dtype=torch.float64, device=device)
    out = torch.empty(shape, dtype=torch.float64, device=device)
    h = _lib.stream_handle()
    for q in range(PROBES):
        _lib.call("ig_mg_probe_vector_f64", _lib.ptr(e), q, nx, ny, nz, h)
        matvec(e.view(-1), out.view(-1))
        _lib.call("ig_mg_probe_read_f64", _lib.ptr(ent), _lib.ptr(out), q, nx, ny, nz, h)
    return ent


def _galerkin(ent):
    """Same-parity Galerkin sums over aggregates of two children per axis.

    ent: (7, 2, 2, 2, *rest) entries of the finer level with the child offsets (ex, ey, ez)
    (0 / 1) leading; returns (7, *rest) indexed by aggregate.  A coupling towards +a from
    a child with e_a = 0 stays inside the aggregate (diagonal), one from e_a = 1 goes to
    aggregate + 1; towards -a the other way round."""
    kids = (0, 1, 2)

    def part(t, axis, e):
        return t.narrow(axis, e, 1).sum(dim=kids)

    d, xm, xp, ym, yp, zm, zp = ent
    diag = d.sum(dim=kids)
    out = []
    for axis, (m_, p_) in enumerate(((xm, xp), (ym, yp), (zm, zp))):
        diag = diag + part(p_, axis, 0) + part(m_, axis, 1)
        out += [part(m_, axis, 0), part(p_, axis, 1)]
    return torch.stack([diag] + out)


def level_one(ent):
    """Level-1 7-point operators (7, 8, cx, cy, cz) from the fine entries (7, nx, ny, nz):
    fine node i = 4 S + 2 e + p (per axis) belongs to aggregate S of parity p.  Raw sums
    (aggregates without interior nodes have a zero diagonal), accumulated slab by slab
    (one strided slab of the fine entries per parity and child offset: no temporaries of
    the grid's size), in a fixed order."""
    _, nx, ny, nz = ent.shape
    c = [(n + 3) // 4 for n in (nx, ny, nz)]
    out = torch.zeros((7, 8) + tuple(c), dtype=ent.dtype, device=ent.device)
    for p in range(8):
        px, py, pz = p >> 2, (p >> 1) & 1, p & 1
        for ex in (0, 1):
            for ey in (0, 1):
                for ez in (0, 1):
                    slab = ent[:, px + 2 * ex::4, py + 2 * ey::4, pz + 2 * ez::4]
                    lx, ly, lz = slab.shape[1:]
                    if lx == 0 or ly == 0 or lz == 0:
                        continue
                    o = out[:, p, :lx, :ly, :lz]
                    # diagonal: the node's own entry and its couplings that stay inside
                    o[0] += slab[0]
                    o[0] += slab[1 if ex else 2]
                    o[0] += slab[3 if ey else 4]
                    o[0] += slab[5 if ez else 6]
                    # couplings that leave the aggregate: -a from e_a = 0, +a from e_a = 1
                    o[2 if ex else 1] += slab[2 if ex else 1]
                    o[4 if ey else 3] += slab[4 if ey else 3]
                    o[6 if ez else 5] += slab[6 if ez else 5]
    return out


def next_level(coef):
    """Level l + 1 operators from the raw level-l ones (7, 8, cx, cy, cz): 2 x 2 x 2
    children per aggregate within each parity."""
    _, _, cx, cy, cz = coef.shape
    d = [(n + 1) // 2 for n in (cx, cy, cz)]
    pad = torch.zeros((7, 8, 2 * d[0], 2 * d[1], 2 * d[2]), dtype=coef.dtype, device=coef.device)
    pad[:, :, :cx, :cy, :cz] = coef
    # (7, p, Sx, ex, Sy, ey, Sz, ez) -> (7, ex, ey, ez, p, Sx, Sy, Sz)
    v = pad.view(7, 8, d[0], 2, d[1], 2, d[2], 2).permute(0, 3, 5, 7, 1, 2, 4, 6).contiguous()
    return _galerkin(v).contiguous()


def finished(coef):
    """The operator the kernels use: aggregates without interior nodes (zero diagonal)
    get diagonal 1 (their right-hand side is always zero, so their value stays zero)."""
    out = coef.clone()
    out[0] = torch.where(coef[0] == 0.0, torch.ones_like(coef[0]), coef[0])
    return out


def dense_blocks(coef):
    """(8, m, m) dense matrices of the eight 7-point operators of one level (m = cx cy cz)."""
    c = coef.detach().cpu().numpy()
    _, _, cx, cy, cz = c.shape
    m = cx * cy * cz
    A = np.zeros((8, m, m))
    idx = np.arange(m).reshape(cx, cy, cz)
    for p in range(8):
        A[p, idx.reshape(-1), idx.reshape(-1)] = c[0, p].reshape(-1)
        for k, (axis, step) in enumerate(((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)), start=1):
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            if step < 0:
                src[axis], dst[axis] = slice(1, None), slice(None, -1)
            else:
                src[axis], dst[axis] = slice(None, -1), slice(1, None)
            rows = idx[tuple(src)].reshape(-1)
            cols = idx[tuple(dst)].reshape(-1)
            A[p, rows, cols] = c[k, p][tuple(src)].reshape(-1)
    return A


def build_levels(ent):
    """All coarse operators (FP64, list of (7, 8, cx, cy, cz)) and the coarsest dense
    inverses (8, m, m) from the fine entries."""
    raw = [level_one(ent)]
    while max(raw[-1].shape[2:]) > COARSEST:
        raw.append(next_level(raw[-1]))
    levels = [finished(c) for c in raw]
    # torch's host LAPACK: a second BLAS thread pool in the process (NumPy's) would slow
    # the eigensolver's small host solves that run later
    inv = torch.linalg.inv(torch.from_numpy(dense_blocks(levels[-1])))
    return levels, inv.to(ent.device)


class PoissonMultigrid:
    """V-cycle preconditioner M ~ A^-1 of the interior monitor Poisson operator.

    ``matvec(x, out)`` applies A (the stencil kernel, interior rows, zero faces) in the
    dtype of x; the hierarchy is FP64 and cast once per dtype for the cycle.  ``entries``
    (7, nx, ny, nz) skips the probing (the fine entries from elsewhere)."""

    def __init__(self, shape, matvec, device, entries=None):
        self.shape = tuple(shape)
        self.n = int(np.prod(self.shape))
        self.matvec = matvec
        self.device = torch.device(device)
        ent = probe_fine_entries(self.shape, matvec, self.device) if entries is None else entries
        diag = ent[0].reshape(-1)
        self.dinv64 = torch.where(diag != 0.0, 1.0 / torch.where(diag != 0.0, diag, 1.0), 0.0)
        self.levels64, self.inv64 = build_levels(ent)
        self.lam = self._spectral_radius()
        self.omega = FINE_OMEGA_SCALE / self.lam
        self._bufs = {}

    def _spectral_radius(self):
        """Power iterations on D^-1 A from a seeded random interior vector (one host read)."""
        g = torch.Generator(device=self.device).manual_seed(1234)
        v = torch.randn(self.n, generator=g, dtype=torch.float64, device=self.device) * (self.dinv64 != 0)
        w = torch.empty_like(v)
        for _ in range(POWER_ITERATIONS):
            v /= torch.linalg.vector_norm(v)
            self.matvec(v, w)
            w *= self.dinv64
            v, w = w, v
        return float(torch.linalg.vector_norm(v).item())

    def buffers(self, dtype):
        """Per-dtype coefficients and work vectors (allocated once)."""
        b = self._bufs.get(dtype)
        if b is None:
            dev = self.device
            dims = [tuple(int(s) for s in c.shape[2:]) for c in self.levels64]
            b = {
                "dw": (self.omega * self.dinv64).to(dtype),
                "t": torch.empty(self.n, dtype=dtype, device=dev),
                "coef": [c.to(dtype).contiguous() for c in self.levels64],
                "inv": self.inv64.to(dtype).contiguous(),
                "dims": dims,
                "rhs": [torch.empty(8 * int(np.prod(d)), dtype=dtype, device=dev) for d in dims],
                "xa": [torch.empty(8 * int(np.prod(d)), dtype=dtype, device=dev) for d in dims],
                "xb": [torch.empty(8 * int(np.prod(d)), dtype=dtype, device=dev) for d in dims],
            }
            self._bufs[dtype] = b
        return b

    def scaled_inverse_diagonal(self, dtype):
        """omega D^-1 (flat, zero on the faces): x = dw * b is the V-cycle's first sweep."""
        return self.buffers(dtype)["dw"]

    @_lib.on_stream
    def finish(self, b, x, stream=None):
        """Complete the V-cycle x ~ A^-1 b from its first fine sweep x = dw * b (in place),
        every launch on ``stream`` if given, else on the current stream."""
        dt = x.dtype
        B = self.buffers(dt)
        sfx = "f64" if dt == torch.float64 else "f32"
        h = _lib.stream_handle()
        nx, ny, nz = self.shape
        cx, cy, cz = B["dims"][0]
        t = B["t"]
        self.matvec(x, t)
        _lib.call(f"ig_mg_fine_restrict_{sfx}", _lib.ptr(b), _lib.ptr(t), nx, ny, nz, _lib.ptr(B["rhs"][0]), cx, cy,
                  cz, h)
        e = self._coarse(B, 0, sfx, h)
        _lib.call(f"ig_mg_fine_prolong_{sfx}", _lib.ptr(x), _lib.ptr(e), nx, ny, nz, cx, cy, cz, h)
        self.matvec(x, t)
        _lib.call(f"ig_mg_smooth_{sfx}", _lib.ptr(x), _lib.ptr(b), _lib.ptr(t), _lib.ptr(B["dw"]), self.n, h)

    @_lib.on_stream
    def apply(self, b, out, stream=None):
        """out = M b (one V-cycle from zero), on ``stream`` if given."""
        torch.mul(b, self.scaled_inverse_diagonal(b.dtype), out=out)
        self.finish(b, out)

    def _coarse(self, B, lev, sfx, h):
        """V-cycle on level lev + 1 (coarse index lev) for the right-hand side in
        B["rhs"][lev]; returns the tensor holding the correction."""
        rhs = B["rhs"][lev]
        if lev == len(B["dims"]) - 1:
            x = B["xa"][lev]
            torch.bmm(B["inv"], rhs.view(8, -1, 1), out=x.view(8, -1, 1))
            return x
        cx, cy, cz = B["dims"][lev]
        dx, dy, dz = B["dims"][lev + 1]
        coef = B["coef"][lev]
        xa, xb = B["xa"][lev], B["xb"][lev]
        _lib.call(f"ig_mg7_jacobi_{sfx}", _lib.ptr(xa), None, _lib.ptr(rhs), _lib.ptr(coef), COARSE_OMEGA, cx, cy,
                  cz, h)
        _lib.call(f"ig_mg7_restrict_{sfx}", _lib.ptr(B["rhs"][lev + 1]), _lib.ptr(rhs), _lib.ptr(xa), _lib.ptr(coef),
                  cx, cy, cz, dx, dy, dz, h)
        e = self._coarse(B, lev + 1, sfx, h)
        _lib.call(f"ig_mg7_prolong_{sfx}", _lib.ptr(xa), _lib.ptr(e), cx, cy, cz, dx, dy, dz, h)
        _lib.call(f"ig_mg7_jacobi_{sfx}", _lib.ptr(xb), _lib.ptr(xa), _lib.ptr(rhs), _lib.ptr(coef), COARSE_OMEGA, cx,
                  cy, cz, h)
        return xb

"""Adjoints of the inner-loop operators behind torch.autograd.Function.

The reference has no derivatives; these are the derivative paths a gradient-based
caller (geometry optimisation, potential inversion, learned functionals) needs, built
on the same HIP kernels as the forward path:

* ``HPsi``            H psi = -1/2 Lap psi + V psi for a block.  Backward uses the exact
                      Euclidean transpose of the stencil (the TR instantiation of the
                      monitor kernel, the transposed legacy kernel): d/dpsi = H^T g,
                      d/dV = sum_m psi_m g_m.
* ``RitzValues``      lowest-k eigenvalues of the fixed-potential operator T + V (block
                      iteration); backward is first-order perturbation theory
                      (Hellmann-Feynman): d eps_i / d V_p = w_p psi_i(p)^2 with Ritz
                      vectors normalised under the weights w that make T + V self-adjoint
                      (cell volumes on monitor grids, (h s_x)(h s_y)(h s_z) on legacy
                      grids), i.e. dE/dv_ext for E = sum_i g_i eps_i.
* ``OrbitalDensity``  rho = sum_i f_i psi_i^2 of the lowest-k states of T + V; backward
                      is the density response: one Sternheimer solve per occupied state
                      ((H - eps_i) z_i = -P g psi_i / w, projected CG on the complement of
                      the occupied states), dL/dV = sum_i 2 f_i w psi_i z_i.
* ``HartreePotential`` v_H = P rho (open-boundary Poisson); backward is the adjoint
                      solve P^T g (transposed stencil in BiCGSTAB, transposed multipole
                      faces and moments), giving dE_H / d rho and the Hartree part of
                      dE/dv_ext chains.
* ``IonicEnergy``     E_ion(R) = E_loc + E_nl + E_II at frozen density and orbitals as a
                      function of the atomic positions; backward gives dE/dR (minus the
                      Hellmann-Feynman forces): ig_vloc_gth_grad_f64 for the local part,
                      projector-position derivative fields for the nonlocal part, the
                      Coulomb sum for E_II.
"""

from __future__ import annotations

import numpy as np
import torch

from . import pseudo
from .device import DeviceGrid
from .kinetic import apply_operator
from .reductions import wdot


class HPsi(torch.autograd.Function):
    """out = -1/2 Lap(psi) + veff * psi on one grid (psi: (k, *shape) or (*shape))."""

    @staticmethod
    def forward(ctx, psi, veff, dg: DeviceGrid, closure="production"):
        ctx.dg = dg
        ctx.closure = closure
        ctx.save_for_backward(psi, veff)
        v = veff.reshape(dg.shape)
        return apply_operator(dg, psi.detach(), scale=-0.5, closure=closure, veff=v.detach())

    @staticmethod
    def backward(ctx, g):
        psi, veff = ctx.saved_tensors
        dg = ctx.dg
        g = g.contiguous()
        grad_psi = grad_v = None
        if ctx.needs_input_grad[0]:
            grad_psi = apply_operator(dg, g, scale=-0.5, closure=ctx.closure,
                                      veff=veff.detach().reshape(dg.shape), transpose=True)
        if ctx.needs_input_grad[1]:
            prod = (psi.detach() * g).reshape(-1, dg.N).sum(0)
            grad_v = prod.reshape(veff.shape)
        return grad_psi, grad_v, None, None


def hpsi(psi, veff, dg, closure="production"):
    return HPsi.apply(psi, veff, dg, closure)


class RitzValues(torch.autograd.Function):
    """eps_1..k of T + V from a starting block; backward: w * sum_i g_i psi_i^2."""

    @staticmethod
    def forward(ctx, veff, dg: DeviceGrid, X0, k, tol=1e-8, closure="trial_fix", max_iter=400):
        from .eigensolver import block_subspace_iteration
        from .hamiltonian import BlockHamiltonian

        H = BlockHamiltonian(dg, veff.detach().reshape(-1), closure=closure)
        res = block_subspace_iteration(H, dg.operator_weights(), X0, k, tol, max_iter)
        if not res.converged:
            raise RuntimeError("the block eigensolver did not converge inside the autograd forward")
        psi = res.orbitals.reshape(k, -1).to(torch.float64)
        ctx.save_for_backward(psi)
        ctx.dg = dg
        ctx.vshape = veff.shape
        return torch.as_tensor(np.asarray(res.eigenvalues), dtype=torch.float64, device=veff.device)

    @staticmethod
    def backward(ctx, g):
        (psi,) = ctx.saved_tensors
        # psi is normalised in the inner product that makes T + V self-adjoint
        w = ctx.dg.operator_weights()
        grad = w * (g.to(psi.device)[:, None] * psi * psi).sum(0)
        return grad.reshape(ctx.vshape), None, None, None, None, None, None


def ritz_values(veff, dg, X0, k, tol=1e-8, closure="trial_fix"):
    return RitzValues.apply(veff, dg, X0, k, tol, closure)


class OrbitalDensity(torch.autograd.Function):
    """rho = sum_i f_i psi_i^2 of the lowest-k states of T + V; backward is the density
    response by Sternheimer solves (recomputed, nothing stored but the orbitals).

    With W the weights that make T + V self-adjoint and psi W-orthonormal, a change dV
    moves psi_i by sum_{j != i} psi_j <psi_j, dV psi_i>_W / (eps_i - eps_j), so for a loss
    L = sum_p g_p rho_p
        dL/dV_p = sum_i 2 f_i w_p psi_i(p) z_i(p),
        (H - eps_i) z_i = -P (g psi_i / w),   z_i in range(P),
    P the W-orthogonal projector off the occupied states (f > 0).  Pairs of occupied
    states contribute (f_i - f_j) / (eps_i - eps_j) <psi_j, g psi_i> w psi_i psi_j, which
    vanishes for equal occupations and is added explicitly otherwise.  On range(P)
    H - eps_i is SPD when the occupied states lie below the rest (a gap), so each z_i
    comes from Jacobi-preconditioned CG in the W inner product, all rows at once with the
    row-wise HIP kernels and one stencil apply per step.
    """

    @staticmethod
    def forward(ctx, veff, dg: DeviceGrid, X0, k, occupations, tol=1e-10, closure="trial_fix", max_iter=400,
                response_tol=1e-10, response_max_iter=2000):
        from .eigensolver import block_subspace_iteration
        from .hamiltonian import BlockHamiltonian

        f = torch.as_tensor(np.asarray(occupations, dtype=np.float64))
        if f.numel() != k or bool((f < 0).any()):
            raise ValueError("occupations must be k non-negative numbers")
        H = BlockHamiltonian(dg, veff.detach().reshape(-1), closure=closure)
        res = block_subspace_iteration(H, dg.operator_weights(), X0, k, tol, max_iter)
        if not res.converged:
            raise RuntimeError("the block eigensolver did not converge inside the autograd forward")
        psi = res.orbitals.reshape(k, -1).to(torch.float64)
        fd = f.to(psi.device)
        ctx.save_for_backward(psi, fd)
        ctx.eps = np.asarray(res.eigenvalues, dtype=np.float64)
        ctx.H, ctx.dg, ctx.vshape = H, dg, veff.shape
        ctx.rtol, ctx.rmax = response_tol, response_max_iter
        return (fd[:, None] * psi * psi).sum(0).reshape(veff.shape)

    @staticmethod
    def backward(ctx, g):
        psi, f = ctx.saved_tensors
        z, info = sternheimer_solve(ctx.H, psi, ctx.eps, f, g.reshape(-1).to(torch.float64), ctx.rtol, ctx.rmax)
        if not info["converged"]:
            raise RuntimeError(f"the Sternheimer solves did not converge: {info}")
        dg = ctx.dg
        w = dg.operator_weights()
        occ = (f > 0).nonzero().flatten()
        grad = 2.0 * w * (f[occ, None] * psi[occ] * z).sum(0)
        # occupied pairs with different occupations
        gf = g.reshape(-1).to(torch.float64)
        fo = f[occ].cpu().numpy()
        eo = ctx.eps[occ.cpu().numpy()]
        for a in range(len(occ)):
            for b in range(a + 1, len(occ)):
                if fo[a] == fo[b]:
                    continue
                if abs(eo[a] - eo[b]) < 1e-12:
                    raise RuntimeError("degenerate occupied levels with different occupations")
                pa, pb = psi[occ[a]], psi[occ[b]]
                coef = 2.0 * (fo[a] - fo[b]) / (eo[a] - eo[b]) * float((gf * pa * pb).sum())
                grad = grad + coef * w * pa * pb
        return grad.reshape(ctx.vshape), None, None, None, None, None, None, None, None, None


RESPONSE_SHIFT = 0.5


def sternheimer_solve(H, psi, eps, f, g, tol=1e-10, max_iter=2000):
    """z_i for the occupied rows (f_i > 0): (H - eps_i) z_i = -P (g psi_i / w) on range(P),
    P = I - Psi_occ Psi_occ^T W.  Projected Jacobi-PCG in the W inner product with
    D = diag(T) + RESPONSE_SHIFT; r stays in range(P), so (r, P D^-1 r)_W = (r, D^-1 r)_W
    and the fused row kernels apply, with the direction projected after each update.
    Returns (z (nocc, N), info)."""
    from .reductions import cg_direction, cg_update, gram, row_dots

    dg = H.dg
    N = dg.N
    shape = dg.shape
    w = dg.operator_weights()
    occ = (f > 0).nonzero().flatten()
    Po = psi[occ].contiguous()
    nocc = Po.shape[0]
    e = torch.as_tensor(np.asarray(eps, dtype=np.float64)[occ.cpu().numpy()], device=psi.device)
    dinv = 1.0 / (-0.5 * dg.kinetic_diagonal() + RESPONSE_SHIFT)

    def project(y):
        C = gram(y, Po, w)                       # (rows, nocc) = y W Po^T
        y.sub_(C @ Po)

    def apply_shifted(pv, out):
        H.apply(pv.view((nocc,) + shape), out.view((nocc,) + shape))
        out.addcmul_(e[:, None], pv, value=-1.0)
        project(out)

    u = (g[None, :] * Po) / w[None, :]
    # convergence is measured against the unprojected right-hand side, so a row whose
    # projection leaves only rounding (e.g. g proportional to w) counts as solved
    thresh = tol * tol * row_dots(u, u * dinv, w)
    b = -u
    del u
    project(b)
    x = torch.zeros_like(b)
    r = b
    pv = r * dinv
    project(pv)
    q = torch.empty_like(b)
    rz = row_dots(r, r * dinv, w)
    it = 0
    converged = bool((rz <= thresh).all())
    while not converged and it < max_iter:
        active = rz > thresh
        apply_shifted(pv, q)
        pq = row_dots(pv, q, w)
        if bool((pq[active] <= 0).any()):
            return x, {"converged": False, "iterations": it, "reason": "operator not positive on the complement"}
        alpha = torch.where(active & (pq > 0), rz / torch.where(pq > 0, pq, 1.0), 0.0)
        rz_new = cg_update(x, r, pv, q, dinv, w, alpha)
        beta = torch.where(active & (rz > 0), rz_new / torch.where(rz > 0, rz, 1.0), 0.0)
        cg_direction(pv, r, dinv, beta)
        project(pv)
        rz = rz_new
        it += 1
        if it % 10 == 0 or it == max_iter:
            converged = bool((rz <= thresh).all())
    rel = torch.sqrt(rz / torch.where(thresh > 0, thresh / (tol * tol), 1.0))
    return x, {"converged": bool((rz <= thresh).all()), "iterations": it, "relative_residual": float(rel.max())}


def orbital_density(veff, dg, X0, k, occupations, tol=1e-10, closure="trial_fix", response_tol=1e-10):
    """sum_i f_i psi_i^2 of the lowest-k states of T + veff, differentiable in veff."""
    return OrbitalDensity.apply(veff, dg, X0, k, occupations, tol, closure, 400, response_tol, 2000)


class HartreePotential(torch.autograd.Function):
    """v_H = P rho, the open-boundary Poisson map (multipole faces + interior solve);
    backward: P^T g by the adjoint solve of poisson.poisson_adjoint_device (BiCGSTAB on
    the transposed stencil, recomputed rather than stored)."""

    @staticmethod
    def forward(ctx, rho, dg: DeviceGrid, multipole_order=2, mode="corrected_moments", tol=1e-10):
        from .poisson import solve_poisson_device

        sol = solve_poisson_device(dg, rho.detach().reshape(-1).to(torch.float64), multipole_order=multipole_order,
                                   tolerance=tol, max_iterations=4000, mode=mode)
        ctx.dg, ctx.order, ctx.mode, ctx.tol = dg, multipole_order, mode, tol
        return sol["potential"].reshape(rho.shape)

    @staticmethod
    def backward(ctx, g):
        from .poisson import poisson_adjoint_device

        grad, info = poisson_adjoint_device(ctx.dg, g, ctx.order, ctx.mode, ctx.tol)
        if not info["converged"]:
            raise RuntimeError(f"the adjoint Poisson solve did not converge: {info}")
        return grad.reshape(g.shape), None, None, None, None


def hartree_potential(rho, dg, multipole_order=2, mode="corrected_moments", tol=1e-10):
    return HartreePotential.apply(rho, dg, multipole_order, mode, tol)


def hartree_energy(rho, dg, multipole_order=2, mode="corrected_moments", tol=1e-10):
    """E_H = 1/2 sum w rho v_H[rho], differentiable in rho (dE_H / d rho)."""
    v = hartree_potential(rho, dg, multipole_order, mode, tol)
    return 0.5 * (dg.weights() * rho.reshape(-1) * v.reshape(-1)).sum()


def _ion_ion(elements, R):
    e = 0.0
    g = np.zeros_like(R)
    for i in range(len(elements)):
        for j in range(i + 1, len(elements)):
            zi = pseudo.GTH_PADE[elements[i]].ionic_charge
            zj = pseudo.GTH_PADE[elements[j]].ionic_charge
            d = R[i] - R[j]
            r = float(np.linalg.norm(d))
            e += zi * zj / r
            g[i] -= zi * zj * d / r ** 3
            g[j] += zi * zj * d / r ** 3
    return e, g


class IonicEnergy(torch.autograd.Function):
    """E_loc + E_nl + E_II at frozen rho / orbitals as a function of positions (natoms, 3)."""

    @staticmethod
    def forward(ctx, positions, dg: DeviceGrid, elements, rho, orbitals, occupations):
        R = positions.detach().to("cpu", torch.float64).numpy()
        atoms = [(el, tuple(float(c) for c in r)) for el, r in zip(elements, R)]
        rho = rho.reshape(-1).to(torch.float64)
        vloc = pseudo.local_potential_device(dg, atoms)
        e_loc = wdot(dg, rho, vloc)
        plan = pseudo.NonlocalPlan(dg, atoms)
        occ = np.asarray(occupations, dtype=np.float64)
        X = orbitals.reshape(-1, dg.N)[:len(occ)].to(torch.float64).contiguous()
        e_nl = float(np.dot(occ, plan.energy_terms(X))) if not plan.empty else 0.0
        e_ii, g_ii = _ion_ion(list(elements), R)
        ctx.dg, ctx.atoms, ctx.elements = dg, atoms, list(elements)
        ctx.plan, ctx.occ, ctx.g_ii = plan, occ, g_ii
        ctx.save_for_backward(rho, X)
        ctx.pdevice = positions.device
        return torch.tensor(e_loc + e_nl + e_ii, dtype=torch.float64, device=positions.device)

    @staticmethod
    def backward(ctx, g):
        rho, X = ctx.saved_tensors
        dg = ctx.dg
        grad = pseudo.local_potential_gradient_device(dg, ctx.atoms, rho * dg.weights()) + ctx.g_ii
        plan = ctx.plan
        if not plan.empty:
            nx, ny, nz = dg.shape
            from . import _lib
            from .reductions import workspace

            lib = _lib.load()
            k = X.shape[0]
            for ap, coef, d in plan.project(X):
                for axis in range(3):
                    dP = ap.field_derivative(dg, axis)
                    work = workspace(dg.device, lib.ig_nonlocal_work_size(k, ap.nproj), "nl")
                    dc = torch.empty((k, ap.nproj), dtype=torch.float64, device=dg.device)
                    dd = torch.empty_like(dc)
                    _lib.call("ig_nonlocal_project_f64", _lib.ptr(X), X.stride(0), k, nx, ny, nz, ap.box,
                              _lib.ptr(dP), ap.nproj, _lib.ptr(dg.weights()), _lib.ptr(ap.H), _lib.ptr(work),
                              _lib.ptr(dc), _lib.ptr(dd), _lib.stream_handle())
                    # d/dR <psi|p> h <p|psi> = 2 <psi|dp> h <p|psi>  (h symmetric)
                    per = 2.0 * (dc * d).sum(dim=1).cpu().numpy()
                    grad[ap.atom_index, axis] += float(np.dot(ctx.occ, per))
        gt = torch.as_tensor(grad, dtype=torch.float64, device=ctx.pdevice) * g
        return gt, None, None, None, None, None


def ionic_energy(positions, dg, elements, rho, orbitals, occupations):
    return IonicEnergy.apply(positions, dg, elements, rho, orbitals, occupations)


def atomic_forces(dg, elements, positions, rho, orbitals, occupations):
    """Hellmann-Feynman forces -dE_ion/dR (natoms, 3) at frozen rho and orbitals."""
    R = torch.as_tensor(np.asarray(positions, dtype=np.float64)).requires_grad_(True)
    e = ionic_energy(R, dg, elements, rho, orbitals, occupations)
    (g,) = torch.autograd.grad(e, R)
    return -g.numpy()

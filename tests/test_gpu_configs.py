"""GPU checks at the benchmark configurations themselves.

c2  96^3 synthetic block, k = 16: FP64 block H psi against the oracle (1e-12, norm-wise
    relative), FP32 (2e-5); the FP64 Poisson solution checked through the oracle's own
    assembled interior operator and boundary values (relative residual 1e-11 at solver
    tolerance 1e-12; boundary values 1e-12); mixed against FP64 (1e-7).
c3  H2O on the 128^3 monitor grid, k = 8: FP32 + FP64 refinement against FP64, every
    eigenvalue within 1e-6 Ha, energy components within 1e-6 Ha at an occupation edge in a gap.
c4  the same H2O setup, adjoints: Hellmann-Feynman forces against central differences of
    the ionic energy at a 1e-3 Bohr step (1e-4 Ha/Bohr) and their Richardson extrapolation
    with the half step (1e-5 Ha/Bohr), and dE/dv_ext of a sum of Ritz values along 16 seeded
    random smooth potential directions (1e-5 relative).
c5  benzene on the 224^3 monitor grid, k = 48: FP64 and FP32 block H psi against the
    oracle on sampled orbitals (1e-12 / 2e-5), and the eigensolve + Poisson in both
    precisions: every eigenvalue within 1e-6 Ha, forces with the molecule's symmetry (net
    force and out-of-plane components 1e-5 Ha/Bohr, the accuracy of orbitals converged to
    residual 3e-4).
"""

import numpy as np
import pytest
import torch

from isogriddft_amd import autograd, configs, eigensolver, kinetic, monitor_grid, poisson, pseudo, workloads, xc
from isogriddft_amd.device import DeviceGrid
from isogriddft_amd.hamiltonian import BlockHamiltonian
from oracle import kinetic as okin
from oracle import poisson as opoi

pytestmark = pytest.mark.gpu


def _nrel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def _hpsi_block_vs_oracle(geom, dg, psi64, veff, sample):
    """FP64 and FP32 block apply against the oracle on the sampled orbitals."""
    ref = np.stack([okin.monitor_kinetic(psi64[m], geom, "trial_fix") + veff * psi64[m] for m in sample])
    out = {}
    for dt in (torch.float64, torch.float32):
        x = torch.from_numpy(psi64).to("cuda", dt)
        v = torch.from_numpy(veff).to("cuda", dt)
        y = kinetic.apply_operator(dg, x, scale=-0.5, closure="trial_fix", veff=v)
        torch.cuda.synchronize()
        got = y[list(sample)].double().cpu().numpy()
        out[dt] = _nrel(got, ref)
    return out


def test_c2_block_hpsi_matches_oracle():
    cfg = configs.C2
    geom = configs.build_synthetic_geometry(cfg)
    dg = DeviceGrid(geom)
    psi = configs.build_synthetic_orbitals(geom, cfg).numpy()
    assert psi.shape == (16, 96, 96, 96)
    veff = -1.0 / np.sqrt(0.3 + geom.x_points ** 2 + geom.y_points ** 2 + (geom.z_points - 0.7) ** 2)
    err = _hpsi_block_vs_oracle(geom, dg, psi, veff, range(16))
    print("c2 H psi relative error", err)
    assert err[torch.float64] <= 1e-12
    assert err[torch.float32] <= 2e-5


def test_c2_poisson_solution_satisfies_the_oracle_system():
    cfg = configs.C2
    geom = configs.build_synthetic_geometry(cfg)
    dg = DeviceGrid(geom)
    rho = configs.build_synthetic_density(geom, cfg)
    sol = poisson.solve_poisson_device(dg, rho.to("cuda"), multipole_order=2, tolerance=1e-12, max_iterations=4000)
    assert sol["converged"]
    u = sol["potential"].cpu().numpy().reshape(dg.shape)
    b_gpu = sol["boundary"].cpu().numpy().reshape(dg.shape)
    rho_np = rho.numpy()
    b, _ = opoi.multipole_boundary(geom, rho_np, 2, "corrected_moments")
    assert np.max(np.abs(b_gpu - b)) <= 1e-12 * np.max(np.abs(b))
    rhs = opoi.FOUR_PI * rho_np + opoi.laplacian(b, geom)
    L, interior = opoi._interior_system(geom)
    A = (-L)[interior][:, interior]
    r = rhs.reshape(-1)[interior]
    res = A @ (u - b).reshape(-1)[interior] - r
    rel = float(np.linalg.norm(res) / np.linalg.norm(r))
    print("c2 Poisson: oracle-operator relative residual", rel, "iterations", sol["iterations"])
    assert rel <= 1e-11
    mixed = poisson.solve_poisson_device(dg, rho.to("cuda"), multipole_order=2, tolerance=1e-8, precision="mixed",
                                         max_iterations=4000)
    um = mixed["potential"].cpu().numpy()
    assert mixed["converged"] and np.max(np.abs(um - u)) <= 1e-7 * np.max(np.abs(u))


def _h2o_setup(n=128, tol=1e-6, k=8):
    mol = configs.H2O
    atoms = pseudo.case_atoms(mol)
    geom = monitor_grid.build_monitor_grid_for_case(
        mol, shape=(n, n, n), box_half_extents=(10.0, 10.0, 10.0),
        element_parameters=monitor_grid.build_default_near_core_element_parameters(mol))
    dg = DeviceGrid(geom)
    rho = workloads.start_density(dg, atoms)
    vloc = pseudo.local_potential_device(dg, atoms)
    _, vxc, _ = xc.lsda_device(0.5 * rho, 0.5 * rho)
    vh = poisson.solve_poisson_device(dg, rho, multipole_order=4, tolerance=1e-10, max_iterations=4000)["potential"]
    veff = vloc + vh.reshape(-1) + vxc.reshape(-1)
    plan = pseudo.NonlocalPlan(dg, atoms)
    H = BlockHamiltonian(dg, veff, closure="trial_fix", nonlocal_plan=plan)
    X0 = eigensolver.default_guess_block(geom, atoms, 4 * k, dg)
    return mol, atoms, geom, dg, rho, veff, H, X0


def test_c3_h2o_mixed_matches_fp64_per_eigenvalue_and_energy_component():
    mol, atoms, geom, dg, rho, veff, H, X0 = _h2o_setup()
    k, nocc = 8, 4
    res = {p: eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-6, 800, precision=p)
           for p in ("fp64", "mixed")}
    assert res["fp64"].converged and res["mixed"].converged
    e64, emx = np.asarray(res["fp64"].eigenvalues), np.asarray(res["mixed"].eigenvalues)
    print("c3 eigenvalues", e64, "abs diff", np.abs(emx - e64))
    assert np.max(np.abs(emx - e64)) <= 1e-6
    n_e, gap = workloads.occupation_edge(e64, nocc)
    print("c3 occupation edge", n_e, gap)
    assert gap is not None and gap >= workloads.EDGE_GAP_HA
    c64 = workloads.energy_components(dg, mol, res["fp64"].orbitals, n_e)
    cmx = workloads.energy_components(dg, mol, res["mixed"].orbitals, n_e)
    print("c3 energy components", c64, {key: abs(cmx[key] - c64[key]) for key in c64})
    for key in ("T_s", "E_loc", "E_nl", "E_H", "E_xc", "E_total"):
        assert abs(cmx[key] - c64[key]) <= 1e-6, key


def test_c4_h2o_adjoint_forces_and_potential_gradient():
    mol, atoms, geom, dg, rho, veff, H, X0 = _h2o_setup(tol=1e-8)
    k = 8
    res = eigensolver.block_subspace_iteration(H, dg.weights(), X0, k, 1e-7, 800)
    assert res.converged
    ev = np.asarray(res.eigenvalues)
    n_e, gap = workloads.occupation_edge(ev, 4)
    assert gap is not None and gap >= workloads.EDGE_GAP_HA
    occ = [2.0] * n_e
    orbs = res.orbitals.reshape(k, -1)[:n_e].to(torch.float64)
    rho_o = (2.0 * orbs * orbs).sum(0)
    elements = [el for el, _ in atoms]
    R0 = np.array([p for _, p in atoms], dtype=np.float64)
    F = autograd.atomic_forces(dg, elements, R0, rho_o, orbs, occ)

    def fd(a, ax, h):
        Rp, Rm = R0.copy(), R0.copy()
        Rp[a, ax] += h
        Rm[a, ax] -= h
        ep = autograd.ionic_energy(torch.from_numpy(Rp), dg, elements, rho_o, orbs, occ).item()
        em = autograd.ionic_energy(torch.from_numpy(Rm), dg, elements, rho_o, orbs, occ).item()
        return -(ep - em) / (2 * h)

    # The central difference at the 1e-3 Bohr step carries an O(h^2) truncation error of
    # ~1e-5 Ha/Bohr at the O atom (its GTH local part varies on a 0.25 Bohr scale); the
    # Richardson combination with the half step removes it, leaving O(h^4).
    h = 1e-3
    worst, worst_plain = 0.0, 0.0
    for a in range(len(elements)):
        for ax in range(3):
            f1, f2 = fd(a, ax, h), fd(a, ax, 0.5 * h)
            rich = (4.0 * f2 - f1) / 3.0
            worst, worst_plain = max(worst, abs(F[a, ax] - rich)), max(worst_plain, abs(F[a, ax] - f1))
            assert abs(F[a, ax] - rich) <= 1e-5, (a, ax, F[a, ax], f1, rich)
            assert abs(F[a, ax] - f1) <= 1e-4, (a, ax, F[a, ax], f1)
    print("c4 forces", F, "worst |adjoint - FD(1e-3)|", worst_plain, "worst |adjoint - Richardson|", worst)

    # dE/dv_ext for E = sum of the n_e lowest Ritz values (a complete set of levels below a
    # gap), warm-started from the converged basis, along 16 seeded smooth directions
    xyz = dg.xyz()
    basis = res.orbitals.reshape(k, -1).to(torch.float64)
    X1 = torch.cat([basis, X0.reshape(X0.shape[0], -1)[:3 * k].to(torch.float64)], 0)
    v = veff.clone().requires_grad_(True)
    wgt = torch.zeros(k, dtype=torch.float64, device="cuda")
    wgt[:n_e] = 1.0
    Hn = H.nonlocal_plan
    assert Hn is not None

    def ritz(vv):
        Hv = BlockHamiltonian(dg, vv.detach(), closure="trial_fix", nonlocal_plan=Hn)
        r = eigensolver.block_subspace_iteration(Hv, dg.weights(), X1, k, 1e-9, 800)
        assert r.converged
        return np.asarray(r.eigenvalues), r

    e0, r0 = ritz(v)
    psi = r0.orbitals.reshape(k, -1).to(torch.float64)
    grad = dg.weights() * (wgt[:, None] * psi * psi).sum(0)  # first-order perturbation theory
    rng = np.random.default_rng(2024)
    worst = 0.0
    for d in range(16):
        c = rng.uniform(-4.0, 4.0, 3)
        a = rng.uniform(0.05, 0.4)
        dv = torch.exp(-a * ((xyz[0] - c[0]) ** 2 + (xyz[1] - c[1]) ** 2 + (xyz[2] - c[2]) ** 2)) * rng.uniform(-1, 1)
        step = 1e-4
        ep, _ = ritz(v + step * dv)
        em, _ = ritz(v - step * dv)
        fd = float(np.dot(wgt.cpu().numpy(), ep - em)) / (2 * step)
        an = float((grad * dv).sum())
        worst = max(worst, abs(an - fd) / max(abs(fd), 1e-8))
        assert abs(an - fd) <= 1e-5 * max(abs(fd), 1e-3), (d, an, fd)
    print("c4 dE/dv worst relative difference", worst)


def test_c5_benzene_block_hpsi_and_mixed_eigensolve():
    mol = configs.BENZENE
    atoms = pseudo.case_atoms(mol)
    geom = monitor_grid.build_monitor_grid_for_case(
        mol, shape=(224, 224, 224), box_half_extents=(14.0, 14.0, 14.0),
        element_parameters=monitor_grid.build_default_near_core_element_parameters(mol))
    dg = DeviceGrid(geom)
    g = np.asarray
    X, Y, Z = g(geom.x_points), g(geom.y_points), g(geom.z_points)
    psi = np.stack([np.exp(-0.3 * (X ** 2 + Y ** 2 + Z ** 2)) * (1 + 0.2 * X * Y),
                    np.exp(-0.5 * ((X - 1.3) ** 2 + Y ** 2 + (Z - 0.2) ** 2)) - 0.5 * np.exp(-0.8 * (X ** 2 + (Y + 2.0) ** 2))])
    veff = pseudo.local_potential_device(dg, atoms).cpu().numpy().reshape(dg.shape)
    err = _hpsi_block_vs_oracle(geom, dg, psi, veff, (0, 1))
    print("c5 H psi relative error", err)
    assert err[torch.float64] <= 1e-12
    assert err[torch.float32] <= 2e-5
    del psi, X, Y, Z
    out = workloads.run("benzene", 224, 48, 192, 3e-4)
    print({key: out[key] for key in ("eig_fp64_s", "eig_mixed_s", "poisson_fp64_s", "poisson_mixed_s",
                                     "eig_mixed_vs_fp64_max_abs_Ha", "energy_occupied_states",
                                     "energy_mixed_vs_fp64_abs_Ha")})
    assert out["eig_fp64_converged"] and out["eig_mixed_converged"]
    assert max(out["eig_mixed_vs_fp64_abs_Ha"]) <= 1e-6
    assert out["poisson_mixed_vs_fp64_max_rel"] <= 1e-6
    F = np.asarray(out["forces_Ha_per_Bohr"])
    # D6h: the net force vanishes and every atom's force is radial in the ring plane, to the
    # accuracy of orbitals converged to residual 3e-4 (converged states locked at 3e-5)
    print("c5 net force", F.sum(0), "max |F_z|", np.max(np.abs(F[:, 2])))
    assert np.max(np.abs(F.sum(0))) <= 1e-5
    assert np.max(np.abs(F[:, 2])) <= 1e-5 * max(np.max(np.abs(F)), 1.0)

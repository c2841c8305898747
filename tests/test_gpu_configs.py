"""The benchmark configurations' own solves on the device against the oracle:

- the Q3 nonlinear heat Jacobian with the tanh law (k(T) = 1 + 0.5 tanh((T - 1700)/50), the
  Gaussian hot spot at (0.5, 0.5, 1), 1% seeded noise, dt = 1e-4, homogeneous Dirichlet), its
  Jacobi-CG solve to 1e-10 (the CG runs on the non-symmetric Jacobian, as configured), at a
  reduced mesh and at the full 64^3 size (vmult within 1e-12, solution within 1e-10);
- the Q2 level-set advection-diffusion operator with the BDF2 tangent (3/(2 dt)) M + A, the
  deformation velocity, artificial diffusion 0.5 h and dt from CFL 0.5, solved with
  Jacobi-GMRES(30) in FP64 and FP32-inner / FP64-refined form to 1e-10.

Bars: iteration counts within +-1 of the oracle's, FP64 operator within 1e-12, solutions within
1e-10 relative of the oracle's converged solution (both FP64 solves stop at the same residual
with the same iterates up to round-off).  The refined solves stop at a 1e-10 relative residual
by a different path, so their solutions are compared against a 1e-13 oracle solve with the
bound kappa * 1e-10 (kappa measured below 10 for these mass-dominated operators: 1e-9)."""
import numpy as np
import pytest

import orc
from mfcases import rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def heat_config(ncell, degree=3):
    """The Q3 heat Jacobian configuration on ncell^3 cells of the unit cube: device scratch,
    operator and fields, and the oracle operator of the same case."""
    import meltpooldg_amd as mf
    s = mf.ScratchData(degree, ncell, stream=torch.cuda.current_stream())
    mask = s.boundary_mask()
    s.set_dirichlet_mask(mask)
    X = s.support_points()
    r2 = ((X - np.array([0.5, 0.5, 1.0])) ** 2).sum(axis=1)
    T = (300.0 + 1500.0 * np.exp(-r2 / (2 * 0.05 ** 2))) * (1.0 + 0.01 * s.random_field(2))
    op = mf.HeatDiffuseOperator(s, law=mf.LAW_TANH, rho_cp=1.0, k=1.0, tanh_amp=0.5,
                                tanh_Tc=1700.0, tanh_width=50.0, t_dependent=True, dt=1e-4)
    T_d = _dev(T)
    op.set_fields(T=T_d)
    b = s.random_field(3)
    b[mask == 1] = 0.0
    m = orc.make_mesh(3, degree, (ncell,) * 3)
    hp = orc.HeatParams(law=orc.LAW_TANH, t_dependent=1, rho_cp=1.0, k=1.0, tanh_amp=0.5,
                        tanh_Tc=1700.0, tanh_width=50.0, dt=1e-4)
    ref = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, mask=mask, T=T, T_old=T)
    return s, op, T_d, b, ref


def _cg_against_oracle(ncell):
    import meltpooldg_amd as mf
    s, op, T_d, b, ref = heat_config(ncell)
    x_ref, it_ref, _, st = ref.solve(b, solver=orc.CG, rel_tol=1e-10,
                                     inv_diag=ref.inverse_diagonal())
    assert st == 0
    pc = mf.JacobiPreconditioner(op)
    pc.update()
    x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    rep = mf.LinearSolver.solve(op, x, _dev(b), solver="CG", rel_tolerance=1e-10,
                                preconditioner=pc)
    assert rep["status"] == 0 and abs(rep["iterations"] - it_ref) <= 1, (rep, it_ref)
    err = rel_err(x.cpu().numpy(), x_ref)
    assert err < 1e-10, err
    # the device solution's true residual under the oracle's operator
    r = b - ref.vmult(x.cpu().numpy())
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(b)
    return s, op, b, rep


def test_heat_jacobian_cg_reduced_mesh():
    """Jacobi-CG on the T-dependent (non-symmetric) Q3 tanh Jacobian, 12^3 cells."""
    _cg_against_oracle(12)


def test_heat_jacobian_full_size_vmult_and_cg():
    """The full 64^3 Q3 configuration (7,189,057 DoFs): one vmult within 1e-12 of the oracle's,
    the Jacobi-CG solution within 1e-10 of the oracle's CG solution with the same iteration
    count (+-1), and the FP32-inner / FP64-refined solve to the same residual."""
    import meltpooldg_amd as mf
    s, op, T_d, b, ref = heat_config(64)
    assert s.n_dofs == 7189057
    src = s.random_field(7)
    dst = torch.full((s.n_dofs,), float("nan"), dtype=torch.float64, device="cuda")
    op.vmult(dst, _dev(src))
    torch.cuda.synchronize()
    err = rel_err(dst.cpu().numpy(), ref.vmult(src))
    assert err < 1e-12, err
    del dst
    x_ref, it_ref, _, st = ref.solve(b, solver=orc.CG, rel_tol=1e-10,
                                     inv_diag=ref.inverse_diagonal())
    assert st == 0
    pc = mf.JacobiPreconditioner(op)
    pc.update()
    b_d = _dev(b)
    x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    rep = mf.LinearSolver.solve(op, x, b_d, solver="CG", rel_tolerance=1e-10, preconditioner=pc)
    assert rep["status"] == 0 and abs(rep["iterations"] - it_ref) <= 1, (rep, it_ref)
    err = rel_err(x.cpu().numpy(), x_ref)
    assert err < 1e-10, err
    x32 = torch.zeros_like(x)
    rep32 = mf.LinearSolver.solve(op, x32, b_d, solver="CG", rel_tolerance=1e-10,
                                  precision="fp32_refined", preconditioner=pc)
    assert rep32["status"] == 0
    assert rep32["final_residual"] <= 1e-10 * rep32["initial_residual"]
    assert rel_err(x32.cpu().numpy(), x_ref) < 1e-9


def levelset_config(ncell):
    """The Q2 level-set advection configuration on ncell^3 cells (BDF2 tangent, deformation
    velocity, diffusion 0.5 h, CFL 0.5, boundary rows Dirichlet)."""
    import meltpooldg_amd as mf
    s = mf.ScratchData(2, ncell, stream=torch.cuda.current_stream())
    mask = s.boundary_mask()
    s.set_dirichlet_mask(mask)
    X = s.support_points()
    h = 1.0 / ncell
    phi0 = np.tanh((0.25 - np.linalg.norm(X - 0.35, axis=1)) / (2 * 1.5 * h))
    x, y, z = X[:, 0], X[:, 1], X[:, 2]
    pi = np.pi
    u = np.stack([2 * np.sin(pi * x) ** 2 * np.sin(2 * pi * y) * np.sin(2 * pi * z),
                  -np.sin(2 * pi * x) * np.sin(pi * y) ** 2 * np.sin(2 * pi * z),
                  -np.sin(2 * pi * x) * np.sin(2 * pi * y) * np.sin(pi * z) ** 2]).ravel()
    dt = 0.5 * h / float(np.abs(u).max())
    op = mf.AdvectionDiffusionOperator(s, theta=1.0, diffusivity=0.5 * h, mass_factor=1.5, dt=dt)
    u_d = _dev(u)
    op.set_fields(velocity=u_d)
    ap = orc.AdvecParams(dt=dt, theta=1.0, diffusivity=0.5 * h, mass_factor=1.5, vel_scale=0.0)
    ref = orc.OracleOperator(orc.OP_ADVEC, orc.make_mesh(3, 2, (ncell,) * 3), advec=ap,
                             vel_kind=orc.VEL_NODAL, mask=mask, vel=u)
    return s, op, u_d, phi0, ref


def test_levelset_bdf2_refined_gmres():
    """Configuration 4's solve (GMRES(30), FP32 inner with FP64 refinement to 1e-10) on 16^3
    cells, against the oracle: right-hand side within 1e-12, FP64 GMRES iterations within +-1
    and solution within 1e-10, refined solution within kappa * 1e-10."""
    import meltpooldg_amd as mf
    s, op, u_d, phi0, ref = levelset_config(16)
    rhs_ref = ref.create_rhs(phi0)
    rhs = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    op.create_rhs(rhs, _dev(phi0))
    assert rel_err(rhs.cpu().numpy(), rhs_ref) < 1e-12
    inv = ref.inverse_diagonal()
    x_ref, it_ref, _, st = ref.solve(rhs_ref, solver=orc.GMRES, rel_tol=1e-10, inv_diag=inv)
    assert st == 0
    x_exact, _, _, st = ref.solve(rhs_ref, solver=orc.GMRES, rel_tol=1e-13, inv_diag=inv)
    assert st == 0
    pc = mf.JacobiPreconditioner(op)
    pc.update()
    x = torch.zeros_like(rhs)
    rep = mf.LinearSolver.solve(op, x, rhs, solver="GMRES", rel_tolerance=1e-10,
                                preconditioner=pc)
    assert rep["status"] == 0 and abs(rep["iterations"] - it_ref) <= 1, (rep, it_ref)
    assert rel_err(x.cpu().numpy(), x_ref) < 1e-10
    x32 = torch.zeros_like(rhs)
    rep32 = mf.LinearSolver.solve(op, x32, rhs, solver="GMRES", rel_tolerance=1e-10,
                                  precision="fp32_refined", preconditioner=pc)
    assert rep32["status"] == 0 and rep32["outer_iterations"] >= 2
    assert rep32["final_residual"] <= 1e-10 * rep32["initial_residual"]
    r = rhs_ref - ref.vmult(x32.cpu().numpy())
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(rhs_ref)
    assert rel_err(x32.cpu().numpy(), x_exact) < 1e-9

"""Device Krylov solves against the oracle's solves of the same system: iteration counts within
+-1, solutions within 1e-10 relative; FP32-inner/FP64-refined solves reach the same residual."""
import numpy as np
import pytest

import orc
from mfcases import Case, device_operator, fields, oracle_operator, rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.tensor(a, dtype=torch.float64, device="cuda")


def _rhs(case, f):
    b = orc.random_field(case.mesh(), 3)
    b[f["mask"] == 1] = 0.0
    return b


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_cg_matches_oracle(degree):
    import meltpooldg_amd as mf
    case = Case(degree=degree, kind="heat", ncell=(6, 5, 4), dt=1e-2)
    f = fields(case)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    b = _rhs(case, f)
    x_ref, it_ref, res_ref, st = ref.solve(b, solver=orc.CG, rel_tol=1e-10,
                                           inv_diag=ref.inverse_diagonal())
    assert st == 0
    x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    rep = mf.LinearSolver.solve(op, x, _dev(b), solver="CG", rel_tolerance=1e-10)
    assert abs(rep["iterations"] - it_ref) <= 1
    assert rel_err(x.cpu().numpy(), x_ref) < 1e-10 * 10
    # true residual
    Ax = ref.vmult(x.cpu().numpy())
    assert np.linalg.norm(b - Ax) <= 1e-10 * np.linalg.norm(b) * 1.5


@pytest.mark.parametrize("case", [Case(degree=2, kind="heat", law="tanh", t_dependent=True,
                                       vel=True, ncell=(5, 4, 4), dt=1e-3),
                                  Case(degree=2, kind="advec", ncell=(5, 5, 4), dt=0.05,
                                       diffusivity=0.01)],
                         ids=["heat-tanh-vel", "advec"])
def test_gmres_matches_oracle(case):
    import meltpooldg_amd as mf
    f = fields(case)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    b = _rhs(case, f)
    x_ref, it_ref, res_ref, st = ref.solve(b, solver=orc.GMRES, rel_tol=1e-10,
                                           inv_diag=ref.inverse_diagonal())
    assert st == 0
    x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    rep = mf.LinearSolver.solve(op, x, _dev(b), solver="GMRES", rel_tolerance=1e-10)
    assert abs(rep["iterations"] - it_ref) <= 1
    assert rel_err(x.cpu().numpy(), x_ref) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "fp32_refined"])
def test_gmres_forced_re_orthogonalization(precision):
    """force_re_orthogonalization runs the second Gram-Schmidt pass for every basis vector (the
    path that the default takes only after a cancellation): the same iterates up to rounding, so
    the solution matches the oracle's (two passes always) and the iteration count stays within
    +-1 of the oracle's and of the default's."""
    import meltpooldg_amd as mf
    case = Case(degree=2, kind="advec", ncell=(5, 5, 4), dt=0.05, diffusivity=0.01)
    f = fields(case)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    b = _rhs(case, f)
    x_ref, it_ref, _, st = ref.solve(b, solver=orc.GMRES, rel_tol=1e-10,
                                     inv_diag=ref.inverse_diagonal())
    assert st == 0
    out = {}
    for force in (False, True):
        x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
        rep = mf.LinearSolver.solve(op, x, _dev(b), solver="GMRES", rel_tolerance=1e-10,
                                    precision=precision, force_re_orthogonalization=force)
        assert rep["status"] == 0
        out[force] = (rep["iterations"], x.cpu().numpy())
        if precision == "fp64":
            assert abs(rep["iterations"] - it_ref) <= 1
        assert rel_err(out[force][1], x_ref) < 1e-9
    assert abs(out[True][0] - out[False][0]) <= 1


@pytest.mark.parametrize("solver", ["CG", "GMRES"])
def test_fp32_refined_reaches_fp64_tolerance(solver):
    import meltpooldg_amd as mf
    case = Case(degree=3, kind="heat", law="tanh", t_dependent=(solver == "GMRES"),
                ncell=(5, 5, 5), dt=1e-3)
    f = fields(case)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    b = _rhs(case, f)
    solver_id = orc.CG if solver == "CG" else orc.GMRES
    x_ref, _, _, st = ref.solve(b, solver=solver_id, rel_tol=1e-12,
                                inv_diag=ref.inverse_diagonal())
    x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    rep = mf.LinearSolver.solve(op, x, _dev(b), solver=solver, rel_tolerance=1e-10,
                                precision="fp32_refined", inner_rel_tolerance=1e-3)
    assert rep["status"] == 0
    assert rep["final_residual"] <= 1e-10 * rep["initial_residual"]
    assert rep["outer_iterations"] >= 2
    assert rel_err(x.cpu().numpy(), x_ref) < 1e-8


@pytest.mark.parametrize("solver", ["CG", "GMRES"])
def test_fp32_refined_converged_start_changes_nothing(solver):
    """ReductionControl semantics of the refined solver: an initial residual within the
    absolute tolerance stops after 0 passes with x untouched (the CG queues its first pass
    before the norm is read and must not apply it); a NaN right-hand side reports
    NoConvergence and leaves x untouched."""
    import meltpooldg_amd as mf
    case = Case(degree=2, kind="heat", ncell=(4, 4, 4), dt=1e-2)
    f = fields(case)
    s, op = device_operator(case, f)
    b = _dev(_rhs(case, f))
    x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    mf.LinearSolver.solve(op, x, b, solver=solver, rel_tolerance=1e-12)
    x0 = x.clone()
    r0 = float(torch.linalg.norm(b))
    rep = mf.LinearSolver.solve(op, x, b, solver=solver, precision="fp32_refined",
                                rel_tolerance=1e-10, abs_tolerance=1e-6 * r0)
    assert rep["status"] == 0 and rep["outer_iterations"] == 0 and rep["iterations"] == 0
    assert rep["final_residual"] == rep["initial_residual"] <= 1e-6 * r0
    assert torch.equal(x, x0)
    # b = 0 from x = 0: |r0| = 0
    z = torch.zeros_like(x)
    rep = mf.LinearSolver.solve(op, z, torch.zeros_like(b), solver=solver,
                                precision="fp32_refined")
    assert rep["outer_iterations"] == 0 and torch.count_nonzero(z) == 0
    bn = b.clone()
    bn[3] = float("nan")
    rep = mf.LinearSolver.solve(op, x, bn, solver=solver, precision="fp32_refined",
                                raise_on_failure=False)
    assert rep["status"] == mf.MF_NO_CONVERGENCE
    assert torch.equal(x, x0)


def test_gmres_restart_beyond_30_is_rejected():
    import meltpooldg_amd as mf
    case = Case(degree=1, kind="heat", ncell=(3, 3, 3), dt=1e-2)
    f = fields(case)
    s, op = device_operator(case, f)
    x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    with pytest.raises(mf.MFError) as e:
        mf.LinearSolver.solve(op, x, _dev(_rhs(case, f)), solver="GMRES", restart=31)
    assert e.value.code == mf.MF_ERR_ARG


def test_no_convergence_is_reported():
    """max_iterations = 1 forces NoConvergence (reference linear_solver.hpp:124-129)."""
    import meltpooldg_amd as mf
    case = Case(degree=2, kind="heat", ncell=(4, 4, 4), dt=1.0)
    f = fields(case)
    s, op = device_operator(case, f)
    b = _rhs(case, f)
    x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    with pytest.raises(mf.SolverNoConvergence):
        mf.LinearSolver.solve(op, x, _dev(b), solver="CG", max_iterations=1)
    rep = mf.LinearSolver.solve(op, x, _dev(b), solver="CG", max_iterations=1,
                                raise_on_failure=False)
    assert rep["status"] == mf.MF_NO_CONVERGENCE and rep["iterations"] == 1


@pytest.mark.parametrize("solver", ["CG", "GMRES"])
def test_jacobi_preconditioner_object(solver):
    """JacobiPreconditioner (preconditioner_jacobi.hpp:29-89): update() computes the inverse
    diagonal once; solves given the object reproduce the solves that recompute it."""
    import meltpooldg_amd as mf
    case = Case(degree=3, kind="heat", law="tanh", t_dependent=(solver == "GMRES"),
                ncell=(5, 4, 4), dt=1e-3)
    f = fields(case)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    b = _dev(_rhs(case, f))
    pc = mf.JacobiPreconditioner(op)
    y = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    with pytest.raises(mf.MFError) as e:
        pc.vmult(y, b)
    assert e.value.code == mf.MF_ERR_STATE
    pc.update()
    d_pc = torch.zeros_like(y)
    pc.inverse_diagonal(d_pc)
    d_op = torch.zeros_like(y)
    op.compute_inverse_diagonal_from_matrixfree(d_op)
    assert torch.equal(d_pc, d_op)
    assert rel_err(d_pc.cpu().numpy(), ref.inverse_diagonal()) < 1e-12
    pc.vmult(y, b)
    assert torch.equal(y, d_op * b)
    for precision in ("fp64", "fp32_refined"):
        x1 = torch.zeros_like(y)
        r1 = mf.LinearSolver.solve(op, x1, b, solver=solver, rel_tolerance=1e-10,
                                   precision=precision)
        x2 = torch.zeros_like(y)
        r2 = mf.LinearSolver.solve(op, x2, b, solver=solver, rel_tolerance=1e-10,
                                   precision=precision, preconditioner=pc)
        assert r1["iterations"] == r2["iterations"] and r1["status"] == r2["status"] == 0
        assert torch.equal(x1, x2)


def test_fp32_refined_default_inner_tolerance():
    """Default inner tolerance sqrt(1e-10) = 1e-5: two defect-correction passes reach the FP64
    tolerance on the tanh Jacobian, within one pass of what 1e-3 needs in inner iterations."""
    import meltpooldg_amd as mf
    case = Case(degree=3, kind="heat", law="tanh", t_dependent=False, ncell=(20, 6, 6), dt=1e-4,
                hi=(2.0, 0.6, 0.6))
    f = fields(case)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    b = _rhs(case, f)
    x_ref, it_ref, _, st = ref.solve(b, solver=orc.CG, rel_tol=1e-10,
                                     inv_diag=ref.inverse_diagonal())
    assert st == 0
    x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    rep = mf.LinearSolver.solve(op, x, _dev(b), solver="CG", rel_tolerance=1e-10,
                                precision="fp32_refined")
    assert rep["status"] == 0 and rep["final_residual"] <= 1e-10 * rep["initial_residual"]
    assert rep["outer_iterations"] <= 3
    assert rep["iterations"] <= 1.5 * it_ref
    assert rel_err(x.cpu().numpy(), x_ref) < 1e-8

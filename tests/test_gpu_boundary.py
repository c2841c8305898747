"""Robin / Neumann boundary faces of the heat operator (tangent_boundary_loop,
rhs_boundary_loop and the boundary operation of compute_diagonal,
source/heat/heat_diffuse_operator.cpp:243-269, 297-391, 719-781, 1088-1129) on the device,
against the oracle's textbook face quadrature.  Bar: FP64 within 1e-12 relative (max-norm), the
FP32 operator within 1e-5; a Jacobi-CG solve with convection sides matches the oracle's CG
(iteration count +-1, solution 1e-10)."""
import numpy as np
import pytest

import orc
from mfcases import Case, device_operator, fields, oracle_operator, rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

ROBIN = dict(convection_sides=[0, 3, 5], radiation_sides=[1, 2], convection_coefficient=3.0,
             convection_temperature_infinity=800.0, emissivity=0.6,
             radiation_temperature_infinity=500.0)
WITH_NEUMANN = dict(ROBIN, neumann_sides={4: 1.5})


def _dev(a):
    return torch.tensor(a, dtype=torch.float64, device="cuda")


def _set(op, spec):
    op.set_boundary(**spec)


CASES = [Case(degree=d, kind="heat", law="tanh", t_dependent=True, mask=mk, ncell=(4, 3, 5))
         for d in (1, 2, 3) for mk in ("none", "random")] + \
        [Case(degree=3, kind="heat", law="tanh", t_dependent=True, vel=True, k_cell=True,
              mask="none", ncell=(37, 2, 3), hi=(3.7, 0.24, 0.27))]


def ids(c):
    return f"Q{c.degree}-{c.mask}-{c.ncell[0]}{'-vel-kc' if c.vel else ''}"


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_boundary_vmult(case):
    f = fields(case)
    ref = oracle_operator(case, f, boundary=WITH_NEUMANN)
    s, op = device_operator(case, f)
    _set(op, WITH_NEUMANN)
    src = orc.random_field(case.mesh(), 7)
    dst = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    op.vmult(dst, _dev(src))
    torch.cuda.synchronize()
    expect = ref.vmult(src)
    assert rel_err(dst.cpu().numpy(), expect) < 1e-12
    # the faces add something: compare with the operator without them
    plain = oracle_operator(case, f).vmult(src)
    assert rel_err(plain, expect) > 1e-6
    d32 = torch.zeros(s.n_dofs, dtype=torch.float32, device="cuda")
    op.vmult_f32(d32, _dev(src).float())
    torch.cuda.synchronize()
    assert rel_err(d32.double().cpu().numpy(), expect) < 1e-5
    # transposed operator: the face term is symmetric, so A^T with faces = A^T without + B
    op.set_transpose(True)
    dt_ = torch.zeros_like(dst)
    op.vmult(dt_, _dev(src))
    op.clear_boundary()
    dt0 = torch.zeros_like(dst)
    op.vmult(dt0, _dev(src))
    op.set_transpose(False)
    torch.cuda.synchronize()
    face = expect - plain
    assert rel_err((dt_ - dt0).cpu().numpy(), face) < 1e-11


@pytest.mark.parametrize("case", CASES[:6], ids=ids)
def test_boundary_inverse_diagonal_and_rhs(case):
    f = fields(case)
    ref = oracle_operator(case, f, boundary=WITH_NEUMANN)
    s, op = device_operator(case, f)
    _set(op, WITH_NEUMANN)
    d = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    op.compute_inverse_diagonal_from_matrixfree(d)
    torch.cuda.synchronize()
    assert rel_err(d.cpu().numpy(), ref.inverse_diagonal()) < 1e-12
    q_s = f["q_s"]
    init = orc.random_field(case.mesh(), 12)
    r = _dev(init)
    op.create_rhs(r, _dev(f["T_old"]), _dev(q_s))
    torch.cuda.synchronize()
    expect = ref.create_rhs(f["T_old"], q_s, dst=init)
    assert rel_err(r.cpu().numpy(), expect) < 1e-12


@pytest.mark.parametrize("degree", [2, 3])
def test_boundary_cg_matches_oracle(degree):
    """Jacobi-CG with convection sides (symmetric positive definite): the fused p.(Ap) of the
    device CG includes the face term."""
    import meltpooldg_amd as mf
    case = Case(degree=degree, kind="heat", ncell=(6, 5, 4), dt=1e-2, mask="none")
    f = fields(case)
    spec = dict(convection_sides=[0, 1, 2, 3, 4, 5], convection_coefficient=40.0)
    ref = oracle_operator(case, f, boundary=spec)
    s, op = device_operator(case, f)
    _set(op, spec)
    b = orc.random_field(case.mesh(), 3)
    x_ref, it_ref, _, st = ref.solve(b, solver=orc.CG, rel_tol=1e-10,
                                     inv_diag=ref.inverse_diagonal())
    assert st == 0
    x = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    rep = mf.LinearSolver.solve(op, x, _dev(b), solver="CG", rel_tolerance=1e-10)
    assert abs(rep["iterations"] - it_ref) <= 1
    assert rel_err(x.cpu().numpy(), x_ref) < 1e-9
    x32 = torch.zeros_like(x)
    rep32 = mf.LinearSolver.solve(op, x32, _dev(b), solver="CG", rel_tolerance=1e-10,
                                  precision="fp32_refined")
    assert rep32["final_residual"] <= 1e-10 * rep32["initial_residual"]
    assert rel_err(x32.cpu().numpy(), x_ref) < 1e-8

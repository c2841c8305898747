"""Device parity: libmeltpooldg_amd (through the C ABI) against the CPU oracle on the same seeded
inputs.  Bar: FP64 operator application within 1e-12 relative (max-norm); the FP32 operator
(used only inside the refined solver) within 1e-5: unit roundoff 6e-8 times the accumulation over
the 18 sum-factorisation sweeps and the tanh-law slope, measured up to 3.2e-6; a misplaced or
missing contribution shows as O(1)."""
import numpy as np
import pytest

import orc
from mfcases import Case, device_operator, fields, oracle_operator, rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5

CASES = [
    Case(degree=1, kind="heat"),
    Case(degree=2, kind="heat"),
    Case(degree=3, kind="heat"),
    Case(degree=2, kind="heat", law="tanh", t_dependent=True),
    Case(degree=3, kind="heat", law="tanh", t_dependent=True, mask="random"),
    Case(degree=3, kind="heat", law="tanh", t_dependent=False, mask="none"),
    Case(degree=2, kind="heat", vel=True, k_cell=True),
    Case(degree=3, kind="heat", law="tanh", t_dependent=True, vel=True, k_cell=True),
    Case(degree=1, kind="advec"),
    Case(degree=2, kind="advec", mask="random"),
    Case(degree=3, kind="advec", theta=1.0, diffusivity=0.0, mass_factor=1.5),
    # rows longer than one workgroup step (16 cells for Q2/Q3, 64 for Q1): the row walk carries
    # the shared node across steps, loads the next step's pencils ahead and ends mid-step
    # (domains keep the cells close to the aspect ratios of the cases above)
    Case(degree=3, kind="heat", law="tanh", t_dependent=True, mask="random", ncell=(37, 2, 3),
         hi=(3.7, 0.24, 0.27)),
    Case(degree=2, kind="advec", ncell=(35, 3, 2), hi=(3.5, 0.36, 0.18)),
    Case(degree=1, kind="heat", vel=True, k_cell=True, ncell=(150, 2, 2), hi=(15.0, 0.24, 0.18)),
    # the mushy-zone law: rho cp(T), the d(rho cp)/dT terms read the old temperature
    Case(degree=2, kind="heat", law="mushy", t_dependent=True, dt=1e-2),
    Case(degree=3, kind="heat", law="mushy", t_dependent=True, vel=True, k_cell=True, dt=1e-2,
         mask="random"),
    Case(degree=1, kind="heat", law="mushy", t_dependent=False, vel=True, dt=1e-2),
    # SUPG: per-cell tau from max_q |w| (coefficient -1) or a fixed coefficient, the Laplacian of
    # u in the tangent and the rhs, the basis Laplacians in the diagonal
    Case(degree=1, kind="advec", supg=True),
    Case(degree=2, kind="advec", supg=True, mask="random", supg_coefficient=0.3),
    Case(degree=3, kind="advec", supg=True, theta=1.0, mass_factor=1.5, diffusivity=0.2),
    Case(degree=2, kind="advec", supg=True, ncell=(35, 3, 2), hi=(3.5, 0.36, 0.18)),
]


def ids(c):
    return f"{c.kind}-Q{c.degree}-{c.law}{'-td' if c.t_dependent else ''}" \
           f"{'-vel' if c.vel else ''}{'-kc' if c.k_cell else ''}{'-supg' if c.supg else ''}" \
           f"-{c.mask}"


def _dev(a):
    return torch.tensor(a, dtype=torch.float64, device="cuda")


@pytest.fixture(params=[0], ids=["pencil"])
def variant(request):
    import meltpooldg_amd as mf
    mf.set_tangent_kernel(request.param)
    yield request.param
    mf.set_tangent_kernel(0)


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_vmult_fp64(case, variant):
    f = fields(case)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    src = orc.random_field(case.mesh(), 7)
    # vmult overwrites dst (zero_dst): a NaN-filled dst shows any entry it fails to store
    dst = torch.full((s.n_dofs,), float("nan"), dtype=torch.float64, device="cuda")
    op.vmult(dst, _dev(src))
    torch.cuda.synchronize()
    expect = ref.vmult(src)
    assert rel_err(dst.cpu().numpy(), expect) < TOL64
    # constrained rows stay zero
    if f["mask"].any():
        assert np.all(dst.cpu().numpy()[f["mask"] == 1] == 0.0)


@pytest.mark.parametrize("case", CASES[:6] + CASES[8:], ids=ids)
def test_vmult_fp32(case, variant):
    f = fields(case)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    src = orc.random_field(case.mesh(), 8)
    dst = torch.full((s.n_dofs,), float("nan"), dtype=torch.float32, device="cuda")
    op.vmult_f32(dst, torch.tensor(src, dtype=torch.float32, device="cuda"))
    torch.cuda.synchronize()
    # the mushy law's d(rho cp)/dT (T - T_old)/dt term takes T - T_old from FP32 copies
    # (ulp(273) = 3e-5 against a difference of order 1): a looser FP32 bound there
    tol = 1e-4 if case.law == "mushy" else TOL32
    assert rel_err(dst.double().cpu().numpy(), ref.vmult(src)) < tol


@pytest.mark.parametrize("case", CASES, ids=ids)
def test_inverse_diagonal(case):
    f = fields(case)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    d = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    op.compute_inverse_diagonal_from_matrixfree(d)
    torch.cuda.synchronize()
    assert rel_err(d.cpu().numpy(), ref.inverse_diagonal()) < TOL64


@pytest.mark.parametrize("case", [c for c in CASES if not (c.kind == "heat" and c.t_dependent
                                                            and c.law == "const")], ids=ids)
def test_create_rhs(case):
    f = fields(case)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    src = f["T_old"] if case.kind == "heat" else orc.random_field(case.mesh(), 9)
    q_s = f["q_s"] if case.kind == "heat" else None
    init = orc.random_field(case.mesh(), 10)
    dst = _dev(init)
    op.create_rhs(dst, _dev(src), _dev(q_s) if q_s is not None else None)
    torch.cuda.synchronize()
    expect = ref.create_rhs(src, q_s, dst=init)
    assert rel_err(dst.cpu().numpy(), expect) < TOL64


@pytest.mark.parametrize("case", [CASES[1], CASES[8]], ids=ids)
def test_no_bc_dof_index(case):
    """reset_dof_index(no-BC): Dirichlet rows/entries are not skipped (Dirichlet lifting)."""
    f = fields(case)
    ref = oracle_operator(case, f, use_mask=0)
    s, op = device_operator(case, f)
    op.reset_dof_index(1)
    src = orc.random_field(case.mesh(), 12)
    dst = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    op.vmult(dst, _dev(src))
    torch.cuda.synchronize()
    assert rel_err(dst.cpu().numpy(), ref.vmult(src)) < TOL64


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_inflow_outflow_rows(degree):
    """Advection with inflow/outflow DoFs (and Dirichlet rows) against the oracle: vmult (FP64,
    NaN-filled dst) and FP32 with the identity rows, the inverse diagonal with 1 on them, and
    the treatment switched off (set_apply_inflow_outflow_constraints_pre_post_matrix_free)."""
    import meltpooldg_amd as mf
    case = Case(degree=degree, kind="advec", mask="random")
    f = fields(case)
    rng = np.random.default_rng(degree)
    io = ((rng.uniform(size=len(f["mask"])) < 0.25) & (f["mask"] == 0)).astype(np.uint8)
    m = case.mesh()
    ap = orc.AdvecParams(dt=case.dt, theta=case.theta, diffusivity=case.diffusivity,
                         mass_factor=case.mass_factor, vel_scale=0.0)
    ref = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL, mask=f["mask"],
                             vel=f["vel"], io_mask=io)
    s, op = device_operator(case, f)
    op.set_inflow_outflow_bc(io)
    src = orc.random_field(m, 21)
    dst = torch.full((s.n_dofs,), float("nan"), dtype=torch.float64, device="cuda")
    op.vmult(dst, _dev(src))
    torch.cuda.synchronize()
    expect = ref.vmult(src)
    assert rel_err(dst.cpu().numpy(), expect) < TOL64
    assert np.array_equal(dst.cpu().numpy()[io == 1], src[io == 1])
    d32 = torch.full((s.n_dofs,), float("nan"), dtype=torch.float32, device="cuda")
    op.vmult_f32(d32, torch.tensor(src, dtype=torch.float32, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(d32.double().cpu().numpy(), expect) < TOL32
    inv = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    op.compute_inverse_diagonal_from_matrixfree(inv)
    torch.cuda.synchronize()
    ref_inv = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL,
                                 mask=f["mask"], vel=f["vel"]).inverse_diagonal()
    ref_inv[io == 1] = 1.0
    assert rel_err(inv.cpu().numpy(), ref_inv) < TOL64
    op.set_apply_inflow_outflow_constraints_pre_post_matrix_free(False)
    op.vmult(dst, _dev(src))
    torch.cuda.synchronize()
    plain = oracle_operator(case, f).vmult(src)
    assert rel_err(dst.cpu().numpy(), plain) < TOL64


def test_supg_rejects_the_transposed_operator():
    import meltpooldg_amd as mf
    case = Case(degree=1, kind="advec", supg=True)
    s, op = device_operator(case, fields(case))
    with pytest.raises(mf.MFError):
        op.set_transpose(True)


@pytest.mark.parametrize("t_dependent", [True, False], ids=["td", "k_only"])
@pytest.mark.parametrize("mask", ["boundary", "random"])
def test_tanh_saturated_rows(t_dependent, mask):
    """T rising along z from far below to far above Tc (300 + 2800 z): the rows of the bottom
    cell layer have tanh = -1 at every quadrature point, those of the top layer +1, the others
    lie in the transition.  The structured Q3 kernel evaluates a saturated wave without T at the
    quadrature points (k = k0 -+ amp, dk/dT = 0); both kinds and the mixed rows must match the
    oracle: FP64 within 1e-12, FP32 within 1e-5."""
    case = Case(degree=3, kind="heat", law="tanh", t_dependent=t_dependent, mask=mask,
                ncell=(12, 4, 12))
    f = fields(case)
    m = case.mesh()
    f["T"] = 300.0 + 2800.0 * orc.node_coords(m)[:, 2] + 5.0 * orc.random_field(m, 3)
    ref = oracle_operator(case, f)
    s, op = device_operator(case, f)
    src = orc.random_field(m, 9)
    expect = ref.vmult(src)
    dst = torch.full((s.n_dofs,), float("nan"), dtype=torch.float64, device="cuda")
    op.vmult(dst, _dev(src))
    torch.cuda.synchronize()
    assert rel_err(dst.cpu().numpy(), expect) < TOL64
    d32 = torch.full((s.n_dofs,), float("nan"), dtype=torch.float32, device="cuda")
    op.vmult_f32(d32, torch.tensor(src, dtype=torch.float32, device="cuda"))
    torch.cuda.synchronize()
    assert rel_err(d32.double().cpu().numpy(), expect) < TOL32
    # the saturated layers alone: a source supported on the bottom and top cell layers
    z = orc.node_coords(m)[:, 2]
    src2 = np.where((z <= 1.0 / 12) | (z >= 11.0 / 12), src, 0.0)
    op.vmult(dst, _dev(src2))
    torch.cuda.synchronize()
    assert rel_err(dst.cpu().numpy(), ref.vmult(src2)) < TOL64

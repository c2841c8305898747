"""The device's one-phase CutFEM heat operator (mf_heat_cut_*) against the oracle
(oracle/orc_cut.c, pinned in tests/test_oracle_cut.py): classification, tangent, diagonal,
residual, cut norm, the Newton step, and the reference's one_phase_cut golden run extruded to
3D; at a larger size, properties that need no oracle (symmetry, the liquid volume)."""
import numpy as np
import pytest

import orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

GOLDEN_ONE_PHASE_CUT = [4.04890e-03, 4.01890e-03, 3.99116e-03, 3.96522e-03, 3.94086e-03]
MUSHY = dict(liquid_k=30.0, liquid_rho=7000.0, liquid_cp=800.0, solid_k=20.0, solid_rho=7500.0,
             solid_cp=600.0, solidus=1.0, liquidus=2.0)


def _dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


# (axis, level, liquid below the level); 0.5 on the 6x4x5 mesh of [0,1]x[0,0.8]x[0,1.25] is a
# node plane for y (0.2 * 2.5) and z (0.25 * 2)
INTERFACES = [(1, 0.43, True), (0, 0.61, False), (2, 0.37, True), (2, 0.5, False),
              (1, 0.4, True), (0, 2.0, True)]


def _case(axis, level, below, ncell=(6, 4, 5), hi=(1.0, 0.8, 1.25), mask_side=None,
          mushy=False, dt=0.3, theta=0.6):
    import meltpooldg_amd as mf
    s = mf.ScratchData(1, ncell, (0.0, 0.0, 0.0), hi, stream=torch.cuda.current_stream())
    m = orc.make_mesh(3, 1, ncell, (0.0, 0.0, 0.0), hi)
    X = orc.node_coords(m)
    phi = (level - X[:, axis]) if below else (X[:, axis] - level)
    mask = None
    if mask_side is not None:
        d, v = mask_side
        mask = (np.abs(X[:, d] - v) < 1e-12).astype(np.uint8)
        s.set_dirichlet_mask(mask)
    kw = dict(theta=theta, gamma_A=1.5, gamma_M=0.75)
    if mushy:
        op = mf.HeatCutOperator(s, law=mf.LAW_MUSHY, mushy=MUSHY, dt=dt, level_set=phi, **kw)
        p = orc.cut_params(dt, mushy=MUSHY, mesh=m, **kw)
    else:
        op = mf.HeatCutOperator(s, rho_cp=3.0, k=2.0, dt=dt, level_set=phi, **kw)
        p = orc.cut_params(dt, rho_cp=3.0, k=2.0, mesh=m, **kw)
    ref = orc.CutOracle(m, phi, p, mask=mask)
    return s, op, ref, m, X


@pytest.mark.parametrize("axis,level,below", INTERFACES)
def test_classification(axis, level, below):
    s, op, ref, m, X = _case(axis, level, below)
    cat, box, act = op.cells()
    np.testing.assert_array_equal(cat, ref.category)
    np.testing.assert_allclose(box, ref.box, atol=1e-15)
    np.testing.assert_array_equal(act, ref.active)
    info = op.info()
    assert info["n_faces"] == len(ref.faces)
    assert info["n_active_dofs"] == int(ref.active.sum())
    assert info["n_intersected"] == int((ref.category == orc.CUT_INTERSECTED).sum())


@pytest.mark.parametrize("axis,level,below", INTERFACES)
@pytest.mark.parametrize("mask_side", [None, (0, 0.0)])
def test_vmult_matches_oracle(axis, level, below, mask_side):
    s, op, ref, m, X = _case(axis, level, below, mask_side=mask_side)
    u = orc.random_field(m, 11)
    want = ref.vmult(u)
    dst = _dev(np.full(s.n_dofs, 7.0))            # vmult overwrites
    op.vmult(dst, _dev(u))
    got = _host(dst)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    # vmult_add adds to what dst holds
    dst = _dev(np.ones(s.n_dofs))
    op.vmult_add(dst, _dev(u))
    # (the fixed rows keep their 1: want is 0 there)
    np.testing.assert_allclose(_host(dst), want + 1.0, rtol=0, atol=1e-12 * np.abs(want).max())


@pytest.mark.parametrize("axis,level,below", INTERFACES[:4])
def test_mushy_vmult_and_residual(axis, level, below):
    s, op, ref, m, X = _case(axis, level, below, mushy=True, dt=0.05, theta=0.5)
    T = 1.5 + 0.4 * orc.random_field(m, 4)      # across the mushy interval and beyond
    To = 1.5 + 0.4 * orc.random_field(m, 5)
    u = orc.random_field(m, 6)
    Td = _dev(T)
    op.set_fields(T=Td)
    dst = _dev(np.zeros(s.n_dofs))
    op.vmult(dst, _dev(u))
    want = ref.vmult(u, T)
    np.testing.assert_allclose(_host(dst), want, rtol=0, atol=1e-12 * np.abs(want).max())
    r = _dev(np.zeros(s.n_dofs))
    op.create_rhs(r, _dev(To))
    want = ref.residual(T, To)
    np.testing.assert_allclose(_host(r), want, rtol=0, atol=1e-12 * np.abs(want).max())
    inv = _dev(np.zeros(s.n_dofs))
    op.compute_inverse_diagonal_from_matrixfree(inv)
    np.testing.assert_allclose(_host(inv), ref.inverse_diagonal(T), rtol=1e-12)


@pytest.mark.parametrize("axis,level,below", INTERFACES)
def test_diagonal_residual_and_norm(axis, level, below):
    s, op, ref, m, X = _case(axis, level, below, mask_side=(1, 0.0))
    inv = _dev(np.zeros(s.n_dofs))
    op.compute_inverse_diagonal_from_matrixfree(inv)
    np.testing.assert_allclose(_host(inv), ref.inverse_diagonal(), rtol=1e-12)
    T = orc.random_field(m, 1)
    To = orc.random_field(m, 2)
    op.set_fields(T=_dev(T))
    r = _dev(np.full(s.n_dofs, 0.5))              # create_rhs adds
    op.create_rhs(r, _dev(To))
    want = ref.residual(T, To)
    np.testing.assert_allclose(_host(r) - 0.5, want, rtol=0, atol=1e-12 * np.abs(want).max())
    for u in (T, X[:, axis] ** 2, np.ones(s.n_dofs)):
        np.testing.assert_allclose(op.compute_cut_norm(_dev(u)), ref.norm(u), rtol=1e-12)


def test_golden_one_phase_cut_extruded():
    """unidirectional_heat_transfer_one_phase_cut (2D Q1, 17^2 cells on [0, 0.1]^2, level set
    0.05 - y, T0 = x, dt 2, theta 0.5, gamma_A 1.5, gamma_M 0.75, GMRES + Diagonal, rel 1e-12,
    abs 1e-10, Newton residual tolerance 1e-9) as one cell of depth 1 in z: the 3D solution does
    not depend on z and its norm equals the 2D norm."""
    import meltpooldg_amd as mf
    s = mf.ScratchData(1, (17, 17, 1), (0.0, 0.0, 0.0), (0.1, 0.1, 1.0),
                       stream=torch.cuda.current_stream())
    X = s.support_points()
    phi = 0.05 - X[:, 1]
    op = mf.HeatCutOperator(s, rho_cp=7850.0 * 460.0, k=55.563, theta=0.5, gamma_A=1.5,
                            gamma_M=0.75, cell_side_length=0.1 / 17, level_set=phi)
    heat = mf.HeatCutOperation(op, newton=dict(residual_tolerance=1e-9), solver="GMRES",
                               rel_tolerance=1e-12, abs_tolerance=1e-10)
    heat.set_initial_condition(_dev(X[:, 0]))
    norms = []
    for _ in range(5):
        heat.init_time_advance(2.0)
        rep = heat.solve()
        assert rep["status"] == 0
        norms.append(heat.compute_L2_norm())
    m = orc.make_mesh(3, 1, (17, 17, 1), (0.0, 0.0, 0.0), (0.1, 0.1, 1.0))
    p = orc.cut_params(2.0, rho_cp=7850.0 * 460.0, k=55.563, theta=0.5, gamma_A=1.5,
                       gamma_M=0.75, h=0.1 / 17)
    _, ref = orc.cut_time_steps(m, phi, p, X[:, 0].copy(), 5)
    np.testing.assert_allclose(norms, ref, rtol=1e-9)
    np.testing.assert_allclose(norms, GOLDEN_ONE_PHASE_CUT, rtol=3e-6)


@pytest.mark.parametrize("solver", ["CG", "GMRES"])
def test_newton_step_mushy_matches_oracle(solver):
    """Two Newton steps of the mushy law (non-symmetric tangent for GMRES; CG runs the fused
    p.Ap path on the symmetric part's problem with a constant law)."""
    import meltpooldg_amd as mf
    mushy = solver == "GMRES"
    s, op, ref, m, X = _case(1, 0.43, True, mushy=mushy, dt=0.05, theta=0.5,
                             mask_side=(0, 0.0))
    T0 = 1.5 + 0.3 * np.cos(3.0 * X[:, 0]) * np.sin(2.0 * X[:, 2]) if mushy else \
        np.cos(3.0 * X[:, 0]) + X[:, 1]
    T = _dev(T0)
    T_old = _dev(T0)
    p = ref.params
    temps, _ = orc.cut_time_steps(m, (0.43 - X[:, 1]), p, T0, 2, mask=ref.mask)
    for k in range(2):
        T_old.copy_(T)
        rep = mf.heat_cut_step(op, T, T_old, newton=dict(residual_tolerance=1e-10,
                                                          correction_tolerance=1e-11),
                               solver=solver, rel_tolerance=1e-13, abs_tolerance=1e-12)
        assert rep["status"] == 0
        got = _host(T)
        act = ref.active == 1
        np.testing.assert_allclose(got[act], temps[k][act], rtol=0,
                                   atol=1e-9 * np.abs(temps[k]).max())


def test_solve_cg_matches_dense():
    import meltpooldg_amd as mf
    s, op, ref, m, X = _case(2, 0.37, True, mask_side=(2, 0.0))
    b = orc.random_field(m, 8)
    fixed = (ref.active == 0) | (ref.mask == 1)
    b[fixed] = 0.0
    A = ref.dense()
    A[fixed, fixed] = 1.0
    want = np.linalg.solve(A, b)
    x = _dev(np.zeros(s.n_dofs))
    rep = mf.LinearSolver.solve(op, x, _dev(b), solver="CG", rel_tolerance=1e-13)
    assert rep["status"] == 0
    np.testing.assert_allclose(_host(x), want, rtol=0, atol=1e-10 * np.abs(want).max())


def test_errors():
    import meltpooldg_amd as mf
    s = mf.ScratchData(1, (4, 4, 4), stream=torch.cuda.current_stream())
    X = s.support_points()
    op = mf.HeatCutOperator(s, rho_cp=1.0, k=1.0, dt=0.1)
    dst = _dev(np.zeros(s.n_dofs))
    with pytest.raises(mf.MFError):            # no level set yet
        op.vmult(dst, dst)
    with pytest.raises(mf.MFError):            # an oblique interface
        op.set_level_set(0.9 - X[:, 0] - X[:, 1])
    op.set_level_set(0.6 - X[:, 2])
    with pytest.raises(mf.MFError):            # no FP32 form
        mf.LinearSolver.solve(op, dst, dst, precision="fp32_refined")
    s2 = mf.ScratchData(2, (4, 4, 4), stream=torch.cuda.current_stream())
    with pytest.raises(mf.MFError):            # FE_Q1 only
        mf.HeatCutOperator(s2, rho_cp=1.0, k=1.0)


def test_large_mesh_properties():
    """128^3 cells (2.1 M DoFs), interface z = 0.4321: symmetry u.(A v) = v.(A u) and
    1^T A 1 = rho cp |Omega_l| (stiffness and ghost penalty annihilate constants)."""
    import meltpooldg_amd as mf
    n = 128
    s = mf.ScratchData(1, (n, n, n), stream=torch.cuda.current_stream())
    X = s.support_points()
    op = mf.HeatCutOperator(s, rho_cp=2.5, k=3.0, dt=0.01, gamma_A=1.5, gamma_M=0.75,
                            level_set=0.4321 - X[:, 2])
    info = op.info()
    assert info["n_intersected"] == n * n and info["n_faces"] == 2 * n * (n - 1) + n * n
    g = torch.Generator(device="cuda").manual_seed(3)
    u = torch.rand(s.n_dofs, dtype=torch.float64, device="cuda", generator=g)
    v = torch.rand(s.n_dofs, dtype=torch.float64, device="cuda", generator=g)
    Au, Av = torch.empty_like(u), torch.empty_like(v)
    op.vmult(Au, u)
    op.vmult(Av, v)
    a, b = float(torch.dot(u, Av)), float(torch.dot(v, Au))
    assert abs(a - b) < 1e-11 * abs(a)
    one = torch.ones_like(u)
    A1 = torch.empty_like(u)
    op.vmult(A1, one)
    np.testing.assert_allclose(float(A1.sum()), 2.5 * 0.4321, rtol=1e-11)
    np.testing.assert_allclose(op.compute_cut_norm(one), np.sqrt(0.4321), rtol=1e-12)

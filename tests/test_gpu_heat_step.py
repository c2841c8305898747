"""The device heat time loop (mf_heat_stepper_*, the loop body of the mp-heat-transfer
application: laser motion, volumetric source, predictor, Newton) against the oracle's
restatement of the same loop (orc.heat_time_steps), step by step."""
import numpy as np
import pytest

import orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

NC = (5, 4, 4)
LASER = dict(power=1000.0, radius=0.3, start=(0.2, 0.5, 0.8), scan_direction=(2.0, 0.0, 0.0),
             scan_speed=5.0)
TANH = dict(rho_cp=1.0, k=1.0, tanh_amp=0.5, tanh_Tc=400.0, tanh_width=50.0)


def _dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def _setup(mf):
    s = mf.ScratchData(2, NC, stream=torch.cuda.current_stream())
    mask = s.boundary_mask()
    s.set_dirichlet_mask(mask)
    T0 = np.full(s.n_dofs, 300.0)
    op = mf.HeatDiffuseOperator(s, law=mf.LAW_TANH, t_dependent=True, dt=1e-2, **TANH)
    return s, mask, T0, op


def test_stepper_matches_oracle_loop():
    import meltpooldg_amd as mf
    s, mask, T0, op = _setup(mf)
    dts = [1e-2, 1e-2, 2e-2, 1e-2]
    st = mf.HeatTimeStepper(op, _dev(T0), laser=LASER, solver="GMRES", rel_tolerance=1e-12)
    got = []
    for dt in dts:
        rep = st.step(dt, dirichlet_values=_dev(T0))
        assert rep["status"] == 0 and rep["iterations"] >= 1
        got.append(st.solution().cpu().numpy())
    hp = orc.HeatParams(law=orc.LAW_TANH, t_dependent=1, dt=1e-2, **TANH)
    ref = orc.heat_time_steps(orc.make_mesh(3, 2, NC), hp, mask, T0, dts, laser=LASER,
                              dirichlet_values=T0)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g - r)) / np.max(np.abs(r - T0)) < 1e-8
    # the laser crossed the conductivity transition: the law is exercised
    assert got[-1].max() > 450.0


def test_stepper_state_and_history():
    import meltpooldg_amd as mf
    s, mask, T0, op = _setup(mf)
    st = mf.HeatTimeStepper(op, _dev(T0), laser=LASER, start_time=0.5, solver="GMRES",
                            rel_tolerance=1e-12)
    st.step(1e-2, dirichlet_values=_dev(T0))
    T1 = st.solution().clone()
    st.step(3e-2, dirichlet_values=_dev(T0))
    assert abs(st.time - 0.54) < 1e-15
    # normalised scan direction: x moves by speed * (sum of dt)
    assert np.allclose(st.laser_position, (0.2 + 5.0 * 0.04, 0.5, 0.8), atol=1e-15, rtol=0)
    assert torch.equal(st.old_solution(), T1)
    X = s.support_points()
    q = orc.laser_volumetric_source(X, LASER["power"], LASER["radius"], st.laser_position)
    assert np.max(np.abs(st.heat_source().cpu().numpy() - q)) <= 1e-13 * q.max()


def test_predictors():
    import meltpooldg_amd as mf
    s, mask, T0, op = _setup(mf)
    out = {}
    for name, p in (("linear", mf.PREDICTOR_LINEAR_EXTRAPOLATION), ("none", mf.PREDICTOR_NONE),
                    ("zero", mf.PREDICTOR_ZERO)):
        st = mf.HeatTimeStepper(op, _dev(T0), laser=LASER, predictor=p, solver="GMRES",
                                rel_tolerance=1e-12)
        sols = []
        for dt in (1e-2, 1e-2):
            st.step(dt, dirichlet_values=_dev(T0))
            sols.append(st.solution())
        out[name] = sols
        del st
    # first step: the old increment is 0, so the linear predictor is the current solution
    assert torch.equal(out["linear"][0], out["none"][0])
    # converged Newton solutions do not depend on the initial guess beyond the tolerances
    for name in ("none", "zero"):
        d = (out[name][1] - out["linear"][1]).abs().max().item()
        assert d / (out["linear"][1] - 300.0).abs().max().item() < 1e-8


def test_stepper_rejects_bad_input():
    import meltpooldg_amd as mf
    s, mask, T0, op = _setup(mf)
    with pytest.raises(mf.MFError):
        mf.HeatTimeStepper(op, _dev(T0), predictor=7)
    bad = dict(LASER, scan_direction=(0.0, 0.0, 0.0))
    with pytest.raises(mf.MFError):
        mf.HeatTimeStepper(op, _dev(T0), laser=bad)
    st = mf.HeatTimeStepper(op, _dev(T0))
    with pytest.raises(mf.MFError):
        st.step(0.0)
    adv = mf.AdvectionDiffusionOperator(s, dt=1e-2, theta=1.0, diffusivity=0.1)
    with pytest.raises(mf.MFError):
        mf.HeatTimeStepper(adv, _dev(T0))


def test_stepper_on_adaptive_mesh_matches_oracle_loop():
    """The same loop on a band-refined mesh with hanging nodes: residual, laser source and
    Newton iterate kept constraint-consistent, against the oracle's adaptive loop."""
    import meltpooldg_amd as mf
    cells = mf.refine_band_cells(4, 1, (0.5, 0.5, 0.6), 0.3, 0.05)
    s = mf.AdaptiveScratchData(2, 4, cells, stream=torch.cuda.current_stream())
    am = orc.AdaptiveMesh(2, 4, cells)
    assert am.hanging.sum() > 0
    X = s.support_points()
    pos = {tuple(np.round(x * 2 ** 24).astype(np.int64)): i for i, x in enumerate(am.xyz)}
    perm = np.array([pos[tuple(np.round(x * 2 ** 24).astype(np.int64))] for x in X])
    op = mf.HeatDiffuseOperator(s, law=mf.LAW_TANH, t_dependent=True, dt=1e-2, **TANH)
    T0 = np.full(s.n_dofs, 300.0)
    st = mf.HeatTimeStepper(op, _dev(T0), laser=LASER, solver="GMRES", rel_tolerance=1e-12)
    dts = [1e-2, 1e-2, 1e-2]
    got = []
    for dt in dts:
        rep = st.step(dt, dirichlet_values=_dev(T0))
        assert rep["status"] == 0
        got.append(st.solution().cpu().numpy())
    hp = orc.HeatParams(law=orc.LAW_TANH, t_dependent=1, dt=1e-2, **TANH)
    T0a = np.full(am.n_nodes, 300.0)
    ref = orc.heat_time_steps(None, hp, am.dirichlet, T0a, dts, laser=LASER,
                              dirichlet_values=T0a, amesh=am)
    for g, r in zip(got, ref):
        r = r[perm]
        assert np.max(np.abs(g - r)) / np.max(np.abs(r - 300.0)) < 1e-8
    assert got[-1].max() > 450.0


def test_least_squares_projection_matches_oracle():
    """mf_predictor_least_squares against the oracle's restatement on the tanh Jacobian (three
    old solutions, general rhs) and on the advection operator (non-symmetric); a dependent old
    solution stops the projection as in the oracle."""
    import meltpooldg_amd as mf
    from mfcases import Case, device_operator, fields, oracle_operator, rel_err
    for case in (Case(degree=2, kind="heat", law="tanh", t_dependent=True),
                 Case(degree=3, kind="advec", mask="random")):
        f = fields(case)
        ref = oracle_operator(case, f)
        s, op = device_operator(case, f)
        m = case.mesh()
        xs = [orc.random_field(m, 30 + i) for i in range(3)]
        rhs = orc.random_field(m, 40)
        exp, used_ref = orc.least_squares_projection(ref.vmult, xs, rhs)
        sol = torch.full((s.n_dofs,), float("nan"), dtype=torch.float64, device="cuda")
        used = mf.predictor_least_squares(op, [_dev(x) for x in xs], _dev(rhs), sol)
        assert used == used_ref == 3
        assert rel_err(sol.cpu().numpy(), exp) < 1e-12
        dep = [xs[0], 2.0 * xs[0]]
        exp, used_ref = orc.least_squares_projection(ref.vmult, dep, rhs)
        used = mf.predictor_least_squares(op, [_dev(x) for x in dep], _dev(rhs), sol)
        assert used == used_ref == 1
        assert rel_err(sol.cpu().numpy(), exp) < 1e-12


def test_stepper_least_squares_predictor_matches_oracle_loop():
    import meltpooldg_amd as mf
    s, mask, T0, op = _setup(mf)
    dts = [1e-2, 1e-2, 2e-2, 1e-2, 1e-2]
    st = mf.HeatTimeStepper(op, _dev(T0), laser=LASER, solver="GMRES", rel_tolerance=1e-12,
                            predictor=mf.PREDICTOR_LEAST_SQUARES_PROJECTION)
    got = []
    for dt in dts:
        rep = st.step(dt, dirichlet_values=_dev(T0))
        assert rep["status"] == 0
        got.append(st.solution().cpu().numpy())
    hp = orc.HeatParams(law=orc.LAW_TANH, t_dependent=1, dt=1e-2, **TANH)
    ref = orc.heat_time_steps(orc.make_mesh(3, 2, NC), hp, mask, T0, dts, laser=LASER,
                              dirichlet_values=T0, predictor="least_squares")
    for g, r in zip(got, ref):
        assert np.max(np.abs(g - r)) / np.max(np.abs(r - T0)) < 1e-8

"""Pins of the oracle's heat time loop (orc.heat_time_steps) and its laser source: the source
integrates to the laser power, the loop reproduces the oracle's implicit-Euler heat steps
(pinned by the golden outputs) for a linear law whatever the predictor, keeps a steady state,
and its steps with a source satisfy the discrete step equation."""
import numpy as np
import pytest

import orc


def _mass_integral(mesh, q):
    hp = orc.HeatParams(law=orc.LAW_CONSTANT, rho_cp=1.0, k=0.0, dt=1.0)
    op = orc.OracleOperator(orc.OP_HEAT, mesh, heat=hp, use_mask=0)
    return float(op.vmult(q).sum())          # 1^T M q = integral of the interpolant


def test_laser_source_integrates_to_power():
    mesh = orc.make_mesh(3, 3, (12, 12, 12))
    X = orc.node_coords(mesh)
    q = orc.laser_volumetric_source(X, 3.0, 0.12, (0.5, 0.45, 0.55))
    assert abs(_mass_integral(mesh, q) / 3.0 - 1.0) < 2e-3
    # a beam cut in half by the top face carries half the power
    q = orc.laser_volumetric_source(X, 3.0, 0.12, (0.5, 0.45, 1.0))
    assert abs(_mass_integral(mesh, q) / 1.5 - 1.0) < 2e-3


@pytest.mark.parametrize("predictor", ["linear", "none", "zero", "least_squares"])
def test_linear_law_matches_implicit_euler(predictor):
    mesh = orc.make_mesh(3, 2, (4, 3, 3))
    mask = orc.boundary_mask(mesh)
    T0 = 1.0 + orc.random_field(mesh, 5)
    T0[mask == 1] = 0.0
    hp = orc.HeatParams(law=orc.LAW_CONSTANT, rho_cp=1.3, k=0.7, dt=1e-2)
    got = orc.heat_time_steps(mesh, hp, mask, T0, [1e-2] * 3, predictor=predictor)
    ref = orc.OracleOperator(orc.OP_HEAT, mesh, heat=hp, mask=mask)
    T_ref, _, st = ref.heat_implicit_euler(T0, 3)
    assert st == 0
    assert np.max(np.abs(got[-1] - T_ref)) < 1e-9 * np.max(np.abs(T0))


def test_steady_state_is_kept():
    mesh = orc.make_mesh(3, 2, (3, 3, 3))
    mask = orc.boundary_mask(mesh)
    T0 = np.full(orc.n_dofs(mesh), 350.0)
    hp = orc.HeatParams(law=orc.LAW_TANH, t_dependent=1, rho_cp=1.0, k=1.0, tanh_amp=0.5,
                        tanh_Tc=400.0, tanh_width=50.0, dt=1e-2)
    out = orc.heat_time_steps(mesh, hp, mask, T0, [1e-2, 2e-2], dirichlet_values=T0)
    assert np.max(np.abs(out[-1] - 350.0)) < 1e-9


def test_step_with_laser_satisfies_step_equation():
    mesh = orc.make_mesh(3, 2, (4, 4, 3))
    mask = orc.boundary_mask(mesh)
    T0 = np.full(orc.n_dofs(mesh), 300.0)
    hp = orc.HeatParams(law=orc.LAW_CONSTANT, rho_cp=1.0, k=1.0, dt=1e-2)
    laser = dict(power=50.0, radius=0.3, start=(0.3, 0.5, 0.7), scan_direction=(2.0, 0.0, 0.0),
                 scan_speed=5.0)
    out = orc.heat_time_steps(mesh, hp, mask, T0, [1e-2, 1e-2], laser=laser, dirichlet_values=T0)
    # the second step: the source at start + 2 dt speed e_x (direction normalised)
    X = orc.node_coords(mesh)
    q = orc.laser_volumetric_source(X, 50.0, 0.3, (0.3 + 2 * 1e-2 * 5.0, 0.5, 0.7))
    op = orc.OracleOperator(orc.OP_HEAT, mesh, heat=hp, mask=mask, T=out[1], T_old=out[0])
    r = op.create_rhs(out[0], q)
    # against the residual of the step's start (T = T_old, i.e. the source term alone)
    op0 = orc.OracleOperator(orc.OP_HEAT, mesh, heat=hp, mask=mask, T=out[0], T_old=out[0])
    r0 = op0.create_rhs(out[0], q)
    assert np.linalg.norm(r) < 1e-9 * np.linalg.norm(r0)
    # a source left at the first step's position does not satisfy it
    q1 = orc.laser_volumetric_source(X, 50.0, 0.3, (0.3 + 1e-2 * 5.0, 0.5, 0.7))
    assert np.linalg.norm(op.create_rhs(out[0], q1)) > 1e-3 * np.linalg.norm(r0)
    # the source heats: the interior warms above the boundary value
    assert out[1][mask == 0].max() > 300.5


def test_adaptive_loop_on_uniform_mesh_equals_structured():
    """The loop on an adaptive mesh of level-1 cells only (no hanging nodes) equals the
    structured loop on the twice finer lattice, node by node."""
    nb = (2, 2, 2)
    cells = np.array([(1, i, j, k) for k in range(4) for j in range(4) for i in range(4)],
                     dtype=np.int32)
    am = orc.AdaptiveMesh(2, nb, cells)
    m = orc.make_mesh(3, 2, (4, 4, 4))
    X = orc.node_coords(m)
    perm = {tuple(np.round(x * 2 ** 24).astype(np.int64)): i for i, x in enumerate(am.xyz)}
    perm = np.array([perm[tuple(np.round(x * 2 ** 24).astype(np.int64))] for x in X])
    hp = orc.HeatParams(law=orc.LAW_TANH, t_dependent=1, rho_cp=1.0, k=1.0, tanh_amp=0.5,
                        tanh_Tc=400.0, tanh_width=50.0, dt=1e-2)
    laser = dict(power=1000.0, radius=0.3, start=(0.2, 0.5, 0.8), scan_speed=5.0)
    mask = orc.boundary_mask(m)
    T0 = np.full(len(X), 300.0)
    ref = orc.heat_time_steps(m, hp, mask, T0, [1e-2, 1e-2], laser=laser, dirichlet_values=T0)
    ma = np.zeros(am.n_nodes, dtype=np.uint8)
    ma[perm] = mask
    T0a = np.full(am.n_nodes, 300.0)
    got = orc.heat_time_steps(None, hp, ma, T0a, [1e-2, 1e-2], laser=laser, dirichlet_values=T0a,
                              amesh=am)
    assert np.max(np.abs(got[-1][perm] - ref[-1])) < 1e-9 * np.max(np.abs(ref[-1] - 300.0))


# ---------------------------------------------------------------- least-squares projection
def _lsq_case(seed=3):
    mesh = orc.make_mesh(3, 2, (3, 3, 2))
    mask = orc.boundary_mask(mesh)
    n = orc.n_dofs(mesh)
    rng = np.random.default_rng(seed)
    T = 1700.0 + 100.0 * rng.uniform(-1, 1, n)
    hp = orc.HeatParams(law=orc.LAW_TANH, t_dependent=1, rho_cp=1.0, k=1.0, tanh_amp=0.5,
                        tanh_Tc=1700.0, tanh_width=50.0, dt=1e-2)
    op = orc.OracleOperator(orc.OP_HEAT, mesh, heat=hp, mask=mask, T=T, T_old=T)
    xs = [rng.uniform(-1, 1, n) for _ in range(3)]
    for x in xs:
        x[mask == 1] = 0.0
    return op, xs, rng


def test_least_squares_reproduces_a_combination_of_old_solutions():
    """rhs = A (a x_0 + b x_1 + c x_2): the projection returns that combination."""
    op, xs, _ = _lsq_case()
    target = 0.7 * xs[0] - 1.3 * xs[1] + 0.2 * xs[2]
    sol, used = orc.least_squares_projection(op.vmult, xs, op.vmult(target))
    assert used == 3
    assert np.max(np.abs(sol - target)) < 1e-12 * np.max(np.abs(target))


def test_least_squares_is_the_least_squares_solution():
    """For a general rhs the coefficients are the minimiser of |W c - rhs| with W = [A x_s],
    here from numpy's SVD-based lstsq (an independent method)."""
    op, xs, rng = _lsq_case(4)
    rhs = rng.uniform(-1, 1, op.n)
    sol, used = orc.least_squares_projection(op.vmult, xs, rhs)
    W = np.stack([op.vmult(x) for x in xs], axis=1)
    c, *_ = np.linalg.lstsq(W, rhs, rcond=None)
    ref = sum(ci * x for ci, x in zip(c, xs))
    assert used == 3
    assert np.max(np.abs(sol - ref)) < 1e-10 * np.max(np.abs(ref))


def test_least_squares_stops_at_a_dependent_old_solution():
    """x_1 = 2 x_0: its orthogonalised image vanishes (below 1e-16 |A x_0|^2), so only x_0
    enters, with the coefficient (A x_0 . rhs) / |A x_0|^2."""
    op, xs, rng = _lsq_case(5)
    rhs = rng.uniform(-1, 1, op.n)
    sol, used = orc.least_squares_projection(op.vmult, [xs[0], 2.0 * xs[0], xs[1]], rhs)
    v0 = op.vmult(xs[0])
    assert used == 1
    assert np.max(np.abs(sol - (v0 @ rhs) / (v0 @ v0) * xs[0])) < 1e-13 * np.max(np.abs(xs[0]))


def test_least_squares_predictor_loop_reaches_the_same_solutions():
    """The heat loop with the least-squares predictor converges to the solutions of the loop
    with linear extrapolation (Newton stops within its tolerances either way)."""
    mesh = orc.make_mesh(3, 2, (4, 3, 3))
    mask = orc.boundary_mask(mesh)
    T0 = np.full(orc.n_dofs(mesh), 300.0)
    hp = orc.HeatParams(law=orc.LAW_TANH, t_dependent=1, rho_cp=1.0, k=1.0, tanh_amp=0.5,
                        tanh_Tc=400.0, tanh_width=50.0, dt=1e-2)
    laser = dict(power=1000.0, radius=0.3, start=(0.2, 0.5, 0.8), scan_direction=(1, 0, 0),
                 scan_speed=5.0)
    kw = dict(laser=laser, dirichlet_values=T0)
    a = orc.heat_time_steps(mesh, hp, mask, T0, [1e-2] * 4, predictor="least_squares", **kw)
    b = orc.heat_time_steps(mesh, hp, mask, T0, [1e-2] * 4, predictor="linear", **kw)
    for x, y in zip(a, b):
        assert np.max(np.abs(x - y)) < 1e-8 * np.max(np.abs(y - 300.0))

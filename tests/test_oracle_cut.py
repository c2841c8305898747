"""Pins of the oracle's one-phase CutFEM heat operator (oracle/orc_cut.c): the reference's golden
run (unidirectional_heat_transfer_one_phase_cut), closed forms of the cut-cell quadrature and of
the ghost penalty, the uncut limit against the (pinned) heat operator, the assembled diagonal,
the residual/tangent consistency and, for the mushy law, the tangent as the derivative of the
residual."""
import numpy as np
import pytest

import orc

GOLDEN_ONE_PHASE_CUT = [4.04890e-03, 4.01890e-03, 3.99116e-03, 3.96522e-03, 3.94086e-03]
CV = 7850.0 * 460.0
MUSHY = dict(liquid_k=30.0, liquid_rho=7000.0, liquid_cp=800.0, solid_k=20.0, solid_rho=7500.0,
             solid_cp=600.0, solidus=1.0, liquidus=2.0)


def _golden_case(dim, lz=1.0):
    """applications/mp-heat-transfer/cases/unidirectional_heat_transfer/tests/
    unidirectional_heat_transfer_one_phase_cut.json: [0, 0.1]^2 in 2^4 | 1 = 17 cells per
    direction (unidirectional_heat_transfer.cpp:161-190), T0 = x (LinearTemp(0, 0.1)), level set
    0.05 - y (HorizontalLevelSet), dt 2, 5 steps, k 55.563, cp 460, rho 7850, gamma_A 1.5,
    gamma_M 0.75, theta 0.5 (heat_data.hpp:34), h = min cell size = 0.1 / 17."""
    if dim == 2:
        mesh = orc.make_mesh(2, 1, (17, 17), (0.0, 0.0), (0.1, 0.1))
    else:
        mesh = orc.make_mesh(3, 1, (17, 17, 1), (0.0, 0.0, 0.0), (0.1, 0.1, lz))
    X = orc.node_coords(mesh)
    p = orc.cut_params(2.0, rho_cp=CV, k=55.563, theta=0.5, gamma_A=1.5, gamma_M=0.75,
                       h=0.1 / 17)
    return mesh, X, 0.05 - X[:, 1], p


def test_golden_one_phase_cut_2d():
    mesh, X, phi, p = _golden_case(2)
    _, norms = orc.cut_time_steps(mesh, phi, p, X[:, 0].copy(), 5)
    # one unit of the sixth printed digit (relative 2.5e-6): an independent dense restatement
    # of the same algorithm gives the same norms to 1e-12; the remaining offset is the same in
    # every step (a static 1e-6 relative shift, not a drift of the dynamics)
    np.testing.assert_allclose(norms, GOLDEN_ONE_PHASE_CUT, rtol=3e-6, atol=0)


@pytest.mark.parametrize("lz", [1.0, 0.1 / 17])
def test_golden_extruded_to_3d(lz):
    mesh2, X2, phi2, p = _golden_case(2)
    _, n2 = orc.cut_time_steps(mesh2, phi2, p, X2[:, 0].copy(), 2)
    mesh3, X3, phi3, _ = _golden_case(3, lz)
    _, n3 = orc.cut_time_steps(mesh3, phi3, p, X3[:, 0].copy(), 2)
    np.testing.assert_allclose(np.array(n3) / np.sqrt(lz), n2, rtol=1e-12)


def test_classification_of_the_golden_mesh():
    mesh, X, phi, p = _golden_case(2)
    op = orc.CutOracle(mesh, phi, p)
    # rows 0-7 liquid, row 8 cut at half height, rows 9-16 gas
    cat = op.category.reshape(17, 17)
    assert np.all(cat[:8] == orc.CUT_LIQUID) and np.all(cat[8] == orc.CUT_INTERSECTED)
    assert np.all(cat[9:] == orc.CUT_GAS)
    np.testing.assert_allclose(op.box.reshape(17, 17, 3, 2)[8, :, 1], [[0.0, 0.5]] * 17,
                               atol=1e-15)
    assert op.active.sum() == 18 * 10          # the node rows of the liquid and cut cells
    # ghost-penalty faces: 17 liquid-cut faces below row 8, 16 cut-cut faces inside it
    assert len(op.faces) == 33
    assert sorted(np.bincount(op.faces[:, 1])) == [16, 17]


def test_interface_on_a_node_plane():
    """A cell touching phi = 0 from the positive side is liquid; from the negative side it is
    intersected with an empty liquid part (source/cut/util.cpp:146-169)."""
    mesh = orc.make_mesh(2, 1, (4, 4), (0.0, 0.0), (1.0, 1.0))
    X = orc.node_coords(mesh)
    op = orc.CutOracle(mesh, 0.5 - X[:, 1], orc.cut_params(1.0, rho_cp=1.0, k=1.0, mesh=mesh))
    cat = op.category.reshape(4, 4)
    assert np.all(cat[:2] == orc.CUT_LIQUID) and np.all(cat[2] == orc.CUT_INTERSECTED)
    assert np.all(cat[3] == orc.CUT_GAS)
    np.testing.assert_allclose(op.box.reshape(4, 4, 3, 2)[2, :, 1], [[0.0, 0.0]] * 4, atol=0)


def test_non_axis_aligned_cut_is_rejected():
    mesh = orc.make_mesh(2, 1, (4, 4), (0.0, 0.0), (1.0, 1.0))
    X = orc.node_coords(mesh)
    with pytest.raises(ValueError):
        orc.CutOracle(mesh, 0.9 - X[:, 0] - X[:, 1], orc.cut_params(1.0, rho_cp=1.0, k=1.0,
                                                                     mesh=mesh))


def _interfaces():
    # (axis, level, liquid below the level)
    return [(1, 0.43, True), (0, 0.61, False), (2, 0.37, True), (2, 0.5, False)]


@pytest.mark.parametrize("axis,level,below", _interfaces())
def test_cut_quadrature_moments(axis, level, below):
    """Mass of the liquid region and its first and second moments (k = 0, no ghost penalty):
    1^T M 1 = |Omega_l|, 1^T M x_d = int x_d, u^T M u = int u^2 via the cut norm."""
    mesh = orc.make_mesh(3, 1, (5, 4, 6), (0.0, 0.0, 0.0), (1.0, 0.8, 1.2))
    X = orc.node_coords(mesh)
    L = np.array([1.0, 0.8, 1.2])
    phi = (level - X[:, axis]) if below else (X[:, axis] - level)
    p = orc.cut_params(1.0, rho_cp=2.5, k=0.0, gamma_A=0.0, gamma_M=0.0, mesh=mesh)
    op = orc.CutOracle(mesh, phi, p)
    lo = np.zeros(3)
    hi = L.copy()
    if below:
        hi[axis] = level
    else:
        lo[axis] = level
    vol = np.prod(hi - lo)
    one = np.ones(len(X))
    np.testing.assert_allclose(op.vmult(one).sum(), 2.5 * vol, rtol=1e-13)
    for d in range(3):
        mom = vol * 0.5 * (lo[d] + hi[d])
        np.testing.assert_allclose(one @ op.vmult(X[:, d]), 2.5 * mom, rtol=1e-13)
        sec = vol * (hi[d] ** 3 - lo[d] ** 3) / (3.0 * (hi[d] - lo[d]))
        np.testing.assert_allclose(op.norm(X[:, d]) ** 2, sec, rtol=1e-13)
    np.testing.assert_allclose(op.norm(one) ** 2, vol, rtol=1e-13)


def test_ghost_penalty_closed_form():
    """u = y^2 interpolated: the normal-derivative jump is 2h on every y-face and 0 on x-faces,
    so u^T G u = gamma (2h)^2 h * (number of y-faces); u = x^2: the x-faces."""
    mesh, X, phi, _ = _golden_case(2)
    h = 0.1 / 17
    gm = 0.75
    p = orc.cut_params(2.0, rho_cp=0.0, k=0.0, gamma_A=0.0, gamma_M=gm, h=h)
    op = orc.CutOracle(mesh, phi, p)
    gamma = gm * h ** 3 / 3.0
    u = X[:, 1] ** 2
    np.testing.assert_allclose(u @ op.vmult(u), gamma * (2 * h) ** 2 * h * 17, rtol=1e-10)
    u = X[:, 0] ** 2
    np.testing.assert_allclose(u @ op.vmult(u), gamma * (2 * h) ** 2 * h * 16, rtol=1e-10)
    # linear fields have no jumps
    scale = gamma / h                           # size of the face matrix entries (2D)
    for u in (X[:, 0], X[:, 1], 1.0 + 0 * X[:, 0]):
        assert np.max(np.abs(op.vmult(u))) < 1e-12 * scale * np.abs(u).max()


def test_uncut_limit_is_the_heat_operator():
    """Without an interface the tangent is cv M + dt theta k K: the oracle's heat tangent with
    rho_cp/dt = cv and k -> dt theta k (pinned by the heat goldens), also with Dirichlet rows."""
    mesh = orc.make_mesh(3, 1, (4, 3, 5), (0.0, 0.0, 0.0), (1.0, 0.6, 1.1))
    X = orc.node_coords(mesh)
    mask = orc.boundary_mask(mesh)
    p = orc.cut_params(0.3, rho_cp=2.0, k=1.7, theta=0.6, mesh=mesh)
    cut = orc.CutOracle(mesh, 5.0 + 0 * X[:, 0], p, mask=mask)
    assert len(cut.faces) == 0 and np.all(cut.category == orc.CUT_LIQUID)
    hp = orc.HeatParams(law=orc.LAW_CONSTANT, rho_cp=2.0, k=0.3 * 0.6 * 1.7, dt=1.0)
    heat = orc.OracleOperator(orc.OP_HEAT, mesh, heat=hp, mask=mask)
    u = orc.random_field(mesh, 3)
    np.testing.assert_allclose(cut.vmult(u), heat.vmult(u), rtol=0, atol=1e-13)


def test_diagonal_symmetry_and_conservation():
    mesh = orc.make_mesh(3, 1, (4, 5, 3), (0.0, 0.0, 0.0), (1.0, 1.0, 0.7))
    X = orc.node_coords(mesh)
    p = orc.cut_params(0.5, rho_cp=3.0, k=2.0, gamma_A=1.5, gamma_M=0.75, mesh=mesh)
    op = orc.CutOracle(mesh, X[:, 1] - 0.33, p)
    A = op.dense()
    free = op.active == 1
    np.testing.assert_allclose(A, A.T, atol=1e-13 * np.abs(A).max())
    d = op.diagonal()
    np.testing.assert_allclose(d[free], np.diag(A)[free], rtol=1e-13)
    assert np.all(d[~free] == 1.0)
    assert np.all(A[~free] == 0.0) and np.all(A[:, ~free] == 0.0)
    # stiffness and ghost penalty annihilate constants: 1^T (A - cv M) = 0
    pk = orc.cut_params(0.5, rho_cp=0.0, k=2.0, gamma_A=1.5, gamma_M=0.75, mesh=mesh)
    opk = orc.CutOracle(mesh, X[:, 1] - 0.33, pk)
    u = orc.random_field(mesh, 9)
    assert abs(opk.vmult(u).sum()) < 1e-13 * np.abs(u).sum()


def test_residual_consistent_with_tangent():
    """Constant law: -R(T) = cv M T_old - dt(1-theta) k K T_old - J T, so -R(T) + J T does not
    depend on T, and -R vanishes at the step's solution."""
    mesh = orc.make_mesh(3, 1, (4, 4, 3), (0.0, 0.0, 0.0), (1.0, 1.0, 0.6))
    X = orc.node_coords(mesh)
    p = orc.cut_params(0.4, rho_cp=3.0, k=2.0, theta=0.7, gamma_A=1.2, gamma_M=0.6, mesh=mesh)
    op = orc.CutOracle(mesh, 0.71 - X[:, 0], p)
    To = orc.random_field(mesh, 1)
    T1, T2 = orc.random_field(mesh, 2), orc.random_field(mesh, 3)
    a = op.residual(T1, To) + op.vmult(T1)
    b = op.residual(T2, To) + op.vmult(T2)
    np.testing.assert_allclose(a, b, atol=1e-12 * np.abs(a).max())
    temps, _ = orc.cut_time_steps(mesh, 0.71 - X[:, 0], p, To, 1)
    assert np.max(np.abs(op.residual(temps[0], To))) < 1e-12 * np.abs(a).max()


def test_mushy_tangent_is_the_residual_derivative():
    """do_solidification: the tangent's d(rho cp)/dT T u and dk/dT grad T u terms
    (tangent_domain_integral, heat_cut_operator.cpp:157-190) are the derivative of the residual
    (rho cp(T) T and k(T) grad T, residual_domain_integral) along any direction."""
    mesh = orc.make_mesh(3, 1, (3, 4, 3), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    X = orc.node_coords(mesh)
    p = orc.cut_params(0.05, theta=0.5, gamma_A=1.5, gamma_M=0.75, mushy=MUSHY, mesh=mesh)
    op = orc.CutOracle(mesh, 0.62 - X[:, 1], p)
    T = 1.5 + 0.3 * orc.random_field(mesh, 4)     # inside the mushy interval
    To = 1.5 + 0.3 * orc.random_field(mesh, 5)
    v = orc.random_field(mesh, 6)
    eps = 1e-6
    fd = -(op.residual(T + eps * v, To) - op.residual(T - eps * v, To)) / (2 * eps)
    Jv = op.vmult(v, T)
    np.testing.assert_allclose(Jv, fd, atol=2e-7 * np.abs(Jv).max())
    # the derivative terms matter: the constant-coefficient tangent at the same T misses it
    pc = orc.cut_params(0.05, rho_cp=7000.0 * 800.0, k=30.0, theta=0.5, gamma_A=1.5,
                        gamma_M=0.75, mesh=mesh)
    assert np.abs(orc.CutOracle(mesh, 0.62 - X[:, 1], pc).vmult(v) - fd).max() > \
        1e-2 * np.abs(Jv).max()

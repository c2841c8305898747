"""Pins of the oracle's adaptive-mesh path (hanging-node constraints, operator, diagonal, solve).

Each check is one that a plausible mistake fails:
- a mesh of level-0 (or of level-1) cells only, given as an adaptive mesh, reproduces the
  structured oracle on the same (or the twice finer) lattice: node identification, per-cell
  geometry and the gather/scatter without constraints;
- the hanging-node constraints reproduce every global polynomial of degree p (a wrong weight,
  master or chain step breaks this at some hanging node);
- the condensed Laplacian annihilates linear functions at interior free nodes and the condensed
  mass matrix integrates 1 to the volume (conformity of the constrained space);
- the operator is symmetric where the bilinear form is, and non-symmetric with convection;
- the diagonal equals the diagonal of the condensed matrix assembled column by column;
- CG on the condensed system matches a dense direct solve.
"""
import numpy as np
import pytest

import orc

MESHES = {
    "q1": (1, (2, 2, (0.1, 0.1, 0.1), 0.2, 0.03)),
    "q2": (2, (3, 2, (0.2, 0.3, 0.25), 0.3, 0.04)),
    "q3": (3, (2, 2, (0.1, 0.1, 0.1), 0.2, 0.03)),
}


def _amesh(name):
    degree, args = MESHES[name]
    cells = orc.refine_band_cells(*args)
    return orc.AdaptiveMesh(degree, args[0], cells), args[0]


def _heat(rho_cp=1.3, k=0.7, dt=1e-2, law=orc.LAW_CONSTANT, td=0):
    return orc.HeatParams(law=law, t_dependent=td, rho_cp=rho_cp, k=k, tanh_amp=0.5,
                          tanh_Tc=1700.0, tanh_width=50.0, dt=dt)


def _coord_map(xa, xb):
    """index map b -> a of two node sets with the same coordinates"""
    def keys(x):
        return [tuple(r) for r in np.round(x * 2 ** 20).astype(np.int64)]
    da = {k: i for i, k in enumerate(keys(xa))}
    return np.array([da[k] for k in keys(xb)])


@pytest.mark.parametrize("level", [0, 1])
@pytest.mark.parametrize("degree", [1, 2, 3])
def test_uniform_adaptive_equals_structured(degree, level):
    nb = (2, 3, 2)
    n = [c << level for c in nb]
    cells = np.array([(level, i, j, k) for k in range(n[2]) for j in range(n[1])
                      for i in range(n[0])], dtype=np.int32)
    am = orc.AdaptiveMesh(degree, nb, cells, hi=(1.0, 1.5, 0.8))
    m = orc.make_mesh(3, degree, tuple(n), hi=(1.0, 1.5, 0.8))
    assert am.n_nodes == orc.n_dofs(m) and am.hanging.sum() == 0
    xs = orc.node_coords(m)
    perm = _coord_map(am.xyz, xs)          # structured index -> adaptive index
    T = 1700 + 150 * orc.random_field(m, 3)
    kc = np.exp(0.3 * np.random.default_rng(1).standard_normal(orc.n_cells(m)))
    # cell order: the adaptive list above is lexicographic like the structured numbering
    hp = _heat(law=orc.LAW_TANH, td=1)
    ref = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, T=T, T_old=T, k_cell=kc,
                             mask=orc.boundary_mask(m))
    Ta = np.empty(am.n_nodes)
    Ta[perm] = T
    mask_a = np.empty(am.n_nodes, dtype=np.uint8)
    mask_a[perm] = orc.boundary_mask(m)
    op = orc.OracleOperator(orc.OP_HEAT, None, heat=hp, T=Ta, T_old=Ta, k_cell=kc, mask=mask_a,
                            amesh=am)
    u = orc.random_field(m, 5)
    ua = np.empty(am.n_nodes)
    ua[perm] = u
    ya = op.vmult(ua)
    y = ref.vmult(u)
    assert np.abs(ya[perm] - y).max() <= 1e-12 * np.abs(y).max()
    da, d = op.diagonal(), ref.diagonal()
    assert np.abs(da[perm] - d).max() <= 1e-12 * np.abs(d).max()


@pytest.mark.parametrize("name", ["q1", "q2", "q3"])
def test_constraints_reproduce_polynomials(name):
    am, _ = _amesh(name)
    p = am.degree
    assert am.hanging.sum() > 0
    x = am.xyz

    def f(x):
        out = 1.0 + 0.5 * x[:, 0] - 0.25 * x[:, 1] + 0.75 * x[:, 2]
        if p >= 2:
            out = out + x[:, 0] * x[:, 1] - 0.5 * x[:, 2] ** 2 + 0.3 * x[:, 0] * x[:, 1] * x[:, 2] * (p >= 3)
        if p >= 3:
            out = out + x[:, 0] ** 3 - 2.0 * x[:, 1] ** 2 * x[:, 2] + 0.7 * x[:, 2] ** 3
        return out

    v = f(x)
    w = am.distribute(np.where(am.hanging == 1, 0.0, v))
    assert np.abs(w - v).max() < 1e-13
    # a polynomial of degree p+1 is not reproduced (the pin is not vacuous)
    g = v + x[:, 0] ** (p + 1) * x[:, 1]
    assert np.abs(am.distribute(np.where(am.hanging == 1, 0.0, g)) - g).max() > 1e-6
    # masters are free nodes
    assert np.all(am.hanging[am.line_col] == 0)


@pytest.mark.parametrize("name", ["q1", "q2", "q3"])
def test_conformity_mass_and_laplace(name):
    am, _ = _amesh(name)
    x = am.xyz
    # mass: 1^T M 1 = (rho cp / dt) |Omega| with no Dirichlet rows
    op_m = orc.OracleOperator(orc.OP_HEAT, None, heat=_heat(rho_cp=1.0, k=0.0, dt=1.0),
                              amesh=am, use_mask=0)
    one = np.ones(am.n_nodes)
    y = op_m.vmult(one)
    assert abs(y[am.hanging == 0].sum() - 1.0) < 1e-12
    assert np.all(y[am.hanging == 1] == 0.0)
    # Laplace of a linear function vanishes at free interior nodes
    op_k = orc.OracleOperator(orc.OP_HEAT, None, heat=_heat(rho_cp=0.0, k=1.0, dt=1.0),
                              amesh=am, use_mask=0)
    lin = 0.3 + x[:, 0] - 2.0 * x[:, 1] + 0.5 * x[:, 2]
    z = op_k.vmult(lin)
    inner = (am.boundary == 0) & (am.hanging == 0)
    assert np.abs(z[inner]).max() < 1e-11
    assert np.abs(z[am.boundary == 1]).max() > 1e-3   # boundary rows carry the flux


@pytest.mark.parametrize("name", ["q2", "q3"])
def test_symmetry_and_diagonal(name):
    am, _ = _amesh(name)
    n = am.n_nodes
    rng = np.random.default_rng(4)
    free = (am.hanging == 0) & (am.dirichlet == 0)
    op = orc.OracleOperator(orc.OP_HEAT, None, heat=_heat(), mask=am.dirichlet, amesh=am)
    u, v = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    u[~free] = 0
    v[~free] = 0
    a, b = v @ op.vmult(u), u @ op.vmult(v)
    assert abs(a - b) < 1e-12 * abs(a)
    vel = rng.uniform(-1, 1, 3 * n)
    opv = orc.OracleOperator(orc.OP_HEAT, None, heat=_heat(), mask=am.dirichlet, amesh=am,
                             vel_kind=orc.VEL_NODAL, vel=vel)
    assert abs(v @ opv.vmult(u) - u @ opv.vmult(v)) > 1e-6 * abs(a)
    # diagonal of the condensed matrix
    idx = np.nonzero(free)[0]
    sel = idx[:: max(1, len(idx) // 60)]
    d = opv.diagonal()
    for j in sel:
        e = np.zeros(n)
        e[j] = 1.0
        assert abs(opv.vmult(e)[j] - d[j]) < 1e-12 * abs(d[j])
    assert np.all(d[am.hanging == 1] == 0) and np.all(d[am.dirichlet == 1] == 0)


def test_cg_on_condensed_system_matches_dense():
    am, _ = _amesh("q1")
    n = am.n_nodes
    free = np.nonzero((am.hanging == 0) & (am.dirichlet == 0))[0]
    op = orc.OracleOperator(orc.OP_HEAT, None, heat=_heat(dt=1e-3), mask=am.dirichlet, amesh=am)
    A = np.zeros((len(free), len(free)))
    for c, j in enumerate(free):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, c] = op.vmult(e)[free]
    b = np.zeros(n)
    b[free] = np.random.default_rng(2).uniform(-1, 1, len(free))
    x, it, res, st = op.solve(b, solver=orc.CG, rel_tol=1e-12, inv_diag=op.inverse_diagonal())
    assert st == orc.SUCCESS and it > 3
    xd = np.linalg.solve(A, b[free])
    assert np.abs(x[free] - xd).max() < 1e-9 * np.abs(xd).max()


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_fe_evaluate_adaptive_reproduces_polynomials(degree):
    """orc.fe_evaluate_adaptive (the transfer check's evaluator): a field of Q_p given by its
    nodal values on a hanging-node mesh is reproduced at arbitrary points, and at the nodes a
    field outside Q_p returns its nodal values."""
    cells = orc.refine_band_cells(2, 1, (0.5, 0.5, 0.5), 0.3, 0.1)
    am = orc.AdaptiveMesh(degree, 2, cells)
    p = degree

    def f(X):
        x, y, z = X[:, 0], X[:, 1], X[:, 2]
        return 1.0 + x - 2.0 * y ** p + 0.5 * x * y * z + (x * z) ** p

    pts = np.random.default_rng(1).uniform(0, 1, (60, 3))
    got = orc.fe_evaluate_adaptive(degree, 2, cells, am.xyz, f(am.xyz), pts)
    assert np.max(np.abs(got - f(pts))) < 1e-12
    g = np.sin(3 * am.xyz[:, 0]) * np.cos(2 * am.xyz[:, 1]) + am.xyz[:, 2] ** 5
    g = am.distribute(g)
    free = am.hanging == 0
    got = orc.fe_evaluate_adaptive(degree, 2, cells, am.xyz, g, am.xyz[free])
    assert np.max(np.abs(got - g[free])) < 1e-12


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_heat_rhs_uniform_adaptive_equals_structured(degree):
    """The adaptive heat residual (orc_create_rhs on an adaptive mesh) on level-1 cells only is
    the structured residual on the twice finer lattice (tanh law, source term, Dirichlet rows)."""
    nb = (2, 2, 1)
    n = [c << 1 for c in nb]
    cells = np.array([(1, i, j, k) for k in range(n[2]) for j in range(n[1])
                      for i in range(n[0])], dtype=np.int32)
    am = orc.AdaptiveMesh(degree, nb, cells)
    m = orc.make_mesh(3, degree, tuple(n))
    perm = _coord_map(am.xyz, orc.node_coords(m))
    T = 1650 + 150 * orc.random_field(m, 3)
    T_old = 1650 + 150 * orc.random_field(m, 4)
    q = orc.random_field(m, 5)
    hp = _heat(law=orc.LAW_TANH, td=1)
    mask = orc.boundary_mask(m)
    ref = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, mask=mask, T=T).create_rhs(T_old, q)
    Ta, Toa, qa, ma = (np.zeros(am.n_nodes) for _ in range(4))
    Ta[perm], Toa[perm], qa[perm] = T, T_old, q
    ma[perm] = mask
    op = orc.OracleOperator(orc.OP_HEAT, None, heat=hp, amesh=am, mask=ma.astype(np.uint8), T=Ta)
    got = op.create_rhs(Toa, qa)
    assert np.max(np.abs(got[perm] - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", list(MESHES))
def test_heat_rhs_consistent_with_condensed_operator(name):
    """Linear law, constraint-consistent fields: rhs(T, T_old = 0, q = 0) = -A T and
    rhs(T = 0, T_old, q) = (rho_cp/dt) M T_old + M q with the condensed (pinned) operators."""
    am, _ = _amesh(name)
    rng = np.random.default_rng(9)
    T = am.distribute(rng.uniform(-1, 1, am.n_nodes))
    T_old = am.distribute(rng.uniform(-1, 1, am.n_nodes))
    q = am.distribute(rng.uniform(-1, 1, am.n_nodes))
    hp = _heat()
    z = np.zeros(am.n_nodes)
    op = orc.OracleOperator(orc.OP_HEAT, None, heat=hp, amesh=am, use_mask=0, T=T)
    A = orc.OracleOperator(orc.OP_HEAT, None, heat=hp, amesh=am, use_mask=0)
    hang = am.hanging == 1
    r = op.create_rhs(z)
    assert np.max(np.abs((r + A.vmult(T))[~hang])) < 1e-12 * np.max(np.abs(r))
    op0 = orc.OracleOperator(orc.OP_HEAT, None, heat=hp, amesh=am, use_mask=0, T=z)
    r = op0.create_rhs(T_old, q)
    M = orc.OracleOperator(orc.OP_HEAT, None, heat=_heat(rho_cp=1.0, k=0.0, dt=1.0), amesh=am,
                           use_mask=0)
    expect = (1.3 / 1e-2) * M.vmult(T_old) + M.vmult(q)
    assert np.max(np.abs((r - expect)[~hang])) < 1e-12 * np.max(np.abs(expect))


def test_advection_steps_uniform_adaptive_equals_structured():
    """The oracle's advection steps (Dirichlet lifting, GMRES) on an adaptive mesh of level-1
    cells only equal the structured steps on the twice finer lattice."""
    nb = (2, 2, 2)
    cells = np.array([(1, i, j, k) for k in range(4) for j in range(4) for i in range(4)],
                     dtype=np.int32)
    am = orc.AdaptiveMesh(2, nb, cells)
    m = orc.make_mesh(3, 2, (4, 4, 4))
    X = orc.node_coords(m)
    perm = _coord_map(am.xyz, X)            # structured index -> adaptive index
    vel = np.stack([2 * (X[:, 2] - 0.5), 0.3 * np.ones(len(X)), -2 * (X[:, 0] - 0.5)])
    phi0 = np.tanh((np.linalg.norm(X - 0.5, axis=1) - 0.25) / 0.1)
    mask = orc.boundary_mask(m)
    ap = orc.AdvecParams(dt=0.05, theta=0.5, diffusivity=0.01, mass_factor=1.0, vel_scale=0.0)
    ref = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL, vel=vel.ravel(),
                             mask=mask)
    phi_s, _, st = ref.advec_theta_steps(phi0, 2, phi0)
    assert st == 0
    na = am.n_nodes
    va, pa, ma = np.zeros((3, na)), np.zeros(na), np.zeros(na, dtype=np.uint8)
    va[:, perm], pa[perm], ma[perm] = vel, phi0, mask
    opa = orc.OracleOperator(orc.OP_ADVEC, None, advec=ap, vel_kind=orc.VEL_NODAL,
                             vel=va.ravel(), mask=ma, amesh=am)
    phi_a, _, st = opa.advec_theta_steps(pa, 2, pa)
    assert st == 0
    assert np.max(np.abs(phi_a[perm] - phi_s)) < 1e-9

"""Pins of the oracle's level-set normal vector and curvature projections.

Each check is one that a plausible mistake fails:
- without a narrow band the projection tangent is the heat tangent with rho cp / dt = 1 and the
  conductivity eta = filter_parameter * h^2 (a wrong damping, a dropped mass or gradient term);
- the sum-factorised tangent equals the brute-force element matrices, with a narrow band too;
- narrow-band extremes: a threshold above every |phi| is no band, one below every |phi| (or 0)
  is the zero operator, whose inverse diagonal is 1 by the reference's rule;
- the normal vector operator is the scalar operator on each block;
- the normal rhs of a linear level set a.x + b is a_d times the integral of each basis function
  (the mass matrix applied to 1), and with normalisation a_d / |a| times it;
- the curvature rhs: zero for a constant normal, invariant under scaling of the nodal normal
  above the tolerance and linear below it, and its sum is minus the flux of the interpolated
  unit normal through the boundary (divergence theorem, evaluated with 1D Lagrange integrals);
- compute_numerical_zero_of_norm at closed-form volumes;
- on adaptive meshes the tangent is the heat tangent with the per-cell conductivity eta(h_cell)
  and the normal rhs of a linear level set is again the mass matrix applied to 1.
"""
import numpy as np
import pytest

import orc

FP = 2.0


def ls_op(m, kind=orc.OP_CURVATURE, fp=FP, band=None, phi=None, normalized=0, amesh=None):
    ls = orc.LsParams(filter_parameter=fp, narrow_band=0 if band is None else 1,
                      narrow_band_threshold=0.0 if band is None else band,
                      compute_normalized_vector=normalized)
    return orc.OracleOperator(kind, m, ls=ls, level_set=phi, use_mask=0, amesh=amesh)


def eta(m, fp=FP):
    h = [(m.hi[d] - m.lo[d]) / m.ncell[d] for d in range(m.dim)]
    return fp * sum(x * x for x in h) / m.dim


def sphere_phi(xyz, c=(0.45, 0.5, 0.52), r=0.3, eps=0.08):
    d = np.linalg.norm(xyz - np.asarray(c)[None, :xyz.shape[1]], axis=1) - r
    return -np.tanh(d / (2 * eps))


MESHES = [(2, 1, (3, 4)), (2, 3, (2, 3)), (3, 1, (3, 2, 2)), (3, 2, (2, 2, 3)), (3, 3, (2, 1, 2))]


@pytest.mark.parametrize("dim,deg,nc", MESHES)
def test_tangent_is_mass_plus_eta_laplace(dim, deg, nc):
    m = orc.make_mesh(dim, deg, nc, (0.0,) * dim, (1.0, 0.7, 1.2)[:dim])
    op = ls_op(m)
    hp = orc.HeatParams(law=orc.LAW_CONSTANT, t_dependent=0, rho_cp=1.0, k=eta(m), dt=1.0)
    heat = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, use_mask=0)
    u = np.random.default_rng(3).uniform(-1, 1, op.n)
    a, b = op.vmult(u), heat.vmult(u)
    assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b))
    # the mass part integrates 1 to the volume, the eta part annihilates constants
    vol = np.prod([m.hi[d] - m.lo[d] for d in range(dim)])
    assert abs(ls_op(m, fp=0.0).vmult(np.ones(op.n)).sum() - vol) < 1e-13
    assert abs(op.vmult(np.ones(op.n)).sum() - vol) < 1e-13


@pytest.mark.parametrize("dim,deg,nc", MESHES)
def test_narrow_band_sum_factorisation_matches_dense(dim, deg, nc):
    m = orc.make_mesh(dim, deg, nc, (0.0,) * dim, (1.0, 0.7, 1.2)[:dim])
    xyz = orc.node_coords(m)[:, :dim]
    phi = sphere_phi(xyz)
    op = ls_op(m, band=0.6, phi=phi)
    u = np.random.default_rng(4).uniform(-1, 1, op.n)
    a, b = op.vmult(u), op.vmult_dense(u)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    # the band changes the operator on this field (the pin above is not vacuous)
    assert np.max(np.abs(a - ls_op(m).vmult(u))) > 1e-3 * np.max(np.abs(a))
    # diagonal = diagonal of the assembled matrix
    n = op.n
    A = np.stack([op.vmult(np.eye(n)[j]) for j in range(n)], axis=1)
    assert np.allclose(op.diagonal(), np.diag(A), rtol=1e-13, atol=1e-16)


def test_narrow_band_extremes():
    m = orc.make_mesh(3, 2, (2, 2, 2), (0, 0, 0), (1, 1, 1))
    phi = sphere_phi(orc.node_coords(m))
    u = np.random.default_rng(5).uniform(-1, 1, orc.n_dofs(m))
    free = ls_op(m).vmult(u)
    assert np.array_equal(ls_op(m, band=1.5, phi=phi).vmult(u), free)
    for thr in (0.0, 1e-300):
        zero = ls_op(m, band=thr, phi=phi + 2.0)
        assert np.all(zero.vmult(u) == 0.0)
        assert np.all(zero.inverse_diagonal() == 1.0)


def test_normal_operator_is_blockwise():
    m = orc.make_mesh(3, 2, (2, 3, 2), (0, 0, 0), (1.0, 1.2, 0.9))
    phi = sphere_phi(orc.node_coords(m))
    n = orc.n_dofs(m)
    nop = ls_op(m, kind=orc.OP_NORMAL, band=0.7, phi=phi)
    sop = ls_op(m, band=0.7, phi=phi)
    assert nop.n == 3 * n
    u = np.random.default_rng(6).uniform(-1, 1, 3 * n)
    y = nop.vmult(u)
    for d in range(3):
        assert np.array_equal(y[d * n:(d + 1) * n], sop.vmult(u[d * n:(d + 1) * n]))
    dg = nop.diagonal()
    assert np.array_equal(dg, np.tile(sop.diagonal(), 3))


@pytest.mark.parametrize("deg", [1, 2, 3])
@pytest.mark.parametrize("normalized", [0, 1])
def test_normal_rhs_linear_level_set(deg, normalized):
    m = orc.make_mesh(3, deg, (2, 3, 2), (0.1, 0.0, -0.2), (1.0, 1.2, 0.9))
    xyz = orc.node_coords(m)
    a = np.array([0.3, -1.1, 0.7])
    phi = xyz @ a + 0.25
    n = orc.n_dofs(m)
    nop = ls_op(m, kind=orc.OP_NORMAL, normalized=normalized)
    rhs = nop.normal_rhs(phi)
    m1 = ls_op(m, fp=0.0).vmult(np.ones(n))  # integral of every basis function
    coef = a / np.linalg.norm(a) if normalized else a
    for d in range(3):
        assert np.allclose(rhs[d * n:(d + 1) * n], coef[d] * m1, rtol=0, atol=1e-14)
    # dst is added to
    assert np.allclose(nop.normal_rhs(phi, dst=rhs), 2 * rhs, rtol=0, atol=1e-14)
    # a band that excludes every point removes the source
    assert np.all(ls_op(m, kind=orc.OP_NORMAL, band=0.01, phi=phi + 10).normal_rhs(phi + 10) == 0)


def test_normal_rhs_zero_gradient_normalisation():
    m = orc.make_mesh(3, 1, (2, 2, 2), (0, 0, 0), (1, 1, 1))
    n = orc.n_dofs(m)
    # |grad phi| <= 1e-16: the reference divides by 1e-16 instead of |grad phi|
    xyz = orc.node_coords(m)
    phi = 1e-18 * xyz[:, 0]
    rhs = ls_op(m, kind=orc.OP_NORMAL, normalized=1).normal_rhs(phi)
    m1 = ls_op(m, fp=0.0).vmult(np.ones(n))
    assert np.allclose(rhs[:n], 1e-2 * m1, rtol=1e-12, atol=0)
    assert np.max(np.abs(rhs[n:])) < 1e-12 * np.max(np.abs(rhs[:n]))  # roundoff of 0 / 1e-16


def lagrange_integrals(K):
    """int_0^1 l_j for the Lagrange basis on the K Gauss-Lobatto points (numpy polynomials)"""
    x = orc.gauss_lobatto(K)
    out = []
    for j in range(K):
        others = np.delete(x, j)
        p = np.poly(others) / np.prod(x[j] - others)
        P = np.polyint(p)
        out.append(np.polyval(P, 1.0) - np.polyval(P, 0.0))
    return np.array(out)


def boundary_flux(m, field):
    """sum over the 6 sides of int_side I_h(field) . nu, field [3][n] nodal"""
    K = m.degree + 1
    W = lagrange_integrals(K)
    NN = [m.ncell[d] * m.degree + 1 for d in range(3)]
    h = [(m.hi[d] - m.lo[d]) / m.ncell[d] for d in range(3)]
    # 1D weights of the nodal interpolant on a line of NN nodes
    def line_w(d):
        w = np.zeros(NN[d])
        for c in range(m.ncell[d]):
            w[c * m.degree:c * m.degree + K] += W * h[d]
        return w
    n = orc.n_dofs(m)
    F = field.reshape(3, NN[2], NN[1], NN[0])  # [comp][z][y][x]
    total = 0.0
    for d in range(3):
        t = [e for e in range(3) if e != d]
        wa, wb = line_w(t[0]), line_w(t[1])
        for upper, sgn in ((0, -1.0), (1, 1.0)):
            idx = NN[d] - 1 if upper else 0
            comp = F[d]
            face = np.take(comp, idx, axis=2 - d)  # remaining axes in order (t[1], t[0])
            total += sgn * np.einsum("ba,a,b->", face, wa, wb)
    return total


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_curvature_rhs_divergence_theorem(deg):
    m = orc.make_mesh(3, deg, (3, 2, 2), (0.0, 0.1, 0.0), (1.0, 0.9, 0.8))
    xyz = orc.node_coords(m)
    n = orc.n_dofs(m)
    c = np.array([0.4, 0.45, -0.3])           # centre below the box: |x - c| >= 0.3 > tol
    v = (xyz - c).T.copy()                     # radial, normalised at the nodes by the operator
    op = ls_op(m)
    rhs = op.curvature_rhs(v.ravel())
    unit = v / np.linalg.norm(v, axis=0)
    assert abs(rhs.sum() + boundary_flux(m, unit.ravel())) < 1e-12
    # div of the unit radial field is 2/r: the rhs approximates -int v 2/r
    m1 = ls_op(m, fp=0.0).vmult(np.ones(n))
    approx = -(2.0 / np.linalg.norm(xyz - c, axis=1)) * m1
    assert abs(rhs.sum() - approx.sum()) < 2e-2 * abs(approx.sum())
    # invariant under positive scaling above the tolerance
    assert np.allclose(op.curvature_rhs(3.7 * v.ravel()), rhs, rtol=0, atol=1e-13)


def test_curvature_rhs_constant_and_small_normals():
    m = orc.make_mesh(3, 2, (2, 2, 2), (0, 0, 0), (1, 1, 1))
    n = orc.n_dofs(m)
    op = ls_op(m)
    const = np.concatenate([np.full(n, 0.3), np.full(n, -0.5), np.full(n, 0.8)])
    assert np.max(np.abs(op.curvature_rhs(const))) < 1e-13
    tol = orc.numerical_zero_of_norm(m)
    assert tol == pytest.approx(1e-3, rel=1e-15)
    xyz = orc.node_coords(m)
    small = 1e-6 * np.concatenate([xyz[:, 0] ** 2, xyz[:, 1], np.zeros(n)])  # |n| < tol
    r1, r2 = op.curvature_rhs(small), op.curvature_rhs(2 * small)
    assert np.max(np.abs(r1)) > 0
    assert np.allclose(r2, 2 * r1, rtol=1e-13, atol=1e-18)
    # below the tolerance n/tol is used: the rhs of small with tol equals that of small/tol
    # with a tolerance of 1, which small/tol (|.| <= 1e-3) is also below
    assert np.allclose(r1, op.curvature_rhs(small / tol, tol=1.0), rtol=1e-13, atol=1e-18)


def test_numerical_zero_of_norm_closed_forms():
    f = orc.lib().orc_numerical_zero_of_norm
    assert f(3, 1.0) == pytest.approx(1e-3)
    assert f(3, 1e-12) == pytest.approx(1e-7)      # side 1e-4
    assert f(3, 1e6) == pytest.approx(1e-2)        # side 100: capped at 1e-2
    assert f(3, 1e-30) == pytest.approx(1e-12)     # side 1e-10: floored at 1e-12
    assert f(2, 0.25) == pytest.approx(1e-4)       # side 0.5 -> 10^-1 * 1e-3


# ---------------------------------------------------------------- adaptive meshes
def _amesh(degree):
    cells = orc.refine_band_cells(2, 2, (0.3, 0.4, 0.35), 0.25, 0.04, hi=(1.0, 1.2, 0.9))
    return orc.AdaptiveMesh(degree, 2, cells, hi=(1.0, 1.2, 0.9)), cells


@pytest.mark.parametrize("degree", [1, 2])
def test_adaptive_tangent_and_normal_rhs(degree):
    am, cells = _amesh(degree)
    m = orc.make_mesh(3, degree, (2, 2, 2), (0, 0, 0), (1.0, 1.2, 0.9))
    op = ls_op(m, amesh=am)
    h0 = np.array([0.5, 0.6, 0.45])
    eta_cell = np.array([FP * np.sum((h0 / 2 ** l) ** 2) / 3 for l in cells[:, 0]])
    hp = orc.HeatParams(law=orc.LAW_CONSTANT, t_dependent=0, rho_cp=1.0, k=1.0, dt=1.0)
    heat = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, k_cell=eta_cell, use_mask=0, amesh=am)
    u = am.distribute(np.random.default_rng(7).uniform(-1, 1, am.n_nodes))
    a, b = op.vmult(u), heat.vmult(u)
    assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b))
    # normal rhs of a linear level set: a_d times the condensed mass matrix applied to 1
    aa = np.array([0.5, 0.2, -0.9])
    phi = am.xyz @ aa - 0.1
    nop = ls_op(m, kind=orc.OP_NORMAL, amesh=am)
    rhs = nop.normal_rhs(phi)
    m1 = ls_op(m, fp=0.0, amesh=am).vmult(np.ones(am.n_nodes))
    n = am.n_nodes
    for d in range(3):
        assert np.allclose(rhs[d * n:(d + 1) * n], aa[d] * m1, rtol=0, atol=1e-14)
    assert np.all(m1[am.hanging == 1] == 0)


# ---------------------------------------------------------------- Olsson reinitialization
def reinit_op(m, normal, dt=0.01, value=0.5, factor=0.0, ttype=orc.THICKNESS_PROPORTIONAL,
              tol=None, amesh=None):
    rp = orc.ReinitParams(dt=dt, thickness_type=ttype, thickness_value=value,
                          tangential_diffusion_factor=factor,
                          normal_tolerance=orc.numerical_zero_of_norm(m) if tol is None else tol)
    return orc.OracleOperator(orc.OP_REINIT, m, vel_kind=orc.VEL_NODAL, vel=normal, use_mask=0,
                              reinit=rp, amesh=amesh)


def cell_size(m):
    h = [(m.hi[d] - m.lo[d]) / m.ncell[d] for d in range(m.dim)]
    return np.sqrt(sum(x * x for x in h) / m.dim)


RMESHES = [(3, 1, (3, 2, 2)), (3, 2, (2, 2, 3)), (3, 3, (2, 1, 2))]


@pytest.mark.parametrize("dim,deg,nc", RMESHES)
def test_reinit_tangent_closed_forms(dim, deg, nc):
    m = orc.make_mesh(dim, deg, nc, (0.0,) * dim, (1.0, 0.7, 1.2))
    n = orc.n_dofs(m)
    u = np.random.default_rng(8).uniform(-1, 1, n)
    eps = 0.5 * cell_size(m)
    dt = 0.02
    # three axis-aligned normals without tangential diffusion sum to 3/dt M + eps K
    total = np.zeros(n)
    for d in range(3):
        e = np.zeros((3, n))
        e[d] = 2.5  # normalised by the operator
        total += reinit_op(m, e.ravel(), dt=dt).vmult(u)
    hp = orc.HeatParams(law=orc.LAW_CONSTANT, t_dependent=0, rho_cp=3.0, k=eps, dt=dt)
    heat = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, use_mask=0)
    ref = heat.vmult(u)
    assert np.max(np.abs(total - ref)) <= 1e-13 * np.max(np.abs(ref))
    # tangential factor 1: isotropic eps whatever the normal field
    rnd = np.random.default_rng(9).uniform(-1, 1, 3 * n)
    hp1 = orc.HeatParams(law=orc.LAW_CONSTANT, t_dependent=0, rho_cp=1.0, k=eps, dt=dt)
    iso = orc.OracleOperator(orc.OP_HEAT, m, heat=hp1, use_mask=0).vmult(u)
    a = reinit_op(m, rnd, dt=dt, factor=1.0).vmult(u)
    assert np.max(np.abs(a - iso)) <= 1e-13 * np.max(np.abs(iso))
    # the three thickness laws agree where they give the same eps
    b1 = reinit_op(m, rnd, dt=dt, value=eps, ttype=orc.THICKNESS_ABSOLUTE).vmult(u)
    b2 = reinit_op(m, rnd, dt=dt, value=3.0, ttype=orc.THICKNESS_CELLS).vmult(u)
    b0 = reinit_op(m, rnd, dt=dt).vmult(u)
    assert np.max(np.abs(b1 - b0)) <= 1e-13 * np.max(np.abs(b0))
    assert np.max(np.abs(b2 - b0)) <= 1e-13 * np.max(np.abs(b0))


@pytest.mark.parametrize("dim,deg,nc", RMESHES)
def test_reinit_sum_factorisation_matches_dense(dim, deg, nc):
    m = orc.make_mesh(dim, deg, nc, (0.0,) * dim, (1.0, 0.7, 1.2))
    n = orc.n_dofs(m)
    rng = np.random.default_rng(10)
    normal = rng.uniform(-1, 1, (3, n))
    normal[:, ::5] *= 1e-5  # some nodal normals far below the tolerance
    op = reinit_op(m, normal.ravel(), factor=0.3)
    u = rng.uniform(-1, 1, n)
    a, b = op.vmult(u), op.vmult_dense(u)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    A = np.stack([op.vmult(np.eye(n)[j]) for j in range(n)], axis=1)
    assert np.allclose(op.diagonal(), np.diag(A), rtol=1e-13, atol=1e-16)
    assert np.allclose(A, A.T, rtol=1e-12, atol=1e-14)  # D is symmetric


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_reinit_rhs_moments(deg):
    m = orc.make_mesh(3, deg, (2, 3, 2), (0.1, 0.0, -0.2), (1.0, 1.2, 0.9))
    xyz = orc.node_coords(m)
    n = orc.n_dofs(m)
    L = np.array([0.9, 1.2, 1.1])
    vol = np.prod(L)
    ex = np.zeros((3, n))
    ex[0] = 1.0
    # constant psi, normal e_x: only the compressive flux; its moments are closed forms
    c = 0.3
    r = reinit_op(m, ex.ravel()).reinit_rhs(np.full(n, c))
    assert abs(r.sum()) < 1e-13
    assert abs(r @ xyz[:, 0] - 0.5 * (1 - c * c) * vol) < 1e-13
    assert abs(r @ xyz[:, 1]) < 1e-13
    # linear psi = a.x + b, isotropic diffusion eps: sum_i x_i r_i = int f_x,
    # f = (1 - psi^2)/2 e_x - eps a
    a = np.array([0.4, -0.3, 0.2])
    b0 = -0.1
    psi = xyz @ a + b0
    eps = 0.5 * cell_size(m)
    r = reinit_op(m, ex.ravel(), factor=1.0).reinit_rhs(psi)
    center = np.array([0.55, 0.6, 0.35])
    int_psi2 = vol * ((a @ center + b0) ** 2 + np.sum(a * a * L * L) / 12.0)
    assert abs(r @ xyz[:, 0] - (0.5 * (vol - int_psi2) - eps * a[0] * vol)) < 1e-12
    assert abs(r @ xyz[:, 1] - (-eps * a[1] * vol)) < 1e-12
    assert abs(r @ xyz[:, 2] - (-eps * a[2] * vol)) < 1e-12
    # without tangential diffusion the gradient part is eps_n (a . n) n = eps a_x e_x
    r0 = reinit_op(m, ex.ravel(), factor=0.0).reinit_rhs(psi)
    assert abs(r0 @ xyz[:, 1]) < 1e-12
    assert abs(r0 @ xyz[:, 0] - (0.5 * (vol - int_psi2) - eps * a[0] * vol)) < 1e-12


@pytest.mark.parametrize("degree", [1, 2])
def test_reinit_adaptive_isotropic_is_heat(degree):
    am, cells = _amesh(degree)
    m = orc.make_mesh(3, degree, (2, 2, 2), (0, 0, 0), (1.0, 1.2, 0.9))
    h0 = np.array([0.5, 0.6, 0.45])
    eps_cell = np.array([0.5 * np.sqrt(np.sum((h0 / 2 ** l) ** 2) / 3) for l in cells[:, 0]])
    rnd = np.random.default_rng(12).uniform(-1, 1, 3 * am.n_nodes)
    op = reinit_op(m, rnd, dt=0.05, factor=1.0, amesh=am)
    hp = orc.HeatParams(law=orc.LAW_CONSTANT, t_dependent=0, rho_cp=1.0, k=1.0, dt=0.05)
    heat = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, k_cell=eps_cell, use_mask=0, amesh=am)
    u = am.distribute(np.random.default_rng(13).uniform(-1, 1, am.n_nodes))
    a, b = op.vmult(u), heat.vmult(u)
    assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b))

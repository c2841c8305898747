"""Pins for every oracle function that the golden outputs leave open: closed forms of the 1D
basis and quadrature, the dense (brute-force) element-matrix path against the sum-factorised
path, invariants of the heat and advection operators, the diagonal against the assembled
matrix, and the Krylov solvers against dense direct solves."""
import itertools

import numpy as np
import pytest

import orc


# ---------------------------------------------------------------- 1D basis and quadrature
@pytest.mark.parametrize("n", range(1, 9))
def test_gauss_exactness(n):
    x, w = orc.gauss(n)
    for k in range(2 * n):
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-14
    # not exact one degree higher (a wrong point set would pass the lower degrees only by luck)
    assert abs(np.dot(w, x ** (2 * n)) - 1.0 / (2 * n + 1)) > 1e-13


def test_gauss_lobatto_closed_forms():
    assert np.allclose(orc.gauss_lobatto(2), [0, 1], atol=0, rtol=0)
    assert np.allclose(orc.gauss_lobatto(3), [0, 0.5, 1], atol=1e-16)
    r5 = 1 / np.sqrt(5)
    assert np.allclose(orc.gauss_lobatto(4), [0, (1 - r5) / 2, (1 + r5) / 2, 1], atol=1e-15)
    r37 = np.sqrt(3 / 7)
    assert np.allclose(orc.gauss_lobatto(5), [0, (1 - r37) / 2, 0.5, (1 + r37) / 2, 1],
                       atol=1e-15)


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
def test_shape_matrices_reproduce_polynomials(degree):
    K = degree + 1
    S, DS, Dq, w = orc.shape_matrices(degree)
    nodes = orc.gauss_lobatto(K)
    xq, _ = orc.gauss(K)
    for k in range(degree + 1):
        f = nodes ** k
        assert np.allclose(S @ f, xq ** k, atol=1e-13)
        df = k * xq ** (k - 1) if k > 0 else np.zeros(K)
        assert np.allclose(DS @ f, df, atol=1e-12)
        assert np.allclose(Dq @ (xq ** k), df, atol=1e-12)
    # partition of unity / Kronecker at the nodes
    assert np.allclose(S.sum(axis=1), 1.0, atol=1e-14)
    assert np.allclose(DS.sum(axis=1), 0.0, atol=1e-12)
    v = np.zeros(K)
    d = np.zeros(K)
    for j, xj in enumerate(nodes):
        orc.lib().orc_lagrange(K, nodes.ctypes.data_as(orc.C.POINTER(orc.C.c_double)),
                               orc.C.c_double(xj), v.ctypes.data_as(orc.C.POINTER(orc.C.c_double)),
                               d.ctypes.data_as(orc.C.POINTER(orc.C.c_double)))
        assert np.allclose(v, np.eye(K)[j], atol=1e-14)


# ---------------------------------------------------------------- operators
def heat_op(m, law="const", td=0, vel=False, kcell=False, mask=None, seed=0, rho_cp=1.3, k=0.7,
            dt=1e-2, use_mask=1):
    n = orc.n_dofs(m)
    rng = np.random.default_rng(seed)
    T = 1700 + 100 * rng.uniform(-1, 1, n) if law == "tanh" else 300 + rng.uniform(-1, 1, n)
    hp = orc.HeatParams(law=orc.LAW_TANH if law == "tanh" else orc.LAW_CONSTANT, t_dependent=td,
                        rho_cp=rho_cp, k=k, tanh_amp=0.5, tanh_Tc=1700, tanh_width=50, dt=dt)
    return orc.OracleOperator(orc.OP_HEAT, m, heat=hp,
                              vel_kind=orc.VEL_NODAL if vel else orc.VEL_NONE, mask=mask, T=T,
                              T_old=T - 1.0, vel=rng.uniform(-1, 1, 3 * n) if vel else None,
                              k_cell=np.exp(rng.normal(0, .3, orc.n_cells(m))) if kcell else None,
                              use_mask=use_mask)


def advec_op(m, mask=None, seed=0, theta=0.5, D=0.1, supg=0, supg_coefficient=-1.0):
    n = orc.n_dofs(m)
    rng = np.random.default_rng(seed)
    ap = orc.AdvecParams(dt=0.1, theta=theta, diffusivity=D, mass_factor=1.0, vel_scale=0,
                         supg=supg, supg_coefficient=supg_coefficient)
    return orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL, mask=mask,
                              vel=rng.uniform(-1, 1, 3 * n))


def assemble(op):
    n = op.n
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = op.vmult(e)
    return A


MESHES = [(1, 1, (4,)), (1, 3, (3,)), (2, 1, (3, 2)), (2, 2, (2, 3)), (2, 3, (2, 2)),
          (3, 1, (2, 3, 2)), (3, 2, (2, 2, 1)), (3, 3, (1, 2, 1))]


@pytest.mark.parametrize("dim,deg,nc", MESHES)
@pytest.mark.parametrize("variant", ["heat", "heat-tanh-td", "heat-vel-kc", "advec",
                                     "advec-supg", "advec-supg-coef"])
def test_sum_factorisation_matches_dense(dim, deg, nc, variant):
    """The dense path builds each local matrix from products of 1D basis values, derivatives
    and second derivatives at each quadrature point; the product path sweeps the 1D
    matrices.  With SUPG the dense path's Laplacian term is rank-one per quadrature point
    (lap u (w . grad v)) and its streamline term a square (sw . grad u)(sw . grad v)."""
    hi = (1.0, 0.8, 1.3)[:dim]
    m = orc.make_mesh(dim, deg, nc, (0.0,) * dim, hi)
    mask = orc.boundary_mask(m)
    if variant == "heat":
        op = heat_op(m, mask=mask)
    elif variant == "heat-tanh-td":
        op = heat_op(m, law="tanh", td=1, mask=mask)
    elif variant == "heat-vel-kc":
        op = heat_op(m, vel=True, kcell=True, td=1, law="tanh")
    elif variant == "advec":
        op = advec_op(m, mask=mask)
    elif variant == "advec-supg":
        op = advec_op(m, mask=mask, supg=1)
    else:
        op = advec_op(m, mask=mask, supg=1, supg_coefficient=0.7, theta=1.0)
    u = np.random.default_rng(1).uniform(-1, 1, op.n)
    a, b = op.vmult(u), op.vmult_dense(u)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))


@pytest.mark.parametrize("dim,deg", [(1, 2), (2, 1), (2, 3), (3, 1), (3, 2), (3, 3)])
def test_mass_sums_to_volume(dim, deg):
    """rho_cp/dt * M applied to 1, summed, equals rho_cp * |Omega| / dt (k = 0)."""
    hi = (1.0, 0.8, 1.3)[:dim]
    m = orc.make_mesh(dim, deg, (3, 2, 2)[:dim], (0.0,) * dim, hi)
    op = heat_op(m, k=0.0, rho_cp=2.5, dt=0.5)
    vol = np.prod(hi)
    assert abs(op.vmult(np.ones(op.n)).sum() - 2.5 * vol / 0.5) < 1e-12 * vol * 5


@pytest.mark.parametrize("dim,deg", [(2, 2), (3, 1), (3, 3)])
def test_laplace_annihilates_constants_and_interior_linears(dim, deg):
    hi = (1.0, 0.8, 1.3)[:dim]
    m = orc.make_mesh(dim, deg, (3, 3, 3)[:dim], (0.0,) * dim, hi)
    op = heat_op(m, rho_cp=0.0, k=1.0)
    assert np.max(np.abs(op.vmult(np.ones(op.n)))) < 1e-12
    xyz = orc.node_coords(m)
    lin = xyz[:, 0] - 2.0 * xyz[:, dim - 1]
    r = op.vmult(lin)
    interior = orc.boundary_mask(m) == 0
    assert np.max(np.abs(r[interior])) < 1e-12
    # and the boundary rows carry the Neumann flux: sum of all rows = boundary integral of dn(u)=0
    assert abs(r.sum()) < 1e-12


def test_symmetry_and_definiteness_constant_law():
    m = orc.make_mesh(3, 2, (3, 2, 2), (0, 0, 0), (1, 0.7, 1.2))
    op = heat_op(m, mask=orc.boundary_mask(m))
    rng = np.random.default_rng(3)
    u, v = rng.uniform(-1, 1, op.n), rng.uniform(-1, 1, op.n)
    free = orc.boundary_mask(m) == 0
    u[~free] = 0
    v[~free] = 0
    assert abs(v @ op.vmult(u) - u @ op.vmult(v)) < 1e-12 * abs(v @ op.vmult(u))
    assert u @ op.vmult(u) > 0


def test_advection_linear_field_and_skew_part():
    """With D = 0, theta = 1 and phi = a.x, the convective part integrates w.a against the test
    functions: A phi - M phi/dt = M_w a.  Check with the dense element matrices' skew part:
    for divergence-free constant w and zero boundary values the convective matrix is skew."""
    m = orc.make_mesh(3, 2, (3, 3, 3))
    n = orc.n_dofs(m)
    w = np.zeros(3 * n)
    w[0:n] = 0.3
    w[n:2 * n] = -0.2
    w[2 * n:] = 0.5
    mask = orc.boundary_mask(m)
    ap = orc.AdvecParams(dt=1.0, theta=1.0, diffusivity=0.0, mass_factor=0.0, vel_scale=0)
    op = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL, mask=mask, vel=w)
    A = assemble(op)
    free = mask == 0
    C = A[np.ix_(free, free)]
    assert np.max(np.abs(C + C.T)) < 1e-13
    # linear field: row sums of (w . grad) phi_j for phi = a.x equal (w.a) * (M 1) in the interior
    xyz = orc.node_coords(m)
    a = np.array([1.0, 2.0, -1.0])
    phi = xyz @ a
    op_nomask = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL, vel=w)
    conv = op_nomask.vmult(phi)
    hp = orc.HeatParams(law=0, t_dependent=0, rho_cp=1.0, k=0.0, dt=1.0)
    mass1 = orc.OracleOperator(orc.OP_HEAT, m, heat=hp).vmult(np.ones(n))
    assert np.allclose(conv, (0.3 * 1.0 - 0.2 * 2.0 + 0.5 * -1.0) * mass1, atol=1e-13)


@pytest.mark.parametrize("variant", ["heat", "heat-tanh-td-vel", "advec", "advec-supg"])
def test_diagonal_matches_assembled(variant):
    m = orc.make_mesh(3, 2, (2, 2, 3), (0, 0, 0), (1, 1.1, 0.9))
    mask = orc.boundary_mask(m)
    mask[5] = 0
    if variant == "heat":
        op = heat_op(m, mask=mask)
    elif variant == "heat-tanh-td-vel":
        op = heat_op(m, law="tanh", td=1, vel=True, kcell=True, mask=mask)
    elif variant == "advec":
        op = advec_op(m, mask=mask)
    else:
        op = advec_op(m, mask=mask, supg=1)
    A = assemble(op)
    d = op.diagonal()
    assert np.max(np.abs(d - np.diag(A))) < 1e-13 * np.max(np.abs(d))
    inv = op.inverse_diagonal()
    free = mask == 0
    assert np.allclose(inv[free], 1.0 / np.diag(A)[free], rtol=1e-14)
    assert np.all(inv[~free] == 1.0)   # zero diagonal of constrained rows inverts to 1


def test_inversion_thresholds():
    d = np.array([1e-11, -2e-10, 0.0, 4.0, 1e-15, 3e-13])
    h = d.copy()
    orc.lib().orc_invert_diagonal(orc.OP_HEAT, orc.C.c_long(len(h)), orc._p(h))
    assert list(h) == [1.0, 1 / -2e-10, 1.0, 0.25, 1.0, 1.0]
    a = d.copy()
    orc.lib().orc_invert_diagonal(orc.OP_ADVEC, orc.C.c_long(len(a)), orc._p(a))
    # threshold 1e-14 * max(1, |d|_inf) = 4e-14
    assert list(a) == [1 / 1e-11, 1 / -2e-10, 1.0, 0.25, 1.0, 1 / 3e-13]


# ---------------------------------------------------------------- solvers
def test_cg_and_gmres_match_dense_solve():
    m = orc.make_mesh(3, 2, (3, 3, 2), (0, 0, 0), (1, 1, 0.7))
    mask = orc.boundary_mask(m)
    op = heat_op(m, mask=mask)
    A = assemble(op)
    free = mask == 0
    rng = np.random.default_rng(5)
    b = rng.uniform(-1, 1, op.n)
    b[~free] = 0
    x_dense = np.zeros(op.n)
    x_dense[free] = np.linalg.solve(A[np.ix_(free, free)], b[free])
    for solver in (orc.CG, orc.GMRES):
        x, it, res, st = op.solve(b, solver=solver, rel_tol=1e-13, inv_diag=op.inverse_diagonal())
        assert st == 0
        assert np.max(np.abs(x - x_dense)) < 1e-11 * np.max(np.abs(x_dense))
        assert res <= 1e-13 * np.linalg.norm(b)


def test_gmres_nonsymmetric_matches_dense_and_finite_termination():
    m = orc.make_mesh(3, 1, (3, 2, 2))
    mask = orc.boundary_mask(m)
    mask[:] = 0
    op = advec_op(m, theta=1.0, D=0.02)
    A = assemble(op)
    b = np.random.default_rng(6).uniform(-1, 1, op.n)
    x_dense = np.linalg.solve(A, b)
    n = op.n
    x, it, res, st = op.solve(b, solver=orc.GMRES, restart=n, rel_tol=1e-12)
    assert st == 0 and it <= n
    assert np.max(np.abs(x - x_dense)) < 1e-9 * np.max(np.abs(x_dense))
    # unpreconditioned CG on a non-symmetric system must not be trusted: GMRES is the fallback
    x, it, res, st = op.solve(b, solver=orc.GMRES, restart=5, rel_tol=1e-12,
                              inv_diag=op.inverse_diagonal())
    assert st == 0
    assert np.max(np.abs(x - x_dense)) < 1e-9 * np.max(np.abs(x_dense))


def test_reduction_control_semantics():
    """Success when |r| <= max(abs_tol, rel_tol*|r0|); max_it reached -> NoConvergence;
    zero rhs and zero guess -> converged at step 0."""
    m = orc.make_mesh(3, 1, (3, 3, 3))
    op = heat_op(m, mask=orc.boundary_mask(m), dt=1.0)
    b = np.random.default_rng(7).uniform(-1, 1, op.n)
    b[orc.boundary_mask(m) == 1] = 0
    x, it, res, st = op.solve(b, solver=orc.CG, max_it=1)
    assert st == orc.NO_CONVERGENCE and it == 1
    x, it, res, st = op.solve(np.zeros(op.n), solver=orc.CG)
    assert st == 0 and it == 0
    x, it, res, st = op.solve(b, solver=orc.CG, abs_tol=1e3, rel_tol=0.0)
    assert st == 0 and it == 0


def test_fe_l2_norm_closed_form():
    """||x||_L2 on [0,1]^3 = 1/sqrt(3), exact for the interpolant of a linear field."""
    m = orc.make_mesh(3, 2, (2, 3, 2))
    xyz = orc.node_coords(m)
    v = orc.fe_l2_norm(m, xyz[:, 0])
    assert abs(v - 1 / np.sqrt(3)) < 1e-7  # per-cell float storage (reference semantics)


def test_random_field_is_numbering_independent_and_uniform():
    m = orc.make_mesh(3, 3, (4, 4, 4))
    a = orc.random_field(m, 1)
    assert a.min() >= -1 and a.max() < 1 and abs(a.mean()) < 0.02
    b = orc.random_field(m, 2)
    assert np.corrcoef(a, b)[0, 1] < 0.05
    # the value depends only on the lattice coordinates and the seed
    m2 = orc.make_mesh(3, 3, (4, 4, 5))
    c = orc.random_field(m2, 1)
    n0 = 13 * 13 * 13
    assert np.array_equal(a[:n0], c[:n0])


# --- boundary faces of the heat operator (Robin / Neumann), textbook face quadrature ---------

def _face_ops(degree=2, ncell=(2, 3, 2), hi=(1.0, 1.5, 0.7), T=None, **bnd):
    m = orc.make_mesh(3, degree, ncell, (0, 0, 0), hi)
    n = orc.n_dofs(m)
    hp = orc.HeatParams(law=orc.LAW_CONSTANT, t_dependent=0, rho_cp=1.0, k=1.0, dt=1e-3)
    b = orc.HeatBoundary()
    for key, val in bnd.items():
        if key == "neumann":
            for side, g in val.items():
                b.neumann_sides |= 1 << side
                b.neumann_flux[side] = g
        elif key.endswith("_sides"):
            setattr(b, key, sum(1 << s_ for s_ in val))
        else:
            setattr(b, key, val)
    T = np.zeros(n) if T is None else T
    kw = dict(heat=hp, mask=np.zeros(n, np.uint8), T=T, T_old=T)
    return m, n, orc.OracleOperator(orc.OP_HEAT, m, **kw), \
        orc.OracleOperator(orc.OP_HEAT, m, boundary=b, **kw)


def _side_area(hi, side):
    d = side // 2
    return np.prod([hi[e] for e in range(3) if e != d])


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_convection_face_term_integrates_constants(degree):
    """sum_i (B 1)_i = alpha |faces|: the Robin term of u = 1 integrates alpha over the sides."""
    hi = (1.0, 1.5, 0.7)
    m, n, op0, opb = _face_ops(degree, hi=hi, convection_sides=[0, 3, 5],
                               convection_coefficient=2.5)
    one = np.ones(n)
    face = opb.vmult(one) - op0.vmult(one)
    area = sum(_side_area(hi, s_) for s_ in (0, 3, 5))
    assert abs(face.sum() - 2.5 * area) < 1e-12 * 2.5 * area


def test_radiation_face_term_uses_T_cubed():
    """Radiation tangent 4 eps sigma T^3 u: with T constant it is a mass term on the side."""
    hi = (1.0, 1.5, 0.7)
    m = orc.make_mesh(3, 2, (2, 3, 2), (0, 0, 0), hi)
    Tc = 1700.0
    _, n, op0, opb = _face_ops(2, hi=hi, T=np.full(orc.n_dofs(m), Tc), radiation_sides=[2],
                               emissivity=0.4)
    one = np.ones(n)
    face = opb.vmult(one) - op0.vmult(one)
    expect = 4 * 0.4 * orc.STEFAN_BOLTZMANN * Tc ** 3 * _side_area(hi, 2)
    assert abs(face.sum() - expect) < 1e-12 * expect


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_face_term_linear_moment(degree):
    """A linear field u = y on the side x = 1: sum_i (B u)_i = alpha * int y ds = alpha A y_c."""
    hi = (1.0, 1.5, 0.7)
    m, n, op0, opb = _face_ops(degree, hi=hi, convection_sides=[1], convection_coefficient=1.7)
    X = orc.node_coords(m)
    u = X[:, 1]
    face = opb.vmult(u) - op0.vmult(u)
    expect = 1.7 * _side_area(hi, 1) * hi[1] / 2
    assert abs(face.sum() - expect) < 1e-12 * expect


def test_face_term_symmetric_and_diagonal_consistent():
    """B is symmetric (v.Bu = u.Bv) and its diagonal (mode of compute_diagonal) equals the
    diagonal of B assembled from unit vectors."""
    m, n, op0, opb = _face_ops(2, ncell=(2, 2, 1), hi=(1.0, 1.0, 0.5),
                               T=1500.0 + 100.0 * orc.random_field(orc.make_mesh(3, 2, (2, 2, 1), (0, 0, 0), (1.0, 1.0, 0.5)), 3),
                               convection_sides=[0, 4], radiation_sides=[3, 5],
                               convection_coefficient=3.0, emissivity=0.8)
    B = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        B[:, i] = opb.vmult(e) - op0.vmult(e)
    assert np.abs(B - B.T).max() < 1e-12 * np.abs(B).max()
    dB = opb.diagonal() - op0.diagonal()
    assert np.abs(dB - np.diag(B)).max() < 1e-12 * np.abs(B).max()


def test_face_rhs_terms():
    """Boundary part of -R(T): zero for T = T_inf (convection and radiation), and
    sum_i = -g |side| on a Neumann side."""
    hi = (1.0, 1.5, 0.7)
    m = orc.make_mesh(3, 2, (2, 3, 2), (0, 0, 0), hi)
    n = orc.n_dofs(m)
    Tinf = 900.0
    T = np.full(n, Tinf)
    _, _, op0, opb = _face_ops(2, hi=hi, T=T, convection_sides=[0, 1], radiation_sides=[2, 3],
                               convection_coefficient=5.0, convection_temperature_infinity=Tinf,
                               emissivity=0.7, radiation_temperature_infinity=Tinf,
                               neumann={5: 2.0})
    r = opb.create_rhs(T) - op0.create_rhs(T)
    # T at the quadrature points is 900 (1 + O(1e-16)); T^4 - T_inf^4 then carries 4 T^4 times
    # that, i.e. ~1e-11 after the eps sigma factor
    assert abs(r.sum() + 2.0 * _side_area(hi, 5)) < 1e-9
    # only the Neumann side contributes
    X = orc.node_coords(m)
    off = np.abs(X[:, 2] - hi[2]) > 1e-12
    assert np.abs(r[off]).max() < 1e-10


# ---------------------------------------------------------------- temperature-dependent laws
def mushy_params(**kw):
    """Water/ice of the reference's solidification_slab case (solidification_slab.json)."""
    p = dict(liquid_k=0.556, liquid_rho=1000.0, liquid_cp=4226.0, solid_k=2.22,
             solid_rho=1000.0, solid_cp=1762.0, solidus=270.0, liquidus=276.0, t_dependent=1,
             dt=1.0)
    p.update(kw)
    return orc.HeatParams(law=orc.LAW_MUSHY, **p)


def test_tanh_law_closed_form():
    """k(T) = k0 + a tanh(x), dk/dT = a (1 - tanh(x)^2) / w with x = (T - Tc)/w, evaluated with
    numpy's tanh (the oracle uses tanh and cosh from libm: an independent form of sech^2)."""
    hp = orc.HeatParams(law=orc.LAW_TANH, rho_cp=1.7, k=1.0, tanh_amp=0.5, tanh_Tc=1700.0,
                        tanh_width=50.0, dt=1.0)
    for T in np.r_[np.linspace(1000.0, 2400.0, 57), 1700.0, 1700.0 + 1e-9]:
        rc, k, drc, dk = orc.material(hp, T)
        t = np.tanh((T - 1700.0) / 50.0)
        assert rc == 1.7 and drc == 0.0
        assert abs(k - (1.0 + 0.5 * t)) < 4e-16 * abs(k)
        assert abs(dk - 0.5 * (1.0 - t * t) / 50.0) <= 1e-15 * 0.5 / 50.0
    # values where the closed form is exact: k(Tc) = k0, dk/dT(Tc) = a/w
    rc, k, drc, dk = orc.material(hp, 1700.0)
    assert k == 1.0 and dk == 0.5 / 50.0


def test_mushy_law_closed_form():
    """Mushy zone: f = clamp((T_l - T)/(T_l - T_s), 0, 1), each property l + (s - l)(3f^2 - 2f^3),
    rho cp = rho cp, derivatives from the chain rule; the pure phases outside [T_s, T_l]; the
    blend is C^1 (the derivative vanishes at both ends); derivatives against central differences
    of the oracle's own values."""
    hp = mushy_params(liquid_rho=900.0)  # unequal densities exercise d(rho)/dT * cp too
    Ts, Tl = 270.0, 276.0
    for T in (250.0, Ts, Tl, 300.0):
        rc, k, drc, dk = orc.material(hp, T)
        solid = T <= Ts
        assert k == (2.22 if solid else 0.556)
        assert rc == (1000.0 * 1762.0 if solid else 900.0 * 4226.0)
        assert drc == 0.0 and dk == 0.0
    for T in np.linspace(270.05, 275.95, 40):
        f = (Tl - T) / (Tl - Ts)
        H = 3 * f ** 2 - 2 * f ** 3
        dHdT = -(6 * f - 6 * f ** 2) / (Tl - Ts)
        k_ = 0.556 + (2.22 - 0.556) * H
        rho = 900.0 + 100.0 * H
        cp = 4226.0 + (1762.0 - 4226.0) * H
        rc, k, drc, dk = orc.material(hp, T)
        assert abs(k - k_) < 1e-15 * 2.22 * 4
        assert abs(rc - rho * cp) < 1e-15 * rc * 4
        assert abs(dk - (2.22 - 0.556) * dHdT) < 1e-13
        assert abs(drc - (100.0 * cp + (1762.0 - 4226.0) * rho) * dHdT) < 1e-9 * abs(rc)
        h = 1e-5
        kp, km = orc.material(hp, T + h), orc.material(hp, T - h)
        assert abs((kp[1] - km[1]) / (2 * h) - dk) < 1e-7 * 2.22
        assert abs((kp[0] - km[0]) / (2 * h) - drc) < 1e-7 * rc
    # C^1 at both ends: the blend's derivative tends to zero
    assert abs(orc.material(hp, Ts + 1e-7)[3]) < 1e-6 and abs(orc.material(hp, Tl - 1e-7)[3]) < 1e-6


def _fd_case(law, td, vel, kcell, seed=11, degree=2, ncell=(3, 2, 2), dim=3):
    m = orc.make_mesh(dim, degree, ncell if dim == 3 else ncell[:dim], hi=(1.0, 0.8, 1.2)[:dim] +
                      (1.0,) * (3 - dim))
    n = orc.n_dofs(m)
    rng = np.random.default_rng(seed)
    if law == "tanh":
        hp = orc.HeatParams(law=orc.LAW_TANH, t_dependent=td, rho_cp=1.3, k=1.0, tanh_amp=0.5,
                            tanh_Tc=1700.0, tanh_width=50.0, dt=1e-2)
        T = 1700.0 + 120.0 * rng.uniform(-1, 1, n)
        T_old = T - 5.0 * rng.uniform(-1, 1, n)
    else:
        hp = mushy_params(t_dependent=td)
        hp.dt = 50.0
        T = 273.0 + 4.0 * rng.uniform(-1, 1, n)   # spans solid, mushy and liquid nodes
        T_old = T + 2.0 * rng.uniform(-1, 1, n)
    kw = {}
    if vel:
        kw.update(vel_kind=orc.VEL_NODAL, vel=rng.uniform(-1, 1, 3 * n))
    if kcell:
        kw.update(k_cell=np.exp(0.3 * rng.standard_normal(orc.n_cells(m))))
    q = rng.uniform(-1, 1, n)
    u = rng.uniform(-1, 1, n)
    return m, hp, T, T_old, q, u, kw


def _residual(m, hp, T, T_old, q, kw):
    """R(T) = -create_rhs(T_old, q) at linearisation point T (no Dirichlet rows)."""
    op = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, T=T, T_old=T_old, use_mask=0, **kw)
    return -op.create_rhs(T_old, q)


@pytest.mark.parametrize("law,vel,kcell", [("tanh", False, False), ("tanh", True, False),
                                           ("tanh", False, True), ("tanh", True, True),
                                           ("mushy", False, False), ("mushy", True, True)])
def test_tangent_is_the_residual_derivative(law, vel, kcell):
    """The Jacobian J = dR/dT of the Newton step: J u equals the central difference
    (R(T + eps u) - R(T - eps u)) / (2 eps) of the residual that create_rhs returns (with its
    sign), at ~1e-8 relative.  This pins every T-dependent term of the tangent (dk/dT grad T u,
    d(rho cp)/dT (T - T_old)/dt u, d(rho cp)/dT (grad T . w) u) against the residual's k(T) grad T,
    rho cp(T)(T - T_old)/dt and rho cp(T) grad T . w, which the reference's own golden outputs
    pin (heat_diffuse_operator.cpp:393-529 against :955-1086)."""
    m, hp, T, T_old, q, u, kw = _fd_case(law, 1, vel, kcell)
    op = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, T=T, T_old=T_old, use_mask=0, **kw)
    Ju = op.vmult(u)
    eps = 1e-4 if law == "tanh" else 2e-5
    fd = (_residual(m, hp, T + eps * u, T_old, q, kw) -
          _residual(m, hp, T - eps * u, T_old, q, kw)) / (2 * eps)
    err = np.abs(Ju - fd).max() / np.abs(Ju).max()
    assert err < 1e-7, err


@pytest.mark.parametrize("law", ["tanh", "mushy"])
def test_fd_check_detects_a_dropped_tangent_term(law):
    """Without the T-derivative terms (t_dependent = 0) the same check fails by orders of
    magnitude: the FD comparison above is sensitive to the dk/dT and d(rho cp)/dT terms."""
    m, hp, T, T_old, q, u, kw = _fd_case(law, 1, True, False)
    hp0 = orc.HeatParams(*(getattr(hp, f) for f, _ in orc.HeatParams._fields_))
    hp0.t_dependent = 0
    op = orc.OracleOperator(orc.OP_HEAT, m, heat=hp0, T=T, T_old=T_old, use_mask=0, **kw)
    Ju = op.vmult(u)
    eps = 1e-4 if law == "tanh" else 2e-5
    fd = (_residual(m, hp, T + eps * u, T_old, q, kw) -
          _residual(m, hp, T - eps * u, T_old, q, kw)) / (2 * eps)
    assert np.abs(Ju - fd).max() / np.abs(Ju).max() > 1e-4


@pytest.mark.parametrize("dim", [1, 2])
def test_tangent_residual_derivative_low_dims(dim):
    """The same pin in 1D and 2D (the dimensions of the reference's golden heat cases)."""
    m, hp, T, T_old, q, u, kw = _fd_case("mushy", 1, False, False, degree=1,
                                         ncell=(7, 5, 1), dim=dim)
    op = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, T=T, T_old=T_old, use_mask=0, **kw)
    Ju = op.vmult(u)
    eps = 2e-5
    fd = (_residual(m, hp, T + eps * u, T_old, q, kw) -
          _residual(m, hp, T - eps * u, T_old, q, kw)) / (2 * eps)
    assert np.abs(Ju - fd).max() / np.abs(Ju).max() < 1e-7


def test_inflow_outflow_rows_are_identity_rows():
    """do_pre_vmult / do_post_vmult: with inflow/outflow DoFs I, vmult applies A' with
    A'[i, j] = A[i, j] off I, zero columns j in I on the other rows, and unit rows on I."""
    m = orc.make_mesh(2, 2, (3, 2), hi=(1.0, 0.7))
    n = orc.n_dofs(m)
    rng = np.random.default_rng(3)
    vel = rng.uniform(-1, 1, 2 * n)
    io = (rng.uniform(size=n) < 0.3).astype(np.uint8)
    ap = orc.AdvecParams(dt=0.1, theta=0.5, diffusivity=0.02, mass_factor=1.0)
    plain = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL, vel=vel)
    iop = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL, vel=vel,
                             io_mask=io)
    E = np.eye(n)
    A = np.stack([plain.vmult(E[:, j]) for j in range(n)], axis=1)
    Ap = np.stack([iop.vmult(E[:, j]) for j in range(n)], axis=1)
    I = io == 1
    expect = A.copy()
    expect[:, I] = 0.0
    expect[I, :] = 0.0
    expect[I, I] = 1.0
    assert np.abs(Ap - expect).max() < 1e-14 * np.abs(A).max()
    # the no-BC index (use_mask = 0) ignores the inflow/outflow set, as the lifting needs
    iop0 = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL, vel=vel,
                              io_mask=io, use_mask=0)
    x = rng.uniform(-1, 1, n)
    assert np.array_equal(iop0.vmult(x), plain.vmult(x))


# ---------------------------------------------------------------- SUPG
def _supg_ops(m, w, D, coef, theta=0.5, **kw):
    """advection operators with constant nodal velocity w, diffusivity D or 0, SUPG on or off"""
    n = orc.n_dofs(m)
    vel = np.concatenate([np.full(n, wi) for wi in w])
    out = {}
    for supg in (0, 1):
        for d in (D, 0.0):
            ap = orc.AdvecParams(dt=0.1, theta=theta, diffusivity=d, mass_factor=1.0,
                                 supg=supg, supg_coefficient=coef)
            out[supg, d] = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL,
                                              vel=vel, **kw)
    return out


# u, its Laplacian integrated against w . grad v for v = x (= w_x int lap u over the unit
# cube), exact in the FE space of the given degree
LAP_CASES = [
    (2, 2, lambda x, y, z: x * x * y * y, 4.0 / 3.0),                     # 2y^2 + 2x^2
    (2, 3, lambda x, y, z: x ** 3 * y ** 3, 1.5),                        # 6xy^3 + 6x^3y
    (3, 2, lambda x, y, z: x * x * y * y * z * z, 2.0 / 3.0),            # sum 2 y^2 z^2 ...
    (3, 3, lambda x, y, z: x ** 3 * y * z + y ** 2, 0.75 + 2.0),         # 6xyz + 2
]


@pytest.mark.parametrize("dim,deg,f,lap_int", LAP_CASES)
def test_supg_laplacian_term_closed_form(dim, deg, f, lap_int):
    """The SUPG Laplacian terms alone: (A_supg(D) - A_supg(0) - A(D) + A(0)) u equals
    -tau theta D (w . grad v, lap u) in the tangent and create_rhs gives +tau (1 - theta) D
    (w . grad v, lap u_old); weighted with the nodal values of v = x it is the closed form
    w_x int lap u over the unit cube.  tau = coefficient * h on square cells."""
    nc = 3
    m = orc.make_mesh(dim, deg, (nc,) * dim)
    X = orc.node_coords(m)
    u = f(X[:, 0], X[:, 1], X[:, 2] if dim == 3 else 0 * X[:, 0])
    w = (0.6, -0.3, 0.2)[:dim]
    D, coef, theta = 0.05, 0.8, 0.5
    ops = _supg_ops(m, w, D, coef, theta)
    tau = coef / nc
    r = (ops[1, D].vmult(u) - ops[1, 0.0].vmult(u) - ops[0, D].vmult(u) + ops[0, 0.0].vmult(u))
    got = X[:, 0] @ r
    assert abs(got - (-tau * theta * D * w[0] * lap_int)) < 1e-12 * max(1.0, abs(got))
    rr = (ops[1, D].create_rhs(u) - ops[1, 0.0].create_rhs(u) - ops[0, D].create_rhs(u) +
          ops[0, 0.0].create_rhs(u))
    got = X[:, 0] @ rr
    assert abs(got - tau * (1 - theta) * D * w[0] * lap_int) < 1e-12 * max(1.0, abs(got))
    # v = 1: w . grad v = 0, no SUPG contribution at all
    assert abs(r.sum()) < 1e-13 and abs(rr.sum()) < 1e-13


def test_supg_streamline_term_closed_form():
    """The streamline term tau (w . grad v)(m u/dt + theta w . grad u) for D = 0 and u = x,
    v = x on the unit square: (A_supg - A) u weighted by v's nodal values is
    tau w_x (m/dt int x + theta w_x) = tau w_x (m/(2 dt) + theta w_x); with the coefficient < 0
    tau = h / |w|."""
    m = orc.make_mesh(2, 2, (4, 4))
    X = orc.node_coords(m)
    w = (0.6, -0.8)
    ops = _supg_ops(m, w, 0.0, -1.0, theta=0.5)
    tau = 0.25 / 1.0
    r = ops[1, 0.0].vmult(X[:, 0]) - ops[0, 0.0].vmult(X[:, 0])
    expect = tau * w[0] * (1.0 / (2 * 0.1) + 0.5 * w[0])
    assert abs(X[:, 0] @ r - expect) < 1e-13
    # below 1e-6 the velocity switches the stabilisation off
    ops = _supg_ops(m, (1e-7, 0.0), 0.0, -1.0)
    assert np.array_equal(ops[1, 0.0].vmult(X[:, 0]), ops[0, 0.0].vmult(X[:, 0]))

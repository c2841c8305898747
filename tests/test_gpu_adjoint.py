"""Adjoint of the implicit heat transient on the device: the transposed operator, the
conductivity-gradient kernel and the checkpointed forward/backward sweep, wrapped as a
torch.autograd.Function.  Checked against (i) the transpose identity <v, A u> = <A^T v, u>,
(ii) a reference gradient from dense matrices assembled with the CPU oracle's operator, and
(iii) central finite differences along seeded random directions."""
import numpy as np
import pytest

import orc

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def setup_case(nc=(4, 4, 4), degree=2, seed=5):
    import meltpooldg_amd as mf
    s = mf.ScratchData(degree, nc, stream=torch.cuda.current_stream())
    mask = s.boundary_mask()
    s.set_dirichlet_mask(mask)
    X = s.support_points()
    vel = np.zeros((3, s.n_dofs))
    vel[0, X[:, 2] > 0.8] = 0.1                       # melt velocity in the top 20 %
    rng = np.random.default_rng(seed)
    kcell = np.exp(0.3 * rng.standard_normal(s.n_cells))
    return s, mask, X, vel.ravel(), kcell


@pytest.mark.parametrize("kind", ["heat", "advec"])
def test_transpose_identity_and_dense(kind):
    import meltpooldg_amd as mf
    s, mask, X, vel, kcell = setup_case((3, 2, 3))
    if kind == "heat":
        op = mf.HeatDiffuseOperator(s, law=mf.LAW_TANH, rho_cp=1.3, k=1.0, tanh_amp=0.5,
                                    tanh_Tc=1700.0, tanh_width=50.0, t_dependent=True, dt=1e-2)
        T = 1700 + 80 * s.random_field(4)
        op.set_fields(T=_dev(T), velocity=_dev(vel * 30), k_cell=_dev(kcell))
    else:
        op = mf.AdvectionDiffusionOperator(s, theta=0.7, diffusivity=0.05, dt=0.05)
        op.set_fields(velocity=_dev(np.random.default_rng(2).uniform(-1, 1, 3 * s.n_dofs)))
    n = s.n_dofs
    u, v = _dev(s.random_field(1)), _dev(s.random_field(2))
    Au, ATv = torch.zeros_like(u), torch.zeros_like(u)
    op.vmult(Au, u)
    op.set_transpose(True)
    op.vmult(ATv, v)
    op.set_transpose(False)
    torch.cuda.synchronize()
    lhs, rhs = float(v @ Au), float(ATv @ u)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # and A^T equals the transpose of the assembled forward operator
    A = np.zeros((n, n))
    e = torch.zeros(n, dtype=torch.float64, device="cuda")
    col = torch.zeros_like(e)
    for j in range(n):
        e.zero_()
        e[j] = 1.0
        op.vmult(col, e)
        A[:, j] = col.cpu().numpy()
    assert np.max(np.abs(ATv.cpu().numpy() - A.T @ v.cpu().numpy())) < 1e-12 * np.abs(A).max()
    if kind == "heat":
        assert np.max(np.abs(A - A.T)) > 1e-6 * np.abs(A).max()   # really non-symmetric


def test_conductivity_gradient_kernel():
    """sum_q JxW grad(lam) . grad(T) per cell equals lam^T K_c T with K_c the stiffness of cell
    c alone (oracle: unit conductivity on cell c, zero elsewhere, no mass)."""
    import meltpooldg_amd as mf
    nc = (2, 3, 2)
    s = mf.ScratchData(3, nc, stream=torch.cuda.current_stream())
    op = mf.HeatDiffuseOperator(s, rho_cp=1.0, k=1.0, dt=1.0)
    lam, T = s.random_field(1), s.random_field(2)
    g = torch.zeros(s.n_cells, dtype=torch.float64, device="cuda")
    mf.conductivity_gradient(op, _dev(lam), _dev(T), g, scale=2.0)
    torch.cuda.synchronize()
    m = orc.make_mesh(3, 3, nc)
    expect = np.zeros(s.n_cells)
    for c in range(s.n_cells):
        ind = np.zeros(s.n_cells)
        ind[c] = 1.0
        hp = orc.HeatParams(law=0, t_dependent=0, rho_cp=0.0, k=1.0, dt=1.0)
        Kc = orc.OracleOperator(orc.OP_HEAT, m, heat=hp, k_cell=ind)
        expect[c] = 2.0 * lam @ Kc.vmult(T)
    assert np.max(np.abs(g.cpu().numpy() - expect)) < 1e-12 * np.abs(expect).max()


class _Transient:
    """Device transient J(k_cell, P, r) = 1/2 (T_N - T*)^T M (T_N - T*)."""

    def __init__(self, nc=(4, 4, 4), degree=2, steps=6, interval=2, dt=1e-3):
        import meltpooldg_amd as mf
        self.mf = mf
        self.s, self.mask, self.X, vel, self.k0 = setup_case(nc, degree)
        self.op = mf.HeatDiffuseOperator(self.s, rho_cp=1.0, k=1.0, dt=dt)
        self.vel = _dev(vel)
        self.steps, self.interval, self.dt = steps, interval, dt
        self.T0 = torch.zeros(self.s.n_dofs, dtype=torch.float64, device="cuda")
        self.target = _dev(np.where(self.mask == 1, 0.0, 0.02 * (1 + self.s.random_field(6))))

    def run(self, kcell, P, r, want_grad=False):
        mf = self.mf
        kc = _dev(kcell) if not torch.is_tensor(kcell) else kcell.contiguous()
        self.op.set_fields(T=self.T0, velocity=self.vel, k_cell=kc)
        tr = mf.HeatTransient(self.op, self.steps, self.dt, P, r, (0.3, 0.5, 1.0),
                              (0.5, 0.0, 0.0), checkpoint_interval=self.interval,
                              solver="GMRES", rel_tolerance=1e-13)
        TN = torch.zeros_like(self.T0)
        tr.forward(self.T0, TN)
        e = TN - self.target
        Me = torch.zeros_like(e)
        mf.mass_vmult(self.s, Me, e)
        J = 0.5 * float(e @ Me)
        if not want_grad:
            return J
        gk = torch.zeros(self.s.n_cells, dtype=torch.float64, device="cuda")
        gP, gR = tr.backward(Me, gk)
        return J, gk, gP, gR


def test_transient_gradient_matches_dense_reference():
    case = _Transient(nc=(3, 3, 3), steps=4, interval=3)
    P, r = 1.0, 0.3
    J, gk, gP, gR = case.run(case.k0, P, r, want_grad=True)
    # dense reference from the oracle's assembled operators
    m = orc.make_mesh(3, 2, (3, 3, 3))
    n = orc.n_dofs(m)
    free = case.mask == 0
    vel = case.vel.cpu().numpy()

    def assemble(op):
        A = np.zeros((n, n))
        for j in np.nonzero(free)[0]:
            e = np.zeros(n)
            e[j] = 1.0
            A[:, j] = op.vmult(e)
        return A[np.ix_(free, free)]

    hp = orc.HeatParams(law=0, t_dependent=0, rho_cp=1.0, k=1.0, dt=case.dt)
    A = assemble(orc.OracleOperator(orc.OP_HEAT, m, heat=hp, vel_kind=orc.VEL_NODAL, vel=vel,
                                    k_cell=case.k0, mask=case.mask))
    hm = orc.HeatParams(law=0, t_dependent=0, rho_cp=1.0, k=0.0, dt=1.0)
    M = assemble(orc.OracleOperator(orc.OP_HEAT, m, heat=hm, mask=case.mask))
    # the residual reads the source plain (read_dof_values_plain): free rows, all columns
    Mq = np.zeros((n, n))
    mop = orc.OracleOperator(orc.OP_HEAT, m, heat=hm)
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        Mq[:, j] = mop.vmult(e)
    Mq = Mq[free, :]
    X = orc.node_coords(m)

    def laser(n_step, d_dr=False):
        pos = np.array([0.3 + 0.5 * n_step * case.dt, 0.5, 1.0])
        d2 = ((X - pos) ** 2).sum(axis=1)
        q = P / (r * np.sqrt(np.pi / 2)) ** 3 * np.exp(-2 * d2 / r ** 2)
        return q * (-3 / r + 4 * d2 / r ** 3) if d_dr else q

    Ts = [np.zeros(free.sum())]
    for k in range(1, case.steps + 1):
        Ts.append(np.linalg.solve(A, M @ Ts[-1] / case.dt + Mq @ laser(k)))
    e = Ts[-1] - case.target.cpu().numpy()[free]
    J_ref = 0.5 * e @ M @ e
    lam = np.linalg.solve(A.T, M @ e)
    gP_ref, gR_ref = 0.0, 0.0
    gk_ref = np.zeros(len(case.k0))
    Kc = []
    for c in range(len(case.k0)):
        ind = np.zeros(len(case.k0))
        ind[c] = 1.0
        hk = orc.HeatParams(law=0, t_dependent=0, rho_cp=0.0, k=1.0, dt=1.0)
        Kc.append(assemble(orc.OracleOperator(orc.OP_HEAT, m, heat=hk, k_cell=ind,
                                              mask=case.mask)))
    for k in range(case.steps, 0, -1):
        if k < case.steps:
            lam = np.linalg.solve(A.T, M @ lam / case.dt)
        gP_ref += lam @ Mq @ laser(k) / P
        gR_ref += lam @ Mq @ laser(k, d_dr=True)
        for c in range(len(case.k0)):
            gk_ref[c] -= lam @ Kc[c] @ Ts[k]
    assert abs(J / J_ref - 1) < 1e-10
    assert abs(gP / gP_ref - 1) < 1e-9
    assert abs(gR / gR_ref - 1) < 1e-9
    assert np.max(np.abs(gk.cpu().numpy() - gk_ref)) < 1e-9 * np.abs(gk_ref).max()


class HeatTransientFunction(torch.autograd.Function):
    """torch.autograd wrapper of the device transient: inputs (k_cell, P, r), output J."""

    @staticmethod
    def forward(ctx, kcell, P, r, case):
        J, gk, gP, gR = case.run(kcell.detach(), float(P), float(r), want_grad=True)
        ctx.save_for_backward(gk, torch.tensor(gP, dtype=torch.float64),
                              torch.tensor(gR, dtype=torch.float64))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        gk, gP, gR = ctx.saved_tensors
        return gJ * gk, gJ * gP, gJ * gR, None


def test_autograd_gradient_vs_central_fd():
    """Central finite differences (h = 1e-6) along seeded random directions in (k_cell, P, r)."""
    case = _Transient(nc=(4, 4, 4), steps=6, interval=2)
    kc = torch.tensor(case.k0, dtype=torch.float64, device="cuda", requires_grad=True)
    P = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    r = torch.tensor(0.25, dtype=torch.float64, requires_grad=True)
    J = HeatTransientFunction.apply(kc, P, r, case)
    J.backward()
    g = np.concatenate([kc.grad.cpu().numpy(), [P.grad.item(), r.grad.item()]])
    rng = np.random.default_rng(7)
    h = 1e-6
    theta = np.concatenate([case.k0, [1.0, 0.25]])
    for _ in range(5):
        d = rng.standard_normal(len(theta))
        d /= np.linalg.norm(d)
        tp, tm = theta + h * d, theta - h * d
        Jp = case.run(tp[:-2], tp[-2], tp[-1])
        Jm = case.run(tm[:-2], tm[-2], tm[-1])
        fd = (Jp - Jm) / (2 * h)
        ad = g @ d
        assert abs(fd - ad) <= 1e-5 * max(abs(ad), 1e-3 * np.abs(g).max()), (fd, ad)


def test_moving_laser_transient_full_size():
    """The full moving-laser case: FE_Q2 on 48^3 cells (912,673 DoFs), per-cell k = exp(g) with g
    ~ N(0, 0.3^2) (seed 5), melt velocity (0.1,0,0) in the top 20 %, Gaussian laser P=1, r=0.05
    moving along x at 0.5 for 20 implicit-Euler steps of dt=1e-3, checkpoints every 5 steps;
    J = 1/2 ||T_20 - T*||^2_L2 with T* the forward solution for k (1 + 0.1 U) (seed 6).
    Adjoint gradient (k_cell, P, r) against central differences (h = 1e-6) along 20 seeded
    directions (seed 7)."""
    import meltpooldg_amd as mf
    nc, steps, interval, dt = 48, 20, 5, 1e-3
    s = mf.ScratchData(2, nc, stream=torch.cuda.current_stream())
    mask = s.boundary_mask()
    s.set_dirichlet_mask(mask)
    X = s.support_points()
    vel = np.zeros((3, s.n_dofs))
    vel[0, X[:, 2] > 0.8] = 0.1
    veld = _dev(vel.ravel())
    k0 = np.exp(0.3 * np.random.default_rng(5).standard_normal(s.n_cells))
    op = mf.HeatDiffuseOperator(s, rho_cp=1.0, k=1.0, dt=dt)
    T0 = torch.zeros(s.n_dofs, dtype=torch.float64, device="cuda")
    start, scan = (0.25, 0.5, 1.0), (0.5, 0.0, 0.0)

    def forward(kcell, P, r):
        op.set_fields(T=T0, velocity=veld, k_cell=_dev(kcell))
        tr = mf.HeatTransient(op, steps, dt, P, r, start, scan, checkpoint_interval=interval,
                              solver="GMRES", rel_tolerance=1e-13)
        TN = torch.zeros_like(T0)
        tr.forward(T0, TN)
        return tr, TN

    _, target = forward(k0 * (1 + 0.1 * np.random.default_rng(6).uniform(-1, 1, s.n_cells)),
                        1.0, 0.05)
    target = target.clone()

    def objective(kcell, P, r, grad=False):
        tr, TN = forward(kcell, P, r)
        e = TN - target
        Me = torch.zeros_like(e)
        mf.mass_vmult(s, Me, e)
        J = 0.5 * float(e @ Me)
        if not grad:
            return J
        gk = torch.zeros(s.n_cells, dtype=torch.float64, device="cuda")
        gP, gR = tr.backward(Me, gk)
        return J, np.concatenate([gk.cpu().numpy(), [gP, gR]])

    J, g = objective(k0, 1.0, 0.05, grad=True)
    assert J > 0 and np.all(np.isfinite(g))
    theta = np.concatenate([k0, [1.0, 0.05]])
    rng = np.random.default_rng(7)
    h = 1e-6
    worst = 0.0
    for _ in range(20):
        d = rng.standard_normal(len(theta))
        d /= np.linalg.norm(d)
        tp, tm = theta + h * d, theta - h * d
        fd = (objective(tp[:-2], tp[-2], tp[-1]) - objective(tm[:-2], tm[-2], tm[-1])) / (2 * h)
        ad = g @ d
        worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-3 * np.abs(g).max()))
    assert worst < 1e-5, worst

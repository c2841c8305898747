"""The CPU oracle against the reference's own golden outputs (whole-path known answers).

Fixtures are the printed FE L2 norms of the reference's application tests (also the
unidirectional_heat_transfer_with_* Robin / velocity / mushy-zone cases and the 1D
solidification_slab case, see their tests below):
  applications/mp-heat-transfer/cases/unidirectional_heat_transfer/tests/
      unidirectional_heat_transfer.mpirun=4.output:3-19 (6 digits)
  applications/mp-advec-diff/cases/tests/advection_diffusion_vortex_mf.mpirun=4.output (10 digits)
  applications/mp-advec-diff/cases/tests/advection_diffusion_vortex_mf_3d.mpirun=4.output
The reference stores per-cell errors in a Vector<float> and accumulates them in single precision
per MPI rank (source/utilities/vector_tools.cpp:137-168), so its 10-digit advection norms carry
~1e-8 relative noise; the advection bound below is 5e-8 (2D: 2e-8)."""
import json
import os

import numpy as np
import pytest

import orc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_norms.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def vortex_setup(dim, refinements):
    n = 2 ** refinements
    lo, hi = (-1.0,) * dim, (1.0,) * dim
    m = orc.make_mesh(dim, 1, n, lo, hi)
    xyz = orc.node_coords(m)
    center = np.array([0.0, 0.5, 0.0]) if dim == 2 else np.array([0.0, 0.0, 0.5])
    h = 2.0 / n
    eps = 0.5 * (h * np.sqrt(dim)) / np.sqrt(dim)   # 0.5 * min cell diameter / sqrt(dim)
    sd = -(np.linalg.norm(xyz - center, axis=1) - 0.25)
    phi0 = np.tanh(sd / (2 * eps))
    x, y = xyz[:, 0], xyz[:, dim - 1]
    tol = 1e-12
    # inflow faces (boundary id 42) of advection_diffusion_vortex.hpp set_boundary_conditions
    mask = (((np.abs(x + 1) < tol) & (y >= 0)) | ((np.abs(x - 1) < tol) & (y <= 0)) |
            ((np.abs(y - 1) < tol) & (x >= 0)) | ((np.abs(y + 1) < tol) & (x <= 0)))
    mask = mask.astype(np.uint8)
    phi0[mask == 1] = -1.0   # constraints distributed into the initial field
    return m, phi0, mask


def test_heat_unidirectional_2d_q1():
    g = golden()["unidirectional_heat_transfer"]
    m = orc.make_mesh(2, 1, 32, lo=(0.0, 0.0), hi=(0.1, 0.1))
    xyz = orc.node_coords(m)
    T0 = 0.0 + (0.1 - 0.0) * (xyz[:, 0] - 0.0) / 0.1   # LinearTemp(0, 0.1)
    hp = orc.HeatParams(law=orc.LAW_CONSTANT, t_dependent=0, rho_cp=1.0, k=1.0, dt=1e-4)
    op = orc.OracleOperator(orc.OP_HEAT, m, heat=hp)
    _, norms, status = op.heat_implicit_euler(T0, 5)
    assert status == 0
    assert [f"{v:.5e}" for v in norms] == g["norms"]


@pytest.mark.parametrize("name,dim,ref,tol", [
    ("advection_diffusion_vortex_mf", 2, 5, 2e-8),
    ("advection_diffusion_vortex_mf_3d", 3, 3, 5e-8)])
def test_advection_vortex(name, dim, ref, tol):
    g = golden()[name]
    m, phi0, mask = vortex_setup(dim, ref)
    ap = orc.AdvecParams(dt=0.01, theta=0.5, diffusivity=0.0, mass_factor=1.0, vel_scale=4.0)
    op = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_VORTEX, mask=mask)
    _, norms, status = op.advec_theta_steps(phi0, 5, -np.ones(len(phi0)))
    assert status == 0
    rel = np.abs(norms / np.array(g["norms"], dtype=float) - 1.0)
    assert rel.max() < tol, rel


def test_advection_vortex_supg():
    """advection_diffusion_vortex_mf_supg: the 2D vortex case with SUPG (coefficient -1:
    tau = h / max_q |w| per cell); without the stabilisation the norms differ by 6e-4."""
    g = golden()["advection_diffusion_vortex_mf_supg"]
    m, phi0, mask = vortex_setup(2, 5)
    ap = orc.AdvecParams(dt=0.01, theta=0.5, diffusivity=0.0, mass_factor=1.0, vel_scale=4.0,
                         supg=1, supg_coefficient=-1.0)
    op = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_VORTEX, mask=mask)
    _, norms, status = op.advec_theta_steps(phi0, 5, -np.ones(len(phi0)))
    assert status == 0
    rel = np.abs(norms / np.array(g["norms"], dtype=float) - 1.0)
    assert rel.max() < 2e-8, rel


# ---------------------------------------------------------------- T-dependent law, Robin faces
def mushy(**kw):
    """HeatParams of the liquid_solid material with the mushy_zone transition."""
    return orc.HeatParams(law=orc.LAW_MUSHY, t_dependent=1, **kw)


def slab_case():
    """solidification_slab (applications/mp-heat-transfer/cases/solidification_slab): 1D Q1,
    [0, 1] with 2^7 cells, ice/water (tests/solidification_slab.json), T0 = 283 with the left
    Dirichlet value 253 distributed into the initial field (HeatDiffuseOperation::
    set_initial_condition ends with constraints.distribute, heat_diffuse_operation.cpp:187),
    dt = 200, Newton residual tolerance 1e-8."""
    hp = mushy(liquid_k=0.556, liquid_rho=1000.0, liquid_cp=4226.0, solid_k=2.22,
               solid_rho=1000.0, solid_cp=1762.0, solidus=270.0, liquidus=276.0)
    return hp, 283.0, 253.0, 200.0, (10, 1e-8, 1e-10)


def test_solidification_slab_1d_q1():
    g = golden()["solidification_slab"]
    hp, T_init, T_hat, dt, newton = slab_case()
    m = orc.make_mesh(1, 1, (128, 1, 1))
    X = orc.node_coords(m)
    mask = (np.abs(X[:, 0]) < 1e-12).astype(np.uint8)
    T0 = np.full(len(X), T_init)
    T0[mask == 1] = T_hat
    out = orc.heat_time_steps(m, hp, mask, T0, [dt] * 5, dirichlet_values=T0, newton=newton)
    assert [f"{orc.fe_l2_norm(m, T):.5e}" for T in out] == g["norms"]


UNIDIRECTIONAL = {
    # name: (convection, radiation, velocity, solidification, dt, Newton residual tolerance)
    "unidirectional_heat_transfer_with_convection": (True, False, 0.0, False, 0.01, 5e-9),
    "unidirectional_heat_transfer_with_radiation": (False, True, 0.0, False, 0.01, 5e-9),
    "unidirectional_heat_transfer_with_radiation_and_convection": (True, True, 0.0, False, 0.01,
                                                                   5e-9),
    "unidirectional_heat_transfer_with_velocity": (True, True, 0.01, False, 0.001, 1e-6),
    "unidirectional_heat_transfer_with_velocity_and_solidification": (False, False, 0.005, True,
                                                                      0.01, 1e-7),
}


def unidirectional_case(name, dim=2):
    """unidirectional_heat_transfer_with_* (applications/mp-heat-transfer/cases/
    unidirectional_heat_transfer): Q1 on [0, 0.1]^2 with 2^5 cells per direction, Robin terms
    on the right side (side 1 of the colourised box), T_inf = 300, alpha = 25, emissivity 0.98;
    T0 = 1000 with Robin terms, LinearTemp(1960, 1980) with solidification; velocity
    (-v, 0) (UnidirectionalVelocityField); steel gas (k 55.563, cp 460, rho 7850) or the
    liquid/solid steel of the solidification case.  Returns the heat parameters, boundary, T0
    as a function of x, the velocity, dt and the Newton tolerance."""
    conv, rad, v, solid, dt, tol = UNIDIRECTIONAL[name]
    b = orc.HeatBoundary()
    b.convection_temperature_infinity = 300.0
    b.radiation_temperature_infinity = 300.0
    if conv:
        b.convection_sides = 1 << 1
        b.convection_coefficient = 25.0
    if rad:
        b.radiation_sides = 1 << 1
        b.emissivity = 0.98
    if solid:
        hp = mushy(liquid_k=55.563, liquid_rho=7850.0, liquid_cp=460.0, solid_k=17.0,
                   solid_rho=7850.0, solid_cp=700.0, solidus=1966.0, liquidus=1974.0)
        T0 = lambda x: 1960.0 + (1980.0 - 1960.0) * x / 0.1  # noqa: E731
    else:
        hp = orc.HeatParams(law=orc.LAW_CONSTANT, t_dependent=0, rho_cp=7850.0 * 460.0,
                            k=55.563)
        T0 = lambda x: np.full_like(x, 1000.0)  # noqa: E731
    return hp, b, T0, v, dt, tol


@pytest.mark.parametrize("name", list(UNIDIRECTIONAL))
def test_unidirectional_robin_velocity_mushy_2d_q1(name):
    g = golden()[name]
    hp, b, T0f, v, dt, tol = unidirectional_case(name)
    m = orc.make_mesh(2, 1, (32, 32, 1), hi=(0.1, 0.1, 1.0))
    X = orc.node_coords(m)
    n = len(X)
    vel = None
    if v:
        vel = np.zeros(2 * n)
        vel[:n] = -v
    out = orc.heat_time_steps(m, hp, np.zeros(n, np.uint8), T0f(X[:, 0]), [dt] * 5, boundary=b,
                              vel=vel, newton=(10, tol, 1e-10))
    assert [f"{orc.fe_l2_norm(m, T):.5e}" for T in out] == g["norms"]


def vortex_inflow_mask(m, dim, scale=4.0):
    """create_inflow_outflow_constraints (advection_diffusion_operation.cpp:368-499) for the
    vortex case with both boundary ids inflow/outflow (advection_diffusion_vortex.hpp
    set_boundary_conditions): boundary support points with n . u <= 0, u = s (y, -x) (2D) or
    s (z, 0, -x) (3D); in 3D the y faces (n . u = 0) are all inflow."""
    xyz = orc.node_coords(m)
    u = np.zeros((len(xyz), 3))
    u[:, 0] = scale * xyz[:, dim - 1]
    u[:, dim - 1] = -scale * xyz[:, 0]
    io = np.zeros(len(xyz), bool)
    for d in range(dim):
        for sgn, val in ((-1.0, -1.0), (1.0, 1.0)):
            io |= (np.abs(xyz[:, d] - val) < 1e-12) & (sgn * u[:, d] <= 0.0)
    return io.astype(np.uint8)


@pytest.mark.parametrize("name,dim,ref,tol", [
    ("advection_diffusion_vortex_mf_inflow_outflow", 2, 5, 2e-8),
    ("advection_diffusion_vortex_mf_inflow_outflow_3d", 3, 3, 5e-8)])
def test_advection_vortex_inflow_outflow(name, dim, ref, tol):
    """Inflow/outflow boundaries (pre/post-vmult identity rows, value -1 lifted): the initial
    field is not constrained (no Dirichlet rows), the inflow set comes from the velocity."""
    g = golden()[name]
    n = 2 ** ref
    m = orc.make_mesh(dim, 1, n, (-1.0,) * dim, (1.0,) * dim)
    xyz = orc.node_coords(m)
    center = np.array([0.0, 0.5, 0.0]) if dim == 2 else np.array([0.0, 0.0, 0.5])
    eps = 0.5 * (2.0 / n)
    phi0 = np.tanh(-(np.linalg.norm(xyz - center, axis=1) - 0.25) / (2 * eps))
    ap = orc.AdvecParams(dt=0.01, theta=0.5, diffusivity=0.0, mass_factor=1.0, vel_scale=4.0)
    op = orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_VORTEX,
                            io_mask=vortex_inflow_mask(m, dim))
    _, norms, status = op.advec_theta_steps(phi0, 5, -np.ones(len(phi0)))
    assert status == 0
    rel = np.abs(norms / np.array(g["norms"], dtype=float) - 1.0)
    assert rel.max() < tol, rel

"""The CPU oracle against the reference's own golden outputs (whole-path known answers).

Fixtures are the printed FE L2 norms of the reference's application tests:
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

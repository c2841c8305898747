"""Shared builders for parity tests: the same seeded case as an oracle operator (CPU, C) and as
a device operator (libmeltpooldg_amd through the C ABI)."""
from __future__ import annotations

import dataclasses

import numpy as np

import orc


@dataclasses.dataclass
class Case:
    degree: int = 2
    ncell: tuple = (3, 4, 5)
    lo: tuple = (0.0, 0.0, 0.0)
    hi: tuple = (1.0, 1.2, 0.9)
    kind: str = "heat"          # heat | advec
    law: str = "const"          # const | tanh | mushy
    t_dependent: bool = False
    vel: bool = False
    k_cell: bool = False
    mask: str = "boundary"      # none | boundary | random
    dt: float = 1e-3
    rho_cp: float = 1.3
    k: float = 0.7
    theta: float = 0.5
    diffusivity: float = 0.05
    mass_factor: float = 1.0
    supg: bool = False
    supg_coefficient: float = -1.0
    seed: int = 11

    def mesh(self):
        return orc.make_mesh(3, self.degree, self.ncell, self.lo, self.hi)


# mushy-zone law of the parity cases: properties of order one so that the mass, stiffness and
# T-derivative terms all show at the 1e-12 bar; the fields span solid, mushy and liquid nodes
MUSHY = dict(liquid_k=0.5, liquid_rho=1.0, liquid_cp=1.5, solid_k=2.0, solid_rho=1.2,
             solid_cp=0.8, solidus=270.0, liquidus=276.0)


def fields(case: Case):
    m = case.mesh()
    n = orc.n_dofs(m)
    rng = np.random.default_rng(case.seed)
    out = {}
    if case.law == "tanh":
        out["T"] = 1700.0 + 50.0 * 3.0 * orc.random_field(m, case.seed + 1)
    elif case.law == "mushy":
        out["T"] = 273.0 + 4.0 * orc.random_field(m, case.seed + 1)
    else:
        out["T"] = 300.0 + orc.random_field(m, case.seed + 1)
    out["T_old"] = out["T"] - (2.0 if case.law == "mushy" else 5.0) * \
        orc.random_field(m, case.seed + 2)
    out["vel"] = rng.uniform(-1, 1, size=3 * n) if (case.vel or case.kind == "advec") else None
    out["k_cell"] = np.exp(0.3 * rng.standard_normal(orc.n_cells(m))) if case.k_cell else None
    if case.mask == "none":
        out["mask"] = np.zeros(n, dtype=np.uint8)
    elif case.mask == "boundary":
        out["mask"] = orc.boundary_mask(m)
    else:
        out["mask"] = (rng.uniform(size=n) < 0.2).astype(np.uint8)
    out["q_s"] = orc.random_field(m, case.seed + 3)
    return out


def oracle_boundary(spec):
    """orc.HeatBoundary from a dict of HeatDiffuseOperator.set_boundary keyword arguments."""
    b = orc.HeatBoundary()
    b.convection_sides = sum(1 << s_ for s_ in spec.get("convection_sides", ()))
    b.radiation_sides = sum(1 << s_ for s_ in spec.get("radiation_sides", ()))
    for side, g in spec.get("neumann_sides", {}).items():
        b.neumann_sides |= 1 << side
        b.neumann_flux[side] = g
    for key in ("convection_coefficient", "convection_temperature_infinity", "emissivity",
                "radiation_temperature_infinity"):
        setattr(b, key, spec.get(key, 0.0))
    return b


def oracle_operator(case: Case, f, use_mask=1, boundary=None):
    m = case.mesh()
    if case.kind == "heat":
        law = {"const": orc.LAW_CONSTANT, "tanh": orc.LAW_TANH, "mushy": orc.LAW_MUSHY}[case.law]
        hp = orc.HeatParams(law=law, t_dependent=int(case.t_dependent), rho_cp=case.rho_cp,
                            k=case.k, tanh_amp=0.5, tanh_Tc=1700.0, tanh_width=50.0, dt=case.dt,
                            **(MUSHY if case.law == "mushy" else {}))
        return orc.OracleOperator(orc.OP_HEAT, m, heat=hp,
                                  vel_kind=orc.VEL_NODAL if case.vel else orc.VEL_NONE,
                                  mask=f["mask"], T=f["T"], T_old=f["T_old"], vel=f["vel"],
                                  k_cell=f["k_cell"], use_mask=use_mask,
                                  boundary=oracle_boundary(boundary) if boundary else None)
    ap = orc.AdvecParams(dt=case.dt, theta=case.theta, diffusivity=case.diffusivity,
                         mass_factor=case.mass_factor, vel_scale=0.0, supg=int(case.supg),
                         supg_coefficient=case.supg_coefficient)
    return orc.OracleOperator(orc.OP_ADVEC, m, advec=ap, vel_kind=orc.VEL_NODAL, mask=f["mask"],
                              vel=f["vel"], use_mask=use_mask)


def device_operator(case: Case, f, dev="cuda"):
    import torch

    import meltpooldg_amd as mf
    s = mf.ScratchData(case.degree, case.ncell, case.lo, case.hi,
                       stream=torch.cuda.current_stream())
    s.set_dirichlet_mask(f["mask"])
    if case.kind == "heat":
        law = {"const": mf.LAW_CONSTANT, "tanh": mf.LAW_TANH, "mushy": mf.LAW_MUSHY}[case.law]
        op = mf.HeatDiffuseOperator(s, law=law, rho_cp=case.rho_cp, k=case.k, tanh_amp=0.5,
                                    tanh_Tc=1700.0, tanh_width=50.0, t_dependent=case.t_dependent,
                                    dt=case.dt, mushy=MUSHY if case.law == "mushy" else None)
    else:
        op = mf.AdvectionDiffusionOperator(s, theta=case.theta, diffusivity=case.diffusivity,
                                           mass_factor=case.mass_factor, dt=case.dt,
                                           supg=case.supg, supg_coefficient=case.supg_coefficient)

    def t(a):
        return None if a is None else torch.tensor(a, dtype=torch.float64, device=dev)

    vel = f["vel"] if (case.vel or case.kind == "advec") else None
    op.set_fields(T=t(f["T"]) if case.kind == "heat" else None, velocity=t(vel),
                  k_cell=t(f["k_cell"]))
    if case.law == "mushy":
        op.set_temperature_old(t(f["T_old"]))
    return s, op


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (scale if scale > 0 else 1.0))

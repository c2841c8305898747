"""ctypes binding of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of bench.py, always as the checker or the timed CPU baseline.  The product
package ``meltpooldg_amd`` never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")

LAW_CONSTANT, LAW_TANH, LAW_MUSHY = 0, 1, 2
VEL_NONE, VEL_NODAL, VEL_VORTEX = 0, 1, 2
OP_HEAT, OP_ADVEC, OP_NORMAL, OP_CURVATURE, OP_REINIT = 0, 1, 2, 3, 4
THICKNESS_PROPORTIONAL, THICKNESS_ABSOLUTE, THICKNESS_CELLS = 0, 1, 2
CG, GMRES = 0, 1
SUCCESS, NO_CONVERGENCE = 0, 1


class Mesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("degree", C.c_int), ("ncell", C.c_int * 3),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3)]


class HeatParams(C.Structure):
    _fields_ = [("law", C.c_int), ("t_dependent", C.c_int), ("rho_cp", C.c_double),
                ("k", C.c_double), ("tanh_amp", C.c_double), ("tanh_Tc", C.c_double),
                ("tanh_width", C.c_double), ("dt", C.c_double),
                ("liquid_k", C.c_double), ("liquid_rho", C.c_double), ("liquid_cp", C.c_double),
                ("solid_k", C.c_double), ("solid_rho", C.c_double), ("solid_cp", C.c_double),
                ("solidus", C.c_double), ("liquidus", C.c_double)]


def material(hp, T):
    """(rho cp, k, d(rho cp)/dT, dk/dT) of the oracle's law at temperature T."""
    out = np.zeros(4)
    lib().orc_material(C.byref(hp), C.c_double(T), _p(out))
    return out


class AdvecParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("theta", C.c_double), ("diffusivity", C.c_double),
                ("mass_factor", C.c_double), ("vel_scale", C.c_double),
                ("supg", C.c_int), ("supg_coefficient", C.c_double)]


class AMeshC(C.Structure):
    _fields_ = [("degree", C.c_int), ("ncell", C.c_int * 3), ("lo", C.c_double * 3),
                ("hi", C.c_double * 3), ("n_cells", C.c_long), ("cells", C.c_void_p)]


class ADataC(C.Structure):
    _fields_ = [("n_nodes", C.c_long), ("cell_nodes", C.POINTER(C.c_long)),
                ("xyz", C.POINTER(C.c_double)), ("boundary", C.POINTER(C.c_uint8)),
                ("line_ptr", C.POINTER(C.c_long)), ("line_col", C.POINTER(C.c_long)),
                ("line_val", C.POINTER(C.c_double))]


class HeatBoundary(C.Structure):
    """Robin / Neumann sides of the heat operator (orc_heat_boundary): side s = 2*direction +
    (0 lower, 1 upper)."""
    _fields_ = [("convection_sides", C.c_uint), ("radiation_sides", C.c_uint),
                ("neumann_sides", C.c_uint), ("convection_coefficient", C.c_double),
                ("convection_temperature_infinity", C.c_double), ("emissivity", C.c_double),
                ("radiation_temperature_infinity", C.c_double), ("neumann_flux", C.c_double * 6)]


STEFAN_BOLTZMANN = 5.670374419e-8


class LsParams(C.Structure):
    """Normal vector / curvature projection parameters (orc_ls_params)."""
    _fields_ = [("filter_parameter", C.c_double), ("narrow_band", C.c_int),
                ("narrow_band_threshold", C.c_double), ("compute_normalized_vector", C.c_int)]


class ReinitParams(C.Structure):
    """Olsson reinitialization parameters (orc_reinit_params)."""
    _fields_ = [("dt", C.c_double), ("thickness_type", C.c_int),
                ("thickness_value", C.c_double), ("tangential_diffusion_factor", C.c_double),
                ("normal_tolerance", C.c_double)]


class Operator(C.Structure):
    _fields_ = [("kind", C.c_int), ("mesh", Mesh), ("heat", HeatParams), ("advec", AdvecParams),
                ("vel_kind", C.c_int), ("mask", C.c_void_p), ("use_mask", C.c_int),
                ("T", C.c_void_p), ("T_old", C.c_void_p), ("vel", C.c_void_p),
                ("k_cell", C.c_void_p), ("nthreads", C.c_int),
                ("amesh", C.POINTER(AMeshC)), ("adata", C.POINTER(ADataC)),
                ("bnd", C.POINTER(HeatBoundary)), ("ls", LsParams),
                ("level_set", C.c_void_p), ("reinit", ReinitParams), ("io_mask", C.c_void_p)]


def build(force: bool = False) -> str:
    """Compile liborc.so with the committed Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_n_dofs.restype = C.c_long
        L.orc_n_cells.restype = C.c_long
        L.orc_fe_l2_norm.restype = C.c_double
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_random_field.argtypes = [C.POINTER(Mesh), C.c_uint64, dp]
        L.orc_solve.argtypes = [C.POINTER(Operator), C.c_int, C.c_int, C.c_int, C.c_double,
                                C.c_double, C.c_void_p, dp, dp, C.POINTER(C.c_int),
                                C.POINTER(C.c_double)]
        L.orc_local_matrix.argtypes = [C.POINTER(Operator), C.c_long, dp]
        L.orc_material.argtypes = [C.POINTER(HeatParams), C.c_double, dp]
        L.orc_amesh_build.argtypes = [C.POINTER(AMeshC), C.POINTER(ADataC)]
        L.orc_adata_free.argtypes = [C.POINTER(ADataC)]
        L.orc_distribute_constraints.argtypes = [C.POINTER(ADataC), dp]
        L.orc_ls_normal_rhs.argtypes = [C.POINTER(Operator), dp, dp]
        L.orc_ls_curvature_rhs.argtypes = [C.POINTER(Operator), dp, C.c_double, dp]
        L.orc_numerical_zero_of_norm.argtypes = [C.c_int, C.c_double]
        L.orc_numerical_zero_of_norm.restype = C.c_double
        L.orc_reinit_rhs.argtypes = [C.POINTER(Operator), dp, dp]
        L.orc_n_rows.argtypes = [C.POINTER(Operator)]
        L.orc_n_rows.restype = C.c_long
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_mesh(dim, degree, ncell, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)) -> Mesh:
    if isinstance(ncell, int):
        ncell = (ncell,) * dim
    nc = list(ncell) + [1] * (3 - len(ncell))
    lo = list(lo) + [0.0] * (3 - len(lo))
    hi = list(hi) + [1.0] * (3 - len(hi))
    return Mesh(dim, degree, (C.c_int * 3)(*nc), (C.c_double * 3)(*lo), (C.c_double * 3)(*hi))


def n_dofs(mesh: Mesh) -> int:
    return int(lib().orc_n_dofs(C.byref(mesh)))


def n_cells(mesh: Mesh) -> int:
    return int(lib().orc_n_cells(C.byref(mesh)))


def gauss(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_gauss(n, _p(x), _p(w))
    return x, w


def gauss_lobatto(n):
    x = np.zeros(n)
    lib().orc_gauss_lobatto(n, _p(x))
    return x


def shape_matrices(degree):
    K = degree + 1
    S, DS, Dq, w = np.zeros((K, K)), np.zeros((K, K)), np.zeros((K, K)), np.zeros(K)
    lib().orc_shape_matrices(degree, _p(S), _p(DS), _p(Dq), _p(w))
    return S, DS, Dq, w


def node_coords(mesh):
    xyz = np.zeros((n_dofs(mesh), 3))
    lib().orc_node_coords(C.byref(mesh), _p(xyz))
    return xyz


def boundary_mask(mesh):
    m = np.zeros(n_dofs(mesh), dtype=np.uint8)
    lib().orc_boundary_mask(C.byref(mesh), m.ctypes.data_as(C.c_void_p))
    return m


def random_field(mesh, seed):
    out = np.zeros(n_dofs(mesh))
    lib().orc_random_field(C.byref(mesh), C.c_uint64(seed), _p(out))
    return out


class AdaptiveMesh:
    """An adaptive octree mesh (cells = int array [n_cells, 4]: level, ix, iy, iz) with the
    oracle's own node identification and hanging-node constraints."""

    def __init__(self, degree, ncell, cells, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)):
        if isinstance(ncell, int):
            ncell = (ncell,) * 3
        self.degree = degree
        self.K = degree + 1
        self.cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
        self.n_cells = len(self.cells)
        self.c = AMeshC(degree, (C.c_int * 3)(*ncell), (C.c_double * 3)(*lo),
                        (C.c_double * 3)(*hi), self.n_cells, self.cells.ctypes.data)
        self.d = ADataC()
        if lib().orc_amesh_build(C.byref(self.c), C.byref(self.d)) != 0:
            raise ValueError("orc_amesh_build failed")
        n = self.n_nodes = int(self.d.n_nodes)
        nloc = self.K ** 3
        self.cell_nodes = np.ctypeslib.as_array(self.d.cell_nodes, (self.n_cells * nloc,)).reshape(
            self.n_cells, nloc).copy()
        self.xyz = np.ctypeslib.as_array(self.d.xyz, (n * 3,)).reshape(n, 3).copy()
        self.boundary = np.ctypeslib.as_array(self.d.boundary, (n,)).copy()
        self.line_ptr = np.ctypeslib.as_array(self.d.line_ptr, (n + 1,)).copy()
        nnz = int(self.line_ptr[-1])
        self.line_col = (np.ctypeslib.as_array(self.d.line_col, (nnz,)).copy() if nnz
                         else np.zeros(0, dtype=np.int64))
        self.line_val = (np.ctypeslib.as_array(self.d.line_val, (nnz,)).copy() if nnz
                         else np.zeros(0))
        self.hanging = (self.line_ptr[1:] > self.line_ptr[:-1]).astype(np.uint8)
        # Dirichlet on the boundary; a hanging node keeps its hanging constraint
        # (AffineConstraints::merge with right_object_wins, constraints.hpp:142-151)
        self.dirichlet = (self.boundary.astype(bool) & ~self.hanging.astype(bool)).astype(np.uint8)

    def distribute(self, x):
        x = np.array(x, dtype=np.float64)
        lib().orc_distribute_constraints(C.byref(self.d), _p(x))
        return x

    def __del__(self):
        try:
            lib().orc_adata_free(C.byref(self.d))
        except Exception:
            pass


class OracleOperator:
    """Holds the numpy arrays an ``orc_operator`` points to, so they stay alive."""

    def __init__(self, kind, mesh, heat=None, advec=None, vel_kind=VEL_NONE, mask=None,
                 T=None, T_old=None, vel=None, k_cell=None, nthreads=0, use_mask=1, amesh=None,
                 boundary=None, ls=None, level_set=None, reinit=None, io_mask=None):
        self._keep = []
        self.amesh = amesh
        if amesh is not None:
            mesh = make_mesh(3, amesh.degree, tuple(amesh.c.ncell), tuple(amesh.c.lo),
                             tuple(amesh.c.hi))
        self.mesh = mesh
        self.n_dofs = amesh.n_nodes if amesh is not None else n_dofs(mesh)
        self.n = self.n_dofs * (mesh.dim if kind == OP_NORMAL else 1)

        def ptr(a, dtype=np.float64):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dtype)
            self._keep.append(a)
            return a.ctypes.data

        self.op = Operator()
        self.op.kind = kind
        self.op.mesh = mesh
        self.op.heat = heat if heat is not None else HeatParams()
        self.op.advec = advec if advec is not None else AdvecParams()
        self.op.vel_kind = vel_kind
        self.op.mask = ptr(mask, np.uint8)
        self.op.use_mask = use_mask
        self.op.T = ptr(T)
        self.op.T_old = ptr(T_old)
        self.op.vel = ptr(vel)
        self.op.k_cell = ptr(k_cell)
        self.op.nthreads = nthreads
        if amesh is not None:
            self.op.amesh = C.pointer(amesh.c)
            self.op.adata = C.pointer(amesh.d)
        if boundary is not None:
            self._bnd = boundary
            self.op.bnd = C.pointer(boundary)
        if ls is not None:
            self.op.ls = ls
        self.op.level_set = ptr(level_set)
        self.op.io_mask = ptr(io_mask, np.uint8)
        if reinit is not None:
            self.op.reinit = reinit

    def vmult(self, src):
        dst = np.zeros(self.n)
        src = np.ascontiguousarray(src, dtype=np.float64)
        lib().orc_vmult(C.byref(self.op), _p(src), _p(dst))
        return dst

    def vmult_dense(self, src):
        dst = np.zeros(self.n)
        src = np.ascontiguousarray(src, dtype=np.float64)
        lib().orc_vmult_dense(C.byref(self.op), _p(src), _p(dst))
        return dst

    def diagonal(self):
        d = np.zeros(self.n)
        lib().orc_diagonal(C.byref(self.op), _p(d))
        return d

    def inverse_diagonal(self):
        d = self.diagonal()
        lib().orc_invert_diagonal(self.op.kind, C.c_long(self.n), _p(d))
        return d

    def create_rhs(self, src, q_s=None, dst=None):
        dst = np.zeros(self.n) if dst is None else np.ascontiguousarray(dst, dtype=np.float64)
        src = np.ascontiguousarray(src, dtype=np.float64)
        qs = None
        if q_s is not None:
            qs = np.ascontiguousarray(q_s, dtype=np.float64)
        lib().orc_create_rhs(C.byref(self.op), _p(src), _p(qs) if qs is not None else None,
                             _p(dst))
        return dst

    def normal_rhs(self, phi, dst=None):
        """NormalVectorOperator::create_rhs: dst [dim][n] += the projected level-set gradient."""
        dst = np.zeros(self.n) if dst is None else np.array(dst, dtype=np.float64)
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        lib().orc_ls_normal_rhs(C.byref(self.op), _p(phi), _p(dst))
        return dst

    def curvature_rhs(self, normal, tol=None, dst=None):
        """CurvatureOperator::create_rhs: dst [n] += -div of the normalised nodal normal."""
        if tol is None:
            tol = numerical_zero_of_norm(self.mesh)
        dst = np.zeros(self.n) if dst is None else np.array(dst, dtype=np.float64)
        normal = np.ascontiguousarray(normal, dtype=np.float64)
        lib().orc_ls_curvature_rhs(C.byref(self.op), _p(normal), tol, _p(dst))
        return dst

    def reinit_rhs(self, psi, dst=None):
        """OlssonOperator::create_rhs (dst += ...; the normal vector is the operator's vel)."""
        dst = np.zeros(self.n) if dst is None else np.array(dst, dtype=np.float64)
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        lib().orc_reinit_rhs(C.byref(self.op), _p(psi), _p(dst))
        return dst

    def local_matrix(self, cell):
        K = self.mesh.degree + 1
        nloc = K ** self.mesh.dim
        A = np.zeros((nloc, nloc))
        lib().orc_local_matrix(C.byref(self.op), cell, _p(A))
        return A

    def solve(self, b, x0=None, solver=CG, restart=30, max_it=10000, abs_tol=1e-20,
              rel_tol=1e-12, inv_diag=None):
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        it, res = C.c_int(0), C.c_double(0.0)
        invd = None
        if inv_diag is not None:
            invd = np.ascontiguousarray(inv_diag, dtype=np.float64)
        status = lib().orc_solve(C.byref(self.op), solver, restart, max_it, abs_tol, rel_tol,
                                 invd.ctypes.data if invd is not None else None, _p(b), _p(x),
                                 C.byref(it), C.byref(res))
        return x, int(it.value), float(res.value), int(status)

    def heat_implicit_euler(self, T0, n_steps):
        T = np.array(T0, dtype=np.float64)
        norms = np.zeros(n_steps)
        st = lib().orc_heat_implicit_euler(C.byref(self.op), n_steps, _p(T), _p(norms))
        return T, norms, st

    def advec_theta_steps(self, phi0, n_steps, dirichlet_values):
        phi = np.array(phi0, dtype=np.float64)
        norms = np.zeros(n_steps)
        dv = np.ascontiguousarray(dirichlet_values, dtype=np.float64)
        st = lib().orc_advec_theta_steps(C.byref(self.op), n_steps, _p(dv), _p(phi), _p(norms))
        return phi, norms, st


def numerical_zero_of_norm(mesh):
    """compute_numerical_zero_of_norm of the box mesh's volume."""
    vol = 1.0
    for d in range(mesh.dim):
        vol *= mesh.hi[d] - mesh.lo[d]
    return float(lib().orc_numerical_zero_of_norm(mesh.dim, vol))


def fe_l2_norm(mesh, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return float(lib().orc_fe_l2_norm(C.byref(mesh), _p(u)))


def refine_band_cells(ncell, levels, center, radius, half_width, lo=(0.0, 0.0, 0.0),
                      hi=(1.0, 1.0, 1.0)):
    """Reference generator of the band-refined test meshes (pure Python, small meshes only):
    cells whose box meets the shell | |x - center| - radius | < half_width are refined, level by
    level, `levels` times; then the mesh is 2:1 balanced across faces, edges and vertices by
    refining the coarser cell.  Returns int array [n_cells, 4] (level, ix, iy, iz), sorted."""
    if isinstance(ncell, int):
        ncell = (ncell,) * 3
    lo, hi, center = np.asarray(lo, float), np.asarray(hi, float), np.asarray(center, float)
    h0 = (hi - lo) / np.asarray(ncell)

    def meets(l, ix):
        a = lo + h0 * np.asarray(ix) / 2 ** l
        b = a + h0 / 2 ** l
        near = np.clip(center, a, b)
        dmin = np.linalg.norm(near - center)
        far = np.where(np.abs(a - center) > np.abs(b - center), a, b)
        dmax = np.linalg.norm(far - center)
        return dmax > radius - half_width and dmin < radius + half_width

    leaves = {(0, i, j, k) for i in range(ncell[0]) for j in range(ncell[1])
              for k in range(ncell[2])}
    for lev in range(levels):
        new = set()
        for c in leaves:
            if c[0] == lev and meets(c[0], c[1:]):
                for d in range(8):
                    new.add((lev + 1, 2 * c[1] + (d & 1), 2 * c[2] + ((d >> 1) & 1),
                             2 * c[3] + (d >> 2)))
            else:
                new.add(c)
        leaves = new

    def covering(l, ix):
        for lc in range(l, -1, -1):
            s = l - lc
            key = (lc, ix[0] >> s, ix[1] >> s, ix[2] >> s)
            if key in leaves:
                return key
        return None

    changed = True
    while changed:
        changed = False
        refine = set()
        for c in leaves:
            l = c[0]
            if l < 2:
                continue
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        nb = (c[1] + dx, c[2] + dy, c[3] + dz)
                        if any(nb[d] < 0 or nb[d] >= ncell[d] << l for d in range(3)):
                            continue
                        cov = covering(l, nb)
                        if cov is not None and cov[0] < l - 1:
                            refine.add(cov)
        if refine:
            changed = True
            for c in refine:
                leaves.discard(c)
                for d in range(8):
                    leaves.add((c[0] + 1, 2 * c[1] + (d & 1), 2 * c[2] + ((d >> 1) & 1),
                                2 * c[3] + (d >> 2)))
    return np.array(sorted(leaves), dtype=np.int32)


def laser_volumetric_source(X, power, radius, position):
    """Nodal values of the Gaussian volumetric laser source (LaserHeatSourceVolumetric::
    compute_volumetric_heat_source, source/heat/laser_heat_source_volumetric.cpp:20-58, with the
    Gaussian profile of include/meltpooldg/heat/laser_intensity_profiles.hpp:175-218):
    P / (r sqrt(pi/2))^3 exp(-2 |x - x_L|^2 / r^2).  Its integral over R^3 is P."""
    d2 = ((np.asarray(X) - np.asarray(position, dtype=np.float64)) ** 2).sum(axis=1)
    return power / (radius * np.sqrt(np.pi / 2)) ** 3 * np.exp(-2.0 * d2 / radius ** 2)


def least_squares_projection(vmult, old_solutions, rhs):
    """LeastSquaresProjection::vmult (include/meltpooldg/linear_algebra/predictor.hpp:36-80):
    v_s = A x_s orthogonalised in place by modified Gram-Schmidt written with add_and_dot
    (v_s += -P[j,s]/P[j,j] v_j, then P[j+1,s] = v_s . v_{j+1}), stopping at the first s with
    P[s,s] < 1e-16 P[0,0]; c_s = v_s . rhs, back substitution c_s = (c_s - sum_{j>s} c_j P[s,j])
    / P[s,s]; solution = c_0 x_0 + c_1 x_1 + ...  vmult: the operator's product.  Returns the
    solution and the number of old solutions used."""
    n_old = len(old_solutions)
    v = [None] * n_old
    P = np.zeros((n_old, n_old))
    step = 0
    while step < n_old:
        v[step] = np.array(vmult(old_solutions[step]), dtype=np.float64)
        P[0, step] = v[step] @ v[0]
        for j in range(step):
            v[step] = v[step] + (-P[j, step] / P[j, j]) * v[j]
            P[j + 1, step] = v[step] @ v[j + 1]
        if P[step, step] < 1e-16 * P[0, 0]:
            break
        step += 1
    c = [v[s] @ rhs for s in range(step)]
    for s in range(step - 1, -1, -1):
        acc = c[s]
        for j in range(s + 1, step):
            acc -= c[j] * P[s, j]
        c[s] = acc / P[s, s]
    sol = c[0] * np.asarray(old_solutions[0], dtype=np.float64)
    for s in range(1, step):
        sol = sol + c[s] * np.asarray(old_solutions[s], dtype=np.float64)
    return sol, step


def heat_time_steps(mesh, heat, mask, T0, dts, laser=None, predictor="linear",
                    dirichlet_values=None, newton=(10, 1e-9, 1e-10), rel_tol=1e-12, amesh=None,
                    boundary=None, vel=None):
    """The heat time loop of the mp-heat-transfer application, restated on the oracle's
    operator (HeatTransferApplication::run, applications/mp-heat-transfer/
    heat_transfer_application.cpp:52-120, heat-diffuse operator, volumetric laser):
      - time increment: the old increment is 0 at the first step and equal to the new one
        afterwards (TimeIterator::compute_next_time_increment, time_iterator.cpp:44-68);
      - LaserOperation::move_laser (source/heat/laser_operation.cpp:236-250): position +=
        normalised scan direction * scan speed * dt, then the volumetric source there;
      - Predictor::vmult (include/meltpooldg/linear_algebra/predictor.hpp:119-161,
        source/linear_algebra/predictor_linear.cpp:76-96), then the history commit;
        predictor "least_squares": the projection over (current, old) with the Jacobian at
        the current solution and rhs = create_rhs(current solution) with the new increment
        and heat source (heat_diffuse_operation.cpp:228-246), once two predictor calls have
        passed, the current solution before (the reference's fallback leaves a zero old
        solution through its buffer swap; here the current solution is kept as the old one);
      - NewtonRaphsonSolver::solve (source/linear_algebra/newton_raphson_solver.cpp:22-80) with
        the residual of HeatDiffuseOperation (heat_diffuse_operation.cpp:86-116): r = rhs(T),
        A d = r by Jacobi-GMRES (T-dependent law) or CG, stop when |r| and |d| are below the
        tolerances, else T += d.
    heat: HeatParams (dt is set per step); laser: dict(power, radius, start, scan_direction,
    scan_speed); boundary: HeatBoundary (Robin / Neumann sides); vel: nodal velocity [dim][n].
    On an adaptive mesh (amesh, mesh ignored) the fields are kept
    constraint-consistent: the heat source, the iterate after the Dirichlet values and after
    every update are distributed (compute_volumetric_heat_source and the Newton solver's
    distribute_constraints).  Returns the temperature after every step."""
    dist = amesh.distribute if amesh is not None else (lambda v: v)
    T = dist(np.array(T0, dtype=np.float64))
    T_old = T.copy()
    X = amesh.xyz if amesh is not None else node_coords(mesh)
    pos = np.array(laser["start"], dtype=np.float64) if laser else None
    if laser:
        sd = np.asarray(laser.get("scan_direction", (1.0, 0.0, 0.0)), dtype=np.float64)
        sd = sd / np.linalg.norm(sd)
    out = []
    for step, dt in enumerate(dts, start=1):
        dt_old = dt if step > 1 else 0.0
        q = None
        if laser:
            if laser.get("scan_speed", 0.0) != 0.0:
                pos = pos + sd * laser["scan_speed"] * dt
            q = dist(laser_volumetric_source(X, laser["power"], laser["radius"], pos))
        hp = HeatParams(*(getattr(heat, f) for f, _ in HeatParams._fields_))
        hp.dt = dt
        if predictor == "none":
            P = T.copy()
        elif predictor == "zero":
            P = np.zeros_like(T)
        elif predictor == "least_squares":
            if step - 1 >= 2:
                lin = OracleOperator(OP_HEAT, mesh, heat=hp, mask=mask, T=T, T_old=T_old,
                                     amesh=amesh, boundary=boundary, vel=vel,
                                     vel_kind=VEL_NONE if vel is None else VEL_NODAL)
                P, _ = least_squares_projection(lin.vmult, [T, T_old], lin.create_rhs(T, q))
            else:
                P = T.copy()
        else:
            P = T.copy() if abs(dt_old) < 1e-12 else (1 + dt / dt_old) * T - (dt / dt_old) * T_old
        T_old, T = T, P
        if dirichlet_values is not None:
            T[mask == 1] = np.asarray(dirichlet_values)[mask == 1]
        T = dist(T)
        converged = False
        for _ in range(newton[0]):
            op = OracleOperator(OP_HEAT, mesh, heat=hp, mask=mask, T=T, T_old=T_old, amesh=amesh,
                                boundary=boundary, vel=vel,
                                vel_kind=VEL_NONE if vel is None else VEL_NODAL)
            r = op.create_rhs(T_old, q)
            nonsym = hp.t_dependent or vel is not None
            d, _, _, st = op.solve(r, solver=GMRES if nonsym else CG, rel_tol=rel_tol,
                                   inv_diag=op.inverse_diagonal())
            if st != SUCCESS:
                raise RuntimeError("oracle linear solve failed")
            if np.linalg.norm(r) < newton[1] and np.linalg.norm(d) < newton[2]:
                converged = True
                break
            T = dist(T + d)
        if not converged:
            raise RuntimeError("oracle Newton did not converge")
        out.append(T.copy())
    return out


def fe_evaluate_adaptive(degree, ncell, cells, node_xyz, values, points, lo=(0.0, 0.0, 0.0),
                         hi=(1.0, 1.0, 1.0)):
    """Values of a continuous FE_Q field on an adaptive mesh at arbitrary points (brute force,
    small meshes only): the leaf containing the point is found among all leaves by its box, its
    K^3 nodal values are looked up by node coordinate, and the tensor Lagrange basis on the
    Gauss-Lobatto nodes is evaluated at the point's local coordinates (the nodal interpolation
    that SolutionTransfer applies for FE_Q, deal.II fe_q_base / solution_transfer)."""
    if isinstance(ncell, int):
        ncell = (ncell,) * 3
    K = degree + 1
    gl = gauss_lobatto(K)
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    h0 = (hi - lo) / np.asarray(ncell, float)
    cells = np.asarray(cells)
    lev = cells[:, 0].astype(float)
    hc = h0[None, :] / (2.0 ** lev)[:, None]
    c_lo = lo[None, :] + hc * cells[:, 1:4]
    key = {tuple(np.round(x * 2 ** 24).astype(np.int64)): i for i, x in enumerate(node_xyz)}

    def lagrange(t):
        out = np.ones(K)
        for j in range(K):
            for m in range(K):
                if m != j:
                    out[j] *= (t - gl[m]) / (gl[j] - gl[m])
        return out

    res = np.zeros(len(points))
    for i, x in enumerate(points):
        tol = 1e-12 * hc
        inside = np.all((x[None, :] >= c_lo - tol) & (x[None, :] <= c_lo + hc + tol), axis=1)
        c = int(np.nonzero(inside)[0][0])
        t = np.clip((x - c_lo[c]) / hc[c], 0.0, 1.0)
        w = [lagrange(t[d]) for d in range(3)]
        s = 0.0
        for e in range(K ** 3):
            a = (e % K, (e // K) % K, e // (K * K))
            xn = c_lo[c] + hc[c] * np.array([gl[a[0]], gl[a[1]], gl[a[2]]])
            s += values[key[tuple(np.round(xn * 2 ** 24).astype(np.int64))]] * \
                w[0][a[0]] * w[1][a[1]] * w[2][a[2]]
        res[i] = s
    return res


# ------------------------------------------------------------------------------------------
# One-phase CutFEM heat operator (orc_cut.c)
# ------------------------------------------------------------------------------------------
CUT_LIQUID, CUT_INTERSECTED, CUT_GAS = 0, 1, 2


class CutParams(C.Structure):
    _fields_ = [("mat", HeatParams), ("dt", C.c_double), ("theta", C.c_double),
                ("gamma_A", C.c_double), ("gamma_M", C.c_double), ("h", C.c_double)]


class CutMeshC(C.Structure):
    _fields_ = [("n_cells", C.c_long), ("n_dofs", C.c_long),
                ("category", C.POINTER(C.c_int8)), ("box", C.POINTER(C.c_double)),
                ("active", C.POINTER(C.c_uint8)), ("n_faces", C.c_long),
                ("faces", C.POINTER(C.c_long))]


def cut_params(dt, rho_cp=None, k=None, theta=0.5, gamma_A=1.0, gamma_M=1.0, h=None,
               mushy=None, mesh=None):
    """CutParams: the constant liquid law (rho_cp, k) or, with mushy = dict(liquid_k, ...,
    solidus, liquidus), the liquid_solid law.  h defaults to the minimal cell diameter /
    sqrt(dim) of `mesh` (ScratchData::get_min_cell_size, source/core/scratch_data.cpp:170)."""
    hp = HeatParams()
    if mushy is None:
        hp.law = LAW_CONSTANT
        hp.rho_cp, hp.k = rho_cp, k
    else:
        hp.law = LAW_MUSHY
        hp.t_dependent = 1
        for f, v in mushy.items():
            setattr(hp, f, v)
    if h is None:
        hs = [(mesh.hi[d] - mesh.lo[d]) / mesh.ncell[d] for d in range(mesh.dim)]
        h = np.sqrt(sum(x * x for x in hs)) / np.sqrt(mesh.dim)
    return CutParams(hp, dt, theta, gamma_A, gamma_M, h)


class CutOracle:
    """The oracle's one-phase CutFEM heat operator on a structured FE_Q1 mesh with the nodal
    level set phi (liquid where phi > 0).  mask: Dirichlet rows (uint8 [n_dofs]) or None."""

    def __init__(self, mesh, phi, params, mask=None):
        L = lib()
        dp = C.POINTER(C.c_double)
        for name in ("orc_cut_vmult", "orc_cut_diagonal", "orc_cut_residual"):
            getattr(L, name).restype = None
        L.orc_cut_norm_l2.restype = C.c_double
        L.orc_cut_norm_l2.argtypes = [C.POINTER(Mesh), C.POINTER(CutMeshC), dp]
        self.mesh, self.params = mesh, params
        self.n = orc_n = n_dofs(mesh)
        self._phi = np.ascontiguousarray(phi, dtype=np.float64)
        assert self._phi.shape == (orc_n,)
        self.cm = CutMeshC()
        st = L.orc_cut_classify(C.byref(mesh), _p(self._phi), C.byref(self.cm))
        if st != 0:
            L.orc_cut_free(C.byref(self.cm))
            raise ValueError("level set not supported by the cut oracle (status %d)" % st)
        nc = self.cm.n_cells
        self.category = np.ctypeslib.as_array(self.cm.category, shape=(nc,)).copy()
        self.box = np.ctypeslib.as_array(self.cm.box, shape=(nc, 3, 2)).copy()
        self.active = np.ctypeslib.as_array(self.cm.active, shape=(orc_n,)).copy()
        nf = self.cm.n_faces
        self.faces = (np.ctypeslib.as_array(self.cm.faces, shape=(nf, 2)).copy() if nf
                      else np.zeros((0, 2), dtype=np.int64))
        self.mask = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)

    def __del__(self):
        try:
            lib().orc_cut_free(C.byref(self.cm))
        except Exception:
            pass

    def _mask(self):
        return None if self.mask is None else self.mask.ctypes.data_as(C.c_void_p)

    def vmult(self, src, T=None):
        src = np.ascontiguousarray(src, dtype=np.float64)
        Tp = None if T is None else _p(np.ascontiguousarray(T, dtype=np.float64))
        dst = np.zeros(self.n)
        lib().orc_cut_vmult(C.byref(self.mesh), C.byref(self.cm), C.byref(self.params),
                            self._mask(), Tp, _p(src), _p(dst))
        return dst

    def diagonal(self, T=None):
        Tp = None if T is None else _p(np.ascontiguousarray(T, dtype=np.float64))
        d = np.zeros(self.n)
        lib().orc_cut_diagonal(C.byref(self.mesh), C.byref(self.cm), C.byref(self.params),
                               self._mask(), Tp, _p(d))
        return d

    def inverse_diagonal(self, T=None):
        """compute_inverse_diagonal_from_matrixfree's inversion (heat_cut_operator.cpp:
        2141-2153): 1/d where |d| > 1e-14 max(1, |d|_inf), else 1."""
        d = self.diagonal(T)
        lim = 1e-14 * max(1.0, np.max(np.abs(d)))
        return np.where(np.abs(d) > lim, 1.0 / np.where(d == 0, 1.0, d), 1.0)

    def residual(self, T, T_old):
        T = np.ascontiguousarray(T, dtype=np.float64)
        To = np.ascontiguousarray(T_old, dtype=np.float64)
        dst = np.zeros(self.n)
        lib().orc_cut_residual(C.byref(self.mesh), C.byref(self.cm), C.byref(self.params),
                               self._mask(), _p(T), _p(To), _p(dst))
        return dst

    def norm(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().orc_cut_norm_l2(C.byref(self.mesh), C.byref(self.cm), _p(u))

    def dense(self, T=None):
        """The assembled operator, column by column (small meshes only)."""
        A = np.zeros((self.n, self.n))
        e = np.zeros(self.n)
        for j in range(self.n):
            e[j] = 1.0
            A[:, j] = self.vmult(e, T)
            e[j] = 0.0
        return A


def cut_time_steps(mesh, phi, params, T0, n_steps, mask=None, newton=(10, 1e-9, 1e-10)):
    """HeatCutOperation's time loop for a fixed interface (heat_cut_operation.cpp:431-482):
    the predictor is the old solution, then NewtonRaphsonSolver::solve
    (source/linear_algebra/newton_raphson_solver.cpp:22-80) with r = create_rhs(T_old) at T,
    J d = r solved directly on the assembled J (small meshes only), converged when |r| and |d|
    are below the tolerances, else T += d.  Returns (temperatures, cut L2 norms) per step."""
    op = CutOracle(mesh, phi, params, mask)
    T = np.array(T0, dtype=np.float64)
    temps, norms = [], []
    for _ in range(n_steps):
        T_old = T.copy()
        for _ in range(newton[0]):
            r = op.residual(T, T_old)
            A = op.dense(T)
            fixed = (op.active == 0) if mask is None else ((op.active == 0) | (op.mask != 0))
            A[fixed, fixed] = 1.0  # the rows of inactive and constrained DoFs: d = 0 there
            d = np.linalg.solve(A, r)
            if np.linalg.norm(r) < newton[1] and np.linalg.norm(d) < newton[2]:
                break
            T = T + d
        else:
            raise RuntimeError("oracle Newton did not converge")
        temps.append(T.copy())
        norms.append(op.norm(T))
    return temps, norms

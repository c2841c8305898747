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

LAW_CONSTANT, LAW_TANH = 0, 1
VEL_NONE, VEL_NODAL, VEL_VORTEX = 0, 1, 2
OP_HEAT, OP_ADVEC = 0, 1
CG, GMRES = 0, 1
SUCCESS, NO_CONVERGENCE = 0, 1


class Mesh(C.Structure):
    _fields_ = [("dim", C.c_int), ("degree", C.c_int), ("ncell", C.c_int * 3),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3)]


class HeatParams(C.Structure):
    _fields_ = [("law", C.c_int), ("t_dependent", C.c_int), ("rho_cp", C.c_double),
                ("k", C.c_double), ("tanh_amp", C.c_double), ("tanh_Tc", C.c_double),
                ("tanh_width", C.c_double), ("dt", C.c_double)]


class AdvecParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("theta", C.c_double), ("diffusivity", C.c_double),
                ("mass_factor", C.c_double), ("vel_scale", C.c_double)]


class Operator(C.Structure):
    _fields_ = [("kind", C.c_int), ("mesh", Mesh), ("heat", HeatParams), ("advec", AdvecParams),
                ("vel_kind", C.c_int), ("mask", C.c_void_p), ("use_mask", C.c_int),
                ("T", C.c_void_p), ("T_old", C.c_void_p), ("vel", C.c_void_p),
                ("k_cell", C.c_void_p), ("nthreads", C.c_int)]


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


class OracleOperator:
    """Holds the numpy arrays an ``orc_operator`` points to, so they stay alive."""

    def __init__(self, kind, mesh, heat=None, advec=None, vel_kind=VEL_NONE, mask=None,
                 T=None, T_old=None, vel=None, k_cell=None, nthreads=0, use_mask=1):
        self._keep = []
        self.mesh = mesh
        self.n = n_dofs(mesh)

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


def fe_l2_norm(mesh, u):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return float(lib().orc_fe_l2_norm(C.byref(mesh), _p(u)))

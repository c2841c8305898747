"""MI355X matrix-free heat / level-set advection operators (host-side Python mirror).

The compute path is ``libmeltpooldg_amd.so`` (hand-written HIP for gfx950) behind the C ABI in
``include/meltpooldg_amd.h``.  This module binds that ABI with ctypes and mirrors the reference's
operator interface (``OperatorMatrixFree``, include/meltpooldg/core/operator_base.hpp:189-283):
``vmult``, ``compute_inverse_diagonal_from_matrixfree``, ``create_rhs``, ``reset_time_increment``,
``reset_dof_index`` and ``LinearSolver.solve`` (include/meltpooldg/linear_algebra/
linear_solver.hpp:22-134).  Vectors are torch device tensors (torch is used for device memory and
streams only).  There is no CPU fallback: if the shared library is missing, importing the
operators raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmeltpooldg_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "meltpooldg_amd.h")

MF_OK, MF_ERR_ARG, MF_ERR_HIP, MF_NO_CONVERGENCE, MF_BREAKDOWN, MF_ERR_STATE = range(6)
LAW_CONSTANT, LAW_TANH, LAW_MUSHY = 0, 1, 2
SOLVER_CG, SOLVER_GMRES = 0, 1
PRECOND_IDENTITY, PRECOND_JACOBI = 0, 1
PRECISION_FP64, PRECISION_FP32_REFINED = 0, 1


class MFError(RuntimeError):
    """Raised for a non-zero status of a C-ABI call (mirrors dealii::ExceptionBase)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class SolverNoConvergence(MFError):
    """Mirrors dealii::SolverControl::NoConvergence."""


class MeshDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("degree", C.c_int), ("ncell", C.c_int * 3),
                ("lo", C.c_double * 3), ("hi", C.c_double * 3)]


class HeatParams(C.Structure):
    _fields_ = [("law", C.c_int), ("t_dependent", C.c_int), ("rho_cp", C.c_double),
                ("k", C.c_double), ("tanh_amp", C.c_double), ("tanh_Tc", C.c_double),
                ("tanh_width", C.c_double),
                ("liquid_k", C.c_double), ("liquid_rho", C.c_double), ("liquid_cp", C.c_double),
                ("solid_k", C.c_double), ("solid_rho", C.c_double), ("solid_cp", C.c_double),
                ("solidus", C.c_double), ("liquidus", C.c_double)]


class HeatCutParams(C.Structure):
    # mf_heat_cut_params: the liquid material (constant or mushy law), theta, gamma_A, gamma_M
    # and the ghost penalty's cell size (<= 0: min cell diameter / sqrt(3))
    _fields_ = [("material", HeatParams), ("theta", C.c_double), ("gamma_A", C.c_double),
                ("gamma_M", C.c_double), ("cell_side_length", C.c_double)]


class AdvecParams(C.Structure):
    _fields_ = [("theta", C.c_double), ("diffusivity", C.c_double), ("mass_factor", C.c_double),
                ("supg", C.c_int), ("supg_coefficient", C.c_double)]


class SolverDesc(C.Structure):
    _fields_ = [("solver", C.c_int), ("restart", C.c_int), ("max_iterations", C.c_int),
                ("abs_tol", C.c_double), ("rel_tol", C.c_double), ("preconditioner", C.c_int),
                ("precision", C.c_int), ("inner_rel_tol", C.c_double), ("max_outer", C.c_int),
                ("force_re_orthogonalization", C.c_int)]


class TransientDesc(C.Structure):
    _fields_ = [("n_steps", C.c_int), ("checkpoint_interval", C.c_int), ("dt", C.c_double),
                ("power", C.c_double), ("radius", C.c_double), ("start", C.c_double * 3),
                ("scan_velocity", C.c_double * 3), ("solver", SolverDesc)]


class NewtonDesc(C.Structure):
    """mf_newton_desc = NonlinearSolverData (nonlinear_solver_data.hpp:13-20)."""
    _fields_ = [("max_iterations", C.c_int), ("residual_tolerance", C.c_double),
                ("correction_tolerance", C.c_double), ("max_iterations_alt", C.c_int),
                ("residual_tolerance_alt", C.c_double),
                ("correction_tolerance_alt", C.c_double),
                ("use_relative_tolerances", C.c_int)]


def newton_desc(newton=None):
    """NewtonDesc from None (the reference defaults), a tuple (max_iterations,
    residual_tolerance, correction_tolerance) or a dict of NewtonDesc fields; unset fields keep
    the reference defaults."""
    d = NewtonDesc(10, 1e-9, 1e-10, 0, 1e-8, 1e-9, 0)
    if newton is None:
        return d
    if isinstance(newton, dict):
        for k, v in newton.items():
            if k not in dict(NewtonDesc._fields_):
                raise ValueError(f"unknown Newton parameter {k}")
            setattr(d, k, type(getattr(d, k))(v))
        return d
    d.max_iterations, d.residual_tolerance, d.correction_tolerance = newton
    return d


_NEWTON_RESIDUAL = C.CFUNCTYPE(C.c_int, C.c_void_p)
_NEWTON_SOLVE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_int))
_NEWTON_NORMS = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_double),
                            C.POINTER(C.c_double), C.POINTER(C.c_double))
_NEWTON_UPDATE = C.CFUNCTYPE(C.c_int, C.c_void_p)


class NewtonCallbacks(C.Structure):
    """mf_newton_callbacks: the residual / solve_with_jacobian / norms / update rules of
    NewtonRaphsonSolver (newton_raphson_solver.hpp)."""
    _fields_ = [("residual", _NEWTON_RESIDUAL), ("solve_with_jacobian", _NEWTON_SOLVE),
                ("norms", _NEWTON_NORMS), ("update", _NEWTON_UPDATE)]


PREDICTOR_NONE, PREDICTOR_ZERO, PREDICTOR_LINEAR_EXTRAPOLATION = 0, 1, 2
PREDICTOR_LEAST_SQUARES_PROJECTION = 3


class HeatStepperDesc(C.Structure):
    _fields_ = [("laser", C.c_int), ("power", C.c_double), ("radius", C.c_double),
                ("start", C.c_double * 3), ("scan_direction", C.c_double * 3),
                ("scan_speed", C.c_double), ("start_time", C.c_double), ("predictor", C.c_int),
                ("solver", SolverDesc), ("newton", NewtonDesc)]


class NewtonReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("linear_iterations", C.c_int),
                ("residual_norm", C.c_double), ("correction_norm", C.c_double),
                ("status", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SolveReport(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer_iterations", C.c_int),
                ("initial_residual", C.c_double), ("final_residual", C.c_double),
                ("status", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LevelSetProjectionParams(C.Structure):
    """mf_levelset_projection_params: NormalVectorData / CurvatureData (filter_parameter,
    narrow_band.enable, narrow_band.level_set_threshold, compute_normalized_vector)."""
    _fields_ = [("filter_parameter", C.c_double), ("narrow_band", C.c_int),
                ("narrow_band_threshold", C.c_double), ("compute_normalized_vector", C.c_int)]


class ReinitParams(C.Structure):
    """mf_reinit_params: interface_thickness_parameter (type, value) and
    hyperbolic.cg.tangential_diffusion_factor of ReinitializationData."""
    _fields_ = [("thickness_type", C.c_int), ("thickness_value", C.c_double),
                ("tangential_diffusion_factor", C.c_double)]


THICKNESS_PROPORTIONAL, THICKNESS_ABSOLUTE, THICKNESS_CELLS = 0, 1, 2


class HeatBoundary(C.Structure):
    """mf_heat_boundary: Robin / Neumann sides of the heat operator, side s = 2*direction +
    (0 lower, 1 upper) (colourised hyper_rectangle boundary ids)."""
    _fields_ = [("convection_sides", C.c_uint), ("radiation_sides", C.c_uint),
                ("neumann_sides", C.c_uint), ("convection_coefficient", C.c_double),
                ("convection_temperature_infinity", C.c_double), ("emissivity", C.c_double),
                ("radiation_temperature_infinity", C.c_double), ("neumann_flux", C.c_double * 6)]


def build(force: bool = True) -> str:
    """Compile libmeltpooldg_amd.so for gfx950 with hipcc (Makefile in csrc/)."""
    args = ["make", "-s", "-j4", "-C", os.path.join(_HERE, "csrc")]
    if force:
        args.insert(2, "-B")
    subprocess.run(args, check=True)
    return LIB_PATH


_lib = None


def header_functions():
    """Names of the functions the C header declares (for the export check)."""
    with open(HEADER_PATH) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", text)))


def lib():
    """Load the HIP library (raises if it is not built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run meltpooldg_amd.build() "
                              "(the device path has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp, i64 = C.c_void_p, C.c_int64
        L.mf_last_error.restype = C.c_char_p
        L.mf_scratch_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(vp)]
        L.mf_scratch_destroy.argtypes = [vp]
        L.mf_scratch_n_dofs.restype = i64
        L.mf_scratch_n_dofs.argtypes = [vp]
        L.mf_scratch_n_cells.restype = i64
        L.mf_scratch_n_cells.argtypes = [vp]
        L.mf_set_stream.argtypes = [vp, vp]
        L.mf_scratch_set_dirichlet_mask.argtypes = [vp, vp]
        L.mf_scratch_shape_info.argtypes = [vp, vp, vp, vp, vp, vp]
        dp3 = C.POINTER(C.c_double)
        L.mf_mesh_refine_band.argtypes = [C.POINTER(C.c_int), dp3, dp3, C.c_int, dp3, C.c_double,
                                          C.c_double, vp, i64, C.POINTER(i64)]
        L.mf_scratch_create_adaptive.argtypes = [C.c_int, C.POINTER(C.c_int), dp3, dp3, vp, i64,
                                                 C.POINTER(vp)]
        L.mf_scratch_is_adaptive.argtypes = [vp]
        L.mf_scratch_support_points.argtypes = [vp, vp]
        L.mf_scratch_default_dirichlet_mask.argtypes = [vp, vp]
        L.mf_scratch_constraint_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i64),
                                                 C.POINTER(C.c_int)]
        L.mf_scratch_constraint_lines.argtypes = [vp, vp, vp, vp, vp]
        L.mf_distribute_constraints.argtypes = [vp, vp]
        L.mf_heat_operator_create.argtypes = [vp, C.POINTER(HeatParams), C.POINTER(vp)]
        L.mf_advection_operator_create.argtypes = [vp, C.POINTER(AdvecParams), C.POINTER(vp)]
        L.mf_operator_destroy.argtypes = [vp]
        L.mf_operator_reset_time_increment.argtypes = [vp, C.c_double]
        L.mf_operator_reset_dof_index.argtypes = [vp, C.c_int]
        L.mf_operator_set_fields.argtypes = [vp, vp, vp, vp]
        L.mf_operator_set_boundary.argtypes = [vp, C.POINTER(HeatBoundary)]
        L.mf_operator_set_temperature_old.argtypes = [vp, vp]
        L.mf_advection_set_inflow_outflow.argtypes = [vp, vp]
        L.mf_operator_set_inflow_outflow_pre_post.argtypes = [vp, C.c_int]
        L.mf_inflow_outflow_mask.argtypes = [vp, vp, C.c_uint, vp]
        L.mf_normal_vector_operator_create.argtypes = [vp, C.POINTER(LevelSetProjectionParams),
                                                       C.POINTER(vp)]
        L.mf_curvature_operator_create.argtypes = [vp, C.POINTER(LevelSetProjectionParams),
                                                   C.POINTER(vp)]
        L.mf_operator_set_level_set.argtypes = [vp, vp]
        L.mf_operator_n_rows.argtypes = [vp]
        L.mf_operator_n_rows.restype = i64
        L.mf_operator_normal_tolerance.argtypes = [vp]
        L.mf_operator_normal_tolerance.restype = C.c_double
        L.mf_write_vtu.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(C.c_char_p),
                                   C.POINTER(C.c_int), C.POINTER(vp)]
        L.mf_restart_save.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(vp), C.POINTER(i64)]
        L.mf_restart_load_scratch.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.mf_restart_load_vectors.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(vp),
                                              C.POINTER(i64)]
        L.mf_reinit_operator_create.argtypes = [vp, C.POINTER(ReinitParams), C.POINTER(vp)]
        L.mf_operator_set_normal_vector.argtypes = [vp, vp]
        L.mf_reinit_min_epsilon.argtypes = [vp]
        L.mf_reinit_min_epsilon.restype = C.c_double
        L.mf_operator_vmult.argtypes = [vp, vp, vp]
        L.mf_operator_vmult_f32.argtypes = [vp, vp, vp]
        L.mf_operator_vmult_add.argtypes = [vp, vp, vp]
        L.mf_operator_compute_inverse_diagonal.argtypes = [vp, vp]
        L.mf_operator_compute_inverse_diagonal_noalloc.argtypes = [vp, vp]
        L.mf_operator_create_rhs.argtypes = [vp, vp, vp, vp]
        L.mf_solve.argtypes = [vp, C.POINTER(SolverDesc), vp, vp, C.POINTER(SolveReport)]
        L.mf_jacobi_create.argtypes = [vp, C.POINTER(vp)]
        L.mf_jacobi_update.argtypes = [vp]
        L.mf_jacobi_vmult.argtypes = [vp, vp, vp]
        L.mf_jacobi_inverse_diagonal.argtypes = [vp, vp]
        L.mf_jacobi_destroy.argtypes = [vp]
        L.mf_solve_preconditioned.argtypes = [vp, C.POINTER(SolverDesc), vp, vp, vp,
                                              C.POINTER(SolveReport)]
        L.mf_set_tangent_kernel.argtypes = [C.c_int]
        L.mf_set_adaptive_setup.argtypes = [C.c_int]
        L.mf_newton_desc_init.argtypes = [C.POINTER(NewtonDesc)]
        L.mf_newton_solve.argtypes = [C.POINTER(NewtonDesc), C.POINTER(NewtonCallbacks), vp,
                                      C.POINTER(NewtonReport)]
        L.mf_heat_newton_step.argtypes = [vp, vp, vp, vp, vp, C.POINTER(SolverDesc),
                                          C.POINTER(NewtonDesc), C.POINTER(NewtonReport)]
        L.mf_advection_step.argtypes = [vp, vp, vp, vp, C.POINTER(SolverDesc),
                                        C.POINTER(SolveReport)]
        L.mf_predictor_linear_extrapolation.argtypes = [vp, vp, vp, vp, C.c_double]
        L.mf_predictor_least_squares.argtypes = [vp, C.c_int, C.POINTER(vp), vp, vp,
                                                 C.POINTER(C.c_int)]
        L.mf_laser_heat_source.argtypes = [vp, C.c_double, C.c_double, C.POINTER(C.c_double),
                                           vp]
        L.mf_compute_norm_l2.argtypes = [vp, vp, C.POINTER(C.c_double)]
        L.mf_operator_set_transpose.argtypes = [vp, C.c_int]
        L.mf_mass_vmult.argtypes = [vp, vp, vp]
        L.mf_vector_dot.argtypes = [vp, vp, vp, C.POINTER(C.c_double)]
        L.mf_heat_conductivity_gradient.argtypes = [vp, vp, vp, C.c_double, vp]
        L.mf_laser_heat_source_radius_derivative.argtypes = [vp, C.c_double, C.c_double,
                                                             C.POINTER(C.c_double), vp]
        L.mf_adaptive_transfer.argtypes = [vp, vp, vp, vp]
        L.mf_heat_stepper_create.argtypes = [vp, C.POINTER(HeatStepperDesc), vp, C.POINTER(vp)]
        L.mf_heat_stepper_destroy.argtypes = [vp]
        L.mf_heat_stepper_copy.argtypes = [vp, C.c_int, vp]
        L.mf_heat_stepper_step.argtypes = [vp, C.c_double, vp, C.POINTER(NewtonReport)]
        L.mf_heat_stepper_state.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.mf_heat_cut_operator_create.argtypes = [vp, C.POINTER(HeatCutParams), C.POINTER(vp)]
        L.mf_heat_cut_set_level_set.argtypes = [vp, vp]
        L.mf_heat_cut_info.argtypes = [vp] + [C.POINTER(i64)] * 5
        L.mf_heat_cut_cells.argtypes = [vp, vp, vp, vp]
        L.mf_heat_cut_norm_l2.argtypes = [vp, vp, C.POINTER(C.c_double)]
        L.mf_heat_cut_step.argtypes = [vp, vp, vp, vp, C.POINTER(SolverDesc),
                                       C.POINTER(NewtonDesc), C.POINTER(NewtonReport)]
        L.mf_transient_create.argtypes = [vp, C.POINTER(TransientDesc), C.POINTER(vp)]
        L.mf_transient_destroy.argtypes = [vp]
        L.mf_transient_forward.argtypes = [vp, vp, vp]
        L.mf_transient_backward.argtypes = [vp, vp, vp, C.POINTER(C.c_double),
                                            C.POINTER(C.c_double)]
        _lib = L
    return _lib


def splitmix64(x):
    import numpy as np
    with np.errstate(over="ignore"):
        z = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def set_adaptive_setup(device):
    """1 (default): the adaptive-mesh setup (band refinement, numbering, hanging-node
    constraints, hanging-cell tables) on the device when one is present, 0: on the host; both
    give the same mesh, numbering and constraints."""
    _check(lib().mf_set_adaptive_setup(int(bool(device))))


def set_tangent_kernel(variant):
    """0 = pencil kernel (default), 1 = slice-resident kernel; same operator."""
    _check(lib().mf_set_tangent_kernel(int(variant)))


def _check(code):
    if code != MF_OK:
        msg = lib().mf_last_error().decode()
        if code == MF_NO_CONVERGENCE:
            raise SolverNoConvergence(code, msg)
        raise MFError(code, msg)


def _current_device():
    import torch
    return torch.cuda.current_device() if torch.cuda.is_available() else -1


def _ptr(t, n=None, dtype="float64", device=None):
    """Device pointer of tensor t after the checks the C ABI cannot make itself: a contiguous
    device tensor of `dtype` with exactly `n` entries (when given) on device index `device`
    (when given).  Every kernel reads and writes the full length it was set up for, so a
    shorter or narrower tensor is rejected here with ValueError, before any C call."""
    if t is None:
        return None
    if not getattr(t, "is_cuda", False):
        raise ValueError("vectors must be device tensors")
    if not t.is_contiguous():
        raise ValueError("vectors must be contiguous")
    got = str(t.dtype).replace("torch.", "")
    if got != dtype:
        raise ValueError(f"vector of dtype {got}, expected {dtype}")
    if n is not None and int(t.numel()) != int(n):
        raise ValueError(f"vector of {int(t.numel())} entries, expected {int(n)}")
    if device is not None and device >= 0 and t.device.index != device:
        raise ValueError(f"vector on device {t.device.index}, expected device {device}")
    return C.c_void_p(t.data_ptr())


class ScratchData:
    """Mesh, DoF numbering, constraints and 1D tables (reference ScratchData + MatrixFree,
    source/core/scratch_data.cpp:194-293) on a structured hex mesh of [lo, hi]."""

    def __init__(self, degree, ncell, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), stream=None):
        if isinstance(ncell, int):
            ncell = (ncell,) * 3
        self.degree = degree
        self.ncell = tuple(ncell)
        self.lo, self.hi = tuple(lo), tuple(hi)
        d = MeshDesc(3, degree, (C.c_int * 3)(*ncell), (C.c_double * 3)(*lo),
                     (C.c_double * 3)(*hi))
        h = C.c_void_p()
        _check(lib().mf_scratch_create(C.byref(d), C.byref(h)))
        self._h = h
        self.device = _current_device()
        self.n_dofs = int(lib().mf_scratch_n_dofs(h))
        self.n_cells = int(lib().mf_scratch_n_cells(h))
        if stream is not None:
            self.set_stream(stream)

    def _v(self, t, n=None, dtype="float64"):
        """Checked pointer of a vector of this scratch (n_dofs entries unless n is given)."""
        return _ptr(t, self.n_dofs if n is None else n, dtype, getattr(self, "device", None))

    def set_stream(self, stream):
        """stream: a torch.cuda.Stream (its HIP handle is passed down) or a raw handle."""
        handle = getattr(stream, "cuda_stream", stream)
        _check(lib().mf_set_stream(self._h, C.c_void_p(handle)))

    def set_dirichlet_mask(self, mask):
        import numpy as np
        if mask is None:
            _check(lib().mf_scratch_set_dirichlet_mask(self._h, None))
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        assert m.size == self.n_dofs
        _check(lib().mf_scratch_set_dirichlet_mask(self._h, m.ctypes.data_as(C.c_void_p)))

    def shape_info(self):
        import numpy as np
        K = self.degree + 1
        S, DS, Dq, w, gl = (np.zeros((K, K)), np.zeros((K, K)), np.zeros((K, K)), np.zeros(K),
                            np.zeros(K))
        _check(lib().mf_scratch_shape_info(self._h, *(a.ctypes.data_as(C.c_void_p)
                                                      for a in (S, DS, Dq, w, gl))))
        return S, DS, Dq, w, gl

    def support_points(self):
        """(n_dofs, 3) coordinates of the DoF support points (Gauss-Lobatto lattice)."""
        import numpy as np
        gl = self.shape_info()[4]
        p = self.degree
        axes = []
        for d in range(3):
            h = (self.hi[d] - self.lo[d]) / self.ncell[d]
            i = np.arange(self.ncell[d] * p + 1)
            c = np.minimum(i // p, self.ncell[d] - 1)
            axes.append(self.lo[d] + h * (c + gl[i - c * p]))
        z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        return np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)

    def default_dirichlet_mask(self):
        import numpy as np
        m = np.zeros(self.n_dofs, dtype=np.uint8)
        _check(lib().mf_scratch_default_dirichlet_mask(self._h, m.ctypes.data_as(C.c_void_p)))
        return m

    def random_field(self, seed):
        """Seeded U(-1,1) per DoF keyed by the node's lattice coordinates (SplitMix64), the
        synthetic-input convention shared with the test oracle."""
        import numpy as np
        nn = [self.ncell[d] * self.degree + 1 for d in range(3)]
        i0, i1, i2 = np.meshgrid(np.arange(nn[0], dtype=np.uint64),
                                 np.arange(nn[1], dtype=np.uint64),
                                 np.arange(nn[2], dtype=np.uint64), indexing="xy")
        # lexicographic x-fastest order: transpose meshgrid output to [z][y][x]
        i0 = np.transpose(i0, (2, 0, 1)).ravel()
        i1 = np.transpose(i1, (2, 0, 1)).ravel()
        i2 = np.transpose(i2, (2, 0, 1)).ravel()
        key = i0 | (i1 << np.uint64(21)) | (i2 << np.uint64(42))
        z = splitmix64(key ^ splitmix64(np.uint64(seed)))
        return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53) - 1.0

    def boundary_mask(self):
        import numpy as np
        if getattr(self, "adaptive", False):
            return self.default_dirichlet_mask()
        nn = [self.ncell[d] * self.degree + 1 for d in range(3)]
        m = np.zeros((nn[2], nn[1], nn[0]), dtype=np.uint8)
        m[0, :, :] = m[-1, :, :] = 1
        m[:, 0, :] = m[:, -1, :] = 1
        m[:, :, 0] = m[:, :, -1] = 1
        return m.ravel()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().mf_scratch_destroy(self._h)
                self._h = None
        except Exception:
            pass


def adaptive_transfer(old_scratch, old_vec, new_scratch, new_vec):
    """Field transfer between two adaptive meshes of the same base mesh and degree
    (mf_adaptive_transfer: the remesh step of AMR::refine_grid with SolutionTransfer)."""
    _check(lib().mf_adaptive_transfer(old_scratch._h, old_scratch._v(old_vec), new_scratch._h,
                                      new_scratch._v(new_vec)))


def refine_band_cells(ncell, levels, center, radius, half_width, lo=(0.0, 0.0, 0.0),
                      hi=(1.0, 1.0, 1.0)):
    """Band-refined, 2:1-balanced octree mesh (mf_mesh_refine_band): int32 array [n, 4] of
    (level, ix, iy, iz) in Morton order."""
    import numpy as np
    if isinstance(ncell, int):
        ncell = (ncell,) * 3
    nc = (C.c_int * 3)(*ncell)
    lo3, hi3, ce3 = ((C.c_double * 3)(*v) for v in (lo, hi, center))
    n = C.c_int64(0)
    _check(lib().mf_mesh_refine_band(nc, lo3, hi3, levels, ce3, radius, half_width, None, 0,
                                     C.byref(n)))
    cells = np.zeros((n.value, 4), dtype=np.int32)
    _check(lib().mf_mesh_refine_band(nc, lo3, hi3, levels, ce3, radius, half_width,
                                     cells.ctypes.data_as(C.c_void_p), n.value, C.byref(n)))
    return cells


class AdaptiveScratchData(ScratchData):
    """ScratchData on an adaptive octree mesh with hanging nodes (mf_scratch_create_adaptive).
    The default Dirichlet rows (non-hanging boundary nodes) are set at creation."""

    def __init__(self, degree, ncell, cells, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), stream=None):
        import numpy as np
        if isinstance(ncell, int):
            ncell = (ncell,) * 3
        self.adaptive = True
        self.degree = degree
        self.ncell = tuple(ncell)
        self.lo, self.hi = tuple(lo), tuple(hi)
        self.cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
        h = C.c_void_p()
        _check(lib().mf_scratch_create_adaptive(degree, (C.c_int * 3)(*ncell),
                                                (C.c_double * 3)(*lo), (C.c_double * 3)(*hi),
                                                self.cells.ctypes.data_as(C.c_void_p),
                                                len(self.cells), C.byref(h)))
        self._h = h
        self.device = _current_device()
        self.n_dofs = int(lib().mf_scratch_n_dofs(h))
        self.n_cells = int(lib().mf_scratch_n_cells(h))
        if stream is not None:
            self.set_stream(stream)

    def support_points(self):
        import numpy as np
        xyz = np.zeros((self.n_dofs, 3))
        _check(lib().mf_scratch_support_points(self._h, xyz.ctypes.data_as(C.c_void_p)))
        return xyz

    def constraint_info(self):
        nh, ne, nc = C.c_int64(0), C.c_int64(0), C.c_int(0)
        _check(lib().mf_scratch_constraint_info(self._h, C.byref(nh), C.byref(ne), C.byref(nc)))
        return {"n_hanging": nh.value, "n_entries": ne.value, "n_colours": nc.value}

    def constraint_lines(self):
        """(node_line[n_dofs], line_ptr, line_col, line_val) as numpy arrays."""
        import numpy as np
        info = self.constraint_info()
        node_line = np.zeros(self.n_dofs, dtype=np.int32)
        ptr = np.zeros(info["n_hanging"] + 1, dtype=np.uint32)
        col = np.zeros(max(info["n_entries"], 1), dtype=np.uint32)
        val = np.zeros(max(info["n_entries"], 1))
        _check(lib().mf_scratch_constraint_lines(self._h, *(a.ctypes.data_as(C.c_void_p)
                                                           for a in (node_line, ptr, col, val))))
        return node_line, ptr, col[:info["n_entries"]], val[:info["n_entries"]]

    def distribute_constraints(self, x):
        """x_h = sum_m w x_m on the hanging nodes of the device vector x (in place)."""
        _check(lib().mf_distribute_constraints(self._h, self._v(x)))

    def random_field(self, seed):
        raise NotImplementedError("lattice-keyed random fields exist on structured meshes only")


class _OperatorBase:
    """Common part of the OperatorMatrixFree mirror."""

    # entries of the source vector of create_rhs, in units of n_dofs
    _rhs_src_blocks = 1

    def __init__(self, scratch, handle):
        self.scratch = scratch
        self._h = handle
        self._fields = ()
        self.n_rows = int(lib().mf_operator_n_rows(handle))

    def _r(self, t, dtype="float64"):
        """Checked pointer of a vector of the operator's rows."""
        return self.scratch._v(t, self.n_rows, dtype)

    def reset_time_increment(self, dt):
        _check(lib().mf_operator_reset_time_increment(self._h, float(dt)))

    def reset_dof_index(self, dof_index):
        _check(lib().mf_operator_reset_dof_index(self._h, int(dof_index)))

    def set_fields(self, T=None, velocity=None, k_cell=None):
        # keep references alive: the library holds the pointers (as the reference holds const refs)
        self._fields = (T, velocity, k_cell)
        sc = self.scratch
        _check(lib().mf_operator_set_fields(self._h, sc._v(T), sc._v(velocity, 3 * sc.n_dofs),
                                            sc._v(k_cell, sc.n_cells)))

    def vmult(self, dst, src):
        _check(lib().mf_operator_vmult(self._h, self._r(dst), self._r(src)))

    def vmult_add(self, dst, src):
        _check(lib().mf_operator_vmult_add(self._h, self._r(dst), self._r(src)))

    def vmult_f32(self, dst, src):
        _check(lib().mf_operator_vmult_f32(self._h, self._r(dst, "float32"),
                                           self._r(src, "float32")))

    def compute_inverse_diagonal_from_matrixfree(self, diag):
        _check(lib().mf_operator_compute_inverse_diagonal(self._h, self._r(diag)))

    def compute_inverse_diagonal_noalloc(self, diag):
        """The same inverse diagonal with no device allocation (static per-GPU scratch; calls
        that overlap on one GPU race)."""
        _check(lib().mf_operator_compute_inverse_diagonal_noalloc(self._h, self._r(diag)))

    def create_rhs(self, dst, src, q_s=None):
        sc = self.scratch
        _check(lib().mf_operator_create_rhs(self._h, self._r(dst),
                                            sc._v(src, self._rhs_src_blocks * sc.n_dofs),
                                            sc._v(q_s)))

    def set_transpose(self, transpose):
        """Apply A^T in vmult / solves (adjoint); the diagonal is unchanged."""
        _check(lib().mf_operator_set_transpose(self._h, int(bool(transpose))))

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().mf_operator_destroy(self._h)
                self._h = None
        except Exception:
            pass


class HeatDiffuseOperator(_OperatorBase):
    """Heat Jacobian (reference HeatDiffuseMultiPhaseOperator, source/heat/
    heat_diffuse_operator.cpp): tangent (rho cp/dt) M + K(k(T)) [+ dk/dT grad T u
    + d(rho cp)/dT ((T - T_old)/dt + grad T . w) u].  law LAW_MUSHY takes `mushy`, a dict of
    liquid_k, liquid_rho, liquid_cp, solid_k, solid_rho, solid_cp, solidus, liquidus (the
    liquid_solid material with the mushy_zone transition)."""

    def __init__(self, scratch, law=LAW_CONSTANT, rho_cp=1.0, k=1.0, tanh_amp=0.0, tanh_Tc=0.0,
                 tanh_width=1.0, t_dependent=False, dt=None, mushy=None):
        p = HeatParams(law, int(bool(t_dependent)), rho_cp, k, tanh_amp, tanh_Tc, tanh_width)
        for f, v in (mushy or {}).items():
            if f not in dict(HeatParams._fields_[7:]):
                raise ValueError(f"unknown mushy-zone parameter {f}")
            setattr(p, f, float(v))
        h = C.c_void_p()
        _check(lib().mf_heat_operator_create(scratch._h, C.byref(p), C.byref(h)))
        super().__init__(scratch, h)
        if dt is not None:
            self.reset_time_increment(dt)

    def set_temperature_old(self, T_old):
        """The old temperature (read by the mushy-zone Jacobian; kept by reference)."""
        self._T_old = T_old
        _check(lib().mf_operator_set_temperature_old(self._h, self.scratch._v(T_old)))

    def set_boundary(self, convection_sides=(), radiation_sides=(), neumann_sides=None,
                     convection_coefficient=0.0, convection_temperature_infinity=0.0,
                     emissivity=0.0, radiation_temperature_infinity=0.0):
        """Robin (convection, radiation) and Neumann sides; neumann_sides maps side -> q.n.
        Sides are 0..5 = 2*direction + (0 lower, 1 upper)."""
        b = HeatBoundary()
        b.convection_sides = sum(1 << int(x) for x in convection_sides)
        b.radiation_sides = sum(1 << int(x) for x in radiation_sides)
        for side, g in (neumann_sides or {}).items():
            b.neumann_sides |= 1 << int(side)
            b.neumann_flux[int(side)] = float(g)
        b.convection_coefficient = convection_coefficient
        b.convection_temperature_infinity = convection_temperature_infinity
        b.emissivity = emissivity
        b.radiation_temperature_infinity = radiation_temperature_infinity
        _check(lib().mf_operator_set_boundary(self._h, C.byref(b)))
        return b

    def clear_boundary(self):
        _check(lib().mf_operator_set_boundary(self._h, None))


class AdvectionDiffusionOperator(_OperatorBase):
    """Level-set advection-diffusion tangent (reference AdvectionDiffusionOperator,
    source/level_set/advection_diffusion_operator.cpp): (m/dt) M + theta (w . grad) + theta D K.
    supg=True adds the SUPG stabilisation (ConvectionStabilizationType::SUPG) with
    supg_coefficient (ConvectionStabilizationData::coefficient, default -1: tau = h / max |w|
    per cell)."""

    def __init__(self, scratch, theta=1.0, diffusivity=0.0, mass_factor=1.0, dt=None, supg=False,
                 supg_coefficient=-1.0):
        p = AdvecParams(theta, diffusivity, mass_factor, int(bool(supg)), supg_coefficient)
        h = C.c_void_p()
        _check(lib().mf_advection_operator_create(scratch._h, C.byref(p), C.byref(h)))
        super().__init__(scratch, h)
        if dt is not None:
            self.reset_time_increment(dt)

    def set_inflow_outflow_bc(self, mask):
        """Inflow/outflow DoFs (uint8 numpy mask of n_dofs, None removes them); the Dirichlet
        mask of the scratch must be set before (AdvectionDiffusionOperator::
        set_inflow_outflow_bc)."""
        import numpy as np
        if mask is None:
            _check(lib().mf_advection_set_inflow_outflow(self._h, None))
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        if m.size != self.scratch.n_dofs:
            raise ValueError(f"mask of {m.size} entries, expected {self.scratch.n_dofs}")
        _check(lib().mf_advection_set_inflow_outflow(self._h, m.ctypes.data_as(C.c_void_p)))

    def set_apply_inflow_outflow_constraints_pre_post_matrix_free(self, enable):
        _check(lib().mf_operator_set_inflow_outflow_pre_post(self._h, int(bool(enable))))


class HeatCutOperator(_OperatorBase):
    """One-phase CutFEM heat tangent (reference HeatCutOperator, source/heat/
    heat_cut_operator.cpp, two_phase false, FE_Q1): over the liquid part {phi > 0} of the mesh
    (cv + dcv/dT T) u v + dt theta (k grad u + dk/dT grad T u) . grad v, plus the ghost penalty
    gamma [d_n u][d_n v] on the faces between an intersected cell and an intersected or liquid
    cell.  law LAW_CONSTANT with rho_cp = rho cp and k of the liquid phase; LAW_MUSHY with
    `mushy` (the liquid_solid material, do_solidification).  Interfaces must be axis-aligned
    planes within each intersected cell."""

    def __init__(self, scratch, rho_cp=1.0, k=1.0, law=LAW_CONSTANT, mushy=None, theta=0.5,
                 gamma_A=1.0, gamma_M=1.0, cell_side_length=0.0, dt=None, level_set=None):
        mat = HeatParams(law, 0, rho_cp, k, 0.0, 0.0, 1.0)
        for f, v in (mushy or {}).items():
            if f not in dict(HeatParams._fields_[7:]):
                raise ValueError(f"unknown mushy-zone parameter {f}")
            setattr(mat, f, float(v))
        p = HeatCutParams(mat, theta, gamma_A, gamma_M, cell_side_length)
        h = C.c_void_p()
        _check(lib().mf_heat_cut_operator_create(scratch._h, C.byref(p), C.byref(h)))
        super().__init__(scratch, h)
        if dt is not None:
            self.reset_time_increment(dt)
        if level_set is not None:
            self.set_level_set(level_set)

    def init_time_advance(self, dt):
        """HeatCutOperator::init_time_advance: the ost factors dt theta and dt (1 - theta)."""
        self.reset_time_increment(dt)

    def set_level_set(self, phi):
        """Nodal level set (numpy, n_dofs): classification, sub-boxes, faces, DoFs."""
        import numpy as np
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        if phi.shape != (self.scratch.n_dofs,):
            raise ValueError("level set must hold one value per node")
        _check(lib().mf_heat_cut_set_level_set(self._h, phi.ctypes.data_as(C.c_void_p)))

    def info(self):
        v = [C.c_int64() for _ in range(5)]
        _check(lib().mf_heat_cut_info(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("n_liquid", "n_intersected", "n_gas", "n_faces", "n_active_dofs"),
                        (x.value for x in v)))

    def cells(self):
        """(category[n_cells], box[n_cells][3][2], active[n_dofs]) host copies."""
        import numpy as np
        sc = self.scratch
        cat = np.zeros(sc.n_cells, dtype=np.int8)
        box = np.zeros((sc.n_cells, 3, 2))
        act = np.zeros(sc.n_dofs, dtype=np.uint8)
        _check(lib().mf_heat_cut_cells(self._h, cat.ctypes.data_as(C.c_void_p),
                                       box.ctypes.data_as(C.c_void_p),
                                       act.ctypes.data_as(C.c_void_p)))
        return cat, box, act

    def compute_cut_norm(self, u):
        """CutUtil::compute_cut_norm (L2) of a device vector over the liquid region."""
        r = C.c_double()
        _check(lib().mf_heat_cut_norm_l2(self._h, self.scratch._v(u), C.byref(r)))
        return r.value


def heat_cut_step(operator, T, T_old, dirichlet_values=None, newton=None, **solver):
    """One HeatCutOperation::solve (Newton over create_rhs(T_old) and the cut tangent).  The
    reference's heat linear solver defaults are GMRES with the Diagonal preconditioner."""
    solver.setdefault("solver", "GMRES")
    rep = NewtonReport()
    sc = operator.scratch
    code = lib().mf_heat_cut_step(operator._h, sc._v(T), sc._v(T_old), sc._v(dirichlet_values),
                                  C.byref(_solver_desc(**solver)), C.byref(newton_desc(newton)),
                                  C.byref(rep))
    _check(code)
    return rep.as_dict()


class HeatCutOperation:
    """HeatCutOperation (source/heat/heat_cut_operation.cpp) for a fixed interface on the
    device: set_initial_condition copies T0, init_time_advance rotates the history and takes
    the old solution as the predictor (:431-455), solve() runs the Newton step (:457-482),
    compute_L2_norm() is the cut L2 norm (:485-496)."""

    def __init__(self, operator, newton=None, **solver):
        import torch
        self.operator = operator
        n = operator.scratch.n_dofs
        dev = torch.device("cuda", torch.cuda.current_device())
        self.T = torch.zeros(n, dtype=torch.float64, device=dev)
        self.T_old = torch.zeros_like(self.T)
        self.newton = newton
        self.solver = solver
        self.last_report = None

    def set_initial_condition(self, T0):
        import torch
        self.T.copy_(torch.as_tensor(T0, dtype=torch.float64))

    def init_time_advance(self, dt):
        self.T_old.copy_(self.T)
        self.operator.init_time_advance(dt)

    def solve(self):
        self.last_report = heat_cut_step(self.operator, self.T, self.T_old, newton=self.newton,
                                         **self.solver)
        return self.last_report

    def compute_L2_norm(self):
        return self.operator.compute_cut_norm(self.T)

    def get_temperature(self):
        return self.T


def inflow_outflow_mask(scratch, velocity, sides=(0, 1, 2, 3, 4, 5)):
    """Inflow DoFs of a box (mf_inflow_outflow_mask): boundary nodes of the given sides (0..5 =
    2*direction + upper) whose nodal velocity points inwards or along the side (n . u <= 0);
    returns a uint8 numpy array of n_dofs entries."""
    import numpy as np
    mask = np.zeros(scratch.n_dofs, dtype=np.uint8)
    bits = sum(1 << int(x) for x in sides)
    _check(lib().mf_inflow_outflow_mask(scratch._h, scratch._v(velocity, 3 * scratch.n_dofs),
                                        bits, mask.ctypes.data_as(C.c_void_p)))
    return mask


class _LevelSetProjection(_OperatorBase):
    """Tangent sum_q m_q (u v + eta grad u . grad v) JxW, eta = filter_parameter *
    (cell diameter / sqrt(3))^2, m_q the narrow-band indicator |phi_q| < threshold."""

    _create = None

    def __init__(self, scratch, filter_parameter=2.0, narrow_band=False,
                 narrow_band_threshold=0.9999999, compute_normalized_vector=False,
                 level_set=None):
        p = LevelSetProjectionParams(float(filter_parameter), int(bool(narrow_band)),
                                     float(narrow_band_threshold),
                                     int(bool(compute_normalized_vector)))
        h = C.c_void_p()
        _check(getattr(lib(), self._create)(scratch._h, C.byref(p), C.byref(h)))
        super().__init__(scratch, h)
        self._level_set = None
        if level_set is not None:
            self.set_level_set(level_set)

    def set_level_set(self, level_set):
        """The level set of the narrow band (kept by reference; call again after it changes)."""
        self._level_set = level_set
        _check(lib().mf_operator_set_level_set(self._h, self.scratch._v(level_set)))


class NormalVectorOperator(_LevelSetProjection):
    """NormalVectorOperator (source/level_set/normal_vector_operator.cpp): a block operator on
    [3][n_dofs] vectors; create_rhs(dst [3][n_dofs], level_set) adds the projected gradient."""
    _create = "mf_normal_vector_operator_create"


class CurvatureOperator(_LevelSetProjection):
    """CurvatureOperator (source/level_set/curvature_operator.cpp); create_rhs(dst, normal
    [3][n_dofs]) adds -div of the nodally normalised normal vector."""
    _create = "mf_curvature_operator_create"
    _rhs_src_blocks = 3

    @property
    def normal_tolerance(self):
        return float(lib().mf_operator_normal_tolerance(self._h))


class ReinitializationOperator(_OperatorBase):
    """OlssonOperator (source/level_set/olsson_CG_operator.cpp): tangent u v / dt +
    (eps_t grad u + (eps_n - eps_t)(grad u . n) n) . grad v with n the normalised normal vector;
    create_rhs(dst, level_set) adds the compressive flux and diffusion terms."""

    def __init__(self, scratch, thickness_type=THICKNESS_PROPORTIONAL, thickness_value=0.5,
                 tangential_diffusion_factor=0.0, dt=None, normal=None):
        p = ReinitParams(int(thickness_type), float(thickness_value),
                         float(tangential_diffusion_factor))
        h = C.c_void_p()
        _check(lib().mf_reinit_operator_create(scratch._h, C.byref(p), C.byref(h)))
        super().__init__(scratch, h)
        self._normal = None
        if dt is not None:
            self.reset_time_increment(dt)
        if normal is not None:
            self.set_normal_vector(normal)

    def set_normal_vector(self, normal):
        self._normal = normal
        _check(lib().mf_operator_set_normal_vector(self._h,
                                                   self.scratch._v(normal, 3 * self.scratch.n_dofs)))

    @property
    def min_epsilon(self):
        return float(lib().mf_reinit_min_epsilon(self._h))

    @property
    def normal_tolerance(self):
        return float(lib().mf_operator_normal_tolerance(self._h))


def _solver_desc(solver="CG", preconditioner="Diagonal", max_iterations=10000,
                 rel_tolerance=1e-12, abs_tolerance=1e-20, restart=30, precision="fp64",
                 inner_rel_tolerance=0.0, max_outer=50, force_re_orthogonalization=False):
    d = SolverDesc()
    d.force_re_orthogonalization = int(bool(force_re_orthogonalization))
    d.solver = {"CG": SOLVER_CG, "GMRES": SOLVER_GMRES}[solver]
    d.restart = restart
    d.max_iterations = max_iterations
    d.abs_tol = abs_tolerance
    d.rel_tol = rel_tolerance
    d.preconditioner = {"Identity": PRECOND_IDENTITY, "Diagonal": PRECOND_JACOBI}[preconditioner]
    d.precision = {"fp64": PRECISION_FP64, "fp32_refined": PRECISION_FP32_REFINED}[precision]
    d.inner_rel_tol = inner_rel_tolerance
    d.max_outer = max_outer
    return d


def compute_norm_l2(scratch, u):
    """FE L2 norm of a device vector (VectorTools::compute_norm semantics)."""
    r = C.c_double()
    _check(lib().mf_compute_norm_l2(scratch._h, scratch._v(u), C.byref(r)))
    return r.value


def laser_heat_source(scratch, q_s, power, radius, position):
    """Nodal Gaussian volumetric laser source (LaserHeatSourceVolumetric)."""
    pos = (C.c_double * 3)(*position)
    _check(lib().mf_laser_heat_source(scratch._h, power, radius, pos, scratch._v(q_s)))


def predictor_least_squares(operator, old_solutions, rhs, solution):
    """LeastSquaresProjection::vmult (include/meltpooldg/linear_algebra/predictor.hpp:17-81):
    solution = sum_s c_s x_s minimising |A sum_s c_s x_s - rhs| over the old solutions x_s
    (device vectors of the operator's rows); returns how many of them entered."""
    ptrs = (C.c_void_p * len(old_solutions))(*[operator._r(x) for x in old_solutions])
    used = C.c_int(0)
    _check(lib().mf_predictor_least_squares(operator._h, len(old_solutions), ptrs,
                                            operator._r(rhs), operator._r(solution),
                                            C.byref(used)))
    return used.value


def predictor_linear_extrapolation(scratch, dst, x_n, x_nm1, tau=1.0):
    _check(lib().mf_predictor_linear_extrapolation(scratch._h, scratch._v(dst), scratch._v(x_n),
                                                   scratch._v(x_nm1), float(tau)))


def newton_solve(residual, solve_with_jacobian, norms, update, newton=None):
    """NewtonRaphsonSolver::solve over Python callables (mf_newton_solve; host only):
    residual() computes rhs = -R(x); solve_with_jacobian() -> linear iterations;
    norms(need_x) -> (|rhs|, |d|, |x|); update() does x += d.  Returns the report dict; raises
    SolverNoConvergence like ExcNewtonDidNotConverge."""
    def wrap(f):
        def g(*a):
            try:
                return f(*a)
            except Exception:  # an exception may not cross the C frames
                return MF_ERR_STATE
        return g

    def res(_):
        residual()
        return MF_OK

    def sol(_, its):
        its[0] = int(solve_with_jacobian())
        return MF_OK

    def nrm(_, need_x, r, d, x):
        r[0], d[0], x[0] = (float(v) for v in norms(bool(need_x)))
        return MF_OK

    def upd(_):
        update()
        return MF_OK

    cb = NewtonCallbacks(_NEWTON_RESIDUAL(wrap(res)), _NEWTON_SOLVE(wrap(sol)),
                         _NEWTON_NORMS(wrap(nrm)), _NEWTON_UPDATE(wrap(upd)))
    rep = NewtonReport()
    code = lib().mf_newton_solve(C.byref(newton_desc(newton)), C.byref(cb), None, C.byref(rep))
    _check(code)
    return rep.as_dict()


def heat_newton_step(operator, T, T_old, q_s=None, dirichlet_values=None, newton=None,
                     **solver):
    """One implicit heat step solved by Newton (HeatDiffuseOperation::solve)."""
    nd = newton_desc(newton)
    rep = NewtonReport()
    sc = operator.scratch
    code = lib().mf_heat_newton_step(operator._h, sc._v(T), sc._v(T_old), sc._v(q_s),
                                     sc._v(dirichlet_values), C.byref(_solver_desc(**solver)),
                                     C.byref(nd), C.byref(rep))
    _check(code)
    return rep.as_dict()


def advection_step(operator, phi, phi_old, dirichlet_values=None, **solver):
    """One advection-diffusion step with Dirichlet lifting (AdvectionDiffusionOperation::solve)."""
    rep = SolveReport()
    sc = operator.scratch
    code = lib().mf_advection_step(operator._h, sc._v(phi), sc._v(phi_old), sc._v(dirichlet_values),
                                   C.byref(_solver_desc(**solver)), C.byref(rep))
    _check(code)
    return rep.as_dict()


def mass_vmult(scratch, dst, src):
    _check(lib().mf_mass_vmult(scratch._h, scratch._v(dst), scratch._v(src)))


def vector_dot(scratch, a, b):
    r = C.c_double()
    _check(lib().mf_vector_dot(scratch._h, scratch._v(a), scratch._v(b), C.byref(r)))
    return r.value


def conductivity_gradient(operator, lam, T, grad_k_cell, scale=1.0):
    sc = operator.scratch
    _check(lib().mf_heat_conductivity_gradient(operator._h, sc._v(lam), sc._v(T), float(scale),
                                               sc._v(grad_k_cell, sc.n_cells)))


def laser_heat_source_radius_derivative(scratch, dq, power, radius, position):
    pos = (C.c_double * 3)(*position)
    _check(lib().mf_laser_heat_source_radius_derivative(scratch._h, power, radius, pos,
                                                        scratch._v(dq)))


class HeatTimeStepper:
    """The heat time loop of the mp-heat-transfer application on the device
    (HeatTransferApplication::run, heat_transfer_application.cpp:52-120, with the heat-diffuse
    operator and a Gaussian volumetric laser; mf_heat_stepper_*): each step() advances the
    time, moves the laser, recomputes the heat source, applies the predictor and runs the
    Newton solve.  T0 is copied; solution(), old_solution() and heat_source() return copies."""

    def __init__(self, operator, T0, laser=None, predictor=PREDICTOR_LINEAR_EXTRAPOLATION,
                 start_time=0.0, newton=None, **solver):
        import torch
        self.operator = operator
        self.n_dofs = operator.scratch.n_dofs
        d = HeatStepperDesc()
        if laser is not None:
            d.laser = 1
            d.power = laser["power"]
            d.radius = laser["radius"]
            d.start = (C.c_double * 3)(*laser["start"])
            d.scan_direction = (C.c_double * 3)(*laser.get("scan_direction", (1.0, 0.0, 0.0)))
            d.scan_speed = laser.get("scan_speed", 0.0)
        d.start_time = start_time
        d.predictor = predictor
        d.solver = _solver_desc(**solver)
        d.newton = newton_desc(newton)
        self._desc = d
        h = C.c_void_p()
        _check(lib().mf_heat_stepper_create(operator._h, C.byref(d), operator.scratch._v(T0),
                                            C.byref(h)))
        self._h = h
        self._torch = torch

    def step(self, dt, dirichlet_values=None):
        rep = NewtonReport()
        code = lib().mf_heat_stepper_step(self._h, float(dt),
                                          self.operator.scratch._v(dirichlet_values),
                                          C.byref(rep))
        _check(code)
        return rep.as_dict()

    def _state(self):
        T, To, q = C.c_void_p(), C.c_void_p(), C.c_void_p()
        t = C.c_double()
        pos = (C.c_double * 3)()
        _check(lib().mf_heat_stepper_state(self._h, C.byref(T), C.byref(To), C.byref(q),
                                           C.byref(t), pos))
        return T.value, To.value, q.value, t.value, tuple(pos)

    def _copy(self, which):
        torch = self._torch
        out = torch.empty(self.n_dofs, dtype=torch.float64, device="cuda")
        _check(lib().mf_heat_stepper_copy(self._h, which, self.operator.scratch._v(out)))
        return out

    def solution(self):
        return self._copy(0)

    def old_solution(self):
        return self._copy(1)

    def heat_source(self):
        return self._copy(2)

    @property
    def time(self):
        return self._state()[3]

    @property
    def laser_position(self):
        return self._state()[4]

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().mf_heat_stepper_destroy(self._h)
                self._h = None
        except Exception:
            pass


class HeatTransient:
    """Moving-laser implicit heat transient with a checkpointed hand-written adjoint
    (C++ forward/backward behind the C ABI; no reference counterpart)."""

    def __init__(self, operator, n_steps, dt, power, radius, start, scan_velocity,
                 checkpoint_interval=5, **solver):
        self.operator = operator
        d = TransientDesc()
        d.n_steps = n_steps
        d.checkpoint_interval = checkpoint_interval
        d.dt = dt
        d.power = power
        d.radius = radius
        d.start = (C.c_double * 3)(*start)
        d.scan_velocity = (C.c_double * 3)(*scan_velocity)
        d.solver = _solver_desc(**solver)
        self.desc = d
        h = C.c_void_p()
        _check(lib().mf_transient_create(operator._h, C.byref(d), C.byref(h)))
        self._h = h

    def forward(self, T0, T_final):
        sc = self.operator.scratch
        _check(lib().mf_transient_forward(self._h, sc._v(T0), sc._v(T_final)))

    def backward(self, dJ_dTN, grad_k_cell=None):
        gp, gr = C.c_double(), C.c_double()
        sc = self.operator.scratch
        _check(lib().mf_transient_backward(self._h, sc._v(dJ_dTN), sc._v(grad_k_cell, sc.n_cells),
                                           C.byref(gp), C.byref(gr)))
        return gp.value, gr.value

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().mf_transient_destroy(self._h)
                self._h = None
        except Exception:
            pass


class JacobiPreconditioner:
    """Mirror of MeltPoolDG::JacobiPreconditioner (preconditioner_jacobi.hpp:29-89): update()
    computes the inverse diagonal of the operator once; solves given this object reuse it."""

    def __init__(self, operator):
        self.operator = operator
        h = C.c_void_p()
        _check(lib().mf_jacobi_create(operator._h, C.byref(h)))
        self._h = h

    def update(self):
        _check(lib().mf_jacobi_update(self._h))

    def vmult(self, dst, src):
        _check(lib().mf_jacobi_vmult(self._h, self.operator._r(dst), self.operator._r(src)))

    def inverse_diagonal(self, dst):
        _check(lib().mf_jacobi_inverse_diagonal(self._h, self.operator._r(dst)))

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().mf_jacobi_destroy(self._h)
                self._h = None
        except Exception:
            pass


class LinearSolver:
    """Mirror of MeltPoolDG::LinearSolver::solve (linear_solver.hpp:22-134).  `preconditioner` is
    "Diagonal" (Jacobi, recomputed in the call), "Identity", or an updated
    JacobiPreconditioner."""

    @staticmethod
    def solve(operator, solution, rhs, solver="CG", preconditioner="Diagonal",
              max_iterations=10000, rel_tolerance=1e-12, abs_tolerance=1e-20, restart=30,
              precision="fp64", inner_rel_tolerance=0.0, max_outer=50, raise_on_failure=True,
              force_re_orthogonalization=False):
        d = SolverDesc()
        d.force_re_orthogonalization = int(bool(force_re_orthogonalization))
        d.solver = {"CG": SOLVER_CG, "GMRES": SOLVER_GMRES}[solver]
        d.restart = restart
        d.max_iterations = max_iterations
        d.abs_tol = abs_tolerance
        d.rel_tol = rel_tolerance
        d.precision = {"fp64": PRECISION_FP64, "fp32_refined": PRECISION_FP32_REFINED}[precision]
        d.inner_rel_tol = inner_rel_tolerance
        d.max_outer = max_outer
        rep = SolveReport()
        if isinstance(preconditioner, JacobiPreconditioner):
            d.preconditioner = PRECOND_JACOBI
            code = lib().mf_solve_preconditioned(operator._h, C.byref(d), preconditioner._h,
                                                 operator._r(solution), operator._r(rhs),
                                                 C.byref(rep))
        else:
            d.preconditioner = {"Identity": PRECOND_IDENTITY,
                                "Diagonal": PRECOND_JACOBI}[preconditioner]
            code = lib().mf_solve(operator._h, C.byref(d), operator._r(solution),
                                  operator._r(rhs), C.byref(rep))
        if code != MF_OK and (raise_on_failure or code not in (MF_NO_CONVERGENCE, MF_BREAKDOWN)):
            _check(code)
        return rep.as_dict()


class NormalVectorOperation:
    """NormalVectorOperation (source/level_set/normal_vector_operation.cpp:36-150) on the device:
    reinit() updates the Jacobi preconditioner (as the reference does at reinit), solve() builds
    the right-hand side from the level set and solves the block system with CG from the current
    solution (PredictorType::none).  Linear solver defaults of NormalVectorData (CG, Diagonal,
    rel 1e-12, abs 1e-20)."""

    def __init__(self, scratch, level_set, filter_parameter=2.0, narrow_band=False,
                 narrow_band_threshold=0.9999999, compute_normalized_vector=False,
                 solver="CG", rel_tolerance=1e-12, abs_tolerance=1e-20, max_iterations=10000,
                 precision="fp64"):
        import torch
        self.scratch = scratch
        self.level_set = level_set
        self.operator = NormalVectorOperator(scratch, filter_parameter, narrow_band,
                                             narrow_band_threshold, compute_normalized_vector,
                                             level_set=level_set)
        n = self.operator.n_rows
        self.solution = torch.zeros(n, dtype=torch.float64, device=level_set.device)
        self.rhs = torch.zeros(n, dtype=torch.float64, device=level_set.device)
        self.preconditioner = JacobiPreconditioner(self.operator)
        self.solver = dict(solver=solver, rel_tolerance=rel_tolerance,
                           abs_tolerance=abs_tolerance, max_iterations=max_iterations,
                           precision=precision)
        self.last_report = None
        self.reinit()

    def reinit(self):
        self.operator.set_level_set(self.level_set)
        self.preconditioner.update()

    def solve(self):
        # the level set's content may have changed: refresh the narrow band's FP32 copy
        self.operator.set_level_set(self.level_set)
        self.rhs.zero_()
        self.operator.create_rhs(self.rhs, self.level_set)
        self.last_report = LinearSolver.solve(self.operator, self.solution, self.rhs,
                                              preconditioner=self.preconditioner, **self.solver)
        return self.last_report["iterations"]

    def get_solution_normal_vector(self):
        return self.solution


class CurvatureOperation:
    """CurvatureOperation (source/level_set/curvature_operation.cpp:13-115, 198-224) on the
    device: solve() computes the normal vector, then the curvature
    kappa = -div(n / |n|) projected with the filter."""

    def __init__(self, scratch, level_set, filter_parameter=2.0, narrow_band=False,
                 narrow_band_threshold=0.9999999, normal_filter_parameter=2.0,
                 normal_narrow_band=False, normal_narrow_band_threshold=0.9999999,
                 compute_normalized_vector=False, solver="CG", rel_tolerance=1e-12,
                 abs_tolerance=1e-20, max_iterations=10000, precision="fp64"):
        import torch
        self.scratch = scratch
        self.level_set = level_set
        self.normal_vector_operation = NormalVectorOperation(
            scratch, level_set, normal_filter_parameter, normal_narrow_band,
            normal_narrow_band_threshold, compute_normalized_vector, solver, rel_tolerance,
            abs_tolerance, max_iterations, precision)
        self.operator = CurvatureOperator(scratch, filter_parameter, narrow_band,
                                          narrow_band_threshold, level_set=level_set)
        n = self.operator.n_rows
        self.solution = torch.zeros(n, dtype=torch.float64, device=level_set.device)
        self.rhs = torch.zeros(n, dtype=torch.float64, device=level_set.device)
        self.preconditioner = JacobiPreconditioner(self.operator)
        self.solver = dict(solver=solver, rel_tolerance=rel_tolerance,
                           abs_tolerance=abs_tolerance, max_iterations=max_iterations,
                           precision=precision)
        self.last_report = None
        self.reinit()

    def reinit(self):
        self.operator.set_level_set(self.level_set)
        self.preconditioner.update()
        self.normal_vector_operation.reinit()

    def update_normal_vector(self):
        return self.normal_vector_operation.solve()

    def solve(self):
        self.normal_vector_operation.solve()
        self.operator.set_level_set(self.level_set)
        self.rhs.zero_()
        self.operator.create_rhs(self.rhs, self.normal_vector_operation.get_solution_normal_vector())
        self.last_report = LinearSolver.solve(self.operator, self.solution, self.rhs,
                                              preconditioner=self.preconditioner, **self.solver)
        return self.last_report["iterations"]

    def get_curvature(self):
        return self.solution

    def get_normal_vector(self):
        return self.normal_vector_operation.get_solution_normal_vector()


class ReinitializationOperation:
    """Olsson reinitialization on the device: ReinitializationHyperbolicCGOperation
    (source/level_set/reinitialization_hyperbolic_CG_operation.cpp:80-280) with the pseudo-time
    loop of LevelSetOperation::do_reinitialization (source/level_set/level_set_operation.cpp:
    507-590).  Each pseudo step assembles the right-hand side from the current level set, solves
    (1/dt M + D) delta = rhs with Jacobi-CG (the previous increment as initial guess), adds delta
    to the level set and stops early when |delta| / |psi| drops below `tolerance`.  The normal
    vector is computed from the initial level set, and after every step when
    `update_normal_vector` is set (the reference does so in its first reinitialization)."""

    def __init__(self, scratch, level_set, thickness_type=THICKNESS_PROPORTIONAL,
                 thickness_value=0.5, tangential_diffusion_factor=0.0, max_n_steps=5,
                 pseudo_time_step_size=-1.0, pseudo_time_step_factor=1.0, tolerance=2.2250738585072014e-308,
                 normal_filter_parameter=2.0, update_normal_vector=False, solver="CG",
                 rel_tolerance=1e-12, abs_tolerance=1e-20, max_iterations=10000,
                 precision="fp64"):
        import torch
        self.scratch = scratch
        self.level_set = level_set
        self.normal_vector_operation = NormalVectorOperation(
            scratch, level_set, normal_filter_parameter, solver=solver,
            rel_tolerance=rel_tolerance, abs_tolerance=abs_tolerance,
            max_iterations=max_iterations, precision=precision)
        self.operator = ReinitializationOperator(
            scratch, thickness_type, thickness_value, tangential_diffusion_factor,
            normal=self.normal_vector_operation.get_solution_normal_vector())
        if pseudo_time_step_size <= 0.0:
            # time increment from epsilon of the smallest cell (level_set_operation.cpp:528-532)
            pseudo_time_step_size = pseudo_time_step_factor * self.operator.min_epsilon
        self.dt = float(pseudo_time_step_size)
        self.operator.reset_time_increment(self.dt)
        n = scratch.n_dofs
        self.delta = torch.zeros(n, dtype=torch.float64, device=level_set.device)
        self.rhs = torch.zeros(n, dtype=torch.float64, device=level_set.device)
        self.preconditioner = JacobiPreconditioner(self.operator)
        self.max_n_steps = max_n_steps
        self.tolerance = tolerance
        self.update_normal_vector = update_normal_vector
        self.solver = dict(solver=solver, rel_tolerance=rel_tolerance,
                           abs_tolerance=abs_tolerance, max_iterations=max_iterations,
                           precision=precision)
        self.max_change_level_set = None
        self.iterations = []

    def _normal(self):
        self.normal_vector_operation.reinit()
        self.normal_vector_operation.solve()
        self.operator.set_normal_vector(self.normal_vector_operation.get_solution_normal_vector())

    def solve(self):
        """One pseudo time step (ReinitializationHyperbolicCGOperation::solve)."""
        self.rhs.zero_()
        self.operator.create_rhs(self.rhs, self.level_set)
        self.preconditioner.update()
        rep = LinearSolver.solve(self.operator, self.delta, self.rhs,
                                 preconditioner=self.preconditioner, **self.solver)
        self.iterations.append(rep["iterations"])
        self.level_set += self.delta
        self.max_change_level_set = float(self.delta.norm() / self.level_set.norm())
        return rep["iterations"]

    def reinitialize(self):
        """The pseudo-time loop; returns the number of steps done."""
        self._normal()
        steps = 0
        for _ in range(self.max_n_steps):
            self.solve()
            steps += 1
            if self.max_change_level_set < self.tolerance:
                break
            if self.update_normal_vector:
                self._normal()
        return steps


def write_vtu(scratch, filename, fields):
    """VTU output of device fields: fields maps a name to a tensor of n_dofs (scalar) or
    3 n_dofs ([3][n_dofs], a vector) entries (mf_write_vtu)."""
    names = list(fields)
    n = scratch.n_dofs
    ncomp = []
    for k in names:
        m = fields[k].numel()
        if m not in (n, 3 * n):
            raise ValueError(f"field {k}: {m} entries, expected {n} or {3 * n}")
        ncomp.append(1 if m == n else 3)
    arr_n = (C.c_char_p * len(names))(*[k.encode() for k in names])
    arr_c = (C.c_int * len(names))(*ncomp)
    arr_f = (C.c_void_p * len(names))(*[scratch._v(fields[k], fields[k].numel()) for k in names])
    _check(lib().mf_write_vtu(scratch._h, str(filename).encode(), len(names), arr_n, arr_c,
                              arr_f))


def restart_save(scratch, path, vectors):
    """Restart file with the mesh description and the given device vectors (mf_restart_save)."""
    arr_v = (C.c_void_p * len(vectors))(*[scratch._v(v, v.numel()) for v in vectors])
    arr_l = (C.c_int64 * len(vectors))(*[v.numel() for v in vectors])
    _check(lib().mf_restart_save(scratch._h, str(path).encode(), len(vectors), arr_v, arr_l))


def restart_load(path, vectors=None, stream=None):
    """Rebuild the scratch of a restart file (same DoF numbering) and read its vectors into the
    given device tensors (mf_restart_load_scratch / mf_restart_load_vectors)."""
    import numpy as np
    h = C.c_void_p()
    _check(lib().mf_restart_load_scratch(str(path).encode(), C.byref(h)))
    raw = np.fromfile(str(path), dtype=np.uint8, count=4096)
    deg, adaptive = np.frombuffer(raw[8:16].tobytes(), dtype=np.int32)
    ncell = tuple(int(x) for x in np.frombuffer(raw[16:28].tobytes(), dtype=np.int32))
    lo = tuple(np.frombuffer(raw[28:52].tobytes(), dtype=np.float64))
    hi = tuple(np.frombuffer(raw[52:76].tobytes(), dtype=np.float64))
    n_cells = int(np.frombuffer(raw[76:84].tobytes(), dtype=np.int64)[0])
    s = object.__new__(AdaptiveScratchData if adaptive else ScratchData)
    s.degree, s.ncell, s.lo, s.hi = int(deg), ncell, lo, hi
    if adaptive:
        s.adaptive = True
        s.cells = np.fromfile(str(path), dtype=np.int32, count=4 * n_cells,
                              offset=84).reshape(-1, 4)
    s._h = h
    s.device = _current_device()
    s.n_dofs = int(lib().mf_scratch_n_dofs(h))
    s.n_cells = int(lib().mf_scratch_n_cells(h))
    if stream is not None:
        s.set_stream(stream)
    if vectors:
        arr_v = (C.c_void_p * len(vectors))(*[s._v(v, v.numel()) for v in vectors])
        arr_l = (C.c_int64 * len(vectors))(*[v.numel() for v in vectors])
        _check(lib().mf_restart_load_vectors(s._h, str(path).encode(), len(vectors), arr_v,
                                             arr_l))
    return s

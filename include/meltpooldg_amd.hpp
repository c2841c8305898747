// meltpooldg_amd.hpp -- header-only C++ host layer over the C ABI (meltpooldg_amd.h).
//
// Mirrors the reference's operator / solver interfaces so that a C++ caller written against
// MeltPoolDG reads the same:
//   OperatorMatrixFree<dim,number>::{vmult, create_rhs, compute_inverse_diagonal_from_matrixfree}
//   OperatorBase::{reset_time_increment, reset_dof_index}      (include/meltpooldg/core/operator_base.hpp:21-84,189-283)
//   LinearSolver::solve                                        (include/meltpooldg/linear_algebra/linear_solver.hpp:22-134)
//   JacobiPreconditioner::{update, vmult}                      (include/meltpooldg/linear_algebra/preconditioner_jacobi.hpp:29-89)
//   NormalVectorOperator / CurvatureOperator                   (source/level_set/normal_vector_operator.cpp,
//                                                               source/level_set/curvature_operator.cpp)
//   OlssonOperator (reinitialization)                          (source/level_set/olsson_CG_operator.cpp)
// Errors are thrown as mf::Exception (the reference throws dealii::ExceptionBase) and
// mf::NoConvergence (dealii::SolverControl::NoConvergence).  DeviceVector is the subset of
// LinearAlgebra::distributed::Vector the solver layer needs, resident in HBM.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "meltpooldg_amd.h"

namespace mf {

class Exception : public std::runtime_error {
 public:
  Exception(int code, const std::string &what) : std::runtime_error(what), code(code) {}
  int code;
};

class NoConvergence : public Exception {
 public:
  using Exception::Exception;
};

inline void check(int code)
{
  if (code == MF_OK) return;
  if (code == MF_NO_CONVERGENCE) throw NoConvergence(code, mf_last_error());
  throw Exception(code, mf_last_error());
}

inline void hip_check(hipError_t e)
{
  if (e != hipSuccess) throw Exception(MF_ERR_HIP, hipGetErrorString(e));
}

// Device-resident vector of n entries (HBM).
template <typename Number>
class DeviceVector {
 public:
  DeviceVector() = default;
  explicit DeviceVector(std::size_t n) { reinit(n); }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  DeviceVector(DeviceVector &&o) noexcept : ptr_(std::exchange(o.ptr_, nullptr)), n_(o.n_) {}
  ~DeviceVector()
  {
    if (ptr_) (void)hipFree(ptr_);
  }
  void reinit(std::size_t n)
  {
    if (ptr_) hip_check(hipFree(ptr_));
    n_ = n;
    hip_check(hipMalloc(&ptr_, sizeof(Number) * (n ? n : 1)));
    hip_check(hipMemset(ptr_, 0, sizeof(Number) * (n ? n : 1)));
  }
  std::size_t size() const { return n_; }
  Number *data() { return ptr_; }
  const Number *data() const { return ptr_; }
  void copy_from_host(const Number *h) { hip_check(hipMemcpy(ptr_, h, sizeof(Number) * n_, hipMemcpyHostToDevice)); }
  void copy_to_host(Number *h) const { hip_check(hipMemcpy(h, ptr_, sizeof(Number) * n_, hipMemcpyDeviceToHost)); }
  // device-to-device copy of a vector of the same size
  void copy_from(const DeviceVector &o)
  {
    if (o.n_ != n_) throw Exception(MF_ERR_ARG, "DeviceVector::copy_from: sizes differ");
    hip_check(hipMemcpy(ptr_, o.ptr_, sizeof(Number) * n_, hipMemcpyDeviceToDevice));
  }
  std::vector<Number> to_host() const
  {
    std::vector<Number> v(n_);
    copy_to_host(v.data());
    return v;
  }

 private:
  Number *ptr_ = nullptr;
  std::size_t n_ = 0;
};

// ScratchData + MatrixFree setup of a structured hex mesh (source/core/scratch_data.cpp:194-293).
class ScratchData {
 public:
  ScratchData(int degree, int nx, int ny, int nz, const double lo[3] = nullptr,
              const double hi[3] = nullptr, hipStream_t stream = nullptr)
  {
    mf_mesh_desc d{};
    d.dim = 3;
    d.degree = degree;
    d.ncell[0] = nx;
    d.ncell[1] = ny;
    d.ncell[2] = nz;
    for (int i = 0; i < 3; ++i) {
      d.lo[i] = lo ? lo[i] : 0.0;
      d.hi[i] = hi ? hi[i] : 1.0;
    }
    check(mf_scratch_create(&d, &h_));
    check(mf_set_stream(h_, stream));
  }
  ScratchData(const ScratchData &) = delete;
  ScratchData &operator=(const ScratchData &) = delete;
  ~ScratchData() { mf_scratch_destroy(h_); }
  std::size_t n_dofs() const { return std::size_t(mf_scratch_n_dofs(h_)); }
  std::size_t n_cells() const { return std::size_t(mf_scratch_n_cells(h_)); }
  void set_dirichlet_mask(const std::vector<std::uint8_t> &mask) { check(mf_scratch_set_dirichlet_mask(h_, mask.data())); }
  template <typename Number>
  void initialize_dof_vector(DeviceVector<Number> &v) const { v.reinit(n_dofs()); }
  mf_scratch *handle() const { return h_; }

 private:
  mf_scratch *h_ = nullptr;
};

// OperatorMatrixFree<3,double> over device vectors.
class OperatorMatrixFree {
 public:
  using VectorType = DeviceVector<double>;
  OperatorMatrixFree(const OperatorMatrixFree &) = delete;
  OperatorMatrixFree &operator=(const OperatorMatrixFree &) = delete;
  virtual ~OperatorMatrixFree() { mf_operator_destroy(h_); }

  void reset_time_increment(double dt) { check(mf_operator_reset_time_increment(h_, dt)); }
  void reset_dof_index(unsigned int dof_idx) { check(mf_operator_reset_dof_index(h_, int(dof_idx))); }
  void vmult(VectorType &dst, const VectorType &src) const { check(mf_operator_vmult(h_, dst.data(), src.data())); }
  void vmult(DeviceVector<float> &dst, const DeviceVector<float> &src) const
  {
    check(mf_operator_vmult_f32(h_, dst.data(), src.data()));
  }
  void compute_inverse_diagonal_from_matrixfree(VectorType &diag) const
  {
    check(mf_operator_compute_inverse_diagonal(h_, diag.data()));
  }
  void create_rhs(VectorType &dst, const VectorType &src, const VectorType *q_s = nullptr) const
  {
    check(mf_operator_create_rhs(h_, dst.data(), src.data(), q_s ? q_s->data() : nullptr));
  }
  mf_operator *handle() const { return h_; }

 protected:
  OperatorMatrixFree() = default;
  mf_operator *h_ = nullptr;
};

// HeatDiffuseMultiPhaseOperator (source/heat/heat_diffuse_operator.cpp) on the device.
class HeatDiffuseOperator : public OperatorMatrixFree {
 public:
  HeatDiffuseOperator(ScratchData &s, const mf_heat_params &p) { check(mf_heat_operator_create(s.handle(), &p, &h_)); }
  // the linearisation vectors the reference wires in at heat_diffuse_operation.cpp:55-70
  void set_fields(const VectorType *T, const VectorType *velocity = nullptr,
                  const VectorType *k_cell = nullptr)
  {
    check(mf_operator_set_fields(h_, T ? T->data() : nullptr, velocity ? velocity->data() : nullptr,
                                 k_cell ? k_cell->data() : nullptr));
  }
  // convection / radiation / Neumann sides (HeatData::convection, ::radiation and the boundary
  // id lists of the reference); nullptr removes them
  void set_boundary(const mf_heat_boundary *b) { check(mf_operator_set_boundary(h_, b)); }
};

// AdvectionDiffusionOperator (source/level_set/advection_diffusion_operator.cpp) on the device.
class AdvectionDiffusionOperator : public OperatorMatrixFree {
 public:
  AdvectionDiffusionOperator(ScratchData &s, const mf_advec_params &p)
  {
    check(mf_advection_operator_create(s.handle(), &p, &h_));
  }
  void set_advection_velocity(const VectorType &velocity)
  {
    check(mf_operator_set_fields(h_, nullptr, velocity.data(), nullptr));
  }
};

// NormalVectorOperator / CurvatureOperator (source/level_set/normal_vector_operator.cpp,
// curvature_operator.cpp): the filtered projections of the level-set gradient and of the
// curvature.  The normal vector operator works on block vectors [3][n_dofs] (n_rows()).
class LevelSetProjectionOperator : public OperatorMatrixFree {
 public:
  void set_level_set(const VectorType *level_set)
  {
    check(mf_operator_set_level_set(h_, level_set ? level_set->data() : nullptr));
  }
  std::size_t n_rows() const { return std::size_t(mf_operator_n_rows(h_)); }
  // create_rhs(dst, src): normal vector: src = level set, dst [3][n_dofs];
  // curvature: src = normal vector [3][n_dofs], dst [n_dofs]
  void create_rhs(VectorType &dst, const VectorType &src) const
  {
    check(mf_operator_create_rhs(h_, dst.data(), src.data(), nullptr));
  }
};

class NormalVectorOperator : public LevelSetProjectionOperator {
 public:
  NormalVectorOperator(ScratchData &s, const mf_levelset_projection_params &p)
  {
    check(mf_normal_vector_operator_create(s.handle(), &p, &h_));
  }
};

class CurvatureOperator : public LevelSetProjectionOperator {
 public:
  CurvatureOperator(ScratchData &s, const mf_levelset_projection_params &p)
  {
    check(mf_curvature_operator_create(s.handle(), &p, &h_));
  }
  double normal_tolerance() const { return mf_operator_normal_tolerance(h_); }
};

// OlssonOperator (source/level_set/olsson_CG_operator.cpp): the hyperbolic CG reinitialization
// step (1/dt M + D) delta = rhs with the normal vector of a NormalVectorOperator solve.
class ReinitializationOperator : public OperatorMatrixFree {
 public:
  ReinitializationOperator(ScratchData &s, const mf_reinit_params &p)
  {
    check(mf_reinit_operator_create(s.handle(), &p, &h_));
  }
  void set_normal_vector(const VectorType *normal)
  {
    check(mf_operator_set_normal_vector(h_, normal ? normal->data() : nullptr));
  }
  // create_rhs(dst, level set): adds the compressive flux and diffusion terms
  void create_rhs(VectorType &dst, const VectorType &level_set) const
  {
    check(mf_operator_create_rhs(h_, dst.data(), level_set.data(), nullptr));
  }
  double min_epsilon() const { return mf_reinit_min_epsilon(h_); }
};

// HeatCutOperator without two_phase (source/heat/heat_cut_operator.cpp), FE_Q1: the tangent
// over the liquid part {phi > 0} plus the ghost penalty.  set_level_set takes the nodal level
// set on the host (classification, sub-boxes, faces); init_time_advance sets the ost factors;
// create_rhs(dst, T_old) adds -R(T) with T from set_temperature.
class HeatCutOperator : public OperatorMatrixFree {
 public:
  HeatCutOperator(ScratchData &s, const mf_heat_cut_params &p)
  {
    check(mf_heat_cut_operator_create(s.handle(), &p, &h_));
  }
  void set_level_set(const double *phi_host) { check(mf_heat_cut_set_level_set(h_, phi_host)); }
  void init_time_advance(double dt) { reset_time_increment(dt); }
  void set_temperature(const VectorType *T)
  {
    check(mf_operator_set_fields(h_, T ? T->data() : nullptr, nullptr, nullptr));
  }
  // CutUtil::compute_cut_norm (L2)
  double compute_cut_norm(const VectorType &u)
  {
    double r = 0.0;
    check(mf_heat_cut_norm_l2(h_, u.data(), &r));
    return r;
  }
  // HeatCutOperation::solve for a fixed interface: T holds the predictor, then the solution;
  // throws NoConvergence (ExcNewtonDidNotConverge)
  mf_newton_report solve(VectorType &T, const VectorType &T_old, const mf_solver_desc &solver,
                         const mf_newton_desc &newton, const VectorType *dirichlet_values = nullptr)
  {
    mf_newton_report rep{};
    check(mf_heat_cut_step(h_, T.data(), T_old.data(),
                           dirichlet_values ? dirichlet_values->data() : nullptr, &solver,
                           &newton, &rep));
    return rep;
  }
};

// JacobiPreconditioner (include/meltpooldg/linear_algebra/preconditioner_jacobi.hpp:29-89)
class JacobiPreconditioner {
 public:
  explicit JacobiPreconditioner(const OperatorMatrixFree &op) { check(mf_jacobi_create(op.handle(), &h_)); }
  JacobiPreconditioner(const JacobiPreconditioner &) = delete;
  JacobiPreconditioner &operator=(const JacobiPreconditioner &) = delete;
  ~JacobiPreconditioner() { mf_jacobi_destroy(h_); }
  void update() { check(mf_jacobi_update(h_)); }
  void vmult(DeviceVector<double> &dst, const DeviceVector<double> &src) const
  {
    check(mf_jacobi_vmult(h_, dst.data(), src.data()));
  }
  mf_jacobi *handle() const { return h_; }

 private:
  mf_jacobi *h_ = nullptr;
};

// LinearSolverData (include/meltpooldg/linear_algebra/linear_solver_data.hpp:37-51)
struct LinearSolverData {
  int solver_type = MF_SOLVER_GMRES;
  int preconditioner_type = MF_PRECOND_JACOBI;
  unsigned int max_iterations = 10000;
  double rel_tolerance = 1e-12;
  double abs_tolerance = 1e-20;
  int precision = MF_PRECISION_FP64;
  double inner_rel_tolerance = 0.0;  // 0: sqrt(rel_tolerance) in [1e-5, 1e-2]
};

struct LinearSolver {
  // returns the iteration count (SolverControl::last_step()); throws NoConvergence
  static int solve(OperatorMatrixFree &op, DeviceVector<double> &solution,
                   const DeviceVector<double> &rhs, const LinearSolverData &data,
                   mf_solve_report *report = nullptr)
  {
    mf_solver_desc d{};
    d.solver = data.solver_type;
    d.restart = 30;
    d.max_iterations = int(data.max_iterations);
    d.abs_tol = data.abs_tolerance;
    d.rel_tol = data.rel_tolerance;
    d.preconditioner = data.preconditioner_type;
    d.precision = data.precision;
    d.inner_rel_tol = data.inner_rel_tolerance;
    mf_solve_report rep{};
    const int code = mf_solve(op.handle(), &d, solution.data(), rhs.data(), &rep);
    if (report) *report = rep;
    check(code);
    return rep.iterations;
  }
  // with a preconditioner updated by the caller (LinearSolver::solve(..., preconditioner))
  static int solve(OperatorMatrixFree &op, DeviceVector<double> &solution,
                   const DeviceVector<double> &rhs, const LinearSolverData &data,
                   const JacobiPreconditioner &preconditioner, mf_solve_report *report = nullptr)
  {
    mf_solver_desc d{};
    d.solver = data.solver_type;
    d.restart = 30;
    d.max_iterations = int(data.max_iterations);
    d.abs_tol = data.abs_tolerance;
    d.rel_tol = data.rel_tolerance;
    d.preconditioner = MF_PRECOND_JACOBI;
    d.precision = data.precision;
    d.inner_rel_tol = data.inner_rel_tolerance;
    mf_solve_report rep{};
    const int code = mf_solve_preconditioned(op.handle(), &d, preconditioner.handle(),
                                             solution.data(), rhs.data(), &rep);
    if (report) *report = rep;
    check(code);
    return rep.iterations;
  }
};

// The heat time loop of the mp-heat-transfer application (HeatTransferApplication::run,
// applications/mp-heat-transfer/heat_transfer_application.cpp:52-120) on the device: each
// step() moves the laser, recomputes the volumetric source, applies the predictor and runs
// the Newton solve; the solution history stays in HBM.
class HeatTimeStepper {
 public:
  HeatTimeStepper(HeatDiffuseOperator &op, const mf_heat_stepper_desc &d,
                  const DeviceVector<double> &T0)
  {
    check(mf_heat_stepper_create(op.handle(), &d, T0.data(), &h_));
  }
  HeatTimeStepper(const HeatTimeStepper &) = delete;
  HeatTimeStepper &operator=(const HeatTimeStepper &) = delete;
  ~HeatTimeStepper() { mf_heat_stepper_destroy(h_); }
  // throws NoConvergence (ExcHeatTransferNoConvergence)
  mf_newton_report step(double dt, const DeviceVector<double> *dirichlet_values = nullptr)
  {
    mf_newton_report rep{};
    check(mf_heat_stepper_step(h_, dt, dirichlet_values ? dirichlet_values->data() : nullptr,
                               &rep));
    return rep;
  }
  // device pointer to the current temperature (valid until the next step)
  const double *solution() const
  {
    const double *T = nullptr;
    check(mf_heat_stepper_state(h_, &T, nullptr, nullptr, nullptr, nullptr));
    return T;
  }
  double time() const
  {
    double t = 0.0;
    check(mf_heat_stepper_state(h_, nullptr, nullptr, nullptr, &t, nullptr));
    return t;
  }

 private:
  mf_heat_stepper *h_ = nullptr;
};

}  // namespace mf

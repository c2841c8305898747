// One time step of the heat-transfer application on the device.
//
// Reference functions replaced (heat-diffuse operator with a volumetric laser):
//   HeatTransferApplication::run, loop body
//     (applications/mp-heat-transfer/heat_transfer_application.cpp:52-120)
//   TimeIterator::compute_next_time_increment (source/time_integration/time_iterator.cpp:44-68)
//   LaserOperation::move_laser (source/heat/laser_operation.cpp:236-250) and
//     compute_heat_source, volumetric model (:316-333)
//   HeatDiffuseOperation::init_time_advance + solve
//     (source/heat/heat_diffuse_operation.cpp:214-290) with Predictor::vmult
//     (include/meltpooldg/linear_algebra/predictor.hpp:119-161,
//     source/linear_algebra/predictor_linear.cpp:76-96)
//
// The stepper owns the solution history (current and recent old temperature) and the heat
// source as device vectors; a step never copies a field to the host.  The history commit of
// the predictor is a rotation of three device buffers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "mf_internal.hpp"

using namespace mf;

struct mf_heat_stepper {
  mf_operator *op = nullptr;
  mf_heat_stepper_desc d{};
  double *T = nullptr;      // current solution
  double *T_old = nullptr;  // recent old solution
  double *buf = nullptr;    // predictor buffer
  double *q = nullptr;      // heat source
  double *rhs = nullptr;    // least-squares predictor: its right-hand side
  double time = 0.0;
  double position[3] = {0.0, 0.0, 0.0};
  int n_steps = 0;
};

extern "C" {

int mf_laser_heat_source(mf_scratch *s, double power, double radius, const double position[3],
                         double *q_s);
int mf_predictor_linear_extrapolation(mf_scratch *s, double *dst, const double *x_n,
                                      const double *x_nm1, double tau);
int mf_operator_reset_time_increment(mf_operator *op, double dt);
int mf_distribute_constraints(mf_scratch *s, double *x);
int mf_operator_set_fields(mf_operator *op, const double *T, const double *vel,
                           const double *k_cell);
int mf_operator_set_temperature_old(mf_operator *op, const double *T_old);
int mf_operator_create_rhs(mf_operator *op, double *dst, const double *src, const double *q_s);

int mf_heat_stepper_create(mf_operator *op, const mf_heat_stepper_desc *d, const double *T0,
                           mf_heat_stepper **out)
{
  MF_REQUIRE(op && d && T0 && out, MF_ERR_ARG, "null argument");
  MF_REQUIRE(op->kind == 0, MF_ERR_ARG, "the heat stepper needs a heat operator");
  MF_REQUIRE(d->predictor == MF_PREDICTOR_NONE || d->predictor == MF_PREDICTOR_ZERO ||
                 d->predictor == MF_PREDICTOR_LINEAR_EXTRAPOLATION ||
                 d->predictor == MF_PREDICTOR_LEAST_SQUARES_PROJECTION,
             MF_ERR_ARG, "unknown predictor type");
  MF_REQUIRE(!d->laser || d->radius > 0, MF_ERR_ARG, "the laser radius must be positive");
  const size_t bytes = sizeof(double) * size_t(op->s->mesh.n_dofs);
  hipStream_t st = op->s->stream;
  auto *h = new mf_heat_stepper();
  h->op = op;
  h->d = *d;
  h->time = d->start_time;
  for (int k = 0; k < 3; ++k) h->position[k] = d->start[k];
  // LaserData::get_scan_direction returns the direction normalised (laser_data.cpp:246-250)
  if (d->laser && d->scan_speed != 0.0) {
    const double nrm = std::sqrt(d->scan_direction[0] * d->scan_direction[0] +
                                 d->scan_direction[1] * d->scan_direction[1] +
                                 d->scan_direction[2] * d->scan_direction[2]);
    if (!(nrm > 0)) {
      delete h;
      set_error("the scan direction must be nonzero");
      return MF_ERR_ARG;
    }
    for (int k = 0; k < 3; ++k) h->d.scan_direction[k] = d->scan_direction[k] / nrm;
  }
  double **bufs[5] = {&h->T, &h->T_old, &h->buf, &h->q, &h->rhs};
  const int n_bufs = d->predictor == MF_PREDICTOR_LEAST_SQUARES_PROJECTION ? 5 : 4;
  for (int i = 0; i < n_bufs; ++i) {
    double **b = bufs[i];
    if (hipMalloc(reinterpret_cast<void **>(b), bytes) != hipSuccess) {
      (void)mf_heat_stepper_destroy(h);
      set_error("hipMalloc failed for the heat stepper's vectors");
      return MF_ERR_HIP;
    }
  }
  int rc = MF_OK;
  if (hipMemcpyAsync(h->T, T0, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
      hipMemsetAsync(h->q, 0, bytes, st) != hipSuccess) {
    set_error("device copy of the initial temperature failed");
    rc = MF_ERR_HIP;
  }
  // on an adaptive mesh the history holds constraint-consistent fields
  if (rc == MF_OK && op->s->ad) rc = mf_distribute_constraints(op->s, h->T);
  if (rc == MF_OK &&
      (hipMemcpyAsync(h->T_old, h->T, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess ||
       hipStreamSynchronize(st) != hipSuccess)) {
    set_error("device copy of the initial temperature failed");
    rc = MF_ERR_HIP;
  }
  if (rc != MF_OK) {
    (void)mf_heat_stepper_destroy(h);
    return rc;
  }
  *out = h;
  return MF_OK;
}

int mf_heat_stepper_destroy(mf_heat_stepper *h)
{
  if (!h) return MF_OK;
  double *bufs[5] = {h->T, h->T_old, h->buf, h->q, h->rhs};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  delete h;
  return MF_OK;
}

int mf_heat_stepper_step(mf_heat_stepper *h, double dt, const double *dirichlet_values,
                         mf_newton_report *report)
{
  MF_REQUIRE(h, MF_ERR_ARG, "null stepper");
  MF_REQUIRE(dt > 0 && std::isfinite(dt), MF_ERR_ARG, "the time increment must be positive");
  mf_operator *op = h->op;
  mf_scratch *s = op->s;
  const int64_t n = s->mesh.n_dofs;
  const size_t bytes = sizeof(double) * size_t(n);
  hipStream_t st = s->stream;
  int rc;
  // TimeIterator: the old increment is 0 at the first step; from the second step on it is set
  // after the new increment is chosen, i.e. equal to it (time_iterator.cpp:58-62)
  ++h->n_steps;
  const double dt_old = h->n_steps > 1 ? dt : 0.0;
  h->time += dt;
  // LaserOperation::move_laser, then the volumetric source into the zeroed heat source
  if (h->d.laser) {
    if (h->d.scan_speed != 0.0)
      for (int k = 0; k < 3; ++k) h->position[k] += h->d.scan_direction[k] * h->d.scan_speed * dt;
    if ((rc = mf_laser_heat_source(s, h->d.power, h->d.radius, h->position, h->q))) return rc;
  }
  // HeatDiffuseOperation::init_time_advance
  if ((rc = mf_operator_reset_time_increment(op, dt))) return rc;
  switch (h->d.predictor) {
    case MF_PREDICTOR_NONE:
      MF_HIP_TRY(hipMemcpyAsync(h->buf, h->T, bytes, hipMemcpyDeviceToDevice, st));
      break;
    case MF_PREDICTOR_ZERO:
      MF_HIP_TRY(hipMemsetAsync(h->buf, 0, bytes, st));
      break;
    case MF_PREDICTOR_LEAST_SQUARES_PROJECTION:
      // Predictor::vmult with n_old_solution_vectors = 2 (the stepper keeps two solutions):
      // the projection once two predictor calls have passed, the current solution before.
      // (The reference's fallback swaps the current solution with the zeroed predictor buffer
      // before the history commit, which leaves a zero old solution; the stepper keeps the
      // current solution as the old one, as for the other predictors.)
      if (h->n_steps - 1 >= 2) {
        // HeatDiffuseOperation::init_time_advance: rhs = create_rhs(current solution) with the
        // operator linearised at the current solution and the new increment and heat source
        if ((rc = mf_operator_set_fields(op, h->T, op->vel, op->k_cell))) return rc;
        if (op->heat.law == MF_LAW_MUSHY && (rc = mf_operator_set_temperature_old(op, h->T_old)))
          return rc;
        MF_HIP_TRY(hipMemsetAsync(h->rhs, 0, bytes, st));
        if ((rc = mf_operator_create_rhs(op, h->rhs, h->T, h->d.laser ? h->q : nullptr)))
          return rc;
        const double *olds[2] = {h->T, h->T_old};
        if ((rc = mf_predictor_least_squares(op, 2, olds, h->rhs, h->buf, nullptr))) return rc;
      } else {
        MF_HIP_TRY(hipMemcpyAsync(h->buf, h->T, bytes, hipMemcpyDeviceToDevice, st));
      }
      break;
    default:
      if (std::fabs(dt_old) < 1e-12) {
        MF_HIP_TRY(hipMemcpyAsync(h->buf, h->T, bytes, hipMemcpyDeviceToDevice, st));
      } else if ((rc = mf_predictor_linear_extrapolation(s, h->buf, h->T, h->T_old, dt / dt_old))) {
        return rc;
      }
      break;
  }
  // SolutionHistory::commit_old_solutions + current.swap(predictor)
  double *const prev_old = h->T_old;
  h->T_old = h->T;
  h->T = h->buf;
  h->buf = prev_old;
  // HeatDiffuseOperation::solve: the Newton solve from the predictor
  return mf_heat_newton_step(op, h->T, h->T_old, h->d.laser ? h->q : nullptr, dirichlet_values,
                             &h->d.solver, &h->d.newton, report);
}

int mf_heat_stepper_state(const mf_heat_stepper *h, const double **T, const double **T_old,
                          const double **heat_source, double *time, double position[3])
{
  MF_REQUIRE(h, MF_ERR_ARG, "null stepper");
  if (T) *T = h->T;
  if (T_old) *T_old = h->T_old;
  if (heat_source) *heat_source = h->q;
  if (time) *time = h->time;
  if (position)
    for (int k = 0; k < 3; ++k) position[k] = h->position[k];
  return MF_OK;
}

int mf_heat_stepper_copy(const mf_heat_stepper *h, int which, double *dst)
{
  MF_REQUIRE(h && dst, MF_ERR_ARG, "null argument");
  MF_REQUIRE(which >= 0 && which <= 2, MF_ERR_ARG, "which must be 0, 1 or 2");
  const double *src = which == 0 ? h->T : (which == 1 ? h->T_old : h->q);
  const size_t bytes = sizeof(double) * size_t(h->op->s->mesh.n_dofs);
  MF_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, h->op->s->stream));
  return MF_OK;
}

}  // extern "C"

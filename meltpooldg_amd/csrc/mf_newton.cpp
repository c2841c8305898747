// Host-side Newton-Raphson driver over caller-supplied callbacks.
//
// Replaces NewtonRaphsonSolver::solve / is_converged / set_tolerances_to_alternative_values
// (source/linear_algebra/newton_raphson_solver.cpp:22-175) with the defaults of
// NonlinearSolverData (include/meltpooldg/linear_algebra/nonlinear_solver_data.hpp:13-20).
// The driver holds no vectors: the residual, the linear solve, the norms and the update are the
// caller's (device vectors for mf_heat_newton_step, host arrays in the unit tests).
#include "mf_internal.hpp"

extern "C" {

int mf_newton_desc_init(mf_newton_desc *d)
{
  MF_REQUIRE(d, MF_ERR_ARG, "null argument");
  d->max_iterations = 10;
  d->residual_tolerance = 1e-9;
  d->correction_tolerance = 1e-10;
  d->max_iterations_alt = 0;
  d->residual_tolerance_alt = 1e-8;
  d->correction_tolerance_alt = 1e-9;
  d->use_relative_tolerances = 0;
  return MF_OK;
}

int mf_newton_solve(const mf_newton_desc *desc, const mf_newton_callbacks *cb, void *user,
                    mf_newton_report *report)
{
  MF_REQUIRE(desc && cb && cb->residual && cb->solve_with_jacobian && cb->norms && cb->update,
             MF_ERR_ARG, "null argument");
  MF_REQUIRE(desc->max_iterations >= 0 && desc->max_iterations_alt >= 0, MF_ERR_ARG,
             "negative Newton iteration limit");
  mf_newton_report rep{};
  const int max_total = desc->max_iterations + desc->max_iterations_alt;
  // every call starts with the standard tolerances
  double tol_r = desc->residual_tolerance, tol_d = desc->correction_tolerance;
  const bool relative = desc->use_relative_tolerances != 0;
  double r0 = 0.0;
  int rc = MF_OK;
  bool converged = false;
  for (int i = 0; i < max_total; ++i) {
    if (i == desc->max_iterations) {
      tol_r = desc->residual_tolerance_alt;
      tol_d = desc->correction_tolerance_alt;
    }
    if ((rc = cb->residual(user))) break;
    int lin = 0;
    rc = cb->solve_with_jacobian(user, &lin);
    rep.linear_iterations += lin;
    if (rc) break;
    double r = 0.0, dx = 0.0, x = 0.0;
    if ((rc = cb->norms(user, relative ? 1 : 0, &r, &dx, &x))) break;
    if (i == 0) r0 = r;
    rep.iterations = i + 1;
    rep.residual_norm = r;
    rep.correction_norm = dx;
    if (relative) {
      // a zero reference norm leaves the tolerance absolute
      if (r0 > 0.0) r /= r0;
      if (x > 0.0) dx /= x;
    }
    if (r < tol_r && dx < tol_d) {
      converged = true;
      break;
    }
    if ((rc = cb->update(user))) break;
  }
  rep.status = rc != MF_OK ? rc : (converged ? MF_OK : MF_NO_CONVERGENCE);
  if (report) *report = rep;
  if (rc == MF_OK && !converged) {
    mf::set_error("Newton-Raphson solver did not converge (ExcNewtonDidNotConverge)");
    return MF_NO_CONVERGENCE;
  }
  return rc;
}

}  // extern "C"

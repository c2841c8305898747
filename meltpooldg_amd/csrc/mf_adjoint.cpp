// Adjoint of the implicit heat transient with a moving Gaussian laser (no reference
// counterpart: the reference has no adjoint).  Forward model, one implicit-Euler step per laser
// position as in HeatDiffuseOperation::solve (source/heat/heat_diffuse_operation.cpp:261-289)
// for a temperature-independent law:
//     A T_n = (rho cp/dt) M T_{n-1} + M q_n(P, r),   A = (rho cp/dt) M + K(k0 kc) + C(w)
// Objective J(T_N) given by its gradient g = dJ/dT_N (a dual vector, e.g. M (T_N - T*)).
// Adjoint:  A^T lambda_N = g,   A^T lambda_n = (rho cp/dt) M lambda_{n+1}
// Gradients: dJ/dkc_c = -k0 sum_n lambda_n^T K_c T_n,  dJ/dP = sum_n lambda_n^T M q_n / P,
//            dJ/dr = sum_n lambda_n^T M dq_n/dr.
// States are checkpointed every `checkpoint_interval` steps in the forward sweep and recomputed
// segment by segment in the backward sweep, so HBM holds O(N/interval + interval) states.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "mf_internal.hpp"
#include "mf_vec.hpp"

namespace mf {
int vec_axpby(double a, const double *x, double b, double *y, int64_t n, hipStream_t st);
template <typename Number>
int launch_cell(int op, int needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                const Number *src, Number *dst, hipStream_t stream);
}  // namespace mf

using namespace mf;

struct mf_transient {
  mf_operator *op = nullptr;
  mf_transient_desc d{};
  std::vector<double *> checkpoints;  // T at n = 0, I, 2I, ... (< N)
  bool have_forward = false;
};

extern "C" {

int mf_operator_vmult(mf_operator *op, double *dst, const double *src);
int mf_operator_create_rhs(mf_operator *op, double *dst, const double *src, const double *q_s);
int mf_operator_set_fields(mf_operator *op, const double *T, const double *vel,
                           const double *k_cell);
int mf_solve(mf_operator *op, const mf_solver_desc *desc, double *x, const double *b,
             mf_solve_report *report);
int mf_laser_heat_source(mf_scratch *s, double power, double radius, const double position[3],
                         double *q_s);

int mf_operator_set_transpose(mf_operator *op, int transpose)
{
  MF_REQUIRE(op, MF_ERR_ARG, "null operator");
  MF_REQUIRE(!transpose || op->kind != 1 || !op->advec.supg, MF_ERR_ARG,
             "the transposed advection operator is not available with SUPG");
  op->transpose = transpose ? 1 : 0;
  return MF_OK;
}

int mf_mass_vmult(mf_scratch *s, double *dst, const double *src)
{
  MF_REQUIRE(s && dst && src, MF_ERR_ARG, "null argument");
  mf_operator m{};
  m.kind = 0;
  m.s = s;
  m.heat.law = MF_LAW_CONSTANT;
  m.heat.rho_cp = 1.0;
  m.heat.k = 0.0;
  m.dt = 1.0;
  m.gather_plain = 1;  // sources are read plain, as the heat residual reads q_s
  return mf_operator_vmult(&m, dst, src);
}

int mf_vector_dot(mf_scratch *s, const double *a, const double *b, double *result)
{
  MF_REQUIRE(s && a && b && result, MF_ERR_ARG, "null argument");
  hipStream_t st = s->stream;
  const int64_t n = s->mesh.n_dofs;
  double *part = nullptr;
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&part), sizeof(double) * (2 * RED_BLOCKS + 1), st));
  int rc = vec_dot<double>(a, b, n, part, part + 2 * RED_BLOCKS, st);
  if (rc == MF_OK) {
    MF_HIP_TRY(hipMemcpyAsync(result, part + 2 * RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, st));
    MF_HIP_TRY(hipStreamSynchronize(st));
  }
  (void)hipFreeAsync(part, st);
  return rc;
}

int mf_heat_conductivity_gradient(mf_operator *op, const double *lambda, const double *T,
                                  double scale, double *grad_k_cell)
{
  MF_REQUIRE(op && lambda && T && grad_k_cell, MF_ERR_ARG, "null argument");
  MF_REQUIRE(op->kind == 0, MF_ERR_ARG, "conductivity gradient needs a heat operator");
  const MeshInfo &m = op->s->mesh;
  CellArgs<double> a;
  std::memset(&a, 0, sizeof(a));
  const Tables &t = op->s->tab;
  for (int i = 0; i < m.K * m.K; ++i) {
    a.S[i] = t.S[i];
    a.DS[i] = t.DS[i];
    a.Dq[i] = t.Dq[i];
  }
  for (int i = 0; i < m.K; ++i) a.w[i] = t.w[i];
  for (int d = 0; d < 3; ++d) {
    a.ncell[d] = m.ncell[d];
    a.h[d] = m.h[d];
    a.inv_h[d] = 1.0 / m.h[d];
  }
  a.NX = uint32_t(m.NN[0]);
  a.NY = uint32_t(m.NN[1]);
  a.n_dofs = uint32_t(m.n_dofs);
  a.n_cells = int(m.n_cells);
  a.T = T;
  a.pgrad_scale = scale;
  return launch_cell<double>(HEAT_PGRAD, false, false, m, a, lambda, grad_k_cell, op->s->stream);
}

int mf_transient_create(mf_operator *op, const mf_transient_desc *d, mf_transient **out)
{
  MF_REQUIRE(op && d && out, MF_ERR_ARG, "null argument");
  MF_REQUIRE(op->kind == 0, MF_ERR_ARG, "the transient needs a heat operator");
  MF_REQUIRE(!op->s->ad, MF_ERR_ARG, "the transient is available on structured meshes only");
  MF_REQUIRE(op->heat.law == MF_LAW_CONSTANT, MF_ERR_ARG,
             "the adjoint transient supports temperature-independent conductivity");
  MF_REQUIRE(d->n_steps >= 1 && d->dt > 0 && d->radius > 0, MF_ERR_ARG, "invalid transient");
  op->dt = d->dt;  // the laser moves with the operator's time increment
  auto *tr = new mf_transient();
  tr->op = op;
  tr->d = *d;
  if (tr->d.checkpoint_interval < 1) tr->d.checkpoint_interval = 1;
  *out = tr;
  return MF_OK;
}

int mf_transient_destroy(mf_transient *tr)
{
  if (!tr) return MF_OK;
  for (double *p : tr->checkpoints) (void)hipFree(p);
  delete tr;
  return MF_OK;
}

}  // extern "C"

namespace {

struct Work {
  hipStream_t st;
  std::vector<double *> bufs;
  ~Work()
  {
    for (double *p : bufs) (void)hipFreeAsync(p, st);
  }
  int alloc(double **p, int64_t n)
  {
    MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(p), sizeof(double) * n, st));
    bufs.push_back(*p);
    return MF_OK;
  }
};

void laser_position(const mf_transient_desc &d, int n, double pos[3])
{
  const double t = n * d.dt;
  for (int i = 0; i < 3; ++i) pos[i] = d.start[i] + d.scan_velocity[i] * t;
}

// T_new = step(T_prev): one linear implicit-Euler step (the Newton iteration of a linear problem)
int forward_step(mf_transient *tr, const double *T_prev, double *T_new, double *q, double *r,
                 int n)
{
  mf_operator *op = tr->op;
  mf_scratch *s = op->s;
  const int64_t N = s->mesh.n_dofs;
  hipStream_t st = s->stream;
  double pos[3];
  laser_position(tr->d, n, pos);
  int rc;
  if ((rc = mf_laser_heat_source(s, tr->d.power, tr->d.radius, pos, q))) return rc;
  MF_HIP_TRY(hipMemcpyAsync(T_new, T_prev, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
  const double *vel = op->vel, *kc = op->k_cell;
  if ((rc = mf_operator_set_fields(op, T_new, vel, kc))) return rc;
  if ((rc = vec_fill<double>(r, 0.0, N, st))) return rc;
  if ((rc = mf_operator_create_rhs(op, r, T_prev, q))) return rc;
  // solve A d = r into q (reuse), T_new += d
  if ((rc = vec_fill<double>(q, 0.0, N, st))) return rc;
  mf_solve_report rep{};
  if ((rc = mf_solve(op, &tr->d.solver, q, r, &rep))) return rc;
  return vec_axpby(1.0, q, 1.0, T_new, N, st);
}

}  // namespace

extern "C" {

int mf_transient_forward(mf_transient *tr, const double *T0, double *T_final)
{
  MF_REQUIRE(tr && T0 && T_final, MF_ERR_ARG, "null argument");
  mf_operator *op = tr->op;
  mf_scratch *s = op->s;
  const int64_t N = s->mesh.n_dofs;
  hipStream_t st = s->stream;
  const int I = tr->d.checkpoint_interval, steps = tr->d.n_steps;
  const int ncp = (steps - 1) / I + 1;  // checkpoints at 0, I, ..., < steps
  if (int(tr->checkpoints.size()) != ncp) {
    for (double *p : tr->checkpoints) MF_HIP_TRY(hipFree(p));
    tr->checkpoints.assign(ncp, nullptr);
    for (auto &p : tr->checkpoints) MF_HIP_TRY(hipMalloc(&p, sizeof(double) * N));
  }
  op->transpose = 0;
  const double *T_user = op->T, *vel = op->vel, *kc = op->k_cell;
  Work w{st, {}};
  double *Ta, *Tb, *q, *r;
  int rc;
  if ((rc = w.alloc(&Ta, N)) || (rc = w.alloc(&Tb, N)) || (rc = w.alloc(&q, N)) ||
      (rc = w.alloc(&r, N)))
    return rc;
  MF_HIP_TRY(hipMemcpyAsync(Ta, T0, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
  for (int n = 1; n <= steps; ++n) {
    if ((n - 1) % I == 0)
      MF_HIP_TRY(hipMemcpyAsync(tr->checkpoints[(n - 1) / I], Ta, sizeof(double) * N,
                                hipMemcpyDeviceToDevice, st));
    if ((rc = forward_step(tr, Ta, Tb, q, r, n))) return rc;
    std::swap(Ta, Tb);
  }
  MF_HIP_TRY(hipMemcpyAsync(T_final, Ta, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
  op->T = T_user;  // the work buffers die here: restore the caller's fields
  op->vel = vel;
  op->k_cell = kc;
  MF_HIP_TRY(hipStreamSynchronize(st));
  tr->have_forward = true;
  return MF_OK;
}

int mf_transient_backward(mf_transient *tr, const double *dJ_dTN, double *grad_k_cell,
                          double *grad_power, double *grad_radius)
{
  MF_REQUIRE(tr && dJ_dTN, MF_ERR_ARG, "null argument");
  MF_REQUIRE(tr->have_forward, MF_ERR_STATE, "mf_transient_backward needs a forward sweep first");
  mf_operator *op = tr->op;
  mf_scratch *s = op->s;
  const int64_t N = s->mesh.n_dofs;
  hipStream_t st = s->stream;
  const int I = tr->d.checkpoint_interval, steps = tr->d.n_steps;
  const double c_mass = op->heat.rho_cp / op->dt;
  Work w{st, {}};
  std::vector<double *> states(I + 1);
  double *lam, *rhs, *q, *r, *dq, *tmp;
  int rc;
  for (auto &p : states)
    if ((rc = w.alloc(&p, N))) return rc;
  if ((rc = w.alloc(&lam, N)) || (rc = w.alloc(&rhs, N)) || (rc = w.alloc(&q, N)) ||
      (rc = w.alloc(&r, N)) || (rc = w.alloc(&dq, N)) || (rc = w.alloc(&tmp, N)))
    return rc;
  double gP = 0.0, gR = 0.0;
  bool first = true;
  const double *T_user = op->T, *vel = op->vel, *kc = op->k_cell;
  for (int seg = (steps - 1) / I; seg >= 0; --seg) {
    const int s0 = seg * I, s1 = std::min(steps, s0 + I);
    // recompute T_{s0..s1} from the checkpoint (forward operator)
    op->transpose = 0;
    MF_HIP_TRY(hipMemcpyAsync(states[0], tr->checkpoints[seg], sizeof(double) * N,
                              hipMemcpyDeviceToDevice, st));
    for (int n = s0 + 1; n <= s1; ++n)
      if ((rc = forward_step(tr, states[n - 1 - s0], states[n - s0], q, r, n))) return rc;
    // adjoint steps n = s1 .. s0+1
    for (int n = s1; n > s0; --n) {
      if (first) {
        MF_HIP_TRY(hipMemcpyAsync(rhs, dJ_dTN, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        first = false;
      } else {
        if ((rc = mf_mass_vmult(s, rhs, lam))) return rc;  // lam holds lambda_{n+1}
        if ((rc = vec_axpby(0.0, rhs, c_mass, rhs, N, st))) return rc;
      }
      op->transpose = 1;
      if ((rc = mf_operator_set_fields(op, states[n - s0], vel, kc))) return rc;
      if ((rc = vec_fill<double>(lam, 0.0, N, st))) return rc;
      mf_solve_report rep{};
      rc = mf_solve(op, &tr->d.solver, lam, rhs, &rep);
      op->transpose = 0;
      if (rc) return rc;
      // parameter gradients with T_n and q_n
      if (grad_k_cell &&
          (rc = mf_heat_conductivity_gradient(op, lam, states[n - s0], -op->heat.k, grad_k_cell)))
        return rc;
      double pos[3];
      laser_position(tr->d, n, pos);
      if ((rc = mf_laser_heat_source(s, tr->d.power, tr->d.radius, pos, q))) return rc;
      if ((rc = mf_mass_vmult(s, tmp, q))) return rc;
      double v;
      if ((rc = mf_vector_dot(s, lam, tmp, &v))) return rc;
      gP += v / tr->d.power;
      if ((rc = mf_laser_heat_source_radius_derivative(s, tr->d.power, tr->d.radius, pos, dq)))
        return rc;
      if ((rc = mf_mass_vmult(s, tmp, dq))) return rc;
      if ((rc = mf_vector_dot(s, lam, tmp, &v))) return rc;
      gR += v;
    }
  }
  MF_HIP_TRY(hipStreamSynchronize(st));
  op->T = T_user;
  op->vel = vel;
  op->k_cell = kc;
  if (grad_power) *grad_power = gP;
  if (grad_radius) *grad_radius = gR;
  return MF_OK;
}

}  // extern "C"

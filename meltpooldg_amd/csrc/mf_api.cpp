// C ABI of libmeltpooldg_amd: scratch (mesh/DoF/constraint setup), operators, solvers.
// Host-side control only; all arithmetic on vectors runs in the HIP kernels of mf_cell.hip and
// mf_vec.hip.  See include/meltpooldg_amd.h for the reference interface of every entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "mf_internal.hpp"
#include "mf_vec.hpp"

namespace mf {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

// ------------------------------------------------------------------------------------------
// 1D tables: FE_Q(p) on Gauss-Lobatto points, QGauss(p+1)
// (source/utilities/fe_util.cpp:18-72, source/core/finite_element_data.cpp:85-89)
// ------------------------------------------------------------------------------------------
static void legendre_pair(int n, double z, double &pn, double &pnm1)
{
  double a = 1.0, b = 0.0;  // P_j, P_{j-1}
  for (int j = 1; j <= n; ++j) {
    const double c = ((2.0 * j - 1.0) * z * a - (j - 1.0) * b) / j;
    b = a;
    a = c;
  }
  pn = a;
  pnm1 = b;
}

static void gauss_points(int n, double *x, double *w)
{
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double pn, pm, dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      legendre_pair(n, z, pn, pm);
      dp = n * (z * pn - pm) / (z * z - 1.0);
      const double dz = pn / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    legendre_pair(n, z, pn, pm);
    dp = n * (z * pn - pm) / (z * z - 1.0);
    x[n - 1 - i] = 0.5 * (1.0 + z);
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
}

static void lobatto_points(int n, double *x)
{
  const int p = n - 1;
  x[0] = 0.0;
  x[p] = 1.0;
  for (int i = 1; i < p; ++i) {
    double z = -std::cos(M_PI * i / p);
    for (int it = 0; it < 100; ++it) {
      double pn, pm;
      legendre_pair(p, z, pn, pm);
      const double d1 = p * (z * pn - pm) / (z * z - 1.0);
      const double d2 = (2.0 * z * d1 - p * (p + 1.0) * pn) / (1.0 - z * z);
      const double dz = d1 / d2;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    x[i] = 0.5 * (1.0 + z);
  }
  for (int i = 0; i < n / 2; ++i) {  // exact symmetry of the point set
    const double a = 0.5 * (x[i] + 1.0 - x[p - i]);
    x[i] = a;
    x[p - i] = 1.0 - a;
  }
  if (n % 2) x[n / 2] = 0.5;
}

// values and derivatives of the Lagrange polynomials on `nodes` at x
static void lagrange_eval(int n, const double *nodes, double x, double *v, double *d)
{
  for (int j = 0; j < n; ++j) {
    double val = 1.0, der = 0.0;
    for (int m = 0; m < n; ++m) {
      if (m == j) continue;
      const double inv = 1.0 / (nodes[j] - nodes[m]);
      der = der * (x - nodes[m]) * inv + val * inv;
      val *= (x - nodes[m]) * inv;
    }
    v[j] = val;
    d[j] = der;
  }
}

void lagrange_basis(int n, const double *nodes, double x, double *v, double *d)
{
  lagrange_eval(n, nodes, x, v, d);
}

void compute_tables(int degree, Tables &t)
{
  const int K = degree + 1;
  std::memset(&t, 0, sizeof(t));
  lobatto_points(K, t.gl);
  gauss_points(K, t.xq, t.w);
  double v[KMAX], d[KMAX];
  for (int q = 0; q < K; ++q) {
    lagrange_eval(K, t.gl, t.xq[q], v, d);
    for (int j = 0; j < K; ++j) {
      t.S[q * K + j] = v[j];
      t.DS[q * K + j] = d[j];
    }
    lagrange_eval(K, t.xq, t.xq[q], v, d);
    for (int j = 0; j < K; ++j) t.Dq[q * K + j] = d[j];
  }
}

// ------------------------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------------------------
template <typename Number>
static CellArgs<Number> make_args(const mf_operator *op)
{
  const mf_scratch *s = op->s;
  const MeshInfo &m = s->mesh;
  CellArgs<Number> a;
  std::memset(&a, 0, sizeof(a));
  const int K = m.K;
  for (int i = 0; i < K * K; ++i) {
    a.S[i] = Number(s->tab.S[i]);
    a.DS[i] = Number(s->tab.DS[i]);
    a.Dq[i] = Number(s->tab.Dq[i]);
  }
  for (int i = 0; i < K; ++i) a.w[i] = Number(s->tab.w[i]);
  for (int d = 0; d < 3; ++d) {
    a.ncell[d] = m.ncell[d];
    a.h[d] = Number(m.h[d]);
    a.inv_h[d] = Number(1.0 / m.h[d]);
  }
  a.NX = uint32_t(m.NN[0]);
  a.NY = uint32_t(m.NN[1]);
  a.n_dofs = uint32_t(m.n_dofs);
  a.n_cells = int(m.n_cells);
  a.mask = s->d_mask;
  a.mask_boundary = s->mask_full_boundary ? 1 : 0;
  // the level-set projections have no Dirichlet rows
  a.use_mask = (op->kind < 2 && op->dof_index == 0) ? 1 : 0;
  a.transpose = op->transpose;
  a.gather_plain = op->gather_plain;
  a.dt_inv = Number(op->dt > 0 ? 1.0 / op->dt : 0.0);
  if (op->kind == 0) {
    a.rho_cp = Number(op->heat.rho_cp);
    a.k0 = Number(op->heat.k);
    a.tanh_amp = Number(op->heat.tanh_amp);
    a.tanh_Tc = Number(op->heat.tanh_Tc);
    a.inv_tanh_w = Number(op->heat.tanh_width != 0 ? 1.0 / op->heat.tanh_width : 0.0);
    a.t_dependent = op->heat.t_dependent;
    if (op->heat.law == MF_LAW_MUSHY) {
      const mf_heat_params &h = op->heat;
      a.mu_k[0] = Number(h.liquid_k);
      a.mu_k[1] = Number(h.solid_k);
      a.mu_rho[0] = Number(h.liquid_rho);
      a.mu_rho[1] = Number(h.solid_rho);
      a.mu_cp[0] = Number(h.liquid_cp);
      a.mu_cp[1] = Number(h.solid_cp);
      a.mu_Tliq = Number(h.liquidus);
      a.mu_inv = Number(1.0 / (h.liquidus - h.solidus));
    }
  } else if (op->kind == 1) {
    a.theta = Number(op->advec.theta);
    a.diffusivity = Number(op->advec.diffusivity);
    a.mass_factor = Number(op->advec.mass_factor);
    if (op->advec.supg) {
      // cell size diameter / sqrt(dim) of a level-0 cell (ScratchData::build,
      // source/core/scratch_data.cpp:256-278)
      const double d2 = m.h[0] * m.h[0] + m.h[1] * m.h[1] + m.h[2] * m.h[2];
      a.supg_cell_size = Number(std::sqrt(d2 / 3.0));
      a.supg_coef = Number(op->advec.supg_coefficient);
      // second derivatives of the 1D basis at the Gauss points: the collocation derivative of
      // the first derivatives (exact: l_j' has degree p - 1)
      for (int q = 0; q < K; ++q)
        for (int j = 0; j < K; ++j) {
          double v = 0.0;
          for (int r = 0; r < K; ++r) v += s->tab.Dq[q * K + r] * s->tab.DS[r * K + j];
          a.DDS[q * K + j] = Number(v);
        }
    }
  } else if (op->kind == 4) {
    double en, et;
    reinit_epsilon(op, en, et, a.re_scale_h);
    a.re_eps_n = Number(en);
    a.re_eps_t = Number(et);
    a.re_tol = Number(op->ls_tol);
    a.re_sign = Number(en > et ? 1.0 : (en < et ? -1.0 : 0.0));
  } else {
    // eta = filter_parameter (diameter / sqrt(3))^2 of a level-0 cell
    const double d2 = m.h[0] * m.h[0] + m.h[1] * m.h[1] + m.h[2] * m.h[2];
    a.ls_eta = Number(op->ls.filter_parameter * d2 / 3.0);
    a.ls_threshold = Number(op->ls.narrow_band_threshold);
  }
  if (const AdaptiveData *ad = s->ad) {
    a.cell_nodes = ad->d.cell_nodes;
    a.cell_hmask = ad->d.cell_hmask;
    a.cell_level = ad->d.cell_level;
    a.cell_hidx = ad->d.cell_hidx;
    a.node_line = ad->d.node_line;
    a.line_ptr = ad->d.line_ptr;
    a.line_col = ad->d.line_col;
    a.list = ad->d.list;
    a.ht_ptr = ad->d.ht_ptr;
    a.tgt_node = ad->d.tgt_node;
    a.tgt_ptr = ad->d.tgt_ptr;
    a.ent_local = ad->d.ent_local;
    a.slot_cell = ad->d.slot_cell;
    a.slot_tgt = ad->d.slot_tgt;
    if constexpr (sizeof(Number) == 8) {
      a.line_val = reinterpret_cast<const Number *>(ad->d.line_val);
      a.ent_w = reinterpret_cast<const Number *>(ad->d.ent_w);
    } else {
      a.line_val = reinterpret_cast<const Number *>(ad->d.line_val32);
      a.ent_w = reinterpret_cast<const Number *>(ad->d.ent_w32);
    }
  }
  if constexpr (sizeof(Number) == 8) {
    a.T = reinterpret_cast<const Number *>(op->T);
    a.vel = reinterpret_cast<const Number *>(op->vel);
    a.k_cell = reinterpret_cast<const Number *>(op->k_cell);
    a.T_old = reinterpret_cast<const Number *>(op->T_old);
  } else {
    a.T = reinterpret_cast<const Number *>(op->T ? op->T32 : nullptr);
    a.vel = reinterpret_cast<const Number *>(op->vel ? op->vel32 : nullptr);
    a.k_cell = reinterpret_cast<const Number *>(op->k_cell ? op->kc32 : nullptr);
    a.T_old = reinterpret_cast<const Number *>(op->T_old ? op->To32 : nullptr);
  }
  return a;
}

// T: the temperature of a T-dependent law (1: tanh, 2: mushy zone), or the level set of a
// narrow band (1)
static int needs_T(const mf_operator *op)
{
  if (op->kind == 4) return 0;
  if (op->kind == 1) return 0;
  if (op->kind >= 2) return op->ls.narrow_band != 0 ? 1 : 0;
  if (op->kind != 0) return 0;
  return op->heat.law == MF_LAW_TANH ? 1 : (op->heat.law == MF_LAW_MUSHY ? 2 : 0);
}

// the cell kernels' field flag: needs_T, or for the advection operator SUPG on
static int cell_flag(const mf_operator *op)
{
  if (op->kind == 1) return op->advec.supg ? 1 : 0;
  return needs_T(op);
}

// the mushy law's tangent reads the old temperature
static bool needs_T_old(const mf_operator *op)
{
  return op->kind == 0 && op->heat.law == MF_LAW_MUSHY && op->heat.t_dependent;
}

// the inflow/outflow treatment applies to this application (advection operator, constrained
// DoF index, pre/post on, a non-empty set)
static bool io_active(const mf_operator *op)
{
  return op->kind == 1 && op->io_mask && op->io_pre_post && op->dof_index == 0;
}

// vector blocks of the operator (BlockVectorType of the normal vector operator)
static int n_blocks(const mf_operator *op) { return op->kind == 2 ? 3 : 1; }
static int64_t n_rows(const mf_operator *op) { return n_blocks(op) * op->s->mesh.n_dofs; }

static int check_ready(const mf_operator *op)
{
  MF_REQUIRE(op && op->s, MF_ERR_ARG, "null operator");
  if (op->kind == 5) {
    MF_REQUIRE(op->dt > 0.0, MF_ERR_STATE,
               "time increment not set (reference AssertThrowZeroTimeIncrement)");
    MF_REQUIRE(cut_ready(op), MF_ERR_STATE, "level set not set (mf_heat_cut_set_level_set)");
    if (op->cut.material.law == MF_LAW_MUSHY)
      MF_REQUIRE(op->T, MF_ERR_STATE,
                 "the mushy-zone law needs T: call mf_operator_set_fields");
    return MF_OK;
  }
  if (op->kind == 4) {
    MF_REQUIRE(op->dt > 0.0, MF_ERR_STATE,
               "time increment not set (reference AssertThrowZeroTimeIncrement)");
    MF_REQUIRE(op->vel, MF_ERR_STATE,
               "the reinitialization operator needs the normal vector (mf_operator_set_normal_vector)");
    return MF_OK;
  }
  if (op->kind >= 2) {
    if (op->ls.narrow_band)
      MF_REQUIRE(op->T, MF_ERR_STATE,
                 "Level set solution vector must not be nullptr for a narrow band computation "
                 "(mf_operator_set_level_set)");
    return MF_OK;
  }
  MF_REQUIRE(op->dt > 0.0, MF_ERR_STATE,
             "time increment not set (reference AssertThrowZeroTimeIncrement)");
  if (needs_T(op)) MF_REQUIRE(op->T, MF_ERR_STATE, "temperature-dependent law needs T: call mf_operator_set_fields");
  if (needs_T_old(op))
    MF_REQUIRE(op->T_old, MF_ERR_STATE,
               "the mushy-zone Jacobian needs the old temperature: call mf_operator_set_temperature_old");
  if (op->kind == 1) MF_REQUIRE(op->vel, MF_ERR_STATE, "advection operator needs a velocity field");
  return MF_OK;
}

int arena_reserve(Arena &a, size_t bytes)
{
  if (a.bytes >= bytes) return MF_OK;
  if (a.p) MF_HIP_TRY(hipFree(a.p));  // synchronises: only when a larger problem comes along
  a.p = nullptr;
  a.bytes = 0;
  MF_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&a.p), bytes));
  a.bytes = bytes;
  return MF_OK;
}

// Sub-allocation of an arena: reserve(sizes...) then take<T>(count) in the same order.
struct Carve {
  std::vector<size_t> sizes;
  char *cur = nullptr;
  template <typename T>
  void plan(size_t count) { sizes.push_back((sizeof(T) * std::max<size_t>(count, 1) + 255) & ~size_t(255)); }
  int reserve(Arena &a)
  {
    size_t total = 0;
    for (size_t b : sizes) total += b;
    int rc = arena_reserve(a, total);
    cur = a.p;
    return rc;
  }
  template <typename T>
  T *take(size_t count)
  {
    T *r = reinterpret_cast<T *>(cur);
    cur += (sizeof(T) * std::max<size_t>(count, 1) + 255) & ~size_t(255);
    return r;
  }
};

// blocks that write fused-dot partials in one operator application: the cell kernel (per
// vector block) and the boundary faces
static int op_tangent_blocks(const mf_operator *op, bool fp32)
{
  if (op->kind == 5) return cut_dot_blocks(op);
  const MeshInfo &m = op->s->mesh;
  // the level-set projections use the pencil kernel like the transposed operator
  const int cells = tangent_blocks(m, op->transpose != 0 || op->kind >= 2, fp32);
  return n_blocks(op) * cells + heat_face_blocks(op) + (io_active(op) ? IO_BLOCKS : 0);
}

// dst = A src (overwrite = true, the reference's vmult with zeroed dst) or dst += A src.
// The structured-mesh kernel overwrites in place (the first colour that writes a node stores
// it); the other kernels accumulate into a zeroed dst.
template <typename Number>
static int apply_tangent(mf_operator *op, Number *dst, const Number *src, double *dot_partial,
                         const int *stop, bool overwrite)
{
  const MeshInfo &m = op->s->mesh;
  hipStream_t st = op->s->stream;
  if (op->kind == 5) {
    if constexpr (sizeof(Number) == 8) {
      return cut_apply(op, dst, src, dot_partial, stop, overwrite);
    } else {
      set_error("the cut heat operator has no FP32 form");
      return MF_ERR_ARG;
    }
  }
  if (n_blocks(op) > 1) {
    // the same scalar operator on every block; the fused dot's partials follow block by block
    const int per = op_tangent_blocks(op, sizeof(Number) == 4) / n_blocks(op);
    const int64_t nd = m.n_dofs;
    mf_operator scalar_view = *op;  // shallow: reads the same fields, owns nothing
    scalar_view.kind = 3;
    for (int b = 0; b < n_blocks(op); ++b) {
      int rc = apply_tangent<Number>(&scalar_view, dst + b * nd, src + b * nd,
                                     dot_partial ? dot_partial + b * per : nullptr, stop, overwrite);
      if (rc) return rc;
    }
    return MF_OK;
  }
  if (overwrite && m.ad) {
    int rc = vec_fill<Number>(dst, Number(0), m.n_dofs, st, stop);
    if (rc) return rc;
  }
  CellArgs<Number> a = make_args<Number>(op);
  // inflow/outflow rows (do_pre_vmult): their src entries read as 0, their rows left to the post
  // pass, like Dirichlet rows
  const bool io = io_active(op);
  if (io) {
    a.mask = op->io_union;
    a.mask_boundary = 0;
  }
  a.dot_partial = dot_partial;
  a.stop = stop;
  a.overwrite = (overwrite && !m.ad) ? 1 : 0;
  const int code = op->kind == 0 ? HEAT_TANGENT
                 : op->kind == 1 ? ADVEC_TANGENT
                 : op->kind == 4 ? REINIT_TANGENT
                                 : LS_TANGENT;
  if (m.ad) {
    if (!op->gather_plain && m.ad->n_hanging > 0) {
      Number *hv;
      if constexpr (sizeof(Number) == 8)
        hv = reinterpret_cast<Number *>(m.ad->d.hang_val);
      else
        hv = reinterpret_cast<Number *>(m.ad->d.hang_val32);
      int rc = launch_hanging_values<Number>(op->s, src, a.use_mask, hv, stop);
      if (rc) return rc;
      a.hang_val = hv;
    }
    int rc = launch_cell_list<Number>(code, cell_flag(op), op->vel != nullptr, m, a, src, dst, st, 0);
    if (rc) return rc;
    // hanging rows hold the unconstrained cell sums: move them to their masters
    return launch_condense<Number>(op->s, dst, stop, a.use_mask != 0);
  }
  // the transposed operator (adjoint) and the level-set projections run on the pencil kernel
  int rc = launch_cell<Number>(code, cell_flag(op), op->vel != nullptr, m, a, src, dst, st);
  if (rc) return rc;
  if (io) {
    // do_post_vmult: dst = src on the inflow/outflow rows; their part of the fused dot follows
    // the cell kernel's partials
    return launch_io_rows<Number>(op->io_mask, src, dst, m.n_dofs,
                                  dot_partial ? dot_partial + tangent_blocks(m, false,
                                                                             sizeof(Number) == 4)
                                              : nullptr,
                                  stop, st);
  }
  // Robin boundary faces (symmetric: the same for the transposed operator); their part of the
  // fused dot follows the cell kernel's partials
  return launch_heat_faces<Number>(
      op, src, dst, 0, stop,
      dot_partial ? dot_partial + tangent_blocks(m, op->transpose != 0 || op->kind >= 2,
                                                 sizeof(Number) == 4)
                  : nullptr);
}

int g_adaptive_setup_device = 1;

}  // namespace mf

using namespace mf;

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
extern "C" {

const char *mf_last_error(void) { return g_last_error.c_str(); }

int mf_version(void) { return 100; }

int mf_set_adaptive_setup(int device)
{
  g_adaptive_setup_device = device ? 1 : 0;
  return MF_OK;
}

int mf_set_tangent_kernel(int variant)
{
  MF_REQUIRE(variant == 0, MF_ERR_ARG, "tangent kernel variant must be 0");
  return MF_OK;
}

int mf_scratch_create(const mf_mesh_desc *desc, mf_scratch **out)
{
  MF_REQUIRE(desc && out, MF_ERR_ARG, "null argument");
  MF_REQUIRE(desc->dim == 3, MF_ERR_ARG, "only dim = 3 is supported on the device");
  MF_REQUIRE(desc->degree >= 1 && desc->degree <= 3, MF_ERR_ARG, "degree must be 1, 2 or 3");
  auto *s = new mf_scratch();
  MeshInfo &m = s->mesh;
  m.degree = desc->degree;
  m.K = desc->degree + 1;
  m.n_dofs = 1;
  m.n_cells = 1;
  for (int d = 0; d < 3; ++d) {
    if (desc->ncell[d] < 1 || !(desc->hi[d] > desc->lo[d])) {
      delete s;
      set_error("invalid mesh extent");
      return MF_ERR_ARG;
    }
    m.ncell[d] = desc->ncell[d];
    m.NN[d] = int64_t(desc->ncell[d]) * desc->degree + 1;
    m.lo[d] = desc->lo[d];
    m.hi[d] = desc->hi[d];
    m.h[d] = (desc->hi[d] - desc->lo[d]) / desc->ncell[d];
    m.n_dofs *= m.NN[d];
    m.n_cells *= m.ncell[d];
  }
  if (m.n_dofs >= (int64_t(1) << 31)) {
    delete s;
    set_error("mesh too large for 32-bit DoF indices");
    return MF_ERR_ARG;
  }
  compute_tables(m.degree, s->tab);
  *out = s;
  return MF_OK;
}

int mf_scratch_destroy(mf_scratch *s)
{
  if (!s) return MF_OK;
  free_adaptive(s);
  if (s->d_mask) (void)hipFree(s->d_mask);
  delete s;
  return MF_OK;
}

int64_t mf_scratch_n_dofs(const mf_scratch *s) { return s ? s->mesh.n_dofs : -1; }
int64_t mf_scratch_n_cells(const mf_scratch *s) { return s ? s->mesh.n_cells : -1; }

int mf_set_stream(mf_scratch *s, void *stream)
{
  MF_REQUIRE(s, MF_ERR_ARG, "null scratch");
  s->stream = reinterpret_cast<hipStream_t>(stream);
  return MF_OK;
}

int mf_scratch_set_dirichlet_mask(mf_scratch *s, const uint8_t *mask_host)
{
  MF_REQUIRE(s, MF_ERR_ARG, "null scratch");
  s->mask_full_boundary = false;
  if (!mask_host) {
    if (s->d_mask) MF_HIP_TRY(hipFree(s->d_mask));
    s->d_mask = nullptr;
    s->h_mask.clear();
    return MF_OK;
  }
  const size_t n = size_t(s->mesh.n_dofs);
  s->h_mask.assign(mask_host, mask_host + n);
  if (!s->mesh.ad) {
    // is the mask exactly the boundary of the box?
    const int64_t NX = s->mesh.NN[0], NY = s->mesh.NN[1], NZ = s->mesh.NN[2];
    bool full = true;
    for (int64_t z = 0, id = 0; z < NZ && full; ++z)
      for (int64_t y = 0; y < NY && full; ++y) {
        const bool face = z == 0 || z == NZ - 1 || y == 0 || y == NY - 1;
        for (int64_t x = 0; x < NX; ++x, ++id) {
          const bool b = face || x == 0 || x == NX - 1;
          if (b != (mask_host[id] != 0)) {
            full = false;
            break;
          }
        }
      }
    s->mask_full_boundary = full;
  }
  if (!s->d_mask) MF_HIP_TRY(hipMalloc(&s->d_mask, n));
  MF_HIP_TRY(hipMemcpy(s->d_mask, mask_host, n, hipMemcpyHostToDevice));
  return MF_OK;
}

int mf_scratch_shape_info(const mf_scratch *s, double *S, double *DS, double *Dq, double *w,
                          double *gl)
{
  MF_REQUIRE(s, MF_ERR_ARG, "null scratch");
  const int K = s->mesh.K;
  if (S) std::memcpy(S, s->tab.S, sizeof(double) * K * K);
  if (DS) std::memcpy(DS, s->tab.DS, sizeof(double) * K * K);
  if (Dq) std::memcpy(Dq, s->tab.Dq, sizeof(double) * K * K);
  if (w) std::memcpy(w, s->tab.w, sizeof(double) * K);
  if (gl) std::memcpy(gl, s->tab.gl, sizeof(double) * K);
  return MF_OK;
}

int mf_heat_operator_create(mf_scratch *s, const mf_heat_params *p, mf_operator **out)
{
  MF_REQUIRE(s && p && out, MF_ERR_ARG, "null argument");
  MF_REQUIRE(p->law == MF_LAW_CONSTANT || p->law == MF_LAW_TANH || p->law == MF_LAW_MUSHY,
             MF_ERR_ARG, "unknown law");
  MF_REQUIRE(p->law != MF_LAW_TANH || p->tanh_width > 0.0, MF_ERR_ARG, "tanh_width must be > 0");
  MF_REQUIRE(p->law != MF_LAW_MUSHY || p->liquidus > p->solidus, MF_ERR_ARG,
             "the liquidus must lie above the solidus");
  if (int rc = ensure_adaptive_on_device(s)) return rc;
  auto *op = new mf_operator();
  op->kind = 0;
  op->s = s;
  op->heat = *p;
  *out = op;
  return MF_OK;
}

int mf_advection_operator_create(mf_scratch *s, const mf_advec_params *p, mf_operator **out)
{
  MF_REQUIRE(s && p && out, MF_ERR_ARG, "null argument");
  MF_REQUIRE(p->diffusivity >= 0.0, MF_ERR_ARG,
             "Advection diffusion operator: diffusivity is smaller than zero!");
  if (int rc = ensure_adaptive_on_device(s)) return rc;
  auto *op = new mf_operator();
  op->kind = 1;
  op->s = s;
  op->advec = *p;
  *out = op;
  return MF_OK;
}

static int levelset_projection_create(mf_scratch *s, const mf_levelset_projection_params *p,
                                      int kind, mf_operator **out)
{
  MF_REQUIRE(s && p && out, MF_ERR_ARG, "null argument");
  MF_REQUIRE(p->filter_parameter >= 0.0, MF_ERR_ARG, "filter_parameter must be >= 0");
  // NormalVectorData::check_input_parameters / CurvatureData (normal_vector_data.cpp:83)
  MF_REQUIRE(!p->narrow_band || p->narrow_band_threshold > 0.0, MF_ERR_ARG,
             "The narrow band threshold must be larger than zero.");
  if (int rc = ensure_adaptive_on_device(s)) return rc;
  auto *op = new mf_operator();
  op->kind = kind;
  op->s = s;
  op->ls = *p;
  // UtilityFunctions::compute_numerical_zero_of_norm of the domain volume
  // (utility_functions.hpp:104-143): 10^floor(log10(V^(1/3))) * 1e-3, clamped to [1e-12, 1e-2]
  const MeshInfo &m = s->mesh;
  const double vol = (m.hi[0] - m.lo[0]) * (m.hi[1] - m.lo[1]) * (m.hi[2] - m.lo[2]);
  const double e = std::floor(std::log10(std::fabs(std::pow(vol, 1.0 / 3.0))));
  op->ls_tol = std::min(1e-2, std::max(std::pow(10.0, e) * 1e-3, 1e-12));
  *out = op;
  return MF_OK;
}

int mf_reinit_operator_create(mf_scratch *s, const mf_reinit_params *p, mf_operator **out)
{
  MF_REQUIRE(s && p && out, MF_ERR_ARG, "null argument");
  MF_REQUIRE(p->thickness_type >= MF_THICKNESS_PROPORTIONAL && p->thickness_type <= MF_THICKNESS_CELLS,
             MF_ERR_ARG, "unknown interface thickness type");
  MF_REQUIRE(p->thickness_value > 0.0, MF_ERR_ARG,
             "Reinitialization: the value of epsilon for the reinitialization function must be "
             "larger than zero!");
  MF_REQUIRE(p->tangential_diffusion_factor >= 0.0, MF_ERR_ARG,
             "tangential_diffusion_factor must be >= 0");
  mf_levelset_projection_params none{};
  int rc = levelset_projection_create(s, &none, 4, out);
  if (rc) return rc;
  (*out)->reinit = *p;
  return MF_OK;
}

int mf_operator_set_normal_vector(mf_operator *op, const double *normal)
{
  MF_REQUIRE(op, MF_ERR_ARG, "null operator");
  MF_REQUIRE(op->kind == 4, MF_ERR_ARG, "only the reinitialization operator takes a normal vector");
  const MeshInfo &m = op->s->mesh;
  op->vel = normal;
  if (normal) {
    if (!op->vel32) MF_HIP_TRY(hipMalloc(&op->vel32, sizeof(float) * 3 * m.n_dofs));
    return vec_cast<double, float>(normal, op->vel32, 1.0, 3 * m.n_dofs, op->s->stream);
  }
  return MF_OK;
}

double mf_reinit_min_epsilon(const mf_operator *op)
{
  if (!op || op->kind != 4) return 0.0;
  double en, et;
  int scale;
  reinit_epsilon(op, en, et, scale);
  if (scale && op->s->mesh.ad) en /= double(int64_t(1) << op->s->mesh.ad->levels);
  return en;
}

int mf_normal_vector_operator_create(mf_scratch *s, const mf_levelset_projection_params *p,
                                     mf_operator **out)
{
  return levelset_projection_create(s, p, 2, out);
}

int mf_curvature_operator_create(mf_scratch *s, const mf_levelset_projection_params *p,
                                 mf_operator **out)
{
  return levelset_projection_create(s, p, 3, out);
}

int mf_operator_set_level_set(mf_operator *op, const double *level_set)
{
  MF_REQUIRE(op, MF_ERR_ARG, "null operator");
  MF_REQUIRE(op->kind == 2 || op->kind == 3, MF_ERR_ARG,
             "only the level-set projections take a level set");
  const MeshInfo &m = op->s->mesh;
  op->T = level_set;
  if (level_set) {
    if (!op->T32) MF_HIP_TRY(hipMalloc(&op->T32, sizeof(float) * m.n_dofs));
    return vec_cast<double, float>(level_set, op->T32, 1.0, m.n_dofs, op->s->stream);
  }
  return MF_OK;
}

int64_t mf_operator_n_rows(const mf_operator *op) { return op && op->s ? n_rows(op) : -1; }

double mf_operator_normal_tolerance(const mf_operator *op) { return op ? op->ls_tol : 0.0; }

int mf_operator_destroy(mf_operator *op)
{
  if (!op) return MF_OK;
  if (op->T32) (void)hipFree(op->T32);
  if (op->vel32) (void)hipFree(op->vel32);
  if (op->kc32) (void)hipFree(op->kc32);
  if (op->To32) (void)hipFree(op->To32);
  if (op->io_mask) (void)hipFree(op->io_mask);
  if (op->io_union) (void)hipFree(op->io_union);
  cut_free(op);
  for (auto &ar : op->arena)
    if (ar.p) (void)hipFree(ar.p);
  delete op;
  return MF_OK;
}

int mf_operator_reset_time_increment(mf_operator *op, double dt)
{
  MF_REQUIRE(op, MF_ERR_ARG, "null operator");
  MF_REQUIRE(dt > 0.0, MF_ERR_ARG, "time increment must be positive");
  op->dt = dt;
  return MF_OK;
}

int mf_operator_reset_dof_index(mf_operator *op, int dof_index)
{
  MF_REQUIRE(op, MF_ERR_ARG, "null operator");
  MF_REQUIRE(dof_index == 0 || dof_index == 1, MF_ERR_ARG, "dof_index must be 0 or 1");
  op->dof_index = dof_index;
  return MF_OK;
}

int mf_operator_set_fields(mf_operator *op, const double *T, const double *vel,
                           const double *k_cell)
{
  MF_REQUIRE(op, MF_ERR_ARG, "null operator");
  MF_REQUIRE(op->kind < 2 || op->kind == 5, MF_ERR_ARG,
             "the level-set projections take their level set from mf_operator_set_level_set");
  const MeshInfo &m = op->s->mesh;
  hipStream_t st = op->s->stream;
  if (op->kind == 5) {
    // the cut operator: T only, no FP32 copies
    MF_REQUIRE(!vel && !k_cell, MF_ERR_ARG,
               "the cut heat operator takes no velocity or conductivity field");
    op->T = T;
    return MF_OK;
  }
  op->T = T;
  op->vel = vel;
  op->k_cell = k_cell;
  if (T) {
    if (!op->T32) MF_HIP_TRY(hipMalloc(&op->T32, sizeof(float) * m.n_dofs));
    int rc = vec_cast<double, float>(T, op->T32, 1.0, m.n_dofs, st);
    if (rc) return rc;
  }
  if (vel) {
    if (!op->vel32) MF_HIP_TRY(hipMalloc(&op->vel32, sizeof(float) * 3 * m.n_dofs));
    int rc = vec_cast<double, float>(vel, op->vel32, 1.0, 3 * m.n_dofs, st);
    if (rc) return rc;
  }
  if (k_cell) {
    if (!op->kc32) MF_HIP_TRY(hipMalloc(&op->kc32, sizeof(float) * m.n_cells));
    int rc = vec_cast<double, float>(k_cell, op->kc32, 1.0, m.n_cells, st);
    if (rc) return rc;
  }
  return MF_OK;
}

int mf_advection_set_inflow_outflow(mf_operator *op, const uint8_t *mask_host)
{
  MF_REQUIRE(op, MF_ERR_ARG, "null operator");
  MF_REQUIRE(op->kind == 1, MF_ERR_ARG, "inflow/outflow boundaries belong to the advection operator");
  MF_REQUIRE(!op->s->ad, MF_ERR_ARG, "inflow/outflow boundaries need a structured mesh");
  const int64_t n = op->s->mesh.n_dofs;
  if (!mask_host) {
    if (op->io_mask) (void)hipFree(op->io_mask);
    if (op->io_union) (void)hipFree(op->io_union);
    op->io_mask = op->io_union = nullptr;
    return MF_OK;
  }
  std::vector<uint8_t> uni(mask_host, mask_host + n);
  bool any = false;
  for (int64_t i = 0; i < n; ++i) {
    any = any || mask_host[i];
    if (!op->s->h_mask.empty() && op->s->h_mask[i]) uni[i] = 1;
  }
  if (!op->io_mask) MF_HIP_TRY(hipMalloc(&op->io_mask, size_t(n)));
  if (!op->io_union) MF_HIP_TRY(hipMalloc(&op->io_union, size_t(n)));
  MF_HIP_TRY(hipMemcpy(op->io_mask, mask_host, size_t(n), hipMemcpyHostToDevice));
  MF_HIP_TRY(hipMemcpy(op->io_union, uni.data(), size_t(n), hipMemcpyHostToDevice));
  if (!any) return mf_advection_set_inflow_outflow(op, nullptr);
  return MF_OK;
}

int mf_operator_set_inflow_outflow_pre_post(mf_operator *op, int enable)
{
  MF_REQUIRE(op, MF_ERR_ARG, "null operator");
  op->io_pre_post = enable ? 1 : 0;
  return MF_OK;
}

int mf_inflow_outflow_mask(mf_scratch *s, const double *vel, unsigned sides, uint8_t *mask_host)
{
  MF_REQUIRE(s && vel && mask_host, MF_ERR_ARG, "null argument");
  MF_REQUIRE(!s->ad, MF_ERR_ARG, "inflow/outflow boundaries need a structured mesh");
  const int64_t n = s->mesh.n_dofs;
  uint8_t *d = nullptr;
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&d), size_t(n), s->stream));
  int rc = launch_inflow_mask(s, vel, sides, d);
  if (rc == MF_OK && hipMemcpyAsync(mask_host, d, size_t(n), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
    rc = MF_ERR_HIP;
  if (rc == MF_OK && hipStreamSynchronize(s->stream) != hipSuccess) rc = MF_ERR_HIP;
  (void)hipFreeAsync(d, s->stream);
  if (rc == MF_ERR_HIP) set_error("mf_inflow_outflow_mask: copy to the host failed");
  return rc;
}

int mf_operator_set_temperature_old(mf_operator *op, const double *T_old)
{
  MF_REQUIRE(op, MF_ERR_ARG, "null operator");
  MF_REQUIRE(op->kind == 0, MF_ERR_ARG, "only the heat operator has an old temperature");
  op->T_old = T_old;
  if (T_old) {
    const MeshInfo &m = op->s->mesh;
    if (!op->To32) MF_HIP_TRY(hipMalloc(&op->To32, sizeof(float) * m.n_dofs));
    return vec_cast<double, float>(T_old, op->To32, 1.0, m.n_dofs, op->s->stream);
  }
  return MF_OK;
}

int mf_operator_vmult(mf_operator *op, double *dst, const double *src)
{
  int rc = check_ready(op);
  if (rc) return rc;
  MF_REQUIRE(dst && src, MF_ERR_ARG, "null vector");
  return apply_tangent<double>(op, dst, src, nullptr, nullptr, true);
}

int mf_operator_vmult_add(mf_operator *op, double *dst, const double *src)
{
  int rc = check_ready(op);
  if (rc) return rc;
  MF_REQUIRE(dst && src, MF_ERR_ARG, "null vector");
  return apply_tangent<double>(op, dst, src, nullptr, nullptr, false);
}

int mf_operator_vmult_f32(mf_operator *op, float *dst, const float *src)
{
  int rc = check_ready(op);
  if (rc) return rc;
  MF_REQUIRE(dst && src, MF_ERR_ARG, "null vector");
  return apply_tangent<float>(op, dst, src, nullptr, nullptr, true);
}

}  // extern "C"

namespace mf {

// The body of mf_operator_compute_inverse_diagonal.  work: RED_BLOCKS doubles for the partial
// maxima of the relative threshold, or nullptr to take them from the operator's small-work arena
// (allocated on first use).  With work given, nothing here allocates device memory.
int compute_inverse_diagonal(mf_operator *op, double *inv_diag, double *work)
{
  int rc = check_ready(op);
  if (rc) return rc;
  MF_REQUIRE(inv_diag, MF_ERR_ARG, "null vector");
  const MeshInfo &m = op->s->mesh;
  hipStream_t st = op->s->stream;
  if (op->kind == 5) {
    // compute_inverse_diagonal_from_matrixfree (heat_cut_operator.cpp:2141-2153): relative
    // threshold 1e-14 max(1, |d|_inf)
    if ((rc = cut_diagonal(op, inv_diag))) return rc;
    if (!work) {
      if ((rc = arena_reserve(op->arena[3], sizeof(double) * RED_BLOCKS))) return rc;
      work = reinterpret_cast<double *>(op->arena[3].p);
    }
    return vec_invert_diagonal(inv_diag, m.n_dofs, true, work, st);
  }
  CellArgs<double> a = make_args<double>(op);
  const int code = op->kind == 0 ? HEAT_DIAG
                 : op->kind == 1 ? ADVEC_DIAG
                 : op->kind == 4 ? REINIT_DIAG
                                 : LS_DIAG;
  if (m.ad) {
    rc = vec_fill<double>(inv_diag, 0.0, m.n_dofs, st);
    if (rc) return rc;
    // cells without hanging nodes: local diagonal; cells with hanging nodes: condensed
    // diagonal c_D . (A_cell c_D) per target node D
    rc = launch_cell_list<double>(code, cell_flag(op), op->vel != nullptr, m, a, nullptr, inv_diag, st, 1);
    if (rc) return rc;
    const int tcode = op->kind == 0 ? HEAT_TANGENT
                      : op->kind == 1 ? ADVEC_TANGENT
                      : op->kind == 4 ? REINIT_TANGENT
                                      : LS_TANGENT;
    rc = launch_cell_list<double>(tcode, cell_flag(op), op->vel != nullptr, m, a, nullptr, inv_diag, st, 2);
  } else {
    a.overwrite = 1;
    rc = launch_cell<double>(code, cell_flag(op), op->vel != nullptr, m, a, nullptr, inv_diag, st);
    if (rc) return rc;
    rc = launch_heat_faces<double>(op, nullptr, inv_diag, 1, nullptr);
  }
  if (rc) return rc;
  // every block of the normal vector operator has the same diagonal
  for (int b = 1; b < n_blocks(op); ++b)
    MF_HIP_TRY(hipMemcpyAsync(inv_diag + b * m.n_dofs, inv_diag, sizeof(double) * m.n_dofs,
                              hipMemcpyDeviceToDevice, st));
  // heat: |d| > 1e-10; advection and the level-set projections: |d| > 1e-14 max(1, |d|_inf)
  const bool relative = op->kind != 0;
  if (relative && !work) {
    if ((rc = arena_reserve(op->arena[3], sizeof(double) * RED_BLOCKS))) return rc;
    work = reinterpret_cast<double *>(op->arena[3].p);
  }
  rc = vec_invert_diagonal(inv_diag, n_rows(op), relative, relative ? work : nullptr, st);
  if (rc || !io_active(op)) return rc;
  // compute_inverse_diagonal_from_matrixfree: 1 on the inflow/outflow rows
  return vec_masked_fill(inv_diag, op->io_mask, 1.0, m.n_dofs, st);
}

}  // namespace mf

extern "C" {

int mf_operator_compute_inverse_diagonal(mf_operator *op, double *inv_diag)
{
  return compute_inverse_diagonal(op, inv_diag, nullptr);
}

int mf_predictor_least_squares(mf_operator *op, int n_old, const double *const *old_solutions,
                               const double *rhs, double *solution, int *n_used)
{
  int rc = check_ready(op);
  if (rc) return rc;
  MF_REQUIRE(old_solutions && rhs && solution, MF_ERR_ARG, "null argument");
  MF_REQUIRE(n_old >= 1 && n_old <= LinComb::MAX, MF_ERR_ARG,
             "least-squares projection over 1..8 old solutions");
  for (int s = 0; s < n_old; ++s) MF_REQUIRE(old_solutions[s], MF_ERR_ARG, "null old solution");
  const int64_t n = n_rows(op);
  hipStream_t st = op->s->stream;
  // internal_vectors (A x_s, orthogonalised in place), the dot partials and their result
  Carve cv;
  for (int s = 0; s < n_old; ++s) cv.plan<double>(n);
  cv.plan<double>(2 * RED_BLOCKS);
  cv.plan<double>(1);
  if ((rc = cv.reserve(op->arena[1]))) return rc;
  std::vector<double *> v(n_old);
  for (int s = 0; s < n_old; ++s) v[s] = cv.take<double>(n);
  double *part = cv.take<double>(2 * RED_BLOCKS), *dres = cv.take<double>(1);
  auto dot = [&](const double *a, const double *b, double &out) -> int {
    int c = vec_dot<double>(a, b, n, part, dres, st);
    return c ? c : read_to_host(&out, dres, sizeof(double), st);
  };
  // projection_matrix P (n_old x n_old, row-major), modified Gram-Schmidt with add_and_dot
  std::vector<double> P(size_t(n_old) * n_old, 0.0);
  auto Pm = [&](int i, int j) -> double & { return P[size_t(i) * n_old + j]; };
  int step = 0;
  for (; step < n_old; ++step) {
    if ((rc = apply_tangent<double>(op, v[step], old_solutions[step], nullptr, nullptr, true)))
      return rc;
    if ((rc = dot(v[step], v[0], Pm(0, step)))) return rc;
    for (int j = 0; j < step; ++j) {
      if ((rc = vec_axpy_dot(v[step], -Pm(j, step) / Pm(j, j), v[j], v[j + 1], n, part, dres, st)))
        return rc;
      if ((rc = read_to_host(&Pm(j + 1, step), dres, sizeof(double), st))) return rc;
    }
    if (Pm(step, step) < 1e-16 * Pm(0, 0)) break;
  }
  MF_REQUIRE(step >= 1, MF_ERR_STATE,
             "least-squares projection: the first old solution has a zero operator image");
  // the least-squares system by back substitution
  std::vector<double> c(step);
  for (int s = 0; s < step; ++s)
    if ((rc = dot(v[s], rhs, c[s]))) return rc;
  for (int s = step - 1; s >= 0; --s) {
    double sum = c[s];
    for (int j = s + 1; j < step; ++j) sum -= c[j] * Pm(s, j);
    c[s] = sum / Pm(s, s);
  }
  LinComb lc{};
  lc.m = step;
  for (int s = 0; s < step; ++s) {
    lc.x[s] = old_solutions[s];
    lc.c[s] = c[s];
  }
  if ((rc = vec_lincomb(solution, lc, n, st))) return rc;
  if (n_used) *n_used = step;
  MF_HIP_TRY(hipStreamSynchronize(st));
  return MF_OK;
}

int mf_operator_create_rhs(mf_operator *op, double *dst, const double *src, const double *q_s)
{
  int rc = check_ready(op);
  if (rc) return rc;
  MF_REQUIRE(dst && src, MF_ERR_ARG, "null vector");
  if (op->kind == 5) {
    MF_REQUIRE(op->T, MF_ERR_STATE, "cut create_rhs needs T (mf_operator_set_fields)");
    MF_REQUIRE(!q_s, MF_ERR_ARG, "the cut heat operator's create_rhs takes no source term");
    return cut_residual(op, dst, src);
  }
  if (op->kind == 0) MF_REQUIRE(op->T, MF_ERR_STATE, "heat create_rhs needs T (mf_operator_set_fields)");
  MF_REQUIRE(op->kind == 0 || !q_s, MF_ERR_ARG, "only the heat create_rhs takes a source term");
  if (op->kind >= 2) {
    // normal vector: src = level set, dst [3][n_dofs]; curvature: src = normal vector
    // [3][n_dofs], dst [n_dofs]; reinitialization: src = level set, dst [n_dofs]
    return launch_levelset_rhs(op, op->kind == 2 ? 0 : (op->kind == 3 ? 1 : 2), src, dst);
  }
  CellArgs<double> a = make_args<double>(op);
  a.aux1 = q_s;
  if (op->s->mesh.ad) {
    // residual on an adaptive mesh: T, src, q_s and the velocity read as stored
    // (constraint-consistent vectors), each cell adds into its own nodes, then the hanging rows
    // (of the cell sums and of what dst held) go to their masters
    int rc3 = launch_cell_list<double>(op->kind == 0 ? HEAT_RHS : ADVEC_RHS, cell_flag(op),
                                       op->vel != nullptr, op->s->mesh, a, src, dst,
                                       op->s->stream, 0);
    if (rc3) return rc3;
    return launch_condense<double>(op->s, dst, nullptr, a.use_mask != 0);
  }
  const int code = op->kind == 0 ? HEAT_RHS : ADVEC_RHS;
  int rc2 = launch_cell<double>(code, cell_flag(op), op->vel != nullptr, op->s->mesh, a, src, dst,
                                op->s->stream);
  if (rc2) return rc2;
  return launch_heat_faces<double>(op, nullptr, dst, 2, nullptr);
}

int mf_operator_set_boundary(mf_operator *op, const mf_heat_boundary *b)
{
  MF_REQUIRE(op, MF_ERR_ARG, "null operator");
  if (!b) {
    op->bnd_set = false;
    return MF_OK;
  }
  MF_REQUIRE(op->kind == 0, MF_ERR_ARG, "boundary terms belong to the heat operator");
  MF_REQUIRE(!op->s->mesh.ad, MF_ERR_ARG, "boundary faces are not available on adaptive meshes");
  const unsigned all = b->convection_sides | b->radiation_sides | b->neumann_sides;
  MF_REQUIRE((all & ~0x3Fu) == 0, MF_ERR_ARG, "side bits beyond the 6 sides of the box");
  MF_REQUIRE((b->neumann_sides & (b->convection_sides | b->radiation_sides)) == 0, MF_ERR_ARG,
             "It is not allowed to specify both Neumann and radiation and/or convection boundary "
             "conditions at the same face.");
  MF_REQUIRE(b->convection_coefficient >= 0.0 && b->emissivity >= 0.0, MF_ERR_ARG,
             "convection coefficient and emissivity must be non-negative");
  op->bnd = *b;
  op->bnd_set = all != 0;
  return MF_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// solvers
// ------------------------------------------------------------------------------------------
namespace mf {


struct SolveOut {
  int first_chunk = 8;  // in: iterations enqueued before the first host check (CG)
  // in (FP32 CG of the defect correction): right-hand side = float(b64_scale * b64), x = 0
  const double *b64 = nullptr;
  double b64_scale = 1.0;
  const double *b64_den = nullptr;  // when set: the scale is 1 / *b64_den, read on the device
  int iterations = 0;
  double res0 = 0, res = 0;
  int status = MF_OK;
};

// Preconditioned CG on the device; T = double (FP64 operator) or float (FP32 operator).
template <typename T>
static int solve_cg(mf_operator *op, T *x, const T *b, const T *inv, int max_it, double abs_tol,
                    double rel_tol, SolveOut &out, bool x_zero)
{
  const int64_t n = n_rows(op);
  hipStream_t st = op->s->stream;
  // per-block partials of the fused p.(Ap): cell kernel blocks, then boundary-face blocks
  const int ncb = op_tangent_blocks(op, sizeof(T) == 4);
  Carve cv;
  cv.plan<T>(n);
  cv.plan<T>(n);
  cv.plan<T>(n);
  for (int i = 0; i < 3; ++i) cv.plan<T>(2 * RED_BLOCKS);
  cv.plan<double>(ncb);
  cv.plan<SolveState>(1);
  int rc = cv.reserve(op->arena[1]);
  if (rc) return rc;
  T *r = cv.take<T>(n), *p = cv.take<T>(n), *q = cv.take<T>(n);
  T *prr = cv.take<T>(2 * RED_BLOCKS), *prz0 = cv.take<T>(2 * RED_BLOCKS),
    *prz1 = cv.take<T>(2 * RED_BLOCKS);
  double *pcell = cv.take<double>(ncb);
  SolveState *sd = cv.take<SolveState>(1);
  // r0 = b - A x (x known to be zero: r0 = b, no operator application)
  if (!x_zero && (rc = apply_tangent<T>(op, q, x, nullptr, nullptr, true))) return rc;
  // cg_check0 initialises every field of the state block
  if constexpr (sizeof(T) == 4) {
    if (out.b64) {
      rc = cg_init_refine(out.b64, out.b64_scale, out.b64_den, x, r, p, inv, n, prr, prz0, sd,
                          abs_tol, rel_tol, st);
      if (rc) return rc;
    }
  }
  if (!(sizeof(T) == 4 && out.b64) &&
      (rc = cg_init<T>(b, x_zero ? nullptr : q, r, p, inv, n, prr, prz0, sd, abs_tol, rel_tol, st)))
    return rc;
  T *prz[2] = {prz0, prz1};
  SolveState h{};
  int it = 0;
  // Iterations are enqueued in chunks and the host reads the state after each chunk; kernels
  // launched after convergence (or after the step-0 check succeeded) exit at once.
  // One iteration is enqueued behind each state read (look-ahead), so the device keeps working
  // while the host waits for the read; after convergence it exits at once.
  constexpr int LOOKAHEAD = 1;
  auto enqueue = [&](int count) -> int {
    for (int k = 0; k < count && it < max_it; ++k) {
      ++it;
      int c;
      if ((c = apply_tangent<T>(op, q, p, pcell, &sd->stop, true))) return c;
      if ((c = cg_step<T>(x, r, p, q, inv, n, prz, prr, pcell, ncb, sd, it, max_it, st))) return c;
    }
    return MF_OK;
  };
  int chunk = std::max(1, out.first_chunk);
  if ((rc = enqueue(chunk))) return rc;
  for (;;) {
    const int it_read = it;
    if ((rc = read_to_host_begin(sd, sizeof(SolveState), st))) return rc;
    if ((rc = enqueue(LOOKAHEAD))) return rc;
    if ((rc = read_to_host_end(&h, sizeof(SolveState)))) return rc;
    if (h.stop) break;
    MF_REQUIRE(it_read < max_it, MF_ERR_STATE, "CG state not stopped at max_it");
    chunk = 8;
    if (h.iterations > 0 && h.res > 0 && h.res < h.res0 && h.tol > 0) {
      // iterations left at the observed average rate (plus one), so that few launches run
      // after convergence and few host round trips are needed
      const double rate = std::log(h.res / h.res0) / h.iterations;
      const double left = std::log(h.tol / h.res) / rate;
      chunk = int(std::min(64.0, std::max(2.0, std::ceil(left) + 1.0)));
    }
    if ((rc = enqueue(std::max(0, chunk - LOOKAHEAD)))) return rc;
  }
  out.iterations = h.iterations;
  out.res0 = h.res0;
  out.res = h.res;
  out.status = h.stop == 1 ? MF_OK : (h.stop == 2 ? MF_NO_CONVERGENCE : MF_BREAKDOWN);
  return MF_OK;
}

// Right-preconditioned GMRES(m) with classical Gram-Schmidt and deal.II's re-orthogonalisation
// rule: one pass, and a second one for this and every later basis vector of the solve once a
// pass has cancelled the vector below 10 sqrt(eps) of its norm (SolverGMRES's default, without
// force_re_orthogonalization).  The pass is fused with the dots of the second pass, so that the
// second pass costs one more read of the basis when it is needed.  Convergence measure: the
// Givens residual estimate |gamma_{j+1}| inside a cycle and the true residual at the start of
// each cycle (deal.II SolverGMRES with right_preconditioning = true, linear_solver.hpp:112-119).
template <typename T>
static int solve_gmres(mf_operator *op, T *x, const T *b, const T *inv, int restart, int max_it,
                       double abs_tol, double rel_tol, SolveOut &out, bool x_zero,
                       bool force_reorth)
{
  const int64_t n = n_rows(op);
  hipStream_t st = op->s->stream;
  const int mm = restart;  // 1..30, checked with the solver description
  Carve cv;
  cv.plan<T>(size_t(mm + 1) * n);
  cv.plan<T>(n);
  cv.plan<T>(size_t(32) * 2 * RED_BLOCKS);
  cv.plan<double>(128);
  int rc = cv.reserve(op->arena[1]);
  if (rc) return rc;
  struct {
    T *p;
  } V{cv.take<T>(size_t(mm + 1) * n)}, z{cv.take<T>(n)}, part{cv.take<T>(size_t(32) * 2 * RED_BLOCKS)};
  struct {
    double *p;
  } dh{cv.take<double>(128)};  // [0..31] h1, [32..63] h2, [64] norm, [65..96] y
  std::vector<double> H(size_t(mm + 1) * mm, 0.0), cs(mm), sn(mm), gam(mm + 1), y(mm);
  double hbuf[128];
  double tol = 0.0, res = 0.0, res0 = 0.0;
  int acc = 0, state = 0;  // 0 iterate, 1 success, 2 failure, 3 breakdown
  auto check = [&](int step, double v) {
    if (step == 0 && acc == 0) tol = std::max(abs_tol, rel_tol * v);
    if (v <= tol) return 1;
    if (step >= max_it || std::isnan(v)) return 2;
    return 0;
  };
  bool first = true;
  bool reorth = force_reorth;  // second Gram-Schmidt pass (sticky within the solve)
  const double cancel = 10.0 * std::sqrt(double(std::numeric_limits<T>::epsilon()));
  for (;;) {
    T *V0 = V.p;
    const bool r_is_b = first && x_zero;  // zero initial guess: r0 = b
    if (!r_is_b && (rc = apply_tangent<T>(op, z.p, x, nullptr, nullptr, true))) return rc;
    if ((rc = vec_residual<T>(b, r_is_b ? nullptr : z.p, V0, n, part.p, dh.p + 64, st))) return rc;
    if ((rc = read_to_host(hbuf + 64, dh.p + 64, sizeof(double), st))) return rc;
    double rho = hbuf[64];
    if (first) {
      res0 = rho;
      first = false;
    }
    res = rho;
    state = check(acc, rho);
    if (state) break;
    if ((rc = vec_scale_dev<T>(V0, V0, n, dh.p + 64, 1, st))) return rc;
    std::fill(gam.begin(), gam.end(), 0.0);
    gam[0] = rho;
    int j = 0;
    for (; j < mm && state == 0; ++j) {
      ++acc;
      T *vj = V.p + size_t(j) * n, *w = V.p + size_t(j + 1) * n;
      if ((rc = vec_precond<T>(z.p, vj, inv, n, st))) return rc;
      if ((rc = apply_tangent<T>(op, w, z.p, nullptr, nullptr, true))) return rc;
      if ((rc = gmres_multidot<T>(V.p, n, j + 1, w, n, part.p, dh.p, st))) return rc;
      if ((rc = gmres_multiaxpy_dot<T>(V.p, n, j + 1, dh.p, w, n, part.p, dh.p + 32, dh.p + 64,
                                       st)))
        return rc;
      if (reorth &&
          (rc = gmres_multiaxpy<T>(V.p, n, j + 1, dh.p + 32, w, n, part.p, dh.p + 64, st)))
        return rc;
      if ((rc = read_to_host(hbuf, dh.p, sizeof(double) * 65, st))) return rc;
      if (!reorth) {
        // |w| before the pass from its parts: |h|^2 + |w - V h|^2
        double start = hbuf[64] * hbuf[64];
        for (int i = 0; i <= j; ++i) start += hbuf[i] * hbuf[i];
        if (hbuf[64] < cancel * std::sqrt(start)) {
          reorth = true;
          if ((rc = gmres_multiaxpy<T>(V.p, n, j + 1, dh.p + 32, w, n, part.p, dh.p + 64, st)))
            return rc;
          if ((rc = read_to_host(hbuf + 64, dh.p + 64, sizeof(double), st))) return rc;
        } else {
          for (int i = 0; i <= j; ++i) hbuf[32 + i] = 0.0;
        }
      }
      std::vector<double> hc(j + 2);
      for (int i = 0; i <= j; ++i) hc[i] = hbuf[i] + hbuf[32 + i];
      const double s = hbuf[64];
      hc[j + 1] = s;
      if (s != 0.0) {
        if ((rc = vec_scale_dev<T>(w, w, n, dh.p + 64, 1, st))) return rc;
      }
      for (int i = 0; i < j; ++i) {
        const double t = cs[i] * hc[i] + sn[i] * hc[i + 1];
        hc[i + 1] = -sn[i] * hc[i] + cs[i] * hc[i + 1];
        hc[i] = t;
      }
      const double r = std::hypot(hc[j], hc[j + 1]);
      if (!(r > 0.0) || std::isnan(r)) {
        state = 3;
        break;
      }
      cs[j] = hc[j] / r;
      sn[j] = hc[j + 1] / r;
      hc[j] = r;
      gam[j + 1] = -sn[j] * gam[j];
      gam[j] = cs[j] * gam[j];
      for (int i = 0; i <= j; ++i) H[size_t(i) * mm + j] = hc[i];
      rho = std::fabs(gam[j + 1]);
      res = rho;
      state = check(acc, rho);
    }
    if (state == 3) break;
    for (int i = j - 1; i >= 0; --i) {
      double sum = gam[i];
      for (int l = i + 1; l < j; ++l) sum -= H[size_t(i) * mm + l] * y[l];
      y[i] = sum / H[size_t(i) * mm + i];
    }
    MF_HIP_TRY(hipMemcpyAsync(dh.p + 65, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, st));
    if ((rc = gmres_update<T>(x, V.p, n, j, dh.p + 65, inv, n, st))) return rc;
    MF_HIP_TRY(hipStreamSynchronize(st));
    if (state) break;
  }
  out.iterations = acc;
  out.res0 = res0;
  out.res = res;
  out.status = state == 1 ? MF_OK : (state == 2 ? MF_NO_CONVERGENCE : MF_BREAKDOWN);
  return MF_OK;
}

template <typename T>
static int krylov(mf_operator *op, const mf_solver_desc &d, T *x, const T *b, const T *inv,
                  double abs_tol, double rel_tol, int max_it, SolveOut &out, bool x_zero = false)
{
  if (d.solver == MF_SOLVER_CG)
    return solve_cg<T>(op, x, b, inv, max_it, abs_tol, rel_tol, out, x_zero);
  return solve_gmres<T>(op, x, b, inv, d.restart > 0 ? d.restart : 30, max_it, abs_tol, rel_tol,
                        out, x_zero, d.force_re_orthogonalization != 0);
}

}  // namespace mf

// Jacobi preconditioner object: the inverse diagonal of an operator, computed by update()
// (JacobiPreconditioner, preconditioner_jacobi.hpp:29-89), with its FP32 copy for the inner
// solves of the refined solver.
struct mf_jacobi {
  mf_operator *op = nullptr;
  int64_t n = 0;
  double *inv = nullptr;
  float *inv32 = nullptr;
  bool valid = false;
};

namespace mf {

// LinearSolver::solve with the inverse diagonal `inv_in` (null: identity) and, optionally, its
// FP32 copy (null: made here when the refined solver needs it)
static int solve_with(mf_operator *op, const mf_solver_desc *desc, const double *inv_in,
                      const float *inv32_in, double *x, const double *b, mf_solve_report *report)
{
  int rc = MF_OK;
  const int64_t n = n_rows(op);
  hipStream_t st = op->s->stream;
  const int max_it = desc->max_iterations > 0 ? desc->max_iterations : 10000;
  struct {
    const double *p = nullptr;
  } inv;
  inv.p = inv_in;
  mf_solve_report rep{};
  // a zero initial guess needs no operator application for r0 = b (SolverCG / SolverGMRES test
  // x.all_zero() the same way)
  if ((rc = arena_reserve(op->arena[3], sizeof(double) * (RED_BLOCKS + 1)))) return rc;
  bool x_zero = false;
  if ((rc = vec_all_zero<double>(x, n, reinterpret_cast<double *>(op->arena[3].p), x_zero, st)))
    return rc;
  if (desc->precision == MF_PRECISION_FP64) {
    SolveOut o;
    if ((rc = krylov<double>(op, *desc, x, b, inv.p, desc->abs_tol, desc->rel_tol, max_it, o,
                             x_zero)))
      return rc;
    rep.iterations = o.iterations;
    rep.initial_residual = o.res0;
    rep.final_residual = o.res;
    rep.status = o.status;
  } else {
    // FP32 inner Krylov inside FP64 defect correction (iterative refinement)
    Carve cv;
    cv.plan<double>(n);
    cv.plan<double>(n);
    cv.plan<double>(2);
    cv.plan<double>(2 * RED_BLOCKS);
    cv.plan<float>(n);
    cv.plan<float>(n);
    cv.plan<float>(n);
    if ((rc = cv.reserve(op->arena[0]))) return rc;
    struct D {
      double *p;
    };
    struct F {
      float *p;
    };
    D r{cv.take<double>(n)}, q{cv.take<double>(n)}, nrm{cv.take<double>(2)},
        part{cv.take<double>(2 * RED_BLOCKS)};
    F r32{cv.take<float>(n)}, e32{cv.take<float>(n)}, inv32{cv.take<float>(n)};
    const float *inv32p = inv32_in;
    if (inv.p && !inv32p) {
      if ((rc = vec_cast<double, float>(inv.p, inv32.p, 1.0, n, st))) return rc;
      inv32p = inv32.p;
    }
    // inner tolerance: by default the square root of the outer one (two passes), but not below
    // 1e-5, about where FP32 Krylov iterates stop improving the FP64 residual
    const double inner_rel = desc->inner_rel_tol > 0
                                 ? desc->inner_rel_tol
                                 : std::min(1e-2, std::max(1e-5, std::sqrt(std::max(desc->rel_tol, 0.0))));
    const int max_outer = desc->max_outer > 0 ? desc->max_outer : 50;
    // r = b - A x and |r| into the device slot nrm[k]
    auto residual = [&](bool xz, double *slot) -> int {
      int c;
      if (!xz && (c = apply_tangent<double>(op, q.p, x, nullptr, nullptr, true))) return c;
      return vec_residual<double>(b, xz ? nullptr : q.p, r.p, n, part.p, slot, st);
    };
    // The inner CG scales its right-hand side and the correction by |r| read on the device, so
    // a pass is enqueued without waiting for the norm; the host reads the norms after the pass
    // (both at once after the first).  GMRES takes the scale from the host.
    const bool dev_scale = desc->solver == MF_SOLVER_CG;
    double res = 0.0, res0 = 0.0, tol = 0.0;
    bool known = false;
    int slot = 0;
    if ((rc = residual(x_zero, nrm.p))) return rc;
    if (!dev_scale) {
      if ((rc = read_to_host(&res, nrm.p, sizeof(double), st))) return rc;
      res0 = res;
      tol = std::max(desc->abs_tol, desc->rel_tol * res0);
      known = true;
    }
    int outer = 0, inner_total = 0, last_inner = 0, status = MF_OK;
    for (;;) {
      if (known) {
        if (res <= tol) break;
        if (outer >= max_outer || std::isnan(res)) {
          status = MF_NO_CONVERGENCE;
          break;
        }
      }
      ++outer;
      SolveOut o;
      if (dev_scale) {
        // the inner CG starts from float(r / |r|) and e = 0 in its first kernel
        o.b64 = r.p;
        o.b64_den = nrm.p + slot;
      } else {
        if ((rc = vec_cast<double, float>(r.p, r32.p, 1.0 / res, n, st))) return rc;
        if ((rc = vec_fill<float>(e32.p, 0.0f, n, st))) return rc;
      }
      // the passes reduce the residual by about the same factor: enqueue the previous pass's
      // iteration count before the first check
      o.first_chunk = inner_total > 0 ? last_inner + 1 : 8;
      if ((rc = krylov<float>(op, *desc, e32.p, r32.p, inv32p, 0.0, inner_rel, max_it, o, true)))
        return rc;
      inner_total += o.iterations;
      last_inner = o.iterations;
      // the first pass of the device-scaled CG was queued before |r0| was known: it changes
      // nothing when |r0| already meets the tolerance (|r0| <= abs_tol, or rel_tol >= 1) or is
      // not finite, and the report then shows the state before it (ReductionControl stops after
      // 0 steps, or fails on a NaN residual)
      const double skip = (dev_scale && !known)
                              ? (desc->rel_tol >= 1.0 ? HUGE_VAL : desc->abs_tol)
                              : -HUGE_VAL;
      auto undo_first_pass = [&](double r0) {
        res0 = res = r0;
        tol = std::max(desc->abs_tol, desc->rel_tol * res0);
        outer = 0;
        inner_total = 0;
        status = std::isfinite(r0) ? MF_OK : MF_NO_CONVERGENCE;
      };
      if (o.status == MF_BREAKDOWN) {
        if (!known) {
          double r0 = 0.0;
          if ((rc = read_to_host(&r0, nrm.p, sizeof(double), st))) return rc;
          if (!(r0 > skip) || !std::isfinite(r0)) {
            undo_first_pass(r0);
            break;
          }
        }
        status = MF_BREAKDOWN;
        break;
      }
      // x += |r| e (skipped on the device in the case above)
      if ((rc = vec_add_scaled_f32(x, e32.p, res, dev_scale ? nrm.p + slot : nullptr, skip, n,
                                   st)))
        return rc;
      if ((rc = residual(false, nrm.p + (slot ^ 1)))) return rc;
      if (!known) {
        double h2[2];
        if ((rc = read_to_host(h2, nrm.p, 2 * sizeof(double), st))) return rc;
        res0 = h2[0];
        res = h2[1];
        tol = std::max(desc->abs_tol, desc->rel_tol * res0);
        known = true;
        if (!(res0 > skip) || !std::isfinite(res0)) {
          undo_first_pass(res0);  // x unchanged
          break;
        }
      } else if ((rc = read_to_host(&res, nrm.p + (slot ^ 1), sizeof(double), st))) {
        return rc;
      }
      slot ^= 1;
    }
    rep.iterations = inner_total;
    rep.outer_iterations = outer;
    rep.initial_residual = res0;
    rep.final_residual = res;
    rep.status = status;
  }
  MF_HIP_TRY(hipStreamSynchronize(st));
  if (report) *report = rep;
  if (rep.status != MF_OK) {
    set_error(rep.status == MF_NO_CONVERGENCE ? "linear solver did not converge (NoConvergence)"
                                              : "Krylov breakdown");
    return rep.status;
  }
  return MF_OK;
}

static int check_solve_args(mf_operator *op, const mf_solver_desc *desc, double *x, const double *b)
{
  int rc = check_ready(op);
  if (rc) return rc;
  MF_REQUIRE(desc && x && b, MF_ERR_ARG, "null argument");
  MF_REQUIRE(desc->solver == MF_SOLVER_CG || desc->solver == MF_SOLVER_GMRES, MF_ERR_ARG,
             "unknown solver type");
  MF_REQUIRE(desc->restart >= 0 && desc->restart <= 30, MF_ERR_ARG,
             "GMRES restart (basis size) must lie in 1..30 (0 = 30)");
  MF_REQUIRE(op->kind != 5 || desc->precision == MF_PRECISION_FP64, MF_ERR_ARG,
             "the cut heat operator has no FP32 form (precision must be MF_PRECISION_FP64)");
  return MF_OK;
}

}  // namespace mf

extern "C" int mf_solve(mf_operator *op, const mf_solver_desc *desc, double *x, const double *b,
                        mf_solve_report *report)
{
  int rc = check_solve_args(op, desc, x, b);
  if (rc) return rc;
  const int64_t n = n_rows(op);
  double *inv = nullptr;
  if (desc->preconditioner == MF_PRECOND_JACOBI) {
    if ((rc = arena_reserve(op->arena[2], sizeof(double) * std::max<int64_t>(n, RED_BLOCKS))))
      return rc;
    inv = reinterpret_cast<double *>(op->arena[2].p);
    if ((rc = mf_operator_compute_inverse_diagonal(op, inv))) return rc;
  }
  return solve_with(op, desc, inv, nullptr, x, b, report);
}

extern "C" int mf_jacobi_create(mf_operator *op, mf_jacobi **out)
{
  MF_REQUIRE(op && out, MF_ERR_ARG, "null argument");
  auto *pc = new mf_jacobi;
  pc->op = op;
  pc->n = n_rows(op);
  // the inversion of the advection diagonal reduces over a few blocks at the end of the array
  const size_t nd = size_t(std::max<int64_t>(pc->n, RED_BLOCKS));
  if (hipMalloc(&pc->inv, sizeof(double) * nd) != hipSuccess ||
      hipMalloc(&pc->inv32, sizeof(float) * size_t(pc->n)) != hipSuccess) {
    if (pc->inv) (void)hipFree(pc->inv);
    delete pc;
    set_error("hipMalloc failed for the Jacobi preconditioner");
    return MF_ERR_HIP;
  }
  *out = pc;
  return MF_OK;
}

extern "C" int mf_jacobi_update(mf_jacobi *pc)
{
  MF_REQUIRE(pc, MF_ERR_ARG, "null preconditioner");
  pc->valid = false;
  int rc = mf_operator_compute_inverse_diagonal(pc->op, pc->inv);
  if (rc) return rc;
  if ((rc = vec_cast<double, float>(pc->inv, pc->inv32, 1.0, pc->n, pc->op->s->stream))) return rc;
  pc->valid = true;
  return MF_OK;
}

extern "C" int mf_jacobi_vmult(mf_jacobi *pc, double *dst, const double *src)
{
  MF_REQUIRE(pc && dst && src, MF_ERR_ARG, "null argument");
  MF_REQUIRE(pc->valid, MF_ERR_STATE, "Jacobi preconditioner not updated (mf_jacobi_update)");
  return vec_precond<double>(dst, src, pc->inv, pc->n, pc->op->s->stream);
}

extern "C" int mf_jacobi_inverse_diagonal(mf_jacobi *pc, double *dst)
{
  MF_REQUIRE(pc && dst, MF_ERR_ARG, "null argument");
  MF_REQUIRE(pc->valid, MF_ERR_STATE, "Jacobi preconditioner not updated (mf_jacobi_update)");
  MF_HIP_TRY(hipMemcpyAsync(dst, pc->inv, sizeof(double) * size_t(pc->n), hipMemcpyDeviceToDevice,
                            pc->op->s->stream));
  return MF_OK;
}

extern "C" int mf_jacobi_destroy(mf_jacobi *pc)
{
  if (!pc) return MF_OK;
  if (pc->inv) (void)hipFree(pc->inv);
  if (pc->inv32) (void)hipFree(pc->inv32);
  delete pc;
  return MF_OK;
}

extern "C" int mf_solve_preconditioned(mf_operator *op, const mf_solver_desc *desc, mf_jacobi *pc,
                                       double *x, const double *b, mf_solve_report *report)
{
  int rc = check_solve_args(op, desc, x, b);
  if (rc) return rc;
  if (pc) {
    MF_REQUIRE(pc->valid, MF_ERR_STATE, "Jacobi preconditioner not updated (mf_jacobi_update)");
    MF_REQUIRE(pc->n == n_rows(op), MF_ERR_ARG,
               "preconditioner and operator have different DoF counts");
  }
  return solve_with(op, desc, pc ? pc->inv : nullptr, pc ? pc->inv32 : nullptr, x, b, report);
}

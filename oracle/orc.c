/*
 * orc.c -- CPU oracle (test infrastructure only; see orc.h for the contract and citations).
 *
 * Structure: one generic cell routine per operator, dimension-generic tensor contractions on
 * K^dim arrays (K = p+1 nodes and quadrature points per direction), a colour-by-parity OpenMP
 * cell loop for the scatter, and the Krylov solvers of the reference's LinearSolver::solve.
 * Gradients are evaluated directly, (DS (x) S (x) S) u etc., not by collocation, so that this
 * path differs from the device kernels in its rounding.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KMAX 8
#define NLOC_MAX (KMAX * KMAX * KMAX)

/* ------------------------------------------------------------------------------------------ */
/* 1D basis                                                                                    */
/* ------------------------------------------------------------------------------------------ */

/* Legendre P_n(z) and P_{n-1}(z) by the three-term recurrence. */
static void legendre(int n, double z, double *pn, double *pnm1)
{
  double p0 = 1.0, p1 = 0.0;
  for (int j = 1; j <= n; ++j) {
    const double p2 = p1;
    p1 = p0;
    p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
  }
  *pn = p0;
  *pnm1 = p1;
}

/* QGauss(n) on [0,1]: Newton iteration on P_n (deal.II QGauss, used as QGauss(p+1) by the
 * reference: source/core/finite_element_data.cpp:85-89). */
void orc_gauss(int n, double *x, double *w)
{
  for (int i = 0; i < n; ++i) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double pn, pm, dp;
    for (int it = 0; it < 100; ++it) {
      legendre(n, z, &pn, &pm);
      dp = n * (z * pn - pm) / (z * z - 1.0);
      const double dz = pn / dp;
      z -= dz;
      if (fabs(dz) < 1e-17) break;
    }
    legendre(n, z, &pn, &pm);
    dp = n * (z * pn - pm) / (z * z - 1.0);
    /* z runs from +1 down to -1; store ascending on [0,1] */
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
}

/* QGaussLobatto(n) on [0,1]: end points plus the roots of P'_{n-1} (FE_Q support points). */
void orc_gauss_lobatto(int n, double *x)
{
  const int p = n - 1;
  x[0] = 0.0;
  x[n - 1] = 1.0;
  for (int i = 1; i < p; ++i) {
    double z = cos(M_PI * i / p);
    for (int it = 0; it < 100; ++it) {
      double pn, pm;
      legendre(p, z, &pn, &pm);
      const double d1 = p * (z * pn - pm) / (z * z - 1.0);
      const double d2 = (2.0 * z * d1 - p * (p + 1.0) * pn) / (1.0 - z * z);
      const double dz = d1 / d2;
      z -= dz;
      if (fabs(dz) < 1e-17) break;
    }
    x[i] = 0.5 * (1.0 - z);
  }
  /* symmetrise exactly like a symmetric point set */
  for (int i = 0; i < n / 2; ++i) {
    const double a = 0.5 * (x[i] + (1.0 - x[n - 1 - i]));
    x[i] = a;
    x[n - 1 - i] = 1.0 - a;
  }
  if (n % 2 == 1) x[n / 2] = 0.5;
}

void orc_lagrange(int n, const double *nodes, double x, double *val, double *der)
{
  for (int j = 0; j < n; ++j) {
    double v = 1.0, d = 0.0;
    for (int m = 0; m < n; ++m) {
      if (m == j) continue;
      const double inv = 1.0 / (nodes[j] - nodes[m]);
      d = d * (x - nodes[m]) * inv + v * inv;
      v *= (x - nodes[m]) * inv;
    }
    val[j] = v;
    der[j] = d;
  }
}

/* second derivatives of the Lagrange polynomials on `nodes` at x, from the product formula
 * l_j'' = sum_{m != j} sum_{k != j, m} 1/((x_j - x_m)(x_j - x_k)) prod_{l != j, m, k}
 * (x - x_l)/(x_j - x_l) */
static void lagrange_d2(int n, const double *nodes, double x, double *d2)
{
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int m = 0; m < n; ++m) {
      if (m == j) continue;
      for (int k = 0; k < n; ++k) {
        if (k == j || k == m) continue;
        double t = 1.0 / ((nodes[j] - nodes[m]) * (nodes[j] - nodes[k]));
        for (int l = 0; l < n; ++l)
          if (l != j && l != m && l != k) t *= (x - nodes[l]) / (nodes[j] - nodes[l]);
        s += t;
      }
    }
    d2[j] = s;
  }
}

/* S[q][j] = l_j(x_q), DS[q][j] = l_j'(x_q), Dq[q][r] = L_r'(x_q) with L the Lagrange basis on the
 * Gauss points (collocation derivative), w = Gauss weights.  K = degree + 1. */
void orc_shape_matrices(int degree, double *S, double *DS, double *Dq, double *w)
{
  const int K = degree + 1;
  double gl[KMAX], xq[KMAX], v[KMAX], d[KMAX];
  orc_gauss_lobatto(K, gl);
  orc_gauss(K, xq, w);
  for (int q = 0; q < K; ++q) {
    orc_lagrange(K, gl, xq[q], v, d);
    for (int j = 0; j < K; ++j) {
      S[q * K + j] = v[j];
      DS[q * K + j] = d[j];
    }
    if (Dq) {
      orc_lagrange(K, xq, xq[q], v, d);
      for (int j = 0; j < K; ++j) Dq[q * K + j] = d[j];
    }
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Mesh                                                                                        */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
  int dim, K, p, nloc;
  long NN[3], stride[3], ncell[3];
  double h[3], lo[3];
  double S[KMAX * KMAX], DS[KMAX * KMAX], wq[KMAX], xq[KMAX], gl[KMAX];
  double DDS[KMAX * KMAX];  /* DDS[q][j] = l_j''(x_q) */
} geo_t;

static void tensor_apply(const geo_t *g, const double *const *mats, int transpose,
                         const double *in, double *out, double *t1, double *t2);

static void geo_init(const orc_mesh *m, geo_t *g)
{
  g->dim = m->dim;
  g->p = m->degree;
  g->K = m->degree + 1;
  g->nloc = 1;
  long s = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < m->dim) {
      g->ncell[d] = m->ncell[d];
      g->NN[d] = (long)m->ncell[d] * m->degree + 1;
      g->h[d] = (m->hi[d] - m->lo[d]) / m->ncell[d];
      g->lo[d] = m->lo[d];
      g->nloc *= g->K;
    } else {
      g->ncell[d] = 1;
      g->NN[d] = 1;
      g->h[d] = 1.0;
      g->lo[d] = 0.0;
    }
    g->stride[d] = s;
    s *= g->NN[d];
  }
  orc_shape_matrices(m->degree, g->S, g->DS, NULL, g->wq);
  double w[KMAX];
  orc_gauss(g->K, g->xq, w);
  orc_gauss_lobatto(g->K, g->gl);
  for (int q = 0; q < g->K; ++q) lagrange_d2(g->K, g->gl, g->xq[q], g->DDS + q * g->K);
}

/* cell size of the stabilisation parameter: diameter / sqrt(dim) (ScratchData::build,
 * source/core/scratch_data.cpp:256-278) */
static double cell_size(const geo_t *g)
{
  double d2 = 0.0;
  for (int d = 0; d < g->dim; ++d) d2 += g->h[d] * g->h[d];
  return sqrt(d2 / g->dim);
}

/* SUPG parameter of a cell from the velocity at its quadrature points w[e][q]
 * (compute_stabilization_parameter, advection_diffusion_operator.cpp:497-530) */
static double supg_tau(const orc_operator *op, const geo_t *g, const double (*w)[NLOC_MAX])
{
  const orc_advec_params *ap = &op->advec;
  double t;
  if (ap->supg_coefficient < 0.0) {
    double mx = 0.0;
    for (int q = 0; q < g->nloc; ++q) {
      double n2 = 0.0;
      for (int e = 0; e < g->dim; ++e) n2 += w[e][q] * w[e][q];
      mx = fmax(mx, sqrt(n2));
    }
    t = mx >= 1e-6 ? 1.0 / mx : 0.0;
  } else {
    t = ap->supg_coefficient;
  }
  return t * cell_size(g);
}

/* Laplacian at the quadrature points of the cell function with nodal values u */
static void eval_laplacian(const geo_t *g, const double *u, double *lap, double *tmp,
                           double *t1, double *t2)
{
  for (int q = 0; q < g->nloc; ++q) lap[q] = 0.0;
  for (int e = 0; e < g->dim; ++e) {
    const double *mats[3];
    for (int d = 0; d < 3; ++d) mats[d] = (d == e) ? g->DDS : g->S;
    tensor_apply(g, mats, 0, u, tmp, t1, t2);
    const double s = 1.0 / (g->h[e] * g->h[e]);
    for (int q = 0; q < g->nloc; ++q) lap[q] += s * tmp[q];
  }
}

long orc_n_dofs(const orc_mesh *m)
{
  long n = 1;
  for (int d = 0; d < m->dim; ++d) n *= (long)m->ncell[d] * m->degree + 1;
  return n;
}

long orc_n_cells(const orc_mesh *m)
{
  long n = 1;
  for (int d = 0; d < m->dim; ++d) n *= m->ncell[d];
  return n;
}

static double node_coord(const geo_t *g, int d, long i)
{
  long c = i / g->p;
  if (c >= g->ncell[d]) c = g->ncell[d] - 1;
  const long a = i - c * g->p;
  return g->lo[d] + g->h[d] * ((double)c + g->gl[a]);
}

void orc_node_coords(const orc_mesh *m, double *xyz)
{
  geo_t g;
  geo_init(m, &g);
  const long n = orc_n_dofs(m);
  for (long idx = 0; idx < n; ++idx) {
    long r = idx;
    for (int d = 0; d < 3; ++d) {
      const long i = r % g.NN[d];
      r /= g.NN[d];
      xyz[3 * idx + d] = d < m->dim ? node_coord(&g, d, i) : 0.0;
    }
  }
}

void orc_boundary_mask(const orc_mesh *m, uint8_t *mask)
{
  geo_t g;
  geo_init(m, &g);
  const long n = orc_n_dofs(m);
  for (long idx = 0; idx < n; ++idx) {
    long r = idx;
    uint8_t b = 0;
    for (int d = 0; d < m->dim; ++d) {
      const long i = r % g.NN[d];
      r /= g.NN[d];
      if (i == 0 || i == g.NN[d] - 1) b = 1;
    }
    mask[idx] = b;
  }
}

uint64_t orc_splitmix64(uint64_t x)
{
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* U(-1,1) keyed by the node's lattice coordinates (i0, i1, i2) and the seed; identical for any
 * DoF numbering. */
void orc_random_field(const orc_mesh *m, uint64_t seed, double *out)
{
  geo_t g;
  geo_init(m, &g);
  const long n = orc_n_dofs(m);
  const uint64_t s = orc_splitmix64(seed);
  for (long idx = 0; idx < n; ++idx) {
    long r = idx;
    uint64_t key = 0;
    for (int d = 0; d < 3; ++d) {
      const uint64_t i = (uint64_t)(r % g.NN[d]);
      r /= g.NN[d];
      key |= i << (21 * d);
    }
    const uint64_t z = orc_splitmix64(key ^ s);
    out[idx] = 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Tensor contractions on K^dim arrays                                                          */
/* ------------------------------------------------------------------------------------------ */

/* out_i = sum_j M[i][j] in_j along direction dir (transpose: out_j = sum_i M[i][j] in_i). */
static void contract(const geo_t *g, int dir, const double *M, int transpose, const double *in,
                     double *out)
{
  const int K = g->K;
  long stride = 1;
  for (int d = 0; d < dir; ++d) stride *= K;
  const long outer = g->nloc / (stride * K);
  for (long o = 0; o < outer; ++o)
    for (long s = 0; s < stride; ++s) {
      const long base = o * stride * K + s;
      for (int i = 0; i < K; ++i) {
        double acc = 0.0;
        for (int j = 0; j < K; ++j)
          acc += (transpose ? M[j * K + i] : M[i * K + j]) * in[base + j * stride];
        out[base + i * stride] = acc;
      }
    }
}

/* Apply one 1D matrix per direction: mats[d] along direction d. */
static void tensor_apply(const geo_t *g, const double *const *mats, int transpose,
                         const double *in, double *out, double *t1, double *t2)
{
  const double *src = in;
  for (int d = 0; d < g->dim; ++d) {
    double *dst = (d == g->dim - 1) ? out : ((d % 2 == 0) ? t1 : t2);
    contract(g, d, mats[d], transpose, src, dst);
    src = dst;
  }
}

typedef struct {
  double t1[NLOC_MAX], t2[NLOC_MAX], t3[NLOC_MAX];
} work_t;

/* values and physical gradients at the quadrature points */
static void eval_vg(const geo_t *g, const double *u, double *val, double *grad, work_t *w)
{
  const double *mats[3];
  for (int d = 0; d < 3; ++d) mats[d] = g->S;
  tensor_apply(g, mats, 0, u, val, w->t1, w->t2);
  for (int e = 0; e < g->dim; ++e) {
    for (int d = 0; d < 3; ++d) mats[d] = (d == e) ? g->DS : g->S;
    double *ge = grad + (long)e * g->nloc;
    tensor_apply(g, mats, 0, u, ge, w->t1, w->t2);
    const double s = 1.0 / g->h[e];
    for (int q = 0; q < g->nloc; ++q) ge[q] *= s;
  }
}

static void eval_v(const geo_t *g, const double *u, double *val, work_t *w)
{
  const double *mats[3] = {g->S, g->S, g->S};
  tensor_apply(g, mats, 0, u, val, w->t1, w->t2);
}

/* out = sum_q val_q phi(q) + sum_q flux_q . grad phi(q); val and flux already carry JxW */
static void integrate(const geo_t *g, const double *val, const double *flux, double *out,
                      work_t *w)
{
  const double *mats[3] = {g->S, g->S, g->S};
  tensor_apply(g, mats, 1, val, out, w->t1, w->t2);
  if (!flux) return;
  for (int e = 0; e < g->dim; ++e) {
    for (int d = 0; d < 3; ++d) mats[d] = (d == e) ? g->DS : g->S;
    tensor_apply(g, mats, 1, flux + (long)e * g->nloc, w->t3, w->t1, w->t2);
    const double s = 1.0 / g->h[e];
    for (int i = 0; i < g->nloc; ++i) out[i] += s * w->t3[i];
  }
}

static void cell_coords(const geo_t *g, long cell, long *c)
{
  long r = cell;
  for (int d = 0; d < 3; ++d) {
    c[d] = r % g->ncell[d];
    r /= g->ncell[d];
  }
}

static void cell_dofs(const geo_t *g, long cell, long *idx)
{
  long c[3];
  cell_coords(g, cell, c);
  const int K = g->K;
  for (int l = 0; l < g->nloc; ++l) {
    int a[3] = {l % K, (l / K) % K, l / (K * K)};
    long gi = 0;
    for (int d = 0; d < g->dim; ++d) gi += (c[d] * g->p + a[d]) * g->stride[d];
    idx[l] = gi;
  }
}

static double cell_volume_weight(const geo_t *g, int q)
{
  const int K = g->K;
  double w = 1.0;
  int r = q;
  for (int d = 0; d < g->dim; ++d) {
    w *= g->wq[r % K] * g->h[d];
    r /= K;
  }
  return w;
}

static void q_point(const geo_t *g, long cell, int q, double *x)
{
  long c[3];
  cell_coords(g, cell, c);
  const int K = g->K;
  int r = q;
  for (int d = 0; d < 3; ++d) {
    x[d] = d < g->dim ? g->lo[d] + g->h[d] * ((double)c[d] + g->xq[r % K]) : 0.0;
    r /= K;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Material law (reference slot: get_material_parameters_and_derivatives,                      */
/* source/heat/heat_diffuse_operator.cpp:1167-1210)                                            */
/* ------------------------------------------------------------------------------------------ */

/* Mushy-zone blend of one property (Material::compute_solid_liquid_phases_property and its
 * temperature derivative, material.templates.hpp:546-600; interpolate_cubic and its derivative,
 * level_set_tools.hpp:72-89): value l + (s - l)(-2 f^3 + 3 f^2) (l at f = 0, s at f = 1 exactly)
 * and d/dT = -(s - l)(-6 f^2 + 6 f) / (T_liq - T_sol), zero at f = 0 and f = 1. */
static void mushy_blend(double f, double inv_mushy, double l, double s, double *v, double *dv)
{
  if (f == 0.0) {
    *v = l;
    *dv = 0.0;
  } else if (f == 1.0) {
    *v = s;
    *dv = 0.0;
  } else {
    *v = l + (s - l) * (-2.0 * f * f * f + 3.0 * f * f);
    *dv = -1.0 * inv_mushy * ((s - l) * (-6.0 * f * f + 6.0 * f));
  }
}

static void material(const orc_heat_params *hp, double T, double *rc, double *k, double *drc,
                     double *dk)
{
  *rc = hp->rho_cp;
  *drc = 0.0;
  if (hp->law == ORC_LAW_TANH) {
    const double x = (T - hp->tanh_Tc) / hp->tanh_width;
    const double t = tanh(x);
    const double c = cosh(x);
    *k = hp->k + hp->tanh_amp * t;
    *dk = hp->tanh_amp / (hp->tanh_width * c * c);
  } else if (hp->law == ORC_LAW_MUSHY) {
    /* compute_temperature_dependent_solid_fraction (material.templates.hpp:748-760) with
     * limit_to_bounds (utility_functions.hpp:62-82) */
    const double inv_mushy = 1.0 / (hp->liquidus - hp->solidus);
    double f = (hp->liquidus - T) * inv_mushy;
    f = f > 1.0 ? 1.0 : f;
    f = f < 0.0 ? 0.0 : f;
    double rho, drho, cp, dcp;
    mushy_blend(f, inv_mushy, hp->liquid_k, hp->solid_k, k, dk);
    mushy_blend(f, inv_mushy, hp->liquid_rho, hp->solid_rho, &rho, &drho);
    mushy_blend(f, inv_mushy, hp->liquid_cp, hp->solid_cp, &cp, &dcp);
    /* get_material_parameters_and_derivatives (heat_diffuse_operator.cpp:1177-1190) */
    *rc = rho * cp;
    *drc = dcp * rho + drho * cp;
  } else {
    *k = hp->k;
    *dk = 0.0;
  }
}

void orc_material(const orc_heat_params *hp, double T, double *out)
{
  material(hp, T, &out[0], &out[1], &out[2], &out[3]);
}

/* ------------------------------------------------------------------------------------------ */
/* Cell operations                                                                             */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
  double u[NLOC_MAX], out[NLOC_MAX];
  double val[NLOC_MAX], grad[3 * NLOC_MAX];
  double Tv[NLOC_MAX], Tg[3 * NLOC_MAX], To[NLOC_MAX];
  double wv[3][NLOC_MAX];
  double fval[NLOC_MAX], flux[3 * NLOC_MAX];
  double tmp[NLOC_MAX], lap[NLOC_MAX];
  long idx[NLOC_MAX];
  work_t w;
} cell_ws_t;

static int has_vel(const orc_operator *op) { return op->vel_kind != ORC_VEL_NONE; }

/* number of DoF entries of the operator's vectors (structured or adaptive mesh) */
static long op_ndofs(const orc_operator *op)
{
  return op->amesh ? op->adata->n_nodes : orc_n_dofs(&op->mesh);
}

static int op_dim(const orc_operator *op) { return op->amesh ? 3 : op->mesh.dim; }

long orc_n_rows(const orc_operator *op)
{
  return (op->kind == ORC_OP_NORMAL ? op_dim(op) : 1) * op_ndofs(op);
}

/* filter parameter eta = scale * (cell diameter / sqrt(dim))^2 of the cell with geometry g
 * (compute_cell_size_dependent_filter_parameter_mf, normal_vector_operator.hpp:95-112: the
 * reference takes max(min cell size, cell size), which is the cell size of every cell) */
static double ls_damping(const orc_operator *op, const geo_t *g)
{
  double d2 = 0.0;
  for (int d = 0; d < g->dim; ++d) d2 += g->h[d] * g->h[d];
  return op->ls.filter_parameter * d2 / g->dim;
}

/* interface thickness eps_n of a cell: compute_interface_thickness_parameter_epsilon of the cell
 * size (diameter / sqrt(dim); n_subdivisions = 1) (reinitialization_data.hpp:181-197,
 * olsson_CG_operator.cpp:733-745) */
static double reinit_eps(const orc_operator *op, const geo_t *g)
{
  double d2 = 0.0;
  for (int d = 0; d < g->dim; ++d) d2 += g->h[d] * g->h[d];
  const double cs = sqrt(d2 / g->dim);
  const orc_reinit_params *r = &op->reinit;
  if (r->thickness_type == ORC_THICKNESS_ABSOLUTE) return r->thickness_value;
  if (r->thickness_type == ORC_THICKNESS_CELLS) return r->thickness_value * cs / 6.0;
  return r->thickness_value * cs;
}

/* normalize<dim>(v, tol) (include/meltpooldg/utilities/dealii_tensor.hpp:173-191) */
static void normalize_tol(double *v, int dim, double tol)
{
  double nrm = 0.0;
  for (int d = 0; d < dim; ++d) nrm += v[d] * v[d];
  nrm = sqrt(nrm);
  for (int d = 0; d < dim; ++d) v[d] = nrm > tol ? v[d] / nrm : v[d] / tol;
}

/* narrow-band indicator (Tools::compute_mask_narrow_band, level_set_tools.hpp:359-368) */
static double ls_mask(const orc_operator *op, double phi)
{
  if (!op->ls.narrow_band) return 1.0;
  return fabs(phi) >= op->ls.narrow_band_threshold ? 0.0 : 1.0;
}

/* velocity at the quadrature points of a cell */
static void eval_velocity(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws)
{
  const long n = op_ndofs(op);
  for (int e = 0; e < 3; ++e)
    for (int q = 0; q < g->nloc; ++q) ws->wv[e][q] = 0.0;
  if (op->vel_kind == ORC_VEL_NODAL && op->vel) {
    for (int e = 0; e < g->dim; ++e) {
      for (int l = 0; l < g->nloc; ++l) ws->tmp[l] = op->vel[e * n + ws->idx[l]];
      eval_v(g, ws->tmp, ws->wv[e], &ws->w);
    }
  } else if (op->vel_kind == ORC_VEL_VORTEX) {
    /* advection_diffusion_vortex.hpp PrescribedVelocityField: w_0 = s*x_{dim-1},
     * w_{dim-1} = -s*x_0 */
    for (int q = 0; q < g->nloc; ++q) {
      double x[3];
      q_point(g, cell, q, x);
      const double s = op->advec.vel_scale;
      double w0 = s * x[g->dim - 1];
      double wl = -s * x[0];
      ws->wv[0][q] = w0;
      ws->wv[g->dim - 1][q] = wl;
    }
  }
}

/* gather with constraints (read_dof_values: Dirichlet entries read as zero) */
static void gather(const orc_operator *op, const geo_t *g, const double *src, cell_ws_t *ws,
                   int use_mask)
{
  for (int l = 0; l < g->nloc; ++l) {
    const long gi = ws->idx[l];
    ws->u[l] = (use_mask && op->mask && op->mask[gi]) ? 0.0 : src[gi];
  }
}

static void scatter(const orc_operator *op, const geo_t *g, double *dst, cell_ws_t *ws,
                    int use_mask)
{
  for (int l = 0; l < g->nloc; ++l) {
    const long gi = ws->idx[l];
    if (use_mask && op->mask && op->mask[gi]) continue;
    dst[gi] += ws->out[l];
  }
}

/* heat Jacobian (reference tangent_local_cell_operation, heat_diffuse_operator.cpp:955-1086) */
static void heat_tangent_cell(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws)
{
  const orc_heat_params *hp = &op->heat;
  const int nq = g->nloc, dim = g->dim;
  const double kc = op->k_cell ? op->k_cell[cell] : 1.0;
  const int td = hp->t_dependent;
  const int hv = has_vel(op);
  eval_vg(g, ws->u, ws->val, ws->grad, &ws->w);
  const int needT = hp->law != ORC_LAW_CONSTANT || td;
  if (needT) {
    for (int l = 0; l < nq; ++l) ws->tmp[l] = op->T[ws->idx[l]];
    eval_vg(g, ws->tmp, ws->Tv, ws->Tg, &ws->w);
  }
  if (td) {
    for (int l = 0; l < nq; ++l) ws->tmp[l] = op->T_old[ws->idx[l]];
    eval_v(g, ws->tmp, ws->To, &ws->w);
  }
  if (hv) eval_velocity(op, g, cell, ws);
  const double dt_inv = 1.0 / hp->dt;
  for (int q = 0; q < nq; ++q) {
    double rc, k, drc, dk;
    material(hp, needT ? ws->Tv[q] : 0.0, &rc, &k, &drc, &dk);
    k *= kc;
    dk *= kc;
    const double u = ws->val[q];
    double val = dt_inv * rc * u;
    double gr[3];
    for (int e = 0; e < dim; ++e) gr[e] = k * ws->grad[e * nq + q];
    if (hv) {
      double s = 0.0;
      for (int e = 0; e < dim; ++e) s += ws->grad[e * nq + q] * ws->wv[e][q];
      val += rc * s;
    }
    if (td) {
      val += dt_inv * drc * (ws->Tv[q] - ws->To[q]) * u;
      for (int e = 0; e < dim; ++e) gr[e] += dk * ws->Tg[e * nq + q] * u;
    }
    if (hv && td) {
      double s = 0.0;
      for (int e = 0; e < dim; ++e) s += ws->Tg[e * nq + q] * ws->wv[e][q];
      val += drc * s * u;
    }
    const double jxw = cell_volume_weight(g, q);
    ws->fval[q] = jxw * val;
    for (int e = 0; e < dim; ++e) ws->flux[e * nq + q] = jxw * gr[e];
  }
  integrate(g, ws->fval, ws->flux, ws->out, &ws->w);
}

/* advection-diffusion tangent (tangent_local_cell_operation, advection_diffusion_operator.cpp:
 * 448-495): values m u/dt + theta w . grad u, gradients theta D grad u, with SUPG
 * + tau w (m u/dt + theta w . grad u - theta D lap u) (the value is submitted before the SUPG
 * residual takes the Laplacian) */
static void advec_tangent_cell(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws)
{
  const orc_advec_params *ap = &op->advec;
  const int nq = g->nloc, dim = g->dim;
  eval_vg(g, ws->u, ws->val, ws->grad, &ws->w);
  eval_velocity(op, g, cell, ws);
  const double m = ap->mass_factor / ap->dt;
  double tau = 0.0;
  if (ap->supg) {
    tau = supg_tau(op, g, ws->wv);
    eval_laplacian(g, ws->u, ws->lap, ws->tmp, ws->w.t1, ws->w.t2);
  }
  for (int q = 0; q < nq; ++q) {
    double vg = 0.0;
    for (int e = 0; e < dim; ++e) vg += ws->wv[e][q] * ws->grad[e * nq + q];
    const double jxw = cell_volume_weight(g, q);
    double res = m * ws->val[q] + ap->theta * vg;
    ws->fval[q] = jxw * res;
    if (ap->supg) res -= ap->theta * ap->diffusivity * ws->lap[q];
    for (int e = 0; e < dim; ++e) {
      double f = ap->theta * ap->diffusivity * ws->grad[e * nq + q];
      if (ap->supg) f += tau * ws->wv[e][q] * res;
      ws->flux[e * nq + q] = jxw * f;
    }
  }
  integrate(g, ws->fval, ws->flux, ws->out, &ws->w);
}

/* normal vector / curvature projection tangent: values m u, gradients m eta grad u
 * (tangent_local_cell_operation, normal_vector_operator.cpp:316-347,
 *  curvature_operator.cpp:339-371); ws->idx holds the cell's nodes */
static void ls_tangent_cell(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws)
{
  (void)cell;
  const int nq = g->nloc, dim = g->dim;
  const double eta = ls_damping(op, g);
  eval_vg(g, ws->u, ws->val, ws->grad, &ws->w);
  if (op->ls.narrow_band) {
    for (int l = 0; l < nq; ++l) ws->tmp[l] = op->level_set[ws->idx[l]];
    eval_v(g, ws->tmp, ws->Tv, &ws->w);
  }
  for (int q = 0; q < nq; ++q) {
    const double m = op->ls.narrow_band ? ls_mask(op, ws->Tv[q]) : 1.0;
    const double jxw = cell_volume_weight(g, q);
    ws->fval[q] = jxw * m * ws->val[q];
    for (int e = 0; e < dim; ++e) ws->flux[e * nq + q] = jxw * m * eta * ws->grad[e * nq + q];
  }
  integrate(g, ws->fval, ws->flux, ws->out, &ws->w);
}

/* Olsson reinitialization tangent: values u/dt, gradients eps_t grad u + (eps_n - eps_t)
 * (grad u . n) n with n the normalised normal vector at q (tangent_cell_operation,
 * olsson_CG_operator.cpp:438-478; the reference caches n per quadrature point in create_rhs,
 * from the same normal vector field) */
static void reinit_tangent_cell(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws)
{
  (void)cell;
  const int nq = g->nloc, dim = g->dim;
  const long n = op_ndofs(op);
  const double en = reinit_eps(op, g);
  const double et = op->reinit.tangential_diffusion_factor * en;
  eval_vg(g, ws->u, ws->val, ws->grad, &ws->w);
  for (int e = 0; e < dim; ++e) {
    for (int l = 0; l < nq; ++l) ws->tmp[l] = op->vel[e * n + ws->idx[l]];
    eval_v(g, ws->tmp, ws->wv[e], &ws->w);
  }
  for (int q = 0; q < nq; ++q) {
    double nv[3] = {0, 0, 0}, gu[3] = {0, 0, 0};
    for (int e = 0; e < dim; ++e) {
      nv[e] = ws->wv[e][q];
      gu[e] = ws->grad[e * nq + q];
    }
    normalize_tol(nv, dim, op->reinit.normal_tolerance);
    double gn = 0.0;
    for (int e = 0; e < dim; ++e) gn += gu[e] * nv[e];
    const double jxw = cell_volume_weight(g, q);
    ws->fval[q] = jxw * ws->val[q] / op->reinit.dt;
    for (int e = 0; e < dim; ++e)
      ws->flux[e * nq + q] = jxw * (et * gu[e] + (en - et) * gn * nv[e]);
  }
  integrate(g, ws->fval, ws->flux, ws->out, &ws->w);
}

static void tangent_cell(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws)
{
  if (op->kind == ORC_OP_HEAT)
    heat_tangent_cell(op, g, cell, ws);
  else if (op->kind == ORC_OP_ADVEC)
    advec_tangent_cell(op, g, cell, ws);
  else if (op->kind == ORC_OP_REINIT)
    reinit_tangent_cell(op, g, cell, ws);
  else
    ls_tangent_cell(op, g, cell, ws);
}

/* the normal vector operator applies the scalar projection to every block */
static int for_blocks(const orc_operator *op, orc_operator *scalar, long *n)
{
  if (op->kind != ORC_OP_NORMAL) return 0;
  *scalar = *op;
  scalar->kind = ORC_OP_CURVATURE;
  *n = op_ndofs(op);
  return op_dim(op);
}

static int nthreads_of(const orc_operator *op)
{
#ifdef _OPENMP
  return op->nthreads > 0 ? op->nthreads : omp_get_max_threads();
#else
  (void)op;
  return 1;
#endif
}

/* Cell loop coloured by the parity of the cell coordinates: cells of one colour share no node,
 * so the scatter needs no synchronisation. */
typedef void (*cell_fn)(const orc_operator *, const geo_t *, long, cell_ws_t *, void *);

static void cell_loop(const orc_operator *op, const geo_t *g, cell_fn fn, void *user)
{
  const int ncol = 1 << g->dim;
  const long ncells = orc_n_cells(&op->mesh);
  const int nt = nthreads_of(op);
  for (int col = 0; col < ncol; ++col) {
#pragma omp parallel num_threads(nt)
    {
      cell_ws_t *ws = (cell_ws_t *)malloc(sizeof(cell_ws_t));
#pragma omp for schedule(static)
      for (long cell = 0; cell < ncells; ++cell) {
        long c[3];
        cell_coords(g, cell, c);
        int cc = 0;
        for (int d = 0; d < g->dim; ++d) cc |= (int)(c[d] & 1) << d;
        if (cc != col) continue;
        fn(op, g, cell, ws, user);
      }
      free(ws);
    }
  }
}

typedef struct {
  const double *src;
  double *dst;
  const double *q_s;
} io_t;

static void vmult_cell(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws, void *u)
{
  io_t *io = (io_t *)u;
  cell_dofs(g, cell, ws->idx);
  gather(op, g, io->src, ws, op->use_mask);
  tangent_cell(op, g, cell, ws);
  scatter(op, g, io->dst, ws, op->use_mask);
}

static void avmult(const orc_operator *op, const double *src, double *dst);
static void adiagonal(const orc_operator *op, double *diag);

/* Boundary faces of the heat operator by textbook face quadrature (QGauss(p+1) in the
 * tangential directions, basis = products of 1D Lagrange values; on a face of the box the
 * normal-direction basis is 1 on the face nodes and 0 elsewhere).
 * mode 0: dst += B src, the Robin tangent alpha u + 4 eps sigma T^3 u
 *         (tangent_boundary_loop / tangent_local_boundary_operation,
 *          heat_diffuse_operator.cpp:243-269, 1088-1129);
 * mode 1: dst += diag(B) (the boundary operation of compute_diagonal, :297-391);
 * mode 2: dst += the boundary part of -R(T): -(q.n) on Neumann sides,
 *         -eps sigma (T^4 - T_inf^4) and -alpha (T - T_inf) (rhs_boundary_loop, :719-781).
 * T is read plain, src with the Dirichlet constraints; constrained rows are skipped. */
#define ORC_STEFAN_BOLTZMANN 5.670374419e-8 /* include/meltpooldg/utilities/physical_constants.hpp:6 */

static void heat_faces(const orc_operator *op, const geo_t *g, const double *src, double *dst,
                       int mode)
{
  const orc_heat_boundary *b = op->bnd;
  if (!b || op->kind != ORC_OP_HEAT) return;
  const int K = g->K, dim = g->dim, P = g->p;
  for (int side = 0; side < 2 * dim; ++side) {
    const unsigned bit = 1u << side;
    const int conv = (b->convection_sides & bit) != 0;
    const int rad = (b->radiation_sides & bit) != 0;
    const int neu = (b->neumann_sides & bit) != 0;
    if (!conv && !rad && !(mode == 2 && neu)) continue;
    const int nd = side / 2, upper = side % 2;
    int td[2], ntd = 0;
    for (int d = 0; d < dim; ++d)
      if (d != nd) td[ntd++] = d;
    long nf = 1;
    for (int t = 0; t < ntd; ++t) nf *= g->ncell[td[t]];
    int nfn = 1;
    for (int t = 0; t < ntd; ++t) nfn *= K;
    for (long f = 0; f < nf; ++f) {
      long c[3] = {0, 0, 0};
      long r = f;
      for (int t = 0; t < ntd; ++t) {
        c[td[t]] = r % g->ncell[td[t]];
        r /= g->ncell[td[t]];
      }
      c[nd] = upper ? g->ncell[nd] - 1 : 0;
      /* face nodes (tangential local indices a[t]) and their global numbers */
      long gi[KMAX * KMAX];
      double uu[KMAX * KMAX], TT[KMAX * KMAX], out[KMAX * KMAX];
      for (int l = 0; l < nfn; ++l) {
        int a[3] = {0, 0, 0};
        int rr = l;
        for (int t = 0; t < ntd; ++t) {
          a[td[t]] = rr % K;
          rr /= K;
        }
        a[nd] = upper ? P : 0;
        long idx = 0;
        for (int d = 0; d < dim; ++d) idx += (c[d] * P + a[d]) * g->stride[d];
        gi[l] = idx;
        const int cons = op->use_mask && op->mask && op->mask[idx];
        uu[l] = (mode == 0 && !cons) ? src[idx] : 0.0;
        TT[l] = op->T ? op->T[idx] : 0.0;
        out[l] = 0.0;
      }
      for (int q = 0; q < nfn; ++q) {
        int qa[2] = {0, 0};
        int rq = q;
        double wq = 1.0;
        for (int t = 0; t < ntd; ++t) {
          qa[t] = rq % K;
          rq /= K;
          wq *= g->wq[qa[t]] * g->h[td[t]];
        }
        double phi[KMAX * KMAX];
        double Tq = 0.0, uq = 0.0;
        for (int l = 0; l < nfn; ++l) {
          double ph = 1.0;
          int rr = l;
          for (int t = 0; t < ntd; ++t) {
            double v[KMAX], dv[KMAX];
            orc_lagrange(K, g->gl, g->xq[qa[t]], v, dv);
            ph *= v[rr % K];
            rr /= K;
          }
          phi[l] = ph;
          Tq += ph * TT[l];
          uq += ph * uu[l];
        }
        const double sb = ORC_STEFAN_BOLTZMANN;
        const double coef = (conv ? b->convection_coefficient : 0.0) +
                            (rad ? 4.0 * b->emissivity * sb * Tq * Tq * Tq : 0.0);
        double val = 0.0;
        if (mode == 2) {
          if (neu) val -= b->neumann_flux[side];
          if (rad) {
            const double Ti = b->radiation_temperature_infinity;
            val -= b->emissivity * sb * (Tq * Tq * Tq * Tq - Ti * Ti * Ti * Ti);
          }
          if (conv) val -= b->convection_coefficient * (Tq - b->convection_temperature_infinity);
        }
        for (int l = 0; l < nfn; ++l) {
          if (mode == 0) out[l] += wq * coef * uq * phi[l];
          else if (mode == 1) out[l] += wq * coef * phi[l] * phi[l];
          else out[l] += wq * val * phi[l];
        }
      }
      for (int l = 0; l < nfn; ++l) {
        if (op->use_mask && op->mask && op->mask[gi[l]]) continue;
        dst[gi[l]] += out[l];
      }
    }
  }
}

void orc_vmult(const orc_operator *op, const double *src, double *dst)
{
  orc_operator sc;
  long nb;
  if (op->io_mask && op->use_mask) {
    /* do_pre_vmult: inflow/outflow entries of src read as 0; do_post_vmult: dst = src there */
    const long n = op_ndofs(op);
    double *s0 = malloc(sizeof(double) * n);
    for (long i = 0; i < n; ++i) s0[i] = op->io_mask[i] ? 0.0 : src[i];
    orc_operator plain = *op;
    plain.io_mask = NULL;
    orc_vmult(&plain, s0, dst);
    for (long i = 0; i < n; ++i)
      if (op->io_mask[i]) dst[i] = src[i];
    free(s0);
    return;
  }
  if (for_blocks(op, &sc, &nb)) {
    for (int d = 0; d < op_dim(op); ++d) orc_vmult(&sc, src + d * nb, dst + d * nb);
    return;
  }
  if (op->amesh) {
    avmult(op, src, dst);
    return;
  }
  geo_t g;
  geo_init(&op->mesh, &g);
  const long n = orc_n_dofs(&op->mesh);
  memset(dst, 0, sizeof(double) * n);
  io_t io = {src, dst, NULL};
  cell_loop(op, &g, vmult_cell, &io);
  heat_faces(op, &g, src, dst, 0);
}

/* ------------------------------------------------------------------------------------------ */
/* Brute force: dense element matrices by textbook quadrature                                   */
/* ------------------------------------------------------------------------------------------ */

/* phi_l(q) and its physical gradient by direct products of 1D functions */
static void basis_at(const geo_t *g, int l, int q, double *phi, double *gphi)
{
  const int K = g->K;
  int a[3] = {l % K, (l / K) % K, l / (K * K)};
  int b[3] = {q % K, (q / K) % K, q / (K * K)};
  double v[3], d[3];
  for (int e = 0; e < 3; ++e) {
    if (e < g->dim) {
      v[e] = g->S[b[e] * K + a[e]];
      d[e] = g->DS[b[e] * K + a[e]] / g->h[e];
    } else {
      v[e] = 1.0;
      d[e] = 0.0;
    }
  }
  *phi = v[0] * v[1] * v[2];
  gphi[0] = d[0] * v[1] * v[2];
  gphi[1] = v[0] * d[1] * v[2];
  gphi[2] = v[0] * v[1] * d[2];
}

/* Laplacian of phi_l at q by direct products of the 1D second derivatives */
static double basis_lap(const geo_t *g, int l, int q)
{
  const int K = g->K;
  int a[3] = {l % K, (l / K) % K, l / (K * K)};
  int b[3] = {q % K, (q / K) % K, q / (K * K)};
  double lap = 0.0;
  for (int e = 0; e < g->dim; ++e) {
    double t = g->DDS[b[e] * K + a[e]] / (g->h[e] * g->h[e]);
    for (int d = 0; d < g->dim; ++d)
      if (d != e) t *= g->S[b[d] * K + a[d]];
    lap += t;
  }
  return lap;
}

/* coefficients of the bilinear form at q:
 *   a(u,v) = sum_q JxW [ c u v + (bu . grad u) v + kq grad u . grad v + u (bg . grad v) ] */
typedef struct {
  double c, k, bu[3], bg[3];
  double dn, nv[3]; /* + dn (nv . grad u)(nv . grad v) */
  double sw[3];     /* + (sw . grad u)(sw . grad v)   (SUPG: sw = sqrt(tau theta) w) */
  double lap[3];    /* + lap(u) (lap . grad v)        (SUPG: -tau theta D w) */
} qcoef_t;

static void q_coefficients(const orc_operator *op, const geo_t *g, long cell, const long *idx,
                           qcoef_t *qc)
{
  const int nq = g->nloc, dim = g->dim;
  const long n = orc_n_dofs(&op->mesh);
  /* velocity at every quadrature point first: the SUPG parameter is a maximum over the cell */
  double wq[3][NLOC_MAX];
  for (int q = 0; q < nq; ++q) {
    double phi[NLOC_MAX], gphi[NLOC_MAX][3];
    for (int l = 0; l < nq; ++l) basis_at(g, l, q, &phi[l], gphi[l]);
    for (int e = 0; e < 3; ++e) wq[e][q] = 0.0;
    if (op->vel_kind == ORC_VEL_NODAL && op->vel) {
      for (int e = 0; e < dim; ++e)
        for (int l = 0; l < nq; ++l) wq[e][q] += op->vel[e * n + idx[l]] * phi[l];
    } else if (op->vel_kind == ORC_VEL_VORTEX) {
      double x[3];
      q_point(g, cell, q, x);
      wq[0][q] = op->advec.vel_scale * x[dim - 1];
      wq[dim - 1][q] = -op->advec.vel_scale * x[0];
    }
  }
  const double tau =
      (op->kind == ORC_OP_ADVEC && op->advec.supg) ? supg_tau(op, g, (const double(*)[NLOC_MAX])wq) : 0.0;
  for (int q = 0; q < nq; ++q) {
    double phi[NLOC_MAX], gphi[NLOC_MAX][3];
    for (int l = 0; l < nq; ++l) basis_at(g, l, q, &phi[l], gphi[l]);
    const double w[3] = {wq[0][q], wq[1][q], wq[2][q]};
    qcoef_t *c = &qc[q];
    memset(c, 0, sizeof(*c));
    if (op->kind == ORC_OP_HEAT) {
      const orc_heat_params *hp = &op->heat;
      double T = 0, To = 0, Tg[3] = {0, 0, 0};
      const int needT = hp->law != ORC_LAW_CONSTANT || hp->t_dependent;
      if (needT)
        for (int l = 0; l < nq; ++l) {
          T += op->T[idx[l]] * phi[l];
          for (int e = 0; e < 3; ++e) Tg[e] += op->T[idx[l]] * gphi[l][e];
        }
      if (hp->t_dependent)
        for (int l = 0; l < nq; ++l) To += op->T_old[idx[l]] * phi[l];
      double rc, k, drc, dk;
      material(hp, T, &rc, &k, &drc, &dk);
      const double kc = op->k_cell ? op->k_cell[cell] : 1.0;
      k *= kc;
      dk *= kc;
      c->c = rc / hp->dt;
      c->k = k;
      if (op->vel_kind != ORC_VEL_NONE)
        for (int e = 0; e < 3; ++e) c->bu[e] = rc * w[e];
      if (hp->t_dependent) {
        c->c += drc / hp->dt * (T - To);
        for (int e = 0; e < 3; ++e) c->bg[e] = dk * Tg[e];
        if (op->vel_kind != ORC_VEL_NONE)
          c->c += drc * (Tg[0] * w[0] + Tg[1] * w[1] + Tg[2] * w[2]);
      }
    } else if (op->kind == ORC_OP_ADVEC) {
      const orc_advec_params *ap = &op->advec;
      c->c = ap->mass_factor / ap->dt;
      c->k = ap->theta * ap->diffusivity;
      for (int e = 0; e < 3; ++e) c->bu[e] = ap->theta * w[e];
      if (ap->supg) {
        /* tau (w . grad v)(m u/dt + theta w . grad u - theta D lap u) */
        const double st = sqrt(tau * ap->theta);
        for (int e = 0; e < 3; ++e) {
          c->bg[e] = tau * ap->mass_factor / ap->dt * w[e];
          c->sw[e] = st * w[e];
          c->lap[e] = -tau * ap->theta * ap->diffusivity * w[e];
        }
      }
    } else if (op->kind == ORC_OP_REINIT) {
      double nv[3] = {0, 0, 0};
      for (int e = 0; e < dim; ++e)
        for (int l = 0; l < nq; ++l) nv[e] += op->vel[e * n + idx[l]] * phi[l];
      normalize_tol(nv, dim, op->reinit.normal_tolerance);
      const double en = reinit_eps(op, g);
      const double et = op->reinit.tangential_diffusion_factor * en;
      c->c = 1.0 / op->reinit.dt;
      c->k = et;
      c->dn = en - et;
      for (int e = 0; e < 3; ++e) c->nv[e] = nv[e];
    } else {
      double ph = 0.0;
      if (op->ls.narrow_band)
        for (int l = 0; l < nq; ++l) ph += op->level_set[idx[l]] * phi[l];
      const double m = ls_mask(op, ph);
      c->c = m;
      c->k = m * ls_damping(op, g);
    }
    const double jxw = cell_volume_weight(g, q);
    c->c *= jxw;
    c->k *= jxw;
    c->dn *= jxw;
    for (int e = 0; e < 3; ++e) {
      c->bu[e] *= jxw;
      c->bg[e] *= jxw;
      c->sw[e] *= sqrt(jxw);
      c->lap[e] *= jxw;
    }
  }
}

void orc_local_matrix(const orc_operator *op, long cell, double *A)
{
  geo_t g;
  geo_init(&op->mesh, &g);
  const int nq = g.nloc;
  long idx[NLOC_MAX];
  cell_dofs(&g, cell, idx);
  qcoef_t *qc = (qcoef_t *)malloc(sizeof(qcoef_t) * nq);
  q_coefficients(op, &g, cell, idx, qc);
  memset(A, 0, sizeof(double) * nq * nq);
  for (int q = 0; q < nq; ++q) {
    double phi[NLOC_MAX], gphi[NLOC_MAX][3], lapj[NLOC_MAX];
    for (int l = 0; l < nq; ++l) {
      basis_at(&g, l, q, &phi[l], gphi[l]);
      lapj[l] = basis_lap(&g, l, q);
    }
    const qcoef_t *c = &qc[q];
    for (int i = 0; i < nq; ++i)
      for (int j = 0; j < nq; ++j) {
        const double gg = gphi[j][0] * gphi[i][0] + gphi[j][1] * gphi[i][1] + gphi[j][2] * gphi[i][2];
        const double bu = c->bu[0] * gphi[j][0] + c->bu[1] * gphi[j][1] + c->bu[2] * gphi[j][2];
        const double bg = c->bg[0] * gphi[i][0] + c->bg[1] * gphi[i][1] + c->bg[2] * gphi[i][2];
        const double ni = c->nv[0] * gphi[i][0] + c->nv[1] * gphi[i][1] + c->nv[2] * gphi[i][2];
        const double nj = c->nv[0] * gphi[j][0] + c->nv[1] * gphi[j][1] + c->nv[2] * gphi[j][2];
        const double si = c->sw[0] * gphi[i][0] + c->sw[1] * gphi[i][1] + c->sw[2] * gphi[i][2];
        const double sj = c->sw[0] * gphi[j][0] + c->sw[1] * gphi[j][1] + c->sw[2] * gphi[j][2];
        const double li = c->lap[0] * gphi[i][0] + c->lap[1] * gphi[i][1] + c->lap[2] * gphi[i][2];
        A[i * nq + j] += c->c * phi[j] * phi[i] + bu * phi[i] + c->k * gg + phi[j] * bg +
                         c->dn * ni * nj + si * sj + lapj[j] * li;
      }
  }
  free(qc);
}

static void dense_cell(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws, void *u)
{
  io_t *io = (io_t *)u;
  const int nq = g->nloc;
  cell_dofs(g, cell, ws->idx);
  gather(op, g, io->src, ws, op->use_mask);
  double *A = (double *)malloc(sizeof(double) * nq * nq);
  orc_local_matrix(op, cell, A);
  for (int i = 0; i < nq; ++i) {
    double s = 0.0;
    for (int j = 0; j < nq; ++j) s += A[i * nq + j] * ws->u[j];
    ws->out[i] = s;
  }
  free(A);
  scatter(op, g, io->dst, ws, op->use_mask);
}

void orc_vmult_dense(const orc_operator *op, const double *src, double *dst)
{
  orc_operator sc;
  long nb;
  if (for_blocks(op, &sc, &nb)) {
    for (int d = 0; d < op_dim(op); ++d) orc_vmult_dense(&sc, src + d * nb, dst + d * nb);
    return;
  }
  geo_t g;
  geo_init(&op->mesh, &g);
  const long n = orc_n_dofs(&op->mesh);
  memset(dst, 0, sizeof(double) * n);
  io_t io = {src, dst, NULL};
  cell_loop(op, &g, dense_cell, &io);
  heat_faces(op, &g, src, dst, 0);
}

/* diagonal of the constrained operator: sum over cells of the local diagonal entries of the
 * bilinear form; constrained rows stay 0 (MatrixFreeTools::compute_diagonal semantics) */
static void diag_cell(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws, void *u)
{
  io_t *io = (io_t *)u;
  const int nq = g->nloc;
  cell_dofs(g, cell, ws->idx);
  qcoef_t *qc = (qcoef_t *)malloc(sizeof(qcoef_t) * nq);
  q_coefficients(op, g, cell, ws->idx, qc);
  for (int i = 0; i < nq; ++i) ws->out[i] = 0.0;
  for (int q = 0; q < nq; ++q) {
    const qcoef_t *c = &qc[q];
    for (int i = 0; i < nq; ++i) {
      double phi, gp[3];
      basis_at(g, i, q, &phi, gp);
      const double gg = gp[0] * gp[0] + gp[1] * gp[1] + gp[2] * gp[2];
      const double bu = c->bu[0] * gp[0] + c->bu[1] * gp[1] + c->bu[2] * gp[2];
      const double bg = c->bg[0] * gp[0] + c->bg[1] * gp[1] + c->bg[2] * gp[2];
      const double ni = c->nv[0] * gp[0] + c->nv[1] * gp[1] + c->nv[2] * gp[2];
      const double si = c->sw[0] * gp[0] + c->sw[1] * gp[1] + c->sw[2] * gp[2];
      const double li = c->lap[0] * gp[0] + c->lap[1] * gp[1] + c->lap[2] * gp[2];
      ws->out[i] += c->c * phi * phi + bu * phi + c->k * gg + phi * bg + c->dn * ni * ni +
                    si * si + basis_lap(g, i, q) * li;
    }
  }
  free(qc);
  scatter(op, g, io->dst, ws, op->use_mask);
}

void orc_diagonal(const orc_operator *op, double *diag)
{
  orc_operator sc;
  long nb;
  if (for_blocks(op, &sc, &nb)) {
    for (int d = 0; d < op_dim(op); ++d) orc_diagonal(&sc, diag + d * nb);
    return;
  }
  if (op->amesh) {
    adiagonal(op, diag);
    return;
  }
  geo_t g;
  geo_init(&op->mesh, &g);
  const long n = orc_n_dofs(&op->mesh);
  memset(diag, 0, sizeof(double) * n);
  io_t io = {NULL, diag, NULL};
  cell_loop(op, &g, diag_cell, &io);
  heat_faces(op, &g, NULL, diag, 1);
}

/* heat: |d| > 1e-10 ? 1/d : 1  (heat_diffuse_operator.cpp:282-283)
 * advection, normal vector, curvature: |d| > 1e-14 max(1, |d|_inf) ? 1/d : 1
 * (advection_diffusion_operator.cpp:434-436, normal_vector_operator.cpp:306-310,
 *  curvature_operator.cpp:331-334) */
void orc_invert_diagonal(int kind, long n, double *diag)
{
  double thr = 1.0e-10;
  if (kind != ORC_OP_HEAT) {
    double mx = 1.0;
    for (long i = 0; i < n; ++i) mx = fmax(mx, fabs(diag[i]));
    thr = 1.0e-14 * mx;
  }
  for (long i = 0; i < n; ++i) diag[i] = fabs(diag[i]) > thr ? 1.0 / diag[i] : 1.0;
}

/* ------------------------------------------------------------------------------------------ */
/* Right-hand sides                                                                            */
/* ------------------------------------------------------------------------------------------ */

/* heat: -R(T) with values (rho cp/dt)(T - T_old) - q_s (+ rho cp grad T . w) and gradients k grad T
 * (rhs_cell_loop, heat_diffuse_operator.cpp:393-529).  src = T_old.
 * advection: values phi/dt - (1-theta) w.grad phi, gradients -(1-theta) D grad phi
 * (create_rhs, advection_diffusion_operator.cpp:274-358).  src = phi_old. */
/* local residual of a cell whose DoF indices are in ws->idx: fields read as stored (the
 * callers pass constraint-consistent vectors), result in ws->out */
static void rhs_local(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws, io_t *io)
{
  const int nq = g->nloc, dim = g->dim;
  if (op->kind == ORC_OP_HEAT) {
    const orc_heat_params *hp = &op->heat;
    for (int l = 0; l < nq; ++l) ws->tmp[l] = op->T[ws->idx[l]];
    eval_vg(g, ws->tmp, ws->Tv, ws->Tg, &ws->w);
    for (int l = 0; l < nq; ++l) ws->tmp[l] = io->src[ws->idx[l]];
    eval_v(g, ws->tmp, ws->To, &ws->w);
    for (int l = 0; l < nq; ++l) ws->tmp[l] = io->q_s ? io->q_s[ws->idx[l]] : 0.0;
    eval_v(g, ws->tmp, ws->val, &ws->w);
    if (has_vel(op)) eval_velocity(op, g, cell, ws);
    const double kc = op->k_cell ? op->k_cell[cell] : 1.0;
    for (int q = 0; q < nq; ++q) {
      double rc, k, drc, dk;
      material(hp, ws->Tv[q], &rc, &k, &drc, &dk);
      k *= kc;
      double val = rc / hp->dt * (ws->Tv[q] - ws->To[q]) - ws->val[q];
      if (has_vel(op)) {
        double s = 0.0;
        for (int e = 0; e < dim; ++e) s += ws->Tg[e * nq + q] * ws->wv[e][q];
        val += rc * s;
      }
      const double jxw = cell_volume_weight(g, q);
      ws->fval[q] = -jxw * val;
      for (int e = 0; e < dim; ++e) ws->flux[e * nq + q] = -jxw * k * ws->Tg[e * nq + q];
    }
  } else {
    const orc_advec_params *ap = &op->advec;
    for (int l = 0; l < nq; ++l) ws->tmp[l] = io->src[ws->idx[l]];
    eval_vg(g, ws->tmp, ws->val, ws->grad, &ws->w);
    /* the Laplacian before eval_velocity reuses ws->tmp */
    if (ap->supg) eval_laplacian(g, ws->tmp, ws->lap, ws->To, ws->w.t1, ws->w.t2);
    eval_velocity(op, g, cell, ws);
    const double tau = ap->supg ? supg_tau(op, g, ws->wv) : 0.0;
    for (int q = 0; q < nq; ++q) {
      double vg = 0.0;
      for (int e = 0; e < dim; ++e) vg += ws->wv[e][q] * ws->grad[e * nq + q];
      const double jxw = cell_volume_weight(g, q);
      double res = ws->val[q] / ap->dt - (1.0 - ap->theta) * vg;
      ws->fval[q] = jxw * res;
      /* SUPG (create_rhs, advection_diffusion_operator.cpp:338-343) */
      if (ap->supg) res += (1.0 - ap->theta) * ap->diffusivity * ws->lap[q];
      for (int e = 0; e < dim; ++e) {
        double f = -(1.0 - ap->theta) * ap->diffusivity * ws->grad[e * nq + q];
        if (ap->supg) f += tau * ws->wv[e][q] * res;
        ws->flux[e * nq + q] = jxw * f;
      }
    }
  }
  integrate(g, ws->fval, ws->flux, ws->out, &ws->w);
}

static void rhs_cell(const orc_operator *op, const geo_t *g, long cell, cell_ws_t *ws, void *u)
{
  io_t *io = (io_t *)u;
  cell_dofs(g, cell, ws->idx);
  rhs_local(op, g, cell, ws, io);
  scatter(op, g, io->dst, ws, op->use_mask);
}

static void arhs(const orc_operator *op, const double *src, const double *q_s, double *dst);

void orc_create_rhs(const orc_operator *op, const double *src, const double *q_s, double *dst)
{
  if (op->amesh) {
    arhs(op, src, q_s, dst);
    return;
  }
  geo_t g;
  geo_init(&op->mesh, &g);
  io_t io = {src, dst, q_s};
  cell_loop(op, &g, rhs_cell, &io);
  heat_faces(op, &g, NULL, dst, 2);
}

/* ------------------------------------------------------------------------------------------ */
/* Krylov solvers (LinearSolver::solve, linear_solver.hpp:22-134; ReductionControl semantics:   */
/* success when |r| <= max(abs_tol, rel_tol*|r_0|), failure when the step count reaches max_it) */
/* ------------------------------------------------------------------------------------------ */

static double dot(long n, const double *a, const double *b)
{
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

typedef struct {
  double tol;
  int max_it;
} control_t;

/* 0 iterate, 1 success, 2 failure */
static int check(control_t *c, int step, double res, double abs_tol, double rel_tol)
{
  if (step == 0) c->tol = fmax(abs_tol, rel_tol * res);
  if (res <= c->tol) return 1;
  if (step >= c->max_it || isnan(res)) return 2;
  return 0;
}

static void precond(long n, const double *inv_diag, const double *r, double *z)
{
  if (inv_diag) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) z[i] = inv_diag[i] * r[i];
  } else {
    memcpy(z, r, sizeof(double) * n);
  }
}

static int solve_cg(const orc_operator *op, long n, int max_it, double abs_tol, double rel_tol,
                    const double *inv_diag, const double *b, double *x, int *iters, double *fres)
{
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *p = malloc(sizeof(double) * n), *q = malloc(sizeof(double) * n);
  control_t c = {0.0, max_it};
  orc_vmult(op, x, q);
  for (long i = 0; i < n; ++i) r[i] = b[i] - q[i];
  double res = sqrt(dot(n, r, r));
  int state = check(&c, 0, res, abs_tol, rel_tol);
  int it = 0;
  double rz_old = 0.0;
  while (state == 0) {
    ++it;
    precond(n, inv_diag, r, z);
    const double rz = dot(n, r, z);
    if (it == 1)
      memcpy(p, z, sizeof(double) * n);
    else {
      const double beta = rz / rz_old;
      for (long i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    rz_old = rz;
    orc_vmult(op, p, q);
    const double alpha = rz / dot(n, p, q);
    for (long i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    res = sqrt(dot(n, r, r));
    state = check(&c, it, res, abs_tol, rel_tol);
  }
  *iters = it;
  *fres = res;
  free(r);
  free(z);
  free(p);
  free(q);
  return state == 1 ? ORC_SUCCESS : ORC_NO_CONVERGENCE;
}

/* GMRES(m) with right preconditioning, classical Gram-Schmidt with one re-orthogonalisation
 * pass, Givens rotations and the estimated residual |gamma_{j+1}| as convergence measure. */
static int solve_gmres(const orc_operator *op, long n, int m, int max_it, double abs_tol,
                       double rel_tol, const double *inv_diag, const double *b, double *x,
                       int *iters, double *fres)
{
  double *V = malloc(sizeof(double) * n * (m + 1));
  double *w = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
  double *H = calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = malloc(sizeof(double) * m), *sn = malloc(sizeof(double) * m);
  double *gam = malloc(sizeof(double) * (m + 1)), *y = malloc(sizeof(double) * m);
  double *hcol = malloc(sizeof(double) * (m + 1));
  control_t c = {0.0, max_it};
  int acc = 0, state = 0;
  double res = 0.0;
  for (;;) {
    orc_vmult(op, x, w);
    for (long i = 0; i < n; ++i) V[i] = b[i] - w[i];
    double rho = sqrt(dot(n, V, V));
    res = rho;
    state = check(&c, acc, rho, abs_tol, rel_tol);
    if (state != 0) break;
    for (long i = 0; i < n; ++i) V[i] /= rho;
    gam[0] = rho;
    int j = 0;
    for (; j < m && state == 0; ++j) {
      ++acc;
      double *vj = V + (long)j * n, *vn = V + (long)(j + 1) * n;
      precond(n, inv_diag, vj, z);
      orc_vmult(op, z, vn);
      for (int i = 0; i <= j + 1; ++i) hcol[i] = 0.0;
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i <= j; ++i) {
          const double hij = dot(n, V + (long)i * n, vn);
          hcol[i] += hij;
          for (long l = 0; l < n; ++l) vn[l] -= hij * V[(long)i * n + l];
        }
      const double s = sqrt(dot(n, vn, vn));
      hcol[j + 1] = s;
      if (s != 0.0)
        for (long l = 0; l < n; ++l) vn[l] /= s;
      for (int i = 0; i < j; ++i) {
        const double t = cs[i] * hcol[i] + sn[i] * hcol[i + 1];
        hcol[i + 1] = -sn[i] * hcol[i] + cs[i] * hcol[i + 1];
        hcol[i] = t;
      }
      const double r = hypot(hcol[j], hcol[j + 1]);
      cs[j] = hcol[j] / r;
      sn[j] = hcol[j + 1] / r;
      hcol[j] = r;
      hcol[j + 1] = 0.0;
      gam[j + 1] = -sn[j] * gam[j];
      gam[j] = cs[j] * gam[j];
      for (int i = 0; i <= j; ++i) H[i * m + j] = hcol[i];
      rho = fabs(gam[j + 1]);
      res = rho;
      state = check(&c, acc, rho, abs_tol, rel_tol);
    }
    /* back substitution and update x += P (V y) */
    for (int i = j - 1; i >= 0; --i) {
      double s = gam[i];
      for (int l = i + 1; l < j; ++l) s -= H[i * m + l] * y[l];
      y[i] = s / H[i * m + i];
    }
    memset(w, 0, sizeof(double) * n);
    for (int i = 0; i < j; ++i)
      for (long l = 0; l < n; ++l) w[l] += y[i] * V[(long)i * n + l];
    precond(n, inv_diag, w, z);
    for (long l = 0; l < n; ++l) x[l] += z[l];
    if (state != 0) break;
  }
  *iters = acc;
  *fres = res;
  free(V);
  free(w);
  free(z);
  free(H);
  free(cs);
  free(sn);
  free(gam);
  free(y);
  free(hcol);
  return state == 1 ? ORC_SUCCESS : ORC_NO_CONVERGENCE;
}

int orc_solve(const orc_operator *op, int solver, int restart, int max_it, double abs_tol,
              double rel_tol, const double *inv_diag, const double *b, double *x, int *iterations,
              double *final_residual)
{
  const long n = orc_n_rows(op);
  if (solver == ORC_CG)
    return solve_cg(op, n, max_it, abs_tol, rel_tol, inv_diag, b, x, iterations, final_residual);
  return solve_gmres(op, n, restart > 0 ? restart : 30, max_it, abs_tol, rel_tol, inv_diag, b, x,
                     iterations, final_residual);
}

/* ------------------------------------------------------------------------------------------ */
/* FE L2 norm: VectorTools::integrate_difference into a Vector<float> of cell errors, then       */
/* compute_global_error (source/utilities/vector_tools.cpp:137-201)                              */
/* ------------------------------------------------------------------------------------------ */

double orc_fe_l2_norm(const orc_mesh *m, const double *u)
{
  geo_t g;
  geo_init(m, &g);
  const long ncells = orc_n_cells(m);
  double total = 0.0;
  cell_ws_t *ws = (cell_ws_t *)malloc(sizeof(cell_ws_t));
  for (long cell = 0; cell < ncells; ++cell) {
    cell_dofs(&g, cell, ws->idx);
    for (int l = 0; l < g.nloc; ++l) ws->tmp[l] = u[ws->idx[l]];
    eval_v(&g, ws->tmp, ws->val, &ws->w);
    double s = 0.0;
    for (int q = 0; q < g.nloc; ++q) s += ws->val[q] * ws->val[q] * cell_volume_weight(&g, q);
    const float e = (float)sqrt(s);
    total += (double)e * (double)e;
  }
  free(ws);
  return sqrt(total);
}

/* ------------------------------------------------------------------------------------------ */
/* Whole-path drivers for the golden outputs                                                    */
/* ------------------------------------------------------------------------------------------ */

/* Implicit Euler heat steps with the Newton loop of NewtonRaphsonSolver::solve
 * (source/linear_algebra/newton_raphson_solver.cpp:22-80): r = -R(T); solve J d = r;
 * converged when |r| < 5e-9 and |d| < 1e-10 (checked before d is added); otherwise T += d.
 * The predictor is the old solution; the converged state does not depend on it. */
int orc_heat_implicit_euler(const orc_operator *op_in, int n_steps, double *T, double *norms)
{
  orc_operator op = *op_in;
  const long n = orc_n_dofs(&op.mesh);
  double *T_old = malloc(sizeof(double) * n), *r = malloc(sizeof(double) * n);
  double *d = malloc(sizeof(double) * n), *inv = malloc(sizeof(double) * n);
  int status = 0;
  for (int step = 0; step < n_steps; ++step) {
    memcpy(T_old, T, sizeof(double) * n);
    op.T = T;
    op.T_old = T_old;
    op.use_mask = 1;
    int converged = 0;
    for (int it = 0; it < 10 && !converged; ++it) {
      memset(r, 0, sizeof(double) * n);
      orc_create_rhs(&op, T_old, NULL, r);
      orc_diagonal(&op, inv);
      orc_invert_diagonal(ORC_OP_HEAT, n, inv);
      memset(d, 0, sizeof(double) * n);
      int iters;
      double fres;
      status |= orc_solve(&op, op.heat.t_dependent ? ORC_GMRES : ORC_CG, 30, 10000, 1e-20,
                          1e-12, inv, r, d, &iters, &fres);
      const double rn = sqrt(dot(n, r, r)), dn = sqrt(dot(n, d, d));
      if (rn < 5e-9 && dn < 1e-10)
        converged = 1;
      else
        for (long i = 0; i < n; ++i) T[i] += d[i];
    }
    if (!converged) status |= 2;
    norms[step] = orc_fe_l2_norm(&op.mesh, T);
  }
  free(T_old);
  free(r);
  free(d);
  free(inv);
  return status;
}

/* theta-scheme advection steps as AdvectionDiffusionOperation::solve
 * (source/level_set/advection_diffusion_operation.cpp:227-367) with the Dirichlet lifting of
 * create_rhs_and_apply_dirichlet_matrixfree (utilities_matrixfree.hpp:33-100):
 *   rhs = -A_noBC x_d + b(phi_old), constrained rows set to zero, GMRES + Jacobi,
 *   constrained entries set to their Dirichlet values afterwards. */
int orc_advec_theta_steps(const orc_operator *op_in, int n_steps, const double *dirichlet_values,
                          double *phi, double *norms)
{
  orc_operator op = *op_in;
  const long n = op_ndofs(&op);
  double *old = malloc(sizeof(double) * n), *rhs = malloc(sizeof(double) * n);
  double *xd = malloc(sizeof(double) * n), *inv = malloc(sizeof(double) * n);
  int status = 0;
  /* constrained entries: Dirichlet rows and the inflow/outflow DoFs (create_inflow_outflow_
   * constraints, advection_diffusion_operation.cpp:368-499, lifted through
   * create_rhs_and_apply_dirichlet_matrixfree's additional inhomogeneous constraints) */
#define ORC_LIFTED(i) ((op.mask && op.mask[i]) || (op.io_mask && op.io_mask[i]))
  for (long i = 0; i < n; ++i) xd[i] = ORC_LIFTED(i) ? dirichlet_values[i] : 0.0;
  /* adaptive mesh: fields kept constraint-consistent (hanging values from their masters) */
  if (op.amesh) orc_distribute_constraints(op.adata, phi);
  for (int step = 0; step < n_steps; ++step) {
    memcpy(old, phi, sizeof(double) * n);
    op.use_mask = 0;
    orc_vmult(&op, xd, rhs);
    for (long i = 0; i < n; ++i) rhs[i] = -rhs[i];
    orc_create_rhs(&op, old, NULL, rhs);
    for (long i = 0; i < n; ++i)
      if (ORC_LIFTED(i)) rhs[i] = 0.0;
    op.use_mask = 1;
    orc_diagonal(&op, inv);
    orc_invert_diagonal(ORC_OP_ADVEC, n, inv);
    /* compute_inverse_diagonal_from_matrixfree: inflow/outflow rows 1 (:437-444) */
    for (long i = 0; i < n; ++i)
      if (op.io_mask && op.io_mask[i]) inv[i] = 1.0;
    int iters;
    double fres;
    status |= orc_solve(&op, ORC_GMRES, 30, 10000, 1e-20, 1e-12, inv, rhs, phi, &iters, &fres);
    for (long i = 0; i < n; ++i)
      if (ORC_LIFTED(i)) phi[i] = dirichlet_values[i];
#undef ORC_LIFTED
    if (op.amesh) orc_distribute_constraints(op.adata, phi);
    norms[step] = op.amesh ? 0.0 : orc_fe_l2_norm(&op.mesh, phi);
  }
  free(old);
  free(rhs);
  free(xd);
  free(inv);
  return status;
}

/* ------------------------------------------------------------------------------------------ */
/* Adaptive meshes: node identification and hanging-node constraints                          */
/* (the constraints deal.II's DoFTools::make_hanging_node_constraints produces for FE_Q on a     */
/* 2:1-balanced mesh, as used by Constraints::make_HNC, constraints.hpp:117-136).                */
/*                                                                                              */
/* Nodes are identified by their coordinates, quantised to 1/1024 of the finest cell: distinct   */
/* Gauss-Lobatto nodes are at least 0.17 of the finest cell apart.  A node is constrained when  */
/* it is not a node of the coarsest cell whose closure contains it; its line is the             */
/* interpolation of that cell's Lagrange basis at the node, with chains resolved.               */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
  long k[3];
  long id;
} hslot_t;

static uint64_t key_hash(const long *k)
{
  uint64_t h = orc_splitmix64((uint64_t)k[0]);
  h = orc_splitmix64(h ^ (uint64_t)k[1]);
  return orc_splitmix64(h ^ (uint64_t)k[2]);
}

static int amesh_levels(const orc_amesh *am)
{
  int L = 0;
  for (long c = 0; c < am->n_cells; ++c)
    if (am->cells[4 * c] > L) L = am->cells[4 * c];
  return L;
}

/* cell geometry as a one-cell structured mesh */
static void acell_geo(const orc_amesh *am, long cell, geo_t *g)
{
  const int32_t *cl = am->cells + 4 * cell;
  orc_mesh m;
  m.dim = 3;
  m.degree = am->degree;
  for (int d = 0; d < 3; ++d) {
    const double h0 = (am->hi[d] - am->lo[d]) / am->ncell[d];
    const double h = h0 / (double)(1L << cl[0]);
    m.ncell[d] = 1;
    m.lo[d] = am->lo[d] + h * cl[1 + d];
    m.hi[d] = m.lo[d] + h;
  }
  geo_init(&m, g);
}

int orc_amesh_build(const orc_amesh *am, orc_adata *ad)
{
  memset(ad, 0, sizeof(*ad));
  const int K = am->degree + 1, nloc = K * K * K;
  const int L = amesh_levels(am);
  const long nc = am->n_cells;
  double gl[KMAX];
  orc_gauss_lobatto(K, gl);
  /* positions in finest-cell units: u = (ix + gl[a]) * 2^(L - level) */
  long cap = 1;
  while (cap < 2 * nc * nloc + 16) cap <<= 1;
  hslot_t *tab = (hslot_t *)malloc(sizeof(hslot_t) * cap);
  for (long i = 0; i < cap; ++i) tab[i].id = -1;
  ad->cell_nodes = (long *)malloc(sizeof(long) * nc * nloc);
  double *upos = (double *)malloc(sizeof(double) * 3 * nc * nloc);
  long nn = 0;
  for (long c = 0; c < nc; ++c) {
    const int32_t *cl = am->cells + 4 * c;
    const double sc = (double)(1L << (L - cl[0]));
    for (int l = 0; l < nloc; ++l) {
      const int a[3] = {l % K, (l / K) % K, l / (K * K)};
      double u[3];
      long k[3];
      for (int d = 0; d < 3; ++d) {
        u[d] = ((double)cl[1 + d] + gl[a[d]]) * sc;
        k[d] = lround(u[d] * 1024.0);
      }
      uint64_t h = key_hash(k) & (uint64_t)(cap - 1);
      for (;;) {
        if (tab[h].id < 0) {
          tab[h].k[0] = k[0];
          tab[h].k[1] = k[1];
          tab[h].k[2] = k[2];
          tab[h].id = nn;
          for (int d = 0; d < 3; ++d) upos[3 * nn + d] = u[d];
          ++nn;
          break;
        }
        if (tab[h].k[0] == k[0] && tab[h].k[1] == k[1] && tab[h].k[2] == k[2]) break;
        h = (h + 1) & (uint64_t)(cap - 1);
      }
      ad->cell_nodes[c * nloc + l] = tab[h].id;
    }
  }
  free(tab);
  ad->n_nodes = nn;
  ad->xyz = (double *)malloc(sizeof(double) * 3 * nn);
  ad->boundary = (uint8_t *)calloc(nn, 1);
  for (long i = 0; i < nn; ++i)
    for (int d = 0; d < 3; ++d) {
      const double ext = (double)am->ncell[d] * (double)(1L << L);
      const double h0 = (am->hi[d] - am->lo[d]) / am->ncell[d];
      ad->xyz[3 * i + d] = am->lo[d] + h0 * upos[3 * i + d] / (double)(1L << L);
      if (fabs(upos[3 * i + d]) < 1e-9 || fabs(upos[3 * i + d] - ext) < 1e-9) ad->boundary[i] = 1;
    }
  /* leaves per base cell for point location */
  const long nb = (long)am->ncell[0] * am->ncell[1] * am->ncell[2];
  if (nb <= 0 || nb > (1L << 40)) return 1;
  long *bptr = (long *)calloc(nb + 1, sizeof(long)), *blist = (long *)malloc(sizeof(long) * nc);
  for (long c = 0; c < nc; ++c) {
    const int32_t *cl = am->cells + 4 * c;
    const long b = (cl[1] >> cl[0]) + am->ncell[0] * ((cl[2] >> cl[0]) + (long)am->ncell[1] * (cl[3] >> cl[0]));
    bptr[b + 1]++;
  }
  for (long b = 0; b < nb; ++b) bptr[b + 1] += bptr[b];
  long *fill = (long *)calloc(nb, sizeof(long));
  for (long c = 0; c < nc; ++c) {
    const int32_t *cl = am->cells + 4 * c;
    const long b = (cl[1] >> cl[0]) + am->ncell[0] * ((cl[2] >> cl[0]) + (long)am->ncell[1] * (cl[3] >> cl[0]));
    blist[bptr[b] + fill[b]++] = c;
  }
  free(fill);
  /* direct constraints */
  const double tol = 1e-9;
  long *dptr = (long *)calloc(nn + 1, sizeof(long));
  long dcap = 1024, dn = 0;
  long *dcol = (long *)malloc(sizeof(long) * dcap);
  double *dval = (double *)malloc(sizeof(double) * dcap);
  const long B = 1L << L;
  for (long i = 0; i < nn; ++i) {
    const double *u = upos + 3 * i;
    long best = -1;
    int blev = 1 << 30;
    long lo_b[3], hi_b[3];
    for (int d = 0; d < 3; ++d) {
      lo_b[d] = (long)floor((u[d] - tol) / (double)B);
      hi_b[d] = (long)floor((u[d] + tol) / (double)B);
      if (lo_b[d] < 0) lo_b[d] = 0;
      if (hi_b[d] > am->ncell[d] - 1) hi_b[d] = am->ncell[d] - 1;
    }
    for (long bz = lo_b[2]; bz <= hi_b[2]; ++bz)
      for (long by = lo_b[1]; by <= hi_b[1]; ++by)
        for (long bx = lo_b[0]; bx <= hi_b[0]; ++bx) {
          const long b = bx + am->ncell[0] * (by + (long)am->ncell[1] * bz);
          for (long t = bptr[b]; t < bptr[b + 1]; ++t) {
            const long c = blist[t];
            const int32_t *cl = am->cells + 4 * c;
            const double sc = (double)(1L << (L - cl[0]));
            int in = 1;
            for (int d = 0; d < 3 && in; ++d)
              in = u[d] >= cl[1 + d] * sc - tol && u[d] <= (cl[1 + d] + 1) * sc + tol;
            if (in && cl[0] < blev) {
              blev = cl[0];
              best = c;
            }
          }
        }
    dptr[i + 1] = dptr[i];
    if (best < 0) continue;
    const int32_t *cl = am->cells + 4 * best;
    const double sc = (double)(1L << (L - cl[0]));
    double wv[3][KMAX];
    int node_of_cell = 1;
    for (int d = 0; d < 3; ++d) {
      const double t = u[d] / sc - cl[1 + d];
      int match = -1;
      for (int j = 0; j < K; ++j)
        if (fabs(t - gl[j]) < 1e-12) match = j;
      if (match >= 0) {
        for (int j = 0; j < K; ++j) wv[d][j] = (j == match) ? 1.0 : 0.0;
      } else {
        double der[KMAX];
        orc_lagrange(K, gl, t, wv[d], der);
        node_of_cell = 0;
      }
    }
    if (node_of_cell) continue;
    for (int l = 0; l < nloc; ++l) {
      const double w = wv[0][l % K] * wv[1][(l / K) % K] * wv[2][l / (K * K)];
      if (w == 0.0) continue;
      if (dn == dcap) {
        dcap *= 2;
        dcol = (long *)realloc(dcol, sizeof(long) * dcap);
        dval = (double *)realloc(dval, sizeof(double) * dcap);
      }
      dcol[dn] = ad->cell_nodes[best * nloc + l];
      dval[dn] = w;
      ++dn;
    }
    dptr[i + 1] = dn;
  }
  free(bptr);
  free(blist);
  free(upos);
  /* resolve chains: expand constrained masters until every master is free */
  long *ptr = dptr, *col = dcol;
  double *val = dval;
  for (int pass = 0; pass < 16; ++pass) {
    int changed = 0;
    long ncap = ptr[nn] + 16, nnz = 0;
    long *nptr = (long *)calloc(nn + 1, sizeof(long));
    long *ncol = (long *)malloc(sizeof(long) * ncap);
    double *nval = (double *)malloc(sizeof(double) * ncap);
    for (long i = 0; i < nn; ++i) {
      const long start = nnz;
      for (long t = ptr[i]; t < ptr[i + 1]; ++t) {
        const long m = col[t];
        if (ptr[m + 1] > ptr[m]) {
          changed = 1;
          for (long s2 = ptr[m]; s2 < ptr[m + 1]; ++s2) {
            if (nnz == ncap) {
              ncap *= 2;
              ncol = (long *)realloc(ncol, sizeof(long) * ncap);
              nval = (double *)realloc(nval, sizeof(double) * ncap);
            }
            ncol[nnz] = col[s2];
            nval[nnz] = val[t] * val[s2];
            ++nnz;
          }
        } else {
          if (nnz == ncap) {
            ncap *= 2;
            ncol = (long *)realloc(ncol, sizeof(long) * ncap);
            nval = (double *)realloc(nval, sizeof(double) * ncap);
          }
          ncol[nnz] = m;
          nval[nnz] = val[t];
          ++nnz;
        }
      }
      /* merge duplicate masters */
      long w = start;
      for (long a = start; a < nnz; ++a) {
        long b = start;
        for (; b < w; ++b)
          if (ncol[b] == ncol[a]) break;
        if (b < w) {
          nval[b] += nval[a];
        } else {
          ncol[w] = ncol[a];
          nval[w] = nval[a];
          ++w;
        }
      }
      nnz = w;
      nptr[i + 1] = nnz;
    }
    free(ptr);
    free(col);
    free(val);
    ptr = nptr;
    col = ncol;
    val = nval;
    if (!changed) break;
  }
  ad->line_ptr = ptr;
  ad->line_col = col;
  ad->line_val = val;
  return 0;
}

void orc_adata_free(orc_adata *ad)
{
  free(ad->cell_nodes);
  free(ad->xyz);
  free(ad->boundary);
  free(ad->line_ptr);
  free(ad->line_col);
  free(ad->line_val);
  memset(ad, 0, sizeof(*ad));
}

void orc_distribute_constraints(const orc_adata *ad, double *x)
{
  for (long i = 0; i < ad->n_nodes; ++i) {
    if (ad->line_ptr[i + 1] == ad->line_ptr[i]) continue;
    double s = 0.0;
    for (long t = ad->line_ptr[i]; t < ad->line_ptr[i + 1]; ++t) s += ad->line_val[t] * x[ad->line_col[t]];
    x[i] = s;
  }
}

static int amasked(const orc_operator *op, long n)
{
  return op->use_mask && op->mask && op->mask[n];
}

/* read_dof_values with hanging-node and Dirichlet constraints */
static void agather(const orc_operator *op, const double *src, int nloc, cell_ws_t *ws)
{
  const orc_adata *ad = op->adata;
  for (int l = 0; l < nloc; ++l) {
    const long i = ws->idx[l];
    if (ad->line_ptr[i + 1] > ad->line_ptr[i]) {
      double s = 0.0;
      for (long t = ad->line_ptr[i]; t < ad->line_ptr[i + 1]; ++t)
        if (!amasked(op, ad->line_col[t])) s += ad->line_val[t] * src[ad->line_col[t]];
      ws->u[l] = s;
    } else {
      ws->u[l] = amasked(op, i) ? 0.0 : src[i];
    }
  }
}

/* distribute_local_to_global: hanging rows go to their masters, constrained rows skipped */
static void ascatter(const orc_operator *op, const double *out, int nloc, const long *idx,
                     double *dst)
{
  const orc_adata *ad = op->adata;
  for (int l = 0; l < nloc; ++l) {
    const long i = idx[l];
    if (ad->line_ptr[i + 1] > ad->line_ptr[i]) {
      for (long t = ad->line_ptr[i]; t < ad->line_ptr[i + 1]; ++t)
        if (!amasked(op, ad->line_col[t])) dst[ad->line_col[t]] += ad->line_val[t] * out[l];
    } else if (!amasked(op, i)) {
      dst[i] += out[l];
    }
  }
}

/* heat residual on an adaptive mesh (HeatDiffuseOperator::create_rhs with the HNC of the
 * constrained DoF index, heat_diffuse_operator.cpp:393-529): T, src (T_old) and q_s read as
 * stored at the cell's nodes (constraint-consistent vectors), the cell residual distributed to
 * the masters of its hanging rows and added to dst; constrained rows untouched */
static void arhs(const orc_operator *op, const double *src, const double *q_s, double *dst)
{
  const orc_amesh *am = op->amesh;
  const int K = am->degree + 1, nloc = K * K * K;
  cell_ws_t *ws = (cell_ws_t *)malloc(sizeof(cell_ws_t));
  io_t io = {src, dst, q_s};
  for (long c = 0; c < am->n_cells; ++c) {
    geo_t g;
    acell_geo(am, c, &g);
    for (int l = 0; l < nloc; ++l) ws->idx[l] = op->adata->cell_nodes[c * nloc + l];
    rhs_local(op, &g, c, ws, &io);
    ascatter(op, ws->out, nloc, ws->idx, dst);
  }
  free(ws);
}

static void avmult(const orc_operator *op, const double *src, double *dst)
{
  const orc_amesh *am = op->amesh;
  const int K = am->degree + 1, nloc = K * K * K;
  memset(dst, 0, sizeof(double) * op->adata->n_nodes);
  cell_ws_t *ws = (cell_ws_t *)malloc(sizeof(cell_ws_t));
  for (long c = 0; c < am->n_cells; ++c) {
    geo_t g;
    acell_geo(am, c, &g);
    for (int l = 0; l < nloc; ++l) ws->idx[l] = op->adata->cell_nodes[c * nloc + l];
    agather(op, src, nloc, ws);
    tangent_cell(op, &g, c, ws);
    ascatter(op, ws->out, nloc, ws->idx, dst);
  }
  free(ws);
}

/* exact diagonal of the condensed operator: for every target DoF D of a cell, the local vector
 * c_D (c_D[l] = 1 if local node l is D, the weight of D if l is a hanging node with master D)
 * gives diag[D] += c_D^T A_cell c_D (MatrixFreeTools::compute_diagonal with constraints). */
static void adiagonal(const orc_operator *op, double *diag)
{
  const orc_amesh *am = op->amesh;
  const orc_adata *ad = op->adata;
  const int K = am->degree + 1, nloc = K * K * K;
  memset(diag, 0, sizeof(double) * ad->n_nodes);
  cell_ws_t *ws = (cell_ws_t *)malloc(sizeof(cell_ws_t));
  double *A = (double *)malloc(sizeof(double) * nloc * nloc);
  long *targets = (long *)malloc(sizeof(long) * nloc * 64);
  double *cv = (double *)malloc(sizeof(double) * nloc);
  for (long c = 0; c < am->n_cells; ++c) {
    geo_t g;
    acell_geo(am, c, &g);
    for (int l = 0; l < nloc; ++l) ws->idx[l] = ad->cell_nodes[c * nloc + l];
    /* local matrix, column by column */
    for (int j = 0; j < nloc; ++j) {
      for (int l = 0; l < nloc; ++l) ws->u[l] = (l == j) ? 1.0 : 0.0;
      tangent_cell(op, &g, c, ws);
      for (int l = 0; l < nloc; ++l) A[l * nloc + j] = ws->out[l];
    }
    int nt = 0;
    for (int l = 0; l < nloc; ++l) {
      const long i = ws->idx[l];
      if (ad->line_ptr[i + 1] > ad->line_ptr[i]) {
        for (long t = ad->line_ptr[i]; t < ad->line_ptr[i + 1]; ++t) {
          int found = 0;
          for (int q = 0; q < nt; ++q) found |= targets[q] == ad->line_col[t];
          if (!found) targets[nt++] = ad->line_col[t];
        }
      } else {
        int found = 0;
        for (int q = 0; q < nt; ++q) found |= targets[q] == i;
        if (!found) targets[nt++] = i;
      }
    }
    for (int q = 0; q < nt; ++q) {
      const long D = targets[q];
      if (amasked(op, D)) continue;
      for (int l = 0; l < nloc; ++l) {
        const long i = ws->idx[l];
        double w = 0.0;
        if (ad->line_ptr[i + 1] > ad->line_ptr[i]) {
          for (long t = ad->line_ptr[i]; t < ad->line_ptr[i + 1]; ++t)
            if (ad->line_col[t] == D) w += ad->line_val[t];
        } else if (i == D) {
          w = 1.0;
        }
        cv[l] = w;
      }
      double s = 0.0;
      for (int a = 0; a < nloc; ++a) {
        if (cv[a] == 0.0) continue;
        double r = 0.0;
        for (int b = 0; b < nloc; ++b) r += A[a * nloc + b] * cv[b];
        s += cv[a] * r;
      }
      diag[D] += s;
    }
  }
  free(ws);
  free(A);
  free(targets);
  free(cv);
}

/* ------------------------------------------------------------------------------------------ */
/* Level-set normal vector and curvature right-hand sides (textbook quadrature: basis values   */
/* and gradients as products of 1D Lagrange functions at every quadrature point)               */
/* ------------------------------------------------------------------------------------------ */

double orc_numerical_zero_of_norm(int dim, double volume)
{
  const double x = pow(volume, 1.0 / dim);
  const double e = floor(log10(fabs(x)));
  return fmin(1e-2, fmax(pow(10.0, e) * 1e-3, 1e-12));
}

/* dst += out over the cell's nodes: structured meshes directly, adaptive meshes through the
 * transposed hanging-node constraints (distribute_local_to_global) */
static void ls_add(const orc_operator *op, const long *idx, int nloc, const double *out,
                   double *dst)
{
  if (op->amesh) {
    const orc_adata *ad = op->adata;
    for (int l = 0; l < nloc; ++l) {
      const long i = idx[l];
      if (ad->line_ptr[i + 1] > ad->line_ptr[i]) {
        for (long t = ad->line_ptr[i]; t < ad->line_ptr[i + 1]; ++t)
          dst[ad->line_col[t]] += ad->line_val[t] * out[l];
      } else {
        dst[i] += out[l];
      }
    }
  } else {
    for (int l = 0; l < nloc; ++l) dst[idx[l]] += out[l];
  }
}

static long ls_cells(const orc_operator *op)
{
  return op->amesh ? op->amesh->n_cells : orc_n_cells(&op->mesh);
}

static void ls_cell_geo(const orc_operator *op, long cell, geo_t *g, long *idx)
{
  if (op->amesh) {
    acell_geo(op->amesh, cell, g);
    for (int l = 0; l < g->nloc; ++l) idx[l] = op->adata->cell_nodes[cell * g->nloc + l];
  } else {
    geo_init(&op->mesh, g);
    cell_dofs(g, cell, idx);
  }
}

void orc_ls_normal_rhs(const orc_operator *op, const double *phi, double *dst)
{
  const long n = op_ndofs(op), ncells = ls_cells(op);
  const int dim = op_dim(op);
  for (long cell = 0; cell < ncells; ++cell) {
    geo_t g;
    long idx[NLOC_MAX];
    ls_cell_geo(op, cell, &g, idx);
    const int nq = g.nloc;
    double out[3][NLOC_MAX];
    memset(out, 0, sizeof(out));
    for (int q = 0; q < nq; ++q) {
      double bv[NLOC_MAX], bg[NLOC_MAX][3];
      double pq = 0.0, gq[3] = {0, 0, 0};
      for (int l = 0; l < nq; ++l) {
        basis_at(&g, l, q, &bv[l], bg[l]);
        const double u = phi[idx[l]];
        pq += u * bv[l];
        for (int e = 0; e < 3; ++e) gq[e] += u * bg[l][e];
      }
      const double m = ls_mask(op, pq);
      double v[3];
      for (int e = 0; e < 3; ++e) v[e] = m * gq[e];
      if (op->ls.compute_normalized_vector) {
        /* normalize<dim>(.., 1e-16), include/meltpooldg/utilities/dealii_tensor.hpp:173-191 */
        const double zero = 1e-16;
        const double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int e = 0; e < 3; ++e) v[e] = nrm > zero ? v[e] / nrm : v[e] / zero;
      }
      const double jxw = cell_volume_weight(&g, q);
      for (int d = 0; d < dim; ++d)
        for (int l = 0; l < nq; ++l) out[d][l] += jxw * v[d] * bv[l];
    }
    for (int d = 0; d < dim; ++d) ls_add(op, idx, nq, out[d], dst + d * n);
  }
}

void orc_ls_curvature_rhs(const orc_operator *op, const double *normal, double tol, double *dst)
{
  const long n = op_ndofs(op), ncells = ls_cells(op);
  const int dim = op_dim(op);
  for (long cell = 0; cell < ncells; ++cell) {
    geo_t g;
    long idx[NLOC_MAX];
    ls_cell_geo(op, cell, &g, idx);
    const int nq = g.nloc;
    /* nodal unit normal: n/|n| where |n| > tol, else n/tol (dim 1: the sign, +1 where n > tol) */
    double nh[3][NLOC_MAX];
    for (int l = 0; l < nq; ++l) {
      double v[3] = {0, 0, 0}, nrm = 0.0;
      for (int d = 0; d < dim; ++d) {
        v[d] = normal[d * n + idx[l]];
        nrm += v[d] * v[d];
      }
      nrm = sqrt(nrm);
      for (int d = 0; d < 3; ++d) {
        if (dim == 1)
          nh[d][l] = d == 0 ? (v[0] > tol ? 1.0 : -1.0) : 0.0;
        else
          nh[d][l] = nrm > tol ? v[d] / nrm : v[d] / tol;
      }
    }
    double out[NLOC_MAX];
    memset(out, 0, sizeof(out));
    for (int q = 0; q < nq; ++q) {
      double bv[NLOC_MAX], bg[NLOC_MAX][3];
      double pq = 0.0, div = 0.0;
      for (int l = 0; l < nq; ++l) {
        basis_at(&g, l, q, &bv[l], bg[l]);
        if (op->ls.narrow_band) pq += op->level_set[idx[l]] * bv[l];
        for (int d = 0; d < dim; ++d) div += nh[d][l] * bg[l][d];
      }
      const double val = -ls_mask(op, pq) * div * cell_volume_weight(&g, q);
      for (int l = 0; l < nq; ++l) out[l] += val * bv[l];
    }
    ls_add(op, idx, nq, out, dst);
  }
}

void orc_reinit_rhs(const orc_operator *op, const double *psi, double *dst)
{
  const long n = op_ndofs(op), ncells = ls_cells(op);
  const int dim = op_dim(op);
  for (long cell = 0; cell < ncells; ++cell) {
    geo_t g;
    long idx[NLOC_MAX];
    ls_cell_geo(op, cell, &g, idx);
    const int nq = g.nloc;
    const double en = reinit_eps(op, &g);
    const double et = op->reinit.tangential_diffusion_factor * en;
    double out[NLOC_MAX];
    memset(out, 0, sizeof(out));
    for (int q = 0; q < nq; ++q) {
      double bv[NLOC_MAX], bg[NLOC_MAX][3];
      double pq = 0.0, gq[3] = {0, 0, 0}, nv[3] = {0, 0, 0};
      for (int l = 0; l < nq; ++l) {
        basis_at(&g, l, q, &bv[l], bg[l]);
        const double u = psi[idx[l]];
        pq += u * bv[l];
        for (int e = 0; e < 3; ++e) gq[e] += u * bg[l][e];
        for (int e = 0; e < dim; ++e) nv[e] += op->vel[e * n + idx[l]] * bv[l];
      }
      normalize_tol(nv, dim, op->reinit.normal_tolerance);
      double gn = 0.0;
      for (int e = 0; e < dim; ++e) gn += gq[e] * nv[e];
      const double cf = 0.5 * (1.0 - pq * pq); /* compressive_flux, olsson_CG_operator.cpp:24-29 */
      double f[3];
      for (int e = 0; e < 3; ++e) f[e] = cf * nv[e] - en * gn * nv[e] - et * (gq[e] - gn * nv[e]);
      const double jxw = cell_volume_weight(&g, q);
      for (int l = 0; l < nq; ++l)
        out[l] += jxw * (f[0] * bg[l][0] + f[1] * bg[l][1] + f[2] * bg[l][2]);
    }
    ls_add(op, idx, nq, out, dst);
  }
}

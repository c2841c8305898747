/*
 * orc_cut.c -- CPU oracle of the one-phase CutFEM heat operator (test infrastructure only; see
 * orc.h for the contract and citations).
 *
 * Written the textbook way on purpose: every cell builds its dense element matrix from the
 * basis functions evaluated at each quadrature point of its (sub-)box, every ghost-penalty
 * face its dense face matrix from the normal derivatives of both cells' basis functions at the
 * face's Gauss points; the device path uses sum factorisation instead.  Serial, double
 * precision, dimension-generic (1D-3D), FE_Q1.
 */
#include "orc.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CUT_NLOC 8 /* 2^dim local DoFs of a Q1 cell, dim <= 3 */

typedef struct {
  int dim, nloc;
  long nc[3], NN[3], stride[3];
  double h[3];
} cgeo_t;

static void cgeo_init(const orc_mesh *m, cgeo_t *g)
{
  g->dim = m->dim;
  g->nloc = 1 << m->dim;
  long s = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < m->dim) {
      g->nc[d] = m->ncell[d];
      g->NN[d] = m->ncell[d] + 1;
      g->h[d] = (m->hi[d] - m->lo[d]) / m->ncell[d];
    } else {
      g->nc[d] = 1;
      g->NN[d] = 1;
      g->h[d] = 1.0;
    }
    g->stride[d] = s;
    s *= g->NN[d];
  }
}

static void cell_index(const cgeo_t *g, long cell, long *ci)
{
  ci[0] = cell % g->nc[0];
  ci[1] = (cell / g->nc[0]) % g->nc[1];
  ci[2] = cell / (g->nc[0] * g->nc[1]);
}

/* global DoF of local vertex l (bit d of l = upper vertex along d) */
static long vertex_dof(const cgeo_t *g, const long *ci, int l)
{
  long n = 0;
  for (int d = 0; d < g->dim; ++d) n += (ci[d] + ((l >> d) & 1)) * g->stride[d];
  return n;
}

/* ------------------------------------------------------------------------------------------ */
/* Classification                                                                              */
/* ------------------------------------------------------------------------------------------ */

int orc_cut_classify(const orc_mesh *m, const double *phi, orc_cut_mesh *c)
{
  memset(c, 0, sizeof(*c));
  if (m->degree != 1) return 2;
  cgeo_t g;
  cgeo_init(m, &g);
  const long nc = orc_n_cells(m), nd = orc_n_dofs(m);
  c->n_cells = nc;
  c->n_dofs = nd;
  c->category = (int8_t *)calloc((size_t)nc, 1);
  c->box = (double *)calloc((size_t)nc * 6, sizeof(double));
  c->active = (uint8_t *)calloc((size_t)nd, 1);
  int status = 0;
  for (long cell = 0; cell < nc; ++cell) {
    long ci[3];
    cell_index(&g, cell, ci);
    double v[CUT_NLOC], mn = INFINITY, mx = -INFINITY, scale = 0.0;
    for (int l = 0; l < g.nloc; ++l) {
      v[l] = phi[vertex_dof(&g, ci, l)];
      mn = fmin(mn, v[l]);
      mx = fmax(mx, v[l]);
      scale = fmax(scale, fabs(v[l]));
    }
    double *b = c->box + 6 * cell;
    for (int d = 0; d < 3; ++d) {
      b[2 * d] = 0.0;
      b[2 * d + 1] = 1.0;
    }
    /* phi >= 0 on the whole cell: liquid; touching zero from the positive side does not cut */
    if (mn >= 0.0) {
      c->category[cell] = ORC_CUT_LIQUID;
      continue;
    }
    if (mx < 0.0) {
      c->category[cell] = ORC_CUT_GAS;
      continue;
    }
    c->category[cell] = ORC_CUT_INTERSECTED;
    /* the axes along which phi varies on this cell */
    int axis = -1, n_var = 0;
    for (int d = 0; d < g.dim; ++d) {
      int varies = 0;
      for (int l = 0; l < g.nloc; ++l)
        if (!((l >> d) & 1) && fabs(v[l | (1 << d)] - v[l]) > 1e-12 * scale) varies = 1;
      if (varies) {
        axis = d;
        ++n_var;
      }
    }
    if (n_var != 1) {
      status = 1;
      continue;
    }
    const double p0 = v[0], p1 = v[1 << axis];
    const double t = p0 / (p0 - p1); /* zero of the linear restriction along the axis */
    if (p0 > p1) {
      b[2 * axis] = 0.0;
      b[2 * axis + 1] = t;
    } else {
      b[2 * axis] = t;
      b[2 * axis + 1] = 1.0;
    }
  }
  /* DoFs: vertices of liquid and intersected cells (FE_Q there, FE_Nothing on gas cells) */
  for (long cell = 0; cell < nc; ++cell) {
    if (c->category[cell] == ORC_CUT_GAS) continue;
    long ci[3];
    cell_index(&g, cell, ci);
    for (int l = 0; l < g.nloc; ++l) c->active[vertex_dof(&g, ci, l)] = 1;
  }
  /* ghost-penalty faces: interior faces with an intersected cell on one side and an
   * intersected or liquid cell on the other (CutUtil::get_face_type, source/cut/util.cpp:
   * 40-98; tangent_inner_face_loop, heat_cut_operator.cpp:1266-1318) */
  c->faces = (long *)malloc(sizeof(long) * 2 * (size_t)(nc * g.dim + 1));
  for (long cell = 0; cell < nc; ++cell) {
    long ci[3];
    cell_index(&g, cell, ci);
    for (int d = 0; d < g.dim; ++d) {
      if (ci[d] + 1 >= g.nc[d]) continue;
      long cs = 1;
      for (int e = 0; e < d; ++e) cs *= g.nc[e];
      const int a = c->category[cell], bcat = c->category[cell + cs];
      const int gp = (a == ORC_CUT_INTERSECTED && bcat != ORC_CUT_GAS) ||
                     (bcat == ORC_CUT_INTERSECTED && a != ORC_CUT_GAS);
      if (!gp) continue;
      c->faces[2 * c->n_faces] = cell;
      c->faces[2 * c->n_faces + 1] = d;
      ++c->n_faces;
    }
  }
  return status;
}

void orc_cut_free(orc_cut_mesh *c)
{
  free(c->category);
  free(c->box);
  free(c->active);
  free(c->faces);
  memset(c, 0, sizeof(*c));
}

/* ------------------------------------------------------------------------------------------ */
/* Element and face matrices                                                                   */
/* ------------------------------------------------------------------------------------------ */

static const double GL1[2] = {0.0, 1.0}; /* FE_Q1 support points */

/* quadrature points (reference coordinates) and JxW of the tensor Gauss rule QGauss<1>(2) on
 * the sub-box b of a cell; returns the number of points */
static int box_quadrature(const cgeo_t *g, const double *b, double (*xi)[3], double *jxw)
{
  double x1[2], w1[2];
  orc_gauss(2, x1, w1);
  int nq = 0;
  for (int q = 0; q < g->nloc; ++q) {
    double w = 1.0;
    for (int d = 0; d < 3; ++d) {
      if (d < g->dim) {
        const double lo = b[2 * d], len = b[2 * d + 1] - b[2 * d];
        const int qd = (q >> d) & 1;
        xi[nq][d] = lo + len * x1[qd];
        w *= len * w1[qd] * g->h[d];
      } else {
        xi[nq][d] = 0.0;
      }
    }
    jxw[nq] = w;
    ++nq;
  }
  return nq;
}

/* values and physical gradients of the 2^dim basis functions at reference point xi */
static void basis(const cgeo_t *g, const double *xi, double *val, double (*grad)[3])
{
  double v1[3][2], d1[3][2];
  for (int d = 0; d < g->dim; ++d) orc_lagrange(2, GL1, xi[d], v1[d], d1[d]);
  for (int l = 0; l < g->nloc; ++l) {
    double v = 1.0;
    for (int d = 0; d < g->dim; ++d) v *= v1[d][(l >> d) & 1];
    val[l] = v;
    for (int e = 0; e < 3; ++e) {
      if (e >= g->dim) {
        grad[l][e] = 0.0;
        continue;
      }
      double gr = d1[e][(l >> e) & 1] / g->h[e];
      for (int d = 0; d < g->dim; ++d)
        if (d != e) gr *= v1[d][(l >> d) & 1];
      grad[l][e] = gr;
    }
  }
}

/* material of the liquid phase at T: rho cp, k and their T-derivatives
 * (internal::get_liquid_material_parameters_and_derivatives, heat_cut_operator.cpp:118-145) */
static void cut_material(const orc_cut_params *p, double T, double *cv, double *k, double *dcv,
                         double *dk)
{
  if (p->mat.law == ORC_LAW_MUSHY) {
    double out[4];
    orc_material(&p->mat, T, out);
    *cv = out[0];
    *k = out[1];
    *dcv = out[2];
    *dk = out[3];
  } else {
    *cv = p->mat.rho_cp;
    *k = p->mat.k;
    *dcv = 0.0;
    *dk = 0.0;
  }
}

/* dense tangent element matrix A[a][b] (row a = test function) of a liquid or intersected
 * cell: internal::tangent_domain_integral (heat_cut_operator.cpp:157-190) at every point of
 * the cell's liquid quadrature, without velocity */
static void element_matrix(const cgeo_t *g, const orc_cut_params *p, const double *b,
                           const double *Tloc, double *A)
{
  double xi[CUT_NLOC][3], jxw[CUT_NLOC];
  const int nq = box_quadrature(g, b, xi, jxw);
  const int n = g->nloc;
  const double ost = p->dt * p->theta;
  memset(A, 0, sizeof(double) * n * n);
  for (int q = 0; q < nq; ++q) {
    double val[CUT_NLOC], grad[CUT_NLOC][3];
    basis(g, xi[q], val, grad);
    double Tq = 0.0, gT[3] = {0.0, 0.0, 0.0};
    for (int l = 0; l < n; ++l) {
      Tq += val[l] * Tloc[l];
      for (int e = 0; e < 3; ++e) gT[e] += grad[l][e] * Tloc[l];
    }
    double cv, k, dcv, dk;
    cut_material(p, Tq, &cv, &k, &dcv, &dk);
    for (int a = 0; a < n; ++a)
      for (int bb = 0; bb < n; ++bb) {
        double gg = 0.0, tg = 0.0;
        for (int e = 0; e < 3; ++e) {
          gg += grad[bb][e] * grad[a][e];
          tg += gT[e] * grad[a][e];
        }
        A[a * n + bb] += jxw[q] * ((cv + dcv * Tq) * val[bb] * val[a] +
                                   ost * (k * gg + dk * tg * val[bb]));
      }
  }
}

/* ghost-penalty face between cell c (lower) and its neighbour along d: the face DoFs
 * (2 * 2^dim entries, both cells' vertices) and the dense matrix
 * F = sum_q gamma JxW j_q j_q^T with j_q = (d_n phi^-_l, -d_n phi^+_l) at the face's Gauss
 * points, gamma = dt theta gamma_A k_l h / 3 + gamma_M h^3 / 3
 * (internal::ghost_penalty_face_integral, heat_cut_operator.cpp:425-450) */
static void face_matrix(const cgeo_t *g, const orc_cut_params *p, long cell, int d, long *dofs,
                        double *F)
{
  const int n = g->nloc, m = 2 * n;
  long ci[3], cp[3];
  cell_index(g, cell, ci);
  memcpy(cp, ci, sizeof(ci));
  cp[d] += 1;
  for (int l = 0; l < n; ++l) {
    dofs[l] = vertex_dof(g, ci, l);
    dofs[n + l] = vertex_dof(g, cp, l);
  }
  const double kl = p->mat.law == ORC_LAW_MUSHY ? p->mat.liquid_k : p->mat.k;
  const double h = p->h;
  const double gamma = p->dt * p->theta * p->gamma_A * kl * h / 3.0 + p->gamma_M * h * h * h / 3.0;
  double x1[2], w1[2];
  orc_gauss(2, x1, w1);
  memset(F, 0, sizeof(double) * m * m);
  const int nt = g->dim - 1; /* transversal directions */
  for (int q = 0; q < (1 << nt); ++q) {
    double xm[3] = {0, 0, 0}, xp[3] = {0, 0, 0}, w = 1.0;
    int t = 0;
    for (int e = 0; e < g->dim; ++e) {
      if (e == d) {
        xm[e] = 1.0;
        xp[e] = 0.0;
        continue;
      }
      const int qe = (q >> t) & 1;
      xm[e] = xp[e] = x1[qe];
      w *= w1[qe] * g->h[e];
      ++t;
    }
    double vm[CUT_NLOC], gm[CUT_NLOC][3], vp[CUT_NLOC], gp[CUT_NLOC][3], j[2 * CUT_NLOC];
    basis(g, xm, vm, gm);
    basis(g, xp, vp, gp);
    for (int l = 0; l < n; ++l) {
      j[l] = gm[l][d];
      j[n + l] = -gp[l][d];
    }
    for (int a = 0; a < m; ++a)
      for (int bb = 0; bb < m; ++bb) F[a * m + bb] += gamma * w * j[a] * j[bb];
  }
}

static int row_free(const orc_cut_mesh *c, const uint8_t *dmask, long i)
{
  return c->active[i] && !(dmask && dmask[i]);
}

void orc_cut_vmult(const orc_mesh *m, const orc_cut_mesh *c, const orc_cut_params *p,
                   const uint8_t *dmask, const double *T, const double *src, double *dst)
{
  cgeo_t g;
  cgeo_init(m, &g);
  const int n = g.nloc;
  memset(dst, 0, sizeof(double) * c->n_dofs);
  for (long cell = 0; cell < c->n_cells; ++cell) {
    if (c->category[cell] == ORC_CUT_GAS) continue;
    long ci[3], idx[CUT_NLOC];
    cell_index(&g, cell, ci);
    double u[CUT_NLOC], Tl[CUT_NLOC], A[CUT_NLOC * CUT_NLOC];
    for (int l = 0; l < n; ++l) {
      idx[l] = vertex_dof(&g, ci, l);
      u[l] = row_free(c, dmask, idx[l]) ? src[idx[l]] : 0.0;
      Tl[l] = T ? T[idx[l]] : 0.0;
    }
    element_matrix(&g, p, c->box + 6 * cell, Tl, A);
    for (int a = 0; a < n; ++a) {
      if (!row_free(c, dmask, idx[a])) continue;
      double s = 0.0;
      for (int bb = 0; bb < n; ++bb) s += A[a * n + bb] * u[bb];
      dst[idx[a]] += s;
    }
  }
  for (long f = 0; f < c->n_faces; ++f) {
    long dofs[2 * CUT_NLOC];
    double F[4 * CUT_NLOC * CUT_NLOC], u[2 * CUT_NLOC];
    face_matrix(&g, p, c->faces[2 * f], (int)c->faces[2 * f + 1], dofs, F);
    const int mm = 2 * n;
    for (int a = 0; a < mm; ++a) u[a] = row_free(c, dmask, dofs[a]) ? src[dofs[a]] : 0.0;
    for (int a = 0; a < mm; ++a) {
      if (!row_free(c, dmask, dofs[a])) continue;
      double s = 0.0;
      for (int bb = 0; bb < mm; ++bb) s += F[a * mm + bb] * u[bb];
      dst[dofs[a]] += s;
    }
  }
}

void orc_cut_diagonal(const orc_mesh *m, const orc_cut_mesh *c, const orc_cut_params *p,
                      const uint8_t *dmask, const double *T, double *diag)
{
  cgeo_t g;
  cgeo_init(m, &g);
  const int n = g.nloc;
  memset(diag, 0, sizeof(double) * c->n_dofs);
  for (long cell = 0; cell < c->n_cells; ++cell) {
    if (c->category[cell] == ORC_CUT_GAS) continue;
    long ci[3], idx[CUT_NLOC];
    cell_index(&g, cell, ci);
    double Tl[CUT_NLOC], A[CUT_NLOC * CUT_NLOC];
    for (int l = 0; l < n; ++l) {
      idx[l] = vertex_dof(&g, ci, l);
      Tl[l] = T ? T[idx[l]] : 0.0;
    }
    element_matrix(&g, p, c->box + 6 * cell, Tl, A);
    for (int a = 0; a < n; ++a) diag[idx[a]] += A[a * n + a];
  }
  /* the face matrix's entries that land on the diagonal of the assembled matrix: pairs of
   * face DoFs with the same global index (a shared face vertex appears on both sides) */
  for (long f = 0; f < c->n_faces; ++f) {
    long dofs[2 * CUT_NLOC];
    double F[4 * CUT_NLOC * CUT_NLOC];
    face_matrix(&g, p, c->faces[2 * f], (int)c->faces[2 * f + 1], dofs, F);
    const int mm = 2 * n;
    for (int a = 0; a < mm; ++a)
      for (int bb = 0; bb < mm; ++bb)
        if (dofs[a] == dofs[bb]) diag[dofs[a]] += F[a * mm + bb];
  }
  for (long i = 0; i < c->n_dofs; ++i)
    if (!row_free(c, dmask, i)) diag[i] = 1.0;
}

void orc_cut_residual(const orc_mesh *m, const orc_cut_mesh *c, const orc_cut_params *p,
                      const uint8_t *dmask, const double *T, const double *T_old, double *dst)
{
  cgeo_t g;
  cgeo_init(m, &g);
  const int n = g.nloc;
  const double oi = p->dt * p->theta, oe = p->dt * (1.0 - p->theta);
  memset(dst, 0, sizeof(double) * c->n_dofs);
  for (long cell = 0; cell < c->n_cells; ++cell) {
    if (c->category[cell] == ORC_CUT_GAS) continue;
    long ci[3], idx[CUT_NLOC];
    cell_index(&g, cell, ci);
    double Tn[CUT_NLOC], To[CUT_NLOC], r[CUT_NLOC] = {0};
    for (int l = 0; l < n; ++l) {
      idx[l] = vertex_dof(&g, ci, l);
      Tn[l] = T[idx[l]];
      To[l] = T_old[idx[l]];
    }
    double xi[CUT_NLOC][3], jxw[CUT_NLOC];
    const int nq = box_quadrature(&g, c->box + 6 * cell, xi, jxw);
    for (int q = 0; q < nq; ++q) {
      double val[CUT_NLOC], grad[CUT_NLOC][3];
      basis(&g, xi[q], val, grad);
      double tn = 0, to = 0, gn[3] = {0, 0, 0}, go[3] = {0, 0, 0};
      for (int l = 0; l < n; ++l) {
        tn += val[l] * Tn[l];
        to += val[l] * To[l];
        for (int e = 0; e < 3; ++e) {
          gn[e] += grad[l][e] * Tn[l];
          go[e] += grad[l][e] * To[l];
        }
      }
      double cvn, kn, cvo, ko, dummy1, dummy2;
      cut_material(p, tn, &cvn, &kn, &dummy1, &dummy2);
      cut_material(p, to, &cvo, &ko, &dummy1, &dummy2);
      /* internal::residual_domain_integral (heat_cut_operator.cpp:194-223), no velocity */
      const double v = cvn * tn - cvo * to;
      double fl[3];
      for (int e = 0; e < 3; ++e) fl[e] = kn * oi * gn[e] + ko * oe * go[e];
      for (int a = 0; a < n; ++a) {
        double s = v * val[a];
        for (int e = 0; e < 3; ++e) s += fl[e] * grad[a][e];
        r[a] -= jxw[q] * s;
      }
    }
    for (int a = 0; a < n; ++a)
      if (row_free(c, dmask, idx[a])) dst[idx[a]] += r[a];
  }
  /* residual_inner_face_loop: the ghost penalty on T with factor -1 */
  for (long f = 0; f < c->n_faces; ++f) {
    long dofs[2 * CUT_NLOC];
    double F[4 * CUT_NLOC * CUT_NLOC];
    face_matrix(&g, p, c->faces[2 * f], (int)c->faces[2 * f + 1], dofs, F);
    const int mm = 2 * n;
    for (int a = 0; a < mm; ++a) {
      if (!row_free(c, dmask, dofs[a])) continue;
      double s = 0.0;
      for (int bb = 0; bb < mm; ++bb) s += F[a * mm + bb] * T[dofs[bb]];
      dst[dofs[a]] -= s;
    }
  }
}

double orc_cut_norm_l2(const orc_mesh *m, const orc_cut_mesh *c, const double *u)
{
  cgeo_t g;
  cgeo_init(m, &g);
  const int n = g.nloc;
  double sum = 0.0;
  for (long cell = 0; cell < c->n_cells; ++cell) {
    if (c->category[cell] == ORC_CUT_GAS) continue;
    long ci[3];
    cell_index(&g, cell, ci);
    double xi[CUT_NLOC][3], jxw[CUT_NLOC];
    const int nq = box_quadrature(&g, c->box + 6 * cell, xi, jxw);
    for (int q = 0; q < nq; ++q) {
      double val[CUT_NLOC], grad[CUT_NLOC][3];
      basis(&g, xi[q], val, grad);
      double uq = 0.0;
      for (int l = 0; l < n; ++l) uq += val[l] * u[vertex_dof(&g, ci, l)];
      sum += uq * uq * jxw[q];
    }
  }
  return sqrt(sum);
}

// One-phase CutFEM heat operator on structured FE_Q1 meshes (HeatCutOperator without two_phase,
// source/heat/heat_cut_operator.cpp; HeatCutOperation, source/heat/heat_cut_operation.cpp).
//
// Cells are liquid (the whole QGauss(2) cell), intersected (QGauss(2) per direction on the
// liquid sub-box of the cell) or gas (no contribution).  The cell kernel walks cell rows along x,
// one workgroup per row, in 4 launches by the parity of (cell y, cell z): rows of one launch
// share no node, the node shared by x-neighbours is combined through LDS, so every node is
// written by one thread per launch with a plain read-modify-write.  A cell's 8 nodes are four
// x-pairs of neighbouring entries (coalesced across the threads of a row).  Values and
// gradients at the 8 quadrature points
// are sum-factorised 2x2 sweeps with the sub-box's 1D tables; the gradient along a direction of
// a Q1 function does not depend on the position along that direction.  The ghost-penalty faces
// (a thin layer around the interface) run one thread per face with atomic adds.
#include "mf_internal.hpp"
#include "mf_vec.hpp"

#include <algorithm>
#include <cmath>
#include <string>

namespace mf {

namespace {

constexpr int32_t CELL_LIQUID = -1;  // cell table: >= 0 = index of the intersected cell's box
constexpr int32_t CELL_GAS = -2;
constexpr int CUT_THREADS = 256;
constexpr int CUT_NORM_BLOCKS = 256;
constexpr double GAUSS_OFF = 0.28867513459481287;  // 1 / (2 sqrt(3)): QGauss(2) on [0, 1]

}  // namespace

struct CutData {
  bool set = false;
  int64_t n_liquid = 0, n_cut = 0, n_gas = 0, n_faces = 0, n_active = 0;
  std::vector<int8_t> category;  // host copies
  std::vector<double> box;
  std::vector<uint8_t> active;
  int32_t *d_cell = nullptr;      // [n_cells] CELL_LIQUID, CELL_GAS or the box index
  double *d_box = nullptr;        // [n_cut][6] liquid sub-box in reference coordinates
  double *d_mats = nullptr;       // [n_cut][3][4] their 1D matrices (m00, m01, m11, kk)
  double *d_faceG = nullptr;      // [n_cut][6][4] work: ghost-penalty face values of one apply
  int64_t *d_face = nullptr;      // [n_faces] (lower cell << 2) | direction
  uint8_t *d_inactive = nullptr;  // [n_dofs] 1 = no DoF (only gas cells around the node)
  double *d_part = nullptr;       // [CUT_NORM_BLOCKS] norm partials
  // node kernel flags (NODE_*), built with the classification on first use
  uint8_t *d_nflag = nullptr;
  int64_t *d_glist = nullptr;     // the nodes with a non-zero flag
  int64_t n_glist = 0;
  // the non-gas cells with a general node among their corners: their element rows, formed once
  // per application by gcell_block and gathered by the general nodes
  int32_t *d_gslot = nullptr;     // [n_cells] slot of a general cell, -1 otherwise
  int64_t *d_gcell = nullptr;     // [n_gcell] the general cells
  double *d_gout = nullptr;       // [n_gcell][8] work: their element rows of one application
  int64_t n_gcell = 0;
  // ghost-penalty faces of each general node (CSR over the general list): entry = index into
  // faceG of the node's corner value | direction << 28 | plane << 30
  int32_t *d_gfptr = nullptr;
  uint32_t *d_gfent = nullptr;
  std::vector<int64_t> faces;     // host copy of the ghost-penalty faces
};

namespace {

struct CutArgs {
  int ncell[3];
  int64_t NX, NXY;  // nodes per x line and per xy plane
  double h[3];
  double cv, k, ost_i, ost_e;  // constant law; dt theta and dt (1 - theta)
  int mushy;
  double mu_k[2], mu_rho[2], mu_cp[2], mu_Tliq, mu_inv;  // liquid / solid, liquidus, 1/interval
  double gamma;            // ghost-penalty coefficient
  const int32_t *cell;
  const double *box;
  const uint8_t *dmask;    // Dirichlet rows (nullable)
  const uint8_t *inactive; // nodes without a DoF
  const double *T;         // linearisation temperature (tangent, diagonal), T^{n+1} (residual)
  const int *stop;
  int colour;
  const uint8_t *nflag;    // node kernel: NODE_GENERAL / NODE_INACTIVE per node
  // per direction: h/6, h/3, 1/h and gamma times the JxW of a face point (0.25 h_d1 h_d2)
  double h6[3], h3[3], ih[3], wq[3];
  const double *mats;      // [n_cut][3][4] the 1D matrices of the intersected cells
  double *faceG;           // [n_cut][6][4] per face of an intersected cell: G_c of the face
  uint64_t mg_nx, mg_nxy;  // ceil(2^64 / d) of the node counts per x line and per xy plane
  const int32_t *gslot;    // general cells: slot per cell (-1: none)
  double *gout;            // [n_gcell][8] element rows of the general cells
  const int32_t *gfptr;    // ghost-penalty faces per general node (CSR)
  const uint32_t *gfent;
};

// n / d for n < 2^32 by multiply-high with m = ceil(2^64 / d), 2 <= d < 2^32
__device__ __forceinline__ uint32_t div_magic(uint32_t n, uint64_t m)
{
  return uint32_t(__umul64hi(uint64_t(n), m));
}

// node flags of the constant-law node kernel: 0 = the node, all its 27 neighbours and its 8
// cells are free and liquid and it is no ghost-penalty node (the uniform stencil applies)
constexpr uint8_t NODE_GENERAL = 1, NODE_INACTIVE = 2;
// a node without a DoF that the node kernel treats itself (not in the general list)
constexpr uint8_t NODE_INACTIVE_ONLY = NODE_INACTIVE;

__device__ __forceinline__ bool fixed_row(const CutArgs &A, int64_t n)
{
  return A.inactive[n] || (A.dmask && A.dmask[n]);
}

// The liquid material and its T-derivatives (get_liquid_material_parameters_and_derivatives,
// heat_cut_operator.cpp:118-145): constant, or the liquid_solid law with the mushy-zone blend
// l + (s - l)(3 f^2 - 2 f^3) of the solid fraction f = clamp((T_liq - T)/(T_liq - T_sol), 0, 1)
// (material.templates.hpp:546-600,748-780), rho cp = rho cp and d(rho cp)/dT = cp' rho + rho' cp.
__device__ __forceinline__ void cut_law(const CutArgs &A, double T, double &cv, double &k,
                                        double &dcv, double &dk)
{
  if (!A.mushy) {
    cv = A.cv;
    k = A.k;
    dcv = 0.0;
    dk = 0.0;
    return;
  }
  double f = (A.mu_Tliq - T) * A.mu_inv;
  f = fmin(fmax(f, 0.0), 1.0);
  const bool pure = f == 0.0 || f == 1.0;
  const double H = f * f * (3.0 - 2.0 * f);
  const double dH = pure ? 0.0 : -A.mu_inv * (6.0 * f - 6.0 * f * f);
  auto blend = [&](const double (&v)[2]) {
    return f == 0.0 ? v[0] : (f == 1.0 ? v[1] : v[0] + (v[1] - v[0]) * H);
  };
  k = blend(A.mu_k);
  dk = (A.mu_k[1] - A.mu_k[0]) * dH;
  const double rho = blend(A.mu_rho), cp = blend(A.mu_cp);
  cv = rho * cp;
  dcv = (A.mu_cp[1] - A.mu_cp[0]) * dH * rho + (A.mu_rho[1] - A.mu_rho[0]) * dH * cp;
}

// 1D tables of one direction of a cell on its liquid interval [a, b] (reference coordinates):
// S[q][i] = l_i(t_q) at the two Gauss points of [a, b], D[i] = l_i' / h, w[q] = physical weight
struct Box1D {
  double S[2][2];
  double D[2];
  double w[2];
};

__device__ __forceinline__ void box_1d(double a, double b, double h, Box1D &o)
{
  const double len = b - a;
  const double t0 = a + len * (0.5 - GAUSS_OFF), t1 = a + len * (0.5 + GAUSS_OFF);
  o.S[0][0] = 1.0 - t0;
  o.S[0][1] = t0;
  o.S[1][0] = 1.0 - t1;
  o.S[1][1] = t1;
  o.D[0] = -1.0 / h;
  o.D[1] = 1.0 / h;
  o.w[0] = o.w[1] = 0.5 * len * h;
}

// Values v[qz][qy][qx] and gradients at the quadrature points of the Q1 function with nodal
// values u[ez][ey][ex]: gx depends on (qz, qy) only, gy on (qz, qx), gz on (qy, qx).
struct QVals {
  double v[2][2][2], gx[2][2], gy[2][2], gz[2][2];
};

__device__ __forceinline__ void eval_q1(const Box1D (&B)[3], const double (&u)[2][2][2], QVals &o)
{
  double vx[2][2][2], dx[2][2];
#pragma unroll
  for (int ez = 0; ez < 2; ++ez)
#pragma unroll
    for (int ey = 0; ey < 2; ++ey) {
#pragma unroll
      for (int q = 0; q < 2; ++q) vx[ez][ey][q] = B[0].S[q][0] * u[ez][ey][0] + B[0].S[q][1] * u[ez][ey][1];
      dx[ez][ey] = B[0].D[0] * u[ez][ey][0] + B[0].D[1] * u[ez][ey][1];
    }
  double vy[2][2][2], gyp[2][2], dxy[2][2];
#pragma unroll
  for (int ez = 0; ez < 2; ++ez) {
#pragma unroll
    for (int qy = 0; qy < 2; ++qy) {
#pragma unroll
      for (int qx = 0; qx < 2; ++qx)
        vy[ez][qy][qx] = B[1].S[qy][0] * vx[ez][0][qx] + B[1].S[qy][1] * vx[ez][1][qx];
      dxy[ez][qy] = B[1].S[qy][0] * dx[ez][0] + B[1].S[qy][1] * dx[ez][1];
    }
#pragma unroll
    for (int qx = 0; qx < 2; ++qx) gyp[ez][qx] = B[1].D[0] * vx[ez][0][qx] + B[1].D[1] * vx[ez][1][qx];
  }
#pragma unroll
  for (int qz = 0; qz < 2; ++qz) {
#pragma unroll
    for (int qy = 0; qy < 2; ++qy) {
#pragma unroll
      for (int qx = 0; qx < 2; ++qx)
        o.v[qz][qy][qx] = B[2].S[qz][0] * vy[0][qy][qx] + B[2].S[qz][1] * vy[1][qy][qx];
      o.gx[qz][qy] = B[2].S[qz][0] * dxy[0][qy] + B[2].S[qz][1] * dxy[1][qy];
    }
#pragma unroll
    for (int qx = 0; qx < 2; ++qx) o.gy[qz][qx] = B[2].S[qz][0] * gyp[0][qx] + B[2].S[qz][1] * gyp[1][qx];
  }
#pragma unroll
  for (int qy = 0; qy < 2; ++qy)
#pragma unroll
    for (int qx = 0; qx < 2; ++qx) o.gz[qy][qx] = B[2].D[0] * vy[0][qy][qx] + B[2].D[1] * vy[1][qy][qx];
}

enum { MODE_TANGENT = 0, MODE_DIAG = 1, MODE_RESIDUAL = 2 };

// The local operation of one liquid or intersected cell.  MODE_TANGENT: tangent_domain_integral
// (heat_cut_operator.cpp:157-190, no velocity) applied to u; MODE_DIAG: the diagonal of that
// element matrix; MODE_RESIDUAL: residual_domain_integral (:194-223) of T^{n+1} = t and T^n = u.
// The 4 quadrature points of the plane qz from the z-interpolated nodal values uz[ey][ex] and
// the z-derivatives dz[ey][ex] (the same on both planes): v[qy][qx], gx[qy], gy[qx], gz[qy][qx].
struct PlaneVals {
  double v[2][2], gx[2], gy[2], gz[2][2];
};

__device__ __forceinline__ void eval_plane(const Box1D (&B)[3], const double (&uz)[2][2],
                                           const double (&dz)[2][2], PlaneVals &o)
{
  double uy[2][2], dzy[2][2], dy[2];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int ex = 0; ex < 2; ++ex) {
      uy[q][ex] = B[1].S[q][0] * uz[0][ex] + B[1].S[q][1] * uz[1][ex];
      dzy[q][ex] = B[1].S[q][0] * dz[0][ex] + B[1].S[q][1] * dz[1][ex];
    }
#pragma unroll
  for (int ex = 0; ex < 2; ++ex) dy[ex] = B[1].D[0] * uz[0][ex] + B[1].D[1] * uz[1][ex];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
#pragma unroll
    for (int qx = 0; qx < 2; ++qx) {
      o.v[q][qx] = B[0].S[qx][0] * uy[q][0] + B[0].S[qx][1] * uy[q][1];
      o.gz[q][qx] = B[0].S[qx][0] * dzy[q][0] + B[0].S[qx][1] * dzy[q][1];
    }
    o.gx[q] = B[0].D[0] * uy[q][0] + B[0].D[1] * uy[q][1];
    o.gy[q] = B[0].S[q][0] * dy[0] + B[0].S[q][1] * dy[1];
  }
}

// out[ez][ey][ex] += the plane qz's part of sum_q fv phi + f . grad phi (transposed sweeps)
__device__ __forceinline__ void integrate_plane(const Box1D (&B)[3], int qz,
                                                const double (&fv)[2][2],
                                                const double (&fx)[2][2],
                                                const double (&fy)[2][2],
                                                const double (&fz)[2][2], double (&out)[2][2][2])
{
  double Vx[2][2], Yx[2][2], Zx[2][2];
#pragma unroll
  for (int qy = 0; qy < 2; ++qy)
#pragma unroll
    for (int ex = 0; ex < 2; ++ex) {
      Vx[qy][ex] = B[0].S[0][ex] * fv[qy][0] + B[0].S[1][ex] * fv[qy][1] +
                   B[0].D[ex] * (fx[qy][0] + fx[qy][1]);
      Yx[qy][ex] = B[0].S[0][ex] * fy[qy][0] + B[0].S[1][ex] * fy[qy][1];
      Zx[qy][ex] = B[0].S[0][ex] * fz[qy][0] + B[0].S[1][ex] * fz[qy][1];
    }
#pragma unroll
  for (int ey = 0; ey < 2; ++ey)
#pragma unroll
    for (int ex = 0; ex < 2; ++ex) {
      const double V = B[1].S[0][ey] * Vx[0][ex] + B[1].S[1][ey] * Vx[1][ex] +
                       B[1].D[ey] * (Yx[0][ex] + Yx[1][ex]);
      const double Z = B[1].S[0][ey] * Zx[0][ex] + B[1].S[1][ey] * Zx[1][ex];
#pragma unroll
      for (int ez = 0; ez < 2; ++ez) out[ez][ey][ex] += B[2].S[qz][ez] * V + B[2].D[ez] * Z;
    }
}

template <int MODE>
__device__ __forceinline__ void cut_cell_local(const CutArgs &A, const Box1D (&B)[3],
                                               const double (&u)[2][2][2],
                                               const double (&t)[2][2][2],
                                               double (&out)[2][2][2])
{
  const bool needT = A.mushy || MODE == MODE_RESIDUAL;
#pragma unroll
  for (int e = 0; e < 8; ++e) out[e >> 2][(e >> 1) & 1][e & 1] = 0.0;
  if (MODE != MODE_DIAG) {
    // one plane of quadrature points at a time: few values live at once
    double du[2][2], dt[2][2];
#pragma unroll
    for (int ey = 0; ey < 2; ++ey)
#pragma unroll
      for (int ex = 0; ex < 2; ++ex) {
        du[ey][ex] = B[2].D[0] * u[0][ey][ex] + B[2].D[1] * u[1][ey][ex];
        dt[ey][ex] = needT ? B[2].D[0] * t[0][ey][ex] + B[2].D[1] * t[1][ey][ex] : 0.0;
      }
#pragma unroll
    for (int qz = 0; qz < 2; ++qz) {
      double uz[2][2], tz[2][2];
#pragma unroll
      for (int ey = 0; ey < 2; ++ey)
#pragma unroll
        for (int ex = 0; ex < 2; ++ex) {
          uz[ey][ex] = B[2].S[qz][0] * u[0][ey][ex] + B[2].S[qz][1] * u[1][ey][ex];
          tz[ey][ex] = needT ? B[2].S[qz][0] * t[0][ey][ex] + B[2].S[qz][1] * t[1][ey][ex] : 0.0;
        }
      PlaneVals U, Tp;
      eval_plane(B, uz, du, U);
      if (needT) eval_plane(B, tz, dt, Tp);
      double fv[2][2], fx[2][2], fy[2][2], fz[2][2];
#pragma unroll
      for (int qy = 0; qy < 2; ++qy)
#pragma unroll
        for (int qx = 0; qx < 2; ++qx) {
          const double jxw = B[0].w[qx] * B[1].w[qy] * B[2].w[qz];
          double cv = A.cv, k = A.k, dcv = 0.0, dk = 0.0;
          double tq = 0.0, gT[3] = {0.0, 0.0, 0.0};
          if (needT) {
            tq = Tp.v[qy][qx];
            gT[0] = Tp.gx[qy];
            gT[1] = Tp.gy[qx];
            gT[2] = Tp.gz[qy][qx];
            if (A.mushy) cut_law(A, tq, cv, k, dcv, dk);
          }
          const double uq = U.v[qy][qx];
          const double g[3] = {U.gx[qy], U.gy[qx], U.gz[qy][qx]};
          if (MODE == MODE_TANGENT) {
            fv[qy][qx] = (cv + dcv * tq) * uq * jxw;
            const double a = A.ost_i * k * jxw, b = A.ost_i * dk * uq * jxw;
            fx[qy][qx] = a * g[0] + b * gT[0];
            fy[qy][qx] = a * g[1] + b * gT[1];
            fz[qy][qx] = a * g[2] + b * gT[2];
          } else {
            // residual: u holds T^n; its material is the constant law or the law at T^n
            double cvo = A.cv, ko = A.k, d1, d2;
            if (A.mushy) cut_law(A, uq, cvo, ko, d1, d2);
            fv[qy][qx] = -(cv * tq - cvo * uq) * jxw;
            fx[qy][qx] = -(A.ost_i * k * gT[0] + A.ost_e * ko * g[0]) * jxw;
            fy[qy][qx] = -(A.ost_i * k * gT[1] + A.ost_e * ko * g[1]) * jxw;
            fz[qy][qx] = -(A.ost_i * k * gT[2] + A.ost_e * ko * g[2]) * jxw;
          }
        }
      integrate_plane(B, qz, fv, fx, fy, fz, out);
    }
    return;
  }
  QVals Tq;
  if (needT) eval_q1(B, t, Tq);
#pragma unroll
  for (int qz = 0; qz < 2; ++qz)
#pragma unroll
    for (int qy = 0; qy < 2; ++qy)
#pragma unroll
      for (int qx = 0; qx < 2; ++qx) {
        const double jxw = B[0].w[qx] * B[1].w[qy] * B[2].w[qz];
        double cv = A.cv, k = A.k, dcv = 0.0, dk = 0.0;
        double tq = 0.0, gT[3] = {0.0, 0.0, 0.0};
        if (needT) {
          tq = Tq.v[qz][qy][qx];
          gT[0] = Tq.gx[qz][qy];
          gT[1] = Tq.gy[qz][qx];
          gT[2] = Tq.gz[qy][qx];
          if (A.mushy) cut_law(A, tq, cv, k, dcv, dk);
        }
        {
          // diagonal: (cv + dcv T) phi_e^2 + dt theta (k |grad phi_e|^2 + dk (grad T . grad phi_e) phi_e)
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const int ex = e & 1, ey = (e >> 1) & 1, ez = e >> 2;
            const double sx = B[0].S[qx][ex], sy = B[1].S[qy][ey], sz = B[2].S[qz][ez];
            const double phi = sx * sy * sz;
            const double g0 = B[0].D[ex] * sy * sz, g1 = sx * B[1].D[ey] * sz,
                         g2 = sx * sy * B[2].D[ez];
            out[ez][ey][ex] +=
                jxw * ((cv + dcv * tq) * phi * phi +
                       A.ost_i * (k * (g0 * g0 + g1 * g1 + g2 * g2) +
                                  dk * (gT[0] * g0 + gT[1] * g1 + gT[2] * g2) * phi));
          }
        }
      }
}

__device__ __forceinline__ void cell_boxes(const CutArgs &A, int32_t cc, Box1D (&B)[3])
{
  if (cc >= 0) {
    const double *b = A.box + 6 * int64_t(cc);
#pragma unroll
    for (int d = 0; d < 3; ++d) box_1d(b[2 * d], b[2 * d + 1], A.h[d], B[d]);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) box_1d(0.0, 1.0, A.h[d], B[d]);
  }
}

// Constant law: the element matrix of a liquid or intersected cell is exactly
//   cv Mx(x)My(x)Mz + dt theta k (Kx(x)My(x)Mz + Mx(x)Ky(x)Mz + Mx(x)My(x)Kz)
// with the 1D matrices of the cell's liquid interval [a, b] (reference coordinates):
// M_ij = h int_a^b l_i l_j dt, K_ij = (b - a) / h (-1)^(i+j) (l_0 = 1 - t, l_1 = t).  QGauss(2) on
// [a, b] integrates these degree-2 products exactly, so this is the quadrature of the reference
// in closed form.
struct Mat1D {
  double m00, m01, m11, kk;
};

__device__ __forceinline__ void cell_mats(const CutArgs &A, int32_t cc, Mat1D (&M)[3])
{
  if (cc >= 0) {
    const double *t = A.mats + 12 * int64_t(cc);  // computed with the classification
#pragma unroll
    for (int d = 0; d < 3; ++d) M[d] = Mat1D{t[4 * d], t[4 * d + 1], t[4 * d + 2], t[4 * d + 3]};
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) M[d] = Mat1D{A.h3[d], A.h6[d], A.h3[d], A.ih[d]};
  }
}

// 2x2 applications along one direction of a [2][2][2] array (index [ez][ey][ex])
template <int DIR>
__device__ __forceinline__ void apply_m(const Mat1D &m, const double (&in)[2][2][2],
                                        double (&out)[2][2][2])
{
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int i0 = p >> 1, i1 = p & 1;  // the two other indices
    auto at = [&](const double (&x)[2][2][2], int e) -> double {
      return DIR == 0 ? x[i0][i1][e] : (DIR == 1 ? x[i0][e][i1] : x[e][i0][i1]);
    };
    auto put = [&](double (&x)[2][2][2], int e, double v) {
      if (DIR == 0) x[i0][i1][e] = v; else if (DIR == 1) x[i0][e][i1] = v; else x[e][i0][i1] = v;
    };
    const double v0 = at(in, 0), v1 = at(in, 1);
    put(out, 0, m.m00 * v0 + m.m01 * v1);
    put(out, 1, m.m01 * v0 + m.m11 * v1);
  }
}

template <int DIR>
__device__ __forceinline__ void apply_k(const Mat1D &m, const double (&in)[2][2][2],
                                        double (&out)[2][2][2])
{
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int i0 = p >> 1, i1 = p & 1;
    auto at = [&](const double (&x)[2][2][2], int e) -> double {
      return DIR == 0 ? x[i0][i1][e] : (DIR == 1 ? x[i0][e][i1] : x[e][i0][i1]);
    };
    auto put = [&](double (&x)[2][2][2], int e, double v) {
      if (DIR == 0) x[i0][i1][e] = v; else if (DIR == 1) x[i0][e][i1] = v; else x[e][i0][i1] = v;
    };
    const double d = m.kk * (at(in, 0) - at(in, 1));
    put(out, 0, d);
    put(out, 1, -d);
  }
}

// out = Mx(x)My(x)Mz p + (Kx(x)My(x)Mz + Mx(x)Ky(x)Mz + Mx(x)My(x)Kz) q
__device__ __forceinline__ void tensor_apply(const Mat1D (&M)[3], const double (&p)[2][2][2],
                                             const double (&q)[2][2][2], double (&out)[2][2][2])
{
  double P[2][2][2], Q[2][2][2], R[2][2][2];
  apply_m<2>(M[2], p, P);
  apply_m<2>(M[2], q, Q);
  apply_k<2>(M[2], q, R);
  double P2[2][2][2], Q2[2][2][2], S2[2][2][2], T2[2][2][2];
  apply_m<1>(M[1], P, P2);
  apply_m<1>(M[1], Q, Q2);
  apply_k<1>(M[1], Q, S2);
  apply_m<1>(M[1], R, T2);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int a = e >> 2, b = (e >> 1) & 1, c = e & 1;
    P2[a][b][c] += S2[a][b][c] + T2[a][b][c];
  }
  double X[2][2][2];
  apply_m<0>(M[0], P2, out);
  apply_k<0>(M[0], Q2, X);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int a = e >> 2, b = (e >> 1) & 1, c = e & 1;
    out[a][b][c] += X[a][b][c];
  }
}

// out = (cv Mx(x)My(x)Mz + c (Kx(x)My(x)Mz + Mx(x)Ky(x)Mz + Mx(x)My(x)Kz)) u, few values live
__device__ __forceinline__ void tensor_apply_1(const Mat1D (&M)[3], double cv, double c,
                                               const double (&u)[2][2][2], double (&out)[2][2][2])
{
  double P[2][2][2], R[2][2][2];
  apply_m<2>(M[2], u, P);
  apply_k<2>(M[2], u, R);
  double P2[2][2][2], S2[2][2][2], T2[2][2][2];
  apply_m<1>(M[1], P, P2);
  apply_k<1>(M[1], P, S2);
  apply_m<1>(M[1], R, T2);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int a = e >> 2, b = (e >> 1) & 1, x = e & 1;
    S2[a][b][x] = cv * P2[a][b][x] + c * (S2[a][b][x] + T2[a][b][x]);
  }
  double X[2][2][2];
  apply_m<0>(M[0], S2, out);
  apply_k<0>(M[0], P2, X);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int a = e >> 2, b = (e >> 1) & 1, x = e & 1;
    out[a][b][x] += c * X[a][b][x];
  }
}

// One launch per (y, z) row-parity colour; one workgroup per cell row, CUT_THREADS cells per
// step.  dst += the cell contributions of the rows of this colour (rows skipped: fixed rows).
// MUSHY: the T-dependent law on the quadrature points; otherwise the closed-form tensor path.
template <int MODE, bool MUSHY>
__global__ __launch_bounds__(CUT_THREADS) void cut_cell_kernel(CutArgs A, const double *src,
                                                               double *dst)
{
  if (A.stop && *A.stop) return;
  const RowColour rc = row_colour(A.ncell, A.colour);
  const int row = blockIdx.x;
  const int cy = rc.yb + 2 * (row % rc.hy), cz = rc.zb + 2 * (row / rc.hy);
  const int ncx = A.ncell[0];
  __shared__ double lds[4][CUT_THREADS];
  __shared__ double carry[4];
  int64_t base[4];  // node row r = ey + 2 ez
#pragma unroll
  for (int r = 0; r < 4; ++r) base[r] = int64_t(cz + (r >> 1)) * A.NXY + int64_t(cy + (r & 1)) * A.NX;
  const int tid = threadIdx.x;
  const bool needT = MUSHY || MODE == MODE_RESIDUAL;
  for (int x0 = 0; x0 < ncx; x0 += CUT_THREADS) {
    const int i = x0 + tid;
    double out[2][2][2];
#pragma unroll
    for (int e = 0; e < 8; ++e) out[e >> 2][(e >> 1) & 1][e & 1] = 0.0;
    if (i < ncx) {
      const int32_t cc = A.cell[(int64_t(cz) * A.ncell[1] + cy) * ncx + i];
      if (cc != CELL_GAS) {
        double u[2][2][2], t[2][2][2];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t n = base[r] + i;
#pragma unroll
          for (int ex = 0; ex < 2; ++ex) {
            double uv = 0.0, tv = 0.0;
            if (MODE == MODE_TANGENT) uv = fixed_row(A, n + ex) ? 0.0 : src[n + ex];
            if (MODE == MODE_RESIDUAL) uv = src[n + ex];  // T^n, read plain
            if (needT) tv = A.T[n + ex];
            u[r >> 1][r & 1][ex] = uv;
            t[r >> 1][r & 1][ex] = tv;
          }
        }
        if constexpr (MUSHY) {
          Box1D B[3];
          cell_boxes(A, cc, B);
          cut_cell_local<MODE>(A, B, u, t, out);
        } else {
          Mat1D M[3];
          cell_mats(A, cc, M);
          const double c = A.ost_i * A.k;
          if (MODE == MODE_TANGENT) {
            tensor_apply_1(M, A.cv, c, u, out);
          } else if (MODE == MODE_RESIDUAL) {
            // -R = -(cv M (T - T^n) + k K (dt theta T + dt (1 - theta) T^n))
            double p[2][2][2], q[2][2][2];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
              const int a = e >> 2, b = (e >> 1) & 1, x = e & 1;
              p[a][b][x] = -A.cv * (t[a][b][x] - u[a][b][x]);
              q[a][b][x] = -A.k * (A.ost_i * t[a][b][x] + A.ost_e * u[a][b][x]);
            }
            tensor_apply(M, p, q, out);
          } else {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
              const int a = e >> 2, b = (e >> 1) & 1, x = e & 1;
              const double mx = x ? M[0].m11 : M[0].m00, my = b ? M[1].m11 : M[1].m00,
                           mz = a ? M[2].m11 : M[2].m00;
              out[a][b][x] = A.cv * mx * my * mz +
                             c * (M[0].kk * my * mz + mx * M[1].kk * mz + mx * my * M[2].kk);
            }
          }
        }
      }
    }
    // the right nodes of each cell go to its x-neighbour
#pragma unroll
    for (int r = 0; r < 4; ++r) lds[r][tid] = out[r >> 1][r & 1][1];
    __syncthreads();
    double left[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) left[r] = tid > 0 ? lds[r][tid - 1] : (x0 > 0 ? carry[r] : 0.0);
    __syncthreads();
    if (tid == CUT_THREADS - 1) {
#pragma unroll
      for (int r = 0; r < 4; ++r) carry[r] = out[r >> 1][r & 1][1];
    }
    if (i < ncx) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t n = base[r] + i;
        if (!fixed_row(A, n)) dst[n] += out[r >> 1][r & 1][0] + left[r];
        if (i == ncx - 1 && !fixed_row(A, n + 1)) dst[n + 1] += out[r >> 1][r & 1][1];
      }
    }
  }
}

// Ghost-penalty faces (ghost_penalty_face_integral, heat_cut_operator.cpp:425-450), one thread
// per face between the lower cell and its neighbour along d.  With node planes p-1, p, p+1
// along d and the 4 corners c of the face, the jump of the normal derivative at the face point
// s is sum_c L_c(s) (2 u_{p,c} - u_{p-1,c} - u_{p+1,c}) / h_d (bilinear L_c), so a test
// function at (plane, c) sees coef_plane L_c(s) / h_d with coef = (-1, 2, -1).
// MODE_TANGENT: dst += G src (fixed entries read as 0); MODE_DIAG: dst += diag G;
// MODE_RESIDUAL: dst -= G T (T read plain).  Fixed rows are skipped.
template <int MODE>
__global__ __launch_bounds__(256) void cut_face_kernel(CutArgs A, const int64_t *faces,
                                                       int64_t n_faces, const double *src,
                                                       double *dst)
{
  if (A.stop && *A.stop) return;
  const int64_t f = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (f >= n_faces) return;
  const int64_t code = faces[f];
  const int d = int(code & 3);
  const int64_t cell = code >> 2;
  const int64_t ncx = A.ncell[0], ncy = A.ncell[1];
  const uint32_t c32 = uint32_t(cell), row = c32 / uint32_t(ncx);
  const int64_t ci = c32 - row * uint32_t(ncx), cj = row % uint32_t(ncy), ck = row / uint32_t(ncy);
  const int64_t n0 = ck * A.NXY + cj * A.NX + ci;
  const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
  const int64_t stride[3] = {1, A.NX, A.NXY};
  const int64_t sd = stride[d], s1 = stride[d1], s2 = stride[d2];
  const double ih = 1.0 / A.h[d];
  const double wq = 0.25 * A.h[d1] * A.h[d2] * A.gamma;  // gamma times JxW of each face point
  const double g[2] = {0.5 - GAUSS_OFF, 0.5 + GAUSS_OFF};
  // L[q][c]: corner c = c1 + 2 c2 at face point q = q1 + 2 q2
  double L[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double a = g[q & 1], b = g[q >> 1];
      L[q][c] = ((c & 1) ? a : 1.0 - a) * ((c >> 1) ? b : 1.0 - b);
    }
  const double coef[3] = {-1.0, 2.0, -1.0};
  if (MODE == MODE_DIAG) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) s += wq * L[q][c] * L[q][c];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const int64_t n = n0 + p * sd + (c & 1) * s1 + (c >> 1) * s2;
        if (!fixed_row(A, n)) unsafeAtomicAdd(dst + n, coef[p] * coef[p] * ih * ih * s);
      }
    }
    return;
  }
  double J[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double v[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t n = n0 + p * sd + (c & 1) * s1 + (c >> 1) * s2;
      v[p] = MODE == MODE_RESIDUAL ? A.T[n] : (fixed_row(A, n) ? 0.0 : src[n]);
    }
    J[c] = (2.0 * v[1] - v[0] - v[2]) * ih;
  }
  double G[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const double jump = L[q][0] * J[0] + L[q][1] * J[1] + L[q][2] * J[2] + L[q][3] * J[3];
#pragma unroll
    for (int c = 0; c < 4; ++c) G[c] += wq * L[q][c] * jump;
  }
  const double sign = MODE == MODE_RESIDUAL ? -1.0 : 1.0;
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t n = n0 + p * sd + (c & 1) * s1 + (c >> 1) * s2;
      if (!fixed_row(A, n)) unsafeAtomicAdd(dst + n, sign * coef[p] * ih * G[c]);
    }
}

__device__ __forceinline__ double block_sum_256(double v, double *red)
{
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

// CutUtil::compute_cut_norm (L2), per-block partial sums of u^2 JxW over the liquid quadrature
__global__ __launch_bounds__(256) void cut_norm_kernel(CutArgs A, const double *u, int64_t n_cells,
                                                       double *part)
{
  __shared__ double red[256];
  double s = 0.0;
  const int64_t ncx = A.ncell[0], ncy = A.ncell[1];
  for (int64_t cell = int64_t(blockIdx.x) * 256 + threadIdx.x; cell < n_cells;
       cell += int64_t(gridDim.x) * 256) {
    const int32_t cc = A.cell[cell];
    if (cc == CELL_GAS) continue;
    Box1D B[3];
    cell_boxes(A, cc, B);
    const uint32_t c32 = uint32_t(cell), row = c32 / uint32_t(ncx);
    const int64_t ci = c32 - row * uint32_t(ncx), cj = row % uint32_t(ncy), ck = row / uint32_t(ncy);
    const int64_t n0 = ck * A.NXY + cj * A.NX + ci;
    double v[2][2][2];
#pragma unroll
    for (int e = 0; e < 8; ++e)
      v[e >> 2][(e >> 1) & 1][e & 1] =
          u[n0 + (e & 1) + ((e >> 1) & 1) * A.NX + (e >> 2) * A.NXY];
    QVals U;
    eval_q1(B, v, U);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int qx = q & 1, qy = (q >> 1) & 1, qz = q >> 2;
      const double uq = U.v[qz][qy][qx];
      s += uq * uq * B[0].w[qx] * B[1].w[qy] * B[2].w[qz];
    }
  }
  const double t = block_sum_256(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// partials of a . b for the fused p.(Ap) of CG (one per block)
__global__ __launch_bounds__(256) void cut_dot_kernel(const double *a, const double *b, int64_t n,
                                                      double *part, const int *stop)
{
  if (stop && *stop) return;
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
    s += a[i] * b[i];
  const double t = block_sum_256(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ------------------------------------------------------------------------------------------
// Constant law: one thread per node, each node written once (no colours, no atomics).
// A node's row of J is the sum over its 8 cells of that cell's row, plus the ghost-penalty
// faces whose 12 nodes contain it.  Nodes whose 8 cells are liquid and whose neighbourhood is
// free (flag 0) take the separable 27-point stencil of the uniform mesh:
//   cv Sx Sy Sz + c (Dx Sy Sz + Sx Dy Sz + Sx Sy Dz),  S = (h/6, 2h/3, h/6), D = (-1, 2, -1)/h.
// The others (flag NODE_GENERAL: near the interface, the box boundary or a Dirichlet node) sum
// their cells' rows with the cells' 1D matrices and the faces explicitly.
// ------------------------------------------------------------------------------------------

// row a = (ax, ay, az) of Mx(x)My(x)Mz p + (Kx(x)My(x)Mz + Mx(x)Ky(x)Mz + Mx(x)My(x)Kz) q
// The general path of node (i, j, k), entry idx of the general list, the share of `lane` (0..7):
// the row of the lane's cell (one of the node's 8) and of the node's ghost-penalty faces lane,
// lane + 8, ...  The 8 shares sum to the node's row.
// MODE_TANGENT: (J src)_n; MODE_RESIDUAL: (-R)_n with src = T^n, A.T = T; MODE_DIAG: J_nn.
template <int MODE>
__device__ __forceinline__ double node_general(const CutArgs &A, const double *src, int64_t i,
                                               int64_t j, int64_t k, int64_t idx, int lane)
{
  const int64_t ncx = A.ncell[0], ncy = A.ncell[1], ncz = A.ncell[2];
  const int64_t NX = A.NX, NXY = A.NXY;
  const double c = A.ost_i * A.k;
  double y = 0.0;
  {
    const int64_t cx = i - 1 + (lane & 1), cy = j - 1 + ((lane >> 1) & 1), cz = k - 1 + (lane >> 2);
    const bool in = cx >= 0 && cx < ncx && cy >= 0 && cy < ncy && cz >= 0 && cz < ncz;
    const int64_t cell = (cz * ncy + cy) * ncx + cx;
    const int ax = int(i - cx), ay = int(j - cy), az = int(k - cz);
    if (MODE == MODE_DIAG) {
      const int32_t cc = in ? A.cell[cell] : CELL_GAS;
      if (cc != CELL_GAS) {
        Mat1D M[3];
        cell_mats(A, cc, M);
        const double mx = ax ? M[0].m11 : M[0].m00, my = ay ? M[1].m11 : M[1].m00,
                     mz = az ? M[2].m11 : M[2].m00;
        y += A.cv * mx * my * mz + c * (M[0].kk * my * mz + mx * M[1].kk * mz + mx * my * M[2].kk);
      }
    } else {
      // the row of this corner, formed once per cell by gcell_block (every non-gas cell
      // of a general node is a general cell; gas cells have no slot)
      const int32_t slot = in ? A.gslot[cell] : -1;
      if (slot >= 0) y += A.gout[8 * int64_t(slot) + ax + 2 * ay + 4 * az];
    }
  }
  // the ghost-penalty faces with this node among their 12 nodes (listed with the level set):
  // the lane's share of them
  const double coef[3] = {-1.0, 2.0, -1.0};
  for (int32_t e = A.gfptr[idx] + lane; e < A.gfptr[idx + 1]; e += 8) {
    const uint32_t code = A.gfent[e];
    const int d = int((code >> 28) & 3u), sp = int(code >> 30);
    const double ih = A.ih[d];
    if (MODE == MODE_DIAG) {
      // sum_q wq L_q(corner)^2 for the bilinear L at the 2 x 2 Gauss points
      const double g0 = 0.5 - GAUSS_OFF, g1 = 0.5 + GAUSS_OFF;
      const double s2 = (g0 * g0 + g1 * g1) * (g0 * g0 + g1 * g1);
      y += coef[sp] * coef[sp] * ih * ih * A.wq[d] * s2;
    } else {
      // G of this corner, computed once per face by face_g_block
      const double G = A.faceG[code & 0x0fffffffu];
      y += (MODE == MODE_RESIDUAL ? -1.0 : 1.0) * coef[sp] * ih * G;
    }
  }
  return y;
}

// Element rows of the general cells, one thread per cell: out[e] (corner e = ex + 2 ey + 4 ez)
// = (Mx(x)My(x)Mz p + (Kx(x)My(x)Mz + Mx(x)Ky(x)Mz + Mx(x)My(x)Kz) q)_e with the cell's 1D
// matrices (sub-box of an intersected cell, the full cell otherwise).  MODE_TANGENT: p = cv u,
// q = dt theta k u (u = src, fixed entries 0); MODE_RESIDUAL: p = -cv (T - T^n),
// q = -k (dt theta T + dt (1 - theta) T^n) with src = T^n.
template <int MODE>
__device__ __forceinline__ void gcell_block(const CutArgs &A, const int64_t *gcell,
                                            int64_t n_gcell, const double *src, int64_t blk)
{
  const int64_t g = blk * 256 + threadIdx.x;
  if (g >= n_gcell) return;
  const int64_t cell = gcell[g];
  const int64_t ncx = A.ncell[0], ncy = A.ncell[1];
  const int64_t cx = cell % ncx, cy = (cell / ncx) % ncy, cz = cell / (ncx * ncy);
  const double c = A.ost_i * A.k;
  Mat1D M[3];
  cell_mats(A, A.cell[cell], M);
  double p[2][2][2], q[2][2][2];
  const int64_t n0 = cz * A.NXY + cy * A.NX + cx;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int ex = e & 1, ey = (e >> 1) & 1, ez = e >> 2;
    const int64_t m = n0 + ex + ey * A.NX + ez * A.NXY;
    if (MODE == MODE_TANGENT) {
      const double u = fixed_row(A, m) ? 0.0 : src[m];
      p[ez][ey][ex] = A.cv * u;
      q[ez][ey][ex] = c * u;
    } else {
      const double t = A.T[m], to = src[m];
      p[ez][ey][ex] = -A.cv * (t - to);
      q[ez][ey][ex] = -A.k * (A.ost_i * t + A.ost_e * to);
    }
  }
  double out[2][2][2];
  tensor_apply(M, p, q, out);
  double *o = A.gout + 8 * g;
#pragma unroll
  for (int e = 0; e < 8; ++e) o[e] = out[e >> 2][(e >> 1) & 1][e & 1];
}

// G_c = sum_q gamma JxW_q L_q(c) [d_n u](q) of every ghost-penalty face (c = its 4 corners),
// one thread per face; stored with the face's intersected cell(s) (slot 2d + 1: the upper face
// of the lower cell, slot 2d: the lower face of the upper cell).  MODE_TANGENT: u = src with
// fixed entries 0; MODE_RESIDUAL: u = T read plain.
template <int MODE>
__device__ __forceinline__ void face_g_block(const CutArgs &A, const int64_t *faces,
                                             int64_t n_faces, const double *src, int64_t blk)
{
  const int64_t f = blk * 256 + threadIdx.x;
  if (f >= n_faces) return;
  const int64_t code = faces[f];
  const int d = int(code & 3);
  const int64_t cell = code >> 2;
  const int64_t ncx = A.ncell[0], ncy = A.ncell[1];
  const uint32_t c32 = uint32_t(cell), row = c32 / uint32_t(ncx);
  const int64_t ci = c32 - row * uint32_t(ncx), cj = row % uint32_t(ncy), ck = row / uint32_t(ncy);
  const int64_t n0 = ck * A.NXY + cj * A.NX + ci;
  const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
  const int64_t stride[3] = {1, A.NX, A.NXY};
  const int64_t sd = stride[d], s1 = stride[d1], s2 = stride[d2];
  const double ih = A.ih[d], wq = A.wq[d];
  const double g[2] = {0.5 - GAUSS_OFF, 0.5 + GAUSS_OFF};
  double J[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double v[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const int64_t n = n0 + p * sd + (c & 1) * s1 + (c >> 1) * s2;
      v[p] = MODE == MODE_RESIDUAL ? A.T[n] : (fixed_row(A, n) ? 0.0 : src[n]);
    }
    J[c] = (2.0 * v[1] - v[0] - v[2]) * ih;
  }
  double G[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double L[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double x = g[q & 1], z = g[q >> 1];
      L[c] = ((c & 1) ? x : 1.0 - x) * ((c >> 1) ? z : 1.0 - z);
    }
    const double jump = L[0] * J[0] + L[1] * J[1] + L[2] * J[2] + L[3] * J[3];
#pragma unroll
    for (int c = 0; c < 4; ++c) G[c] += wq * L[c] * jump;
  }
  const int64_t sc[3] = {1, ncx, ncx * ncy};
  const int32_t a = A.cell[cell], b = A.cell[cell + sc[d]];
  if (a >= 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) A.faceG[24 * int64_t(a) + 4 * (2 * d + 1) + c] = G[c];
  }
  if (b >= 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) A.faceG[24 * int64_t(b) + 4 * (2 * d) + c] = G[c];
  }
}

// 1D stencils of one direction at a node with lower / upper cell present (l, r in {0, 1}):
// S = (l h/6, (l + r) h/3, r h/6), D = (-l, l + r, -r) / h; the offsets of the neighbours, a
// missing one read at the node itself (its weight is 0)
struct Stencil1D {
  double S[3], D[3];
  int64_t off[3];
};

__device__ __forceinline__ Stencil1D stencil_1d(uint32_t idx, uint32_t last, int64_t stride,
                                                double h6, double h3, double ih)
{
  const double l = idx > 0 ? 1.0 : 0.0, r = idx < last ? 1.0 : 0.0;
  Stencil1D s;
  s.S[0] = l * h6;
  s.S[1] = (l + r) * h3;
  s.S[2] = r * h6;
  s.D[0] = -l * ih;
  s.D[1] = (l + r) * ih;
  s.D[2] = -r * ih;
  s.off[0] = idx > 0 ? -stride : 0;
  s.off[1] = 0;
  s.off[2] = idx < last ? stride : 0;
  return s;
}

// The y part of the stencil in one layer at node n (row j): with Sy = (l h/6, (l + r) h/3, r h/6)
// and Dy = (-l, l + r, -r)/h, tangent: ys = Sy u, yd = Dy u (u = src with Dirichlet entries 0);
// residual: ys = Sy p, yd = Dy q, yq = Sy q with p = -cv (T - T^n), q = -k (dt theta T +
// dt (1 - theta) T^n), src = T^n.  The centre value u (src[n]) is returned for the fused dot.
struct LayerY {
  double ys, yd, yq, uc;
};

template <int MODE>
__device__ __forceinline__ LayerY layer_y(const CutArgs &A, const double *src, int64_t n,
                                          const Stencil1D &sy)
{
  LayerY o{0.0, 0.0, 0.0, 0.0};
  if (MODE == MODE_TANGENT) {
    double u[3];
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      const int64_t mm = n + sy.off[dy];
      u[dy] = (A.dmask && A.dmask[mm]) ? 0.0 : src[mm];
    }
    o.ys = sy.S[0] * u[0] + sy.S[1] * u[1] + sy.S[2] * u[2];
    o.yd = sy.D[0] * u[0] + sy.D[1] * u[1] + sy.D[2] * u[2];
    o.uc = src[n];
  } else if (MODE == MODE_RESIDUAL) {
    double pv[3], qv[3];
#pragma unroll
    for (int dy = 0; dy < 3; ++dy) {
      const int64_t mm = n + sy.off[dy];
      const double t = A.T[mm], to = src[mm];
      pv[dy] = -A.cv * (t - to);
      qv[dy] = -A.k * (A.ost_i * t + A.ost_e * to);
    }
    o.ys = sy.S[0] * pv[0] + sy.S[1] * pv[1] + sy.S[2] * pv[2];
    o.yd = sy.D[0] * qv[0] + sy.D[1] * qv[1] + sy.D[2] * qv[2];
    o.yq = sy.S[0] * qv[0] + sy.S[1] * qv[1] + sy.S[2] * qv[2];
  }
  return o;
}

// MODE_TANGENT: dst = J src (overwrite) or dst += J src, with partials of src . (J src) per
// block when dot_partial is set; MODE_DIAG: dst = diag J (1 on fixed rows); MODE_RESIDUAL:
// dst += -R(T) with src = T^n.  A block covers NODE_CHUNK = 254 consecutive nodes of a layer
// (chunks in XCD-major order) and NODE_LAYERS layers, marching along z: thread t keeps the y
// parts (layer_y) of its in-layer node base + t - 1 for the layers z - 1, z, z + 1, combines
// them with the z stencil, stages the result in LDS, and threads 1..254 do the x-pass of their
// node from the staged neighbours (the first and last thread only provide the halo).  Each
// layer is read once per block, plus the two layers before the block's first.  At the box
// boundary a direction has one cell only: its stencils are the one-sided ones; across the end
// of an x line the weight is 0.
constexpr int NODE_CHUNK = 254;
constexpr int NODE_LAYERS = 16;

template <int MODE>
__device__ __forceinline__ void node_block(const CutArgs &A, const double *src, double *dst,
                                           int overwrite, double *dot_partial, int64_t blk,
                                           int64_t nblk)
{
  __shared__ double red[256];
  __shared__ double lm[256], ls[256];
  const int64_t NX = A.NX, NXY = A.NXY;
  const uint32_t nx = uint32_t(NX);
  const uint32_t lastx = nx - 1, lasty = uint32_t(A.ncell[1]), lastz = uint32_t(A.ncell[2]);
  const int64_t n_layer_chunks = (NXY + NODE_CHUNK - 1) / NODE_CHUNK;
  const int64_t n_z_blocks = (int64_t(lastz) + NODE_LAYERS) / NODE_LAYERS;
  const double c = A.ost_i * A.k;
  const double *h6 = A.h6, *h3 = A.h3, *ih = A.ih;
  const int tid = threadIdx.x;
  // consecutive chunks on one XCD (the hardware deals blocks to the 8 XCDs round-robin): the
  // y neighbours a block reads were then fetched into the same L2 by its neighbours
  const int64_t nb8 = nblk / 8;
  const int64_t chunk = (blk % 8) * nb8 + blk / 8;
  double dot = 0.0;
  if (chunk < n_layer_chunks * n_z_blocks) {
    const int64_t xy_chunk = chunk % n_layer_chunks;
    const uint32_t z0 = uint32_t(chunk / n_layer_chunks) * NODE_LAYERS;
    const uint32_t z1 = min(z0 + NODE_LAYERS, lastz + 1);
    const int64_t q = xy_chunk * NODE_CHUNK + tid - 1;  // node in the layer
    const bool in = q >= 0 && q < NXY;
    const bool own = tid >= 1 && tid <= NODE_CHUNK && in;
    // 32-bit index arithmetic (the node count is below 2^32: checked with the level set)
    const uint32_t q32 = uint32_t(in ? q : xy_chunk * NODE_CHUNK);
    const uint32_t j = div_magic(q32, A.mg_nx), i = q32 - j * nx;
    const Stencil1D sy = stencil_1d(j, lasty, NX, h6[1], h3[1], ih[1]);
    const Stencil1D sx = stencil_1d(i, lastx, 1, h6[0], h3[0], ih[0]);
    // y parts of the layers z - 1 (lo), z (mid), z + 1 (hi)
    LayerY lo{0.0, 0.0, 0.0, 0.0}, mid{0.0, 0.0, 0.0, 0.0}, hi{0.0, 0.0, 0.0, 0.0};
    if (MODE != MODE_DIAG) {
      if (z0 > 0) lo = layer_y<MODE>(A, src, int64_t(q32) + int64_t(z0 - 1) * NXY, sy);
      mid = layer_y<MODE>(A, src, int64_t(q32) + int64_t(z0) * NXY, sy);
    }
    for (uint32_t k = z0; k < z1; ++k) {
      const int64_t n = int64_t(q32) + int64_t(k) * NXY;
      const uint8_t f = in ? A.nflag[n] : NODE_INACTIVE;
      if (MODE != MODE_DIAG) {
        if (k < lastz) hi = layer_y<MODE>(A, src, n + NXY, sy);
        const Stencil1D sz = stencil_1d(k, lastz, NXY, h6[2], h3[2], ih[2]);
        double m = 0.0, s = 0.0;
        if (f != NODE_INACTIVE) {
          if (MODE == MODE_TANGENT) {
            // sa = Sz Sy u, da = Sz Dy u, sb = Dz Sy u
            const double sa = sz.S[0] * lo.ys + sz.S[1] * mid.ys + sz.S[2] * hi.ys;
            const double da = sz.S[0] * lo.yd + sz.S[1] * mid.yd + sz.S[2] * hi.yd;
            const double sb = sz.D[0] * lo.ys + sz.D[1] * mid.ys + sz.D[2] * hi.ys;
            m = A.cv * sa + c * (da + sb);
            s = c * sa;
          } else {
            const double pa = sz.S[0] * lo.ys + sz.S[1] * mid.ys + sz.S[2] * hi.ys;
            const double qa = sz.S[0] * lo.yq + sz.S[1] * mid.yq + sz.S[2] * hi.yq;
            const double qk = sz.S[0] * lo.yd + sz.S[1] * mid.yd + sz.S[2] * hi.yd;
            const double qb = sz.D[0] * lo.yq + sz.D[1] * mid.yq + sz.D[2] * hi.yq;
            m = pa + qk + qb;
            s = qa;
          }
        }
        lm[tid] = m;
        ls[tid] = s;
        __syncthreads();
      }
      if (own) {
        const bool fixed = f == NODE_INACTIVE || (A.dmask && A.dmask[n]);
        if (fixed) {
          // a node without a free DoF: 0 (tangent), 1 (diagonal), unchanged (residual, add)
          if (MODE == MODE_DIAG)
            dst[n] = 1.0;
          else if (MODE == MODE_TANGENT && overwrite)
            dst[n] = 0.0;
        } else if (f == 0) {
          double val;
          if (MODE == MODE_DIAG) {
            const Stencil1D sz = stencil_1d(k, lastz, NXY, h6[2], h3[2], ih[2]);
            val = A.cv * sx.S[1] * sy.S[1] * sz.S[1] +
                  c * (sx.D[1] * sy.S[1] * sz.S[1] + sx.S[1] * sy.D[1] * sz.S[1] +
                       sx.S[1] * sy.S[1] * sz.D[1]);
          } else {
            val = sx.S[0] * lm[tid - 1] + sx.S[1] * lm[tid] + sx.S[2] * lm[tid + 1] +
                  sx.D[0] * ls[tid - 1] + sx.D[1] * ls[tid] + sx.D[2] * ls[tid + 1];
          }
          if (overwrite)
            dst[n] = val;
          else
            dst[n] += val;
          if (MODE == MODE_TANGENT && dot_partial) dot += mid.uc * val;
        }
      }
      if (MODE != MODE_DIAG) {
        __syncthreads();  // lm, ls are rewritten by the next layer
        lo = mid;
        mid = hi;
      }
    }
  }
  if (MODE == MODE_TANGENT && dot_partial) {
    const double t = block_sum_256(dot, red);
    if (tid == 0) dot_partial[blk] = t;
  }
}


// The general nodes (flag != 0, a list built on the host): fixed rows (no DoF, Dirichlet) get
// 0 (tangent), 1 (diagonal) or stay (residual, vmult_add); the others sum their cells' rows
// and their ghost-penalty faces.  Fused-dot partials follow the node kernel's.
template <int MODE>
__device__ __forceinline__ void general_block(const CutArgs &A, const int64_t *list,
                                              int64_t n_list, const double *src, double *dst,
                                              int overwrite, double *dot_partial, int64_t blk,
                                              int64_t nblk)
{
  __shared__ double red[256];
  const int64_t NX = A.NX, NXY = A.NXY;
  const int lane = threadIdx.x & 7;  // 8 consecutive lanes share a node
  double dot = 0.0;
  for (int64_t idx = (blk * 256 + threadIdx.x) >> 3; idx < n_list; idx += nblk * 32) {
    const int64_t n = list[idx];
    double val;
    bool write = true;
    if (A.dmask && A.dmask[n]) {
      val = MODE == MODE_DIAG ? 1.0 : 0.0;
      write = MODE == MODE_DIAG || (MODE == MODE_TANGENT && overwrite);
    } else {
      const uint32_t n32 = uint32_t(n), nxy = uint32_t(NXY), nx = uint32_t(NX);
      const uint32_t k32 = div_magic(n32, A.mg_nxy), r32 = n32 - k32 * nxy;
      const uint32_t j32 = div_magic(r32, A.mg_nx);
      val = node_general<MODE>(A, src, r32 - j32 * nx, j32, k32, idx, lane);
      val += __shfl_xor(val, 1);
      val += __shfl_xor(val, 2);
      val += __shfl_xor(val, 4);
      if (MODE == MODE_TANGENT && dot_partial && lane == 0) dot += src[n] * val;
    }
    if (write && lane == 0) {
      if (overwrite)
        dst[n] = val;
      else
        dst[n] += val;
    }
  }
  if (MODE == MODE_TANGENT && dot_partial) {
    const double t = block_sum_256(dot, red);
    if (threadIdx.x == 0) dot_partial[blk] = t;
  }
}

// One application in two launches.  The first forms the ghost-penalty face values (blocks
// [0, fb)) and the element rows of the general cells ([fb, fb + gb)); the second runs the
// general nodes, which read both (blocks [0, genb)), beside the node stencil (the remaining nb
// blocks).  genb is a multiple of 8, so that the node blocks keep their XCD-major chunk order.
template <int MODE>
__global__ __launch_bounds__(256) void cut_pre_kernel(CutArgs A, const double *src,
                                                      const int64_t *faces, int64_t n_faces,
                                                      int64_t fb, const int64_t *gcell,
                                                      int64_t n_gcell)
{
  if (A.stop && *A.stop) return;  // uniform over the grid
  const int64_t blk = blockIdx.x;
  if (blk < fb)
    face_g_block<MODE>(A, faces, n_faces, src, blk);
  else
    gcell_block<MODE>(A, gcell, n_gcell, src, blk - fb);
}

template <int MODE>
__global__ __launch_bounds__(256) void cut_apply_kernel(CutArgs A, const double *src,
                                                        double *dst, int overwrite,
                                                        double *dot_partial,
                                                        const int64_t *list, int64_t n_list,
                                                        int64_t genb, int64_t nb)
{
  if (A.stop && *A.stop) return;  // uniform over the grid
  const int64_t blk = blockIdx.x;
  if (blk < genb)
    general_block<MODE>(A, list, n_list, src, dst, overwrite,
                        dot_partial ? dot_partial + nb : nullptr, blk, genb);
  else
    node_block<MODE>(A, src, dst, overwrite, dot_partial, blk - genb, nb);
}

CutArgs make_cut_args(const mf_operator *op)
{
  const MeshInfo &m = op->s->mesh;
  const mf_heat_cut_params &p = op->cut;
  const mf_heat_params &mp = p.material;
  CutArgs A{};
  for (int d = 0; d < 3; ++d) {
    A.ncell[d] = m.ncell[d];
    A.h[d] = m.h[d];
  }
  A.NX = m.NN[0];
  A.NXY = m.NN[0] * m.NN[1];
  A.mushy = mp.law == MF_LAW_MUSHY ? 1 : 0;
  A.cv = mp.rho_cp;
  A.k = mp.k;
  A.mu_k[0] = mp.liquid_k;
  A.mu_k[1] = mp.solid_k;
  A.mu_rho[0] = mp.liquid_rho;
  A.mu_rho[1] = mp.solid_rho;
  A.mu_cp[0] = mp.liquid_cp;
  A.mu_cp[1] = mp.solid_cp;
  A.mu_Tliq = mp.liquidus;
  A.mu_inv = A.mushy ? 1.0 / (mp.liquidus - mp.solidus) : 0.0;
  A.ost_i = op->dt * p.theta;
  A.ost_e = op->dt * (1.0 - p.theta);
  // cell_side_length = ScratchData::get_min_cell_size: min cell diameter / sqrt(dim)
  const double h = p.cell_side_length > 0.0
                       ? p.cell_side_length
                       : std::sqrt((m.h[0] * m.h[0] + m.h[1] * m.h[1] + m.h[2] * m.h[2]) / 3.0);
  const double kl = A.mushy ? mp.liquid_k : mp.k;
  A.gamma = A.ost_i * p.gamma_A * kl * h / 3.0 + p.gamma_M * h * h * h / 3.0;
  for (int d = 0; d < 3; ++d) {
    A.h6[d] = m.h[d] / 6.0;
    A.h3[d] = m.h[d] / 3.0;
    A.ih[d] = 1.0 / m.h[d];
    const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
    A.wq[d] = 0.25 * m.h[d1] * m.h[d2] * A.gamma;
  }
  A.mg_nx = ~uint64_t(0) / uint64_t(m.NN[0]) + 1;
  A.mg_nxy = ~uint64_t(0) / uint64_t(m.NN[0] * m.NN[1]) + 1;
  A.mats = op->cutd->d_mats;
  A.faceG = op->cutd->d_faceG;
  A.cell = op->cutd->d_cell;
  A.box = op->cutd->d_box;
  A.dmask = op->dof_index == 0 ? op->s->d_mask : nullptr;
  A.inactive = op->cutd->d_inactive;
  A.T = op->T;
  return A;
}

template <int MODE>
int launch_cut(const mf_operator *op, CutArgs A, const double *src, double *dst)
{
  const MeshInfo &m = op->s->mesh;
  hipStream_t st = op->s->stream;
  for (int colour = 0; colour < 4; ++colour) {
    const RowColour rc = row_colour(m.ncell, colour);
    const int64_t rows = int64_t(rc.hy) * rc.hz;
    if (rows == 0) continue;
    A.colour = colour;
    if (A.mushy)
      cut_cell_kernel<MODE, true><<<dim3(unsigned(rows)), dim3(CUT_THREADS), 0, st>>>(A, src, dst);
    else
      cut_cell_kernel<MODE, false><<<dim3(unsigned(rows)), dim3(CUT_THREADS), 0, st>>>(A, src, dst);
    MF_HIP_TRY(hipGetLastError());
  }
  const CutData *cd = op->cutd;
  if (cd->n_faces > 0) {
    const unsigned blocks = unsigned((cd->n_faces + 255) / 256);
    cut_face_kernel<MODE><<<dim3(blocks), dim3(256), 0, st>>>(A, cd->d_face, cd->n_faces, src, dst);
    MF_HIP_TRY(hipGetLastError());
  }
  return MF_OK;
}

void free_tables(CutData *cd)
{
  if (cd->d_cell) (void)hipFree(cd->d_cell);
  if (cd->d_box) (void)hipFree(cd->d_box);
  if (cd->d_mats) (void)hipFree(cd->d_mats);
  if (cd->d_faceG) (void)hipFree(cd->d_faceG);
  cd->d_mats = nullptr;
  cd->d_faceG = nullptr;
  if (cd->d_face) (void)hipFree(cd->d_face);
  if (cd->d_inactive) (void)hipFree(cd->d_inactive);
  if (cd->d_part) (void)hipFree(cd->d_part);
  if (cd->d_nflag) (void)hipFree(cd->d_nflag);
  if (cd->d_glist) (void)hipFree(cd->d_glist);
  if (cd->d_gslot) (void)hipFree(cd->d_gslot);
  if (cd->d_gcell) (void)hipFree(cd->d_gcell);
  if (cd->d_gout) (void)hipFree(cd->d_gout);
  if (cd->d_gfptr) (void)hipFree(cd->d_gfptr);
  if (cd->d_gfent) (void)hipFree(cd->d_gfent);
  cd->d_gfptr = nullptr;
  cd->d_gfent = nullptr;
  cd->d_gslot = nullptr;
  cd->d_gcell = nullptr;
  cd->d_gout = nullptr;
  cd->n_gcell = 0;
  cd->d_nflag = nullptr;
  cd->d_glist = nullptr;
  cd->n_glist = 0;
  cd->d_cell = nullptr;
  cd->d_box = nullptr;
  cd->d_face = nullptr;
  cd->d_inactive = nullptr;
  cd->d_part = nullptr;
  cd->set = false;
}

// Flags of the node kernel: NODE_INACTIVE for the nodes without a DoF, NODE_GENERAL for the
// nodes of intersected and gas cells and of ghost-penalty faces; built on the host with the
// level set.  (The Dirichlet mask is read by the kernels themselves.)
int ensure_node_flags(mf_operator *op)
{
  CutData *cd = op->cutd;
  const mf_scratch *sc = op->s;
  if (cd->d_nflag) return MF_OK;
  const MeshInfo &m = sc->mesh;
  const int64_t nx = m.ncell[0], ny = m.ncell[1], nz = m.ncell[2];
  const int64_t NX = m.NN[0], NY = m.NN[1], NZ = m.NN[2], NXY = NX * NY;
  std::vector<uint8_t> f(size_t(m.n_dofs), 0);
  auto mark = [&](int64_t n) { f[size_t(n)] |= NODE_GENERAL; };
  // the nodes of intersected and gas cells
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i)
        if (cd->category[size_t((k * ny + j) * nx + i)] != 0)
          for (int e = 0; e < 8; ++e)
            mark((k + (e >> 2)) * NXY + (j + ((e >> 1) & 1)) * NX + i + (e & 1));
  // the 12 nodes of every ghost-penalty face
  for (int64_t code : cd->faces) {
    const int d = int(code & 3);
    const int64_t c = code >> 2;
    const int64_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    const int64_t sn[3] = {1, NX, NXY};
    const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
    const int64_t n0 = k * NXY + j * NX + i;
    for (int pl = 0; pl < 3; ++pl)
      for (int cr = 0; cr < 4; ++cr) mark(n0 + pl * sn[d] + (cr & 1) * sn[d1] + (cr >> 1) * sn[d2]);
  }
  // nodes without a DoF: the node kernel writes their fixed value (flag NODE_INACTIVE alone)
  for (int64_t n = 0; n < m.n_dofs; ++n)
    if (!cd->active[size_t(n)]) f[size_t(n)] = NODE_INACTIVE;
  std::vector<int64_t> list;
  for (int64_t n = 0; n < m.n_dofs; ++n)
    if (f[size_t(n)] & NODE_GENERAL) list.push_back(n);
  // the non-gas cells with a general node among their corners, and their slots
  std::vector<int32_t> gslot(size_t(nx * ny * nz), -1);
  std::vector<int64_t> gcell;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t cl = (k * ny + j) * nx + i;
        if (cd->category[size_t(cl)] == 2) continue;
        bool g = false;
        for (int e = 0; e < 8 && !g; ++e)
          g = (f[size_t((k + (e >> 2)) * NXY + (j + ((e >> 1) & 1)) * NX + i + (e & 1))] &
               NODE_GENERAL) != 0;
        if (g) {
          MF_REQUIRE(gcell.size() < size_t(INT32_MAX), MF_ERR_ARG,
                     "too many general cells for the 32-bit slot table");
          gslot[size_t(cl)] = int32_t(gcell.size());
          gcell.push_back(cl);
        }
      }
  // the ghost-penalty faces of each general node: the 36 candidates (3 directions x 3 planes of
  // the lower cell x 4 transversal positions), kept where the face joins two non-gas cells one
  // of which is intersected; the corner value sits with the intersected cell (upper face of the
  // lower cell, slot 2 d + 1, or lower face of the upper cell, slot 2 d)
  std::vector<int32_t> cutidx(size_t(nx * ny * nz), -1);
  {
    int32_t nc = 0;
    for (size_t cl = 0; cl < cutidx.size(); ++cl)
      if (cd->category[cl] == 1) cutidx[cl] = nc++;
  }
  MF_REQUIRE(24 * cd->n_cut < (int64_t(1) << 28), MF_ERR_ARG,
             "too many intersected cells for the face tables (24 n_cut >= 2^28)");
  std::vector<int32_t> gfptr(1, 0);
  std::vector<uint32_t> gfent;
  {
    const int64_t lim[3] = {nx, ny, nz}, scl[3] = {1, nx, nx * ny};
    for (int64_t n : list) {
      const int64_t nidx[3] = {n % NX, (n / NX) % NY, n / NXY};
      for (int fc = 0; fc < 36; ++fc) {
        const int d = fc / 12, sp = (fc % 12) >> 2, c1 = fc & 1, c2 = (fc >> 1) & 1;
        const int d1 = d == 0 ? 1 : 0, d2 = d == 2 ? 1 : 2;
        const int64_t cm = nidx[d] - sp, t1 = nidx[d1] - c1, t2 = nidx[d2] - c2;
        if (cm < 0 || cm + 1 >= lim[d] || t1 < 0 || t1 >= lim[d1] || t2 < 0 || t2 >= lim[d2])
          continue;
        const int64_t lower = cm * scl[d] + t1 * scl[d1] + t2 * scl[d2], upper = lower + scl[d];
        const int8_t ca = cd->category[size_t(lower)], cb = cd->category[size_t(upper)];
        if (!((ca == 1 && cb != 2) || (cb == 1 && ca != 2))) continue;
        const int cn = c1 + 2 * c2;
        const int64_t gi = ca == 1 ? 24 * int64_t(cutidx[size_t(lower)]) + 4 * (2 * d + 1) + cn
                                   : 24 * int64_t(cutidx[size_t(upper)]) + 4 * (2 * d) + cn;
        gfent.push_back(uint32_t(gi) | (uint32_t(d) << 28) | (uint32_t(sp) << 30));
      }
      MF_REQUIRE(gfent.size() < size_t(INT32_MAX), MF_ERR_ARG,
                 "too many ghost-penalty face entries for the 32-bit node face lists");
      gfptr.push_back(int32_t(gfent.size()));
    }
  }
  MF_HIP_TRY(hipStreamSynchronize(sc->stream));
  if (!cd->d_nflag) MF_HIP_TRY(hipMalloc(&cd->d_nflag, size_t(m.n_dofs)));
  MF_HIP_TRY(hipMemcpy(cd->d_nflag, f.data(), size_t(m.n_dofs), hipMemcpyHostToDevice));
  if (cd->d_glist) MF_HIP_TRY(hipFree(cd->d_glist));
  cd->d_glist = nullptr;
  MF_HIP_TRY(hipMalloc(&cd->d_glist, sizeof(int64_t) * std::max<size_t>(list.size(), 1)));
  if (!list.empty())
    MF_HIP_TRY(hipMemcpy(cd->d_glist, list.data(), sizeof(int64_t) * list.size(),
                         hipMemcpyHostToDevice));
  cd->n_glist = int64_t(list.size());
  if (!cd->d_gslot) MF_HIP_TRY(hipMalloc(&cd->d_gslot, sizeof(int32_t) * gslot.size()));
  MF_HIP_TRY(hipMemcpy(cd->d_gslot, gslot.data(), sizeof(int32_t) * gslot.size(),
                       hipMemcpyHostToDevice));
  if (cd->d_gcell) MF_HIP_TRY(hipFree(cd->d_gcell));
  if (cd->d_gout) MF_HIP_TRY(hipFree(cd->d_gout));
  cd->d_gcell = nullptr;
  cd->d_gout = nullptr;
  MF_HIP_TRY(hipMalloc(&cd->d_gcell, sizeof(int64_t) * std::max<size_t>(gcell.size(), 1)));
  MF_HIP_TRY(hipMalloc(&cd->d_gout, sizeof(double) * 8 * std::max<size_t>(gcell.size(), 1)));
  if (!gcell.empty())
    MF_HIP_TRY(hipMemcpy(cd->d_gcell, gcell.data(), sizeof(int64_t) * gcell.size(),
                         hipMemcpyHostToDevice));
  cd->n_gcell = int64_t(gcell.size());
  if (cd->d_gfptr) MF_HIP_TRY(hipFree(cd->d_gfptr));
  if (cd->d_gfent) MF_HIP_TRY(hipFree(cd->d_gfent));
  cd->d_gfptr = nullptr;
  cd->d_gfent = nullptr;
  MF_HIP_TRY(hipMalloc(&cd->d_gfptr, sizeof(int32_t) * gfptr.size()));
  MF_HIP_TRY(hipMalloc(&cd->d_gfent, sizeof(uint32_t) * std::max<size_t>(gfent.size(), 1)));
  MF_HIP_TRY(hipMemcpy(cd->d_gfptr, gfptr.data(), sizeof(int32_t) * gfptr.size(),
                       hipMemcpyHostToDevice));
  if (!gfent.empty())
    MF_HIP_TRY(hipMemcpy(cd->d_gfent, gfent.data(), sizeof(uint32_t) * gfent.size(),
                         hipMemcpyHostToDevice));
  return MF_OK;
}

// blocks of the general nodes: 32 nodes per block and step, at most 1024 blocks, a multiple of 8
int general_blocks(const CutData *cd)
{
  return int((std::max<int64_t>(1, std::min<int64_t>(1024, (cd->n_glist + 31) / 32)) + 7) / 8 * 8);
}

// blocks of the node kernel: one per NODE_CHUNK nodes of a layer and NODE_LAYERS layers,
// rounded up to a multiple of the 8 XCDs
int node_blocks(const MeshInfo &m)
{
  const int64_t nxy = m.NN[0] * m.NN[1];
  const int64_t chunks = (nxy + NODE_CHUNK - 1) / NODE_CHUNK;
  const int64_t zb = (m.NN[2] + NODE_LAYERS - 1) / NODE_LAYERS;
  return int((chunks * zb + 7) / 8 * 8);
}

template <int MODE>
int launch_node(mf_operator *op, CutArgs A, const double *src, double *dst, int overwrite,
                double *dot_partial)
{
  int rc = ensure_node_flags(op);
  if (rc) return rc;
  const CutData *cd = op->cutd;
  A.nflag = cd->d_nflag;
  const int nb = node_blocks(op->s->mesh);
  A.gslot = cd->d_gslot;
  A.gout = cd->d_gout;
  A.gfptr = cd->d_gfptr;
  A.gfent = cd->d_gfent;
  // blocks of the face values and of the general cells, rounded up to multiples of 8
  auto blocks8 = [](int64_t n) { return n > 0 ? ((n + 255) / 256 + 7) / 8 * 8 : int64_t(0); };
  const int64_t fb = MODE != MODE_DIAG ? blocks8(cd->n_faces) : 0;
  const int64_t gb = MODE != MODE_DIAG ? blocks8(cd->n_gcell) : 0;
  if (fb + gb > 0) {
    cut_pre_kernel<MODE><<<dim3(unsigned(fb + gb)), dim3(256), 0, op->s->stream>>>(
        A, src, cd->d_face, cd->n_faces, fb, cd->d_gcell, cd->n_gcell);
    MF_HIP_TRY(hipGetLastError());
  }
  const int64_t genb = general_blocks(cd);
  cut_apply_kernel<MODE><<<dim3(unsigned(genb + nb)), dim3(256), 0, op->s->stream>>>(
      A, src, dst, overwrite, dot_partial, cd->d_glist, cd->n_glist, genb, nb);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

}  // namespace

bool cut_ready(const mf_operator *op) { return op->cutd && op->cutd->set; }

int cut_dot_blocks(const mf_operator *op)
{
  return node_blocks(op->s->mesh) + general_blocks(op->cutd);
}

void cut_free(mf_operator *op)
{
  if (!op->cutd) return;
  free_tables(op->cutd);
  delete op->cutd;
  op->cutd = nullptr;
}

int cut_apply(mf_operator *op, double *dst, const double *src, double *dot_partial,
              const int *stop, bool overwrite)
{
  const int64_t n = op->s->mesh.n_dofs;
  hipStream_t st = op->s->stream;
  int rc;
  CutArgs A = make_cut_args(op);
  A.stop = stop;
  // constant law: one node-centric launch
  if (!A.mushy) return launch_node<MODE_TANGENT>(op, A, src, dst, overwrite ? 1 : 0, dot_partial);
  if (overwrite && (rc = vec_fill<double>(dst, 0.0, n, st, stop))) return rc;
  if ((rc = launch_cut<MODE_TANGENT>(op, A, src, dst))) return rc;
  if (dot_partial) {
    cut_dot_kernel<<<dim3(unsigned(cut_dot_blocks(op))), dim3(256), 0, st>>>(src, dst, n,
                                                                           dot_partial, stop);
    MF_HIP_TRY(hipGetLastError());
  }
  return MF_OK;
}

int cut_diagonal(mf_operator *op, double *diag)
{
  const int64_t n = op->s->mesh.n_dofs;
  hipStream_t st = op->s->stream;
  int rc;
  CutArgs A = make_cut_args(op);
  if (!A.mushy) return launch_node<MODE_DIAG>(op, A, nullptr, diag, 1, nullptr);
  if ((rc = vec_fill<double>(diag, 0.0, n, st))) return rc;
  if ((rc = launch_cut<MODE_DIAG>(op, A, nullptr, diag))) return rc;
  // rows without a free DoF: 1 (the identity rows of the solver)
  if ((rc = vec_masked_fill(diag, op->cutd->d_inactive, 1.0, n, st))) return rc;
  if (A.dmask && (rc = vec_masked_fill(diag, A.dmask, 1.0, n, st))) return rc;
  return MF_OK;
}

int cut_residual(mf_operator *op, double *dst, const double *T_old)
{
  CutArgs A = make_cut_args(op);
  if (!A.mushy) return launch_node<MODE_RESIDUAL>(op, A, T_old, dst, 0, nullptr);
  return launch_cut<MODE_RESIDUAL>(op, A, T_old, dst);
}

}  // namespace mf

using namespace mf;

extern "C" {

int mf_heat_cut_operator_create(mf_scratch *s, const mf_heat_cut_params *p, mf_operator **out)
{
  MF_REQUIRE(s && p && out, MF_ERR_ARG, "null argument");
  MF_REQUIRE(!s->ad, MF_ERR_ARG, "the cut heat operator needs a structured mesh");
  MF_REQUIRE(s->mesh.degree == 1, MF_ERR_ARG,
             "only degree 1 is supported for now (HeatCutOperator, heat_cut_operator.cpp:498)");
  const mf_heat_params &mp = p->material;
  MF_REQUIRE(mp.law == MF_LAW_CONSTANT || mp.law == MF_LAW_MUSHY, MF_ERR_ARG,
             "the cut heat operator takes the constant liquid law or the mushy-zone law");
  if (mp.law == MF_LAW_MUSHY)
    MF_REQUIRE(mp.liquidus > mp.solidus, MF_ERR_ARG, "mushy law: liquidus must exceed solidus");
  MF_REQUIRE(p->theta >= 0.0 && p->theta <= 1.0, MF_ERR_ARG, "theta must lie in [0, 1]");
  auto *op = new mf_operator();
  op->kind = 5;
  op->s = s;
  op->cut = *p;
  op->cutd = new CutData();
  *out = op;
  return MF_OK;
}

int mf_heat_cut_set_level_set(mf_operator *op, const double *phi)
{
  MF_REQUIRE(op && phi, MF_ERR_ARG, "null argument");
  MF_REQUIRE(op->kind == 5 && op->cutd, MF_ERR_ARG, "not a cut heat operator");
  const MeshInfo &m = op->s->mesh;
  const int64_t nx = m.ncell[0], ny = m.ncell[1], nz = m.ncell[2];
  const int64_t NX = m.NN[0], NXY = m.NN[0] * m.NN[1];
  const int64_t nc = m.n_cells, nd = m.n_dofs;
  CutData tmp;
  tmp.category.assign(size_t(nc), 0);
  tmp.box.assign(size_t(nc) * 6, 0.0);
  tmp.active.assign(size_t(nd), 0);
  std::vector<int32_t> cell(static_cast<size_t>(nc));
  std::vector<double> boxes, mats;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t c = (k * ny + j) * nx + i;
        const int64_t n0 = k * NXY + j * NX + i;
        double v[8], mn = HUGE_VAL, mx = -HUGE_VAL, scale = 0.0;
        for (int e = 0; e < 8; ++e) {
          v[e] = phi[n0 + (e & 1) + ((e >> 1) & 1) * NX + (e >> 2) * NXY];
          mn = std::min(mn, v[e]);
          mx = std::max(mx, v[e]);
          scale = std::max(scale, std::fabs(v[e]));
        }
        double *b = &tmp.box[size_t(c) * 6];
        for (int d = 0; d < 3; ++d) {
          b[2 * d] = 0.0;
          b[2 * d + 1] = 1.0;
        }
        // MeshClassifier's location (positive: "outside" = liquid), zero on the positive side
        // does not cut (CutUtil::location_from_quadratures, source/cut/util.cpp:146-169)
        if (mn >= 0.0) {
          tmp.category[c] = 0;
          cell[c] = CELL_LIQUID;
          ++tmp.n_liquid;
          continue;
        }
        if (mx < 0.0) {
          tmp.category[c] = 2;
          cell[c] = CELL_GAS;
          ++tmp.n_gas;
          continue;
        }
        int axis = -1, n_var = 0;
        for (int d = 0; d < 3; ++d) {
          bool varies = false;
          for (int e = 0; e < 8; ++e)
            if (!((e >> d) & 1) && std::fabs(v[e | (1 << d)] - v[e]) > 1e-12 * scale) varies = true;
          if (varies) {
            axis = d;
            ++n_var;
          }
        }
        if (n_var != 1) {
          set_error("cut cell (" + std::to_string(i) + ", " + std::to_string(j) + ", " +
                    std::to_string(k) +
                    ") has a level set that varies along more than one axis: only axis-aligned "
                    "planar interfaces are supported");
          return MF_ERR_ARG;
        }
        const double p0 = v[0], p1 = v[1 << axis];
        const double t = p0 / (p0 - p1);
        b[2 * axis] = p0 > p1 ? 0.0 : t;
        b[2 * axis + 1] = p0 > p1 ? t : 1.0;
        tmp.category[c] = 1;
        cell[c] = int32_t(tmp.n_cut);
        boxes.insert(boxes.end(), b, b + 6);
        // the 1D matrices of the cell's liquid intervals (M_ij = h int_a^b l_i l_j, K = (b-a)/h)
        for (int d = 0; d < 3; ++d) {
          const double a = b[2 * d], e = b[2 * d + 1], h = m.h[d], oa = 1.0 - a, oe = 1.0 - e;
          mats.push_back(h * (oa * oa * oa - oe * oe * oe) / 3.0);
          mats.push_back(h * ((e * e - a * a) / 2.0 - (e * e * e - a * a * a) / 3.0));
          mats.push_back(h * (e * e * e - a * a * a) / 3.0);
          mats.push_back((e - a) / h);
        }
        ++tmp.n_cut;
      }
  MF_REQUIRE(tmp.n_cut < (int64_t(1) << 31), MF_ERR_ARG, "too many intersected cells");
  MF_REQUIRE(nd < (int64_t(1) << 32), MF_ERR_ARG,
             "the cut heat operator takes meshes of fewer than 2^32 nodes");
  // DoFs: the vertices of liquid and intersected cells; ghost-penalty faces: intersected on one
  // side, intersected or liquid on the other (CutUtil::get_face_type, source/cut/util.cpp:40-98)
  std::vector<int64_t> faces;
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t c = (k * ny + j) * nx + i;
        const int a = tmp.category[c];
        if (a == 2) continue;
        const int64_t n0 = k * NXY + j * NX + i;
        for (int e = 0; e < 8; ++e) tmp.active[size_t(n0 + (e & 1) + ((e >> 1) & 1) * NX + (e >> 2) * NXY)] = 1;
        const int64_t idx[3] = {i, j, k}, lim[3] = {nx, ny, nz}, cs[3] = {1, nx, nx * ny};
        for (int d = 0; d < 3; ++d) {
          if (idx[d] + 1 >= lim[d]) continue;
          const int bcat = tmp.category[c + cs[d]];
          if ((a == 1 && bcat != 2) || (bcat == 1 && a != 2)) faces.push_back((c << 2) | d);
        }
      }
  for (int64_t i = 0; i < nd; ++i) tmp.n_active += tmp.active[i];
  tmp.n_faces = int64_t(faces.size());
  std::vector<uint8_t> inactive(static_cast<size_t>(nd));
  for (int64_t i = 0; i < nd; ++i) inactive[i] = tmp.active[i] ? 0 : 1;
  // device tables (replacing the previous classification)
  CutData *cd = op->cutd;
  hipStream_t st = op->s->stream;
  MF_HIP_TRY(hipStreamSynchronize(st));  // work queued on the old tables has finished
  free_tables(cd);
  MF_HIP_TRY(hipMalloc(&cd->d_cell, sizeof(int32_t) * size_t(nc)));
  MF_HIP_TRY(hipMalloc(&cd->d_box, sizeof(double) * std::max<size_t>(boxes.size(), 1)));
  MF_HIP_TRY(hipMalloc(&cd->d_mats, sizeof(double) * std::max<size_t>(mats.size(), 1)));
  MF_HIP_TRY(hipMalloc(&cd->d_faceG, sizeof(double) * 24 * std::max<size_t>(size_t(tmp.n_cut), 1)));
  MF_HIP_TRY(hipMalloc(&cd->d_face, sizeof(int64_t) * std::max<size_t>(faces.size(), 1)));
  MF_HIP_TRY(hipMalloc(&cd->d_inactive, size_t(nd)));
  MF_HIP_TRY(hipMalloc(&cd->d_part, sizeof(double) * CUT_NORM_BLOCKS));
  MF_HIP_TRY(hipMemcpy(cd->d_cell, cell.data(), sizeof(int32_t) * size_t(nc), hipMemcpyHostToDevice));
  if (!boxes.empty()) {
    MF_HIP_TRY(hipMemcpy(cd->d_box, boxes.data(), sizeof(double) * boxes.size(), hipMemcpyHostToDevice));
    MF_HIP_TRY(hipMemcpy(cd->d_mats, mats.data(), sizeof(double) * mats.size(), hipMemcpyHostToDevice));
  }
  if (!faces.empty())
    MF_HIP_TRY(hipMemcpy(cd->d_face, faces.data(), sizeof(int64_t) * faces.size(), hipMemcpyHostToDevice));
  MF_HIP_TRY(hipMemcpy(cd->d_inactive, inactive.data(), size_t(nd), hipMemcpyHostToDevice));
  cd->n_liquid = tmp.n_liquid;
  cd->n_cut = tmp.n_cut;
  cd->n_gas = tmp.n_gas;
  cd->n_faces = tmp.n_faces;
  cd->n_active = tmp.n_active;
  cd->faces = faces;
  cd->category.swap(tmp.category);
  cd->box.swap(tmp.box);
  cd->active.swap(tmp.active);
  cd->set = true;
  return ensure_node_flags(op);
}

int mf_heat_cut_info(const mf_operator *op, int64_t *n_liquid, int64_t *n_intersected,
                     int64_t *n_gas, int64_t *n_faces, int64_t *n_active_dofs)
{
  MF_REQUIRE(op && op->kind == 5 && op->cutd, MF_ERR_ARG, "not a cut heat operator");
  MF_REQUIRE(op->cutd->set, MF_ERR_STATE, "level set not set (mf_heat_cut_set_level_set)");
  const CutData *cd = op->cutd;
  if (n_liquid) *n_liquid = cd->n_liquid;
  if (n_intersected) *n_intersected = cd->n_cut;
  if (n_gas) *n_gas = cd->n_gas;
  if (n_faces) *n_faces = cd->n_faces;
  if (n_active_dofs) *n_active_dofs = cd->n_active;
  return MF_OK;
}

int mf_heat_cut_cells(const mf_operator *op, int8_t *category, double *box, uint8_t *active)
{
  MF_REQUIRE(op && op->kind == 5 && op->cutd, MF_ERR_ARG, "not a cut heat operator");
  MF_REQUIRE(op->cutd->set, MF_ERR_STATE, "level set not set (mf_heat_cut_set_level_set)");
  const CutData *cd = op->cutd;
  if (category) std::copy(cd->category.begin(), cd->category.end(), category);
  if (box) std::copy(cd->box.begin(), cd->box.end(), box);
  if (active) std::copy(cd->active.begin(), cd->active.end(), active);
  return MF_OK;
}

int mf_heat_cut_norm_l2(mf_operator *op, const double *u, double *result)
{
  MF_REQUIRE(op && u && result, MF_ERR_ARG, "null argument");
  MF_REQUIRE(op->kind == 5 && op->cutd, MF_ERR_ARG, "not a cut heat operator");
  MF_REQUIRE(op->cutd->set, MF_ERR_STATE, "level set not set (mf_heat_cut_set_level_set)");
  hipStream_t st = op->s->stream;
  CutArgs A = make_cut_args(op);
  cut_norm_kernel<<<dim3(CUT_NORM_BLOCKS), dim3(256), 0, st>>>(A, u, op->s->mesh.n_cells,
                                                              op->cutd->d_part);
  MF_HIP_TRY(hipGetLastError());
  double part[CUT_NORM_BLOCKS];
  MF_HIP_TRY(hipMemcpyAsync(part, op->cutd->d_part, sizeof(part), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  double s = 0.0;
  for (double v : part) s += v;
  *result = std::sqrt(s);
  return MF_OK;
}

}  // extern "C"

// One time step of HeatCutOperation::solve (heat_cut_operation.cpp:431-482) through the host
// Newton driver: r = create_rhs(T_old) at T (the operator's T field), J d = r, T += d.
namespace {
struct CutNewton {
  mf_operator *op;
  double *T;
  const double *T_old;
  const mf_solver_desc *desc;
  double *r, *d, *nrm, *part;
  int64_t n;
};

int cut_newton_residual(void *u)
{
  auto *h = static_cast<CutNewton *>(u);
  int rc;
  if ((rc = mf_operator_set_fields(h->op, h->T, nullptr, nullptr))) return rc;
  if ((rc = vec_fill<double>(h->r, 0.0, h->n, h->op->s->stream))) return rc;
  return mf_operator_create_rhs(h->op, h->r, h->T_old, nullptr);
}

int cut_newton_solve(void *u, int *lin_its)
{
  auto *h = static_cast<CutNewton *>(u);
  int rc;
  if ((rc = vec_fill<double>(h->d, 0.0, h->n, h->op->s->stream))) return rc;
  mf_solve_report lin{};
  rc = mf_solve(h->op, h->desc, h->d, h->r, &lin);
  *lin_its = lin.iterations;
  return rc;
}

int cut_newton_norms(void *u, int need_x, double *rn, double *dn, double *xn)
{
  auto *h = static_cast<CutNewton *>(u);
  hipStream_t st = h->op->s->stream;
  int rc;
  if ((rc = vec_norm<double>(h->r, h->n, h->part, h->nrm, st))) return rc;
  if ((rc = vec_norm<double>(h->d, h->n, h->part, h->nrm + 1, st))) return rc;
  if (need_x && (rc = vec_norm<double>(h->T, h->n, h->part, h->nrm + 2, st))) return rc;
  double v[3] = {0.0, 0.0, 0.0};
  if ((rc = read_to_host(v, h->nrm, sizeof(double) * (need_x ? 3 : 2), st))) return rc;
  *rn = v[0];
  *dn = v[1];
  *xn = v[2];
  return MF_OK;
}

int cut_newton_update(void *u)
{
  auto *h = static_cast<CutNewton *>(u);
  return vec_axpby(1.0, h->d, 1.0, h->T, h->n, h->op->s->stream);
}
}  // namespace

extern "C" int mf_heat_cut_step(mf_operator *op, double *T, const double *T_old,
                                const double *dirichlet_values, const mf_solver_desc *desc,
                                const mf_newton_desc *newton, mf_newton_report *report)
{
  MF_REQUIRE(op && T && T_old && desc && newton, MF_ERR_ARG, "null argument");
  MF_REQUIRE(op->kind == 5, MF_ERR_ARG, "mf_heat_cut_step needs a cut heat operator");
  MF_REQUIRE(op->cutd && op->cutd->set, MF_ERR_STATE,
             "level set not set (mf_heat_cut_set_level_set)");
  mf_scratch *s = op->s;
  const int64_t n = s->mesh.n_dofs;
  hipStream_t st = s->stream;
  CutNewton h{op, T, T_old, desc, nullptr, nullptr, nullptr, nullptr, n};
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&h.r), sizeof(double) * n, st));
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&h.d), sizeof(double) * n, st));
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&h.nrm), sizeof(double) * 3, st));
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&h.part), sizeof(double) * 2 * RED_BLOCKS, st));
  int rc = MF_OK;
  // the Dirichlet values of the iterate (AffineConstraints::distribute)
  if (dirichlet_values && s->d_mask)
    rc = vec_masked_assign(T, s->d_mask, dirichlet_values, true, n, st);
  mf_newton_report rep{};
  if (rc == MF_OK) {
    const mf_newton_callbacks cb{cut_newton_residual, cut_newton_solve, cut_newton_norms,
                                 cut_newton_update};
    rc = mf_newton_solve(newton, &cb, &h, &rep);
  }
  (void)hipFreeAsync(h.r, st);
  (void)hipFreeAsync(h.d, st);
  (void)hipFreeAsync(h.nrm, st);
  (void)hipFreeAsync(h.part, st);
  // the operator's temperature is the final iterate
  if (rc == MF_OK || rc == MF_NO_CONVERGENCE) {
    const int rc2 = mf_operator_set_fields(op, T, nullptr, nullptr);
    if (rc == MF_OK) rc = rc2;
  }
  if (rc == MF_OK) MF_HIP_TRY(hipStreamSynchronize(st));
  rep.status = rc;
  if (report) *report = rep;
  if (rc == MF_NO_CONVERGENCE)
    set_error("Newton-Raphson solver did not converge (ExcNewtonDidNotConverge)");
  return rc;
}

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
  int64_t *d_face = nullptr;      // [n_faces] (lower cell << 2) | direction
  uint8_t *d_inactive = nullptr;  // [n_dofs] 1 = no DoF (only gas cells around the node)
  double *d_part = nullptr;       // [CUT_NORM_BLOCKS] norm partials
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
};

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

// out[ez][ey][ex] = sum_q fv phi + fx d_x phi + fy d_y phi + fz d_z phi (the transposed sweeps)
__device__ __forceinline__ void integrate_q1(const Box1D (&B)[3], const double (&fv)[2][2][2],
                                             const double (&fx)[2][2][2],
                                             const double (&fy)[2][2][2],
                                             const double (&fz)[2][2][2], double (&out)[2][2][2])
{
  double Vz[2][2][2], Xz[2][2][2], Yz[2][2][2];
#pragma unroll
  for (int ez = 0; ez < 2; ++ez)
#pragma unroll
    for (int qy = 0; qy < 2; ++qy)
#pragma unroll
      for (int qx = 0; qx < 2; ++qx) {
        Vz[ez][qy][qx] = B[2].S[0][ez] * fv[0][qy][qx] + B[2].S[1][ez] * fv[1][qy][qx] +
                         B[2].D[ez] * (fz[0][qy][qx] + fz[1][qy][qx]);
        Xz[ez][qy][qx] = B[2].S[0][ez] * fx[0][qy][qx] + B[2].S[1][ez] * fx[1][qy][qx];
        Yz[ez][qy][qx] = B[2].S[0][ez] * fy[0][qy][qx] + B[2].S[1][ez] * fy[1][qy][qx];
      }
  double Vy[2][2][2], Xy[2][2][2];
#pragma unroll
  for (int ez = 0; ez < 2; ++ez)
#pragma unroll
    for (int ey = 0; ey < 2; ++ey)
#pragma unroll
      for (int qx = 0; qx < 2; ++qx) {
        Vy[ez][ey][qx] = B[1].S[0][ey] * Vz[ez][0][qx] + B[1].S[1][ey] * Vz[ez][1][qx] +
                         B[1].D[ey] * (Yz[ez][0][qx] + Yz[ez][1][qx]);
        Xy[ez][ey][qx] = B[1].S[0][ey] * Xz[ez][0][qx] + B[1].S[1][ey] * Xz[ez][1][qx];
      }
#pragma unroll
  for (int ez = 0; ez < 2; ++ez)
#pragma unroll
    for (int ey = 0; ey < 2; ++ey)
#pragma unroll
      for (int ex = 0; ex < 2; ++ex)
        out[ez][ey][ex] = B[0].S[0][ex] * Vy[ez][ey][0] + B[0].S[1][ex] * Vy[ez][ey][1] +
                          B[0].D[ex] * (Xy[ez][ey][0] + Xy[ez][ey][1]);
}

enum { MODE_TANGENT = 0, MODE_DIAG = 1, MODE_RESIDUAL = 2 };

// The local operation of one liquid or intersected cell.  MODE_TANGENT: tangent_domain_integral
// (heat_cut_operator.cpp:157-190, no velocity) applied to u; MODE_DIAG: the diagonal of that
// element matrix; MODE_RESIDUAL: residual_domain_integral (:194-223) of T^{n+1} = t and T^n = u.
template <int MODE>
__device__ __forceinline__ void cut_cell_local(const CutArgs &A, const Box1D (&B)[3],
                                               const double (&u)[2][2][2],
                                               const double (&t)[2][2][2],
                                               double (&out)[2][2][2])
{
  QVals U, Tq;
  if (MODE != MODE_DIAG) eval_q1(B, u, U);
  if (A.mushy || MODE == MODE_RESIDUAL) eval_q1(B, t, Tq);
  if (MODE == MODE_DIAG) {
#pragma unroll
    for (int e = 0; e < 8; ++e) out[e >> 2][(e >> 1) & 1][e & 1] = 0.0;
  }
  double fv[2][2][2], fx[2][2][2], fy[2][2][2], fz[2][2][2];
#pragma unroll
  for (int qz = 0; qz < 2; ++qz)
#pragma unroll
    for (int qy = 0; qy < 2; ++qy)
#pragma unroll
      for (int qx = 0; qx < 2; ++qx) {
        const double jxw = B[0].w[qx] * B[1].w[qy] * B[2].w[qz];
        double cv = A.cv, k = A.k, dcv = 0.0, dk = 0.0;
        double tq = 0.0, gT[3] = {0.0, 0.0, 0.0};
        if (A.mushy || MODE == MODE_RESIDUAL) {
          tq = Tq.v[qz][qy][qx];
          gT[0] = Tq.gx[qz][qy];
          gT[1] = Tq.gy[qz][qx];
          gT[2] = Tq.gz[qy][qx];
          if (A.mushy) cut_law(A, tq, cv, k, dcv, dk);
        }
        if (MODE == MODE_TANGENT) {
          const double uq = U.v[qz][qy][qx];
          const double g[3] = {U.gx[qz][qy], U.gy[qz][qx], U.gz[qy][qx]};
          fv[qz][qy][qx] = (cv + dcv * tq) * uq * jxw;
          const double a = A.ost_i * k * jxw, b = A.ost_i * dk * uq * jxw;
          fx[qz][qy][qx] = a * g[0] + b * gT[0];
          fy[qz][qy][qx] = a * g[1] + b * gT[1];
          fz[qz][qy][qx] = a * g[2] + b * gT[2];
        } else if (MODE == MODE_RESIDUAL) {
          // T^n at the point and its material (the constant law or the law at T^n)
          const double to = U.v[qz][qy][qx];
          const double go[3] = {U.gx[qz][qy], U.gy[qz][qx], U.gz[qy][qx]};
          double cvo = A.cv, ko = A.k, d1, d2;
          if (A.mushy) cut_law(A, to, cvo, ko, d1, d2);
          fv[qz][qy][qx] = -(cv * tq - cvo * to) * jxw;
          fx[qz][qy][qx] = -(A.ost_i * k * gT[0] + A.ost_e * ko * go[0]) * jxw;
          fy[qz][qy][qx] = -(A.ost_i * k * gT[1] + A.ost_e * ko * go[1]) * jxw;
          fz[qz][qy][qx] = -(A.ost_i * k * gT[2] + A.ost_e * ko * go[2]) * jxw;
        } else {
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
  if (MODE != MODE_DIAG) integrate_q1(B, fv, fx, fy, fz, out);
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

// One launch per (y, z) row-parity colour; one workgroup per cell row, CUT_THREADS cells per
// step.  dst += the cell contributions of the rows of this colour (rows skipped: fixed rows).
template <int MODE>
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
  for (int x0 = 0; x0 < ncx; x0 += CUT_THREADS) {
    const int i = x0 + tid;
    double out[2][2][2];
#pragma unroll
    for (int e = 0; e < 8; ++e) out[e >> 2][(e >> 1) & 1][e & 1] = 0.0;
    if (i < ncx) {
      const int32_t cc = A.cell[(int64_t(cz) * A.ncell[1] + cy) * ncx + i];
      if (cc != CELL_GAS) {
        Box1D B[3];
        cell_boxes(A, cc, B);
        double u[2][2][2], t[2][2][2];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t n = base[r] + i;
#pragma unroll
          for (int ex = 0; ex < 2; ++ex) {
            double uv = 0.0, tv = 0.0;
            if (MODE == MODE_TANGENT) uv = fixed_row(A, n + ex) ? 0.0 : src[n + ex];
            if (MODE == MODE_RESIDUAL) uv = src[n + ex];  // T^n, read plain
            if (A.mushy || MODE == MODE_RESIDUAL) tv = A.T[n + ex];
            u[r >> 1][r & 1][ex] = uv;
            t[r >> 1][r & 1][ex] = tv;
          }
        }
        cut_cell_local<MODE>(A, B, u, t, out);
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
  const int64_t ci = cell % ncx, cj = (cell / ncx) % ncy, ck = cell / (ncx * ncy);
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
    const int64_t ci = cell % ncx, cj = (cell / ncx) % ncy, ck = cell / (ncx * ncy);
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

// partials of a . b for the fused p.(Ap) of CG (CUT_DOT_BLOCKS blocks)
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
    cut_cell_kernel<MODE><<<dim3(unsigned(rows)), dim3(CUT_THREADS), 0, st>>>(A, src, dst);
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
  if (cd->d_face) (void)hipFree(cd->d_face);
  if (cd->d_inactive) (void)hipFree(cd->d_inactive);
  if (cd->d_part) (void)hipFree(cd->d_part);
  cd->d_cell = nullptr;
  cd->d_box = nullptr;
  cd->d_face = nullptr;
  cd->d_inactive = nullptr;
  cd->d_part = nullptr;
  cd->set = false;
}

}  // namespace

bool cut_ready(const mf_operator *op) { return op->cutd && op->cutd->set; }

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
  if (overwrite && (rc = vec_fill<double>(dst, 0.0, n, st, stop))) return rc;
  CutArgs A = make_cut_args(op);
  A.stop = stop;
  if ((rc = launch_cut<MODE_TANGENT>(op, A, src, dst))) return rc;
  if (dot_partial) {
    cut_dot_kernel<<<dim3(CUT_DOT_BLOCKS), dim3(256), 0, st>>>(src, dst, n, dot_partial, stop);
    MF_HIP_TRY(hipGetLastError());
  }
  return MF_OK;
}

int cut_diagonal(mf_operator *op, double *diag)
{
  const int64_t n = op->s->mesh.n_dofs;
  hipStream_t st = op->s->stream;
  int rc;
  if ((rc = vec_fill<double>(diag, 0.0, n, st))) return rc;
  CutArgs A = make_cut_args(op);
  if ((rc = launch_cut<MODE_DIAG>(op, A, nullptr, diag))) return rc;
  // rows without a free DoF: 1 (the identity rows of the solver)
  if ((rc = vec_masked_fill(diag, op->cutd->d_inactive, 1.0, n, st))) return rc;
  if (A.dmask && (rc = vec_masked_fill(diag, A.dmask, 1.0, n, st))) return rc;
  return MF_OK;
}

int cut_residual(mf_operator *op, double *dst, const double *T_old)
{
  CutArgs A = make_cut_args(op);
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
  std::vector<double> boxes;
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
        ++tmp.n_cut;
      }
  MF_REQUIRE(tmp.n_cut < (int64_t(1) << 31), MF_ERR_ARG, "too many intersected cells");
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
  MF_HIP_TRY(hipMalloc(&cd->d_face, sizeof(int64_t) * std::max<size_t>(faces.size(), 1)));
  MF_HIP_TRY(hipMalloc(&cd->d_inactive, size_t(nd)));
  MF_HIP_TRY(hipMalloc(&cd->d_part, sizeof(double) * CUT_NORM_BLOCKS));
  MF_HIP_TRY(hipMemcpy(cd->d_cell, cell.data(), sizeof(int32_t) * size_t(nc), hipMemcpyHostToDevice));
  if (!boxes.empty())
    MF_HIP_TRY(hipMemcpy(cd->d_box, boxes.data(), sizeof(double) * boxes.size(), hipMemcpyHostToDevice));
  if (!faces.empty())
    MF_HIP_TRY(hipMemcpy(cd->d_face, faces.data(), sizeof(int64_t) * faces.size(), hipMemcpyHostToDevice));
  MF_HIP_TRY(hipMemcpy(cd->d_inactive, inactive.data(), size_t(nd), hipMemcpyHostToDevice));
  cd->n_liquid = tmp.n_liquid;
  cd->n_cut = tmp.n_cut;
  cd->n_gas = tmp.n_gas;
  cd->n_faces = tmp.n_faces;
  cd->n_active = tmp.n_active;
  cd->category.swap(tmp.category);
  cd->box.swap(tmp.box);
  cd->active.swap(tmp.active);
  cd->set = true;
  return MF_OK;
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

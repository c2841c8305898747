// Right-hand sides of the level-set normal vector and curvature projections.
//
// Reference functions replaced:
//   NormalVectorOperator::create_rhs (source/level_set/normal_vector_operator.cpp:158-237):
//     dst_d += sum_q v m_q (grad phi)_d JxW, or of grad phi / |grad phi| (normalize<dim> with
//     zero = 1e-16, include/meltpooldg/utilities/dealii_tensor.hpp:173-191)
//   CurvatureOperator::create_rhs (source/level_set/curvature_operator.cpp:185-255):
//     dst += -sum_q v m_q div(n_hat) JxW, n_hat = the nodal normal normalised with tolerance tol
//   OlssonOperator::create_rhs (source/level_set/olsson_CG_operator.cpp:197-278):
//     dst += sum_q JxW f . grad v, f = (1 - psi^2)/2 n - eps_n (grad psi . n) n
//            - eps_t (grad psi - (grad psi . n) n), n = the interpolated normal normalised with tol
// with m_q the narrow-band indicator of the level set (level_set_tools.hpp:359-368).  Sources
// are read plain (read_dof_values_plain); the result is added through the transposed
// hanging-node constraints on adaptive meshes (distribute_local_to_global).
//
// Each right-hand side is assembled once per solve.  A 256-thread block serves 256 / K^3 cells
// (32 for Q1, 9 for Q2, 4 for Q3); thread t owns local node / quadrature point
// t % K^3 = i + K j + K^2 k of cell t / K^3, the 1D sweeps run through LDS, gradients at the
// Gauss points use the collocation derivative there, and the nodal results are added with
// atomics (nodes are shared by up to 8 cells).
#include "mf_internal.hpp"

namespace mf {

struct LsRhsArgs {
  double S[KMAX * KMAX];   // S[q][j] = l_j(x_q)
  double Dq[KMAX * KMAX];  // collocation derivative on the Gauss points
  double w[KMAX];
  int ncell[3];
  uint32_t NX, NY;
  uint32_t n_dofs;
  double h[3];
  int n_cells;
  int narrow_band;
  double threshold;
  int normalized;
  double tol;                 // normalisation tolerance of the normal vector
  double eps_n, eps_t;        // reinitialization: of a level-0 cell
  int scale_h;                // reinitialization: eps scales with the cell size
  const double *normal;       // reinitialization: normal vector [3][n]
  const double *level_set;    // narrow band (nullable when narrow_band == 0)
  const uint32_t *cell_nodes; // adaptive meshes (null on structured meshes)
  const uint8_t *cell_level;
  const uint64_t *cell_hmask;
  const int32_t *node_line;
  const uint32_t *line_ptr, *line_col;
  const double *line_val;
};

constexpr int LS_THREADS = 256;

// out[t] = sum_m M[a_dir][m] in[t with a_dir = m]  (TR: M^T), a_dir the DIR index of the
// thread's local point tl (t = cell slot * K^3 + tl)
template <int K, int DIR, bool TR>
__device__ __forceinline__ double sweep1(const double *M, const double *in, int t, int tl)
{
  constexpr int ST = DIR == 0 ? 1 : (DIR == 1 ? K : K * K);
  const int a = (tl / ST) % K;
  const int base = t - a * ST;
  double acc = 0.0;
#pragma unroll
  for (int m = 0; m < K; ++m) acc += (TR ? M[m * K + a] : M[a * K + m]) * in[base + m * ST];
  return acc;
}

// NF fields of L[0..NF) through S (forward: nodes -> points) or S^T (backward) in all three
// directions, in place
template <int K, int NF, bool TR>
__device__ __forceinline__ void sweep3(const double *S, double (*L)[LS_THREADS], int t, int tl,
                                       bool act)
{
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double r[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const double *in = L[f];
      if (!act) {
        r[f] = 0.0;
      } else if (d == 0) {
        r[f] = TR ? sweep1<K, 2, true>(S, in, t, tl) : sweep1<K, 0, false>(S, in, t, tl);
      } else if (d == 1) {
        r[f] = TR ? sweep1<K, 1, true>(S, in, t, tl) : sweep1<K, 1, false>(S, in, t, tl);
      } else {
        r[f] = TR ? sweep1<K, 0, true>(S, in, t, tl) : sweep1<K, 2, false>(S, in, t, tl);
      }
    }
    __syncthreads();  // every thread has read L before it is overwritten
#pragma unroll
    for (int f = 0; f < NF; ++f) L[f][t] = r[f];
    __syncthreads();
  }
}

__device__ __forceinline__ void add_through_constraints(const LsRhsArgs &A, double *dst,
                                                        uint32_t node, bool hanging, double v)
{
  if (!hanging) {
    atomicAdd(dst + node, v);
    return;
  }
  const int32_t line = A.node_line[node];
  for (uint32_t e = A.line_ptr[line]; e < A.line_ptr[line + 1]; ++e)
    atomicAdd(dst + A.line_col[e], A.line_val[e] * v);
}

// MODE 0: normal vector rhs (src = level set, dst [3][n]); 1: curvature (src = [3][n], dst [n]);
// 2: reinitialization (src = level set, A.normal = [3][n], dst [n])
template <int K, int MODE>
__global__ void __launch_bounds__(LS_THREADS) levelset_rhs_kernel(const LsRhsArgs A,
                                                                  const double *__restrict__ src,
                                                                  double *__restrict__ dst)
{
  constexpr int NL = K * K * K;
  constexpr int CPB = LS_THREADS / NL;
  __shared__ double L[8][LS_THREADS];
  const int t = int(threadIdx.x);
  const int slot = t / NL, tl = t - slot * NL;
  const int cell = int(blockIdx.x) * CPB + slot;
  const bool act = slot < CPB && cell < A.n_cells;
  const int i = tl % K, j = (tl / K) % K, k = tl / (K * K);
  const size_t n = A.n_dofs;

  // node and geometry of the cell
  uint32_t node = 0;
  bool hanging = false;
  double hs = 1.0;
  if (A.cell_nodes) {
    if (act) {
      node = A.cell_nodes[size_t(cell) * NL + tl];
      hanging = (A.cell_hmask[cell] >> tl) & 1u;
      hs = ldexp(1.0, -int(A.cell_level[cell]));
    }
  } else if (act) {
    const int cx = cell % A.ncell[0], cy = (cell / A.ncell[0]) % A.ncell[1],
              cz = cell / (A.ncell[0] * A.ncell[1]);
    constexpr int P = K - 1;
    node = uint32_t(cx * P + i) + A.NX * (uint32_t(cy * P + j) + A.NY * uint32_t(cz * P + k));
  }
  const double h[3] = {A.h[0] * hs, A.h[1] * hs, A.h[2] * hs};
  const double jxw = act ? A.w[i] * A.w[j] * A.w[k] * h[0] * h[1] * h[2] : 0.0;

  if constexpr (MODE == 0) {
    L[0][t] = act ? src[node] : 0.0;
    __syncthreads();
    sweep3<K, 1, false>(A.S, L, t, tl, act);  // phi at the quadrature points
    double g[3] = {0, 0, 0}, pq = 0.0;
    if (act) {
      pq = L[0][t];
      g[0] = sweep1<K, 0, false>(A.Dq, L[0], t, tl) / h[0];
      g[1] = sweep1<K, 1, false>(A.Dq, L[0], t, tl) / h[1];
      g[2] = sweep1<K, 2, false>(A.Dq, L[0], t, tl) / h[2];
    }
    double msk = 1.0;
    if (A.narrow_band) msk = fabs(pq) >= A.threshold ? 0.0 : 1.0;  // phi itself gives the band
#pragma unroll
    for (int d = 0; d < 3; ++d) g[d] *= msk;
    if (A.normalized) {
      const double zero = 1e-16;
      const double nrm = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
      const double div = nrm > zero ? nrm : zero;
#pragma unroll
      for (int d = 0; d < 3; ++d) g[d] /= div;
    }
    __syncthreads();  // L[0] read by the gradient sweeps of every thread
#pragma unroll
    for (int d = 0; d < 3; ++d) L[d][t] = jxw * g[d];
    __syncthreads();
    sweep3<K, 3, true>(A.S, L, t, tl, act);
    if (act) {
#pragma unroll
      for (int d = 0; d < 3; ++d) add_through_constraints(A, dst + d * n, node, hanging, L[d][t]);
    }
  } else if constexpr (MODE == 2) {
    L[0][t] = act ? src[node] : 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) L[1 + d][t] = act ? A.normal[d * n + node] : 0.0;
    __syncthreads();
    sweep3<K, 4, false>(A.S, L, t, tl, act);  // psi and the normal vector at the quadrature points
    double f[3] = {0, 0, 0};
    if (act) {
      const double pq = L[0][t];
      const double g[3] = {sweep1<K, 0, false>(A.Dq, L[0], t, tl) / h[0],
                           sweep1<K, 1, false>(A.Dq, L[0], t, tl) / h[1],
                           sweep1<K, 2, false>(A.Dq, L[0], t, tl) / h[2]};
      double nv[3] = {L[1][t], L[2][t], L[3][t]};
      const double nrm = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
      const double den = nrm > A.tol ? nrm : A.tol;
#pragma unroll
      for (int d = 0; d < 3; ++d) nv[d] /= den;
      const double sc = A.scale_h ? hs : 1.0;
      const double en = A.eps_n * sc, et = A.eps_t * sc;
      const double gn = g[0] * nv[0] + g[1] * nv[1] + g[2] * nv[2];
      const double cf = 0.5 * (1.0 - pq * pq);  // compressive flux
#pragma unroll
      for (int d = 0; d < 3; ++d)
        f[d] = jxw * (cf * nv[d] - en * gn * nv[d] - et * (g[d] - gn * nv[d])) / h[d];
    }
    __syncthreads();  // L read by the gradient sweeps of every thread
#pragma unroll
    for (int d = 0; d < 3; ++d) L[d][t] = f[d];
    __syncthreads();
    // sum_d Dq_d^T (JxW f_d / h_d), then S^T in every direction
    double r = 0.0;
    if (act)
      r = sweep1<K, 0, true>(A.Dq, L[0], t, tl) + sweep1<K, 1, true>(A.Dq, L[1], t, tl) +
          sweep1<K, 2, true>(A.Dq, L[2], t, tl);
    __syncthreads();
    L[0][t] = r;
    __syncthreads();
    sweep3<K, 1, true>(A.S, L, t, tl, act);
    if (act) add_through_constraints(A, dst, node, hanging, L[0][t]);
  } else {
    // nodal unit normal: n / |n| where |n| > tol, else n / tol
    double v[3] = {0, 0, 0};
    if (act) {
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = src[d * n + node];
    }
    const double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double den = nrm > A.tol ? nrm : A.tol;
#pragma unroll
    for (int d = 0; d < 3; ++d) L[d][t] = v[d] / den;
    L[3][t] = (act && A.narrow_band) ? A.level_set[node] : 0.0;
    __syncthreads();
    sweep3<K, 4, false>(A.S, L, t, tl, act);
    double div = 0.0, pq = 0.0;
    if (act) {
      div = sweep1<K, 0, false>(A.Dq, L[0], t, tl) / h[0] + sweep1<K, 1, false>(A.Dq, L[1], t, tl) / h[1] +
            sweep1<K, 2, false>(A.Dq, L[2], t, tl) / h[2];
      pq = L[3][t];
    }
    const double msk = (A.narrow_band && fabs(pq) >= A.threshold) ? 0.0 : 1.0;
    __syncthreads();
    L[0][t] = -msk * div * jxw;
    __syncthreads();
    sweep3<K, 1, true>(A.S, L, t, tl, act);
    if (act) add_through_constraints(A, dst, node, hanging, L[0][t]);
  }
}

template <int K>
static int launch_k(const LsRhsArgs &a, int mode, int64_t n_cells, const double *src, double *dst,
                    hipStream_t st)
{
  constexpr int CPB = LS_THREADS / (K * K * K);
  const dim3 grid(unsigned((n_cells + CPB - 1) / CPB));
  if (mode == 0)
    hipLaunchKernelGGL((levelset_rhs_kernel<K, 0>), grid, dim3(LS_THREADS), 0, st, a, src, dst);
  else if (mode == 2)
    hipLaunchKernelGGL((levelset_rhs_kernel<K, 2>), grid, dim3(LS_THREADS), 0, st, a, src, dst);
  else
    hipLaunchKernelGGL((levelset_rhs_kernel<K, 1>), grid, dim3(LS_THREADS), 0, st, a, src, dst);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

int launch_levelset_rhs(const mf_operator *op, int mode, const double *src, double *dst)
{
  const mf_scratch *s = op->s;
  const MeshInfo &m = s->mesh;
  if (m.n_cells == 0) return MF_OK;
  LsRhsArgs a;
  std::memset(&a, 0, sizeof(a));
  const int K = m.K;
  for (int q = 0; q < K * K; ++q) {
    a.S[q] = s->tab.S[q];
    a.Dq[q] = s->tab.Dq[q];
  }
  for (int q = 0; q < K; ++q) a.w[q] = s->tab.w[q];
  for (int d = 0; d < 3; ++d) {
    a.ncell[d] = m.ncell[d];
    a.h[d] = m.h[d];
  }
  a.NX = uint32_t(m.NN[0]);
  a.NY = uint32_t(m.NN[1]);
  a.n_dofs = uint32_t(m.n_dofs);
  a.n_cells = int(m.n_cells);
  a.narrow_band = op->ls.narrow_band ? 1 : 0;
  a.threshold = op->ls.narrow_band_threshold;
  a.normalized = op->ls.compute_normalized_vector ? 1 : 0;
  a.tol = op->ls_tol;
  a.level_set = op->T;
  if (op->kind == 4) {
    reinit_epsilon(op, a.eps_n, a.eps_t, a.scale_h);
    a.normal = op->vel;
  }
  if (const AdaptiveData *ad = s->ad) {
    a.cell_nodes = ad->d.cell_nodes;
    a.cell_level = ad->d.cell_level;
    a.cell_hmask = ad->d.cell_hmask;
    a.node_line = ad->d.node_line;
    a.line_ptr = ad->d.line_ptr;
    a.line_col = ad->d.line_col;
    a.line_val = ad->d.line_val;
  }
  switch (K) {
  case 2: return launch_k<2>(a, mode, m.n_cells, src, dst, s->stream);
  case 3: return launch_k<3>(a, mode, m.n_cells, src, dst, s->stream);
  case 4: return launch_k<4>(a, mode, m.n_cells, src, dst, s->stream);
  default: set_error("degree must be 1, 2 or 3"); return MF_ERR_ARG;
  }
}

}  // namespace mf

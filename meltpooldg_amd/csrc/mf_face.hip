// Boundary faces of the heat operator on structured meshes: the Robin tangent (convection
// alpha u + radiation 4 eps sigma T^3 u), its diagonal, and the boundary part of the residual
// -R(T) (Neumann flux, radiation, convection).
//
// Reference functions replaced:
//   HeatDiffuseMultiPhaseOperator::tangent_boundary_loop + tangent_local_boundary_operation
//     (source/heat/heat_diffuse_operator.cpp:243-269, 1088-1129)
//   the boundary operation of internal_compute_diagonal_or_system_matrix (:297-391)
//   HeatDiffuseMultiPhaseOperator::rhs_boundary_loop (:719-781)
//
// One thread per boundary face of an active side; the K x K face nodes are the cell's nodes on
// that side (Gauss-Lobatto nodes include the end points), values at the K x K face quadrature
// points come from two 1D sweeps of S, and the integration applies S^T twice.  Nodes on face
// edges belong to up to 4 faces, so the results are added with atomics (a few hundred thousand
// per application when all six sides are active; nothing runs when none is).
#include "mf_internal.hpp"

namespace mf {

template <typename Number>
struct FaceArgs {
  Number S[KMAX * KMAX];
  Number w[KMAX];
  Number h[3];
  int ncell[3];
  uint32_t NX, NY;
  unsigned conv, rad, neu;
  Number alpha, Tinf_c, eps_sigma, Tinf_r;
  Number gflux[6];
  const Number *T;
  const uint8_t *mask;
  int use_mask;
  const int *stop;
  double *dot_partial;  // MODE 0: per-block src . (B src) partials (nullable)
  int off[7];  // first face of each side in the launch (prefix sums over the active sides)
};

// MODE 0: dst += B src; 1: dst += diag(B); 2: dst += boundary part of -R(T)
template <typename Number, int K, int MODE>
__global__ void __launch_bounds__(256) heat_face_kernel(const FaceArgs<Number> A, const Number *src,
                                                        Number *dst)
{
  if (A.stop && *A.stop) return;  // uniform over the grid
  const int f = int(blockIdx.x * blockDim.x + threadIdx.x);
  double ldot = 0.0;
  if (f < A.off[6]) {
  int side = 0;
  while (f >= A.off[side + 1]) ++side;
  const int lf = f - A.off[side];
  const int nd = side >> 1, upper = side & 1;
  const int t0 = nd == 0 ? 1 : 0, t1 = nd == 2 ? 1 : 2;  // tangential directions, t0 < t1
  int c[3];
  c[nd] = upper ? A.ncell[nd] - 1 : 0;
  c[t0] = lf % A.ncell[t0];
  c[t1] = lf / A.ncell[t0];
  constexpr int P = K - 1;
  const bool conv = (A.conv >> side) & 1u, rad = (A.rad >> side) & 1u, neu = (A.neu >> side) & 1u;

  uint32_t node[K][K];
  Number u[K][K], T[K][K];
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = 0; b < K; ++b) {
      int x[3];
      x[nd] = c[nd] * P + (upper ? P : 0);
      x[t0] = c[t0] * P + a;
      x[t1] = c[t1] * P + b;
      const uint32_t n = uint32_t(x[0]) + A.NX * (uint32_t(x[1]) + A.NY * uint32_t(x[2]));
      node[a][b] = n;
      const bool m = A.use_mask && A.mask && A.mask[n];
      u[a][b] = (MODE == 0 && !m) ? src[n] : Number(0);  // read_dof_values: constrained as 0
      T[a][b] = A.T ? A.T[n] : Number(0);                 // read plain
    }
  // values at the face quadrature points (qa along t0, qb along t1)
  Number uq[K][K], Tq[K][K];
#pragma unroll
  for (int qa = 0; qa < K; ++qa)
#pragma unroll
    for (int b = 0; b < K; ++b) {
      Number su = 0, st = 0;
#pragma unroll
      for (int a = 0; a < K; ++a) {
        su = fma(A.S[qa * K + a], u[a][b], su);
        st = fma(A.S[qa * K + a], T[a][b], st);
      }
      uq[qa][b] = su;
      Tq[qa][b] = st;
    }
#pragma unroll
  for (int qa = 0; qa < K; ++qa) {
    Number ru[K], rt[K];
#pragma unroll
    for (int qb = 0; qb < K; ++qb) {
      Number su = 0, st = 0;
#pragma unroll
      for (int b = 0; b < K; ++b) {
        su = fma(A.S[qb * K + b], uq[qa][b], su);
        st = fma(A.S[qb * K + b], Tq[qa][b], st);
      }
      ru[qb] = su;
      rt[qb] = st;
    }
#pragma unroll
    for (int qb = 0; qb < K; ++qb) {
      uq[qa][qb] = ru[qb];
      Tq[qa][qb] = rt[qb];
    }
  }
  // quadrature-point terms (JxW of the face: w_qa w_qb h_t0 h_t1)
  Number vq[K][K];
#pragma unroll
  for (int qa = 0; qa < K; ++qa)
#pragma unroll
    for (int qb = 0; qb < K; ++qb) {
      const Number wq = A.w[qa] * A.h[t0] * A.w[qb] * A.h[t1];
      const Number t = Tq[qa][qb];
      const Number coef = (conv ? A.alpha : Number(0)) +
                          (rad ? Number(4) * A.eps_sigma * t * t * t : Number(0));
      if constexpr (MODE == 0) {
        vq[qa][qb] = wq * coef * uq[qa][qb];
      } else if constexpr (MODE == 1) {
        vq[qa][qb] = wq * coef;
      } else {
        Number val = 0;
        if (neu) val -= A.gflux[side];
        if (rad) val -= A.eps_sigma * (t * t * t * t - A.Tinf_r * A.Tinf_r * A.Tinf_r * A.Tinf_r);
        if (conv) val -= A.alpha * (t - A.Tinf_c);
        vq[qa][qb] = wq * val;
      }
    }
  // integrate: out[a][b] = sum_q S[qa][a] S[qb][b] vq (diagonal: with the squares)
  Number tq[K][K];
#pragma unroll
  for (int qa = 0; qa < K; ++qa)
#pragma unroll
    for (int b = 0; b < K; ++b) {
      Number s = 0;
#pragma unroll
      for (int qb = 0; qb < K; ++qb) {
        const Number sb = A.S[qb * K + b];
        s = fma(MODE == 1 ? sb * sb : sb, vq[qa][qb], s);
      }
      tq[qa][b] = s;
    }
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = 0; b < K; ++b) {
      Number s = 0;
#pragma unroll
      for (int qa = 0; qa < K; ++qa) {
        const Number sa = A.S[qa * K + a];
        s = fma(MODE == 1 ? sa * sa : sa, tq[qa][b], s);
      }
      if constexpr (MODE == 0) ldot += double(u[a][b] * s);  // constrained u entries are 0
      const uint32_t n = node[a][b];
      if (A.use_mask && A.mask && A.mask[n]) continue;  // constrained rows are skipped
      atomicAdd(dst + n, s);
    }
  }  // f < faces
  if constexpr (MODE == 0) {
    if (A.dot_partial) {
      __shared__ double red[4];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) ldot += __shfl_xor(ldot, o, 64);
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ldot;
      __syncthreads();
      if (threadIdx.x == 0) A.dot_partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
    }
  }
}

template <typename Number, int K, int MODE>
static void launch_faces_k(const FaceArgs<Number> &a, const Number *src, Number *dst,
                           hipStream_t st)
{
  const int nf = a.off[6];
  hipLaunchKernelGGL((heat_face_kernel<Number, K, MODE>), dim3((nf + 255) / 256), dim3(256), 0,
                     st, a, src, dst);
}

// faces of the sides active for the tangent (0 when none)
static int tangent_faces(const mf_operator *op)
{
  if (!op->bnd_set || op->kind != 0 || op->s->mesh.ad) return 0;
  const unsigned sides = op->bnd.convection_sides | op->bnd.radiation_sides;
  const MeshInfo &m = op->s->mesh;
  int nf = 0;
  for (int side = 0; side < 6; ++side) {
    const int nd = side >> 1;
    const int t0 = nd == 0 ? 1 : 0, t1 = nd == 2 ? 1 : 2;
    if ((sides >> side) & 1u) nf += m.ncell[t0] * m.ncell[t1];
  }
  return nf;
}

int heat_face_blocks(const mf_operator *op) { return (tangent_faces(op) + 255) / 256; }

template <typename Number>
int launch_heat_faces(const mf_operator *op, const Number *src, Number *dst, int mode,
                      const int *stop, double *dot_partial)
{
  if (!op->bnd_set || op->kind != 0) return MF_OK;
  const mf_scratch *s = op->s;
  const MeshInfo &m = s->mesh;
  MF_REQUIRE(!m.ad, MF_ERR_ARG, "boundary faces are not available on adaptive meshes");
  const mf_heat_boundary &b = op->bnd;
  const unsigned tangent_sides = b.convection_sides | b.radiation_sides;
  const unsigned sides = mode == 2 ? (tangent_sides | b.neumann_sides) : tangent_sides;
  if (!sides) return MF_OK;
  if (b.radiation_sides) MF_REQUIRE(op->T, MF_ERR_STATE, "radiation boundary needs T (mf_operator_set_fields)");
  FaceArgs<Number> a;
  std::memset(&a, 0, sizeof(a));
  const int K = m.K;
  for (int i = 0; i < K * K; ++i) a.S[i] = Number(s->tab.S[i]);
  for (int i = 0; i < K; ++i) a.w[i] = Number(s->tab.w[i]);
  for (int d = 0; d < 3; ++d) {
    a.h[d] = Number(m.h[d]);
    a.ncell[d] = m.ncell[d];
  }
  a.NX = uint32_t(m.NN[0]);
  a.NY = uint32_t(m.NN[1]);
  a.conv = b.convection_sides;
  a.rad = b.radiation_sides;
  a.neu = mode == 2 ? b.neumann_sides : 0u;
  a.alpha = Number(b.convection_coefficient);
  a.Tinf_c = Number(b.convection_temperature_infinity);
  a.eps_sigma = Number(b.emissivity * MF_STEFAN_BOLTZMANN);
  a.Tinf_r = Number(b.radiation_temperature_infinity);
  for (int i = 0; i < 6; ++i) a.gflux[i] = Number(b.neumann_flux[i]);
  if constexpr (sizeof(Number) == 8)
    a.T = reinterpret_cast<const Number *>(op->T);
  else
    a.T = reinterpret_cast<const Number *>(op->T ? op->T32 : nullptr);
  a.mask = s->d_mask;
  a.use_mask = op->dof_index == 0 ? 1 : 0;
  a.stop = stop;
  a.dot_partial = mode == 0 ? dot_partial : nullptr;
  a.off[0] = 0;
  for (int side = 0; side < 6; ++side) {
    const int nd = side >> 1;
    const int t0 = nd == 0 ? 1 : 0, t1 = nd == 2 ? 1 : 2;
    a.off[side + 1] = a.off[side] + (((sides >> side) & 1u) ? m.ncell[t0] * m.ncell[t1] : 0);
  }
  hipStream_t st = s->stream;
  switch (K * 4 + mode) {
  case 2 * 4 + 0: launch_faces_k<Number, 2, 0>(a, src, dst, st); break;
  case 2 * 4 + 1: launch_faces_k<Number, 2, 1>(a, src, dst, st); break;
  case 2 * 4 + 2: launch_faces_k<Number, 2, 2>(a, src, dst, st); break;
  case 3 * 4 + 0: launch_faces_k<Number, 3, 0>(a, src, dst, st); break;
  case 3 * 4 + 1: launch_faces_k<Number, 3, 1>(a, src, dst, st); break;
  case 3 * 4 + 2: launch_faces_k<Number, 3, 2>(a, src, dst, st); break;
  case 4 * 4 + 0: launch_faces_k<Number, 4, 0>(a, src, dst, st); break;
  case 4 * 4 + 1: launch_faces_k<Number, 4, 1>(a, src, dst, st); break;
  case 4 * 4 + 2: launch_faces_k<Number, 4, 2>(a, src, dst, st); break;
  default: set_error("boundary faces: degree must be 1, 2 or 3"); return MF_ERR_ARG;
  }
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

template int launch_heat_faces<double>(const mf_operator *, const double *, double *, int,
                                       const int *, double *);
template int launch_heat_faces<float>(const mf_operator *, const float *, float *, int,
                                      const int *, double *);

}  // namespace mf

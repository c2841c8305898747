// Tangent (vmult) kernel with slice-resident sum factorisation: the hot path of the heat
// Jacobian and the advection-diffusion operator.
//
// Reference: HeatDiffuseMultiPhaseOperator::tangent_cell_loop / tangent_local_cell_operation
// (source/heat/heat_diffuse_operator.cpp:199-241, 955-1086) and
// AdvectionDiffusionOperator::vmult / tangent_local_cell_operation
// (source/level_set/advection_diffusion_operator.cpp:236-272, 448-495).
//
// Decomposition.  K threads serve one cell; a block holds CPB cells, thread (c, t) with the cell
// index c fastest in threadIdx.x.  Two register layouts of a cell tensor X(i, j, k):
//   XY layout: thread t owns the z-slice k = t, X(:, :, t) as a K x K register array [j][i];
//              the x and y sweeps run in registers.
//   ZY layout: thread t owns the x-slice i = t, X(t, :, :) as [j][k]; the z sweeps and the
//              y / z collocation derivatives run in registers.
// One LDS transpose moves data XY -> ZY for evaluation and one ZY -> XY for integration:
//   evaluate:  a = S_y u;  A = S_x a, B = DS_x a   (XY)  -> LDS ->
//              u_q = S_z A, d_x u = S_z B, d_y u = Dq_y u_q, d_z u = Dq_z u_q   (ZY)
//   integrate: P = S_z^T (v + Dq_y^T f_y + Dq_z^T f_z), Q = S_z^T f_x   (ZY) -> LDS ->
//              out = S_y^T (S_x^T P + DS_x^T Q)   (XY)
// so each field crosses LDS twice per application (8 B written and read per entry and
// direction) instead of once per 1D sweep.  d_x uses the nodal derivative matrix DS (the x
// index is not resident in the ZY layout); d_y, d_z use the collocation derivative Dq on the
// Gauss points (Dq S = DS exactly for K points).
//
// Temperature-dependent conductivity (tanh law): T is gathered and interpolated for every cell
// and k(T_q) evaluated at every quadrature point.  The dk/dT grad T u term needs grad T only
// where sech^2 is non-zero; tanh_sech2 returns exactly zero beyond |x| = 20, so when every
// quadrature point of the block is beyond it the gradient of T is not formed (the term it
// multiplies is exactly zero); otherwise it is computed with one more transpose.
#include "mf_internal.hpp"

namespace mf {

template <int K>
constexpr int slab_cells_per_block()
{
  return K == 4 ? 32 : 64;
}

// out[r][a] = sum_b M[a][b] in[r][b]   (apply along the fast index)
template <typename Number, int K, bool TR>
__device__ __forceinline__ void apply_fast(const Number *M, const Number (&in)[K][K],
                                           Number (&out)[K][K])
{
#pragma unroll
  for (int r = 0; r < K; ++r)
#pragma unroll
    for (int a = 0; a < K; ++a) {
      Number s = 0;
#pragma unroll
      for (int b = 0; b < K; ++b) s = fma(TR ? M[b * K + a] : M[a * K + b], in[r][b], s);
      out[r][a] = s;
    }
}

// out[a][r] = sum_b M[a][b] in[b][r]   (apply along the slow index)
template <typename Number, int K, bool TR>
__device__ __forceinline__ void apply_slow(const Number *M, const Number (&in)[K][K],
                                           Number (&out)[K][K])
{
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int r = 0; r < K; ++r) {
      Number s = 0;
#pragma unroll
      for (int b = 0; b < K; ++b) s = fma(TR ? M[b * K + a] : M[a * K + b], in[b][r], s);
      out[a][r] = s;
    }
}

// LDS element (i, j, k) of cell c
template <int K, int CPB>
__device__ __forceinline__ int lidx(int i, int j, int k, int c)
{
  return (i + K * j + K * K * k) * CPB + c;
}

template <typename Number>
__device__ __forceinline__ void tanh_sech2_s(Number x, Number &t, Number &s2)
{
  const Number ax = fabs(x);
  const Number cut = sizeof(Number) == 8 ? Number(20) : Number(10);
  if (ax < cut) {
    const Number em = expm1(Number(-2) * ax);
    const Number inv = Number(1) / (Number(2) + em);
    t = -em * inv;
    s2 = Number(4) * (Number(1) + em) * inv * inv;
    if (x < 0) t = -t;
  } else {
    t = x < 0 ? Number(-1) : Number(1);
    s2 = 0;
  }
}

__device__ __forceinline__ void atomic_add_s(double *a, double v) { unsafeAtomicAdd(a, v); }
__device__ __forceinline__ void atomic_add_s(float *a, float v) { unsafeAtomicAdd(a, v); }

// XY-layout gather of the z-slice k = t of a nodal field, then a = S_y v, A = S_x a (and
// optionally B = DS_x a) into LDS
template <typename Number, int K, int CPB, bool WITH_B>
__device__ __forceinline__ void xy_interp_to_lds(const CellArgs<Number> &A, const Number (&v)[K][K],
                                                 Number *bufA, Number *bufB, int t, int c)
{
  Number a[K][K], o[K][K];
  apply_slow<Number, K, false>(A.S, v, a);
  apply_fast<Number, K, false>(A.S, a, o);
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < K; ++i) bufA[lidx<K, CPB>(i, j, t, c)] = o[j][i];
  if constexpr (WITH_B) {
    apply_fast<Number, K, false>(A.DS, a, o);
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
      for (int i = 0; i < K; ++i) bufB[lidx<K, CPB>(i, j, t, c)] = o[j][i];
  }
}

// ZY-layout read of x-slice i = t and z interpolation: q[j][k] = sum_m S[k][m] X(t, j, m)
template <typename Number, int K, int CPB>
__device__ __forceinline__ void zy_interp_from_lds(const Number *S, const Number *buf,
                                                   Number (&q)[K][K], int t, int c)
{
  Number x[K][K];
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int k = 0; k < K; ++k) x[j][k] = buf[lidx<K, CPB>(t, j, k, c)];
  apply_fast<Number, K, false>(S, x, q);
}

// XY-layout gather of the z-slice k = t of a nodal field: v[j][i] = f(cx*P+i, cy*P+j, cz*P+t)
template <typename Number, int K>
__device__ __forceinline__ void gather_slice(const Number *f, uint32_t zbase, uint32_t NX,
                                             bool active, Number (&v)[K][K])
{
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < K; ++i) v[j][i] = active ? f[zbase + j * NX + i] : Number(0);
}

// XY phase that produces d_x of a field: a = S_y v, bufB <- DS_x a
template <typename Number, int K, int CPB>
__device__ __forceinline__ void xy_dx_to_lds(const CellArgs<Number> &A, const Number (&v)[K][K],
                                             Number *bufB, int t, int c)
{
  Number a[K][K];
  apply_slow<Number, K, false>(A.S, v, a);
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < K; ++i) {
      Number s = 0;
#pragma unroll
      for (int m = 0; m < K; ++m) s = fma(A.DS[i * K + m], a[j][m], s);
      bufB[lidx<K, CPB>(i, j, t, c)] = s;
    }
}

template <typename Number, int K, int OP, bool NEEDT, bool VEL>
__global__ void __launch_bounds__(slab_cells_per_block<K>() * K)
    slab_tangent_kernel(const CellArgs<Number> A, const Number *__restrict__ src,
                        Number *__restrict__ dst)
{
  constexpr int CPB = slab_cells_per_block<K>();
  constexpr int NL = K * K * K;
  constexpr int P = K - 1;
  constexpr bool HEAT = (OP == HEAT_TANGENT);
  __shared__ Number bufA[NL * CPB];
  __shared__ Number bufB[NL * CPB];
  __shared__ double red[8];

  if (A.stop && *A.stop) return;  // uniform over the grid

  const int c = threadIdx.x % CPB;
  const int t = threadIdx.x / CPB;
  const int cell = blockIdx.x * CPB + c;
  const bool active = cell < A.n_cells;
  const int cx = cell % A.ncell[0];
  const int cy = (cell / A.ncell[0]) % A.ncell[1];
  const int cz = cell / (A.ncell[0] * A.ncell[1]);
  // first node of the z-slice k = t of this cell; row j at + j*NX
  const uint32_t zbase =
      uint32_t(cx * P) + A.NX * (uint32_t(cy * P) + A.NY * uint32_t(cz * P + t));
  const uint32_t NX = A.NX;
  const Number kc = (A.k_cell && active) ? A.k_cell[cell] : Number(1);
  // h0 h1 h2 w_t: the part of JxW at q = (t, j, k) that does not depend on j, k
  const Number wt = A.w[t] * A.h[0] * A.h[1] * A.h[2];

  // ---------------- u: gather with constraints, evaluate (XY -> LDS -> ZY) ----------------
  {
    Number v[K][K];
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
      for (int i = 0; i < K; ++i) {
        Number x = 0;
        if (active) {
          const uint32_t gi = zbase + j * NX + i;
          x = src[gi];
          if (A.use_mask && !A.gather_plain && A.mask && A.mask[gi]) x = 0;
        }
        v[j][i] = x;
      }
    xy_interp_to_lds<Number, K, CPB, true>(A, v, bufA, bufB, t, c);
  }
  __syncthreads();
  Number uq[K][K], gx[K][K], gy[K][K], gz[K][K];
  zy_interp_from_lds<Number, K, CPB>(A.S, bufA, uq, t, c);
  zy_interp_from_lds<Number, K, CPB>(A.S, bufB, gx, t, c);
  apply_slow<Number, K, false>(A.Dq, uq, gy);
  apply_fast<Number, K, false>(A.Dq, uq, gz);
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      gx[j][k] *= A.inv_h[0];
      gy[j][k] *= A.inv_h[1];
      gz[j][k] *= A.inv_h[2];
    }

  // ---------------- velocity: conv = w . grad u, one component at a time ----------------
  Number conv[K][K];
  if constexpr (VEL) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      __syncthreads();  // previous readers of bufA are done
      {
        Number v[K][K];
        gather_slice<Number, K>(A.vel + size_t(d) * A.n_dofs, zbase, NX, active, v);
        xy_interp_to_lds<Number, K, CPB, false>(A, v, bufA, bufB, t, c);
      }
      __syncthreads();
      Number wd[K][K];
      zy_interp_from_lds<Number, K, CPB>(A.S, bufA, wd, t, c);
#pragma unroll
      for (int j = 0; j < K; ++j)
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const Number g = d == 0 ? gx[j][k] : (d == 1 ? gy[j][k] : gz[j][k]);
          conv[j][k] = d == 0 ? wd[j][k] * g : fma(wd[j][k], g, conv[j][k]);
        }
    }
  }

  // ---------------- fluxes: f = k(T) kc grad u [+ kc (dk/dT) grad T u] ----------------
  if constexpr (HEAT && NEEDT) {
    __syncthreads();
    {
      Number v[K][K];
      gather_slice<Number, K>(A.T, zbase, NX, active, v);
      xy_interp_to_lds<Number, K, CPB, false>(A, v, bufA, bufB, t, c);
    }
    __syncthreads();
    Number Tq[K][K];
    zy_interp_from_lds<Number, K, CPB>(A.S, bufA, Tq, t, c);
    int nonzero = 0;
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        Number th, s2;
        tanh_sech2_s((Tq[j][k] - A.tanh_Tc) * A.inv_tanh_w, th, s2);
        nonzero |= (s2 != Number(0));
      }
    const bool has_dk = __syncthreads_or(A.t_dependent && nonzero) != 0;
    if (has_dk) {
      // d_x T through a second transpose; d_y T, d_z T by collocation; one component at a time
      {
        Number v[K][K];
        gather_slice<Number, K>(A.T, zbase, NX, active, v);
        xy_dx_to_lds<Number, K, CPB>(A, v, bufB, t, c);
      }
      __syncthreads();
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        Number gT[K][K];
        if (d == 0) {
          zy_interp_from_lds<Number, K, CPB>(A.S, bufB, gT, t, c);
        } else if (d == 1) {
          apply_slow<Number, K, false>(A.Dq, Tq, gT);
        } else {
          apply_fast<Number, K, false>(A.Dq, Tq, gT);
        }
#pragma unroll
        for (int j = 0; j < K; ++j)
#pragma unroll
          for (int k = 0; k < K; ++k) {
            Number th, s2;
            tanh_sech2_s((Tq[j][k] - A.tanh_Tc) * A.inv_tanh_w, th, s2);
            const Number keff = (A.k0 + A.tanh_amp * th) * kc;
            const Number dku = A.tanh_amp * A.inv_tanh_w * s2 * kc * uq[j][k];
            Number &g = d == 0 ? gx[j][k] : (d == 1 ? gy[j][k] : gz[j][k]);
            g = keff * g + dku * (gT[j][k] * A.inv_h[d]);
          }
      }
    } else {
#pragma unroll
      for (int j = 0; j < K; ++j)
#pragma unroll
        for (int k = 0; k < K; ++k) {
          Number th, s2;
          tanh_sech2_s((Tq[j][k] - A.tanh_Tc) * A.inv_tanh_w, th, s2);
          const Number keff = (A.k0 + A.tanh_amp * th) * kc;
          gx[j][k] *= keff;
          gy[j][k] *= keff;
          gz[j][k] *= keff;
        }
    }
  } else {
    const Number keff = HEAT ? A.k0 * kc : A.theta * A.diffusivity;
#pragma unroll
    for (int j = 0; j < K; ++j)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        gx[j][k] *= keff;
        gy[j][k] *= keff;
        gz[j][k] *= keff;
      }
  }

  // ---------------- values and quadrature weights (in place) ----------------
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const Number jxw = wt * A.w[j] * A.w[k];
      Number val;
      if constexpr (HEAT) {
        val = A.dt_inv * A.rho_cp * uq[j][k];
        if constexpr (VEL) val += A.rho_cp * conv[j][k];
      } else {
        val = A.mass_factor * A.dt_inv * uq[j][k] + A.theta * conv[j][k];
      }
      uq[j][k] = jxw * val;
      gx[j][k] *= jxw * A.inv_h[0];
      gy[j][k] *= jxw * A.inv_h[1];
      gz[j][k] *= jxw * A.inv_h[2];
    }

  // ---------------- integrate: ZY part (P = S_z^T (v + Dq_y^T f_y + Dq_z^T f_z), Q = S_z^T f_x)
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int r = 0; r < K; ++r) {
      // uq[a][r] is only read here, gy and gz are not written: in place into uq is safe
      Number s = uq[a][r];
#pragma unroll
      for (int b = 0; b < K; ++b) s = fma(A.Dq[b * K + a], gy[b][r], s);
#pragma unroll
      for (int b = 0; b < K; ++b) s = fma(A.Dq[b * K + r], gz[a][b], s);
      uq[a][r] = s;
    }
  __syncthreads();  // every thread has finished reading bufA / bufB
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int k = 0; k < K; ++k) {
      Number p = 0, q = 0;
#pragma unroll
      for (int m = 0; m < K; ++m) {
        p = fma(A.S[m * K + k], uq[j][m], p);
        q = fma(A.S[m * K + k], gx[j][m], q);
      }
      bufA[lidx<K, CPB>(t, j, k, c)] = p;
      bufB[lidx<K, CPB>(t, j, k, c)] = q;
    }
  __syncthreads();

  // ---------------- integrate: XY part (out = S_y^T (S_x^T P + DS_x^T Q)) and scatter -------
  Number r1[K][K];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    Number pp[K], qq[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      pp[i] = bufA[lidx<K, CPB>(i, j, t, c)];
      qq[i] = bufB[lidx<K, CPB>(i, j, t, c)];
    }
#pragma unroll
    for (int a = 0; a < K; ++a) {
      Number s = 0;
#pragma unroll
      for (int i = 0; i < K; ++i) s = fma(A.S[i * K + a], pp[i], fma(A.DS[i * K + a], qq[i], s));
      r1[j][a] = s;
    }
  }
  double local_dot = 0;
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int i = 0; i < K; ++i) {
      Number o = 0;
#pragma unroll
      for (int b = 0; b < K; ++b) o = fma(A.S[b * K + j], r1[b][i], o);
      if (active) {
        const uint32_t gi = zbase + j * NX + i;
        if (!(A.use_mask && A.mask && A.mask[gi])) {
          atomic_add_s(dst + gi, o);
          if (A.dot_partial) local_dot += double(src[gi]) * double(o);
        }
      }
    }
  if (A.dot_partial) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) local_dot += __shfl_xor(local_dot, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = local_dot;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0;
      for (int w = 0; w < int(blockDim.x >> 6); ++w) s += red[w];
      A.dot_partial[blockIdx.x] = s;
    }
  }
}

template <typename Number, int K, int OP, bool NEEDT, bool VEL>
static int launch_slab_k(const MeshInfo &m, const CellArgs<Number> &a, const Number *src,
                         Number *dst, hipStream_t st)
{
  constexpr int CPB = slab_cells_per_block<K>();
  const int blocks = int((m.n_cells + CPB - 1) / CPB);
  hipLaunchKernelGGL((slab_tangent_kernel<Number, K, OP, NEEDT, VEL>), dim3(blocks),
                     dim3(CPB * K), 0, st, a, src, dst);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

template <typename Number, int K>
static int launch_slab_deg(int op, bool needT, bool vel, const MeshInfo &m,
                           const CellArgs<Number> &a, const Number *src, Number *dst,
                           hipStream_t st)
{
  if (op == ADVEC_TANGENT) return launch_slab_k<Number, K, ADVEC_TANGENT, false, true>(m, a, src, dst, st);
  if (needT) {
    if (vel) return launch_slab_k<Number, K, HEAT_TANGENT, true, true>(m, a, src, dst, st);
    return launch_slab_k<Number, K, HEAT_TANGENT, true, false>(m, a, src, dst, st);
  }
  if (vel) return launch_slab_k<Number, K, HEAT_TANGENT, false, true>(m, a, src, dst, st);
  return launch_slab_k<Number, K, HEAT_TANGENT, false, false>(m, a, src, dst, st);
}

int slab_blocks(const MeshInfo &m)
{
  const int cpb = m.K == 4 ? slab_cells_per_block<4>() : slab_cells_per_block<2>();
  return int((m.n_cells + cpb - 1) / cpb);
}

template <typename Number>
int launch_slab_tangent(int op, bool needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                        const Number *src, Number *dst, hipStream_t st)
{
  switch (m.K) {
  case 2: return launch_slab_deg<Number, 2>(op, needT, vel, m, a, src, dst, st);
  case 3: return launch_slab_deg<Number, 3>(op, needT, vel, m, a, src, dst, st);
  case 4: return launch_slab_deg<Number, 4>(op, needT, vel, m, a, src, dst, st);
  default: set_error("degree must be 1, 2 or 3"); return MF_ERR_ARG;
  }
}

template int launch_slab_tangent<double>(int, bool, bool, const MeshInfo &,
                                         const CellArgs<double> &, const double *, double *,
                                         hipStream_t);
template int launch_slab_tangent<float>(int, bool, bool, const MeshInfo &, const CellArgs<float> &,
                                        const float *, float *, hipStream_t);

}  // namespace mf

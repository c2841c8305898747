// Device kernels and C-ABI entry points for the callers on either side of the operator:
// one advection-diffusion time step with Dirichlet lifting, one Newton-solved implicit heat step,
// the linear-extrapolation predictor, the volumetric Gaussian laser source and the FE L2 norm.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mf_internal.hpp"
#include "mf_vec.hpp"

namespace mf {

// ------------------------------------------------------------------------------------------
// small vector kernels
// ------------------------------------------------------------------------------------------
// y = a x + b y
__global__ void axpby_kernel(double a, const double *x, double b, double *y, int64_t n)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    y[i] = a * x[i] + b * y[i];
}

// dst[i] = mask[i] ? (vals ? vals[i] : 0) : keep (when keep_free) or 0
__global__ void masked_assign_kernel(double *dst, const uint8_t *mask, const double *vals,
                                     int keep_free, int64_t n)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (mask && mask[i])
      dst[i] = vals ? vals[i] : 0.0;
    else if (!keep_free)
      dst[i] = 0.0;
  }
}

static dim3 vgrid(int64_t n)
{
  int64_t b = (n + VEC_THREADS - 1) / VEC_THREADS;
  return dim3(unsigned(std::max<int64_t>(1, std::min<int64_t>(b, RED_BLOCKS))));
}

int vec_axpby(double a, const double *x, double b, double *y, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(axpby_kernel, vgrid(n), dim3(VEC_THREADS), 0, st, a, x, b, y, n);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

int vec_masked_assign(double *dst, const uint8_t *mask, const double *vals, bool keep_free,
                      int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(masked_assign_kernel, vgrid(n), dim3(VEC_THREADS), 0, st, dst, mask, vals,
                     keep_free ? 1 : 0, n);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

__global__ void masked_fill_kernel(double *dst, const uint8_t *mask, double value, int64_t n)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    if (mask[i]) dst[i] = value;
}

int vec_masked_fill(double *dst, const uint8_t *mask, double value, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(masked_fill_kernel, vgrid(n), dim3(VEC_THREADS), 0, st, dst, mask, value, n);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

// do_post_vmult (advection_diffusion_operator.cpp:561-578): dst = src on the inflow/outflow
// rows; with `part`, each of the IO_BLOCKS blocks writes its part of sum src_i^2 over them (the
// identity rows' share of a fused src . (A src))
template <typename Number>
__global__ void io_rows_kernel(const uint8_t *__restrict__ io, const Number *__restrict__ src,
                               Number *__restrict__ dst, int64_t n, double *part, const int *stop)
{
  if (stop && *stop) return;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (io[i]) {
      const Number v = src[i];
      dst[i] = v;
      acc += double(v) * double(v);
    }
  }
  if (!part) return;
  __shared__ double red[VEC_THREADS / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < VEC_THREADS / 64; ++w) s += red[w];
    part[blockIdx.x] = s;
  }
}

template <typename Number>
int launch_io_rows(const uint8_t *io_mask, const Number *src, Number *dst, int64_t n,
                   double *dot_partial, const int *stop, hipStream_t st)
{
  hipLaunchKernelGGL(io_rows_kernel<Number>, dim3(IO_BLOCKS), dim3(VEC_THREADS), 0, st, io_mask,
                     src, dst, n, dot_partial, stop);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

template int launch_io_rows<double>(const uint8_t *, const double *, double *, int64_t, double *,
                                    const int *, hipStream_t);
template int launch_io_rows<float>(const uint8_t *, const float *, float *, int64_t, double *,
                                   const int *, hipStream_t);

// the inflow DoFs of create_inflow_outflow_constraints (advection_diffusion_operation.cpp:
// 368-499) on a box: a node on side s (direction d = s/2, upper = s%2) with the outer normal
// n = (2 upper - 1) e_d is an inflow DoF when n . u <= 0 for its nodal velocity u
__global__ void inflow_mask_kernel(const double *__restrict__ vel, unsigned sides, int64_t NX,
                                   int64_t NY, int64_t NZ, uint8_t *__restrict__ mask)
{
  const int64_t n = NX * NY * NZ;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t c[3] = {i % NX, (i / NX) % NY, i / (NX * NY)};
    const int64_t NN[3] = {NX, NY, NZ};
    uint8_t in = 0;
    for (int d = 0; d < 3; ++d) {
      const double u = vel[int64_t(d) * n + i];
      if (((sides >> (2 * d)) & 1u) && c[d] == 0 && -u <= 0.0) in = 1;
      if (((sides >> (2 * d + 1)) & 1u) && c[d] == NN[d] - 1 && u <= 0.0) in = 1;
    }
    mask[i] = in;
  }
}

int launch_inflow_mask(const ::mf_scratch *s, const double *vel, unsigned sides, uint8_t *mask)
{
  const MeshInfo &m = s->mesh;
  hipLaunchKernelGGL(inflow_mask_kernel, vgrid(m.n_dofs), dim3(VEC_THREADS), 0, s->stream, vel,
                     sides, m.NN[0], m.NN[1], m.NN[2], mask);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

// ------------------------------------------------------------------------------------------
// laser heat source: nodal interpolation of P/(r sqrt(pi/2))^3 exp(-2 (d/r)^2)
// (GaussVolumetricIntensityProfile, include/meltpooldg/heat/laser_intensity_profiles.hpp:175-218;
//  LaserHeatSourceVolumetric::compute_volumetric_heat_source,
//  source/heat/laser_heat_source_volumetric.cpp:20-58)
// ------------------------------------------------------------------------------------------
struct NodeGeo {
  double lo[3], h[3];
  int64_t NN[3];
  int ncell[3];
  int p;
  double gl[KMAX];
};

__device__ __forceinline__ double node_coord(const NodeGeo &g, int d, int64_t i)
{
  int64_t c = i / g.p;
  if (c >= g.ncell[d]) c = g.ncell[d] - 1;
  return g.lo[d] + g.h[d] * (double(c) + g.gl[i - c * g.p]);
}

__global__ void laser_kernel(NodeGeo g, double peak, double inv_r2, double x0, double y0,
                             double z0, double *q, int64_t n, double radius, int d_dr)
{
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t ix = idx % g.NN[0], iy = (idx / g.NN[0]) % g.NN[1], iz = idx / (g.NN[0] * g.NN[1]);
    const double dx = node_coord(g, 0, ix) - x0, dy = node_coord(g, 1, iy) - y0,
                 dz = node_coord(g, 2, iz) - z0;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double v = peak * exp(-2.0 * d2 * inv_r2);
    // d/dr of P/(r sqrt(pi/2))^3 exp(-2 d^2/r^2) = q (-3/r + 4 d^2/r^3)
    q[idx] = d_dr ? v * (-3.0 / radius + 4.0 * d2 * inv_r2 / radius) : v;
  }
}

// the same source at the support points of an adaptive mesh ([n][3])
__global__ void laser_points_kernel(const double *__restrict__ xyz, double peak, double inv_r2,
                                    double x0, double y0, double z0, double *q, int64_t n,
                                    double radius, int d_dr)
{
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < n;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const double dx = xyz[3 * idx] - x0, dy = xyz[3 * idx + 1] - y0, dz = xyz[3 * idx + 2] - z0;
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double v = peak * exp(-2.0 * d2 * inv_r2);
    q[idx] = d_dr ? v * (-3.0 / radius + 4.0 * d2 * inv_r2 / radius) : v;
  }
}

// ------------------------------------------------------------------------------------------
// FE L2 norm: per cell sqrt(sum_q u_q^2 JxW) stored in single precision (the reference's
// integrate_difference writes a Vector<float>, source/utilities/vector_tools.cpp:137-168), then
// the square root of the sum of squares.  One thread per cell, tensor in registers.
// ------------------------------------------------------------------------------------------
template <int K>
__global__ void l2_norm_kernel(CellArgs<double> A, const double *u, double *part)
{
  constexpr int P = K - 1;
  const int cell = blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (cell < A.n_cells) {
    const int cx = cell % A.ncell[0], cy = (cell / A.ncell[0]) % A.ncell[1],
              cz = cell / (A.ncell[0] * A.ncell[1]);
    double v[K][K][K], w1[K][K][K];
    // adaptive mesh: the cell's nodes from its table (lexicographic local order)
    const uint32_t *cn = A.cell_nodes ? A.cell_nodes + size_t(cell) * (K * K * K) : nullptr;
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int j = 0; j < K; ++j)
#pragma unroll
        for (int i = 0; i < K; ++i)
          v[k][j][i] = cn ? u[cn[i + K * (j + K * k)]]
                          : u[uint32_t(cx * P + i) +
                              A.NX * (uint32_t(cy * P + j) + A.NY * uint32_t(cz * P + k))];
    // x, y, z interpolation
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int j = 0; j < K; ++j)
#pragma unroll
        for (int a = 0; a < K; ++a) {
          double s = 0;
#pragma unroll
          for (int b = 0; b < K; ++b) s += A.S[a * K + b] * v[k][j][b];
          w1[k][j][a] = s;
        }
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
      for (int a = 0; a < K; ++a)
#pragma unroll
        for (int i = 0; i < K; ++i) {
          double s = 0;
#pragma unroll
          for (int b = 0; b < K; ++b) s += A.S[a * K + b] * w1[k][b][i];
          v[k][a][i] = s;
        }
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
      for (int j = 0; j < K; ++j)
#pragma unroll
        for (int i = 0; i < K; ++i) {
          double s = 0;
#pragma unroll
          for (int b = 0; b < K; ++b) s += A.S[a * K + b] * v[b][j][i];
          acc += s * s * (A.w[i] * A.w[j] * A.w[a]);
        }
    acc *= A.h[0] * A.h[1] * A.h[2];
    if (cn) acc = ldexp(acc, -3 * int(A.cell_level[cell]));  // cell size h / 2^level
    const float ef = float(sqrt(acc));
    e2 = double(ef) * double(ef);
  }
  // deterministic block partial
  __shared__ double red[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e2 += __shfl_xor(e2, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e2;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void sum_sqrt_kernel(const double *part, int n, double *out)
{
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += part[i];
    *out = sqrt(s);
  }
}

}  // namespace mf

using namespace mf;

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
namespace mf {

// x[h] = sum_m w x[m] for every hanging node h; masters are free nodes, so the order of the
// updates does not matter
__global__ void distribute_kernel(const int32_t *__restrict__ node_line,
                                  const uint32_t *__restrict__ line_ptr,
                                  const uint32_t *__restrict__ line_col,
                                  const double *__restrict__ line_val, double *x, int64_t n)
{
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t l = node_line[i];
    if (l < 0) continue;
    double v = 0.0;
    for (uint32_t t = line_ptr[l]; t < line_ptr[l + 1]; ++t) v += line_val[t] * x[line_col[t]];
    x[i] = v;
  }
}

// y[m] += sum_h w_hm y[h] for every master m: 16 lanes per master share its entries and
// combine them with shuffles (one group per master, no conflicts)
template <typename Number>
__global__ void condense_masters_kernel(const uint32_t *__restrict__ cm_node,
                                        const uint32_t *__restrict__ cm_ptr,
                                        const uint32_t *__restrict__ cm_h,
                                        const Number *__restrict__ cm_w,
                                        const uint8_t *__restrict__ mask, int64_t n_masters,
                                        Number *y, const int *stop)
{
  if (stop && *stop) return;
  const int lane = threadIdx.x & 15;
  const int64_t groups = int64_t(gridDim.x) * (blockDim.x / 16);
  for (int64_t i = int64_t(blockIdx.x) * (blockDim.x / 16) + (threadIdx.x >> 4); i < n_masters;
       i += groups) {
    const uint32_t m = cm_node[i];
    Number v = 0;
    for (uint32_t t = cm_ptr[i] + uint32_t(lane); t < cm_ptr[i + 1]; t += 16u)
      v += cm_w[t] * y[cm_h[t]];
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 16);
    if (lane == 0 && !(mask && mask[m])) y[m] += v;
  }
}

// hv[l] = sum_t w_t src[m_t], 16 lanes per line
template <typename Number>
__global__ void hanging_values_kernel(const uint32_t *__restrict__ line_ptr,
                                      const uint32_t *__restrict__ line_col,
                                      const Number *__restrict__ line_val,
                                      const uint8_t *__restrict__ mask, const Number *__restrict__ src,
                                      Number *__restrict__ hv, int64_t n_lines, const int *stop)
{
  if (stop && *stop) return;
  const int lane = threadIdx.x & 15;
  const int64_t groups = int64_t(gridDim.x) * (blockDim.x / 16);
  for (int64_t l = int64_t(blockIdx.x) * (blockDim.x / 16) + (threadIdx.x >> 4); l < n_lines;
       l += groups) {
    Number v = 0;
    for (uint32_t t = line_ptr[l] + uint32_t(lane); t < line_ptr[l + 1]; t += 16u) {
      const uint32_t m = line_col[t];
      if (!(mask && mask[m])) v += line_val[t] * src[m];
    }
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) v += __shfl_xor(v, off, 16);
    if (lane == 0) hv[l] = v;
  }
}

template <typename Number>
int launch_hanging_values(const mf_scratch *s, const Number *src, int use_mask, Number *hv,
                          const int *stop)
{
  const AdaptiveData *A = s->ad;
  if (!A || A->n_hanging == 0) return MF_OK;
  const Number *w;
  if constexpr (sizeof(Number) == 8)
    w = reinterpret_cast<const Number *>(A->d.line_val);
  else
    w = reinterpret_cast<const Number *>(A->d.line_val32);
  const int64_t nl = A->n_hanging;
  const int b = int(std::min<int64_t>((nl + 15) / 16, 16384));
  hipLaunchKernelGGL(hanging_values_kernel<Number>, dim3(b), dim3(256), 0, s->stream,
                     A->d.line_ptr, A->d.line_col, w, use_mask ? s->d_mask : nullptr, src, hv, nl,
                     stop);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

template int launch_hanging_values<double>(const mf_scratch *, const double *, int, double *,
                                           const int *);
template int launch_hanging_values<float>(const mf_scratch *, const float *, int, float *,
                                          const int *);

template <typename Number>
__global__ void zero_hanging_kernel(const uint32_t *__restrict__ hang, int64_t n_hang, Number *y,
                                    const int *stop)
{
  if (stop && *stop) return;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n_hang;
       i += int64_t(gridDim.x) * blockDim.x)
    y[hang[i]] = Number(0);
}

template <typename Number>
int launch_condense(const mf_scratch *s, Number *y, const int *stop, bool use_mask)
{
  const AdaptiveData *A = s->ad;
  if (!A || A->hang_nodes.empty()) return MF_OK;
  const int64_t nm = int64_t(A->cm_node.size()), nh = int64_t(A->hang_nodes.size());
  const Number *w;
  if constexpr (sizeof(Number) == 8)
    w = reinterpret_cast<const Number *>(A->d.cm_w);
  else
    w = reinterpret_cast<const Number *>(A->d.cm_w32);
  const int bm = int(std::min<int64_t>((nm + 15) / 16, 16384));
  hipLaunchKernelGGL(condense_masters_kernel<Number>, dim3(bm), dim3(256), 0, s->stream,
                     A->d.cm_node, A->d.cm_ptr, A->d.cm_h, w, use_mask ? s->d_mask : nullptr,
                     nm, y, stop);
  MF_HIP_TRY(hipGetLastError());
  const int bh = int(std::min<int64_t>((nh + 255) / 256, 4096));
  hipLaunchKernelGGL(zero_hanging_kernel<Number>, dim3(bh), dim3(256), 0, s->stream,
                     A->d.hang_nodes, nh, y, stop);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

template int launch_condense<double>(const mf_scratch *, double *, const int *, bool);
template int launch_condense<float>(const mf_scratch *, float *, const int *, bool);

int launch_distribute(const int32_t *node_line, const uint32_t *line_ptr, const uint32_t *line_col,
                      const double *line_val, double *x, int64_t n, hipStream_t st)
{
  const int blocks = int(std::min<int64_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(distribute_kernel, dim3(blocks), dim3(256), 0, st, node_line, line_ptr,
                     line_col, line_val, x, n);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

}  // namespace mf

extern "C" {

int mf_operator_vmult(mf_operator *op, double *dst, const double *src);
int mf_operator_create_rhs(mf_operator *op, double *dst, const double *src, const double *q_s);
int mf_distribute_constraints(mf_scratch *s, double *x);
int mf_operator_reset_dof_index(mf_operator *op, int dof_index);
int mf_operator_set_fields(mf_operator *op, const double *T, const double *vel,
                           const double *k_cell);
int mf_solve(mf_operator *op, const mf_solver_desc *desc, double *x, const double *b,
             mf_solve_report *report);

int mf_predictor_linear_extrapolation(mf_scratch *s, double *dst, const double *x_n,
                                      const double *x_nm1, double tau)
{
  MF_REQUIRE(s && dst && x_n && x_nm1, MF_ERR_ARG, "null argument");
  const int64_t n = s->mesh.n_dofs;
  MF_HIP_TRY(hipMemcpyAsync(dst, x_n, sizeof(double) * n, hipMemcpyDeviceToDevice, s->stream));
  // dst = (1 + tau) x_n - tau x_{n-1}
  int rc = vec_axpby(-tau, x_nm1, 1.0 + tau, dst, n, s->stream);
  return rc;
}

static int laser_impl(mf_scratch *s, double power, double radius, const double position[3],
                      double *q_s, int d_dr)
{
  MF_REQUIRE(s, MF_ERR_ARG, "null scratch");
  MF_REQUIRE(s && position && q_s, MF_ERR_ARG, "null argument");
  MF_REQUIRE(radius > 0.0, MF_ERR_ARG, "The laser beam radius must be greater than zero! Abort..");
  const MeshInfo &m = s->mesh;
  if (s->ad) {
    // nodal values at the support points, then the hanging values from their masters
    // (compute_volumetric_heat_source ends with constraints.distribute)
    AdaptiveData *A = s->ad;
    int rc = ensure_adaptive_on_device(s);
    if (rc) return rc;
    if (!A->d.xyz) {
      MF_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&A->d.xyz), sizeof(double) * A->xyz.size()));
      MF_HIP_TRY(hipMemcpy(A->d.xyz, A->xyz.data(), sizeof(double) * A->xyz.size(),
                           hipMemcpyHostToDevice));
    }
    const double a = radius * std::sqrt(0.5 * M_PI);
    hipLaunchKernelGGL(laser_points_kernel, vgrid(m.n_dofs), dim3(VEC_THREADS), 0, s->stream,
                       A->d.xyz, power / (a * a * a), 1.0 / (radius * radius), position[0],
                       position[1], position[2], q_s, m.n_dofs, radius, d_dr);
    MF_HIP_TRY(hipGetLastError());
    return mf_distribute_constraints(s, q_s);
  }
  NodeGeo g{};
  for (int d = 0; d < 3; ++d) {
    g.lo[d] = m.lo[d];
    g.h[d] = m.h[d];
    g.NN[d] = m.NN[d];
    g.ncell[d] = m.ncell[d];
  }
  g.p = m.degree;
  for (int i = 0; i < m.K; ++i) g.gl[i] = s->tab.gl[i];
  const double a = radius * std::sqrt(0.5 * M_PI);
  const double peak = power / (a * a * a);
  hipLaunchKernelGGL(laser_kernel, vgrid(m.n_dofs), dim3(VEC_THREADS), 0, s->stream, g, peak,
                     1.0 / (radius * radius), position[0], position[1], position[2], q_s,
                     m.n_dofs, radius, d_dr);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

int mf_laser_heat_source(mf_scratch *s, double power, double radius, const double position[3],
                         double *q_s)
{
  return laser_impl(s, power, radius, position, q_s, 0);
}

int mf_laser_heat_source_radius_derivative(mf_scratch *s, double power, double radius,
                                           const double position[3], double *dq_dr)
{
  return laser_impl(s, power, radius, position, dq_dr, 1);
}

int mf_compute_norm_l2(mf_scratch *s, const double *u, double *result)
{
  MF_REQUIRE(s && u && result, MF_ERR_ARG, "null argument");
  const MeshInfo &m = s->mesh;
  CellArgs<double> a;
  std::memset(&a, 0, sizeof(a));
  if (s->ad) {
    // adaptive mesh: u read as stored at each cell's nodes (a constraint-consistent vector)
    int rc = ensure_adaptive_on_device(s);
    if (rc) return rc;
    a.cell_nodes = s->ad->d.cell_nodes;
    a.cell_level = s->ad->d.cell_level;
  }
  for (int i = 0; i < m.K * m.K; ++i) a.S[i] = s->tab.S[i];
  for (int i = 0; i < m.K; ++i) a.w[i] = s->tab.w[i];
  for (int d = 0; d < 3; ++d) {
    a.ncell[d] = m.ncell[d];
    a.h[d] = m.h[d];
  }
  a.NX = uint32_t(m.NN[0]);
  a.NY = uint32_t(m.NN[1]);
  a.n_cells = int(m.n_cells);
  const int threads = 256;
  const int blocks = int((m.n_cells + threads - 1) / threads);
  double *part = nullptr, *out = nullptr;
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&part), sizeof(double) * (blocks + 1), s->stream));
  out = part + blocks;
  switch (m.K) {
  case 2: hipLaunchKernelGGL(l2_norm_kernel<2>, dim3(blocks), dim3(threads), 0, s->stream, a, u, part); break;
  case 3: hipLaunchKernelGGL(l2_norm_kernel<3>, dim3(blocks), dim3(threads), 0, s->stream, a, u, part); break;
  default: hipLaunchKernelGGL(l2_norm_kernel<4>, dim3(blocks), dim3(threads), 0, s->stream, a, u, part); break;
  }
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(sum_sqrt_kernel, dim3(1), dim3(64), 0, s->stream, part, blocks, out);
  MF_HIP_TRY(hipGetLastError());
  MF_HIP_TRY(hipMemcpyAsync(result, out, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  MF_HIP_TRY(hipFreeAsync(part, s->stream));
  MF_HIP_TRY(hipStreamSynchronize(s->stream));
  return MF_OK;
}

// One time step of AdvectionDiffusionOperation::solve
// (source/level_set/advection_diffusion_operation.cpp:227-367) with the Dirichlet lifting of
// create_rhs_and_apply_dirichlet_matrixfree (include/meltpooldg/linear_algebra/
// utilities_matrixfree.hpp:33-100): rhs = b(old) - A_noBC x_d, constrained rows zeroed; solve;
// constrained entries set to their values (AffineConstraints::distribute).
int mf_advection_step(mf_operator *op, double *phi, const double *phi_old,
                      const double *dirichlet_values, const mf_solver_desc *desc,
                      mf_solve_report *report)
{
  MF_REQUIRE(op && phi && phi_old && desc, MF_ERR_ARG, "null argument");
  MF_REQUIRE(op->kind == 1, MF_ERR_ARG, "mf_advection_step needs an advection operator");
  mf_scratch *s = op->s;
  const int64_t n = s->mesh.n_dofs;
  hipStream_t st = s->stream;
  double *rhs = nullptr, *xd = nullptr;
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&rhs), sizeof(double) * n, st));
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&xd), sizeof(double) * n, st));
  int rc = MF_OK;
  const int saved = op->dof_index;
  // constrained entries: Dirichlet rows and the inflow/outflow DoFs (their values lifted as the
  // additional inhomogeneous constraints of create_rhs_and_apply_dirichlet_matrixfree)
  const bool io = op->io_mask && op->io_pre_post;
  const uint8_t *cmask = io ? op->io_union : s->d_mask;
  MF_REQUIRE(!io || dirichlet_values, MF_ERR_ARG,
             "inflow/outflow DoFs need their values (dirichlet_values)");
  do {
    // x_d: the values on constrained entries, zero elsewhere
    if ((rc = vec_masked_assign(xd, cmask, dirichlet_values, false, n, st))) break;
    op->dof_index = 1;
    if ((rc = mf_operator_vmult(op, rhs, xd))) break;
    if ((rc = vec_axpby(0.0, rhs, -1.0, rhs, n, st))) break;
    if ((rc = mf_operator_create_rhs(op, rhs, phi_old, nullptr))) break;
    op->dof_index = 0;
    if ((rc = vec_masked_assign(rhs, cmask, nullptr, true, n, st))) break;
    rc = mf_solve(op, desc, phi, rhs, report);
    if (rc != MF_OK && rc != MF_NO_CONVERGENCE && rc != MF_BREAKDOWN) break;
    const int solve_rc = rc;
    if ((rc = vec_masked_assign(phi, cmask, dirichlet_values, true, n, st))) break;
    // adaptive mesh: the hanging values from their masters (constraints.distribute)
    if (s->ad && (rc = mf_distribute_constraints(s, phi))) break;
    rc = solve_rc;
  } while (false);
  op->dof_index = saved;
  (void)hipFreeAsync(rhs, st);
  (void)hipFreeAsync(xd, st);
  return rc;
}

// One implicit-Euler heat step solved with NewtonRaphsonSolver::solve
// (source/linear_algebra/newton_raphson_solver.cpp:22-175, wired as in
// source/heat/heat_diffuse_operation.cpp:86-116,261-289) through the host driver
// mf_newton_solve: residual r = -R(T) (create_rhs, with T as the linearisation field), J d = r
// by the linear solver, the l2 norms of r, d (and T for relative tolerances) read back once per
// iteration, T += d with the constraints distributed.  T holds the predictor on entry and the
// converged temperature on exit.
namespace {
struct HeatNewtonState {
  mf_operator *op;
  double *T;
  const double *T_old, *q_s;
  const mf_solver_desc *desc;
  double *r, *d, *nrm, *part;
  const double *vel, *kc;
  int64_t n;
};

int heat_residual(void *u)
{
  auto *h = static_cast<HeatNewtonState *>(u);
  hipStream_t st = h->op->s->stream;
  int rc;
  if ((rc = mf_operator_set_fields(h->op, h->T, h->vel, h->kc))) return rc;
  if ((rc = vec_fill<double>(h->r, 0.0, h->n, st))) return rc;
  return mf_operator_create_rhs(h->op, h->r, h->T_old, h->q_s);
}

int heat_jacobian_solve(void *u, int *lin_its)
{
  auto *h = static_cast<HeatNewtonState *>(u);
  int rc;
  if ((rc = vec_fill<double>(h->d, 0.0, h->n, h->op->s->stream))) return rc;
  mf_solve_report lin{};
  rc = mf_solve(h->op, h->desc, h->d, h->r, &lin);
  *lin_its = lin.iterations;
  return rc;
}

int heat_norms(void *u, int need_x, double *rn, double *dn, double *xn)
{
  auto *h = static_cast<HeatNewtonState *>(u);
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

int heat_update(void *u)
{
  auto *h = static_cast<HeatNewtonState *>(u);
  int rc;
  if ((rc = vec_axpby(1.0, h->d, 1.0, h->T, h->n, h->op->s->stream))) return rc;
  // adaptive mesh: the iterate stays constraint-consistent (distribute_constraints)
  if (h->op->s->ad) return mf_distribute_constraints(h->op->s, h->T);
  return MF_OK;
}
}  // namespace

int mf_heat_newton_step(mf_operator *op, double *T, const double *T_old, const double *q_s,
                        const double *dirichlet_values, const mf_solver_desc *desc,
                        const mf_newton_desc *newton, mf_newton_report *report)
{
  MF_REQUIRE(op && T && T_old && desc && newton, MF_ERR_ARG, "null argument");
  MF_REQUIRE(op->kind == 0, MF_ERR_ARG, "mf_heat_newton_step needs a heat operator");
  mf_scratch *s = op->s;
  const int64_t n = s->mesh.n_dofs;
  hipStream_t st = s->stream;
  HeatNewtonState h{op, T, T_old, q_s, desc, nullptr, nullptr, nullptr, nullptr, op->vel,
                    op->k_cell, n};
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&h.r), sizeof(double) * n, st));
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&h.d), sizeof(double) * n, st));
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&h.nrm), sizeof(double) * 3, st));
  MF_HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&h.part), sizeof(double) * 2 * RED_BLOCKS, st));
  int rc = MF_OK;
  // constraints of the predictor (AffineConstraints::distribute)
  if (dirichlet_values) rc = vec_masked_assign(T, s->d_mask, dirichlet_values, true, n, st);
  // on an adaptive mesh the iterate is kept constraint-consistent (hanging values from their
  // masters): NewtonRaphsonSolver's distribute_constraints after every update
  if (rc == MF_OK && s->ad) rc = mf_distribute_constraints(s, T);
  // the mushy-zone Jacobian reads the old temperature (the operator's temperature_old)
  if (rc == MF_OK && op->heat.law == MF_LAW_MUSHY) rc = mf_operator_set_temperature_old(op, T_old);
  mf_newton_report rep{};
  if (rc == MF_OK) {
    const mf_newton_callbacks cb{heat_residual, heat_jacobian_solve, heat_norms, heat_update};
    rc = mf_newton_solve(newton, &cb, &h, &rep);
  }
  (void)hipFreeAsync(h.r, st);
  (void)hipFreeAsync(h.d, st);
  (void)hipFreeAsync(h.nrm, st);
  (void)hipFreeAsync(h.part, st);
  // the operator's linearisation field is the final temperature
  if (rc == MF_OK || rc == MF_NO_CONVERGENCE) {
    const int rc2 = mf_operator_set_fields(op, T, h.vel, h.kc);
    if (rc == MF_OK) rc = rc2;
  }
  if (rc == MF_OK) MF_HIP_TRY(hipStreamSynchronize(st));
  rep.status = rc;
  if (report) *report = rep;
  if (rc == MF_NO_CONVERGENCE)
    set_error("Newton-Raphson solver did not converge (ExcNewtonDidNotConverge)");
  return rc;
}

}  // extern "C"

// Heat Jacobian tangent on structured meshes, FE_Q3 / QGauss(4), with the cell tensors held as
// planes in registers and moved between lanes with permlane swaps (no LDS tensors, no barriers
// inside the sum factorisation).
//
// Reference functions replaced (the cell-level arithmetic of deal.II FEEvaluation driven by):
//   HeatDiffuseMultiPhaseOperator::vmult / tangent_cell_loop / tangent_local_cell_operation
//     (source/heat/heat_diffuse_operator.cpp:179-241, 955-1086)
//   get_material_parameters_and_derivatives (heat_diffuse_operator.cpp:1167-1210) for a constant
//     conductivity or k(T) = k0 + a tanh((T - Tc) / w)
//
// Work decomposition.  A cell's 4 x 4 x 4 tensor lives in 4 lanes, one 4 x 4 plane each: lane
// x + 16 R of a wave serves cell x of the wave's 16 consecutive cells along a row, and holds the
// plane R.  Two layouts are used:
//   P_z: lane R holds the plane with z index R, registers v[a][b] = (x index a, y index b);
//   P_y: lane R holds the plane with y index R, registers v[a][b] = (x index a, z index b).
// Sweeps along a (x) and b (y in P_z, z in P_y) are 4 x 4 matrix-vector products on registers.
// Between the layouts, for each a, the 4 registers v[a][0..3] and the 4 lanes R of a cell form a
// 4 x 4 block that is transposed with two v_permlane32_swap and two v_permlane16_swap (these
// swap the upper half of one register with the lower half of another, and the odd 16-lane rows
// of one with the even rows of another).  A workgroup holds one whole row of cells (16 per
// wave), so the node shared by x-neighbours is combined inside the workgroup; the rows are
// launched in the 4 (y, z) parity colours of the general cell kernel, and the first colour that
// writes a node stores it (dst = A src), the later ones add.
//
// Per quadrature point q (P_y layout) the tangent is
//   (rho cp / dt) u v + k(T) grad u . grad v + dk/dT(T) u grad T . grad v        (times JxW),
// evaluated as val = c_q u and flux_d = A_q du_d + B_qd u with A_q = JxW k(T), B_qd = JxW dk/dT
// dT_d; the derivative sweeps of u and T use Dq / h_d^2 (both 1/h factors of grad u . grad v),
// and the integration uses Dq^T and S^T.
#include "mf_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace mf {

template <typename Number>
struct PlaneArgs {
  Number S[16];       // S[q][j] = l_j(x_q)
  Number Dq[16];      // collocation derivative on the Gauss points
  Number DqS[3][16];  // Dq / h_d^2
  Number ww[16];      // w[a] * w[b]: the Gauss weights of the two register indices
  Number wr[4];       // w[r] * cell volume: the weight of the lane's plane index
  Number mass;        // rho cp / dt
  Number k0, amp, Tc, inv_w, dk0;  // k(T) = k0 + amp tanh((T - Tc) inv_w), dk0 = amp inv_w
  int tdep;           // the dk/dT u grad T . grad v term
  // saturation test of a wave: every quadrature-point T of the wave lies within
  // leb * (max - min) / 2 of the midpoint of its nodal T range (leb = (max_q sum_j |S_qj|)^3),
  // and tanh is +-1 to the last bit and sech^2 is 0 beyond |T - Tc| inv_w >= sat
  Number leb, sat;
  int ncell[3];
  uint32_t NX, NY;
  const uint8_t *mask;  // Dirichlet rows (null: none)
  int mask_boundary;    // the Dirichlet rows are exactly the boundary nodes of the box (computed
                        // from the node coordinates, mask not read)
  int gather_plain;     // read src without the Dirichlet constraints
  int overwrite;        // dst = A src (first writer stores) instead of dst += A src
  int colour, part_off;
  const Number *T;
  const Number *k_cell;
  double *dot_partial;  // fused src . (A src) per workgroup (nullable)
  const int *stop;      // early-exit flag (nullable)
};

// 4 x 4 transpose between the lane rows R (16 lanes each) and 4 registers: on entry register b
// of row R holds M[R][b], on exit M[b][R]
__device__ __forceinline__ void t4(uint32_t &v0, uint32_t &v1, uint32_t &v2, uint32_t &v3)
{
  const auto a = __builtin_amdgcn_permlane32_swap(v0, v2, false, false);
  const auto c = __builtin_amdgcn_permlane32_swap(v1, v3, false, false);
  const auto d = __builtin_amdgcn_permlane16_swap(a[0], c[0], false, false);
  const auto e = __builtin_amdgcn_permlane16_swap(a[1], c[1], false, false);
  v0 = d[0];
  v1 = d[1];
  v2 = e[0];
  v3 = e[1];
}

__device__ __forceinline__ void t4(float &f0, float &f1, float &f2, float &f3)
{
  uint32_t v0 = __float_as_uint(f0), v1 = __float_as_uint(f1), v2 = __float_as_uint(f2),
           v3 = __float_as_uint(f3);
  t4(v0, v1, v2, v3);
  f0 = __uint_as_float(v0);
  f1 = __uint_as_float(v1);
  f2 = __uint_as_float(v2);
  f3 = __uint_as_float(v3);
}

__device__ __forceinline__ void t4(double &d0, double &d1, double &d2, double &d3)
{
  uint64_t u0 = __double_as_longlong(d0), u1 = __double_as_longlong(d1),
           u2 = __double_as_longlong(d2), u3 = __double_as_longlong(d3);
  uint32_t l0 = uint32_t(u0), l1 = uint32_t(u1), l2 = uint32_t(u2), l3 = uint32_t(u3);
  uint32_t h0 = uint32_t(u0 >> 32), h1 = uint32_t(u1 >> 32), h2 = uint32_t(u2 >> 32),
           h3 = uint32_t(u3 >> 32);
  t4(l0, l1, l2, l3);
  t4(h0, h1, h2, h3);
  d0 = __longlong_as_double(int64_t((uint64_t(h0) << 32) | l0));
  d1 = __longlong_as_double(int64_t((uint64_t(h1) << 32) | l1));
  d2 = __longlong_as_double(int64_t((uint64_t(h2) << 32) | l2));
  d3 = __longlong_as_double(int64_t((uint64_t(h3) << 32) | l3));
}

// P_z <-> P_y: lane index and register index b exchange, for every a
template <typename Number>
__device__ __forceinline__ void transpose_planes(Number (&v)[4][4])
{
#pragma unroll
  for (int a = 0; a < 4; ++a) t4(v[a][0], v[a][1], v[a][2], v[a][3]);
}

// M (row-major 4 x 4, or its transpose) along index a, for every b
template <typename Number, bool TR>
__device__ __forceinline__ void sweep_a(const Number *M, Number (&v)[4][4])
{
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const Number in[4] = {v[0][b], v[1][b], v[2][b], v[3][b]};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      Number acc = Number(0);
#pragma unroll
      for (int m = 0; m < 4; ++m) acc = fma(TR ? M[m * 4 + a] : M[a * 4 + m], in[m], acc);
      v[a][b] = acc;
    }
  }
}

// M along index b, for every a
template <typename Number, bool TR>
__device__ __forceinline__ void sweep_b(const Number *M, Number (&v)[4][4])
{
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const Number in[4] = {v[a][0], v[a][1], v[a][2], v[a][3]};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      Number acc = Number(0);
#pragma unroll
      for (int m = 0; m < 4; ++m) acc = fma(TR ? M[m * 4 + b] : M[b * 4 + m], in[m], acc);
      v[a][b] = acc;
    }
  }
}

template <typename Number>
__device__ __forceinline__ void copy_planes(const Number (&s)[4][4], Number (&d)[4][4])
{
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) d[a][b] = s[a][b];
}

// two consecutive entries (8-byte aligned) in one 16-byte access
typedef double plane_d2 __attribute__((ext_vector_type(2)));
typedef float plane_f2 __attribute__((ext_vector_type(2)));
typedef plane_d2 plane_d2u __attribute__((aligned(8)));
typedef plane_f2 plane_f2u __attribute__((aligned(4)));
template <typename Number>
struct Pair;
template <>
struct Pair<double> {
  typedef plane_d2 v;
  typedef plane_d2u u;
};
template <>
struct Pair<float> {
  typedef plane_f2 v;
  typedef plane_f2u u;
};
template <typename Number>
__device__ __forceinline__ void load2(const Number *p, Number &a, Number &b)
{
  const typename Pair<Number>::v v = *reinterpret_cast<const typename Pair<Number>::u *>(p);
  a = v.x;
  b = v.y;
}
template <typename Number>
__device__ __forceinline__ void store2(Number *p, Number a, Number b)
{
  typename Pair<Number>::v v;
  v.x = a;
  v.y = b;
  *reinterpret_cast<typename Pair<Number>::u *>(p) = v;
}

// tanh(x) and sech^2(x) from e = exp(-2|x|) (see the general cell kernel: beyond |x| = 20 in
// double and 10 in float, tanh is +-1 to the last bit and the sech^2 term is below rounding)
__device__ __forceinline__ double plane_rcp(double d)
{
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  return fma(fma(-d, r, 1.0), r, r);
}
__device__ __forceinline__ float plane_rcp(float d)
{
  const float r = __builtin_amdgcn_rcpf(d);
  return fmaf(fmaf(-d, r, 1.0f), r, r);
}
__device__ __forceinline__ double plane_exp_neg2(double ax) { return exp(-2.0 * ax); }
__device__ __forceinline__ float plane_exp_neg2(float ax)
{
  return __builtin_amdgcn_exp2f(-2.8853900817779268f * ax);  // 2 log2(e)
}
template <typename Number>
__device__ __forceinline__ void plane_tanh_sech2(Number x, Number &t, Number &s2)
{
  const Number ax = fabs(x);
  const Number cut = sizeof(Number) == 8 ? Number(20) : Number(10);
  if (ax < cut) {
    const Number e = plane_exp_neg2(ax);
    const Number inv = plane_rcp(Number(1) + e);
    t = (Number(1) - e) * inv;
    s2 = Number(4) * e * inv * inv;
    if (x < 0) t = -t;
  } else {
    t = x < 0 ? Number(-1) : Number(1);
    s2 = 0;
  }
}

template <typename Number>
__device__ __forceinline__ Number wave_sum(Number v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// The loads of one cell row for one lane: T (when read), the source and the Dirichlet bits of
// the lane's plane.
template <typename Number, bool TLAW>
struct PlaneRaw {
  Number t[TLAW ? 16 : 1];
  Number u[16];
  uint32_t mb;  // bit i + 4 j: node (i, j) of the plane is a Dirichlet row
};

// Node rows j of the lane's plane (z node 3 cz + R) of row `row` of the launch's colour start at
// x node 3 cx (the row start for idle lanes, so that every load is in bounds).
struct PlaneRow {
  int cy, cz;
  uint32_t g[4];
};

__device__ __forceinline__ PlaneRow plane_row(const int ncell[3], uint32_t NX, uint32_t NY,
                                              const RowColour &rc, int row, int R, int cx,
                                              bool active)
{
  PlaneRow r;
  r.cy = 2 * (row % rc.hy) + rc.yb;
  r.cz = 2 * (row / rc.hy) + rc.zb;
#pragma unroll
  for (int j = 0; j < 4; ++j)
    r.g[j] = NX * (uint32_t(3 * r.cy + j) + NY * uint32_t(3 * r.cz + R)) + uint32_t(active ? 3 * cx : 0);
  return r;
}

// Issue the loads of a row: T first, then the source, then the Dirichlet bits (read, or from the
// node coordinates when the mask is the box boundary).  Nothing is selected here, so that all
// loads go out together; the T phase of the row then runs while the source values may still be in
// flight.  A pencil's 4 nodes are read as two 16-byte pairs.
template <typename Number, bool TLAW>
__device__ __forceinline__ void plane_issue(const PlaneArgs<Number> &A, const Number *src,
                                            const PlaneRow &r, int R, int cx,
                                            PlaneRaw<Number, TLAW> &raw)
{
  if constexpr (TLAW) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      load2(A.T + r.g[j], raw.t[4 * j], raw.t[4 * j + 1]);
      load2(A.T + r.g[j] + 2, raw.t[4 * j + 2], raw.t[4 * j + 3]);
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    load2(src + r.g[j], raw.u[4 * j], raw.u[4 * j + 1]);
    load2(src + r.g[j] + 2, raw.u[4 * j + 2], raw.u[4 * j + 3]);
  }
  uint32_t mb = 0;
  if (A.mask_boundary) {
    // node (3 cx + i, 3 cy + j, 3 cz + R) on a face of the box: bit i + 4 j
    if ((R == 0 && r.cz == 0) || (R == 3 && r.cz == A.ncell[2] - 1)) mb = 0xffffu;
    if (r.cy == 0) mb |= 0x000fu;
    if (r.cy == A.ncell[1] - 1) mb |= 0xf000u;
    if (cx == 0) mb |= 0x1111u;
    if (cx == A.ncell[0] - 1) mb |= 0x8888u;
  } else if (A.mask) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) mb |= uint32_t(A.mask[r.g[j] + i] != 0) << (i + 4 * j);
  }
  raw.mb = mb;
}

// A workgroup serves one cell row of the launch's colour (16 cells per wave).  TLAW: k(T) from
// the tanh law (T read), otherwise k = k0 (T not read).
template <typename Number, bool TLAW, bool DOT>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2)))
    plane_tangent_kernel(const PlaneArgs<Number> A, const Number *__restrict__ src,
                         Number *__restrict__ dst)
{
  __shared__ Number xb[8][4][4];  // per wave: the outputs of its last cell at the node i = 3
  __shared__ double red[8];
  // fused dot: the gathered source (constrained entries zero), entry k of thread t at k * 512 + t
  __shared__ Number us[DOT ? 16 * 512 : 1];
  if (A.stop && *A.stop) return;  // uniform over the grid

  const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
  const int nw = int(blockDim.x) >> 6;
  const int x = lane & 15, R = lane >> 4;
  const RowColour rc = row_colour(A.ncell, A.colour);
  const int ncx = A.ncell[0];
  const int cx = wave * 16 + x;
  const bool active = cx < ncx;
  const bool last = cx == ncx - 1;
  // the Gauss weight of the lane's plane index times the cell volume
  Number wrR = A.wr[0];
#pragma unroll
  for (int r = 1; r < 4; ++r) wrR = R == r ? A.wr[r] : wrR;

  const int row = int(blockIdx.x);
  const PlaneRow cur = plane_row(A.ncell, A.NX, A.NY, rc, row, R, cx, active);
  PlaneRaw<Number, TLAW> raw;
  plane_issue<Number, TLAW>(A, src, cur, R, cx, raw);
  {
    const int cy = cur.cy, cz = cur.cz;
    [[maybe_unused]] Number t[4][4];
    Number u[4][4];
    const uint32_t mbits = raw.mb;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if constexpr (TLAW) t[i][j] = active ? raw.t[i + 4 * j] : Number(0);
        u[i][j] = raw.u[i + 4 * j];
      }
    const Number kc = (A.k_cell && active) ? A.k_cell[cx + ncx * (cy + A.ncell[1] * cz)]
                                           : Number(1);

    // ---------------- T: values and gradients at the quadrature points (P_y) ----------------
    // Aq = JxW k(T) kc, Bd = JxW dk/dT(T) kc dT/dx_d / h_d^2.  A saturated wave (tanh = +-1 at
    // every quadrature point of its 16 cells, decided from the nodal T range) needs no T at the
    // quadrature points: k = k0 +- amp and dk/dT = 0, so the T sweeps and the B terms vanish
    // exactly; without the dk/dT term (tdep = 0) the T gradients are not formed either.
    Number Aq[4][4];
    [[maybe_unused]] Number Bx[4][4], By[4][4], Bz[4][4];
    if constexpr (TLAW) {
      // nodal T range of the lane's cell (its 4 planes sit in the lanes x + 16 R); a NaN or
      // infinite T gives an infinite range and the general path.  The wave takes the shortcut
      // when all its active cells are saturated (each with its own sign).
      Number tmin = raw.t[0], tmax = raw.t[0], probe = raw.t[0] * Number(0);
#pragma unroll
      for (int e = 1; e < 16; ++e) {
        tmin = fmin(tmin, raw.t[e]);
        tmax = fmax(tmax, raw.t[e]);
        probe += raw.t[e] * Number(0);
      }
      if (!(probe == Number(0))) {
        tmin = -Number(INFINITY);
        tmax = Number(INFINITY);
      }
#pragma unroll
      for (int off = 16; off < 64; off <<= 1) {
        tmin = fmin(tmin, __shfl_xor(tmin, off, 64));
        tmax = fmax(tmax, __shfl_xor(tmax, off, 64));
      }
      const Number mid = Number(0.5) * (tmin + tmax), rad = Number(0.5) * (tmax - tmin);
      const bool cell_saturated = (fabs(mid - A.Tc) - A.leb * rad) * A.inv_w >= A.sat;
      const bool saturated = __all(cell_saturated || !active);
      if (saturated) {
        const Number th = mid < A.Tc ? Number(-1) : Number(1);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            Aq[a][b] = (A.ww[a * 4 + b] * wrR) * ((A.k0 + A.amp * th) * kc);
            Bx[a][b] = By[a][b] = Bz[a][b] = Number(0);
          }
      } else {
        const bool with_b = A.tdep != 0;
        sweep_a<Number, false>(A.S, t);
        sweep_b<Number, false>(A.S, t);
        transpose_planes(t);
        sweep_b<Number, false>(A.S, t);  // T at the quadrature points
        if (with_b) {
          copy_planes(t, Bx);
          sweep_a<Number, false>(A.DqS[0], Bx);
          copy_planes(t, Bz);
          sweep_b<Number, false>(A.DqS[2], Bz);
          copy_planes(t, By);
          transpose_planes(By);
          sweep_b<Number, false>(A.DqS[1], By);
          transpose_planes(By);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            const Number jxw = A.ww[a * 4 + b] * wrR;
            Number th, s2;
            plane_tanh_sech2((t[a][b] - A.Tc) * A.inv_w, th, s2);
            Aq[a][b] = jxw * ((A.k0 + A.amp * th) * kc);
            if (with_b) {
              const Number dk = jxw * (A.dk0 * kc) * s2;
              Bx[a][b] *= dk;
              By[a][b] *= dk;
              Bz[a][b] *= dk;
            } else {
              Bx[a][b] = By[a][b] = Bz[a][b] = Number(0);
            }
          }
      }
    } else {
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) Aq[a][b] = A.ww[a * 4 + b] * wrR * (A.k0 * kc);
    }

    // read_dof_values: Dirichlet entries of the source (and everything on idle lanes) read as 0
    const uint32_t zero_bits = active ? (A.gather_plain ? 0u : mbits) : 0xffffu;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if ((zero_bits >> (i + 4 * j)) & 1u) u[i][j] = Number(0);
    if constexpr (DOT) {
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) us[(i + 4 * j) * 512 + int(threadIdx.x)] = u[i][j];
    }

    // ---------------- u: values and gradients at the quadrature points, fluxes ----------------
    // Each flux is integrated (Dq^T) and summed as soon as it is formed, so that the T
    // coefficients and the flux planes die early:
    // acc = Dq_x^T fx + Dq_z^T fz + (rho cp / dt) JxW u + Dq_y^T fy.
    sweep_a<Number, false>(A.S, u);
    sweep_b<Number, false>(A.S, u);
    transpose_planes(u);
    sweep_b<Number, false>(A.S, u);  // u at the quadrature points (P_y)
    Number acc[4][4], f[4][4];
    copy_planes(u, acc);
    sweep_a<Number, false>(A.DqS[0], acc);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if constexpr (TLAW)
          acc[a][b] = fma(Aq[a][b], acc[a][b], Bx[a][b] * u[a][b]);
        else
          acc[a][b] *= Aq[a][b];
      }
    sweep_a<Number, true>(A.Dq, acc);
    copy_planes(u, f);
    sweep_b<Number, false>(A.DqS[2], f);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if constexpr (TLAW)
          f[a][b] = fma(Aq[a][b], f[a][b], Bz[a][b] * u[a][b]);
        else
          f[a][b] *= Aq[a][b];
      }
    sweep_b<Number, true>(A.Dq, f);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += f[a][b];
    copy_planes(u, f);
    transpose_planes(f);
    sweep_b<Number, false>(A.DqS[1], f);
    transpose_planes(f);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if constexpr (TLAW)
          f[a][b] = fma(Aq[a][b], f[a][b], By[a][b] * u[a][b]);
        else
          f[a][b] *= Aq[a][b];
        acc[a][b] = fma(A.ww[a * 4 + b] * wrR * A.mass, u[a][b], acc[a][b]);  // (rho cp/dt) JxW u
      }

    const PlaneRow &cur_row = cur;

    transpose_planes(f);
    sweep_b<Number, true>(A.Dq, f);
    transpose_planes(f);

    // ---------------- integrate: S^T in every direction ----------------
    // the dst entries the scatter adds to, read while the last sweeps run (the rows of one
    // colour own their nodes during the launch; not needed where this colour is the first
    // writer of an overwriting application)
    const bool shz = (R == 0 && cz > 0) || (R == 3 && cz < A.ncell[2] - 1);
    uint32_t store_rows = 0;  // bit j: node row j is stored, not added to
    Number d[4][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const bool shy = (j == 0 && cy > 0) || (j == 3 && cy < A.ncell[1] - 1);
      const bool first = (!shy || rc.yb == 0) && (!shz || rc.zb == 0);
      const bool store = A.overwrite && first;
      store_rows |= uint32_t(store) << j;
      d[0][j] = d[1][j] = d[2][j] = d[3][j] = Number(0);
      if (active && !store) {
        load2(dst + cur_row.g[j], d[0][j], d[1][j]);
        d[2][j] = dst[cur_row.g[j] + 2];
        if (last) d[3][j] = dst[cur_row.g[j] + 3];
      }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) u[a][b] = acc[a][b] + f[a][b];
    sweep_a<Number, true>(A.S, u);
    sweep_b<Number, true>(A.S, u);
    transpose_planes(u);
    sweep_b<Number, true>(A.S, u);  // u[i][j]: the cell's contribution at node (i, j, R) (P_z)

    // ---------------- fused src . (A src) over the cell's nodes ----------------
    if constexpr (DOT) {
      Number local = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) local += us[(i + 4 * j) * 512 + int(threadIdx.x)] * u[i][j];
      const double ws = wave_sum<double>(double(local));
      if (lane == 0) red[wave] = ws;
    }

    // ---------------- scatter ----------------
    // the node shared with the previous cell in x: from the previous lane of the 16-lane row,
    // and for the wave's first cell from the previous wave (through LDS)
    if (x == 15) {
#pragma unroll
      for (int j = 0; j < 4; ++j) xb[wave][R][j] = u[3][j];
    }
    __syncthreads();
    if constexpr (DOT) {
      if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < nw; ++w) s += red[w];
        A.dot_partial[A.part_off + row] = s;  // one partial per row
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const Number prev = __shfl_up(u[3][j], 1, 16);
      u[0][j] += x > 0 ? prev : (wave > 0 ? xb[wave - 1][R][j] : Number(0));
    }
    if (active) {
      // every node the cell writes (i < 3, and i = 3 at the row end) gets its final value:
      // constrained rows are zero where this colour stores them (dst = A src) and keep their
      // value where it adds (distribute_local_to_global skips them)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bool store = (store_rows >> j) & 1u;
        Number v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const bool con = (mbits >> (i + 4 * j)) & 1u;
          v[i] = con ? (store ? Number(0) : d[i][j]) : (store ? u[i][j] : d[i][j] + u[i][j]);
        }
        Number *q = dst + cur_row.g[j];
        store2(q, v[0], v[1]);
        q[2] = v[2];
        if (last) q[3] = v[3];
      }
    }
  }
}

// whether the plane kernel serves this application (structured FE_Q3 heat tangent, constant or
// tanh conductivity, no velocity, not transposed, rows of at most 128 cells)
template <typename Number>
bool plane_applicable(const MeshInfo &m, int needT, bool vel, const CellArgs<Number> &a)
{
  static const int disabled = [] {
    const char *e = std::getenv("MF_PLANE_KERNEL");
    return (e && e[0] == '0') ? 1 : 0;
  }();
  return !disabled && !m.ad && m.K == 4 && (needT == 0 || needT == 1) && !vel && !a.transpose &&
         a.hang_val == nullptr && m.ncell[0] >= 1 && m.ncell[0] <= 128;
}

template <typename Number>
int launch_heat_plane(const MeshInfo &m, int needT, const CellArgs<Number> &a, const Number *src,
                      Number *dst, hipStream_t st)
{
  PlaneArgs<Number> p;
  std::memset(&p, 0, sizeof(p));
  for (int i = 0; i < 16; ++i) {
    p.S[i] = a.S[i];
    p.Dq[i] = a.Dq[i];
    for (int d = 0; d < 3; ++d) p.DqS[d][i] = Number(double(a.Dq[i]) / (m.h[d] * m.h[d]));
  }
  const double vol = m.h[0] * m.h[1] * m.h[2];
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < 4; ++j) p.ww[i * 4 + j] = Number(double(a.w[i]) * double(a.w[j]));
    p.wr[i] = Number(double(a.w[i]) * vol);
  }
  p.mass = Number(double(a.rho_cp) * double(a.dt_inv));
  p.k0 = a.k0;
  p.amp = a.tanh_amp;
  p.Tc = a.tanh_Tc;
  p.inv_w = a.inv_tanh_w;
  p.dk0 = Number(double(a.tanh_amp) * double(a.inv_tanh_w));
  p.tdep = a.t_dependent;
  {
    double m = 0.0;
    for (int q = 0; q < 4; ++q) {
      double r = 0.0;
      for (int j = 0; j < 4; ++j) r += std::fabs(double(a.S[q * 4 + j]));
      m = std::max(m, r);
    }
    // a small margin over the cut of plane_tanh_sech2 covers the rounding of the sweeps
    p.leb = Number(m * m * m * (1.0 + 1e-6));
    p.sat = sizeof(Number) == 8 ? Number(20.0 * (1.0 + 1e-6)) : Number(10.0 * (1.0 + 1e-3));
  }
  for (int d = 0; d < 3; ++d) p.ncell[d] = m.ncell[d];
  p.NX = uint32_t(m.NN[0]);
  p.NY = uint32_t(m.NN[1]);
  p.mask = (a.use_mask && a.mask) ? a.mask : nullptr;
  p.mask_boundary = (p.mask && a.mask_boundary) ? 1 : 0;
  p.gather_plain = a.gather_plain;
  p.overwrite = a.overwrite;
  p.T = a.T;
  p.k_cell = a.k_cell;
  p.dot_partial = a.dot_partial;
  p.stop = a.stop;
  const int waves = (m.ncell[0] + 15) / 16;
  int part_off = 0;
  for (int colour = 0; colour < 4; ++colour) {
    const RowColour rc = row_colour(m.ncell, colour);
    const int rows = rc.hy * rc.hz;
    if (rows == 0) continue;
    const int blocks = rows;
    p.colour = colour;
    p.part_off = part_off;
    const bool tl = needT == 1, dot = a.dot_partial != nullptr;
    if (tl && dot)
      hipLaunchKernelGGL((plane_tangent_kernel<Number, true, true>), dim3(blocks), dim3(64 * waves), 0, st, p, src, dst);
    else if (tl)
      hipLaunchKernelGGL((plane_tangent_kernel<Number, true, false>), dim3(blocks), dim3(64 * waves), 0, st, p, src, dst);
    else if (dot)
      hipLaunchKernelGGL((plane_tangent_kernel<Number, false, true>), dim3(blocks), dim3(64 * waves), 0, st, p, src, dst);
    else
      hipLaunchKernelGGL((plane_tangent_kernel<Number, false, false>), dim3(blocks), dim3(64 * waves), 0, st, p, src, dst);
    MF_HIP_TRY(hipGetLastError());
    part_off += rows;
  }
  return MF_OK;
}

template bool plane_applicable<double>(const MeshInfo &, int, bool, const CellArgs<double> &);
template bool plane_applicable<float>(const MeshInfo &, int, bool, const CellArgs<float> &);
template int launch_heat_plane<double>(const MeshInfo &, int, const CellArgs<double> &,
                                       const double *, double *, hipStream_t);
template int launch_heat_plane<float>(const MeshInfo &, int, const CellArgs<float> &,
                                      const float *, float *, hipStream_t);

}  // namespace mf

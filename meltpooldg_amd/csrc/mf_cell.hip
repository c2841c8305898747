// Matrix-free cell kernels for gfx950: gather -> sum-factorised evaluate -> quadrature-point
// physics -> sum-factorised integrate -> scatter-add, for the heat Jacobian / residual /
// diagonal and the advection-diffusion tangent / rhs / diagonal.
//
// Reference functions replaced (cell-level arithmetic of deal.II FEEvaluation driven by):
//   HeatDiffuseMultiPhaseOperator::tangent_cell_loop + tangent_local_cell_operation
//     (source/heat/heat_diffuse_operator.cpp:199-241, 955-1086)
//   HeatDiffuseMultiPhaseOperator::rhs_cell_loop (source/heat/heat_diffuse_operator.cpp:393-529)
//   internal_compute_diagonal_or_system_matrix (source/heat/heat_diffuse_operator.cpp:297-391)
//   AdvectionDiffusionOperator::vmult / create_rhs / compute_inverse_diagonal_from_matrixfree /
//     tangent_local_cell_operation (source/level_set/advection_diffusion_operator.cpp:236-495)
//
// Work decomposition.  A workgroup holds CPB cells; each cell is served by K*K threads, thread
// (cell c, pencil p) with c fastest in threadIdx.x so that the 16-lane groups of a wave touch
// 16 different cells (consecutive LDS words, consecutive cells in x in HBM).  Each 1D sweep is a
// K x K matrix applied to K-long pencils: a thread reads its pencil from LDS into registers,
// multiplies, and writes it back; the pencil direction changes between phases, separated by
// barriers.  Cell tensors live in LDS as [K^3][CPB] arrays.  Gradients at the quadrature
// points use the collocation derivative Dq on the Gauss points (grad = Dq (S u)), which equals
// (DS u) exactly in exact arithmetic because K points determine the degree-p polynomial.
// Scatter without atomics: a workgroup walks one row of cells along x, CPB cells per step, and
// the rows are launched in 4 colours by the parity of (cell y, cell z).  Rows of one colour share
// no node, so each adds its nodes with a plain read-modify-write; the node shared by
// x-neighbours is combined with a lane shuffle (or carried to the next step) and written once.
// The 4 launches run in stream order.
#include "mf_internal.hpp"

#include <type_traits>

namespace mf {

#ifndef MF_CPB_F32_K4
#define MF_CPB_F32_K4 16
#endif
#ifndef MF_CPB_F64_K4
#define MF_CPB_F64_K4 16
#endif
#ifndef MF_PREFETCH_F32
#define MF_PREFETCH_F32 1
#endif
#ifndef MF_PREFETCH_F64
#define MF_PREFETCH_F64 0
#endif
#ifndef MF_WPE_F64_K4
#define MF_WPE_F64_K4 4
#endif
#ifndef MF_WPE_F32_K4
#define MF_WPE_F32_K4 1
#endif

// Cells per workgroup step (one x row segment); K*K threads per cell.
template <typename Number, int K>
constexpr int cells_per_block()
{
  if constexpr (sizeof(Number) == 4 && K == 4) return MF_CPB_F32_K4;
  if constexpr (sizeof(Number) == 8 && K == 4) return MF_CPB_F64_K4;
  return K == 2 ? 64 : 16;
}

// Minimum waves per SIMD the register allocation must allow.  The FP64 Q3 tangent kernels
// without a velocity field fit 4 waves (128 VGPRs) without spilling; with the velocity, and the
// diagonal kernels, they would spill.
template <typename Number, int K, bool VEL, int OP>
constexpr int min_waves_per_eu()
{
  if constexpr (sizeof(Number) == 4 && K == 4) return MF_WPE_F32_K4;
  if constexpr (sizeof(Number) == 8 && K == 4) return (VEL || !is_tangent_op(OP)) ? 1 : MF_WPE_F64_K4;
  return 1;
}

// Whether the row walk loads the next step's pencils during the current step (costs 2K+1
// registers per thread; pays where it does not lower the occupancy).
template <typename Number, int K>
constexpr bool prefetch_pencils()
{
  if constexpr (sizeof(Number) == 4) return MF_PREFETCH_F32;
  return MF_PREFETCH_F64;
}

template <int K>
__device__ __forceinline__ int elem(int i, int j, int k)
{
  return i + K * j + K * K * k;
}

// element e of pencil p (p = a + K*b) along direction DIR
template <int K, int DIR>
__device__ __forceinline__ int pencil(int p, int e)
{
  const int a = p % K, b = p / K;
  if constexpr (DIR == 0) return elem<K>(e, a, b);
  if constexpr (DIR == 1) return elem<K>(a, e, b);
  return elem<K>(a, b, e);
}

// LDS layout of a cell tensor: [K^3][CPB] (cell fastest, so that the 16-lane groups of a wave,
// which serve 16 different cells, hit 16 different 8-byte bank pairs).  FP32 accesses are served
// in 32-lane groups, two pencils of 16 cells; for x pencils those sit 4 element rows apart, in
// the same banks, so the FP32 layout swaps element rows e and e ^ 1 where bit 2 of e is set
// (a bijection that puts such pairs into opposite bank halves and keeps y and z pencils apart).
template <typename Number, int K, int CPB_>
struct Layout {
  static constexpr int CPB = CPB_;
  static constexpr bool SWZ = sizeof(Number) == 4 && K == 4;
  __device__ static __forceinline__ int idx(int e, int c)
  {
    if constexpr (SWZ) return (e ^ ((e >> 2) & 1)) * CPB + c;
    return e * CPB + c;
  }
};

template <typename Number, int K, class LY, int DIR>
__device__ __forceinline__ void load_pencil(const Number *buf, int p, int c, Number (&v)[K])
{
#pragma unroll
  for (int e = 0; e < K; ++e) v[e] = buf[LY::idx(pencil<K, DIR>(p, e), c)];
}

template <typename Number, int K, class LY, int DIR>
__device__ __forceinline__ void store_pencil(Number *buf, int p, int c, const Number (&v)[K])
{
#pragma unroll
  for (int e = 0; e < K; ++e) buf[LY::idx(pencil<K, DIR>(p, e), c)] = v[e];
}

// A 1D matrix (S or Dq, K x K row-major) as the sweeps use it.  VALU form: the matrix itself,
// read from the kernel arguments.
template <typename Number, int K>
struct Mat1D {
  const Number *M;
  __device__ explicit Mat1D(const Number *m) : M(m) {}
};

// MFMA form (FP32): v_mfma_f32_4x4x1f32 treats the wave as 16 blocks of 4 lanes and adds
// A[4b + i] * B[4b + j] to register i of lane 4b + j.  With B = element k of the lane's own
// pencil and A = M[i][k] on lane 4b + i, the 4-register accumulator of every lane becomes M times
// its pencil after K instructions (one per k).  n[k] = M[r][k] and t[k] = M[k][r] with r = lane
// mod 4 are the A operands of M and M^T; rows r >= K are zero.
template <int K>
struct Mat1DMfma {
  float n[K], t[K];
  __device__ explicit Mat1DMfma(const float *m)
  {
    const int r = int(threadIdx.x) & 3;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      n[k] = r < K ? m[r * K + k] : 0.0f;
      t[k] = r < K ? m[k * K + r] : 0.0f;
    }
  }
};

// out = M in (TRANSPOSE: out = M^T in), M row-major K x K
template <typename Number, int K, bool TRANSPOSE>
__device__ __forceinline__ void matvec(const Mat1D<Number, K> &m, const Number (&in)[K],
                                       Number (&out)[K])
{
  const Number *M = m.M;
#pragma unroll
  for (int i = 0; i < K; ++i) {
    Number acc = 0;
#pragma unroll
    for (int mm = 0; mm < K; ++mm) acc = fma(TRANSPOSE ? M[mm * K + i] : M[i * K + mm], in[mm], acc);
    out[i] = acc;
  }
}

typedef float mf_v4f __attribute__((ext_vector_type(4)));

template <typename Number, int K, bool TRANSPOSE>
__device__ __forceinline__ void matvec(const Mat1DMfma<K> &m, const float (&in)[K], float (&out)[K])
{
  mf_v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < K; ++k)
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(TRANSPOSE ? m.t[k] : m.n[k], in[k], acc, 0, 0, 0);
#pragma unroll
  for (int i = 0; i < K; ++i) out[i] = acc[i];
}

template <typename Number, int K, class LY, int DIR, bool TRANSPOSE, class M1>
__device__ __forceinline__ void sweep_inplace(const M1 &M, Number *buf, int p, int c)
{
  Number v[K], o[K];
  load_pencil<Number, K, LY, DIR>(buf, p, c, v);
  matvec<Number, K, TRANSPOSE>(M, v, o);
  store_pencil<Number, K, LY, DIR>(buf, p, c, o);
}

template <typename Number, int K, class LY, int DIR, bool TRANSPOSE, class M1>
__device__ __forceinline__ void sweep(const M1 &M, const Number *in, Number *out, int p, int c)
{
  Number v[K], o[K];
  load_pencil<Number, K, LY, DIR>(in, p, c, v);
  matvec<Number, K, TRANSPOSE>(M, v, o);
  store_pencil<Number, K, LY, DIR>(out, p, c, o);
}

// Two independent sweeps with the LDS reads of both issued before either computes, so that one
// LDS round trip serves both (in place is allowed: each thread rewrites only its own pencils).
template <typename Number, int K, class LY, int DA, bool TA, int DB, bool TB, class M1>
__device__ __forceinline__ void sweep_pair(const M1 &MA, const Number *inA, Number *outA,
                                           const M1 &MB, const Number *inB, Number *outB, int p,
                                           int c)
{
  Number va[K], vb[K], oa[K], ob[K];
  load_pencil<Number, K, LY, DA>(inA, p, c, va);
  load_pencil<Number, K, LY, DB>(inB, p, c, vb);
  matvec<Number, K, TA>(MA, va, oa);
  matvec<Number, K, TB>(MB, vb, ob);
  store_pencil<Number, K, LY, DA>(outA, p, c, oa);
  store_pencil<Number, K, LY, DB>(outB, p, c, ob);
}

// a[idx] for a thread-dependent idx < K without a memory access (a is uniform)
template <typename Number, int K>
__device__ __forceinline__ Number select_k(const Number *a, int idx)
{
  Number v = a[0];
#pragma unroll
  for (int k = 1; k < K; ++k) v = idx == k ? a[k] : v;
  return v;
}

// Sweeps of the FP32 kernels with K = 3, 4 run on the matrix cores (the FP32 MFMA rate is twice
// the unpacked FP32 VALU rate, and the VALU is left to the quadrature-point work); FP64 sweeps
// stay on the VALU (same FP64 rate on both units here, no lane shuffles needed).
template <typename Number, int K>
struct SweepMat {
  static constexpr bool MFMA = sizeof(Number) == 4 && (K == 3 || K == 4);
  using type = typename std::conditional<MFMA, Mat1DMfma<K>, Mat1D<Number, K>>::type;
};

// tanh(x) and sech^2(x) from e = exp(-2|x|): tanh|x| = (1 - e)/(1 + e), sech^2 x = 4e/(1 + e)^2.
// 1/(1 + e) lies in [1/2, 1], so the hardware reciprocal estimate refined by Newton steps (two
// in double, one in float) is accurate to the last bit or two without special cases.  For
// small |x| the (1 - e) cancellation costs relative accuracy of tanh but not absolute accuracy,
// which is what k = k0 + a tanh needs.  Beyond |x| = 20 (double) / 10 (float) tanh is +-1 to the
// last bit and the sech^2 term is below 1e-16 (1e-8) of the dk/dT scale, so it is taken as 0.
__device__ __forceinline__ double fast_rcp(double d)
{
  double r = __builtin_amdgcn_rcp(d);
  r = fma(fma(-d, r, 1.0), r, r);
  return fma(fma(-d, r, 1.0), r, r);
}
__device__ __forceinline__ float fast_rcp(float d)
{
  const float r = __builtin_amdgcn_rcpf(d);
  return fmaf(fmaf(-d, r, 1.0f), r, r);
}
__device__ __forceinline__ double exp_neg2(double ax) { return exp(-2.0 * ax); }
__device__ __forceinline__ float exp_neg2(float ax)
{
  return __builtin_amdgcn_exp2f(-2.8853900817779268f * ax);  // 2 log2(e)
}

// the matrix handle of one step: the hoisted MFMA operands, or the VALU matrix as read from the
// kernel arguments in this step
template <typename Number, int K>
__device__ __forceinline__ typename SweepMat<Number, K>::type step_mat(
    const typename SweepMat<Number, K>::type &hoisted, const Number *M)
{
  if constexpr (SweepMat<Number, K>::MFMA) {
    return hoisted;
  } else {
    return Mat1D<Number, K>(M);
  }
}

template <typename Number>
__device__ __forceinline__ void tanh_sech2(Number x, Number &t, Number &s2)
{
  const Number ax = fabs(x);
  const Number cut = sizeof(Number) == 8 ? Number(20) : Number(10);
  if (ax < cut) {
    const Number e = exp_neg2(ax);
    const Number inv = fast_rcp(Number(1) + e);
    t = (Number(1) - e) * inv;
    s2 = Number(4) * e * inv * inv;
    if (x < 0) t = -t;
  } else {
    t = x < 0 ? Number(-1) : Number(1);
    s2 = 0;
  }
}

// Mushy-zone law (MaterialTypes::liquid_solid with the mushy_zone transition): solid fraction
// f = clamp((T_liq - T) / (T_liq - T_sol), 0, 1) (compute_temperature_dependent_solid_fraction,
// material.templates.hpp:748-760); each property l + (s - l) H(f), H = 3 f^2 - 2 f^3, exactly l
// at f = 0 and s at f = 1, with d/dT = -(s - l)(6 f - 6 f^2) / (T_liq - T_sol) inside the
// mushy interval and 0 at its ends (material.templates.hpp:546-600); rho cp = rho * cp and
// d(rho cp)/dT = cp' rho + rho' cp (heat_diffuse_operator.cpp:1177-1190).
template <typename Number>
__device__ __forceinline__ void mushy_law(const CellArgs<Number> &A, Number T, Number &k,
                                          Number &dk, Number &rc, Number &drc)
{
  Number f = (A.mu_Tliq - T) * A.mu_inv;
  f = f > Number(1) ? Number(1) : f;
  f = f < Number(0) ? Number(0) : f;
  const bool liquid = f == Number(0), solid = f == Number(1);
  const Number H = f * f * (Number(3) - Number(2) * f);
  const Number dH = (liquid || solid) ? Number(0) : -A.mu_inv * (Number(6) * f - Number(6) * f * f);
  auto blend = [&](const Number (&v)[2]) {
    return liquid ? v[0] : (solid ? v[1] : v[0] + (v[1] - v[0]) * H);
  };
  k = blend(A.mu_k);
  dk = (A.mu_k[1] - A.mu_k[0]) * dH;
  const Number rho = blend(A.mu_rho), cp = blend(A.mu_cp);
  rc = rho * cp;
  drc = (A.mu_cp[1] - A.mu_cp[0]) * dH * rho + (A.mu_rho[1] - A.mu_rho[0]) * dH * cp;
}

template <typename Number>
__device__ __forceinline__ Number block_reduce_sum(Number v, Number *red)
{
  // wave reduction with shuffles, then across waves through LDS
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  Number s = 0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) / 64;
    for (int i = 0; i < nw; ++i) s += red[i];
  }
  return s;
}

// One colour launch of the cell kernel.
struct LaunchPart {
  int colour;     // structured meshes: row colour 0..3
  int part_off;   // offset of this launch's blocks in dot_partial
  int64_t start;  // adaptive meshes: first list entry (cell or slot) of the launch
  int64_t count;  // number of list entries
};

// Node access of the structured mesh.  A workgroup walks one row of cells along x in steps of
// CPB cells; thread (c, p) = c + CPB*p serves the x pencil p = j + K*k of cell c of the step,
// i.e. the nodes (cx*P + i, cy*P + j, cz*P + k), i = 0..K-1, which are contiguous in memory.
// The pencil loads (source, linearisation field and Dirichlet mask) of the next step are issued
// while the current step computes, so that a step does not wait for a full memory round trip.
template <typename Number, int K, int CPB>
struct GridAccess {
  static constexpr bool COOP = false;
  static constexpr bool RAW = true;  // gathers through issue / take
  static constexpr bool PREFETCH = prefetch_pencils<Number, K>();
  static constexpr int P = K - 1;
  uint32_t rowbase, gbase;
  int c, ncx, cellrow, nsub;
  int cx, cell;
  const uint8_t *mask;  // Dirichlet mask (null when not applied)
  uint32_t mbits;       // bit i: node i of the pencil is constrained (this step)
  bool active;
  bool first;    // this colour is the first to write the pencil's nodes
  Number carry;  // output of the previous step's last cell at the node it shares with cell 0

  __device__ static void map(int t, int &c_, int &p_)
  {
    c_ = t % CPB;
    p_ = t / CPB;
  }

  __device__ GridAccess(const CellArgs<Number> &A, const LaunchPart &lp, int c_, int p)
  {
    const RowColour rc = row_colour(A.ncell, lp.colour);
    const int cy = 2 * (int(blockIdx.x) % rc.hy) + rc.yb;
    const int cz = 2 * (int(blockIdx.x) / rc.hy) + rc.zb;
    c = c_;
    ncx = A.ncell[0];
    nsub = (ncx + CPB - 1) / CPB;
    rowbase = A.NX * (uint32_t(cy * P + p % K) + A.NY * uint32_t(cz * P + p / K));
    cellrow = ncx * (cy + A.ncell[1] * cz);
    // A node on the face between rows cy and cy+1 (or layers cz, cz+1) is written by both
    // parities; the launch order is colour 0..3 = (y parity) + 2 (z parity), so the first
    // writer of a shared node is the even one.
    const int j = p % K, k = p / K;
    const bool shared_y = (j == 0 && cy > 0) || (j == P && cy < A.ncell[1] - 1);
    const bool shared_z = (k == 0 && cz > 0) || (k == P && cz < A.ncell[2] - 1);
    first = (!shared_y || rc.yb == 0) && (!shared_z || rc.zb == 0);
    carry = Number(0);
    mask = (A.use_mask && A.mask) ? A.mask : nullptr;
    at(0);
  }
  __device__ int iterations() const { return nsub; }
  // cell size relative to the mesh's base cell
  __device__ Number hscale() const { return Number(1); }
  __device__ void at(int it)
  {
    cx = it * CPB + c;
    active = cx < ncx;
    cell = cellrow + cx;
    gbase = rowbase + uint32_t(cx * P);
    // the mask bits of the step come with the raw pencil (take)
  }
  // entry i with the constraints applied (read_dof_values: Dirichlet entries read as zero)
  __device__ Number value(const CellArgs<Number> &, const Number *f, int i) const
  {
    const Number v = f[(active ? gbase : 0u) + uint32_t(i)];
    return (active && !((mbits >> i) & 1u)) ? v : Number(0);
  }
  // entry i of field f as stored (read_dof_values_plain), zero past the row end
  __device__ Number plain(const Number *f, int i) const
  {
    const Number v = f[(active ? gbase : 0u) + uint32_t(i)];
    return active ? v : Number(0);
  }
  // Raw pencil of step `it` (loaded at an in-bounds index for every lane: node i of the row
  // start when the cell is past the end of the row); nothing is selected here, so the loads of
  // a step go out together.
  struct Raw {
    Number u[K], t[K];
    uint32_t mb;
  };
  template <bool LU, bool LT>
  __device__ void issue(int it, const Number *fu, const Number *ft, Raw &r) const
  {
    const int cxn = it * CPB + c;
    const uint32_t g = cxn < ncx ? rowbase + uint32_t(cxn * P) : 0u;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if constexpr (LU) r.u[i] = fu[g + i];
      if constexpr (LT) r.t[i] = ft[g + i];
    }
    r.mb = 0u;
    if (mask) {
#pragma unroll
      for (int i = 0; i < K; ++i) r.mb |= uint32_t(mask[g + i] != 0) << i;
    }
  }
  // take the raw pencil of the current step: constrained entries read as zero when `cons`
  // (read_dof_values), as stored otherwise (read_dof_values_plain); zero past the row end
  template <bool LU, bool LT>
  __device__ void take(const Raw &r, bool cons, Number (&u)[K], Number (&t)[K])
  {
    mbits = r.mb;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if constexpr (LU) u[i] = (active && !(cons && ((mbits >> i) & 1u))) ? r.u[i] : Number(0);
      if constexpr (LT) t[i] = active ? r.t[i] : Number(0);
    }
  }
  // current dst entries of the pencil (not needed where this launch is the first writer of an
  // overwriting application); rows of one colour own their nodes during the launch, so these
  // can be read before the cell is computed
  __device__ void prefetch(const CellArgs<Number> &A, const Number *dst, Number (&d)[K]) const
  {
    const bool need = active && !(A.overwrite && first);
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = need ? dst[gbase + i] : Number(0);
  }
  // Add the K outputs of the x pencil into dst.  Cell c-1 of the step sits in the previous
  // lane, so its last node (shared with this cell's first) is fetched with a shuffle and added
  // here; cell 0 takes it from the previous step's last cell.  A cell writes nodes 0..K-2 and,
  // at the end of the row, node K-1.  Constrained rows are skipped
  // (distribute_local_to_global semantics); an overwriting application (dst = A src) stores
  // where it is the first writer and writes zeros on constrained rows there.
  __device__ void scatter(const CellArgs<Number> &A, Number (&o)[K], const Number (&d)[K], int,
                          Number *dst, Number *)
  {
    if (!active) {
#pragma unroll
      for (int i = 0; i < K; ++i) o[i] = Number(0);
    }
    const Number prev = __shfl_up(o[K - 1], 1, CPB);
    const Number wrap = __shfl(o[K - 1], CPB - 1, CPB);
    o[0] += (c > 0) ? prev : carry;
    carry = wrap;
    if (!active) return;
    const bool last = (cx == ncx - 1);
    const bool store = A.overwrite && first;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (i == K - 1 && !last) break;
      if ((mbits >> i) & 1u) {
        if (store) dst[gbase + i] = Number(0);  // constrained rows of dst = A src are zero
        continue;
      }
      dst[gbase + i] = store ? o[i] : d[i] + o[i];
    }
  }
};

// Node access of an adaptive mesh: a workgroup serves CPB cells of one colour from the cell
// list (cells of one colour share no node).  Node values are gathered cooperatively: the
// block's CPB*K^3 (cell, local node) slots are spread over its threads so that consecutive
// lanes read the nodes of one cell (whose numbers are close: nodes are numbered in order of
// first appearance over the Morton-ordered cells), and staged in LDS in the x-pencil layout.
// A hanging node reads the interpolation of its masters (read_dof_values with hanging-node
// constraints), evaluated for all hanging nodes by a pass before the cell launches.
template <typename Number, int K, int CPB>
struct ListAccess {
  static constexpr bool COOP = true;
  static constexpr bool RAW = false;
  static constexpr bool PREFETCH = false;
  struct Raw {};
  __device__ static void map(int t, int &c, int &p)
  {
    c = t % CPB;
    p = t / CPB;
  }
  static constexpr int NLOC = K * K * K;
  int cell, level;     // the cell of the thread's pencils
  bool active;
  uint32_t node[K];    // global node of each cooperative slot of the thread
  uint32_t hbits = 0;  // slot i is a hanging node
  uint32_t abits = 0;  // slot i belongs to an existing cell

  __device__ ListAccess(const CellArgs<Number> &A, const LaunchPart &lp, int c, int)
  {
    const int64_t slot = int64_t(blockIdx.x) * CPB + c;
    active = slot < lp.count;
    cell = active ? int(A.list[lp.start + slot]) : 0;
    level = active ? int(A.cell_level[cell]) : 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int j = int(threadIdx.x) + i * int(blockDim.x);
      const int64_t cs = int64_t(blockIdx.x) * CPB + j / NLOC;
      const int e = j % NLOC;
      node[i] = 0u;
      if (cs < lp.count) {
        const uint32_t g = A.list[lp.start + cs];
        node[i] = A.cell_nodes[size_t(g) * NLOC + e];
        abits |= 1u << i;
        if ((A.cell_hmask[g] >> e) & 1ull) hbits |= 1u << i;
      }
    }
  }
  __device__ int iterations() const { return 1; }
  __device__ void at(int) {}
  __device__ Number hscale() const { return ldexp(Number(1), -level); }
  // LDS position ([local node][cell] layout) of cooperative slot i
  __device__ int coop_pos(int i) const
  {
    const int j = int(threadIdx.x) + i * int(blockDim.x);
    return Layout<Number, K, CPB>::idx(j % NLOC, j / NLOC);
  }
  __device__ bool masked(const CellArgs<Number> &A, uint32_t n) const
  {
    return A.use_mask && A.mask && A.mask[n];
  }
  // Values of field f at the thread's slots.  The value, mask and constraint-line loads of all
  // slots are issued before any is used (unused slots read node 0); only a hanging node's
  // interpolated value waits for its line index.
  __device__ void coop_load(const CellArgs<Number> &A, const Number *f, bool cons,
                            Number (&v)[K]) const
  {
    if (!f) {
#pragma unroll
      for (int i = 0; i < K; ++i) v[i] = Number(0);
      return;
    }
    Number raw[K];
#pragma unroll
    for (int i = 0; i < K; ++i) raw[i] = f[node[i]];
    if (!cons) {
#pragma unroll
      for (int i = 0; i < K; ++i) v[i] = ((abits >> i) & 1u) ? raw[i] : Number(0);
      return;
    }
    const bool um = A.use_mask && A.mask;
    uint32_t mb = 0u;
    int32_t line[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (um) mb |= uint32_t(A.mask[node[i]] != 0) << i;
      line[i] = ((hbits >> i) & 1u) ? A.node_line[node[i]] : -1;
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
      // hanging node: its interpolated value, computed for the whole source beforehand
      const Number h = line[i] >= 0 ? A.hang_val[line[i]] : Number(0);
      const bool keep = ((abits >> i) & 1u) && !((mb >> i) & 1u);
      v[i] = !((abits >> i) & 1u) ? Number(0) : (line[i] >= 0 ? h : (keep ? raw[i] : Number(0)));
    }
  }
  __device__ void coop_store(const Number (&v)[K], Number *T0, Number *T1) const
  {
#pragma unroll
    for (int i = 0; i < K; ++i) {
      T0[coop_pos(i)] = v[i];
      if (T1) T1[coop_pos(i)] = v[i];
    }
  }
  __device__ void prefetch(const CellArgs<Number> &, const Number *dst, Number (&d)[K]) const
  {
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = ((abits >> i) & 1u) ? dst[node[i]] : Number(0);
  }
  // Every cell adds into its own nodes, hanging ones included (cells of one colour share no
  // node); the hanging rows are moved to their masters by the condensation that follows the
  // colour launches.  Dirichlet rows are skipped.  The outputs go through LDS back to the
  // cooperative slots.
  __device__ void scatter(const CellArgs<Number> &A, Number (&o)[K], const Number (&d)[K], int c,
                          Number *dst, Number *stage)
  {
    const int p = int(threadIdx.x) / CPB;
    __syncthreads();  // the stage may still be read by other threads
    store_pencil<Number, K, Layout<Number, K, CPB>, 0>(stage, p, c, o);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (!((abits >> i) & 1u)) continue;
      const uint32_t n = node[i];
      if (!masked(A, n)) dst[n] = d[i] + stage[coop_pos(i)];
    }
  }
};

// Condensed diagonal on cells with hanging nodes: a slot is (cell, target node D).  The local
// source is the coefficient vector c_D of D in the cell's constrained basis (1 on the free
// local node D, the weight of D on each hanging local node with master D); the slot adds
// c_D . (A_cell c_D) to diag[D] (MatrixFreeTools::compute_diagonal with constraints).  Slots of
// different cells may share D, so the addition is atomic (this runs once per diagonal).
template <typename Number, int K, int CPB>
struct SlotAccess {
  static constexpr bool COOP = false;
  static constexpr bool RAW = false;
  static constexpr bool PREFETCH = false;
  struct Raw {};
  __device__ static void map(int t, int &c, int &p)
  {
    c = t % CPB;
    p = t / CPB;
  }
  static constexpr int NLOC = K * K * K;
  const uint32_t *nodes;
  uint32_t tgt;
  int c, p, cell, level;
  bool active;
  Number cv[K];

  __device__ SlotAccess(const CellArgs<Number> &A, const LaunchPart &lp, int c_, int p_)
  {
    c = c_;
    p = p_;
    const int64_t slot = int64_t(blockIdx.x) * CPB + c;
    active = slot < lp.count;
    cell = active ? int(A.slot_cell[lp.start + slot]) : 0;
    tgt = active ? A.slot_tgt[lp.start + slot] : 0u;
    nodes = A.cell_nodes + size_t(cell) * NLOC + K * p;
    level = active ? int(A.cell_level[cell]) : 0;
#pragma unroll
    for (int i = 0; i < K; ++i) cv[i] = Number(0);
    if (active) {
      for (uint32_t q = A.tgt_ptr[tgt]; q < A.tgt_ptr[tgt + 1]; ++q) {
        const int e = int(A.ent_local[q]);
        if (e / K == p) {
#pragma unroll
          for (int i = 0; i < K; ++i)
            if (e % K == i) cv[i] += A.ent_w[q];
        }
      }
    }
  }
  __device__ int iterations() const { return 1; }
  __device__ void at(int) {}
  __device__ Number hscale() const { return ldexp(Number(1), -level); }
  __device__ Number plain(const Number *f, int i) const { return active ? f[nodes[i]] : Number(0); }
  __device__ Number value(const CellArgs<Number> &, const Number *, int i) const { return cv[i]; }
  __device__ void prefetch(const CellArgs<Number> &, const Number *, Number (&d)[K]) const
  {
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = Number(0);
  }
  __device__ void scatter(const CellArgs<Number> &A, Number (&o)[K], const Number (&)[K], int,
                          Number *dst, Number *stage)
  {
    Number part = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) part += cv[i] * o[i];
    __syncthreads();
    stage[p * CPB + c] = part;
    __syncthreads();
    if (p == 0 && active) {
      Number sum = 0;
      for (int j = 0; j < K * K; ++j) sum += stage[j * CPB + c];
      const uint32_t n = A.tgt_node[tgt];
      if (!(A.use_mask && A.mask && A.mask[n])) atomicAdd(dst + n, sum);
    }
  }
};

constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <typename Number, int K, int OP, int NEEDT, bool VEL, bool DOT, class Access>
__global__ void __launch_bounds__((cells_per_block<Number, K>() * K * K))
    __attribute__((amdgpu_waves_per_eu(min_waves_per_eu<Number, K, VEL, OP>())))
    cell_kernel(const CellArgs<Number> A0, const Number *__restrict__ src, Number *__restrict__ dst,
                const LaunchPart lp)
{
  constexpr int CPB = cells_per_block<Number, K>();
  constexpr int NL = K * K * K;
  constexpr int BUF = NL * CPB;
  constexpr bool LS = (OP == LS_TANGENT || OP == LS_DIAG);
  constexpr bool REINIT = (OP == REINIT_TANGENT || OP == REINIT_DIAG);
  constexpr bool IS_DIAG = (OP == HEAT_DIAG || OP == ADVEC_DIAG || OP == LS_DIAG || OP == REINIT_DIAG);
  constexpr bool HAS_U = !IS_DIAG;
  constexpr bool PGRAD = (OP == HEAT_PGRAD);
  // T: the linearisation temperature, or for the level-set projections the level set whose
  // values give the narrow band (NEEDT = band enabled); only the heat operator needs grad T
  constexpr bool USE_T = (NEEDT && (OP == HEAT_TANGENT || OP == HEAT_DIAG || LS)) || PGRAD;
  constexpr bool TGRAD = USE_T && !LS;
  constexpr bool RHS_HEAT = (OP == HEAT_RHS);
  constexpr bool TANGENT = is_tangent_op(OP);
  // mushy-zone law: rho cp varies with T; the tangent and its diagonal read the old temperature
  // for the d(rho cp)/dT (T - T_old)/dt term
  constexpr bool MUSHY = NEEDT == 2 && (OP == HEAT_TANGENT || OP == HEAT_DIAG || OP == HEAT_RHS);
  constexpr bool TOLD = MUSHY && !RHS_HEAT;
  // advection with SUPG (NEEDT set): a per-cell tau from max_q |w|, and the Laplacian of u at
  // the quadrature points (tangent, rhs) or of the basis functions (diagonal)
  constexpr bool SUPG = NEEDT && (OP == ADVEC_TANGENT || OP == ADVEC_RHS || OP == ADVEC_DIAG);

  // LDS cell tensors [K^3][CPB].  While values are interpolated: u in L0, T in L1, the velocity
  // components and the rhs fields after it.  Then the x, y gradients of T take L2, L3 and those
  // of u L0, L1 (with T) or L1, L2 (without).  The diagonal needs five coefficient tensors.
  constexpr int I_T = 1;
  constexpr int I_W = 1 + (USE_T ? 1 : 0);
  constexpr int I_A = I_W + (VEL ? 3 : 0);
  constexpr int I_O = I_A + (RHS_HEAT ? 2 : 0);
  constexpr int N_EVAL = I_O + (TOLD ? 1 : 0);
  constexpr int N_GRAD = TGRAD ? 4 : 3;
  // with the fused dot, the gathered source is kept: in registers (FP32 x pencils), or in one
  // more tensor when it was gathered cooperatively or when K more registers would cost the FP64
  // kernel a wave per SIMD (it sits just under 128)
  constexpr bool FDOT = DOT && TANGENT;
  constexpr bool FDOT_LDS = FDOT && (Access::COOP || sizeof(Number) == 8);
  constexpr int I_SRC = cmax(cmax(N_EVAL, N_GRAD), IS_DIAG ? 5 : 0);
  constexpr int NBUF = I_SRC + (FDOT_LDS ? 1 : 0);
  __shared__ Number lds[NBUF * BUF];
  Number *const L0 = lds;
  Number *const L1 = lds + BUF;
  Number *const L2 = lds + 2 * BUF;

  if (A0.stop && *A0.stop) return;  // uniform over the grid

  using LY = Layout<Number, K, CPB>;
  int c, p;
  Access::map(int(threadIdx.x), c, p);
  const int qi = p % K, qj = p / K;
  Access ac(A0, lp, c, p);
  double dot_acc = 0;
  // The arguments are re-read from the kernel-argument segment in every step (the empty asm
  // hides the pointer from loop-invariant hoisting), so the 1D matrices and coefficients live
  // in scalar registers only while a step uses them.
  typedef __attribute__((address_space(4))) const CellArgs<Number> KernArgs;
  KernArgs *kargs = (KernArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  // MFMA operands are built once and stay in registers
  using MH = typename SweepMat<Number, K>::type;
  const MH S0(A0.S), Dq0(A0.Dq);
  // product of the 1D Gauss weights of the thread's z pencil (qi, qj)
  const Number wqq = select_k<Number, K>(A0.w, qi) * select_k<Number, K>(A0.w, qj);
  // pencils of the next step, in flight while the current step computes
  constexpr bool PF = Access::PREFETCH;
  typename Access::Raw raw;
  if constexpr (PF) ac.template issue<HAS_U, USE_T>(0, RHS_HEAT ? A0.T : src, A0.T, raw);

  for (int it = 0; it < ac.iterations(); ++it) {
  if (it > 0) {
    __syncthreads();  // LDS of the previous step fully consumed
    ac.at(it);
  }
  asm volatile("" : "+s"(kargs));
  const CellArgs<Number> &A = *(const CellArgs<Number> *)kargs;
  const MH S = step_mat<Number, K>(S0, A.S);
  const MH Dq = step_mat<Number, K>(Dq0, A.Dq);
  const Number hs = ac.hscale();
  const Number ih[3] = {A.inv_h[0] / hs, A.inv_h[1] / hs, A.inv_h[2] / hs};

  // ---------------- E1: x pencils, gather from HBM ----------------
  Number dpre[K];  // dst entries of the pencil, read while the integration sweeps run
  [[maybe_unused]] Number usrc[K];  // fused dot: the gathered source x pencil
  if constexpr (Access::COOP) {
    // cooperative gather into the LDS tensors (all fields' loads in flight together), then the
    // x interpolation in place
    Number gu[K], gt[K];
    if constexpr (HAS_U)
      ac.coop_load(A, RHS_HEAT ? A.T : src, TANGENT && !A.gather_plain, gu);
    if constexpr (USE_T) ac.coop_load(A, A.T, false, gt);
    if constexpr (HAS_U) ac.coop_store(gu, L0, FDOT_LDS ? lds + I_SRC * BUF : nullptr);
    if constexpr (USE_T) ac.coop_store(gt, lds + I_T * BUF, nullptr);
    if constexpr (VEL) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        ac.coop_load(A, A.vel + size_t(d) * A.n_dofs, false, gu);
        ac.coop_store(gu, lds + (I_W + d) * BUF, nullptr);
      }
    }
    if constexpr (RHS_HEAT) {
      ac.coop_load(A, src, false, gu);  // T_old
      ac.coop_load(A, A.aux1, false, gt);
      ac.coop_store(gu, lds + I_A * BUF, nullptr);
      ac.coop_store(gt, lds + (I_A + 1) * BUF, nullptr);
    }
    if constexpr (TOLD) {
      ac.coop_load(A, A.T_old, false, gu);
      ac.coop_store(gu, lds + I_O * BUF, nullptr);
    }
    __syncthreads();
    if constexpr (HAS_U) sweep_inplace<Number, K, LY, 0, false>(S, L0, p, c);
    if constexpr (USE_T) sweep_inplace<Number, K, LY, 0, false>(S, lds + I_T * BUF, p, c);
    if constexpr (VEL) {
#pragma unroll
      for (int d = 0; d < 3; ++d)
        sweep_inplace<Number, K, LY, 0, false>(S, lds + (I_W + d) * BUF, p, c);
    }
    if constexpr (RHS_HEAT) {
      sweep_inplace<Number, K, LY, 0, false>(S, lds + I_A * BUF, p, c);
      sweep_inplace<Number, K, LY, 0, false>(S, lds + (I_A + 1) * BUF, p, c);
    }
    if constexpr (TOLD) sweep_inplace<Number, K, LY, 0, false>(S, lds + I_O * BUF, p, c);
  } else {
  {
    Number o[K];
    const Number *in = RHS_HEAT ? A.T : src;
    const bool cons = TANGENT && !A.gather_plain;
    Number u_loc[K], t_loc[K];
    if constexpr (PF) {
      ac.template take<HAS_U, USE_T>(raw, cons, u_loc, t_loc);
      if (it + 1 < ac.iterations()) ac.template issue<HAS_U, USE_T>(it + 1, in, A.T, raw);
    } else if constexpr (Access::RAW) {
      // all loads of the step (values and mask bits) go out together, selected afterwards
      typename Access::Raw rw;
      ac.template issue<HAS_U, USE_T>(it, in, A.T, rw);
      ac.template take<HAS_U, USE_T>(rw, cons, u_loc, t_loc);
    } else {
#pragma unroll
      for (int i = 0; i < K; ++i) {
        if constexpr (HAS_U) u_loc[i] = cons ? ac.value(A, in, i) : ac.plain(in, i);
        if constexpr (USE_T) t_loc[i] = ac.plain(A.T, i);
      }
    }
    if constexpr (HAS_U) {
      matvec<Number, K, false>(S, u_loc, o);
      store_pencil<Number, K, LY, 0>(L0, p, c, o);
      if constexpr (FDOT_LDS) {
        store_pencil<Number, K, LY, 0>(lds + I_SRC * BUF, p, c, u_loc);
      } else if constexpr (FDOT) {
#pragma unroll
        for (int i = 0; i < K; ++i) usrc[i] = u_loc[i];
      }
    }
    if constexpr (USE_T) {
      matvec<Number, K, false>(S, t_loc, o);
      store_pencil<Number, K, LY, 0>(lds + I_T * BUF, p, c, o);
    }
    if constexpr (VEL) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        Number v[K];
#pragma unroll
        for (int i = 0; i < K; ++i) v[i] = ac.plain(A.vel + size_t(d) * A.n_dofs, i);
        matvec<Number, K, false>(S, v, o);
        store_pencil<Number, K, LY, 0>(lds + (I_W + d) * BUF, p, c, o);
      }
    }
    if constexpr (RHS_HEAT) {
      Number v[K];
#pragma unroll
      for (int i = 0; i < K; ++i) v[i] = ac.plain(src, i);  // T_old
      matvec<Number, K, false>(S, v, o);
      store_pencil<Number, K, LY, 0>(lds + I_A * BUF, p, c, o);
#pragma unroll
      for (int i = 0; i < K; ++i) v[i] = A.aux1 ? ac.plain(A.aux1, i) : Number(0);
      matvec<Number, K, false>(S, v, o);
      store_pencil<Number, K, LY, 0>(lds + (I_A + 1) * BUF, p, c, o);
    }
    if constexpr (TOLD) {
      Number v[K];
#pragma unroll
      for (int i = 0; i < K; ++i) v[i] = A.T_old ? ac.plain(A.T_old, i) : Number(0);
      matvec<Number, K, false>(S, v, o);
      store_pencil<Number, K, LY, 0>(lds + I_O * BUF, p, c, o);
    }
  }
  }
  __syncthreads();

  // ---------------- E2: y pencils ----------------
  if constexpr (HAS_U && USE_T) {
    sweep_pair<Number, K, LY, 1, false, 1, false>(S, L0, L0, S, lds + I_T * BUF, lds + I_T * BUF,
                                                  p, c);
  } else {
    if constexpr (HAS_U) sweep_inplace<Number, K, LY, 1, false>(S, L0, p, c);
    if constexpr (USE_T) sweep_inplace<Number, K, LY, 1, false>(S, lds + I_T * BUF, p, c);
  }
  if constexpr (VEL) {
#pragma unroll
    for (int d = 0; d < 3; ++d) sweep_inplace<Number, K, LY, 1, false>(S, lds + (I_W + d) * BUF, p, c);
  }
  if constexpr (RHS_HEAT) {
    sweep_inplace<Number, K, LY, 1, false>(S, lds + I_A * BUF, p, c);
    sweep_inplace<Number, K, LY, 1, false>(S, lds + (I_A + 1) * BUF, p, c);
  }
  if constexpr (TOLD) sweep_inplace<Number, K, LY, 1, false>(S, lds + I_O * BUF, p, c);
  __syncthreads();

  // ---------------- E3: z pencils (values at q and d/dz kept in registers) ----------------
  Number uq[K], guz[K], Tq[K], gTz[K], wq[3][K], a1q[K], a2q[K];
  [[maybe_unused]] Number Toq[K];
  {
    Number vu[K], vt[K];
    if constexpr (HAS_U) load_pencil<Number, K, LY, 2>(L0, p, c, vu);
    if constexpr (USE_T) load_pencil<Number, K, LY, 2>(lds + I_T * BUF, p, c, vt);
    if constexpr (HAS_U) {
      matvec<Number, K, false>(S, vu, uq);
      matvec<Number, K, false>(Dq, uq, guz);
    }
    if constexpr (USE_T) {
      matvec<Number, K, false>(S, vt, Tq);
      matvec<Number, K, false>(Dq, Tq, gTz);
    }
    if constexpr (HAS_U) store_pencil<Number, K, LY, 2>(L0, p, c, uq);
    if constexpr (USE_T) store_pencil<Number, K, LY, 2>(lds + I_T * BUF, p, c, Tq);
  }
  if constexpr (VEL) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      Number v[K];
      load_pencil<Number, K, LY, 2>(lds + (I_W + d) * BUF, p, c, v);
      matvec<Number, K, false>(S, v, wq[d]);
    }
  }
  if constexpr (RHS_HEAT) {
    Number v[K];
    load_pencil<Number, K, LY, 2>(lds + I_A * BUF, p, c, v);
    matvec<Number, K, false>(S, v, a1q);
    load_pencil<Number, K, LY, 2>(lds + (I_A + 1) * BUF, p, c, v);
    matvec<Number, K, false>(S, v, a2q);
  }
  if constexpr (TOLD) {
    Number v[K];
    load_pencil<Number, K, LY, 2>(lds + I_O * BUF, p, c, v);
    matvec<Number, K, false>(S, v, Toq);
  }
  __syncthreads();

  // SUPG parameter of the cell: the maximum of |w|^2 over its quadrature points, reduced over
  // the cell's K^2 z pencils through the (free) velocity tensor
  [[maybe_unused]] Number tau = 0;
  if constexpr (SUPG) {
    Number m2 = 0;
#pragma unroll
    for (int z = 0; z < K; ++z) {
      const Number n2 = wq[0][z] * wq[0][z] + wq[1][z] * wq[1][z] + wq[2][z] * wq[2][z];
      m2 = n2 > m2 ? n2 : m2;
    }
    Number *red = lds + I_W * BUF;
    red[LY::idx(p, c)] = m2;
    __syncthreads();
    Number mx = 0;
    for (int i = 0; i < K * K; ++i) {
      const Number v = red[LY::idx(i, c)];
      mx = v > mx ? v : mx;
    }
    __syncthreads();
    mx = sqrt(mx);
    const Number coef = A.supg_coef < Number(0) ? (mx >= Number(1e-6) ? Number(1) / mx : Number(0))
                                                : A.supg_coef;
    tau = A.supg_cell_size * hs * coef;
  }

  const Number kc = (A.k_cell && ac.active) ? A.k_cell[ac.cell] : Number(1);

  // ---------------- x and y gradients of T and u at the quadrature points ----------------
  // With T and u (FUSED): the x and y pencils of T (L1) and of u (L0) are read, and after a
  // barrier the T gradients go to L2, L3 and the u gradients to L0 (x, over the u values, which
  // stay in registers from E3) and L1 (y, over T, whose values also stay in registers): one
  // barrier less than two separate sweep phases, and the per-point T coefficients are formed in
  // the quadrature loop from these tensors instead of being held in registers across phases.
  constexpr bool FUSED = TGRAD && HAS_U;
  Number *const UGX = FUSED ? L0 : L1;  // tensors of the u gradients
  Number *const UGY = FUSED ? L1 : L2;
  if constexpr (FUSED) {
    Number tx[K], ty[K], ux[K], uy[K], o[K];
    load_pencil<Number, K, LY, 0>(lds + I_T * BUF, p, c, tx);
    load_pencil<Number, K, LY, 1>(lds + I_T * BUF, p, c, ty);
    load_pencil<Number, K, LY, 0>(L0, p, c, ux);
    load_pencil<Number, K, LY, 1>(L0, p, c, uy);
    __syncthreads();
    matvec<Number, K, false>(Dq, tx, o);
    store_pencil<Number, K, LY, 0>(L2, p, c, o);
    matvec<Number, K, false>(Dq, ty, o);
    store_pencil<Number, K, LY, 1>(lds + 3 * BUF, p, c, o);
    matvec<Number, K, false>(Dq, ux, o);
    store_pencil<Number, K, LY, 0>(L0, p, c, o);
    matvec<Number, K, false>(Dq, uy, o);
    store_pencil<Number, K, LY, 1>(L1, p, c, o);
    __syncthreads();
  } else if constexpr (TGRAD) {
    // the diagonal: T gradients only
    sweep_pair<Number, K, LY, 0, false, 1, false>(Dq, lds + I_T * BUF, L2, Dq, lds + I_T * BUF,
                                                  lds + 3 * BUF, p, c);
    __syncthreads();
  } else if constexpr (HAS_U) {
    // ---------------- E4: d/dx and d/dy of u at the quadrature points ----------------
    sweep_pair<Number, K, LY, 0, false, 1, false>(Dq, L0, L1, Dq, L0, L2, p, c);
    __syncthreads();
  }
  // SUPG: d2u/dx2 + d2u/dy2 (scaled) in L3 by the collocation derivative of the first
  // derivatives (exact for Q_p); d2u/dz2 (unscaled) in registers
  [[maybe_unused]] Number uzz[K];
  if constexpr (SUPG && HAS_U) {
    Number *const L3 = lds + 3 * BUF;
    {
      Number v[K], o[K];
      load_pencil<Number, K, LY, 0>(L1, p, c, v);
      matvec<Number, K, false>(Dq, v, o);
#pragma unroll
      for (int i = 0; i < K; ++i) o[i] *= ih[0] * ih[0];
      store_pencil<Number, K, LY, 0>(L3, p, c, o);
    }
    matvec<Number, K, false>(Dq, guz, uzz);
    __syncthreads();
    {
      Number v[K], o[K], a[K];
      load_pencil<Number, K, LY, 1>(L2, p, c, v);
      matvec<Number, K, false>(Dq, v, o);
      load_pencil<Number, K, LY, 1>(L3, p, c, a);
#pragma unroll
      for (int i = 0; i < K; ++i) a[i] += o[i] * (ih[1] * ih[1]);
      store_pencil<Number, K, LY, 1>(L3, p, c, a);
    }
    __syncthreads();
  }

  // ---------------- Q: quadrature-point operation (z pencil (i,j), q = (i,j,z)) ----------------
  const Number wij = wqq * (A.h[0] * hs) * (A.h[1] * hs) * (A.h[2] * hs);
  Number rq[K], fz[K];
#pragma unroll
  for (int z = 0; z < K; ++z) {
    const int e = LY::idx(elem<K>(qi, qj, z), c);
    const Number jxw = wij * A.w[z];
    Number g[3] = {0, 0, 0};
    if constexpr (HAS_U) {
      g[0] = UGX[e] * ih[0];
      g[1] = UGY[e] * ih[1];
      g[2] = guz[z] * ih[2];
    }
    // T coefficients of the point (read before this iteration overwrites L2, L3 at e):
    // heat tangent / diagonal kq = k(T) kc and dkg = dk/dT kc grad T; PGRAD dkg = grad T; the
    // mushy law's rho cp(T) (rcq) and d(rho cp)/dT ((T - T_old)/dt + grad T . w) (cxq)
    [[maybe_unused]] Number kqz = 0, dkg[3] = {0, 0, 0}, rcq = 0, cxq = 0;
    if constexpr (TGRAD) {
      const Number g0 = L2[e] * ih[0], g1 = lds[3 * BUF + e] * ih[1], g2 = gTz[z] * ih[2];
      if constexpr (PGRAD) {
        dkg[0] = g0;
        dkg[1] = g1;
        dkg[2] = g2;
      } else if constexpr (MUSHY) {
        Number k, dk, rc, drc;
        mushy_law(A, Tq[z], k, dk, rc, drc);
        kqz = k * kc;
        dk *= kc;
        dkg[0] = dk * g0;
        dkg[1] = dk * g1;
        dkg[2] = dk * g2;
        rcq = rc;
        Number cx = drc * (Tq[z] - Toq[z]) * A.dt_inv;
        if constexpr (VEL) cx += drc * (g0 * wq[0][z] + g1 * wq[1][z] + g2 * wq[2][z]);
        cxq = A.t_dependent ? cx : Number(0);
      } else {
        Number t, s2;
        tanh_sech2((Tq[z] - A.tanh_Tc) * A.inv_tanh_w, t, s2);
        kqz = (A.k0 + A.tanh_amp * t) * kc;
        const Number dk = A.tanh_amp * A.inv_tanh_w * s2 * kc;
        dkg[0] = dk * g0;
        dkg[1] = dk * g1;
        dkg[2] = dk * g2;
      }
    }
    Number w3[3] = {0, 0, 0};
    if constexpr (VEL) {
      w3[0] = wq[0][z];
      w3[1] = wq[1][z];
      w3[2] = wq[2][z];
    }
    Number val = 0, f[3] = {0, 0, 0};
    Number cc = 0, kk = 0, bb[3] = {0, 0, 0};
    if constexpr (PGRAD) {
      val = g[0] * dkg[0] + g[1] * dkg[1] + g[2] * dkg[2];  // JxW applied below
    } else if constexpr (OP == HEAT_TANGENT || OP == HEAT_DIAG) {
      Number k = A.k0 * kc;
      if constexpr (USE_T) k = kqz;
      const bool td = USE_T && A.t_dependent;
      Number rho_cp = A.rho_cp;
      if constexpr (MUSHY) rho_cp = rcq;
      if constexpr (OP == HEAT_TANGENT) {
        const Number u = uq[z];
        val = A.dt_inv * rho_cp * u;
        // mushy law: d(rho cp)/dT ((T - T_old)/dt + grad T . w) u v (symmetric in u, v)
        if constexpr (MUSHY) val += cxq * u;
#pragma unroll
        for (int d = 0; d < 3; ++d) f[d] = k * g[d];
        if (!A.transpose) {
          // (rho cp w . grad u) v + dk/dT u (grad T . grad v)
          if constexpr (VEL) val += rho_cp * (g[0] * w3[0] + g[1] * w3[1] + g[2] * w3[2]);
          if constexpr (USE_T) {
            if (td) {
#pragma unroll
              for (int d = 0; d < 3; ++d) f[d] += dkg[d] * u;
            }
          }
        } else {
          // transposed: (rho cp w u) . grad v + dk/dT (grad T . grad u) v
          if constexpr (VEL) {
#pragma unroll
            for (int d = 0; d < 3; ++d) f[d] += rho_cp * w3[d] * u;
          }
          if constexpr (USE_T) {
            if (td) val += dkg[0] * g[0] + dkg[1] * g[1] + dkg[2] * g[2];
          }
        }
      } else {
        cc = A.dt_inv * rho_cp;
        if constexpr (MUSHY) cc += cxq;
        kk = k;
        if constexpr (VEL) {
#pragma unroll
          for (int d = 0; d < 3; ++d) bb[d] = rho_cp * w3[d];
        }
        if constexpr (USE_T) {
          if (td) {
#pragma unroll
            for (int d = 0; d < 3; ++d) bb[d] += dkg[d];
          }
        }
      }
    } else if constexpr (OP == HEAT_RHS) {
      // U holds T: values uq, gradient g
      Number k = A.k0, rho_cp = A.rho_cp;
      if constexpr (MUSHY) {
        Number dk, drc;
        mushy_law(A, uq[z], k, dk, rho_cp, drc);
      } else if constexpr (NEEDT) {
        Number t, s2;
        tanh_sech2((uq[z] - A.tanh_Tc) * A.inv_tanh_w, t, s2);
        k = A.k0 + A.tanh_amp * t;
      }
      k *= kc;
      Number v = A.dt_inv * rho_cp * (uq[z] - a1q[z]) - a2q[z];
      if constexpr (VEL) v += rho_cp * (g[0] * w3[0] + g[1] * w3[1] + g[2] * w3[2]);
      val = -v;
#pragma unroll
      for (int d = 0; d < 3; ++d) f[d] = -k * g[d];
    } else if constexpr (OP == ADVEC_TANGENT) {
      const Number vg = g[0] * w3[0] + g[1] * w3[1] + g[2] * w3[2];
      val = A.mass_factor * A.dt_inv * uq[z];
#pragma unroll
      for (int d = 0; d < 3; ++d) f[d] = A.theta * A.diffusivity * g[d];
      if (!A.transpose) {
        val += A.theta * vg;
      } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) f[d] += A.theta * w3[d] * uq[z];
      }
      if constexpr (SUPG) {
        // tau w (m u/dt + theta w . grad u - theta D lap u) against grad v (not transposed)
        const Number lap = lds[3 * BUF + e] + uzz[z] * (ih[2] * ih[2]);
        const Number res = val - A.theta * A.diffusivity * lap;
#pragma unroll
        for (int d = 0; d < 3; ++d) f[d] += tau * w3[d] * res;
      }
    } else if constexpr (OP == ADVEC_RHS) {
      const Number vg = g[0] * w3[0] + g[1] * w3[1] + g[2] * w3[2];
      val = A.dt_inv * uq[z] - (Number(1) - A.theta) * vg;
#pragma unroll
      for (int d = 0; d < 3; ++d) f[d] = -(Number(1) - A.theta) * A.diffusivity * g[d];
      if constexpr (SUPG) {
        // tau w (u_old/dt - (1-theta) w . grad u_old + (1-theta) D lap u_old)
        const Number lap = lds[3 * BUF + e] + uzz[z] * (ih[2] * ih[2]);
        const Number res = val + (Number(1) - A.theta) * A.diffusivity * lap;
#pragma unroll
        for (int d = 0; d < 3; ++d) f[d] += tau * w3[d] * res;
      }
    } else if constexpr (REINIT) {
      // u v / dt + (eps_t grad u + (eps_n - eps_t) (grad u . n) n) . grad v, n = the normal
      // vector (held in the velocity slot) normalised with the tolerance
      Number nv[3] = {w3[0], w3[1], w3[2]};
      const Number nrm = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
      const Number den = nrm > A.re_tol ? nrm : A.re_tol;
#pragma unroll
      for (int d = 0; d < 3; ++d) nv[d] /= den;
      const Number sc = A.re_scale_h ? hs : Number(1);
      const Number en = A.re_eps_n * sc, et = A.re_eps_t * sc;
      if constexpr (OP == REINIT_TANGENT) {
        val = A.dt_inv * uq[z];
        const Number gn = g[0] * nv[0] + g[1] * nv[1] + g[2] * nv[2];
#pragma unroll
        for (int d = 0; d < 3; ++d) f[d] = et * g[d] + (en - et) * gn * nv[d];
      } else {
        // the anisotropic part enters the diagonal as sign (b . grad phi)^2 with
        // b = sqrt(|eps_n - eps_t| JxW) n (stored without the JxW factor applied below)
        cc = A.dt_inv;
        kk = et;
        const Number sb = sqrt((en > et ? en - et : et - en) / jxw);
#pragma unroll
        for (int d = 0; d < 3; ++d) bb[d] = sb * nv[d];
      }
    } else if constexpr (LS) {
      // m (u v + eta grad u . grad v); m = 0 where |level set| >= threshold (a NaN keeps 1,
      // as the reference's test does)
      Number msk = 1;
      if constexpr (USE_T) {
        const Number at = Tq[z] < Number(0) ? -Tq[z] : Tq[z];
        msk = (at >= A.ls_threshold) ? Number(0) : Number(1);
      }
      const Number eta = A.ls_eta * hs * hs;
      if constexpr (OP == LS_TANGENT) {
        val = msk * uq[z];
#pragma unroll
        for (int d = 0; d < 3; ++d) f[d] = msk * eta * g[d];
      } else {
        cc = msk;
        kk = msk * eta;
      }
    } else {  // ADVEC_DIAG (SUPG: bb = w, the scalar factors applied in the diagonal loop)
      cc = A.mass_factor * A.dt_inv;
      kk = A.theta * A.diffusivity;
#pragma unroll
      for (int d = 0; d < 3; ++d) bb[d] = SUPG ? w3[d] : A.theta * w3[d];
    }
    if constexpr (IS_DIAG) {
      L0[e] = jxw * cc;
      L1[e] = jxw * kk;
      L2[e] = jxw * bb[0];
      lds[3 * BUF + e] = jxw * bb[1];
      lds[4 * BUF + e] = jxw * bb[2];
    } else {
      rq[z] = jxw * val;
      L1[e] = jxw * f[0] * ih[0];
      L2[e] = jxw * f[1] * ih[1];
      fz[z] = jxw * f[2] * ih[2];
    }
  }

  if constexpr (PGRAD) {
    // per-cell sum of the K^2 z-pencil partials (L0 is free after E4)
    Number part = 0;
#pragma unroll
    for (int z = 0; z < K; ++z) part += rq[z];
    L0[LY::idx(p, c)] = part;
    __syncthreads();
    if (p == 0 && ac.active) {
      double sum = 0;
      for (int i = 0; i < K * K; ++i) sum += double(L0[LY::idx(i, c)]);
      dst[ac.cell] += Number(A.pgrad_scale * sum);
    }
    continue;
  }

  if constexpr (!PGRAD) ac.prefetch(A, dst, dpre);

  if constexpr (IS_DIAG) {
    __syncthreads();
    // A_ii = sum_q [cc phi^2 + kk |grad phi|^2 + (bb . grad phi) phi] for the thread's x pencil
    const int jb = p % K, kb = p / K;
    Number diag[K];
#pragma unroll
    for (int ia = 0; ia < K; ++ia) diag[ia] = 0;
    // SUPG: (theta + tau m/dt)(w . grad phi) phi + tau theta (w . grad phi)^2
    // - tau theta D (w . grad phi) lap phi, with b = JxW w stored
    [[maybe_unused]] Number s_beta = 0, s_gam = 0, s_del = 0, s_vol = 0;
    if constexpr (SUPG) {
      s_beta = A.theta + tau * A.mass_factor * A.dt_inv;
      s_gam = tau * A.theta;
      s_del = -tau * A.theta * A.diffusivity;
      s_vol = (A.h[0] * hs) * (A.h[1] * hs) * (A.h[2] * hs);
    }
    for (int q3 = 0; q3 < K; ++q3)
      for (int q2 = 0; q2 < K; ++q2) {
        const Number sy = A.S[q2 * K + jb], dy = A.DS[q2 * K + jb] * ih[1];
        const Number sz = A.S[q3 * K + kb], dz = A.DS[q3 * K + kb] * ih[2];
#pragma unroll
        for (int q1 = 0; q1 < K; ++q1) {
          const int e = LY::idx(elem<K>(q1, q2, q3), c);
          const Number cq = L0[e], kq2 = L1[e], b0 = L2[e], b1 = lds[3 * BUF + e],
                       b2 = lds[4 * BUF + e];
#pragma unroll
          for (int ia = 0; ia < K; ++ia) {
            const Number sx = A.S[q1 * K + ia], dx = A.DS[q1 * K + ia] * ih[0];
            const Number phi = sx * sy * sz;
            const Number gx = dx * sy * sz, gy = sx * dy * sz, gz = sx * sy * dz;
            if constexpr (SUPG) {
              const Number jxw = A.w[q1] * A.w[q2] * A.w[q3] * s_vol;
              const Number lap = A.DDS[q1 * K + ia] * (ih[0] * ih[0]) * sy * sz +
                                 sx * A.DDS[q2 * K + jb] * (ih[1] * ih[1]) * sz +
                                 sx * sy * A.DDS[q3 * K + kb] * (ih[2] * ih[2]);
              const Number wg = b0 * gx + b1 * gy + b2 * gz;
              diag[ia] += cq * phi * phi + kq2 * (gx * gx + gy * gy + gz * gz) +
                          wg * (s_beta * phi + s_gam * wg / jxw + s_del * lap);
            } else if constexpr (OP == REINIT_DIAG) {
              const Number bg = b0 * gx + b1 * gy + b2 * gz;
              diag[ia] += cq * phi * phi + kq2 * (gx * gx + gy * gy + gz * gz) + A.re_sign * bg * bg;
            } else {
              diag[ia] += cq * phi * phi + kq2 * (gx * gx + gy * gy + gz * gz) +
                          (b0 * gx + b1 * gy + b2 * gz) * phi;
            }
          }
        }
      }
    ac.scatter(A, diag, dpre, c, dst, L0);
  } else {
    // r = val + Dq_z^T fz (z direction in registers)
    Number tz[K];
    matvec<Number, K, true>(Dq, fz, tz);
#pragma unroll
    for (int z = 0; z < K; ++z) rq[z] += tz[z];
    store_pencil<Number, K, LY, 2>(L0, p, c, rq);
    __syncthreads();

    // ---------------- I1: Dq^T along x (on fx) and y (on fy) ----------------
    sweep_pair<Number, K, LY, 0, true, 1, true>(Dq, L1, L1, Dq, L2, L2, p, c);
    __syncthreads();

    // ---------------- I2: sum and S^T along z ----------------
    {
      Number a[K], b[K], d[K], o[K];
      load_pencil<Number, K, LY, 2>(L0, p, c, a);
      load_pencil<Number, K, LY, 2>(L1, p, c, b);
      load_pencil<Number, K, LY, 2>(L2, p, c, d);
#pragma unroll
      for (int e = 0; e < K; ++e) a[e] += b[e] + d[e];
      matvec<Number, K, true>(S, a, o);
      store_pencil<Number, K, LY, 2>(L0, p, c, o);
    }
    __syncthreads();

    // ---------------- I3: S^T along y ----------------
    sweep_inplace<Number, K, LY, 1, true>(S, L0, p, c);
    __syncthreads();

    // ---------------- I4: S^T along x and scatter-add ----------------
    Number o[K];
    {
      Number v[K];
      load_pencil<Number, K, LY, 0>(L0, p, c, v);
      matvec<Number, K, true>(S, v, o);
    }
    if constexpr (FDOT) {
      // src . (A src) over the cell's nodes (constrained entries of the gathered source are 0),
      // before the x-shared node is combined across cells
      Number u[K];
      if constexpr (FDOT_LDS) {
        load_pencil<Number, K, LY, 0>(lds + I_SRC * BUF, p, c, u);
      } else {
#pragma unroll
        for (int i = 0; i < K; ++i) u[i] = usrc[i];
      }
      Number local_dot = 0;
#pragma unroll
      for (int i = 0; i < K; ++i) local_dot += u[i] * o[i];
      dot_acc += double(local_dot);
    }
    ac.scatter(A, o, dpre, c, dst, L0);
  }
  }  // steps along the row

  if constexpr (FDOT) {
    __syncthreads();  // L0 is reused for the wave partials
    const double s = block_reduce_sum<double>(dot_acc, reinterpret_cast<double *>(lds));
    if (threadIdx.x == 0) A0.dot_partial[lp.part_off + blockIdx.x] = s;
  }
}

static int row_blocks(const MeshInfo &m, int colour)
{
  const RowColour s = row_colour(m.ncell, colour);
  return s.hy * s.hz;
}

template <typename Number>
static int cell_blocks_t(const MeshInfo &m)
{
  constexpr int CPBS[5] = {0, 0, cells_per_block<Number, 2>(), cells_per_block<Number, 3>(),
                           cells_per_block<Number, 4>()};
  if (m.ad) {
    int total = 0;
    for (int col = 0; col < m.ad->n_colours; ++col)
      total += int((m.ad->colour_start[col + 1] - m.ad->colour_start[col] + CPBS[m.K] - 1) / CPBS[m.K]);
    return total;
  }
  int total = 0;
  for (int colour = 0; colour < 4; ++colour) total += row_blocks(m, colour);
  return total;
}

int cell_blocks(const MeshInfo &m, bool fp32)
{
  return fp32 ? cell_blocks_t<float>(m) : cell_blocks_t<double>(m);
}

template <typename Number, int K, int OP, int NEEDT, bool VEL, bool DOT>
static int launch_colours(const MeshInfo &m, const CellArgs<Number> &a, const Number *src,
                          Number *dst, hipStream_t st)
{
  constexpr int CPB = cells_per_block<Number, K>();
  int part_off = 0;
  for (int colour = 0; colour < 4; ++colour) {
    const int blocks = row_blocks(m, colour);
    if (blocks == 0) continue;
    const LaunchPart lp{colour, part_off, 0, 0};
    hipLaunchKernelGGL((cell_kernel<Number, K, OP, NEEDT, VEL, DOT, GridAccess<Number, K, CPB>>),
                       dim3(blocks), dim3(CPB * K * K), 0, st, a, src, dst, lp);
    MF_HIP_TRY(hipGetLastError());
    part_off += blocks;
  }
  return MF_OK;
}

template <typename Number, int K, int OP, int NEEDT, bool VEL>
static int launch_k(const MeshInfo &m, const CellArgs<Number> &a, const Number *src, Number *dst,
                    hipStream_t st)
{
  if constexpr (is_tangent_op(OP)) {
    if (a.dot_partial) return launch_colours<Number, K, OP, NEEDT, VEL, true>(m, a, src, dst, st);
  }
  return launch_colours<Number, K, OP, NEEDT, VEL, false>(m, a, src, dst, st);
}

template <typename Number, int K, int OP, int NEEDT, bool VEL, bool DOT, class Access>
static int launch_parts(const std::vector<int64_t> &start, const std::vector<int64_t> &end,
                        const CellArgs<Number> &a, const Number *src, Number *dst,
                        hipStream_t st)
{
  constexpr int CPB = cells_per_block<Number, K>();
  int part_off = 0;
  for (size_t col = 0; col < start.size(); ++col) {
    const int64_t count = end[col] - start[col];
    if (count <= 0) continue;
    const int blocks = int((count + CPB - 1) / CPB);
    const LaunchPart lp{int(col), part_off, start[col], count};
    hipLaunchKernelGGL((cell_kernel<Number, K, OP, NEEDT, VEL, DOT, Access>), dim3(blocks),
                       dim3(CPB * K * K), 0, st, a, src, dst, lp);
    MF_HIP_TRY(hipGetLastError());
    part_off += blocks;
  }
  return MF_OK;
}

// colour ranges of an adaptive launch: mode 0 all cells, 1 cells without hanging nodes,
// 2 condensed-diagonal slots
static void list_ranges(const AdaptiveData &ad, int mode, std::vector<int64_t> &start,
                        std::vector<int64_t> &end)
{
  if (mode == 2) {
    start.assign(1, ad.slot_start[0]);
    end.assign(1, ad.slot_start[1]);
    return;
  }
  const int nc = ad.n_colours;
  start.resize(nc);
  end.resize(nc);
  for (int col = 0; col < nc; ++col) {
    start[col] = ad.colour_start[col];
    end[col] = mode == 1 ? ad.colour_hang[col] : ad.colour_start[col + 1];
  }
}

template <typename Number, int K, int OP, int NEEDT, bool VEL>
static int launch_list_k(const MeshInfo &m, const CellArgs<Number> &a, const Number *src,
                         Number *dst, hipStream_t st, int mode)
{
  constexpr int CPB = cells_per_block<Number, K>();
  std::vector<int64_t> start, end;
  list_ranges(*m.ad, mode, start, end);
  if (mode == 2) {
    if constexpr (is_tangent_op(OP))
      return launch_parts<Number, K, OP, NEEDT, VEL, false, SlotAccess<Number, K, CPB>>(
          start, end, a, src, dst, st);
    return MF_ERR_ARG;
  }
  if constexpr (is_tangent_op(OP)) {
    if (a.dot_partial)
      return launch_parts<Number, K, OP, NEEDT, VEL, true, ListAccess<Number, K, CPB>>(
          start, end, a, src, dst, st);
  }
  return launch_parts<Number, K, OP, NEEDT, VEL, false, ListAccess<Number, K, CPB>>(
      start, end, a, src, dst, st);
}

template <typename Number, int K, int OP>
static int launch_list_op(int needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                          const Number *src, Number *dst, hipStream_t st, int mode)
{
  if constexpr (OP == ADVEC_TANGENT || OP == ADVEC_RHS || OP == ADVEC_DIAG) {
    if (needT) return launch_list_k<Number, K, OP, 1, true>(m, a, src, dst, st, mode);
    return launch_list_k<Number, K, OP, false, true>(m, a, src, dst, st, mode);
  } else if constexpr (OP == LS_TANGENT || OP == LS_DIAG) {
    if (needT) return launch_list_k<Number, K, OP, true, false>(m, a, src, dst, st, mode);
    return launch_list_k<Number, K, OP, false, false>(m, a, src, dst, st, mode);
  } else if constexpr (OP == REINIT_TANGENT || OP == REINIT_DIAG) {
    return launch_list_k<Number, K, OP, false, true>(m, a, src, dst, st, mode);
  } else {
    if (needT == 2) {
      if (vel) return launch_list_k<Number, K, OP, 2, true>(m, a, src, dst, st, mode);
      return launch_list_k<Number, K, OP, 2, false>(m, a, src, dst, st, mode);
    }
    if (needT) {
      if (vel) return launch_list_k<Number, K, OP, 1, true>(m, a, src, dst, st, mode);
      return launch_list_k<Number, K, OP, 1, false>(m, a, src, dst, st, mode);
    }
    if (vel) return launch_list_k<Number, K, OP, false, true>(m, a, src, dst, st, mode);
    return launch_list_k<Number, K, OP, false, false>(m, a, src, dst, st, mode);
  }
}

template <typename Number, int K>
static int launch_list_deg(int op, int needT, bool vel, const MeshInfo &m,
                           const CellArgs<Number> &a, const Number *src, Number *dst,
                           hipStream_t st, int mode)
{
  switch (op) {
  case HEAT_TANGENT:
    return launch_list_op<Number, K, HEAT_TANGENT>(needT, vel, m, a, src, dst, st, mode);
  case ADVEC_TANGENT:
    return launch_list_op<Number, K, ADVEC_TANGENT>(needT, vel, m, a, src, dst, st, mode);
  case HEAT_DIAG:
    if constexpr (sizeof(Number) == 8)
      return launch_list_op<Number, K, HEAT_DIAG>(needT, vel, m, a, src, dst, st, mode);
    break;
  case ADVEC_DIAG:
    if constexpr (sizeof(Number) == 8)
      return launch_list_op<Number, K, ADVEC_DIAG>(needT, vel, m, a, src, dst, st, mode);
    break;
  case HEAT_RHS:
    if constexpr (sizeof(Number) == 8)
      return launch_list_op<Number, K, HEAT_RHS>(needT, vel, m, a, src, dst, st, mode);
    break;
  case ADVEC_RHS:
    if constexpr (sizeof(Number) == 8)
      return launch_list_op<Number, K, ADVEC_RHS>(needT, vel, m, a, src, dst, st, mode);
    break;
  case LS_TANGENT:
    return launch_list_op<Number, K, LS_TANGENT>(needT, vel, m, a, src, dst, st, mode);
  case LS_DIAG:
    if constexpr (sizeof(Number) == 8)
      return launch_list_op<Number, K, LS_DIAG>(needT, vel, m, a, src, dst, st, mode);
    break;
  case REINIT_TANGENT:
    return launch_list_op<Number, K, REINIT_TANGENT>(needT, vel, m, a, src, dst, st, mode);
  case REINIT_DIAG:
    if constexpr (sizeof(Number) == 8)
      return launch_list_op<Number, K, REINIT_DIAG>(needT, vel, m, a, src, dst, st, mode);
    break;
  default:
    break;
  }
  set_error("operation not available on adaptive meshes");
  return MF_ERR_ARG;
}

template <typename Number>
int launch_cell_list(int op, int needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                     const Number *src, Number *dst, hipStream_t st, int mode)
{
  if (!m.ad) {
    set_error("not an adaptive mesh");
    return MF_ERR_ARG;
  }
  switch (m.K) {
  case 2: return launch_list_deg<Number, 2>(op, needT, vel, m, a, src, dst, st, mode);
  case 3: return launch_list_deg<Number, 3>(op, needT, vel, m, a, src, dst, st, mode);
  case 4: return launch_list_deg<Number, 4>(op, needT, vel, m, a, src, dst, st, mode);
  default: set_error("degree must be 1, 2 or 3"); return MF_ERR_ARG;
  }
}

template int launch_cell_list<double>(int, int, bool, const MeshInfo &, const CellArgs<double> &,
                                      const double *, double *, hipStream_t, int);
template int launch_cell_list<float>(int, int, bool, const MeshInfo &, const CellArgs<float> &,
                                     const float *, float *, hipStream_t, int);

template <typename Number, int K, int OP>
static int launch_op(int needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                     const Number *src, Number *dst, hipStream_t st)
{
  if constexpr (OP == ADVEC_TANGENT || OP == ADVEC_RHS || OP == ADVEC_DIAG) {
    if (needT) return launch_k<Number, K, OP, 1, true>(m, a, src, dst, st);
    return launch_k<Number, K, OP, false, true>(m, a, src, dst, st);
  } else if constexpr (OP == LS_TANGENT || OP == LS_DIAG) {
    if (needT) return launch_k<Number, K, OP, true, false>(m, a, src, dst, st);
    return launch_k<Number, K, OP, false, false>(m, a, src, dst, st);
  } else if constexpr (OP == REINIT_TANGENT || OP == REINIT_DIAG) {
    return launch_k<Number, K, OP, false, true>(m, a, src, dst, st);
  } else {
    if (needT == 2) {
      if (vel) return launch_k<Number, K, OP, 2, true>(m, a, src, dst, st);
      return launch_k<Number, K, OP, 2, false>(m, a, src, dst, st);
    }
    if (needT) {
      if (vel) return launch_k<Number, K, OP, 1, true>(m, a, src, dst, st);
      return launch_k<Number, K, OP, 1, false>(m, a, src, dst, st);
    }
    if (vel) return launch_k<Number, K, OP, false, true>(m, a, src, dst, st);
    return launch_k<Number, K, OP, false, false>(m, a, src, dst, st);
  }
}

template <typename Number, int K>
static int launch_deg(int op, int needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                      const Number *src, Number *dst, hipStream_t st)
{
  switch (op) {
  case HEAT_TANGENT: return launch_op<Number, K, HEAT_TANGENT>(needT, vel, m, a, src, dst, st);
  case HEAT_RHS: return launch_op<Number, K, HEAT_RHS>(needT, vel, m, a, src, dst, st);
  case HEAT_DIAG: return launch_op<Number, K, HEAT_DIAG>(needT, vel, m, a, src, dst, st);
  case ADVEC_TANGENT: return launch_op<Number, K, ADVEC_TANGENT>(needT, vel, m, a, src, dst, st);
  case ADVEC_RHS: return launch_op<Number, K, ADVEC_RHS>(needT, vel, m, a, src, dst, st);
  case ADVEC_DIAG: return launch_op<Number, K, ADVEC_DIAG>(needT, vel, m, a, src, dst, st);
  case HEAT_PGRAD: return launch_k<Number, K, HEAT_PGRAD, false, false>(m, a, src, dst, st);
  case LS_TANGENT: return launch_op<Number, K, LS_TANGENT>(needT, vel, m, a, src, dst, st);
  case LS_DIAG:
    if constexpr (sizeof(Number) == 8) return launch_op<Number, K, LS_DIAG>(needT, vel, m, a, src, dst, st);
    break;
  case REINIT_TANGENT: return launch_op<Number, K, REINIT_TANGENT>(needT, vel, m, a, src, dst, st);
  case REINIT_DIAG:
    if constexpr (sizeof(Number) == 8) return launch_op<Number, K, REINIT_DIAG>(needT, vel, m, a, src, dst, st);
    break;
  default: break;
  }
  set_error("unknown cell operation");
  return MF_ERR_ARG;
}

template <typename Number>
int launch_cell(int op, int needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                const Number *src, Number *dst, hipStream_t st)
{
  if (m.ad) {
    set_error("operation not available on adaptive meshes (structured-mesh kernel)");
    return MF_ERR_ARG;
  }
  if (op == HEAT_TANGENT && plane_applicable<Number>(m, needT, vel, a))
    return launch_heat_plane<Number>(m, needT, a, src, dst, st);
  switch (m.K) {
  case 2: return launch_deg<Number, 2>(op, needT, vel, m, a, src, dst, st);
  case 3: return launch_deg<Number, 3>(op, needT, vel, m, a, src, dst, st);
  case 4: return launch_deg<Number, 4>(op, needT, vel, m, a, src, dst, st);
  default: set_error("degree must be 1, 2 or 3"); return MF_ERR_ARG;
  }
}

template int launch_cell<double>(int, int, bool, const MeshInfo &, const CellArgs<double> &,
                                 const double *, double *, hipStream_t);
template int launch_cell<float>(int, int, bool, const MeshInfo &, const CellArgs<float> &,
                                const float *, float *, hipStream_t);

}  // namespace mf

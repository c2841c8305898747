// Vector-layer and Krylov kernels: fills and casts, Jacobi inversion, the fused
// preconditioned-CG update kernels with device-side convergence control, and the GMRES
// orthogonalisation kernels.  Replaces the LA::distributed::Vector operations deal.II's
// SolverCG / SolverGMRES perform inside LinearSolver::solve
// (include/meltpooldg/linear_algebra/linear_solver.hpp:106-119) and JacobiPreconditioner::vmult
// (include/meltpooldg/linear_algebra/preconditioner_jacobi.hpp:51-56).
//
// Reductions are deterministic: every reduction kernel writes one partial per block
// (RED_BLOCKS blocks, grid-stride), and consumers re-reduce the partials in a fixed order.
// FP64 partials are plain sums; FP32 partials carry a Neumaier compensation term, combined with
// two-sum at every level (wave, block, partials), so FP32 dot products keep ~FP64-like accuracy.
#include <utility>

#include "mf_vec.hpp"

namespace mf {

// ------------------------------------------------------------------------------------------
// compensated accumulator
// ------------------------------------------------------------------------------------------
template <typename T>
struct Acc;

template <>
struct Acc<double> {
  double s = 0, c = 0;
  __device__ void add(double x) { s += x; }
  __device__ void merge(double os, double oc) { s += os; c += oc; }
  __device__ double value() const { return s + c; }
};

template <>
struct Acc<float> {
  float s = 0, c = 0;
  __device__ void add(float x)
  {
    const float t = s + x;
    c += fabsf(s) >= fabsf(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  __device__ void merge(float os, float oc)
  {
    const float t = s + os;
    c += (fabsf(s) >= fabsf(os) ? (s - t) + os : (os - t) + s) + oc;
    s = t;
  }
  __device__ float value() const { return s + c; }
};

// block-wide reduction of an accumulator; result valid in thread 0
template <typename T>
__device__ Acc<T> block_reduce(Acc<T> a)
{
  __shared__ T rs[16], rc[16];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const T os = __shfl_xor(a.s, off, 64);
    const T oc = __shfl_xor(a.c, off, 64);
    a.merge(os, oc);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    rs[wid] = a.s;
    rc[wid] = a.c;
  }
  __syncthreads();
  Acc<T> r;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r.merge(rs[i], rc[i]);
  }
  return r;
}

// every block reduces the same partial array (2*n values: sums then compensations) and returns
// the same value to all its threads
template <typename T>
__device__ T reduce_partials(const T *part, int n)
{
  __shared__ T result;
  Acc<T> a;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a.merge(part[i], part[n + i]);
  a = block_reduce(a);
  if (threadIdx.x == 0) result = a.value();
  __syncthreads();
  return result;
}

// 16-byte chunks of the CG vectors: N entries per load/store (used when every pointer is
// 16-byte aligned; the entries past the last whole chunk take the scalar loop)
template <typename T>
struct Chunk;
template <>
struct Chunk<double> {
  static constexpr int N = 2;
  using V = double2;
};
template <>
struct Chunk<float> {
  static constexpr int N = 4;
  using V = float4;
};
template <typename T>
union ChunkU {
  typename Chunk<T>::V v;
  T a[Chunk<T>::N];
};

__device__ __forceinline__ bool aligned16(const void *p) { return (uintptr_t(p) & 15u) == 0; }

// three reduce_partials in one pass (the same arithmetic, one set of block barriers)
template <typename T>
__device__ void reduce_partials3(const T *pa, const T *pb, const T *pc, int n, T &ra, T &rb, T &rc)
{
  __shared__ T result[3];
  __shared__ T rs[3][16], rcmp[3][16];
  Acc<T> a[3];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    a[0].merge(pa[i], pa[n + i]);
    a[1].merge(pb[i], pb[n + i]);
    a[2].merge(pc[i], pc[n + i]);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const T os = __shfl_xor(a[k].s, off, 64);
      const T oc = __shfl_xor(a[k].c, off, 64);
      a[k].merge(os, oc);
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      rs[k][wid] = a[k].s;
      rcmp[k][wid] = a[k].c;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Acc<T> r;
      for (int i = 0; i < nw; ++i) r.merge(rs[k][i], rcmp[k][i]);
      result[k] = r.value();
    }
  }
  __syncthreads();
  ra = result[0];
  rb = result[1];
  rc = result[2];
}

template <typename T>
__device__ void write_partial(T *part, int n, Acc<T> a)
{
  a = block_reduce(a);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = a.s;
    part[n + blockIdx.x] = a.c;
  }
}

// ------------------------------------------------------------------------------------------
// elementwise kernels
// ------------------------------------------------------------------------------------------
template <typename T>
__global__ void fill_kernel(T *x, T v, int64_t n, const int *stop)
{
  if (stop && *stop) return;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[i] = v;
}

template <typename Ta, typename Tb>
__global__ void cast_kernel(const Ta *a, Tb *b, double scale, int64_t n)
{
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t i0 = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int64_t tail = 0;
  if constexpr (sizeof(Ta) == 8 && sizeof(Tb) == 4) {
    // 4 entries per step: two 16-byte loads, one 16-byte store
    if (aligned16(a) && aligned16(b)) {
      const int64_t nc = n / 4;
      for (int64_t c = i0; c < nc; c += stride) {
        const double2 u = reinterpret_cast<const double2 *>(a)[2 * c];
        const double2 v = reinterpret_cast<const double2 *>(a)[2 * c + 1];
        reinterpret_cast<float4 *>(b)[c] = make_float4(float(u.x * scale), float(u.y * scale),
                                                       float(v.x * scale), float(v.y * scale));
      }
      tail = nc * 4;
    }
  }
  for (int64_t i = tail + i0; i < n; i += stride) b[i] = Tb(double(a[i]) * scale);
}

// x += scale * e  (FP64 solution update of the defect correction)
__global__ void add_scaled_f32_kernel(double *x, const float *e, double scale,
                                      const double *scale_dev, double skip_at_or_below, int64_t n)
{
  if (scale_dev) scale = *scale_dev;
  // a scale at or below the threshold, or not finite, leaves x as it is
  if (!(scale > skip_at_or_below) || !isfinite(scale)) return;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t i0 = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int64_t tail = 0;
  if (aligned16(x) && aligned16(e)) {
    const int64_t nc = n / 4;
    for (int64_t c = i0; c < nc; c += stride) {
      const float4 d = reinterpret_cast<const float4 *>(e)[c];
      double2 u = reinterpret_cast<const double2 *>(x)[2 * c];
      double2 v = reinterpret_cast<const double2 *>(x)[2 * c + 1];
      u.x += scale * double(d.x);
      u.y += scale * double(d.y);
      v.x += scale * double(d.z);
      v.y += scale * double(d.w);
      reinterpret_cast<double2 *>(x)[2 * c] = u;
      reinterpret_cast<double2 *>(x)[2 * c + 1] = v;
    }
    tail = nc * 4;
  }
  for (int64_t i = tail + i0; i < n; i += stride) x[i] += scale * double(e[i]);
}

// r = b - q, partials of r.r
template <typename T>
__global__ void residual_kernel(const T *__restrict__ b, const T *__restrict__ q,
                                T *__restrict__ r, int64_t n, T *part)
{
  Acc<T> a;
  constexpr int N = Chunk<T>::N;
  using V = typename Chunk<T>::V;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t i0 = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int64_t tail = 0;
  if (aligned16(b) && aligned16(r) && (!q || aligned16(q))) {
    const int64_t nc = n / N;
#pragma unroll 2
    for (int64_t c = i0; c < nc; c += stride) {
      ChunkU<T> vb, vq;
      vb.v = reinterpret_cast<const V *>(b)[c];
      if (q) vq.v = reinterpret_cast<const V *>(q)[c];
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const T v = q ? vb.a[k] - vq.a[k] : vb.a[k];
        vb.a[k] = v;
        a.add(v * v);
      }
      reinterpret_cast<V *>(r)[c] = vb.v;
    }
    tail = nc * N;
  }
#pragma unroll 4
  for (int64_t i = tail + i0; i < n; i += stride) {
    const T v = q ? b[i] - q[i] : b[i];  // q = null: A x = 0 (zero initial guess)
    r[i] = v;
    a.add(v * v);
  }
  write_partial(part, gridDim.x, a);
}

template <typename T>
__global__ void dot_kernel(const T *__restrict__ a, const T *__restrict__ b, int64_t n, T *part)
{
  Acc<T> acc;
#pragma unroll 4
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    acc.add(a[i] * b[i]);
  write_partial(part, gridDim.x, acc);
}

// x += a v, then partials of x . w (LA::distributed::Vector::add_and_dot)
__global__ void axpy_dot_kernel(double *__restrict__ x, double a, const double *__restrict__ v,
                                const double *w, int64_t n, double *part)
{
  Acc<double> acc;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double xi = x[i] + a * v[i];
    x[i] = xi;
    acc.add(xi * w[i]);
  }
  write_partial(part, gridDim.x, acc);
}

// dst = c_0 x_0, then dst += c_s x_s for s = 1..m-1 (Vector::equ followed by Vector::add)
__global__ void lincomb_kernel(double *__restrict__ dst, LinComb lc, int64_t n)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double v = lc.c[0] * lc.x[0][i];
    for (int s = 1; s < lc.m; ++s) v += lc.c[s] * lc.x[s][i];
    dst[i] = v;
  }
}

// max |d| partials (no compensation needed)
template <typename T>
__global__ void absmax_kernel(const T *d, int64_t n, double *part)
{
  __shared__ double red[16];
  double m = 0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    m = fmax(m, fabs(double(d[i])));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < int(blockDim.x >> 6); ++i) m = fmax(m, red[i]);
    part[blockIdx.x] = m;
  }
}

// |d| > thr ? 1/d : 1; thr = abs_thr or rel_thr * max(1, max_i |d_i|) from the partials
__global__ void invert_diag_kernel(double *d, int64_t n, double abs_thr, double rel_thr,
                                   const double *max_part, int nparts)
{
  __shared__ double thr;
  if (threadIdx.x == 0) {
    double t = abs_thr;
    if (max_part) {
      double m = 1.0;
      for (int i = 0; i < nparts; ++i) m = fmax(m, max_part[i]);
      t = rel_thr * m;
    }
    thr = t;
  }
  __syncthreads();
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const double v = d[i];
    d[i] = fabs(v) > thr ? 1.0 / v : 1.0;
  }
}

// ------------------------------------------------------------------------------------------
// fused PCG (deal.II SolverCG semantics with ReductionControl)
// ------------------------------------------------------------------------------------------
// After q = A x:  r = b - q, z = D r, p = z, partial r.r and r.z (slot 0).
template <typename T>
__global__ void cg_init_kernel(const T *__restrict__ b, const T *__restrict__ q, T *__restrict__ r,
                               T *__restrict__ p, const T *__restrict__ inv, int64_t n,
                               T *part_rr, T *part_rz0)
{
  Acc<T> arr, arz;
  constexpr int N = Chunk<T>::N;
  using V = typename Chunk<T>::V;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t i0 = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int64_t tail = 0;
  if (aligned16(b) && aligned16(r) && aligned16(p) && (!q || aligned16(q)) &&
      (!inv || aligned16(inv))) {
    const int64_t nc = n / N;
#pragma unroll 2
    for (int64_t c = i0; c < nc; c += stride) {
      ChunkU<T> vb, vq, vd, vz;
      vb.v = reinterpret_cast<const V *>(b)[c];
      if (q) vq.v = reinterpret_cast<const V *>(q)[c];
      if (inv) vd.v = reinterpret_cast<const V *>(inv)[c];
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const T v = q ? vb.a[k] - vq.a[k] : vb.a[k];
        const T z = inv ? vd.a[k] * v : v;
        vb.a[k] = v;
        vz.a[k] = z;
        arr.add(v * v);
        arz.add(v * z);
      }
      reinterpret_cast<V *>(r)[c] = vb.v;
      reinterpret_cast<V *>(p)[c] = vz.v;
    }
    tail = nc * N;
  }
#pragma unroll 4
  for (int64_t i = tail + i0; i < n; i += stride) {
    const T v = q ? b[i] - q[i] : b[i];  // q = null: A x = 0 (zero initial guess)
    const T z = inv ? inv[i] * v : v;
    r[i] = v;
    p[i] = z;
    arr.add(v * v);
    arz.add(v * z);
  }
  write_partial(part_rr, gridDim.x, arr);
  write_partial(part_rz0, gridDim.x, arz);
}

// Start of an FP32 inner CG of the defect correction from the FP64 residual: r = float(s b64),
// z = D r, p = z, x = 0, partial r.r and r.z (the cast, the fill of x and cg_init in one pass)
__global__ void cg_init_refine_kernel(const double *__restrict__ b64, double s,
                                      const double *__restrict__ den, float *__restrict__ x,
                                      float *__restrict__ r, float *__restrict__ p,
                                      const float *__restrict__ inv, int64_t n, float *part_rr,
                                      float *part_rz0)
{
  // scale 1/|b64| from the device (a zero norm gives a zero right-hand side)
  if (den) s = *den > 0.0 ? 1.0 / *den : 0.0;
  Acc<float> arr, arz;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t i0 = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int64_t tail = 0;
  if (aligned16(b64) && aligned16(x) && aligned16(r) && aligned16(p) && (!inv || aligned16(inv))) {
    const int64_t nc = n / 4;
    for (int64_t c = i0; c < nc; c += stride) {
      const double2 u = reinterpret_cast<const double2 *>(b64)[2 * c];
      const double2 w = reinterpret_cast<const double2 *>(b64)[2 * c + 1];
      ChunkU<float> vr, vz, vd;
      vr.a[0] = float(u.x * s);
      vr.a[1] = float(u.y * s);
      vr.a[2] = float(w.x * s);
      vr.a[3] = float(w.y * s);
      if (inv) vd.v = reinterpret_cast<const float4 *>(inv)[c];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const float v = vr.a[k];
        const float z = inv ? vd.a[k] * v : v;
        vz.a[k] = z;
        arr.add(v * v);
        arz.add(v * z);
      }
      reinterpret_cast<float4 *>(r)[c] = vr.v;
      reinterpret_cast<float4 *>(p)[c] = vz.v;
      reinterpret_cast<float4 *>(x)[c] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    tail = nc * 4;
  }
  for (int64_t i = tail + i0; i < n; i += stride) {
    const float v = float(b64[i] * s);
    const float z = inv ? inv[i] * v : v;
    r[i] = v;
    p[i] = z;
    x[i] = 0.f;
    arr.add(v * v);
    arz.add(v * z);
  }
  write_partial(part_rr, gridDim.x, arr);
  write_partial(part_rz0, gridDim.x, arz);
}

// step-0 check: tol = max(abs, rel*|r0|), success if |r0| <= tol
template <typename T>
__global__ void cg_check0_kernel(const T *part_rr, int nparts, SolveState *st, double abs_tol,
                                 double rel_tol)
{
  const T rr = reduce_partials(part_rr, nparts);
  if (threadIdx.x == 0) {
    const double res = sqrt(double(rr));
    st->res0 = res;
    st->res = res;
    st->tol = fmax(abs_tol, rel_tol * res);
    st->iterations = 0;
    st->breakdown = 0;
    st->pAp = 0.0;
    st->alpha = 0.0;
    st->stop = (res <= st->tol) ? 1 : (isnan(res) ? 3 : 0);
  }
}


// 1-block reduction of the cell kernel's fused p.(Ap) partials into st->pAp (measured: cheaper
// than every update block reducing the partials itself)
__global__ void reduce_cell_dot_kernel(const double *part, int n, SolveState *st)
{
  if (st->stop) return;
  Acc<double> a;
  for (int i = threadIdx.x; i < n; i += blockDim.x) a.add(part[i]);
  a = block_reduce(a);
  if (threadIdx.x == 0) st->pAp = a.value();
}

// alpha = rz/pAp; r -= alpha q; partial r.r, r.(D r) into slot it%2.  x is updated by the
// direction kernel, which reads p anyway.
template <typename T>
__global__ void cg_update_kernel(T *__restrict__ r, const T *__restrict__ q,
                                 const T *__restrict__ inv, int64_t n,
                                 const T *part_rz_old, T *part_rz_new, T *part_rr, int nparts,
                                 SolveState *st)
{
  if (st->stop) return;
  const T rz = reduce_partials(part_rz_old, nparts);
  const double pAp = st->pAp;
  if (!(pAp != 0.0) || isnan(pAp)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->breakdown = 1;
    return;
  }
  const T alpha = T(double(rz) / pAp);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha = double(alpha);
  Acc<T> arr, arz;
  constexpr int N = Chunk<T>::N;
  using V = typename Chunk<T>::V;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  int64_t i0 = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int64_t tail = 0;
  if (aligned16(r) && aligned16(q) && (!inv || aligned16(inv))) {
    const int64_t nc = n / N;
#pragma unroll 2
    for (int64_t c = i0; c < nc; c += stride) {
      ChunkU<T> vr, vq, vd;
      vr.v = reinterpret_cast<const V *>(r)[c];
      vq.v = reinterpret_cast<const V *>(q)[c];
      if (inv) vd.v = reinterpret_cast<const V *>(inv)[c];
#pragma unroll
      for (int k = 0; k < N; ++k) {
        const T v = vr.a[k] - alpha * vq.a[k];
        vr.a[k] = v;
        arr.add(v * v);
        arz.add(v * (inv ? vd.a[k] * v : v));
      }
      reinterpret_cast<V *>(r)[c] = vr.v;
    }
    tail = nc * N;
  }
#pragma unroll 4
  for (int64_t i = tail + i0; i < n; i += stride) {
    const T v = r[i] - alpha * q[i];
    r[i] = v;
    arr.add(v * v);
    arz.add(v * (inv ? inv[i] * v : v));
  }
  write_partial(part_rr, nparts, arr);
  write_partial(part_rz_new, nparts, arz);
}

// x += alpha p (this iteration's step), convergence check for step `it`, then
// p = D r + beta p
template <typename T>
__global__ void cg_direction_kernel(T *__restrict__ x, T *__restrict__ p, const T *__restrict__ r,
                                    const T *__restrict__ inv, int64_t n,
                                    const T *part_rz_new, const T *part_rz_old, const T *part_rr,
                                    int nparts, SolveState *st, int it, int max_it)
{
  if (st->stop) return;
  if (st->breakdown) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->stop = 3;
      st->iterations = it;
    }
    return;
  }
  const T alpha = T(st->alpha);
  T rr, rz_new, rz_old;
  reduce_partials3(part_rr, part_rz_new, part_rz_old, nparts, rr, rz_new, rz_old);
  const double res = sqrt(double(rr));
  const double tol = st->tol;
  const bool done = res <= tol;
  const bool fail = !done && (it >= max_it || isnan(res));
  if (done || fail) {
  #pragma unroll 4
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
         i += int64_t(gridDim.x) * blockDim.x)
      x[i] += alpha * p[i];
    // every block sees the same values; block 0 records
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->res = res;
      st->iterations = it;
      st->stop = done ? 1 : 2;
    }
    return;
  }
  const T beta = T(double(rz_new) / double(rz_old));
  constexpr int N = Chunk<T>::N;
  using V = typename Chunk<T>::V;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const int64_t i0 = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  int64_t tail = 0;
  if (aligned16(x) && aligned16(p) && aligned16(r) && (!inv || aligned16(inv))) {
    const int64_t nc = n / N;
#pragma unroll 2
    for (int64_t c = i0; c < nc; c += stride) {
      ChunkU<T> vx, vp, vr, vd;
      vx.v = reinterpret_cast<const V *>(x)[c];
      vp.v = reinterpret_cast<const V *>(p)[c];
      vr.v = reinterpret_cast<const V *>(r)[c];
      if (inv) vd.v = reinterpret_cast<const V *>(inv)[c];
#pragma unroll
      for (int k = 0; k < N; ++k) {
        vx.a[k] += alpha * vp.a[k];
        const T z = inv ? vd.a[k] * vr.a[k] : vr.a[k];
        vp.a[k] = z + beta * vp.a[k];
      }
      reinterpret_cast<V *>(x)[c] = vx.v;
      reinterpret_cast<V *>(p)[c] = vp.v;
    }
    tail = nc * N;
  }
#pragma unroll 4
  for (int64_t i = tail + i0; i < n; i += stride) {
    const T pi = p[i];
    x[i] += alpha * pi;
    const T z = inv ? inv[i] * r[i] : r[i];
    p[i] = z + beta * pi;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->res = res;
    st->iterations = it;
  }
}

// ------------------------------------------------------------------------------------------
// GMRES helpers (classical Gram-Schmidt, two passes)
// ------------------------------------------------------------------------------------------
// partials of h_i = V_i . w for i < M (V row-major [M][n]); part layout [M][2*nparts].  M is a
// template parameter so that the M accumulators stay in registers.
template <typename T, int M>
__global__ void multidot_kernel(const T *V, int64_t ld, const T *w, int64_t n, T *part)
{
  Acc<T> acc[M];
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const T wi = w[i];
#pragma unroll
    for (int k = 0; k < M; ++k) acc[k].add(V[k * ld + i] * wi);
  }
#pragma unroll
  for (int k = 0; k < M; ++k) write_partial(part + size_t(k) * 2 * gridDim.x, gridDim.x, acc[k]);
}

constexpr int GMRES_MAX_BASIS = 30;

template <typename T, int... Ms>
static int launch_multidot(int m, const T *V, int64_t ld, const T *w, int64_t n, T *part,
                           hipStream_t st, std::integer_sequence<int, Ms...>)
{
  using Fn = void (*)(const T *, int64_t, const T *, int64_t, T *);
  static constexpr Fn table[] = {multidot_kernel<T, Ms + 1>...};
  if (m < 1 || m > int(sizeof...(Ms))) {
    set_error("GMRES basis beyond the multi-dot kernels (at most 30 vectors)");
    return MF_ERR_ARG;
  }
  hipLaunchKernelGGL(table[m - 1], dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, V, ld, w, n, part);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

// reduce [m][2*nparts] partials into h[m] (block k reduces h_k)
template <typename T>
__global__ void reduce_multi_kernel(const T *part, int nparts, double *h)
{
  const int k = int(blockIdx.x);
  Acc<T> a;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
    a.merge(part[size_t(k) * 2 * nparts + i], part[size_t(k) * 2 * nparts + nparts + i]);
  a = block_reduce(a);
  if (threadIdx.x == 0) h[k] = double(a.value());
}

// w -= sum_k h_k V_k, then partial of w.w
template <typename T>
__global__ void multiaxpy_kernel(const T *V, int64_t ld, int m, const double *h, T *w, int64_t n,
                                 T *part_ww)
{
  __shared__ T hs[32];
  if (threadIdx.x < m) hs[threadIdx.x] = T(h[threadIdx.x]);
  __syncthreads();
  Acc<T> a;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    T v = w[i];
    for (int k = 0; k < m; ++k) v -= hs[k] * V[k * ld + i];
    w[i] = v;
    a.add(v * v);
  }
  write_partial(part_ww, gridDim.x, a);
}

// A Gram-Schmidt pass fused with the dots of a second one: w -= sum_k h_k V_k, then partials
// of V_k . w for k < M (rows 0..M-1 of part) and of w . w (row M).  Each V_k is read once for
// both.  The update is the one of multiaxpy_kernel and the dots run in multidot_kernel's
// grid-stride order and partial layout, so the dots are bit for bit those of a separate
// multi-dot.
template <typename T, int M>
__global__ void __launch_bounds__(VEC_THREADS)
    multiaxpy_dot_kernel(const T *V, int64_t ld, const double *h, T *w, int64_t n, T *part)
{
  __shared__ T hs[32];
  if (threadIdx.x < M) hs[threadIdx.x] = T(h[threadIdx.x]);
  __syncthreads();
  Acc<T> acc[M], aww;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    T vk[M];
#pragma unroll
    for (int k = 0; k < M; ++k) vk[k] = V[k * ld + i];
    T v = w[i];
#pragma unroll
    for (int k = 0; k < M; ++k) v -= hs[k] * vk[k];
    w[i] = v;
#pragma unroll
    for (int k = 0; k < M; ++k) acc[k].add(vk[k] * v);
    aww.add(v * v);
  }
#pragma unroll
  for (int k = 0; k < M; ++k) write_partial(part + size_t(k) * 2 * gridDim.x, gridDim.x, acc[k]);
  write_partial(part + size_t(M) * 2 * gridDim.x, gridDim.x, aww);
}

// reduce [m + 1][2*nparts] partials: h[k] for k < m (block k), and the square root of row m into
// *norm (block m)
template <typename T>
__global__ void reduce_multi_norm_kernel(const T *part, int nparts, int m, double *h, double *norm)
{
  const int k = int(blockIdx.x);
  Acc<T> a;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
    a.merge(part[size_t(k) * 2 * nparts + i], part[size_t(k) * 2 * nparts + nparts + i]);
  a = block_reduce(a);
  if (threadIdx.x == 0) {
    if (k < m)
      h[k] = double(a.value());
    else
      *norm = sqrt(double(a.value()));
  }
}

template <typename T, int... Ms>
static int launch_multiaxpy_dot(int m, const T *V, int64_t ld, const double *h, T *w, int64_t n,
                                T *part, hipStream_t st, std::integer_sequence<int, Ms...>)
{
  using Fn = void (*)(const T *, int64_t, const double *, T *, int64_t, T *);
  static constexpr Fn table[] = {multiaxpy_dot_kernel<T, Ms + 1>...};
  if (m < 1 || m > int(sizeof...(Ms))) {
    set_error("GMRES basis beyond the multi-dot kernels (at most 30 vectors)");
    return MF_ERR_ARG;
  }
  hipLaunchKernelGGL(table[m - 1], dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, V, ld, h, w, n,
                     part);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

// dst = scale * src, scale read on the device (1/s for normalisation)
template <typename T>
__global__ void scale_kernel(T *dst, const T *src, int64_t n, const double *s, int invert)
{
  const double v = *s;
  const T f = T(invert ? 1.0 / v : v);
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    dst[i] = f * src[i];
}

// x += D (sum_k y_k V_k)   (right-preconditioned update)
template <typename T>
__global__ void gmres_update_kernel(T *x, const T *V, int64_t ld, int m, const double *y,
                                    const T *inv, int64_t n)
{
  __shared__ T ys[32];
  if (threadIdx.x < m) ys[threadIdx.x] = T(y[threadIdx.x]);
  __syncthreads();
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    T v = 0;
    for (int k = 0; k < m; ++k) v += ys[k] * V[k * ld + i];
    x[i] += inv ? inv[i] * v : v;
  }
}

// z = D v
template <typename T>
__global__ void precond_kernel(T *z, const T *v, const T *inv, int64_t n)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    z[i] = inv ? inv[i] * v[i] : v[i];
}

// 1-block: s = sum of partials
template <typename T>
__global__ void sum_from_partials_kernel(const T *part, int nparts, double *out)
{
  const T s = reduce_partials(part, nparts);
  if (threadIdx.x == 0) *out = double(s);
}

// 1-block: s = sqrt(sum of partials)
template <typename T>
__global__ void norm_from_partials_kernel(const T *part, int nparts, double *out)
{
  const T s = reduce_partials(part, nparts);
  if (threadIdx.x == 0) *out = sqrt(double(s));
}

// ------------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------------
static inline dim3 grid_for(int64_t n)
{
  int64_t b = (n + VEC_THREADS - 1) / VEC_THREADS;
  if (b > RED_BLOCKS) b = RED_BLOCKS;
  if (b < 1) b = 1;
  return dim3(unsigned(b));
}

template <typename T>
int vec_fill(T *x, T v, int64_t n, hipStream_t st, const int *stop)
{
  hipLaunchKernelGGL(fill_kernel<T>, grid_for(n), dim3(VEC_THREADS), 0, st, x, v, n, stop);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int vec_fill<double>(double *, double, int64_t, hipStream_t, const int *);
template int vec_fill<float>(float *, float, int64_t, hipStream_t, const int *);

template <typename Ta, typename Tb>
int vec_cast(const Ta *a, Tb *b, double scale, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL((cast_kernel<Ta, Tb>), grid_for(n), dim3(VEC_THREADS), 0, st, a, b, scale, n);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int vec_cast<double, float>(const double *, float *, double, int64_t, hipStream_t);
template int vec_cast<float, double>(const float *, double *, double, int64_t, hipStream_t);
template int vec_cast<double, double>(const double *, double *, double, int64_t, hipStream_t);

int vec_add_scaled_f32(double *x, const float *e, double scale, const double *scale_dev,
                       double skip_at_or_below, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(add_scaled_f32_kernel, grid_for(n), dim3(VEC_THREADS), 0, st, x, e, scale,
                     scale_dev, skip_at_or_below, n);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

namespace {
// pinned staging buffer and event of the calling thread on the current device
struct Stage {
  int dev = -1;
  void *buf = nullptr;
  hipEvent_t ev = nullptr;
};
thread_local Stage stage;
}  // namespace

int read_to_host_begin(const void *src, size_t bytes, hipStream_t st)
{
  MF_REQUIRE(bytes <= 4096, MF_ERR_ARG, "read_to_host: at most 4 KiB");
  int dev = 0;
  MF_HIP_TRY(hipGetDevice(&dev));
  if (stage.dev != dev) {
    if (stage.buf) (void)hipHostFree(stage.buf);
    if (stage.ev) (void)hipEventDestroy(stage.ev);
    stage = Stage{};
    MF_HIP_TRY(hipHostMalloc(&stage.buf, 4096));
    MF_HIP_TRY(hipEventCreateWithFlags(&stage.ev, hipEventDisableTiming));
    stage.dev = dev;
  }
  MF_HIP_TRY(hipMemcpyAsync(stage.buf, src, bytes, hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipEventRecord(stage.ev, st));
  return MF_OK;
}

int read_to_host_end(void *dst, size_t bytes)
{
  hipError_t e;
  while ((e = hipEventQuery(stage.ev)) == hipErrorNotReady) {
  }
  MF_HIP_TRY(e);
  std::memcpy(dst, stage.buf, bytes);
  return MF_OK;
}

int read_to_host(void *dst, const void *src, size_t bytes, hipStream_t st)
{
  // staged through pinned memory and awaited by polling an event: the host sees the result as
  // soon as the copy ends instead of waking from a blocking stream synchronisation
  const int rc = read_to_host_begin(src, bytes, st);
  if (rc) return rc;
  return read_to_host_end(dst, bytes);
}

int vec_invert_diagonal(double *d, int64_t n, bool relative, double *work, hipStream_t st)
{
  if (relative) {
    hipLaunchKernelGGL(absmax_kernel<double>, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, d, n, work);
    MF_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(invert_diag_kernel, grid_for(n), dim3(VEC_THREADS), 0, st, d, n, 0.0,
                       1.0e-14, work, RED_BLOCKS);
  } else {
    hipLaunchKernelGGL(invert_diag_kernel, grid_for(n), dim3(VEC_THREADS), 0, st, d, n, 1.0e-10,
                       0.0, (const double *)nullptr, 0);
  }
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

template <typename T>
int vec_residual(const T *b, const T *q, T *r, int64_t n, T *part, double *norm_out,
                 hipStream_t st)
{
  hipLaunchKernelGGL(residual_kernel<T>, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, b, q, r, n,
                     part);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(norm_from_partials_kernel<T>, dim3(1), dim3(VEC_THREADS), 0, st, part,
                     RED_BLOCKS, norm_out);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int vec_residual<double>(const double *, const double *, double *, int64_t, double *,
                                  double *, hipStream_t);
template int vec_residual<float>(const float *, const float *, float *, int64_t, float *, double *,
                                 hipStream_t);

// max over the partials of absmax_kernel
__global__ void max_from_partials_kernel(const double *part, int nparts, double *out)
{
  __shared__ double red[16];
  double m = 0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) m = fmax(m, part[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < int(blockDim.x >> 6); ++i) m = fmax(m, red[i]);
    *out = m;
  }
}

// whether x is zero everywhere (Vector::all_zero); work holds RED_BLOCKS + 1 doubles
template <typename T>
int vec_all_zero(const T *x, int64_t n, double *work, bool &zero, hipStream_t st)
{
  hipLaunchKernelGGL(absmax_kernel<T>, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, x, n, work);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(max_from_partials_kernel, dim3(1), dim3(VEC_THREADS), 0, st, work,
                     RED_BLOCKS, work + RED_BLOCKS);
  MF_HIP_TRY(hipGetLastError());
  double m = 1.0;
  if (int rc = read_to_host(&m, work + RED_BLOCKS, sizeof(double), st)) return rc;
  zero = (m == 0.0);
  return MF_OK;
}
template int vec_all_zero<double>(const double *, int64_t, double *, bool &, hipStream_t);
template int vec_all_zero<float>(const float *, int64_t, double *, bool &, hipStream_t);

template <typename T>
int vec_norm(const T *a, int64_t n, T *part, double *norm_out, hipStream_t st)
{
  hipLaunchKernelGGL(dot_kernel<T>, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, a, a, n, part);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(norm_from_partials_kernel<T>, dim3(1), dim3(VEC_THREADS), 0, st, part,
                     RED_BLOCKS, norm_out);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int vec_norm<double>(const double *, int64_t, double *, double *, hipStream_t);
template int vec_norm<float>(const float *, int64_t, float *, double *, hipStream_t);

int vec_axpy_dot(double *x, double a, const double *v, const double *w, int64_t n, double *part,
                 double *out, hipStream_t st)
{
  hipLaunchKernelGGL(axpy_dot_kernel, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, x, a, v, w, n,
                     part);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(sum_from_partials_kernel<double>, dim3(1), dim3(VEC_THREADS), 0, st, part,
                     RED_BLOCKS, out);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

int vec_lincomb(double *dst, const LinComb &lc, int64_t n, hipStream_t st)
{
  MF_REQUIRE(lc.m >= 1 && lc.m <= LinComb::MAX, MF_ERR_ARG, "linear combination of 1..8 vectors");
  hipLaunchKernelGGL(lincomb_kernel, grid_for(n), dim3(VEC_THREADS), 0, st, dst, lc, n);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}

template <typename T>
int vec_dot(const T *a, const T *b, int64_t n, T *part, double *out, hipStream_t st)
{
  hipLaunchKernelGGL(dot_kernel<T>, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, a, b, n, part);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(sum_from_partials_kernel<T>, dim3(1), dim3(VEC_THREADS), 0, st, part,
                     RED_BLOCKS, out);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int vec_dot<double>(const double *, const double *, int64_t, double *, double *,
                             hipStream_t);
template int vec_dot<float>(const float *, const float *, int64_t, float *, double *, hipStream_t);

template <typename T>
int cg_init(const T *b, T *q, T *r, T *p, const T *inv, int64_t n, T *part_rr, T *part_rz0,
            SolveState *sd, double abs_tol, double rel_tol, hipStream_t st)
{
  hipLaunchKernelGGL(cg_init_kernel<T>, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, b, q, r, p,
                     inv, n, part_rr, part_rz0);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(cg_check0_kernel<T>, dim3(1), dim3(VEC_THREADS), 0, st, part_rr, RED_BLOCKS,
                     sd, abs_tol, rel_tol);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int cg_init<double>(const double *, double *, double *, double *, const double *, int64_t,
                             double *, double *, SolveState *, double, double, hipStream_t);

int cg_init_refine(const double *b64, double s, const double *den, float *x, float *r, float *p,
                   const float *inv, int64_t n, float *part_rr, float *part_rz0, SolveState *sd,
                   double abs_tol, double rel_tol, hipStream_t st)
{
  hipLaunchKernelGGL(cg_init_refine_kernel, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, b64, s, den,
                     x, r, p, inv, n, part_rr, part_rz0);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(cg_check0_kernel<float>, dim3(1), dim3(VEC_THREADS), 0, st, part_rr,
                     RED_BLOCKS, sd, abs_tol, rel_tol);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int cg_init<float>(const float *, float *, float *, float *, const float *, int64_t,
                            float *, float *, SolveState *, double, double, hipStream_t);


template <typename T>
int cg_step(T *x, T *r, T *p, T *q, const T *inv, int64_t n, T *part_rz[2], T *part_rr,
            const double *cell_part, int ncell_part, SolveState *sd, int it, int max_it,
            hipStream_t st)
{
  T *rz_new = part_rz[it & 1], *rz_old = part_rz[(it - 1) & 1];
  hipLaunchKernelGGL(reduce_cell_dot_kernel, dim3(1), dim3(1024), 0, st, cell_part, ncell_part, sd);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(cg_update_kernel<T>, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, r, q, inv,
                     n, rz_old, rz_new, part_rr, RED_BLOCKS, sd);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(cg_direction_kernel<T>, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, x, p, r,
                     inv, n, rz_new, rz_old, part_rr, RED_BLOCKS, sd, it, max_it);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int cg_step<double>(double *, double *, double *, double *, const double *, int64_t,
                             double *[2], double *, const double *, int, SolveState *, int, int,
                             hipStream_t);
template int cg_step<float>(float *, float *, float *, float *, const float *, int64_t, float *[2],
                            float *, const double *, int, SolveState *, int, int, hipStream_t);

template <typename T>
int gmres_multidot(const T *V, int64_t ld, int m, const T *w, int64_t n, T *part, double *h,
                   hipStream_t st)
{
  if (int rc = launch_multidot<T>(m, V, ld, w, n, part, st,
                                  std::make_integer_sequence<int, GMRES_MAX_BASIS>{}))
    return rc;
  hipLaunchKernelGGL(reduce_multi_kernel<T>, dim3(m), dim3(VEC_THREADS), 0, st, part, RED_BLOCKS,
                     h);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int gmres_multidot<double>(const double *, int64_t, int, const double *, int64_t, double *,
                                    double *, hipStream_t);
template int gmres_multidot<float>(const float *, int64_t, int, const float *, int64_t, float *,
                                   double *, hipStream_t);

template <typename T>
int gmres_multiaxpy(const T *V, int64_t ld, int m, const double *h, T *w, int64_t n, T *part,
                    double *norm_out, hipStream_t st)
{
  hipLaunchKernelGGL(multiaxpy_kernel<T>, dim3(RED_BLOCKS), dim3(VEC_THREADS), 0, st, V, ld, m, h,
                     w, n, part);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(norm_from_partials_kernel<T>, dim3(1), dim3(VEC_THREADS), 0, st, part,
                     RED_BLOCKS, norm_out);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template <typename T>
int gmres_multiaxpy_dot(const T *V, int64_t ld, int m, const double *h, T *w, int64_t n, T *part,
                        double *h2, double *norm_out, hipStream_t st)
{
  if (int rc = launch_multiaxpy_dot<T>(m, V, ld, h, w, n, part, st,
                                       std::make_integer_sequence<int, GMRES_MAX_BASIS>{}))
    return rc;
  hipLaunchKernelGGL(reduce_multi_norm_kernel<T>, dim3(m + 1), dim3(VEC_THREADS), 0, st, part,
                     RED_BLOCKS, m, h2, norm_out);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int gmres_multiaxpy_dot<double>(const double *, int64_t, int, const double *, double *,
                                         int64_t, double *, double *, double *, hipStream_t);
template int gmres_multiaxpy_dot<float>(const float *, int64_t, int, const double *, float *,
                                        int64_t, float *, double *, double *, hipStream_t);

template int gmres_multiaxpy<double>(const double *, int64_t, int, const double *, double *,
                                     int64_t, double *, double *, hipStream_t);
template int gmres_multiaxpy<float>(const float *, int64_t, int, const double *, float *, int64_t,
                                    float *, double *, hipStream_t);

template <typename T>
int vec_scale_dev(T *dst, const T *src, int64_t n, const double *s, int invert, hipStream_t st)
{
  hipLaunchKernelGGL(scale_kernel<T>, grid_for(n), dim3(VEC_THREADS), 0, st, dst, src, n, s, invert);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int vec_scale_dev<double>(double *, const double *, int64_t, const double *, int,
                                   hipStream_t);
template int vec_scale_dev<float>(float *, const float *, int64_t, const double *, int, hipStream_t);

template <typename T>
int gmres_update(T *x, const T *V, int64_t ld, int m, const double *y, const T *inv, int64_t n,
                 hipStream_t st)
{
  hipLaunchKernelGGL(gmres_update_kernel<T>, grid_for(n), dim3(VEC_THREADS), 0, st, x, V, ld, m, y,
                     inv, n);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int gmres_update<double>(double *, const double *, int64_t, int, const double *,
                                  const double *, int64_t, hipStream_t);
template int gmres_update<float>(float *, const float *, int64_t, int, const double *,
                                 const float *, int64_t, hipStream_t);

template <typename T>
int vec_precond(T *z, const T *v, const T *inv, int64_t n, hipStream_t st)
{
  hipLaunchKernelGGL(precond_kernel<T>, grid_for(n), dim3(VEC_THREADS), 0, st, z, v, inv, n);
  MF_HIP_TRY(hipGetLastError());
  return MF_OK;
}
template int vec_precond<double>(double *, const double *, const double *, int64_t, hipStream_t);
template int vec_precond<float>(float *, const float *, const float *, int64_t, hipStream_t);

}  // namespace mf

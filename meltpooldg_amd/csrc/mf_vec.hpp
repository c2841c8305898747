// Vector-layer / Krylov kernel launchers (mf_vec.hip).
#pragma once

#include "mf_internal.hpp"

namespace mf {

constexpr int VEC_THREADS = 256;
constexpr int RED_BLOCKS = 1024;  // blocks of every reduction kernel (= partials per reduction)

// Device-resident solver control block (ReductionControl state).
struct SolveState {
  int stop;        // 0 iterate, 1 success, 2 max iterations, 3 breakdown / NaN
  int iterations;
  int breakdown;
  int pad;
  double res0, tol, res;
  double pAp;      // reduced p.(A p) of the current iteration
  double alpha;    // step length of the current iteration (x is updated one kernel later)
};

// Device -> host copy of a few bytes that the caller waits for (<= 4 KiB)
int read_to_host(void *dst, const void *src, size_t bytes, hipStream_t st);
// the same in two halves: begin enqueues the copy (more work may be enqueued behind it), end
// waits for it; no other read_to_host of the thread may come in between
int read_to_host_begin(const void *src, size_t bytes, hipStream_t st);
int read_to_host_end(void *dst, size_t bytes);

template <typename T> int vec_fill(T *x, T v, int64_t n, hipStream_t st, const int *stop = nullptr);
// dst[i] = mask[i] ? (vals ? vals[i] : 0) : (keep_free ? dst[i] : 0)   (mf_ops.hip)
int vec_masked_assign(double *dst, const uint8_t *mask, const double *vals, bool keep_free,
                      int64_t n, hipStream_t st);
// y = a x + b y   (mf_ops.hip)
int vec_axpby(double a, const double *x, double b, double *y, int64_t n, hipStream_t st);
// dst[i] = value where mask[i]   (mf_ops.hip)
int vec_masked_fill(double *dst, const uint8_t *mask, double value, int64_t n, hipStream_t st);
template <typename Ta, typename Tb> int vec_cast(const Ta *a, Tb *b, double scale, int64_t n, hipStream_t st);
// x += scale * e, the scale read from scale_dev when that is set; nothing is added when the
// scale is at or below skip_at_or_below or not finite
int vec_add_scaled_f32(double *x, const float *e, double scale, const double *scale_dev,
                       double skip_at_or_below, int64_t n, hipStream_t st);
int vec_invert_diagonal(double *d, int64_t n, bool relative, double *work, hipStream_t st);
template <typename T> int vec_residual(const T *b, const T *q, T *r, int64_t n, T *part, double *norm_out, hipStream_t st);
template <typename T> int vec_dot(const T *a, const T *b, int64_t n, T *part, double *out, hipStream_t st);
// x += a v and out = x . w (on the device; partials in part, 2 RED_BLOCKS entries)
int vec_axpy_dot(double *x, double a, const double *v, const double *w, int64_t n, double *part,
                 double *out, hipStream_t st);
// dst = sum_s c_s x_s, s = 0..m-1, accumulated in order of s
struct LinComb {
  static constexpr int MAX = 8;
  const double *x[MAX];
  double c[MAX];
  int m;
};
int vec_lincomb(double *dst, const LinComb &lc, int64_t n, hipStream_t st);
template <typename T> int vec_norm(const T *a, int64_t n, T *part, double *norm_out, hipStream_t st);
template <typename T> int vec_all_zero(const T *x, int64_t n, double *work, bool &zero, hipStream_t st);
template <typename T> int vec_precond(T *z, const T *v, const T *inv, int64_t n, hipStream_t st);
template <typename T> int vec_scale_dev(T *dst, const T *src, int64_t n, const double *s, int invert, hipStream_t st);

template <typename T>
int cg_init(const T *b, T *q, T *r, T *p, const T *inv, int64_t n, T *part_rr, T *part_rz0,
            SolveState *sd, double abs_tol, double rel_tol, hipStream_t st);
// FP32 inner CG of the defect correction started from the FP64 residual b64 scaled by s (by
// 1/*den when den is set, 0 for a zero norm), x = 0
int cg_init_refine(const double *b64, double s, const double *den, float *x, float *r, float *p,
                   const float *inv, int64_t n, float *part_rr, float *part_rz0, SolveState *sd,
                   double abs_tol, double rel_tol, hipStream_t st);
// one CG iteration after q = A p: reduction of the operator's p.(Ap) block partials
// cell_part[0..ncell_part), update and direction kernels
template <typename T>
int cg_step(T *x, T *r, T *p, T *q, const T *inv, int64_t n, T *part_rz[2], T *part_rr,
            const double *cell_part, int ncell_part, SolveState *sd, int it, int max_it,
            hipStream_t st);

template <typename T>
int gmres_multidot(const T *V, int64_t ld, int m, const T *w, int64_t n, T *part, double *h, hipStream_t st);
template <typename T>
int gmres_multiaxpy(const T *V, int64_t ld, int m, const double *h, T *w, int64_t n, T *part, double *norm_out, hipStream_t st);
// w -= V h (a Gram-Schmidt pass), then h2 = V^T w (the dots of a second pass) and |w| into
// *norm_out, in one read of V (part: (m + 1) x 2 RED_BLOCKS partials)
template <typename T>
int gmres_multiaxpy_dot(const T *V, int64_t ld, int m, const double *h, T *w, int64_t n, T *part, double *h2, double *norm_out, hipStream_t st);
template <typename T>
int gmres_update(T *x, const T *V, int64_t ld, int m, const double *y, const T *inv, int64_t n, hipStream_t st);

}  // namespace mf

// mf_operator_compute_inverse_diagonal_noalloc: the inverse diagonal of an operator without any
// device allocation, for callers that cannot allocate (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).  The relative inversion threshold of the advection
// and level-set operators reduces max |d| over RED_BLOCKS per-block partials; here those
// partials live in a file-scope device array instead of the operator's arena.
//
// Each GPU holds its own copy of the array (8 KiB) until the process ends.  Its device address is
// looked up once per device, with that device current, on the first call there.  Every call on a
// GPU shares that GPU's array: the reduction writes every partial before it reads any, so nothing
// needs clearing between calls, but calls that overlap on one GPU (different streams or threads
// without ordering between them) race on the partials.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mf_internal.hpp"
#include "mf_vec.hpp"

namespace {

constexpr int SCRATCH_DOUBLES = 1024;  // the maximum: one partial per reduction block
static_assert(mf::RED_BLOCKS <= SCRATCH_DOUBLES, "reduction partials exceed the scratch array");

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_lock;
double *g_scratch[MAX_DEVICES];  // device address of `scratch` per device (null: not looked up)

// the current device's copy of the array; the host symbol itself is no device address
int scratch_on_current_device(double **out)
{
  int dev = -1;
  MF_HIP_TRY(hipGetDevice(&dev));
  MF_REQUIRE(dev >= 0 && dev < MAX_DEVICES, MF_ERR_ARG,
             "device ordinal beyond the no-allocation scratch table (64 devices)");
  std::lock_guard<std::mutex> guard(g_lock);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      mf::set_error(std::string("hipGetSymbolAddress of the no-allocation scratch array failed: ") +
                    hipGetErrorString(e));
      return MF_ERR_HIP;
    }
    g_scratch[dev] = static_cast<double *>(p);
  }
  *out = g_scratch[dev];
  return MF_OK;
}

}  // namespace

extern "C" int mf_operator_compute_inverse_diagonal_noalloc(mf_operator *op, double *inv_diag)
{
  MF_REQUIRE(op && op->s && inv_diag, MF_ERR_ARG, "null argument");
  // the partials the reduction needs against the array's size (independent of the vector length)
  MF_REQUIRE(mf::RED_BLOCKS <= SCRATCH_DOUBLES, MF_ERR_ARG,
             "reduction partials exceed the no-allocation scratch array (1024 doubles)");
  double *work = nullptr;
  if (int rc = scratch_on_current_device(&work)) return rc;
  return mf::compute_inverse_diagonal(op, inv_diag, work);
}

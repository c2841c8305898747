// Node identification of an adaptive octree mesh on the device.
//
// Reference function replaced: FiniteElementUtils::distribute_dofs
// (include/meltpooldg/utilities/fe_util.hpp:21-81) for the adaptive meshes of the benchmark (the
// host path in mf_adaptive.cpp restates it; this is its device form, used when a GPU is present).
//
// Every (cell, local node) slot s = c K^3 + e gets the three axis keys of its node (the same keys
// as the host path: dyadic nodes by their coordinate in half-finest-cell units, interior
// Gauss-Lobatto nodes of Q3 by (level, lattice index)).  The slots are sorted by a 64-bit hash
// of the keys (stable radix sort, so equal hashes keep increasing slot order); runs of equal
// hashes are one node, whose full keys are compared to rule out collisions.  A node is numbered
// by the rank of its first slot in slot order (an exclusive prefix sum over first-slot flags),
// which is the host's first-appearance numbering.  Support points come from the node's last
// slot (the host overwrites them cell by cell) and the boundary flag is set by any slot.
#include <hipcub/hipcub.hpp>

#include <vector>

#include "mf_internal.hpp"

namespace mf {

namespace {

__device__ __forceinline__ uint64_t mix64_d(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct NodeArgs {
  int p, K, L;
  int64_t n_cells;
  int ncell[3];
  double lo[3], h[3], gl[KMAX];
};

__device__ __forceinline__ uint64_t axis_key_d(int p, int L, int l, int64_t ix, int i)
{
  const int64_t g = ix * p + i;
  const bool dyadic = (i == 0 || i == p || (p == 2 && i == 1));
  if (dyadic) return uint64_t((g << (L - l + 1)) / p);
  return (1ull << 62) | (uint64_t(l) << 48) | uint64_t(g);
}

__global__ void slot_keys_kernel(const NodeArgs a, const int32_t *__restrict__ cells,
                                 uint64_t *__restrict__ keys, uint64_t *__restrict__ hkey,
                                 uint32_t *__restrict__ sval, int64_t n_slots)
{
  const int nloc = a.K * a.K * a.K;
  for (int64_t s = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; s < n_slots;
       s += int64_t(gridDim.x) * blockDim.x) {
    const int64_t c = s / nloc;
    const int e = int(s - c * nloc);
    const int32_t *cl = cells + 4 * c;
    const int ai[3] = {e % a.K, (e / a.K) % a.K, e / (a.K * a.K)};
    uint64_t k[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) k[d] = axis_key_d(a.p, a.L, cl[0], cl[1 + d], ai[d]);
    keys[3 * s] = k[0];
    keys[3 * s + 1] = k[1];
    keys[3 * s + 2] = k[2];
    hkey[s] = mix64_d(k[0] ^ mix64_d(k[1] ^ mix64_d(k[2])));
    sval[s] = uint32_t(s);
  }
}

// head flags of the sorted runs (position of the run head, 0 elsewhere, for a max-scan), the
// first slot of every run, and a collision check of the full keys inside each run
__global__ void run_heads_kernel(const uint64_t *__restrict__ hs, const uint32_t *__restrict__ ss,
                                 const uint64_t *__restrict__ keys, int32_t *__restrict__ headpos,
                                 int32_t *__restrict__ first, int *collision, int64_t n)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const bool head = (i == 0) || hs[i] != hs[i - 1];
    headpos[i] = head ? int32_t(i) : 0;
    if (head) {
      first[ss[i]] = 1;
    } else {
      const uint64_t *ka = keys + 3 * size_t(ss[i]), *kb = keys + 3 * size_t(ss[i - 1]);
      if (ka[0] != kb[0] || ka[1] != kb[1] || ka[2] != kb[2]) atomicOr(collision, 1);
    }
  }
}

__global__ void assign_kernel(const NodeArgs a, const int32_t *__restrict__ cells,
                              const uint32_t *__restrict__ ss, const int32_t *__restrict__ grp,
                              const int32_t *__restrict__ nid, uint32_t *__restrict__ cell_nodes,
                              double *__restrict__ xyz, uint8_t *__restrict__ boundary, int64_t n)
{
  const int nloc = a.K * a.K * a.K;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int32_t g = grp[i];
    const uint32_t s = ss[i];
    const uint32_t id = uint32_t(nid[ss[g]]);
    cell_nodes[s] = id;
    const int64_t c = int64_t(s) / nloc;
    const int e = int(int64_t(s) - c * nloc);
    const int32_t *cl = cells + 4 * c;
    const int ai[3] = {e % a.K, (e / a.K) % a.K, e / (a.K * a.K)};
    bool bnd = false;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int64_t gg = int64_t(cl[1 + d]) * a.p + ai[d];
      bnd |= (ai[d] == 0 && gg == 0) || (ai[d] == a.p && gg == (int64_t(a.ncell[d]) << cl[0]) * a.p);
    }
    if (bnd) boundary[id] = 1;
    // the last slot of the run (largest slot: the host's last writer) sets the support point
    const bool tail = (i + 1 == n) || grp[i + 1] != g;
    if (tail) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double h = a.h[d] / double(int64_t(1) << cl[0]);
        xyz[3 * size_t(id) + d] = a.lo[d] + h * (double(cl[1 + d]) + a.gl[ai[d]]);
      }
    }
  }
}

struct Max32 {
  __device__ __forceinline__ int32_t operator()(int32_t x, int32_t y) const { return x > y ? x : y; }
};

// device buffers released on every return path
struct DevBufs {
  std::vector<void *> p;
  template <typename T>
  hipError_t alloc(T **ptr, size_t count)
  {
    void *q = nullptr;
    const hipError_t e = hipMalloc(&q, sizeof(T) * (count ? count : 1));
    if (e == hipSuccess) p.push_back(q);
    *ptr = static_cast<T *>(q);
    return e;
  }
  ~DevBufs()
  {
    for (void *q : p) (void)hipFree(q);
  }
};

}  // namespace

int device_node_identification(const mf_scratch *s, const int32_t *cells, int64_t n_cells, int L,
                               AdaptiveData *A, std::vector<uint64_t> &nkeys, int64_t &nn)
{
  const MeshInfo &m = s->mesh;
  const int p = m.degree, K = p + 1, nloc = K * K * K;
  const int64_t n_slots = n_cells * nloc;
  MF_REQUIRE(n_slots < (int64_t(1) << 31), MF_ERR_ARG, "adaptive mesh too large for 32-bit slots");
  hipStream_t st = s->stream;
  NodeArgs a{};
  a.p = p;
  a.K = K;
  a.L = L;
  a.n_cells = n_cells;
  for (int d = 0; d < 3; ++d) {
    a.ncell[d] = m.ncell[d];
    a.lo[d] = m.lo[d];
    a.h[d] = m.h[d];
  }
  for (int i = 0; i < K; ++i) a.gl[i] = s->tab.gl[i];
  DevBufs b;
  int32_t *d_cells, *headpos, *grp, *first, *nid;
  uint64_t *keys, *hkey, *hkey_s;
  uint32_t *sval, *sval_s, *cnodes;
  int *collision;
  MF_HIP_TRY(b.alloc(&d_cells, size_t(4 * n_cells)));
  MF_HIP_TRY(b.alloc(&keys, size_t(3 * n_slots)));
  MF_HIP_TRY(b.alloc(&hkey, size_t(n_slots)));
  MF_HIP_TRY(b.alloc(&hkey_s, size_t(n_slots)));
  MF_HIP_TRY(b.alloc(&sval, size_t(n_slots)));
  MF_HIP_TRY(b.alloc(&sval_s, size_t(n_slots)));
  MF_HIP_TRY(b.alloc(&headpos, size_t(n_slots)));
  MF_HIP_TRY(b.alloc(&grp, size_t(n_slots)));
  MF_HIP_TRY(b.alloc(&first, size_t(n_slots) + 1));
  MF_HIP_TRY(b.alloc(&nid, size_t(n_slots) + 1));
  MF_HIP_TRY(b.alloc(&cnodes, size_t(n_slots)));
  MF_HIP_TRY(b.alloc(&collision, 1));
  MF_HIP_TRY(hipMemcpyAsync(d_cells, cells, sizeof(int32_t) * 4 * size_t(n_cells),
                            hipMemcpyHostToDevice, st));
  MF_HIP_TRY(hipMemsetAsync(first, 0, sizeof(int32_t) * (size_t(n_slots) + 1), st));
  MF_HIP_TRY(hipMemsetAsync(collision, 0, sizeof(int), st));
  const dim3 grid(unsigned(std::min<int64_t>((n_slots + 255) / 256, 65536))), block(256);
  hipLaunchKernelGGL(slot_keys_kernel, grid, block, 0, st, a, d_cells, keys, hkey, sval, n_slots);
  MF_HIP_TRY(hipGetLastError());
  // stable sort of the slots by hash
  size_t tmp_bytes = 0, tmp2 = 0, tmp3 = 0;
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, hkey, hkey_s, sval, sval_s,
                                                int(n_slots), 0, 64, st));
  MF_HIP_TRY(hipcub::DeviceScan::InclusiveScan(nullptr, tmp2, headpos, grp, Max32(), int(n_slots), st));
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp3, first, nid, int(n_slots) + 1, st));
  tmp_bytes = std::max(tmp_bytes, std::max(tmp2, tmp3));
  void *tmp;
  MF_HIP_TRY(b.alloc(reinterpret_cast<char **>(&tmp), tmp_bytes));
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, hkey, hkey_s, sval, sval_s,
                                                int(n_slots), 0, 64, st));
  hipLaunchKernelGGL(run_heads_kernel, grid, block, 0, st, hkey_s, sval_s, keys, headpos, first,
                     collision, n_slots);
  MF_HIP_TRY(hipGetLastError());
  MF_HIP_TRY(hipcub::DeviceScan::InclusiveScan(tmp, tmp_bytes, headpos, grp, Max32(), int(n_slots), st));
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, first, nid, int(n_slots) + 1, st));
  int32_t total = 0;
  int coll = 0;
  MF_HIP_TRY(hipMemcpyAsync(&total, nid + n_slots, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipMemcpyAsync(&coll, collision, sizeof(int), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  if (coll) {
    set_error("adaptive mesh: 64-bit key hash collision in the device node identification");
    return MF_ERR_STATE;  // the caller falls back to the host path
  }
  nn = total;
  double *xyz;
  uint8_t *bnd;
  // the support points stay on the device as the scratch's copy (released on every failure)
  struct Kept {
    void *p = nullptr;
    ~Kept()
    {
      if (p) (void)hipFree(p);
    }
  } kept;
  MF_HIP_TRY(hipMalloc(&kept.p, sizeof(double) * 3 * size_t(nn)));
  xyz = static_cast<double *>(kept.p);
  MF_HIP_TRY(b.alloc(&bnd, size_t(nn)));
  MF_HIP_TRY(hipMemsetAsync(bnd, 0, size_t(nn), st));
  hipLaunchKernelGGL(assign_kernel, grid, block, 0, st, a, d_cells, sval_s, grp, nid, cnodes, xyz,
                     bnd, n_slots);
  MF_HIP_TRY(hipGetLastError());
  // node keys are not kept: the device constraint pass recomputes them per slot, and the host
  // fallback rebuilds them from the cells (nkeys left empty)
  A->cell_nodes.resize(size_t(n_slots));
  nkeys.clear();
  A->xyz.resize(size_t(3 * nn));
  A->boundary.resize(size_t(nn));
  MF_HIP_TRY(hipMemcpyAsync(A->cell_nodes.data(), cnodes, sizeof(uint32_t) * size_t(n_slots),
                            hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipMemcpyAsync(A->xyz.data(), xyz, sizeof(double) * size_t(3 * nn),
                            hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipMemcpyAsync(A->boundary.data(), bnd, size_t(nn), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  if (A->d.xyz) (void)hipFree(A->d.xyz);
  A->d.xyz = xyz;
  kept.p = nullptr;
  return MF_OK;
}

// Transposed hanging-node constraints: for every master, the hanging nodes it receives from and
// their weights, masters ascending and, per master, hanging nodes ascending (the order the host
// path builds).  Entries are keyed by (master << 32 | hanging node) and radix-sorted.
namespace {
__global__ void transpose_keys_kernel(const uint32_t *__restrict__ line_ptr,
                                      const uint32_t *__restrict__ line_col,
                                      const uint32_t *__restrict__ hang_of_line, int64_t n_lines,
                                      uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
  for (int64_t l = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; l < n_lines;
       l += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t h = hang_of_line[l];
    for (uint32_t t = line_ptr[l]; t < line_ptr[l + 1]; ++t) {
      key[t] = (uint64_t(line_col[t]) << 32) | h;
      val[t] = t;
    }
  }
}

__global__ void transpose_gather_kernel(const uint64_t *__restrict__ key_s,
                                        const uint32_t *__restrict__ val_s,
                                        const double *__restrict__ line_val,
                                        uint32_t *__restrict__ master, uint32_t *__restrict__ hang,
                                        double *__restrict__ w, int64_t n)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    master[i] = uint32_t(key_s[i] >> 32);
    hang[i] = uint32_t(key_s[i] & 0xffffffffull);
    w[i] = line_val[val_s[i]];
  }
}
}  // namespace

int device_transpose_constraints(AdaptiveData *A, const std::vector<uint32_t> &hang_of_line,
                                 hipStream_t st)
{
  const int64_t n_lines = int64_t(hang_of_line.size());
  const int64_t E = int64_t(A->line_col.size());
  A->cm_node.clear();
  A->cm_ptr.assign(1, 0u);
  A->cm_h.clear();
  A->cm_w.clear();
  if (E == 0) return MF_OK;
  MF_REQUIRE(E < (int64_t(1) << 31), MF_ERR_ARG, "too many constraint entries");
  DevBufs b;
  uint32_t *ptr, *col, *hol, *val, *val_s, *master, *hang;
  uint64_t *key, *key_s;
  double *lv, *w;
  MF_HIP_TRY(b.alloc(&ptr, size_t(n_lines + 1)));
  MF_HIP_TRY(b.alloc(&col, size_t(E)));
  MF_HIP_TRY(b.alloc(&hol, size_t(n_lines)));
  MF_HIP_TRY(b.alloc(&lv, size_t(E)));
  MF_HIP_TRY(b.alloc(&key, size_t(E)));
  MF_HIP_TRY(b.alloc(&key_s, size_t(E)));
  MF_HIP_TRY(b.alloc(&val, size_t(E)));
  MF_HIP_TRY(b.alloc(&val_s, size_t(E)));
  MF_HIP_TRY(b.alloc(&master, size_t(E)));
  MF_HIP_TRY(b.alloc(&hang, size_t(E)));
  MF_HIP_TRY(b.alloc(&w, size_t(E)));
  MF_HIP_TRY(hipMemcpyAsync(ptr, A->line_ptr.data(), sizeof(uint32_t) * size_t(n_lines + 1),
                            hipMemcpyHostToDevice, st));
  MF_HIP_TRY(hipMemcpyAsync(col, A->line_col.data(), sizeof(uint32_t) * size_t(E),
                            hipMemcpyHostToDevice, st));
  MF_HIP_TRY(hipMemcpyAsync(hol, hang_of_line.data(), sizeof(uint32_t) * size_t(n_lines),
                            hipMemcpyHostToDevice, st));
  MF_HIP_TRY(hipMemcpyAsync(lv, A->line_val.data(), sizeof(double) * size_t(E),
                            hipMemcpyHostToDevice, st));
  const dim3 block(256);
  hipLaunchKernelGGL(transpose_keys_kernel, dim3(unsigned(std::min<int64_t>((n_lines + 255) / 256, 65536))),
                     block, 0, st, ptr, col, hol, n_lines, key, val);
  MF_HIP_TRY(hipGetLastError());
  size_t tmp_bytes = 0;
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, key_s, val, val_s, int(E),
                                                0, 64, st));
  char *tmp;
  MF_HIP_TRY(b.alloc(&tmp, tmp_bytes));
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, key_s, val, val_s, int(E), 0,
                                                64, st));
  hipLaunchKernelGGL(transpose_gather_kernel, dim3(unsigned(std::min<int64_t>((E + 255) / 256, 65536))),
                     block, 0, st, key_s, val_s, lv, master, hang, w, E);
  MF_HIP_TRY(hipGetLastError());
  std::vector<uint32_t> hm(static_cast<size_t>(E));
  A->cm_h.resize(size_t(E));
  A->cm_w.resize(size_t(E));
  MF_HIP_TRY(hipMemcpyAsync(hm.data(), master, sizeof(uint32_t) * size_t(E), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipMemcpyAsync(A->cm_h.data(), hang, sizeof(uint32_t) * size_t(E),
                            hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipMemcpyAsync(A->cm_w.data(), w, sizeof(double) * size_t(E), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  for (int64_t i = 0; i < E; ++i) {
    if (i == 0 || hm[size_t(i)] != hm[size_t(i - 1)]) {
      if (i > 0) A->cm_ptr.push_back(uint32_t(i));
      A->cm_node.push_back(hm[size_t(i)]);
    }
  }
  A->cm_ptr.push_back(uint32_t(E));
  return MF_OK;
}

// Hanging-node constraints on the device, the same lines as the host path: a node of a level-l
// cell c is constrained when a level-(l-1) leaf touches c at an entity holding the node and the
// node is not one of that leaf's nodes; its line interpolates the leaf's Lagrange basis at the
// node.  The host takes, for every node, the first (cell, neighbour direction, local node) in
// loop order that constrains it, and numbers the lines in that order; here every candidate
// claims its node with an atomic minimum of that order key, and the lines are numbered by
// sorting the winning keys.  Chains (masters that are constrained themselves) are resolved by
// the caller.
namespace {
__device__ __forceinline__ uint64_t leaf_key_d(int l, int64_t x, int64_t y, int64_t z)
{
  return (uint64_t(l) << 60) | (uint64_t(x) << 40) | (uint64_t(y) << 20) | uint64_t(z);
}

// index of the leaf with the key, or -1 (binary search in the sorted leaf keys)
__device__ int64_t find_leaf(const uint64_t *keys, const uint32_t *idx, int64_t n, uint64_t k)
{
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && keys[lo] == k) ? int64_t(idx[lo]) : -1;
}

__global__ void leaf_keys_kernel(const int32_t *__restrict__ cells, int64_t n, uint64_t *key,
                                 uint32_t *val)
{
  for (int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; c < n;
       c += int64_t(gridDim.x) * blockDim.x) {
    const int32_t *cl = cells + 4 * c;
    key[c] = leaf_key_d(cl[0], cl[1], cl[2], cl[3]);
    val[c] = uint32_t(c);
  }
}

__global__ void fill_u64_kernel(uint64_t *x, uint64_t v, int64_t n)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[i] = v;
}

struct HncArgs {
  int p, K, L;
  int64_t n_cells;
  int ncell[3];
  double gl[KMAX];
};

// the coarse leaf of direction dir of cell c when it is one level coarser (else -1; -2 when it
// is two or more levels coarser), with its lattice index cc
__device__ int64_t coarse_neighbour(const HncArgs &a, const int32_t *cl, int dir,
                                    const uint64_t *lkeys, const uint32_t *lidx, int64_t cc[3])
{
  const int l = cl[0];
  const int del[3] = {dir % 3 - 1, (dir / 3) % 3 - 1, dir / 9 - 1};
  int64_t nb[3];
  for (int d = 0; d < 3; ++d) {
    nb[d] = int64_t(cl[1 + d]) + del[d];
    if (nb[d] < 0 || nb[d] >= (int64_t(a.ncell[d]) << l)) return -1;
  }
  // covering leaf: the finest existing leaf at or above level l containing the cell nb
  for (int lc = l; lc >= 0; --lc) {
    const int sh = l - lc;
    const int64_t f = find_leaf(lkeys, lidx, a.n_cells, leaf_key_d(lc, nb[0] >> sh, nb[1] >> sh, nb[2] >> sh));
    if (f >= 0) {
      if (lc < l - 1) return -2;  // not 2:1 balanced
      if (lc != l - 1) return -1;
      for (int d = 0; d < 3; ++d) cc[d] = nb[d] >> 1;
      return f;
    }
  }
  return -1;
}

// 1D weights of the constrained node of local coordinate ai along axis d: one-hot if the node
// is a node of the coarse cell along d, else the coarse Lagrange basis at its position (the
// products in the host's order, uncontracted, so the weights are the host's bit for bit); returns
// whether it is a coarse node along d
__device__ bool axis_weights(const HncArgs &a, int l, int64_t ix, int ai, int64_t ccd, double *wv)
{
#pragma clang fp contract(off)
  const uint64_t kk = axis_key_d(a.p, a.L, l, ix, ai);
  int match = -1;
  for (int i = 0; i < a.K; ++i)
    if (axis_key_d(a.p, a.L, l - 1, ccd, i) == kk) match = i;
  if (match >= 0) {
    for (int i = 0; i < a.K; ++i) wv[i] = (i == match) ? 1.0 : 0.0;
    return true;
  }
  const double t = 0.5 * (double(ix) + a.gl[ai]) - double(ccd);
  for (int j = 0; j < a.K; ++j) {
    double val = 1.0;
    for (int m = 0; m < a.K; ++m) {
      if (m == j) continue;
      const double inv = 1.0 / (a.gl[j] - a.gl[m]);
      val *= (t - a.gl[m]) * inv;
    }
    wv[j] = val;
  }
  return false;
}

__device__ __forceinline__ bool on_entity(int dir, const int ai[3], int p)
{
  const int del[3] = {dir % 3 - 1, (dir / 3) % 3 - 1, dir / 9 - 1};
  bool on = true;
  for (int d = 0; d < 3; ++d) on &= (del[d] == 0) || (del[d] < 0 && ai[d] == 0) || (del[d] > 0 && ai[d] == p);
  return on;
}

__global__ void hnc_claim_kernel(const HncArgs a, const int32_t *__restrict__ cells,
                                 const uint32_t *__restrict__ cell_nodes,
                                 const uint64_t *__restrict__ lkeys, const uint32_t *__restrict__ lidx,
                                 unsigned long long *claim, int *unbalanced)
{
  const int nloc = a.K * a.K * a.K;
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < a.n_cells * 27;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t c = t / 27;
    const int dir = int(t - c * 27);
    if (dir == 13) continue;  // the cell itself
    const int32_t *cl = cells + 4 * c;
    if (cl[0] == 0) continue;
    int64_t cc[3];
    const int64_t ci = coarse_neighbour(a, cl, dir, lkeys, lidx, cc);
    if (ci == -2) *unbalanced = 1;
    if (ci < 0) continue;
    for (int e = 0; e < nloc; ++e) {
      const int ai[3] = {e % a.K, (e / a.K) % a.K, e / (a.K * a.K)};
      if (!on_entity(dir, ai, a.p)) continue;
      double wv[KMAX];
      bool coarse_node = true;
      for (int d = 0; d < 3; ++d) coarse_node &= axis_weights(a, cl[0], cl[1 + d], ai[d], cc[d], wv);
      if (coarse_node) continue;
      const unsigned long long key = (unsigned long long)(t) * 64ull + unsigned(e);
      atomicMin(claim + cell_nodes[size_t(c) * nloc + e], key);
    }
  }
}

// hanging nodes: claim keys and node ids compacted in node order
__global__ void hnc_flag_kernel(const unsigned long long *claim, int64_t nn, int32_t *flag)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nn;
       i += int64_t(gridDim.x) * blockDim.x)
    flag[i] = claim[i] != ~0ull ? 1 : 0;
}

__global__ void hnc_compact_kernel(const unsigned long long *claim, const int32_t *flag,
                                   const int32_t *pos, int64_t nn, uint64_t *key, uint32_t *node)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nn;
       i += int64_t(gridDim.x) * blockDim.x)
    if (flag[i]) {
      key[pos[i]] = claim[i];
      node[pos[i]] = uint32_t(i);
    }
}

// line l of the sorted claims: its node's line index, then the entry count (mode 0) or the
// entries (mode 1)
template <int MODE>
__global__ void hnc_line_kernel(const HncArgs a, const int32_t *__restrict__ cells,
                                const uint32_t *__restrict__ cell_nodes,
                                const uint64_t *__restrict__ lkeys, const uint32_t *__restrict__ lidx,
                                const uint64_t *__restrict__ skey, const uint32_t *__restrict__ snode,
                                int64_t n_lines, int32_t *line_of, uint32_t *count,
                                const uint32_t *ptr, uint32_t *col, double *val)
{
  const int nloc = a.K * a.K * a.K;
  for (int64_t l = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; l < n_lines;
       l += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t k = skey[l];
    const int e = int(k % 64);
    const int64_t t = int64_t(k / 64);
    const int64_t c = t / 27;
    const int dir = int(t - c * 27);
    const int32_t *cl = cells + 4 * c;
    int64_t cc[3];
    const int64_t ci = coarse_neighbour(a, cl, dir, lkeys, lidx, cc);
    const int ai[3] = {e % a.K, (e / a.K) % a.K, e / (a.K * a.K)};
    double wv[3][KMAX];
    for (int d = 0; d < 3; ++d) axis_weights(a, cl[0], cl[1 + d], ai[d], cc[d], wv[d]);
    if (MODE == 0) line_of[snode[l]] = int32_t(l);
    uint32_t q = MODE == 1 ? ptr[l] : 0u;
    for (int f = 0; f < nloc; ++f) {
      const double w = wv[0][f % a.K] * wv[1][(f / a.K) % a.K] * wv[2][f / (a.K * a.K)];
      if (w == 0.0) continue;
      if (MODE == 1) {
        col[q] = cell_nodes[size_t(ci) * nloc + f];
        val[q] = w;
      }
      ++q;
    }
    if (MODE == 0) count[l] = q;
  }
}
}  // namespace

int device_hanging_constraints(const mf_scratch *s, const int32_t *cells, int64_t n_cells, int L,
                               const std::vector<uint32_t> &cell_nodes, int64_t nn,
                               std::vector<int32_t> &line_of, std::vector<uint32_t> &line_ptr,
                               std::vector<uint32_t> &line_col, std::vector<double> &line_val)
{
  const MeshInfo &m = s->mesh;
  const int p = m.degree, K = p + 1, nloc = K * K * K;
  hipStream_t st = s->stream;
  MF_REQUIRE(n_cells * 27 < (int64_t(1) << 57), MF_ERR_ARG, "adaptive mesh too large");
  HncArgs a{};
  a.p = p;
  a.K = K;
  a.L = L;
  a.n_cells = n_cells;
  for (int d = 0; d < 3; ++d) a.ncell[d] = m.ncell[d];
  for (int i = 0; i < K; ++i) a.gl[i] = s->tab.gl[i];
  DevBufs b;
  int32_t *d_cells, *flag, *pos, *d_line_of;
  uint32_t *d_nodes, *lval_idx, *lidx, *snode0, *snode, *cnt, *d_ptr, *d_col;
  uint64_t *lkey0, *lkeys, *skey0, *skey;
  unsigned long long *claim;
  double *d_val;
  MF_HIP_TRY(b.alloc(&d_cells, size_t(4 * n_cells)));
  MF_HIP_TRY(b.alloc(&d_nodes, size_t(n_cells) * nloc));
  MF_HIP_TRY(b.alloc(&lkey0, size_t(n_cells)));
  MF_HIP_TRY(b.alloc(&lkeys, size_t(n_cells)));
  MF_HIP_TRY(b.alloc(&lval_idx, size_t(n_cells)));
  MF_HIP_TRY(b.alloc(&lidx, size_t(n_cells)));
  MF_HIP_TRY(b.alloc(&claim, size_t(nn)));
  MF_HIP_TRY(b.alloc(&flag, size_t(nn) + 1));
  MF_HIP_TRY(b.alloc(&pos, size_t(nn) + 1));
  MF_HIP_TRY(b.alloc(&d_line_of, size_t(nn)));
  int *unbal;
  MF_HIP_TRY(b.alloc(&unbal, 1));
  MF_HIP_TRY(hipMemcpyAsync(d_cells, cells, sizeof(int32_t) * 4 * size_t(n_cells),
                            hipMemcpyHostToDevice, st));
  MF_HIP_TRY(hipMemcpyAsync(d_nodes, cell_nodes.data(), sizeof(uint32_t) * cell_nodes.size(),
                            hipMemcpyHostToDevice, st));
  const dim3 block(256);
  auto grid = [](int64_t n) { return dim3(unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)))); };
  hipLaunchKernelGGL(leaf_keys_kernel, grid(n_cells), block, 0, st, d_cells, n_cells, lkey0, lval_idx);
  MF_HIP_TRY(hipGetLastError());
  size_t tb1 = 0, tb2 = 0, tb3 = 0;
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb1, lkey0, lkeys, lval_idx, lidx,
                                                int(n_cells), 0, 64, st));
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, flag, pos, int(nn) + 1, st));
  char *tmp;
  const size_t tmax = std::max(tb1, tb2);
  MF_HIP_TRY(b.alloc(&tmp, tmax));
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb1, lkey0, lkeys, lval_idx, lidx,
                                                int(n_cells), 0, 64, st));
  hipLaunchKernelGGL(fill_u64_kernel, grid(nn), block, 0, st, reinterpret_cast<uint64_t *>(claim),
                     ~0ull, nn);
  MF_HIP_TRY(hipGetLastError());
  MF_HIP_TRY(hipMemsetAsync(d_line_of, 0xff, sizeof(int32_t) * size_t(nn), st));
  MF_HIP_TRY(hipMemsetAsync(flag + nn, 0, sizeof(int32_t), st));
  MF_HIP_TRY(hipMemsetAsync(unbal, 0, sizeof(int), st));
  hipLaunchKernelGGL(hnc_claim_kernel, grid(n_cells * 27), block, 0, st, a, d_cells, d_nodes,
                     lkeys, lidx, claim, unbal);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(hnc_flag_kernel, grid(nn), block, 0, st, claim, nn, flag);
  MF_HIP_TRY(hipGetLastError());
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tb2, flag, pos, int(nn) + 1, st));
  int32_t n_lines32 = 0;
  int unbalanced = 0;
  MF_HIP_TRY(hipMemcpyAsync(&n_lines32, pos + nn, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipMemcpyAsync(&unbalanced, unbal, sizeof(int), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  MF_REQUIRE(!unbalanced, MF_ERR_ARG, "adaptive mesh is not 2:1 balanced");
  const int64_t n_lines = n_lines32;
  line_of.assign(size_t(nn), -1);
  line_ptr.assign(1, 0u);
  line_col.clear();
  line_val.clear();
  if (n_lines == 0) return MF_OK;
  MF_HIP_TRY(b.alloc(&skey0, size_t(n_lines)));
  MF_HIP_TRY(b.alloc(&skey, size_t(n_lines)));
  MF_HIP_TRY(b.alloc(&snode0, size_t(n_lines)));
  MF_HIP_TRY(b.alloc(&snode, size_t(n_lines)));
  MF_HIP_TRY(b.alloc(&cnt, size_t(n_lines) + 1));
  MF_HIP_TRY(b.alloc(&d_ptr, size_t(n_lines) + 1));
  hipLaunchKernelGGL(hnc_compact_kernel, grid(nn), block, 0, st, claim, flag, pos, nn, skey0, snode0);
  MF_HIP_TRY(hipGetLastError());
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb3, skey0, skey, snode0, snode,
                                                int(n_lines), 0, 64, st));
  size_t tb4 = 0;
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb4, cnt, d_ptr, int(n_lines) + 1, st));
  char *tmp2;
  MF_HIP_TRY(b.alloc(&tmp2, std::max(tb3, tb4)));
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp2, tb3, skey0, skey, snode0, snode,
                                                int(n_lines), 0, 64, st));
  MF_HIP_TRY(hipMemsetAsync(cnt + n_lines, 0, sizeof(uint32_t), st));
  hipLaunchKernelGGL((hnc_line_kernel<0>), grid(n_lines), block, 0, st, a, d_cells, d_nodes, lkeys,
                     lidx, skey, snode, n_lines, d_line_of, cnt, (const uint32_t *)nullptr,
                     (uint32_t *)nullptr, (double *)nullptr);
  MF_HIP_TRY(hipGetLastError());
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp2, tb4, cnt, d_ptr, int(n_lines) + 1, st));
  uint32_t n_entries = 0;
  MF_HIP_TRY(hipMemcpyAsync(&n_entries, d_ptr + n_lines, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  MF_HIP_TRY(b.alloc(&d_col, size_t(n_entries)));
  MF_HIP_TRY(b.alloc(&d_val, size_t(n_entries)));
  hipLaunchKernelGGL((hnc_line_kernel<1>), grid(n_lines), block, 0, st, a, d_cells, d_nodes, lkeys,
                     lidx, skey, snode, n_lines, (int32_t *)nullptr, (uint32_t *)nullptr, d_ptr,
                     d_col, d_val);
  MF_HIP_TRY(hipGetLastError());
  line_ptr.resize(size_t(n_lines) + 1);
  line_col.resize(n_entries);
  line_val.resize(n_entries);
  MF_HIP_TRY(hipMemcpyAsync(line_of.data(), d_line_of, sizeof(int32_t) * size_t(nn),
                            hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipMemcpyAsync(line_ptr.data(), d_ptr, sizeof(uint32_t) * (size_t(n_lines) + 1),
                            hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipMemcpyAsync(line_col.data(), d_col, sizeof(uint32_t) * n_entries,
                            hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipMemcpyAsync(line_val.data(), d_val, sizeof(double) * n_entries,
                            hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  return MF_OK;
}

// Hanging masks of the cells (bit e: local node e is hanging) and the scatter targets of the
// hanging cells, the tables the cell kernel's hanging pass reads: for the
// h-th hanging cell (cells in index order) every distinct node it writes through its constraint
// lines, ascending, with the (local node, weight) entries that sum into it, in the host's order
// (a stable sort of the cell's entries by node).  Here all entries of all hanging cells are
// keyed (h << 32 | node) and sorted stably in one radix sort, whose runs are the targets.
namespace {
__global__ void tgt_count_kernel(const uint32_t *__restrict__ cell_nodes,
                                 const int32_t *__restrict__ node_line,
                                 const uint32_t *__restrict__ line_ptr, int64_t n_cells, int nloc,
                                 uint64_t *hmask, uint32_t *hflag, uint32_t *ecount)
{
  for (int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; c < n_cells;
       c += int64_t(gridDim.x) * blockDim.x) {
    uint32_t cnt = 0;
    uint64_t mask = 0;
    for (int e = 0; e < nloc; ++e) {
      const int32_t l = node_line[cell_nodes[size_t(c) * nloc + e]];
      if (l >= 0) {
        mask |= 1ull << e;
        cnt += line_ptr[l + 1] - line_ptr[l];
      } else {
        cnt += 1;
      }
    }
    hmask[c] = mask;
    hflag[c] = mask ? 1u : 0u;
    ecount[c] = mask ? cnt : 0u;
  }
}

__global__ void tgt_fill_kernel(const uint32_t *__restrict__ cell_nodes,
                                const int32_t *__restrict__ node_line,
                                const uint32_t *__restrict__ line_ptr,
                                const uint32_t *__restrict__ line_col,
                                const double *__restrict__ line_val, int64_t n_cells, int nloc,
                                const uint32_t *__restrict__ hflag, const uint32_t *__restrict__ hpos,
                                const uint32_t *__restrict__ eoff, int32_t *cell_hidx,
                                uint32_t *hcell, uint64_t *key, uint32_t *idx, uint8_t *loc, double *w)
{
  for (int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; c < n_cells;
       c += int64_t(gridDim.x) * blockDim.x) {
    if (!hflag[c]) {
      cell_hidx[c] = -1;
      continue;
    }
    const uint32_t h = hpos[c];
    cell_hidx[c] = int32_t(h);
    hcell[h] = uint32_t(c);
    uint32_t q = eoff[c];
    for (int e = 0; e < nloc; ++e) {
      const uint32_t id = cell_nodes[size_t(c) * nloc + e];
      const int32_t l = node_line[id];
      if (l >= 0) {
        for (uint32_t t = line_ptr[l]; t < line_ptr[l + 1]; ++t, ++q) {
          key[q] = (uint64_t(h) << 32) | line_col[t];
          idx[q] = q;
          loc[q] = uint8_t(e);
          w[q] = line_val[t];
        }
      } else {
        key[q] = (uint64_t(h) << 32) | id;
        idx[q] = q;
        loc[q] = uint8_t(e);
        w[q] = 1.0;
        ++q;
      }
    }
  }
}

__global__ void tgt_gather_kernel(const uint64_t *__restrict__ key_s, const uint32_t *__restrict__ idx_s,
                                  const uint8_t *__restrict__ loc, const double *__restrict__ w,
                                  int64_t n, uint8_t *ent_local, double *ent_w, float *ent_w32,
                                  uint32_t *head)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint32_t j = idx_s[i];
    ent_local[i] = loc[j];
    ent_w[i] = w[j];
    ent_w32[i] = float(w[j]);
    head[i] = (i == 0 || key_s[i] != key_s[i - 1]) ? 1u : 0u;
  }
}

__global__ void tgt_heads_kernel(const uint64_t *__restrict__ key_s, const uint32_t *__restrict__ head,
                                 const uint32_t *__restrict__ tpos, const uint32_t *__restrict__ hcell,
                                 int64_t n, uint32_t *ht_ptr, uint32_t *tgt_node, uint32_t *tgt_ptr,
                                 uint32_t *slot_cell, uint32_t *slot_tgt)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (!head[i]) continue;
    const uint32_t t = tpos[i];
    const uint32_t h = uint32_t(key_s[i] >> 32);
    tgt_node[t] = uint32_t(key_s[i] & 0xffffffffull);
    tgt_ptr[t] = uint32_t(i);
    slot_tgt[t] = t;
    slot_cell[t] = hcell[h];
    if (i == 0 || uint32_t(key_s[i - 1] >> 32) != h) ht_ptr[h] = t;
  }
}

__global__ void tgt_tail_kernel(uint32_t *ht_ptr, uint32_t *tgt_ptr, int64_t n_hc, uint32_t n_t,
                                uint32_t n_e)
{
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ht_ptr[n_hc] = n_t;
    tgt_ptr[n_t] = n_e;
  }
}

template <typename T>
hipError_t keep_on_device(T **d, const std::vector<T> &h, hipStream_t st)
{
  if (*d || h.empty()) return hipSuccess;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(d), sizeof(T) * h.size());
  if (e != hipSuccess) return e;
  return hipMemcpyAsync(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st);
}

template <typename T>
hipError_t alloc_kept(T **d, size_t n)
{
  return hipMalloc(reinterpret_cast<void **>(d), sizeof(T) * (n ? n : 1));
}
}  // namespace

int device_hanging_targets(AdaptiveData *A, int64_t n_cells, int nloc, hipStream_t st)
{
  // the inputs stay on the device as the operator's tables
  MF_HIP_TRY(keep_on_device(&A->d.cell_nodes, A->cell_nodes, st));
  MF_HIP_TRY(keep_on_device(&A->d.node_line, A->node_line, st));
  MF_HIP_TRY(keep_on_device(&A->d.line_ptr, A->line_ptr, st));
  MF_HIP_TRY(keep_on_device(&A->d.line_col, A->line_col, st));
  MF_HIP_TRY(keep_on_device(&A->d.line_val, A->line_val, st));
  MF_HIP_TRY(alloc_kept(&A->d.cell_hidx, size_t(n_cells)));
  MF_HIP_TRY(alloc_kept(&A->d.cell_hmask, size_t(n_cells)));
  DevBufs b;
  uint32_t *hflag, *hpos, *ecount, *eoff;
  MF_HIP_TRY(b.alloc(&hflag, size_t(n_cells) + 1));
  MF_HIP_TRY(b.alloc(&hpos, size_t(n_cells) + 1));
  MF_HIP_TRY(b.alloc(&ecount, size_t(n_cells) + 1));
  MF_HIP_TRY(b.alloc(&eoff, size_t(n_cells) + 1));
  const dim3 block(256);
  auto grid = [](int64_t n) { return dim3(unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)))); };
  MF_HIP_TRY(hipMemsetAsync(hflag + n_cells, 0, sizeof(uint32_t), st));
  MF_HIP_TRY(hipMemsetAsync(ecount + n_cells, 0, sizeof(uint32_t), st));
  // a cell without hanging nodes reads an empty line table: node_line is then all -1
  hipLaunchKernelGGL(tgt_count_kernel, grid(n_cells), block, 0, st, A->d.cell_nodes, A->d.node_line,
                     A->d.line_ptr, n_cells, nloc, A->d.cell_hmask, hflag, ecount);
  MF_HIP_TRY(hipGetLastError());
  A->cell_hmask.resize(size_t(n_cells));
  MF_HIP_TRY(hipMemcpyAsync(A->cell_hmask.data(), A->d.cell_hmask, sizeof(uint64_t) * size_t(n_cells),
                            hipMemcpyDeviceToHost, st));
  size_t tb = 0;
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, hflag, hpos, int(n_cells) + 1, st));
  char *tmp;
  MF_HIP_TRY(b.alloc(&tmp, tb));
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tb, hflag, hpos, int(n_cells) + 1, st));
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tb, ecount, eoff, int(n_cells) + 1, st));
  uint32_t tot[2] = {0, 0};
  MF_HIP_TRY(hipMemcpyAsync(&tot[0], hpos + n_cells, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipMemcpyAsync(&tot[1], eoff + n_cells, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  const int64_t n_hc = tot[0], n_e = tot[1];
  MF_REQUIRE(n_e < (int64_t(1) << 31), MF_ERR_ARG, "too many hanging-cell entries");
  uint32_t *hcell, *idx, *idx_s, *head, *tpos;
  uint64_t *key, *key_s;
  uint8_t *loc;
  double *w;
  MF_HIP_TRY(b.alloc(&hcell, size_t(n_hc)));
  MF_HIP_TRY(b.alloc(&key, size_t(n_e)));
  MF_HIP_TRY(b.alloc(&key_s, size_t(n_e)));
  MF_HIP_TRY(b.alloc(&idx, size_t(n_e)));
  MF_HIP_TRY(b.alloc(&idx_s, size_t(n_e)));
  MF_HIP_TRY(b.alloc(&loc, size_t(n_e)));
  MF_HIP_TRY(b.alloc(&w, size_t(n_e)));
  MF_HIP_TRY(b.alloc(&head, size_t(n_e) + 1));
  MF_HIP_TRY(b.alloc(&tpos, size_t(n_e) + 1));
  hipLaunchKernelGGL(tgt_fill_kernel, grid(n_cells), block, 0, st, A->d.cell_nodes, A->d.node_line,
                     A->d.line_ptr, A->d.line_col, A->d.line_val, n_cells, nloc, hflag, hpos, eoff,
                     A->d.cell_hidx, hcell, key, idx, loc, w);
  MF_HIP_TRY(hipGetLastError());
  A->slot_start.assign(2, 0);
  if (n_e == 0) return MF_OK;
  int hbits = 1;
  while ((int64_t(1) << hbits) <= n_hc) ++hbits;
  size_t tb2 = 0, tb3 = 0;
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, key, key_s, idx, idx_s, int(n_e), 0,
                                                32 + hbits, st));
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, head, tpos, int(n_e) + 1, st));
  char *tmp2;
  MF_HIP_TRY(b.alloc(&tmp2, std::max(tb2, tb3)));
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp2, tb2, key, key_s, idx, idx_s, int(n_e), 0,
                                                32 + hbits, st));
  MF_HIP_TRY(alloc_kept(&A->d.ent_local, size_t(n_e)));
  MF_HIP_TRY(alloc_kept(&A->d.ent_w, size_t(n_e)));
  MF_HIP_TRY(alloc_kept(&A->d.ent_w32, size_t(n_e)));
  MF_HIP_TRY(hipMemsetAsync(head + n_e, 0, sizeof(uint32_t), st));
  hipLaunchKernelGGL(tgt_gather_kernel, grid(n_e), block, 0, st, key_s, idx_s, loc, w, n_e,
                     A->d.ent_local, A->d.ent_w, A->d.ent_w32, head);
  MF_HIP_TRY(hipGetLastError());
  MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp2, tb3, head, tpos, int(n_e) + 1, st));
  uint32_t n_t32 = 0;
  MF_HIP_TRY(hipMemcpyAsync(&n_t32, tpos + n_e, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  const int64_t n_t = n_t32;
  MF_HIP_TRY(alloc_kept(&A->d.ht_ptr, size_t(n_hc) + 1));
  MF_HIP_TRY(alloc_kept(&A->d.tgt_node, size_t(n_t)));
  MF_HIP_TRY(alloc_kept(&A->d.tgt_ptr, size_t(n_t) + 1));
  MF_HIP_TRY(alloc_kept(&A->d.slot_cell, size_t(n_t)));
  MF_HIP_TRY(alloc_kept(&A->d.slot_tgt, size_t(n_t)));
  hipLaunchKernelGGL(tgt_heads_kernel, grid(n_e), block, 0, st, key_s, head, tpos, hcell, n_e,
                     A->d.ht_ptr, A->d.tgt_node, A->d.tgt_ptr, A->d.slot_cell, A->d.slot_tgt);
  MF_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(tgt_tail_kernel, dim3(1), dim3(64), 0, st, A->d.ht_ptr, A->d.tgt_ptr, n_hc,
                     uint32_t(n_t), uint32_t(n_e));
  MF_HIP_TRY(hipGetLastError());
  MF_HIP_TRY(hipStreamSynchronize(st));
  A->slot_start[1] = n_t;
  return MF_OK;
}

// Band refinement of the base mesh with 2:1 balance across faces, edges and vertices, the same
// leaves in the same (Morton) order as the host path: every level, the leaves of that level
// that meet the shell are split into their 8 children (a count, a scan and a fill); then, until
// nothing changes, every leaf of level l >= 2 marks each neighbouring leaf of level < l - 1
// (found by binary search of the radix-sorted leaf keys), and the marked leaves are split.
namespace {
struct BandArgs {
  int ncell[3];
  double lo[3], hi[3], center[3], radius, half_width;
};

__device__ __forceinline__ int key_level(uint64_t k) { return int(k >> 60); }
__device__ __forceinline__ int64_t key_x(uint64_t k, int d)
{
  return int64_t((k >> (40 - 20 * d)) & 0xFFFFFull);
}

// whether the cell (l, ix) meets the shell |x - center| in (radius - half_width, radius +
// half_width): the host's formula, uncontracted, so the decision is the host's bit for bit
__device__ bool meets_band_d(const BandArgs &b, int l, const int64_t *ix)
{
#pragma clang fp contract(off)
  double dmin2 = 0, dmax2 = 0;
  for (int d = 0; d < 3; ++d) {
    const double h = (b.hi[d] - b.lo[d]) / b.ncell[d] / double(int64_t(1) << l);
    const double a = b.lo[d] + h * double(ix[d]), e = a + h;
    const double nr = fmin(fmax(b.center[d], a), e);
    dmin2 += (nr - b.center[d]) * (nr - b.center[d]);
    const double fr = fabs(a - b.center[d]) > fabs(e - b.center[d]) ? a : e;
    dmax2 += (fr - b.center[d]) * (fr - b.center[d]);
  }
  return sqrt(dmax2) > b.radius - b.half_width && sqrt(dmin2) < b.radius + b.half_width;
}

__global__ void base_leaves_kernel(BandArgs b, uint64_t *keys, int64_t n)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const int64_t x = i % b.ncell[0], y = (i / b.ncell[0]) % b.ncell[1], z = i / (int64_t(b.ncell[0]) * b.ncell[1]);
    keys[i] = leaf_key_d(0, x, y, z);
  }
}

__global__ void band_mark_kernel(BandArgs b, const uint64_t *keys, int64_t n, int lev, uint32_t *cnt)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    const int64_t ix[3] = {key_x(k, 0), key_x(k, 1), key_x(k, 2)};
    cnt[i] = (key_level(k) == lev && meets_band_d(b, lev, ix)) ? 8u : 1u;
  }
}

// cnt[i] = 8: split leaf i into its children at pos[i], else copy it
__global__ void split_kernel(const uint64_t *keys, const uint32_t *cnt, const uint32_t *pos,
                             int64_t n, uint64_t *out)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    if (cnt[i] == 1u) {
      out[pos[i]] = k;
      continue;
    }
    const int l = key_level(k);
    for (int c = 0; c < 8; ++c)
      out[pos[i] + c] = leaf_key_d(l + 1, 2 * key_x(k, 0) + (c & 1), 2 * key_x(k, 1) + ((c >> 1) & 1),
                                   2 * key_x(k, 2) + (c >> 2));
  }
}

__device__ int64_t find_key(const uint64_t *keys, int64_t n, uint64_t k)
{
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && keys[lo] == k) ? lo : -1;
}

// sorted keys: cnt[j] = 8 for every leaf j that is two or more levels coarser than a leaf it
// touches, else 1; *any = 1 when some leaf is marked
__global__ void balance_mark_kernel(BandArgs b, const uint64_t *keys, int64_t n, uint32_t *cnt,
                                    int *any)
{
  for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n * 27;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = t / 27;
    const int dir = int(t - i * 27);
    const uint64_t k = keys[i];
    const int l = key_level(k);
    if (l < 2 || dir == 13) continue;
    const int del[3] = {dir % 3 - 1, (dir / 3) % 3 - 1, dir / 9 - 1};
    int64_t nb[3];
    bool inside = true;
    for (int d = 0; d < 3; ++d) {
      nb[d] = key_x(k, d) + del[d];
      inside &= nb[d] >= 0 && nb[d] < (int64_t(b.ncell[d]) << l);
    }
    if (!inside) continue;
    for (int lc = l; lc >= 0; --lc) {
      const int sh = l - lc;
      const int64_t j = find_key(keys, n, leaf_key_d(lc, nb[0] >> sh, nb[1] >> sh, nb[2] >> sh));
      if (j >= 0) {
        if (lc < l - 1) {
          cnt[j] = 8u;
          *any = 1;
        }
        break;
      }
    }
  }
}

__global__ void fill_u32_kernel(uint32_t *x, uint32_t v, int64_t n)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[i] = v;
}

__global__ void morton_kernel(const uint64_t *keys, int64_t n, int L, uint64_t *code)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    const int sh = L - key_level(k);
    uint64_t c = 0;
    for (int bit = 0; bit < 21; ++bit)
      for (int d = 0; d < 3; ++d) c |= ((uint64_t(key_x(k, d) << sh) >> bit) & 1ull) << (3 * bit + d);
    code[i] = c;
  }
}

__global__ void cells_kernel(const uint64_t *keys, int64_t n, int32_t *cells)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t k = keys[i];
    cells[4 * i] = key_level(k);
    for (int d = 0; d < 3; ++d) cells[4 * i + 1 + d] = int32_t(key_x(k, d));
  }
}

__global__ void max_level_kernel(const uint64_t *keys, int64_t n, int *lmax)
{
  int m = 0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    m = max(m, key_level(keys[i]));
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(lmax, m);
}
}  // namespace

int device_refine_band(const int ncell[3], const double lo[3], const double hi[3], int levels,
                       const double center[3], double radius, double half_width,
                       std::vector<int32_t> &cells_out)
{
  BandArgs b{};
  for (int d = 0; d < 3; ++d) {
    b.ncell[d] = ncell[d];
    b.lo[d] = lo[d];
    b.hi[d] = hi[d];
    b.center[d] = center[d];
  }
  b.radius = radius;
  b.half_width = half_width;
  hipStream_t st = nullptr;
  MF_HIP_TRY(hipStreamCreate(&st));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } guard{st};
  DevBufs b0;
  const dim3 block(256);
  auto grid = [](int64_t n) { return dim3(unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)))); };
  int64_t n = int64_t(ncell[0]) * ncell[1] * ncell[2];
  // buffers that grow: reallocated (with the old ones freed) when a step needs more
  struct Grow {
    void *p = nullptr;
    size_t bytes = 0;
    ~Grow()
    {
      if (p) (void)hipFree(p);
    }
    hipError_t need(size_t nb)
    {
      if (nb <= bytes) return hipSuccess;
      if (p) (void)hipFree(p);
      p = nullptr;
      bytes = 0;
      const hipError_t e = hipMalloc(&p, nb);
      if (e == hipSuccess) bytes = nb;
      return e;
    }
  } gkeys, gout, gcnt, gpos, gtmp, gcode;
  int *any;
  MF_HIP_TRY(b0.alloc(&any, 2));
  MF_HIP_TRY(gkeys.need(sizeof(uint64_t) * size_t(n)));
  hipLaunchKernelGGL(base_leaves_kernel, grid(n), block, 0, st, b, static_cast<uint64_t *>(gkeys.p), n);
  MF_HIP_TRY(hipGetLastError());
  // split the leaves with cnt = 8: scan, fill into gout, swap with gkeys
  auto split = [&]() -> int {
    // every leaf may become 8: the 32-bit counts and their scan stay exact below 2^28 leaves
    MF_REQUIRE(n < (int64_t(1) << 28), MF_ERR_ARG, "adaptive mesh too large");
    uint32_t *cnt = static_cast<uint32_t *>(gcnt.p), *pos = static_cast<uint32_t *>(gpos.p);
    size_t tb = 0;
    MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, pos, int(n) + 1, st));
    MF_HIP_TRY(gtmp.need(tb));
    MF_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(gtmp.p, tb, cnt, pos, int(n) + 1, st));
    uint32_t total = 0;
    MF_HIP_TRY(hipMemcpyAsync(&total, pos + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    MF_HIP_TRY(hipStreamSynchronize(st));
    MF_REQUIRE(int64_t(total) < (int64_t(1) << 31), MF_ERR_ARG, "adaptive mesh too large");
    MF_HIP_TRY(gout.need(sizeof(uint64_t) * size_t(total)));
    hipLaunchKernelGGL(split_kernel, grid(n), block, 0, st, static_cast<const uint64_t *>(gkeys.p), cnt,
                       pos, n, static_cast<uint64_t *>(gout.p));
    MF_HIP_TRY(hipGetLastError());
    std::swap(gkeys.p, gout.p);
    std::swap(gkeys.bytes, gout.bytes);
    n = int64_t(total);
    return MF_OK;
  };
  for (int lev = 0; lev < levels; ++lev) {
    MF_HIP_TRY(gcnt.need(sizeof(uint32_t) * (size_t(n) + 1)));
    MF_HIP_TRY(gpos.need(sizeof(uint32_t) * (size_t(n) + 1)));
    MF_HIP_TRY(hipMemsetAsync(static_cast<uint32_t *>(gcnt.p) + n, 0, sizeof(uint32_t), st));
    hipLaunchKernelGGL(band_mark_kernel, grid(n), block, 0, st, b, static_cast<const uint64_t *>(gkeys.p),
                       n, lev, static_cast<uint32_t *>(gcnt.p));
    MF_HIP_TRY(hipGetLastError());
    const int rc = split();
    if (rc != MF_OK) return rc;
  }
  // sort the leaf keys (into gout, then swapped back)
  auto sort_keys = [&]() -> int {
    MF_HIP_TRY(gout.need(sizeof(uint64_t) * size_t(n)));
    size_t tb = 0;
    MF_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, static_cast<const uint64_t *>(gkeys.p),
                                                 static_cast<uint64_t *>(gout.p), int(n), 0, 64, st));
    MF_HIP_TRY(gtmp.need(tb));
    MF_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(gtmp.p, tb, static_cast<const uint64_t *>(gkeys.p),
                                                 static_cast<uint64_t *>(gout.p), int(n), 0, 64, st));
    std::swap(gkeys.p, gout.p);
    std::swap(gkeys.bytes, gout.bytes);
    return MF_OK;
  };
  for (int pass = 0;; ++pass) {
    MF_REQUIRE(pass < 64, MF_ERR_ARG, "2:1 balance does not terminate");
    int rc = sort_keys();
    if (rc != MF_OK) return rc;
    MF_HIP_TRY(gcnt.need(sizeof(uint32_t) * (size_t(n) + 1)));
    MF_HIP_TRY(gpos.need(sizeof(uint32_t) * (size_t(n) + 1)));
    hipLaunchKernelGGL(fill_u32_kernel, grid(n), block, 0, st, static_cast<uint32_t *>(gcnt.p), 1u, n);
    MF_HIP_TRY(hipGetLastError());
    MF_HIP_TRY(hipMemsetAsync(static_cast<uint32_t *>(gcnt.p) + n, 0, sizeof(uint32_t), st));
    MF_HIP_TRY(hipMemsetAsync(any, 0, sizeof(int), st));
    hipLaunchKernelGGL(balance_mark_kernel, grid(n * 27), block, 0, st, b,
                       static_cast<const uint64_t *>(gkeys.p), n, static_cast<uint32_t *>(gcnt.p), any);
    MF_HIP_TRY(hipGetLastError());
    int h_any = 0;
    MF_HIP_TRY(hipMemcpyAsync(&h_any, any, sizeof(int), hipMemcpyDeviceToHost, st));
    MF_HIP_TRY(hipStreamSynchronize(st));
    if (!h_any) break;
    rc = split();
    if (rc != MF_OK) return rc;
  }
  // Morton order of the cell origins on the finest lattice
  MF_HIP_TRY(hipMemsetAsync(any + 1, 0, sizeof(int), st));
  hipLaunchKernelGGL(max_level_kernel, grid(n), block, 0, st, static_cast<const uint64_t *>(gkeys.p), n,
                     any + 1);
  MF_HIP_TRY(hipGetLastError());
  int L = 0;
  MF_HIP_TRY(hipMemcpyAsync(&L, any + 1, sizeof(int), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  MF_HIP_TRY(gcode.need(sizeof(uint64_t) * 2 * size_t(n)));
  uint64_t *code = static_cast<uint64_t *>(gcode.p), *code_s = code + n;
  hipLaunchKernelGGL(morton_kernel, grid(n), block, 0, st, static_cast<const uint64_t *>(gkeys.p), n, L,
                     code);
  MF_HIP_TRY(hipGetLastError());
  MF_HIP_TRY(gout.need(sizeof(uint64_t) * size_t(n)));
  size_t tb = 0;
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, code, code_s, static_cast<const uint64_t *>(gkeys.p),
                                                static_cast<uint64_t *>(gout.p), int(n), 0, 63, st));
  MF_HIP_TRY(gtmp.need(tb));
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(gtmp.p, tb, code, code_s, static_cast<const uint64_t *>(gkeys.p),
                                                static_cast<uint64_t *>(gout.p), int(n), 0, 63, st));
  int32_t *dcells;
  MF_HIP_TRY(b0.alloc(&dcells, 4 * size_t(n)));
  hipLaunchKernelGGL(cells_kernel, grid(n), block, 0, st, static_cast<const uint64_t *>(gout.p), n, dcells);
  MF_HIP_TRY(hipGetLastError());
  cells_out.resize(4 * size_t(n));
  MF_HIP_TRY(hipMemcpyAsync(cells_out.data(), dcells, sizeof(int32_t) * 4 * size_t(n),
                            hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  return MF_OK;
}

// Transfer of a finite-element field between two adaptive meshes of the same base mesh and
// degree: every node of the new mesh takes the value of the old (constraint-consistent) field
// at its support point, found in the old leaf that contains it.  For FE_Q this is what
// SolutionTransfer does on refinement (prolongation interpolates the parent's function) and
// on coarsening (the parent's nodal values are the children's function at its nodes); the
// caller distributes the new mesh's constraints afterwards, as the reference's post() does.
namespace {
struct TransferArgs {
  int p, K, L;
  int ncell[3];
  double lo[3], h[3], gl[KMAX];
};

__global__ void transfer_eval_kernel(const TransferArgs a, const uint64_t *__restrict__ keys,
                                     const uint32_t *__restrict__ idx, int64_t n_leaves,
                                     const uint32_t *__restrict__ cell_nodes,
                                     const double *__restrict__ val, const double *__restrict__ xyz,
                                     int64_t n, double *__restrict__ out)
{
  const int nloc = a.K * a.K * a.K;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    double u[3];
    for (int d = 0; d < 3; ++d) u[d] = (xyz[3 * i + d] - a.lo[d]) / a.h[d];  // base-cell units
    int64_t c = -1, ix[3] = {0, 0, 0};
    int lev = 0;
    for (int l = a.L; l >= 0 && c < 0; --l) {
      const double sc = double(int64_t(1) << l);
      for (int d = 0; d < 3; ++d) {
        const int64_t top = (int64_t(a.ncell[d]) << l) - 1;
        ix[d] = min(max(int64_t(floor(u[d] * sc)), int64_t(0)), top);
      }
      c = find_leaf(keys, idx, n_leaves, leaf_key_d(l, ix[0], ix[1], ix[2]));
      lev = l;
    }
    if (c < 0) {  // cannot happen for a point of the domain: the old leaves tile it
      out[i] = 0.0;
      continue;
    }
    double wv[3][KMAX];
    for (int d = 0; d < 3; ++d) {
      const double t = fmin(fmax(u[d] * double(int64_t(1) << lev) - double(ix[d]), 0.0), 1.0);
      for (int j = 0; j < a.K; ++j) {
        double v = 1.0;
        for (int m = 0; m < a.K; ++m)
          if (m != j) v *= (t - a.gl[m]) / (a.gl[j] - a.gl[m]);
        wv[d][j] = v;
      }
    }
    double s = 0.0;
    for (int e = 0; e < nloc; ++e)
      s += val[cell_nodes[size_t(c) * nloc + e]] * (wv[0][e % a.K] * wv[1][(e / a.K) % a.K] *
                                                    wv[2][e / (a.K * a.K)]);
    out[i] = s;
  }
}
}  // namespace

int device_transfer_field(const mf_scratch *old_s, const double *old_val, const mf_scratch *new_s,
                          double *new_vec)
{
  const AdaptiveData *A = old_s->ad;
  const MeshInfo &m = old_s->mesh;
  hipStream_t st = new_s->stream;
  const int64_t n_leaves = m.n_cells, n_new = new_s->mesh.n_dofs;
  TransferArgs a{};
  a.p = m.degree;
  a.K = m.degree + 1;
  a.L = A->levels;
  for (int d = 0; d < 3; ++d) {
    a.ncell[d] = m.ncell[d];
    a.lo[d] = m.lo[d];
    a.h[d] = m.h[d];
  }
  for (int i = 0; i < a.K; ++i) a.gl[i] = old_s->tab.gl[i];
  DevBufs b;
  int32_t *d_cells;
  uint64_t *k0, *k1;
  uint32_t *v0, *v1;
  double *xyz;
  MF_HIP_TRY(b.alloc(&d_cells, 4 * size_t(n_leaves)));
  MF_HIP_TRY(b.alloc(&k0, size_t(n_leaves)));
  MF_HIP_TRY(b.alloc(&k1, size_t(n_leaves)));
  MF_HIP_TRY(b.alloc(&v0, size_t(n_leaves)));
  MF_HIP_TRY(b.alloc(&v1, size_t(n_leaves)));
  MF_HIP_TRY(hipMemcpyAsync(d_cells, A->cells.data(), sizeof(int32_t) * 4 * size_t(n_leaves),
                            hipMemcpyHostToDevice, st));
  // the new mesh's support points: its device copy when node identification ran there
  const double *xyz_new = new_s->ad->d.xyz;
  if (!xyz_new) {
    MF_HIP_TRY(b.alloc(&xyz, 3 * size_t(n_new)));
    MF_HIP_TRY(hipMemcpyAsync(xyz, new_s->ad->xyz.data(), sizeof(double) * 3 * size_t(n_new),
                              hipMemcpyHostToDevice, st));
    xyz_new = xyz;
  }
  const dim3 block(256);
  auto grid = [](int64_t n) { return dim3(unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536)))); };
  hipLaunchKernelGGL(leaf_keys_kernel, grid(n_leaves), block, 0, st, d_cells, n_leaves, k0, v0);
  MF_HIP_TRY(hipGetLastError());
  size_t tb = 0;
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, int(n_leaves), 0, 64, st));
  char *tmp;
  MF_HIP_TRY(b.alloc(&tmp, tb));
  MF_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, int(n_leaves), 0, 64, st));
  hipLaunchKernelGGL(transfer_eval_kernel, grid(n_new), block, 0, st, a, k1, v1, n_leaves,
                     A->d.cell_nodes, old_val, xyz_new, n_new, new_vec);
  MF_HIP_TRY(hipGetLastError());
  MF_HIP_TRY(hipStreamSynchronize(st));
  return MF_OK;
}

}  // namespace mf

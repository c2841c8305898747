// Adaptive octree meshes with hanging nodes: mesh generation (band refinement + 2:1 balance),
// node numbering, hanging-node constraints, cell colouring and the device tables the cell
// kernels read.
//
// Reference functions replaced:
//   FiniteElementUtils::distribute_dofs (include/meltpooldg/utilities/fe_util.hpp:21-81)
//   Constraints::make_HNC / merge_HNC_into_DBC (include/meltpooldg/utilities/constraints.hpp:
//     117-151; deal.II DoFTools::make_hanging_node_constraints, AffineConstraints::close)
//   the AMR band refinement of the mp-heat-transfer application
//     (applications/mp-heat-transfer/heat_transfer_application.cpp:484-527) for the benchmark
//     mesh: cells meeting a spherical shell are refined level by level, then 2:1 balanced.
//
// Nodes.  Node i of axis d of a level-l cell with index ix sits at lattice index g = ix*P + i of
// that level.  Vertex-type nodes (i = 0, P, and the midpoint for P = 2) are dyadic and keyed by
// their coordinate in units of half a finest cell; the interior Gauss-Lobatto nodes of P = 3 are
// irrational and only coincide within one level, so they are keyed by (level, g).  Two cell
// nodes are the same DoF iff their three axis keys agree.
//
// Constraints.  With the mesh 2:1 balanced across faces, edges and vertices, a node of a level-l
// cell F is constrained exactly when a level-(l-1) leaf C touches F at an entity containing the
// node and the node is not a node of C; its line interpolates C's Lagrange basis at the node.
// Masters that are constrained themselves are substituted until all masters are free.
// Dirichlet rows are boundary nodes that are not hanging (a hanging line wins the merge).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "mf_internal.hpp"

namespace mf {

namespace {

inline uint64_t leaf_key(int l, int64_t x, int64_t y, int64_t z)
{
  return (uint64_t(l) << 60) | (uint64_t(x) << 40) | (uint64_t(y) << 20) | uint64_t(z);
}

inline uint64_t mix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct LeafSet {
  std::unordered_set<uint64_t> s;
  bool has(int l, int64_t x, int64_t y, int64_t z) const { return s.count(leaf_key(l, x, y, z)); }
  // leaf covering the level-l cell (x,y,z): its level, or -1 (region refined below l)
  int covering(int l, int64_t x, int64_t y, int64_t z) const
  {
    for (int lc = l; lc >= 0; --lc) {
      const int sh = l - lc;
      if (has(lc, x >> sh, y >> sh, z >> sh)) return lc;
    }
    return -1;
  }
};

struct BandSpec {
  int ncell[3];
  double lo[3], hi[3], center[3], radius, half_width;
};

bool meets_band(const BandSpec &b, int l, const int64_t *ix)
{
  double dmin2 = 0, dmax2 = 0;
  for (int d = 0; d < 3; ++d) {
    const double h = (b.hi[d] - b.lo[d]) / b.ncell[d] / double(int64_t(1) << l);
    const double a = b.lo[d] + h * double(ix[d]), e = a + h;
    const double near = std::min(std::max(b.center[d], a), e);
    dmin2 += (near - b.center[d]) * (near - b.center[d]);
    const double far = std::fabs(a - b.center[d]) > std::fabs(e - b.center[d]) ? a : e;
    dmax2 += (far - b.center[d]) * (far - b.center[d]);
  }
  return std::sqrt(dmax2) > b.radius - b.half_width && std::sqrt(dmin2) < b.radius + b.half_width;
}

void refine_into(LeafSet &L, int l, int64_t x, int64_t y, int64_t z)
{
  L.s.erase(leaf_key(l, x, y, z));
  for (int c = 0; c < 8; ++c)
    L.s.insert(leaf_key(l + 1, 2 * x + (c & 1), 2 * y + ((c >> 1) & 1), 2 * z + (c >> 2)));
}

void balance(LeafSet &L, const int ncell[3])
{
  for (;;) {
    std::vector<uint64_t> todo;
    for (uint64_t k : L.s) {
      const int l = int(k >> 60);
      if (l < 2) continue;
      const int64_t x = (k >> 40) & 0xFFFFF, y = (k >> 20) & 0xFFFFF, z = k & 0xFFFFF;
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const int64_t nx = x + dx, ny = y + dy, nz = z + dz;
            if (nx < 0 || ny < 0 || nz < 0 || nx >= (int64_t(ncell[0]) << l) ||
                ny >= (int64_t(ncell[1]) << l) || nz >= (int64_t(ncell[2]) << l))
              continue;
            const int lc = L.covering(l, nx, ny, nz);
            if (lc >= 0 && lc < l - 1) {
              const int sh = l - lc;
              todo.push_back(leaf_key(lc, nx >> sh, ny >> sh, nz >> sh));
            }
          }
    }
    if (todo.empty()) return;
    std::sort(todo.begin(), todo.end());
    todo.erase(std::unique(todo.begin(), todo.end()), todo.end());
    for (uint64_t k : todo)
      if (L.s.count(k))
        refine_into(L, int(k >> 60), (k >> 40) & 0xFFFFF, (k >> 20) & 0xFFFFF, k & 0xFFFFF);
  }
}

// Morton code of the cell origin on the finest lattice (locality of the cell order)
uint64_t morton(int L, int l, const int64_t *ix)
{
  uint64_t code = 0;
  const int sh = L - l;
  for (int b = 0; b < 21; ++b)
    for (int d = 0; d < 3; ++d) code |= ((uint64_t(ix[d] << sh) >> b) & 1ull) << (3 * b + d);
  return code;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// tables of an adaptive scratch
// ------------------------------------------------------------------------------------------
int build_adaptive(mf_scratch *s, const int32_t *cells, int64_t n_cells)
{
  MeshInfo &m = s->mesh;
  const int p = m.degree, K = p + 1, nloc = K * K * K;
  auto *A = new AdaptiveData();
  s->ad = A;
  A->cells.assign(cells, cells + 4 * n_cells);
  int L = 0;
  for (int64_t c = 0; c < n_cells; ++c) L = std::max(L, int(cells[4 * c]));
  A->levels = L;
  if (L > 12) {
    set_error("adaptive mesh: more than 12 refinement levels");
    return MF_ERR_ARG;
  }
  LeafSet leaves;
  for (int64_t c = 0; c < n_cells; ++c) {
    const int32_t *cl = cells + 4 * c;
    for (int d = 0; d < 3; ++d)
      if (cl[0] < 0 || cl[1 + d] < 0 || int64_t(cl[1 + d]) >= (int64_t(m.ncell[d]) << cl[0])) {
        set_error("adaptive mesh: cell index outside the base mesh");
        return MF_ERR_ARG;
      }
    if (!leaves.s.insert(leaf_key(cl[0], cl[1], cl[2], cl[3])).second) {
      set_error("adaptive mesh: duplicate cell");
      return MF_ERR_ARG;
    }
  }
  // --- node identification ---------------------------------------------------------------
  // on the device when one is present (mf_adaptive_dev.hip, the same numbering), else here
  auto axis_key = [&](int l, int64_t ix, int i) -> uint64_t {
    const int64_t g = ix * p + i;
    const bool dyadic = (i == 0 || i == p || (p == 2 && i == 1));
    if (dyadic) return uint64_t((g << (L - l + 1)) / p);
    return (1ull << 62) | (uint64_t(l) << 48) | uint64_t(g);
  };
  std::vector<uint64_t> nkeys;  // 3 per node
  int64_t nn = 0;
  bool on_device = false;
  int n_devices = 0;
  if (g_adaptive_setup_device && hipGetDeviceCount(&n_devices) == hipSuccess && n_devices > 0)
    on_device = device_node_identification(s, cells, n_cells, L, A, nkeys, nn) == MF_OK;
  if (!on_device) {
  nkeys.reserve(size_t(n_cells) * (p == 1 ? 3 : 3 * p * p * p) + 64);
  size_t cap = 1;
  while (cap < size_t(2) * size_t(n_cells) * nloc / std::max(1, p) + 64) cap <<= 1;
  std::vector<int32_t> table(cap, -1);
  A->cell_nodes.resize(size_t(n_cells) * nloc);
  for (int64_t c = 0; c < n_cells; ++c) {
    const int32_t *cl = cells + 4 * c;
    uint64_t ak[3][KMAX];
    for (int d = 0; d < 3; ++d)
      for (int i = 0; i < K; ++i) ak[d][i] = axis_key(cl[0], cl[1 + d], i);
    for (int e = 0; e < nloc; ++e) {
      const uint64_t k0 = ak[0][e % K], k1 = ak[1][(e / K) % K], k2 = ak[2][e / (K * K)];
      size_t h = size_t(mix64(k0 ^ mix64(k1 ^ mix64(k2)))) & (cap - 1);
      for (;;) {
        const int32_t id = table[h];
        if (id < 0) {
          if (nn >= (int64_t(1) << 31) - 1) {
            set_error("adaptive mesh too large for 32-bit DoF indices");
            return MF_ERR_ARG;
          }
          table[h] = int32_t(nn);
          nkeys.push_back(k0);
          nkeys.push_back(k1);
          nkeys.push_back(k2);
          A->cell_nodes[size_t(c) * nloc + e] = uint32_t(nn);
          ++nn;
          if (size_t(nn) * 2 > cap) {  // grow and rehash
            cap <<= 1;
            table.assign(cap, -1);
            for (int64_t q = 0; q < nn; ++q) {
              size_t hq = size_t(mix64(nkeys[3 * q] ^ mix64(nkeys[3 * q + 1] ^ mix64(nkeys[3 * q + 2])))) &
                          (cap - 1);
              while (table[hq] >= 0) hq = (hq + 1) & (cap - 1);
              table[hq] = int32_t(q);
            }
          }
          break;
        }
        if (nkeys[3 * size_t(id)] == k0 && nkeys[3 * size_t(id) + 1] == k1 &&
            nkeys[3 * size_t(id) + 2] == k2) {
          A->cell_nodes[size_t(c) * nloc + e] = uint32_t(id);
          break;
        }
        h = (h + 1) & (cap - 1);
      }
    }
  }
  table.clear();
  table.shrink_to_fit();
  // support points and boundary flags
  A->xyz.resize(size_t(nn) * 3);
  A->boundary.assign(size_t(nn), 0);
  for (int64_t c = 0; c < n_cells; ++c) {
    const int32_t *cl = cells + 4 * c;
    for (int e = 0; e < nloc; ++e) {
      const uint32_t id = A->cell_nodes[size_t(c) * nloc + e];
      const int a[3] = {e % K, (e / K) % K, e / (K * K)};
      for (int d = 0; d < 3; ++d) {
        const double h = m.h[d] / double(int64_t(1) << cl[0]);
        A->xyz[3 * size_t(id) + d] = m.lo[d] + h * (double(cl[1 + d]) + s->tab.gl[a[d]]);
        const int64_t g = int64_t(cl[1 + d]) * p + a[d];
        if ((a[d] == 0 && g == 0) || (a[d] == p && g == (int64_t(m.ncell[d]) << cl[0]) * p))
          A->boundary[id] = 1;
      }
    }
  }
  }  // host node identification
  m.n_dofs = nn;
  m.n_cells = n_cells;
  // --- hanging-node constraints ---------------------------------------------------------
  // lines before chain resolution as CSR, on the device when node identification ran there
  std::vector<int32_t> line_of;
  std::vector<uint32_t> lptr, lcol;
  std::vector<double> lval;
  const bool hnc_on_device =
      on_device && device_hanging_constraints(s, cells, n_cells, L, A->cell_nodes, nn, line_of, lptr,
                                              lcol, lval) == MF_OK;
  if (!hnc_on_device) {
  if (nkeys.empty()) {  // numbered on the device: the node keys from the cells
    nkeys.assign(size_t(3 * nn), 0ull);
    for (int64_t c = 0; c < n_cells; ++c) {
      const int32_t *cl = cells + 4 * c;
      for (int e = 0; e < nloc; ++e) {
        const int a[3] = {e % K, (e / K) % K, e / (K * K)};
        const uint32_t id = A->cell_nodes[size_t(c) * nloc + e];
        for (int d = 0; d < 3; ++d) nkeys[3 * size_t(id) + d] = axis_key(cl[0], cl[1 + d], a[d]);
      }
    }
  }
  std::unordered_map<uint64_t, int64_t> leaf_index;
  leaf_index.reserve(size_t(n_cells) * 2);
  for (int64_t c = 0; c < n_cells; ++c)
    leaf_index[leaf_key(cells[4 * c], cells[4 * c + 1], cells[4 * c + 2], cells[4 * c + 3])] = c;
  line_of.assign(size_t(nn), -1);
  std::vector<std::vector<std::pair<uint32_t, double>>> lines;
  for (int64_t c = 0; c < n_cells; ++c) {
    const int32_t *cl = cells + 4 * c;
    const int l = cl[0];
    if (l == 0) continue;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (!dx && !dy && !dz) continue;
          const int del[3] = {dx, dy, dz};
          int64_t nb[3];
          bool inside = true;
          for (int d = 0; d < 3; ++d) {
            nb[d] = int64_t(cl[1 + d]) + del[d];
            inside &= nb[d] >= 0 && nb[d] < (int64_t(m.ncell[d]) << l);
          }
          if (!inside) continue;
          const int cov = leaves.covering(l, nb[0], nb[1], nb[2]);
          if (cov >= 0 && cov < l - 1) {
            set_error("adaptive mesh is not 2:1 balanced");
            return MF_ERR_ARG;
          }
          if (cov != l - 1) continue;
          const int64_t cc[3] = {nb[0] >> 1, nb[1] >> 1, nb[2] >> 1};  // the coarse leaf
          const int64_t ci = leaf_index.at(leaf_key(l - 1, cc[0], cc[1], cc[2]));
          uint64_t ck[3][KMAX];
          for (int d = 0; d < 3; ++d)
            for (int i = 0; i < K; ++i) ck[d][i] = axis_key(l - 1, cc[d], i);
          for (int e = 0; e < nloc; ++e) {
            const int a[3] = {e % K, (e / K) % K, e / (K * K)};
            bool on = true;
            for (int d = 0; d < 3; ++d)
              on &= (del[d] == 0) || (del[d] < 0 && a[d] == 0) || (del[d] > 0 && a[d] == p);
            if (!on) continue;
            const uint32_t id = A->cell_nodes[size_t(c) * nloc + e];
            if (line_of[id] >= 0) continue;
            double wv[3][KMAX];
            bool node_of_coarse = true;
            for (int d = 0; d < 3; ++d) {
              const uint64_t kk = nkeys[3 * size_t(id) + d];
              int match = -1;
              for (int i = 0; i < K; ++i)
                if (ck[d][i] == kk) match = i;
              if (match >= 0) {
                for (int i = 0; i < K; ++i) wv[d][i] = (i == match) ? 1.0 : 0.0;
              } else {
                node_of_coarse = false;
                // position of the node in the coarse cell's unit coordinates
                const double t = 0.5 * (double(cl[1 + d]) + s->tab.gl[a[d]]) - double(cc[d]);
                double der[KMAX];
                lagrange_basis(K, s->tab.gl, t, wv[d], der);
              }
            }
            if (node_of_coarse) continue;
            line_of[id] = int32_t(lines.size());
            lines.emplace_back();
            auto &ln = lines.back();
            for (int f = 0; f < nloc; ++f) {
              const double w = wv[0][f % K] * wv[1][(f / K) % K] * wv[2][f / (K * K)];
              if (w != 0.0) ln.emplace_back(A->cell_nodes[size_t(ci) * nloc + f], w);
            }
          }
        }
  }
  lptr.assign(1, 0u);
  for (const auto &ln : lines) {
    for (const auto &mw : ln) {
      lcol.push_back(mw.first);
      lval.push_back(mw.second);
    }
    lptr.push_back(uint32_t(lcol.size()));
  }
  }  // host hanging-node constraints
  // chains: substitute constrained masters until every master is free.  Lines are rewritten in
  // place in line order, a rewritten line kept in (rcol, rval) and used by the later lines
  const int64_t n_lines = int64_t(lptr.size()) - 1;
  std::vector<int32_t> ov(size_t(n_lines), -1);
  std::vector<std::vector<uint32_t>> rcol;
  std::vector<std::vector<double>> rval;
  auto span = [&](int64_t l, const uint32_t *&c, const double *&v) -> size_t {
    if (ov[size_t(l)] >= 0) {
      c = rcol[size_t(ov[size_t(l)])].data();
      v = rval[size_t(ov[size_t(l)])].data();
      return rcol[size_t(ov[size_t(l)])].size();
    }
    c = lcol.data() + lptr[size_t(l)];
    v = lval.data() + lptr[size_t(l)];
    return lptr[size_t(l) + 1] - lptr[size_t(l)];
  };
  std::vector<std::pair<uint32_t, double>> out;
  for (int pass = 0; pass < 32; ++pass) {
    bool changed = false;
    for (int64_t l = 0; l < n_lines; ++l) {
      const uint32_t *c;
      const double *v;
      const size_t n = span(l, c, v);
      bool any = false;
      for (size_t i = 0; i < n && !any; ++i) any = line_of[c[i]] >= 0;
      if (!any) continue;
      changed = true;
      out.clear();
      for (size_t i = 0; i < n; ++i) {
        if (line_of[c[i]] >= 0) {
          const uint32_t *c2;
          const double *v2;
          const size_t n2 = span(line_of[c[i]], c2, v2);
          for (size_t j = 0; j < n2; ++j) out.emplace_back(c2[j], v[i] * v2[j]);
        } else {
          out.emplace_back(c[i], v[i]);
        }
      }
      std::sort(out.begin(), out.end(),
                [](const std::pair<uint32_t, double> &a, const std::pair<uint32_t, double> &b) {
                  return a.first < b.first;
                });
      std::vector<uint32_t> mc;
      std::vector<double> mv;
      for (const auto &mw : out) {
        if (!mc.empty() && mc.back() == mw.first)
          mv.back() += mw.second;
        else {
          mc.push_back(mw.first);
          mv.push_back(mw.second);
        }
      }
      if (ov[size_t(l)] < 0) {
        ov[size_t(l)] = int32_t(rcol.size());
        rcol.emplace_back();
        rval.emplace_back();
      }
      rcol[size_t(ov[size_t(l)])].swap(mc);
      rval[size_t(ov[size_t(l)])].swap(mv);
    }
    if (!changed) break;
    if (pass == 31) {
      set_error("adaptive mesh: constraint chains do not terminate");
      return MF_ERR_ARG;
    }
  }
  A->n_hanging = n_lines;
  A->node_line = line_of;
  if (rcol.empty()) {
    A->line_ptr.swap(lptr);
    A->line_col.swap(lcol);
    A->line_val.swap(lval);
  } else {
    A->line_ptr.assign(1, 0u);
    A->line_col.clear();
    A->line_val.clear();
    for (int64_t l = 0; l < n_lines; ++l) {
      const uint32_t *c;
      const double *v;
      const size_t n = span(l, c, v);
      A->line_col.insert(A->line_col.end(), c, c + n);
      A->line_val.insert(A->line_val.end(), v, v + n);
      A->line_ptr.push_back(uint32_t(A->line_col.size()));
    }
  }
  // default Dirichlet rows: boundary nodes that are not hanging
  s->h_mask.assign(size_t(nn), 0);
  for (int64_t i = 0; i < nn; ++i) s->h_mask[i] = (A->boundary[i] && line_of[i] < 0) ? 1 : 0;
  // --- per-cell hanging masks and colouring ------------------------------------------------
  // A cell writes its own K^3 nodes (hanging ones included); the hanging rows are moved to
  // their masters afterwards by a separate gather over the masters.  So cells conflict only
  // when they share a node, as on a conforming mesh.  The colour of a cell is its class
  // (level parity, index parity): cells of one class never share a node (same level: two
  // apart along some axis; levels two or more apart never touch on a 2:1-balanced mesh, which
  // the constraint pass checks), so the classes are conflict-free colours.
  A->cell_level.resize(size_t(n_cells));
  for (int64_t c = 0; c < n_cells; ++c) A->cell_level[c] = uint8_t(cells[4 * c]);
  if (on_device) {
    // the hanging masks with the hanging-cell tables, all kept on the device
    const int rc = device_hanging_targets(A, n_cells, nloc, s->stream);
    if (rc != MF_OK) return rc;
  } else {
    A->cell_hmask.assign(size_t(n_cells), 0);
    for (int64_t c = 0; c < n_cells; ++c)
      for (int e = 0; e < nloc; ++e)
        if (line_of[A->cell_nodes[size_t(c) * nloc + e]] >= 0) A->cell_hmask[c] |= 1ull << e;
  }
  auto cls = [&](int64_t c) {
    const int32_t *cl = cells + 4 * c;
    return (cl[0] & 1) * 8 + (cl[1] & 1) + 2 * (cl[2] & 1) + 4 * (cl[3] & 1);
  };
  int64_t cls_count[16] = {};
  for (int64_t c = 0; c < n_cells; ++c) ++cls_count[cls(c)];
  int cmap[16];
  int n_col = 0;
  for (int k = 0; k < 16; ++k) cmap[k] = cls_count[k] ? n_col++ : -1;
  // transposed constraints: for every master, the hanging nodes (and weights) it receives from
  A->hang_nodes.clear();
  for (int64_t i = 0; i < nn; ++i)
    if (line_of[i] >= 0) A->hang_nodes.push_back(uint32_t(i));
  bool transposed = false;
  if (on_device) {
    std::vector<uint32_t> hang_of_line(static_cast<size_t>(A->n_hanging));
    for (uint32_t h : A->hang_nodes) hang_of_line[size_t(line_of[h])] = h;
    transposed = device_transpose_constraints(A, hang_of_line, s->stream) == MF_OK;
  }
  if (!transposed) {
    std::vector<std::pair<uint32_t, std::pair<uint32_t, double>>> tr;  // master, (hanging, w)
    tr.reserve(A->line_col.size());
    for (uint32_t h : A->hang_nodes) {
      const int32_t l = line_of[h];
      for (uint32_t t = A->line_ptr[l]; t < A->line_ptr[l + 1]; ++t)
        tr.push_back({A->line_col[t], {h, A->line_val[t]}});
    }
    std::stable_sort(tr.begin(), tr.end(),
                     [](const auto &a, const auto &b) { return a.first < b.first; });
    A->cm_node.clear();
    A->cm_ptr.assign(1, 0u);
    A->cm_h.clear();
    A->cm_w.clear();
    for (size_t a = 0; a < tr.size();) {
      size_t b = a;
      A->cm_node.push_back(tr[a].first);
      while (b < tr.size() && tr[b].first == tr[a].first) {
        A->cm_h.push_back(tr[b].second.first);
        A->cm_w.push_back(tr[b].second.second);
        ++b;
      }
      A->cm_ptr.push_back(uint32_t(A->cm_h.size()));
      a = b;
    }
  }
  // cell lists: by colour, non-hanging cells first, each part in cell order
  std::vector<int64_t> off(size_t(2 * n_col) + 1, 0);
  for (int64_t c = 0; c < n_cells; ++c) ++off[size_t(2 * cmap[cls(c)] + (A->cell_hmask[c] != 0)) + 1];
  for (size_t k = 1; k < off.size(); ++k) off[k] += off[k - 1];
  A->colour_start.assign(n_col + 1, 0);
  A->colour_hang.assign(n_col, 0);
  for (int col = 0; col < n_col; ++col) {
    A->colour_start[col] = off[size_t(2 * col)];
    A->colour_hang[col] = off[size_t(2 * col + 1)];
  }
  A->colour_start[n_col] = n_cells;
  A->list.resize(size_t(n_cells));
  for (int64_t c = 0; c < n_cells; ++c)
    A->list[size_t(off[size_t(2 * cmap[cls(c)] + (A->cell_hmask[c] != 0))]++)] = uint32_t(c);
  A->n_colours = n_col;
  // hanging-cell scatter targets: each distinct node a hanging cell writes, with the local
  // entries (local node, weight) that sum into it, and the slots of the condensed diagonal,
  // one per (hanging cell, target), added with atomics (built above when on the device)
  if (!on_device) {
    A->cell_hidx.assign(size_t(n_cells), -1);
    A->ht_ptr.assign(1, 0u);
    A->tgt_ptr.assign(1, 0u);
    std::vector<std::pair<uint32_t, std::pair<int, double>>> ent;  // node, (local, weight)
    for (int64_t c = 0; c < n_cells; ++c) {
      if (!A->cell_hmask[c]) continue;
      A->cell_hidx[c] = int32_t(A->ht_ptr.size() - 1);
      ent.clear();
      for (int e = 0; e < nloc; ++e) {
        const uint32_t id = A->cell_nodes[size_t(c) * nloc + e];
        if (line_of[id] >= 0) {
          for (uint32_t t = A->line_ptr[line_of[id]]; t < A->line_ptr[line_of[id] + 1]; ++t)
            ent.push_back({A->line_col[t], {e, A->line_val[t]}});
        } else {
          ent.push_back({id, {e, 1.0}});
        }
      }
      std::stable_sort(ent.begin(), ent.end(),
                       [](const auto &a, const auto &b) { return a.first < b.first; });
      for (size_t a = 0; a < ent.size();) {
        size_t b = a;
        A->tgt_node.push_back(ent[a].first);
        while (b < ent.size() && ent[b].first == ent[a].first) {
          A->ent_local.push_back(uint8_t(ent[b].second.first));
          A->ent_w.push_back(ent[b].second.second);
          ++b;
        }
        A->tgt_ptr.push_back(uint32_t(A->ent_w.size()));
        a = b;
      }
      A->ht_ptr.push_back(uint32_t(A->tgt_node.size()));
    }
    // slots of the condensed diagonal: (hanging cell, target); added with atomics
    A->slot_start.assign(2, 0);
    for (int64_t c = 0; c < n_cells; ++c) {
      const int32_t hi = A->cell_hidx[c];
      if (hi < 0) continue;
      for (uint32_t t = A->ht_ptr[hi]; t < A->ht_ptr[hi + 1]; ++t) {
        A->slot_cell.push_back(uint32_t(c));
        A->slot_tgt.push_back(t);
      }
    }
    A->slot_start[1] = int64_t(A->slot_cell.size());
  }
  m.ad = A;
  return MF_OK;  // device tables are uploaded when the first operator is created
}

int ensure_adaptive_on_device(mf_scratch *s)
{
  if (!s->ad || s->ad->uploaded) return MF_OK;
  const int rc = upload_adaptive(s);
  if (rc == MF_OK) s->ad->uploaded = true;
  return rc;
}

template <typename T>
static int upload(T **d, const std::vector<T> &h)
{
  if (*d) return MF_OK;  // built on the device
  if (h.empty()) {
    *d = nullptr;
    return MF_OK;
  }
  MF_HIP_TRY(hipMalloc(reinterpret_cast<void **>(d), sizeof(T) * h.size()));
  MF_HIP_TRY(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return MF_OK;
}

int upload_adaptive(mf_scratch *s)
{
  AdaptiveData *A = s->ad;
  int rc;
  std::vector<float> lv32(A->line_val.begin(), A->line_val.end());
  std::vector<float> ew32(A->ent_w.begin(), A->ent_w.end());
  std::vector<float> cw32(A->cm_w.begin(), A->cm_w.end());
  if ((rc = upload(&A->d.cell_nodes, A->cell_nodes)) || (rc = upload(&A->d.cell_hmask, A->cell_hmask)) ||
      (rc = upload(&A->d.cell_level, A->cell_level)) || (rc = upload(&A->d.cell_hidx, A->cell_hidx)) ||
      (rc = upload(&A->d.node_line, A->node_line)) || (rc = upload(&A->d.line_ptr, A->line_ptr)) ||
      (rc = upload(&A->d.line_col, A->line_col)) || (rc = upload(&A->d.line_val, A->line_val)) ||
      (rc = upload(&A->d.line_val32, lv32)) || (rc = upload(&A->d.list, A->list)) ||
      (rc = upload(&A->d.ht_ptr, A->ht_ptr)) || (rc = upload(&A->d.tgt_node, A->tgt_node)) ||
      (rc = upload(&A->d.tgt_ptr, A->tgt_ptr)) || (rc = upload(&A->d.ent_local, A->ent_local)) ||
      (rc = upload(&A->d.ent_w, A->ent_w)) || (rc = upload(&A->d.ent_w32, ew32)) ||
      (rc = upload(&A->d.slot_cell, A->slot_cell)) || (rc = upload(&A->d.slot_tgt, A->slot_tgt)) ||
      (rc = upload(&A->d.hang_nodes, A->hang_nodes)) || (rc = upload(&A->d.cm_node, A->cm_node)) ||
      (rc = upload(&A->d.cm_ptr, A->cm_ptr)) || (rc = upload(&A->d.cm_h, A->cm_h)) ||
      (rc = upload(&A->d.cm_w, A->cm_w)) || (rc = upload(&A->d.cm_w32, cw32)))
    return rc;
  if (A->n_hanging > 0) {
    MF_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&A->d.hang_val), sizeof(double) * A->n_hanging));
    MF_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&A->d.hang_val32), sizeof(float) * A->n_hanging));
  }
  const size_t n = size_t(s->mesh.n_dofs);
  if (!s->d_mask && !s->h_mask.empty()) {
    MF_HIP_TRY(hipMalloc(&s->d_mask, n));
    MF_HIP_TRY(hipMemcpy(s->d_mask, s->h_mask.data(), n, hipMemcpyHostToDevice));
  }
  return MF_OK;
}

void free_adaptive(mf_scratch *s)
{
  AdaptiveData *A = s->ad;
  if (!A) return;
  void *ptrs[] = {A->d.cell_nodes, A->d.cell_hmask, A->d.cell_level, A->d.cell_hidx,
                  A->d.node_line,  A->d.line_ptr,   A->d.line_col,   A->d.line_val,
                  A->d.line_val32, A->d.list,       A->d.ht_ptr,     A->d.tgt_node,
                  A->d.tgt_ptr,    A->d.ent_local,  A->d.ent_w,      A->d.ent_w32,
                  A->d.slot_cell,  A->d.slot_tgt,   A->d.hang_nodes, A->d.cm_node,
                  A->d.cm_ptr,     A->d.cm_h,       A->d.cm_w,       A->d.cm_w32,
                  A->d.hang_val,   A->d.hang_val32, A->d.xyz};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  delete A;
  s->ad = nullptr;
}

}  // namespace mf

using namespace mf;

extern "C" {

int mf_mesh_refine_band(const int ncell[3], const double lo[3], const double hi[3], int levels,
                        const double center[3], double radius, double half_width, int32_t *cells,
                        int64_t capacity, int64_t *n_cells)
{
  MF_REQUIRE(ncell && lo && hi && center && n_cells, MF_ERR_ARG, "null argument");
  MF_REQUIRE(levels >= 0 && levels <= 8, MF_ERR_ARG, "levels must be in [0, 8]");
  BandSpec b;
  for (int d = 0; d < 3; ++d) {
    MF_REQUIRE(ncell[d] >= 1 && (int64_t(ncell[d]) << levels) < (1 << 20) && hi[d] > lo[d],
               MF_ERR_ARG, "invalid base mesh");
    b.ncell[d] = ncell[d];
    b.lo[d] = lo[d];
    b.hi[d] = hi[d];
    b.center[d] = center[d];
  }
  b.radius = radius;
  b.half_width = half_width;
  int n_devices = 0;
  if (g_adaptive_setup_device && hipGetDeviceCount(&n_devices) == hipSuccess && n_devices > 0) {
    std::vector<int32_t> dev_cells;
    const int rc = device_refine_band(ncell, lo, hi, levels, center, radius, half_width, dev_cells);
    if (rc != MF_OK) return rc;
    *n_cells = int64_t(dev_cells.size() / 4);
    if (!cells) return MF_OK;
    MF_REQUIRE(capacity >= *n_cells, MF_ERR_ARG, "cell buffer too small");
    std::memcpy(cells, dev_cells.data(), sizeof(int32_t) * dev_cells.size());
    return MF_OK;
  }
  LeafSet L;
  for (int64_t z = 0; z < ncell[2]; ++z)
    for (int64_t y = 0; y < ncell[1]; ++y)
      for (int64_t x = 0; x < ncell[0]; ++x) L.s.insert(leaf_key(0, x, y, z));
  for (int lev = 0; lev < levels; ++lev) {
    std::vector<uint64_t> todo;
    for (uint64_t k : L.s) {
      if (int(k >> 60) != lev) continue;
      const int64_t ix[3] = {int64_t((k >> 40) & 0xFFFFF), int64_t((k >> 20) & 0xFFFFF),
                             int64_t(k & 0xFFFFF)};
      if (meets_band(b, lev, ix)) todo.push_back(k);
    }
    for (uint64_t k : todo) refine_into(L, lev, (k >> 40) & 0xFFFFF, (k >> 20) & 0xFFFFF, k & 0xFFFFF);
  }
  balance(L, ncell);
  int Lmax = 0;
  for (uint64_t k : L.s) Lmax = std::max(Lmax, int(k >> 60));
  std::vector<std::pair<uint64_t, uint64_t>> order;
  order.reserve(L.s.size());
  for (uint64_t k : L.s) {
    const int64_t ix[3] = {int64_t((k >> 40) & 0xFFFFF), int64_t((k >> 20) & 0xFFFFF),
                           int64_t(k & 0xFFFFF)};
    order.push_back({morton(Lmax, int(k >> 60), ix), k});
  }
  std::sort(order.begin(), order.end());
  *n_cells = int64_t(order.size());
  if (!cells) return MF_OK;
  MF_REQUIRE(capacity >= int64_t(order.size()), MF_ERR_ARG, "cell buffer too small");
  for (size_t i = 0; i < order.size(); ++i) {
    const uint64_t k = order[i].second;
    cells[4 * i] = int32_t(k >> 60);
    cells[4 * i + 1] = int32_t((k >> 40) & 0xFFFFF);
    cells[4 * i + 2] = int32_t((k >> 20) & 0xFFFFF);
    cells[4 * i + 3] = int32_t(k & 0xFFFFF);
  }
  return MF_OK;
}

int mf_scratch_create_adaptive(int degree, const int ncell[3], const double lo[3],
                               const double hi[3], const int32_t *cells, int64_t n_cells,
                               mf_scratch **out)
{
  MF_REQUIRE(ncell && lo && hi && cells && out, MF_ERR_ARG, "null argument");
  MF_REQUIRE(degree >= 1 && degree <= 3, MF_ERR_ARG, "degree must be 1, 2 or 3");
  MF_REQUIRE(n_cells >= 1 && n_cells < (int64_t(1) << 31), MF_ERR_ARG, "invalid cell count");
  auto *s = new mf_scratch();
  MeshInfo &m = s->mesh;
  m.degree = degree;
  m.K = degree + 1;
  for (int d = 0; d < 3; ++d) {
    if (ncell[d] < 1 || !(hi[d] > lo[d])) {
      delete s;
      set_error("invalid mesh extent");
      return MF_ERR_ARG;
    }
    m.ncell[d] = ncell[d];
    m.lo[d] = lo[d];
    m.hi[d] = hi[d];
    m.h[d] = (hi[d] - lo[d]) / ncell[d];
    m.NN[d] = 0;
  }
  compute_tables(degree, s->tab);
  const int rc = build_adaptive(s, cells, n_cells);
  if (rc) {
    free_adaptive(s);
    if (s->d_mask) (void)hipFree(s->d_mask);
    delete s;
    return rc;
  }
  *out = s;
  return MF_OK;
}

int mf_scratch_is_adaptive(const mf_scratch *s) { return (s && s->ad) ? 1 : 0; }

int mf_scratch_support_points(const mf_scratch *s, double *xyz)
{
  MF_REQUIRE(s && xyz, MF_ERR_ARG, "null argument");
  const MeshInfo &m = s->mesh;
  if (s->ad) {
    std::memcpy(xyz, s->ad->xyz.data(), sizeof(double) * 3 * size_t(m.n_dofs));
    return MF_OK;
  }
  const int p = m.degree;
  for (int64_t idx = 0; idx < m.n_dofs; ++idx) {
    int64_t r = idx;
    for (int d = 0; d < 3; ++d) {
      const int64_t i = r % m.NN[d];
      r /= m.NN[d];
      const int64_t c = std::min<int64_t>(i / p, m.ncell[d] - 1);
      xyz[3 * idx + d] = m.lo[d] + m.h[d] * (double(c) + s->tab.gl[i - c * p]);
    }
  }
  return MF_OK;
}

int mf_scratch_default_dirichlet_mask(const mf_scratch *s, uint8_t *mask)
{
  MF_REQUIRE(s && mask, MF_ERR_ARG, "null argument");
  const MeshInfo &m = s->mesh;
  if (s->ad) {
    for (int64_t i = 0; i < m.n_dofs; ++i)
      mask[i] = (s->ad->boundary[i] && s->ad->node_line[i] < 0) ? 1 : 0;
    return MF_OK;
  }
  for (int64_t idx = 0; idx < m.n_dofs; ++idx) {
    int64_t r = idx;
    uint8_t b = 0;
    for (int d = 0; d < 3; ++d) {
      const int64_t i = r % m.NN[d];
      r /= m.NN[d];
      b |= (i == 0 || i == m.NN[d] - 1) ? 1 : 0;
    }
    mask[idx] = b;
  }
  return MF_OK;
}

int mf_scratch_constraint_info(const mf_scratch *s, int64_t *n_hanging, int64_t *n_entries,
                               int *n_colours)
{
  MF_REQUIRE(s, MF_ERR_ARG, "null scratch");
  const AdaptiveData *A = s->ad;
  if (n_hanging) *n_hanging = A ? A->n_hanging : 0;
  if (n_entries) *n_entries = A ? int64_t(A->line_col.size()) : 0;
  if (n_colours) *n_colours = A ? A->n_colours : 4;
  return MF_OK;
}

int mf_distribute_constraints(mf_scratch *s, double *x)
{
  MF_REQUIRE(s && x, MF_ERR_ARG, "null argument");
  if (!s->ad || s->ad->n_hanging == 0) return MF_OK;
  int rc = ensure_adaptive_on_device(s);
  if (rc) return rc;
  return launch_distribute(s->ad->d.node_line, s->ad->d.line_ptr, s->ad->d.line_col,
                           s->ad->d.line_val, x, s->mesh.n_dofs, s->stream);
}

int mf_adaptive_transfer(mf_scratch *old_s, const double *old_vec, mf_scratch *new_s,
                         double *new_vec)
{
  MF_REQUIRE(old_s && old_vec && new_s && new_vec, MF_ERR_ARG, "null argument");
  MF_REQUIRE(old_s->ad && new_s->ad, MF_ERR_ARG, "the transfer needs two adaptive scratches");
  const MeshInfo &a = old_s->mesh, &b = new_s->mesh;
  bool same = a.degree == b.degree;
  for (int d = 0; d < 3; ++d)
    same = same && a.ncell[d] == b.ncell[d] && a.lo[d] == b.lo[d] && a.h[d] == b.h[d];
  MF_REQUIRE(same, MF_ERR_ARG, "the meshes must share the base mesh and the degree");
  int rc = ensure_adaptive_on_device(old_s);
  if (rc) return rc;
  if ((rc = ensure_adaptive_on_device(new_s))) return rc;
  // the old field made constraint-consistent (hanging values from their masters)
  double *tmp = nullptr;
  const size_t bytes = sizeof(double) * size_t(a.n_dofs);
  MF_HIP_TRY(hipStreamSynchronize(new_s->stream));
  MF_HIP_TRY(hipMalloc(reinterpret_cast<void **>(&tmp), bytes));
  rc = hipMemcpyAsync(tmp, old_vec, bytes, hipMemcpyDeviceToDevice, old_s->stream) == hipSuccess
           ? MF_OK
           : MF_ERR_HIP;
  if (rc == MF_OK) rc = mf_distribute_constraints(old_s, tmp);
  if (rc == MF_OK && hipStreamSynchronize(old_s->stream) != hipSuccess) rc = MF_ERR_HIP;
  if (rc == MF_OK) rc = device_transfer_field(old_s, tmp, new_s, new_vec);
  (void)hipFree(tmp);
  if (rc != MF_OK) {
    if (rc == MF_ERR_HIP) set_error("device error in the field transfer");
    return rc;
  }
  return mf_distribute_constraints(new_s, new_vec);
}

int mf_scratch_constraint_lines(const mf_scratch *s, int32_t *node_line, uint32_t *line_ptr,
                                uint32_t *line_col, double *line_val)
{
  MF_REQUIRE(s && s->ad, MF_ERR_ARG, "not an adaptive scratch");
  const AdaptiveData *A = s->ad;
  if (node_line) std::memcpy(node_line, A->node_line.data(), sizeof(int32_t) * A->node_line.size());
  if (line_ptr) std::memcpy(line_ptr, A->line_ptr.data(), sizeof(uint32_t) * A->line_ptr.size());
  if (line_col) std::memcpy(line_col, A->line_col.data(), sizeof(uint32_t) * A->line_col.size());
  if (line_val) std::memcpy(line_val, A->line_val.data(), sizeof(double) * A->line_val.size());
  return MF_OK;
}

}  // extern "C"

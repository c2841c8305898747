// Formats around the device path: VTU output of device-resident fields and restart files.
//
// Reference functions replaced:
//   GenericDataOut + Postprocessor::process / DataOut::write_vtu (source/post_processing/
//     generic_data_out.cpp, postprocessor.cpp): here every FE cell of degree p is written as p^3
//     trilinear hexahedra on its Gauss-Lobatto nodes (deal.II's DataOut with n_subdivisions = p
//     samples equidistant points instead; the node values are written as they are, so nothing
//     is interpolated), one point per DoF, point data = the fields.
//   Restart::serialize_internal / deserialize_internal (source/utilities/restart.cpp:115-186):
//     the reference saves the p4est triangulation and the vectors through SolutionTransfer; here
//     one binary file holds the mesh description (box, degree, cells of an adaptive mesh, the
//     Dirichlet mask) and the vectors, and loading rebuilds the same DoF numbering from it.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "mf_internal.hpp"

using namespace mf;

namespace {

struct FileCloser {
  void operator()(std::FILE *f) const
  {
    if (f) std::fclose(f);
  }
};
using File = std::unique_ptr<std::FILE, FileCloser>;

int copy_to_host(std::vector<double> &h, const double *d, int64_t n, hipStream_t st)
{
  h.resize(size_t(n));
  MF_HIP_TRY(hipMemcpyAsync(h.data(), d, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, st));
  MF_HIP_TRY(hipStreamSynchronize(st));
  return MF_OK;
}

// global node of local node (i, j, k) of cell c
struct NodeMap {
  const mf_scratch *s;
  int64_t operator()(int64_t c, int i, int j, int k) const
  {
    const MeshInfo &m = s->mesh;
    const int K = m.K;
    if (s->ad) return int64_t(s->ad->cell_nodes[size_t(c) * K * K * K + i + K * (j + K * k)]);
    const int64_t cx = c % m.ncell[0], cy = (c / m.ncell[0]) % m.ncell[1],
                  cz = c / (int64_t(m.ncell[0]) * m.ncell[1]);
    const int p = m.degree;
    return (cx * p + i) + m.NN[0] * ((cy * p + j) + m.NN[1] * (cz * p + k));
  }
};

constexpr uint32_t RESTART_MAGIC = 0x3152464du;  // "MFR1"

}  // namespace

extern "C" {

int mf_write_vtu(const mf_scratch *s, const char *filename, int n_fields, const char *const *names,
                 const int *n_components, const double *const *fields)
{
  MF_REQUIRE(s && filename, MF_ERR_ARG, "null argument");
  MF_REQUIRE(n_fields >= 0 && (n_fields == 0 || (names && n_components && fields)), MF_ERR_ARG,
             "field arguments missing");
  const MeshInfo &m = s->mesh;
  const int64_t n = m.n_dofs;
  const int p = m.degree;
  for (int f = 0; f < n_fields; ++f) {
    MF_REQUIRE(names[f] && fields[f], MF_ERR_ARG, "null field");
    MF_REQUIRE(n_components[f] == 1 || n_components[f] == 3, MF_ERR_ARG,
               "fields have 1 or 3 components");
  }
  std::vector<double> xyz(static_cast<size_t>(3 * n));
  if (int rc = mf_scratch_support_points(s, xyz.data())) return rc;
  // p^3 hexahedra per cell, VTK corner order: (0,0,0) (1,0,0) (1,1,0) (0,1,0), then z + 1
  const int64_t n_sub = m.n_cells * int64_t(p) * p * p;
  std::vector<int64_t> conn(static_cast<size_t>(8 * n_sub));
  std::vector<int64_t> offsets(static_cast<size_t>(n_sub));
  std::vector<uint8_t> types(size_t(n_sub), uint8_t(12));  // VTK_HEXAHEDRON
  const NodeMap node{s};
  int64_t q = 0;
  for (int64_t c = 0; c < m.n_cells; ++c)
    for (int k = 0; k < p; ++k)
      for (int j = 0; j < p; ++j)
        for (int i = 0; i < p; ++i) {
          const int di[8] = {0, 1, 1, 0, 0, 1, 1, 0}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                    dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
          for (int v = 0; v < 8; ++v) conn[size_t(8 * q + v)] = node(c, i + di[v], j + dj[v], k + dk[v]);
          offsets[size_t(q)] = 8 * (q + 1);
          ++q;
        }
  std::vector<std::vector<double>> data(static_cast<size_t>(n_fields));
  for (int f = 0; f < n_fields; ++f) {
    std::vector<double> h;
    if (int rc = copy_to_host(h, fields[f], n * n_components[f], s->stream)) return rc;
    if (n_components[f] == 1) {
      data[size_t(f)] = std::move(h);
    } else {  // [3][n] -> interleaved xyz
      data[size_t(f)].resize(size_t(3 * n));
      for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) data[size_t(f)][size_t(3 * i + d)] = h[size_t(d * n + i)];
    }
  }
  // appended raw data: each block is a UInt64 byte count followed by the bytes
  struct Block {
    const void *ptr;
    uint64_t bytes;
  };
  std::vector<Block> blocks;
  std::string xml;
  uint64_t offset = 0;
  auto array = [&](const char *type, const char *name, int ncomp, const void *ptr, uint64_t bytes) {
    xml += std::string("        <DataArray type=\"") + type + "\" Name=\"" + name +
           "\" NumberOfComponents=\"" + std::to_string(ncomp) +
           "\" format=\"appended\" offset=\"" + std::to_string(offset) + "\"/>\n";
    blocks.push_back({ptr, bytes});
    offset += 8 + bytes;
  };
  xml += "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" "
         "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n  <UnstructuredGrid>\n";
  xml += "    <Piece NumberOfPoints=\"" + std::to_string(n) + "\" NumberOfCells=\"" +
         std::to_string(n_sub) + "\">\n";
  xml += "      <PointData>\n";
  for (int f = 0; f < n_fields; ++f)
    array("Float64", names[f], n_components[f], data[size_t(f)].data(),
          uint64_t(sizeof(double) * data[size_t(f)].size()));
  xml += "      </PointData>\n      <Points>\n";
  array("Float64", "Points", 3, xyz.data(), uint64_t(sizeof(double) * xyz.size()));
  xml += "      </Points>\n      <Cells>\n";
  array("Int64", "connectivity", 1, conn.data(), uint64_t(sizeof(int64_t) * conn.size()));
  array("Int64", "offsets", 1, offsets.data(), uint64_t(sizeof(int64_t) * offsets.size()));
  array("UInt8", "types", 1, types.data(), uint64_t(types.size()));
  xml += "      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n  <AppendedData encoding=\"raw\">\n   _";
  File fp(std::fopen(filename, "wb"));
  MF_REQUIRE(fp, MF_ERR_ARG, std::string("cannot open ") + filename);
  bool ok = std::fwrite(xml.data(), 1, xml.size(), fp.get()) == xml.size();
  for (const Block &b : blocks) {
    ok = ok && std::fwrite(&b.bytes, sizeof(uint64_t), 1, fp.get()) == 1;
    ok = ok && std::fwrite(b.ptr, 1, size_t(b.bytes), fp.get()) == size_t(b.bytes);
  }
  const char *tail = "\n  </AppendedData>\n</VTKFile>\n";
  ok = ok && std::fwrite(tail, 1, std::strlen(tail), fp.get()) == std::strlen(tail);
  MF_REQUIRE(ok, MF_ERR_ARG, std::string("write failed: ") + filename);
  return MF_OK;
}

// File layout (little endian): magic, version, degree, adaptive flag, ncell[3] (int32), lo[3],
// hi[3] (double), n_cells (int64), adaptive: cells [n_cells][4] int32; n_dofs (int64), mask flag
// (int32) and mask [n_dofs] uint8; n_vectors (int32), each: length (int64) + doubles.
int mf_restart_save(const mf_scratch *s, const char *path, int n_vectors,
                    const double *const *vectors, const int64_t *lengths)
{
  MF_REQUIRE(s && path, MF_ERR_ARG, "null argument");
  MF_REQUIRE(n_vectors >= 0 && (n_vectors == 0 || (vectors && lengths)), MF_ERR_ARG,
             "vector arguments missing");
  const MeshInfo &m = s->mesh;
  File fp(std::fopen(path, "wb"));
  MF_REQUIRE(fp, MF_ERR_ARG, std::string("cannot open ") + path);
  bool ok = true;
  auto put = [&](const void *p, size_t bytes) {
    ok = ok && std::fwrite(p, 1, bytes, fp.get()) == bytes;
  };
  const uint32_t head[2] = {RESTART_MAGIC, 1u};
  put(head, sizeof(head));
  const int32_t deg = m.degree, adaptive = s->ad ? 1 : 0;
  put(&deg, 4);
  put(&adaptive, 4);
  const int32_t nc[3] = {m.ncell[0], m.ncell[1], m.ncell[2]};
  put(nc, sizeof(nc));
  put(m.lo, sizeof(m.lo));
  put(m.hi, sizeof(m.hi));
  const int64_t ncells = m.n_cells;
  put(&ncells, 8);
  if (s->ad) put(s->ad->cells.data(), sizeof(int32_t) * s->ad->cells.size());
  const int64_t ndofs = m.n_dofs;
  put(&ndofs, 8);
  const int32_t has_mask = s->h_mask.empty() ? 0 : 1;
  put(&has_mask, 4);
  if (has_mask) put(s->h_mask.data(), s->h_mask.size());
  const int32_t nv = n_vectors;
  put(&nv, 4);
  for (int v = 0; v < n_vectors; ++v) {
    MF_REQUIRE(vectors[v] && lengths[v] >= 0, MF_ERR_ARG, "null vector");
    std::vector<double> h;
    if (int rc = copy_to_host(h, vectors[v], lengths[v], s->stream)) return rc;
    put(&lengths[v], 8);
    put(h.data(), sizeof(double) * h.size());
  }
  MF_REQUIRE(ok, MF_ERR_ARG, std::string("write failed: ") + path);
  return MF_OK;
}

}  // extern "C"

namespace {
// bytes left in the file after the current position (-1 when unknown)
int64_t bytes_left(std::FILE *f)
{
  const long pos = std::ftell(f);
  if (pos < 0 || std::fseek(f, 0, SEEK_END) != 0) return -1;
  const long end = std::ftell(f);
  if (end < 0 || std::fseek(f, pos, SEEK_SET) != 0) return -1;
  return int64_t(end) - int64_t(pos);
}
}  // namespace

extern "C" {

int mf_restart_load_scratch(const char *path, mf_scratch **out)
{
  MF_REQUIRE(path && out, MF_ERR_ARG, "null argument");
  File fp(std::fopen(path, "rb"));
  MF_REQUIRE(fp, MF_ERR_ARG, std::string("cannot open ") + path);
  bool ok = true;
  auto get = [&](void *p, size_t bytes) {
    ok = ok && std::fread(p, 1, bytes, fp.get()) == bytes;
  };
  uint32_t head[2] = {0, 0};
  get(head, sizeof(head));
  MF_REQUIRE(ok && head[0] == RESTART_MAGIC && head[1] == 1u, MF_ERR_ARG,
             "not a restart file of this library");
  int32_t deg = 0, adaptive = 0, nc[3] = {0, 0, 0};
  double lo[3], hi[3];
  int64_t ncells = 0;
  get(&deg, 4);
  get(&adaptive, 4);
  get(nc, sizeof(nc));
  get(lo, sizeof(lo));
  get(hi, sizeof(hi));
  get(&ncells, 8);
  MF_REQUIRE(ok && ncells >= 1 && ncells < (int64_t(1) << 31), MF_ERR_ARG, "corrupt restart file");
  // the declared counts must fit in what the file holds before anything is sized from them
  MF_REQUIRE(!adaptive || bytes_left(fp.get()) >= int64_t(sizeof(int32_t)) * 4 * ncells,
             MF_ERR_ARG, "corrupt restart file (shorter than its cell list)");
  mf_scratch *s = nullptr;
  int rc;
  if (adaptive) {
    std::vector<int32_t> cells(static_cast<size_t>(4 * ncells));
    get(cells.data(), sizeof(int32_t) * cells.size());
    MF_REQUIRE(ok, MF_ERR_ARG, "corrupt restart file");
    const int ncell[3] = {nc[0], nc[1], nc[2]};
    rc = mf_scratch_create_adaptive(deg, ncell, lo, hi, cells.data(), ncells, &s);
  } else {
    mf_mesh_desc d{};
    d.dim = 3;
    d.degree = deg;
    for (int i = 0; i < 3; ++i) {
      d.ncell[i] = nc[i];
      d.lo[i] = lo[i];
      d.hi[i] = hi[i];
    }
    rc = mf_scratch_create(&d, &s);
  }
  if (rc) return rc;
  int64_t ndofs = 0;
  int32_t has_mask = 0;
  get(&ndofs, 8);
  get(&has_mask, 4);
  if (!ok || ndofs != s->mesh.n_dofs) {
    mf_scratch_destroy(s);
    set_error("restart file does not match the rebuilt mesh");
    return MF_ERR_ARG;
  }
  if (has_mask) {
    if (bytes_left(fp.get()) < ndofs) {
      mf_scratch_destroy(s);
      set_error("corrupt restart file (shorter than its Dirichlet mask)");
      return MF_ERR_ARG;
    }
    std::vector<uint8_t> mask(static_cast<size_t>(ndofs));
    get(mask.data(), mask.size());
    if (!ok || (rc = mf_scratch_set_dirichlet_mask(s, mask.data()))) {
      mf_scratch_destroy(s);
      if (!ok) set_error("corrupt restart file");
      return rc ? rc : MF_ERR_ARG;
    }
  }
  *out = s;
  return MF_OK;
}

int mf_restart_load_vectors(const mf_scratch *s, const char *path, int n_vectors,
                            double *const *vectors, const int64_t *lengths)
{
  MF_REQUIRE(s && path, MF_ERR_ARG, "null argument");
  MF_REQUIRE(n_vectors >= 0 && (n_vectors == 0 || (vectors && lengths)), MF_ERR_ARG,
             "vector arguments missing");
  File fp(std::fopen(path, "rb"));
  MF_REQUIRE(fp, MF_ERR_ARG, std::string("cannot open ") + path);
  bool ok = true;
  auto get = [&](void *p, size_t bytes) {
    ok = ok && std::fread(p, 1, bytes, fp.get()) == bytes;
  };
  auto skip = [&](int64_t bytes) { ok = ok && std::fseek(fp.get(), long(bytes), SEEK_CUR) == 0; };
  uint32_t head[2] = {0, 0};
  get(head, sizeof(head));
  MF_REQUIRE(ok && head[0] == RESTART_MAGIC && head[1] == 1u, MF_ERR_ARG,
             "not a restart file of this library");
  int32_t deg = 0, adaptive = 0, nc[3];
  double lo[3], hi[3];
  int64_t ncells = 0, ndofs = 0;
  get(&deg, 4);
  get(&adaptive, 4);
  get(nc, sizeof(nc));
  get(lo, sizeof(lo));
  get(hi, sizeof(hi));
  get(&ncells, 8);
  const MeshInfo &m = s->mesh;
  MF_REQUIRE(ok && deg == m.degree && (adaptive != 0) == (s->ad != nullptr) &&
                 ncells == m.n_cells && nc[0] == m.ncell[0] && nc[1] == m.ncell[1] &&
                 nc[2] == m.ncell[2],
             MF_ERR_ARG, "restart file belongs to another mesh");
  for (int d = 0; d < 3; ++d)
    MF_REQUIRE(lo[d] == m.lo[d] && hi[d] == m.hi[d], MF_ERR_ARG,
               "restart file belongs to another mesh (box differs)");
  if (adaptive) {
    // the same counts on another band mesh would load in another numbering: compare the cells
    MF_REQUIRE(bytes_left(fp.get()) >= int64_t(sizeof(int32_t)) * 4 * ncells, MF_ERR_ARG,
               "corrupt restart file (shorter than its cell list)");
    std::vector<int32_t> cells(static_cast<size_t>(4 * ncells));
    get(cells.data(), sizeof(int32_t) * cells.size());
    MF_REQUIRE(ok && cells.size() == s->ad->cells.size() &&
                   std::equal(cells.begin(), cells.end(), s->ad->cells.begin()),
               MF_ERR_ARG, "restart file belongs to another mesh (cells differ)");
  }
  int32_t has_mask = 0, nv = 0;
  get(&ndofs, 8);
  get(&has_mask, 4);
  MF_REQUIRE(ok && ndofs == m.n_dofs, MF_ERR_ARG, "restart file belongs to another mesh");
  if (has_mask) skip(ndofs);
  get(&nv, 4);
  MF_REQUIRE(ok && nv == n_vectors, MF_ERR_ARG,
             "restart file holds " + std::to_string(nv) + " vectors, " +
                 std::to_string(n_vectors) + " requested");
  for (int v = 0; v < n_vectors; ++v) {
    int64_t len = -1;
    get(&len, 8);
    MF_REQUIRE(ok && len == lengths[v] && vectors[v], MF_ERR_ARG,
               "vector " + std::to_string(v) + " has length " + std::to_string(len));
    MF_REQUIRE(len >= 0 && bytes_left(fp.get()) >= int64_t(sizeof(double)) * len, MF_ERR_ARG,
               "corrupt restart file (shorter than its vectors)");
    std::vector<double> h(static_cast<size_t>(len));
    get(h.data(), sizeof(double) * h.size());
    MF_REQUIRE(ok, MF_ERR_ARG, "corrupt restart file");
    MF_HIP_TRY(hipMemcpyAsync(vectors[v], h.data(), sizeof(double) * h.size(),
                              hipMemcpyHostToDevice, s->stream));
    MF_HIP_TRY(hipStreamSynchronize(s->stream));
  }
  return MF_OK;
}

}  // extern "C"

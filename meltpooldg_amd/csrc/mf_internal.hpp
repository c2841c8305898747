// Internal declarations shared by the HIP translation units of libmeltpooldg_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/meltpooldg_amd.h"

struct mf_scratch;
struct mf_operator;

namespace mf {

void set_error(const std::string &msg);

#define MF_HIP_TRY(expr)                                                                      \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      mf::set_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));                \
      return MF_ERR_HIP;                                                                      \
    }                                                                                         \
  } while (0)

#define MF_REQUIRE(cond, code, msg)                                                           \
  do {                                                                                        \
    if (!(cond)) {                                                                            \
      mf::set_error(msg);                                                                     \
      return code;                                                                            \
    }                                                                                         \
  } while (0)

constexpr int KMAX = 4;  // degree <= 3

// Operation selected at compile time in the cell kernel.
enum CellOp : int {
  HEAT_TANGENT = 0,  // heat_diffuse_operator.cpp:955-1086
  HEAT_RHS = 1,      // heat_diffuse_operator.cpp:393-529
  HEAT_DIAG = 2,     // heat_diffuse_operator.cpp:297-391 (diagonal of the tangent)
  ADVEC_TANGENT = 3, // advection_diffusion_operator.cpp:448-495
  ADVEC_RHS = 4,     // advection_diffusion_operator.cpp:274-358
  ADVEC_DIAG = 5,    // advection_diffusion_operator.cpp:399-446
  HEAT_PGRAD = 6,    // per-cell sum_q JxW grad(src) . grad(T): conductivity gradient (adjoint)
  LS_TANGENT = 7,    // normal_vector_operator.cpp:316-347, curvature_operator.cpp:339-371
  LS_DIAG = 8,       // their diagonal (normal_vector_operator.cpp:256-311)
  REINIT_TANGENT = 9,  // olsson_CG_operator.cpp:438-478
  REINIT_DIAG = 10     // olsson_CG_operator.cpp:357-436
};

constexpr bool is_tangent_op(int op)
{
  return op == HEAT_TANGENT || op == ADVEC_TANGENT || op == LS_TANGENT || op == REINIT_TANGENT;
}

// Everything a cell kernel needs besides src/dst; passed by value (kernel argument segment).
template <typename Number>
struct CellArgs {
  Number S[KMAX * KMAX];   // S[q][j] = l_j(x_q)
  Number DS[KMAX * KMAX];  // DS[q][j] = l_j'(x_q)
  Number Dq[KMAX * KMAX];  // collocation derivative on the Gauss points
  Number w[KMAX];          // Gauss weights on [0,1]
  int ncell[3];
  uint32_t NX, NY;         // nodes per x / y line
  uint32_t n_dofs;
  int n_cells;
  Number h[3], inv_h[3];
  const uint8_t *mask;     // Dirichlet mask (nullable)
  int use_mask;
  // heat
  Number rho_cp, k0, tanh_amp, tanh_Tc, inv_tanh_w;
  int t_dependent;
  // mushy-zone law: liquid / solid conductivity, density and heat capacity, the liquidus and
  // 1 / (liquidus - solidus)
  Number mu_k[2], mu_rho[2], mu_cp[2], mu_Tliq, mu_inv;
  const Number *T_old;     // old temperature (mushy law, t_dependent)
  // advection; SUPG: the coefficient and the cell size diameter / sqrt(3) of a level-0 cell
  Number theta, diffusivity, mass_factor;
  Number dt_inv;
  Number supg_coef, supg_cell_size;
  Number DDS[KMAX * KMAX];  // DDS[q][j] = l_j''(x_q) (SUPG diagonal)
  // level-set projections: eta of a level-0 cell (scaled by h^2 on refined cells) and the
  // narrow-band threshold on |T| (T holds the level set)
  Number ls_eta, ls_threshold;
  // Olsson reinitialization (normal vector in vel): eps_n, eps_t of a level-0 cell (scaled by
  // the cell size on refined cells when re_scale_h), the normalisation tolerance and
  // sign(eps_n - eps_t) for the diagonal
  Number re_eps_n, re_eps_t, re_tol, re_sign;
  int re_scale_h;
  // fields
  const Number *T;         // linearisation temperature (heat tangent / diagonal)
  const Number *vel;       // nodal velocity [3][n_dofs]
  const Number *k_cell;    // per-cell conductivity factor
  const Number *aux1;      // heat rhs: q_s (nullable)
  int transpose;           // tangent ops: apply the transposed operator (adjoint)
  int gather_plain;        // tangent ops: read src without constraints (rows still skipped)
  double pgrad_scale;      // HEAT_PGRAD: dst[cell] += pgrad_scale * sum_q JxW grad src . grad T
  double *dot_partial;     // fused src . (A src) per block (nullable)
  const int *stop;         // early-exit flag set by the solver (nullable)
  // adaptive meshes (cell lists, hanging-node constraints); null on structured meshes
  const uint32_t *cell_nodes;  // [n_cells][K^3] global node of each local node
  const uint64_t *cell_hmask;  // [n_cells] bit e: local node e is a hanging node
  const uint8_t *cell_level;   // [n_cells] refinement level: h = h_base / 2^level
  const int32_t *cell_hidx;    // [n_cells] index into the hanging-cell target table (-1: none)
  const int32_t *node_line;    // [n_dofs] constraint line of a hanging node (-1: free)
  const uint32_t *line_ptr;    // [n_lines + 1]
  const uint32_t *line_col;    // masters (free nodes)
  const Number *line_val;      // weights
  const uint32_t *list;        // cells in colour order
  const uint32_t *ht_ptr;      // [n_hanging_cells + 1] targets of each hanging cell
  const uint32_t *tgt_node;    // node of each target
  const uint32_t *tgt_ptr;     // [n_targets + 1] entries of each target
  const uint8_t *ent_local;    // local node of each entry
  const Number *ent_w;         // weight of each entry
  const uint32_t *slot_cell;   // condensed diagonal: (cell, target) per slot
  const uint32_t *slot_tgt;
  const Number *hang_val;      // [n_hanging] interpolated source values of the hanging nodes
  int overwrite;               // structured meshes: dst = A src (first writer of a node stores)
  int mask_boundary;           // structured meshes: mask is exactly the boundary of the box
};

// Row decomposition of a structured mesh for colour `colour` (0..3): bit 0 = parity of the cell
// row y, bit 1 = parity of the cell layer z.  Rows of one colour share no node.
struct RowColour {
  int yb, zb;  // parities
  int hy, hz;  // number of rows / layers of this colour
};

__host__ __device__ inline RowColour row_colour(const int ncell[3], int colour)
{
  RowColour s;
  s.yb = colour & 1;
  s.zb = (colour >> 1) & 1;
  s.hy = (ncell[1] - s.yb + 1) / 2;
  s.hz = (ncell[2] - s.zb + 1) / 2;
  return s;
}

struct AdaptiveData;

struct MeshInfo {
  const AdaptiveData *ad = nullptr;  // adaptive mesh tables (null: structured mesh)
  int degree = 0, K = 0;
  int ncell[3] = {0, 0, 0};
  int64_t NN[3] = {0, 0, 0};
  double h[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int64_t n_dofs = 0, n_cells = 0;
};

struct Tables {
  double S[KMAX * KMAX], DS[KMAX * KMAX], Dq[KMAX * KMAX], w[KMAX], gl[KMAX], xq[KMAX];
};

void compute_tables(int degree, Tables &t);
// Lagrange polynomials on `nodes` (n of them): values v and derivatives d at x
void lagrange_basis(int n, const double *nodes, double x, double *v, double *d);

// Adaptive mesh: host tables and their device copies (mf_adaptive.cpp)
struct AdaptiveData {
  int levels = 0, n_colours = 0;
  int64_t n_hanging = 0;
  bool uploaded = false;
  std::vector<int32_t> cells;          // [n_cells][4] level, ix, iy, iz
  std::vector<double> xyz;             // [n_dofs][3]
  std::vector<uint8_t> boundary;       // [n_dofs]
  std::vector<uint32_t> cell_nodes;    // [n_cells][K^3]
  std::vector<uint64_t> cell_hmask;
  std::vector<uint8_t> cell_level;
  std::vector<int32_t> cell_hidx;
  std::vector<int32_t> node_line;
  std::vector<uint32_t> line_ptr, line_col;
  std::vector<double> line_val;
  std::vector<uint32_t> list;          // cells by colour, non-hanging cells first
  std::vector<int64_t> colour_start;   // [n_colours + 1] offsets into list
  std::vector<int64_t> colour_hang;    // [n_colours] first hanging cell of each colour
  std::vector<uint32_t> ht_ptr, tgt_node, tgt_ptr;
  std::vector<uint8_t> ent_local;
  std::vector<double> ent_w;
  std::vector<uint32_t> slot_cell, slot_tgt;
  std::vector<int64_t> slot_start;     // [2]: all slots in one launch
  std::vector<uint32_t> hang_nodes;    // hanging nodes
  std::vector<uint32_t> cm_node, cm_ptr, cm_h;  // per master: the hanging nodes it receives from
  std::vector<double> cm_w;
  struct {
    uint32_t *cell_nodes = nullptr;
    uint64_t *cell_hmask = nullptr;
    uint8_t *cell_level = nullptr;
    int32_t *cell_hidx = nullptr, *node_line = nullptr;
    uint32_t *line_ptr = nullptr, *line_col = nullptr;
    double *line_val = nullptr;
    float *line_val32 = nullptr;
    uint32_t *list = nullptr, *ht_ptr = nullptr, *tgt_node = nullptr, *tgt_ptr = nullptr;
    uint8_t *ent_local = nullptr;
    double *ent_w = nullptr;
    float *ent_w32 = nullptr;
    uint32_t *slot_cell = nullptr, *slot_tgt = nullptr;
    uint32_t *hang_nodes = nullptr, *cm_node = nullptr, *cm_ptr = nullptr, *cm_h = nullptr;
    double *cm_w = nullptr;
    float *cm_w32 = nullptr;
    double *xyz = nullptr;        // support points, uploaded by the first call that needs them
    double *hang_val = nullptr;   // work: hanging-node values of the current source (FP64)
    float *hang_val32 = nullptr;  // and FP32
  } d;
};

int build_adaptive(::mf_scratch *s, const int32_t *cells, int64_t n_cells);
// node identification of an adaptive mesh on the device (mf_adaptive_dev.hip): cell_nodes,
// xyz and boundary, numbered as the host path numbers them; nkeys is left empty (the node keys
// stay on the device)
int device_node_identification(const ::mf_scratch *s, const int32_t *cells, int64_t n_cells,
                               int L, AdaptiveData *A, std::vector<uint64_t> &nkeys, int64_t &nn);
// transposed hanging-node constraints (cm_node, cm_ptr, cm_h, cm_w) on the device
// (mf_adaptive_dev.hip); hang_of_line[l] = the hanging node of constraint line l
int device_transpose_constraints(AdaptiveData *A, const std::vector<uint32_t> &hang_of_line,
                                 hipStream_t st);
// hanging-node constraint lines before chain resolution, on the device (mf_adaptive_dev.hip):
// line_of[node] = its line or -1, lines numbered in the host path's order, CSR (ptr, col, val)
int device_hanging_constraints(const ::mf_scratch *s, const int32_t *cells, int64_t n_cells, int L,
                               const std::vector<uint32_t> &cell_nodes, int64_t nn,
                               std::vector<int32_t> &line_of, std::vector<uint32_t> &line_ptr,
                               std::vector<uint32_t> &line_col, std::vector<double> &line_val);
// hanging-cell scatter targets and condensed-diagonal slots on the device, written into A->d
// (cell_hidx, ht_ptr, tgt_node, tgt_ptr, ent_local, ent_w, ent_w32, slot_cell, slot_tgt) with
// A->slot_start; uploads and keeps cell_nodes, node_line and the constraint lines there too
int device_hanging_targets(AdaptiveData *A, int64_t n_cells, int nloc, hipStream_t st);
// band-refined, 2:1-balanced leaves in Morton order on the device (mf_adaptive_dev.hip), the
// host path's cells: [n][4] level, ix, iy, iz
int device_refine_band(const int ncell[3], const double lo[3], const double hi[3], int levels,
                       const double center[3], double radius, double half_width,
                       std::vector<int32_t> &cells_out);
// new_vec[i] = the old field (constraint-consistent values old_val on old_s, whose tables are
// on the device) at the support point of node i of new_s (mf_adaptive_dev.hip)
int device_transfer_field(const ::mf_scratch *old_s, const double *old_val,
                          const ::mf_scratch *new_s, double *new_vec);
// 1: adaptive setup steps run on the device when one is present (default), 0: on the host
extern int g_adaptive_setup_device;
int upload_adaptive(::mf_scratch *s);
int ensure_adaptive_on_device(::mf_scratch *s);
// y[m] += sum_h w_hm y[h] over the hanging nodes h with master m, then y[h] = 0 (the transpose
// of the constraints applied to the unconstrained cell sums); with use_mask, Dirichlet masters
// keep their (zero) rows.  mf_ops.hip
template <typename Number>
int launch_condense(const ::mf_scratch *s, Number *y, const int *stop, bool use_mask);
// hv[l] = sum_t w_t src[m_t] over the line of every hanging node (constrained masters read as 0
// when use_mask), the source values hanging nodes take in read_dof_values, mf_ops.hip
template <typename Number>
int launch_hanging_values(const ::mf_scratch *s, const Number *src, int use_mask, Number *hv,
                          const int *stop);
// x[h] = sum_m w x[m] on every hanging node h (AffineConstraints::distribute), mf_ops.hip
int launch_distribute(const int32_t *node_line, const uint32_t *line_ptr, const uint32_t *line_col,
                      const double *line_val, double *x, int64_t n, hipStream_t st);
void free_adaptive(::mf_scratch *s);

// launches (mf_cell.hip)
template <typename Number>
int launch_cell(int op, int needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                const Number *src, Number *dst, hipStream_t stream);
// structured FE_Q3 heat tangent with the cell tensors as register planes (mf_plane.hip): whether it
// serves an application, and its 4 colour launches (one workgroup per cell row)
template <typename Number>
bool plane_applicable(const MeshInfo &m, int needT, bool vel, const CellArgs<Number> &a);
template <typename Number>
int launch_heat_plane(const MeshInfo &m, int needT, const CellArgs<Number> &a, const Number *src,
                      Number *dst, hipStream_t stream);
// adaptive meshes: mode 0 = all cells (tangent ops), 1 = cells without hanging nodes (diagonal
// ops), 2 = condensed diagonal of the cells with hanging nodes (op = a tangent op)
template <typename Number>
int launch_cell_list(int op, int needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                     const Number *src, Number *dst, hipStream_t stream, int mode);

// Stefan-Boltzmann constant in W/(m^2 K^4) (include/meltpooldg/utilities/physical_constants.hpp:6)
#define MF_STEFAN_BOLTZMANN 5.670374419e-8

// boundary faces of the heat operator (mf_face.hip): mode 0 dst += B src, 1 dst += diag(B),
// 2 dst += boundary part of -R(T); nothing when no side is active.  Mode 0 writes
// heat_face_blocks(op) partials of src . (B src) when dot_partial is given.
template <typename Number>
int launch_heat_faces(const mf_operator *op, const Number *src, Number *dst, int mode,
                      const int *stop, double *dot_partial = nullptr);
int heat_face_blocks(const mf_operator *op);

// number of blocks the cell kernel uses (size of dot_partial) in FP64 or FP32
int cell_blocks(const MeshInfo &m, bool fp32);

// level-set right-hand sides (mf_levelset.hip): mode 0 normal vector (src = level set, dst
// [3][n_dofs]), 1 curvature (src = normal vector [3][n_dofs], dst [n_dofs]), 2 Olsson
// reinitialization (src = level set, normal vector = op->vel, dst [n_dofs]); dst += ...
int launch_levelset_rhs(const mf_operator *op, int mode, const double *src, double *dst);

// advection inflow/outflow rows after a tangent application (mf_ops.hip): dst[i] = src[i] on the
// rows of io_mask (do_post_vmult); with dot_partial, IO_BLOCKS partials of src . src over them
constexpr int IO_BLOCKS = 64;
template <typename Number>
int launch_io_rows(const uint8_t *io_mask, const Number *src, Number *dst, int64_t n,
                   double *dot_partial, const int *stop, hipStream_t st);
// mask[i] = 1 on boundary nodes of the given sides (bit 2*direction + upper) where the outer
// normal times the nodal velocity is <= 0 (mf_ops.hip, structured meshes)
int launch_inflow_mask(const ::mf_scratch *s, const double *vel, unsigned sides, uint8_t *mask);

// blocks of one tangent application (size of the fused-dot partials)
inline int tangent_blocks(const MeshInfo &m, bool, bool fp32) { return cell_blocks(m, fp32); }

}  // namespace mf

struct mf_scratch {
  mf::MeshInfo mesh;
  mf::Tables tab;
  hipStream_t stream = nullptr;
  uint8_t *d_mask = nullptr;
  std::vector<uint8_t> h_mask;
  // structured meshes: the Dirichlet mask is exactly the boundary of the box (the kernels may
  // then compute it from the node coordinates instead of reading it)
  bool mask_full_boundary = false;
  mf::AdaptiveData *ad = nullptr;  // adaptive mesh (null: structured)
};

namespace mf {
// device memory an operator keeps between solves (grown on demand, freed with the operator)
struct Arena {
  char *p = nullptr;
  size_t bytes = 0;
};
}  // namespace mf

namespace mf {
struct CutData;  // one-phase CutFEM heat operator (mf_cut.hip)
}

struct mf_operator {
  mf::Arena arena[4];  // 0: refinement vectors, 1: Krylov vectors, 2: inverse diagonal, 3: small work
  // 0 heat, 1 advection, 2 normal vector (3 blocks), 3 curvature, 4 reinit, 5 CutFEM heat
  int kind = 0;
  mf_heat_cut_params cut{};            // kind 5
  mf::CutData *cutd = nullptr;         // kind 5: classification and its device tables
  mf_reinit_params reinit{};           // kind 4
  mf_levelset_projection_params ls{};  // kinds 2, 3
  double ls_tol = 0.0;                 // curvature: nodal normalisation tolerance
  mf_scratch *s = nullptr;
  mf_heat_params heat{};
  mf_advec_params advec{};
  double dt = 0.0;
  int dof_index = 0;
  const double *T = nullptr, *vel = nullptr, *k_cell = nullptr, *T_old = nullptr;
  float *T32 = nullptr, *vel32 = nullptr, *kc32 = nullptr, *To32 = nullptr;  // FP32 copies
  bool fields32_valid = false;
  int transpose = 0;  // vmult applies A^T (adjoint solves)
  mf_heat_boundary bnd{};  // Robin / Neumann sides (mf_operator_set_boundary)
  bool bnd_set = false;
  int gather_plain = 0;  // vmult reads src without constraints (source mass products)
  // advection: inflow/outflow DoFs (pre/post-vmult identity rows) and, for the cell kernels,
  // their union with the Dirichlet rows; io_pre_post switches the treatment in vmult
  uint8_t *io_mask = nullptr, *io_union = nullptr;
  int io_pre_post = 1;
};

namespace mf {
// eps_n and eps_t of a level-0 cell of the Olsson operator (compute_interface_thickness_parameter_
// epsilon of the cell size diameter / sqrt(3), reinitialization_data.hpp:181-197); scale_h: eps
// scales with the cell size on refined cells
inline void reinit_epsilon(const mf_operator *op, double &eps_n, double &eps_t, int &scale_h)
{
  const MeshInfo &m = op->s->mesh;
  const double cs = std::sqrt((m.h[0] * m.h[0] + m.h[1] * m.h[1] + m.h[2] * m.h[2]) / 3.0);
  const mf_reinit_params &r = op->reinit;
  if (r.thickness_type == MF_THICKNESS_ABSOLUTE) {
    eps_n = r.thickness_value;
    scale_h = 0;
  } else {
    eps_n = r.thickness_type == MF_THICKNESS_CELLS ? r.thickness_value * cs / 6.0
                                                   : r.thickness_value * cs;
    scale_h = 1;
  }
  eps_t = r.tangential_diffusion_factor * eps_n;
}

// One-phase CutFEM heat operator (mf_cut.hip).  cut_apply: dst = J src (overwrite) or
// dst += J src, with cut_dot_blocks(op) partials of src . (J src) when dot_partial is given;
// cut_diagonal: the diagonal of J, 1 on the rows without a free DoF; cut_residual: dst += -R(T)
// with T = op->T and T_old; cut_ready: the level set is set; cut_free: the device tables.
int cut_dot_blocks(const mf_operator *op);
int cut_apply(mf_operator *op, double *dst, const double *src, double *dot_partial,
              const int *stop, bool overwrite);
int cut_diagonal(mf_operator *op, double *diag);
int cut_residual(mf_operator *op, double *dst, const double *T_old);
bool cut_ready(const mf_operator *op);
void cut_free(mf_operator *op);

// mf_operator_compute_inverse_diagonal with the RED_BLOCKS doubles of the reduction partials
// given by the caller (work = nullptr: the operator's own arena)
int compute_inverse_diagonal(mf_operator *op, double *inv_diag, double *work);
}  // namespace mf

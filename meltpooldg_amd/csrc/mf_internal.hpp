// Internal declarations shared by the HIP translation units of libmeltpooldg_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/meltpooldg_amd.h"

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
  HEAT_PGRAD = 6     // per-cell sum_q JxW grad(src) . grad(T): conductivity gradient (adjoint)
};

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
  // advection
  Number theta, diffusivity, mass_factor;
  Number dt_inv;
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
};

struct MeshInfo {
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

// launches (mf_cell.hip)
template <typename Number>
int launch_cell(int op, bool needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                const Number *src, Number *dst, hipStream_t stream);

// number of blocks the cell kernel uses (size of dot_partial)
int cell_blocks(const MeshInfo &m);

// slice-resident tangent kernel (mf_slab.hip)
template <typename Number>
int launch_slab_tangent(int op, bool needT, bool vel, const MeshInfo &m, const CellArgs<Number> &a,
                        const Number *src, Number *dst, hipStream_t stream);
int slab_blocks(const MeshInfo &m);

// tangent kernel selection: 0 = pencil kernel (mf_cell.hip), 1 = slice-resident (mf_slab.hip)
extern int g_tangent_variant;
inline int tangent_blocks(const MeshInfo &m, bool transpose = false)
{
  return (g_tangent_variant == 1 && !transpose) ? slab_blocks(m) : cell_blocks(m);
}

}  // namespace mf

struct mf_scratch {
  mf::MeshInfo mesh;
  mf::Tables tab;
  hipStream_t stream = nullptr;
  uint8_t *d_mask = nullptr;
  std::vector<uint8_t> h_mask;
};

struct mf_operator {
  int kind = 0;  // 0 heat, 1 advection
  mf_scratch *s = nullptr;
  mf_heat_params heat{};
  mf_advec_params advec{};
  double dt = 0.0;
  int dof_index = 0;
  const double *T = nullptr, *vel = nullptr, *k_cell = nullptr;
  float *T32 = nullptr, *vel32 = nullptr, *kc32 = nullptr;  // FP32 copies of the fields
  bool fields32_valid = false;
  int transpose = 0;  // vmult applies A^T (adjoint solves)
  int gather_plain = 0;  // vmult reads src without constraints (source mass products)
};

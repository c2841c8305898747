// A C++ caller of the one-phase CutFEM heat operator through meltpooldg_amd.hpp, written as the
// reference's HeatCutOperation time loop reads: the case of
// applications/mp-heat-transfer/cases/unidirectional_heat_transfer/tests/
// unidirectional_heat_transfer_one_phase_cut (17 x 17 cells of [0, 0.1]^2, level set 0.05 - y,
// T0 = x, dt 2, theta 0.5, gamma_A 1.5, gamma_M 0.75, Jacobi-GMRES rel 1e-12 abs 1e-10, Newton
// residual tolerance 1e-9), extruded to one cell of depth 1 in z.
//
//   cut_driver <out.bin>
// out.bin: int32 number of steps; the cut L2 norm after each step (float64)
#include <cstdint>
#include <cstdio>
#include <vector>

#include "meltpooldg_amd.hpp"

int main(int argc, char **argv)
{
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s out.bin\n", argv[0]);
    return 2;
  }
  try {
    const int nc = 17;
    const double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.1, 0.1, 1.0};
    mf::ScratchData scratch(1, nc, nc, 1, lo, hi);
    const size_t n = scratch.n_dofs();
    const int NX = nc + 1;
    std::vector<double> phi(n), x(n);
    for (size_t i = 0; i < n; ++i) {
      const int ix = int(i % NX), iy = int((i / NX) % NX);
      x[i] = 0.1 * ix / nc;
      phi[i] = 0.05 - 0.1 * iy / nc;
    }
    mf_heat_cut_params cp{};
    cp.material.law = MF_LAW_CONSTANT;
    cp.material.rho_cp = 7850.0 * 460.0;
    cp.material.k = 55.563;
    cp.theta = 0.5;
    cp.gamma_A = 1.5;
    cp.gamma_M = 0.75;
    cp.cell_side_length = 0.1 / nc;  // the 2D case's minimal cell size
    mf::HeatCutOperator heat(scratch, cp);
    heat.set_level_set(phi.data());
    mf_solver_desc gmres{MF_SOLVER_GMRES, 30, 0, 1e-10, 1e-12, MF_PRECOND_JACOBI,
                         MF_PRECISION_FP64, 0.0, 0};
    mf_newton_desc newton;
    mf::check(mf_newton_desc_init(&newton));
    newton.residual_tolerance = 1e-9;
    mf::DeviceVector<double> T(n), T_old(n);
    T.copy_from_host(x.data());
    const int n_steps = 5;
    std::vector<double> norms;
    for (int step = 0; step < n_steps; ++step) {
      T_old.copy_from(T);  // the predictor is the old solution
      heat.init_time_advance(2.0);
      heat.solve(T, T_old, gmres, newton);
      norms.push_back(heat.compute_cut_norm(T));
      std::printf("cut_driver: step %d ||T|| = %.5e\n", step + 1, norms.back());
    }
    std::FILE *out = std::fopen(argv[1], "wb");
    if (!out) return 2;
    const int32_t ns = n_steps;
    bool ok = std::fwrite(&ns, 4, 1, out) == 1 &&
              std::fwrite(norms.data(), 8, norms.size(), out) == norms.size();
    ok = (std::fclose(out) == 0) && ok;
    return ok ? 0 : 2;
  } catch (const mf::Exception &e) {
    std::fprintf(stderr, "cut_driver: mf::Exception %d: %s\n", e.code, e.what());
    return 1;
  }
}

// A C++ caller of libmeltpooldg_amd through the header-only layer meltpooldg_amd.hpp, written
// the way a MeltPoolDG operator/solver user reads (ScratchData, HeatDiffuseOperator::vmult,
// JacobiPreconditioner::update, LinearSolver::solve, NoConvergence).
//
//   hpp_driver <in.bin> <out.bin>
// in.bin:  int32 degree, nx, ny, nz; then n = n_dofs entries each of the Dirichlet mask (uint8),
//          the linearisation temperature T, a source vector src and a right-hand side b (float64)
// out.bin: int64 n; dst = A src [n]; x with A x = b (Jacobi-GMRES, rel. tol 1e-10) [n];
//          int32 GMRES iterations; int32 1 if max_iterations = 1 threw mf::NoConvergence
// The operator is the tanh-law heat Jacobian with its T-derivative terms
// (k = 1 + 0.5 tanh((T - 1700)/50), rho cp = 1, dt = 1e-3).
#include <cstdint>
#include <cstdio>
#include <vector>

#include "meltpooldg_amd.hpp"

namespace {
template <typename T>
bool read_n(std::FILE *f, T *p, size_t n)
{
  return std::fread(p, sizeof(T), n, f) == n;
}
}  // namespace

int main(int argc, char **argv)
{
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 2;
  }
  std::FILE *in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int32_t hdr[4];
  if (!read_n(in, hdr, 4)) return 2;
  try {
    mf::ScratchData s(hdr[0], hdr[1], hdr[2], hdr[3]);
    const size_t n = s.n_dofs();
    std::vector<uint8_t> mask(n);
    std::vector<double> T(n), src(n), b(n);
    if (!read_n(in, mask.data(), n) || !read_n(in, T.data(), n) || !read_n(in, src.data(), n) ||
        !read_n(in, b.data(), n))
      return 2;
    std::fclose(in);
    s.set_dirichlet_mask(mask);

    mf_heat_params p{};
    p.law = MF_LAW_TANH;
    p.t_dependent = 1;
    p.rho_cp = 1.0;
    p.k = 1.0;
    p.tanh_amp = 0.5;
    p.tanh_Tc = 1700.0;
    p.tanh_width = 50.0;
    mf::HeatDiffuseOperator op(s, p);
    op.reset_time_increment(1e-3);
    mf::DeviceVector<double> Td(n), x(n), dst(n), srcd(n), bd(n);
    Td.copy_from_host(T.data());
    srcd.copy_from_host(src.data());
    bd.copy_from_host(b.data());
    op.set_fields(&Td);

    op.vmult(dst, srcd);

    mf::JacobiPreconditioner pc(op);
    pc.update();
    mf::LinearSolverData data;
    data.solver_type = MF_SOLVER_GMRES;
    data.rel_tolerance = 1e-10;
    const int iterations = mf::LinearSolver::solve(op, x, bd, data, pc);

    int threw = 0;
    mf::DeviceVector<double> y(n);
    data.max_iterations = 1;
    try {
      mf::LinearSolver::solve(op, y, bd, data, pc);
    } catch (const mf::NoConvergence &) {
      threw = 1;
    }

    const std::vector<double> hd = dst.to_host(), hx = x.to_host();
    std::FILE *out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    const int64_t nn = int64_t(n);
    const int32_t tail[2] = {iterations, threw};
    bool ok = std::fwrite(&nn, 8, 1, out) == 1 && std::fwrite(hd.data(), 8, n, out) == n &&
              std::fwrite(hx.data(), 8, n, out) == n && std::fwrite(tail, 4, 2, out) == 2;
    ok = (std::fclose(out) == 0) && ok;
    std::printf("hpp_driver: n_dofs %zu, GMRES iterations %d, NoConvergence thrown %d\n", n,
                iterations, threw);
    return ok ? 0 : 2;
  } catch (const mf::Exception &e) {
    std::fprintf(stderr, "hpp_driver: mf::Exception %d: %s\n", e.code, e.what());
    return 1;
  }
}

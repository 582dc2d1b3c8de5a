// C++ test of the public interface (include/bpx.hpp): a solver_t written
// against the reference's interface (include/solver.hpp:47-49) is solved by
// cg::Solve (include/linearSolver.hpp:49-54), and the result is compared
// with the library's own BP5 solve of the same problem. Exit status: 0 ok,
// 1 failure, 77 no HIP device (the build and the link were still checked).
#include <cmath>
#include <cstdio>
#include <vector>

#include "bpx.hpp"

namespace {

// BP5's operator applied through solver_t::Operator, counting applications
class bp5_operator_t : public libbpx::solver_t {
 public:
  explicit bp5_operator_t(libbpx::bp_t& b) : bp(b) {}
  void Operator(libbpx::deviceMemory<double>& o_q, libbpx::deviceMemory<double>& o_Aq) override {
    ++applications;
    bp.Operator(o_q, o_Aq);
  }
  libbpx::bp_t& bp;
  int applications = 0;
};

// an operator that fails: the exception must reach the caller of Solve
class failing_operator_t : public libbpx::solver_t {
 public:
  void Operator(libbpx::deviceMemory<double>&, libbpx::deviceMemory<double>&) override {
    throw std::runtime_error("operator failed on purpose");
  }
};

int failures = 0;
void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL: %s\n", what);
    ++failures;
  }
}

}  // namespace

int main() {
  int ndev = 0;
  bpx_device_count(&ndev);
  if (ndev == 0) {
    std::printf("cg_solver_test: no HIP device; header and link checked\n");
    return 77;
  }
  try {
    libbpx::platform_t platform(0);
    bpx_mesh_opts o = {};
    o.N = 5;
    o.nx = o.ny = o.nz = 4;
    for (int c = 0; c < 3; ++c) {
      o.lo[c] = -1.0;
      o.dim[c] = 2.0;
    }
    o.perturb = 0.1;
    o.seed = 3;
    libbpx::mesh_t mesh(o);
    libbpx::bp_t bp(platform, mesh, 5, 1.0);
    const int N = bp.Ngather;
    std::vector<double> b((size_t)N);
    bpx_random_normal(7, 0, N, b.data());

    // cg::Solve through solver_t
    const int maxit = 60;
    auto o_x = platform.malloc<double>(std::vector<double>((size_t)N, 0.0));
    auto o_r = platform.malloc<double>(b);
    libbpx::cg solver(platform, N, 0);
    bp5_operator_t op(bp);
    const int its = solver.Solve(op, o_x, o_r, 0.0, maxit, 0);
    const std::vector<double> hist = solver.residualHistory();
    expect(its == maxit, "cg::Solve ran MAXIT iterations");
    expect(op.applications == maxit + 1, "one Operator per iteration plus the initial residual");

    // the library's own solve of the same problem
    auto o_x2 = platform.malloc<double>(std::vector<double>((size_t)N, 0.0));
    auto o_r2 = platform.malloc<double>(b);
    std::vector<double> hist2;
    const int its2 = bp.Solve(o_x2, o_r2, 0.0, maxit, &hist2);
    expect(its2 == maxit && hist.size() == hist2.size(), "same iteration counts");
    double hmax = 0;
    for (size_t k = 0; k < hist.size() && k < hist2.size(); ++k)
      hmax = std::fmax(hmax, std::fabs(hist[k] - hist2[k]) / hist2[k]);
    std::vector<double> x, x2;
    o_x.copyTo(x);
    o_x2.copyTo(x2);
    double dx = 0, nx = 0;
    for (int n = 0; n < N; ++n) {
      dx = std::fmax(dx, std::fabs(x[n] - x2[n]));
      nx = std::fmax(nx, std::fabs(x2[n]));
    }
    std::printf("cg::Solve via solver_t: %d its, residual history vs bp solve %.2e, solution %.2e\n", its, hmax,
                dx / nx);
    expect(hmax <= 1e-10, "residual histories agree to 1e-10");
    expect(dx <= 1e-9 * nx, "solutions agree to 1e-9");

    // linAlg_t and the ogs / halo objects of the same interface
    libbpx::linAlg_t la = platform.linAlg();
    const double rr = la.innerProd(N, o_r, o_r);
    expect(std::fabs(std::sqrt(rr) - hist.back()) <= 1e-6 * hist.back(), "innerProd(r, r) = last residual^2");
    libbpx::ogs::ogs_t ogs;
    ogs.SetupMasked(mesh, 1, platform);
    libbpx::ogs::halo_t halo;
    halo.SetupFromGather(ogs);
    halo.ExchangeStart(o_x, 1);
    halo.ExchangeFinish(o_x, 1);
    expect(ogs.Ngather == N && halo.Nhalo == 0, "ogs rows and an empty halo on one GPU");

    // an exception from the operator reaches the caller
    failing_operator_t bad;
    bool caught = false;
    try {
      auto o_r3 = platform.malloc<double>(b);
      solver.Solve(bad, o_x, o_r3, 0.0, 5, 0);
    } catch (const std::runtime_error& e) {
      caught = std::string(e.what()) == "operator failed on purpose";
    }
    expect(caught, "the operator's exception is rethrown by Solve");
  } catch (const std::exception& e) {
    std::printf("FAIL: exception %s\n", e.what());
    return 1;
  }
  if (failures) return 1;
  std::printf("cg_solver_test ok\n");
  return 0;
}

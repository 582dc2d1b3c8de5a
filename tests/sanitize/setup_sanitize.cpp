// AddressSanitizer / UndefinedBehaviorSanitizer run of the host code: the
// product's host setup (1D basis, box mesh, geometric and cubature factors,
// masked global ids, ogs T maps, GlobalToLocal; benchparanumal_amd/csrc/
// host/{basis,mesh,ogs}.cpp) and the CPU oracle (oracle/*.cpp), built with
// g++ -fsanitize=address,undefined (host code only; no GPU code is built or
// run). For each configuration the product's arrays are compared with the
// oracle's (maps exactly, factors to 1e-12), and the oracle's Ax, gather
// and CG run once, so every loop of both setups and of the oracle's
// operators executes under the sanitizers. Exit status 0 = clean and equal.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../benchparanumal_amd/csrc/host/bpx.hpp"
#include "../../oracle/oracle.hpp"

namespace {

int failures = 0;

template <typename A, typename B>
void same(const char* what, const A& a, const B& b) {
  bool ok = a.size() == b.size();
  for (size_t i = 0; ok && i < a.size(); ++i) ok = (a[i] == (typename A::value_type)b[i]);
  if (!ok) {
    std::printf("  MISMATCH %s (sizes %zu, %zu)\n", what, a.size(), b.size());
    ++failures;
  }
}

void close(const char* what, const std::vector<double>& a, const std::vector<double>& b, double tol) {
  double err = 0, mx = 0;
  bool ok = a.size() == b.size();
  for (size_t i = 0; ok && i < a.size(); ++i) {
    err = std::fmax(err, std::fabs(a[i] - b[i]));
    mx = std::fmax(mx, std::fabs(b[i]));
  }
  if (!ok || err > tol * std::fmax(mx, 1e-300)) {
    std::printf("  MISMATCH %s: max |diff| %.3e of max %.3e (sizes %zu, %zu)\n", what, err, mx, a.size(), b.size());
    ++failures;
  }
}

struct Config {
  int bp, N, nx, ny, nz, affine;
  double perturb;
  int elementType;  // 12 hex, 4 quad, 3 triangle, 6 tetrahedron
};

}  // namespace

int main() {
  const Config configs[] = {{5, 1, 1, 1, 1, 0, 0.0, 12}, {5, 2, 2, 3, 2, 0, 0.1, 12}, {3, 3, 2, 2, 1, 0, 0.1, 12},
                            {1, 4, 3, 2, 2, 1, 0.0, 12}, {6, 5, 2, 1, 2, 0, 0.2, 12}, {4, 7, 2, 2, 2, 0, 0.1, 12},
                            {5, 8, 1, 2, 1, 1, 0.0, 12}, {2, 11, 1, 1, 2, 0, 0.1, 12}, {5, 15, 1, 1, 1, 0, 0.1, 12},
                            {5, 4, 3, 2, 1, 0, 0.1, 4},  {3, 3, 2, 3, 1, 0, 0.1, 4},   {2, 2, 2, 2, 1, 0, 0.0, 4},
                            {5, 5, 3, 2, 1, 0, 0.1, 3},  {6, 3, 2, 2, 1, 0, 0.2, 3},   {5, 3, 2, 2, 2, 0, 0.1, 6},
                            {4, 2, 1, 2, 1, 0, 0.1, 6}};
  for (const Config& c : configs) {
    std::printf("BP%d p=%d %dx%dx%d affine=%d perturb=%.1f element %d\n", c.bp, c.N, c.nx, c.ny, c.nz, c.affine,
                c.perturb, c.elementType);
    const bool hex = c.elementType == 12;
    const bool cub = c.bp <= 4 && (hex || c.elementType == 4);
    const int Nfields = (c.bp % 2 == 0) ? (c.elementType == 3 || c.elementType == 4 ? 2 : 3) : 1;
    // product host setup
    bpx::MeshOpts o;
    o.N = c.N;
    o.nx = c.nx;
    o.ny = c.ny;
    o.nz = c.nz;
    o.affine = c.affine;
    o.cubature = cub ? 1 : 0;
    o.perturb = c.perturb;
    o.seed = 7;
    o.elementType = c.elementType;
    bpx::mesh_t m;
    if (m.Setup(o)) {
      std::printf("  product mesh setup failed: %s\n", bpx::last_error());
      return 2;
    }
    std::vector<bpx::hlong> ids;
    m.MaskedGlobalIds(Nfields, ids);
    bpx::ogs::ogs_t g;
    if (g.Setup((bpx::dlong)ids.size(), ids.data())) {
      std::printf("  product ogs setup failed: %s\n", bpx::last_error());
      return 2;
    }
    std::vector<bpx::dlong> G2L(ids.size());
    g.SetupGlobalToLocalMapping(G2L.data());
    std::vector<double> tri;
    if (hex) m.TrilinearCoefficients(tri);
    const bool trilinear = !hex || m.NodesAreTrilinear();
    if (!trilinear && !c.affine) {
      std::printf("  box nodes not trilinear\n");
      ++failures;
    }

    // oracle setup
    orc::Problem P;
    P.bp = c.bp;
    P.N = c.N;
    P.nx = c.nx;
    P.ny = c.ny;
    P.nz = c.nz;
    P.affine = c.affine;
    P.lambda = 1.0;
    P.perturb = c.perturb;
    P.seed = 7;
    P.elementType = c.elementType;
    if (orc::problem_setup(P)) {
      std::printf("  oracle setup failed: %s\n", P.error.c_str());
      return 2;
    }
    same("rowStartsT", g.rowStartsT, P.rowStartsT);
    same("colIdsT", g.colIdsT, P.colIdsT);
    same("blockRowStartsT", g.blockRowStartsT, P.blockRowStartsT);
    same("GlobalToLocal", G2L, P.GlobalToLocal);
    close("x", m.x, P.x, 1e-13);
    if (c.elementType == 3 || c.elementType == 6) {
      close("MM", m.MM, P.MM, 1e-12);
      close("Dsim", m.Dsim, P.Dsim, 1e-11);
      close("Ssim", m.Ssim, P.Ssim, 1e-11);
    }
    if (cub && !c.affine) {
      close("cubwJ", m.cubwJ, P.cubwJ, 1e-12);
      if (c.bp >= 3) close("cubggeo", m.cubggeo, P.cubggeo, 1e-12);
    } else {
      close("wJ", m.wJ, P.wJ, 1e-12);
      close("ggeo", m.ggeo, P.ggeo, 1e-12);
    }

    // the oracle's operators and CG
    std::vector<double> q(P.Ngather), AqL(P.NLocal), Aq(P.Ngather), r(P.Ngather), x(P.Ngather, 0.0);
    for (bpx::dlong n = 0; n < P.Ngather; ++n) q[n] = orc::rng_normal(3, n);
    orc::ax_local(P, q.data(), AqL.data());
    orc::ogs_gather(P, AqL.data(), Aq.data());
    std::vector<double> back(P.NLocal);
    orc::ogs_scatter(P, Aq.data(), back.data());
    r = q;
    std::vector<double> hist(21), alpha(20);
    const int its = orc::cg_solve(P, x.data(), r.data(), 0.0, 20, hist.data(), alpha.data());
    if (its != 20 && P.Ngather > 20) {
      std::printf("  oracle CG ran %d iterations\n", its);
      ++failures;
    }
  }
  std::printf("%s (%d mismatches)\n", failures ? "FAILED" : "sanitized setup ok", failures);
  return failures ? 1 : 0;
}

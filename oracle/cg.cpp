// ORACLE — test infrastructure only (see oracle.hpp).
// Restates libs/core/linAlg.cpp:38-89 with the two-stage block reductions of
// libs/core/okl/linAlg{InnerProd,Norm2}.okl, the fused update of
// libs/core/okl/linearSolverUpdateCG.okl and cg::Solve/UpdateCG of
// libs/core/linearSolverCG.cpp:62-156. The reduction order (per-thread
// strided partials, then an LDS tree, then a one-block pass over the block
// partials) is reproduced exactly.
#include <cmath>
#include <vector>

#include "oracle.hpp"

namespace orc {

static const int LINALG_BLOCKSIZE = 256;  // libs/core/linAlgSetup.cpp:33
static const int CG_BLOCKSIZE = 512;      // libs/core/linearSolverCG.cpp:31

// The LDS tree of the OKL reductions: halve while t < s, then s[0]+s[1].
static double tree(std::vector<double>& s, int bs) {
  for (int w = bs / 2; w >= 2; w /= 2)
    for (int t = 0; t < w; ++t) s[t] += s[t + w];
  return s[0] + s[1];
}

template <class F>
static double two_stage(dlong N, int bs, F term) {
  int Nblocks = (N + bs - 1) / bs;
  Nblocks = (Nblocks > bs) ? bs : Nblocks;
  std::vector<double> partial(Nblocks);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < Nblocks; ++b) {
    std::vector<double> s(bs);
    for (int t = 0; t < bs; ++t) {
      double acc = 0.0;
      for (long long id = t + (long long)b * bs; id < N; id += (long long)bs * Nblocks) acc += term(id);
      s[t] = acc;
    }
    partial[b] = tree(s, bs);
  }
  std::vector<double> s(bs);
  for (int t = 0; t < bs; ++t) {
    double acc = 0.0;
    for (int id = t; id < Nblocks; id += bs) acc += partial[id];
    s[t] = acc;
  }
  return tree(s, bs);
}

// libs/core/okl/linAlgSet.okl:29-36
void la_set(dlong N, double alpha, double* x) {
#pragma omp parallel for
  for (dlong n = 0; n < N; ++n) x[n] = alpha;
}

// libs/core/okl/linAlgAXPY.okl:29-38
void la_axpy(dlong N, double alpha, const double* x, double beta, double* y) {
#pragma omp parallel for
  for (dlong n = 0; n < N; ++n) y[n] = alpha * x[n] + beta * y[n];
}

// libs/core/linAlg.cpp:71-89 + libs/core/okl/linAlgInnerProd.okl:28-100
double la_inner_prod(dlong N, const double* x, const double* y) {
  return two_stage(N, LINALG_BLOCKSIZE, [&](long long id) { return x[id] * y[id]; });
}

// libs/core/linAlg.cpp:50-67 + libs/core/okl/linAlgNorm2.okl:27-96
double la_norm2(dlong N, const double* x) {
  return std::sqrt(two_stage(N, LINALG_BLOCKSIZE, [&](long long id) { return x[id] * x[id]; }));
}

// libs/core/linearSolverCG.cpp:131-156 + okl/linearSolverUpdateCG.okl:29-103
// (r <- r - alpha*Ap with r.r; the x update is done by the caller).
double la_update_cg(dlong N, const double* Ap, double alpha, double* r) {
#pragma omp parallel for
  for (dlong n = 0; n < N; ++n) r[n] = r[n] - alpha * Ap[n];
  return two_stage(N, CG_BLOCKSIZE, [&](long long id) { return r[id] * r[id]; });
}

// bpN_t::Operator (BP/BP5/src/bp5Operator.cpp:29-85): Ax over the local
// elements then the T gather.
static void Operator(const Problem& P, const double* q, double* Aq, std::vector<double>& AqL) {
  ax_local(P, q, AqL.data());
  ogs_gather(P, AqL.data(), Aq);
}

// libs/core/linearSolverCG.cpp:62-129. rnorm_hist[0] is the initial
// residual norm, rnorm_hist[it] the norm after iteration it; alpha_hist[it-1]
// the step of iteration it. Returns the number of iterations.
int cg_solve(const Problem& P, double* x, double* r, double tol, int maxit, double* rnorm_hist,
             double* alpha_hist) {
  const dlong N = P.Ngather;
  std::vector<double> p(N, 0.0), Ap(N, 0.0), AqL(P.NLocal);
  double rdotr1 = 0.0, rdotr2 = 0.0, alpha = 0.0, beta = 0.0, pAp = 0.0, rdotr = 0.0;
  Operator(P, x, Ap.data(), AqL);
  la_axpy(N, -1.0, Ap.data(), 1.0, r);
  rdotr = la_norm2(N, r);
  rdotr = rdotr * rdotr;
  const double TOL = std::max(tol * tol * rdotr, tol * tol);
  if (rnorm_hist) rnorm_hist[0] = std::sqrt(rdotr);
  int iter;
  for (iter = 0; iter < maxit; ++iter) {
    if (rdotr <= TOL) break;
    rdotr2 = rdotr1;
    rdotr1 = rdotr;
    beta = (iter == 0) ? 0.0 : rdotr1 / rdotr2;
    la_axpy(N, 1.0, r, beta, p.data());
    Operator(P, p.data(), Ap.data(), AqL);
    pAp = la_inner_prod(N, p.data(), Ap.data());
    alpha = rdotr1 / pAp;
    rdotr = la_update_cg(N, Ap.data(), alpha, r);
    la_axpy(N, alpha, p.data(), 1.0, x);
    if (rnorm_hist) rnorm_hist[iter + 1] = std::sqrt(rdotr);
    if (alpha_hist) alpha_hist[iter] = alpha;
  }
  return iter;
}

}  // namespace orc

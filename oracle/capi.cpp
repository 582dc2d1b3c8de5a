// ORACLE — test infrastructure only (see oracle.hpp).
// Flat C entry points so the Python tests (ctypes) and bench.py's CPU
// baseline leg can drive the restatement.
#include <omp.h>

#include <cstring>
#include <string>
#include <vector>

#include "oracle.hpp"

using namespace orc;

static thread_local std::string g_err;

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

int orc_jacobi_gq(double alpha, double beta, int N, double* x, double* w) {
  return JacobiGQ(alpha, beta, N, x, w);
}
int orc_jacobi_gll(int N, double* x, double* w) { return JacobiGLL(N, x, w); }
double orc_jacobi_p(double a, double alpha, double beta, int N) { return JacobiP(a, alpha, beta, N); }
void orc_vandermonde1d(int N, int Npoints, const double* r, double* V) {
  Vandermonde1D(N, Npoints, r, V);
}
int orc_mass_matrix1d(int N, const double* r, double* MM) {
  std::vector<double> V((N + 1) * (N + 1));
  Vandermonde1D(N, N + 1, r, V.data());
  return MassMatrix1D(N + 1, V.data(), MM);
}
int orc_dmatrix1d(int N, int Nout, const double* rin, const double* rout, double* D) {
  return Dmatrix1D(N, Nout, rin, rout, D);
}
int orc_interp1d(int N, int Nout, const double* rin, const double* rout, double* I) {
  return InterpolationMatrix1D(N, Nout, rin, rout, I);
}
unsigned long long orc_rng_u64(unsigned long long seed, unsigned long long counter) {
  return rng_u64(seed, counter);
}
double orc_rng_u01(unsigned long long seed, unsigned long long counter) { return rng_u01(seed, counter); }
double orc_rng_normal(unsigned long long seed, unsigned long long counter) {
  return rng_normal(seed, counter);
}

// Problem ------------------------------------------------------------------
void* orc_problem_create(int bp, int N, int nx, int ny, int nz, int affine, double lambda,
                         const double* lo, const double* dim, double perturb,
                         unsigned long long seed) {
  Problem* P = new Problem;
  P->bp = bp;
  P->N = N;
  P->nx = nx;
  P->ny = ny;
  P->nz = nz;
  P->affine = affine;
  P->lambda = lambda;
  for (int c = 0; c < 3; ++c) {
    if (lo) P->lo[c] = lo[c];
    if (dim) P->dim[c] = dim[c];
  }
  P->perturb = perturb;
  P->seed = seed;
  const int st = problem_setup(*P);
  if (st) {
    g_err = P->error.empty() ? "setup failed" : P->error;
    delete P;
    return nullptr;
  }
  return P;
}

// a quad (4), triangle (3) or tetrahedron (6) box problem
void* orc_problem_create_family(int elementType, int bp, int N, int nx, int ny, int nz, double lambda,
                                const double* lo, const double* dim, double perturb, unsigned long long seed) {
  Problem* P = new Problem;
  P->elementType = elementType;
  P->bp = bp;
  P->N = N;
  P->nx = nx;
  P->ny = ny;
  P->nz = nz;
  P->lambda = lambda;
  for (int c = 0; c < 3; ++c) {
    if (lo) P->lo[c] = lo[c];
    if (dim) P->dim[c] = dim[c];
  }
  P->perturb = perturb;
  P->seed = seed;
  const int st = problem_setup(*P);
  if (st) {
    g_err = P->error.empty() ? "setup failed" : P->error;
    delete P;
    return nullptr;
  }
  return P;
}

void orc_problem_destroy(void* h) { delete static_cast<Problem*>(h); }

// ints: bp N Nq Np cubNq cubNp Nelements Nfields affine NLocal Ngather nnzT NrowBlocksT
int orc_problem_info(void* h, long long* out) {
  const Problem& P = *static_cast<Problem*>(h);
  const long long v[] = {P.bp, P.N, P.Nq, P.Np, P.cubNq, P.cubNp, P.Nelements, P.Nfields,
                         P.affine, P.NLocal, P.Ngather, P.nnzT, P.NrowBlocksT};
  std::memcpy(out, v, sizeof(v));
  return (int)(sizeof(v) / sizeof(v[0]));
}

// dtype: 0 = f64, 1 = i32, 2 = i64
const void* orc_problem_array(void* h, const char* name, long long* n, int* dtype) {
  Problem& P = *static_cast<Problem*>(h);
  const std::string s(name);
#define F64(field)                       \
  if (s == #field) {                     \
    *n = (long long)P.field.size();      \
    *dtype = 0;                          \
    return P.field.data();               \
  }
#define I32(field)                       \
  if (s == #field) {                     \
    *n = (long long)P.field.size();      \
    *dtype = 1;                          \
    return P.field.data();               \
  }
#define I64(field)                       \
  if (s == #field) {                     \
    *n = (long long)P.field.size();      \
    *dtype = 2;                          \
    return P.field.data();               \
  }
  F64(r) F64(s) F64(t) F64(gllz) F64(gllw) F64(D) F64(MM) F64(V) F64(cubr) F64(cubw)
  F64(cubInterp) F64(cubProject) F64(cubD) F64(invV) F64(EX) F64(EY) F64(EZ) F64(x) F64(y)
  F64(z) F64(wJ) F64(ggeo) F64(cubwJ) F64(cubggeo) F64(Dsim) F64(Ssim)
  I32(faceNodes) I32(EToE) I32(EToF) I32(vmapM) I32(vmapP) I32(elementList)
  I32(rowStartsT) I32(colIdsT) I32(blockRowStartsT) I32(GlobalToLocal)
  I64(EToV) I64(globalIds) I64(maskedIds)
#undef F64
#undef I32
#undef I64
  *n = 0;
  *dtype = -1;
  return nullptr;
}

void orc_ax_local(void* h, const double* q, double* AqL) { ax_local(*static_cast<Problem*>(h), q, AqL); }
void orc_gather(void* h, const double* vL, double* v) { ogs_gather(*static_cast<Problem*>(h), vL, v); }
void orc_scatter(void* h, const double* v, double* vL) { ogs_scatter(*static_cast<Problem*>(h), v, vL); }
void orc_rhs_forcing_local(void* h, double* rhsL) { rhs_forcing_local(*static_cast<Problem*>(h), rhsL); }

void orc_operator(void* h, const double* q, double* Aq) {
  Problem& P = *static_cast<Problem*>(h);
  std::vector<double> AqL(P.NLocal);
  ax_local(P, q, AqL.data());
  ogs_gather(P, AqL.data(), Aq);
}

void orc_set_lambda(void* h, double lambda) { static_cast<Problem*>(h)->lambda = lambda; }

// variable coefficient: kappa on the node lattice (nullptr = constant 1)
int orc_set_kappa(void* h, const double* kappa) {
  Problem& P = *static_cast<Problem*>(h);
  const int st = set_kappa(P, kappa);
  if (st) g_err = P.error;
  return st;
}
long long orc_lattice_nodes(void* h) { return lattice_nodes(*static_cast<Problem*>(h)); }
void orc_lattice_index(void* h, long long* out) {
  const Problem& P = *static_cast<Problem*>(h);
  for (dlong e = 0; e < P.Nelements; ++e)
    for (int n = 0; n < P.Np; ++n) out[(size_t)e * P.Np + n] = lattice_index(P, e, n);
}
// dL/dkappa on the lattice from the global solution u and adjoint lam
void orc_kappa_grad(void* h, const double* u, const double* lam, double* grad) {
  const Problem& P = *static_cast<Problem*>(h);
  std::vector<double> gL(P.NLocal);
  kappa_grad_local(P, u, lam, gL.data());
  lattice_sum(P, gL.data(), grad);
}

// Rebuild the ogs maps without (mask=0) or with (mask=1) the Dirichlet mask.
int orc_problem_set_mask(void* h, int mask) {
  Problem& P = *static_cast<Problem*>(h);
  P.mask = mask;
  const int st = problem_setup(P);
  if (st) g_err = P.error;
  return st;
}

// libs/ogs/ogsOperator.cpp:474-546 on a caller-given rowStarts[Nrows+1];
// blockStarts needs Nrows+1 entries. Returns the block count, or -1.
int orc_block_rows(int Nrows, const int* rowStarts, int* blockStarts) {
  std::vector<dlong> rs(rowStarts, rowStarts + Nrows + 1), bs;
  dlong Nb = 0;
  if (blockRows(Nrows, rs, Nb, bs)) return -1;
  for (size_t i = 0; i < bs.size(); ++i) blockStarts[i] = bs[i];
  return Nb;
}

int orc_cg_solve(void* h, double* x, double* r, double tol, int maxit, double* rnorm_hist,
                 double* alpha_hist) {
  return cg_solve(*static_cast<Problem*>(h), x, r, tol, maxit, rnorm_hist, alpha_hist);
}

void orc_la_axpy(int N, double alpha, const double* x, double beta, double* y) {
  la_axpy(N, alpha, x, beta, y);
}
double orc_la_inner_prod(int N, const double* x, const double* y) { return la_inner_prod(N, x, y); }
double orc_la_norm2(int N, const double* x) { return la_norm2(N, x); }
double orc_la_update_cg(int N, const double* Ap, double alpha, double* r) {
  return la_update_cg(N, Ap, alpha, r);
}

}  // extern "C"

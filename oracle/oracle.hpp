// ============================================================================
// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement (C++17 + OpenMP) of paranumal/benchparanumal's BK/BP hex
// hot path: 1D basis construction, box mesh, node connectivity, masked ogs
// maps, geometric factors, the BP1-BP6 hex Ax loop nests, the ogs gather and
// scatter, the linAlg reductions and the CG solver. Every function cites the
// reference file:line it restates.
//
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
// load this library, and only as the checker (or as the timed CPU baseline),
// never as part of the product path.
//
// Pinning (what checks each part; see tests/test_oracle_*.py):
//   basis   : numpy Legendre roots / Gauss-Legendre, closed-form GLL weights
//             and D, polynomial exactness of D, I, MM, invV.
//   geometry: closed form on the undistorted box, finite differences of the
//             trilinear map, sum(wJ) = volume.
//   Ax      : dense element matrices assembled by brute force (numpy),
//             symmetry, affine == non-affine on the undistorted box,
//             BP2/4/6 == 3 x BP1/3/5, A*1 = lambda*M*1 without the mask.
//   ogs     : brute-force first-appearance grouping, closed-form counts and
//             multiplicities, hand-derived blockRows answers.
//   CG      : README.md:109-113 residual norms (soft, 1e-6), textbook
//             properties (A-orthogonality, residual recurrence).
//   quad/tri/tet (other.cpp): node sets (edge nodes = GLL, symmetry),
//             MM and D exact on polynomials (numpy monomial integrals and
//             derivatives), S = D^T MM D, areas/volumes, node grouping =
//             coincident coordinates, Ax against numpy-assembled element
//             matrices.
// ============================================================================
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace orc {

typedef int dlong;         // include/types.h:32-69 (device index width)
typedef long long hlong;   // include/types.h (global id width)

// ---------------------------------------------------------------- basis --
double JacobiP(double a, double alpha, double beta, int N);
double GradJacobiP(double a, double alpha, double beta, int N);
int JacobiGQ(double alpha, double beta, int N, double* x, double* w);
int JacobiGLL(int N, double* x, double* w);
void Vandermonde1D(int N, int Npoints, const double* r, double* V);
void GradVandermonde1D(int N, int Npoints, const double* r, double* Vr);
int MatrixInverse(int n, double* A);
int MatrixRightSolve(int NrowsA, int NcolsA, const double* A, int NrowsB,
                     int NcolsB, const double* B, double* C);
int MassMatrix1D(int Np, const double* V, double* MM);
int Dmatrix1D(int N, int NpointsOut, const double* rIn, const double* rOut,
              double* D);
int InterpolationMatrix1D(int N, int NpointsOut, const double* rIn,
                          const double* rOut, double* I);

// -------------------------------------------------------- counter RNG --
// SplitMix64: output number `counter` (0-based) of the stream seeded `seed`.
uint64_t rng_u64(uint64_t seed, uint64_t counter);
double rng_u01(uint64_t seed, uint64_t counter);
double rng_normal(uint64_t seed, uint64_t counter);

// -------------------------------------------------------------- problem --
struct Problem {
  // configuration
  int bp = 5, N = 1, Nq = 2, Np = 8, Nfp = 4;
  int cubN = 2, cubNq = 3, cubNp = 27;
  int nx = 1, ny = 1, nz = 1, Nelements = 1, Nfields = 1, affine = 0;
  double lambda = 1.0;
  double lo[3] = {-1, -1, -1}, dim[3] = {2, 2, 2};
  double perturb = 0.0;
  uint64_t seed = 0;
  int mask = 1;  // 0: keep boundary nodes (no Dirichlet mask)
  int Nverts = 8, Nfaces = 6;
  // element family (libs/mesh/meshSetElementType.cpp codes): 12 hex,
  // 4 quad, 3 triangle, 6 tetrahedron (other.cpp); dimension 2 or 3
  int elementType = 12, dimension = 3;
  // triangles / tetrahedra: packed D (dimension x Np x Np) and stiffness
  // (3 or 6 x Np x Np); MM then holds the Np x Np mass matrix
  std::vector<double> Dsim, Ssim;

  // reference element
  std::vector<double> r, s, t;
  std::vector<double> gllz, gllw, D, MM, V;
  std::vector<double> cubr, cubw, cubInterp, cubProject, cubD, invV;
  std::vector<int> faceNodes;

  // mesh
  std::vector<hlong> EToV;
  std::vector<double> EX, EY, EZ;
  std::vector<dlong> EToE;
  std::vector<int> EToF;
  std::vector<dlong> vmapM, vmapP;
  std::vector<double> x, y, z;
  std::vector<double> wJ, ggeo, cubwJ, cubggeo;
  std::vector<hlong> globalIds, maskedIds;
  std::vector<dlong> elementList;

  // ogs (T maps only; the N maps depend on a random pick, off the path)
  dlong NLocal = 0, Ngather = 0, nnzT = 0, NrowBlocksT = 0;
  std::vector<dlong> rowStartsT, colIdsT, blockRowStartsT, GlobalToLocal;

  // variable coefficient (extension, BP5 only): kappa per local node,
  // scaling the stiffness part; empty = 1
  std::vector<double> kappaL;

  std::string error;
};

int problem_setup(Problem& P);
int problem_setup_other(Problem& P);  // quad / tri / tet (other.cpp)
int ogs_setup_maps(Problem& P);       // masked ids -> T maps (mesh.cpp)
int blockRows(dlong Nrows, const std::vector<dlong>& rowStarts, dlong& Nblocks,
              std::vector<dlong>& blockStarts);

// ----------------------------------------------------------- operators --
void ax_local(const Problem& P, const double* q, double* AqL);
void ogs_gather(const Problem& P, const double* vL, double* v);
void ogs_scatter(const Problem& P, const double* v, double* vL);
void rhs_forcing_local(const Problem& P, double* rhsL);
void ax_local_other(const Problem& P, const double* q, double* AqL);
void rhs_forcing_local_other(const Problem& P, double* rhsL);

// ------------------------------------- variable coefficient and adjoint --
hlong lattice_index(const Problem& P, dlong e, int n);
int set_kappa(Problem& P, const double* kappa);  // lattice-ordered; nullptr clears
hlong lattice_nodes(const Problem& P);
void kappa_grad_local(const Problem& P, const double* u, const double* lam, double* gradL);
void lattice_sum(const Problem& P, const double* vL, double* v);

// --------------------------------------------------------------- linAlg --
void la_set(dlong N, double alpha, double* x);
void la_axpy(dlong N, double alpha, const double* x, double beta, double* y);
double la_inner_prod(dlong N, const double* x, const double* y);
double la_norm2(dlong N, const double* x);
double la_update_cg(dlong N, const double* Ap, double alpha, double* r);

int cg_solve(const Problem& P, double* x, double* r, double tol, int maxit,
             double* rnorm_hist, double* alpha_hist);

}  // namespace orc

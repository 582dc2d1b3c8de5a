// ORACLE — test infrastructure only (see oracle.hpp).
// Restates libs/mesh/meshBasis1D.cpp and the dense helpers of
// libs/core/linAlgMatrix{Eig,Inverse,RightSolve}.cpp. The reference calls
// LAPACK (dgeev, dgetrf/dgetri, dgesv); this file replaces those calls with
// self-contained routines (cyclic Jacobi eigen-solve, Gauss-Jordan inverse,
// LU with partial pivoting) that solve the same dense problems.
#include <algorithm>
#include <cmath>
#include <vector>

#include "oracle.hpp"

namespace orc {

// libs/mesh/meshBasis1D.cpp:170-202 (orthonormal Jacobi polynomial by
// three-term recurrence).
double JacobiP(double a, double alpha, double beta, int N) {
  const double ax = a;
  std::vector<double> P(N + 1);
  const double gamma0 = std::pow(2, (alpha + beta + 1)) / (alpha + beta + 1) *
                        std::tgamma(1 + alpha) * std::tgamma(1 + beta) /
                        std::tgamma(1 + alpha + beta);
  const double p0 = 1.0 / std::sqrt(gamma0);
  if (N == 0) return p0;
  P[0] = p0;
  const double gamma1 = (alpha + 1) * (beta + 1) / (alpha + beta + 3) * gamma0;
  const double p1 =
      ((alpha + beta + 2) * ax / 2 + (alpha - beta) / 2) / std::sqrt(gamma1);
  if (N == 1) return p1;
  P[1] = p1;
  double aold = 2 / (2 + alpha + beta) *
                std::sqrt((alpha + 1.) * (beta + 1.) / (alpha + beta + 3.));
  for (int i = 1; i <= N - 1; ++i) {
    double h1 = 2. * i + alpha + beta;
    double anew = 2. / (h1 + 2.) *
                  std::sqrt((i + 1.) * (i + 1. + alpha + beta) * (i + 1 + alpha) *
                            (i + 1 + beta) / (h1 + 1) / (h1 + 3));
    double bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2);
    P[i + 1] = 1. / anew * (-aold * P[i - 1] + (ax - bnew) * P[i]);
    aold = anew;
  }
  return P[N];
}

// libs/mesh/meshBasis1D.cpp:204-213
double GradJacobiP(double a, double alpha, double beta, int N) {
  double PNr = 0;
  if (N > 0)
    PNr = std::sqrt(N * (N + alpha + beta + 1.)) *
          JacobiP(a, alpha + 1.0, beta + 1.0, N - 1);
  return PNr;
}

// Symmetric eigen-decomposition A = V diag(lam) V^T by cyclic Jacobi
// rotations. Stands in for the dgeev call of
// libs/core/linAlgMatrixEig.cpp:29-64 (eigenvectors stored column-wise,
// VR[row*n+col], unit 2-norm).
static void SymEig(int n, std::vector<double> A, double* lam, double* VR) {
  for (int i = 0; i < n * n; ++i) VR[i] = 0;
  for (int i = 0; i < n; ++i) VR[i * n + i] = 1;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p * n + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double theta = (A[q * n + q] - A[p * n + p]) / (2 * apq);
        const double tt = (theta >= 0 ? 1.0 : -1.0) /
                          (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(tt * tt + 1), s = tt * c;
        for (int k = 0; k < n; ++k) {  // A <- A J
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {  // V <- V J
          const double vkp = VR[k * n + p], vkq = VR[k * n + q];
          VR[k * n + p] = c * vkp - s * vkq;
          VR[k * n + q] = s * vkp + c * vkq;
        }
      }
    }
  }
  for (int i = 0; i < n; ++i) lam[i] = A[i * n + i];
}

// libs/mesh/meshBasis1D.cpp:266-338 (Golub-Welsch: eigenvalues of the
// Jacobi matrix, weights from the first eigenvector components, then the
// "sloppy sort").
int JacobiGQ(double alpha, double beta, int N, double* x, double* w) {
  if (N == 0) {
    x[0] = (alpha - beta) / (alpha + beta + 2);
    w[0] = 2;
  }
  const int n1 = N + 1;
  std::vector<double> J(n1 * n1, 0.0), h1(n1);
  for (int n = 0; n <= N; ++n) h1[n] = 2 * n + alpha + beta;
  for (int n = 0; n <= N; ++n) {
    J[n * n1 + n] += -0.5 * (alpha * alpha - beta * beta) /
                     ((h1[n] + 2) * h1[n]) * 2;
    if (n < N) {
      const double v = (2. / (h1[n] + 2.)) *
                       std::sqrt((n + 1) * (n + 1 + alpha + beta) * (n + 1 + alpha) *
                                 (n + 1 + beta) / ((h1[n] + 1) * (h1[n] + 3)));
      J[n * n1 + n + 1] += v;
      J[(n + 1) * n1 + n] += v;
    }
  }
  double eps = 1;
  while (1 + eps > 1) eps = eps / 2.;
  if (alpha + beta < 10 * eps) J[0] = 0;
  // J is the symmetric tridiagonal Jacobi matrix (the reference writes the
  // MATLAB "J = J + J'" out as a doubled diagonal plus each off-diagonal
  // entry placed on both sides, meshBasis1D.cpp:288-302).
  std::vector<double> VR(n1 * n1);
  SymEig(n1, J, x, VR.data());
  for (int n = 0; n <= N; ++n) {
    w[n] = std::pow(VR[0 * n1 + n], 2) *
           (std::pow(2, alpha + beta + 1) / (alpha + beta + 1)) *
           std::tgamma(alpha + 1) * std::tgamma(beta + 1) /
           std::tgamma(alpha + beta + 1);
  }
  for (int n = 0; n <= N; ++n) {
    for (int m = n + 1; m <= N; ++m) {
      if (x[n] > x[m]) {
        std::swap(x[n], x[m]);
        std::swap(w[n], w[m]);
      }
    }
  }
  return 0;
}

// libs/mesh/meshBasis1D.cpp:58-89
void Vandermonde1D(int N, int Npoints, const double* r, double* V) {
  const int Np = N + 1;
  for (int n = 0; n < Npoints; n++)
    for (int i = 0; i < Np; i++) V[n * Np + i] = JacobiP(r[n], 0, 0, i);
}

void GradVandermonde1D(int N, int Npoints, const double* r, double* Vr) {
  const int Np = N + 1;
  for (int n = 0; n < Npoints; n++)
    for (int i = 0; i < Np; i++) Vr[n * Np + i] = GradJacobiP(r[n], 0, 0, i);
}

// In-place inverse (Gauss-Jordan, partial pivoting); stands in for
// libs/core/linAlgMatrixInverse.cpp:29-54 (dgetrf + dgetri).
int MatrixInverse(int n, double* A) {
  std::vector<double> M(A, A + n * n), Inv(n * n, 0.0);
  for (int i = 0; i < n; ++i) Inv[i * n + i] = 1;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(M[r * n + c]) > std::fabs(M[piv * n + c])) piv = r;
    if (M[piv * n + c] == 0) return -1;
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(M[c * n + k], M[piv * n + k]);
        std::swap(Inv[c * n + k], Inv[piv * n + k]);
      }
    const double d = M[c * n + c];
    for (int k = 0; k < n; ++k) {
      M[c * n + k] /= d;
      Inv[c * n + k] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const double f = M[r * n + c];
      if (f == 0) continue;
      for (int k = 0; k < n; ++k) {
        M[r * n + k] -= f * M[c * n + k];
        Inv[r * n + k] -= f * Inv[c * n + k];
      }
    }
  }
  std::copy(Inv.begin(), Inv.end(), A);
  return 0;
}

// C = A / B, i.e. C*B = A (A: NrowsA x NcolsA, B: NrowsB x NcolsB square).
// Stands in for libs/core/linAlgMatrixRightSolve.cpp:139-169 (dgesv on the
// transposed system). Solved as C = A * inv(B).
int MatrixRightSolve(int NrowsA, int NcolsA, const double* A, int NrowsB,
                     int NcolsB, const double* B, double* C) {
  if (NrowsB != NcolsB || NcolsA != NcolsB) return -1;
  const int n = NrowsB;
  std::vector<double> Binv(B, B + n * n);
  if (MatrixInverse(n, Binv.data())) return -1;
  for (int i = 0; i < NrowsA; ++i)
    for (int j = 0; j < n; ++j) {
      double acc = 0;
      for (int k = 0; k < n; ++k) acc += A[i * NcolsA + k] * Binv[k * n + j];
      C[i * n + j] = acc;
    }
  return 0;
}

// libs/mesh/meshBasis1D.cpp:96-112 (MM = inv(V V^T)).
int MassMatrix1D(int Np, const double* V, double* MM) {
  for (int n = 0; n < Np; ++n)
    for (int m = 0; m < Np; ++m) {
      double res = 0;
      for (int i = 0; i < Np; ++i) res += V[n * Np + i] * V[m * Np + i];
      MM[n * Np + m] = res;
    }
  return MatrixInverse(Np, MM);
}

// libs/mesh/meshBasis1D.cpp:232-261 (GLL: interior nodes from
// JacobiGQ(1,1,N-2); weights = row sums of the exact mass matrix).
int JacobiGLL(int N, double* x, double* w) {
  x[0] = -1.;
  x[N] = 1.;
  if (N > 1) {
    std::vector<double> wtmp(N - 1);
    JacobiGQ(1, 1, N - 2, x + 1, wtmp.data());
  }
  if (w) {
    const int n1 = N + 1;
    std::vector<double> V(n1 * n1), MM(n1 * n1);
    Vandermonde1D(N, n1, x, V.data());
    if (MassMatrix1D(n1, V.data(), MM.data())) return -1;
    for (int n = 0; n <= N; ++n) {
      double res = 0;
      for (int m = 0; m <= N; ++m) res += MM[n * n1 + m];
      w[n] = res;
    }
  }
  return 0;
}

// libs/mesh/meshBasis1D.cpp:114-136 (D = Vr / V).
int Dmatrix1D(int N, int NpointsOut, const double* rIn, const double* rOut,
              double* D) {
  const int Np = N + 1;
  std::vector<double> V(Np * Np), Vr(NpointsOut * Np);
  Vandermonde1D(N, Np, rIn, V.data());
  GradVandermonde1D(N, NpointsOut, rOut, Vr.data());
  return MatrixRightSolve(NpointsOut, Np, Vr.data(), Np, Np, V.data(), D);
}

// libs/mesh/meshBasis1D.cpp:138-159 (I = VOut / VIn).
int InterpolationMatrix1D(int N, int NpointsOut, const double* rIn,
                          const double* rOut, double* I) {
  const int Np = N + 1;
  std::vector<double> VIn(Np * Np), VOut(NpointsOut * Np);
  Vandermonde1D(N, Np, rIn, VIn.data());
  Vandermonde1D(N, NpointsOut, rOut, VOut.data());
  return MatrixRightSolve(NpointsOut, Np, VOut.data(), Np, Np, VIn.data(), I);
}

// ------------------------------------------------------------------ RNG --
// SplitMix64 (Steele, Lea, Flood 2014): the stream seeded `seed` has
// output k = mix(seed + (k+1)*golden). Counter-based, so CPU and GPU can
// draw the same value for the same (seed, counter).
uint64_t rng_u64(uint64_t seed, uint64_t counter) {
  uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double rng_u01(uint64_t seed, uint64_t counter) {
  return (double)(rng_u64(seed, counter) >> 11) * 0x1.0p-53;
}

// Box-Muller on counters 2k, 2k+1.
double rng_normal(uint64_t seed, uint64_t counter) {
  const double u1 = 1.0 - rng_u01(seed, 2 * counter);
  const double u2 = rng_u01(seed, 2 * counter + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

}  // namespace orc

// Reference-element operators from closed forms.
//
// The reference builds these with LAPACK (libs/mesh/meshBasis1D.cpp:
// Golub-Welsch eigen-solve for nodes, D = Vr/V by dgesv, MM = inv(V V^T),
// libs/mesh/meshCubatureSetupHex3D.cpp: cubD, invV). Here every matrix comes
// from a closed form evaluated in double precision, which avoids the
// Vandermonde solves (their conditioning grows with N):
//   GLL nodes     : Newton on P_N'(x) (interior), weights 2/(N(N+1) P_N(x)^2)
//   Gauss nodes   : Newton on P_n(x),  weights 2/((1-x^2) P_n'(x)^2)
//   D (GLL)       : P_N(x_i)/(P_N(x_j)(x_i-x_j)), D_00 = -N(N+1)/4, D_NN = -D_00
//   D (any nodes) : barycentric l_j'(x_i), diagonal = -(row sum)
//   I             : barycentric Lagrange interpolation
//   MM            : exact Gram matrix of the GLL Lagrange basis (Gauss rule)
//   invV          : Legendre coefficients of the Lagrange basis (Gauss rule)
//   modal cubD    : d/dr on orthonormal Legendre coefficients (Gauss rule)
#include "basis.hpp"

#include <cmath>

namespace bpx {
namespace basis {

void legendre(int n, double x, double& P, double& dP) {
  double p0 = 1.0, p1 = x;
  if (n == 0) {
    P = 1.0;
    dP = 0.0;
    return;
  }
  for (int k = 2; k <= n; ++k) {
    const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = pk;
  }
  P = p1;
  // n P_{n-1} = (derivative relation) (1-x^2) P_n' = n (P_{n-1} - x P_n)
  if (std::fabs(1.0 - x * x) > 1e-300)
    dP = n * (p0 - x * p1) / (1.0 - x * x);
  else
    dP = (x > 0 ? 1.0 : (n % 2 ? 1.0 : -1.0)) * 0.5 * n * (n + 1.0);
}

double orthonormal_legendre(int n, double x, double* deriv) {
  double P, dP;
  legendre(n, x, P, dP);
  const double s = std::sqrt((2.0 * n + 1.0) / 2.0);
  if (deriv) *deriv = s * dP;
  return s * P;
}

void gll(int N, std::vector<double>& x, std::vector<double>& w) {
  x.assign(N + 1, 0.0);
  w.assign(N + 1, 0.0);
  x[0] = -1.0;
  x[N] = 1.0;
  for (int i = 1; i < N; ++i) {
    double r = -std::cos(M_PI * i / N);  // Chebyshev-Gauss-Lobatto guess
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre(N, r, P, dP);
      // P'' from the Legendre equation (1-x^2)P'' = 2x P' - N(N+1) P
      const double d2P = (2.0 * r * dP - N * (N + 1.0) * P) / (1.0 - r * r);
      const double dr = dP / d2P;
      r -= dr;
      if (std::fabs(dr) < 1e-16) break;
    }
    x[i] = r;
  }
  for (int i = 0; i <= N; ++i) {
    double P, dP;
    legendre(N, x[i], P, dP);
    w[i] = 2.0 / (N * (N + 1.0) * P * P);
  }
}

void gauss(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double r = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double P = 0, dP = 1;
    for (int it = 0; it < 100; ++it) {
      legendre(n, r, P, dP);
      const double dr = P / dP;
      r -= dr;
      if (std::fabs(dr) < 1e-16) break;
    }
    legendre(n, r, P, dP);
    x[i] = r;
    w[i] = 2.0 / ((1.0 - r * r) * dP * dP);
  }
}

static std::vector<double> bary_weights(const std::vector<double>& z) {
  const int n = (int)z.size();
  std::vector<double> lam(n, 1.0);
  for (int j = 0; j < n; ++j)
    for (int m = 0; m < n; ++m)
      if (m != j) lam[j] /= (z[j] - z[m]);
  return lam;
}

void lagrange_at(const std::vector<double>& z, double x, std::vector<double>& l) {
  const int n = (int)z.size();
  l.assign(n, 0.0);
  const std::vector<double> lam = bary_weights(z);
  for (int j = 0; j < n; ++j)
    if (x == z[j]) {
      l[j] = 1.0;
      return;
    }
  double den = 0;
  for (int j = 0; j < n; ++j) den += lam[j] / (x - z[j]);
  for (int j = 0; j < n; ++j) l[j] = (lam[j] / (x - z[j])) / den;
}

void dmatrix_gll(int N, const std::vector<double>& x, std::vector<double>& D) {
  const int n = N + 1;
  D.assign(n * n, 0.0);
  std::vector<double> P(n);
  for (int i = 0; i < n; ++i) {
    double dP;
    legendre(N, x[i], P[i], dP);
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (i != j) D[i * n + j] = P[i] / (P[j] * (x[i] - x[j]));
  D[0] = -N * (N + 1.0) / 4.0;
  D[N * n + N] = N * (N + 1.0) / 4.0;
}

void dmatrix(const std::vector<double>& z, std::vector<double>& D) {
  const int n = (int)z.size();
  const std::vector<double> lam = bary_weights(z);
  D.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) {
    double diag = 0;
    for (int j = 0; j < n; ++j)
      if (j != i) {
        const double v = (lam[j] / lam[i]) / (z[i] - z[j]);
        D[i * n + j] = v;
        diag -= v;
      }
    D[i * n + i] = diag;
  }
}

void interp(const std::vector<double>& zin, const std::vector<double>& zout, std::vector<double>& I) {
  const int nin = (int)zin.size(), nout = (int)zout.size();
  I.assign(nout * nin, 0.0);
  std::vector<double> l;
  for (int q = 0; q < nout; ++q) {
    lagrange_at(zin, zout[q], l);
    for (int j = 0; j < nin; ++j) I[q * nin + j] = l[j];
  }
}

void mass_gll(int N, const std::vector<double>& x, std::vector<double>& MM) {
  const int n = N + 1;
  std::vector<double> g, gw, l;
  gauss(n, g, gw);  // exact for degree 2N
  MM.assign(n * n, 0.0);
  for (size_t q = 0; q < g.size(); ++q) {
    lagrange_at(x, g[q], l);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) MM[i * n + j] += gw[q] * l[i] * l[j];
  }
}

void inv_vandermonde(int N, const std::vector<double>& x, std::vector<double>& invV) {
  const int n = N + 1;
  std::vector<double> g, gw, l;
  gauss(n, g, gw);
  invV.assign(n * n, 0.0);
  for (size_t q = 0; q < g.size(); ++q) {
    lagrange_at(x, g[q], l);
    for (int j = 0; j < n; ++j) {
      const double Pj = orthonormal_legendre(j, g[q], nullptr);
      for (int i = 0; i < n; ++i) invV[j * n + i] += gw[q] * Pj * l[i];
    }
  }
}

void modal_dmatrix(int N, std::vector<double>& Dm) {
  const int n = N + 1;
  std::vector<double> g, gw;
  gauss(n, g, gw);
  Dm.assign(n * n, 0.0);
  for (size_t q = 0; q < g.size(); ++q)
    for (int a = 0; a < n; ++a) {
      const double Pa = orthonormal_legendre(a, g[q], nullptr);
      for (int m = 0; m < n; ++m) {
        double dPm;
        orthonormal_legendre(m, g[q], &dPm);
        Dm[a * n + m] += gw[q] * Pa * dPm;
      }
    }
}

}  // namespace basis
}  // namespace bpx

// Reference-element operators of the triangle and the tetrahedron.
//
// Reference behaviour (libs/mesh/meshBasisTri2D.cpp, meshBasisTet3D.cpp,
// meshReferenceNodesTri2D.cpp, meshReferenceNodesTet3D.cpp):
//  - nodes: the equispaced lattice (r index fastest, then s, then t) moved by
//    Warburton's warp & blend map (T. Warburton, "An explicit construction of
//    interpolation nodes on the simplex", J. Eng. Math. 56 (2006) 247-262)
//    with the blend parameters alpha optimised per degree given there;
//  - modes: the orthonormal Koornwinder-Dubiner polynomials
//    (Hesthaven & Warburton, Nodal Discontinuous Galerkin Methods, 2008,
//    sections 6.1 and 10.1), ordered i (the collapsed r index) outermost;
//  - MM = (V V^T)^{-1}, D_a = V_a V^{-1}, and the packed stiffness matrices
//    S_ab = D_a^T MM D_b (+ D_b^T MM D_a for a != b): rr, rs, ss in 2D and
//    rr, rs, rt, ss, st, tt in 3D.
// The inverses come from one Gauss-Jordan elimination of V with partial
// pivoting (MM = V^{-T} V^{-1}), where the reference calls LAPACK.
#include <cmath>
#include <stdexcept>

#include "basis.hpp"

namespace bpx {
namespace simplex {

namespace {

// orthonormal Jacobi polynomial of degree n for the weight (1-x)^a (1+x)^b
double jacobi(double x, double a, double b, int n) {
  const double g0 = std::pow(2.0, a + b + 1.0) / (a + b + 1.0) * std::tgamma(a + 1.0) * std::tgamma(b + 1.0) /
                    std::tgamma(a + b + 1.0);
  double p0 = 1.0 / std::sqrt(g0);
  if (n == 0) return p0;
  const double g1 = (a + 1.0) * (b + 1.0) / (a + b + 3.0) * g0;
  double p1 = ((a + b + 2.0) * x / 2.0 + (a - b) / 2.0) / std::sqrt(g1);
  double aold = 2.0 / (2.0 + a + b) * std::sqrt((a + 1.0) * (b + 1.0) / (a + b + 3.0));
  for (int i = 1; i < n; ++i) {
    const double h1 = 2.0 * i + a + b;
    const double anew = 2.0 / (h1 + 2.0) *
                        std::sqrt((i + 1.0) * (i + 1.0 + a + b) * (i + 1.0 + a) * (i + 1.0 + b) / (h1 + 1.0) /
                                  (h1 + 3.0));
    const double bnew = -(a * a - b * b) / h1 / (h1 + 2.0);
    const double p2 = (-aold * p0 + (x - bnew) * p1) / anew;
    p0 = p1;
    p1 = p2;
    aold = anew;
  }
  return p1;
}

double grad_jacobi(double x, double a, double b, int n) {
  return n == 0 ? 0.0 : std::sqrt(n * (n + a + b + 1.0)) * jacobi(x, a + 1.0, b + 1.0, n - 1);
}

// the warp of the equispaced 1D lattice onto the GLL nodes, evaluated at r
// and divided by 1 - r^2 (the blend of the edge warp into the interior)
double warp_factor(int N, const std::vector<double>& gll, double r) {
  double w = 0;
  for (int i = 0; i <= N; ++i) {
    const double ri = -1.0 + 2.0 * i / N;
    double l = 1.0;  // equispaced Lagrange polynomial i at r
    for (int j = 0; j <= N; ++j)
      if (j != i) l *= (r - (-1.0 + 2.0 * j / N)) / (ri - (-1.0 + 2.0 * j / N));
    w += l * (gll[i] - ri);
  }
  // zero at the ends of the edge, where the blend vanishes
  return std::fabs(r) < 1.0 - 1e-10 ? w / (1.0 - r * r) : 0.0;
}

const double kAlphaTri[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                              1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
const double kAlphaTet[15] = {0.0000, 0.0000, 0.00000, 0.1002, 1.1332, 1.5608, 1.3413, 1.2577,
                              1.1603, 1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655};

// tangential warp of a face with barycentric coordinates (L1, L2, L3), in
// the face's equilateral frame
void face_warp(int N, double alpha, const std::vector<double>& gll, double L1, double L2, double L3, double& w1,
               double& w2) {
  const double f1 = warp_factor(N, gll, L3 - L2), f2 = warp_factor(N, gll, L1 - L3),
               f3 = warp_factor(N, gll, L2 - L1);
  const double a1 = 4.0 * L2 * L3 * f1 * (1.0 + alpha * alpha * L1 * L1);
  const double a2 = 4.0 * L3 * L1 * f2 * (1.0 + alpha * alpha * L2 * L2);
  const double a3 = 4.0 * L1 * L2 * f3 * (1.0 + alpha * alpha * L3 * L3);
  w1 = a1 + std::cos(2.0 * M_PI / 3.0) * a2 + std::cos(4.0 * M_PI / 3.0) * a3;
  w2 = std::sin(2.0 * M_PI / 3.0) * a2 + std::sin(4.0 * M_PI / 3.0) * a3;
}

// the collapsed coordinates of a point of the triangle / tetrahedron
void collapse_tri(double r, double s, double& a, double& b) {
  a = (s != 1.0) ? 2.0 * (1.0 + r) / (1.0 - s) - 1.0 : -1.0;
  b = s;
}
void collapse_tet(double r, double s, double t, double& a, double& b, double& c) {
  a = (std::fabs(s + t) > 1e-8) ? 2.0 * (1.0 + r) / (-s - t) - 1.0 : -1.0;
  b = (std::fabs(t - 1.0) > 1e-8) ? 2.0 * (1.0 + s) / (1.0 - t) - 1.0 : -1.0;
  c = t;
}

}  // namespace

void nodes_tri(int N, std::vector<double>& r, std::vector<double>& s, std::vector<int>& idx) {
  const int Np = (N + 1) * (N + 2) / 2;
  r.assign(Np, 0.0);
  s.assign(Np, 0.0);
  idx.assign(2 * Np, 0);
  std::vector<double> gll, w;
  basis::gll(N, gll, w);
  const double alpha = N < 16 ? kAlphaTri[N - 1] : 5.0 / 3.0;
  int n = 0;
  for (int j = 0; j <= N; ++j)
    for (int i = 0; i <= N - j; ++i, ++n) {
      idx[2 * n + 0] = i;
      idx[2 * n + 1] = j;
      const double re = -1.0 + 2.0 * i / N, se = -1.0 + 2.0 * j / N;
      // barycentric coordinates and the equilateral triangle
      const double L1 = 0.5 * (1.0 + se), L2 = -0.5 * (re + se), L3 = 0.5 * (1.0 + re);
      double x = -L2 + L3, y = (-L2 - L3 + 2.0 * L1) / std::sqrt(3.0);
      double w1, w2;
      face_warp(N, alpha, gll, L1, L2, L3, w1, w2);
      x += w1;
      y += w2;
      // back to (r, s)
      const double B1 = (std::sqrt(3.0) * y + 1.0) / 3.0;
      const double B2 = (-3.0 * x - std::sqrt(3.0) * y + 2.0) / 6.0;
      const double B3 = (3.0 * x - std::sqrt(3.0) * y + 2.0) / 6.0;
      r[n] = -B2 + B3 - B1;
      s[n] = -B2 - B3 + B1;
    }
}

void nodes_tet(int N, std::vector<double>& r, std::vector<double>& s, std::vector<double>& t, std::vector<int>& idx) {
  const int Np = (N + 1) * (N + 2) * (N + 3) / 6;
  r.assign(Np, 0.0);
  s.assign(Np, 0.0);
  t.assign(Np, 0.0);
  idx.assign(3 * Np, 0);
  std::vector<double> gll, w;
  basis::gll(N, gll, w);
  const double alpha = N < 16 ? kAlphaTet[N - 1] : 1.0;
  // the equilateral tetrahedron and the two unit tangents of each face
  const double v[4][3] = {{-1.0, -1.0 / std::sqrt(3.0), -1.0 / std::sqrt(6.0)},
                          {1.0, -1.0 / std::sqrt(3.0), -1.0 / std::sqrt(6.0)},
                          {0.0, 2.0 / std::sqrt(3.0), -1.0 / std::sqrt(6.0)},
                          {0.0, 0.0, 3.0 / std::sqrt(6.0)}};
  double t1[4][3], t2[4][3];
  for (int d = 0; d < 3; ++d) {
    t1[0][d] = v[1][d] - v[0][d];
    t1[1][d] = v[1][d] - v[0][d];
    t1[2][d] = v[2][d] - v[1][d];
    t1[3][d] = v[2][d] - v[0][d];
    t2[0][d] = v[2][d] - 0.5 * (v[0][d] + v[1][d]);
    t2[1][d] = v[3][d] - 0.5 * (v[0][d] + v[1][d]);
    t2[2][d] = v[3][d] - 0.5 * (v[1][d] + v[2][d]);
    t2[3][d] = v[3][d] - 0.5 * (v[0][d] + v[2][d]);
  }
  for (int f = 0; f < 4; ++f) {
    const double n1 = std::sqrt(t1[f][0] * t1[f][0] + t1[f][1] * t1[f][1] + t1[f][2] * t1[f][2]);
    const double n2 = std::sqrt(t2[f][0] * t2[f][0] + t2[f][1] * t2[f][1] + t2[f][2] * t2[f][2]);
    for (int d = 0; d < 3; ++d) {
      t1[f][d] /= n1;
      t2[f][d] /= n2;
    }
  }
  // (r, s, t) from a point of the equilateral tetrahedron: solve
  // 0.5 [v1-v0, v2-v0, v3-v0] rst = xyz - 0.5 (v1 + v2 + v3 - v0)
  double A[3][3];
  for (int d = 0; d < 3; ++d) {
    A[d][0] = 0.5 * (v[1][d] - v[0][d]);
    A[d][1] = 0.5 * (v[2][d] - v[0][d]);
    A[d][2] = 0.5 * (v[3][d] - v[0][d]);
  }
  const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                     A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                     A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
  int n = 0;
  for (int k = 0; k <= N; ++k)
    for (int j = 0; j <= N - k; ++j)
      for (int i = 0; i <= N - j - k; ++i, ++n) {
        idx[3 * n + 0] = i;
        idx[3 * n + 1] = j;
        idx[3 * n + 2] = k;
        const double re = -1.0 + 2.0 * i / N, se = -1.0 + 2.0 * j / N, te = -1.0 + 2.0 * k / N;
        const double L[4] = {-0.5 * (1.0 + re + se + te), 0.5 * (1.0 + re), 0.5 * (1.0 + se), 0.5 * (1.0 + te)};
        // L[0..3] weigh v0..v3; the faces opposite v3, v2, v0, v1 in turn
        double X[3];
        for (int d = 0; d < 3; ++d) X[d] = L[0] * v[0][d] + L[1] * v[1][d] + L[2] * v[2][d] + L[3] * v[3][d];
        double shift[3] = {0, 0, 0};
        // (La: the vertex off the face; Lb, Lc, Ld: the face's own)
        const int face[4][4] = {{3, 2, 0, 1}, {2, 3, 0, 1}, {0, 3, 1, 2}, {1, 3, 0, 2}};
        for (int f = 0; f < 4; ++f) {
          const double La = L[face[f][0]], Lb = L[face[f][1]], Lc = L[face[f][2]], Ld = L[face[f][3]];
          double w1, w2;
          face_warp(N, alpha, gll, Lb, Lc, Ld, w1, w2);
          double blend = Lb * Lc * Ld;
          const double denom = (Lb + 0.5 * La) * (Lc + 0.5 * La) * (Ld + 0.5 * La);
          if (denom > 1e-10) blend = (1.0 + (alpha * La) * (alpha * La)) * blend / denom;
          for (int d = 0; d < 3; ++d) shift[d] += blend * w1 * t1[f][d] + blend * w2 * t2[f][d];
          // on the face's boundary edges only the face's own warp counts
          if (La < 1e-10 && (Lb < 1e-10 || Lc < 1e-10 || Ld < 1e-10))
            for (int d = 0; d < 3; ++d) shift[d] = w1 * t1[f][d] + w2 * t2[f][d];
        }
        double b[3];
        for (int d = 0; d < 3; ++d) b[d] = X[d] + shift[d] - 0.5 * (v[1][d] + v[2][d] + v[3][d] - v[0][d]);
        // Cramer's rule
        double rst[3];
        for (int c = 0; c < 3; ++c) {
          double M[3][3];
          for (int rr = 0; rr < 3; ++rr)
            for (int cc = 0; cc < 3; ++cc) M[rr][cc] = (cc == c) ? b[rr] : A[rr][cc];
          rst[c] = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                    M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) /
                   det;
        }
        r[n] = rst[0];
        s[n] = rst[1];
        t[n] = rst[2];
      }
}

void vandermonde_tri(int N, const std::vector<double>& r, const std::vector<double>& s, std::vector<double>& V,
                     std::vector<double>& Vr, std::vector<double>& Vs) {
  const int Np = (N + 1) * (N + 2) / 2, P = (int)r.size();
  V.assign((size_t)P * Np, 0.0);
  Vr.assign((size_t)P * Np, 0.0);
  Vs.assign((size_t)P * Np, 0.0);
  for (int n = 0; n < P; ++n) {
    double a, b;
    collapse_tri(r[n], s[n], a, b);
    int m = 0;
    for (int i = 0; i <= N; ++i)
      for (int j = 0; j <= N - i; ++j, ++m) {
        const double fa = jacobi(a, 0, 0, i), dfa = grad_jacobi(a, 0, 0, i);
        const double gb = jacobi(b, 2 * i + 1, 0, j), dgb = grad_jacobi(b, 2 * i + 1, 0, j);
        const double h = 0.5 * (1.0 - b);
        V[(size_t)n * Np + m] = std::sqrt(2.0) * fa * gb * std::pow(1.0 - b, i);
        // d/dr = 2/(1-b) d/da; d/ds = (1+a)/(1-b) d/da + d/db
        double pr = dfa * gb * (i > 0 ? std::pow(h, i - 1) : 1.0);
        double ps = dfa * gb * 0.5 * (1.0 + a) * (i > 0 ? std::pow(h, i - 1) : 1.0);
        double tmp = dgb * std::pow(h, i);
        if (i > 0) tmp -= 0.5 * i * gb * std::pow(h, i - 1);
        ps += fa * tmp;
        const double sc = std::pow(2.0, i + 0.5);
        Vr[(size_t)n * Np + m] = pr * sc;
        Vs[(size_t)n * Np + m] = ps * sc;
      }
  }
}

void vandermonde_tet(int N, const std::vector<double>& r, const std::vector<double>& s, const std::vector<double>& t,
                     std::vector<double>& V, std::vector<double>& Vr, std::vector<double>& Vs,
                     std::vector<double>& Vt) {
  const int Np = (N + 1) * (N + 2) * (N + 3) / 6, P = (int)r.size();
  V.assign((size_t)P * Np, 0.0);
  Vr.assign((size_t)P * Np, 0.0);
  Vs.assign((size_t)P * Np, 0.0);
  Vt.assign((size_t)P * Np, 0.0);
  for (int n = 0; n < P; ++n) {
    double a, b, c;
    collapse_tet(r[n], s[n], t[n], a, b, c);
    int m = 0;
    for (int i = 0; i <= N; ++i)
      for (int j = 0; j <= N - i; ++j)
        for (int k = 0; k <= N - i - j; ++k, ++m) {
          const double p1 = jacobi(a, 0, 0, i), p2 = jacobi(b, 2 * i + 1, 0, j),
                       p3 = jacobi(c, 2 * (i + j) + 2, 0, k);
          const double d1 = grad_jacobi(a, 0, 0, i), d2 = grad_jacobi(b, 2 * i + 1, 0, j),
                       d3 = grad_jacobi(c, 2 * (i + j) + 2, 0, k);
          const double hb = 0.5 * (1.0 - b), hc = 0.5 * (1.0 - c);
          V[(size_t)n * Np + m] = 2.0 * std::sqrt(2.0) * p1 * p2 * p3 * std::pow(1.0 - b, i) * std::pow(1.0 - c, i + j);
          double pr = d1 * p2 * p3;
          if (i > 0) pr *= std::pow(hb, i - 1);
          if (i + j > 0) pr *= std::pow(hc, i + j - 1);
          double ps = 0.5 * (1.0 + a) * pr;
          double tmp = d2 * std::pow(hb, i);
          if (i > 0) tmp += -0.5 * i * p2 * std::pow(hb, i - 1);
          if (i + j > 0) tmp *= std::pow(hc, i + j - 1);
          tmp *= p1 * p3;
          ps += tmp;
          double pt = 0.5 * (1.0 + a) * pr + 0.5 * (1.0 + b) * tmp;
          double tmp2 = d3 * std::pow(hc, i + j);
          if (i + j > 0) tmp2 -= 0.5 * (i + j) * p3 * std::pow(hc, i + j - 1);
          tmp2 *= p1 * p2 * std::pow(hb, i);
          pt += tmp2;
          const double sc = std::pow(2.0, 2 * i + j + 1.5);
          Vr[(size_t)n * Np + m] = pr * sc;
          Vs[(size_t)n * Np + m] = ps * sc;
          Vt[(size_t)n * Np + m] = pt * sc;
        }
  }
}

// inverse of the n x n row-major A (Gauss-Jordan, partial pivoting)
void invert(int n, std::vector<double>& A) {
  std::vector<double> I((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) I[(size_t)i * n + i] = 1.0;
  for (int c = 0; c < n; ++c) {
    int p = c;
    for (int rr = c + 1; rr < n; ++rr)
      if (std::fabs(A[(size_t)rr * n + c]) > std::fabs(A[(size_t)p * n + c])) p = rr;
    if (A[(size_t)p * n + c] == 0.0) throw std::runtime_error("singular Vandermonde matrix");
    if (p != c)
      for (int k = 0; k < n; ++k) {
        std::swap(A[(size_t)p * n + k], A[(size_t)c * n + k]);
        std::swap(I[(size_t)p * n + k], I[(size_t)c * n + k]);
      }
    const double d = 1.0 / A[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      A[(size_t)c * n + k] *= d;
      I[(size_t)c * n + k] *= d;
    }
#pragma omp parallel for
    for (int rr = 0; rr < n; ++rr) {
      if (rr == c) continue;
      const double f = A[(size_t)rr * n + c];
      if (f == 0.0) continue;
      for (int k = 0; k < n; ++k) {
        A[(size_t)rr * n + k] -= f * A[(size_t)c * n + k];
        I[(size_t)rr * n + k] -= f * I[(size_t)c * n + k];
      }
    }
  }
  A.swap(I);
}

void operators(int Np, int dim, const std::vector<double>& V, const std::vector<const std::vector<double>*>& Vg,
               std::vector<double>& MM, std::vector<double>& D, std::vector<double>& S) {
  std::vector<double> Vi(V);
  invert(Np, Vi);
  const size_t NN = (size_t)Np * Np;
  // MM = V^{-T} V^{-1}
  MM.assign(NN, 0.0);
#pragma omp parallel for
  for (int a = 0; a < Np; ++a)
    for (int b = 0; b < Np; ++b) {
      double acc = 0;
      for (int k = 0; k < Np; ++k) acc += Vi[(size_t)k * Np + a] * Vi[(size_t)k * Np + b];
      MM[(size_t)a * Np + b] = acc;
    }
  // D_d = V_d V^{-1}
  D.assign(dim * NN, 0.0);
  for (int d = 0; d < dim; ++d) {
    const std::vector<double>& G = *Vg[d];
#pragma omp parallel for
    for (int a = 0; a < Np; ++a)
      for (int b = 0; b < Np; ++b) {
        double acc = 0;
        for (int k = 0; k < Np; ++k) acc += G[(size_t)a * Np + k] * Vi[(size_t)k * Np + b];
        D[d * NN + (size_t)a * Np + b] = acc;
      }
  }
  // MD_d = MM D_d, then S_ab = D_a^T (MM D_b) (+ its transpose for a != b)
  std::vector<double> MD(dim * NN, 0.0);
  for (int d = 0; d < dim; ++d)
#pragma omp parallel for
    for (int a = 0; a < Np; ++a)
      for (int b = 0; b < Np; ++b) {
        double acc = 0;
        for (int k = 0; k < Np; ++k) acc += MM[(size_t)a * Np + k] * D[d * NN + (size_t)k * Np + b];
        MD[d * NN + (size_t)a * Np + b] = acc;
      }
  const int nS = dim == 2 ? 3 : 6;
  S.assign(nS * NN, 0.0);
  int sIdx = 0;
  for (int p = 0; p < dim; ++p)
    for (int q = p; q < dim; ++q, ++sIdx)
#pragma omp parallel for
      for (int a = 0; a < Np; ++a)
        for (int b = 0; b < Np; ++b) {
          double acc = 0;
          for (int k = 0; k < Np; ++k) {
            acc += D[p * NN + (size_t)k * Np + a] * MD[q * NN + (size_t)k * Np + b];
            if (p != q) acc += D[q * NN + (size_t)k * Np + a] * MD[p * NN + (size_t)k * Np + b];
          }
          S[sIdx * NN + (size_t)a * Np + b] = acc;
        }
}

}  // namespace simplex
}  // namespace bpx

// ORACLE — test infrastructure only (see oracle.hpp).
//
// Restatement of the reference's quadrilateral, triangular and tetrahedral
// box problems: reference nodes and operators, box meshes, connectivity,
// node numbering, masked ids, geometric (and for quads cubature) factors,
// the Ax loop nests and the forcing. Each function cites the reference
// file:line it follows.
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <numeric>

#include "oracle.hpp"

namespace orc {

namespace {

// --------------------------------------------------- 2D / 3D simplices --

// libs/mesh/meshBasisTri2D.cpp:465-510 (Warpfactor): the equispaced-to-GLL
// displacement interpolated with the 1D Vandermonde, scaled by 1/(1-r^2)
void Warpfactor(int N, int Npoints, const double* r, double* warp) {
  std::vector<double> gll(N + 1), w(N + 1), req(N + 1), I((size_t)Npoints * (N + 1));
  JacobiGLL(N, gll.data(), w.data());
  for (int i = 0; i <= N; ++i) req[i] = -1.0 + 2.0 * i / N;  // meshBasis1D.cpp EquispacedNodes1D
  InterpolationMatrix1D(N, Npoints, req.data(), r, I.data());
  for (int n = 0; n < Npoints; n++) {
    warp[n] = 0.0;
    for (int i = 0; i < N + 1; i++) warp[n] += I[(size_t)n * (N + 1) + i] * (gll[i] - req[i]);
    const double zerof = (std::abs(r[n]) < 1.0 - 1.0e-10) ? 1 : 0;
    const double sf = 1.0 - (zerof * r[n]) * (zerof * r[n]);
    warp[n] = warp[n] / sf + warp[n] * (zerof - 1);
  }
}

// libs/mesh/meshBasisTri2D.cpp:200-217 (EquispacedNodesTri2D), :512-560
// (WarpBlendTransformTri2D) and the xytors helper before it
void NodesTri2D(int N, std::vector<double>& r, std::vector<double>& s) {
  const int Np = (N + 1) * (N + 2) / 2;
  r.resize(Np);
  s.resize(Np);
  int sk = 0;
  for (int n = 0; n < N + 1; n++)
    for (int m = 0; m < N + 1 - n; m++) {
      r[sk] = -1.0 + 2.0 * m / N;
      s[sk] = -1.0 + 2.0 * n / N;
      sk++;
    }
  const double alpopt[15] = {0.0000, 0.0000, 1.4152, 0.1001, 0.2751, 0.9800, 1.0999, 1.2832,
                             1.3648, 1.4773, 1.4959, 1.5743, 1.5770, 1.6223, 1.6258};
  const double alpha = N < 16 ? alpopt[N - 1] : 5. / 3.;
  std::vector<double> L1(Np), L2(Np), L3(Np), d32(Np), d13(Np), d21(Np), x(Np), y(Np), w1(Np), w2(Np), w3(Np);
  for (int n = 0; n < Np; n++) {
    L1[n] = 0.5 * (1. + s[n]);
    L2[n] = -0.5 * (r[n] + s[n]);
    L3[n] = 0.5 * (1. + r[n]);
    d32[n] = L3[n] - L2[n];
    d13[n] = L1[n] - L3[n];
    d21[n] = L2[n] - L1[n];
    x[n] = -L2[n] + L3[n];
    y[n] = (-L2[n] - L3[n] + 2. * L1[n]) / sqrt(3.0);
  }
  Warpfactor(N, Np, d32.data(), w1.data());
  Warpfactor(N, Np, d13.data(), w2.data());
  Warpfactor(N, Np, d21.data(), w3.data());
  for (int n = 0; n < Np; n++) {
    const double b1 = 4.0 * L2[n] * L3[n], b2 = 4.0 * L3[n] * L1[n], b3 = 4.0 * L1[n] * L2[n];
    const double warp1 = b1 * w1[n] * (1.0 + alpha * alpha * L1[n] * L1[n]);
    const double warp2 = b2 * w2[n] * (1.0 + alpha * alpha * L2[n] * L2[n]);
    const double warp3 = b3 * w3[n] * (1.0 + alpha * alpha * L3[n] * L3[n]);
    x[n] += 1. * warp1 + cos(2. * M_PI / 3.) * warp2 + cos(4. * M_PI / 3.) * warp3;
    y[n] += 0. * warp1 + sin(2. * M_PI / 3.) * warp2 + sin(4. * M_PI / 3.) * warp3;
  }
  for (int n = 0; n < Np; n++) {
    const double A1 = (sqrt(3.0) * y[n] + 1.0) / 3.0;
    const double A2 = (-3.0 * x[n] - sqrt(3.0) * y[n] + 2.0) / 6.0;
    const double A3 = (3.0 * x[n] - sqrt(3.0) * y[n] + 2.0) / 6.0;
    r[n] = -A2 + A3 - A1;
    s[n] = -A2 - A3 + A1;
  }
}

// libs/mesh/meshBasisTet3D.cpp:604-645 (WarpShiftFace3D)
void WarpShiftFace3D(int N, double alpha, int Np, const double* L1, const double* L2, const double* L3, double* w1,
                     double* w2) {
  std::vector<double> d32(Np), d13(Np), d21(Np), f1(Np), f2(Np), f3(Np);
  for (int n = 0; n < Np; n++) {
    d32[n] = L3[n] - L2[n];
    d13[n] = L1[n] - L3[n];
    d21[n] = L2[n] - L1[n];
  }
  Warpfactor(N, Np, d32.data(), f1.data());
  Warpfactor(N, Np, d13.data(), f2.data());
  Warpfactor(N, Np, d21.data(), f3.data());
  for (int n = 0; n < Np; n++) {
    const double b1 = 4.0 * L2[n] * L3[n], b2 = 4.0 * L3[n] * L1[n], b3 = 4.0 * L1[n] * L2[n];
    const double warp1 = b1 * f1[n] * (1.0 + alpha * alpha * L1[n] * L1[n]);
    const double warp2 = b2 * f2[n] * (1.0 + alpha * alpha * L2[n] * L2[n]);
    const double warp3 = b3 * f3[n] * (1.0 + alpha * alpha * L3[n] * L3[n]);
    w1[n] = 1. * warp1 + cos(2. * M_PI / 3.) * warp2 + cos(4. * M_PI / 3.) * warp3;
    w2[n] = 0. * warp1 + sin(2. * M_PI / 3.) * warp2 + sin(4. * M_PI / 3.) * warp3;
  }
}

// libs/mesh/meshBasisTet3D.cpp:227-248 (EquispacedNodesTet3D), :647-760
// (WarpBlendTransformTet3D) and :564-602 (xyztorst)
void NodesTet3D(int N, std::vector<double>& r, std::vector<double>& s, std::vector<double>& t) {
  const int Np = (N + 1) * (N + 2) * (N + 3) / 6;
  r.resize(Np);
  s.resize(Np);
  t.resize(Np);
  int sk = 0;
  for (int k = 0; k < N + 1; k++)
    for (int n = 0; n < N + 1 - k; n++)
      for (int m = 0; m < N + 1 - n - k; m++) {
        r[sk] = -1.0 + 2.0 * m / N;
        s[sk] = -1.0 + 2.0 * n / N;
        t[sk] = -1.0 + 2.0 * k / N;
        sk++;
      }
  const double alpopt[15] = {0.0000, 0.0000, 0.00000, 0.1002, 1.1332, 1.5608, 1.3413, 1.2577,
                             1.1603, 1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655};
  const double alpha = N < 16 ? alpopt[N - 1] : 1.;
  const double v1[3] = {-1.0, -1. / sqrt(3.), -1. / sqrt(6.)}, v2[3] = {1.0, -1. / sqrt(3.), -1. / sqrt(6.)},
               v3[3] = {0.0, 2. / sqrt(3.), -1. / sqrt(6.)}, v4[3] = {0.0, 0., 3. / sqrt(6.)};
  double t1[4][3], t2[4][3];
  for (int v = 0; v < 3; v++) {
    t1[0][v] = v2[v] - v1[v];
    t1[1][v] = v2[v] - v1[v];
    t1[2][v] = v3[v] - v2[v];
    t1[3][v] = v3[v] - v1[v];
    t2[0][v] = v3[v] - 0.5 * (v1[v] + v2[v]);
    t2[1][v] = v4[v] - 0.5 * (v1[v] + v2[v]);
    t2[2][v] = v4[v] - 0.5 * (v2[v] + v3[v]);
    t2[3][v] = v4[v] - 0.5 * (v1[v] + v3[v]);
  }
  for (int v = 0; v < 4; v++) {
    const double n1 = sqrt(t1[v][0] * t1[v][0] + t1[v][1] * t1[v][1] + t1[v][2] * t1[v][2]);
    const double n2 = sqrt(t2[v][0] * t2[v][0] + t2[v][1] * t2[v][1] + t2[v][2] * t2[v][2]);
    for (int i = 0; i < 3; i++) {
      t1[v][i] /= n1;
      t2[v][i] /= n2;
    }
  }
  std::vector<double> L1(Np), L2(Np), L3(Np), L4(Np), x(Np), y(Np), z(Np), sx(Np, 0.0), sy(Np, 0.0), sz(Np, 0.0);
  for (int n = 0; n < Np; n++) {
    L1[n] = 0.5 * (1. + t[n]);
    L2[n] = 0.5 * (1. + s[n]);
    L3[n] = -0.5 * (1.0 + r[n] + s[n] + t[n]);
    L4[n] = 0.5 * (1. + r[n]);
    x[n] = L3[n] * v1[0] + L4[n] * v2[0] + L2[n] * v3[0] + L1[n] * v4[0];
    y[n] = L3[n] * v1[1] + L4[n] * v2[1] + L2[n] * v3[1] + L1[n] * v4[1];
    z[n] = L3[n] * v1[2] + L4[n] * v2[2] + L2[n] * v3[2] + L1[n] * v4[2];
  }
  std::vector<double> w1(Np), w2(Np);
  for (int f = 0; f < 4; f++) {
    const double *La, *Lb, *Lc, *Ld;
    if (f == 0) { La = L1.data(); Lb = L2.data(); Lc = L3.data(); Ld = L4.data(); }
    if (f == 1) { La = L2.data(); Lb = L1.data(); Lc = L3.data(); Ld = L4.data(); }
    if (f == 2) { La = L3.data(); Lb = L1.data(); Lc = L4.data(); Ld = L2.data(); }
    if (f == 3) { La = L4.data(); Lb = L1.data(); Lc = L3.data(); Ld = L2.data(); }
    WarpShiftFace3D(N, alpha, Np, Lb, Lc, Ld, w1.data(), w2.data());
    for (int n = 0; n < Np; n++) {
      double blend = Lb[n] * Lc[n] * Ld[n];
      const double denom = (Lb[n] + .5 * La[n]) * (Lc[n] + .5 * La[n]) * (Ld[n] + .5 * La[n]);
      if (denom > 1.e-10) blend = (1 + (alpha * La[n]) * (alpha * La[n])) * blend / denom;
      sx[n] += (blend * w1[n]) * t1[f][0] + (blend * w2[n]) * t2[f][0];
      sy[n] += (blend * w1[n]) * t1[f][1] + (blend * w2[n]) * t2[f][1];
      sz[n] += (blend * w1[n]) * t1[f][2] + (blend * w2[n]) * t2[f][2];
      if ((La[n] < 1.e-10) && ((Lb[n] < 1.e-10) || (Lc[n] < 1.e-10) || (Ld[n] < 1.e-10))) {
        sx[n] = w1[n] * t1[f][0] + w2[n] * t2[f][0];
        sy[n] = w1[n] * t1[f][1] + w2[n] * t2[f][1];
        sz[n] = w1[n] * t1[f][2] + w2[n] * t2[f][2];
      }
    }
  }
  // xyztorst: solve A rst = xyz - 0.5 (v2 + v3 + v4 - v1), A = 0.5 [v2-v1, v3-v1, v4-v1]
  double A[9];
  for (int i = 0; i < 3; i++) {
    A[0 * 3 + i] = 0.5 * (v2[i] - v1[i]);
    A[1 * 3 + i] = 0.5 * (v3[i] - v1[i]);
    A[2 * 3 + i] = 0.5 * (v4[i] - v1[i]);
  }
  // A is stored by columns (A[c*3+i] = column c, row i): invert its transpose
  double M[9];
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) M[i * 3 + c] = A[c * 3 + i];
  MatrixInverse(3, M);
  for (int n = 0; n < Np; n++) {
    const double X[3] = {x[n] + sx[n] - 0.5 * (v2[0] + v3[0] + v4[0] - v1[0]),
                         y[n] + sy[n] - 0.5 * (v2[1] + v3[1] + v4[1] - v1[1]),
                         z[n] + sz[n] - 0.5 * (v2[2] + v3[2] + v4[2] - v1[2])};
    r[n] = M[0] * X[0] + M[1] * X[1] + M[2] * X[2];
    s[n] = M[3] * X[0] + M[4] * X[1] + M[5] * X[2];
    t[n] = M[6] * X[0] + M[7] * X[1] + M[8] * X[2];
  }
}

// libs/mesh/meshBasisTri2D.cpp:253-303 (OrthonormalBasisTri2D, GradOrthonormalBasisTri2D)
void BasisTri2D(double r, double s, int i, int j, double& P, double& Pr, double& Ps) {
  const double a = (s != 1.) ? 2. * (1. + r) / (1. - s) - 1. : -1.;
  const double b = s;
  P = sqrt(2.0) * JacobiP(a, 0, 0, i) * JacobiP(b, 2 * i + 1, 0, j) * pow(1. - b, i);
  const double fa = JacobiP(a, 0, 0, i), dfa = GradJacobiP(a, 0, 0, i);
  const double gb = JacobiP(b, 2 * i + 1, 0, j), dgb = GradJacobiP(b, 2 * i + 1, 0, j);
  Pr = dfa * gb;
  if (i > 0) Pr = Pr * pow(0.5 * (1 - b), i - 1);
  Ps = dfa * (gb * (0.5 * (1 + a)));
  if (i > 0) Ps = Ps * pow(0.5 * (1 - b), i - 1);
  double tmp = dgb * pow(0.5 * (1 - b), i);
  if (i > 0) tmp = tmp - 0.5 * i * gb * pow(0.5 * (1 - b), i - 1);
  Ps = Ps + fa * tmp;
  Pr *= pow(2, i + 0.5);
  Ps *= pow(2, i + 0.5);
}

// libs/mesh/meshBasisTet3D.cpp:318-396 (OrthonormalBasisTet3D, GradOrthonormalBasisTet3D)
void BasisTet3D(double r, double s, double t, int i, int j, int k, double& P, double& Pr, double& Ps, double& Pt) {
  const double a = (fabs(s + t) > 1e-8) ? 2.0 * (1. + r) / (-s - t) - 1.0 : -1.0;
  const double b = (fabs(t - 1) > 1e-8) ? 2.0 * (1 + s) / (1. - t) - 1.0 : -1.;
  const double c = t;
  const double p1 = JacobiP(a, 0, 0, i), p2 = JacobiP(b, 2 * i + 1, 0, j), p3 = JacobiP(c, 2 * (i + j) + 2, 0, k);
  P = 2. * sqrt(2.0) * p1 * p2 * p3 * pow(1.0 - b, i) * pow(1.0 - c, i + j);
  const double p1a = GradJacobiP(a, 0, 0, i), p2b = GradJacobiP(b, 2 * i + 1, 0, j),
               p3c = GradJacobiP(c, 2 * (i + j) + 2, 0, k);
  Pr = p1a * p2 * p3;
  if (i > 0) Pr *= pow(0.5 * (1.0 - b), i - 1);
  if (i + j > 0) Pr *= pow(0.5 * (1.0 - c), i + j - 1);
  Ps = 0.5 * (1.0 + a) * (Pr);
  double tmp = p2b * pow(0.5 * (1.0 - b), i);
  if (i > 0) tmp += -0.5 * i * p2 * pow(0.5 * (1.0 - b), i - 1);
  if (i + j > 0) tmp *= pow(0.5 * (1.0 - c), i + j - 1);
  tmp *= p1 * p3;
  Ps += tmp;
  Pt = 0.5 * (1.0 + a) * (Pr) + 0.5 * (1.0 + b) * tmp;
  tmp = p3c * pow(0.5 * (1 - c), i + j);
  if (i + j > 0) tmp -= 0.5 * (i + j) * (p3 * pow(0.5 * (1.0 - c), i + j - 1));
  tmp *= p1 * p2 * pow(0.5 * (1 - b), i);
  Pt += tmp;
  Pr *= pow(2, 2 * i + j + 1.5);
  Ps *= pow(2, 2 * i + j + 1.5);
  Pt *= pow(2, 2 * i + j + 1.5);
}

// meshBasisTri2D.cpp:305-430 / meshBasisTet3D.cpp:398-527: V, grad V, MM =
// inv(V V^T), D = V_a / V (right solve), and S with the reference's loop
int SimplexOperators(Problem& P) {
  const int N = P.N, Np = P.Np, dim = P.dimension;
  std::vector<double> V((size_t)Np * Np), G[3];
  for (int d = 0; d < dim; ++d) G[d].assign((size_t)Np * Np, 0.0);
  for (int n = 0; n < Np; n++) {
    int sk = 0;
    if (dim == 2) {
      for (int i = 0; i < N + 1; i++)
        for (int j = 0; j < N + 1 - i; j++, sk++)
          BasisTri2D(P.r[n], P.s[n], i, j, V[(size_t)n * Np + sk], G[0][(size_t)n * Np + sk],
                     G[1][(size_t)n * Np + sk]);
    } else {
      for (int i = 0; i < N + 1; i++)
        for (int j = 0; j < N + 1 - i; j++)
          for (int k = 0; k < N + 1 - i - j; k++, sk++)
            BasisTet3D(P.r[n], P.s[n], P.t[n], i, j, k, V[(size_t)n * Np + sk], G[0][(size_t)n * Np + sk],
                       G[1][(size_t)n * Np + sk], G[2][(size_t)n * Np + sk]);
    }
  }
  P.V = V;
  P.MM.assign((size_t)Np * Np, 0.0);
  for (int n = 0; n < Np; ++n)
    for (int m = 0; m < Np; ++m) {
      double res = 0;
      for (int i = 0; i < Np; ++i) res += V[(size_t)n * Np + i] * V[(size_t)m * Np + i];
      P.MM[(size_t)n * Np + m] = res;
    }
  if (MatrixInverse(Np, P.MM.data())) return -1;
  P.Dsim.assign((size_t)dim * Np * Np, 0.0);
  for (int d = 0; d < dim; ++d)
    if (MatrixRightSolve(Np, Np, G[d].data(), Np, Np, V.data(), P.Dsim.data() + (size_t)d * Np * Np)) return -1;
  // SmatrixTri2D (:406-430) / SmatrixTet3D (:494-527):
  // S_ab[m + n Np] += D_a[n + l Np] MM[k + l Np] D_b[m + k Np] (+ the swap)
  const int nS = dim == 2 ? 3 : 6;
  const int pa[6] = {0, 0, 0, 1, 1, 2}, pb[6] = {0, 1, 2, 1, 2, 2};
  const int pa2[3] = {0, 0, 1}, pb2[3] = {0, 1, 1};
  P.Ssim.assign((size_t)nS * Np * Np, 0.0);
  const double* Dd = P.Dsim.data();
  const size_t NN = (size_t)Np * Np;
#pragma omp parallel for
  for (int n = 0; n < Np; n++)
    for (int c = 0; c < nS; ++c) {
      const int a = dim == 2 ? pa2[c] : pa[c], b = dim == 2 ? pb2[c] : pb[c];
      for (int m = 0; m < Np; m++) {
        double acc = 0;
        for (int k = 0; k < Np; k++)
          for (int l = 0; l < Np; l++) {
            acc += Dd[a * NN + n + (size_t)l * Np] * P.MM[k + (size_t)l * Np] * Dd[b * NN + m + (size_t)k * Np];
            if (a != b)
              acc += Dd[b * NN + n + (size_t)l * Np] * P.MM[k + (size_t)l * Np] * Dd[a * NN + m + (size_t)k * Np];
          }
        P.Ssim[c * NN + m + (size_t)n * Np] = acc;
      }
    }
  return 0;
}

// ------------------------------------------------------------- meshes --

// libs/mesh/meshSetupBoxQuad2D.cpp, meshSetupBoxTri2D.cpp:63-112,
// meshSetupBoxTet3D.cpp:79-178 (one rank: X0 = lo, DIM = dim); interior
// vertices moved as in the hex box
void SetupBoxOther(Problem& P) {
  const bool d3 = P.elementType == 6;
  const int nx = P.nx, ny = P.ny, nz = d3 ? P.nz : 1;
  const hlong NnX = nx + 1, NnY = ny + 1;
  const double dx = P.dim[0] / nx, dy = P.dim[1] / ny, dz = d3 ? P.dim[2] / nz : 0.0;
  const int per = P.elementType == 3 ? 2 : d3 ? 6 : 1;
  P.Nelements = per * nx * ny * nz;
  P.EToV.assign((size_t)P.Nelements * P.Nverts, 0);
  P.EX.assign((size_t)P.Nelements * P.Nverts, 0.0);
  P.EY.assign((size_t)P.Nelements * P.Nverts, 0.0);
  P.EZ.assign(d3 ? (size_t)P.Nelements * P.Nverts : 0, 0.0);
  // corners: (di, dj, dk) offsets of each element's vertices, in the
  // reference's order
  const int quad[1][4][3] = {{{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}}};
  const int tri[2][3][3] = {{{0, 0, 0}, {1, 0, 0}, {1, 1, 0}}, {{0, 0, 0}, {1, 1, 0}, {0, 1, 0}}};
  const int tet[6][4][3] = {{{0, 0, 0}, {1, 1, 0}, {0, 1, 0}, {1, 1, 1}}, {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {1, 1, 1}},
                            {{0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {1, 1, 1}}, {{0, 0, 0}, {0, 1, 1}, {0, 0, 1}, {1, 1, 1}},
                            {{0, 0, 0}, {1, 0, 1}, {1, 0, 0}, {1, 1, 1}}, {{0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}}};
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int q = 0; q < per; ++q) {
          const dlong e = per * (i + j * nx + k * nx * ny) + q;
          const double x0 = P.lo[0] + dx * i, y0 = P.lo[1] + dy * j, z0 = d3 ? P.lo[2] + dz * k : 0.0;
          for (int v = 0; v < P.Nverts; ++v) {
            const int* o = P.elementType == 4 ? quad[0][v] : P.elementType == 3 ? tri[q][v] : tet[q][v];
            const hlong I = i + o[0], J = j + o[1], K = k + o[2];
            const hlong vid = I + J * NnX + K * NnX * NnY;
            double X = o[0] ? x0 + dx : x0, Y = o[1] ? y0 + dy : y0, Z = o[2] ? z0 + dz : z0;
            const bool interior = I > 0 && I < nx && J > 0 && J < ny && (!d3 || (K > 0 && K < nz));
            if (P.perturb != 0.0 && interior) {
              X += P.perturb * dx * (2.0 * rng_u01(P.seed, 3 * (uint64_t)vid + 0) - 1.0);
              Y += P.perturb * dy * (2.0 * rng_u01(P.seed, 3 * (uint64_t)vid + 1) - 1.0);
              if (d3) Z += P.perturb * dz * (2.0 * rng_u01(P.seed, 3 * (uint64_t)vid + 2) - 1.0);
            }
            P.EToV[(size_t)e * P.Nverts + v] = vid;
            P.EX[(size_t)e * P.Nverts + v] = X;
            P.EY[(size_t)e * P.Nverts + v] = Y;
            if (d3) P.EZ[(size_t)e * P.Nverts + v] = Z;
          }
        }
}

// faces by their vertices: libs/mesh/meshSetElementType.cpp:38-73
const int kTriFaces[3][3] = {{0, 1, -1}, {1, 2, -1}, {2, 0, -1}};
const int kQuadFaces[4][3] = {{0, 1, -1}, {1, 2, -1}, {2, 3, -1}, {3, 0, -1}};
const int kTetFaces[4][3] = {{0, 1, 2}, {0, 3, 1}, {1, 3, 2}, {0, 2, 3}};

// libs/mesh/meshConnect.cpp (faces matched by sorted vertex ids),
// meshConnectFaceNodes.cpp (matching face nodes; here by their physical
// coordinates, which is what the reference's rotation tables encode) and
// meshConnectNodes.cpp:32-82 (iterated min over the face traces)
void ConnectOther(Problem& P) {
  const dlong Nel = P.Nelements;
  const int Nf = P.Nfaces, nfv = P.elementType == 6 ? 3 : 2, Np = P.Np, Nfp = P.Nfp;
  const int(*FV)[3] = P.elementType == 3 ? kTriFaces : P.elementType == 4 ? kQuadFaces : kTetFaces;
  P.EToE.assign((size_t)Nel * Nf, -1);
  P.EToF.assign((size_t)Nel * Nf, -1);
  std::map<std::array<hlong, 3>, std::pair<dlong, int>> open;
  for (dlong e = 0; e < Nel; ++e)
    for (int f = 0; f < Nf; ++f) {
      std::array<hlong, 3> key = {-1, -1, -1};
      for (int n = 0; n < nfv; ++n) key[n] = P.EToV[(size_t)e * P.Nverts + FV[f][n]];
      std::sort(key.begin(), key.begin() + nfv);
      auto it = open.find(key);
      if (it == open.end()) {
        open.emplace(key, std::make_pair(e, f));
      } else {
        const dlong eP = it->second.first;
        const int fP = it->second.second;
        P.EToE[(size_t)e * Nf + f] = eP;
        P.EToF[(size_t)e * Nf + f] = fP;
        P.EToE[(size_t)eP * Nf + fP] = e;
        P.EToF[(size_t)eP * Nf + fP] = f;
        open.erase(it);
      }
    }
  P.vmapM.assign((size_t)Nel * Nf * Nfp, 0);
  P.vmapP.assign((size_t)Nel * Nf * Nfp, 0);
  const bool d3 = P.elementType == 6;
#pragma omp parallel for
  for (dlong eM = 0; eM < Nel; ++eM)
    for (int fM = 0; fM < Nf; ++fM) {
      const dlong eP = P.EToE[(size_t)eM * Nf + fM];
      const int fP = P.EToF[(size_t)eM * Nf + fM];
      for (int nM = 0; nM < Nfp; ++nM) {
        const int idM = P.faceNodes[fM * Nfp + nM];
        const size_t id = ((size_t)eM * Nf + fM) * Nfp + nM;
        P.vmapM[id] = idM + eM * Np;
        P.vmapP[id] = idM + eM * Np;
        if (eP < 0) continue;
        const size_t gM = (size_t)eM * Np + idM;
        double best = 1e300;
        for (int nP = 0; nP < Nfp; ++nP) {
          const int idP = P.faceNodes[fP * Nfp + nP];
          const size_t gP = (size_t)eP * Np + idP;
          double d2 = (P.x[gM] - P.x[gP]) * (P.x[gM] - P.x[gP]) + (P.y[gM] - P.y[gP]) * (P.y[gM] - P.y[gP]);
          if (d3) d2 += (P.z[gM] - P.z[gP]) * (P.z[gM] - P.z[gP]);
          if (d2 < best) {
            best = d2;
            P.vmapP[id] = idP + eP * Np;
          }
        }
      }
    }
  P.globalIds.resize((size_t)Nel * Np);
  for (size_t id = 0; id < (size_t)Nel * Np; ++id) P.globalIds[id] = 1 + (hlong)id;
  hlong change = 1;
  while (change > 0) {
    change = 0;
    for (size_t id = 0; id < P.vmapM.size(); ++id) {
      const dlong idM = P.vmapM[id], idP = P.vmapP[id];
      if (P.globalIds[idP] < P.globalIds[idM]) {
        ++change;
        P.globalIds[idM] = P.globalIds[idP];
      }
    }
  }
}

// meshPhysicalNodesQuad2D.cpp / Tri2D.cpp / Tet3D.cpp
void PhysicalNodesOther(Problem& P) {
  const dlong Nel = P.Nelements;
  const int Np = P.Np, Nv = P.Nverts;
  const bool d3 = P.elementType == 6;
  P.x.assign((size_t)Nel * Np, 0.0);
  P.y.assign((size_t)Nel * Np, 0.0);
  P.z.assign(d3 ? (size_t)Nel * Np : 0, 0.0);
  for (dlong e = 0; e < Nel; ++e) {
    const double* X = &P.EX[(size_t)e * Nv];
    const double* Y = &P.EY[(size_t)e * Nv];
    for (int n = 0; n < Np; ++n) {
      const double rn = P.r[n], sn = P.s[n];
      const size_t id = (size_t)e * Np + n;
      if (P.elementType == 4) {
        P.x[id] = 0.25 * ((1 - rn) * (1 - sn) * X[0] + (1 + rn) * (1 - sn) * X[1] + (1 + rn) * (1 + sn) * X[2] +
                          (1 - rn) * (1 + sn) * X[3]);
        P.y[id] = 0.25 * ((1 - rn) * (1 - sn) * Y[0] + (1 + rn) * (1 - sn) * Y[1] + (1 + rn) * (1 + sn) * Y[2] +
                          (1 - rn) * (1 + sn) * Y[3]);
      } else if (!d3) {
        P.x[id] = -0.5 * (rn + sn) * X[0] + 0.5 * (1 + rn) * X[1] + 0.5 * (1 + sn) * X[2];
        P.y[id] = -0.5 * (rn + sn) * Y[0] + 0.5 * (1 + rn) * Y[1] + 0.5 * (1 + sn) * Y[2];
      } else {
        const double* Z = &P.EZ[(size_t)e * Nv];
        const double tn = P.t[n];
        P.x[id] = -0.5 * (1 + rn + sn + tn) * X[0] + 0.5 * (1 + rn) * X[1] + 0.5 * (1 + sn) * X[2] + 0.5 * (1 + tn) * X[3];
        P.y[id] = -0.5 * (1 + rn + sn + tn) * Y[0] + 0.5 * (1 + rn) * Y[1] + 0.5 * (1 + sn) * Y[2] + 0.5 * (1 + tn) * Y[3];
        P.z[id] = -0.5 * (1 + rn + sn + tn) * Z[0] + 0.5 * (1 + rn) * Z[1] + 0.5 * (1 + sn) * Z[2] + 0.5 * (1 + tn) * Z[3];
      }
    }
  }
}

// meshGeometricFactorsQuad2D.cpp:77-127 (per node), meshCubatureSetupQuad2D.cpp
// (cubature), meshGeometricFactorsTri2D.cpp:46-90 / Tet3D.cpp:54-107 (affine)
int GeometricFactorsOther(Problem& P) {
  const dlong Nel = P.Nelements;
  const int Np = P.Np, Nq = P.Nq;
  if (P.elementType == 4) {
    P.wJ.assign((size_t)Nel * Np, 0.0);
    P.ggeo.assign((size_t)Nel * Np * 3, 0.0);
    for (dlong e = 0; e < Nel; ++e)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int n = i + j * Nq;
          double xr = 0, xs = 0, yr = 0, ys = 0;
          for (int m = 0; m < Nq; ++m) {
            const size_t idr = (size_t)e * Np + j * Nq + m, ids = (size_t)e * Np + m * Nq + i;
            xr += P.D[i * Nq + m] * P.x[idr];
            xs += P.D[j * Nq + m] * P.x[ids];
            yr += P.D[i * Nq + m] * P.y[idr];
            ys += P.D[j * Nq + m] * P.y[ids];
          }
          const double J = xr * ys - xs * yr;
          if (J < 1e-8) {
            P.error = "negative J";
            return -1;
          }
          const double rx = ys / J, ry = -xs / J, sx = -yr / J, sy = xr / J;
          const double JW = J * P.gllw[i] * P.gllw[j];
          P.wJ[(size_t)Np * e + n] = JW;
          P.ggeo[3 * ((size_t)Np * e + n) + 0] = JW * (rx * rx + ry * ry);
          P.ggeo[3 * ((size_t)Np * e + n) + 1] = JW * (rx * sx + ry * sy);
          P.ggeo[3 * ((size_t)Np * e + n) + 2] = JW * (sx * sx + sy * sy);
        }
    if (P.bp <= 4) {
      const int cq = P.cubNq;
      P.cubwJ.assign((size_t)Nel * cq * cq, 0.0);
      P.cubggeo.assign((size_t)Nel * cq * cq * 3, 0.0);
      std::vector<double> xre(Np), xse(Np), yre(Np), yse(Np), xre1(cq * Nq), xse1(cq * Nq), yre1(cq * Nq),
          yse1(cq * Nq);
      for (dlong e = 0; e < Nel; ++e) {
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i) {
            const int n = i + j * Nq;
            xre[n] = xse[n] = yre[n] = yse[n] = 0.0;
            for (int m = 0; m < Nq; ++m) {
              const size_t idr = (size_t)e * Np + j * Nq + m, ids = (size_t)e * Np + m * Nq + i;
              xre[n] += P.D[i * Nq + m] * P.x[idr];
              xse[n] += P.D[j * Nq + m] * P.x[ids];
              yre[n] += P.D[i * Nq + m] * P.y[idr];
              yse[n] += P.D[j * Nq + m] * P.y[ids];
            }
          }
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < cq; ++i) {
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
            for (int n = 0; n < Nq; ++n) {
              a0 += P.cubInterp[i * Nq + n] * xre[j * Nq + n];
              a1 += P.cubInterp[i * Nq + n] * xse[j * Nq + n];
              a2 += P.cubInterp[i * Nq + n] * yre[j * Nq + n];
              a3 += P.cubInterp[i * Nq + n] * yse[j * Nq + n];
            }
            xre1[j * cq + i] = a0;
            xse1[j * cq + i] = a1;
            yre1[j * cq + i] = a2;
            yse1[j * cq + i] = a3;
          }
        for (int j = 0; j < cq; ++j)
          for (int i = 0; i < cq; ++i) {
            double xr = 0, xs = 0, yr = 0, ys = 0;
            for (int n = 0; n < Nq; ++n) {
              xr += P.cubInterp[j * Nq + n] * xre1[n * cq + i];
              xs += P.cubInterp[j * Nq + n] * xse1[n * cq + i];
              yr += P.cubInterp[j * Nq + n] * yre1[n * cq + i];
              ys += P.cubInterp[j * Nq + n] * yse1[n * cq + i];
            }
            const double J = xr * ys - xs * yr;
            if (J < 1e-8) {
              P.error = "negative J";
              return -1;
            }
            const double rx = ys / J, ry = -xs / J, sx = -yr / J, sy = xr / J;
            const double JW = J * P.cubw[i] * P.cubw[j];
            const size_t g = (size_t)e * cq * cq + i + j * cq;
            P.cubwJ[g] = JW;
            P.cubggeo[3 * g + 0] = JW * (rx * rx + ry * ry);
            P.cubggeo[3 * g + 1] = JW * (rx * sx + ry * sy);
            P.cubggeo[3 * g + 2] = JW * (sx * sx + sy * sy);
          }
      }
    }
    return 0;
  }
  const bool d3 = P.elementType == 6;
  const int NG = d3 ? 6 : 3;
  P.wJ.assign(Nel, 0.0);
  P.ggeo.assign((size_t)Nel * NG, 0.0);
  for (dlong e = 0; e < Nel; ++e) {
    const size_t id = (size_t)e * P.Nverts;
    if (!d3) {
      const double xr = 0.5 * (P.EX[id + 1] - P.EX[id + 0]), xs = 0.5 * (P.EX[id + 2] - P.EX[id + 0]);
      const double yr = 0.5 * (P.EY[id + 1] - P.EY[id + 0]), ys = 0.5 * (P.EY[id + 2] - P.EY[id + 0]);
      const double J = xr * ys - xs * yr;
      if (J < 1e-8) {
        P.error = "negative J";
        return -1;
      }
      const double rx = ys / J, ry = -xs / J, sx = -yr / J, sy = xr / J;
      P.wJ[e] = J;
      P.ggeo[3 * (size_t)e + 0] = J * (rx * rx + ry * ry);
      P.ggeo[3 * (size_t)e + 1] = J * (rx * sx + ry * sy);
      P.ggeo[3 * (size_t)e + 2] = J * (sx * sx + sy * sy);
    } else {
      const double xe1 = P.EX[id + 0], ye1 = P.EY[id + 0], ze1 = P.EZ[id + 0];
      const double xe2 = P.EX[id + 1], ye2 = P.EY[id + 1], ze2 = P.EZ[id + 1];
      const double xe3 = P.EX[id + 2], ye3 = P.EY[id + 2], ze3 = P.EZ[id + 2];
      const double xe4 = P.EX[id + 3], ye4 = P.EY[id + 3], ze4 = P.EZ[id + 3];
      const double xr = 0.5 * (xe2 - xe1), xs = 0.5 * (xe3 - xe1), xt = 0.5 * (xe4 - xe1);
      const double yr = 0.5 * (ye2 - ye1), ys = 0.5 * (ye3 - ye1), yt = 0.5 * (ye4 - ye1);
      const double zr = 0.5 * (ze2 - ze1), zs = 0.5 * (ze3 - ze1), zt = 0.5 * (ze4 - ze1);
      const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
      const double rx = (ys * zt - zs * yt) / J, ry = -(xs * zt - zs * xt) / J, rz = (xs * yt - ys * xt) / J;
      const double sx = -(yr * zt - zr * yt) / J, sy = (xr * zt - zr * xt) / J, sz = -(xr * yt - yr * xt) / J;
      const double tx = (yr * zs - zr * ys) / J, ty = -(xr * zs - zr * xs) / J, tz = (xr * ys - yr * xs) / J;
      if (J < 0) {
        P.error = "negative J";
        return -1;
      }
      P.wJ[e] = J;
      double* g = &P.ggeo[6 * (size_t)e];
      g[0] = J * (rx * rx + ry * ry + rz * rz);
      g[1] = J * (rx * sx + ry * sy + rz * sz);
      g[2] = J * (rx * tx + ry * ty + rz * tz);
      g[3] = J * (sx * sx + sy * sy + sz * sz);
      g[4] = J * (sx * tx + sy * ty + sz * tz);
      g[5] = J * (tx * tx + ty * ty + tz * tz);
    }
  }
  return 0;
}

// meshMaskedGatherScatterSetup.cpp:31-77: face nodes of faces without a
// neighbour are masked
void MaskedIdsOther(Problem& P) {
  const dlong Nel = P.Nelements;
  const int Np = P.Np, Nfp = P.Nfp, Nf = P.Nfields, Nfaces = P.Nfaces;
  P.maskedIds.resize((size_t)Nel * Np * Nf);
  for (dlong e = 0; e < Nel; e++)
    for (int n = 0; n < Np; n++) {
      const hlong id = P.globalIds[(size_t)e * Np + n];
      for (int f = 0; f < Nf; f++) P.maskedIds[(size_t)e * Np * Nf + f * Np + n] = id * Nf + f;
    }
  if (!P.mask) return;
  for (dlong e = 0; e < Nel; e++)
    for (int f = 0; f < Nfaces; f++)
      if (P.EToE[(size_t)f + e * Nfaces] < 0)
        for (int n = 0; n < Nfp; n++) {
          const int fid = P.faceNodes[n + f * Nfp];
          for (int nf = 0; nf < Nf; nf++) P.maskedIds[(size_t)e * Np * Nf + nf * Np + fid] = 0;
        }
}

}  // namespace

int ogs_setup_maps(Problem& P);  // mesh.cpp

// mesh_t setup for the quad, triangle and tetrahedron boxes
// (libs/mesh/meshSetup.cpp:31-85 with the per-family files above)
int problem_setup_other(Problem& P) {
  const int et = P.elementType;
  if (et != 3 && et != 4 && et != 6) {
    P.error = "element type must be 3, 4, 6 or 12";
    return -1;
  }
  const bool d3 = et == 6;
  if (P.N < 1 || P.N > 15 || P.nx < 1 || P.ny < 1 || (d3 && P.nz < 1)) {
    P.error = "invalid problem size";
    return -1;
  }
  if (!d3) P.nz = 1;
  P.dimension = d3 ? 3 : 2;
  P.Nfields = (P.bp == 2 || P.bp == 4 || P.bp == 6) ? P.dimension : 1;
  P.Nq = P.N + 1;
  const int N = P.N;
  P.Nverts = et == 3 ? 3 : 4;
  P.Nfaces = et == 3 ? 3 : 4;
  // reference nodes and operators
  if (et == 4) {
    // meshReferenceNodesQuad2D.cpp: GLL tensor nodes (r fastest), 1D D, MM
    P.Np = P.Nq * P.Nq;
    P.Nfp = P.Nq;
    P.gllz.resize(P.Nq);
    P.gllw.resize(P.Nq);
    JacobiGLL(N, P.gllz.data(), P.gllw.data());
    P.r.resize(P.Np);
    P.s.resize(P.Np);
    for (int j = 0; j < P.Nq; ++j)
      for (int i = 0; i < P.Nq; ++i) {
        P.r[i + j * P.Nq] = P.gllz[i];
        P.s[i + j * P.Nq] = P.gllz[j];
      }
    P.D.resize(P.Nq * P.Nq);
    Dmatrix1D(N, P.Nq, P.gllz.data(), P.gllz.data(), P.D.data());
    // meshBasisQuad2D.cpp FaceNodesQuad2D: s = -1, r = 1, s = 1, r = -1
    P.faceNodes.resize(4 * P.Nfp);
    for (int i = 0; i < P.Nq; ++i) {
      P.faceNodes[0 * P.Nfp + i] = i;
      P.faceNodes[1 * P.Nfp + i] = N + i * P.Nq;
      P.faceNodes[2 * P.Nfp + i] = i + N * P.Nq;
      P.faceNodes[3 * P.Nfp + i] = i * P.Nq;
    }
    if (P.bp <= 4) {
      // meshCubatureSetupQuad2D.cpp:32-70 (non-affine): N+2 Gauss points
      P.cubN = N + 1;
      P.cubNq = P.cubN + 1;
      P.cubNp = P.cubNq * P.cubNq;
      P.cubr.resize(P.cubNq);
      P.cubw.resize(P.cubNq);
      JacobiGQ(0, 0, P.cubN, P.cubr.data(), P.cubw.data());
      P.cubInterp.resize(P.cubNq * P.Nq);
      InterpolationMatrix1D(N, P.cubNq, P.gllz.data(), P.cubr.data(), P.cubInterp.data());
      P.cubD.resize(P.cubNq * P.cubNq);
      Dmatrix1D(P.cubN, P.cubNq, P.cubr.data(), P.cubr.data(), P.cubD.data());
    }
  } else {
    P.Np = d3 ? (N + 1) * (N + 2) * (N + 3) / 6 : (N + 1) * (N + 2) / 2;
    P.Nfp = d3 ? (N + 1) * (N + 2) / 2 : N + 1;
    if (d3)
      NodesTet3D(N, P.r, P.s, P.t);
    else
      NodesTri2D(N, P.r, P.s);
    if (SimplexOperators(P)) {
      P.error = "singular Vandermonde matrix";
      return -1;
    }
    // FaceNodesTri2D (meshBasisTri2D.cpp:41-62), FaceNodesTet3D (meshBasisTet3D.cpp:42-66)
    const double NODETOL = 1.0e-5;
    P.faceNodes.assign((size_t)P.Nfaces * P.Nfp, 0);
    int cnt[4] = {0, 0, 0, 0};
    for (int n = 0; n < P.Np; n++) {
      if (!d3) {
        if (fabs(P.s[n] + 1) < NODETOL) P.faceNodes[0 * P.Nfp + (cnt[0]++)] = n;
        if (fabs(P.r[n] + P.s[n]) < NODETOL) P.faceNodes[1 * P.Nfp + (cnt[1]++)] = n;
        if (fabs(P.r[n] + 1) < NODETOL) P.faceNodes[2 * P.Nfp + (cnt[2]++)] = n;
      } else {
        if (fabs(P.t[n] + 1) < NODETOL) P.faceNodes[0 * P.Nfp + (cnt[0]++)] = n;
        if (fabs(P.s[n] + 1) < NODETOL) P.faceNodes[1 * P.Nfp + (cnt[1]++)] = n;
        if (fabs(P.r[n] + P.s[n] + P.t[n] + 1.0) < NODETOL) P.faceNodes[2 * P.Nfp + (cnt[2]++)] = n;
        if (fabs(P.r[n] + 1) < NODETOL) P.faceNodes[3 * P.Nfp + (cnt[3]++)] = n;
      }
    }
  }
  SetupBoxOther(P);
  PhysicalNodesOther(P);
  ConnectOther(P);
  if (int st = GeometricFactorsOther(P)) return st;
  P.elementList.resize(P.Nelements);
  std::iota(P.elementList.begin(), P.elementList.end(), 0);
  MaskedIdsOther(P);
  P.NLocal = P.Nelements * P.Np * P.Nfields;
  return ogs_setup_maps(P);
}

// ------------------------------------------------------------ operators --

// okl/bp5AxQuad2D.okl (GLL: D^T G D q + lambda wJ q), okl/bp1AxQuad2D.okl
// (I^T cubwJ I q), okl/bp3AxQuad2D.okl (I^T (cubD^T G cubD + lambda cubwJ) I q),
// per field; okl/bp5AxAffineTri2D.okl / Tet3D.okl and bp3AxAffine*.okl
// (sum_ab G_ab S_ab q + lambda J MM q), bp1AxAffine*.okl (J MM q)
void ax_local_other(const Problem& P, const double* q, double* AqL) {
  const dlong Nel = P.Nelements;
  const int Np = P.Np, Nq = P.Nq, NF = P.Nfields;
#pragma omp parallel for
  for (dlong e = 0; e < Nel; ++e) {
    std::vector<double> qe(Np), a(Np), b(Np), u, t1, fr, fs;
    for (int f = 0; f < NF; ++f) {
      const size_t lb = ((size_t)e * NF + f) * Np;
      for (int n = 0; n < Np; ++n) {
        const dlong id = P.GlobalToLocal[lb + n];
        qe[n] = (id != -1) ? q[id] : 0.0;
      }
      if (P.elementType == 4 && (P.bp == 5 || P.bp == 6)) {
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i) {
            double qr = 0, qs = 0;
            for (int m = 0; m < Nq; ++m) {
              qr += P.D[i * Nq + m] * qe[j * Nq + m];
              qs += P.D[j * Nq + m] * qe[m * Nq + i];
            }
            const size_t gb = (size_t)e * Np + i + j * Nq;
            const double G00 = P.ggeo[3 * gb + 0], G01 = P.ggeo[3 * gb + 1], G11 = P.ggeo[3 * gb + 2];
            a[j * Nq + i] = G00 * qr + G01 * qs;
            b[j * Nq + i] = G01 * qr + G11 * qs;
          }
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i) {
            double r = 0;
            for (int m = 0; m < Nq; ++m) {
              r += P.D[m * Nq + i] * a[j * Nq + m];
              r += P.D[m * Nq + j] * b[m * Nq + i];
            }
            const size_t gb = (size_t)e * Np + i + j * Nq;
            AqL[lb + i + j * Nq] = r + P.lambda * P.wJ[gb] * qe[j * Nq + i];
          }
      } else if (P.elementType == 4) {
        const int cq = P.cubNq;
        u.assign(cq * cq, 0.0);
        t1.assign(Nq * cq, 0.0);
        for (int j = 0; j < Nq; ++j)
          for (int c = 0; c < cq; ++c) {
            double acc = 0;
            for (int i = 0; i < Nq; ++i) acc += P.cubInterp[c * Nq + i] * qe[j * Nq + i];
            t1[j * cq + c] = acc;
          }
        for (int d = 0; d < cq; ++d)
          for (int c = 0; c < cq; ++c) {
            double acc = 0;
            for (int j = 0; j < Nq; ++j) acc += P.cubInterp[d * Nq + j] * t1[j * cq + c];
            u[d * cq + c] = acc;
          }
        std::vector<double> res(cq * cq);
        const size_t cb = (size_t)e * cq * cq;
        if (P.bp <= 2) {
          for (int g = 0; g < cq * cq; ++g) res[g] = P.cubwJ[cb + g] * u[g];
        } else {
          fr.assign(cq * cq, 0.0);
          fs.assign(cq * cq, 0.0);
          for (int d = 0; d < cq; ++d)
            for (int c = 0; c < cq; ++c) {
              double ur = 0, us = 0;
              for (int m = 0; m < cq; ++m) {
                ur += P.cubD[c * cq + m] * u[d * cq + m];
                us += P.cubD[d * cq + m] * u[m * cq + c];
              }
              const size_t g = cb + d * cq + c;
              const double G00 = P.cubggeo[3 * g + 0], G01 = P.cubggeo[3 * g + 1], G11 = P.cubggeo[3 * g + 2];
              fr[d * cq + c] = G00 * ur + G01 * us;
              fs[d * cq + c] = G01 * ur + G11 * us;
            }
          for (int d = 0; d < cq; ++d)
            for (int c = 0; c < cq; ++c) {
              double r = P.lambda * P.cubwJ[cb + d * cq + c] * u[d * cq + c];
              for (int m = 0; m < cq; ++m) {
                r += P.cubD[m * cq + c] * fr[d * cq + m];
                r += P.cubD[m * cq + d] * fs[m * cq + c];
              }
              res[d * cq + c] = r;
            }
        }
        for (int j = 0; j < Nq; ++j)
          for (int c = 0; c < cq; ++c) {
            double acc = 0;
            for (int d = 0; d < cq; ++d) acc += P.cubInterp[d * Nq + j] * res[d * cq + c];
            t1[j * cq + c] = acc;
          }
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i) {
            double acc = 0;
            for (int c = 0; c < cq; ++c) acc += P.cubInterp[c * Nq + i] * t1[j * cq + c];
            AqL[lb + i + j * Nq] = acc;
          }
      } else {
        const int NG = P.dimension == 3 ? 6 : 3;
        const size_t NN = (size_t)Np * Np;
        const double J = P.wJ[e];
        for (int n = 0; n < Np; ++n) {
          double qM = 0, qS[6] = {0, 0, 0, 0, 0, 0};
          for (int k = 0; k < Np; ++k) {
            qM += P.MM[n + (size_t)k * Np] * qe[k];
            for (int c = 0; c < NG; ++c) qS[c] += P.Ssim[n + (size_t)k * Np + c * NN] * qe[k];
          }
          double out;
          if (P.bp <= 2) {
            out = J * qM;
          } else {
            out = 0;
            for (int c = 0; c < NG; ++c) out += P.ggeo[(size_t)NG * e + c] * qS[c];
            out += J * P.lambda * qM;
          }
          AqL[lb + n] = out;
        }
      }
    }
  }
}

// okl/rhsQuad2D.okl (rhs = wJ f), okl/rhsTri2D.okl / rhsTet3D.okl
// (rhs = MM (J f)); BP1/BP2 are forced with lambda = 0 (bp1Run.cpp)
void rhs_forcing_local_other(const Problem& P, double* rhsL) {
  const dlong Nel = P.Nelements;
  const int Np = P.Np, NF = P.Nfields;
  const double lam = P.bp <= 2 ? 0.0 : P.lambda;
  const double PI = M_PI;
  auto forcing = [&](double x, double y, double z, double* f) {
    if (P.dimension == 2) {
      if (NF == 1) {
        f[0] = (2 * PI * PI + lam) * sin(PI * x) * sin(PI * y);
      } else {
        f[0] = (3 * PI * PI + lam) * sin(2.0 * PI * x) * sin(PI * y);
        f[1] = (3 * PI * PI + lam) * sin(PI * x) * sin(2.0 * PI * y);
      }
    } else if (NF == 1) {
      f[0] = (3 * PI * PI + lam) * sin(PI * x) * sin(PI * y) * sin(PI * z);
    } else {
      f[0] = (4 * PI * PI + lam) * sin(2.0 * PI * x) * sin(PI * y) * sin(PI * z);
      f[1] = (4 * PI * PI + lam) * sin(PI * x) * sin(2.0 * PI * y) * sin(PI * z);
      f[2] = (4 * PI * PI + lam) * sin(PI * x) * sin(PI * y) * sin(2.0 * PI * z);
    }
  };
  for (dlong e = 0; e < Nel; ++e) {
    std::vector<double> s((size_t)NF * Np);
    for (int n = 0; n < Np; ++n) {
      const size_t id = (size_t)e * Np + n;
      double f[3];
      forcing(P.x[id], P.y[id], P.dimension == 3 ? P.z[id] : 0.0, f);
      const double W = P.elementType == 4 ? P.wJ[id] : P.wJ[e];
      for (int c = 0; c < NF; ++c) s[(size_t)c * Np + n] = W * f[c];
    }
    for (int n = 0; n < Np; ++n)
      for (int c = 0; c < NF; ++c) {
        double r = s[(size_t)c * Np + n];
        if (P.elementType != 4) {
          r = 0;
          for (int m = 0; m < Np; ++m) r += P.MM[n + (size_t)m * Np] * s[(size_t)c * Np + m];
        }
        rhsL[((size_t)e * NF + c) * Np + n] = r;
      }
  }
}

}  // namespace orc

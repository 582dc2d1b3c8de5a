// ORACLE — test infrastructure only (see oracle.hpp).
// Restates the hex Ax loop nests of okl/bp{1..6}Ax{,Affine}Hex3D.okl (the
// HIP/Serial branches), okl/rhsHex3D.okl and libs/ogs/okl/ogsKernels.okl.
// Each element is one OKL @outer iteration; @inner loops run serially in
// the order the OKL writes them. Multi-field kernels (BP2/4/6) use the
// local layout [element][field][node] of okl/bp6AxHex3D.okl:600-606.
#include <cmath>
#include <vector>

#include "oracle.hpp"

namespace orc {

namespace {

inline double loadq(const dlong* G2L, const double* q, size_t idx) {
  const dlong id = G2L[idx];
  return (id != -1) ? q[id] : 0.0;
}

// okl/bp5AxHex3D.okl:415-569 (3D-LDS variant; the p>8 slice variant at
// :189-390 forms the same sums in a different order). Affine:
// okl/bp5AxAffineHex3D.okl:419-579.
void bp5_element(const Problem& P, dlong e, int f, const double* q, double* Aq) {
  const int Nq = P.Nq, Np = P.Np, Nf = P.Nfields;
  const double* D = P.D.data();
  std::vector<double> sq(Np), Gqr(Np), Gqs(Np), Gqt(Np), rwJ(Np);
  const size_t lbase = (size_t)e * Np * Nf + (size_t)f * Np;
  for (int n = 0; n < Np; ++n) sq[n] = loadq(P.GlobalToLocal.data(), q, lbase + n);
  for (int k = 0; k < Nq; ++k)
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i) {
        const int n = i + j * Nq + k * Nq * Nq;
        double G00, G01, G02, G11, G12, G22, wj;
        if (P.affine) {
          const double* g = &P.ggeo[(size_t)6 * e];
          G00 = g[0]; G01 = g[1]; G02 = g[2]; G11 = g[3]; G12 = g[4]; G22 = g[5];
          wj = P.wJ[e];
        } else {
          const size_t gb = (size_t)e * Np + n;
          const double* g = &P.ggeo[6 * gb];
          G00 = g[0]; G01 = g[1]; G02 = g[2]; G11 = g[3]; G12 = g[4]; G22 = g[5];
          wj = P.wJ[gb];
        }
        if (!P.kappaL.empty()) {  // variable coefficient: kappa scales G
          const double kap = P.kappaL[(size_t)e * Np + n];
          G00 *= kap; G01 *= kap; G02 *= kap; G11 *= kap; G12 *= kap; G22 *= kap;
        }
        double tmp = 0.0;
        for (int m = 0; m < Nq; ++m) tmp += D[k * Nq + m] * sq[i + j * Nq + m * Nq * Nq];
        double gr = G02 * tmp, gs = G12 * tmp, gt = G22 * tmp;
        tmp = 0;
        for (int m = 0; m < Nq; ++m) tmp += D[i * Nq + m] * sq[m + j * Nq + k * Nq * Nq];
        gr += G00 * tmp; gs += G01 * tmp; gt += G02 * tmp;
        tmp = 0;
        for (int m = 0; m < Nq; ++m) tmp += D[j * Nq + m] * sq[i + m * Nq + k * Nq * Nq];
        gr += G01 * tmp; gs += G11 * tmp; gt += G12 * tmp;
        if (P.affine) {
          const double w = P.gllw[i] * P.gllw[j] * P.gllw[k];
          gr *= w; gs *= w; gt *= w; wj *= w;
        }
        Gqr[n] = gr; Gqs[n] = gs; Gqt[n] = gt; rwJ[n] = wj;
      }
  for (int k = 0; k < Nq; ++k)
    for (int j = 0; j < Nq; ++j)
      for (int i = 0; i < Nq; ++i) {
        const int n = i + j * Nq + k * Nq * Nq;
        double acc = sq[n] * P.lambda * rwJ[n];
        for (int m = 0; m < Nq; ++m) {
          acc += D[m * Nq + i] * Gqr[m + j * Nq + k * Nq * Nq];
          acc += D[m * Nq + j] * Gqs[i + m * Nq + k * Nq * Nq];
        }
        for (int m = 0; m < Nq; ++m) acc += D[m * Nq + k] * Gqt[i + j * Nq + m * Nq * Nq];
        Aq[lbase + n] = acc;
      }
}

// Apply a 1D operator A (rows x cols, row-major) along one direction of a
// 3D tensor u[c2][c1][c0] -> v. dir 0 contracts index c0, 1 -> c1, 2 -> c2.
void apply1d(const double* A, int rows, int cols, int dir, const int nin[3],
             const double* u, double* v) {
  int nout[3] = {nin[0], nin[1], nin[2]};
  nout[dir] = rows;
  for (int c2 = 0; c2 < nout[2]; ++c2)
    for (int c1 = 0; c1 < nout[1]; ++c1)
      for (int c0 = 0; c0 < nout[0]; ++c0) {
        const int o[3] = {c0, c1, c2};
        double acc = 0;
        for (int m = 0; m < cols; ++m) {
          int ii[3] = {c0, c1, c2};
          ii[dir] = m;
          acc += A[o[dir] * cols + m] * u[ii[0] + nin[0] * (ii[1] + nin[1] * ii[2])];
        }
        v[c0 + nout[0] * (c1 + nout[1] * c2)] = acc;
      }
}

void transpose(const double* A, int rows, int cols, double* AT) {
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) AT[c * rows + r] = A[r * cols + c];
}

// okl/bp1AxHex3D.okl:61-325: interpolate GLL->Gauss in b, a, c; scale by
// cubwJ; project back in b, a, c. Affine: okl/bp1AxAffineHex3D.okl:61-191,
// J * (MM x MM x MM) q applied in r, s, t.
void bp1_element(const Problem& P, dlong e, int f, const double* q, double* Aq) {
  const int Nq = P.Nq, Np = P.Np, Nf = P.Nfields, cNq = P.cubNq;
  const size_t lbase = (size_t)e * Np * Nf + (size_t)f * Np;
  std::vector<double> u(Np);
  for (int n = 0; n < Np; ++n) u[n] = loadq(P.GlobalToLocal.data(), q, lbase + n);
  if (P.affine) {
    std::vector<double> a(Np), b(Np), MMT(Nq * Nq);
    transpose(P.MM.data(), Nq, Nq, MMT.data());  // kernel uses s_MM[m][i]
    const int n3[3] = {Nq, Nq, Nq};
    apply1d(MMT.data(), Nq, Nq, 0, n3, u.data(), a.data());
    apply1d(MMT.data(), Nq, Nq, 1, n3, a.data(), b.data());
    apply1d(MMT.data(), Nq, Nq, 2, n3, b.data(), a.data());
    const double J = P.cubwJ[e];
    for (int n = 0; n < Np; ++n) Aq[lbase + n] = J * a[n];
    return;
  }
  const double* I = P.cubInterp.data();
  std::vector<double> IT(Nq * cNq);
  transpose(I, cNq, Nq, IT.data());
  std::vector<double> t1(Nq * cNq * Nq), t2(cNq * cNq * Nq), t3(cNq * cNq * cNq);
  const int s0[3] = {Nq, Nq, Nq};
  apply1d(I, cNq, Nq, 1, s0, u.data(), t1.data());
  const int s1[3] = {Nq, cNq, Nq};
  apply1d(I, cNq, Nq, 0, s1, t1.data(), t2.data());
  const int s2[3] = {cNq, cNq, Nq};
  apply1d(I, cNq, Nq, 2, s2, t2.data(), t3.data());
  const double* W = &P.cubwJ[(size_t)e * P.cubNp];
  for (int n = 0; n < P.cubNp; ++n) t3[n] *= W[n];
  const int s3[3] = {cNq, cNq, cNq};
  apply1d(IT.data(), Nq, cNq, 1, s3, t3.data(), t2.data());
  const int s4[3] = {cNq, Nq, cNq};
  apply1d(IT.data(), Nq, cNq, 0, s4, t2.data(), t1.data());
  const int s5[3] = {Nq, Nq, cNq};
  std::vector<double> out(Np);
  apply1d(IT.data(), Nq, cNq, 2, s5, t1.data(), out.data());
  for (int n = 0; n < Np; ++n) Aq[lbase + n] = out[n];
}

// okl/bp3AxHex3D.okl:369-712: interpolate to the (p+2)^3 Gauss points,
// differentiate with cubD, apply cubggeo, add lambda*cubwJ*q, weak
// derivative transposes, project back. Affine: okl/bp3AxAffineHex3D.okl:
// 332-646 (nodal -> modal by invV, modal derivative cubD, per-element G,
// identity modal mass, back by invV^T).
void bp3_element(const Problem& P, dlong e, int f, const double* q, double* Aq) {
  const int Nq = P.Nq, Np = P.Np, Nf = P.Nfields;
  const size_t lbase = (size_t)e * Np * Nf + (size_t)f * Np;
  std::vector<double> u(Np);
  for (int n = 0; n < Np; ++n) u[n] = loadq(P.GlobalToLocal.data(), q, lbase + n);
  int Mq;  // points per direction where the operator is applied
  std::vector<double> uq;
  const double* Dd = P.cubD.data();
  if (P.affine) {
    Mq = Nq;
    const int n3[3] = {Nq, Nq, Nq};
    std::vector<double> a(Np), b(Np);
    apply1d(P.invV.data(), Nq, Nq, 2, n3, u.data(), a.data());
    apply1d(P.invV.data(), Nq, Nq, 0, n3, a.data(), b.data());
    uq.resize(Np);
    apply1d(P.invV.data(), Nq, Nq, 1, n3, b.data(), uq.data());
  } else {
    Mq = P.cubNq;
    const int cNq = P.cubNq;
    const double* I = P.cubInterp.data();
    std::vector<double> t1(Nq * cNq * Nq), t2(cNq * cNq * Nq);
    uq.resize(P.cubNp);
    const int s0[3] = {Nq, Nq, Nq};
    apply1d(I, cNq, Nq, 1, s0, u.data(), t1.data());
    const int s1[3] = {Nq, cNq, Nq};
    apply1d(I, cNq, Nq, 0, s1, t1.data(), t2.data());
    const int s2[3] = {cNq, cNq, Nq};
    apply1d(I, cNq, Nq, 2, s2, t2.data(), uq.data());
  }
  const int Mp = Mq * Mq * Mq;
  std::vector<double> Gr(Mp), Gs(Mp), Gt(Mp), res(Mp);
  for (int k = 0; k < Mq; ++k)
    for (int j = 0; j < Mq; ++j)
      for (int i = 0; i < Mq; ++i) {
        const int n = i + j * Mq + k * Mq * Mq;
        double dr = 0, ds = 0, dt = 0;
        for (int m = 0; m < Mq; ++m) {
          dr += Dd[i * Mq + m] * uq[m + j * Mq + k * Mq * Mq];
          ds += Dd[j * Mq + m] * uq[i + m * Mq + k * Mq * Mq];
          dt += Dd[k * Mq + m] * uq[i + j * Mq + m * Mq * Mq];
        }
        const double* g;
        double wj;
        if (P.affine) {
          g = &P.ggeo[(size_t)6 * e];
          wj = P.wJ[e];
        } else {
          const size_t cb = (size_t)e * Mp + n;
          g = &P.cubggeo[6 * cb];
          wj = P.cubwJ[cb];
        }
        Gr[n] = g[0] * dr + g[1] * ds + g[2] * dt;
        Gs[n] = g[1] * dr + g[3] * ds + g[4] * dt;
        Gt[n] = g[2] * dr + g[4] * ds + g[5] * dt;
        res[n] = P.lambda * wj * uq[n];
      }
  for (int k = 0; k < Mq; ++k)
    for (int j = 0; j < Mq; ++j)
      for (int i = 0; i < Mq; ++i) {
        const int n = i + j * Mq + k * Mq * Mq;
        double acc = 0;
        for (int m = 0; m < Mq; ++m) {
          acc += Dd[m * Mq + i] * Gr[m + j * Mq + k * Mq * Mq];
          acc += Dd[m * Mq + j] * Gs[i + m * Mq + k * Mq * Mq];
          acc += Dd[m * Mq + k] * Gt[i + j * Mq + m * Mq * Mq];
        }
        res[n] += acc;
      }
  std::vector<double> out(Np);
  if (P.affine) {
    std::vector<double> VT(Nq * Nq), a(Np), b(Np);
    transpose(P.invV.data(), Nq, Nq, VT.data());
    const int n3[3] = {Nq, Nq, Nq};
    apply1d(VT.data(), Nq, Nq, 1, n3, res.data(), a.data());
    apply1d(VT.data(), Nq, Nq, 0, n3, a.data(), b.data());
    apply1d(VT.data(), Nq, Nq, 2, n3, b.data(), out.data());
  } else {
    const int cNq = P.cubNq;
    std::vector<double> IT(Nq * cNq), t1(cNq * Nq * cNq), t2(Nq * Nq * cNq);
    transpose(P.cubInterp.data(), cNq, Nq, IT.data());
    const int s3[3] = {cNq, cNq, cNq};
    apply1d(IT.data(), Nq, cNq, 1, s3, res.data(), t1.data());
    const int s4[3] = {cNq, Nq, cNq};
    apply1d(IT.data(), Nq, cNq, 0, s4, t1.data(), t2.data());
    const int s5[3] = {Nq, Nq, cNq};
    apply1d(IT.data(), Nq, cNq, 2, s5, t2.data(), out.data());
  }
  for (int n = 0; n < Np; ++n) Aq[lbase + n] = out[n];
}

}  // namespace

// BK/BKn run the Ax kernel over all elements (BK/BK5/src/bk5Run.cpp:54-89);
// BPn split the same elements into two launches (BP/BP5/src/bp5Operator.cpp:
// 29-85), which writes the same AqL.
void ax_local(const Problem& P, const double* q, double* AqL) {
  if (P.elementType != 12) return ax_local_other(P, q, AqL);
  const dlong Nel = P.Nelements;
#pragma omp parallel for schedule(static)
  for (dlong ee = 0; ee < Nel; ++ee) {
    const dlong e = P.elementList[ee];
    for (int f = 0; f < P.Nfields; ++f) {
      switch (P.bp) {
        case 1: case 2: bp1_element(P, e, f, q, AqL); break;
        case 3: case 4: bp3_element(P, e, f, q, AqL); break;
        default: bp5_element(P, e, f, q, AqL); break;
      }
    }
  }
}

// libs/ogs/okl/ogsKernels.okl:86-122 (gather, Add): rows summed from 0 in
// ascending colId order.
void ogs_gather(const Problem& P, const double* vL, double* v) {
#pragma omp parallel for schedule(static)
  for (dlong r = 0; r < P.Ngather; ++r) {
    double gq = 0.0;
    for (dlong g = P.rowStartsT[r]; g < P.rowStartsT[r + 1]; ++g) gq += vL[P.colIdsT[g]];
    v[r] = gq;
  }
}

// libs/ogs/okl/ogsKernels.okl:127-164 (scatter); masked entries left as set.
void ogs_scatter(const Problem& P, const double* v, double* vL) {
#pragma omp parallel for schedule(static)
  for (dlong r = 0; r < P.Ngather; ++r)
    for (dlong g = P.rowStartsT[r]; g < P.rowStartsT[r + 1]; ++g) vL[P.colIdsT[g]] = v[r];
}

// okl/rhsHex3D.okl:28-128 (forcing times wJ at the local nodes)
void rhs_forcing_local(const Problem& P, double* rhsL) {
  if (P.elementType != 12) return rhs_forcing_local_other(P, rhsL);
  const dlong Nel = P.Nelements;
  const int Nq = P.Nq, Np = P.Np, Nf = P.Nfields;
  const double PI = M_PI, lambda = P.lambda;
#pragma omp parallel for
  for (dlong e = 0; e < Nel; ++e)
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int n = i + j * Nq + k * Nq * Nq;
          const size_t id = (size_t)e * Np + n;
          const double W = P.affine ? P.wJ[e] * P.gllw[i] * P.gllw[j] * P.gllw[k] : P.wJ[id];
          const double X = P.x[id], Y = P.y[id], Z = P.z[id];
          if (Nf == 1) {
            const double f = (3 * PI * PI + lambda) * std::sin(PI * X) * std::sin(PI * Y) * std::sin(PI * Z);
            rhsL[id] = W * f;
          } else {
            const double c = (4 * PI * PI + lambda);
            const double fx = c * std::sin(2.0 * PI * X) * std::sin(PI * Y) * std::sin(PI * Z);
            const double fy = c * std::sin(PI * X) * std::sin(2.0 * PI * Y) * std::sin(PI * Z);
            const double fz = c * std::sin(PI * X) * std::sin(PI * Y) * std::sin(2.0 * PI * Z);
            const size_t base = (size_t)e * Np * Nf + n;
            rhsL[base + 0 * Np] = W * fx;
            rhsL[base + 1 * Np] = W * fy;
            rhsL[base + 2 * Np] = W * fz;
          }
        }
}

}  // namespace orc

// Oracle (test infrastructure): variable-coefficient BP5 and the adjoint
// gradient. No reference counterpart: the reference's BP5 has a constant
// coefficient (okl/bp5AxHex3D.okl:415-569). This file restates the
// extension's definition on top of the reference's operator:
//
//   A(kappa) u = sum_e Z_e^T [ D^T (kappa_n G_n) D + lambda wJ_n ] Z_e u,
//
// kappa a nodal field on the box's node lattice (one value per lattice
// point, shared by every element copy of the node), scaling the geometric
// factors G_n of okl/bp5AxHex3D.okl at each GLL node n. For
// L(u(kappa)) with A(kappa) u = b and the adjoint A lam = dL/du (A is
// symmetric), dL/dkappa_p = -lam^T (dA/dkappa_p) u
//                         = -sum_{local copies l of p} (D lam)_l . G_l (D u)_l.
#include <vector>

#include "oracle.hpp"

namespace orc {

// Lattice point of local node n of element e. Elements are numbered
// e = ex + nx*(ey + ny*ez) (libs/mesh/meshSetupBoxHex3D.cpp:31-131), nodes
// n = a + b*Nq + c*Nq^2; the lattice has (nx*N+1) x (ny*N+1) x (nz*N+1) points.
hlong lattice_index(const Problem& P, dlong e, int n) {
  const int N = P.N, Nq = P.Nq;
  const hlong ex = e % P.nx, ey = (e / P.nx) % P.ny, ez = e / ((hlong)P.nx * P.ny);
  const int a = n % Nq, b = (n / Nq) % Nq, c = n / (Nq * Nq);
  const hlong LX = (hlong)P.nx * N + 1, LY = (hlong)P.ny * N + 1;
  return (ex * N + a) + LX * ((ey * N + b) + LY * (ez * N + c));
}

hlong lattice_nodes(const Problem& P) {
  return ((hlong)P.nx * P.N + 1) * ((hlong)P.ny * P.N + 1) * ((hlong)P.nz * P.N + 1);
}

int set_kappa(Problem& P, const double* kappa) {
  if (!kappa) {
    P.kappaL.clear();
    return 0;
  }
  if (P.bp != 5) {
    P.error = "kappa: BP5 only";
    return -1;
  }
  P.kappaL.assign((size_t)P.NLocal, 0.0);
  for (dlong e = 0; e < P.Nelements; ++e)
    for (int n = 0; n < P.Np; ++n) P.kappaL[(size_t)e * P.Np + n] = kappa[lattice_index(P, e, n)];
  return 0;
}

// gradL[e*Np + n] = -(D lam)_n . G_n (D u)_n with G_n the unscaled factors
// (affine: J G_e w_a w_b w_c), u and lam global (masked nodes read as 0).
void kappa_grad_local(const Problem& P, const double* u, const double* lam, double* gradL) {
  const int Nq = P.Nq, Np = P.Np;
  const double* D = P.D.data();
#pragma omp parallel for schedule(static)
  for (dlong e = 0; e < P.Nelements; ++e) {
    std::vector<double> su(Np), sl(Np);
    for (int n = 0; n < Np; ++n) {
      const dlong id = P.GlobalToLocal[(size_t)e * Np + n];
      su[n] = id >= 0 ? u[id] : 0.0;
      sl[n] = id >= 0 ? lam[id] : 0.0;
    }
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int n = i + j * Nq + k * Nq * Nq;
          double ur = 0, us = 0, ut = 0, lr = 0, ls = 0, lt = 0;
          for (int m = 0; m < Nq; ++m) {
            ur += D[i * Nq + m] * su[m + j * Nq + k * Nq * Nq];
            us += D[j * Nq + m] * su[i + m * Nq + k * Nq * Nq];
            ut += D[k * Nq + m] * su[i + j * Nq + m * Nq * Nq];
            lr += D[i * Nq + m] * sl[m + j * Nq + k * Nq * Nq];
            ls += D[j * Nq + m] * sl[i + m * Nq + k * Nq * Nq];
            lt += D[k * Nq + m] * sl[i + j * Nq + m * Nq * Nq];
          }
          double G00, G01, G02, G11, G12, G22;
          if (P.affine) {
            const double* g = &P.ggeo[(size_t)6 * e];
            const double w = P.gllw[i] * P.gllw[j] * P.gllw[k];
            G00 = g[0] * w; G01 = g[1] * w; G02 = g[2] * w; G11 = g[3] * w; G12 = g[4] * w; G22 = g[5] * w;
          } else {
            const double* g = &P.ggeo[6 * ((size_t)e * Np + n)];
            G00 = g[0]; G01 = g[1]; G02 = g[2]; G11 = g[3]; G12 = g[4]; G22 = g[5];
          }
          const double gr = G00 * ur + G01 * us + G02 * ut;
          const double gs = G01 * ur + G11 * us + G12 * ut;
          const double gt = G02 * ur + G12 * us + G22 * ut;
          gradL[(size_t)e * Np + n] = -(lr * gr + ls * gs + lt * gt);
        }
  }
}

// v[p] = sum of vL over the local copies of lattice point p, ascending local index
void lattice_sum(const Problem& P, const double* vL, double* v) {
  const hlong NL = lattice_nodes(P);
  for (hlong p = 0; p < NL; ++p) v[p] = 0.0;
  for (dlong e = 0; e < P.Nelements; ++e)
    for (int n = 0; n < P.Np; ++n) v[lattice_index(P, e, n)] += vL[(size_t)e * P.Np + n];
}

}  // namespace orc

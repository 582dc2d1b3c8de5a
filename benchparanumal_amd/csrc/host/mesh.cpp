// Hex box mesh on one GPU's rank: vertices, nodal coordinates, global node
// numbering, geometric and cubature factors.
//
// Reference behaviour followed:
//  - box vertices and element numbering: libs/mesh/meshSetupBoxHex3D.cpp:31-131
//    (element e = i + j*nx + k*nx*ny, vertex order of meshSetElementType.cpp)
//  - trilinear nodal coordinates: libs/mesh/meshPhysicalNodesHex3D.cpp:31-116
//  - geometric factors: libs/mesh/meshGeometricFactorsHex3D.cpp:31-153
//    (6 entries of G = J w (grad r)(grad r)^T per node plus wJ, or J and J G
//    per element on affine meshes; J < 1e-12 (non-affine) / J < 0 (affine)
//    is an error)
//  - cubature factors: libs/mesh/meshCubatureSetupHex3D.cpp:97-257
//    (coordinate derivatives interpolated to the p+2 Gauss points)
// The global node ids are the closed form of the structured lattice instead
// of the reference's iterated face-min propagation
// (libs/mesh/meshConnectNodes.cpp:32-82); only the grouping they induce
// matters to ogs, and it is the same.
#include <omp.h>

#include <algorithm>
#include <cmath>

#include "basis.hpp"
#include "bpx.hpp"
#include "geometry.hpp"
#include "rng.hpp"

namespace bpx {

namespace {

// reference coordinates of the 8 vertices of an element, in vertex order
const double kVr[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, kVs[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
             kVt[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

}  // namespace

void mesh_t::TrilinearCoefficients(std::vector<double>& c) const {
  c.assign((size_t)Nelements * 24, 0.0);
  for (dlong e = 0; e < Nelements; ++e)
    for (int d = 0; d < 3; ++d) {
      const double* V = (d == 0 ? EX.data() : d == 1 ? EY.data() : EZ.data()) + (size_t)e * 8;
      double* ce = &c[(size_t)e * 24 + 8 * d];
      for (int v = 0; v < 8; ++v) {
        const double X = 0.125 * V[v];
        ce[0] += kVr[v] * X;
        ce[1] += kVs[v] * X;
        ce[2] += kVt[v] * X;
        ce[3] += kVr[v] * kVs[v] * X;
        ce[4] += kVr[v] * kVt[v] * X;
        ce[5] += kVs[v] * kVt[v] * X;
        ce[6] += kVr[v] * kVs[v] * kVt[v] * X;
      }
    }
}

bool mesh_t::NodesAreTrilinear() const {
  std::vector<double> c;
  TrilinearCoefficients(c);
  const double scale = std::max(opts.dim[0], std::max(opts.dim[1], opts.dim[2]));
  const double tol = 1e-12 * scale;
  bool ok = true;
#pragma omp parallel for reduction(&& : ok)
  for (dlong e = 0; e < Nelements; ++e) {
    const double* ce = &c[(size_t)e * 24];
    for (int k = 0; k < Nq && ok; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const double r = gllz[i], s = gllz[j], t = gllz[k];
          const size_t id = (size_t)e * Np + i + j * Nq + k * Nq * Nq;
          const double* X[3] = {x.data(), y.data(), z.data()};
          for (int d = 0; d < 3; ++d) {
            const double* cd = ce + 8 * d;
            // the constant term: the mean of the 8 vertices
            const double* V = (d == 0 ? EX.data() : d == 1 ? EY.data() : EZ.data()) + (size_t)e * 8;
            double c0 = 0;
            for (int v = 0; v < 8; ++v) c0 += 0.125 * V[v];
            const double m = c0 + cd[0] * r + cd[1] * s + cd[2] * t + cd[3] * r * s + cd[4] * r * t +
                             cd[5] * s * t + cd[6] * r * s * t;
            if (std::fabs(m - X[d][id]) > tol) ok = false;
          }
        }
  }
  return ok;
}

int mesh_t::SetupVertices(const MeshOpts& o) {
  if (o.elementType != 0 && o.elementType != ELEM_HEX)
    return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: the vertex setup shared with the device covers hexahedra");
  elementType = ELEM_HEX;
  dim = 3;
  Nverts = 8;
  Nggeo = 6;
  boundary.clear();
  if (o.N < 1 || o.N > 15) return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: degree N must be in 1..15");
  if (o.nx < 1 || o.ny < 1 || o.nz < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: nx, ny, nz must be >= 1");
  if (!(o.dim[0] > 0 && o.dim[1] > 0 && o.dim[2] > 0))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: box size must be positive");
  if (!(o.perturb >= 0 && o.perturb < 0.5))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: perturb must be in [0, 0.5)");
  const long long nel = (long long)o.nx * o.ny * o.nz;
  const long long nlocal = nel * (o.N + 1) * (o.N + 1) * (o.N + 1);
  if (nlocal * 6 > 0x7fffffffLL)
    return set_error(BPX_ERR_UNSUPPORTED, "mesh: local node count exceeds 32-bit indexing");
  opts = o;
  N = o.N;
  Nq = N + 1;
  Np = Nq * Nq * Nq;
  Nelements = (dlong)nel;

  onDevice = false;
  basis::gll(N, gllz, gllw);
  basis::dmatrix_gll(N, gllz, D);
  basis::mass_gll(N, gllz, MM);
  // 1D operators of the cubature (BP1-BP4)
  cubNq = Nq + 1;
  cubNp = cubNq * cubNq * cubNq;
  cubr.clear(); cubw.clear(); cubInterp.clear(); cubD.clear(); invV.clear();
  if (o.cubature) {
    basis::gauss(cubNq, cubr, cubw);
    basis::interp(gllz, cubr, cubInterp);
    if (o.affine) {
      basis::inv_vandermonde(N, gllz, invV);
      basis::modal_dmatrix(N, cubD);
    } else {
      basis::dmatrix(cubr, cubD);
    }
  }
  x.clear(); y.clear(); z.clear(); globalIds.clear();
  wJ.clear(); ggeo.clear(); cubwJ.clear(); cubggeo.clear();

  // ---- vertices (meshSetupBoxHex3D.cpp:88-126) + seeded perturbation
  const int nx = o.nx, ny = o.ny, nz = o.nz;
  const long long NnX = nx + 1, NnY = ny + 1;
  EX.assign((size_t)Nelements * 8, 0.0);
  EY.assign((size_t)Nelements * 8, 0.0);
  EZ.assign((size_t)Nelements * 8, 0.0);
  const double dx = o.dim[0] / nx, dy = o.dim[1] / ny, dz = o.dim[2] / nz;
  static const int vi[8] = {0, 1, 1, 0, 0, 1, 1, 0}, vj[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                   vk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
#pragma omp parallel for collapse(3)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t e = (size_t)i + (size_t)j * nx + (size_t)k * nx * ny;
        const double x0 = o.lo[0] + dx * i, y0 = o.lo[1] + dy * j, z0 = o.lo[2] + dz * k;
        for (int v = 0; v < 8; ++v) {
          double X = vi[v] ? x0 + dx : x0, Y = vj[v] ? y0 + dy : y0, Z = vk[v] ? z0 + dz : z0;
          const long long I = i + vi[v], Jv = j + vj[v], K = k + vk[v];
          const bool interior = I > 0 && I < nx && Jv > 0 && Jv < ny && K > 0 && K < nz;
          if (o.perturb != 0.0 && interior) {
            const uint64_t vid = (uint64_t)(I + Jv * NnX + K * NnX * NnY);
            X += o.perturb * dx * (2.0 * rng::u01(o.seed, 3 * vid + 0) - 1.0);
            Y += o.perturb * dy * (2.0 * rng::u01(o.seed, 3 * vid + 1) - 1.0);
            Z += o.perturb * dz * (2.0 * rng::u01(o.seed, 3 * vid + 2) - 1.0);
          }
          EX[e * 8 + v] = X;
          EY[e * 8 + v] = Y;
          EZ[e * 8 + v] = Z;
        }
      }
  return BPX_OK;
}

int mesh_t::Setup(const MeshOpts& o) {
  if (o.elementType != 0 && o.elementType != ELEM_HEX) return SetupOther(o);
  if (int st = SetupVertices(o)) return st;
  const int nx = o.nx, ny = o.ny;

  // ---- nodal coordinates (trilinear map)
  const size_t NL = (size_t)Nelements * Np;
  x.assign(NL, 0.0);
  y.assign(NL, 0.0);
  z.assign(NL, 0.0);
#pragma omp parallel for
  for (dlong e = 0; e < Nelements; ++e)
    for (int n = 0; n < Np; ++n) {
      const double r = gllz[n % Nq], s = gllz[(n / Nq) % Nq], t = gllz[n / (Nq * Nq)];
      double X = 0, Y = 0, Z = 0;
      for (int v = 0; v < 8; ++v) {
        const double c = vertex_weight(v, r, s, t);
        X += c * EX[(size_t)e * 8 + v];
        Y += c * EY[(size_t)e * 8 + v];
        Z += c * EZ[(size_t)e * 8 + v];
      }
      x[(size_t)e * Np + n] = X;
      y[(size_t)e * Np + n] = Y;
      z[(size_t)e * Np + n] = Z;
    }

  // ---- global ids: lattice position of each node
  globalIds.assign(NL, 0);
  const long long LX = (long long)nx * N + 1, LY = (long long)ny * N + 1;
#pragma omp parallel for
  for (dlong e = 0; e < Nelements; ++e) {
    const long long ex = e % nx, ey = (e / nx) % ny, ez = e / ((long long)nx * ny);
    for (int n = 0; n < Np; ++n) {
      const long long X = ex * N + n % Nq, Y = ey * N + (n / Nq) % Nq, Z = ez * N + n / (Nq * Nq);
      globalIds[(size_t)e * Np + n] = 1 + X + Y * LX + Z * LX * LY;
    }
  }

  // ---- geometric factors
  int bad = 0;
  if (o.affine) {
    wJ.assign(Nelements, 0.0);
    ggeo.assign((size_t)Nelements * 6, 0.0);
#pragma omp parallel for reduction(| : bad)
    for (dlong e = 0; e < Nelements; ++e) {
      const size_t b = (size_t)e * 8;
      Geo9 d;
      d.xr = 0.5 * (EX[b + 1] - EX[b]); d.xs = 0.5 * (EX[b + 3] - EX[b]); d.xt = 0.5 * (EX[b + 4] - EX[b]);
      d.yr = 0.5 * (EY[b + 1] - EY[b]); d.ys = 0.5 * (EY[b + 3] - EY[b]); d.yt = 0.5 * (EY[b + 4] - EY[b]);
      d.zr = 0.5 * (EZ[b + 1] - EZ[b]); d.zs = 0.5 * (EZ[b + 3] - EZ[b]); d.zt = 0.5 * (EZ[b + 4] - EZ[b]);
      double J, g[6];
      metric(d, J, g);
      if (!(J >= 0)) bad = 1;
      wJ[e] = J;
      for (int c = 0; c < 6; ++c) ggeo[(size_t)e * 6 + c] = g[c];
    }
  } else {
    wJ.assign(NL, 0.0);
    ggeo.assign(NL * 6, 0.0);
#pragma omp parallel for reduction(| : bad)
    for (dlong e = 0; e < Nelements; ++e) {
      const double* X = &x[(size_t)e * Np];
      const double* Y = &y[(size_t)e * Np];
      const double* Z = &z[(size_t)e * Np];
      for (int n = 0; n < Np; ++n) {
        const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
        Geo9 d{0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int m = 0; m < Nq; ++m) {
          const int ir = k * Nq * Nq + j * Nq + m, is = k * Nq * Nq + m * Nq + i, it = m * Nq * Nq + j * Nq + i;
          const double Di = D[i * Nq + m], Dj = D[j * Nq + m], Dk = D[k * Nq + m];
          d.xr += Di * X[ir]; d.xs += Dj * X[is]; d.xt += Dk * X[it];
          d.yr += Di * Y[ir]; d.ys += Dj * Y[is]; d.yt += Dk * Y[it];
          d.zr += Di * Z[ir]; d.zs += Dj * Z[is]; d.zt += Dk * Z[it];
        }
        double J, g[6];
        metric(d, J, g);
        if (!(J >= 1e-12)) bad = 1;
        const double w = gllw[i] * gllw[j] * gllw[k];
        wJ[(size_t)e * Np + n] = J * w;
        for (int c = 0; c < 6; ++c) ggeo[((size_t)e * Np + n) * 6 + c] = g[c] * w;
      }
    }
  }
  if (bad) return set_error(BPX_ERR_NEGATIVE_JACOBIAN, "mesh: negative or vanishing Jacobian found");

  // ---- cubature factors (BP1-BP4, non-affine)
  if (o.cubature && !o.affine) {
    {
      const int cq = cubNq;
      cubwJ.assign((size_t)Nelements * cubNp, 0.0);
      cubggeo.assign((size_t)Nelements * cubNp * 6, 0.0);
      const std::vector<double>& I = cubInterp;
#pragma omp parallel for reduction(| : bad)
      for (dlong e = 0; e < Nelements; ++e) {
        // 9 coordinate derivatives at the GLL nodes
        std::vector<double> d0(9 * (size_t)Np), d1(9 * (size_t)Nq * Nq * cq), d2(9 * (size_t)Nq * cq * cq);
        const double* C[3] = {&x[(size_t)e * Np], &y[(size_t)e * Np], &z[(size_t)e * Np]};
        for (int n = 0; n < Np; ++n) {
          const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
          for (int c = 0; c < 3; ++c) {
            double ar = 0, as = 0, at = 0;
            for (int m = 0; m < Nq; ++m) {
              ar += D[i * Nq + m] * C[c][k * Nq * Nq + j * Nq + m];
              as += D[j * Nq + m] * C[c][k * Nq * Nq + m * Nq + i];
              at += D[k * Nq + m] * C[c][m * Nq * Nq + j * Nq + i];
            }
            d0[(size_t)(3 * c + 0) * Np + n] = ar;
            d0[(size_t)(3 * c + 1) * Np + n] = as;
            d0[(size_t)(3 * c + 2) * Np + n] = at;
          }
        }
        // interpolate to the Gauss points, direction r, then s, then t
        for (int f = 0; f < 9; ++f) {
          const double* a0 = &d0[(size_t)f * Np];
          double* a1 = &d1[(size_t)f * Nq * Nq * cq];
          double* a2 = &d2[(size_t)f * Nq * cq * cq];
          for (int k = 0; k < Nq; ++k)
            for (int j = 0; j < Nq; ++j)
              for (int i = 0; i < cq; ++i) {
                double acc = 0;
                for (int m = 0; m < Nq; ++m) acc += I[i * Nq + m] * a0[k * Nq * Nq + j * Nq + m];
                a1[(k * Nq + j) * cq + i] = acc;
              }
          for (int k = 0; k < Nq; ++k)
            for (int j = 0; j < cq; ++j)
              for (int i = 0; i < cq; ++i) {
                double acc = 0;
                for (int m = 0; m < Nq; ++m) acc += I[j * Nq + m] * a1[(k * Nq + m) * cq + i];
                a2[(k * cq + j) * cq + i] = acc;
              }
        }
        for (int k = 0; k < cq; ++k)
          for (int j = 0; j < cq; ++j)
            for (int i = 0; i < cq; ++i) {
              double v[9];
              for (int f = 0; f < 9; ++f) {
                const double* a2 = &d2[(size_t)f * Nq * cq * cq];
                double acc = 0;
                for (int m = 0; m < Nq; ++m) acc += I[k * Nq + m] * a2[(m * cq + j) * cq + i];
                v[f] = acc;
              }
              const Geo9 d{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]};
              double J, g[6];
              metric(d, J, g);
              if (!(J >= 1e-8)) bad = 1;
              const double w = cubw[i] * cubw[j] * cubw[k];
              const size_t cid = (size_t)e * cubNp + (size_t)(k * cq + j) * cq + i;
              cubwJ[cid] = J * w;
              for (int c = 0; c < 6; ++c) cubggeo[cid * 6 + c] = g[c] * w;
            }
      }
      if (bad) return set_error(BPX_ERR_NEGATIVE_JACOBIAN, "mesh: negative Jacobian at a cubature point");
    }
  }
  return BPX_OK;
}

void mesh_t::MaskedGlobalIds(int Nfields, std::vector<hlong>& ids) const {
  const size_t NL = (size_t)Nelements * Np;
  ids.assign(NL * Nfields, 0);
  if (elementType != ELEM_HEX) {  // boundary flags per node
#pragma omp parallel for
    for (dlong e = 0; e < Nelements; ++e)
      for (int n = 0; n < Np; ++n) {
        const size_t l = (size_t)e * Np + n;
        for (int f = 0; f < Nfields; ++f)
          ids[(size_t)e * Np * Nfields + (size_t)f * Np + n] = boundary[l] ? 0 : globalIds[l] * Nfields + f;
      }
    return;
  }
  const long long LX = (long long)opts.nx * N + 1, LY = (long long)opts.ny * N + 1,
                  LZ = (long long)opts.nz * N + 1;
#pragma omp parallel for
  for (dlong e = 0; e < Nelements; ++e)
    for (int n = 0; n < Np; ++n) {
      const hlong g = globalIds[(size_t)e * Np + n] - 1;
      const long long X = g % LX, Y = (g / LX) % LY, Z = g / (LX * LY);
      const bool boundary = X == 0 || X == LX - 1 || Y == 0 || Y == LY - 1 || Z == 0 || Z == LZ - 1;
      for (int f = 0; f < Nfields; ++f)
        ids[(size_t)e * Np * Nfields + (size_t)f * Np + n] = boundary ? 0 : (g + 1) * Nfields + f;
    }
}

}  // namespace bpx

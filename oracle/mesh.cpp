// ORACLE — test infrastructure only (see oracle.hpp).
// Restates the hex setup path of libs/mesh and the single-rank T-map
// construction of libs/ogs (ogsSetup.cpp, ogsOperator.cpp).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <map>
#include <numeric>

#include "oracle.hpp"

namespace orc {

// libs/mesh/meshSetElementType.cpp:84-86
static const int kFaceVertices[6][4] = {{0, 1, 2, 3}, {0, 4, 5, 1}, {1, 5, 6, 2},
                                        {2, 6, 7, 3}, {0, 3, 7, 4}, {4, 7, 6, 5}};

// libs/mesh/meshReferenceNodesHex3D.cpp:31-70, meshBasisHex3D.cpp:34-93
static int ReferenceNodes(Problem& P) {
  P.Nq = P.N + 1;
  P.Nfp = P.Nq * P.Nq;
  P.Np = P.Nq * P.Nq * P.Nq;
  const int Nq = P.Nq, Np = P.Np, Nfp = P.Nfp;
  P.gllz.resize(Nq);
  P.gllw.resize(Nq);
  if (JacobiGLL(P.N, P.gllz.data(), P.gllw.data())) return -1;
  P.r.resize(Np);
  P.s.resize(Np);
  P.t.resize(Np);
  for (int k = 0; k < Nq; k++)
    for (int j = 0; j < Nq; j++)
      for (int i = 0; i < Nq; i++) {
        P.r[i + j * Nq + k * Nq * Nq] = P.gllz[i];
        P.s[i + j * Nq + k * Nq * Nq] = P.gllz[j];
        P.t[i + j * Nq + k * Nq * Nq] = P.gllz[k];
      }
  double deps = 1.;
  while ((1. + deps) > 1.) deps *= 0.5;
  const double NODETOL = 1000. * deps;
  P.faceNodes.assign(6 * Nfp, 0);
  int cnt[6] = {0, 0, 0, 0, 0, 0};
  for (int n = 0; n < Np; n++) {
    if (std::fabs(P.t[n] + 1) < NODETOL) P.faceNodes[0 * Nfp + (cnt[0]++)] = n;
    if (std::fabs(P.s[n] + 1) < NODETOL) P.faceNodes[1 * Nfp + (cnt[1]++)] = n;
    if (std::fabs(P.r[n] - 1) < NODETOL) P.faceNodes[2 * Nfp + (cnt[2]++)] = n;
    if (std::fabs(P.s[n] - 1) < NODETOL) P.faceNodes[3 * Nfp + (cnt[3]++)] = n;
    if (std::fabs(P.r[n] + 1) < NODETOL) P.faceNodes[4 * Nfp + (cnt[4]++)] = n;
    if (std::fabs(P.t[n] - 1) < NODETOL) P.faceNodes[5 * Nfp + (cnt[5]++)] = n;
  }
  P.V.resize(Nq * Nq);
  Vandermonde1D(P.N, Nq, P.gllz.data(), P.V.data());
  P.MM.resize(Nq * Nq);
  if (MassMatrix1D(Nq, P.V.data(), P.MM.data())) return -1;
  P.D.resize(Nq * Nq);
  if (Dmatrix1D(P.N, Nq, P.gllz.data(), P.gllz.data(), P.D.data())) return -1;
  return 0;
}

// libs/mesh/meshSetupBoxHex3D.cpp:31-131 on one rank. The seeded vertex
// perturbation has no reference counterpart: interior vertex (I,J,K) moves
// by perturb*h_c*(2u-1) in coordinate c, u = rng_u01(seed, 3*vid+c), vid
// the global vertex number. Boundary vertices stay fixed.
static void SetupBox(Problem& P) {
  const int nx = P.nx, ny = P.ny, nz = P.nz;
  const hlong NX = nx, NY = ny, NZ = nz;
  const hlong NnX = NX + 1, NnY = NY + 1, NnZ = NZ + 1;
  P.Nelements = nx * ny * nz;
  P.EToV.resize((size_t)P.Nelements * 8);
  P.EX.resize((size_t)P.Nelements * 8);
  P.EY.resize((size_t)P.Nelements * 8);
  P.EZ.resize((size_t)P.Nelements * 8);
  const double X0 = P.lo[0], Y0 = P.lo[1], Z0 = P.lo[2];
  const double dx = P.dim[0] / nx, dy = P.dim[1] / ny, dz = P.dim[2] / nz;
#pragma omp parallel for collapse(3)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const dlong e = i + j * nx + k * nx * ny;
        const hlong i0 = i, i1 = (i + 1) % NnX, j0 = j, j1 = (j + 1) % NnY;
        const hlong k0 = k, k1 = (k + 1) % NnZ;
        hlong* v = P.EToV.data() + (size_t)e * 8;
        v[0] = i0 + j0 * NnX + k0 * NnX * NnY;
        v[1] = i1 + j0 * NnX + k0 * NnX * NnY;
        v[2] = i1 + j1 * NnX + k0 * NnX * NnY;
        v[3] = i0 + j1 * NnX + k0 * NnX * NnY;
        v[4] = i0 + j0 * NnX + k1 * NnX * NnY;
        v[5] = i1 + j0 * NnX + k1 * NnX * NnY;
        v[6] = i1 + j1 * NnX + k1 * NnX * NnY;
        v[7] = i0 + j1 * NnX + k1 * NnX * NnY;
        const double x0 = X0 + dx * i, y0 = Y0 + dy * j, z0 = Z0 + dz * k;
        double* ex = P.EX.data() + (size_t)e * 8;
        double* ey = P.EY.data() + (size_t)e * 8;
        double* ez = P.EZ.data() + (size_t)e * 8;
        ex[0] = x0;      ey[0] = y0;      ez[0] = z0;
        ex[1] = x0 + dx; ey[1] = y0;      ez[1] = z0;
        ex[2] = x0 + dx; ey[2] = y0 + dy; ez[2] = z0;
        ex[3] = x0;      ey[3] = y0 + dy; ez[3] = z0;
        ex[4] = x0;      ey[4] = y0;      ez[4] = z0 + dz;
        ex[5] = x0 + dx; ey[5] = y0;      ez[5] = z0 + dz;
        ex[6] = x0 + dx; ey[6] = y0 + dy; ez[6] = z0 + dz;
        ex[7] = x0;      ey[7] = y0 + dy; ez[7] = z0 + dz;
        if (P.perturb != 0.0) {
          for (int n = 0; n < 8; ++n) {
            const hlong vid = v[n];
            const hlong I = vid % NnX, J = (vid / NnX) % NnY, K = vid / (NnX * NnY);
            if (I == 0 || I == NX || J == 0 || J == NY || K == 0 || K == NZ) continue;
            ex[n] += P.perturb * dx * (2.0 * rng_u01(P.seed, 3 * (uint64_t)vid + 0) - 1.0);
            ey[n] += P.perturb * dy * (2.0 * rng_u01(P.seed, 3 * (uint64_t)vid + 1) - 1.0);
            ez[n] += P.perturb * dz * (2.0 * rng_u01(P.seed, 3 * (uint64_t)vid + 2) - 1.0);
          }
        }
      }
}

// libs/mesh/meshConnect.cpp:51-292 on one rank: faces are matched by their
// sorted vertex ids.
static void Connect(Problem& P) {
  const dlong Nel = P.Nelements;
  P.EToE.assign((size_t)Nel * 6, -1);
  P.EToF.assign((size_t)Nel * 6, -1);
  std::map<std::array<hlong, 4>, std::pair<dlong, int>> open;
  for (dlong e = 0; e < Nel; ++e)
    for (int f = 0; f < 6; ++f) {
      std::array<hlong, 4> key;
      for (int n = 0; n < 4; ++n) key[n] = P.EToV[(size_t)e * 8 + kFaceVertices[f][n]];
      std::sort(key.begin(), key.end());
      auto it = open.find(key);
      if (it == open.end()) {
        open.emplace(key, std::make_pair(e, f));
      } else {
        const dlong eP = it->second.first;
        const int fP = it->second.second;
        P.EToE[(size_t)e * 6 + f] = eP;
        P.EToF[(size_t)e * 6 + f] = fP;
        P.EToE[(size_t)eP * 6 + fP] = e;
        P.EToF[(size_t)eP * 6 + fP] = f;
        open.erase(it);
      }
    }
}

// libs/mesh/meshConnectFaceNodes.cpp:32-94. Every box element has the same
// vertex orientation, so the matching node on the neighbour face is the one
// at the same tangential (i,j,k) position with the normal index mirrored
// (0 <-> N); tests check the match against the physical coordinates.
static void ConnectFaceNodes(Problem& P) {
  const dlong Nel = P.Nelements;
  const int Nq = P.Nq, Np = P.Np, Nfp = P.Nfp;
  P.vmapM.resize((size_t)Nel * 6 * Nfp);
  P.vmapP.resize((size_t)Nel * 6 * Nfp);
  // normal direction of each face: 0:t 1:s 2:r 3:s 4:r 5:t
  const int normal[6] = {2, 1, 0, 1, 0, 2};
#pragma omp parallel for
  for (dlong eM = 0; eM < Nel; ++eM)
    for (int fM = 0; fM < 6; ++fM) {
      const dlong eP = P.EToE[(size_t)eM * 6 + fM];
      for (int nM = 0; nM < Nfp; ++nM) {
        const int idM = P.faceNodes[fM * Nfp + nM];
        const size_t id = (size_t)eM * 6 * Nfp + fM * Nfp + nM;
        P.vmapM[id] = idM + eM * Np;
        if (eP < 0) {
          P.vmapP[id] = idM + eM * Np;
        } else {
          int ijk[3] = {idM % Nq, (idM / Nq) % Nq, idM / (Nq * Nq)};
          ijk[normal[fM]] = P.N - ijk[normal[fM]];
          const int idP = ijk[0] + ijk[1] * Nq + ijk[2] * Nq * Nq;
          P.vmapP[id] = idP + eP * Np;
        }
      }
    }
}

// libs/mesh/meshConnectNodes.cpp:32-82 (iterated min over face traces).
static void ConnectNodes(Problem& P) {
  const dlong Nel = P.Nelements;
  const int Np = P.Np, Nfp = P.Nfp;
  P.globalIds.resize((size_t)Nel * Np);
  for (dlong e = 0; e < Nel; ++e)
    for (int n = 0; n < Np; ++n) {
      const dlong id = e * Np + n;
      P.globalIds[id] = 1 + (hlong)id;
    }
  hlong gatherChange = 1;
  while (gatherChange > 0) {
    gatherChange = 0;
    for (dlong e = 0; e < Nel; ++e)
      for (int n = 0; n < Nfp * 6; ++n) {
        const size_t id = (size_t)e * Nfp * 6 + n;
        const dlong idM = P.vmapM[id], idP = P.vmapP[id];
        const hlong gidM = P.globalIds[idM], gidP = P.globalIds[idP];
        if (gidP < gidM) {
          ++gatherChange;
          P.globalIds[idM] = gidP;
        }
      }
  }
}

// libs/mesh/meshPhysicalNodesHex3D.cpp:31-116 (trilinear map).
static void PhysicalNodes(Problem& P) {
  const dlong Nel = P.Nelements;
  const int Np = P.Np;
  P.x.resize((size_t)Nel * Np);
  P.y.resize((size_t)Nel * Np);
  P.z.resize((size_t)Nel * Np);
#pragma omp parallel for
  for (dlong e = 0; e < Nel; ++e) {
    const double* X = P.EX.data() + (size_t)e * 8;
    const double* Y = P.EY.data() + (size_t)e * 8;
    const double* Z = P.EZ.data() + (size_t)e * 8;
    for (int n = 0; n < Np; ++n) {
      const double rn = P.r[n], sn = P.s[n], tn = P.t[n];
      const double c[8] = {0.125 * (1 - rn) * (1 - sn) * (1 - tn),
                           0.125 * (1 + rn) * (1 - sn) * (1 - tn),
                           0.125 * (1 + rn) * (1 + sn) * (1 - tn),
                           0.125 * (1 - rn) * (1 + sn) * (1 - tn),
                           0.125 * (1 - rn) * (1 - sn) * (1 + tn),
                           0.125 * (1 + rn) * (1 - sn) * (1 + tn),
                           0.125 * (1 + rn) * (1 + sn) * (1 + tn),
                           0.125 * (1 - rn) * (1 + sn) * (1 + tn)};
      double xx = 0, yy = 0, zz = 0;
      for (int v = 0; v < 8; ++v) {
        xx += c[v] * X[v];
        yy += c[v] * Y[v];
        zz += c[v] * Z[v];
      }
      P.x[(size_t)e * Np + n] = xx;
      P.y[(size_t)e * Np + n] = yy;
      P.z[(size_t)e * Np + n] = zz;
    }
  }
}

// libs/mesh/meshGeometricFactorsHex3D.cpp:31-153
static int GeometricFactors(Problem& P) {
  const dlong Nel = P.Nelements;
  const int Nq = P.Nq, Np = P.Np;
  int bad = 0;
  if (P.affine) {
    P.wJ.resize(Nel);
    P.ggeo.resize((size_t)Nel * 6);
#pragma omp parallel for reduction(| : bad)
    for (dlong e = 0; e < Nel; ++e) {
      const size_t id = (size_t)e * 8;
      const double xe1 = P.EX[id + 0], ye1 = P.EY[id + 0], ze1 = P.EZ[id + 0];
      const double xe2 = P.EX[id + 1], ye2 = P.EY[id + 1], ze2 = P.EZ[id + 1];
      const double xe3 = P.EX[id + 3], ye3 = P.EY[id + 3], ze3 = P.EZ[id + 3];
      const double xe4 = P.EX[id + 4], ye4 = P.EY[id + 4], ze4 = P.EZ[id + 4];
      const double xr = 0.5 * (xe2 - xe1), xs = 0.5 * (xe3 - xe1), xt = 0.5 * (xe4 - xe1);
      const double yr = 0.5 * (ye2 - ye1), ys = 0.5 * (ye3 - ye1), yt = 0.5 * (ye4 - ye1);
      const double zr = 0.5 * (ze2 - ze1), zs = 0.5 * (ze3 - ze1), zt = 0.5 * (ze4 - ze1);
      const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
      const double rx = (ys * zt - zs * yt) / J, ry = -(xs * zt - zs * xt) / J, rz = (xs * yt - ys * xt) / J;
      const double sx = -(yr * zt - zr * yt) / J, sy = (xr * zt - zr * xt) / J, sz = -(xr * yt - yr * xt) / J;
      const double tx = (yr * zs - zr * ys) / J, ty = -(xr * zs - zr * xs) / J, tz = (xr * ys - yr * xs) / J;
      if (J < 0) bad = 1;
      P.wJ[e] = J;
      double* g = P.ggeo.data() + (size_t)6 * e;
      g[0] = J * (rx * rx + ry * ry + rz * rz);
      g[1] = J * (rx * sx + ry * sy + rz * sz);
      g[2] = J * (rx * tx + ry * ty + rz * tz);
      g[3] = J * (sx * sx + sy * sy + sz * sz);
      g[4] = J * (sx * tx + sy * ty + sz * tz);
      g[5] = J * (tx * tx + ty * ty + tz * tz);
    }
  } else {
    P.wJ.resize((size_t)Nel * Np);
    P.ggeo.resize((size_t)Nel * Np * 6);
#pragma omp parallel for reduction(| : bad)
    for (dlong e = 0; e < Nel; ++e)
      for (int k = 0; k < Nq; ++k)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i) {
            const int n = i + j * Nq + k * Nq * Nq;
            double xr = 0, xs = 0, xt = 0, yr = 0, ys = 0, yt = 0, zr = 0, zs = 0, zt = 0;
            for (int m = 0; m < Nq; ++m) {
              const size_t idr = (size_t)e * Np + k * Nq * Nq + j * Nq + m;
              const size_t ids = (size_t)e * Np + k * Nq * Nq + m * Nq + i;
              const size_t idt = (size_t)e * Np + m * Nq * Nq + j * Nq + i;
              xr += P.D[i * Nq + m] * P.x[idr];
              xs += P.D[j * Nq + m] * P.x[ids];
              xt += P.D[k * Nq + m] * P.x[idt];
              yr += P.D[i * Nq + m] * P.y[idr];
              ys += P.D[j * Nq + m] * P.y[ids];
              yt += P.D[k * Nq + m] * P.y[idt];
              zr += P.D[i * Nq + m] * P.z[idr];
              zs += P.D[j * Nq + m] * P.z[ids];
              zt += P.D[k * Nq + m] * P.z[idt];
            }
            const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
            if (J < 1e-12) bad = 1;
            const double rx = (ys * zt - zs * yt) / J, ry = -(xs * zt - zs * xt) / J, rz = (xs * yt - ys * xt) / J;
            const double sx = -(yr * zt - zr * yt) / J, sy = (xr * zt - zr * xt) / J, sz = -(xr * yt - yr * xt) / J;
            const double tx = (yr * zs - zr * ys) / J, ty = -(xr * zs - zr * xs) / J, tz = (xr * ys - yr * xs) / J;
            const double JW = J * P.gllw[i] * P.gllw[j] * P.gllw[k];
            P.wJ[(size_t)Np * e + n] = JW;
            double* g = P.ggeo.data() + (size_t)6 * (Np * (size_t)e + n);
            g[0] = JW * (rx * rx + ry * ry + rz * rz);
            g[1] = JW * (rx * sx + ry * sy + rz * sz);
            g[2] = JW * (rx * tx + ry * ty + rz * tz);
            g[3] = JW * (sx * sx + sy * sy + sz * sz);
            g[4] = JW * (sx * tx + sy * ty + sz * tz);
            g[5] = JW * (tx * tx + ty * ty + tz * tz);
          }
  }
  if (bad) {
    P.error = "Negative J found";
    return -2;
  }
  return 0;
}

// libs/mesh/meshCubatureSetupHex3D.cpp:31-260
static int CubatureSetup(Problem& P) {
  const int N = P.N, Nq = P.Nq, Np = P.Np;
  P.cubN = N + 1;
  P.cubNq = P.cubN + 1;
  P.cubNp = P.cubNq * P.cubNq * P.cubNq;
  const int cubNq = P.cubNq, cubNp = P.cubNp;
  P.cubr.resize(cubNq);
  P.cubw.resize(cubNq);
  JacobiGQ(0, 0, P.cubN, P.cubr.data(), P.cubw.data());
  P.cubInterp.resize(cubNq * Nq);
  if (InterpolationMatrix1D(N, cubNq, P.gllz.data(), P.cubr.data(), P.cubInterp.data())) return -1;
  P.cubProject.resize(cubNq * Nq);
  for (int i = 0; i < cubNq; ++i)
    for (int n = 0; n < Nq; ++n) P.cubProject[n * cubNq + i] = P.cubInterp[i * Nq + n];
  if (P.affine) {
    std::vector<double> Vr(Nq * Nq);
    GradVandermonde1D(N, Nq, P.gllz.data(), Vr.data());
    P.invV.resize(Nq * Nq);
    Vandermonde1D(N, Nq, P.gllz.data(), P.invV.data());
    if (MatrixInverse(Nq, P.invV.data())) return -1;
    P.cubD.resize(Nq * Nq);
    for (int n = 0; n < Nq; ++n)
      for (int m = 0; m < Nq; ++m) {
        double acc = 0;
        for (int k = 0; k < Nq; ++k) acc += P.invV[n * Nq + k] * Vr[k * Nq + m];
        P.cubD[n * Nq + m] = acc;
      }
    P.cubwJ = P.wJ;
    P.cubggeo = P.ggeo;
    return 0;
  }
  P.cubD.resize(cubNq * cubNq);
  if (Dmatrix1D(P.cubN, cubNq, P.cubr.data(), P.cubr.data(), P.cubD.data())) return -1;
  const dlong Nel = P.Nelements;
  P.cubwJ.resize((size_t)Nel * cubNp);
  P.cubggeo.resize((size_t)Nel * cubNp * 6);
  int bad = 0;
  const std::vector<double>& I = P.cubInterp;
#pragma omp parallel for reduction(| : bad)
  for (dlong e = 0; e < Nel; ++e) {
    std::vector<double> d0(9 * Np), d1(9 * Nq * Nq * cubNq), d2(9 * Nq * cubNq * cubNq);
    const double* X[3] = {&P.x[(size_t)e * Np], &P.y[(size_t)e * Np], &P.z[(size_t)e * Np]};
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int n = i + j * Nq + k * Nq * Nq;
          for (int c = 0; c < 3; ++c) {
            double ar = 0, as = 0, at = 0;
            for (int m = 0; m < Nq; ++m) {
              ar += P.D[i * Nq + m] * X[c][k * Nq * Nq + j * Nq + m];
              as += P.D[j * Nq + m] * X[c][k * Nq * Nq + m * Nq + i];
              at += P.D[k * Nq + m] * X[c][m * Nq * Nq + j * Nq + i];
            }
            d0[(3 * c + 0) * Np + n] = ar;
            d0[(3 * c + 1) * Np + n] = as;
            d0[(3 * c + 2) * Np + n] = at;
          }
        }
    for (int f = 0; f < 9; ++f) {
      const double* a0 = &d0[(size_t)f * Np];
      double* a1 = &d1[(size_t)f * Nq * Nq * cubNq];
      double* a2 = &d2[(size_t)f * Nq * cubNq * cubNq];
      for (int k = 0; k < Nq; ++k)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < cubNq; ++i) {
            double acc = 0.0;
            for (int n = 0; n < Nq; ++n) acc += I[i * Nq + n] * a0[k * Nq * Nq + j * Nq + n];
            a1[k * Nq * cubNq + j * cubNq + i] = acc;
          }
      for (int k = 0; k < Nq; ++k)
        for (int j = 0; j < cubNq; ++j)
          for (int i = 0; i < cubNq; ++i) {
            double acc = 0.0;
            for (int n = 0; n < Nq; ++n) acc += I[j * Nq + n] * a1[k * Nq * cubNq + n * cubNq + i];
            a2[k * cubNq * cubNq + j * cubNq + i] = acc;
          }
    }
    for (int k = 0; k < cubNq; ++k)
      for (int j = 0; j < cubNq; ++j)
        for (int i = 0; i < cubNq; ++i) {
          double v[9];
          for (int f = 0; f < 9; ++f) {
            double acc = 0.0;
            const double* a2 = &d2[(size_t)f * Nq * cubNq * cubNq];
            for (int n = 0; n < Nq; ++n) acc += I[k * Nq + n] * a2[n * cubNq * cubNq + j * cubNq + i];
            v[f] = acc;
          }
          const double xr = v[0], xs = v[1], xt = v[2], yr = v[3], ys = v[4], yt = v[5];
          const double zr = v[6], zs = v[7], zt = v[8];
          const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
          if (J < 1e-8) bad = 1;
          const double rx = (ys * zt - zs * yt) / J, ry = -(xs * zt - zs * xt) / J, rz = (xs * yt - ys * xt) / J;
          const double sx = -(yr * zt - zr * yt) / J, sy = (xr * zt - zr * xt) / J, sz = -(xr * yt - yr * xt) / J;
          const double tx = (yr * zs - zr * ys) / J, ty = -(xr * zs - zr * xs) / J, tz = (xr * ys - yr * xs) / J;
          const double JW = J * P.cubw[i] * P.cubw[j] * P.cubw[k];
          const size_t cid = (size_t)cubNp * e + i + j * cubNq + k * cubNq * cubNq;
          P.cubwJ[cid] = JW;
          double* g = P.cubggeo.data() + 6 * cid;
          g[0] = JW * (rx * rx + ry * ry + rz * rz);
          g[1] = JW * (rx * sx + ry * sy + rz * sz);
          g[2] = JW * (rx * tx + ry * ty + rz * tz);
          g[3] = JW * (sx * sx + sy * sy + sz * sz);
          g[4] = JW * (sx * tx + sy * ty + sz * tz);
          g[5] = JW * (tx * tx + ty * ty + tz * tz);
        }
  }
  if (bad) {
    P.error = "Negative J found (cubature)";
    return -2;
  }
  return 0;
}

// libs/mesh/meshMaskedGatherScatterSetup.cpp:31-77
static void MaskedIds(Problem& P) {
  const dlong Nel = P.Nelements;
  const int Np = P.Np, Nfp = P.Nfp, Nf = P.Nfields;
  P.maskedIds.resize((size_t)Nel * Np * Nf);
  for (dlong e = 0; e < Nel; e++)
    for (int n = 0; n < Np; n++) {
      const hlong id = P.globalIds[(size_t)e * Np + n];
      for (int f = 0; f < Nf; f++) P.maskedIds[(size_t)e * Np * Nf + f * Np + n] = id * Nf + f;
    }
  if (!P.mask) return;
  for (dlong e = 0; e < Nel; e++)
    for (int f = 0; f < 6; f++)
      if (P.EToE[(size_t)f + e * 6] < 0)
        for (int n = 0; n < Nfp; n++) {
          const int fid = P.faceNodes[n + f * Nfp];
          for (int nf = 0; nf < Nf; nf++) P.maskedIds[(size_t)e * Np * Nf + nf * Np + fid] = 0;
        }
}

// libs/ogs/ogsOperator.cpp:452-468 (first index in [first,last) with v > val)
static dlong upperBound(dlong first, dlong last, const dlong* v, dlong val) {
  dlong count = last - first;
  while (count > 0) {
    dlong step = count / 2;
    dlong it = first + step;
    if (!(val < v[it])) {
      first = ++it;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  return first;
}

static const int gatherNodesPerBlock = 512;  // include/ogs/ogsOperator.hpp:38

// libs/ogs/ogsOperator.cpp:474-546
int blockRows(dlong Nrows, const std::vector<dlong>& rowStarts, dlong& Nblocks,
                     std::vector<dlong>& blockStarts) {
  Nblocks = 0;
  blockStarts.clear();
  if (!Nrows) return 0;
  dlong maxRowSize = 0;
  for (dlong n = 0; n < Nrows; ++n) maxRowSize = std::max(maxRowSize, rowStarts[n + 1] - rowStarts[n]);
  if (maxRowSize > gatherNodesPerBlock) return -1;
  dlong maxNblocks = 1;
  while (maxNblocks < Nrows) maxNblocks *= 2;
  std::vector<dlong> oldS(maxNblocks + 1), newS(maxNblocks + 1), sizes(maxNblocks);
  Nblocks = 1;
  oldS[0] = 0;
  oldS[1] = Nrows;
  newS[0] = 0;
  sizes[0] = rowStarts[Nrows];
  dlong maxSize = sizes[0];
  while (maxSize > gatherNodesPerBlock) {
    // The reference sizes these buffers for log2(Nrows) passes; unbalanced
    // row sizes can need one more, which writes past them there. Grow them
    // so the bisection completes as intended.
    if ((dlong)newS.size() < 2 * Nblocks + 1) {
      oldS.resize(2 * Nblocks + 1);
      newS.resize(2 * Nblocks + 1);
      sizes.resize(2 * Nblocks);
    }
    newS[2 * Nblocks] = Nrows;
    for (dlong n = 0; n < Nblocks; ++n) {
      const dlong start = oldS[n], end = oldS[n + 1];
      const dlong rowBlockSize = rowStarts[end] - rowStarts[start];
      if (rowBlockSize > gatherNodesPerBlock) {
        const dlong midSize = rowStarts[start] + (rowBlockSize + 1) / 2;
        dlong mid = upperBound(start, end, rowStarts.data(), midSize);
        if (mid == end) --mid;
        newS[2 * n] = start;
        newS[2 * n + 1] = mid;
        sizes[2 * n] = rowStarts[mid] - rowStarts[start];
        sizes[2 * n + 1] = rowStarts[end] - rowStarts[mid];
      } else {
        newS[2 * n] = start;
        newS[2 * n + 1] = end;
        sizes[2 * n] = rowBlockSize;
        sizes[2 * n + 1] = 0;
      }
    }
    oldS.swap(newS);
    Nblocks *= 2;
    maxSize = 0;
    for (dlong n = 0; n < Nblocks; ++n) maxSize = std::max(maxSize, sizes[n]);
  }
  // prim::unique over the Nblocks+1 sorted starts
  blockStarts.push_back(oldS[0]);
  for (dlong n = 1; n <= Nblocks; ++n)
    if (oldS[n] != blockStarts.back()) blockStarts.push_back(oldS[n]);
  Nblocks = (dlong)blockStarts.size() - 1;
  return 0;
}

// libs/ogs/ogsSetup.cpp:87-290 (single rank, Signed, unique=true).
// With one rank every group is local and FindSharedGroups (:292-469) keeps
// exactly one member of each group positive, so every group counts as a
// "positive" group in EnumerateGatherGroups and the T maps do not depend on
// which member the random pick made positive.
static int OgsSetup(Problem& P) {
  const dlong N = (dlong)P.maskedIds.size();
  // keep non-zero ids, in local order (prim::select, :121-139)
  std::vector<dlong> colIds;
  std::vector<hlong> baseIds;
  for (dlong n = 0; n < N; ++n)
    if (P.maskedIds[n] != 0) {
      colIds.push_back(n);
      baseIds.push_back(P.maskedIds[n]);
    }
  const dlong Nids = (dlong)colIds.size();
  // EnumerateGatherGroups (:492-559): stable sort by |id|, run-length encode,
  // number groups by first appearance.
  std::vector<dlong> sortIds(Nids);
  std::iota(sortIds.begin(), sortIds.end(), 0);
  std::stable_sort(sortIds.begin(), sortIds.end(), [&](dlong a, dlong b) {
    return std::llabs(baseIds[a]) < std::llabs(baseIds[b]);
  });
  std::vector<dlong> groupOffsets;
  for (dlong i = 0; i < Nids; ++i)
    if (i == 0 || std::llabs(baseIds[sortIds[i]]) != std::llabs(baseIds[sortIds[i - 1]]))
      groupOffsets.push_back(i);
  const dlong NgroupsT = (dlong)groupOffsets.size();
  groupOffsets.push_back(Nids);
  std::vector<dlong> gids(Nids, 0);
  for (dlong g = 0; g < NgroupsT; ++g) gids[sortIds[groupOffsets[g]]] = 1;
  std::vector<dlong> gatherIds;
  for (dlong n = 0; n < Nids; ++n)
    if (gids[n] == 1) gatherIds.push_back(n);
  for (dlong g = 0; g < NgroupsT; ++g) gids[gatherIds[g]] = g;
  for (dlong g = 0; g < NgroupsT; ++g) {
    const dlong gid = gids[sortIds[groupOffsets[g]]];
    for (dlong i = groupOffsets[g] + 1; i < groupOffsets[g + 1]; ++i) gids[sortIds[i]] = gid;
  }
  // SortByRow (:561-579): stable sort of (row, colId) by row.
  std::vector<dlong> order(Nids);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](dlong a, dlong b) { return gids[a] < gids[b]; });
  P.Ngather = NgroupsT;
  P.nnzT = Nids;
  P.colIdsT.resize(Nids);
  std::vector<dlong> rows(Nids);
  for (dlong i = 0; i < Nids; ++i) {
    P.colIdsT[i] = colIds[order[i]];
    rows[i] = gids[order[i]];
  }
  // runLengthEncodeConsecutive (libs/ogs/ogsOperator.cpp:560-561)
  P.rowStartsT.assign(NgroupsT + 1, 0);
  for (dlong i = 0; i < Nids; ++i) P.rowStartsT[rows[i] + 1]++;
  for (dlong g = 0; g < NgroupsT; ++g) P.rowStartsT[g + 1] += P.rowStartsT[g];
  if (blockRows(NgroupsT, P.rowStartsT, P.NrowBlocksT, P.blockRowStartsT)) {
    P.error = "Multiplicity of a global node is too large";
    return -3;
  }
  // SetupGlobalToLocalMapping (:833-849)
  P.GlobalToLocal.assign(N, -1);
  for (dlong g = 0; g < NgroupsT; ++g)
    for (dlong i = P.rowStartsT[g]; i < P.rowStartsT[g + 1]; ++i) P.GlobalToLocal[P.colIdsT[i]] = g;
  return 0;
}

int ogs_setup_maps(Problem& P) { return OgsSetup(P); }

int problem_setup(Problem& P) {
  if (P.elementType != 12) return problem_setup_other(P);
  P.dimension = 3;
  if (P.N < 1 || P.N > 15 || P.nx < 1 || P.ny < 1 || P.nz < 1) {
    P.error = "invalid problem size";
    return -1;
  }
  P.Nfields = (P.bp == 2 || P.bp == 4 || P.bp == 6) ? 3 : 1;
  if (ReferenceNodes(P)) return -1;
  SetupBox(P);
  Connect(P);
  ConnectFaceNodes(P);
  ConnectNodes(P);
  PhysicalNodes(P);
  int st = GeometricFactors(P);
  if (st) return st;
  if (P.bp >= 1 && P.bp <= 4) {
    st = CubatureSetup(P);
    if (st) return st;
  }
  P.elementList.resize(P.Nelements);
  std::iota(P.elementList.begin(), P.elementList.end(), 0);
  MaskedIds(P);
  P.NLocal = P.Nelements * P.Np * P.Nfields;
  return OgsSetup(P);
}

}  // namespace orc

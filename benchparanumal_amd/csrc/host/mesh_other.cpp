// Quadrilateral, triangular and tetrahedral box meshes on one rank: vertices,
// nodes, global node numbering, boundary flags, geometric and cubature
// factors.
//
// Reference behaviour followed:
//  - box vertices and element order: libs/mesh/meshSetupBoxQuad2D.cpp
//    (counter-clockwise quads), meshSetupBoxTri2D.cpp (two triangles per
//    square: (0,1,3) then (0,3,2) of its corners), meshSetupBoxTet3D.cpp (six
//    tetrahedra per cube around the diagonal from corner 0 to corner 7);
//  - nodes: the bilinear map at the GLL nodes (meshPhysicalNodesQuad2D.cpp),
//    the affine map at the warp & blend nodes (meshPhysicalNodesTri2D.cpp,
//    meshPhysicalNodesTet3D.cpp);
//  - geometric factors: meshGeometricFactorsQuad2D.cpp (per node: wJ and G
//    times the GLL weights), meshGeometricFactorsTri2D.cpp and
//    meshGeometricFactorsTet3D.cpp (affine: J and J G per element); cubature
//    factors of quads: meshCubatureSetupQuad2D.cpp.
// Global numbering: quads by their lattice position, as the hex box; on
// simplices a node is named by the global vertices it lies between and its
// lattice weights on them (the warp & blend node set of a face or edge is
// symmetric, so two elements sharing it place these nodes at the same
// points), which gives the grouping of the reference's node matching
// (libs/mesh/meshConnectNodes.cpp:32-82) without comparing coordinates.
// Quads on an affine mesh keep per-node factors (the affine map's, equal to
// the reference's per-element J and G times the weights); their operators
// run the non-affine kernels.
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>

#include "basis.hpp"
#include "bpx.hpp"
#include "geometry.hpp"
#include "rng.hpp"

namespace bpx {

namespace {

// the seeded displacement of interior box vertex (I, J[, K]) (as the hex box)
inline void perturb_vertex(const MeshOpts& o, long long vid, bool interior, const double h[3], double* X, int dim) {
  if (o.perturb == 0.0 || !interior) return;
  for (int d = 0; d < dim; ++d) X[d] += o.perturb * h[d] * (2.0 * rng::u01(o.seed, 3 * (uint64_t)vid + d) - 1.0);
}

}  // namespace

int mesh_t::SetupOther(const MeshOpts& o) {
  const int et = o.elementType;
  if (et != ELEM_QUAD && et != ELEM_TRI && et != ELEM_TET)
    return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: element type must be 3 (tri), 4 (quad), 6 (tet) or 12 (hex)");
  if (o.N < 1 || o.N > 15) return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: degree N must be in 1..15");
  const bool d3 = et == ELEM_TET;
  if (o.nx < 1 || o.ny < 1 || (d3 && o.nz < 1))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: nx, ny (, nz) must be >= 1");
  if (!(o.dim[0] > 0 && o.dim[1] > 0 && (!d3 || o.dim[2] > 0)))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: box size must be positive");
  if (!(o.perturb >= 0 && o.perturb < 0.5))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: perturb must be in [0, 0.5)");
  opts = o;
  opts.nz = d3 ? o.nz : 1;
  elementType = et;
  dim = d3 ? 3 : 2;
  Nverts = et == ELEM_QUAD ? 4 : (d3 ? 4 : 3);
  Nggeo = d3 ? 6 : 3;
  onDevice = false;
  N = o.N;
  Nq = N + 1;
  Np = et == ELEM_QUAD ? Nq * Nq : (d3 ? (N + 1) * (N + 2) * (N + 3) / 6 : (N + 1) * (N + 2) / 2);
  const int nx = o.nx, ny = o.ny, nz = opts.nz;
  const long long nel = (long long)nx * ny * nz * (et == ELEM_TRI ? 2 : d3 ? 6 : 1);
  if (nel * Np * 6 > 0x7fffffffLL)
    return set_error(BPX_ERR_UNSUPPORTED, "mesh: local node count exceeds 32-bit indexing");
  Nelements = (dlong)nel;

  basis::gll(N, gllz, gllw);
  basis::dmatrix_gll(N, gllz, D);
  cubNq = Nq + 1;
  cubNp = et == ELEM_QUAD ? cubNq * cubNq : 0;
  cubr.clear(); cubw.clear(); cubInterp.clear(); cubD.clear(); invV.clear();
  rN.clear(); sN.clear(); tN.clear(); Dsim.clear(); Ssim.clear(); nodeIdx.clear();
  if (et == ELEM_QUAD) {
    basis::mass_gll(N, gllz, MM);
    if (o.cubature) {
      basis::gauss(cubNq, cubr, cubw);
      basis::interp(gllz, cubr, cubInterp);
      basis::dmatrix(cubr, cubD);
    }
  } else {
    std::vector<double> V, Vr, Vs, Vt;
    try {
      if (d3) {
        simplex::nodes_tet(N, rN, sN, tN, nodeIdx);
        simplex::vandermonde_tet(N, rN, sN, tN, V, Vr, Vs, Vt);
        simplex::operators(Np, 3, V, {&Vr, &Vs, &Vt}, MM, Dsim, Ssim);
      } else {
        simplex::nodes_tri(N, rN, sN, nodeIdx);
        simplex::vandermonde_tri(N, rN, sN, V, Vr, Vs);
        simplex::operators(Np, 2, V, {&Vr, &Vs}, MM, Dsim, Ssim);
      }
    } catch (const std::exception& ex) {
      return set_error(BPX_ERR_INVALID_ARGUMENT, std::string("mesh: ") + ex.what());
    }
  }

  // ---- vertices (+ seeded perturbation of the interior box vertices)
  const long long NnX = nx + 1, NnY = ny + 1;
  const double h[3] = {o.dim[0] / nx, o.dim[1] / ny, d3 ? o.dim[2] / nz : 1.0};
  EX.assign((size_t)Nelements * Nverts, 0.0);
  EY.assign((size_t)Nelements * Nverts, 0.0);
  EZ.assign(d3 ? (size_t)Nelements * Nverts : 0, 0.0);
  std::vector<long long> EToV((size_t)Nelements * Nverts);
  // corner c of a cell: bit 0 = +x, bit 1 = +y, bit 2 = +z
  auto corner = [&](int i, int j, int k, int c, double* X, long long& vid) {
    const int I = i + (c & 1), J = j + ((c >> 1) & 1), K = k + ((c >> 2) & 1);
    // the cell's corner, then + h (the reference's arithmetic)
    const double x0 = o.lo[0] + h[0] * i, y0 = o.lo[1] + h[1] * j, z0 = d3 ? o.lo[2] + h[2] * k : 0.0;
    X[0] = (c & 1) ? x0 + h[0] : x0;
    X[1] = ((c >> 1) & 1) ? y0 + h[1] : y0;
    X[2] = ((c >> 2) & 1) ? z0 + h[2] : z0;
    vid = I + J * NnX + (long long)K * NnX * NnY;
    const bool interior = I > 0 && I < nx && J > 0 && J < ny && (!d3 || (K > 0 && K < nz));
    perturb_vertex(o, vid, interior, h, X, dim);
  };
  // element corner lists per cell (cell corners by the bit code above)
  static const int kQuad[1][4] = {{0, 1, 3, 2}};
  static const int kTri[2][3] = {{0, 1, 3}, {0, 3, 2}};
  static const int kTet[6][4] = {{0, 3, 2, 7}, {0, 1, 3, 7}, {0, 2, 6, 7}, {0, 6, 4, 7}, {0, 5, 1, 7}, {0, 4, 5, 7}};
  const int per = et == ELEM_TRI ? 2 : d3 ? 6 : 1;
#pragma omp parallel for collapse(3)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i)
        for (int q = 0; q < per; ++q) {
          const size_t e = (size_t)per * (i + (size_t)j * nx + (size_t)k * nx * ny) + q;
          for (int v = 0; v < Nverts; ++v) {
            const int c = et == ELEM_QUAD ? kQuad[0][v] : et == ELEM_TRI ? kTri[q][v] : kTet[q][v];
            double X[3];
            long long vid;
            corner(i, j, k, c, X, vid);
            EX[e * Nverts + v] = X[0];
            EY[e * Nverts + v] = X[1];
            if (d3) EZ[e * Nverts + v] = X[2];
            EToV[e * Nverts + v] = vid;
          }
        }

  // ---- nodes
  const size_t NL = (size_t)Nelements * Np;
  x.assign(NL, 0.0);
  y.assign(NL, 0.0);
  z.assign(d3 ? NL : 0, 0.0);
#pragma omp parallel for
  for (dlong e = 0; e < Nelements; ++e) {
    const double* VX = &EX[(size_t)e * Nverts];
    const double* VY = &EY[(size_t)e * Nverts];
    for (int n = 0; n < Np; ++n) {
      const size_t id = (size_t)e * Np + n;
      if (et == ELEM_QUAD) {
        const double r = gllz[n % Nq], s = gllz[n / Nq];
        const double w[4] = {0.25 * (1 - r) * (1 - s), 0.25 * (1 + r) * (1 - s), 0.25 * (1 + r) * (1 + s),
                             0.25 * (1 - r) * (1 + s)};
        x[id] = w[0] * VX[0] + w[1] * VX[1] + w[2] * VX[2] + w[3] * VX[3];
        y[id] = w[0] * VY[0] + w[1] * VY[1] + w[2] * VY[2] + w[3] * VY[3];
      } else if (!d3) {
        const double r = rN[n], s = sN[n];
        x[id] = -0.5 * (r + s) * VX[0] + 0.5 * (1 + r) * VX[1] + 0.5 * (1 + s) * VX[2];
        y[id] = -0.5 * (r + s) * VY[0] + 0.5 * (1 + r) * VY[1] + 0.5 * (1 + s) * VY[2];
      } else {
        const double* VZ = &EZ[(size_t)e * Nverts];
        const double r = rN[n], s = sN[n], t = tN[n];
        const double w0 = -0.5 * (1 + r + s + t), w1 = 0.5 * (1 + r), w2 = 0.5 * (1 + s), w3 = 0.5 * (1 + t);
        x[id] = w0 * VX[0] + w1 * VX[1] + w2 * VX[2] + w3 * VX[3];
        y[id] = w0 * VY[0] + w1 * VY[1] + w2 * VY[2] + w3 * VY[3];
        z[id] = w0 * VZ[0] + w1 * VZ[1] + w2 * VZ[2] + w3 * VZ[3];
      }
    }
  }

  // ---- global ids (1-based) and boundary flags
  globalIds.assign(NL, 0);
  boundary.assign(NL, 0);
  if (et == ELEM_QUAD) {
    const long long LX = (long long)nx * N + 1, LY = (long long)ny * N + 1;
#pragma omp parallel for
    for (dlong e = 0; e < Nelements; ++e) {
      const long long ex = e % nx, ey = e / nx;
      for (int n = 0; n < Np; ++n) {
        const long long X = ex * N + n % Nq, Y = ey * N + n / Nq;
        globalIds[(size_t)e * Np + n] = 1 + X + Y * LX;
        boundary[(size_t)e * Np + n] = X == 0 || X == LX - 1 || Y == 0 || Y == LY - 1;
      }
    }
  } else {
    // key: (vertex id, lattice weight) pairs of the node's support, sorted by vertex id
    const int nv = dim + 1;
    typedef std::array<long long, 8> Key;
    std::vector<std::pair<Key, size_t>> keys(NL);
#pragma omp parallel for
    for (dlong e = 0; e < Nelements; ++e)
      for (int n = 0; n < Np; ++n) {
        int wgt[4];
        int sum = 0;
        for (int d = 0; d < dim; ++d) sum += (wgt[d + 1] = nodeIdx[(size_t)n * dim + d]);
        wgt[0] = N - sum;
        std::array<std::pair<long long, int>, 4> pv;
        for (int v = 0; v < nv; ++v) pv[v] = {wgt[v] > 0 ? EToV[(size_t)e * Nverts + v] : -1, wgt[v]};
        std::sort(pv.begin(), pv.begin() + nv);
        Key K;
        K.fill(-1);
        for (int v = 0; v < nv; ++v) {
          K[2 * v] = pv[v].first;
          K[2 * v + 1] = pv[v].first < 0 ? 0 : pv[v].second;
        }
        keys[(size_t)e * Np + n] = {K, (size_t)e * Np + n};
      }
    std::sort(keys.begin(), keys.end());
    long long g = 0;
    for (size_t i = 0; i < NL; ++i) {
      if (i == 0 || keys[i].first != keys[i - 1].first) ++g;
      globalIds[keys[i].second] = g;
    }
    // Dirichlet mask, as libs/mesh/meshMaskedGatherScatterSetup.cpp:44-58:
    // a local node is masked when it lies on a face of its element that has
    // no neighbour (all of the face's vertices on one side of the box).
    // Unlike on the hex box, a simplex can touch the boundary with a vertex
    // or an edge alone; those copies stay unmasked there, and so does the
    // node if any copy is, as in the reference.
    static const int kTriF[3][3] = {{0, 1, -1}, {1, 2, -1}, {2, 0, -1}};
    static const int kTetF[4][3] = {{0, 1, 2}, {0, 3, 1}, {1, 3, 2}, {0, 2, 3}};
    const long long NnZ = nz + 1;
    auto on_side = [&](long long vid, int side) {
      const long long I = vid % NnX, J = (vid / NnX) % NnY, K = vid / (NnX * NnY);
      switch (side) {
        case 0: return I == 0;
        case 1: return I == nx;
        case 2: return J == 0;
        case 3: return J == ny;
        case 4: return d3 && K == 0;
        default: return d3 && K == NnZ - 1;
      }
    };
    const int nfaces = d3 ? 4 : 3, nfv = d3 ? 3 : 2;
#pragma omp parallel for
    for (dlong e = 0; e < Nelements; ++e) {
      bool bface[4] = {false, false, false, false};
      for (int f = 0; f < nfaces; ++f)
        for (int side = 0; side < 6 && !bface[f]; ++side) {
          bool all = true;
          for (int v = 0; v < nfv; ++v)
            if (!on_side(EToV[(size_t)e * Nverts + (d3 ? kTetF[f][v] : kTriF[f][v])], side)) all = false;
          bface[f] = all;
        }
      for (int n = 0; n < Np; ++n) {
        int wgt[4];
        int sum = 0;
        for (int d = 0; d < dim; ++d) sum += (wgt[d + 1] = nodeIdx[(size_t)n * dim + d]);
        wgt[0] = N - sum;
        bool bnd = false;
        for (int f = 0; f < nfaces && !bnd; ++f) {
          if (!bface[f]) continue;
          // on face f: no weight on the vertex off the face
          bool onf = true;
          for (int v = 0; v < nv; ++v) {
            bool inF = false;
            for (int k = 0; k < nfv; ++k) inF |= (d3 ? kTetF[f][k] : kTriF[f][k]) == v;
            if (!inF && wgt[v] != 0) onf = false;
          }
          bnd = onf;
        }
        boundary[(size_t)e * Np + n] = bnd;
      }
    }
  }

  // ---- geometric factors
  int bad = 0;
  wJ.clear(); ggeo.clear(); cubwJ.clear(); cubggeo.clear();
  if (et == ELEM_QUAD) {
    wJ.assign(NL, 0.0);
    ggeo.assign(NL * 3, 0.0);
#pragma omp parallel for reduction(| : bad)
    for (dlong e = 0; e < Nelements; ++e)
      for (int n = 0; n < Np; ++n) {
        const int i = n % Nq, j = n / Nq;
        double xr = 0, xs = 0, yr = 0, ys = 0;
        for (int m = 0; m < Nq; ++m) {
          const size_t ir = (size_t)e * Np + j * Nq + m, is = (size_t)e * Np + m * Nq + i;
          xr += D[i * Nq + m] * x[ir];
          xs += D[j * Nq + m] * x[is];
          yr += D[i * Nq + m] * y[ir];
          ys += D[j * Nq + m] * y[is];
        }
        const double J = xr * ys - xs * yr;
        if (!(J >= 1e-8)) bad = 1;
        const double rx = ys / J, ry = -xs / J, sx = -yr / J, sy = xr / J;
        const double JW = J * gllw[i] * gllw[j];
        const size_t id = (size_t)e * Np + n;
        wJ[id] = JW;
        ggeo[3 * id + 0] = JW * (rx * rx + ry * ry);
        ggeo[3 * id + 1] = JW * (rx * sx + ry * sy);
        ggeo[3 * id + 2] = JW * (sx * sx + sy * sy);
      }
    if (!bad && o.cubature) {
      const int cq = cubNq;
      cubwJ.assign((size_t)Nelements * cubNp, 0.0);
      cubggeo.assign((size_t)Nelements * cubNp * 3, 0.0);
      const std::vector<double>& I = cubInterp;
#pragma omp parallel for reduction(| : bad)
      for (dlong e = 0; e < Nelements; ++e) {
        std::vector<double> d0(4 * (size_t)Np), d1(4 * (size_t)Nq * cq);
        for (int n = 0; n < Np; ++n) {
          const int i = n % Nq, j = n / Nq;
          double v[4] = {0, 0, 0, 0};
          for (int m = 0; m < Nq; ++m) {
            const size_t ir = (size_t)e * Np + j * Nq + m, is = (size_t)e * Np + m * Nq + i;
            v[0] += D[i * Nq + m] * x[ir];
            v[1] += D[j * Nq + m] * x[is];
            v[2] += D[i * Nq + m] * y[ir];
            v[3] += D[j * Nq + m] * y[is];
          }
          for (int f = 0; f < 4; ++f) d0[(size_t)f * Np + n] = v[f];
        }
        for (int f = 0; f < 4; ++f)
          for (int j = 0; j < Nq; ++j)
            for (int a = 0; a < cq; ++a) {
              double acc = 0;
              for (int m = 0; m < Nq; ++m) acc += I[a * Nq + m] * d0[(size_t)f * Np + j * Nq + m];
              d1[(size_t)f * Nq * cq + j * cq + a] = acc;
            }
        for (int b = 0; b < cq; ++b)
          for (int a = 0; a < cq; ++a) {
            double v[4];
            for (int f = 0; f < 4; ++f) {
              double acc = 0;
              for (int m = 0; m < Nq; ++m) acc += I[b * Nq + m] * d1[(size_t)f * Nq * cq + m * cq + a];
              v[f] = acc;
            }
            const double xr = v[0], xs = v[1], yr = v[2], ys = v[3];
            const double J = xr * ys - xs * yr;
            if (!(J >= 1e-8)) bad = 1;
            const double rx = ys / J, ry = -xs / J, sx = -yr / J, sy = xr / J;
            const double JW = J * cubw[a] * cubw[b];
            const size_t cid = (size_t)e * cubNp + b * cq + a;
            cubwJ[cid] = JW;
            cubggeo[3 * cid + 0] = JW * (rx * rx + ry * ry);
            cubggeo[3 * cid + 1] = JW * (rx * sx + ry * sy);
            cubggeo[3 * cid + 2] = JW * (sx * sx + sy * sy);
          }
      }
    }
  } else {
    wJ.assign(Nelements, 0.0);
    ggeo.assign((size_t)Nelements * Nggeo, 0.0);
#pragma omp parallel for reduction(| : bad)
    for (dlong e = 0; e < Nelements; ++e) {
      const size_t b = (size_t)e * Nverts;
      if (!d3) {
        const double xr = 0.5 * (EX[b + 1] - EX[b]), xs = 0.5 * (EX[b + 2] - EX[b]);
        const double yr = 0.5 * (EY[b + 1] - EY[b]), ys = 0.5 * (EY[b + 2] - EY[b]);
        const double J = xr * ys - xs * yr;
        if (!(J >= 1e-8)) bad = 1;
        const double rx = ys / J, ry = -xs / J, sx = -yr / J, sy = xr / J;
        wJ[e] = J;
        ggeo[3 * (size_t)e + 0] = J * (rx * rx + ry * ry);
        ggeo[3 * (size_t)e + 1] = J * (rx * sx + ry * sy);
        ggeo[3 * (size_t)e + 2] = J * (sx * sx + sy * sy);
      } else {
        Geo9 d;
        d.xr = 0.5 * (EX[b + 1] - EX[b]); d.xs = 0.5 * (EX[b + 2] - EX[b]); d.xt = 0.5 * (EX[b + 3] - EX[b]);
        d.yr = 0.5 * (EY[b + 1] - EY[b]); d.ys = 0.5 * (EY[b + 2] - EY[b]); d.yt = 0.5 * (EY[b + 3] - EY[b]);
        d.zr = 0.5 * (EZ[b + 1] - EZ[b]); d.zs = 0.5 * (EZ[b + 2] - EZ[b]); d.zt = 0.5 * (EZ[b + 3] - EZ[b]);
        double J, g[6];
        metric(d, J, g);
        if (!(J >= 0)) bad = 1;
        wJ[e] = J;
        for (int c = 0; c < 6; ++c) ggeo[6 * (size_t)e + c] = g[c];
      }
    }
  }
  if (bad) return set_error(BPX_ERR_NEGATIVE_JACOBIAN, "mesh: negative or vanishing Jacobian found");
  return BPX_OK;
}

}  // namespace bpx

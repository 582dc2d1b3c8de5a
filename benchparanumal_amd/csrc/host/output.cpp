// Output and verification: the VTU writer of mesh_t::PlotFields
// (libs/mesh/meshPlotFields3D.cpp:31-121, called from
// BP/BP5/src/bp5Run.cpp:220-238) and the error of a solution against the
// forcing's manufactured solution.
#include <cmath>
#include <cstdio>
#include <memory>
#include <string>
#include <vector>

#include "bpx.hpp"

namespace bpx {

// The 6 tetrahedra of the unit cube that share its main diagonal (corner
// bits: 1 = +i, 2 = +j, 4 = +k): one per ordering of the three axes, the
// path 0 -> a -> a+b -> 7 along them. Each is listed with positive volume.
static void cube_tets(int tets[6][4]) {
  const int axes[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int p = 0; p < 6; ++p) {
    int v0 = 0, v1 = 1 << axes[p][0], v2 = v1 | (1 << axes[p][1]), v3 = 7;
    // signed volume of (v1 - v0, v2 - v0, v3 - v0): the determinant of the
    // corner bit vectors; swap two vertices where it is negative
    auto bit = [](int v, int a) { return (double)((v >> a) & 1); };
    double m[3][3];
    const int vs[3] = {v1, v2, v3};
    for (int r = 0; r < 3; ++r)
      for (int a = 0; a < 3; ++a) m[r][a] = bit(vs[r], a) - bit(v0, a);
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (det < 0) std::swap(v1, v2);
    tets[p][0] = v0;
    tets[p][1] = v1;
    tets[p][2] = v2;
    tets[p][3] = v3;
  }
}

int mesh_t::PlotFields(const double* QL, int Nfields, const std::string& fileName) const {
  if (!QL || Nfields < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "plot: no field data");
  if (elementType != ELEM_HEX) return set_error(BPX_ERR_UNSUPPORTED, "plot: the VTU writer covers hexahedra");
  std::unique_ptr<FILE, int (*)(FILE*)> fp(std::fopen(fileName.c_str(), "w"), &std::fclose);
  if (!fp) return set_error(BPX_ERR_INVALID_ARGUMENT, "plot: cannot open " + fileName);
  FILE* f = fp.get();
  const int N1 = N, plotCells = 6 * N1 * N1 * N1;
  std::fprintf(f, "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"BigEndian\">\n");
  std::fprintf(f, "  <UnstructuredGrid>\n");
  std::fprintf(f, "    <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", (long long)Nelements * Np,
               (long long)Nelements * plotCells);
  std::fprintf(f, "      <Points>\n");
  std::fprintf(f, "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" Format=\"ascii\">\n");
  for (size_t n = 0; n < (size_t)Nelements * Np; ++n) std::fprintf(f, "       %g %g %g\n", x[n], y[n], z[n]);
  std::fprintf(f, "        </DataArray>\n      </Points>\n");
  std::fprintf(f, "      <PointData Scalars=\"scalars\">\n");
  std::fprintf(f, "        <DataArray type=\"Float32\" Name=\"Fields\" NumberOfComponents=\"%d\" Format=\"ascii\">\n",
               Nfields);
  for (dlong e = 0; e < Nelements; ++e)
    for (int n = 0; n < Np; ++n) {
      std::fprintf(f, "       ");
      for (int c = 0; c < Nfields; ++c) std::fprintf(f, "%g ", QL[((size_t)e * Nfields + c) * Np + n]);
      std::fprintf(f, "\n");
    }
  std::fprintf(f, "        </DataArray>\n      </PointData>\n");
  // cells: sub-cube (i, j, k) of each element, split into 6 tetrahedra
  int tets[6][4];
  cube_tets(tets);
  std::fprintf(f, "      <Cells>\n");
  std::fprintf(f, "        <DataArray type=\"Int32\" Name=\"connectivity\" Format=\"ascii\">\n");
  for (dlong e = 0; e < Nelements; ++e)
    for (int k = 0; k < N1; ++k)
      for (int j = 0; j < N1; ++j)
        for (int i = 0; i < N1; ++i)
          for (int t = 0; t < 6; ++t) {
            std::fprintf(f, "       ");
            for (int v = 0; v < 4; ++v) {
              const int c = tets[t][v];
              const int node = (i + (c & 1)) + (j + ((c >> 1) & 1)) * Nq + (k + ((c >> 2) & 1)) * Nq * Nq;
              std::fprintf(f, "%lld ", (long long)e * Np + node);
            }
            std::fprintf(f, "\n");
          }
  std::fprintf(f, "        </DataArray>\n");
  std::fprintf(f, "        <DataArray type=\"Int32\" Name=\"offsets\" Format=\"ascii\">\n");
  long long cnt = 0;
  for (long long c = 0; c < (long long)Nelements * plotCells; ++c) {
    cnt += 4;
    std::fprintf(f, "       %lld\n", cnt);
  }
  std::fprintf(f, "        </DataArray>\n");
  std::fprintf(f, "        <DataArray type=\"Int32\" Name=\"types\" Format=\"ascii\">\n");
  for (long long c = 0; c < (long long)Nelements * plotCells; ++c) std::fprintf(f, "10\n");  // VTK_TETRA
  std::fprintf(f, "        </DataArray>\n      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n");
  if (std::ferror(f)) return set_error(BPX_ERR_INVALID_ARGUMENT, "plot: write failed for " + fileName);
  return BPX_OK;
}

int bp_t::PlotSolution(const dfloat* o_u, const std::string& fileName) {
  // ogs.Scatter(o_xL, o_x) then mesh.PlotFields (BP/BP5/src/bp5Run.cpp:224-237);
  // Dirichlet nodes (no row) are written as 0
  if (int st = mesh->EnsureHost()) return st;  // node coordinates on the host
  if (mesh->elementType != ELEM_HEX) return set_error(BPX_ERR_UNSUPPORTED, "plot: the VTU writer covers hexahedra");
  hipStream_t s = platform->stream;
  if (int st = launch_set<double>(s, NLocal, 0.0, o_AqL.ptr())) return st;
  if (int st = ogs.Scatter<double>(*platform, o_AqL.ptr(), o_u, 1)) return st;
  std::vector<double> xL(NLocal);
  BPX_HIP(hipMemcpyAsync(xL.data(), o_AqL.ptr(), NLocal * sizeof(double), hipMemcpyDeviceToHost, s));
  if (int st = platform->finish()) return st;
  return mesh->PlotFields(xL.data(), Nfields, fileName);
}

int bp_t::ErrorNorms(const dfloat* o_u, double* out) {
  if (Nfields != 1) return set_error(BPX_ERR_UNSUPPORTED, "error norms: one-field problems (BP1, BP3, BP5)");
  if (mesh->elementType != ELEM_HEX) return set_error(BPX_ERR_UNSUPPORTED, "error norms: hexahedral meshes");
  if (int st = mesh->EnsureHost()) return st;  // wJ on the host
  const mesh_t& m = *mesh;
  hipStream_t s = platform->stream;
  // the forcing's solution: u = sin sin sin for the stiffness problems; the
  // mass problems (BP1) are forced with lambda = 0, so M u = M 3 pi^2 u0
  const double c = (bp <= 2) ? 3 * M_PI * M_PI : 1.0;
  if (int st = NodeCoordinates()) return st;
  // GLL weight times Jacobian per local node (affine meshes store J per element)
  std::vector<double> W((size_t)NLocal);
  const int Nq = m.Nq, Np = m.Np;
  for (dlong e = 0; e < m.Nelements; ++e)
    for (int n = 0; n < Np; ++n) {
      const size_t id = (size_t)e * Np + n;
      W[id] = m.opts.affine ? m.wJ[e] * m.gllw[n % Nq] * m.gllw[(n / Nq) % Nq] * m.gllw[n / (Nq * Nq)] : m.wJ[id];
    }
  deviceBuffer<double> o_W, o_part;
  if (int st = o_W.upload(W, s)) return st;
  if (int st = o_part.alloc(3 * (size_t)kRedBlocks)) return st;
  if (int st = launch_error_norms(s, NLocal, o_GlobalToLocal.ptr(), o_u, o_x.ptr(), o_y.ptr(), o_z.ptr(), o_W.ptr(), c,
                                  o_part.ptr()))
    return st;
  std::vector<double> part(3 * (size_t)kRedBlocks);
  BPX_HIP(hipMemcpyAsync(part.data(), o_part.ptr(), part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  if (int st = platform->finish()) return st;
  double mx = 0, e2 = 0, u2 = 0;
  for (int b = 0; b < kRedBlocks; ++b) {
    mx = std::max(mx, part[3 * b]);
    e2 += part[3 * b + 1];
    u2 += part[3 * b + 2];
  }
  out[0] = mx;
  out[1] = std::sqrt(e2);
  out[2] = u2 > 0 ? std::sqrt(e2 / u2) : 0.0;
  return BPX_OK;
}

}  // namespace bpx

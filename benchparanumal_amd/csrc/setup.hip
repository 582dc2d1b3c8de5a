// GPU-side setup of the hex box problem: the nodes, global ids, geometric and
// cubature factors of mesh_t; the masked ids, T maps, row blocks and
// GlobalToLocal of ogs_t; the solve-space maps of the fused BP5 operator.
//
// Reference behaviour followed (the host versions in host/mesh.cpp,
// host/ogs.cpp and host/solver.cpp state the same rules):
//  - nodes on the trilinear map: libs/mesh/meshPhysicalNodesHex3D.cpp:31-116
//  - geometric factors: libs/mesh/meshGeometricFactorsHex3D.cpp:31-153
//  - cubature factors: libs/mesh/meshCubatureSetupHex3D.cpp:97-257
//  - masked ids: libs/mesh/meshMaskedGatherScatterSetup.cpp:31-77
//  - ogs T maps: libs/ogs/ogsSetup.cpp:87-290 (nodes grouped by |id|, rows
//    numbered by the group's first appearance, each row's nodes ascending)
//    and ogsSetup.cpp:833-849 (GlobalToLocal)
//  - row blocks: libs/ogs/ogsOperator.cpp:474-546 (bisection at the first
//    row start past the block's midpoint until every block holds at most 512
//    entries)
//
// Every floating-point value is formed by the operations of the host code in
// the same order (host/geometry.hpp holds the shared point formulas, and this
// file, like the host code, is built without floating-point contraction), so
// the device arrays equal the host arrays bit for bit. The maps come from a
// stable radix sort of (|id|, local index) pairs, prefix sums and one kernel
// per bisection level; they too equal the host maps exactly.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "host/bpx.hpp"
#include "host/geometry.hpp"

namespace bpx {

namespace {

constexpr int kT = 256;
inline unsigned blocks_for(size_t n) { return (unsigned)((n + kT - 1) / kT); }

template <typename T>
__global__ void k_fill(size_t n, T v, T* __restrict__ a) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

template <typename T>
int fill(hipStream_t s, size_t n, T v, T* a) {
  if (n == 0) return BPX_OK;
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks_for(n)), dim3(kT), 0, s, n, v, a);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int read_value(hipStream_t s, const T* d, T& h) {
  BPX_HIP(hipMemcpyAsync(&h, d, sizeof(T), hipMemcpyDeviceToHost, s));
  BPX_HIP(hipStreamSynchronize(s));
  return BPX_OK;
}

// hipcub call with temporary storage grown on demand: f(temp, bytes)
template <typename F>
int cub_call(deviceBuffer<char>& tmp, F f) {
  size_t bytes = 0;
  BPX_HIP(f(nullptr, bytes));
  if (bytes > tmp.length())
    if (int st = tmp.alloc(bytes)) return st;
  BPX_HIP(f(tmp.ptr(), bytes));
  return BPX_OK;
}

// ------------------------------------------------------------- mesh -----

// nodal coordinates (trilinear map of the 8 vertices) and global ids (1 +
// lattice index), one thread per local node
__global__ void k_mesh_nodes(dlong Nel, int Nq, int nx, int ny, const double* __restrict__ gllz,
                             const double* __restrict__ EX, const double* __restrict__ EY,
                             const double* __restrict__ EZ, double* __restrict__ x, double* __restrict__ y,
                             double* __restrict__ z, hlong* __restrict__ gid) {
  const int Np = Nq * Nq * Nq, N = Nq - 1;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)Nel * Np) return;
  const dlong e = (dlong)(id / Np);
  const int n = (int)(id - (size_t)e * Np);
  const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
  const double r = gllz[i], s = gllz[j], t = gllz[k];
  double X = 0, Y = 0, Z = 0;
  for (int v = 0; v < 8; ++v) {
    const double c = vertex_weight(v, r, s, t);
    X += c * EX[(size_t)e * 8 + v];
    Y += c * EY[(size_t)e * 8 + v];
    Z += c * EZ[(size_t)e * 8 + v];
  }
  x[id] = X;
  y[id] = Y;
  z[id] = Z;
  const long long LX = (long long)nx * N + 1, LY = (long long)ny * N + 1;
  const long long ex = e % nx, ey = (e / nx) % ny, ez = e / ((long long)nx * ny);
  const long long Xl = ex * N + i, Yl = ey * N + j, Zl = ez * N + k;
  gid[id] = 1 + Xl + Yl * LX + Zl * LX * LY;
}

// the 9 coordinate derivatives at GLL node (i, j, k) of an element, each a
// sum over m in increasing order (host/mesh.cpp's order)
__device__ inline Geo9 node_derivatives(int Nq, int i, int j, int k, const double* __restrict__ D,
                                        const double* __restrict__ X, const double* __restrict__ Y,
                                        const double* __restrict__ Z) {
  Geo9 d{0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int m = 0; m < Nq; ++m) {
    const int ir = k * Nq * Nq + j * Nq + m, is = k * Nq * Nq + m * Nq + i, it = m * Nq * Nq + j * Nq + i;
    const double Di = D[i * Nq + m], Dj = D[j * Nq + m], Dk = D[k * Nq + m];
    d.xr += Di * X[ir]; d.xs += Dj * X[is]; d.xt += Dk * X[it];
    d.yr += Di * Y[ir]; d.ys += Dj * Y[is]; d.yt += Dk * Y[it];
    d.zr += Di * Z[ir]; d.zs += Dj * Z[is]; d.zt += Dk * Z[it];
  }
  return d;
}

// non-affine geometric factors: wJ and the 6 entries of G per node, both
// times the GLL weight; bad = 1 where J < 1e-12
__global__ void k_geo_nodes(dlong Nel, int Nq, const double* __restrict__ D, const double* __restrict__ gllw,
                            const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, double* __restrict__ wJ, double* __restrict__ ggeo,
                            int* __restrict__ bad) {
  const int Np = Nq * Nq * Nq;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)Nel * Np) return;
  const dlong e = (dlong)(id / Np);
  const int n = (int)(id - (size_t)e * Np);
  const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
  const size_t b = (size_t)e * Np;
  const Geo9 d = node_derivatives(Nq, i, j, k, D, x + b, y + b, z + b);
  double J, g[6];
  metric(d, J, g);
  if (!(J >= 1e-12)) *bad = 1;
  const double w = gllw[i] * gllw[j] * gllw[k];
  wJ[id] = J * w;
  for (int c = 0; c < 6; ++c) ggeo[id * 6 + c] = g[c] * w;
}

// affine geometric factors: J and J G per element from the vertex edges;
// bad = 1 where J < 0
__global__ void k_geo_affine(dlong Nel, const double* __restrict__ EX, const double* __restrict__ EY,
                             const double* __restrict__ EZ, double* __restrict__ wJ, double* __restrict__ ggeo,
                             int* __restrict__ bad) {
  const dlong e = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= Nel) return;
  const size_t b = (size_t)e * 8;
  Geo9 d;
  d.xr = 0.5 * (EX[b + 1] - EX[b]); d.xs = 0.5 * (EX[b + 3] - EX[b]); d.xt = 0.5 * (EX[b + 4] - EX[b]);
  d.yr = 0.5 * (EY[b + 1] - EY[b]); d.ys = 0.5 * (EY[b + 3] - EY[b]); d.yt = 0.5 * (EY[b + 4] - EY[b]);
  d.zr = 0.5 * (EZ[b + 1] - EZ[b]); d.zs = 0.5 * (EZ[b + 3] - EZ[b]); d.zt = 0.5 * (EZ[b + 4] - EZ[b]);
  double J, g[6];
  metric(d, J, g);
  if (!(J >= 0)) *bad = 1;
  wJ[e] = J;
  for (int c = 0; c < 6; ++c) ggeo[(size_t)e * 6 + c] = g[c];
}

// Cubature factors, in four passes over a chunk of elements (e0 ..
// e0 + Nc - 1) with three scratch arrays, as the host does per element:
// the 9 derivatives at the GLL nodes (a0: [el][f][Np]), interpolated to the
// Gauss points along r (a1: [el][f][Nq][Nq][cq]), s (a2: [el][f][Nq][cq][cq])
// and t, where the last pass forms J and G.
__global__ void k_cub_d0(dlong e0, dlong Nc, int Nq, const double* __restrict__ D, const double* __restrict__ x,
                         const double* __restrict__ y, const double* __restrict__ z, double* __restrict__ a0) {
  const int Np = Nq * Nq * Nq;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)Nc * Np) return;
  const dlong el = (dlong)(id / Np);
  const int n = (int)(id - (size_t)el * Np);
  const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
  const size_t b = (size_t)(e0 + el) * Np;
  const double* C[3] = {x + b, y + b, z + b};
  double* out = a0 + (size_t)el * 9 * Np;
  for (int c = 0; c < 3; ++c) {
    double ar = 0, as = 0, at = 0;
    for (int m = 0; m < Nq; ++m) {
      ar += D[i * Nq + m] * C[c][k * Nq * Nq + j * Nq + m];
      as += D[j * Nq + m] * C[c][k * Nq * Nq + m * Nq + i];
      at += D[k * Nq + m] * C[c][m * Nq * Nq + j * Nq + i];
    }
    out[(size_t)(3 * c + 0) * Np + n] = ar;
    out[(size_t)(3 * c + 1) * Np + n] = as;
    out[(size_t)(3 * c + 2) * Np + n] = at;
  }
}

__global__ void k_cub_r(dlong Nc, int Nq, int cq, const double* __restrict__ I, const double* __restrict__ a0,
                        double* __restrict__ a1) {
  const int Np = Nq * Nq * Nq, n1 = Nq * Nq * cq;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)Nc * 9 * n1) return;
  const size_t ef = id / n1;  // element-in-chunk * 9 + field
  const int o = (int)(id - ef * n1);
  const int i = o % cq, kj = o / cq;  // kj = k * Nq + j
  const double* in = a0 + ef * Np + (size_t)kj * Nq;
  double acc = 0;
  for (int m = 0; m < Nq; ++m) acc += I[i * Nq + m] * in[m];
  a1[ef * n1 + o] = acc;
}

__global__ void k_cub_s(dlong Nc, int Nq, int cq, const double* __restrict__ I, const double* __restrict__ a1,
                        double* __restrict__ a2) {
  const int n1 = Nq * Nq * cq, n2 = Nq * cq * cq;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)Nc * 9 * n2) return;
  const size_t ef = id / n2;
  const int o = (int)(id - ef * n2);
  const int i = o % cq, j = (o / cq) % cq, k = o / (cq * cq);
  const double* in = a1 + ef * n1;
  double acc = 0;
  for (int m = 0; m < Nq; ++m) acc += I[j * Nq + m] * in[(k * Nq + m) * cq + i];
  a2[ef * n2 + o] = acc;
}

__global__ void k_cub_t(dlong e0, dlong Nc, int Nq, int cq, const double* __restrict__ I,
                        const double* __restrict__ cubw, const double* __restrict__ a2,
                        double* __restrict__ cubwJ, double* __restrict__ cubggeo, int* __restrict__ bad) {
  const int n2 = Nq * cq * cq, cubNp = cq * cq * cq;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)Nc * cubNp) return;
  const dlong el = (dlong)(id / cubNp);
  const int o = (int)(id - (size_t)el * cubNp);
  const int i = o % cq, j = (o / cq) % cq, k = o / (cq * cq);
  double v[9];
  for (int f = 0; f < 9; ++f) {
    const double* in = a2 + ((size_t)el * 9 + f) * n2;
    double acc = 0;
    for (int m = 0; m < Nq; ++m) acc += I[k * Nq + m] * in[(m * cq + j) * cq + i];
    v[f] = acc;
  }
  const Geo9 d{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], v[8]};
  double J, g[6];
  metric(d, J, g);
  if (!(J >= 1e-8)) *bad = 1;
  const double w = cubw[i] * cubw[j] * cubw[k];
  const size_t cid = (size_t)(e0 + el) * cubNp + o;
  cubwJ[cid] = J * w;
  for (int c = 0; c < 6; ++c) cubggeo[cid * 6 + c] = g[c] * w;
}

// masked ids of Nfields fields ([e][f][n]): 0 on the box boundary, else
// (global id) * Nfields + f
__global__ void k_masked_ids(dlong Nel, int Nq, int Nfields, long long LX, long long LY, long long LZ,
                             const hlong* __restrict__ gid, hlong* __restrict__ ids) {
  const int Np = Nq * Nq * Nq;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)Nel * Np) return;
  const dlong e = (dlong)(id / Np);
  const int n = (int)(id - (size_t)e * Np);
  const hlong g = gid[id] - 1;
  const long long X = g % LX, Y = (g / LX) % LY, Z = g / (LX * LY);
  const bool boundary = X == 0 || X == LX - 1 || Y == 0 || Y == LY - 1 || Z == 0 || Z == LZ - 1;
  for (int f = 0; f < Nfields; ++f)
    ids[(size_t)e * Np * Nfields + (size_t)f * Np + n] = boundary ? 0 : (g + 1) * Nfields + f;
}

// bad = 1 where a node is off its element's trilinear map by more than tol
// (c: 24 coefficients per element, as mesh_t::TrilinearCoefficients)
__global__ void k_check_trilinear(dlong Nel, int Nq, double tol, const double* __restrict__ gllz,
                                  const double* __restrict__ c, const double* __restrict__ EX,
                                  const double* __restrict__ EY, const double* __restrict__ EZ,
                                  const double* __restrict__ x, const double* __restrict__ y,
                                  const double* __restrict__ z, int* __restrict__ bad) {
  const int Np = Nq * Nq * Nq;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (size_t)Nel * Np) return;
  const dlong e = (dlong)(id / Np);
  const int n = (int)(id - (size_t)e * Np);
  const double r = gllz[n % Nq], s = gllz[(n / Nq) % Nq], t = gllz[n / (Nq * Nq)];
  const double* V[3] = {EX + (size_t)e * 8, EY + (size_t)e * 8, EZ + (size_t)e * 8};
  const double* X[3] = {x, y, z};
  for (int d = 0; d < 3; ++d) {
    const double* cd = c + (size_t)e * 24 + 8 * d;
    double c0 = 0;
    for (int v = 0; v < 8; ++v) c0 += 0.125 * V[d][v];
    const double m = c0 + cd[0] * r + cd[1] * s + cd[2] * t + cd[3] * r * s + cd[4] * r * t + cd[5] * s * t +
                     cd[6] * r * s * t;
    if (fabs(m - X[d][id]) > tol) *bad = 1;
  }
}

// -------------------------------------------------------------- ogs -----

__global__ void k_ogs_keys(dlong N, const hlong* __restrict__ ids, unsigned long long* __restrict__ key,
                           dlong* __restrict__ val) {
  const dlong n = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (n >= N) return;
  const hlong g = ids[n];
  key[n] = (unsigned long long)(g < 0 ? -g : g);
  val[n] = n;
}

// over the sorted keys: head[i] = 1 at the first entry of each nonzero key;
// first[0] = index of the first nonzero key (left at N when there is none)
__global__ void k_ogs_heads(dlong N, const unsigned long long* __restrict__ key, dlong* __restrict__ head,
                            dlong* __restrict__ first) {
  const dlong i = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const unsigned long long k = key[i], kp = i > 0 ? key[i - 1] : 0ull;
  head[i] = (k != 0 && (i == 0 || k != kp)) ? 1 : 0;
  if (k != 0 && (i == 0 || kp == 0)) first[0] = i;
}

// gstart[g] = sorted position of group g's first entry; firstFlag[n] = 1 for
// the smallest local index of each group (the sort keeps indices ascending)
__global__ void k_ogs_groups(dlong N, const dlong* __restrict__ head, const dlong* __restrict__ gnum,
                             const dlong* __restrict__ val, dlong* __restrict__ gstart,
                             dlong* __restrict__ firstFlag) {
  const dlong i = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  if (head[i]) {
    gstart[gnum[i] - 1] = i;
    firstFlag[val[i]] = 1;
  }
}

// row of each group = rank of its first member in local order; cnt[row] =
// the group's size (cnt[Ngroups] = 0 for the exclusive sum)
__global__ void k_ogs_group_rows(dlong Ngroups, dlong N, const dlong* __restrict__ gstart,
                                 const dlong* __restrict__ val, const dlong* __restrict__ rowAt,
                                 dlong* __restrict__ grow, dlong* __restrict__ cnt) {
  const dlong g = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (g > Ngroups) return;
  if (g == Ngroups) {
    cnt[Ngroups] = 0;
    return;
  }
  const dlong s = gstart[g], e = (g + 1 < Ngroups) ? gstart[g + 1] : N;
  const dlong row = rowAt[val[s]];
  grow[g] = row;
  cnt[row] = e - s;
}

__global__ void k_ogs_cols(dlong N, dlong z, const dlong* __restrict__ gnum, const dlong* __restrict__ gstart,
                           const dlong* __restrict__ grow, const dlong* __restrict__ rowStarts,
                           const dlong* __restrict__ val, dlong* __restrict__ colIds, dlong* __restrict__ G2L) {
  const dlong i = z + (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= N) return;
  const dlong g = gnum[i] - 1, row = grow[g];
  colIds[rowStarts[row] + (i - gstart[g])] = val[i];
  if (G2L) G2L[val[i]] = row;
}

// SetupRows: key = row (masked nodes: Nrows, sorted last); bad = 1 for a row
// index >= Nrows
__global__ void k_rows_keys(dlong N, dlong Nrows, const dlong* __restrict__ rowOf, unsigned* __restrict__ key,
                            dlong* __restrict__ val, int* __restrict__ bad) {
  const dlong n = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (n >= N) return;
  const dlong r = rowOf[n];
  if (r >= Nrows) *bad = 1;
  key[n] = (unsigned)(r < 0 || r >= Nrows ? Nrows : r);
  val[n] = n;
}

// row starts from the sorted row keys: entry i opens rows key[i-1]+1 ..
// key[i]; nnz = first index whose key is Nrows
__global__ void k_rows_starts(dlong N, dlong Nrows, const unsigned* __restrict__ key,
                              dlong* __restrict__ rowStarts) {
  const dlong i = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i > N) return;
  const dlong kc = i < N ? (dlong)key[i] : Nrows;
  const dlong kp = i > 0 ? (dlong)key[i - 1] : -1;
  for (dlong r = kp + 1; r <= kc; ++r) rowStarts[r] = i;
}

__global__ void k_row_sizes(dlong Nrows, const dlong* __restrict__ rowStarts, dlong* __restrict__ size) {
  const dlong r = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r < Nrows) size[r] = rowStarts[r + 1] - rowStarts[r];
}

// one bisection level: block b = rows [cur[b], cur[b+1]); an oversized block
// splits at the first row in (s, e) whose start exceeds start + ceil(size/2)
// (the last row when none does)
__global__ void k_bisect(dlong nb, const dlong* __restrict__ cur, const dlong* __restrict__ rs, dlong maxNnz,
                         dlong* __restrict__ cnt, dlong* __restrict__ mid) {
  const dlong b = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (b >= nb) return;
  const dlong s = cur[b], e = cur[b + 1];
  const dlong size = rs[e] - rs[s];
  if (size > maxNnz) {
    const dlong target = rs[s] + (size + 1) / 2;
    dlong lo = s, hi = e;  // upper bound of target in rs[s .. e-1]
    while (lo < hi) {
      const dlong m = lo + (hi - lo) / 2;
      if (rs[m] <= target)
        lo = m + 1;
      else
        hi = m;
    }
    if (lo == e) --lo;
    mid[b] = lo;
    cnt[b] = 2;
  } else {
    cnt[b] = 1;
  }
}

__global__ void k_bisect_emit(dlong nb, const dlong* __restrict__ cur, const dlong* __restrict__ cnt,
                              const dlong* __restrict__ mid, const dlong* __restrict__ off,
                              dlong* __restrict__ next) {
  const dlong b = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (b >= nb) return;
  if (b == 0) next[0] = 0;
  const dlong o = 1 + off[b];
  if (cnt[b] == 2) {
    next[o] = mid[b];
    next[o + 1] = cur[b + 1];
  } else {
    next[o] = cur[b + 1];
  }
}

__global__ void k_set2(dlong a, dlong b, dlong* __restrict__ out) {
  out[0] = a;
  out[1] = b;
}

}  // namespace

// ------------------------------------------------------------- mesh -----

int mesh_t::SetupDevice(platform_t& p, const MeshOpts& o) {
  if (int st = SetupVertices(o)) return st;
  BPX_HIP(hipSetDevice(p.device));
  hipStream_t s = p.stream;
  onDevice = true;
  devPlatform = &p;
  const size_t NL = (size_t)Nelements * Np;
  deviceBuffer<double> dEX, dEY, dEZ, dz, dw, dD;
  deviceBuffer<int> bad;
  if (int st = dEX.upload(EX, s)) return st;
  if (int st = dEY.upload(EY, s)) return st;
  if (int st = dEZ.upload(EZ, s)) return st;
  if (int st = dz.upload(gllz, s)) return st;
  if (int st = dw.upload(gllw, s)) return st;
  if (int st = dD.upload(D, s)) return st;
  if (int st = bad.alloc(1)) return st;
  if (int st = fill(s, 1, 0, bad.ptr())) return st;
  if (int st = o_x.alloc(NL)) return st;
  if (int st = o_y.alloc(NL)) return st;
  if (int st = o_z.alloc(NL)) return st;
  if (int st = o_globalIds.alloc(NL)) return st;
  hipLaunchKernelGGL(k_mesh_nodes, dim3(blocks_for(NL)), dim3(kT), 0, s, Nelements, Nq, o.nx, o.ny, dz.ptr(),
                     dEX.ptr(), dEY.ptr(), dEZ.ptr(), o_x.ptr(), o_y.ptr(), o_z.ptr(), o_globalIds.ptr());
  BPX_HIP(hipGetLastError());
  if (o.affine) {
    if (int st = o_wJ.alloc(Nelements)) return st;
    if (int st = o_ggeo.alloc((size_t)Nelements * 6)) return st;
    hipLaunchKernelGGL(k_geo_affine, dim3(blocks_for(Nelements)), dim3(kT), 0, s, Nelements, dEX.ptr(), dEY.ptr(),
                       dEZ.ptr(), o_wJ.ptr(), o_ggeo.ptr(), bad.ptr());
  } else {
    if (int st = o_wJ.alloc(NL)) return st;
    if (int st = o_ggeo.alloc(NL * 6)) return st;
    hipLaunchKernelGGL(k_geo_nodes, dim3(blocks_for(NL)), dim3(kT), 0, s, Nelements, Nq, dD.ptr(), dw.ptr(),
                       o_x.ptr(), o_y.ptr(), o_z.ptr(), o_wJ.ptr(), o_ggeo.ptr(), bad.ptr());
  }
  BPX_HIP(hipGetLastError());
  int hbad = 0;
  if (int st = read_value(s, bad.ptr(), hbad)) return st;
  if (hbad) return set_error(BPX_ERR_NEGATIVE_JACOBIAN, "mesh: negative or vanishing Jacobian found");

  if (o.cubature && !o.affine) {
    const int cq = cubNq;
    const size_t n0 = (size_t)9 * Np, n1 = (size_t)9 * Nq * Nq * cq, n2 = (size_t)9 * Nq * cq * cq;
    // elements per chunk: about 512 MiB of scratch
    const size_t perEl = (n0 + n1 + n2) * sizeof(double);
    const dlong chunk = (dlong)std::max<size_t>(1, std::min<size_t>((size_t)Nelements, ((size_t)512 << 20) / perEl));
    deviceBuffer<double> a0, a1, a2, dI, dcw;
    if (int st = a0.alloc(n0 * chunk)) return st;
    if (int st = a1.alloc(n1 * chunk)) return st;
    if (int st = a2.alloc(n2 * chunk)) return st;
    if (int st = dI.upload(cubInterp, s)) return st;
    if (int st = dcw.upload(cubw, s)) return st;
    if (int st = o_cubwJ.alloc((size_t)Nelements * cubNp)) return st;
    if (int st = o_cubggeo.alloc((size_t)Nelements * cubNp * 6)) return st;
    for (dlong e0 = 0; e0 < Nelements; e0 += chunk) {
      const dlong Nc = std::min(chunk, Nelements - e0);
      hipLaunchKernelGGL(k_cub_d0, dim3(blocks_for((size_t)Nc * Np)), dim3(kT), 0, s, e0, Nc, Nq, dD.ptr(),
                         o_x.ptr(), o_y.ptr(), o_z.ptr(), a0.ptr());
      hipLaunchKernelGGL(k_cub_r, dim3(blocks_for((size_t)Nc * n1)), dim3(kT), 0, s, Nc, Nq, cq, dI.ptr(),
                         a0.ptr(), a1.ptr());
      hipLaunchKernelGGL(k_cub_s, dim3(blocks_for((size_t)Nc * n2)), dim3(kT), 0, s, Nc, Nq, cq, dI.ptr(),
                         a1.ptr(), a2.ptr());
      hipLaunchKernelGGL(k_cub_t, dim3(blocks_for((size_t)Nc * cubNp)), dim3(kT), 0, s, e0, Nc, Nq, cq, dI.ptr(),
                         dcw.ptr(), a2.ptr(), o_cubwJ.ptr(), o_cubggeo.ptr(), bad.ptr());
      BPX_HIP(hipGetLastError());
    }
    if (int st = read_value(s, bad.ptr(), hbad)) return st;
    if (hbad) return set_error(BPX_ERR_NEGATIVE_JACOBIAN, "mesh: negative Jacobian at a cubature point");
  }
  BPX_HIP(hipStreamSynchronize(s));
  return BPX_OK;
}

int mesh_t::EnsureHost() {
  if (!onDevice || !x.empty() || Nelements == 0) return BPX_OK;
  BPX_HIP(hipSetDevice(devPlatform->device));
  hipStream_t s = devPlatform->stream;
  if (int st = o_x.download(x, s)) return st;
  if (int st = o_y.download(y, s)) return st;
  if (int st = o_z.download(z, s)) return st;
  if (int st = o_wJ.download(wJ, s)) return st;
  if (int st = o_ggeo.download(ggeo, s)) return st;
  if (int st = o_cubwJ.download(cubwJ, s)) return st;
  if (int st = o_cubggeo.download(cubggeo, s)) return st;
  return o_globalIds.download(globalIds, s);
}

bool mesh_t::NodesTrilinear() {
  if (!onDevice) return NodesAreTrilinear();
  if (hipSetDevice(devPlatform->device) != hipSuccess) return false;
  hipStream_t s = devPlatform->stream;
  std::vector<double> c;
  TrilinearCoefficients(c);
  deviceBuffer<double> dc, dEX, dEY, dEZ, dz;
  deviceBuffer<int> bad;
  if (dc.upload(c, s) || dEX.upload(EX, s) || dEY.upload(EY, s) || dEZ.upload(EZ, s) || dz.upload(gllz, s) ||
      bad.alloc(1) || fill(s, 1, 0, bad.ptr()))
    return false;
  const double scale = std::max(opts.dim[0], std::max(opts.dim[1], opts.dim[2]));
  const size_t NL = (size_t)Nelements * Np;
  if (NL > 0)
    hipLaunchKernelGGL(k_check_trilinear, dim3(blocks_for(NL)), dim3(kT), 0, s, Nelements, Nq, 1e-12 * scale,
                       dz.ptr(), dc.ptr(), dEX.ptr(), dEY.ptr(), dEZ.ptr(), o_x.ptr(), o_y.ptr(), o_z.ptr(),
                       bad.ptr());
  int hbad = 1;
  if (hipGetLastError() != hipSuccess || read_value(s, bad.ptr(), hbad)) return false;
  return hbad == 0;
}

int mesh_t::MaskedGlobalIdsDevice(int Nfields, hlong* o_ids) const {
  if (!onDevice) return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: not set up on the device");
  const size_t NL = (size_t)Nelements * Np;
  if (NL == 0) return BPX_OK;
  const long long LX = (long long)opts.nx * N + 1, LY = (long long)opts.ny * N + 1,
                  LZ = (long long)opts.nz * N + 1;
  hipLaunchKernelGGL(k_masked_ids, dim3(blocks_for(NL)), dim3(kT), 0, devPlatform->stream, Nelements, Nq, Nfields,
                     LX, LY, LZ, o_globalIds.ptr(), o_ids);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

// -------------------------------------------------------------- ogs -----
namespace ogs {

namespace {

constexpr dlong kBlockNnz = 512;  // include/ogs/ogsOperator.hpp:38

// row blocks of at most kBlockNnz entries by level-by-level bisection
int block_rows_device(hipStream_t s, deviceBuffer<char>& tmp, dlong Nrows, const dlong* rs,
                      deviceBuffer<dlong>& out, dlong& Nblocks) {
  out.release();
  Nblocks = 0;
  if (Nrows == 0) return BPX_OK;
  deviceBuffer<dlong> cur, next, cnt, mid, off, total;
  if (int st = cur.alloc(2)) return st;
  hipLaunchKernelGGL(k_set2, dim3(1), dim3(1), 0, s, (dlong)0, Nrows, cur.ptr());
  if (int st = total.alloc(1)) return st;
  dlong nb = 1;
  // each level at least halves the largest block: 2^31 entries need < 32
  for (int level = 0;; ++level) {
    if (level > 64) return set_error(BPX_ERR_OGS, "ogs: row blocks did not converge (row starts not ascending)");
    if (int st = cnt.alloc(nb)) return st;
    if (int st = mid.alloc(nb)) return st;
    if (int st = off.alloc(nb)) return st;
    hipLaunchKernelGGL(k_bisect, dim3(blocks_for(nb)), dim3(kT), 0, s, nb, cur.ptr(), rs, kBlockNnz, cnt.ptr(),
                       mid.ptr());
    BPX_HIP(hipGetLastError());
    if (int st = cub_call(tmp, [&](void* t, size_t& b) {
          return hipcub::DeviceScan::ExclusiveSum(t, b, cnt.ptr(), off.ptr(), (int)nb, s);
        }))
      return st;
    dlong lastOff = 0, lastCnt = 0;
    if (int st = read_value(s, off.ptr() + nb - 1, lastOff)) return st;
    if (int st = read_value(s, cnt.ptr() + nb - 1, lastCnt)) return st;
    const dlong nnext = lastOff + lastCnt;
    if (nnext == nb) break;
    if (int st = next.alloc((size_t)nnext + 1)) return st;
    hipLaunchKernelGGL(k_bisect_emit, dim3(blocks_for(nb)), dim3(kT), 0, s, nb, cur.ptr(), cnt.ptr(), mid.ptr(),
                       off.ptr(), next.ptr());
    BPX_HIP(hipGetLastError());
    cur = std::move(next);
    nb = nnext;
  }
  out = std::move(cur);
  Nblocks = nb;
  return BPX_OK;
}

// the largest row size; multiplicity above kBlockNnz is an error
int check_rows(hipStream_t s, deviceBuffer<char>& tmp, dlong Nrows, const dlong* rs) {
  if (Nrows == 0) return BPX_OK;
  deviceBuffer<dlong> size, mx;
  if (int st = size.alloc(Nrows)) return st;
  if (int st = mx.alloc(1)) return st;
  hipLaunchKernelGGL(k_row_sizes, dim3(blocks_for(Nrows)), dim3(kT), 0, s, Nrows, rs, size.ptr());
  BPX_HIP(hipGetLastError());
  if (int st = cub_call(tmp, [&](void* t, size_t& b) {
        return hipcub::DeviceReduce::Max(t, b, size.ptr(), mx.ptr(), (int)Nrows, s);
      }))
    return st;
  dlong m = 0;
  if (int st = read_value(s, mx.ptr(), m)) return st;
  if (m > kBlockNnz) return set_error(BPX_ERR_OGS, "ogs: a global node has multiplicity above 512");
  return BPX_OK;
}

int bits_for(unsigned long long v) {
  int b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

}  // namespace

int ogs_t::SetupDevice(platform_t& p, dlong N_, const hlong* o_ids, dlong* o_G2L) {
  if (N_ < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: N < 0");
  BPX_HIP(hipSetDevice(p.device));
  hipStream_t s = p.stream;
  N = N_;
  onDevice = true;
  rowStartsT.clear();
  colIdsT.clear();
  blockRowStartsT.clear();
  Ngather = nnzT = NrowBlocksT = 0;
  o_colIdsT.release();
  o_blockRowStartsT.release();
  if (int st = o_rowStartsT.alloc(1)) return st;
  if (int st = fill(s, 1, (dlong)0, o_rowStartsT.ptr())) return st;
  if (o_G2L)
    if (int st = fill(s, (size_t)N, (dlong)-1, o_G2L)) return st;
  if (N == 0) return BPX_OK;

  deviceBuffer<char> tmp;
  deviceBuffer<unsigned long long> k0, k1, kmax;
  deviceBuffer<dlong> v0, v1, head, gnum, first;
  if (int st = k0.alloc(N)) return st;
  if (int st = k1.alloc(N)) return st;
  if (int st = kmax.alloc(1)) return st;
  if (int st = v0.alloc(N)) return st;
  if (int st = v1.alloc(N)) return st;
  hipLaunchKernelGGL(k_ogs_keys, dim3(blocks_for(N)), dim3(kT), 0, s, N, o_ids, k0.ptr(), v0.ptr());
  BPX_HIP(hipGetLastError());
  if (int st = cub_call(tmp, [&](void* t, size_t& b) {
        return hipcub::DeviceReduce::Max(t, b, k0.ptr(), kmax.ptr(), (int)N, s);
      }))
    return st;
  unsigned long long maxKey = 0;
  if (int st = read_value(s, kmax.ptr(), maxKey)) return st;
  if (maxKey == 0) return BPX_OK;  // every node masked
  const int endBit = bits_for(maxKey);
  // stable: equal keys keep their local indices in ascending order
  if (int st = cub_call(tmp, [&](void* t, size_t& b) {
        return hipcub::DeviceRadixSort::SortPairs(t, b, k0.ptr(), k1.ptr(), v0.ptr(), v1.ptr(), (int)N, 0, endBit,
                                                  s);
      }))
    return st;
  k0.release();
  v0.release();
  if (int st = head.alloc(N)) return st;
  if (int st = gnum.alloc(N)) return st;
  if (int st = first.alloc(1)) return st;
  if (int st = fill(s, 1, N, first.ptr())) return st;
  hipLaunchKernelGGL(k_ogs_heads, dim3(blocks_for(N)), dim3(kT), 0, s, N, k1.ptr(), head.ptr(), first.ptr());
  BPX_HIP(hipGetLastError());
  if (int st = cub_call(tmp, [&](void* t, size_t& b) {
        return hipcub::DeviceScan::InclusiveSum(t, b, head.ptr(), gnum.ptr(), (int)N, s);
      }))
    return st;
  dlong Ngroups = 0, z = 0;
  if (int st = read_value(s, gnum.ptr() + N - 1, Ngroups)) return st;
  if (int st = read_value(s, first.ptr(), z)) return st;
  k1.release();

  deviceBuffer<dlong> gstart, firstFlag, rowAt, grow, cnt;
  if (int st = gstart.alloc(Ngroups)) return st;
  if (int st = firstFlag.alloc(N)) return st;
  if (int st = fill(s, (size_t)N, (dlong)0, firstFlag.ptr())) return st;
  hipLaunchKernelGGL(k_ogs_groups, dim3(blocks_for(N)), dim3(kT), 0, s, N, head.ptr(), gnum.ptr(), v1.ptr(),
                     gstart.ptr(), firstFlag.ptr());
  BPX_HIP(hipGetLastError());
  if (int st = rowAt.alloc(N)) return st;
  if (int st = cub_call(tmp, [&](void* t, size_t& b) {
        return hipcub::DeviceScan::ExclusiveSum(t, b, firstFlag.ptr(), rowAt.ptr(), (int)N, s);
      }))
    return st;
  if (int st = grow.alloc(Ngroups)) return st;
  if (int st = cnt.alloc((size_t)Ngroups + 1)) return st;
  hipLaunchKernelGGL(k_ogs_group_rows, dim3(blocks_for((size_t)Ngroups + 1)), dim3(kT), 0, s, Ngroups, N,
                     gstart.ptr(), v1.ptr(), rowAt.ptr(), grow.ptr(), cnt.ptr());
  BPX_HIP(hipGetLastError());
  if (int st = o_rowStartsT.alloc((size_t)Ngroups + 1)) return st;
  if (int st = cub_call(tmp, [&](void* t, size_t& b) {
        return hipcub::DeviceScan::ExclusiveSum(t, b, cnt.ptr(), o_rowStartsT.ptr(), (int)Ngroups + 1, s);
      }))
    return st;
  if (int st = check_rows(s, tmp, Ngroups, o_rowStartsT.ptr())) return st;
  const dlong nnz = N - z;
  if (int st = o_colIdsT.alloc(nnz)) return st;
  hipLaunchKernelGGL(k_ogs_cols, dim3(blocks_for(nnz)), dim3(kT), 0, s, N, z, gnum.ptr(), gstart.ptr(), grow.ptr(),
                     o_rowStartsT.ptr(), v1.ptr(), o_colIdsT.ptr(), o_G2L);
  BPX_HIP(hipGetLastError());
  Ngather = Ngroups;
  nnzT = nnz;
  if (int st = block_rows_device(s, tmp, Ngroups, o_rowStartsT.ptr(), o_blockRowStartsT, NrowBlocksT)) return st;
  BPX_HIP(hipStreamSynchronize(s));
  return BPX_OK;
}

int ogs_t::SetupRowsDevice(platform_t& p, dlong N_, dlong Nrows, const dlong* o_rowOf) {
  if (N_ < 0 || Nrows < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: N or Nrows < 0");
  BPX_HIP(hipSetDevice(p.device));
  hipStream_t s = p.stream;
  N = N_;
  onDevice = true;
  rowStartsT.clear();
  colIdsT.clear();
  blockRowStartsT.clear();
  Ngather = Nrows;
  nnzT = NrowBlocksT = 0;
  o_colIdsT.release();
  o_blockRowStartsT.release();
  if (int st = o_rowStartsT.alloc((size_t)Nrows + 1)) return st;
  deviceBuffer<char> tmp;
  deviceBuffer<unsigned> k0, k1;
  deviceBuffer<dlong> v0, v1;
  deviceBuffer<int> bad;
  if (int st = bad.alloc(1)) return st;
  if (int st = fill(s, 1, 0, bad.ptr())) return st;
  if (N > 0) {
    if (int st = k0.alloc(N)) return st;
    if (int st = k1.alloc(N)) return st;
    if (int st = v0.alloc(N)) return st;
    if (int st = v1.alloc(N)) return st;
    hipLaunchKernelGGL(k_rows_keys, dim3(blocks_for(N)), dim3(kT), 0, s, N, Nrows, o_rowOf, k0.ptr(), v0.ptr(),
                       bad.ptr());
    BPX_HIP(hipGetLastError());
    int hbad = 0;
    if (int st = read_value(s, bad.ptr(), hbad)) return st;
    if (hbad) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: row index out of range");
    const int endBit = bits_for((unsigned long long)Nrows);
    if (int st = cub_call(tmp, [&](void* t, size_t& b) {
          return hipcub::DeviceRadixSort::SortPairs(t, b, k0.ptr(), k1.ptr(), v0.ptr(), v1.ptr(), (int)N, 0, endBit,
                                                    s);
        }))
      return st;
  }
  hipLaunchKernelGGL(k_rows_starts, dim3(blocks_for((size_t)N + 1)), dim3(kT), 0, s, N, Nrows, k1.ptr(),
                     o_rowStartsT.ptr());
  BPX_HIP(hipGetLastError());
  dlong nnz = 0;
  if (int st = read_value(s, o_rowStartsT.ptr() + Nrows, nnz)) return st;
  nnzT = nnz;
  if (int st = o_colIdsT.copy_device(v1.ptr(), nnz, s)) return st;
  if (int st = check_rows(s, tmp, Nrows, o_rowStartsT.ptr())) return st;
  if (int st = block_rows_device(s, tmp, Nrows, o_rowStartsT.ptr(), o_blockRowStartsT, NrowBlocksT)) return st;
  BPX_HIP(hipStreamSynchronize(s));
  return BPX_OK;
}

int ogs_t::EnsureHost(platform_t& p) {
  if (!onDevice || !rowStartsT.empty()) return BPX_OK;
  BPX_HIP(hipSetDevice(p.device));
  if (int st = o_rowStartsT.download(rowStartsT, p.stream)) return st;
  if (int st = o_colIdsT.download(colIdsT, p.stream)) return st;
  return o_blockRowStartsT.download(blockRowStartsT, p.stream);
}

}  // namespace ogs

// ------------------------------------------------- fused BP5 maps -------
namespace {

// shared[r] = row r has two or more local copies
__global__ void k_fused_shared(dlong Ng, const dlong* __restrict__ rs, dlong* __restrict__ shared) {
  const dlong r = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r < Ng) shared[r] = (rs[r + 1] - rs[r]) >= 2 ? 1 : 0;
}

// the fused form needs every interior node on a single-copy row and every
// surface node masked or on a shared row
__global__ void k_fused_check(dlong Nel, int Nq, const dlong* __restrict__ G2L, const dlong* __restrict__ shared,
                              int* __restrict__ bad) {
  const int Np = Nq * Nq * Nq;
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= (size_t)Nel * Np) return;
  const int n = (int)(l % Np);
  const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
  const dlong row = G2L[l];
  const bool interior = fused_surface_index(Nq, i, j, k) < 0;
  if (interior ? (row < 0 || shared[row]) : (row >= 0 && !shared[row])) *bad = 1;
}

// solve rows: shared rows 0..Nshared-1 in reference order (sidx = exclusive
// sum of shared), then each element's interior run; rowOfSlot = solve row
// of each surface slot
__global__ void k_fused_rows(dlong Nel, int Nq, long long S0, int NIP, int NSP, const dlong* __restrict__ G2L,
                             const dlong* __restrict__ shared, const dlong* __restrict__ sidx,
                             dlong* __restrict__ srow, dlong* __restrict__ rowOfSlot) {
  const int Np = Nq * Nq * Nq;
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= (size_t)Nel * Np) return;
  const dlong e = (dlong)(l / Np);
  const int n = (int)(l - (size_t)e * Np);
  const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
  const dlong row = G2L[l];
  const int sp = fused_surface_index(Nq, i, j, k);
  if (sp < 0) {
    srow[row] = (dlong)(S0 + (long long)NIP * e + fused_interior_index(Nq, i, j, k));
  } else if (row >= 0) {
    rowOfSlot[(size_t)e * NSP + sp] = sidx[row];
  }
  (void)shared;
}

__global__ void k_fused_shared_rows(dlong Ng, const dlong* __restrict__ shared, const dlong* __restrict__ sidx,
                                    dlong* __restrict__ srow) {
  const dlong r = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r < Ng && shared[r]) srow[r] = sidx[r];
}

__global__ void k_fused_perm(dlong Ng, const dlong* __restrict__ srow, dlong* __restrict__ perm) {
  const dlong r = (dlong)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r < Ng && srow[r] >= 0) perm[srow[r]] = r;
}

__global__ void k_fused_g2ls(size_t NL, const dlong* __restrict__ G2L, const dlong* __restrict__ srow,
                             dlong* __restrict__ G2Ls) {
  const size_t l = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (l < NL) G2Ls[l] = G2L[l] < 0 ? -1 : srow[G2L[l]];
}

}  // namespace

int bp_t::PrepareFusedDevice() {
  fused = false;
  permuted = false;
  Ns = ogs.Ngather;
  if (bp != 5 || Nfields != 1) return BPX_OK;
  hipStream_t s = platform->stream;
  const dlong Ng = ogs.Ngather, Nel = mesh->Nelements;
  const int Nq = mesh->Nq;
  const long long NIP = fused_interior_pad(Nq), NSP = fused_surface_pad(Nq);
  const size_t NL = (size_t)NLocal;
  if (Ng == 0 || NL == 0) return BPX_OK;
  deviceBuffer<char> tmp;
  deviceBuffer<dlong> shared, sidx, srow;
  deviceBuffer<int> bad;
  if (int st = shared.alloc(Ng)) return st;
  if (int st = sidx.alloc(Ng)) return st;
  if (int st = bad.alloc(1)) return st;
  if (int st = fill(s, 1, 0, bad.ptr())) return st;
  hipLaunchKernelGGL(k_fused_shared, dim3(blocks_for(Ng)), dim3(kT), 0, s, Ng, ogs.o_rowStartsT.ptr(),
                     shared.ptr());
  hipLaunchKernelGGL(k_fused_check, dim3(blocks_for(NL)), dim3(kT), 0, s, Nel, Nq, o_GlobalToLocal.ptr(),
                     shared.ptr(), bad.ptr());
  BPX_HIP(hipGetLastError());
  int hbad = 0;
  if (int st = read_value(s, bad.ptr(), hbad)) return st;
  if (hbad) return BPX_OK;
  if (int st = cub_call(tmp, [&](void* t, size_t& b) {
        return hipcub::DeviceScan::ExclusiveSum(t, b, shared.ptr(), sidx.ptr(), (int)Ng, s);
      }))
    return st;
  dlong lastIdx = 0, lastShared = 0;
  if (int st = read_value(s, sidx.ptr() + Ng - 1, lastIdx)) return st;
  if (int st = read_value(s, shared.ptr() + Ng - 1, lastShared)) return st;
  const dlong Nshared = lastIdx + lastShared;
  const long long S0 = fused_pad(Nshared);
  const long long ns = S0 + NIP * Nel, nslots = NSP * Nel;
  if (ns >= (1LL << 31) - 1 || nslots >= (1LL << 31) - 1) return BPX_OK;  // 32-bit row and slot numbers
  deviceBuffer<dlong> rowOfSlot, perm;
  if (int st = srow.alloc(Ng)) return st;
  if (int st = fill(s, (size_t)Ng, (dlong)-1, srow.ptr())) return st;
  if (int st = rowOfSlot.alloc((size_t)nslots)) return st;
  if (int st = fill(s, (size_t)nslots, (dlong)-1, rowOfSlot.ptr())) return st;
  hipLaunchKernelGGL(k_fused_shared_rows, dim3(blocks_for(Ng)), dim3(kT), 0, s, Ng, shared.ptr(), sidx.ptr(),
                     srow.ptr());
  hipLaunchKernelGGL(k_fused_rows, dim3(blocks_for(NL)), dim3(kT), 0, s, Nel, Nq, S0, (int)NIP, (int)NSP,
                     o_GlobalToLocal.ptr(), shared.ptr(), sidx.ptr(), srow.ptr(), rowOfSlot.ptr());
  BPX_HIP(hipGetLastError());
  Ns = (dlong)ns;
  if (int st = perm.alloc((size_t)Ns)) return st;
  if (int st = fill(s, (size_t)Ns, (dlong)-1, perm.ptr())) return st;
  hipLaunchKernelGGL(k_fused_perm, dim3(blocks_for(Ng)), dim3(kT), 0, s, Ng, srow.ptr(), perm.ptr());
  if (int st = o_G2Ls.alloc(NL)) return st;
  hipLaunchKernelGGL(k_fused_g2ls, dim3(blocks_for(NL)), dim3(kT), 0, s, NL, o_GlobalToLocal.ptr(), srow.ptr(),
                     o_G2Ls.ptr());
  BPX_HIP(hipGetLastError());
  if (int st = solveOgs.SetupRowsDevice(*platform, NLocal, Ns, o_G2Ls.ptr())) return st;
  if (int st = sharedOgs.SetupRowsDevice(*platform, (dlong)nslots, Nshared, rowOfSlot.ptr())) return st;
  o_sperm = std::move(perm);
  if (int st = o_slots.alloc((size_t)nslots)) return st;
  if (int st = o_sx.alloc(Ns)) return st;
  if (int st = o_sr.alloc(Ns)) return st;
  naxParts = bp5_workgroups(mesh->N, Nel);
  if (int st = o_partAx.alloc(std::max(naxParts, 1))) return st;
  fusedS0 = (dlong)S0;
  permuted = true;
  fused = true;
  BPX_HIP(hipStreamSynchronize(s));
  return BPX_OK;
}

}  // namespace bpx

// Forcing term of the BP drivers at the local nodes.
//
// Semantics: okl/rhsHex3D.okl:28-128. One field:
//   rhs = W * (3 pi^2 + lambda) sin(pi x) sin(pi y) sin(pi z);
// three fields ([element][field][node]):
//   (4 pi^2 + lambda) * {sin(2 pi x) sin(pi y) sin(pi z), ...}.
// W = wJ per node, or J * w_i w_j w_k on affine meshes.
#include <hip/hip_runtime.h>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

constexpr double kPi = 3.14159265358979323846;

__global__ __launch_bounds__(256) void k_rhs(int Nq, int Nfields, int affine, long long Nlocal_nodes,
                                             const double* __restrict__ wJ, const double* __restrict__ gllw,
                                             const double* __restrict__ x, const double* __restrict__ y,
                                             const double* __restrict__ z, double lambda,
                                             double* __restrict__ rhs) {
  const int Np = Nq * Nq * Nq;
  for (long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x; id < Nlocal_nodes;
       id += (long long)gridDim.x * blockDim.x) {
    const long long e = id / Np;
    const int n = (int)(id - e * Np);
    const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
    const double W = affine ? wJ[e] * gllw[i] * gllw[j] * gllw[k] : wJ[id];
    const double X = x[id], Y = y[id], Z = z[id];
    if (Nfields == 1) {
      rhs[id] = W * ((3 * kPi * kPi + lambda) * sin(kPi * X) * sin(kPi * Y) * sin(kPi * Z));
    } else {
      const double c = 4 * kPi * kPi + lambda;
      const long long base = e * Np * Nfields + n;
      rhs[base + 0 * Np] = W * (c * sin(2.0 * kPi * X) * sin(kPi * Y) * sin(kPi * Z));
      rhs[base + 1 * Np] = W * (c * sin(kPi * X) * sin(2.0 * kPi * Y) * sin(kPi * Z));
      rhs[base + 2 * Np] = W * (c * sin(kPi * X) * sin(kPi * Y) * sin(2.0 * kPi * Z));
    }
  }
}

}  // namespace

int launch_rhs(hipStream_t s, int N, int Nfields, int affine, int Nelements, const double* wJ,
               const double* gllw, const double* x, const double* y, const double* z, double lambda,
               double* rhs) {
  if (N < 1 || N > 15 || (Nfields != 1 && Nfields != 3) || Nelements < 0)
    return set_error(BPX_ERR_INVALID_ARGUMENT, "rhs: bad N / Nfields / Nelements");
  if (Nelements == 0) return BPX_OK;
  if (!wJ || !x || !y || !z || !rhs || (affine && !gllw))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "rhs: null device pointer");
  const int Nq = N + 1;
  const long long nodes = (long long)Nelements * Nq * Nq * Nq;
  long long blocks = (nodes + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_rhs, dim3((unsigned)blocks), dim3(256), 0, s, Nq, Nfields, affine, nodes, wJ, gllw,
                     x, y, z, lambda, rhs);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

}  // namespace bpx

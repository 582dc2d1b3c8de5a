// Error of a discrete solution against the manufactured solution of the
// forcing kernel (okl/rhsHex3D.okl:30-33): for one field, -lap u + lambda u
// = (3 pi^2 + lambda) u with u = sin(pi x) sin(pi y) sin(pi z), which
// vanishes on the reference's box boundary (and on [0,1]^3).
//
// One pass over the local nodes: node n of element e reads the solution
// through GlobalToLocal (-1 = Dirichlet node, value 0), compares it with
// c * u(x, y, z) at the node and accumulates max |diff| and the quadrature
// sums sum wJ diff^2 and sum wJ (c u)^2 (wJ = GLL weights times the
// Jacobian, so the sums are the integrals of each element's interpolants).
// A fixed grid of kRedBlocks workgroups writes one partial per workgroup;
// the host folds the partials in order, so the result does not depend on
// timing.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

constexpr int VT = 256;

__global__ __launch_bounds__(VT) void k_error_norms(long long NLocal, const int* __restrict__ G2L,
                                                    const double* __restrict__ u, const double* __restrict__ x,
                                                    const double* __restrict__ y, const double* __restrict__ z,
                                                    const double* __restrict__ wJ, double c,
                                                    double* __restrict__ part) {
  __shared__ double s_max[VT], s_e2[VT], s_u2[VT];
  double mx = 0, e2 = 0, u2 = 0;
  const long long stride = (long long)gridDim.x * VT;
  for (long long n = (long long)blockIdx.x * VT + threadIdx.x; n < NLocal; n += stride) {
    const int id = G2L[n];
    const double uh = id >= 0 ? u[id] : 0.0;
    const double ex = c * sin(M_PI * x[n]) * sin(M_PI * y[n]) * sin(M_PI * z[n]);
    const double d = uh - ex;
    mx = fmax(mx, fabs(d));
    e2 += wJ[n] * d * d;
    u2 += wJ[n] * ex * ex;
  }
  s_max[threadIdx.x] = mx;
  s_e2[threadIdx.x] = e2;
  s_u2[threadIdx.x] = u2;
  __syncthreads();
  for (int h = VT / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + h]);
      s_e2[threadIdx.x] += s_e2[threadIdx.x + h];
      s_u2[threadIdx.x] += s_u2[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = s_max[0];
    part[3 * blockIdx.x + 1] = s_e2[0];
    part[3 * blockIdx.x + 2] = s_u2[0];
  }
}

}  // namespace

int launch_error_norms(hipStream_t s, long long NLocal, const int* G2L, const double* u, const double* x,
                       const double* y, const double* z, const double* wJ, double c, double* part) {
  if (NLocal < 0 || !part || (NLocal > 0 && (!G2L || !u || !x || !y || !z || !wJ)))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "error norms: bad arguments");
  hipLaunchKernelGGL(k_error_norms, dim3(kRedBlocks), dim3(VT), 0, s, NLocal, G2L, u, x, y, z, wJ, c, part);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

}  // namespace bpx

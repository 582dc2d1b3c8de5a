// Variable-coefficient BP5 (kappa on the node lattice) and the adjoint
// gradient: geometry scaling. No reference counterpart (the reference's BP5
// has a constant coefficient, okl/bp5AxHex3D.okl:415-569).
//
// The operator with a nodal coefficient is BP5's with G_n replaced by
// kappa_n G_n, so the solve reuses the BP5 Ax kernel on a scaled copy of
// the geometric factors, built once per kappa by k_kappa_geometry (one pass
// over ggeo). An affine mesh stores G per element, which cannot carry a
// nodal kappa, so there the copy is expanded to per-node factors
// (J G_e w_a w_b w_c kappa_n and J w_a w_b w_c) for the non-affine kernel.
#include <hip/hip_runtime.h>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

__global__ __launch_bounds__(256) void k_kappa_geometry(long long NLocal, int Nq, int affine,
                                                        const int* __restrict__ lattice,
                                                        const double* __restrict__ kappa,
                                                        const double* __restrict__ ggeo,
                                                        const double* __restrict__ wJ,
                                                        const double* __restrict__ gllw,
                                                        double* __restrict__ ggeoK, double* __restrict__ wJK) {
  const int Np = Nq * Nq * Nq;
  for (long long l = (long long)blockIdx.x * blockDim.x + threadIdx.x; l < NLocal;
       l += (long long)gridDim.x * blockDim.x) {
    const double kap = kappa[lattice[l]];
    if (affine) {
      const long long e = l / Np;
      const int n = (int)(l - e * Np);
      const int a = n % Nq, b = (n / Nq) % Nq, c = n / (Nq * Nq);
      const double w = gllw[a] * gllw[b] * gllw[c];
      for (int m = 0; m < 6; ++m) ggeoK[6 * l + m] = kap * (ggeo[6 * e + m] * w);
      wJK[l] = wJ[e] * w;
    } else {
      for (int m = 0; m < 6; ++m) ggeoK[6 * l + m] = kap * ggeo[6 * l + m];
    }
  }
}

}  // namespace

int launch_kappa_geometry(hipStream_t s, long long NLocal, int Nq, int affine, const int* lattice,
                          const double* kappa, const double* ggeo, const double* wJ, const double* gllw,
                          double* ggeoK, double* wJK) {
  if (NLocal <= 0) return BPX_OK;
  if (!lattice || !kappa || !ggeo || !ggeoK || (affine && (!wJ || !gllw || !wJK)))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "kappa geometry: null pointer");
  long long nb = (NLocal + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_kappa_geometry, dim3((unsigned)nb), dim3(256), 0, s, NLocal, Nq, affine, lattice, kappa,
                     ggeo, wJ, gllw, ggeoK, wJK);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

}  // namespace bpx

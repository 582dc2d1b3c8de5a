// BP1-BP6 Ax and the forcing on quadrilaterals, triangles and tetrahedra
// (FP64).
//
// Reference semantics:
//  - quads: okl/bp5AxQuad2D.okl (GLL collocation: Aq = D^T G D q +
//    lambda wJ q with per-node G and wJ), okl/bp1AxQuad2D.okl (mass on the
//    (N+2)-point Gauss grid: Aq = I^T W I q), okl/bp3AxQuad2D.okl (stiffness
//    on the Gauss grid: I^T (cubD^T G cubD + lambda W) I q); BP2, BP4, BP6
//    apply the same to each of their 2 fields ([element][field][node]).
//  - triangles and tetrahedra (affine): okl/bp5AxAffineTri2D.okl,
//    bp5AxAffineTet3D.okl, bp3AxAffine*.okl (Aq = sum_ab G_ab S_ab q +
//    lambda J MM q with the packed stiffness matrices S and per-element G
//    and J) and bp1AxAffine*.okl (Aq = J MM q); BP2/4/6 per field (2 fields
//    on triangles, 3 on tetrahedra).
//  - forcing: okl/rhsQuad2D.okl (rhs = wJ f at the nodes), okl/rhsTri2D.okl
//    and rhsTet3D.okl (rhs = MM (J f)), with f = (2 pi^2 + lambda) sin(pi x)
//    sin(pi y) in 2D (two fields: (3 pi^2 + lambda) {sin(2 pi x) sin(pi y),
//    sin(pi x) sin(2 pi y)}) and the hex box's forcing in 3D.
//
// Kernels: a quad's element runs in one workgroup, one lane per point of
// its GLL or Gauss grid with the grid values in LDS. The simplex operator is
// one rocBLAS DGEMM over all elements between a gather and a combine kernel
// (launch_simplex_ax_gemm).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------- simplices ---

// The simplex operator as one library GEMM: the element values gathered
// into Q (Np x Nel*NF, column e*NF + f), Z = Sall Q with Sall the reference
// matrices stacked by rows (the NG stiffness blocks A_c[n][k] = S_c[k][n],
// then MM; MM alone for BP1/BP2), then each element's factors applied to
// its column of Z. Each dense matrix entry is read once per GEMM tile of
// elements instead of once per element.
__global__ __launch_bounds__(256) void k_simplex_gather(long long n, int Np, int NF, const int* __restrict__ elementList,
                                                        const int* __restrict__ G2L, const double* __restrict__ q,
                                                        double* __restrict__ Q) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long col = i / Np;
    const int node = (int)(i - col * Np);
    const int eo = (int)(col / NF), f = (int)(col - (long long)eo * NF);
    const int id = G2L[((size_t)elementList[eo] * NF + f) * Np + node];
    Q[i] = id >= 0 ? q[id] : 0.0;
  }
}

template <int NG, bool STIFF>
__global__ __launch_bounds__(256) void k_simplex_combine(long long n, int Np, int NF, const int* __restrict__ elementList,
                                                         const double* __restrict__ wJ,
                                                         const double* __restrict__ ggeo, double lambda,
                                                         const double* __restrict__ Z, double* __restrict__ Aq) {
  const long long ldz = (long long)(STIFF ? NG + 1 : 1) * Np;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const long long col = i / Np;
    const int node = (int)(i - col * Np);
    const int eo = (int)(col / NF), f = (int)(col - (long long)eo * NF);
    const int e = elementList[eo];
    const double* z = Z + col * ldz + node;
    const double J = wJ[e];
    double out;
    if (STIFF) {
      out = 0;
#pragma unroll
      for (int c = 0; c < NG; ++c) out += ggeo[(size_t)e * NG + c] * z[(size_t)c * Np];
      out += J * lambda * z[(size_t)NG * Np];
    } else {
      out = J * z[0];
    }
    Aq[((size_t)e * NF + f) * Np + node] = out;
  }
}

// --------------------------------------------------------------- quads ---
enum { QUAD_BP5 = 0, QUAD_BP1 = 1, QUAD_BP3 = 2 };

// one element per workgroup, lanes over the MQ x MQ grid (MQ = Nq for the
// GLL forms, Nq + 1 for the Gauss forms); LDS: the Nq x Nq values, two
// MQ x MQ (or Nq x MQ) work tiles
template <int MODE>
__global__ __launch_bounds__(320) void k_quad_ax(int Nq, int NF, const int* __restrict__ elementList,
                                                 const int* __restrict__ G2L, const double* __restrict__ W,
                                                 const double* __restrict__ G, const double* __restrict__ Dm,
                                                 const double* __restrict__ I, double lambda,
                                                 const double* __restrict__ q, double* __restrict__ Aq) {
  __shared__ double s_q[16 * 16], s_a[17 * 17], s_b[17 * 17], s_c[17 * 17];
  // the 1D matrices, staged once: the inner loops then read LDS instead of
  // issuing one global (cache) load per multiply-add
  __shared__ double s_Dm[17 * 17], s_I[17 * 16];
  const int MQ = MODE == QUAD_BP5 ? Nq : Nq + 1;
  const int Np = Nq * Nq, MP = MQ * MQ;
  const int e = elementList[blockIdx.x];
  const int t = threadIdx.x;
  for (int n = t; n < MQ * MQ; n += blockDim.x) s_Dm[n] = Dm[n];
  if (MODE != QUAD_BP5)
    for (int n = t; n < MQ * Nq; n += blockDim.x) s_I[n] = I[n];
  __syncthreads();
  const int a = t % MQ, b = t / MQ;  // the lane's grid point (a along r, b along s)
  const bool on = t < MP;
  for (int f = 0; f < NF; ++f) {
    const size_t lb = ((size_t)e * NF + f) * Np;
    if (t < Np) {
      const int id = G2L[lb + t];
      s_q[t] = id >= 0 ? q[id] : 0.0;
    }
    __syncthreads();
    if (MODE == QUAD_BP5) {
      // qr, qs at node (a, b); G q -> tiles; Aq = D^T (G q) + lambda wJ q
      double gr = 0, gs = 0, qn = 0;
      if (on) {
        double qr = 0, qs = 0;
        for (int m = 0; m < Nq; ++m) {
          qr += s_Dm[a * Nq + m] * s_q[b * Nq + m];
          qs += s_Dm[b * Nq + m] * s_q[m * Nq + a];
        }
        const size_t gb = (size_t)e * Np + t;
        const double G00 = G[3 * gb + 0], G01 = G[3 * gb + 1], G11 = G[3 * gb + 2];
        gr = G00 * qr + G01 * qs;
        gs = G01 * qr + G11 * qs;
        qn = s_q[t] * W[gb];
        s_a[t] = gr;
        s_b[t] = gs;
      }
      __syncthreads();
      if (on) {
        double acc = lambda * qn;
        for (int m = 0; m < Nq; ++m) {
          acc += s_Dm[m * Nq + a] * s_a[b * Nq + m];
          acc += s_Dm[m * Nq + b] * s_b[m * Nq + a];
        }
        Aq[lb + t] = acc;
      }
    } else {
      // interpolate along r: s_a[j][a] (Nq x MQ), then along s: u(a, b)
      if (t < Nq * MQ) {
        const int j = t / MQ, aa = t % MQ;
        double acc = 0;
        for (int i = 0; i < Nq; ++i) acc += s_I[aa * Nq + i] * s_q[j * Nq + i];
        s_a[j * MQ + aa] = acc;
      }
      __syncthreads();
      double u = 0;
      if (on) {
        for (int j = 0; j < Nq; ++j) u += s_I[b * Nq + j] * s_a[j * MQ + a];
        s_b[t] = u;
      }
      __syncthreads();
      double res = 0;
      const size_t cb = (size_t)e * MP + t;
      if (MODE == QUAD_BP1) {
        if (on) res = W[cb] * u;
      } else {
        // gradient on the Gauss grid, fluxes, weak divergence
        double fr = 0, fs = 0;
        if (on) {
          double ur = 0, us = 0;
          for (int m = 0; m < MQ; ++m) {
            ur += s_Dm[a * MQ + m] * s_b[b * MQ + m];
            us += s_Dm[b * MQ + m] * s_b[m * MQ + a];
          }
          const double G00 = G[3 * cb + 0], G01 = G[3 * cb + 1], G11 = G[3 * cb + 2];
          fr = G00 * ur + G01 * us;
          fs = G01 * ur + G11 * us;
        }
        __syncthreads();  // every lane has read s_b
        if (on) {
          s_c[t] = fr;
          s_a[t] = fs;  // s_a is free: the interpolation along s has read it
        }
        __syncthreads();
        if (on) {
          res = lambda * W[cb] * u;
          for (int m = 0; m < MQ; ++m) {
            res += s_Dm[m * MQ + a] * s_c[b * MQ + m];
            res += s_Dm[m * MQ + b] * s_a[m * MQ + a];
          }
        }
      }
      __syncthreads();  // the tiles are reused below
      if (on) s_b[t] = res;
      __syncthreads();
      // back along s: s_a[j][a] = sum_b I[b][j] res(a, b)
      if (t < Nq * MQ) {
        const int j = t / MQ, aa = t % MQ;
        double acc = 0;
        for (int bb = 0; bb < MQ; ++bb) acc += s_I[bb * Nq + j] * s_b[bb * MQ + aa];
        s_a[j * MQ + aa] = acc;
      }
      __syncthreads();
      // back along r: Aq[j][i] = sum_a I[a][i] s_a[j][a]
      if (t < Np) {
        const int j = t / Nq, i = t % Nq;
        double acc = 0;
        for (int aa = 0; aa < MQ; ++aa) acc += s_I[aa * Nq + i] * s_a[j * MQ + aa];
        Aq[lb + t] = acc;
      }
    }
    __syncthreads();  // the next field rewrites the tiles
  }
}

// ------------------------------------------------------------ forcing ---

__device__ inline void forcing(int dim, int NF, double x, double y, double z, double lambda, double* f) {
  if (dim == 2) {
    if (NF == 1) {
      f[0] = (2 * kPi * kPi + lambda) * sin(kPi * x) * sin(kPi * y);
    } else {
      f[0] = (3 * kPi * kPi + lambda) * sin(2.0 * kPi * x) * sin(kPi * y);
      f[1] = (3 * kPi * kPi + lambda) * sin(kPi * x) * sin(2.0 * kPi * y);
    }
  } else if (NF == 1) {
    f[0] = (3 * kPi * kPi + lambda) * sin(kPi * x) * sin(kPi * y) * sin(kPi * z);
  } else {
    const double c = 4 * kPi * kPi + lambda;
    f[0] = c * sin(2.0 * kPi * x) * sin(kPi * y) * sin(kPi * z);
    f[1] = c * sin(kPi * x) * sin(2.0 * kPi * y) * sin(kPi * z);
    f[2] = c * sin(kPi * x) * sin(kPi * y) * sin(2.0 * kPi * z);
  }
}

// quads: rhs = wJ f per node
__global__ __launch_bounds__(256) void k_rhs_quad(long long nodes, int Np, int NF, const double* __restrict__ wJ,
                                                  const double* __restrict__ x, const double* __restrict__ y,
                                                  double lambda, double* __restrict__ rhs) {
  const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= nodes) return;
  const long long e = id / Np;
  const int n = (int)(id - e * Np);
  double f[3];
  forcing(2, NF, x[id], y[id], 0.0, lambda, f);
  for (int c = 0; c < NF; ++c) rhs[(e * NF + c) * Np + n] = wJ[id] * f[c];
}

// simplices: rhs = MM (J f), one workgroup per element
__global__ __launch_bounds__(1024) void k_rhs_simplex(int Np, int dim, int NF, const double* __restrict__ wJ,
                                                      const double* __restrict__ MM, const double* __restrict__ x,
                                                      const double* __restrict__ y, const double* __restrict__ z,
                                                      double lambda, double* __restrict__ rhs) {
  extern __shared__ double s_f[];  // NF x Np
  const int e = blockIdx.x, n = threadIdx.x;
  if (n < Np) {
    const size_t id = (size_t)e * Np + n;
    double f[3];
    forcing(dim, NF, x[id], y[id], dim == 3 ? z[id] : 0.0, lambda, f);
    for (int c = 0; c < NF; ++c) s_f[c * Np + n] = wJ[e] * f[c];
  }
  __syncthreads();
  if (n >= Np) return;
  for (int c = 0; c < NF; ++c) {
    double acc = 0;
    for (int m = 0; m < Np; ++m) acc += MM[(size_t)m * Np + n] * s_f[c * Np + m];
    rhs[((size_t)e * NF + c) * Np + n] = acc;
  }
}

inline int round64(int n) { return (n + 63) / 64 * 64; }

}  // namespace

int simplex_blas_create(hipStream_t s, void** handle) {
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success)
    return set_error(BPX_ERR_HIP, "simplex Ax: rocblas_create_handle failed");
  if (rocblas_set_stream(h, s) != rocblas_status_success) {
    rocblas_destroy_handle(h);
    return set_error(BPX_ERR_HIP, "simplex Ax: rocblas_set_stream failed");
  }
  *handle = h;
  return BPX_OK;
}

void simplex_blas_destroy(void* handle) {
  if (handle) rocblas_destroy_handle((rocblas_handle)handle);
}

int launch_simplex_ax_gemm(hipStream_t s, void* blas, int bp, int dim, int Np, int Nfields, int Nelements,
                           const int* elementList, const int* G2L, const double* wJ, const double* ggeo,
                           const double* Sall, double* Q, double* Z, double lambda, const double* q, double* Aq) {
  if (Nelements <= 0) return BPX_OK;
  const bool stiff = bp >= 3;
  const int NG = dim == 2 ? 3 : 6;
  const int M = (stiff ? NG + 1 : 1) * Np;
  const long long cols = (long long)Nelements * Nfields, n = cols * Np;
  const unsigned grid = (unsigned)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_simplex_gather, dim3(grid), dim3(256), 0, s, n, Np, Nfields, elementList, G2L, q, Q);
  BPX_HIP(hipGetLastError());
  const double one = 1.0, zero = 0.0;
  if (rocblas_set_stream((rocblas_handle)blas, s) != rocblas_status_success ||
      rocblas_dgemm((rocblas_handle)blas, rocblas_operation_none, rocblas_operation_none, M, (rocblas_int)cols, Np,
                    &one, Sall, M, Q, Np, &zero, Z, M) != rocblas_status_success)
    return set_error(BPX_ERR_HIP, "simplex Ax: rocblas_dgemm failed");
  if (dim == 2) {
    if (stiff)
      hipLaunchKernelGGL((k_simplex_combine<3, true>), dim3(grid), dim3(256), 0, s, n, Np, Nfields, elementList, wJ,
                         ggeo, lambda, Z, Aq);
    else
      hipLaunchKernelGGL((k_simplex_combine<3, false>), dim3(grid), dim3(256), 0, s, n, Np, Nfields, elementList, wJ,
                         ggeo, lambda, Z, Aq);
  } else {
    if (stiff)
      hipLaunchKernelGGL((k_simplex_combine<6, true>), dim3(grid), dim3(256), 0, s, n, Np, Nfields, elementList, wJ,
                         ggeo, lambda, Z, Aq);
    else
      hipLaunchKernelGGL((k_simplex_combine<6, false>), dim3(grid), dim3(256), 0, s, n, Np, Nfields, elementList, wJ,
                         ggeo, lambda, Z, Aq);
  }
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}


int launch_quad_ax(hipStream_t s, int bp, int N, int Nfields, int Nelements, const int* elementList, const int* G2L,
                   const double* W, const double* G, const double* Dm, const double* I, double lambda,
                   const double* q, double* Aq) {
  if (Nelements <= 0) return BPX_OK;
  if (N < 1 || N > 15) return set_error(BPX_ERR_INVALID_ARGUMENT, "quad Ax: degree N must be in 1..15");
  const int Nq = N + 1;
  const dim3 grid(Nelements);
  if (bp == 5 || bp == 6) {
    hipLaunchKernelGGL(k_quad_ax<QUAD_BP5>, grid, dim3(round64(Nq * Nq)), 0, s, Nq, Nfields, elementList, G2L, W, G,
                       Dm, I, lambda, q, Aq);
  } else {
    const int lanes = round64((Nq + 1) * (Nq + 1));
    if (bp == 1 || bp == 2)
      hipLaunchKernelGGL(k_quad_ax<QUAD_BP1>, grid, dim3(lanes), 0, s, Nq, Nfields, elementList, G2L, W, G, Dm, I,
                         lambda, q, Aq);
    else
      hipLaunchKernelGGL(k_quad_ax<QUAD_BP3>, grid, dim3(lanes), 0, s, Nq, Nfields, elementList, G2L, W, G, Dm, I,
                         lambda, q, Aq);
  }
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_rhs_other(hipStream_t s, int elementType, int N, int Np, int Nfields, int Nelements, const double* wJ,
                     const double* MM, const double* x, const double* y, const double* z, double lambda,
                     double* rhs) {
  if (Nelements <= 0) return BPX_OK;
  if (elementType == 4) {
    const long long nodes = (long long)Nelements * Np;
    hipLaunchKernelGGL(k_rhs_quad, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, s, nodes, Np, Nfields, wJ, x,
                       y, lambda, rhs);
  } else {
    const int dim = elementType == 6 ? 3 : 2;
    if (Np > 1024) return set_error(BPX_ERR_UNSUPPORTED, "simplex forcing: more than 1024 nodes per element");
    hipLaunchKernelGGL(k_rhs_simplex, dim3(Nelements), dim3(round64(Np)), (size_t)Nfields * Np * sizeof(double), s,
                       Np, dim, Nfields, wJ, MM, x, y, z, lambda, rhs);
  }
  (void)N;
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

}  // namespace bpx

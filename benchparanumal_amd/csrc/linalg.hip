// linAlg vector kernels and the fused, device-resident CG steps.
//
// Semantics: libs/core/okl/linAlgSet.okl, linAlgAXPY.okl (y = a*x + b*y),
// linAlgInnerProd.okl / linAlgNorm2.okl (dot products), and cg::Solve /
// cg::UpdateCG of libs/core/linearSolverCG.cpp:62-156 with
// okl/linearSolverUpdateCG.okl (r -= alpha*Ap fused with r.r).
//
// Design (gfx950, 8 XCDs with private L2s): a reduction never synchronises
// workgroups inside a kernel. A streaming kernel with a fixed grid of
// kRedBlocks workgroups writes one FP64 partial per workgroup; the next
// kernel of the sequence (which needs the scalar anyway) folds the partials
// at its start, every workgroup in the same fixed order. Kernel boundaries
// provide the coherence, so there are no device-scope fences or atomics,
// and with a grid that depends only on kRedBlocks the results are
// bit-reproducible. The CG scalars (alpha, beta, stopping test, history)
// are computed on the device; the host only queues kernels, passing the
// iteration index as an argument, so an iteration never waits on the host
// (the reference waits 8 times per iteration, linearSolverCG.cpp:146 et al.).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

constexpr int RT = kRedThreads;
constexpr int RB = kRedBlocks;
constexpr int ST = 256;  // threads of the plain streaming kernels

inline int stream_blocks(dlong N) {
  long long nb = ((long long)N + ST - 1) / ST;
  if (nb < 1) nb = 1;
  if (nb > 16384) nb = 16384;
  return (int)nb;
}

// Fixed-order sum over the workgroup; every thread gets the total.
__device__ inline double block_allsum(double v, double* s_w) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  __syncthreads();  // s_w may still be read from a previous call
  if (lane == 0) s_w[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < nw; ++i) t += s_w[i];
    s_w[32] = t;
  }
  __syncthreads();
  return s_w[32];
}

// Sum of partials[0..n) in a fixed order, by one wavefront: lane l folds
// l, l+64, ... into four accumulators (so its loads are in flight
// together), then an xor butterfly (each step adds the same two operands on
// both lanes of a pair, so every lane ends with the same bits). Every
// wavefront that folds the same partials gets the same value, whatever
// kernel or workgroup size it runs in.
__device__ inline double wave_fold(const double* __restrict__ part, int n) {
  const int lane = threadIdx.x & 63;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  int i = lane;
  for (; i + 192 < n; i += 256) {
    a0 += part[i];
    a1 += part[i + 64];
    a2 += part[i + 128];
    a3 += part[i + 192];
  }
  for (; i < n; i += 64) a0 += part[i];
  double v = (a0 + a1) + (a2 + a3);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// wave_fold by wavefront 0, broadcast to the workgroup through LDS
__device__ inline double fold_partials(const double* __restrict__ part, int n, double* s_w) {
  if (threadIdx.x < 64) {
    const double v = wave_fold(part, n);
    if (threadIdx.x == 0) s_w[32] = v;
  }
  __syncthreads();
  return s_w[32];
}

// Grid-strided sum with four independent accumulators per lane.
template <typename F>
__device__ inline double strided_sum(dlong N, F term) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  for (; n + 3 * stride < N; n += 4 * stride) {
    a0 += term(n);
    a1 += term(n + stride);
    a2 += term(n + 2 * stride);
    a3 += term(n + 3 * stride);
  }
  for (; n < N; n += stride) a0 += term(n);
  return (a0 + a1) + (a2 + a3);
}

// Grid-strided loop, four independent elements per lane and step.
template <typename F>
__device__ inline void strided_for(dlong N, F body) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; n + 3 * stride < N; n += 4 * stride) {
    body(n);
    body(n + stride);
    body(n + 2 * stride);
    body(n + 3 * stride);
  }
  for (; n < N; n += stride) body(n);
}

// 16-byte vector view of a float stream: V4 floats per lane and access,
// used when every pointer of the stream is 16-byte aligned
template <typename T>
struct vec4;
template <>
struct vec4<float> {
  typedef float4 type;
};
template <>
struct vec4<double> {
  typedef double4 type;
};

template <typename T>
__global__ __launch_bounds__(ST) void k_set(dlong N, T alpha, T* __restrict__ x) {
  for (dlong n = blockIdx.x * ST + threadIdx.x; n < N; n += gridDim.x * ST) x[n] = alpha;
}

template <typename T>
__global__ __launch_bounds__(ST) void k_axpy(dlong N, T alpha, const T* __restrict__ x, T beta,
                                             T* __restrict__ y) {
  for (dlong n = blockIdx.x * ST + threadIdx.x; n < N; n += gridDim.x * ST) y[n] = alpha * x[n] + beta * y[n];
}

__global__ __launch_bounds__(ST) void k_axpy_f32_f64(dlong N, double alpha, const float* __restrict__ x,
                                                     double beta, double* __restrict__ y) {
  for (dlong n = blockIdx.x * ST + threadIdx.x; n < N; n += gridDim.x * ST)
    y[n] = alpha * (double)x[n] + beta * y[n];
}

__global__ __launch_bounds__(ST) void k_copy_f64_f32(dlong N, double scale, const double* __restrict__ x,
                                                     float* __restrict__ y) {
  for (dlong n = blockIdx.x * ST + threadIdx.x; n < N; n += gridDim.x * ST) y[n] = (float)(scale * x[n]);
}

template <typename T>
__global__ __launch_bounds__(RT) void k_dot_partials(dlong N, const T* __restrict__ x, const T* __restrict__ y,
                                                     double* __restrict__ part) {
  __shared__ double s_w[33];
  const double v = strided_sum(N, [&](long long n) { return (double)x[n] * (double)y[n]; });
  const double t = block_allsum(v, s_w);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(RT) void k_fold(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double s_w[33];
  const double t = fold_partials(part, n, s_w);
  if (threadIdx.x == 0) *out = t;
}

// r <- r - Ap (residual of the initial guess) with r.r partials
template <typename T>
__global__ __launch_bounds__(RT) void k_cg_init(dlong N, const T* __restrict__ Ap, T* __restrict__ r,
                                                double* __restrict__ part_rr) {
  __shared__ double s_w[33];
  const double v = strided_sum(N, [&](long long n) {
    const T rn = T(-1) * Ap[n] + T(1) * r[n];
    r[n] = rn;
    return (double)rn * (double)rn;
  });
  const double t = block_allsum(v, s_w);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = t;
}

// Start of iteration k: rr_k = r.r from the partials, stopping test,
// beta = rr_k / rr_{k-1}, p <- r + beta p (p <- r at k = 0). With pending
// (k > 0), first the x update of iteration k-1, x <- x + alpha_{k-1} p,
// which reads the same p (it is deferred to here from the end of iteration
// k-1 so that x and p share one pass). VEC: float4 accesses (all pointers
// 16-byte aligned, float only).
template <typename T, bool VEC>
__global__ __launch_bounds__(RT) void k_cg_update_p(dlong N, int k, const T* __restrict__ r, T* __restrict__ p,
                                                    T* __restrict__ x, const double* __restrict__ part_rr, int nrr,
                                                    CgState* st, double* __restrict__ hist_rr,
                                                    const double* __restrict__ hist_alpha, double tol, int pending) {
  __shared__ double s_w[33];
  if (k > 0 && (st->done || (st->stop_at != 0 && st->stop_at <= k))) return;
  typedef typename vec4<T>::type V;
  const V* r4 = reinterpret_cast<const V*>(r);
  V* p4 = reinterpret_cast<V*>(p);
  V* x4 = reinterpret_cast<V*>(x);
  // VEC: the lane's first four items are loaded before the fold of r.r
  const long long stride = (long long)gridDim.x * blockDim.x, N4 = N / 4;
  const long long n0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  V rq[VEC ? 4 : 1], pq[VEC ? 4 : 1], xq[VEC ? 4 : 1];
  if (VEC) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (n0 + j * stride < N4) {
        rq[j] = r4[n0 + j * stride];
        pq[j] = p4[n0 + j * stride];
        xq[j] = x4[n0 + j * stride];
      }
  }
  double rr = fold_partials(part_rr, nrr, s_w);
  double tol2;
  if (k == 0) {
    const double nrm = sqrt(rr);  // rdotr = norm2(r)^2 (linearSolverCG.cpp:84-85)
    rr = nrm * nrm;
    tol2 = fmax(tol * tol * rr, tol * tol);
  } else {
    tol2 = st->tol2;
  }
  const bool stop = !(rr > tol2) || !isfinite(rr);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    hist_rr[k] = rr;
    if (k == 0) {
      st->tol2 = tol2;
      st->breakdown = 0;
      st->done = 0;
      st->stop_at = 0;
    }
    if (stop) {
      st->iter = k;
      if (!isfinite(rr)) st->breakdown = 1;
      st->stop_at = k + 1;
    }
  }
  const bool apply = pending && k > 0;
  const T a = apply ? (T)hist_alpha[k - 1] : T(0);
  const bool first = (k == 0);
  const bool upd = !stop;
  const T b = first ? T(0) : (T)(rr / hist_rr[k - 1]);
  if (!apply && !upd) return;
  if (VEC) {
    auto step = [&](const V rv, V pv, V xv, long long n) {
      if (apply) {
        xv.x = a * pv.x + T(1) * xv.x; xv.y = a * pv.y + T(1) * xv.y;
        xv.z = a * pv.z + T(1) * xv.z; xv.w = a * pv.w + T(1) * xv.w;
        x4[n] = xv;
      }
      if (upd) {
        if (first) {
          pv = rv;
        } else {
          pv.x = T(1) * rv.x + b * pv.x;
          pv.y = T(1) * rv.y + b * pv.y;
          pv.z = T(1) * rv.z + b * pv.z;
          pv.w = T(1) * rv.w + b * pv.w;
        }
        p4[n] = pv;
      }
    };
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (n0 + j * stride < N4) step(rq[j], pq[j], xq[j], n0 + j * stride);
    for (long long n = n0 + 4 * stride; n < N4; n += stride) step(r4[n], p4[n], x4[n], n);
    const long long tail = (long long)(N / 4) * 4;
    const long long n = tail + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) {
      const T pv = p[n];
      if (apply) x[n] = a * pv + T(1) * x[n];
      if (upd) p[n] = first ? r[n] : T(1) * r[n] + b * pv;
    }
  } else {
    strided_for(N, [&](long long n) {
      const T pv = p[n];
      if (apply) x[n] = a * pv + T(1) * x[n];
      if (upd) p[n] = first ? r[n] : T(1) * r[n] + b * pv;
    });
  }
}

// End of iteration k: alpha = rr_k / p.Ap, r <- r - alpha Ap, r.r partials
// for iteration k+1 (x <- x + alpha p is left to the next k_cg_update_p,
// or to k_cg_flush_x after the last iteration). VEC: float4 accesses
// (float only).
template <typename T, bool VEC>
__global__ __launch_bounds__(RT) void k_cg_update_r(dlong N, int k, const T* __restrict__ Ap, T* __restrict__ r,
                                                    const double* __restrict__ part_pAp, int npAp,
                                                    double* __restrict__ part_rr, CgState* st,
                                                    const double* __restrict__ hist_rr,
                                                    double* __restrict__ hist_alpha) {
  __shared__ double s_w[33];
  if (st->done || st->stop_at) return;
  typedef typename vec4<T>::type V;
  const V* A4 = reinterpret_cast<const V*>(Ap);
  V* r4 = reinterpret_cast<V*>(r);
  // VEC: the lane's first four items are loaded before the fold of p.Ap,
  // so the stream is in flight while one wavefront folds the partials
  const long long stride = (long long)gridDim.x * blockDim.x, N4 = N / 4;
  const long long n0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  V rp[VEC ? 4 : 1], Ap4[VEC ? 4 : 1];
  if (VEC) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (n0 + j * stride < N4) {
        rp[j] = r4[n0 + j * stride];
        Ap4[j] = A4[n0 + j * stride];
      }
  }
  const double pAp = fold_partials(part_pAp, npAp, s_w);
  const double alpha = hist_rr[k] / pAp;
  if (!isfinite(alpha) || !(pAp > 0.0)) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->breakdown = 1;
      st->done = 1;
      st->iter = k;
    }
    return;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) hist_alpha[k] = alpha;
  const T a = (T)alpha;
  double v;
  if (VEC) {
    auto step = [&](V rv, const V Av, long long n) {
      rv.x = rv.x - a * Av.x; rv.y = rv.y - a * Av.y; rv.z = rv.z - a * Av.z; rv.w = rv.w - a * Av.w;
      r4[n] = rv;
      return ((double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y) +
             ((double)rv.z * (double)rv.z + (double)rv.w * (double)rv.w);
    };
    double acc[4] = {0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (n0 + j * stride < N4) acc[j] += step(rp[j], Ap4[j], n0 + j * stride);
    for (long long n = n0 + 4 * stride; n < N4; n += stride) acc[0] += step(r4[n], A4[n], n);
    v = (acc[0] + acc[1]) + (acc[2] + acc[3]);
    const long long tail = (long long)(N / 4) * 4;
    const long long n = tail + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < N) {
      const T rn = r[n] - a * Ap[n];
      r[n] = rn;
      v += (double)rn * (double)rn;
    }
  } else {
    v = strided_sum(N, [&](long long n) {
      const T rn = r[n] - a * Ap[n];
      r[n] = rn;
      return (double)rn * (double)rn;
    });
  }
  const double t = block_allsum(v, s_w);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = t;
}

// The deferred x update of iteration k: x <- x + alpha_k p. when_running:
// only if the recurrence has not stopped (a stop inside k_cg_update_p has
// applied it already, a breakdown leaves its iteration out).
template <typename T>
__global__ __launch_bounds__(ST) void k_cg_flush_x(dlong N, int k, const T* __restrict__ p, T* __restrict__ x,
                                                   const double* __restrict__ hist_alpha, const CgState* st,
                                                   int when_running) {
  if (when_running && (st->done || st->stop_at)) return;
  const T a = (T)hist_alpha[k];
  for (dlong n = blockIdx.x * ST + threadIdx.x; n < N; n += gridDim.x * ST) x[n] = a * p[n] + T(1) * x[n];
}

// After the last queued iteration kend: record rr_kend and the state.
__global__ __launch_bounds__(RT) void k_cg_finish(int kend, int maxit, const double* __restrict__ part_rr,
                                                  int nrr, CgState* st, double* __restrict__ hist_rr) {
  __shared__ double s_w[33];
  if (st->done) return;
  if (st->stop_at) {  // stopped inside the chunk: iter is already set
    if (threadIdx.x == 0) st->done = 1;
    return;
  }
  const double rr = fold_partials(part_rr, nrr, s_w);
  if (threadIdx.x == 0) {
    hist_rr[kend] = rr;
    st->iter = kend;
    st->maxit = maxit;
    if (!isfinite(rr)) st->breakdown = 1;
    if (!(rr > st->tol2) || !isfinite(rr) || kend >= maxit) st->done = 1;
  }
}

// Reliable update, part 1: x64 += x32 (FP32 correction), x32 = 0.
__global__ __launch_bounds__(ST) void k_accumulate_f32(dlong N, float* c32, double* __restrict__ x64) {
  for (dlong n = blockIdx.x * ST + threadIdx.x; n < N; n += gridDim.x * ST) {
    x64[n] += (double)c32[n];
    c32[n] = 0.0f;
  }
}

// Row permutations between the reference row order and a solver's own
// order: perm[i] is the reference row of row i, or -1 for a padding row.
__global__ __launch_bounds__(ST) void k_permute_in(dlong N, const dlong* __restrict__ perm,
                                                   const double* __restrict__ ref, double* __restrict__ out) {
  for (dlong n = blockIdx.x * ST + threadIdx.x; n < N; n += gridDim.x * ST) {
    const dlong r = perm[n];
    out[n] = (r >= 0) ? ref[r] : 0.0;
  }
}

__global__ __launch_bounds__(ST) void k_permute_out(dlong N, const dlong* __restrict__ perm,
                                                    const double* __restrict__ v, double* __restrict__ ref) {
  for (dlong n = blockIdx.x * ST + threadIdx.x; n < N; n += gridDim.x * ST) {
    const dlong r = perm[n];
    if (r >= 0) ref[r] = v[n];
  }
}

// Reliable update, part 2: true residual r = b - Ax in FP64, handed to the
// FP32 recurrence as r32 = float(r), with partials of r.r (FP64, for the
// stopping test) and of r32.r32 (the recurrence's own r.r).
__global__ __launch_bounds__(RT) void k_true_residual(dlong N, const double* __restrict__ b,
                                                      const double* __restrict__ Ax, float* __restrict__ r32,
                                                      double* __restrict__ part_true,
                                                      double* __restrict__ part_rr32) {
  __shared__ double s_w[33];
  double t0 = 0, t1 = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < N; n += stride) {
    const double rn = b[n] - Ax[n];
    const float rf = (float)rn;
    r32[n] = rf;
    t0 += rn * rn;
    t1 += (double)rf * (double)rf;
  }
  const double s0 = block_allsum(t0, s_w);
  const double s1 = block_allsum(t1, s_w);
  if (threadIdx.x == 0) {
    part_true[blockIdx.x] = s0;
    part_rr32[blockIdx.x] = s1;
  }
}

// cg::UpdateCG as one pass: r <- r - alpha Ap, x <- x + alpha p, r.r partials
template <typename T>
__global__ __launch_bounds__(RT) void k_update_cg(dlong N, T alpha, const T* __restrict__ p,
                                                  const T* __restrict__ Ap, T* __restrict__ x, T* __restrict__ r,
                                                  double* __restrict__ part) {
  __shared__ double s_w[33];
  const double v = strided_sum(N, [&](long long n) {
    const T rn = r[n] - alpha * Ap[n];
    r[n] = rn;
    x[n] = alpha * p[n] + T(1) * x[n];
    return (double)rn * (double)rn;
  });
  const double t = block_allsum(v, s_w);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

}  // namespace

template <typename T>
int launch_set(hipStream_t s, dlong N, T alpha, T* x) {
  if (N < 0 || (N > 0 && !x)) return set_error(BPX_ERR_INVALID_ARGUMENT, "set: bad arguments");
  if (N == 0) return BPX_OK;
  hipLaunchKernelGGL(k_set<T>, dim3(stream_blocks(N)), dim3(ST), 0, s, N, alpha, x);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_axpy(hipStream_t s, dlong N, T alpha, const T* x, T beta, T* y) {
  if (N < 0 || (N > 0 && (!x || !y))) return set_error(BPX_ERR_INVALID_ARGUMENT, "axpy: bad arguments");
  if (N == 0) return BPX_OK;
  hipLaunchKernelGGL(k_axpy<T>, dim3(stream_blocks(N)), dim3(ST), 0, s, N, alpha, x, beta, y);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_axpy_f32_to_f64(hipStream_t s, dlong N, double alpha, const float* x, double beta, double* y) {
  if (N <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_axpy_f32_f64, dim3(stream_blocks(N)), dim3(ST), 0, s, N, alpha, x, beta, y);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_copy_f64_to_f32(hipStream_t s, dlong N, double scale, const double* x, float* y) {
  if (N <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_copy_f64_f32, dim3(stream_blocks(N)), dim3(ST), 0, s, N, scale, x, y);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_dot_partials(hipStream_t s, dlong N, const T* x, const T* y, double* part) {
  if (N < 0 || !part || (N > 0 && (!x || !y))) return set_error(BPX_ERR_INVALID_ARGUMENT, "dot: bad arguments");
  hipLaunchKernelGGL(k_dot_partials<T>, dim3(RB), dim3(RT), 0, s, N, x, y, part);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_fold(hipStream_t s, const double* part, int n, double* out) {
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(RT), 0, s, part, n, out);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_dot(hipStream_t s, dlong N, const T* x, const T* y, double* part, double* out) {
  if (int st = launch_dot_partials<T>(s, N, x, y, part)) return st;
  return launch_fold(s, part, RB, out);
}

template <typename T>
int launch_cg_init(hipStream_t s, dlong N, const T* Ap, T* r, double* part_rr) {
  hipLaunchKernelGGL(k_cg_init<T>, dim3(RB), dim3(RT), 0, s, N, Ap, r, part_rr);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

inline bool aligned16(const void* a) { return ((uintptr_t)a & 15) == 0; }

// float streams use 16-byte accesses when every pointer allows it; FP64
// keeps the scalar form, whose partial sums the parity tests pin
template <typename T>
int launch_cg_update_p(hipStream_t s, dlong N, int k, const T* r, T* p, T* x, const double* part_rr, int nrr,
                       CgState* st, double* hist_rr, const double* hist_alpha, double tol, bool pending) {
  if (std::is_same<T, float>::value && aligned16(r) && aligned16(p) && aligned16(x))
    hipLaunchKernelGGL((k_cg_update_p<T, true>), dim3(RB), dim3(RT), 0, s, N, k, r, p, x, part_rr, nrr, st,
                       hist_rr, hist_alpha, tol, (int)pending);
  else
    hipLaunchKernelGGL((k_cg_update_p<T, false>), dim3(RB), dim3(RT), 0, s, N, k, r, p, x, part_rr, nrr, st,
                       hist_rr, hist_alpha, tol, (int)pending);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_cg_update_r(hipStream_t s, dlong N, int k, const T* Ap, T* r, const double* part_pAp, int npAp,
                       double* part_rr, CgState* st, const double* hist_rr, double* hist_alpha) {
  if (std::is_same<T, float>::value && aligned16(Ap) && aligned16(r))
    hipLaunchKernelGGL((k_cg_update_r<T, true>), dim3(RB), dim3(RT), 0, s, N, k, Ap, r, part_pAp, npAp,
                       part_rr, st, hist_rr, hist_alpha);
  else
    hipLaunchKernelGGL((k_cg_update_r<T, false>), dim3(RB), dim3(RT), 0, s, N, k, Ap, r, part_pAp, npAp,
                       part_rr, st, hist_rr, hist_alpha);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_cg_flush_x(hipStream_t s, dlong N, int k, const T* p, T* x, const double* hist_alpha,
                      const CgState* st, bool when_running) {
  if (N <= 0 || k < 0) return BPX_OK;
  hipLaunchKernelGGL(k_cg_flush_x<T>, dim3(stream_blocks(N)), dim3(ST), 0, s, N, k, p, x, hist_alpha, st,
                     (int)when_running);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_cg_finish(hipStream_t s, int kend, int maxit, const double* part_rr, int nrr, CgState* st,
                     double* hist_rr) {
  hipLaunchKernelGGL(k_cg_finish, dim3(1), dim3(RT), 0, s, kend, maxit, part_rr, nrr, st, hist_rr);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_accumulate_f32(hipStream_t s, dlong N, float* c32, double* x64) {
  if (N <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_accumulate_f32, dim3(stream_blocks(N)), dim3(ST), 0, s, N, c32, x64);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_permute(hipStream_t s, dlong N, const dlong* perm, const double* in, double* out, bool to_ref) {
  if (N <= 0) return BPX_OK;
  if (!perm || !in || !out) return set_error(BPX_ERR_INVALID_ARGUMENT, "permute: null pointer");
  if (to_ref)
    hipLaunchKernelGGL(k_permute_out, dim3(stream_blocks(N)), dim3(ST), 0, s, N, perm, in, out);
  else
    hipLaunchKernelGGL(k_permute_in, dim3(stream_blocks(N)), dim3(ST), 0, s, N, perm, in, out);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_true_residual(hipStream_t s, dlong N, const double* b, const double* Ax, float* r32, double* part_true,
                         double* part_rr32) {
  hipLaunchKernelGGL(k_true_residual, dim3(RB), dim3(RT), 0, s, N, b, Ax, r32, part_true, part_rr32);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_update_cg(hipStream_t s, dlong N, double alpha, const double* p, const double* Ap, double* x,
                     double* r, double* part, double* out) {
  if (N < 0 || !part || !out || (N > 0 && (!p || !Ap || !x || !r)))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "update_cg: bad arguments");
  hipLaunchKernelGGL(k_update_cg<double>, dim3(RB), dim3(RT), 0, s, N, alpha, p, Ap, x, r, part);
  BPX_HIP(hipGetLastError());
  return launch_fold(s, part, RB, out);
}

#define BPX_LINALG_INST(T)                                                                                   \
  template int launch_set<T>(hipStream_t, dlong, T, T*);                                                      \
  template int launch_axpy<T>(hipStream_t, dlong, T, const T*, T, T*);                                        \
  template int launch_dot_partials<T>(hipStream_t, dlong, const T*, const T*, double*);                       \
  template int launch_dot<T>(hipStream_t, dlong, const T*, const T*, double*, double*);                       \
  template int launch_cg_init<T>(hipStream_t, dlong, const T*, T*, double*);                                  \
  template int launch_cg_update_p<T>(hipStream_t, dlong, int, const T*, T*, T*, const double*, int, CgState*, \
                                     double*, const double*, double, bool);                                   \
  template int launch_cg_update_r<T>(hipStream_t, dlong, int, const T*, T*, const double*, int, double*,      \
                                     CgState*, const double*, double*);                                       \
  template int launch_cg_flush_x<T>(hipStream_t, dlong, int, const T*, T*, const double*, const CgState*, bool);
BPX_LINALG_INST(double)
BPX_LINALG_INST(float)

}  // namespace bpx

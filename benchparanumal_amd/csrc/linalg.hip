// linAlg vector kernels and the fused, device-resident CG steps.
//
// Semantics: libs/core/okl/linAlgSet.okl, linAlgAXPY.okl (y = a*x + b*y),
// linAlgInnerProd.okl / linAlgNorm2.okl (dot products), and cg::Solve /
// cg::UpdateCG of libs/core/linearSolverCG.cpp:62-156 with
// okl/linearSolverUpdateCG.okl (r -= alpha*Ap fused with r.r).
//
// Design (gfx950): every reduction is one launch. Each workgroup folds a
// grid-strided slice (four independent accumulators per lane to keep four
// loads in flight), reduces across the wave with xor shuffles and across
// waves through LDS in a fixed order, stores its partial, and takes a
// ticket; the last workgroup sums the partials in index order and runs the
// scalar epilogue (alpha, beta, stopping test) on the device. The grid
// size depends only on N, so results are bit-reproducible. CG scalars live
// in a device CgState, so an iteration never waits on the host (the
// reference waits 8 times per iteration, linearSolverCG.cpp:146 et al.).
#include <hip/hip_runtime.h>

#include <cmath>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

constexpr int BS = kReduceThreads;
constexpr int kWaves = BS / 64;

struct ReduceScratch {
  double partial[kReduceBlocks];
  unsigned int ticket;
};

inline int reduce_blocks(dlong N) {
  long long nb = ((long long)N + BS * 4 - 1) / (BS * 4);
  if (nb < 1) nb = 1;
  if (nb > kReduceBlocks) nb = kReduceBlocks;
  return (int)nb;
}

inline int stream_blocks(dlong N) {
  long long nb = ((long long)N + BS - 1) / BS;
  if (nb < 1) nb = 1;
  if (nb > 8192) nb = 8192;
  return (int)nb;
}

// fixed-order block sum; the total is valid in thread 0
__device__ inline double block_sum(double v, double* s_w) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) s_w[w] = v;
  __syncthreads();
  double tot = 0;
  if (threadIdx.x == 0)
#pragma unroll
    for (int i = 0; i < kWaves; ++i) tot += s_w[i];
  return tot;
}

// Publish this block's partial; returns true in every thread of the last
// block, after which all partials are visible.
__device__ inline bool finish_partials(ReduceScratch* rs, double blocksum, bool* s_last) {
  if (threadIdx.x == 0) {
    rs->partial[blockIdx.x] = blocksum;
    __threadfence();
    const unsigned t = atomicAdd(&rs->ticket, 1u);
    *s_last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!*s_last) return false;
  __threadfence();
  return true;
}

// Sum of the partials in index order (valid in thread 0); resets the ticket.
__device__ inline double total_of_partials(ReduceScratch* rs, double* s_w) {
  double v = 0;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += BS)
    v += __hip_atomic_load(&rs->partial[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const double tot = block_sum(v, s_w);
  if (threadIdx.x == 0) __hip_atomic_store(&rs->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return tot;
}

template <typename T>
__global__ __launch_bounds__(BS) void k_set(dlong N, T alpha, T* __restrict__ x) {
  for (dlong n = blockIdx.x * BS + threadIdx.x; n < N; n += gridDim.x * BS) x[n] = alpha;
}

template <typename T>
__global__ __launch_bounds__(BS) void k_axpy(dlong N, T alpha, const T* __restrict__ x, T beta,
                                             T* __restrict__ y) {
  for (dlong n = blockIdx.x * BS + threadIdx.x; n < N; n += gridDim.x * BS) y[n] = alpha * x[n] + beta * y[n];
}

__global__ __launch_bounds__(BS) void k_axpy_f32_f64(dlong N, double alpha, const float* __restrict__ x,
                                                     double beta, double* __restrict__ y) {
  for (dlong n = blockIdx.x * BS + threadIdx.x; n < N; n += gridDim.x * BS)
    y[n] = alpha * (double)x[n] + beta * y[n];
}

__global__ __launch_bounds__(BS) void k_copy_f64_f32(dlong N, double scale, const double* __restrict__ x,
                                                     float* __restrict__ y) {
  for (dlong n = blockIdx.x * BS + threadIdx.x; n < N; n += gridDim.x * BS) y[n] = (float)(scale * x[n]);
}

// four independent strided accumulators per lane
template <typename F>
__device__ inline double strided_sum(dlong N, F term) {
  const long long stride = (long long)gridDim.x * BS;
  long long n = (long long)blockIdx.x * BS + threadIdx.x;
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

template <typename T>
__global__ __launch_bounds__(BS) void k_dot(dlong N, const T* __restrict__ x, const T* __restrict__ y,
                                            ReduceScratch* rs, double* out) {
  __shared__ double s_w[kWaves];
  __shared__ bool s_last;
  const double v = strided_sum(N, [&](long long n) { return (double)x[n] * (double)y[n]; });
  const double bsum = block_sum(v, s_w);
  if (!finish_partials(rs, bsum, &s_last)) return;
  const double tot = total_of_partials(rs, s_w);
  if (threadIdx.x == 0) *out = tot;
}

// r <- r - Ap (x0 residual) and r.r; sets up the stopping threshold
// (linearSolverCG.cpp:79-87: rdotr = norm2(r)^2, TOL = max(tol^2 rdotr, tol^2)).
template <typename T>
__global__ __launch_bounds__(BS) void k_cg_init(dlong N, const T* __restrict__ Ap, T* __restrict__ r,
                                                ReduceScratch* rs, CgState* st, double tol, int maxit,
                                                double* hist_rr) {
  __shared__ double s_w[kWaves];
  __shared__ bool s_last;
  const double v = strided_sum(N, [&](long long n) {
    const T rn = T(-1) * Ap[n] + T(1) * r[n];
    r[n] = rn;
    return (double)rn * (double)rn;
  });
  const double bsum = block_sum(v, s_w);
  if (!finish_partials(rs, bsum, &s_last)) return;
  const double tot = total_of_partials(rs, s_w);
  if (threadIdx.x == 0) {
    const double nrm = sqrt(tot);
    const double rdotr = nrm * nrm;
    st->rdotr = rdotr;
    st->rdotr_prev = 0.0;
    st->alpha = 0.0;
    st->pAp = 0.0;
    st->tol2 = fmax(tol * tol * rdotr, tol * tol);
    st->iter = 0;
    st->maxit = maxit;
    st->breakdown = isfinite(rdotr) ? 0 : 1;
    st->done = (rdotr <= st->tol2 || maxit <= 0 || st->breakdown) ? 1 : 0;
    if (hist_rr) hist_rr[0] = rdotr;
  }
}

// p <- r + beta p, beta = rdotr/rdotr_prev (0 on the first iteration)
template <typename T>
__global__ __launch_bounds__(BS) void k_cg_update_p(dlong N, const T* __restrict__ r, T* __restrict__ p,
                                                    const CgState* st) {
  if (st->done) return;
  // first iteration: beta = 0, p = r (also when p holds stale values)
  const bool first = (st->iter == 0);
  const T b = first ? T(0) : (T)(st->rdotr / st->rdotr_prev);
  for (dlong n = blockIdx.x * BS + threadIdx.x; n < N; n += gridDim.x * BS)
    p[n] = first ? r[n] : T(1) * r[n] + b * p[n];
}

// p.Ap, then alpha = rdotr / pAp on the device
template <typename T>
__global__ __launch_bounds__(BS) void k_cg_pAp(dlong N, const T* __restrict__ p, const T* __restrict__ Ap,
                                               ReduceScratch* rs, CgState* st) {
  __shared__ double s_w[kWaves];
  __shared__ bool s_last;
  if (st->done) return;
  const double v = strided_sum(N, [&](long long n) { return (double)p[n] * (double)Ap[n]; });
  const double bsum = block_sum(v, s_w);
  if (!finish_partials(rs, bsum, &s_last)) return;
  const double tot = total_of_partials(rs, s_w);
  if (threadIdx.x == 0) {
    st->pAp = tot;
    const double alpha = st->rdotr / tot;
    st->alpha = alpha;
    if (!isfinite(alpha) || !(tot > 0.0)) {
      st->breakdown = 1;
      st->done = 1;
    }
  }
}

// r <- r - alpha Ap, x <- x + alpha p, r.r; advances the iteration
template <typename T>
__global__ __launch_bounds__(BS) void k_cg_update_xr(dlong N, const T* __restrict__ p, const T* __restrict__ Ap,
                                                     T* __restrict__ x, T* __restrict__ r, ReduceScratch* rs,
                                                     CgState* st, double* hist_rr, double* hist_alpha) {
  __shared__ double s_w[kWaves];
  __shared__ bool s_last;
  if (st->done) return;
  const double alpha = st->alpha;
  const T a = (T)alpha;
  const double v = strided_sum(N, [&](long long n) {
    const T rn = r[n] - a * Ap[n];
    r[n] = rn;
    x[n] = a * p[n] + T(1) * x[n];
    return (double)rn * (double)rn;
  });
  const double bsum = block_sum(v, s_w);
  if (!finish_partials(rs, bsum, &s_last)) return;
  const double tot = total_of_partials(rs, s_w);
  if (threadIdx.x == 0) {
    st->rdotr_prev = st->rdotr;
    st->rdotr = tot;
    const int it = st->iter + 1;
    st->iter = it;
    if (hist_rr) hist_rr[it] = tot;
    if (hist_alpha) hist_alpha[it - 1] = alpha;
    if (!isfinite(tot)) st->breakdown = 1;
    st->done = (tot <= st->tol2 || it >= st->maxit || st->breakdown) ? 1 : 0;
  }
}

}  // namespace

size_t reduce_scratch_bytes() { return sizeof(ReduceScratch); }

template <typename T>
int launch_set(hipStream_t s, dlong N, T alpha, T* x) {
  if (N < 0 || (N > 0 && !x)) return set_error(BPX_ERR_INVALID_ARGUMENT, "set: bad arguments");
  if (N == 0) return BPX_OK;
  hipLaunchKernelGGL(k_set<T>, dim3(stream_blocks(N)), dim3(BS), 0, s, N, alpha, x);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_axpy(hipStream_t s, dlong N, T alpha, const T* x, T beta, T* y) {
  if (N < 0 || (N > 0 && (!x || !y))) return set_error(BPX_ERR_INVALID_ARGUMENT, "axpy: bad arguments");
  if (N == 0) return BPX_OK;
  hipLaunchKernelGGL(k_axpy<T>, dim3(stream_blocks(N)), dim3(BS), 0, s, N, alpha, x, beta, y);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_axpy_f32_to_f64(hipStream_t s, dlong N, double alpha, const float* x, double beta, double* y) {
  if (N <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_axpy_f32_f64, dim3(stream_blocks(N)), dim3(BS), 0, s, N, alpha, x, beta, y);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_copy_f64_to_f32(hipStream_t s, dlong N, double scale, const double* x, float* y) {
  if (N <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_copy_f64_f32, dim3(stream_blocks(N)), dim3(BS), 0, s, N, scale, x, y);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_dot(hipStream_t s, dlong N, const T* x, const T* y, void* scratch, double* out) {
  if (N < 0 || !out || !scratch || (N > 0 && (!x || !y)))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "dot: bad arguments");
  hipLaunchKernelGGL(k_dot<T>, dim3(reduce_blocks(N)), dim3(BS), 0, s, N, x, y,
                     static_cast<ReduceScratch*>(scratch), out);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_cg_init(hipStream_t s, dlong N, const T* Ap, T* r, void* scratch, CgState* st, double tol,
                   int maxit, double* hist_rr) {
  hipLaunchKernelGGL(k_cg_init<T>, dim3(reduce_blocks(N)), dim3(BS), 0, s, N, Ap, r,
                     static_cast<ReduceScratch*>(scratch), st, tol, maxit, hist_rr);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_cg_update_p(hipStream_t s, dlong N, const T* r, T* p, const CgState* st) {
  hipLaunchKernelGGL(k_cg_update_p<T>, dim3(stream_blocks(N)), dim3(BS), 0, s, N, r, p, st);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_cg_pAp(hipStream_t s, dlong N, const T* p, const T* Ap, void* scratch, CgState* st) {
  hipLaunchKernelGGL(k_cg_pAp<T>, dim3(reduce_blocks(N)), dim3(BS), 0, s, N, p, Ap,
                     static_cast<ReduceScratch*>(scratch), st);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_cg_update_xr(hipStream_t s, dlong N, const T* p, const T* Ap, T* x, T* r, void* scratch,
                        CgState* st, double* hist_rr, double* hist_alpha) {
  hipLaunchKernelGGL(k_cg_update_xr<T>, dim3(reduce_blocks(N)), dim3(BS), 0, s, N, p, Ap, x, r,
                     static_cast<ReduceScratch*>(scratch), st, hist_rr, hist_alpha);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

#define BPX_LINALG_INST(T)                                                                         \
  template int launch_set<T>(hipStream_t, dlong, T, T*);                                            \
  template int launch_axpy<T>(hipStream_t, dlong, T, const T*, T, T*);                              \
  template int launch_dot<T>(hipStream_t, dlong, const T*, const T*, void*, double*);               \
  template int launch_cg_init<T>(hipStream_t, dlong, const T*, T*, void*, CgState*, double, int,    \
                                 double*);                                                          \
  template int launch_cg_update_p<T>(hipStream_t, dlong, const T*, T*, const CgState*);             \
  template int launch_cg_pAp<T>(hipStream_t, dlong, const T*, const T*, void*, CgState*);           \
  template int launch_cg_update_xr<T>(hipStream_t, dlong, const T*, const T*, T*, T*, void*,        \
                                      CgState*, double*, double*);
BPX_LINALG_INST(double)
BPX_LINALG_INST(float)

}  // namespace bpx

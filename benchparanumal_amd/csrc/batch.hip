// Ragged batch CG: many independent BP5 problems (any sizes, degrees)
// solved together, one CG per problem. No reference counterpart; each
// problem runs cg::Solve's recurrence (libs/core/linearSolverCG.cpp:62-129).
//
// Layout: the problems' global vectors are concatenated (problem i owns
// [off_i, off_i + Ngather_i)). The streaming work is cut into chunks that
// never straddle two problems, and the gather's super-blocks likewise, so
// every reduction partial belongs to one problem: a problem's scalars are
// folds over its own contiguous range of partials, done by every workgroup
// that needs them (no atomics, no inter-workgroup waits, bit-reproducible).
//
// Per-problem state lives on the device, so iterations never wait on the
// host and a whole block of iterations can be captured in one hipGraph and
// replayed. Writers and readers of each state field are split across
// kernels so no kernel reads a field it also writes:
//   update_p  (iteration k): reads rr_prev, iter; first chunk writes rr_k,
//             tol2 (k = 0), and done/rnorm when the problem stops;
//   update_xr (iteration k): reads rr_k; first chunk writes rr_prev = rr_k,
//             iter = k + 1, and done/breakdown on a non-positive p.Ap;
// every chunk recomputes the scalars it acts on, so all chunks of a
// problem take the same branch. A converged problem's elements, rows and
// chunks return at once in every kernel.
#include <hip/hip_runtime.h>

#include <cmath>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

constexpr int BT = 256;  // threads of the streaming kernels

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// fixed-order fold of part[a..b) by wavefront 0, broadcast through LDS
__device__ inline double fold_range(const double* __restrict__ part, int a, int b, double* s_w) {
  if (threadIdx.x < 64) {
    double a0 = 0, a1 = 0;
    int i = a + (int)threadIdx.x;
    for (; i + 64 < b; i += 128) {
      a0 += part[i];
      a1 += part[i + 64];
    }
    if (i < b) a0 += part[i];
    const double v = wave_sum(a0 + a1);
    if (threadIdx.x == 0) s_w[8] = v;
  }
  __syncthreads();
  return s_w[8];
}

__device__ inline double block_sum(double v, double* s_w) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < BT / 64; ++w) t += s_w[w];
  return t;
}

// r = b, x = 0, partial r.r per chunk; state reset
__global__ __launch_bounds__(BT) void k_batch_init(const BatchChunk* __restrict__ chunks, const double* __restrict__ b,
                                                   double* __restrict__ x, double* __restrict__ r,
                                                   double* __restrict__ part_rr, BatchState* st, int maxit) {
  __shared__ double s_w[16];
  const BatchChunk c = chunks[blockIdx.x];
  double v = 0;
  for (int n = c.n0 + threadIdx.x; n < c.n1; n += BT) {
    const double rn = b[n];
    r[n] = rn;
    x[n] = 0.0;
    v += rn * rn;
  }
  const double t = block_sum(v, s_w);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = t;
    if (c.first) {
      BatchState& s = st[c.pid];
      s.rr_k = 0;
      s.rr_prev = 0;
      s.tol2 = 0;
      s.rnorm = 0;
      s.iter = 0;
      s.done = 0;
      s.breakdown = 0;
      s.maxit = maxit;
    }
  }
}

__global__ __launch_bounds__(BT) void k_batch_update_p(const BatchChunk* __restrict__ chunks,
                                                       const int2* __restrict__ rrRange, const double* __restrict__ r,
                                                       double* __restrict__ p, const double* __restrict__ part_rr,
                                                       BatchState* st, double tol) {
  __shared__ double s_w[16];
  const BatchChunk c = chunks[blockIdx.x];
  BatchState* s = st + c.pid;
  if (s->done) return;
  const int2 range = rrRange[c.pid];
  double rr = fold_range(part_rr, range.x, range.y, s_w);
  const int k = s->iter;
  double tol2;
  if (k == 0) {
    const double nrm = sqrt(rr);  // rdotr = norm2(r)^2 (linearSolverCG.cpp:84-85)
    rr = nrm * nrm;
    tol2 = fmax(tol * tol * rr, tol * tol);
  } else {
    tol2 = s->tol2;
  }
  const bool stop = !(rr > tol2) || !isfinite(rr) || k >= s->maxit;
  if (c.first && threadIdx.x == 0) {
    s->rr_k = rr;
    if (k == 0) s->tol2 = tol2;
    if (stop) {
      s->rnorm = sqrt(rr);
      s->done = 1;
      if (!isfinite(rr)) s->breakdown = 1;
    }
  }
  if (stop) return;
  const double beta = k == 0 ? 0.0 : rr / s->rr_prev;
  for (int n = c.n0 + threadIdx.x; n < c.n1; n += BT) p[n] = k == 0 ? r[n] : 1.0 * r[n] + beta * p[n];
}

__global__ __launch_bounds__(BT) void k_batch_update_xr(const BatchChunk* __restrict__ chunks,
                                                        const int2* __restrict__ pApRange,
                                                        const double* __restrict__ p, const double* __restrict__ Ap,
                                                        double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ part_pAp,
                                                        double* __restrict__ part_rr, BatchState* st) {
  __shared__ double s_w[16];
  const BatchChunk c = chunks[blockIdx.x];
  BatchState* s = st + c.pid;
  if (s->done) return;
  const int2 range = pApRange[c.pid];
  const double pAp = fold_range(part_pAp, range.x, range.y, s_w);
  const double rr = s->rr_k;
  const double alpha = rr / pAp;
  if (!isfinite(alpha) || !(pAp > 0.0)) {
    if (c.first && threadIdx.x == 0) {
      s->breakdown = 1;
      s->done = 1;
      s->rnorm = sqrt(rr);
    }
    return;
  }
  double v = 0;
  for (int n = c.n0 + threadIdx.x; n < c.n1; n += BT) {
    const double rn = r[n] - alpha * Ap[n];
    r[n] = rn;
    x[n] = alpha * p[n] + 1.0 * x[n];
    v += rn * rn;
  }
  const double t = block_sum(v, s_w);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = t;
    if (c.first) {
      s->rr_prev = rr;
      s->iter = s->iter + 1;
    }
  }
}

// Gather with one p.Ap partial per super-block; super-blocks never straddle
// problems and a converged problem's super-blocks return at once.
constexpr int GT = 512;
constexpr int GMAX = 4096;  // entries staged per super-block (8 row blocks of <= 512)

// T: the vectors' type (sums in T, the p.Ap partials in FP64); DOT: write
// the p.Ap partials
template <typename T, bool DOT>
__global__ __launch_bounds__(GT) void k_batch_gather(const BatchSuper* __restrict__ supers,
                                                     const int* __restrict__ rowStarts,
                                                     const int* __restrict__ colIds, const T* __restrict__ q,
                                                     T* __restrict__ gq, const T* __restrict__ p,
                                                     double* __restrict__ part, const BatchState* __restrict__ st) {
  __shared__ T s_v[GMAX];
  __shared__ double s_w[GT / 64];
  const BatchSuper sb = supers[blockIdx.x];
  if (st[sb.pid].done) return;
  const int r0 = sb.r0, r1 = sb.r1;
  const int g0 = rowStarts[r0], g1 = rowStarts[r1];
  const bool staged = (g1 - g0) <= GMAX && (r1 - r0) <= GMAX;
  constexpr int PL = GMAX / GT;  // entries (and rows) per lane when staged
  double dot = 0;
  if (staged) {
    // all of a lane's loads are issued before any is used, from indices
    // clamped into the super-block (lanes past it discard what they read):
    // a load under a branch makes the compiler wait for it in place
    int cid[PL], ra[PL], re[PL];
    T v[PL], pr[PL];
#pragma unroll
    for (int m = 0; m < PL; ++m) {
      cid[m] = colIds[max(min(g0 + (int)threadIdx.x + m * GT, g1 - 1), 0)];
      const int rc = max(min(r0 + (int)threadIdx.x + m * GT, r1 - 1), 0);
      ra[m] = rowStarts[rc];
      re[m] = rowStarts[rc + 1];
      pr[m] = DOT ? p[rc] : T(0);
    }
#pragma unroll
    for (int m = 0; m < PL; ++m) v[m] = q[cid[m]];
#pragma unroll
    for (int m = 0; m < PL; ++m)
      if (g0 + (int)threadIdx.x + m * GT < g1) s_v[threadIdx.x + m * GT] = v[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < PL; ++m) {
      const int row = r0 + threadIdx.x + m * GT;
      if (row < r1) {
        T acc = 0;
        for (int g = ra[m]; g < re[m]; ++g) acc += s_v[g - g0];
        gq[row] = acc;
        dot += (double)pr[m] * (double)acc;
      }
    }
  } else {
    __syncthreads();  // the same barrier count on every path (uniform: staged is per block)
    for (int row = r0 + threadIdx.x; row < r1; row += GT) {
      const int a = rowStarts[row], e = rowStarts[row + 1];
      T acc = 0;
      for (int g = a; g < e; ++g) acc += q[colIds[g]];
      gq[row] = acc;
      if (DOT) dot += (double)p[row] * (double)acc;
    }
  }
  if (!DOT) return;
  dot = wave_sum(dot);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = dot;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < GT / 64; ++w) t += s_w[w];
    part[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(256) void k_batch_count_done(int nprob, const BatchState* __restrict__ st,
                                                          int* __restrict__ count) {
  __shared__ int s_c[4];
  int c = 0;
  for (int i = threadIdx.x; i < nprob; i += 256) c += st[i].done ? 1 : 0;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0) s_c[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) *count = s_c[0] + s_c[1] + s_c[2] + s_c[3];
}

// ---- mixed batch: FP32 CG per problem with FP64 corrections ------------
// Two states per problem. inner: the FP32 recurrence (rr_k, rr_prev; iter
// = iteration within the current recurrence; done = the recurrence is
// paused: it stopped by itself (r.r = 0, p.Ap <= 0, the iteration cap) or
// the problem has finished). outer: the problem (rr_k = ||b - A x||^2 at
// the last correction, tol2 = tol^2 ||b||^2, rnorm = ||b - A x||, iter =
// FP32 iterations, updates = corrections, done, breakdown, maxit). The
// FP32 kernels skip problems whose inner state is done, the FP64 ones
// those whose outer state is; a correction restarts a paused recurrence
// from the true residual unless the problem has finished. The field rules
// of the FP64 batch hold: no kernel reads a state field that it writes.

__global__ __launch_bounds__(BT) void k_to_f32(size_t n, const double* __restrict__ a, float* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * BT + threadIdx.x;
  if (i < n) out[i] = (float)a[i];
}

// x = 0, r = float(b), c = 0; partials of b.b (FP64) and of r.r (the FP32
// residual, summed in FP64); both states reset
__global__ __launch_bounds__(BT) void k_bm_init(const BatchChunk* __restrict__ chunks, const double* __restrict__ b,
                                                double* __restrict__ x, float* __restrict__ r, float* __restrict__ c,
                                                double* __restrict__ part_true, double* __restrict__ part_rr,
                                                BatchState* inner, BatchState* outer, int maxit) {
  __shared__ double s_w[16];
  const BatchChunk ch = chunks[blockIdx.x];
  double vt = 0, vr = 0;
  for (int n = ch.n0 + threadIdx.x; n < ch.n1; n += BT) {
    const double bn = b[n];
    const float rf = (float)bn;
    x[n] = 0.0;
    r[n] = rf;
    c[n] = 0.0f;
    vt += bn * bn;
    vr += (double)rf * (double)rf;
  }
  const double tt = block_sum(vt, s_w);
  const double tr = block_sum(vr, s_w);
  if (threadIdx.x == 0) {
    part_true[blockIdx.x] = tt;
    part_rr[blockIdx.x] = tr;
    if (ch.first) {
      const BatchState z{0, 0, 0, 0, 0, 0, 0, maxit, 0, 0};
      inner[ch.pid] = z;
      outer[ch.pid] = z;
    }
  }
}

// one thread per problem: the true residual's norm (a fixed-order sum of
// the problem's partials) against the tolerance; init: the first one (b.b)
// sets the threshold
__global__ __launch_bounds__(256) void k_bm_decide(int nprob, const int2* __restrict__ rrRange,
                                                   const double* __restrict__ part_true, BatchState* inner,
                                                   BatchState* outer, double tol, int init) {
  const int pid = blockIdx.x * 256 + threadIdx.x;
  if (pid >= nprob) return;
  BatchState& o = outer[pid];
  BatchState& in = inner[pid];
  if (o.done) return;
  const int2 range = rrRange[pid];
  double rn2 = 0;
  for (int i = range.x; i < range.y; ++i) rn2 += part_true[i];
  o.rr_k = rn2;
  o.rnorm = sqrt(rn2);
  if (init) {
    o.tol2 = tol * tol * rn2;
    in.tol2 = rn2;  // b.b (for the relative residual reported at the end)
    if (!(rn2 > 0)) {  // b = 0: x = 0 is the solution
      o.done = 1;
      in.done = 1;
      if (!isfinite(rn2)) o.breakdown = 1;
    }
    return;
  }
  o.updates = o.updates + 1;
  if (!isfinite(rn2)) {
    o.breakdown = 1;
    o.done = 1;
    in.done = 1;
  } else if (!(rn2 > o.tol2) || o.iter >= o.maxit) {
    o.done = 1;
    in.done = 1;
  } else if (in.done) {  // restart the paused recurrence from the true residual
    in.done = 0;
    in.iter = 0;
  }
}

// p = r + beta p in FP32 (p = r at the recurrence's first iteration)
__global__ __launch_bounds__(BT) void k_bm_update_p(const BatchChunk* __restrict__ chunks,
                                                    const int2* __restrict__ rrRange, const float* __restrict__ r,
                                                    float* __restrict__ p, const double* __restrict__ part_rr,
                                                    BatchState* inner, const BatchState* __restrict__ outer) {
  __shared__ double s_w[16];
  const BatchChunk ch = chunks[blockIdx.x];
  BatchState* s = inner + ch.pid;
  if (s->done) return;
  const int2 range = rrRange[ch.pid];
  const double rr = fold_range(part_rr, range.x, range.y, s_w);
  const int k = s->iter;
  const bool stop = !(rr > 0) || !isfinite(rr) || outer[ch.pid].iter >= outer[ch.pid].maxit;
  if (ch.first && threadIdx.x == 0) {
    s->rr_k = rr;
    if (stop) s->done = 1;
  }
  if (stop) return;
  const float beta = k == 0 ? 0.0f : (float)(rr / s->rr_prev);
  for (int n = ch.n0 + threadIdx.x; n < ch.n1; n += BT) p[n] = k == 0 ? r[n] : fmaf(beta, p[n], r[n]);
}

// c += alpha p, r -= alpha Ap in FP32; partial r.r in FP64
__global__ __launch_bounds__(BT) void k_bm_update_xr(const BatchChunk* __restrict__ chunks,
                                                     const int2* __restrict__ pApRange, const float* __restrict__ p,
                                                     const float* __restrict__ Ap, float* __restrict__ c,
                                                     float* __restrict__ r, const double* __restrict__ part_pAp,
                                                     double* __restrict__ part_rr, BatchState* inner,
                                                     BatchState* outer) {
  __shared__ double s_w[16];
  const BatchChunk ch = chunks[blockIdx.x];
  BatchState* s = inner + ch.pid;
  if (s->done) return;
  const int2 range = pApRange[ch.pid];
  const double pAp = fold_range(part_pAp, range.x, range.y, s_w);
  const double rr = s->rr_k;
  const double alpha = rr / pAp;
  if (!isfinite(alpha) || !(pAp > 0.0)) {  // the recurrence stops; the next correction restarts it
    if (ch.first && threadIdx.x == 0) s->done = 1;
    return;
  }
  const float a = (float)alpha;
  double v = 0;
  for (int n = ch.n0 + threadIdx.x; n < ch.n1; n += BT) {
    const float rn = fmaf(-a, Ap[n], r[n]);
    r[n] = rn;
    c[n] = fmaf(a, p[n], c[n]);
    v += (double)rn * (double)rn;
  }
  const double t = block_sum(v, s_w);
  if (threadIdx.x == 0) {
    part_rr[blockIdx.x] = t;
    if (ch.first) {
      s->rr_prev = rr;
      s->iter = s->iter + 1;
      outer[ch.pid].iter = outer[ch.pid].iter + 1;
    }
  }
}

// x += c (FP64), c = 0
__global__ __launch_bounds__(BT) void k_bm_flush(const BatchChunk* __restrict__ chunks, float* __restrict__ c,
                                                 double* __restrict__ x, const BatchState* __restrict__ outer) {
  const BatchChunk ch = chunks[blockIdx.x];
  if (outer[ch.pid].done) return;
  for (int n = ch.n0 + threadIdx.x; n < ch.n1; n += BT) {
    x[n] += (double)c[n];
    c[n] = 0.0f;
  }
}

// r = float(b - A x); partials of the FP64 residual's and of r's squares
__global__ __launch_bounds__(BT) void k_bm_true_residual(const BatchChunk* __restrict__ chunks,
                                                         const double* __restrict__ b, const double* __restrict__ Ax,
                                                         float* __restrict__ r, double* __restrict__ part_true,
                                                         double* __restrict__ part_rr,
                                                         const BatchState* __restrict__ outer) {
  __shared__ double s_w[16];
  const BatchChunk ch = chunks[blockIdx.x];
  if (outer[ch.pid].done) return;
  double vt = 0, vr = 0;
  for (int n = ch.n0 + threadIdx.x; n < ch.n1; n += BT) {
    const double rd = b[n] - Ax[n];
    const float rf = (float)rd;
    r[n] = rf;
    vt += rd * rd;
    vr += (double)rf * (double)rf;
  }
  const double tt = block_sum(vt, s_w);
  const double tr = block_sum(vr, s_w);
  if (threadIdx.x == 0) {
    part_true[blockIdx.x] = tt;
    part_rr[blockIdx.x] = tr;
  }
}

}  // namespace

int launch_to_f32(hipStream_t s, size_t n, const double* a, float* out) {
  if (n == 0) return BPX_OK;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + BT - 1) / BT)), dim3(BT), 0, s, n, a, out);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_bm_init(hipStream_t s, int nchunks, const BatchChunk* chunks, const double* b, double* x, float* r,
                   float* c, double* part_true, double* part_rr, BatchState* inner, BatchState* outer, int maxit) {
  if (nchunks <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_bm_init, dim3(nchunks), dim3(BT), 0, s, chunks, b, x, r, c, part_true, part_rr, inner, outer,
                     maxit);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_bm_decide(hipStream_t s, int nprob, const int2* rrRange, const double* part_true, BatchState* inner,
                     BatchState* outer, double tol, int init) {
  if (nprob <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_bm_decide, dim3((nprob + 255) / 256), dim3(256), 0, s, nprob, rrRange, part_true, inner, outer,
                     tol, init);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_bm_update_p(hipStream_t s, int nchunks, const BatchChunk* chunks, const int2* rrRange, const float* r,
                       float* p, const double* part_rr, BatchState* inner, const BatchState* outer) {
  if (nchunks <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_bm_update_p, dim3(nchunks), dim3(BT), 0, s, chunks, rrRange, r, p, part_rr, inner, outer);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_bm_update_xr(hipStream_t s, int nchunks, const BatchChunk* chunks, const int2* pApRange, const float* p,
                        const float* Ap, float* c, float* r, const double* part_pAp, double* part_rr,
                        BatchState* inner, BatchState* outer) {
  if (nchunks <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_bm_update_xr, dim3(nchunks), dim3(BT), 0, s, chunks, pApRange, p, Ap, c, r, part_pAp, part_rr,
                     inner, outer);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_bm_flush(hipStream_t s, int nchunks, const BatchChunk* chunks, float* c, double* x,
                    const BatchState* outer) {
  if (nchunks <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_bm_flush, dim3(nchunks), dim3(BT), 0, s, chunks, c, x, outer);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_bm_true_residual(hipStream_t s, int nchunks, const BatchChunk* chunks, const double* b, const double* Ax,
                            float* r, double* part_true, double* part_rr, const BatchState* outer) {
  if (nchunks <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_bm_true_residual, dim3(nchunks), dim3(BT), 0, s, chunks, b, Ax, r, part_true, part_rr, outer);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_batch_init(hipStream_t s, int nchunks, const BatchChunk* chunks, const double* b, double* x, double* r,
                      double* part_rr, BatchState* st, int maxit) {
  if (nchunks <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_batch_init, dim3(nchunks), dim3(BT), 0, s, chunks, b, x, r, part_rr, st, maxit);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_batch_update_p(hipStream_t s, int nchunks, const BatchChunk* chunks, const int2* rrRange, const double* r,
                          double* p, const double* part_rr, BatchState* st, double tol) {
  if (nchunks <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_batch_update_p, dim3(nchunks), dim3(BT), 0, s, chunks, rrRange, r, p, part_rr, st, tol);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_batch_update_xr(hipStream_t s, int nchunks, const BatchChunk* chunks, const int2* pApRange,
                           const double* p, const double* Ap, double* x, double* r, const double* part_pAp,
                           double* part_rr, BatchState* st) {
  if (nchunks <= 0) return BPX_OK;
  hipLaunchKernelGGL(k_batch_update_xr, dim3(nchunks), dim3(BT), 0, s, chunks, pApRange, p, Ap, x, r, part_pAp,
                     part_rr, st);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_batch_gather(hipStream_t s, int nsupers, const BatchSuper* supers, const int* rowStarts,
                        const int* colIds, const double* q, double* gq, const double* p, double* part,
                        const BatchState* st) {
  if (nsupers <= 0) return BPX_OK;
  hipLaunchKernelGGL((k_batch_gather<double, true>), dim3(nsupers), dim3(GT), 0, s, supers, rowStarts, colIds, q, gq,
                     p, part, st);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_batch_gather_f32(hipStream_t s, int nsupers, const BatchSuper* supers, const int* rowStarts,
                            const int* colIds, const float* q, float* gq, const float* p, double* part,
                            const BatchState* st) {
  if (nsupers <= 0) return BPX_OK;
  hipLaunchKernelGGL((k_batch_gather<float, true>), dim3(nsupers), dim3(GT), 0, s, supers, rowStarts, colIds, q, gq,
                     p, part, st);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_batch_gather_only(hipStream_t s, int nsupers, const BatchSuper* supers, const int* rowStarts,
                             const int* colIds, const double* q, double* gq, const BatchState* st) {
  if (nsupers <= 0) return BPX_OK;
  hipLaunchKernelGGL((k_batch_gather<double, false>), dim3(nsupers), dim3(GT), 0, s, supers, rowStarts, colIds, q,
                     gq, (const double*)nullptr, (double*)nullptr, st);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

int launch_batch_count_done(hipStream_t s, int nprob, const BatchState* st, int* count) {
  hipLaunchKernelGGL(k_batch_count_done, dim3(1), dim3(256), 0, s, nprob, st, count);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

}  // namespace bpx

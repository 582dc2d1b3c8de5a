// ogs gather / scatter / gather-scatter on the T maps (Add).
//
// Semantics: libs/ogs/okl/ogsKernels.okl:86-122 (gather: each row summed
// from 0 in ascending column order), :127-164 (scatter), :32-81 (gather
// then scatter back). K vectors are interleaved: entry k of node n at
// k + n*K.
//
// Design (gfx950): atomic-free segmented sums over the sorted segment lists
// (rowStarts/colIds). A workgroup owns kSuper consecutive row blocks of at
// most 512 entries (the setup's blockStarts, so no row straddles two
// workgroups): its lanes load the column ids with unit stride and the
// gathered values in column order into LDS (all loads independent, so they
// are in flight together), then each lane sums whole rows out of LDS in
// column order, which reproduces the reference's summation order bit for
// bit. The gather can also emit one FP64 partial of p.Ap per workgroup
// (fused CG dot product), saving a pass over Ap.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

constexpr int kThreads = 512;
constexpr int kSuper = 8;            // row blocks per workgroup (scatter, gather-scatter)
constexpr int kMaxNnz = 512 * kSuper;
constexpr int kGatherSuper = 8;      // row blocks per gather step (and per p.Ap partial)
constexpr int kGatherNnz = 512 * kGatherSuper;
constexpr int kGatherGrid = 1024;    // most workgroups of the gather (a multiple of 8)

// Super-block handled by workgroup w at its step `it`. Mode 0: strided
// (w, w + G, ...). Mode 1: XCD-blocked, so that neighbouring super-blocks,
// which share cache lines of the local vector, meet in one L2: workgroups
// are dealt to the 8 XCDs round-robin (w -> w % 8), and XCD x walks the
// x-th eighth of the super-blocks.
__device__ inline int gather_sb(int mode, int w, int it, int G, int nsb) {
  if (mode == 0) return w + it * G;
  const int per = (nsb + 7) / 8;
  const int x = w & 7, j = w >> 3, gx = G >> 3;
  const int sb = x * per + j + it * gx;
  return sb < min((x + 1) * per, nsb) ? sb : nsb;
}

// Gather of one vector (K = 1), optionally with one partial of
// p.(gathered vector) per super-block (so the partials do not depend on
// which workgroup handled which super-block). Per super-block, phase 1
// issues every global load at once (column ids, the gathered values into
// LDS, and for the lane's own rows their rowStarts and p), phase 2 sums
// rows out of LDS in column order.
// FOLD: super-block sb also folds its share of the fused Ax kernel's
// partials (axpart) into its own.
template <typename T, bool DOT, bool FOLD>
__global__ __launch_bounds__(kThreads) void k_gather(int Nblocks, int mode, const int* __restrict__ blockStarts,
                                                     const int* __restrict__ rowStarts,
                                                     const int* __restrict__ colIds, const T* __restrict__ q,
                                                     T* __restrict__ gq, const T* __restrict__ p,
                                                     double* __restrict__ part,
                                                     const double* __restrict__ axpart, int nax) {
  constexpr int RPL = kGatherNnz / kThreads;  // rows per lane when the rows fit the staging
  constexpr int EPT = kGatherNnz / kThreads;  // entries per lane likewise
  __shared__ T s_v[kGatherNnz];
  __shared__ double s_w[kThreads / 64];
  const int nsb = (Nblocks + kGatherSuper - 1) / kGatherSuper;
  for (int it = 0;; ++it) {
    const int sb = gather_sb(mode, blockIdx.x, it, gridDim.x, nsb);
    if (sb >= nsb) break;
    const int b0 = sb * kGatherSuper;
    const int b1 = min(b0 + kGatherSuper, Nblocks);
    const int r0 = blockStarts[b0], r1 = blockStarts[b1];
    const int g0 = rowStarts[r0], g1 = rowStarts[r1];
    // caller-built blocks may exceed 512 entries or hold empty rows: then
    // sum straight from memory
    const bool staged = (g1 - g0) <= kGatherNnz && (r1 - r0) <= kGatherNnz;
    int ra[RPL], re[RPL];
    T pr[RPL];
    if (staged) {
      // two rounds of independent loads: column ids and row starts; then
      // the gathered values and p. Every load is unconditional, from an
      // index clamped into the block (a load under a branch makes the
      // compiler wait for it, and for every earlier load, in place); lanes
      // past the block discard what they read.
      int cid[EPT];
#pragma unroll
      for (int m = 0; m < EPT; ++m) {
        const int g = g0 + threadIdx.x + m * kThreads;
        const int v = colIds[max(min(g, g1 - 1), 0)];
        cid[m] = (g < g1) ? v : -1;
      }
#pragma unroll
      for (int m = 0; m < RPL; ++m) {
        const int rc = max(min(r0 + (int)threadIdx.x + m * kThreads, r1 - 1), 0);
        ra[m] = rowStarts[rc];
        re[m] = rowStarts[rc + 1];
      }
      T val[EPT];
#pragma unroll
      for (int m = 0; m < EPT; ++m) {
        const T v = q[max(cid[m], 0)];
        val[m] = (cid[m] >= 0) ? v : T(0);
      }
      if (DOT) {
#pragma unroll
        for (int m = 0; m < RPL; ++m) pr[m] = p[max(min(r0 + (int)threadIdx.x + m * kThreads, r1 - 1), 0)];
      }
#pragma unroll
      for (int m = 0; m < EPT; ++m)
        if (cid[m] >= 0) s_v[g0 + threadIdx.x + m * kThreads - g0] = val[m];
    }
    __syncthreads();
    double dot = 0;
    if (staged) {
#pragma unroll
      for (int m = 0; m < RPL; ++m) {
        const int row = r0 + threadIdx.x + m * kThreads;
        if (row < r1) {
          T acc = T(0);
          for (int i = ra[m] - g0; i < re[m] - g0; ++i) acc += s_v[i];
          gq[row] = acc;
          if (DOT) dot += (double)pr[m] * (double)acc;
        }
      }
    } else {
      for (int row = r0 + threadIdx.x; row < r1; row += kThreads) {
        T acc = T(0);
        for (int g = rowStarts[row]; g < rowStarts[row + 1]; ++g) acc += q[colIds[g]];
        gq[row] = acc;
        if (DOT) dot += (double)p[row] * (double)acc;
      }
    }
    if (DOT) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
      if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = dot;
    }
    __syncthreads();  // s_v is restaged by the next super-block; s_w is complete
    if (DOT && threadIdx.x == 0) {
      double t = 0;
      for (int i = 0; i < kThreads / 64; ++i) t += s_w[i];
      if (FOLD) {
        const int a0 = (int)((long long)sb * nax / nsb), a1 = (int)((long long)(sb + 1) * nax / nsb);
        for (int i = a0; i < a1; ++i) t += axpart[i];
      }
      part[sb] = t;
    }
  }
}

// K > 1 interleaved vectors: lane <-> (row, k) with k fastest, so the K
// entries of one node are read by neighbouring lanes as one contiguous run
// of K*sizeof(T) bytes; the column ids are read once per row for all K
template <typename T>
__global__ __launch_bounds__(kThreads) void k_gather_k(int Nblocks, int K, const int* __restrict__ blockStarts,
                                                       const int* __restrict__ rowStarts,
                                                       const int* __restrict__ colIds, const T* __restrict__ q,
                                                       T* __restrict__ gq) {
  const int b0 = blockIdx.x * kSuper;
  const int b1 = min(b0 + kSuper, Nblocks);
  const int r0 = blockStarts[b0], r1 = blockStarts[b1];
  const long long n = (long long)(r1 - r0) * K;
  for (long long t = threadIdx.x; t < n; t += kThreads) {
    const int row = r0 + (int)(t / K), k = (int)(t % K);
    const int a = rowStarts[row], e = rowStarts[row + 1];
    T acc = T(0);
    for (int g = a; g < e; ++g) acc += q[k + (size_t)colIds[g] * K];
    gq[k + (size_t)row * K] = acc;
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_scatter_k(int Nblocks, int K, const int* __restrict__ blockStarts,
                                                        const int* __restrict__ rowStarts,
                                                        const int* __restrict__ colIds, const T* __restrict__ gq,
                                                        T* __restrict__ q) {
  const int b0 = blockIdx.x * kSuper;
  const int b1 = min(b0 + kSuper, Nblocks);
  const int r0 = blockStarts[b0], r1 = blockStarts[b1];
  const long long n = (long long)(r1 - r0) * K;
  for (long long t = threadIdx.x; t < n; t += kThreads) {
    const int row = r0 + (int)(t / K), k = (int)(t % K);
    const T v = gq[k + (size_t)row * K];
    const int a = rowStarts[row], e = rowStarts[row + 1];
    for (int g = a; g < e; ++g) q[k + (size_t)colIds[g] * K] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_scatter(int Nblocks, int K, const int* __restrict__ blockStarts,
                                                      const int* __restrict__ rowStarts,
                                                      const int* __restrict__ colIds, const T* __restrict__ gq,
                                                      T* __restrict__ q) {
  const int k = blockIdx.y;
  const int b0 = blockIdx.x * kSuper;
  const int b1 = min(b0 + kSuper, Nblocks);
  const int r0 = blockStarts[b0], r1 = blockStarts[b1];
  for (int row = r0 + threadIdx.x; row < r1; row += kThreads) {
    const T v = gq[k + (size_t)row * K];
    const int a = rowStarts[row], e = rowStarts[row + 1];
    for (int g = a; g < e; ++g) q[k + (size_t)colIds[g] * K] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_gather_scatter(int Nblocks, int K, const int* __restrict__ blockStarts,
                                                             const int* __restrict__ rowStarts,
                                                             const int* __restrict__ colIds, T* q) {
  __shared__ T s_v[kMaxNnz];
  const int k = blockIdx.y;
  const int b0 = blockIdx.x * kSuper;
  const int b1 = min(b0 + kSuper, Nblocks);
  const int r0 = blockStarts[b0], r1 = blockStarts[b1];
  const int g0 = rowStarts[r0], g1 = rowStarts[r1];
  if ((g1 - g0) > kMaxNnz) {  // caller-built oversized blocks: rows from memory
    for (int row = r0 + threadIdx.x; row < r1; row += kThreads) {
      const int a = rowStarts[row], e = rowStarts[row + 1];
      T acc = T(0);
      for (int g = a; g < e; ++g) acc += q[k + (size_t)colIds[g] * K];
      for (int g = a; g < e; ++g) q[k + (size_t)colIds[g] * K] = acc;
    }
    return;
  }
  for (int g = g0 + threadIdx.x; g < g1; g += kThreads) s_v[g - g0] = q[k + (size_t)colIds[g] * K];
  __syncthreads();
  for (int row = r0 + threadIdx.x; row < r1; row += kThreads) {
    const int a = rowStarts[row] - g0, e = rowStarts[row + 1] - g0;
    if (e <= a) continue;
    T acc = T(0);
    for (int i = a; i < e; ++i) acc += s_v[i];
    s_v[a] = acc;  // park the sum; entries of one row belong to one lane
  }
  __syncthreads();
  for (int row = r0 + threadIdx.x; row < r1; row += kThreads) {
    const int a = rowStarts[row] - g0, e = rowStarts[row + 1] - g0;
    if (e <= a) continue;
    const T v = s_v[a];
    for (int i = a; i < e; ++i) q[k + (size_t)colIds[g0 + i] * K] = v;
  }
}

int check(int Nblocks, int K, const void* bs, const void* rs, const void* ci, const void* a, const void* b) {
  if (Nblocks < 0 || K < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: Nblocks < 0 or K < 1");
  if (K > 65535) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: K above 65535");
  if (Nblocks > 0 && (!bs || !rs || !ci || !a || !b))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: null device pointer");
  return BPX_OK;
}

inline int super_blocks(int Nblocks) { return (Nblocks + kSuper - 1) / kSuper; }

inline int gather_steps(int Nblocks) { return (Nblocks + kGatherSuper - 1) / kGatherSuper; }
inline int gather_grid(int Nblocks) { return std::min((gather_steps(Nblocks) + 7) / 8 * 8, kGatherGrid); }

}  // namespace

int ogs_gather_partials(int Nblocks) { return Nblocks > 0 ? gather_steps(Nblocks) : 0; }

template <typename T>
int launch_ogs_gather(hipStream_t s, int Nblocks, int K, const int* blockStarts, const int* rowStarts,
                      const int* colIds, const T* q, T* gq, int mode) {
  if (int st = check(Nblocks, K, blockStarts, rowStarts, colIds, q, gq)) return st;
  if (Nblocks == 0) return BPX_OK;
  if (K > 1)
    hipLaunchKernelGGL(k_gather_k<T>, dim3(super_blocks(Nblocks)), dim3(kThreads), 0, s, Nblocks, K,
                       blockStarts, rowStarts, colIds, q, gq);
  else
    hipLaunchKernelGGL((k_gather<T, false, false>), dim3(gather_grid(Nblocks)), dim3(kThreads), 0, s, Nblocks,
                       mode, blockStarts, rowStarts, colIds, q, gq, (const T*)nullptr, (double*)nullptr,
                       (const double*)nullptr, 0);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_ogs_gather_dot(hipStream_t s, int Nblocks, const int* blockStarts, const int* rowStarts,
                          const int* colIds, const T* q, T* gq, const T* p, double* part, int mode) {
  if (int st = check(Nblocks, 1, blockStarts, rowStarts, colIds, q, gq)) return st;
  if (Nblocks == 0) return BPX_OK;
  hipLaunchKernelGGL((k_gather<T, true, false>), dim3(gather_grid(Nblocks)), dim3(kThreads), 0, s, Nblocks, mode,
                     blockStarts, rowStarts, colIds, q, gq, p, part, (const double*)nullptr, 0);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_ogs_gather_dot_fold(hipStream_t s, int Nblocks, const int* blockStarts, const int* rowStarts,
                               const int* colIds, const T* q, T* gq, const T* p, double* part, const double* axpart,
                               int nax, int mode) {
  if (int st = check(Nblocks, 1, blockStarts, rowStarts, colIds, q, gq)) return st;
  if (Nblocks == 0) return BPX_OK;
  if (!part || (nax > 0 && !axpart)) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs fold: null pointer");
  hipLaunchKernelGGL((k_gather<T, true, true>), dim3(gather_grid(Nblocks)), dim3(kThreads), 0, s, Nblocks, mode,
                     blockStarts, rowStarts, colIds, q, gq, p, part, axpart, nax);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_ogs_scatter(hipStream_t s, int Nblocks, int K, const int* blockStarts, const int* rowStarts,
                       const int* colIds, const T* gq, T* q) {
  if (int st = check(Nblocks, K, blockStarts, rowStarts, colIds, gq, q)) return st;
  if (Nblocks == 0) return BPX_OK;
  if (K > 1)
    hipLaunchKernelGGL(k_scatter_k<T>, dim3(super_blocks(Nblocks)), dim3(kThreads), 0, s, Nblocks, K,
                       blockStarts, rowStarts, colIds, gq, q);
  else
    hipLaunchKernelGGL(k_scatter<T>, dim3(super_blocks(Nblocks), 1), dim3(kThreads), 0, s, Nblocks, 1,
                       blockStarts, rowStarts, colIds, gq, q);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_ogs_gather_scatter(hipStream_t s, int Nblocks, int K, const int* blockStarts, const int* rowStarts,
                              const int* colIds, T* q) {
  if (int st = check(Nblocks, K, blockStarts, rowStarts, colIds, q, q)) return st;
  if (Nblocks == 0) return BPX_OK;
  hipLaunchKernelGGL(k_gather_scatter<T>, dim3(super_blocks(Nblocks), K), dim3(kThreads), 0, s, Nblocks, K,
                     blockStarts, rowStarts, colIds, q);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template int launch_ogs_gather<double>(hipStream_t, int, int, const int*, const int*, const int*,
                                       const double*, double*, int);
template int launch_ogs_gather<float>(hipStream_t, int, int, const int*, const int*, const int*,
                                      const float*, float*, int);
template int launch_ogs_gather_dot<double>(hipStream_t, int, const int*, const int*, const int*, const double*,
                                           double*, const double*, double*, int);
template int launch_ogs_gather_dot<float>(hipStream_t, int, const int*, const int*, const int*, const float*,
                                          float*, const float*, double*, int);
template int launch_ogs_gather_dot_fold<double>(hipStream_t, int, const int*, const int*, const int*, const double*,
                                                double*, const double*, double*, const double*, int, int);
template int launch_ogs_gather_dot_fold<float>(hipStream_t, int, const int*, const int*, const int*, const float*,
                                               float*, const float*, double*, const double*, int, int);
template int launch_ogs_scatter<double>(hipStream_t, int, int, const int*, const int*, const int*,
                                        const double*, double*);
template int launch_ogs_scatter<float>(hipStream_t, int, int, const int*, const int*, const int*,
                                       const float*, float*);
template int launch_ogs_gather_scatter<double>(hipStream_t, int, int, const int*, const int*, const int*,
                                               double*);

}  // namespace bpx

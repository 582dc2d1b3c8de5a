// ogs gather / scatter / gather-scatter on the T maps (Add).
//
// Semantics: libs/ogs/okl/ogsKernels.okl:86-122 (gather: each row summed
// from 0 in ascending column order), :127-164 (scatter), :32-81 (gather
// then scatter back). K vectors are interleaved: entry k of node n at
// k + n*K.
//
// Design (gfx950): atomic-free segmented sums over the precomputed sorted
// segment lists (rowStarts/colIds). A workgroup takes one of the <=512-nnz
// row blocks built at setup (blockStarts); each lane owns whole rows and
// sums its segment in column order, so every row is summed in exactly the
// reference order and the result is bit-reproducible. Consecutive rows
// have adjacent segments, so a wave's index loads cover a few cache lines
// and the value loads follow the first-appearance numbering (mostly
// ascending local ids).
#include <hip/hip_runtime.h>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

constexpr int kOgsThreads = 256;

template <typename T>
__global__ __launch_bounds__(kOgsThreads) void k_gather(int Nblocks, int K, const int* __restrict__ blockStarts,
                                                        const int* __restrict__ rowStarts,
                                                        const int* __restrict__ colIds,
                                                        const T* __restrict__ q, T* __restrict__ gq) {
  const int blk = blockIdx.x;
  const int k = blockIdx.y;
  const int r0 = blockStarts[blk], r1 = blockStarts[blk + 1];
  for (int row = r0 + threadIdx.x; row < r1; row += kOgsThreads) {
    const int g0 = rowStarts[row], g1 = rowStarts[row + 1];
    T acc = T(0);
    for (int g = g0; g < g1; ++g) acc += q[k + (size_t)colIds[g] * K];
    gq[k + (size_t)row * K] = acc;
  }
}

template <typename T>
__global__ __launch_bounds__(kOgsThreads) void k_scatter(int Nblocks, int K, const int* __restrict__ blockStarts,
                                                         const int* __restrict__ rowStarts,
                                                         const int* __restrict__ colIds,
                                                         const T* __restrict__ gq, T* __restrict__ q) {
  const int blk = blockIdx.x;
  const int k = blockIdx.y;
  const int r0 = blockStarts[blk], r1 = blockStarts[blk + 1];
  for (int row = r0 + threadIdx.x; row < r1; row += kOgsThreads) {
    const T v = gq[k + (size_t)row * K];
    const int g0 = rowStarts[row], g1 = rowStarts[row + 1];
    for (int g = g0; g < g1; ++g) q[k + (size_t)colIds[g] * K] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(kOgsThreads) void k_gather_scatter(int Nblocks, int K,
                                                                const int* __restrict__ blockStarts,
                                                                const int* __restrict__ rowStarts,
                                                                const int* __restrict__ colIds, T* q) {
  const int blk = blockIdx.x;
  const int k = blockIdx.y;
  const int r0 = blockStarts[blk], r1 = blockStarts[blk + 1];
  for (int row = r0 + threadIdx.x; row < r1; row += kOgsThreads) {
    const int g0 = rowStarts[row], g1 = rowStarts[row + 1];
    T acc = T(0);
    for (int g = g0; g < g1; ++g) acc += q[k + (size_t)colIds[g] * K];
    for (int g = g0; g < g1; ++g) q[k + (size_t)colIds[g] * K] = acc;
  }
}

int check(int Nblocks, int K, const void* bs, const void* rs, const void* ci, const void* a, const void* b) {
  if (Nblocks < 0 || K < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: Nblocks < 0 or K < 1");
  if (K > 65535) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: K above 65535");
  if (Nblocks > 0 && (!bs || !rs || !ci || !a || !b))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: null device pointer");
  return BPX_OK;
}

}  // namespace

template <typename T>
int launch_ogs_gather(hipStream_t s, int Nblocks, int K, const int* blockStarts, const int* rowStarts,
                      const int* colIds, const T* q, T* gq) {
  if (int st = check(Nblocks, K, blockStarts, rowStarts, colIds, q, gq)) return st;
  if (Nblocks == 0) return BPX_OK;
  hipLaunchKernelGGL(k_gather<T>, dim3(Nblocks, K), dim3(kOgsThreads), 0, s, Nblocks, K, blockStarts,
                     rowStarts, colIds, q, gq);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_ogs_scatter(hipStream_t s, int Nblocks, int K, const int* blockStarts, const int* rowStarts,
                       const int* colIds, const T* gq, T* q) {
  if (int st = check(Nblocks, K, blockStarts, rowStarts, colIds, gq, q)) return st;
  if (Nblocks == 0) return BPX_OK;
  hipLaunchKernelGGL(k_scatter<T>, dim3(Nblocks, K), dim3(kOgsThreads), 0, s, Nblocks, K, blockStarts,
                     rowStarts, colIds, gq, q);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T>
int launch_ogs_gather_scatter(hipStream_t s, int Nblocks, int K, const int* blockStarts,
                              const int* rowStarts, const int* colIds, T* q) {
  if (int st = check(Nblocks, K, blockStarts, rowStarts, colIds, q, q)) return st;
  if (Nblocks == 0) return BPX_OK;
  hipLaunchKernelGGL(k_gather_scatter<T>, dim3(Nblocks, K), dim3(kOgsThreads), 0, s, Nblocks, K,
                     blockStarts, rowStarts, colIds, q);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template int launch_ogs_gather<double>(hipStream_t, int, int, const int*, const int*, const int*,
                                       const double*, double*);
template int launch_ogs_gather<float>(hipStream_t, int, int, const int*, const int*, const int*,
                                      const float*, float*);
template int launch_ogs_scatter<double>(hipStream_t, int, int, const int*, const int*, const int*,
                                        const double*, double*);
template int launch_ogs_scatter<float>(hipStream_t, int, int, const int*, const int*, const int*,
                                       const float*, float*);
template int launch_ogs_gather_scatter<double>(hipStream_t, int, int, const int*, const int*,
                                               const int*, double*);

}  // namespace bpx

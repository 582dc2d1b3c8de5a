// BP5 / BP6 hex Ax (collocated GLL stiffness + lambda*mass), affine and
// non-affine, FP64 and FP32.
//
// Semantics: okl/bp5AxHex3D.okl:415-569 and okl/bp5AxAffineHex3D.okl:419-579
// (Aq = D^T G D q + lambda*wJ*q per element, q read through GlobalToLocal
// with -1 -> 0); BP6 is the same operator on 3 fields, local layout
// [element][field][node] (okl/bp6AxHex3D.okl:600-606).
//
// Design (gfx950): one workgroup handles EPB elements; each element is an
// Nq x Nq sheet of lanes, lane (i,j) owning the pencil q(i,j,0..Nq-1) in
// registers. The element is swept in k-slabs: the slab is staged in LDS for
// the r and s contractions, while the t contraction and its transpose stay
// in registers, so a node costs 2*Nq LDS reads for the gradient and 2*Nq for
// the weak divergence. For Nq <= 10 each lane also keeps row/column i and j
// of D in registers. The 7 geometric factors of slab k+1 are loaded while
// slab k computes, keeping the HBM stream of ggeo/wJ (56 B per node in
// FP64) in flight.
//
// k right-hand sides (an extension; the reference applies one vector per
// launch): vectors are interleaved as in the ogs K form, and each element is
// swept K/chunk times with chunk right-hand sides in registers; the sweeps
// of one element run back to back in one workgroup, so its geometry comes
// from HBM once and from L2 after that.
//
// This file is compiled once per Nq (-DBPX_NQ=n) so the instantiations
// build in parallel; ax_dispatch.cpp picks the instantiation at run time.
#include <hip/hip_runtime.h>

#include "ax_common.hpp"

#ifndef BPX_NQ
#error "compile with -DBPX_NQ=<Nq>"
#endif

namespace bpx {
namespace {

// All arithmetic is written as explicit fused multiply-adds and the file is
// compiled with -ffp-contract=off, so every instantiation (one or several
// right-hand sides, any workgroup shape) rounds identically and gives the
// same bits for the same input.
__device__ inline double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ inline float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// elements per workgroup: ~256 lanes for small Nq, one element from Nq=8
template <int NQ>
constexpr int bp5_epb() {
  return NQ >= 8 ? 1 : 256 / (NQ * NQ);
}

// occupancy target handed to the register allocator
template <int NQ, int NF, bool KRHS>
constexpr int bp5_waves() {
  if (KRHS) return NF <= 2 && NQ <= 8 ? 2 : 1;
  return NF == 1 ? (NQ <= 8 ? 3 : 2) : (NQ <= 8 ? 2 : 1);
}

// D rows/columns in registers (saves 4*Nq LDS reads per node) for one
// field per sweep; with several the registers hold the extra pencils
template <int NQ, int NF>
constexpr bool bp5_dreg() {
  return NQ <= 10 && NF == 1;
}

template <typename T, int NQ, int NF, int EPB, bool AFFINE, int WAVES, bool KRHS>
__global__ __launch_bounds__(NQ* NQ* EPB) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_bp5_ax(
    int Nelements, int K, const int* __restrict__ elementList, const int* __restrict__ G2L,
    const T* __restrict__ wJ, const T* __restrict__ ggeo, const T* __restrict__ gllw,
    const T* __restrict__ D, T lambda, const T* __restrict__ q, T* __restrict__ Aq) {
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  constexpr int LD = NQ + 1;  // padded LDS row
  constexpr bool DREG = bp5_dreg<NQ, NF>();

  __shared__ T s_D[NQ][LD];
  __shared__ T s_q[NF][EPB][NQ][LD];
  __shared__ T s_r[NF][EPB][NQ][LD];
  __shared__ T s_s[NF][EPB][NQ][LD];

  const int t = threadIdx.x;
  const int es = t / NQ2;
  const int ij = t - es * NQ2;
  const int i = ij % NQ;
  const int j = ij / NQ;
  const int eo = blockIdx.x * EPB + es;
  const bool active = eo < Nelements;
  const int e = active ? elementList[eo] : 0;

  for (int n = t; n < NQ2; n += NQ2 * EPB) s_D[n / NQ][n % NQ] = D[n];

  // KRHS: the pencil's gather rows, read once for all sweeps
  int gid[KRHS ? NQ : 1];
  if (KRHS) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) gid[k] = active ? G2L[(size_t)e * NP + k * NQ2 + ij] : -1;
  }

  const int nsweep = KRHS ? K / NF : 1;
  for (int c = 0; c < nsweep; ++c) {
  T rq[NF][NQ];
  T rA[NF][NQ];
  if (KRHS) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const T* src = q + (size_t)max(gid[k], 0) * K + c * NF;
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        rq[f][k] = (gid[k] >= 0) ? src[f] : T(0);
        rA[f][k] = T(0);
      }
    }
  } else {
    const size_t lbase = (size_t)e * NP * NF + ij;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        const int id = active ? G2L[lbase + f * NP + k * NQ2] : -1;
        rq[f][k] = (id >= 0) ? q[id] : T(0);
        rA[f][k] = T(0);
      }
  }

  // geometry of the current slab (non-affine: prefetched one slab ahead)
  T g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0, g5 = 0, gw = 0;
  T wij = 0;
  if (AFFINE) {
    const T* ge = ggeo + (size_t)6 * e;
    g0 = ge[0]; g1 = ge[1]; g2 = ge[2]; g3 = ge[3]; g4 = ge[4]; g5 = ge[5];
    gw = wJ[e];
    wij = gllw[i] * gllw[j];
  } else {
    const T* gp = ggeo + 6 * ((size_t)e * NP + ij);
    g0 = gp[0]; g1 = gp[1]; g2 = gp[2]; g3 = gp[3]; g4 = gp[4]; g5 = gp[5];
    gw = wJ[(size_t)e * NP + ij];
  }
  __syncthreads();

  T Di[DREG ? NQ : 1], Dj[DREG ? NQ : 1], DTi[DREG ? NQ : 1], DTj[DREG ? NQ : 1];
  if (DREG) {
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      Di[m] = s_D[i][m];
      Dj[m] = s_D[j][m];
      DTi[m] = s_D[m][i];
      DTj[m] = s_D[m][j];
    }
  }

#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    // keep the accumulators materialised at slab boundaries so the
    // compiler does not sink their updates past later slabs
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int m = 0; m < NQ; ++m) asm volatile("" : "+v"(rA[f][m]));

    T G00, G01, G02, G11, G12, G22, W;
    if (AFFINE) {
      const T w = wij * gllw[k];
      G00 = g0 * w; G01 = g1 * w; G02 = g2 * w; G11 = g3 * w; G12 = g4 * w; G22 = g5 * w;
      W = gw * w;
    } else {
      G00 = g0; G01 = g1; G02 = g2; G11 = g3; G12 = g4; G22 = g5; W = gw;
      if (k + 1 < NQ) {
        const size_t gb = (size_t)e * NP + (k + 1) * NQ2 + ij;
        const T* gp = ggeo + 6 * gb;
        g0 = gp[0]; g1 = gp[1]; g2 = gp[2]; g3 = gp[3]; g4 = gp[4]; g5 = gp[5];
        gw = wJ[gb];
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f) s_q[f][es][j][i] = rq[f][k];
    __syncthreads();
    asm volatile("" ::: "memory");

#pragma unroll
    for (int f = 0; f < NF; ++f) {
      T qr = 0, qs = 0, qt = 0;
#pragma unroll
      for (int m = 0; m < NQ; ++m) {
        qr = fmad((DREG ? Di[m] : s_D[i][m]), s_q[f][es][j][m], qr);
        qs = fmad((DREG ? Dj[m] : s_D[j][m]), s_q[f][es][m][i], qs);
        qt = fmad(s_D[k][m], rq[f][m], qt);
      }
      const T Gr = fmad(G02, qt, fmad(G01, qs, G00 * qr));
      const T Gs = fmad(G12, qt, fmad(G11, qs, G01 * qr));
      const T Gt = fmad(G22, qt, fmad(G12, qs, G02 * qr));
      s_r[f][es][j][i] = Gr;
      s_s[f][es][j][i] = Gs;
      rA[f][k] = fmad(lambda * W, rq[f][k], rA[f][k]);
#pragma unroll
      for (int m = 0; m < NQ; ++m) rA[f][m] = fmad(s_D[k][m], Gt, rA[f][m]);
    }
    __syncthreads();
    asm volatile("" ::: "memory");

#pragma unroll
    for (int f = 0; f < NF; ++f) {
      T acc = 0;
#pragma unroll
      for (int m = 0; m < NQ; ++m) {
        acc = fmad((DREG ? DTi[m] : s_D[m][i]), s_r[f][es][j][m], acc);
        acc = fmad((DREG ? DTj[m] : s_D[m][j]), s_s[f][es][m][i], acc);
      }
      rA[f][k] += acc;
    }
  }

  if (active) {
    if (KRHS) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        T* out = Aq + ((size_t)e * NP + k * NQ2 + ij) * K + c * NF;
#pragma unroll
        for (int f = 0; f < NF; ++f) out[f] = rA[f][k];
      }
    } else {
      T* out = Aq + (size_t)e * NP * NF + ij;
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int k = 0; k < NQ; ++k) out[f * NP + k * NQ2] = rA[f][k];
    }
  }
  }  // sweeps
}

template <typename T, int NQ, int NF, bool KRHS>
int launch_bp5(hipStream_t s, const AxArgs<T>& a) {
  constexpr int EPB = bp5_epb<NQ>();
  constexpr int WV = bp5_waves<NQ, NF, KRHS>();
  const int blocks = (a.Nelements + EPB - 1) / EPB;
  if (blocks == 0) return BPX_OK;
  if (a.affine)
    hipLaunchKernelGGL((k_bp5_ax<T, NQ, NF, EPB, true, WV, KRHS>), dim3(blocks), dim3(NQ * NQ * EPB), 0,
                       s, a.Nelements, a.K, a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.gllw, a.D,
                       a.lambda, a.q, a.Aq);
  else
    hipLaunchKernelGGL((k_bp5_ax<T, NQ, NF, EPB, false, WV, KRHS>), dim3(blocks), dim3(NQ * NQ * EPB), 0,
                       s, a.Nelements, a.K, a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.gllw, a.D,
                       a.lambda, a.q, a.Aq);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

}  // namespace

template <typename T, int NQ, int NF>
int launch_bp5_nq(hipStream_t s, const AxArgs<T>& a) {
  return launch_bp5<T, NQ, NF, false>(s, a);
}

template <typename T, int NQ>
int launch_bp5_krhs_nq(hipStream_t s, const AxArgs<T>& a, int chunk) {
  switch (chunk) {
    case 1: return launch_bp5<T, NQ, 1, true>(s, a);
    case 2: return launch_bp5<T, NQ, 2, true>(s, a);
    case 4: return launch_bp5<T, NQ, 4, true>(s, a);
    default: return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 k-RHS Ax: chunk must be 1, 2 or 4");
  }
}

template int launch_bp5_nq<double, BPX_NQ, 1>(hipStream_t, const AxArgs<double>&);
template int launch_bp5_nq<double, BPX_NQ, 3>(hipStream_t, const AxArgs<double>&);
template int launch_bp5_nq<float, BPX_NQ, 1>(hipStream_t, const AxArgs<float>&);
template int launch_bp5_krhs_nq<double, BPX_NQ>(hipStream_t, const AxArgs<double>&, int);

}  // namespace bpx

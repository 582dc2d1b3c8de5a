// BP5 hex Ax at Nq = 16 (p = 15), FP64, one field, with the r and s
// contractions on the matrix cores.
//
// Semantics: okl/bp5AxHex3D.okl:415-569 and okl/bp5AxAffineHex3D.okl:419-579
// (Aq = D^T G D q + lambda*wJ*q per element, q read through GlobalToLocal
// with -1 -> 0), the same operator as k_bp5_ax; only the order of the
// additions differs (results agree to rounding, not bit for bit).
//
// Design (gfx950): one wavefront per element. The element is 16 slabs
// (k) of 16 x 16 nodes (j, i); a slab is handled in the accumulator layout
// of v_mfma_f64_16x16x4_f64: lane (g, c) = (l / 16, l % 16) owns nodes
// (k, j = 4v + g, i = c), v = 0..3, so the 64 lanes' v-th nodes are 64
// consecutive nodes of the slab. q of the whole element sits in LDS
// (rows padded to 17 doubles, 35 KB per wavefront, 4 wavefronts per CU);
// Aq of the whole element stays in registers (64 accumulators per lane).
//   s contraction (along j, the layout's rows): D * Q_k, 4 MFMA with D as
//     the register operand and the slab as read from LDS;
//   r contraction (along i, the columns): Q_k * D^T, 4 MFMA whose slab
//     operand is the slab read transposed;
//   t contraction (across slabs): 16 FMA per node with D[k][m] uniform
//     over the wavefront and q of the 16 slabs read from LDS; the weak
//     divergence's t part is the matching scatter into the 16 slab
//     accumulators, all in registers.
// The weak divergence's r and s parts are the same MFMA with D^T, the r
// part's operand transposed through a 2 KB wave-private tile. A node costs
// 21 LDS accesses (the vector-ALU kernel at Nq = 16: ~160, which bound it)
// and D takes 8 registers per lane (D[c][4v+g] and D[4v+g][c]). The
// geometric factors of slab k+1 are loaded while slab k computes.
#include <hip/hip_runtime.h>

#include "bpx_internal.hpp"

namespace bpx {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

// C (16x16) += A (16x4) * B (4x16); lane l supplies A[l%16][l/16] and
// B[l/16][l%16] and holds C[4v + l/16][l%16] in component v (FP64 rows
// interleave by component)
__device__ inline d4 mfma_16x16x4(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int NQ = 16;
constexpr int NQ2 = NQ * NQ;
constexpr int NP = NQ2 * NQ;
constexpr int LW = NQ + 1;  // padded row of the transpose tile

// FUSED: the fused CG operator's output (launch_bp5_fused): interior nodes
// to Ap rows ldq + e*fused_interior_pad + (interior index), surface nodes
// to Aq + e*fused_surface_pad, both runs written whole with zero padding,
// and the interior p.Ap sum to part[workgroup]. p.Ap is summed in the order
// of the vector-ALU fused kernel (256 lanes: lane t takes interior indices
// t, t+256, ..., then a butterfly within each wavefront and the 4 sums in
// order); this wavefront plays its 4 wavefronts, so the partials have the
// same bits as well.
template <bool AFFINE, bool FUSED>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_bp5_ax_mfma16(
    const int* __restrict__ elementList, const int* __restrict__ G2L, const double* __restrict__ wJ,
    const double* __restrict__ ggeo, const double* __restrict__ gllw, const double* __restrict__ D,
    double lambda, const double* __restrict__ q, double* __restrict__ Aq, long long ldq,
    double* __restrict__ Ap, double* __restrict__ part) {
  __shared__ double s_q[NQ][NQ][LW];  // q(k, j, i)
  __shared__ double s_t[NQ][LW];      // transpose tile
  const int l = threadIdx.x;
  const int c = l & 15;  // i of every node of the lane
  const int g = l >> 4;  // j = 4v + g
  const int e = elementList[blockIdx.x];
  // node (k, 4v+g, c) of the element: k*256 + 64v + 16g + c
  const size_t lane0 = (size_t)e * NP + l;

  // geometry of slab 0 first, so its loads overlap the gather of q
  double G[7][4], Gn[7][4];
  double ga[6] = {0, 0, 0, 0, 0, 0}, gwe = 0, wij[4] = {0, 0, 0, 0};
  if (AFFINE) {
#pragma unroll
    for (int n = 0; n < 6; ++n) ga[n] = ggeo[6 * (size_t)e + n];
    gwe = wJ[e];
#pragma unroll
    for (int v = 0; v < 4; ++v) wij[v] = gllw[c] * gllw[4 * v + g];
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const size_t n = lane0 + 64 * v;
      const double* gp = ggeo + 6 * n;
#pragma unroll
      for (int f = 0; f < 6; ++f) Gn[f][v] = gp[f];
      Gn[6][v] = wJ[n];
    }
  }

  double DA[4], DT[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    DA[v] = D[c * NQ + 4 * v + g];    // D[c][4v+g]
    DT[v] = D[(4 * v + g) * NQ + c];  // D[4v+g][c]
  }

  // gather q into LDS, 8 slabs per batch: every load of a batch is issued
  // before the first is used (a load under a branch, stored straight to
  // LDS, would wait out one memory round trip per node). Masked nodes
  // (id < 0) read wJ[0], which is never empty (q may be), and store 0.
#pragma unroll
  for (int k0 = 0; k0 < NQ; k0 += 8) {
    int id[8][4];
    double qv[8][4];
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) id[k][v] = G2L[lane0 + (k0 + k) * NQ2 + 64 * v];
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) qv[k][v] = *(id[k][v] >= 0 ? q + id[k][v] : wJ);
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) s_q[k0 + k][4 * v + g][c] = id[k][v] >= 0 ? qv[k][v] : 0.0;
  }
  d4 acc[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) acc[k] = d4{0.0, 0.0, 0.0, 0.0};
  __syncthreads();

#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    // slabs stay in program order: hoisting later slabs' scalar and
    // geometry loads overflows the registers
    __builtin_amdgcn_sched_barrier(0);
    // this slab's geometric factors; start loading the next slab's
    if (AFFINE) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double w = wij[v] * gllw[k];
#pragma unroll
        for (int f = 0; f < 6; ++f) G[f][v] = ga[f] * w;
        G[6][v] = gwe * w;
      }
    } else {
#pragma unroll
      for (int f = 0; f < 7; ++f)
#pragma unroll
        for (int v = 0; v < 4; ++v) G[f][v] = Gn[f][v];
      if (k + 1 < NQ) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const size_t n = lane0 + (k + 1) * NQ2 + 64 * v;
          const double* gp = ggeo + 6 * n;
#pragma unroll
          for (int f = 0; f < 6; ++f) Gn[f][v] = gp[f];
          Gn[6][v] = wJ[n];
        }
      }
    }

    double rk[4], qa[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      rk[v] = s_q[k][4 * v + g][c];  // Q_k[4v+g][c]
      qa[v] = s_q[k][c][4 * v + g];  // Q_k[c][4v+g]
    }
    // qs = D * Q_k (rows j), qr = Q_k * D^T (columns i)
    d4 qs = {0.0, 0.0, 0.0, 0.0}, qr = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qs = mfma_16x16x4(DA[v], rk[v], qs);
      qr = mfma_16x16x4(qa[v], DA[v], qr);
    }

    // qt = sum_m D[k][m] Q_m (across slabs)
    // row k of D, uniform over the wavefront: scalar registers
    const double* Dr = D + k * NQ;
    double Dk[NQ];
#pragma unroll
    for (int m = 0; m < NQ; ++m) Dk[m] = Dr[m];
    double qt[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double a = 0.0;
#pragma unroll
      for (int m = 0; m < NQ; ++m) a = __builtin_fma(Dk[m], s_q[m][4 * v + g][c], a);
      qt[v] = a;
    }

    double gr[4], gs[4], gt[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const double G00 = G[0][v], G01 = G[1][v], G02 = G[2][v], G11 = G[3][v], G12 = G[4][v],
                   G22 = G[5][v], W = G[6][v];
      gr[v] = __builtin_fma(G02, qt[v], __builtin_fma(G01, qs[v], G00 * qr[v]));
      gs[v] = __builtin_fma(G12, qt[v], __builtin_fma(G11, qs[v], G01 * qr[v]));
      gt[v] = __builtin_fma(G22, qt[v], __builtin_fma(G12, qs[v], G02 * qr[v]));
      acc[k][v] = __builtin_fma(lambda * W, rk[v], acc[k][v]);
    }

    // t part of the weak divergence: slab m gains D[k][m] gt
#pragma unroll
    for (int m = 0; m < NQ; ++m)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[m][v] = __builtin_fma(Dk[m], gt[v], acc[m][v]);

    // s part: D^T * GS_k; r part: GR_k * D through the transpose tile
#pragma unroll
    for (int v = 0; v < 4; ++v) s_t[4 * v + g][c] = gr[v];
    __syncthreads();
    double gra[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) gra[v] = s_t[c][4 * v + g];  // GR_k[c][4v+g]
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      acc[k] = mfma_16x16x4(DT[v], gs[v], acc[k]);
      acc[k] = mfma_16x16x4(gra[v], DT[v], acc[k]);
    }
    __syncthreads();  // the tile is rewritten by the next slab
  }

  if (!FUSED) {
    double* out = Aq + lane0;
#pragma unroll
    for (int k = 0; k < NQ; ++k)
#pragma unroll
      for (int v = 0; v < 4; ++v) out[k * NQ2 + 64 * v] = acc[k][v];
    return;
  }

  // fused: Aq of the element to LDS in node order (q is no longer read)
  constexpr int NI = fused_interior(NQ);
  constexpr int NIP = fused_interior_pad(NQ);
  constexpr int NS = NP - NI;
  constexpr int NSP = fused_surface_pad(NQ);
  static_assert(NP <= NQ * NQ * LW, "the node-order copy fits in the q tile");
  // p of the lane's interior rows (indices l + 64w + 256t), all requested
  // before the staging so the loads overlap it (clamped: past the run the
  // value is not used)
  constexpr int NT = (NIP + 255) / 256;
  const double* pin = q + ldq + (size_t)e * NIP;  // interior rows of p
  double pv[4][NT];
#pragma unroll
  for (int w = 0; w < 4; ++w)
#pragma unroll
    for (int t = 0; t < NT; ++t) pv[w][t] = pin[min(l + 64 * w + 256 * t, NI - 1)];
  __syncthreads();  // every lane has read its last q from the tile
  double* s_out = &s_q[0][0][0];
#pragma unroll
  for (int k = 0; k < NQ; ++k)
#pragma unroll
    for (int v = 0; v < 4; ++v) s_out[k * NQ2 + 64 * v + l] = acc[k][v];
  __syncthreads();
  double* inner = Ap + ldq + (size_t)e * NIP;
  double pd[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int w = 0; w < 4; ++w)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int n = l + 64 * w + 256 * t;
      if (n < NIP) {
        double val = 0.0;
        if (n < NI) {
          val = s_out[fused_interior_node(NQ, n)];
          pd[w] = fma(pv[w][t], val, pd[w]);
        }
        inner[n] = val;
      }
    }
  double* surf = Aq + (size_t)e * NSP;
  for (int n = l; n < NSP; n += 64) surf[n] = (n < NS) ? s_out[fused_surface_node(NQ, n)] : 0.0;
  double tot = 0.0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    double r = pd[w];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
    tot += r;
  }
  if (l == 0) part[blockIdx.x] = tot;
}

}  // namespace

namespace {
template <bool FUSED>
int launch_mfma16(hipStream_t s, const AxArgs<double>& a, double* Ap, double* part) {
  if (a.N != 15 || a.Nfields != 1)
    return set_error(BPX_ERR_UNSUPPORTED, "bp5 matrix-core Ax: degree 15, one field only");
  if (a.Nelements == 0) return BPX_OK;
  if (a.affine)
    launch_stamped(k_bp5_ax_mfma16<true, FUSED>, dim3(a.Nelements), dim3(64), s, a.evStart, a.evStop,
                   a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.gllw, a.D, a.lambda, a.q, a.Aq, a.ldq, Ap,
                   part);
  else
    launch_stamped(k_bp5_ax_mfma16<false, FUSED>, dim3(a.Nelements), dim3(64), s, a.evStart, a.evStop,
                   a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.gllw, a.D, a.lambda, a.q, a.Aq, a.ldq, Ap,
                   part);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}
}  // namespace

int launch_bp5_ax_mfma16(hipStream_t s, const AxArgs<double>& a) {
  return launch_mfma16<false>(s, a, nullptr, nullptr);
}

int launch_bp5_fused_mfma16(hipStream_t s, const AxArgs<double>& a, double* Ap, double* part) {
  return launch_mfma16<true>(s, a, Ap, part);
}

}  // namespace bpx

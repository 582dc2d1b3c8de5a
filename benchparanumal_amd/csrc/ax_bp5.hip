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
// launch): the right-hand sides are stored one after another (leading
// dimensions ldq, ldA), and each element is swept K/chunk times with chunk
// right-hand sides in registers; the sweeps of one element run back to back
// in one workgroup, so its geometry comes from HBM once and from cache
// after that, while every sweep reads and writes whole cache lines.
//
// This file is compiled once per Nq (-DBPX_NQ=n) so the instantiations
// build in parallel; ax_dispatch.cpp picks the instantiation at run time.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "ax_common.hpp"
#include "buffer_rsrc.hpp"

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

// the one-field FP32 Ax at Nq = 8 fits 7 waves per SIMD without spilling
// (68 VGPRs with D in LDS); measured 52.0 against 53.0 us at 24^3. The
// FP64 one at Nq >= 13 is held to 168 VGPRs (3 waves instead of 2, spills
// of at most 28 B): more elements, so more bytes in flight, per CU;
// measured 570 against 682 us at p=12 and 1139 against 1194 us at p=15.
template <typename T, int NQ, int NF, int VAR>
constexpr int bp5_waves_v() {
  if (sizeof(T) == 4 && NF == 1 && VAR == 0 && NQ == 8) return 7;
  if (sizeof(T) == 8 && NF == 1 && VAR == 0 && NQ >= 13) return 3;
  return bp5_waves<NQ, NF, VAR == 1>();
}

// occupancy target of the FP32 kernels with the trilinear geometry (the
// plain one at Nq = 8 spills at 7 and at 5 waves; 4 hold it)
template <int NQ, int VAR>
constexpr int bp5_waves_tri() {
  return (VAR == 0 && NQ == 8) ? 4 : bp5_waves_v<float, NQ, 1, VAR>();
}

// D rows/columns in registers (saves 4*Nq LDS reads per node) for one
// field per sweep in FP64 up to Nq = 8; with several fields the registers
// hold the extra pencils. Measured at 24^3, p=7: the FP64 plain Ax is
// faster with them (102.6 against 107.8 us); the FP32 kernels and the FP64
// fused one are faster without, at higher occupancy (FP32 fused Ax 55.7
// against 59.5 us, FP64 fused 107.6 against 134.8 us, which spilled; with
// the branch-free gather it spills 5 registers, and the FP64 CG still ran
// 232.5 against 226.5 ms with D in registers), and
// so is the FP64 plain one at Nq = 9 (3 waves instead of 2: 100.5 against
// 110.7 us at 20^3, p=8; even at Nq = 10). The batch and k-RHS variants
// keep D in registers up to Nq = 10 (the ragged batch ran 4% slower
// without).
template <typename T, int NQ, int NF, int VAR>
constexpr bool bp5_dreg() {
  return sizeof(T) == 8 && NF == 1 && VAR != 3 && (NQ <= 8 || (NQ <= 10 && VAR != 0));
}

// KRHS: K right-hand sides stored one after another, r-th at q + r*ldq and
// Aq + r*ldA; the element is swept K/NF times, NF right-hand sides per
// sweep, re-reading its geometry (from cache: the sweeps of one element run
// back to back in one workgroup).
// BATCH: elements of many problems; an element whose problem
// (elemPid[e]) is done is skipped.
// FUSED (one field), rows numbered so that element e's interior nodes are
// rows S0 + e*fused_interior_pad(NQ) + (interior index in local order), the
// only copies of their rows: their results go straight to Ap (a one-entry
// gather row sums to the value itself) with the p.Ap terms summed into
// part[workgroup]. The element's surface nodes go, in local order, to
// Aq + e*fused_surface_pad(NQ) for the gather of the shared rows. S0 comes
// in ldq. Padding entries of both runs are written with zeros, so every
// cache line of Ap and Aq is written whole by one workgroup.
enum { VAR_PLAIN = 0, VAR_KRHS = 1, VAR_BATCH = 2, VAR_FUSED = 3 };

// k-RHS workgroup kernel: geometry loads non-temporal (1) or cached (0);
// elements dealt to the XCDs round-robin (0) or in contiguous eighths (1)
#ifndef BPX_KRHS_GEO_NT
#define BPX_KRHS_GEO_NT 1
#endif
#ifndef BPX_KRHS_XCD_BLOCK
#define BPX_KRHS_XCD_BLOCK 0
#endif

// Geometry loads of the fused (CG) operator are non-temporal: an FP32 CG
// iteration cycles ~340 MB through the 256 MB Infinity Cache, ~200 MB of it
// the geometric factors, read once per iteration. Streaming them past the
// cache keeps the CG vectors, the gather maps and the slots resident
// (measured at 24^3, p=7: mixed solve 128.8 -> 122.7 ms, FP64 CG 236.0 ->
// 226.1 ms; the Ax itself unchanged). The plain kernel keeps cached loads:
// launched back to back (the Ax benchmark) it reuses geometry from the cache.
// The fused epilogue's run index -> element-local node maps as tables (one
// 16-bit entry per run position; 0xFFFF = padding): a lane looks its node
// up with one load instead of forming it with integer divisions and the
// surface run's case analysis (about 250 of the FP32 operator's 1,194 VALU
// instructions per element at Nq = 8 went to that).
template <int NQ>
struct FusedNodeTables {
  unsigned short inner[fused_interior_pad(NQ) > 0 ? fused_interior_pad(NQ) : 1];
  unsigned short surf[fused_surface_pad(NQ)];
  constexpr FusedNodeTables() : inner(), surf() {
    for (int n = 0; n < fused_interior_pad(NQ); ++n)
      inner[n] = n < fused_interior(NQ) ? (unsigned short)fused_interior_node(NQ, n) : (unsigned short)0xFFFF;
    for (int n = 0; n < fused_surface_pad(NQ); ++n)
      surf[n] = n < NQ * NQ * NQ - fused_interior(NQ) ? (unsigned short)fused_surface_node(NQ, n)
                                                       : (unsigned short)0xFFFF;
  }
};
__constant__ FusedNodeTables<BPX_NQ> kFusedNodes = FusedNodeTables<BPX_NQ>();

template <bool STREAM, typename T>
__device__ inline T geo_load(const T* p) {
  if constexpr (STREAM) return __builtin_nontemporal_load(p);
  return *p;
}

// TRI (FP32, one field, non-affine; plain or fused): the geometric factors
// are computed in the kernel from the element's trilinear map instead of
// read from memory. ggeo then holds 24 values per element, for x, y and z
// in turn the coefficients of r, s, t, rs, rt, st, rst of
// x(r,s,t) = sum_v (1+r_v r)(1+s_v s)(1+t_v t) X_v / 8 and one pad, and wJ
// holds the Nq GLL nodes. Lane (i,j) keeps the derivatives' pencil
// constants (x_r and x_s are linear in t along the pencil, x_t is
// constant); each node then forms the adjugate of the Jacobian, J and
// G = adj adj^T w / J (the formula of the stored factors), ~55 VALU
// operations in place of 28 B of HBM traffic per node.
// FOLD (FUSED only): the CG direction update of the element's interior rows
// (rows >= S0 = ldq, each the only copy of its row) is done here, as
// k_cg_update_p does it for the rows below S0: x <- alpha_{k-1} p + x and
// p <- r + beta p (p <- r at k = 0), the same fused multiply-adds, so the
// solve gives the bits it gives with the update in its own kernel. q is
// then the direction p itself (fold.p, updated in place: no other lane
// reads an interior row).
template <typename T, int NQ, int NF, int EPB, bool AFFINE, int WAVES, int VAR, bool TRI = false, bool FOLD = false,
          int FS = 1>
__global__ __launch_bounds__(NQ* NQ* EPB) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_bp5_ax(
    int Nelements, int K, const int* __restrict__ elementList, const int* __restrict__ G2L,
    const T* __restrict__ wJ, const T* __restrict__ ggeo, const T* __restrict__ gllw,
    const T* __restrict__ D, T lambda, const T* q, T* __restrict__ Aq, long long ldq,
    long long ldA, const int* __restrict__ elemPid, const BatchState* __restrict__ bst, T* __restrict__ Ap,
    double* __restrict__ part, CgFold<T> fold) {
  constexpr bool KRHS = VAR == VAR_KRHS;
  constexpr bool BATCH = VAR == VAR_BATCH;
  constexpr bool FUSED = VAR == VAR_FUSED;
  static_assert(!FUSED || NF == 1, "the fused operator covers one field");
  static_assert(!FOLD || FUSED, "the folded direction update is a form of the fused operator");
  static_assert(!TRI || (sizeof(T) == 4 && NF == 1 && !AFFINE && (VAR == VAR_PLAIN || FUSED)),
                "TRI: FP32, one field, non-affine");
  static_assert(FS == 1 || (NF == 1 && EPB == 1 && VAR == VAR_PLAIN && !TRI),
                "field split: the plain one-field sweep, one element per workgroup");
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  constexpr int LD = NQ + 1;  // padded LDS row
  // TRI moves ~77 MB per launch at 24^3, p = 7 instead of ~274 MB, and
  // the LDS reads become the bound: D rows/columns in registers and the
  // broadcast D[k][m] from scalar (constant address space) loads
  // (measured Ax 44.1 -> 38.2 -> 36.5 us, the mixed solve 111.1 -> 101.6 ms).
  // Both only at Nq = 8 and 16. With D in registers and D[k][m] from LDS
  // the FP32 operator came out 20-42% off at Nq = 5..15 except 8 (right at
  // 2, 3, 4): the compiler's IR load/store vectorizer merges the load of
  // s_D[i][1] with the lane's column load s_D[0][i] as if the two were
  // adjacent (they are only for i = 0), so lanes i > 0 use a wrong D entry;
  // the same code built with that pass off is right at every Nq. Those Nq
  // keep D in LDS, which the pass leaves intact.
  constexpr bool DREG = (TRI && NQ % 8 == 0) || bp5_dreg<T, NQ, NF, VAR>();
  constexpr bool SDK = TRI && NQ % 8 == 0;
  typedef const __attribute__((address_space(4))) T* cptr;
  const cptr Dc = (cptr)D;

  __shared__ T s_D[NQ][LD];
  __shared__ T s_q[NF][EPB][NQ][LD];
  __shared__ T s_r[NF][EPB][NQ][LD];
  __shared__ T s_s[NF][EPB][NQ][LD];

  const int t = threadIdx.x;
  const int es = t / NQ2;
  const int ij = t - es * NQ2;
  const int i = ij % NQ;
  const int j = ij / NQ;
  // FS > 1: each workgroup applies the operator to one field of an
  // FS-field layout ([element][field][node]). Block b takes field
  // (b >> 3) % FS of element ((b >> 3) / FS) * 8 + b % 8, so the FS
  // workgroups of an element run one after another on one XCD (block mod
  // 8), and the second and third read its geometric factors from that L2.
  int fsel = 0;
  int eo = blockIdx.x * EPB + es;
  if constexpr (FS > 1) {
    const int r = (int)(blockIdx.x >> 3);
    fsel = r % FS;
    eo = (r / FS) * 8 + (int)(blockIdx.x & 7);
  }
  bool active = eo < Nelements;
  const int e = active ? elementList[eo] : 0;
  if (BATCH) {
    if (active && bst[elemPid[e]].done) active = false;
    if (EPB == 1 && !active) return;  // one element per workgroup: uniform exit
  }

  for (int n = t; n < NQ2; n += NQ2 * EPB) s_D[n / NQ][n % NQ] = D[n];

  // FUSED: the epilogue's node tables, loaded now so their latency hides
  // behind the sweep (positions past a run's end read its last entry)
  constexpr int FUI = (FUSED && fused_interior_pad(NQ) > 0) ? (fused_interior_pad(NQ) + NQ2 - 1) / NQ2 : 1;
  constexpr int FUS = FUSED ? (fused_surface_pad(NQ) + NQ2 - 1) / NQ2 : 1;
  int fni[FUI], fns[FUS];
  if constexpr (FUSED) {
#pragma unroll
    for (int u = 0; u < FUI; ++u)
      fni[u] = fused_interior_pad(NQ) > 0 ? kFusedNodes.inner[min(ij + u * NQ2, fused_interior_pad(NQ) - 1)] : 0xFFFF;
#pragma unroll
    for (int u = 0; u < FUS; ++u) fns[u] = kFusedNodes.surf[min(ij + u * NQ2, fused_surface_pad(NQ) - 1)];
  }

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
  // Every load below is issued unconditionally from a valid address (a
  // masked node reads wJ[0], never empty, and stores 0): a load under a
  // branch makes the compiler wait out each G2L -> q round trip in turn.
  if (KRHS) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const T* src = q + (size_t)(c * NF) * ldq + max(gid[k], 0);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const T v = *(gid[k] >= 0 ? src + (size_t)f * ldq : wJ);
        rq[f][k] = (gid[k] >= 0) ? v : T(0);
        rA[f][k] = T(0);
      }
    }
  } else {
    const size_t lbase = ((size_t)e * FS + fsel) * NP * NF + ij;  // e = 0 on an idle sheet: in bounds
    int id[NF][NQ];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 0; k < NQ; ++k) id[f][k] = active ? G2L[lbase + f * NP + k * NQ2] : -1;
    // FOLD: which of the CG updates of iteration fold.k apply (the stopping
    // test and its flags are k_cg_update_p's, written before this launch)
    bool fx = false, fp = false, ffirst = false;
    T fa = T(0), fb = T(0);
    T fr[FOLD ? NQ : 1], fxv[FOLD ? NQ : 1];
    // the lane's pencil crosses interior rows at k = 1 .. NQ-2 when it is an
    // interior column of the element
    const bool icol = FOLD && active && i >= 1 && i <= NQ - 2 && j >= 1 && j <= NQ - 2;
    if constexpr (FOLD) {
      const int kk = fold.k;
      const int stop_at = fold.st->stop_at;
      const bool done_before = kk > 0 && (fold.st->done || (stop_at != 0 && stop_at <= kk));
      if (!done_before) {
        fx = fold.pending && kk > 0;
        fp = stop_at != kk + 1;
        ffirst = kk == 0;
        if (fx) fa = (T)fold.hist_alpha[kk - 1];
        if (fp && !ffirst) fb = (T)(fold.hist_rr[kk] / fold.hist_rr[kk - 1]);
      }
      if (icol && (fx || fp)) {
#pragma unroll
        for (int k = 1; k < NQ - 1; ++k) {
          fr[k] = fold.r[id[0][k]];
          fxv[k] = fold.x[id[0][k]];
        }
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        const T v = *(id[f][k] >= 0 ? q + id[f][k] : wJ);
        rq[f][k] = (id[f][k] >= 0) ? v : T(0);
        rA[f][k] = T(0);
      }
    if constexpr (FOLD) {
      if (icol && (fx || fp)) {
#pragma unroll
        for (int k = 1; k < NQ - 1; ++k) {
          const T pv = rq[0][k];
          if (fx) fold.x[id[0][k]] = fmad(fa, pv, fxv[k]);
          if (fp) {
            const T pn = ffirst ? fr[k] : fmad(fb, pv, fr[k]);
            rq[0][k] = pn;
            fold.p[id[0][k]] = pn;
          }
        }
      }
    }
  }

  // geometry of the current slab (non-affine: prefetched one slab ahead)
  T g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0, g5 = 0, gw = 0;
  T wij = 0;
  // TRI: along the pencil, x_r = Ar + Br t, x_s = As + Bs t, x_t = Ct, so
  // the adjugate rows J grad r = x_s x x_t = Pr + Qr t and J grad s =
  // x_t x x_r = Ps + Qs t are linear in t, J grad t = x_r x x_s = T0 + T1 t
  // + T2 t^2 is quadratic, and so is J = x_t . (J grad t) = j0 + j1 t + j2 t^2
  // (14 operations per node instead of 27 for the derivatives, adjugate and J)
  T Pr[TRI ? 3 : 1], Qr[TRI ? 3 : 1], Ps[TRI ? 3 : 1], Qs[TRI ? 3 : 1];
  T T0[TRI ? 3 : 1], T1[TRI ? 3 : 1], T2[TRI ? 3 : 1], jt[TRI ? 3 : 1];
  if (TRI) {
    const T r = wJ[i], sn = wJ[j];
    const T* ce = ggeo + (size_t)24 * e;
    T Ar[3], Br[3], As[3], Bs[3], Ct[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const T c_r = ce[8 * d + 0], c_s = ce[8 * d + 1], c_t = ce[8 * d + 2], c_rs = ce[8 * d + 3];
      const T c_rt = ce[8 * d + 4], c_st = ce[8 * d + 5], c_rst = ce[8 * d + 6];
      Ar[d] = fmad(c_rs, sn, c_r);
      Br[d] = fmad(c_rst, sn, c_rt);
      As[d] = fmad(c_rs, r, c_s);
      Bs[d] = fmad(c_rst, r, c_st);
      Ct[d] = fmad(Br[d], r, fmad(c_st, sn, c_t));
    }
    // (a x b)_d = a_{d+1} b_{d+2} - a_{d+2} b_{d+1}
    auto cross = [](const T* a, const T* b, int d) {
      const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
      return fmad(a[d1], b[d2], -(a[d2] * b[d1]));
    };
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      Pr[d] = cross(As, Ct, d);
      Qr[d] = cross(Bs, Ct, d);
      Ps[d] = cross(Ct, Ar, d);
      Qs[d] = cross(Ct, Br, d);
      T0[d] = cross(Ar, As, d);
      T1[d] = cross(Ar, Bs, d) + cross(Br, As, d);
      T2[d] = cross(Br, Bs, d);
    }
    jt[0] = fmad(Ct[0], T0[0], fmad(Ct[1], T0[1], Ct[2] * T0[2]));
    jt[1] = fmad(Ct[0], T1[0], fmad(Ct[1], T1[1], Ct[2] * T1[2]));
    jt[2] = fmad(Ct[0], T2[0], fmad(Ct[1], T2[1], Ct[2] * T2[2]));
    wij = gllw[i] * gllw[j];
  } else if (AFFINE) {
    const T* ge = ggeo + (size_t)6 * e;
    g0 = ge[0]; g1 = ge[1]; g2 = ge[2]; g3 = ge[3]; g4 = ge[4]; g5 = ge[5];
    gw = wJ[e];
    wij = gllw[i] * gllw[j];
  } else {
    const T* gp = ggeo + 6 * ((size_t)e * NP + ij);
    g0 = geo_load<FUSED>(gp + 0); g1 = geo_load<FUSED>(gp + 1); g2 = geo_load<FUSED>(gp + 2);
    g3 = geo_load<FUSED>(gp + 3); g4 = geo_load<FUSED>(gp + 4); g5 = geo_load<FUSED>(gp + 5);
    gw = geo_load<FUSED>(wJ + (size_t)e * NP + ij);
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

    T G00 = 0, G01 = 0, G02 = 0, G11 = 0, G12 = 0, G22 = 0, W;
    // TRI: rows of the adjugate (J grad r, J grad s, J grad t) and w/J; G is
    // applied as adj (w/J) adj^T without forming it (21 operations per node
    // instead of 24 to form G and 9 to apply it)
    T arx = 0, ary = 0, arz = 0, asx = 0, asy = 0, asz = 0, atx = 0, aty = 0, atz = 0, sc = 0;
    if (TRI) {
      const T tk = wJ[k];
      arx = fmad(Qr[0], tk, Pr[0]); ary = fmad(Qr[1], tk, Pr[1]); arz = fmad(Qr[2], tk, Pr[2]);
      asx = fmad(Qs[0], tk, Ps[0]); asy = fmad(Qs[1], tk, Ps[1]); asz = fmad(Qs[2], tk, Ps[2]);
      atx = fmad(fmad(T2[0], tk, T1[0]), tk, T0[0]);
      aty = fmad(fmad(T2[1], tk, T1[1]), tk, T0[1]);
      atz = fmad(fmad(T2[2], tk, T1[2]), tk, T0[2]);
      const T J = fmad(fmad(jt[2], tk, jt[1]), tk, jt[0]);
      const T w = wij * gllw[k];
      // hardware reciprocal (1 ulp) in place of the correctly rounded
      // division, which takes ~10 instructions (Ax 44.8 -> 44.1 us)
      sc = w * __builtin_amdgcn_rcpf(J);
      W = J * w;
    } else if (AFFINE) {
      const T w = wij * gllw[k];
      G00 = g0 * w; G01 = g1 * w; G02 = g2 * w; G11 = g3 * w; G12 = g4 * w; G22 = g5 * w;
      W = gw * w;
    } else {
      G00 = g0; G01 = g1; G02 = g2; G11 = g3; G12 = g4; G22 = g5; W = gw;
      if (k + 1 < NQ) {
        const size_t gb = (size_t)e * NP + (k + 1) * NQ2 + ij;
        const T* gp = ggeo + 6 * gb;
        g0 = geo_load<FUSED>(gp + 0); g1 = geo_load<FUSED>(gp + 1); g2 = geo_load<FUSED>(gp + 2);
        g3 = geo_load<FUSED>(gp + 3); g4 = geo_load<FUSED>(gp + 4); g5 = geo_load<FUSED>(gp + 5);
        gw = geo_load<FUSED>(wJ + gb);
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
        qt = fmad((SDK ? Dc[k * NQ + m] : s_D[k][m]), rq[f][m], qt);
      }
      T Gr, Gs, Gt;
      if (TRI) {
        const T gx = sc * fmad(atx, qt, fmad(asx, qs, arx * qr));
        const T gy = sc * fmad(aty, qt, fmad(asy, qs, ary * qr));
        const T gz = sc * fmad(atz, qt, fmad(asz, qs, arz * qr));
        Gr = fmad(arz, gz, fmad(ary, gy, arx * gx));
        Gs = fmad(asz, gz, fmad(asy, gy, asx * gx));
        Gt = fmad(atz, gz, fmad(aty, gy, atx * gx));
      } else {
        Gr = fmad(G02, qt, fmad(G01, qs, G00 * qr));
        Gs = fmad(G12, qt, fmad(G11, qs, G01 * qr));
        Gt = fmad(G22, qt, fmad(G12, qs, G02 * qr));
      }
      s_r[f][es][j][i] = Gr;
      s_s[f][es][j][i] = Gs;
      rA[f][k] = fmad(lambda * W, rq[f][k], rA[f][k]);
#pragma unroll
      for (int m = 0; m < NQ; ++m) rA[f][m] = fmad((SDK ? Dc[k * NQ + m] : s_D[k][m]), Gt, rA[f][m]);
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

  if (FUSED) {
    // the element's results go through LDS (every lane writes all of its
    // nodes, as the plain kernel does to the local vector); then the
    // element's lanes copy its interior run to Ap and its surface run to
    // Aq, each as one contiguous, coalesced sweep with zero padding
    constexpr int NI = fused_interior(NQ);
    constexpr int NIP = fused_interior_pad(NQ);
    constexpr int NS = NP - NI;
    constexpr int NSP = fused_surface_pad(NQ);
    __shared__ T s_out[FUSED ? EPB : 1][FUSED ? NP : 1];
    __shared__ T s_pin[FUSED ? EPB : 1][FUSED ? NP : 1];
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      s_out[es][k * NQ2 + ij] = rA[0][k];
      s_pin[es][k * NQ2 + ij] = rq[0][k];
    }
    __syncthreads();
    double pd = 0;
    if (active) {
      T* inner = Ap + ldq + (size_t)e * NIP;
#pragma unroll
      for (int u = 0; u < FUI; ++u) {
        const int n = ij + u * NQ2;
        if (n < NIP) {
          T v = T(0);
          if (fni[u] != 0xFFFF) {
            v = s_out[es][fni[u]];
            pd = fma((double)s_pin[es][fni[u]], (double)v, pd);
          }
          inner[n] = v;
        }
      }
      T* surf = Aq + (size_t)e * NSP;
#pragma unroll
      for (int u = 0; u < FUS; ++u) {
        const int n = ij + u * NQ2;
        if (n < NSP) surf[n] = (fns[u] != 0xFFFF) ? s_out[es][fns[u]] : T(0);
      }
    }
    // workgroup sum of the p.Ap terms (fixed order). When NQ^2*EPB is not a
    // multiple of 64 the last wavefront is partial, and the registers of
    // its missing lanes hold whatever an earlier wave left there: a lane
    // adds its butterfly partner's value only where that lane exists (full
    // wavefronts add every value, as a plain butterfly does)
    const int lane = t & 63;
    const int nlive = min(64, NQ2 * EPB - (t & ~63));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double o = __shfl_xor(pd, off, 64);
      if ((lane ^ off) < nlive) pd += o;
    }
    if (EPB * NQ2 <= 64) {
      if (t == 0) part[blockIdx.x] = pd;
    } else {
      __shared__ double s_pd[(NQ2 * EPB + 63) / 64];
      if ((t & 63) == 0) s_pd[t >> 6] = pd;
      __syncthreads();
      if (t == 0) {
        double acc = 0;
        for (int w = 0; w < (NQ2 * EPB + 63) / 64; ++w) acc += s_pd[w];
        part[blockIdx.x] = acc;
      }
    }
  } else if (active) {
    if (KRHS) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) {
        T* out = Aq + (size_t)(c * NF) * ldA + (size_t)e * NP + k * NQ2 + ij;
#pragma unroll
        for (int f = 0; f < NF; ++f) out[(size_t)f * ldA] = rA[f][k];
      }
    } else {
      T* out = Aq + ((size_t)e * FS + fsel) * NP * NF + ij;
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int k = 0; k < NQ; ++k) out[f * NP + k * NQ2] = rA[f][k];
    }
  }
  }  // sweeps
}

// k right-hand sides with the geometric factors read once per workgroup
// (non-affine): a workgroup holds one element as W lane sheets, sheet s
// applying the operator to NF right-hand sides of its own; the 7 factors of
// slab k+1 are loaded by all W*Nq^2 lanes together (about 7/W loads per
// lane), staged in an LDS double buffer (component-major, so each node's
// reads are conflict-free) and read from there by every sheet. A pass
// covers W*NF right-hand sides; K needs ceil(K / (W*NF)) passes per element,
// and the blocks of one element's passes are placed on the same XCD (block
// b runs on XCD b mod 8), so a pass after the first finds the factors in
// that XCD's L2. The arithmetic is the single-vector kernel's, term by term
// (D from LDS, the same fma order), so each right-hand side gets the bits
// k_bp5_ax gives it.
template <int NQ>
constexpr int krhs_wg_max_sheets() {
  return (1024 / (NQ * NQ)) < 16 ? (1024 / (NQ * NQ)) : 16;
}

template <typename T, int NQ, int NF, int W, bool DR>
__global__ __launch_bounds__(NQ* NQ* W) __attribute__((amdgpu_waves_per_eu(DR ? 2 : 4))) void k_bp5_krhs_wg(
    int Nelements, int K, int npass, const int* __restrict__ elementList, const int* __restrict__ G2L,
    const T* __restrict__ wJ, const T* __restrict__ ggeo, const T* __restrict__ D, T lambda,
    const T* __restrict__ q, T* __restrict__ Aq, long long ldq, long long ldA) {
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  constexpr int LD = NQ + 1;
  constexpr int NG = 7 * NQ2;  // factors of one slab
  constexpr int NT = W * NQ2;  // lanes
  // D[k][m] is the same for every lane: scalar loads (constant address space)
  typedef const __attribute__((address_space(4))) T* cptr;
  const cptr Dc = (cptr)D;
  // block -> (element, pass): the passes of element eo share its XCD
  const int xcd = blockIdx.x & 7;
  const int m = blockIdx.x >> 3;
  const int pass = m % npass;
#if BPX_KRHS_XCD_BLOCK
  // each XCD a contiguous eighth of the elements (neighbours share q lines in its L2)
  const int eo = xcd * (int)((Nelements + 7) / 8) + m / npass;
#else
  const int eo = (m / npass) * 8 + xcd;
#endif
  if (eo >= Nelements) return;  // uniform over the workgroup, before any barrier
  const int e = elementList[eo];

  extern __shared__ __align__(16) unsigned char krhs_smem[];
  T* s_D = reinterpret_cast<T*>(krhs_smem);  // [NQ][LD]
  T* s_geo = s_D + NQ * LD;                  // [2][7][NQ2]
  T* s_work = s_geo + 2 * NG;                // per sheet: s_q, s_r, s_s [NF][NQ][LD] each

  const int t = threadIdx.x;
  const int sh = t / NQ2;
  const int ij = t - sh * NQ2;
  const int i = ij % NQ;
  const int j = ij / NQ;
  const int r0 = (pass * W + sh) * NF;  // first right-hand side of this sheet
  const bool live = r0 < K;
  T* s_q = s_work + (size_t)sh * 3 * NF * NQ * LD;
  T* s_r = s_q + NF * NQ * LD;
  T* s_s = s_r + NF * NQ * LD;
  auto SQ = [&](T* base, int f, int a, int b) -> T& { return base[(f * NQ + a) * LD + b]; };

  for (int n = t; n < NQ2; n += NT) s_D[(n / NQ) * LD + n % NQ] = D[n];

  // cooperative load of the factors of slab k into registers, then LDS
  constexpr int GMAX = (NG + NT - 1) / NT;  // loads per lane
  T gv[GMAX];
  auto load_slab = [&](int k) {
    const T* gs = ggeo + ((size_t)e * NP + (size_t)k * NQ2) * 6;
    const T* ws = wJ + (size_t)e * NP + (size_t)k * NQ2;
#pragma unroll
    for (int u = 0; u < GMAX; ++u) {
      const int f = t + u * NT;
#if BPX_KRHS_GEO_NT
      gv[u] = (f < 6 * NQ2) ? __builtin_nontemporal_load(gs + f) : (f < NG ? __builtin_nontemporal_load(ws + (f - 6 * NQ2)) : T(0));
#else
      gv[u] = (f < 6 * NQ2) ? gs[f] : (f < NG ? ws[f - 6 * NQ2] : T(0));
#endif
    }
  };
  auto store_slab = [&](int buf) {
    T* g = s_geo + buf * NG;
#pragma unroll
    for (int u = 0; u < GMAX; ++u) {
      const int f = t + u * NT;
      if (f < 6 * NQ2) g[(f % 6) * NQ2 + f / 6] = gv[u];
      else if (f < NG) g[6 * NQ2 + (f - 6 * NQ2)] = gv[u];
    }
  };

  // the pencil's gather rows and the sheet's right-hand sides (branch-free:
  // a masked node, or an idle sheet, reads wJ[0] and keeps 0)
  int gid[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) gid[k] = G2L[(size_t)e * NP + k * NQ2 + ij];
  load_slab(0);
  T rq[NF][NQ], rA[NF][NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const bool on = live && gid[k] >= 0;
    const T* src = q + (size_t)(live ? r0 : 0) * ldq + max(gid[k], 0);
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      const T v = *(on ? src + (size_t)f * ldq : wJ);
      rq[f][k] = on ? v : T(0);
      rA[f][k] = T(0);
    }
  }
  store_slab(0);
  __syncthreads();
  // DR: row and column i and j of D in registers (4 Nq LDS reads fewer per
  // node and right-hand side), at fewer waves per SIMD
  T Di[DR ? NQ : 1], Dj[DR ? NQ : 1], DTi[DR ? NQ : 1], DTj[DR ? NQ : 1];
  if constexpr (DR) {
#pragma unroll
    for (int mm = 0; mm < NQ; ++mm) {
      Di[mm] = s_D[i * LD + mm];
      Dj[mm] = s_D[j * LD + mm];
      DTi[mm] = s_D[mm * LD + i];
      DTj[mm] = s_D[mm * LD + j];
    }
  }

#pragma unroll
  for (int k = 0; k < NQ; ++k) {
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int mm = 0; mm < NQ; ++mm) asm volatile("" : "+v"(rA[f][mm]));
    if (k + 1 < NQ) load_slab(k + 1);
#pragma unroll
    for (int f = 0; f < NF; ++f) SQ(s_q, f, j, i) = rq[f][k];
    __syncthreads();
    asm volatile("" ::: "memory");

    const T* g = s_geo + (k & 1) * NG;
    const T G00 = g[0 * NQ2 + ij], G01 = g[1 * NQ2 + ij], G02 = g[2 * NQ2 + ij];
    const T G11 = g[3 * NQ2 + ij], G12 = g[4 * NQ2 + ij], G22 = g[5 * NQ2 + ij];
    const T Wk = g[6 * NQ2 + ij];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      T qr = 0, qs = 0, qt = 0;
#pragma unroll
      for (int mm = 0; mm < NQ; ++mm) {
        qr = fmad(DR ? Di[mm] : s_D[i * LD + mm], SQ(s_q, f, j, mm), qr);
        qs = fmad(DR ? Dj[mm] : s_D[j * LD + mm], SQ(s_q, f, mm, i), qs);
        qt = fmad(Dc[k * NQ + mm], rq[f][mm], qt);
      }
      const T Gr = fmad(G02, qt, fmad(G01, qs, G00 * qr));
      const T Gs = fmad(G12, qt, fmad(G11, qs, G01 * qr));
      const T Gt = fmad(G22, qt, fmad(G12, qs, G02 * qr));
      SQ(s_r, f, j, i) = Gr;
      SQ(s_s, f, j, i) = Gs;
      rA[f][k] = fmad(lambda * Wk, rq[f][k], rA[f][k]);
#pragma unroll
      for (int mm = 0; mm < NQ; ++mm) rA[f][mm] = fmad(Dc[k * NQ + mm], Gt, rA[f][mm]);
    }
    // slab k+1's factors into the other buffer (its last readers, in slab
    // k-1, finished before the barrier above)
    if (k + 1 < NQ) store_slab((k + 1) & 1);
    __syncthreads();
    asm volatile("" ::: "memory");

#pragma unroll
    for (int f = 0; f < NF; ++f) {
      T acc = 0;
#pragma unroll
      for (int mm = 0; mm < NQ; ++mm) {
        acc = fmad(DR ? DTi[mm] : s_D[mm * LD + i], SQ(s_r, f, j, mm), acc);
        acc = fmad(DR ? DTj[mm] : s_D[mm * LD + j], SQ(s_s, f, mm, i), acc);
      }
      rA[f][k] += acc;
    }
  }
  if (live) {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      T* out = Aq + (size_t)r0 * ldA + (size_t)e * NP + k * NQ2 + ij;
#pragma unroll
      for (int f = 0; f < NF; ++f) out[(size_t)f * ldA] = rA[f][k];
    }
  }
}

// k right-hand sides, pencil form: one element per Nq^2 lanes (one
// wavefront at Nq = 8; floor(64 / Nq^2) elements per workgroup below Nq = 6).
// The element's geometric factors are loaded once, into registers (a lane
// keeps the 7 factors of its pencil along k), and the workgroup applies the
// operator to the K right-hand sides one after another, with q loaded two
// right-hand sides ahead (gathered through bounds-checked buffer loads, so
// a masked node costs no compare or select). Each
// one-dimensional contraction runs along a pencil that a single lane holds,
// with D[a][b] the same for every lane (scalar loads): the t derivative on
// the lane's pencil along k, the r and s derivatives on pencils along i and
// along j, to which q moves through LDS (lane (a, b) = (t % Nq, t / Nq) of
// an element holds node (k, b, a) for k = 0..Nq-1, the pencil along i of
// (j, k) = (a, b), or the pencil along j of (i, k) = (a, b)). A node's value
// crosses LDS once per move: 16 LDS accesses per node and right-hand side
// (q out twice and in twice, qr and qs back, Gr and Gs out, their divergence
// parts back), against 42 in the sheet kernels, where every contraction
// reads its Nq operands from LDS. Two LDS buffers, one per moving direction:
// s_b (i pencils) holds node (k, j, i) at k*LK + j*Nq + (i + j) mod Nq,
// s_c (j pencils) at k*LK + j*Nq + i; at Nq = 8 with LK = 72 each 16 lanes
// of an access fall on distinct 8-byte banks in both layouts of each
// buffer, and every offset but the i pencils' is a lane base plus a
// constant. The divergence's r and s parts are summed apart (each along its
// own pencil), so results agree with k_bp5_ax to rounding, not bit for bit.
template <int NQ>
constexpr int krhs_pencil_epb() {
  return NQ * NQ >= 32 ? 1 : 64 / (NQ * NQ);
}

template <typename T, int NQ, bool AFFINE>
__global__ __launch_bounds__(NQ* NQ* krhs_pencil_epb<NQ>()) __attribute__((amdgpu_waves_per_eu(2))) void k_bp5_krhs_pencil(
    int Nelements, int K, const int* __restrict__ elementList, const int* __restrict__ G2L,
    const T* __restrict__ wJ, const T* __restrict__ ggeo, const T* __restrict__ gllw, const T* __restrict__ D,
    T lambda, const T* __restrict__ q, T* __restrict__ Aq, long long ldq, long long ldA) {
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  constexpr int EPB = krhs_pencil_epb<NQ>();
  constexpr int LK = NQ2 + 8;  // slab stride in LDS
  constexpr int LE = NQ * LK;  // element stride in LDS
  __shared__ T s_b[EPB * LE];  // i pencils: q, qr, Gr, the divergence's r part
  __shared__ T s_c[EPB * LE];  // j pencils: q, qs, Gs, the divergence's s part
  typedef const __attribute__((address_space(4))) T* cptr;
  cptr Dc = (cptr)D;
  // D[a][b] is read with scalar loads where it is used: an opaque update of
  // the pointer before each row keeps the compiler from holding all Nq^2
  // entries in scalar registers (128 at Nq = 8, more than a wavefront has)
#define BPX_DROW()                     \
  do {                                 \
    __builtin_amdgcn_sched_barrier(0); \
    asm volatile("" : "+s"(Dc));       \
  } while (0)
  const int t = threadIdx.x;
  const int es = t / NQ2;
  const int tl = t - es * NQ2;
  const int a = tl % NQ;
  const int b = tl / NQ;
  // EPB elements per workgroup; each XCD (block mod 8) takes a contiguous
  // eighth of the element groups, so face neighbours share q lines in its L2
  const int groups = (Nelements + EPB - 1) / EPB;
  const int per_xcd = (groups + 7) / 8;
  const int grp = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (grp >= groups) return;  // uniform over the workgroup
  const int eo = grp * EPB + es;
  const bool active = eo < Nelements;  // the last group may be partly empty
  const int e = elementList[active ? eo : 0];
  const size_t base = (size_t)e * NP + tl;  // node (k, b, a) is base + k*NQ2

  // LDS slots. Along k (node (k, b, a)): kb + k*LK in s_b, kc + k*LK in s_c;
  // along i (node (b, a, m)): oi[m] in s_b; along j (node (b, m, a)):
  // jc + m*NQ in s_c
  const int kb = es * LE + b * NQ + (a + b) % NQ;
  const int kc = es * LE + b * NQ + a;
  const int jc = es * LE + b * LK + a;
  int oi[NQ];
#pragma unroll
  for (int m = 0; m < NQ; ++m) oi[m] = es * LE + b * LK + a * NQ + (m + a) % NQ;

  // q is read through a buffer resource of kBufRange bytes per right-hand
  // side: a masked node (or an idle lane of a partly empty group, which
  // reads element elementList[0] and stores nothing) gets an offset beyond
  // that range, which the hardware answers with 0 without touching memory
  // (no compare or select per load). Rows must lie below kBufRange / sizeof(T).
  int qoff[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int v = G2L[base + k * NQ2];
    qoff[k] = (active && v >= 0) ? v * (int)sizeof(T) : kBufMasked;
  }
  // byte offset of node (0, b, a) in a right-hand side of Aq (the same range)
  const int aoff = active ? (int)(base * sizeof(T)) : kBufMasked;
  T g[7][AFFINE ? 1 : NQ];
  T ga[6] = {0, 0, 0, 0, 0, 0}, gw = 0, wab = 0;
  if constexpr (AFFINE) {
#pragma unroll
    for (int f = 0; f < 6; ++f) ga[f] = ggeo[6 * (size_t)e + f];
    gw = wJ[e];
    wab = gllw[a] * gllw[b];
  } else {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      const T* gp = ggeo + 6 * (base + k * NQ2);
#pragma unroll
      for (int f = 0; f < 6; ++f) g[f][k] = __builtin_nontemporal_load(gp + f);
      g[6][k] = __builtin_nontemporal_load(wJ + base + k * NQ2);
    }
  }

  // rows of D through scalar registers, row n+1 loaded while row n is in
  // use; body(n, row) for n = 0..Nq-1
  auto for_rows = [&](auto&& body) {
    T d[2][NQ];
    BPX_DROW();
#pragma unroll
    for (int m = 0; m < NQ; ++m) d[0][m] = Dc[m];
#pragma unroll
    for (int n = 0; n < NQ; ++n) {
      if (n + 1 < NQ) {
        BPX_DROW();
#pragma unroll
        for (int m = 0; m < NQ; ++m) d[(n + 1) & 1][m] = Dc[(n + 1) * NQ + m];
      }
      body(n, d[n & 1]);
    }
  };

  // q of right-hand side r along the lane's pencil
  auto load_q = [&](T(&dst)[NQ], int r) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(q + (size_t)r * ldq), (short)0, kBufRange, kBufFlags);
#pragma unroll
    for (int k = 0; k < NQ; ++k) dst[k] = buffer_load<T>(rs, qoff[k]);
  };

  // the operator on right-hand side r, q along the lane's pencil in rq
  auto apply = [&](const T(&rq)[NQ], int r) {
    // (each lane writes and, after the last barrier of the previous
    // right-hand side, last read only its own slots along k)
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      s_b[k * LK + kb] = rq[k];
      s_c[k * LK + kc] = rq[k];
    }
    __syncthreads();

    // r derivative along i at (j, k) = (a, b); s derivative along j at
    // (i, k) = (a, b); each lane overwrites only its own pencil
    {
      T v[NQ];
#pragma unroll
      for (int m = 0; m < NQ; ++m) v[m] = s_b[oi[m]];
      for_rows([&](int n, const T* dr) {
        T acc = 0;
#pragma unroll
        for (int m = 0; m < NQ; ++m) acc = fmad(dr[m], v[m], acc);
        s_b[oi[n]] = acc;
      });
#pragma unroll
      for (int m = 0; m < NQ; ++m) v[m] = s_c[jc + m * NQ];
      for_rows([&](int n, const T* dr) {
        T acc = 0;
#pragma unroll
        for (int m = 0; m < NQ; ++m) acc = fmad(dr[m], v[m], acc);
        s_c[jc + n * NQ] = acc;
      });
    }
    __syncthreads();

    // along k: t derivative, G, the mass term and the divergence's t part
    T rA[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) rA[k] = T(0);
    for_rows([&](int k, const T* dr) {
      T qt = 0;
#pragma unroll
      for (int m = 0; m < NQ; ++m) qt = fmad(dr[m], rq[m], qt);
      const T qr = s_b[k * LK + kb];
      const T qs = s_c[k * LK + kc];
      T G00, G01, G02, G11, G12, G22, W;
      if constexpr (AFFINE) {
        const T w = wab * gllw[k];
        G00 = ga[0] * w; G01 = ga[1] * w; G02 = ga[2] * w; G11 = ga[3] * w; G12 = ga[4] * w; G22 = ga[5] * w;
        W = gw * w;
      } else {
        G00 = g[0][k]; G01 = g[1][k]; G02 = g[2][k]; G11 = g[3][k]; G12 = g[4][k]; G22 = g[5][k]; W = g[6][k];
      }
      s_b[k * LK + kb] = fmad(G02, qt, fmad(G01, qs, G00 * qr));
      s_c[k * LK + kc] = fmad(G12, qt, fmad(G11, qs, G01 * qr));
      const T Gt = fmad(G22, qt, fmad(G12, qs, G02 * qr));
      rA[k] = fmad(lambda * W, rq[k], rA[k]);
#pragma unroll
      for (int m = 0; m < NQ; ++m) rA[m] = fmad(dr[m], Gt, rA[m]);
    });
    __syncthreads();

    // the divergence's r part along i, its s part along j (rows of D;
    // each output summed over m in order)
    {
      T v[NQ], acc[NQ];
#pragma unroll
      for (int m = 0; m < NQ; ++m) v[m] = s_b[oi[m]];
#pragma unroll
      for (int n = 0; n < NQ; ++n) acc[n] = T(0);
      for_rows([&](int m, const T* dr) {
#pragma unroll
        for (int n = 0; n < NQ; ++n) acc[n] = fmad(dr[n], v[m], acc[n]);
      });
#pragma unroll
      for (int n = 0; n < NQ; ++n) s_b[oi[n]] = acc[n];
#pragma unroll
      for (int m = 0; m < NQ; ++m) v[m] = s_c[jc + m * NQ];
#pragma unroll
      for (int n = 0; n < NQ; ++n) acc[n] = T(0);
      for_rows([&](int m, const T* dr) {
#pragma unroll
        for (int n = 0; n < NQ; ++n) acc[n] = fmad(dr[n], v[m], acc[n]);
      });
#pragma unroll
      for (int n = 0; n < NQ; ++n) s_c[jc + n * NQ] = acc[n];
    }
    __syncthreads();

    // buffer stores, no branch: an idle lane's offset lies beyond the range
    // and its stores are dropped (every path then has the same stores in
    // flight, so the next load's wait need not cover them)
    const __amdgpu_buffer_rsrc_t ro =
        __builtin_amdgcn_make_buffer_rsrc(Aq + (size_t)r * ldA, (short)0, kBufRange, kBufFlags);
#pragma unroll
    for (int k = 0; k < NQ; ++k)
      buffer_store<T>(rA[k] + (s_b[k * LK + kb] + s_c[k * LK + kc]), ro, aoff + k * NQ2 * (int)sizeof(T));
  };

  // three register sets for q, each loaded two right-hand sides ahead: a
  // load is waited for only where its values are first used, past the
  // stores of the right-hand sides before (the first right-hand side is
  // peeled so that the loop is entered with the same loads and stores in
  // flight as its back edge carries; the early exits leave the loop)
  T q0[NQ], q1[NQ], q2[NQ];
  load_q(q0, 0);
  load_q(q1, min(1, K - 1));
  load_q(q2, min(2, K - 1));
  apply(q0, 0);
  for (int r = 1; r < K; r += 3) {
    load_q(q0, min(r + 2, K - 1));
    apply(q1, r);
    if (r + 1 >= K) break;
    load_q(q1, min(r + 3, K - 1));
    apply(q2, r + 1);
    if (r + 2 >= K) break;
    load_q(q2, min(r + 4, K - 1));
    apply(q0, r + 2);
  }
#undef BPX_DROW
}

// Adjoint gradient of the variable-coefficient BP5 (an extension):
// gL[e*Np + n] = -(D lam)_n . G_n (D u)_n with the unscaled factors G_n,
// u and lam global vectors read through GlobalToLocal (-1 -> 0). Same lane
// sheet and slab sweep as the Ax kernel, without the weak divergence.
template <int NQ, int EPB, bool AFFINE>
__global__ __launch_bounds__(NQ* NQ* EPB) void k_bp5_kappa_grad(int Nelements, const int* __restrict__ G2L,
                                                               const double* __restrict__ ggeo,
                                                               const double* __restrict__ gllw,
                                                               const double* __restrict__ D,
                                                               const double* __restrict__ u,
                                                               const double* __restrict__ lam,
                                                               double* __restrict__ gL) {
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  constexpr int LD = NQ + 1;
  __shared__ double s_D[NQ][LD];
  __shared__ double s_u[EPB][NQ][LD];
  __shared__ double s_l[EPB][NQ][LD];

  const int t = threadIdx.x;
  const int es = t / NQ2;
  const int ij = t - es * NQ2;
  const int i = ij % NQ;
  const int j = ij / NQ;
  const int e = blockIdx.x * EPB + es;
  const bool active = e < Nelements;
  for (int n = t; n < NQ2; n += NQ2 * EPB) s_D[n / NQ][n % NQ] = D[n];

  // loads from valid addresses without branches (masked: ggeo[0], stored 0)
  double ru[NQ], rl[NQ];
  int id[NQ];
#pragma unroll
  for (int k = 0; k < NQ; ++k) id[k] = G2L[(size_t)(active ? e : 0) * NP + k * NQ2 + ij];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const bool on = active && id[k] >= 0;
    const double vu = *(on ? u + id[k] : ggeo);
    const double vl = *(on ? lam + id[k] : ggeo);
    ru[k] = on ? vu : 0.0;
    rl[k] = on ? vl : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    s_u[es][j][i] = ru[k];
    s_l[es][j][i] = rl[k];
    __syncthreads();
    double ur = 0, us = 0, ut = 0, lr = 0, ls = 0, lt = 0;
#pragma unroll
    for (int m = 0; m < NQ; ++m) {
      ur = fmad(s_D[i][m], s_u[es][j][m], ur);
      us = fmad(s_D[j][m], s_u[es][m][i], us);
      ut = fmad(s_D[k][m], ru[m], ut);
      lr = fmad(s_D[i][m], s_l[es][j][m], lr);
      ls = fmad(s_D[j][m], s_l[es][m][i], ls);
      lt = fmad(s_D[k][m], rl[m], lt);
    }
    if (active) {
      double G00, G01, G02, G11, G12, G22;
      if (AFFINE) {
        const double* g = ggeo + (size_t)6 * e;
        const double w = gllw[i] * gllw[j] * gllw[k];
        G00 = g[0] * w; G01 = g[1] * w; G02 = g[2] * w; G11 = g[3] * w; G12 = g[4] * w; G22 = g[5] * w;
      } else {
        const double* g = ggeo + 6 * ((size_t)e * NP + k * NQ2 + ij);
        G00 = g[0]; G01 = g[1]; G02 = g[2]; G11 = g[3]; G12 = g[4]; G22 = g[5];
      }
      const double gr = fmad(G02, ut, fmad(G01, us, G00 * ur));
      const double gs = fmad(G12, ut, fmad(G11, us, G01 * ur));
      const double gt = fmad(G22, ut, fmad(G12, us, G02 * ur));
      gL[(size_t)e * NP + k * NQ2 + ij] = -fmad(lt, gt, fmad(ls, gs, lr * gr));
    }
    __syncthreads();
  }
}

template <typename T, int NQ, int NF, int VAR, bool FOLD>
int launch_bp5_v(hipStream_t s, const AxArgs<T>& a, T* Ap, double* part) {
  constexpr int EPB = bp5_epb<NQ>();
  constexpr int WV = bp5_waves_v<T, NQ, NF, VAR>();
  const int blocks = (a.Nelements + EPB - 1) / EPB;
  if (blocks == 0) return BPX_OK;
  const int* noPid = nullptr;
  const BatchState* noState = nullptr;
  if constexpr (std::is_same<T, float>::value && NF == 1 && (VAR == VAR_PLAIN || VAR == VAR_FUSED)) {
    if (a.tri && !a.affine) {
      constexpr int WT = bp5_waves_tri<NQ, VAR>();
      launch_stamped(k_bp5_ax<T, NQ, 1, EPB, false, WT, VAR, true, FOLD>, dim3(blocks), dim3(NQ * NQ * EPB), s,
                     a.evStart, a.evStop, a.Nelements, a.K, a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.gllw,
                     a.D, a.lambda, a.q, a.Aq, a.ldq, a.ldAq, noPid, noState, Ap, part, a.fold);
      BPX_HIP(hipGetLastError());
      return BPX_OK;
    }
  }
  if constexpr (NF == 3 && VAR == VAR_PLAIN && EPB == 1) {
    // three fields, non-affine: one workgroup per field running the
    // one-field sweep (3 waves per SIMD in FP64 at Nq = 8, where the
    // three-field sweep needs 256 VGPRs and spills at 2)
    if (!a.affine) {
      constexpr int W1 = bp5_waves_v<T, NQ, 1, VAR>();
      const unsigned fblocks = 8u * 3u * (unsigned)((a.Nelements + 7) / 8);
      launch_stamped(k_bp5_ax<T, NQ, 1, 1, false, W1, VAR, false, false, 3>, dim3(fblocks), dim3(NQ * NQ), s,
                     a.evStart, a.evStop, a.Nelements, a.K, a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.gllw,
                     a.D, a.lambda, a.q, a.Aq, a.ldq, a.ldAq, noPid, noState, Ap, part, a.fold);
      BPX_HIP(hipGetLastError());
      return BPX_OK;
    }
  }
  if (a.affine)
    launch_stamped(k_bp5_ax<T, NQ, NF, EPB, true, WV, VAR, false, FOLD>, dim3(blocks), dim3(NQ * NQ * EPB), s,
                   a.evStart, a.evStop, a.Nelements, a.K, a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.gllw,
                   a.D, a.lambda, a.q, a.Aq, a.ldq, a.ldAq, noPid, noState, Ap, part, a.fold);
  else
    launch_stamped(k_bp5_ax<T, NQ, NF, EPB, false, WV, VAR, false, FOLD>, dim3(blocks), dim3(NQ * NQ * EPB), s,
                   a.evStart, a.evStop, a.Nelements, a.K, a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.gllw,
                   a.D, a.lambda, a.q, a.Aq, a.ldq, a.ldAq, noPid, noState, Ap, part, a.fold);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <typename T, int NQ, int NF, int VAR>
int launch_bp5(hipStream_t s, const AxArgs<T>& a, T* Ap = nullptr, double* part = nullptr) {
  if constexpr (VAR == VAR_FUSED) {
    if (a.fold.on) {
      if (!a.fold.p || !a.fold.r || !a.fold.x || !a.fold.hist_rr || !a.fold.hist_alpha || !a.fold.st ||
          a.fold.p != a.q)
        return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 fused Ax: incomplete CG fold arguments");
      return launch_bp5_v<T, NQ, NF, VAR, true>(s, a, Ap, part);
    }
  }
  return launch_bp5_v<T, NQ, NF, VAR, false>(s, a, Ap, part);
}

}  // namespace

template <typename T, int NQ, int NF>
int launch_bp5_nq(hipStream_t s, const AxArgs<T>& a) {
  return launch_bp5<T, NQ, NF, VAR_PLAIN>(s, a);
}

// k_bp5_krhs_wg: W sheets (at most 16, 1024 lanes and 64 KB of LDS) with NF
// right-hand sides each, W no larger than the K/NF the element needs
template <typename T, int NQ, int NF>
constexpr int krhs_wg_sheets() {
  constexpr size_t fixed = (size_t)NQ * (NQ + 1) + 2 * 7 * NQ * NQ;
  constexpr size_t per_sheet = (size_t)3 * NF * NQ * (NQ + 1);
  constexpr int w_lds = (int)(((65536 / sizeof(T)) - fixed) / per_sheet);
  return w_lds < krhs_wg_max_sheets<NQ>() ? w_lds : krhs_wg_max_sheets<NQ>();
}
// the form is used where a sheet fits and its registers fit four waves
// per SIMD (FP64 with two right-hand sides per sheet spills at Nq >= 8)
template <typename T, int NQ, int NF>
constexpr bool krhs_wg_fits() {
  return krhs_wg_sheets<T, NQ, NF>() >= 1 && !(NF == 2 && sizeof(T) == 8 && NQ >= 8);
}
// four sheets (measured at 20^3, p = 7, FP64: 1.42x over K single launches
// at K = 32 with D in registers, against 1.31x with two sheets and 1.18x
// with eight), fewer when K needs fewer
inline int krhs_wg_default_sheets(int need) { return need <= 1 ? 1 : need <= 2 ? 2 : 4; }

template <typename T, int NQ, int NF, int W, bool DR>
int launch_bp5_krhs_wg_w(hipStream_t s, const AxArgs<T>& a) {
  constexpr int NQ2 = NQ * NQ;
  constexpr int LD = NQ + 1;
  constexpr size_t fixed = (size_t)NQ * LD + 2 * 7 * NQ2;
  constexpr size_t per_sheet = (size_t)3 * NF * NQ * LD;
  static_assert(W <= krhs_wg_sheets<T, NQ, NF>(), "k-RHS workgroup: sheets beyond the LDS or lane limit");
  const int need = (a.K + NF - 1) / NF;
  const int npass = (need + W - 1) / W;
  const long long groups = (a.Nelements + 7) / 8;
  if (groups * 8 * npass > 0x7fffffffLL) return set_error(BPX_ERR_UNSUPPORTED, "bp5 k-RHS Ax: grid too large");
  const size_t smem = (fixed + W * per_sheet) * sizeof(T);
  launch_stamped_lds(k_bp5_krhs_wg<T, NQ, NF, W, DR>, dim3((unsigned)(groups * 8 * npass)), dim3(W * NQ2), smem, s,
                     a.evStart, a.evStop, a.Nelements, a.K, npass, a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.D,
                     a.lambda, a.q, a.Aq, (long long)a.ldq, (long long)a.ldAq);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

// sheets per workgroup: a.krhs_sheets when set (1, 2, 4, 8 or 16, within
// the LDS and lane limits), else the default below
template <typename T, int NQ, int NF>
int launch_bp5_krhs_wg(hipStream_t s, const AxArgs<T>& a) {
  constexpr int w_max = krhs_wg_sheets<T, NQ, NF>();
  const int need = (a.K + NF - 1) / NF;
  int W = a.krhs_sheets;
  if (W == 0) W = krhs_wg_default_sheets(need);
  while (W > w_max || W > 16) W /= 2;
  // D in registers: FP64, one right-hand side per sheet, Nq <= 10 (default)
  constexpr bool dr_fits = sizeof(T) == 8 && NF == 1 && NQ <= 10;
  const bool dr = dr_fits && a.krhs_dreg >= 0;
  switch (W) {
    case 16:
      if constexpr (w_max >= 16) return launch_bp5_krhs_wg_w<T, NQ, NF, 16, false>(s, a);
      break;
    case 8:
      if constexpr (w_max >= 8) {
        if constexpr (dr_fits) {
          if (dr) return launch_bp5_krhs_wg_w<T, NQ, NF, 8, true>(s, a);
        }
        return launch_bp5_krhs_wg_w<T, NQ, NF, 8, false>(s, a);
      }
      break;
    case 4:
      if constexpr (w_max >= 4) {
        if constexpr (dr_fits) {
          if (dr) return launch_bp5_krhs_wg_w<T, NQ, NF, 4, true>(s, a);
        }
        return launch_bp5_krhs_wg_w<T, NQ, NF, 4, false>(s, a);
      }
      break;
    case 2:
      if constexpr (w_max >= 2) {
        if constexpr (dr_fits) {
          if (dr) return launch_bp5_krhs_wg_w<T, NQ, NF, 2, true>(s, a);
        }
        return launch_bp5_krhs_wg_w<T, NQ, NF, 2, false>(s, a);
      }
      break;
    case 1:
      if constexpr (dr_fits) {
        if (dr) return launch_bp5_krhs_wg_w<T, NQ, NF, 1, true>(s, a);
      }
      return launch_bp5_krhs_wg_w<T, NQ, NF, 1, false>(s, a);
    default:
      break;
  }
  return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 k-RHS Ax: sheets per workgroup must be 1, 2, 4, 8 or 16");
}

// the pencil form keeps the 7 factors of a pencil along k in registers:
// up to Nq = 10 (FP64: 140 of its 256 registers at two waves per SIMD)
template <int NQ>
constexpr bool krhs_pencil_fits() {
  return NQ >= 2 && NQ <= 10;
}

template <typename T, int NQ>
int launch_bp5_krhs_pencil(hipStream_t s, const AxArgs<T>& a) {
  constexpr int EPB = krhs_pencil_epb<NQ>();
  const long long groups = (a.Nelements + EPB - 1) / EPB;
  const long long blocks = (groups + 7) / 8 * 8;
  if (blocks > 0x7fffffffLL) return set_error(BPX_ERR_UNSUPPORTED, "bp5 k-RHS Ax: grid too large");
  if (a.affine)
    launch_stamped(k_bp5_krhs_pencil<T, NQ, true>, dim3((unsigned)blocks), dim3(NQ * NQ * EPB), s, a.evStart,
                   a.evStop, a.Nelements, a.K, a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.gllw, a.D, a.lambda,
                   a.q, a.Aq, (long long)a.ldq, (long long)a.ldAq);
  else
    launch_stamped(k_bp5_krhs_pencil<T, NQ, false>, dim3((unsigned)blocks), dim3(NQ * NQ * EPB), s, a.evStart,
                   a.evStop, a.Nelements, a.K, a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.gllw, a.D, a.lambda,
                   a.q, a.Aq, (long long)a.ldq, (long long)a.ldAq);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

// chunk: 0 = the default form, -2 = the pencil form, -1 = the sheet
// workgroup form (non-affine), 1, 2 or 4 = the sweep form with that many
// right-hand sides per sweep
template <typename T, int NQ>
int launch_bp5_krhs_nq(hipStream_t s, const AxArgs<T>& a, int chunk) {
  if (chunk == 0 || chunk == -2) {
    // q's rows lie below the local node count (every gather row has a local copy)
    const bool in_range = (double)a.Nelements * NQ * NQ * NQ * sizeof(T) < (double)kBufRange;
    if constexpr (krhs_pencil_fits<NQ>()) {
      if (in_range) return launch_bp5_krhs_pencil<T, NQ>(s, a);
    }
    if (chunk == -2)
      return set_error(BPX_ERR_UNSUPPORTED,
                       "bp5 k-RHS Ax: the pencil form covers N <= 9 and Nelements*Np*sizeof(dfloat) < 2^31");
  }
  // non-affine: the factors are read once per workgroup (shared by its
  // sheets); affine factors are per element, so the sweep kernel serves
  if (chunk == 0 || chunk == -1) {
    if constexpr (krhs_wg_fits<T, NQ, 1>()) {
      if (!a.affine) return launch_bp5_krhs_wg<T, NQ, 1>(s, a);
    }
    if (chunk == -1)
      return set_error(BPX_ERR_UNSUPPORTED, "bp5 k-RHS Ax: the sheet workgroup form is for non-affine meshes");
    chunk = (a.K % 2 == 0) ? 2 : 1;  // the sweep form's default
  }
  switch (chunk) {
    case 1: return launch_bp5<T, NQ, 1, VAR_KRHS>(s, a);
    case 2: return launch_bp5<T, NQ, 2, VAR_KRHS>(s, a);
    case 4: return launch_bp5<T, NQ, 4, VAR_KRHS>(s, a);
    default: return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 k-RHS Ax: chunk must be 1, 2 or 4");
  }
}

template <int NQ>
int launch_bp5_kappa_grad_nq(hipStream_t s, int Nelements, int affine, const int* G2L, const double* ggeo,
                             const double* gllw, const double* D, const double* u, const double* lam,
                             double* gL) {
  constexpr int EPB = bp5_epb<NQ>();
  const int blocks = (Nelements + EPB - 1) / EPB;
  if (blocks == 0) return BPX_OK;
  if (affine)
    hipLaunchKernelGGL((k_bp5_kappa_grad<NQ, EPB, true>), dim3(blocks), dim3(NQ * NQ * EPB), 0, s, Nelements,
                       G2L, ggeo, gllw, D, u, lam, gL);
  else
    hipLaunchKernelGGL((k_bp5_kappa_grad<NQ, EPB, false>), dim3(blocks), dim3(NQ * NQ * EPB), 0, s, Nelements,
                       G2L, ggeo, gllw, D, u, lam, gL);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

// Batch variants for the setup-time tuner: workgroup shapes that round
// identically (the default elements per workgroup, twice that, half of
// it), so the tuner's choice never changes a result bit.
template <int NQ, int V>
constexpr int batch_epb() {
  return V == 0 ? bp5_epb<NQ>() : V == 1 ? 2 * bp5_epb<NQ>() : bp5_epb<NQ>() / 2;
}

template <int NQ, int V>
int launch_batch_v(hipStream_t s, int Nelements, const int* elementList, const int* G2L, const double* wJ,
                   const double* ggeo, const double* D, double lambda, const double* q, double* Aq,
                   const int* elemPid, const BatchState* done) {
  constexpr int EPB = batch_epb<NQ, V>();
  static_assert(EPB >= 1 && NQ * NQ * EPB <= 1024, "workgroup shape");
  constexpr int WV = bp5_waves<NQ, 1, false>();
  const int blocks = (Nelements + EPB - 1) / EPB;
  if (blocks == 0) return BPX_OK;
  hipLaunchKernelGGL((k_bp5_ax<double, NQ, 1, EPB, false, WV, VAR_BATCH>), dim3(blocks), dim3(NQ * NQ * EPB), 0,
                     s, Nelements, 1, elementList, G2L, wJ, ggeo, (const double*)nullptr, D, lambda, q, Aq, 0LL,
                     0LL, elemPid, done, (double*)nullptr, (double*)nullptr, CgFold<double>());
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template <int NQ>
int launch_bp5_batch_nq(hipStream_t s, int variant, int Nelements, const int* elementList, const int* G2L,
                        const double* wJ, const double* ggeo, const double* D, double lambda, const double* q,
                        double* Aq, const int* elemPid, const BatchState* done) {
  if (variant == 0)
    return launch_batch_v<NQ, 0>(s, Nelements, elementList, G2L, wJ, ggeo, D, lambda, q, Aq, elemPid, done);
  if constexpr (NQ * NQ * 2 * bp5_epb<NQ>() <= 1024) {
    if (variant == 1)
      return launch_batch_v<NQ, 1>(s, Nelements, elementList, G2L, wJ, ggeo, D, lambda, q, Aq, elemPid, done);
  }
  if constexpr (bp5_epb<NQ>() >= 2) {
    if (variant == 2)
      return launch_batch_v<NQ, 2>(s, Nelements, elementList, G2L, wJ, ggeo, D, lambda, q, Aq, elemPid, done);
  }
  return set_error(BPX_ERR_INVALID_ARGUMENT, "batch Ax: variant not available for this degree");
}

// FP32 batch Ax (the mixed batch's inner operator): the default shape only
template <int NQ>
int launch_bp5_batch_f32_nq(hipStream_t s, int Nelements, const int* elementList, const int* G2L, const float* wJ,
                            const float* ggeo, const float* D, float lambda, const float* q, float* Aq,
                            const int* elemPid, const BatchState* done) {
  constexpr int EPB = bp5_epb<NQ>();
  constexpr int WV = bp5_waves<NQ, 1, false>();
  const int blocks = (Nelements + EPB - 1) / EPB;
  if (blocks == 0) return BPX_OK;
  hipLaunchKernelGGL((k_bp5_ax<float, NQ, 1, EPB, false, WV, VAR_BATCH>), dim3(blocks), dim3(NQ * NQ * EPB), 0, s,
                     Nelements, 1, elementList, G2L, wJ, ggeo, (const float*)nullptr, D, lambda, q, Aq, 0LL, 0LL,
                     elemPid, done, (float*)nullptr, (double*)nullptr, CgFold<float>());
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

template int launch_bp5_batch_f32_nq<BPX_NQ>(hipStream_t, int, const int*, const int*, const float*, const float*,
                                             const float*, float, const float*, float*, const int*,
                                             const BatchState*);
template int launch_bp5_batch_nq<BPX_NQ>(hipStream_t, int, int, const int*, const int*, const double*,
                                         const double*, const double*, double, const double*, double*, const int*,
                                         const BatchState*);
template int launch_bp5_kappa_grad_nq<BPX_NQ>(hipStream_t, int, int, const int*, const double*, const double*,
                                              const double*, const double*, const double*, double*);
template <typename T, int NQ>
int launch_bp5_fused_nq(hipStream_t s, const AxArgs<T>& a, T* Ap, double* part) {
  return launch_bp5<T, NQ, 1, VAR_FUSED>(s, a, Ap, part);
}

template int launch_bp5_fused_nq<double, BPX_NQ>(hipStream_t, const AxArgs<double>&, double*, double*);
template int launch_bp5_fused_nq<float, BPX_NQ>(hipStream_t, const AxArgs<float>&, float*, double*);
template int launch_bp5_nq<double, BPX_NQ, 1>(hipStream_t, const AxArgs<double>&);
template int launch_bp5_nq<double, BPX_NQ, 3>(hipStream_t, const AxArgs<double>&);
template int launch_bp5_nq<float, BPX_NQ, 1>(hipStream_t, const AxArgs<float>&);
template int launch_bp5_krhs_nq<double, BPX_NQ>(hipStream_t, const AxArgs<double>&, int);
template int launch_bp5_krhs_nq<float, BPX_NQ>(hipStream_t, const AxArgs<float>&, int);

}  // namespace bpx

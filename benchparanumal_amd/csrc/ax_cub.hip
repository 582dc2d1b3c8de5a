// BP1/BP2 (mass) and BP3/BP4 (over-integrated stiffness + lambda*mass) hex
// Ax, affine and non-affine.
//
// Semantics:
//  BP1  okl/bp1AxHex3D.okl:61-325        Aq = I^T (cubwJ .* (I q)), I = cubInterp
//  BP1A okl/bp1AxAffineHex3D.okl:61-191  Aq = J (MM x MM x MM) q
//  BP3  okl/bp3AxHex3D.okl:369-712       Aq = I^T (cubD^T G cubD + lambda cubwJ) I q
//  BP3A okl/bp3AxAffineHex3D.okl:668-924 same in the modal basis: I = invV,
//       cubD = modal derivative, per-element G and J (no weights)
//  BP2/BP4: three fields, local layout [element][field][node].
//
// Design (gfx950): one element per workgroup, an MQ x MQ sheet of lanes
// (MQ = points per direction of the quadrature grid: p+2 Gauss points, or
// Nq for the affine modal/mass forms). Lane (a,b) keeps its pencil along the
// third direction in registers. The tensor interpolation runs direction by
// direction: the pencil direction in registers, the other two slab by slab
// through two LDS slabs. BP3 then runs the slab-swept gradient/weak
// divergence on the quadrature grid (the same scheme as the BP5 kernel),
// and the projection back mirrors the interpolation. Only MQ x (MQ+1)
// slabs live in LDS, so p = 15 fits.
#include <hip/hip_runtime.h>

#include "ax_common.hpp"
#include "buffer_rsrc.hpp"

#ifndef BPX_NQ
#error "compile with -DBPX_NQ=<Nq>"
#endif

namespace bpx {
namespace {

enum { MODE_BP1 = 0, MODE_BP1A = 1, MODE_BP3 = 2, MODE_BP3A = 3 };

// The FP64 BP1 kernel on the 9-point grid spills some 56 registers at 2
// waves per SIMD; at 1 it keeps everything in registers (measured 175
// against 182 us at 24^3, p=7). BP3 is faster at 2 despite its spills.
template <typename T, int MQ, int NF, int MODE>
constexpr int cub_waves() {
  if (sizeof(T) == 8 && MODE == MODE_BP1 && NF == 1 && MQ == 9) return 1;
  return NF == 1 ? (MQ <= 9 ? 2 : 1) : 1;
}

// elements per workgroup: an MQ x MQ sheet of lanes each, ~256 lanes (at
// MQ = 9 one sheet is 81 lanes, which would leave 47 of the second wave's
// 64 lanes idle; three sheets fill 243 of 256)
template <int MQ>
constexpr int cub_epb() {
  return MQ * MQ >= 128 ? 1 : 256 / (MQ * MQ);
}

// The FP32 BP3 kernel (the mixed-precision solve's operator) loads its
// geometric factors non-temporally, so the CG vectors and gather maps stay
// in the Infinity Cache across the iteration (measured at 24^3, p=7: mixed
// BP3 solve 209.8 -> 200.8 ms). The FP64 BP3 solve and both BP1 solves ran
// slower or within 2.5% that way, so they keep cached loads.
template <bool STREAM, typename T>
__device__ inline T geo_load(const T* p) {
  if constexpr (STREAM) return __builtin_nontemporal_load(p);
  return *p;
}

template <typename T, int NQ, int MQ, int NF, int MODE, int WAVES, int EPB>
__global__ __launch_bounds__(MQ* MQ* EPB) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_cub_ax(
    int Nelements, const int* __restrict__ elementList, const int* __restrict__ G2L,
    const T* __restrict__ W, const T* __restrict__ G, const T* __restrict__ Dq,
    const T* __restrict__ I, T lambda, const T* __restrict__ q, T* __restrict__ Aq) {
  constexpr int NP = NQ * NQ * NQ;
  constexpr int MP = MQ * MQ * MQ;
  constexpr int LD = MQ + 1;
  constexpr bool STIFF = (MODE == MODE_BP3 || MODE == MODE_BP3A);
  constexpr bool PROJECT = (MODE != MODE_BP1A);
  constexpr bool SG = (MODE == MODE_BP3) && sizeof(T) == 4;  // stream the geometry

  __shared__ T s_I[MQ][NQ + 1];
  __shared__ T s_D[STIFF ? MQ : 1][LD];
  __shared__ T s_a[NF][EPB][MQ][LD];
  __shared__ T s_b[NF][EPB][MQ][LD];
  __shared__ T s_c[STIFF ? NF : 1][EPB][MQ][LD];

  const int t = threadIdx.x;
  const int es = t / (MQ * MQ);
  const int ab = t - es * MQ * MQ;
  const int a = ab % MQ;
  const int b = ab / MQ;
  const int eo = blockIdx.x * EPB + es;
  const bool active = eo < Nelements;
  const int e = active ? elementList[eo] : 0;  // an idle sheet computes element 0 and stores nothing
  const bool in_nodes = active && (a < NQ) && (b < NQ);

  for (int n = t; n < MQ * NQ; n += MQ * MQ * EPB) {
    const int g = n / NQ, c = n % NQ;
    // BP1A applies MM^T as okl/bp1AxAffineHex3D.okl does (s_MM[m][i])
    s_I[g][c] = (MODE == MODE_BP1A) ? I[c * NQ + g] : I[g * NQ + c];
  }
  if (STIFF)
    for (int n = t; n < MQ * MQ; n += MQ * MQ * EPB) s_D[n / MQ][n % MQ] = Dq[n];

  T u[NF][MQ];
  {
    // loads from valid addresses without branches (lanes off the node grid
    // read node (min(a), min(b)); masked nodes read W[0], never empty, and
    // store 0): a load under a branch makes the compiler wait out each
    // G2L -> q round trip in turn
    T r[NF][NQ];
    int id[NF][NQ];
    const size_t nb = (size_t)e * NP * NF + min(b, NQ - 1) * NQ + min(a, NQ - 1);
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int c = 0; c < NQ; ++c) id[f][c] = in_nodes ? G2L[nb + f * NP + c * NQ * NQ] : -1;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int c = 0; c < NQ; ++c) {
        const T v = *(id[f][c] >= 0 ? q + id[f][c] : W);
        r[f][c] = (id[f][c] >= 0) ? v : T(0);
      }
    __syncthreads();
    // interpolate along c in registers
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int g = 0; g < MQ; ++g) {
        T v = 0;
#pragma unroll
        for (int c = 0; c < NQ; ++c) v += s_I[g][c] * r[f][c];
        u[f][g] = v;
      }
  }

  // interpolate along b, then a, slab by slab
#pragma unroll
  for (int g = 0; g < MQ; ++g) {
#pragma unroll
    for (int f = 0; f < NF; ++f) asm volatile("" : "+v"(u[f][g]));
    if (in_nodes)
#pragma unroll
      for (int f = 0; f < NF; ++f) s_a[f][es][b][a] = u[f][g];
    __syncthreads();
    asm volatile("" ::: "memory");
    if (a < NQ)
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        T v = 0;
#pragma unroll
        for (int m = 0; m < NQ; ++m) v += s_I[b][m] * s_a[f][es][m][a];
        s_b[f][es][b][a] = v;
      }
    __syncthreads();
    asm volatile("" ::: "memory");
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      T v = 0;
#pragma unroll
      for (int m = 0; m < NQ; ++m) v += s_I[a][m] * s_b[f][es][b][m];
      u[f][g] = v;
    }
  }

  if (MODE == MODE_BP1) {
    const T* Wp = W + (size_t)e * MP + b * MQ + a;
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int g = 0; g < MQ; ++g) u[f][g] *= Wp[g * MQ * MQ];
  } else if (MODE == MODE_BP1A) {
    const T J = W[e];
    if (active)
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int g = 0; g < NQ; ++g)
          Aq[(size_t)e * NP * NF + f * NP + g * NQ * NQ + b * NQ + a] = J * u[f][g];
    return;
  } else {
    // gradient, G, weak divergence on the quadrature grid, g-slab by g-slab
    T res[NF][MQ];
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int g = 0; g < MQ; ++g) res[f][g] = 0;
    // BP3A: one set per element; BP3: slab g's factors, slab g+1's loaded
    // while slab g computes
    T e0 = 0, e1 = 0, e2 = 0, e3 = 0, e4 = 0, e5 = 0, eW = 0;
    if (MODE == MODE_BP3A) {
      const T* ge = G + (size_t)6 * e;
      e0 = ge[0]; e1 = ge[1]; e2 = ge[2]; e3 = ge[3]; e4 = ge[4]; e5 = ge[5];
      eW = W[e];
    } else {
      const size_t cb = (size_t)e * MP + b * MQ + a;
      const T* gp = G + 6 * cb;
      e0 = geo_load<SG>(gp + 0); e1 = geo_load<SG>(gp + 1); e2 = geo_load<SG>(gp + 2);
      e3 = geo_load<SG>(gp + 3); e4 = geo_load<SG>(gp + 4); e5 = geo_load<SG>(gp + 5);
      eW = geo_load<SG>(W + cb);
    }
#pragma unroll
    for (int g = 0; g < MQ; ++g) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int m = 0; m < MQ; ++m) asm volatile("" : "+v"(res[f][m]));
      const T G00 = e0, G01 = e1, G02 = e2, G11 = e3, G12 = e4, G22 = e5, WW = eW;
      if (MODE == MODE_BP3 && g + 1 < MQ) {
        const size_t cb = (size_t)e * MP + (g + 1) * MQ * MQ + b * MQ + a;
        const T* gp = G + 6 * cb;
        e0 = geo_load<SG>(gp + 0); e1 = geo_load<SG>(gp + 1); e2 = geo_load<SG>(gp + 2);
        e3 = geo_load<SG>(gp + 3); e4 = geo_load<SG>(gp + 4); e5 = geo_load<SG>(gp + 5);
        eW = geo_load<SG>(W + cb);
      }
#pragma unroll
      for (int f = 0; f < NF; ++f) s_a[f][es][b][a] = u[f][g];
      __syncthreads();
      asm volatile("" ::: "memory");
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        T ur = 0, us = 0, ut = 0;
#pragma unroll
        for (int m = 0; m < MQ; ++m) {
          ur += s_D[a][m] * s_a[f][es][b][m];
          us += s_D[b][m] * s_a[f][es][m][a];
          ut += s_D[g][m] * u[f][m];
        }
        const T Gr = G00 * ur + G01 * us + G02 * ut;
        const T Gs = G01 * ur + G11 * us + G12 * ut;
        const T Gt = G02 * ur + G12 * us + G22 * ut;
        s_b[f][es][b][a] = Gr;
        s_c[f][es][b][a] = Gs;
        res[f][g] += lambda * WW * u[f][g];
#pragma unroll
        for (int m = 0; m < MQ; ++m) res[f][m] += s_D[g][m] * Gt;
      }
      __syncthreads();
      asm volatile("" ::: "memory");
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        T acc = 0;
#pragma unroll
        for (int m = 0; m < MQ; ++m) {
          acc += s_D[m][a] * s_b[f][es][b][m];
          acc += s_D[m][b] * s_c[f][es][m][a];
        }
        res[f][g] += acc;
      }
    }
#pragma unroll
    for (int f = 0; f < NF; ++f)
#pragma unroll
      for (int g = 0; g < MQ; ++g) u[f][g] = res[f][g];
  }

  if (PROJECT) {
    // project back along a, then b, slab by slab
#pragma unroll
    for (int g = 0; g < MQ; ++g) {
#pragma unroll
      for (int f = 0; f < NF; ++f) asm volatile("" : "+v"(u[f][g]));
      __syncthreads();
#pragma unroll
      for (int f = 0; f < NF; ++f) s_a[f][es][b][a] = u[f][g];
      __syncthreads();
      asm volatile("" ::: "memory");
      if (a < NQ)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          T v = 0;
#pragma unroll
          for (int m = 0; m < MQ; ++m) v += s_I[m][a] * s_a[f][es][b][m];
          s_b[f][es][b][a] = v;
        }
      __syncthreads();
      asm volatile("" ::: "memory");
      if (in_nodes)
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          T v = 0;
#pragma unroll
          for (int m = 0; m < MQ; ++m) v += s_I[m][b] * s_b[f][es][m][a];
          u[f][g] = v;
        }
    }
    // project back along c in registers
    if (in_nodes) {
#pragma unroll
      for (int f = 0; f < NF; ++f)
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
          T v = 0;
#pragma unroll
          for (int g = 0; g < MQ; ++g) v += s_I[g][c] * u[f][g];
          Aq[(size_t)e * NP * NF + f * NP + c * NQ * NQ + b * NQ + a] = v;
        }
    }
  }
}

// BP1/BP2 non-affine (Aq = I^T (cubwJ .* (I q)), okl/bp1AxHex3D.okl:61-325)
// for NQ <= 8, one wavefront per element. The element passes through two
// LDS tiles, x[k][j][a] (after interpolating along i) and y[k][b][a] (after
// j); each 1D pass gives every lane whole rows or columns, with their NQ or
// MQ inputs in registers:
//   rows (k,j):    x = I q along i          (the q gather reads these rows)
//   columns (k,a): y = I x along j
//   columns (b,a): along k, times cubwJ, and back along k, in registers
//   columns (k,a): x = I^T y along j
//   rows (k,j):    Aq = I^T x along i       (written to the local vector)
// I (MQ x NQ) is uniform over the wavefront and read through scalar loads.
// A lane holds at most 2 of the MQ^2 cubature columns, whose cubwJ values
// are requested at the start. The tiles take 9.8 KB at NQ = 8 (16
// wavefronts per CU). This replaces the sheet kernel, whose FP64 form
// spilled at 2 waves/SIMD and held 3 elements per CU at 1.
// FPAR (NF > 1): one workgroup per field instead of the fields one after
// another; the NF workgroups of an element are consecutive on one XCD
// (wave_element), so the second and third read its factors from that L2.
// The same operations per field either way.
template <int NF, bool FPAR>
__device__ inline int wave_element(int Nelements, int& f0, int& f1) {
  f0 = 0;
  f1 = NF;
  if (!FPAR) return blockIdx.x;
  const int b = blockIdx.x, r = b >> 3;
  f0 = r % NF;
  f1 = f0 + 1;
  const int eo = (r / NF) * 8 + (b & 7);
  return eo < Nelements ? eo : -1;
}

template <int NF, bool FPAR>
inline unsigned wave_blocks(int Nelements) {
  return FPAR ? (unsigned)(8 * NF * ((Nelements + 7) / 8)) : (unsigned)Nelements;
}

template <typename T, int NQ, int NF, bool FPAR>
__global__ __launch_bounds__(64) void k_bp1_wave(int Nelements, const int* __restrict__ elementList,
                                                 const int* __restrict__ G2L, const T* __restrict__ W,
                                                 const T* __restrict__ I, const T* __restrict__ q,
                                                 T* __restrict__ Aq) {
  constexpr int MQ = NQ + 1;
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  constexpr int MQ2 = MQ * MQ;
  constexpr int MP = MQ2 * MQ;
  constexpr int LX = MQ | 1;  // odd row stride of x: conflict-free row reads
  constexpr int C3 = (MQ2 + 63) / 64;  // cubature columns per lane
  static_assert(NQ2 <= 64 && NQ * MQ <= 128 && C3 <= 2, "one or two rows/columns per lane");
  __shared__ T s_x[NQ * NQ * LX];
  __shared__ T s_y[NQ * MQ * MQ];
  typedef const __attribute__((address_space(4))) T* cptr;
  const cptr Ic = (cptr)I;  // I[g][c] = Ic[g * NQ + c]
  // entries of I are read with scalar loads where they are used: an opaque
  // update of the pointer before each output keeps the compiler from holding
  // all MQ * NQ of them in scalar registers (144 at NQ = 8), which spilled
#define BPX_IROW() \
  cptr Ir = Ic;    \
  asm volatile("" : "+s"(Ir))
  const int l = threadIdx.x;
  int f0, f1;
  const int eo = wave_element<NF, FPAR>(Nelements, f0, f1);
  if (eo < 0) return;  // uniform over the workgroup
  const int e = elementList[eo];

  // cubwJ of the lane's cubature columns (b, a) = col, all g
  T wv[C3][MQ];
#pragma unroll
  for (int c = 0; c < C3; ++c) {
    const int col = min(l + 64 * c, MQ2 - 1);
#pragma unroll
    for (int g = 0; g < MQ; ++g) wv[c][g] = W[(size_t)e * MP + g * MQ2 + col];
  }

  for (int f = f0; f < f1; ++f) {
    const size_t lb = (size_t)e * NP * NF + (size_t)f * NP;
    // rows (k, j) = row: gather q, interpolate along i
    if (l < NQ2) {
      int id[NQ];
#pragma unroll
      for (int i = 0; i < NQ; ++i) id[i] = G2L[lb + l * NQ + i];
      T v[NQ];
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        const T t = *(id[i] >= 0 ? q + id[i] : W);  // masked: a valid address, value 0
        v[i] = id[i] >= 0 ? t : T(0);
      }
#pragma unroll
      for (int a = 0; a < MQ; ++a) {
        BPX_IROW();
        T acc = 0;
#pragma unroll
        for (int i = 0; i < NQ; ++i) acc += Ir[a * NQ + i] * v[i];
        s_x[l * LX + a] = acc;
      }
    }
    __syncthreads();
    // columns (k, a): along j
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int col = l + 64 * c;
      if (col < NQ * MQ) {
        const int k = col / MQ, a = col - k * MQ;
        T v[NQ];
#pragma unroll
        for (int j = 0; j < NQ; ++j) v[j] = s_x[(k * NQ + j) * LX + a];
#pragma unroll
        for (int b = 0; b < MQ; ++b) {
          BPX_IROW();
          T acc = 0;
#pragma unroll
          for (int j = 0; j < NQ; ++j) acc += Ir[b * NQ + j] * v[j];
          s_y[(k * MQ + b) * MQ + a] = acc;
        }
      }
    }
    __syncthreads();
    // columns (b, a): along k, times cubwJ, back along k (in place)
#pragma unroll
    for (int c = 0; c < C3; ++c) {
      const int col = l + 64 * c;
      if (col < MQ2) {
        T v[NQ];
#pragma unroll
        for (int k = 0; k < NQ; ++k) v[k] = s_y[k * MQ2 + col];
        T z[MQ];
#pragma unroll
        for (int g = 0; g < MQ; ++g) {
          BPX_IROW();
          T acc = 0;
#pragma unroll
          for (int k = 0; k < NQ; ++k) acc += Ir[g * NQ + k] * v[k];
          z[g] = acc * wv[c][g];
        }
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
          BPX_IROW();
          T acc = 0;
#pragma unroll
          for (int g = 0; g < MQ; ++g) acc += Ir[g * NQ + k] * z[g];
          s_y[k * MQ2 + col] = acc;
        }
      }
    }
    __syncthreads();
    // columns (k, a): back along j
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int col = l + 64 * c;
      if (col < NQ * MQ) {
        const int k = col / MQ, a = col - k * MQ;
        T v[MQ];
#pragma unroll
        for (int b = 0; b < MQ; ++b) v[b] = s_y[(k * MQ + b) * MQ + a];
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
          BPX_IROW();
          T acc = 0;
#pragma unroll
          for (int b = 0; b < MQ; ++b) acc += Ir[b * NQ + j] * v[b];
          s_x[(k * NQ + j) * LX + a] = acc;
        }
      }
    }
    __syncthreads();
    // rows (k, j): back along i, to the local vector
    if (l < NQ2) {
      T v[MQ];
#pragma unroll
      for (int a = 0; a < MQ; ++a) v[a] = s_x[l * LX + a];
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        BPX_IROW();
        T acc = 0;
#pragma unroll
        for (int a = 0; a < MQ; ++a) acc += Ir[a * NQ + i] * v[a];
        Aq[lb + l * NQ + i] = acc;
      }
    }
    __syncthreads();  // the next field rewrites the tiles
  }
#undef BPX_IROW
}

// BP3/BP4 non-affine (Aq = I^T (cubD^T G cubD + lambda cubwJ) I q,
// okl/bp3AxHex3D.okl:369-712) for NQ <= 8, two wavefronts per element (one
// row or column per lane: 64 wide, the MQ^2 = 81 columns fitted in one pass
// of 128 lanes instead of two of 64, the second 17 wide), the BP1 scheme
// with the stiffness on the cubature grid in between. Three LDS
// tiles of MQ^3: tu holds y[k][b][a] and then u[g][b][a]; tr holds the
// r-fluxes Gr (and x[k][j][a] before and after); ts the s-fluxes Gs.
//   rows (k,j):      x = I q along i                      -> tr
//   columns (k,a):   y = I x along j                      -> tu
//   columns (b,a):   u = I y along k, in registers and    -> tu
//   columns (b,a):   ur, us from tu, ut in registers; G;
//                    Gr -> tr, Gs -> ts; res = lambda W u + D^T Gt (registers)
//   columns (b,a):   res += D^T Gr (from tr) + D^T Gs (from ts); back along k -> tu
//   columns (k,a):   back along j                         -> tr
//   rows (k,j):      back along i                         -> Aq
// I and the t-direction D rows are uniform over the wavefront (scalar
// loads); the lane's own D rows and columns come from global memory.
template <typename T, int NQ, int NF, bool FPAR>
__global__ __launch_bounds__(128) void k_bp3_wave(int Nelements, const int* __restrict__ elementList,
                                                 const int* __restrict__ G2L, const T* __restrict__ W,
                                                 const T* __restrict__ G, const T* __restrict__ Dq,
                                                 const T* __restrict__ I, T lambda, const T* __restrict__ q,
                                                 T* __restrict__ Aq) {
  constexpr int MQ = NQ + 1;
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  constexpr int MQ2 = MQ * MQ;
  constexpr int MP = MQ2 * MQ;
  constexpr int LX = MQ | 1;
  static_assert(NQ2 <= 128 && NQ * MQ <= 128 && MQ2 <= 128, "one row/column per lane");
  static_assert(NQ * NQ * LX <= MP, "x fits in an MQ^3 tile");
  __shared__ T tu[MP], tr[MP], ts[MP];
  typedef const __attribute__((address_space(4))) T* cptr;
  const cptr Ic = (cptr)I;   // I[g][c] = Ic[g * NQ + c]
  const cptr Dc = (cptr)Dq;  // D[r][c] = Dc[r * MQ + c]
  const int l = threadIdx.x;
  int f0, f1;
  const int eo = wave_element<NF, FPAR>(Nelements, f0, f1);
  if (eo < 0) return;  // uniform over the workgroup
  const int e = elementList[eo];

  for (int f = f0; f < f1; ++f) {
    const size_t lb = (size_t)e * NP * NF + (size_t)f * NP;
    // rows (k, j): gather q, along i -> x in tr
    if (l < NQ2) {
      int id[NQ];
#pragma unroll
      for (int i = 0; i < NQ; ++i) id[i] = G2L[lb + l * NQ + i];
      T v[NQ];
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        const T t = *(id[i] >= 0 ? q + id[i] : W);  // masked: a valid address, value 0
        v[i] = id[i] >= 0 ? t : T(0);
      }
#pragma unroll
      for (int a = 0; a < MQ; ++a) {
        T acc = 0;
#pragma unroll
        for (int i = 0; i < NQ; ++i) acc += Ic[a * NQ + i] * v[i];
        tr[l * LX + a] = acc;
      }
    }
    __syncthreads();
    // columns (k, a): along j -> y in tu
    {
      constexpr int c = 0;
      const int col = l;
      if (col < NQ * MQ) {
        const int k = col / MQ, a = col - k * MQ;
        T v[NQ];
#pragma unroll
        for (int j = 0; j < NQ; ++j) v[j] = tr[(k * NQ + j) * LX + a];
#pragma unroll
        for (int b = 0; b < MQ; ++b) {
          T acc = 0;
#pragma unroll
          for (int j = 0; j < NQ; ++j) acc += Ic[b * NQ + j] * v[j];
          tu[k * MQ2 + b * MQ + a] = acc;
        }
      }
    }
    __syncthreads();
    // columns (b, a): along k -> u, kept and written in place
    T u[1][MQ];  // [column of the lane]
    {
      constexpr int c = 0;
      const int col = l;
      if (col < MQ2) {
        T v[NQ];
#pragma unroll
        for (int k = 0; k < NQ; ++k) v[k] = tu[k * MQ2 + col];
#pragma unroll
        for (int g = 0; g < MQ; ++g) {
          T acc = 0;
#pragma unroll
          for (int k = 0; k < NQ; ++k) acc += Ic[g * NQ + k] * v[k];
          u[c][g] = acc;
        }
      }
    }
    __syncthreads();  // every column has read its y before u overwrites it
    {
      constexpr int c = 0;
      const int col = l;
      if (col < MQ2)
#pragma unroll
        for (int g = 0; g < MQ; ++g) tu[g * MQ2 + col] = u[c][g];
    }
    __syncthreads();
    // columns (b, a): gradient, G, fluxes; res = lambda W u + D^T Gt
    T res[1][MQ];  // [column of the lane]
    {
      constexpr int c = 0;
      const int col = l;
      if (col < MQ2) {
        const int b = col / MQ, a = col - b * MQ;
        T Da[MQ], Db[MQ];
#pragma unroll
        for (int m = 0; m < MQ; ++m) {
          Da[m] = Dq[a * MQ + m];
          Db[m] = Dq[b * MQ + m];
        }
        T Gt[MQ];
#pragma unroll
        for (int g = 0; g < MQ; ++g) {
          T ur = 0, us = 0, ut = 0;
#pragma unroll
          for (int m = 0; m < MQ; ++m) {
            ur += Da[m] * tu[g * MQ2 + b * MQ + m];
            us += Db[m] * tu[g * MQ2 + m * MQ + a];
            ut += Dc[g * MQ + m] * u[c][m];
          }
          // FP32 (the mixed solve's operator): non-temporal, as in k_cub_ax
          constexpr bool SG = sizeof(T) == 4;
          const size_t cb = (size_t)e * MP + g * MQ2 + col;
          const T* gp = G + 6 * cb;
          const T G00 = geo_load<SG>(gp + 0), G01 = geo_load<SG>(gp + 1), G02 = geo_load<SG>(gp + 2),
                  G11 = geo_load<SG>(gp + 3), G12 = geo_load<SG>(gp + 4), G22 = geo_load<SG>(gp + 5);
          const T WW = geo_load<SG>(W + cb);
          tr[g * MQ2 + col] = G00 * ur + G01 * us + G02 * ut;
          ts[g * MQ2 + col] = G01 * ur + G11 * us + G12 * ut;
          Gt[g] = G02 * ur + G12 * us + G22 * ut;
          res[c][g] = lambda * WW * u[c][g];
        }
#pragma unroll
        for (int g = 0; g < MQ; ++g) {
          T acc = 0;
#pragma unroll
          for (int m = 0; m < MQ; ++m) acc += Dc[m * MQ + g] * Gt[m];
          res[c][g] += acc;
        }
      }
    }
    __syncthreads();
    // columns (b, a): r and s parts of the weak divergence; back along k
    {
      constexpr int c = 0;
      const int col = l;
      if (col < MQ2) {
        const int b = col / MQ, a = col - b * MQ;
        T DTa[MQ], DTb[MQ];
#pragma unroll
        for (int m = 0; m < MQ; ++m) {
          DTa[m] = Dq[m * MQ + a];
          DTb[m] = Dq[m * MQ + b];
        }
#pragma unroll
        for (int g = 0; g < MQ; ++g) {
          T acc = 0;
#pragma unroll
          for (int m = 0; m < MQ; ++m) {
            acc += DTa[m] * tr[g * MQ2 + b * MQ + m];
            acc += DTb[m] * ts[g * MQ2 + m * MQ + a];
          }
          res[c][g] += acc;
        }
#pragma unroll
        for (int k = 0; k < NQ; ++k) {
          T acc = 0;
#pragma unroll
          for (int g = 0; g < MQ; ++g) acc += Ic[g * NQ + k] * res[c][g];
          tu[k * MQ2 + col] = acc;
        }
      }
    }
    __syncthreads();
    // columns (k, a): back along j -> x in tr
    {
      constexpr int c = 0;
      const int col = l;
      if (col < NQ * MQ) {
        const int k = col / MQ, a = col - k * MQ;
        T v[MQ];
#pragma unroll
        for (int b = 0; b < MQ; ++b) v[b] = tu[k * MQ2 + b * MQ + a];
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
          T acc = 0;
#pragma unroll
          for (int b = 0; b < MQ; ++b) acc += Ic[b * NQ + j] * v[b];
          tr[(k * NQ + j) * LX + a] = acc;
        }
      }
    }
    __syncthreads();
    // rows (k, j): back along i -> Aq
    if (l < NQ2) {
      T v[MQ];
#pragma unroll
      for (int a = 0; a < MQ; ++a) v[a] = tr[l * LX + a];
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        T acc = 0;
#pragma unroll
        for (int a = 0; a < MQ; ++a) acc += Ic[a * NQ + i] * v[a];
        Aq[lb + l * NQ + i] = acc;
      }
    }
    __syncthreads();  // the next field rewrites the tiles
  }
}

// BP1-BP4 non-affine, pencil form (NQ <= 8; the default there). A slot is
// one field of one element; a 256-lane workgroup holds S = 256 / MQ^2 slots
// (three at p = 7: the three fields of a BP2/BP4 element, or three BP1/BP3
// elements), so the 64-, 72- and 81-line passes of three slots fill 192,
// 216 and 243 lanes instead of taking two passes of one wavefront each.
// Every one-dimensional contraction runs along a line (pencil) that one
// lane holds in registers, with the 1D matrix uniform over the workgroup
// (scalar loads, one row in use and the next in flight); values move
// between line directions through LDS once per move. LDS tiles per slot
// with line stride PR = MQ | 1 (odd strides keep the strided accesses free
// of bank conflicts) and slab stride PS = MQ * PR: A and B for BP3/BP4; one
// tile for BP1/BP2 (A and B below name the same one; a pass that rewrites
// it holds its outputs in registers across a barrier):
//   i-lines (k,j):  gather q, x = I q along i                 -> A
//   j-lines (k,a):  y = I x along j                           -> B
//   t-lines (b,a):  u = I y along k, in registers
//     BP1/BP2:      times cubwJ, back along k (in place)      -> B
//     BP3/BP4:      u                                         -> A
//   r-lines (g,b):  ur = D u along a                          -> B
//   t-lines (b,a):  ur into registers
//   s-lines (g,a):  us = D u along b                          -> B
//   t-lines (b,a):  ut = D u along g (registers); G: Gr -> A, Gs -> B;
//                   res = lambda W u + D^T Gt (registers)
//   r-lines, s-lines: D^T Gr (A), D^T Gs (B), in place
//   t-lines (b,a):  res += both; back along g                 -> A
//   j-lines (k,a):  back along j: z (A; B for BP1/BP2)        -> x (B; A)
//   i-lines (k,j):  back along i                              -> Aq
// A node crosses LDS 4 (BP1) or 16 (BP3) times, against 9 values read per
// output of a gradient or divergence in the wavefront forms. The geometric
// factors of the first PF cubature slabs are requested when u is formed,
// so they are in flight through the r and s derivatives. With the host's
// even-odd tables (EO) each contraction runs in even-odd form; the r, s and
// t parts of the divergence are summed apart. Results agree with the
// wavefront forms to rounding, not bit for bit (the row-by-row BP1/BP2
// form adds in their order). q is gathered through a bounds-checked buffer
// resource: a masked node (G2L = -1) or an idle slot reads 0 without a
// compare or select per load; BP2 gathers and stores in node order.
template <typename T>
__device__ inline T cfma(T a, T b, T c) {
  if constexpr (sizeof(T) == 8)
    return __builtin_fma(a, b, c);
  else
    return __builtin_fmaf(a, b, c);
}

// rows of a uniform R x C row-major matrix through scalar registers, row
// n + 1 requested while row n is in use: body(n, row) for n = 0..R-1. The
// opaque update of the pointer before each row keeps the compiler from
// holding the whole matrix in scalar registers.
template <int R, int C, typename T, typename F>
__device__ inline void uniform_rows(const T* M, F&& body) {
  typedef const __attribute__((address_space(4))) T* cptr;
  cptr Mc = (cptr)M;
  T d[2][C];
  asm volatile("" : "+s"(Mc));
#pragma unroll
  for (int m = 0; m < C; ++m) d[0][m] = Mc[m];
#pragma unroll
  for (int n = 0; n < R; ++n) {
    if (n + 1 < R) {
      __builtin_amdgcn_sched_barrier(0);
      asm volatile("" : "+s"(Mc));
#pragma unroll
      for (int m = 0; m < C; ++m) d[(n + 1) & 1][m] = Mc[(n + 1) * C + m];
    }
    body(n, d[n & 1]);
  }
}

// Even-odd form of y = A x for a 1D operator A (N x M) with A[N-1-n][M-1-m]
// = SIGN * A[n][m] (the interpolation matrix between symmetric point sets
// and its transpose: SIGN = 1; the derivative on symmetric points and its
// transpose: SIGN = -1). With xe[m] = x[m] + x[M-1-m], xo[m] = x[m] -
// x[M-1-m] (m < M/2) and Ae = (A[n][m] + A[n][M-1-m]) / 2, Ao = (A[n][m] -
// A[n][M-1-m]) / 2: ye[n] = Ae xe (+ A[n][M/2] x[M/2] for odd M), yo[n] =
// Ao xo, y[n] = ye + yo and y[N-1-n] = SIGN (ye - yo), which halves the
// multiply-adds (Kopriva's even-odd decomposition). Table row n (n <=
// (N-1)/2): Ae[n][0..M/2-1], A[n][M/2] (odd M), Ao[n][0..M/2-1]; in the
// middle row of an odd N only the part that does not vanish is used.
template <int N, int M>
struct EoShape {
  static constexpr int hM = M / 2, oM = M & 1, LE = hM + oM, RL = LE + hM;
  static constexpr int hN = N / 2, oN = N & 1, rows = hN + oN, size = rows * RL;
};

template <int N, int M, int SIGN, typename T, typename P>
__device__ inline void eo_apply(const T* tab, const T (&x)[M], P&& put) {  // put(n, y[n])
  using E = EoShape<N, M>;
  T xe[E::hM > 0 ? E::hM : 1], xo[E::hM > 0 ? E::hM : 1];
#pragma unroll
  for (int m = 0; m < E::hM; ++m) {
    xe[m] = x[m] + x[M - 1 - m];
    xo[m] = x[m] - x[M - 1 - m];
  }
  uniform_rows<E::rows, E::RL>(tab, [&](int n, const T* r) {
    const bool mid = E::oN && n == E::hN;
    T ye = 0, yo = 0;
    if (!mid || SIGN > 0) {
#pragma unroll
      for (int m = 0; m < E::hM; ++m) ye = cfma(r[m], xe[m], ye);
      if constexpr (E::oM) ye = cfma(r[E::hM], x[E::hM], ye);
    }
    if (!mid || SIGN < 0) {
#pragma unroll
      for (int m = 0; m < E::hM; ++m) yo = cfma(r[E::LE + m], xo[m], yo);
    }
    if (mid) {
      put(n, SIGN > 0 ? ye : yo);
    } else {
      put(n, ye + yo);
      put(N - 1 - n, SIGN > 0 ? ye - yo : yo - ye);
    }
  });
}

// offsets of the four tables in the kernel's even-odd array: I (MQ x NQ),
// I^T, D (MQ x MQ, the cubature derivative) and D^T
template <int NQ>
struct EoTables {
  static constexpr int MQ = NQ + 1;
  static constexpr int I = 0, IT = I + EoShape<MQ, NQ>::size, D = IT + EoShape<NQ, MQ>::size,
                       DT = D + EoShape<MQ, MQ>::size, size = DT + EoShape<MQ, MQ>::size;
};

template <int MQ>
constexpr int pencil_slots() {
  return MQ * MQ >= 256 ? 1 : 256 / (MQ * MQ);
}

// LDS tiles per slot (values). BP3/BP4: A holds x (NQ^2 lines of PR), u,
// Gr and z (MQ slabs), B holds y, ur, us and Gs. BP1/BP2: one tile holds
// x, y, z and x in turn (NQ^2 lines or NQ slabs), the values of a pass
// that rewrites it staged in registers across a barrier.
template <int MQ, bool STIFF>
constexpr int pencil_tile_a() {
  constexpr int x = (MQ - 1) * (MQ - 1) * (MQ | 1), y = (MQ - 1) * MQ * (MQ | 1);
  return STIFF ? MQ * MQ * (MQ | 1) : (x > y ? x : y);
}
template <int MQ, bool STIFF>
constexpr int pencil_tile_b() {
  return STIFF ? MQ * MQ * (MQ | 1) : 0;
}

template <typename T, int MQ, bool STIFF>
constexpr int pencil_lds_bytes() {
  return pencil_slots<MQ>() * (pencil_tile_a<MQ, STIFF>() + pencil_tile_b<MQ, STIFF>()) * (int)sizeof(T);
}

// waves per SIMD that the 160 KB of LDS per CU allows (a workgroup is one
// wavefront on each of the 4 SIMDs), at most 8; BP3/BP4 at most 4 (FP64)
// or 5 (FP32), above which they spill
template <typename T, int MQ, bool STIFF>
constexpr int pencil_waves() {
  constexpr int w = (160 * 1024) / pencil_lds_bytes<T, MQ, STIFF>();
  constexpr int cap = STIFF ? (sizeof(T) == 8 ? 4 : 5) : 8;
  return w < 1 ? 1 : (w > cap ? cap : w);
}

// cubature slabs whose geometric factors are requested ahead (BP3/BP4)
template <typename T, int MQ>
constexpr int pencil_prefetch() {
  return sizeof(T) == 8 ? (MQ < 4 ? MQ : 4) : (MQ < 6 ? MQ : 6);
}

template <typename T, int NQ, int NF, bool STIFF, bool EO>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(pencil_waves<T, NQ + 1, STIFF>()))) void
k_cub_pencil(int Nelements, const int* __restrict__ elementList, const int* __restrict__ G2L,
             const T* __restrict__ W, const T* __restrict__ G, const T* __restrict__ Dq,
             const T* __restrict__ I, const T* __restrict__ eo, T lambda, const T* __restrict__ q,
             T* __restrict__ Aq) {
  constexpr int MQ = NQ + 1;
  constexpr int NQ2 = NQ * NQ;
  constexpr int NP = NQ2 * NQ;
  constexpr int MQ2 = MQ * MQ;
  constexpr int MP = MQ2 * MQ;
  constexpr int PR = MQ | 1;  // line stride in LDS
  constexpr int PS = MQ * PR;  // slab stride in LDS
  constexpr int SA = pencil_tile_a<MQ, STIFF>();
  constexpr int SB = STIFF ? pencil_tile_b<MQ, STIFF>() : SA;
  constexpr int S = pencil_slots<MQ>();
  // (the FP64 even-odd form holds one slab less: with four it spills)
  constexpr int PF = STIFF ? (EO && sizeof(T) == 8 ? (MQ < 3 ? MQ : 3) : pencil_prefetch<T, MQ>()) : 1;
  constexpr bool SG = sizeof(T) == 4;  // FP32: geometry loads non-temporal
  static_assert(S * MQ2 <= 256, "one line per lane and pass");
  static_assert(NQ * NQ * PR <= SA && NQ * PS <= SB, "x and y fit their tiles");
  __shared__ T s_a[S * SA];
  __shared__ T s_bt[STIFF ? S * SB : 1];
  T* const s_b = STIFF ? s_bt : s_a;  // BP1/BP2: one tile

  const int pairs = Nelements * NF;
  const int groups = (pairs + S - 1) / S;
  const int per_xcd = (groups + 7) / 8;
  // each XCD (block mod 8) takes a contiguous run of slot groups, so face
  // neighbours share q lines in its L2
  const int grp = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
  if (grp >= groups) return;  // uniform over the workgroup
  const int t = threadIdx.x;

  // the element and field of slot s (an idle slot reads element 0's data
  // and stores nothing)
  // The group's slots cover at most NE consecutive entries of elementList
  // from eo0; where NE is small they are read with scalar loads (uniform
  // over the workgroup, through the scalar cache), which takes a vector
  // memory round trip off the chain element -> G2L -> q.
  constexpr int NE = (S + NF - 1) / NF + 1;
  const int eo0 = (grp * S) / NF;
  int E[NE <= 4 ? NE : 1];
  if constexpr (NE <= 4) {
    typedef const __attribute__((address_space(4))) int* ciptr;
    const ciptr el = (ciptr)elementList;
#pragma unroll
    for (int j = 0; j < NE; ++j) E[j] = el[min(eo0 + j, Nelements - 1)];
  }
  auto slot = [&](int s, int& e, int& f) -> bool {
    const int p = grp * S + s;
    const bool act = s < S && p < pairs;
    const int eo = act ? p / NF : eo0;
    f = act ? p - eo * NF : 0;
    if constexpr (NE <= 4) {
      const int rel = eo - eo0;  // in [0, NE)
      e = E[0];
#pragma unroll
      for (int j = 1; j < NE; ++j) e = rel == j ? E[j] : e;
    } else {
      e = elementList[eo];
    }
    return act;
  };

  // i-line lanes (passes 1 and 9): line (k, j) = li of slot si
  const int si = t / NQ2, li = t - si * NQ2;
  const bool oni = si < S;
  // BP2 (three fields of a mass operator, the most gather-bound form)
  // gathers q and stores Aq in node order through LDS; the others gather
  // and store along the i-lines (faster for them: BP1 already streams at
  // the device-copy rate, and the extra pass costs BP3/BP4 more than the
  // gather gains)
  constexpr bool NODE_ORDER = NF > 1 && !STIFF;
  int ei, fi;
  const bool acti = slot(si, ei, fi);
  const size_t lbi = ((size_t)ei * NF + fi) * NP + (size_t)li * NQ;
  // j-line lanes (passes 2 and 8): line (k, a)
  const int sj = t / (NQ * MQ), lj = t - sj * (NQ * MQ);
  const bool onj = sj < S;
  const int jk = lj / MQ, ja = lj - jk * MQ;
  // t-line lanes: line (b, a) = lt of slot st; the r- and s-line passes
  // use the same split, line (g, c) = lt
  const int st = t / MQ2, lt = t - st * MQ2;
  const bool ont = st < S;
  const int tb = lt / MQ, ta = lt - tb * MQ;
  const int tp = tb * PR + ta;  // LDS offset of point (., b, a) in a slab
  int et, ft;
  slot(st, et, ft);
  T* const At = s_a + st * SA;
  T* const Bt = s_b + st * SB;

  // the 1D contractions: row by row from I and D, or (EO) in even-odd form
  // from the tables in eo
  using ET = EoTables<NQ>;
  auto interp = [&](const T(&x)[NQ], T(&y)[MQ]) {  // y = I x
    if constexpr (EO) {
      eo_apply<MQ, NQ, 1>(eo + ET::I, x, [&](int n, T v) { y[n] = v; });
    } else {
      uniform_rows<MQ, NQ>(I, [&](int n, const T* r) {
        T acc = 0;
#pragma unroll
        for (int m = 0; m < NQ; ++m) acc = cfma(r[m], x[m], acc);
        y[n] = acc;
      });
    }
  };
  auto project = [&](const T(&x)[MQ], T(&y)[NQ]) {  // y = I^T x
    if constexpr (EO) {
      eo_apply<NQ, MQ, 1>(eo + ET::IT, x, [&](int n, T v) { y[n] = v; });
    } else {
#pragma unroll
      for (int n = 0; n < NQ; ++n) y[n] = T(0);
      uniform_rows<MQ, NQ>(I, [&](int m, const T* r) {
#pragma unroll
        for (int n = 0; n < NQ; ++n) y[n] = cfma(r[n], x[m], y[n]);
      });
    }
  };

  T wv[STIFF ? 1 : MQ];
  if constexpr (!STIFF) {
    // cubwJ of the t-line, requested first
    if (ont) {
#pragma unroll
      for (int g = 0; g < MQ; ++g) wv[g] = W[(size_t)et * MP + g * MQ2 + lt];
    }
  }

  // The slots' nodes in node order: node c of slot s is lane t of round r,
  // c + s * NP = r * 256 + t, so consecutive lanes read consecutive G2L
  // entries and write consecutive Aq entries, and a q gather touches a few
  // lattice rows per instruction instead of one cache line per lane. Row
  // (k, j) of a slot sits in LDS at (k * NQ + j) * PR (the i-lines' rows).
  // Bounds-checked buffer accesses: an idle slot or lane gets an offset
  // beyond the range and reads 0 or writes nothing.
  constexpr int R = NODE_ORDER ? (S * NP + 255) / 256 : 1;
  static_assert(STIFF ? pencil_tile_b<MQ, STIFF>() == SA : true, "the output rows use the same offsets");
  // round r: the local node index (kBufMasked: none) and its LDS offset
  // (-1: none); recomputed for the stores rather than held in registers
  // (offsets fit in 32 bits: pencil_in_range bounds the local vector).
  // Up to 4 slots, each slot's local base is uniform, formed once and
  // selected per round; beyond that each round looks its slot up.
  constexpr bool FEW = S <= 4;
  int sbase[FEW ? S : 1];
  if constexpr (NODE_ORDER && FEW) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      int e, f;
      sbase[s] = slot(s, e, f) ? (e * NF + f) * NP : -1;
    }
  }
  auto node = [&](int r, int lane, int& nlds) -> int {
    if constexpr (FEW && NP % 256 == 0) {
      // a round lies inside one slot (Nq = 8): the slot and its base are
      // fixed per round
      const int s = (r * 256) / NP, c = (r * 256) % NP + lane;
      nlds = s < S ? s * SA + (c / NQ) * PR + c % NQ : -1;
      const int b = s < S ? sbase[s < S ? s : 0] : -1;
      return b >= 0 ? b + c : kBufMasked;
    }
    const int n = r * 256 + lane;
    const int s = n / NP, c = n - s * NP;
    nlds = s < S ? s * SA + (c / NQ) * PR + c % NQ : -1;
    if constexpr (FEW) {
      int b = -1;
#pragma unroll
      for (int j = 0; j < S; ++j) b = s == j ? sbase[j] : b;
      return b >= 0 ? b + c : kBufMasked;
    } else {
      int e, f;
      const bool act = slot(s, e, f);
      return act ? (e * NF + f) * NP + c : kBufMasked;
    }
  };
  if constexpr (NODE_ORDER) {
    int noff[R], nlds[R];
#pragma unroll
    for (int r = 0; r < R; ++r) noff[r] = node(r, t, nlds[r]);
    const __amdgpu_buffer_rsrc_t rg = buffer_rsrc(G2L);
    const __amdgpu_buffer_rsrc_t rq = buffer_rsrc(q);
    int id[R];
#pragma unroll
    for (int r = 0; r < R; ++r) id[r] = buffer_load<int>(rg, noff[r] == kBufMasked ? kBufMasked : noff[r] * 4);
    T qv[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
      qv[r] = buffer_load<T>(rq, (noff[r] != kBufMasked && id[r] >= 0) ? id[r] * (int)sizeof(T) : kBufMasked);
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (nlds[r] >= 0) s_a[nlds[r]] = qv[r];
    __syncthreads();
  }

  // pass 1: i-lines, interpolate q along i -> A (node order: in place,
  // each lane reads its row before rewriting it; else gathered here)
  if (oni) {
    T* const x = s_a + si * SA + li * PR;
    T v[NQ];
    if constexpr (NODE_ORDER) {
#pragma unroll
      for (int i = 0; i < NQ; ++i) v[i] = x[i];
    } else {
      const __amdgpu_buffer_rsrc_t rq = buffer_rsrc(q);
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        const int id = G2L[lbi + i];
        v[i] = buffer_load<T>(rq, (acti && id >= 0) ? id * (int)sizeof(T) : kBufMasked);
      }
    }
    T xv[MQ];
    interp(v, xv);
#pragma unroll
    for (int a = 0; a < MQ; ++a) x[a] = xv[a];
  }
  __syncthreads();
  // pass 2: j-lines, along j -> B (y[k][b][a] at k*PS + b*PR + a)
  {
    T v[NQ], yv[MQ];
    if (onj) {
      const T* x = s_a + sj * SA + jk * NQ * PR + ja;
#pragma unroll
      for (int j = 0; j < NQ; ++j) v[j] = x[j * PR];
      interp(v, yv);
    }
    if constexpr (!STIFF) __syncthreads();  // y overwrites x in the one tile
    if (onj) {
      T* const y = s_b + sj * SB + jk * PS + ja;
#pragma unroll
      for (int b = 0; b < MQ; ++b) y[b * PR] = yv[b];
    }
  }
  __syncthreads();
  // pass 3: t-lines, along k
  T u[MQ];
  T gpre[PF][7];
  if (ont) {
    T v[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) v[k] = Bt[tp + k * PS];
    if constexpr (STIFF) {
      // the first PF slabs' geometric factors, in flight from here
#pragma unroll
      for (int g = 0; g < PF; ++g) {
        const size_t cb = (size_t)et * MP + g * MQ2 + lt;
#pragma unroll
        for (int c = 0; c < 6; ++c) gpre[g][c] = geo_load<SG>(G + 6 * cb + c);
        gpre[g][6] = geo_load<SG>(W + cb);
      }
    }
    interp(v, u);
    if constexpr (!STIFF) {
      // times cubwJ, back along k, in place
      T zg[MQ], z[NQ];
#pragma unroll
      for (int g = 0; g < MQ; ++g) zg[g] = u[g] * wv[g];
      project(zg, z);
#pragma unroll
      for (int k = 0; k < NQ; ++k) Bt[tp + k * PS] = z[k];
    } else {
#pragma unroll
      for (int g = 0; g < MQ; ++g) At[tp + g * PS] = u[g];
    }
  }
  __syncthreads();

  if constexpr (STIFF) {
    // contraction of the line at p (stride ls) with D (rows n) or D^T
    // (rows m), written back to the same points
    auto line_d = [&](const T* src, T* dst, int ls) {
      T v[MQ];
#pragma unroll
      for (int m = 0; m < MQ; ++m) v[m] = src[m * ls];
      if constexpr (EO) {
        // outputs held until the rows are done: an LDS store inside the row
        // loop would join each row's wait for its scalar load
        T o[MQ];
        eo_apply<MQ, MQ, -1>(eo + ET::D, v, [&](int n, T y) { o[n] = y; });
#pragma unroll
        for (int n = 0; n < MQ; ++n) dst[n * ls] = o[n];
      } else {
        uniform_rows<MQ, MQ>(Dq, [&](int n, const T* r) {
          T acc = 0;
#pragma unroll
          for (int m = 0; m < MQ; ++m) acc = cfma(r[m], v[m], acc);
          dst[n * ls] = acc;
        });
      }
    };
    auto line_dt = [&](T* p, int ls) {
      T v[MQ], o[MQ];
#pragma unroll
      for (int m = 0; m < MQ; ++m) v[m] = p[m * ls];
      if constexpr (EO) {
        eo_apply<MQ, MQ, -1>(eo + ET::DT, v, [&](int n, T y) { o[n] = y; });
      } else {
#pragma unroll
        for (int n = 0; n < MQ; ++n) o[n] = T(0);
        uniform_rows<MQ, MQ>(Dq, [&](int m, const T* r) {
#pragma unroll
          for (int n = 0; n < MQ; ++n) o[n] = cfma(r[n], v[m], o[n]);
        });
      }
#pragma unroll
      for (int n = 0; n < MQ; ++n) p[n * ls] = o[n];
    };
    // the r-line (g, b) = (tb, ta) runs along a at tb*PS + ta*PR; the
    // s-line (g, a) = (tb, ta) along b at tb*PS + ta
    const int ro = tb * PS + ta * PR, so = tb * PS + ta;
    // pass 4: r-lines, ur along a -> B
    if (ont) line_d(At + ro, Bt + ro, 1);
    __syncthreads();
    // pass 5: t-lines take their ur
    T ur[MQ];
    if (ont) {
#pragma unroll
      for (int g = 0; g < MQ; ++g) ur[g] = Bt[tp + g * PS];
    }
    __syncthreads();
    // pass 6: s-lines, us along b -> B
    if (ont) line_d(At + so, Bt + so, PR);
    __syncthreads();
    // pass 7: t-lines: ut, G (Gr -> A, Gs -> B), the mass term and the
    // divergence's t part
    T res[MQ];
    if (ont) {
#pragma unroll
      for (int m = 0; m < MQ; ++m) res[m] = T(0);
      uniform_rows<MQ, MQ>(Dq, [&](int g, const T* r) {
        T ut = 0;
#pragma unroll
        for (int m = 0; m < MQ; ++m) ut = cfma(r[m], u[m], ut);
        const T us = Bt[tp + g * PS];
        T gg[7];
        if (g < PF) {
#pragma unroll
          for (int c = 0; c < 7; ++c) gg[c] = gpre[g < PF ? g : 0][c];
        } else {
          const size_t cb = (size_t)et * MP + g * MQ2 + lt;
#pragma unroll
          for (int c = 0; c < 6; ++c) gg[c] = geo_load<SG>(G + 6 * cb + c);
          gg[6] = geo_load<SG>(W + cb);
        }
        const T G00 = gg[0], G01 = gg[1], G02 = gg[2], G11 = gg[3], G12 = gg[4], G22 = gg[5], WW = gg[6];
        At[tp + g * PS] = cfma(G02, ut, cfma(G01, us, G00 * ur[g]));
        Bt[tp + g * PS] = cfma(G12, ut, cfma(G11, us, G01 * ur[g]));
        const T Gt = cfma(G22, ut, cfma(G12, us, G02 * ur[g]));
        res[g] = cfma(lambda * WW, u[g], res[g]);
#pragma unroll
        for (int m = 0; m < MQ; ++m) res[m] = cfma(r[m], Gt, res[m]);
      });
    }
    __syncthreads();
    // pass 8: r-lines D^T Gr (A), s-lines D^T Gs (B), in place
    if (ont) {
      line_dt(At + ro, 1);
      line_dt(Bt + so, PR);
    }
    __syncthreads();
    // pass 9: t-lines: add the r and s parts, back along g -> A
    if (ont) {
#pragma unroll
      for (int g = 0; g < MQ; ++g) res[g] += At[tp + g * PS] + Bt[tp + g * PS];
      T z[NQ];
      project(res, z);
#pragma unroll
      for (int k = 0; k < NQ; ++k) At[tp + k * PS] = z[k];
    }
    __syncthreads();
  }

  // j-lines, back along j: z (B; A for BP3/BP4) -> x (A; B)
  {
    T* const Z = STIFF ? s_a + sj * SA : s_b + sj * SB;
    T* const X = STIFF ? s_b + sj * SB : s_a + sj * SA;
    T v[MQ], o[NQ];
    if (onj) {
      const T* z = Z + jk * PS + ja;
#pragma unroll
      for (int b = 0; b < MQ; ++b) v[b] = z[b * PR];
      project(v, o);
    }
    if constexpr (!STIFF) __syncthreads();  // x overwrites z in the one tile
    if (onj) {
      T* const x = X + jk * NQ * PR + ja;
#pragma unroll
      for (int j = 0; j < NQ; ++j) x[j * PR] = o[j];
    }
  }
  __syncthreads();
  // i-lines, back along i: to Aq, or (node order) in place
  T* const Xo = STIFF ? s_b : s_a;
  if (NODE_ORDER ? oni : acti) {
    T* const x = Xo + si * (STIFF ? SB : SA) + li * PR;
    T v[MQ], o[NQ];
#pragma unroll
    for (int a = 0; a < MQ; ++a) v[a] = x[a];
    project(v, o);
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      if constexpr (NODE_ORDER)
        x[i] = o[i];
      else
        Aq[lbi + i] = o[i];
    }
  }
  // Aq in node order (the lane index made opaque, so the compiler
  // recomputes the offsets instead of keeping the gather's live)
  if constexpr (NODE_ORDER) {
    __syncthreads();
    const __amdgpu_buffer_rsrc_t ra = buffer_rsrc(Aq);
    int lane = t;
    asm volatile("" : "+v"(lane));
#pragma unroll
    for (int r = 0; r < R; ++r) {
      int nl;
      const int no = node(r, lane, nl);
      const T v = nl >= 0 ? Xo[nl] : T(0);
      buffer_store<T>(v, ra, no == kBufMasked ? kBufMasked : no * (int)sizeof(T));
    }
  }
}

// the 3-field wavefront forms split their fields over workgroups (set to 0
// to run them one after another in one workgroup; same results)
#ifndef BPX_FIELD_SPLIT
#define BPX_FIELD_SPLIT 1
#endif

template <typename T, int NQ, int MQ, int NF, int MODE>
int launch_mode(hipStream_t s, const AxArgs<T>& x, const T* W, const T* G, const T* Dq, const T* I) {
  constexpr int WV = cub_waves<T, MQ, NF, MODE>();
  constexpr int EPB = cub_epb<MQ>();
  const int blocks = (x.Nelements + EPB - 1) / EPB;
  launch_stamped(k_cub_ax<T, NQ, MQ, NF, MODE, WV, EPB>, dim3(blocks), dim3(MQ * MQ * EPB), s, x.evStart, x.evStop,
                 x.Nelements, x.elementList, x.GlobalToLocal, W, G, Dq, I, x.lambda, x.q, x.Aq);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

// the pencil form gathers q at byte offsets below kBufRange: every
// gathered index lies below the number of local nodes
template <typename T, int NQ, int NF>
bool pencil_in_range(const AxArgs<T>& a) {
  return (double)a.Nelements * NF * NQ * NQ * NQ * sizeof(T) < (double)kBufRange;
}

// even-odd tables (a.eo, formed by the host from I and D) where given: the
// problem objects' operators; the kernel-level entry points take I and D
// only and run the row-by-row form
template <typename T, int NQ, int NF, bool STIFF>
int launch_cub_pencil(hipStream_t s, const AxArgs<T>& a, const T* G, const T* Dq) {
  constexpr int S = pencil_slots<NQ + 1>();
  const long long groups = ((long long)a.Nelements * NF + S - 1) / S;
  const unsigned blocks = 8u * (unsigned)((groups + 7) / 8);
  if (a.eo && a.cubForm != CUB_FORM_PENCIL_ROWS)
    launch_stamped(k_cub_pencil<T, NQ, NF, STIFF, true>, dim3(blocks), dim3(256), s, a.evStart, a.evStop,
                   a.Nelements, a.elementList, a.GlobalToLocal, a.wJ, G, Dq, a.I, a.eo, a.lambda, a.q, a.Aq);
  else
    launch_stamped(k_cub_pencil<T, NQ, NF, STIFF, false>, dim3(blocks), dim3(256), s, a.evStart, a.evStop,
                   a.Nelements, a.elementList, a.GlobalToLocal, a.wJ, G, Dq, a.I, a.eo, a.lambda, a.q, a.Aq);
  BPX_HIP(hipGetLastError());
  return BPX_OK;
}

}  // namespace

template <typename T, int NQ, int NF>
int launch_bp1_nq(hipStream_t s, const AxArgs<T>& a) {
  if (a.affine) return launch_mode<T, NQ, NQ, NF, MODE_BP1A>(s, a, a.wJ, nullptr, nullptr, a.MM);
  if constexpr (NQ <= 8) {
    if (a.cubForm != CUB_FORM_WAVE && pencil_in_range<T, NQ, NF>(a))
      return launch_cub_pencil<T, NQ, NF, false>(s, a, nullptr, nullptr);
    constexpr bool FP = NF > 1 && BPX_FIELD_SPLIT;
    launch_stamped(k_bp1_wave<T, NQ, NF, FP>, dim3(wave_blocks<NF, FP>(a.Nelements)), dim3(64), s, a.evStart,
                   a.evStop, a.Nelements, a.elementList, a.GlobalToLocal, a.wJ, a.I, a.q, a.Aq);
    BPX_HIP(hipGetLastError());
    return BPX_OK;
  }
  return launch_mode<T, NQ, NQ + 1, NF, MODE_BP1>(s, a, a.wJ, nullptr, nullptr, a.I);
}

template <typename T, int NQ, int NF>
int launch_bp3_nq(hipStream_t s, const AxArgs<T>& a) {
  if (a.affine) return launch_mode<T, NQ, NQ, NF, MODE_BP3A>(s, a, a.wJ, a.ggeo, a.D, a.invV);
  if constexpr (NQ <= 8) {
    if (a.cubForm != CUB_FORM_WAVE && pencil_in_range<T, NQ, NF>(a))
      return launch_cub_pencil<T, NQ, NF, true>(s, a, a.ggeo, a.D);
    constexpr bool FP = NF > 1 && BPX_FIELD_SPLIT;
    launch_stamped(k_bp3_wave<T, NQ, NF, FP>, dim3(wave_blocks<NF, FP>(a.Nelements)), dim3(128), s, a.evStart,
                   a.evStop, a.Nelements, a.elementList, a.GlobalToLocal, a.wJ, a.ggeo, a.D, a.I, a.lambda, a.q,
                   a.Aq);
    BPX_HIP(hipGetLastError());
    return BPX_OK;
  }
  return launch_mode<T, NQ, NQ + 1, NF, MODE_BP3>(s, a, a.wJ, a.ggeo, a.D, a.I);
}

template int launch_bp1_nq<double, BPX_NQ, 1>(hipStream_t, const AxArgs<double>&);
template int launch_bp1_nq<double, BPX_NQ, 3>(hipStream_t, const AxArgs<double>&);
template int launch_bp1_nq<float, BPX_NQ, 1>(hipStream_t, const AxArgs<float>&);
template int launch_bp3_nq<double, BPX_NQ, 1>(hipStream_t, const AxArgs<double>&);
template int launch_bp3_nq<double, BPX_NQ, 3>(hipStream_t, const AxArgs<double>&);
template int launch_bp3_nq<float, BPX_NQ, 1>(hipStream_t, const AxArgs<float>&);

}  // namespace bpx

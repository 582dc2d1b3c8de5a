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

}  // namespace

template <typename T, int NQ, int NF>
int launch_bp1_nq(hipStream_t s, const AxArgs<T>& a) {
  if (a.affine) return launch_mode<T, NQ, NQ, NF, MODE_BP1A>(s, a, a.wJ, nullptr, nullptr, a.MM);
  if constexpr (NQ <= 8) {
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

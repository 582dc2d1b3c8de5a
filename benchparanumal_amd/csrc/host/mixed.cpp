// FP32 inner CG inside an FP64 defect-correction loop (no reference
// counterpart; the reference solves in dfloat = double only).
//
//   r = b - A x                     (FP64 operator, FP64 vectors)
//   while ||r|| / ||b|| >= tol:
//     r32 = float(r / ||r||)        (scaled so FP32 sees O(1) values)
//     solve A32 c = r32 by CG in FP32 to a relative reduction inner_tol
//     x  += ||r|| * c               (FP64 update)
//     r   = b - A x                 (FP64 residual)
//
// The inner CG uses the FP32 operator (geometric factors, D and the
// interpolation matrices rounded to float, FP32 arithmetic) and FP32
// vectors; its dot products accumulate in FP64. FP64 work is limited to the
// outer residual, its norm and the solution update.
#include <cmath>

#include "bpx.hpp"
#include "cg_impl.hpp"

namespace bpx {

int bp_t::SolveMixed(dfloat* o_x, const dfloat* o_b, double tol, double inner_tol, int max_outer, int max_inner,
                     double* stats) {
  if (Nfields != 1) return set_error(BPX_ERR_UNSUPPORTED, "mixed solve: one-field problems (BP1/3/5) only");
  if (!(tol > 0) || !(inner_tol > 0) || max_outer < 1 || max_inner < 1)
    return set_error(BPX_ERR_INVALID_ARGUMENT, "mixed solve: tol, inner_tol > 0 and iteration caps >= 1");
  if (int st = Prepare32()) return st;
  const dlong N = ogs.Ngather;
  hipStream_t s = platform->stream;
  linAlg_t la(*platform);
  deviceBuffer<double> r64, Ax64;
  deviceBuffer<float> r32, c32;
  if (int st = r64.alloc(N)) return st;
  if (int st = Ax64.alloc(N)) return st;
  if (int st = r32.alloc(N)) return st;
  if (int st = c32.alloc(N)) return st;
  cg_t<float> inner;
  if (int st = inner.Setup(*platform, N)) return st;

  double nb = 0;
  if (int st = la.norm2(N, o_b, &nb)) return st;
  if (nb == 0) {
    if (int st = la.set(N, 0.0, o_x)) return st;
    if (stats) stats[0] = stats[1] = stats[2] = 0;
    return platform->finish();
  }
  auto residual = [&](double* rn) -> int {
    if (int st = Operator(o_x, Ax64.ptr())) return st;
    BPX_HIP(hipMemcpyAsync(r64.ptr(), o_b, N * sizeof(double), hipMemcpyDeviceToDevice, s));
    if (int st = la.axpy(N, -1.0, Ax64.ptr(), 1.0, r64.ptr())) return st;
    return la.norm2(N, r64.ptr(), rn);
  };
  double rn = 0;
  if (int st = residual(&rn)) return st;
  int outer = 0, inner_total = 0;
  auto op32 = [&](const float* q, float* Aq) { return Operator32(q, Aq); };
  while (outer < max_outer && rn / nb >= tol) {
    if (int st = launch_copy_f64_to_f32(s, N, 1.0 / rn, r64.ptr(), r32.ptr())) return st;
    if (int st = launch_set<float>(s, N, 0.0f, c32.ptr())) return st;
    int its = 0;
    const int st_in = inner.Solve(op32, c32.ptr(), r32.ptr(), inner_tol, max_inner, &its);
    if (st_in && st_in != BPX_ERR_BREAKDOWN) return st_in;
    inner_total += its;
    if (int st = launch_axpy_f32_to_f64(s, N, rn, c32.ptr(), 1.0, o_x)) return st;
    const double prev = rn;
    if (int st = residual(&rn)) return st;
    ++outer;
    if (!(rn < prev)) break;  // no progress (FP32 floor or breakdown)
  }
  if (stats) {
    stats[0] = outer;
    stats[1] = inner_total;
    stats[2] = rn / nb;
  }
  return platform->finish();
}

}  // namespace bpx

// FP32 CG with FP64 reliable updates (no reference counterpart; the
// reference solves in dfloat = double only, include/types.h).
//
// The CG recurrence runs in FP32: FP32 vectors (r, p, Ap and the
// correction c) and the FP32 operator (geometric factors, D and the
// interpolation matrices rounded to float, FP32 arithmetic); its dot
// products accumulate in FP64. Every `interval` iterations an FP64 defect
// correction replaces the drifting FP32 residual by the true one:
//
//   x   += c;  c = 0                (FP64 solution update)
//   r64  = b - A x                  (FP64 operator and residual)
//   r    = float(r64)               (the recurrence continues from it)
//
// and the stopping test ||r64|| / ||b|| < tol is made on that FP64
// residual. Unlike restarted defect correction, the Krylov recurrence is
// kept across the updates (beta continues from the previous r.r, as in
// reliable-update CG), so the FP32 iteration count stays close to the FP64
// one while each iteration moves about half the bytes.
#include <cmath>

#include "bpx.hpp"

namespace bpx {

int bp_t::SolveMixed(dfloat* o_x, const dfloat* o_b, double tol, int interval, int maxit, double* stats) {
  if (!(tol > 0) || interval < 1 || maxit < 1)
    return set_error(BPX_ERR_INVALID_ARGUMENT, "mixed solve: tol > 0, interval >= 1 and maxit >= 1 required");
  if (int st = Prepare32()) return st;
  const dlong N = ogs.Ngather;
  hipStream_t s = platform->stream;
  MixedWork& w = mixed;
  if (w.N != N || w.maxit < maxit) {
    int st;
    if ((st = w.Ax64.alloc(N)) || (st = w.r32.alloc(N)) || (st = w.p32.alloc(N)) || (st = w.Ap32.alloc(N)) ||
        (st = w.c32.alloc(N)) || (st = w.part_true.alloc(kRedBlocks)) || (st = w.part_rr.alloc(kRedBlocks)) ||
        (st = w.part_pAp.alloc(std::max(GatherPartials(), kRedBlocks))) || (st = w.hist_rr.alloc(maxit + 1)) ||
        (st = w.hist_alpha.alloc(maxit + 1)) || (st = w.state.alloc(1)) || (st = w.norm.alloc(1)))
      return st;
    w.N = N;
    w.maxit = maxit;
  }
  linAlg_t la(*platform);
  double nb = 0;
  if (int st = la.norm2(N, o_b, &nb)) return st;
  if (nb == 0) {
    if (int st = la.set(N, 0.0, o_x)) return st;
    if (stats) stats[0] = stats[1] = stats[2] = 0;
    return platform->finish();
  }
  // true residual of the initial guess
  auto true_residual = [&](double* rn) -> int {
    if (int st = Operator(o_x, w.Ax64.ptr())) return st;
    if (int st = launch_true_residual(s, N, o_b, w.Ax64.ptr(), w.r32.ptr(), w.part_true.ptr(), w.part_rr.ptr()))
      return st;
    if (int st = launch_fold(s, w.part_true.ptr(), kRedBlocks, w.norm.ptr())) return st;
    BPX_HIP(hipMemcpyAsync(platform->h_scalar, w.norm.ptr(), sizeof(double), hipMemcpyDeviceToHost, s));
    if (int st = platform->finish()) return st;
    *rn = std::sqrt(*platform->h_scalar);
    return BPX_OK;
  };
  if (int st = launch_set<float>(s, N, 0.0f, w.c32.ptr())) return st;
  double rn = 0;
  if (int st = true_residual(&rn)) return st;
  int k = 0, updates = 0;
  CgState* st_d = w.state.ptr();
  while (rn > tol * nb && k < maxit) {
    const int n = std::min(interval, maxit - k);
    for (int j = 0; j < n; ++j, ++k) {
      // tol = 0: the recurrence never stops itself; the FP64 test decides
      if (int st = launch_cg_update_p<float>(s, N, k, w.r32.ptr(), w.p32.ptr(), w.part_rr.ptr(), kRedBlocks, st_d,
                                             w.hist_rr.ptr(), 0.0))
        return st;
      int npart = 0;
      if (int st = OperatorDot32(w.p32.ptr(), w.Ap32.ptr(), w.part_pAp.ptr(), &npart)) return st;
      if (int st = launch_cg_update_xr<float>(s, N, k, w.p32.ptr(), w.Ap32.ptr(), w.c32.ptr(), w.r32.ptr(),
                                              w.part_pAp.ptr(), npart, w.part_rr.ptr(), st_d, w.hist_rr.ptr(),
                                              w.hist_alpha.ptr()))
        return st;
    }
    // reliable update: x += c, r = b - A x in FP64, r32 = float(r)
    if (int st = launch_accumulate_f32(s, N, w.c32.ptr(), o_x)) return st;
    if (int st = true_residual(&rn)) return st;
    ++updates;
    CgState hst;
    BPX_HIP(hipMemcpy(&hst, st_d, sizeof(CgState), hipMemcpyDeviceToHost));
    if (hst.breakdown || !std::isfinite(rn))
      return set_error(BPX_ERR_BREAKDOWN, "mixed solve: breakdown in the FP32 recurrence");
  }
  if (stats) {
    stats[0] = updates;
    stats[1] = k;
    stats[2] = rn / nb;
  }
  return BPX_OK;
}

}  // namespace bpx

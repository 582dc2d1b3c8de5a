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
// one while each iteration moves about half the bytes. Only when the FP32
// recurrence has stopped by itself (r.r = 0, or p.Ap <= 0 once its residual
// has fallen out of FP32 range) is it restarted from the true residual.
#include <cmath>

#include "bpx.hpp"

namespace bpx {

int bp_t::SolveMixed(dfloat* o_x, const dfloat* o_b, double tol, int interval, int maxit, double* stats) {
  if (!(tol > 0) || interval < 1 || maxit < 1)
    return set_error(BPX_ERR_INVALID_ARGUMENT, "mixed solve: tol > 0, interval >= 1 and maxit >= 1 required");
  if (int st = Prepare32()) return st;
  const dlong N = Ns;  // the recurrence runs in the solve space
  hipStream_t s = platform->stream;
  MixedWork& w = mixed;
  if (w.N != N || w.maxit < maxit) {
    int st;
    if ((st = w.Ax64.alloc(N)) || (st = w.r32.alloc(N)) || (st = w.p32.alloc(N)) || (st = w.Ap32.alloc(N)) ||
        (st = w.xs.alloc(N)) || (st = w.bs.alloc(N)) ||
        (st = w.c32.alloc(N)) || (st = w.part_true.alloc(kRedBlocks)) || (st = w.part_rr.alloc(kRedBlocks)) ||
        (st = w.part_pAp.alloc(std::max(GatherPartials(), kRedBlocks))) || (st = w.hist_rr.alloc(maxit + 1)) ||
        (st = w.hist_alpha.alloc(maxit + 1)) || (st = w.state.alloc(1)) || (st = w.norm.alloc(1)))
      return st;
    // padding rows of the solve space are never written by the fused
    // operator: they must hold zeros from the start
    if (N > 0) {
      BPX_HIP(hipMemsetAsync(w.Ap32.ptr(), 0, (size_t)N * sizeof(float), s));
      BPX_HIP(hipMemsetAsync(w.p32.ptr(), 0, (size_t)N * sizeof(float), s));
    }
    w.N = N;
    w.maxit = maxit;
  }
  linAlg_t la(*platform);
  double nb = 0;
  if (int st = la.norm2(ogs.Ngather, o_b, &nb)) return st;
  if (nb == 0) {
    if (int st = la.set(ogs.Ngather, 0.0, o_x)) return st;
    if (stats) stats[0] = stats[1] = stats[2] = 0;
    return platform->finish();
  }
  double* xs = w.xs.ptr();
  const double* bs = w.bs.ptr();
  if (int st = ToSolve(o_x, xs)) return st;
  if (int st = ToSolve(o_b, w.bs.ptr())) return st;
  // true residual of the initial guess
  auto true_residual = [&](double* rn) -> int {
    if (int st = OperatorS(xs, w.Ax64.ptr())) return st;
    if (int st = launch_true_residual(s, N, bs, w.Ax64.ptr(), w.r32.ptr(), w.part_true.ptr(), w.part_rr.ptr()))
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
  if (!std::isfinite(rn) || !std::isfinite(nb)) return set_error(BPX_ERR_BREAKDOWN, "mixed solve: non-finite residual");
  int k = 0, kseg = 0, updates = 0;  // kseg: iteration within the current recurrence
  CgState* st_d = w.state.ptr();
  // with the fused FP32 operator the direction update of the interior rows
  // is done inside it (CgFold); k_cg_update_p covers the rows below
  const dlong frows = FoldRows(32);
  while (rn > tol * nb && k < maxit) {
    const int n = std::min(interval, maxit - k);
    for (int j = 0; j < n; ++j, ++k, ++kseg) {
      // tol = 0: the recurrence never stops itself; the FP64 test decides.
      // c += alpha p of the previous iteration rides along (not across a
      // correction, before which it is flushed)
      if (int st = launch_cg_update_p<float>(s, frows > 0 ? frows : N, kseg, w.r32.ptr(), w.p32.ptr(), w.c32.ptr(),
                                             w.part_rr.ptr(), kRedBlocks, st_d, w.hist_rr.ptr(), w.hist_alpha.ptr(),
                                             0.0, j > 0))
        return st;
      int npart = 0;
      if (frows > 0) {
        CgFold<float> f;
        f.on = 1;
        f.k = kseg;
        f.pending = j > 0;
        f.p = w.p32.ptr();
        f.r = w.r32.ptr();
        f.x = w.c32.ptr();
        f.hist_rr = w.hist_rr.ptr();
        f.hist_alpha = w.hist_alpha.ptr();
        f.st = st_d;
        if (int st = OperatorDot32(w.p32.ptr(), w.Ap32.ptr(), w.part_pAp.ptr(), &npart, &f)) return st;
      } else if (int st = OperatorDot32(w.p32.ptr(), w.Ap32.ptr(), w.part_pAp.ptr(), &npart)) {
        return st;
      }
      if (int st = launch_cg_update_r<float>(s, N, kseg, w.Ap32.ptr(), w.r32.ptr(), w.part_pAp.ptr(), npart,
                                             w.part_rr.ptr(), st_d, w.hist_rr.ptr(), w.hist_alpha.ptr()))
        return st;
    }
    // the last iteration's c update (unless the recurrence stopped, which
    // has applied or dropped it already)
    if (int st = launch_cg_flush_x<float>(s, N, kseg - 1, w.p32.ptr(), w.c32.ptr(), w.hist_alpha.ptr(), st_d, true))
      return st;
    // reliable update: x += c, r = b - A x in FP64, r32 = float(r)
    if (int st = launch_accumulate_f32(s, N, w.c32.ptr(), xs)) return st;
    if (int st = true_residual(&rn)) return st;
    ++updates;
    if (!std::isfinite(rn)) return set_error(BPX_ERR_BREAKDOWN, "mixed solve: non-finite residual");
    CgState hst;
    BPX_HIP(hipMemcpy(&hst, st_d, sizeof(CgState), hipMemcpyDeviceToHost));
    // the FP32 recurrence stopped (r.r = 0 or p.Ap <= 0, e.g. after it has
    // resolved the residual below FP32 range): restart it from the true
    // residual unless that has converged
    if (hst.done || hst.stop_at || hst.breakdown) kseg = 0;
  }
  if (int st = FromSolve(xs, o_x)) return st;
  if (int st = platform->finish()) return st;
  if (stats) {
    stats[0] = updates;
    stats[1] = k;
    stats[2] = rn / nb;
  }
  return BPX_OK;
}

}  // namespace bpx

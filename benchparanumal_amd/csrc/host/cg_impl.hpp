// Definition of cg_t<T>::Solve (a template over the operator callable, so
// it lives in a header that every caller includes).
#pragma once
#include <algorithm>

#include "bpx.hpp"

namespace bpx {

template <typename T>
template <class OP>
int cg_t<T>::Solve(OP&& apply, T* o_x, T* o_r, double tol, int MAXIT, int* iterations) {
  if (MAXIT < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "cg: MAXIT < 0");
  hipStream_t s = platform->stream;
  if (hist_len < MAXIT + 1) {
    if (int st = o_hist_rr.alloc(MAXIT + 1)) return st;
    if (int st = o_hist_alpha.alloc(MAXIT + 1)) return st;
    hist_len = MAXIT + 1;
  }
  last_maxit = MAXIT;
  CgState* st_d = o_state.ptr();
  void* scr = platform->reduce_scratch;
  // r <- b - A x, rdotr, stopping threshold (linearSolverCG.cpp:79-87)
  if (int st = apply(o_x, o_Ap.ptr())) return st;
  if (int st = launch_cg_init<T>(s, N, o_Ap.ptr(), o_r, scr, st_d, tol, MAXIT, o_hist_rr.ptr())) return st;
  // Iterations are queued without host waits; with tol > 0 the host polls
  // the device stop flag between chunks so a converged solve stops early.
  const int chunk = (tol > 0) ? 16 : (MAXIT > 0 ? MAXIT : 1);
  CgState hst;
  int queued = 0;
  while (queued < MAXIT) {
    const int n = std::min(chunk, MAXIT - queued);
    for (int k = 0; k < n; ++k) {
      if (int st = launch_cg_update_p<T>(s, N, o_r, o_p.ptr(), st_d)) return st;
      if (int st = apply(o_p.ptr(), o_Ap.ptr())) return st;
      if (int st = launch_cg_pAp<T>(s, N, o_p.ptr(), o_Ap.ptr(), scr, st_d)) return st;
      if (int st = launch_cg_update_xr<T>(s, N, o_p.ptr(), o_Ap.ptr(), o_x, o_r, scr, st_d,
                                          o_hist_rr.ptr(), o_hist_alpha.ptr()))
        return st;
    }
    queued += n;
    if (queued < MAXIT) {
      BPX_HIP(hipMemcpyAsync(&hst, st_d, sizeof(CgState), hipMemcpyDeviceToHost, s));
      BPX_HIP(hipStreamSynchronize(s));
      if (hst.done) break;
    }
  }
  BPX_HIP(hipMemcpyAsync(&hst, st_d, sizeof(CgState), hipMemcpyDeviceToHost, s));
  BPX_HIP(hipStreamSynchronize(s));
  if (iterations) *iterations = hst.iter;
  if (hst.breakdown) return set_error(BPX_ERR_BREAKDOWN, "cg: breakdown (non-finite or non-positive p.Ap / r.r)");
  return BPX_OK;
}

}  // namespace bpx

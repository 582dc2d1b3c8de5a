// Ragged batch of BP5 problems solved by one device-resident CG per
// problem, replayed from a captured hipGraph. No reference counterpart;
// each problem is set up exactly as bp5_t (BP/BP5/src/bp5Setup.cpp:29-95:
// box mesh, masked ogs, GlobalToLocal) and iterates cg::Solve
// (libs/core/linearSolverCG.cpp:62-129).
//
// Layout (see batch.hip): global vectors concatenated by problem; local
// vectors concatenated by degree group (problems of one degree share one Ax
// launch, their elements back to back); the gather maps of all problems
// concatenated with offsets, so one gather launch serves the batch and each
// row still sums its entries in the reference's ascending order.
#include <algorithm>
#include <cmath>
#include <map>

#include "bpx.hpp"

namespace bpx {

namespace {
constexpr int kChunk = 4096;        // vector entries per streaming workgroup
constexpr int kSuperBlocks = 8;     // row blocks per gather super-block
constexpr int kGraphIters = 8;      // CG iterations per captured graph
}  // namespace

bp_batch_t::~bp_batch_t() {
  if (graphExec) (void)hipGraphExecDestroy(graphExec);
  if (graphMixed) (void)hipGraphExecDestroy(graphMixed);
  if (h_count) (void)hipHostFree(h_count);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

int bp_batch_t::Setup(platform_t& p, const std::vector<MeshOpts>& opts, dfloat lambda_) {
  platform = &p;
  lambda = lambda_;
  nprob = (int)opts.size();
  if (nprob < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "batch: no problems");
  BPX_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  BPX_HIP(hipEventCreateWithFlags(&ev0, hipEventDisableTiming));
  BPX_HIP(hipEventCreateWithFlags(&ev1, hipEventDisableTiming));
  BPX_HIP(hipHostMalloc((void**)&h_count, sizeof(int), hipHostMallocDefault));

  // ---- per-problem host setup (bp5_t: mesh, masked ogs, GlobalToLocal)
  std::vector<mesh_t> meshes(nprob);
  std::vector<ogs::ogs_t> ogss(nprob);
  std::vector<std::vector<dlong>> g2l(nprob);
  for (int i = 0; i < nprob; ++i) {
    if (opts[i].affine) return set_error(BPX_ERR_UNSUPPORTED, "batch: non-affine (perturbed) meshes only");
    MeshOpts o = opts[i];
    o.cubature = 0;
    if (int st = meshes[i].Setup(o)) return st;
    std::vector<hlong> ids;
    meshes[i].MaskedGlobalIds(1, ids);
    const dlong NL = (dlong)ids.size();
    if (int st = ogss[i].Setup(NL, ids.data())) return st;
    g2l[i].resize(NL);
    ogss[i].SetupGlobalToLocalMapping(g2l[i].data());
  }

  // ---- global offsets (problem order)
  off.assign(nprob + 1, 0);
  for (int i = 0; i < nprob; ++i) {
    const long long next = (long long)off[i] + ogss[i].Ngather;
    if (next >= (1LL << 31)) return set_error(BPX_ERR_UNSUPPORTED, "batch: more than 2^31 unknowns");
    off[i + 1] = (dlong)next;
  }
  Ntot = off[nprob];

  // ---- degree groups; local layout group by group, problems in order
  std::map<int, int> groupOf;
  for (int i = 0; i < nprob; ++i)
    if (!groupOf.count(meshes[i].N)) {
      const int g = (int)groups.size();
      groupOf[meshes[i].N] = g;
      groups.emplace_back();
      groups[g].N = meshes[i].N;
      groups[g].Nq = meshes[i].Nq;
    }
  std::vector<long long> localOff(nprob);  // start of problem i in the local vector
  long long loff = 0;
  for (auto& G : groups) {
    G.loff = loff;
    G.Nel = 0;
    for (int i = 0; i < nprob; ++i)
      if (meshes[i].N == G.N) {
        localOff[i] = loff + (long long)G.Nel * meshes[i].Np;
        G.Nel += meshes[i].Nelements;
      }
    loff += (long long)G.Nel * G.Nq * G.Nq * G.Nq;
  }
  if (loff >= (1LL << 31)) return set_error(BPX_ERR_UNSUPPORTED, "batch: more than 2^31 local nodes");
  NLocalTot = (dlong)loff;

  hipStream_t s = stream;
  for (auto& G : groups) {
    const int Np = G.Nq * G.Nq * G.Nq;
    std::vector<double> ggeo, wJ;
    std::vector<dlong> G2L, pid, list(G.Nel);
    ggeo.reserve((size_t)G.Nel * Np * 6);
    wJ.reserve((size_t)G.Nel * Np);
    G2L.reserve((size_t)G.Nel * Np);
    pid.reserve(G.Nel);
    const std::vector<double>* D = nullptr;
    for (int i = 0; i < nprob; ++i) {
      if (meshes[i].N != G.N) continue;
      const mesh_t& m = meshes[i];
      D = &m.D;
      ggeo.insert(ggeo.end(), m.ggeo.begin(), m.ggeo.end());
      wJ.insert(wJ.end(), m.wJ.begin(), m.wJ.end());
      for (dlong l = 0; l < (dlong)g2l[i].size(); ++l) G2L.push_back(g2l[i][l] >= 0 ? g2l[i][l] + off[i] : -1);
      for (dlong e = 0; e < m.Nelements; ++e) pid.push_back(i);
    }
    for (dlong e = 0; e < G.Nel; ++e) list[e] = e;
    if (int st = G.ggeo.upload(ggeo, s)) return st;
    if (int st = G.wJ.upload(wJ, s)) return st;
    if (int st = G.G2L.upload(G2L, s)) return st;
    if (int st = G.elemPid.upload(pid, s)) return st;
    if (int st = G.elementList.upload(list, s)) return st;
    if (int st = G.D.upload(*D, s)) return st;
  }

  // ---- concatenated gather maps, super-blocks and chunks
  std::vector<dlong> rowStarts(1, 0), colIds;
  std::vector<BatchSuper> supers;
  std::vector<BatchChunk> chunks;
  std::vector<int2> pApRange(nprob), rrRange(nprob);
  colIds.reserve(NLocalTot);
  for (int i = 0; i < nprob; ++i) {
    const ogs::ogs_t& o = ogss[i];
    const dlong nnz0 = (dlong)colIds.size();
    for (dlong r = 0; r < o.Ngather; ++r) rowStarts.push_back(nnz0 + o.rowStartsT[r + 1]);
    for (dlong g = 0; g < o.nnzT; ++g) colIds.push_back((dlong)(localOff[i] + o.colIdsT[g]));
    pApRange[i].x = (int)supers.size();
    for (dlong b = 0; b < o.NrowBlocksT; b += kSuperBlocks) {
      const dlong b1 = std::min<dlong>(b + kSuperBlocks, o.NrowBlocksT);
      supers.push_back({i, off[i] + o.blockRowStartsT[b], off[i] + o.blockRowStartsT[b1], 0});
    }
    pApRange[i].y = (int)supers.size();
    rrRange[i].x = (int)chunks.size();
    const dlong n = o.Ngather;
    if (n == 0) chunks.push_back({i, off[i], off[i], 1});  // still initialises the state
    for (dlong c = 0; c < n; c += kChunk)
      chunks.push_back({i, off[i] + c, off[i] + std::min<dlong>(c + kChunk, n), c == 0 ? 1 : 0});
    rrRange[i].y = (int)chunks.size();
  }
  nsupers = (int)supers.size();
  nchunks = (int)chunks.size();
  if (int st = o_rowStarts.upload(rowStarts, s)) return st;
  if (int st = o_colIds.upload(colIds, s)) return st;
  if (int st = o_supers.upload(supers, s)) return st;
  if (int st = o_chunks.upload(chunks, s)) return st;
  if (int st = o_pApRange.upload(pApRange, s)) return st;
  if (int st = o_rrRange.upload(rrRange, s)) return st;
  if (int st = o_state.alloc(nprob)) return st;
  if (int st = o_r.alloc(std::max<dlong>(Ntot, 1))) return st;
  if (int st = o_p.alloc(std::max<dlong>(Ntot, 1))) return st;
  if (int st = o_Ap.alloc(std::max<dlong>(Ntot, 1))) return st;
  if (int st = o_AqL.alloc(std::max<dlong>(NLocalTot, 1))) return st;
  if (int st = o_part_rr.alloc(std::max(nchunks, 1))) return st;
  if (int st = o_part_pAp.alloc(std::max(nsupers, 1))) return st;
  if (int st = o_count.alloc(1)) return st;
  BPX_HIP(hipMemsetAsync(o_state.ptr(), 0, sizeof(BatchState) * nprob, s));
  BPX_HIP(hipMemsetAsync(o_p.ptr(), 0, sizeof(double) * std::max<dlong>(Ntot, 1), s));
  BPX_HIP(hipStreamSynchronize(s));
  return Tune();
}

int bp_batch_t::AxGroup(const Group& G, int variant, const double* q, double* AqL) const {
  return launch_bp5_batch(stream, G.N, variant, (int)G.Nel, G.elementList.ptr(), G.G2L.ptr(), G.wJ.ptr(),
                          G.ggeo.ptr(), G.D.ptr(), lambda, q, AqL + G.loff, G.elemPid.ptr(), o_state.ptr());
}

// Times every workgroup shape of each degree group's Ax on the group's own
// elements (all variants give identical results) and keeps the fastest.
int bp_batch_t::Tune() {
  hipEvent_t t0, t1;
  BPX_HIP(hipEventCreate(&t0));
  BPX_HIP(hipEventCreate(&t1));
  int st = BPX_OK;
  for (auto& G : groups) {
    float best = 1e30f;
    G.variant = 0;
    for (int v = 0; v < 3 && st == BPX_OK; ++v) {
      G.us[v] = -1;
      if (bp5_batch_epb(G.Nq, v) == 0) continue;
      if ((st = AxGroup(G, v, o_p.ptr(), o_AqL.ptr()))) break;  // warm-up
      float ms = 1e30f;
      for (int rep = 0; rep < 3 && st == BPX_OK; ++rep) {
        float t = 0;
        if (hipEventRecord(t0, stream) != hipSuccess || (st = AxGroup(G, v, o_p.ptr(), o_AqL.ptr())) ||
            hipEventRecord(t1, stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
            hipEventElapsedTime(&t, t0, t1) != hipSuccess) {
          if (!st) st = set_error(BPX_ERR_HIP, "batch tune: event timing failed");
          break;
        }
        ms = std::min(ms, t);
      }
      G.us[v] = ms * 1e3;
      if (ms < best) {
        best = ms;
        G.variant = v;
      }
    }
    if (st) break;
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return st;
}

// One CG iteration for every problem still running (queued on `stream`).
int bp_batch_t::Iteration(double* o_x, double tol) {
  BatchState* stt = o_state.ptr();
  if (int st = launch_batch_update_p(stream, nchunks, o_chunks.ptr(), o_rrRange.ptr(), o_r.ptr(), o_p.ptr(),
                                     o_part_rr.ptr(), stt, tol))
    return st;
  for (const auto& G : groups)
    if (int st = AxGroup(G, G.variant, o_p.ptr(), o_AqL.ptr())) return st;
  if (int st = launch_batch_gather(stream, nsupers, o_supers.ptr(), o_rowStarts.ptr(), o_colIds.ptr(), o_AqL.ptr(),
                                   o_Ap.ptr(), o_p.ptr(), o_part_pAp.ptr(), stt))
    return st;
  return launch_batch_update_xr(stream, nchunks, o_chunks.ptr(), o_pApRange.ptr(), o_p.ptr(), o_Ap.ptr(), o_x,
                                o_r.ptr(), o_part_pAp.ptr(), o_part_rr.ptr(), stt);
}

int bp_batch_t::Capture(double* o_x, double tol) {
  if (graphExec && graphX == o_x && graphTol == tol) return BPX_OK;
  if (graphExec) {
    (void)hipGraphExecDestroy(graphExec);
    graphExec = nullptr;
  }
  BPX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  int st = BPX_OK;
  for (int k = 0; k < kGraphIters && st == BPX_OK; ++k) st = Iteration(o_x, tol);
  if (st == BPX_OK) st = launch_batch_count_done(stream, nprob, o_state.ptr(), o_count.ptr());
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(stream, &graph);
  if (st) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
  const hipError_t e2 = hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e2 != hipSuccess) {
    graphExec = nullptr;
    return hip_error(e2, "hipGraphInstantiate");
  }
  graphX = o_x;
  graphTol = tol;
  return BPX_OK;
}

int bp_batch_t::Solve(double* o_x, const double* o_b, double tol, int maxit, int* iterations, double* rnorm,
                      int* breakdown) {
  if (maxit < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "batch: maxit < 0");
  if (Ntot > 0 && (!o_x || !o_b)) return set_error(BPX_ERR_INVALID_ARGUMENT, "batch: null vector");
  // order after the caller's work on the platform stream
  BPX_HIP(hipEventRecord(ev0, platform->stream));
  BPX_HIP(hipStreamWaitEvent(stream, ev0, 0));
  if (int st = launch_batch_init(stream, nchunks, o_chunks.ptr(), o_b, o_x, o_r.ptr(), o_part_rr.ptr(),
                                 o_state.ptr(), maxit))
    return st;
  if (int st = Capture(o_x, tol)) return st;
  // every replay runs kGraphIters iterations of the problems still going;
  // maxit/kGraphIters + 1 replays let the last one reach its cap and stop
  const long long maxReplays = (long long)maxit / kGraphIters + 2;
  replays = 0;
  for (long long rep = 0; rep < maxReplays; ++rep) {
    BPX_HIP(hipGraphLaunch(graphExec, stream));
    BPX_HIP(hipMemcpyAsync(h_count, o_count.ptr(), sizeof(int), hipMemcpyDeviceToHost, stream));
    BPX_HIP(hipStreamSynchronize(stream));
    ++replays;
    if (*h_count >= nprob) break;
  }
  std::vector<BatchState> hs(nprob);
  BPX_HIP(hipMemcpy(hs.data(), o_state.ptr(), sizeof(BatchState) * nprob, hipMemcpyDeviceToHost));
  int anyBreak = 0;
  for (int i = 0; i < nprob; ++i) {
    if (iterations) iterations[i] = hs[i].iter;
    if (rnorm) rnorm[i] = hs[i].rnorm;
    if (breakdown) breakdown[i] = hs[i].breakdown;
    anyBreak |= hs[i].breakdown;
  }
  BPX_HIP(hipEventRecord(ev1, stream));
  BPX_HIP(hipStreamWaitEvent(platform->stream, ev1, 0));
  if (anyBreak) return set_error(BPX_ERR_BREAKDOWN, "batch: CG breakdown in at least one problem");
  return BPX_OK;
}

// ---- FP32 CG with FP64 corrections, every problem at once -------------
// The recurrence of bp_t::SolveMixed per problem: FP32 vectors and
// operator (factors rounded to FP32), FP64 dot products, and every
// `interval` iterations x += c and r = float(b - A x) with the FP64
// operator, where the stopping test is made. Problems advance in lockstep
// segments (interval FP32 iterations, then the correction), captured once
// as a hipGraph; a finished problem's elements, rows and chunks return at
// once in every kernel of later segments.

int bp_batch_t::Prepare32() {
  if (have32) return BPX_OK;
  hipStream_t s = stream;
  for (auto& G : groups) {
    int st;
    if ((st = G.wJ32.alloc(G.wJ.length())) || (st = G.ggeo32.alloc(G.ggeo.length())) ||
        (st = G.D32.alloc(G.D.length())))
      return st;
    if ((st = launch_to_f32(s, G.wJ.length(), G.wJ.ptr(), G.wJ32.ptr())) ||
        (st = launch_to_f32(s, G.ggeo.length(), G.ggeo.ptr(), G.ggeo32.ptr())) ||
        (st = launch_to_f32(s, G.D.length(), G.D.ptr(), G.D32.ptr())))
      return st;
  }
  const size_t n = std::max<dlong>(Ntot, 1), nl = std::max<dlong>(NLocalTot, 1);
  int st;
  if ((st = o_r32.alloc(n)) || (st = o_p32.alloc(n)) || (st = o_Ap32.alloc(n)) || (st = o_c32.alloc(n)) ||
      (st = o_AqL32.alloc(nl)) || (st = o_Ax64.alloc(n)) || (st = o_part_true.alloc(std::max(nchunks, 1))) ||
      (st = o_inner.alloc(nprob)))
    return st;
  BPX_HIP(hipStreamSynchronize(s));
  have32 = true;
  return BPX_OK;
}

// interval FP32 iterations, then the FP64 correction and the stopping test
int bp_batch_t::Segment(double* o_x, const double* o_b, double tol, int interval) {
  BatchState* in = o_inner.ptr();
  BatchState* out = o_state.ptr();
  const float lam32 = (float)lambda;
  for (int j = 0; j < interval; ++j) {
    if (int st = launch_bm_update_p(stream, nchunks, o_chunks.ptr(), o_rrRange.ptr(), o_r32.ptr(), o_p32.ptr(),
                                    o_part_rr.ptr(), in, out))
      return st;
    for (const auto& G : groups)
      if (int st = launch_bp5_batch_f32(stream, G.N, (int)G.Nel, G.elementList.ptr(), G.G2L.ptr(), G.wJ32.ptr(),
                                        G.ggeo32.ptr(), G.D32.ptr(), lam32, o_p32.ptr(), o_AqL32.ptr() + G.loff,
                                        G.elemPid.ptr(), in))
        return st;
    if (int st = launch_batch_gather_f32(stream, nsupers, o_supers.ptr(), o_rowStarts.ptr(), o_colIds.ptr(),
                                         o_AqL32.ptr(), o_Ap32.ptr(), o_p32.ptr(), o_part_pAp.ptr(), in))
      return st;
    if (int st = launch_bm_update_xr(stream, nchunks, o_chunks.ptr(), o_pApRange.ptr(), o_p32.ptr(), o_Ap32.ptr(),
                                     o_c32.ptr(), o_r32.ptr(), o_part_pAp.ptr(), o_part_rr.ptr(), in, out))
      return st;
  }
  if (int st = launch_bm_flush(stream, nchunks, o_chunks.ptr(), o_c32.ptr(), o_x, out)) return st;
  for (const auto& G : groups)
    if (int st = launch_bp5_batch(stream, G.N, G.variant, (int)G.Nel, G.elementList.ptr(), G.G2L.ptr(), G.wJ.ptr(),
                                  G.ggeo.ptr(), G.D.ptr(), lambda, o_x, o_AqL.ptr() + G.loff, G.elemPid.ptr(), out))
      return st;
  if (int st = launch_batch_gather_only(stream, nsupers, o_supers.ptr(), o_rowStarts.ptr(), o_colIds.ptr(),
                                        o_AqL.ptr(), o_Ax64.ptr(), out))
    return st;
  if (int st = launch_bm_true_residual(stream, nchunks, o_chunks.ptr(), o_b, o_Ax64.ptr(), o_r32.ptr(),
                                       o_part_true.ptr(), o_part_rr.ptr(), out))
    return st;
  if (int st = launch_bm_decide(stream, nprob, o_rrRange.ptr(), o_part_true.ptr(), in, out, tol, 0)) return st;
  return launch_batch_count_done(stream, nprob, out, o_count.ptr());
}

int bp_batch_t::CaptureMixed(double* o_x, const double* o_b, double tol, int interval) {
  if (graphMixed && graphMixedX == o_x && graphMixedB == o_b && graphMixedTol == tol &&
      graphMixedInterval == interval)
    return BPX_OK;
  if (graphMixed) {
    (void)hipGraphExecDestroy(graphMixed);
    graphMixed = nullptr;
  }
  BPX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  const int st = Segment(o_x, o_b, tol, interval);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(stream, &graph);
  if (st) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
  const hipError_t e2 = hipGraphInstantiate(&graphMixed, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e2 != hipSuccess) {
    graphMixed = nullptr;
    return hip_error(e2, "hipGraphInstantiate");
  }
  graphMixedX = o_x;
  graphMixedB = o_b;
  graphMixedTol = tol;
  graphMixedInterval = interval;
  return BPX_OK;
}

int bp_batch_t::SolveMixed(double* o_x, const double* o_b, double tol, int interval, int maxit, int* iterations,
                           int* updates, double* relres, int* breakdown) {
  if (!(tol > 0) || interval < 1 || maxit < 1)
    return set_error(BPX_ERR_INVALID_ARGUMENT, "batch mixed solve: tol > 0, interval >= 1 and maxit >= 1 required");
  if (Ntot > 0 && (!o_x || !o_b)) return set_error(BPX_ERR_INVALID_ARGUMENT, "batch: null vector");
  if (int st = Prepare32()) return st;
  BPX_HIP(hipEventRecord(ev0, platform->stream));
  BPX_HIP(hipStreamWaitEvent(stream, ev0, 0));
  BatchState* in = o_inner.ptr();
  BatchState* out = o_state.ptr();
  if (int st = launch_bm_init(stream, nchunks, o_chunks.ptr(), o_b, o_x, o_r32.ptr(), o_c32.ptr(),
                              o_part_true.ptr(), o_part_rr.ptr(), in, out, maxit))
    return st;
  if (int st = launch_bm_decide(stream, nprob, o_rrRange.ptr(), o_part_true.ptr(), in, out, tol, 1)) return st;
  if (int st = CaptureMixed(o_x, o_b, tol, interval)) return st;
  // a segment adds up to `interval` FP32 iterations per running problem;
  // once a problem reaches maxit its recurrence pauses and the next
  // correction ends it
  const long long maxReplays = (long long)maxit / interval + 2;
  replays = 0;
  for (long long rep = 0; rep < maxReplays; ++rep) {
    BPX_HIP(hipGraphLaunch(graphMixed, stream));
    BPX_HIP(hipMemcpyAsync(h_count, o_count.ptr(), sizeof(int), hipMemcpyDeviceToHost, stream));
    BPX_HIP(hipStreamSynchronize(stream));
    ++replays;
    if (*h_count >= nprob) break;
  }
  std::vector<BatchState> ho(nprob), hi(nprob);
  BPX_HIP(hipMemcpy(ho.data(), out, sizeof(BatchState) * nprob, hipMemcpyDeviceToHost));
  BPX_HIP(hipMemcpy(hi.data(), in, sizeof(BatchState) * nprob, hipMemcpyDeviceToHost));
  int anyBreak = 0;
  for (int i = 0; i < nprob; ++i) {
    if (iterations) iterations[i] = ho[i].iter;
    if (updates) updates[i] = ho[i].updates;
    if (relres) relres[i] = hi[i].tol2 > 0 ? ho[i].rnorm / std::sqrt(hi[i].tol2) : 0.0;
    if (breakdown) breakdown[i] = ho[i].breakdown;
    anyBreak |= ho[i].breakdown;
  }
  BPX_HIP(hipEventRecord(ev1, stream));
  BPX_HIP(hipStreamWaitEvent(platform->stream, ev1, 0));
  if (anyBreak) return set_error(BPX_ERR_BREAKDOWN, "batch mixed solve: non-finite residual in at least one problem");
  return BPX_OK;
}

}  // namespace bpx

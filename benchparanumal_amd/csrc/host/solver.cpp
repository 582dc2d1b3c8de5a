// platform_t, linAlg_t, cg and bp_t.
#include <cmath>
#include <cstdio>
#include <cstring>

#include <algorithm>

#include <type_traits>

#include "bpx.hpp"

namespace bpx {

// Even-odd tables of a 1D operator A (N x M, row-major) with
// A[N-1-n][M-1-m] = sign * A[n][m]: row n (n <= (N-1)/2) holds
// (A[n][m] + A[n][M-1-m]) / 2 for m < M/2, A[n][M/2] for odd M, then
// (A[n][m] - A[n][M-1-m]) / 2 for m < M/2 (the layout of eo_apply in
// ax_cub.hip). False if A is not (anti)symmetric to rounding.
static bool even_odd_rows(const double* A, int N, int M, int sign, std::vector<double>& out) {
  double amax = 0;
  for (int i = 0; i < N * M; ++i) amax = std::max(amax, std::fabs(A[i]));
  for (int n = 0; n < N; ++n)
    for (int m = 0; m < M; ++m)
      if (std::fabs(A[(N - 1 - n) * M + (M - 1 - m)] - sign * A[n * M + m]) > 1e-12 * amax) return false;
  const int hM = M / 2, hN = N / 2;
  for (int n = 0; n < hN + (N & 1); ++n) {
    for (int m = 0; m < hM; ++m) out.push_back(0.5 * (A[n * M + m] + A[n * M + M - 1 - m]));
    if (M & 1) out.push_back(A[n * M + hM]);
    for (int m = 0; m < hM; ++m) out.push_back(0.5 * (A[n * M + m] - A[n * M + M - 1 - m]));
  }
  return true;
}

// the pencil kernel's tables (BP1-BP4, non-affine hex, N <= 7): I (cubNq x
// Nq), I^T, D and D^T of the cubature grid, one after another; empty where
// they do not apply
static std::vector<double> cub_even_odd_tables(const mesh_t& m) {
  std::vector<double> t;
  const int nq = m.Nq, mq = m.cubNq;
  if (m.elementType != ELEM_HEX || m.opts.affine || nq > 8 || mq != nq + 1 ||
      m.cubInterp.size() != (size_t)mq * nq || m.cubD.size() != (size_t)mq * mq)
    return t;
  std::vector<double> IT((size_t)nq * mq), DT((size_t)mq * mq);
  for (int a = 0; a < mq; ++a)
    for (int i = 0; i < nq; ++i) IT[(size_t)i * mq + a] = m.cubInterp[(size_t)a * nq + i];
  for (int a = 0; a < mq; ++a)
    for (int b = 0; b < mq; ++b) DT[(size_t)b * mq + a] = m.cubD[(size_t)a * mq + b];
  if (!even_odd_rows(m.cubInterp.data(), mq, nq, 1, t) || !even_odd_rows(IT.data(), nq, mq, 1, t) ||
      !even_odd_rows(m.cubD.data(), mq, mq, -1, t) || !even_odd_rows(DT.data(), mq, mq, -1, t))
    t.clear();
  return t;
}


// ------------------------------------------------------------- errors ---
static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

const char* last_error() { return g_last_error.c_str(); }

int hip_error(hipError_t e, const char* what) {
  const int code = (e == hipErrorOutOfMemory) ? BPX_ERR_OUT_OF_MEMORY
                   : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? BPX_ERR_NO_DEVICE
                                                                          : BPX_ERR_HIP;
  return set_error(code, std::string(what) + ": " + hipGetErrorString(e));
}

// ----------------------------------------------------------- platform ---
platform_t::~platform_t() {
  if (device >= 0) (void)hipSetDevice(device);
  if (partials) (void)hipFree(partials);
  if (d_scalar) (void)hipFree(d_scalar);
  if (h_scalar) (void)hipHostFree(h_scalar);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

int platform_t::Create(int dev, platform_t** out) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return set_error(BPX_ERR_NO_DEVICE, "platform: no HIP device visible");
  if (dev < 0 || dev >= count) return set_error(BPX_ERR_INVALID_ARGUMENT, "platform: device index out of range");
  platform_t* p = new platform_t;
  p->device = dev;
  int st = BPX_OK;
  auto fail = [&](int s) {
    delete p;
    return s;
  };
  if ((e = hipSetDevice(dev)) != hipSuccess) return fail(hip_error(e, "hipSetDevice"));
  if ((e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess)
    return fail(hip_error(e, "hipStreamCreate"));
  p->own_stream = true;
  if ((e = hipMalloc(&p->partials, kRedBlocks * sizeof(double))) != hipSuccess)
    return fail(hip_error(e, "hipMalloc(partials)"));
  if ((e = hipMalloc(&p->d_scalar, sizeof(double))) != hipSuccess) return fail(hip_error(e, "hipMalloc"));
  if ((e = hipHostMalloc(&p->h_scalar, sizeof(double), hipHostMallocDefault)) != hipSuccess)
    return fail(hip_error(e, "hipHostMalloc"));
  (void)st;
  *out = p;
  return BPX_OK;
}

int platform_t::finish() {
  BPX_HIP(hipStreamSynchronize(stream));
  return BPX_OK;
}

int platform_t::setStream(hipStream_t s) {
  if (int st = finish()) return st;
  if (own_stream && stream) (void)hipStreamDestroy(stream);
  own_stream = false;
  stream = s;
  return BPX_OK;
}

// ------------------------------------------------------------- linAlg ---
int linAlg_t::set(dlong N, dfloat alpha, dfloat* o_x) { return launch_set<double>(platform.stream, N, alpha, o_x); }

int linAlg_t::axpy(dlong N, dfloat alpha, const dfloat* o_x, dfloat beta, dfloat* o_y) {
  return launch_axpy<double>(platform.stream, N, alpha, o_x, beta, o_y);
}

int linAlg_t::innerProd(dlong N, const dfloat* o_x, const dfloat* o_y, dfloat* result) {
  if (int st = launch_dot<double>(platform.stream, N, o_x, o_y, platform.partials, platform.d_scalar))
    return st;
  BPX_HIP(hipMemcpyAsync(platform.h_scalar, platform.d_scalar, sizeof(double), hipMemcpyDeviceToHost,
                         platform.stream));
  if (int st = platform.finish()) return st;
  *result = *platform.h_scalar;
  return BPX_OK;
}

int linAlg_t::norm2(dlong N, const dfloat* o_x, dfloat* result) {
  double d;
  if (int st = innerProd(N, o_x, o_x, &d)) return st;
  *result = std::sqrt(d);
  return BPX_OK;
}

// ----------------------------------------------------------------- cg ---
template <typename T>
int cg_t<T>::Setup(platform_t& p, dlong N_, int npart_max) {
  platform = &p;
  N = N_;
  if (int st = o_p.alloc(N)) return st;
  if (int st = o_Ap.alloc(N)) return st;
  if (int st = o_state.alloc(1)) return st;
  if (int st = o_part_rr.alloc(kRedBlocks)) return st;
  if (int st = o_part_pAp.alloc(std::max(npart_max, kRedBlocks))) return st;
  if (N > 0) {
    BPX_HIP(hipMemsetAsync(o_p.ptr(), 0, N * sizeof(T), p.stream));
    BPX_HIP(hipMemsetAsync(o_Ap.ptr(), 0, N * sizeof(T), p.stream));
  }
  return BPX_OK;
}

// cg::Solve (libs/core/linearSolverCG.cpp:62-129). The host queues
// iterations without waiting; with tol > 0 it reads the device stop flag
// after every chunk of 16 iterations, so a converged solve stops early.
template <typename T>
int cg_t<T>::Solve(const Apply& apply, const ApplyDot& apply_dot, T* o_x, T* o_r, double tol, int MAXIT,
                   int* iterations, const ApplyFold& apply_fold, dlong fold_rows) {
  if (MAXIT < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "cg: MAXIT < 0");
  if (hist_len < MAXIT + 1) {
    if (int st = o_hist_rr.alloc(MAXIT + 1)) return st;
    if (int st = o_hist_alpha.alloc(MAXIT + 1)) return st;
    hist_len = MAXIT + 1;
  }
  CgWork<T> w;
  w.p = o_p.ptr();
  w.Ap = o_Ap.ptr();
  w.state = o_state.ptr();
  w.part_rr = o_part_rr.ptr();
  w.part_pAp = o_part_pAp.ptr();
  w.hist_rr = o_hist_rr.ptr();
  w.hist_alpha = o_hist_alpha.ptr();
  return cg_run<T>(platform->stream, N, w, apply, apply_dot, o_x, o_r, tol, MAXIT, iterations, apply_fold,
                   fold_rows);
}

template <typename T>
int cg_run(hipStream_t s, dlong N, const CgWork<T>& w, const typename cg_t<T>::Apply& apply,
           const typename cg_t<T>::ApplyDot& apply_dot, T* o_x, T* o_r, double tol, int MAXIT, int* iterations,
           const typename cg_t<T>::ApplyFold& apply_fold, dlong fold_rows) {
  if (MAXIT < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "cg: MAXIT < 0");
  CgState* st_d = w.state;
  double* prr = w.part_rr;
  double* ppAp = w.part_pAp;
  // r <- b - A x (linearSolverCG.cpp:79-82)
  if (int st = apply(o_x, w.Ap)) return st;
  if (int st = launch_cg_init<T>(s, N, w.Ap, o_r, prr)) return st;
  CgState hst;
  if (MAXIT == 0) {  // only the initial residual and the stopping test
    if (int st = launch_cg_update_p<T>(s, N, 0, o_r, w.p, o_x, prr, kRedBlocks, st_d, w.hist_rr, w.hist_alpha,
                                       tol, false))
      return st;
  }
  const int chunk = (tol > 0) ? 16 : std::max(MAXIT, 1);
  int queued = 0;
  while (queued < MAXIT) {
    const int n = std::min(chunk, MAXIT - queued);
    for (int k = queued; k < queued + n; ++k) {
      int npart = kRedBlocks;
      if (apply_fold) {
        // the direction update of the rows below fold_rows here, of the
        // others inside the operator
        if (int st = launch_cg_update_p<T>(s, fold_rows, k, o_r, w.p, o_x, prr, kRedBlocks, st_d, w.hist_rr,
                                           w.hist_alpha, tol, k > 0))
          return st;
        CgFold<T> f;
        f.on = 1;
        f.k = k;
        f.pending = k > 0;
        f.p = w.p;
        f.r = o_r;
        f.x = o_x;
        f.hist_rr = w.hist_rr;
        f.hist_alpha = w.hist_alpha;
        f.st = st_d;
        if (int st = apply_fold(f, w.Ap, ppAp, &npart)) return st;
        if (int st = launch_cg_update_r<T>(s, N, k, w.Ap, o_r, ppAp, npart, prr, st_d, w.hist_rr, w.hist_alpha))
          return st;
        continue;
      }
      // also x += alpha_{k-1} p, the x update of iteration k-1
      if (int st = launch_cg_update_p<T>(s, N, k, o_r, w.p, o_x, prr, kRedBlocks, st_d, w.hist_rr, w.hist_alpha,
                                         tol, k > 0))
        return st;
      if (apply_dot) {
        if (int st = apply_dot(w.p, w.Ap, ppAp, &npart)) return st;
      } else {
        if (int st = apply(w.p, w.Ap)) return st;
        if (int st = launch_dot_partials<T>(s, N, w.p, w.Ap, ppAp)) return st;
      }
      if (int st = launch_cg_update_r<T>(s, N, k, w.Ap, o_r, ppAp, npart, prr, st_d, w.hist_rr, w.hist_alpha))
        return st;
    }
    queued += n;
    if (int st = launch_cg_finish(s, queued, MAXIT, prr, kRedBlocks, st_d, w.hist_rr)) return st;
    if (queued < MAXIT) {
      BPX_HIP(hipMemcpyAsync(&hst, st_d, sizeof(CgState), hipMemcpyDeviceToHost, s));
      BPX_HIP(hipStreamSynchronize(s));
      if (hst.done) break;
    }
  }
  BPX_HIP(hipMemcpyAsync(&hst, st_d, sizeof(CgState), hipMemcpyDeviceToHost, s));
  BPX_HIP(hipStreamSynchronize(s));
  // stopped after its last queued iteration (converged there, or MAXIT):
  // that iteration's x update is still to be applied
  if (MAXIT > 0 && hst.done && !hst.breakdown && hst.iter == queued) {
    if (int st = launch_cg_flush_x<T>(s, N, queued - 1, w.p, o_x, w.hist_alpha, st_d, false)) return st;
    BPX_HIP(hipStreamSynchronize(s));
  }
  if (iterations) *iterations = (MAXIT == 0) ? 0 : hst.iter;
  if (MAXIT > 0 && hst.breakdown) return set_error(BPX_ERR_BREAKDOWN, "cg: breakdown (non-finite or non-positive p.Ap / r.r)");
  return BPX_OK;
}

template <typename T>
int cg_t<T>::History(std::vector<double>& rr, std::vector<double>& alpha) {
  CgState hst;
  BPX_HIP(hipMemcpy(&hst, o_state.ptr(), sizeof(CgState), hipMemcpyDeviceToHost));
  rr.assign(hst.iter + 1, 0.0);
  alpha.assign(hst.iter, 0.0);
  if (hist_len == 0) return BPX_OK;
  BPX_HIP(hipMemcpy(rr.data(), o_hist_rr.ptr(), rr.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (hst.iter > 0)
    BPX_HIP(hipMemcpy(alpha.data(), o_hist_alpha.ptr(), alpha.size() * sizeof(double), hipMemcpyDeviceToHost));
  return BPX_OK;
}

template class cg_t<double>;
template class cg_t<float>;
template int cg_run<double>(hipStream_t, dlong, const CgWork<double>&, const cg_t<double>::Apply&,
                            const cg_t<double>::ApplyDot&, double*, double*, double, int, int*,
                            const cg_t<double>::ApplyFold&, dlong);
template int cg_run<float>(hipStream_t, dlong, const CgWork<float>&, const cg_t<float>::Apply&,
                           const cg_t<float>::ApplyDot&, float*, float*, double, int, int*,
                           const cg_t<float>::ApplyFold&, dlong);

cg::cg(platform_t& p, dlong N, dlong Nhalo, int npart_max) { status_ = impl.Setup(p, N + Nhalo, npart_max); }

int cg::Solve(solver_t& solver, dfloat* o_x, dfloat* o_r, dfloat tol, int MAXIT, int verbose, int* iterations) {
  if (status_) return status_;
  int its = 0;
  cg_t<double>::ApplyDot fused;
  if (solver.HasOperatorDot())
    fused = [&](const double* q, double* Aq, double* part, int* n) { return solver.OperatorDot(q, Aq, part, n); };
  cg_t<double>::ApplyFold fold;
  const dlong foldRows = solver.FoldRows();
  if (foldRows > 0)
    fold = [&](const CgFold<double>& f, double* Aq, double* part, int* n) {
      return solver.OperatorDotFold(f, Aq, part, n);
    };
  const int st = impl.Solve([&](const double* q, double* Aq) { return solver.Operator(q, Aq); }, fused, o_x, o_r,
                            tol, MAXIT, &its, fold, foldRows);
  if (iterations) *iterations = its;
  if (verbose) {
    std::vector<double> rr, al;
    if (impl.History(rr, al) == BPX_OK) {
      printf("CG: initial res norm %12.12f \n", std::sqrt(rr[0]));
      for (int it = 1; it < (int)rr.size(); ++it)
        printf("CG: it %d, r norm %12.12le, alpha = %le \n", it, std::sqrt(rr[it]), al[it - 1]);
    }
  }
  return st;
}

// ----------------------------------------------------------------- BP ---
int bp_t::Setup(platform_t& p, mesh_t& m, int bp_, dfloat lambda_) {
  if (bp_ < 1 || bp_ > 6) return set_error(BPX_ERR_INVALID_ARGUMENT, "bp: bp must be in 1..6");
  const bool needCub = bp_ <= 4;
  const bool simplexMesh = m.elementType == ELEM_TRI || m.elementType == ELEM_TET;
  if (needCub && !simplexMesh && m.cubD.empty())
    return set_error(BPX_ERR_INVALID_ARGUMENT, "bp: BP1-BP4 need a mesh built with cubature=1");
  platform = &p;
  mesh = &m;
  bp = bp_;
  lambda = lambda_;
  Nfields = (bp == 2 || bp == 4 || bp == 6) ? m.dim : 1;
  if (m.elementType != ELEM_HEX && m.onDevice)
    return set_error(BPX_ERR_UNSUPPORTED, "bp: the device setup covers hexahedra");
  BPX_HIP(hipSetDevice(p.device));
  hipStream_t s = p.stream;

  // masked ogs + GlobalToLocal (BP/BP5/src/bp5Setup.cpp:43-52): on the
  // device for a mesh set up there, else on the host
  NLocal = (dlong)((size_t)m.Nelements * m.Np * Nfields);
  GlobalToLocal.clear();
  if (m.onDevice) {
    deviceBuffer<hlong> ids;
    if (int st = ids.alloc(NLocal)) return st;
    if (int st = m.MaskedGlobalIdsDevice(Nfields, ids.ptr())) return st;
    if (int st = o_GlobalToLocal.alloc(NLocal)) return st;
    if (int st = ogs.SetupDevice(p, NLocal, ids.ptr(), o_GlobalToLocal.ptr())) return st;
  } else {
    std::vector<hlong> ids;
    m.MaskedGlobalIds(Nfields, ids);
    if (int st = ogs.Setup(NLocal, ids.data())) return st;
    GlobalToLocal.resize(NLocal);
    ogs.SetupGlobalToLocalMapping(GlobalToLocal.data());
    if (int st = ogs.Upload(p)) return st;
    if (int st = o_GlobalToLocal.upload(GlobalToLocal, s)) return st;
  }
  if (int st = gHalo.SetupFromGather(ogs)) return st;
  std::vector<dlong> list(m.Nelements);
  for (dlong e = 0; e < m.Nelements; ++e) list[e] = e;
  if (int st = o_elementList.upload(list, s)) return st;
  if (int st = o_AqL.alloc(NLocal)) return st;

  // operator data (BP/BPn/src/bpnOperator.cpp argument lists); the factors
  // of a device mesh are copied on the device
  const int affine = m.opts.affine;
  auto factors = [&](deviceBuffer<double>& dst, const std::vector<double>& h, const deviceBuffer<double>& d) {
    return m.onDevice ? dst.copy_device(d.ptr(), d.length(), s) : dst.upload(h, s);
  };
  if (simplexMesh) {  // per-element J and G, the reference matrices
    if (int st = o_wJ.upload(m.wJ, s)) return st;
    if (int st = o_ggeo.upload(m.ggeo, s)) return st;
    if (int st = o_MM.upload(m.MM, s)) return st;
    // the GEMM form: Sall column-major with the NG stiffness blocks
    // A_c[n][k] = S_c[k][n] stacked over MM (MM alone for BP1/BP2)
    const int np = m.Np, ng = m.dim == 2 ? 3 : 6, blocks = bp >= 3 ? ng + 1 : 1;
    const size_t M = (size_t)blocks * np, NN = (size_t)np * np;
    std::vector<double> sall(M * np);
    for (int k = 0; k < np; ++k)
      for (int n = 0; n < np; ++n) {
        for (int c = 0; c + 1 < blocks; ++c) sall[(size_t)c * np + n + k * M] = m.Ssim[c * NN + (size_t)k * np + n];
        sall[(size_t)(blocks - 1) * np + n + k * M] = m.MM[(size_t)k * np + n];
      }
    const size_t cols = (size_t)m.Nelements * Nfields;
    void* h = nullptr;
    if (int st = simplex_blas_create(s, &h)) return st;
    blas = std::shared_ptr<void>(h, simplex_blas_destroy);
    if (int st = o_Sall.upload(sall, s)) return st;
    if (int st = o_Qs.alloc(std::max<size_t>(1, cols * np))) return st;
    if (int st = o_Zs.alloc(std::max<size_t>(1, cols * M))) return st;
  } else if (m.elementType == ELEM_QUAD) {  // per-node factors; the Gauss forms for BP1-BP4
    if (bp == 5 || bp == 6) {
      if (int st = o_wJ.upload(m.wJ, s)) return st;
      if (int st = o_ggeo.upload(m.ggeo, s)) return st;
      if (int st = o_D.upload(m.D, s)) return st;
    } else {
      if (int st = o_wJ.upload(m.cubwJ, s)) return st;
      if (int st = o_ggeo.upload(m.cubggeo, s)) return st;
      if (int st = o_D.upload(m.cubD, s)) return st;
      if (int st = o_I.upload(m.cubInterp, s)) return st;
    }
  } else if (bp == 5 || bp == 6) {
    if (int st = factors(o_wJ, m.wJ, m.o_wJ)) return st;
    if (int st = factors(o_ggeo, m.ggeo, m.o_ggeo)) return st;
    if (int st = o_gllw.upload(m.gllw, s)) return st;
    if (int st = o_D.upload(m.D, s)) return st;
  } else if (bp == 1 || bp == 2) {
    if (int st = affine ? factors(o_wJ, m.wJ, m.o_wJ) : factors(o_wJ, m.cubwJ, m.o_cubwJ)) return st;
    if (int st = o_I.upload(m.cubInterp, s)) return st;
    if (int st = o_MM.upload(m.MM, s)) return st;
    const std::vector<double> eo = cub_even_odd_tables(m);
    if (!eo.empty())
      if (int st = o_eo.upload(eo, s)) return st;
  } else {
    if (int st = affine ? factors(o_wJ, m.wJ, m.o_wJ) : factors(o_wJ, m.cubwJ, m.o_cubwJ)) return st;
    if (int st = affine ? factors(o_ggeo, m.ggeo, m.o_ggeo) : factors(o_ggeo, m.cubggeo, m.o_cubggeo)) return st;
    if (int st = o_D.upload(m.cubD, s)) return st;
    if (int st = o_I.upload(m.cubInterp, s)) return st;
    if (affine)
      if (int st = o_invV.upload(m.invV, s)) return st;
    const std::vector<double> eo = cub_even_odd_tables(m);
    if (!eo.empty())
      if (int st = o_eo.upload(eo, s)) return st;
  }
  if (int st = m.onDevice ? PrepareFusedDevice() : PrepareFused()) return st;
  if (int st = Tune()) return st;
  return p.finish();
}

template <typename T>
int bp_t::ax(const T* q, T* AqL, const T* wJ, const T* ggeo, const T* gllw, const T* D, const T* I, const T* MM,
             const T* invV, const dlong* G2L) {
  AxArgs<T> a;
  a.N = mesh->N;
  a.Nelements = mesh->Nelements;
  a.Nfields = Nfields;
  a.affine = haveKappa ? 0 : mesh->opts.affine;  // kappa-scaled factors are per node
  a.elementList = o_elementList.ptr();
  a.GlobalToLocal = G2L ? G2L : o_GlobalToLocal.ptr();
  a.wJ = wJ;
  a.ggeo = ggeo;
  a.gllw = gllw;
  a.D = D;
  a.I = I;
  a.MM = MM;
  a.invV = invV;
  a.lambda = (T)lambda;
  a.q = q;
  a.Aq = AqL;
  if constexpr (std::is_same<T, double>::value) a.engine = axEngine;
  a.cubForm = cubForm;
  if (bp <= 4 && !haveKappa) {
    if constexpr (std::is_same<T, double>::value)
      a.eo = o_eo.length() ? o_eo.ptr() : nullptr;
    else
      a.eo = o_eo32.length() ? o_eo32.ptr() : nullptr;
  }
  if constexpr (std::is_same<T, float>::value) a.tri = tri32;
  a.evStart = axEvStart;
  a.evStop = axEvStop;
  if (mesh->elementType != ELEM_HEX) {
    if constexpr (std::is_same<T, double>::value) {
      if (mesh->elementType == ELEM_QUAD)
        return launch_quad_ax(platform->stream, bp, mesh->N, Nfields, mesh->Nelements, a.elementList, a.GlobalToLocal,
                              wJ, ggeo, D, I, lambda, q, AqL);
      return launch_simplex_ax_gemm(platform->stream, blas.get(), bp, mesh->dim, mesh->Np, Nfields, mesh->Nelements,
                                    a.elementList, a.GlobalToLocal, wJ, ggeo, o_Sall.ptr(), o_Qs.ptr(),
                                    o_Zs.ptr(), lambda, q, AqL);
    } else {
      return set_error(BPX_ERR_UNSUPPORTED, "bp: the FP32 operators cover hexahedra");
    }
  }
  if (bp == 5 || bp == 6) return launch_bp5_ax<T>(platform->stream, a);
  if (bp == 1 || bp == 2) return launch_bp1_ax<T>(platform->stream, a);
  return launch_bp3_ax<T>(platform->stream, a);
}

int bp_t::AxLocal(const dfloat* o_q, dfloat* o_AqL, const dlong* G2L) {
  if (haveKappa)
    return ax<double>(o_q, o_AqL, mesh->opts.affine ? o_wJK.ptr() : o_wJ.ptr(), o_ggeoK.ptr(), o_gllw.ptr(),
                      o_D.ptr(), nullptr, nullptr, nullptr, G2L);
  return ax<double>(o_q, o_AqL, o_wJ.ptr(), o_ggeo.ptr(), o_gllw.ptr(), o_D.ptr(), o_I.ptr(), o_MM.ptr(),
                    o_invV.ptr(), G2L);
}

hlong bp_t::LatticeNodes() const {
  const MeshOpts& o = mesh->opts;
  return ((hlong)o.nx * mesh->N + 1) * ((hlong)o.ny * mesh->N + 1) * ((hlong)o.nz * mesh->N + 1);
}

int bp_t::SetKappa(const dfloat* o_kappa) {
  if (!o_kappa) {
    haveKappa = false;
    return BPX_OK;
  }
  if (bp != 5 || mesh->elementType != ELEM_HEX)
    return set_error(BPX_ERR_UNSUPPORTED, "bp: a variable coefficient is implemented for BP5 on hexahedra");
  hipStream_t s = platform->stream;
  const mesh_t& m = *mesh;
  const hlong Nnodes = LatticeNodes();
  if (Nnodes >= ((hlong)1 << 31)) return set_error(BPX_ERR_UNSUPPORTED, "bp: lattice above 2^31 nodes");
  if (!o_lattice.ptr()) {
    // lattice point of each local node: the mesh's global ids are 1 + lattice index
    if (int st = mesh->EnsureHost()) return st;
    std::vector<dlong> lat(NLocal);
    for (dlong l = 0; l < NLocal; ++l) lat[l] = (dlong)(m.globalIds[l] - 1);
    if (int st = o_lattice.upload(lat, s)) return st;
    if (int st = latticeOgs.SetupRows(NLocal, (dlong)Nnodes, lat.data())) return st;
    if (int st = latticeOgs.Upload(*platform)) return st;
    if (int st = o_ggeoK.alloc((size_t)NLocal * 6)) return st;
    if (m.opts.affine)
      if (int st = o_wJK.alloc(NLocal)) return st;
    if (int st = o_gradL.alloc(NLocal)) return st;
  }
  if (int st = launch_kappa_geometry(s, NLocal, m.Nq, m.opts.affine, o_lattice.ptr(), o_kappa, o_ggeo.ptr(),
                                     o_wJ.ptr(), o_gllw.ptr(), o_ggeoK.ptr(), o_wJK.ptr()))
    return st;
  haveKappa = true;
  return BPX_OK;
}

int bp_t::KappaGradient(const dfloat* o_u, const dfloat* o_lam, dfloat* o_grad) {
  if (bp != 5) return set_error(BPX_ERR_UNSUPPORTED, "bp: the kappa gradient is implemented for BP5");
  if (!o_lattice.ptr()) return set_error(BPX_ERR_INVALID_ARGUMENT, "bp: set a coefficient (kappa) first");
  hipStream_t s = platform->stream;
  if (int st = launch_bp5_kappa_grad(s, mesh->N, mesh->Nelements, mesh->opts.affine, o_GlobalToLocal.ptr(),
                                     o_ggeo.ptr(), o_gllw.ptr(), o_D.ptr(), o_u, o_lam, o_gradL.ptr()))
    return st;
  return latticeOgs.Gather<double>(*platform, o_grad, o_gradL.ptr(), 1);
}

// bpN_t::Operator (BP/BP5/src/bp5Operator.cpp:29-85). On one GPU the halo
// exchanges move nothing, so the two element halves become one launch.
int bp_t::Operator(const dfloat* o_q, dfloat* o_Aq) {
  // the reference splits Ax into the elements that need no ghost values
  // and those that do, around the halo exchange (bp5Operator.cpp:31-84);
  // with no halo one launch covers every element
  if (int st = gHalo.ExchangeStart(o_q, 1)) return st;
  if (int st = gHalo.ExchangeFinish(o_q, 1)) return st;
  if (int st = AxLocal(o_q, o_AqL.ptr())) return st;
  return ogs.Gather<double>(*platform, o_Aq, o_AqL.ptr(), 1);
}

// The operator with the q.Aq partials on reference-order vectors: through
// the solve space when the rows are renumbered (the Aq rows are the same
// sums either way, bit for bit).
int bp_t::OperatorDot(const dfloat* o_q, dfloat* o_Aq, double* part, int* npart) {
  if (!permuted) return OperatorDotUnfusedS(o_q, o_Aq, part, npart);
  if (int st = ToSolve(o_q, o_sx.ptr())) return st;
  if (int st = OperatorDotS(o_sx.ptr(), o_sr.ptr(), part, npart)) return st;
  return FromSolve(o_sr.ptr(), o_Aq);
}

int bp_t::ToSolve(const dfloat* o_ref, dfloat* o_s) {
  if (!permuted) {
    if (o_s != o_ref)
      BPX_HIP(hipMemcpyAsync(o_s, o_ref, (size_t)Ns * sizeof(double), hipMemcpyDeviceToDevice, platform->stream));
    return BPX_OK;
  }
  return launch_permute(platform->stream, Ns, o_sperm.ptr(), o_ref, o_s, false);
}

int bp_t::FromSolve(const dfloat* o_s, dfloat* o_ref) {
  if (!permuted) {
    if (o_s != o_ref)
      BPX_HIP(hipMemcpyAsync(o_ref, o_s, (size_t)Ns * sizeof(double), hipMemcpyDeviceToDevice, platform->stream));
    return BPX_OK;
  }
  return launch_permute(platform->stream, Ns, o_sperm.ptr(), o_s, o_ref, true);
}

int bp_t::OperatorS(const dfloat* o_q, dfloat* o_Aq) {
  if (int st = AxLocal(o_q, o_AqL.ptr(), SolveG2L())) return st;
  const ogs::ogs_t& G = SolveOgs();
  return launch_ogs_gather<double>(platform->stream, G.NrowBlocksT, 1, G.o_blockRowStartsT.ptr(),
                                   G.o_rowStartsT.ptr(), G.o_colIdsT.ptr(), o_AqL.ptr(), o_Aq, ogs.gatherMode);
}

int bp_t::OperatorDotS(const dfloat* o_q, dfloat* o_Aq, double* part, int* npart) {
  if (fused && tuning.fused64) return OperatorDotFused<double>(o_q, o_Aq, part, npart);
  return OperatorDotUnfusedS(o_q, o_Aq, part, npart);
}

int bp_t::OperatorDotUnfusedS(const dfloat* o_q, dfloat* o_Aq, double* part, int* npart) {
  if (int st = AxLocal(o_q, o_AqL.ptr(), SolveG2L())) return st;
  const ogs::ogs_t& G = SolveOgs();
  *npart = ogs_gather_partials(G.NrowBlocksT);
  return launch_ogs_gather_dot<double>(platform->stream, G.NrowBlocksT, G.o_blockRowStartsT.ptr(),
                                       G.o_rowStartsT.ptr(), G.o_colIdsT.ptr(), o_AqL.ptr(), o_Aq, o_q, part,
                                       ogs.gatherMode);
}

int bp_t::SolveCG(cg& solver, dfloat* o_x, dfloat* o_r, dfloat tol, int maxit, int verbose, int* iterations) {
  if (!permuted) return solver.Solve(*this, o_x, o_r, tol, maxit, verbose, iterations);
  if (int st = ToSolve(o_x, o_sx.ptr())) return st;
  if (int st = ToSolve(o_r, o_sr.ptr())) return st;
  bp_solve_view_t view(*this);
  const int st = solver.Solve(view, o_sx.ptr(), o_sr.ptr(), tol, maxit, verbose, iterations);
  if (int s2 = FromSolve(o_sx.ptr(), o_x)) return s2;
  if (int s2 = FromSolve(o_sr.ptr(), o_r)) return s2;
  return st;
}

// The solve space and the fused operator's maps (BP5, one field), built
// when every element is regular: its interior nodes are the only copies of
// their rows and each surface node is masked or shared (true of the box
// mesh, all of whose boundary is masked). Every row's gather lists its
// copies in ascending local order, as the reference maps do, and an
// interior row's sum is its one value, so each row of Aq is the same sum in
// either order and the fused operator equals the unfused one bit for bit.
int bp_t::PrepareFused() {
  fused = false;
  permuted = false;
  Ns = ogs.Ngather;
  if (bp != 5 || Nfields != 1 || mesh->elementType != ELEM_HEX) return BPX_OK;
  const dlong Ng = ogs.Ngather, Nel = mesh->Nelements;
  const int Nq = mesh->Nq, Np = mesh->Np;
  const long long NIP = fused_interior_pad(Nq), NSP = fused_surface_pad(Nq);
  std::vector<char> shared(Ng);
  for (dlong r = 0; r < Ng; ++r) shared[r] = (ogs.rowStartsT[r + 1] - ogs.rowStartsT[r]) >= 2;
  for (dlong e = 0; e < Nel; ++e)
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const dlong row = GlobalToLocal[(size_t)e * Np + (k * Nq + j) * Nq + i];
          const bool interior = fused_surface_index(Nq, i, j, k) < 0;
          if (interior ? (row < 0 || shared[row]) : (row >= 0 && !shared[row])) return BPX_OK;
        }
  std::vector<dlong> srow(Ng, -1);
  dlong Nshared = 0;
  for (dlong r = 0; r < Ng; ++r)
    if (shared[r]) srow[r] = Nshared++;
  const long long S0 = fused_pad(Nshared);
  const long long ns = S0 + NIP * Nel, nslots = NSP * Nel;
  if (ns >= (1LL << 31) - 1 || nslots >= (1LL << 31) - 1) return BPX_OK;  // 32-bit row and slot numbers
  std::vector<dlong> G2Ls(NLocal), rowOfSlot(nslots, -1);
  for (dlong e = 0; e < Nel; ++e)
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const size_t l = (size_t)e * Np + (k * Nq + j) * Nq + i;
          const dlong row = GlobalToLocal[l];
          const int sp = fused_surface_index(Nq, i, j, k);
          if (sp < 0) {
            srow[row] = (dlong)(S0 + NIP * e + fused_interior_index(Nq, i, j, k));
          } else if (row >= 0) {
            rowOfSlot[(size_t)e * NSP + sp] = srow[row];  // a shared row: solve rows 0..Nshared-1
          }
        }
  Ns = (dlong)ns;
  std::vector<dlong> perm(Ns, -1);
  for (dlong r = 0; r < Ng; ++r) perm[srow[r]] = r;
  for (size_t l = 0; l < (size_t)NLocal; ++l) G2Ls[l] = GlobalToLocal[l] < 0 ? -1 : srow[GlobalToLocal[l]];
  if (int st = solveOgs.SetupRows(NLocal, Ns, G2Ls.data())) return st;
  if (int st = solveOgs.Upload(*platform)) return st;
  if (int st = sharedOgs.SetupRows((dlong)nslots, Nshared, rowOfSlot.data())) return st;
  if (int st = sharedOgs.Upload(*platform)) return st;
  hipStream_t s = platform->stream;
  if (int st = o_sperm.upload(perm, s)) return st;
  if (int st = o_G2Ls.upload(G2Ls, s)) return st;
  if (int st = o_slots.alloc(nslots)) return st;
  if (int st = o_sx.alloc(Ns)) return st;
  if (int st = o_sr.alloc(Ns)) return st;
  naxParts = bp5_workgroups(mesh->N, Nel);
  if (int st = o_partAx.alloc(std::max(naxParts, 1))) return st;
  fusedS0 = (dlong)S0;
  permuted = true;
  fused = true;
  return BPX_OK;
}

dlong bp_t::FoldRows(int precision) const {
  if (!(cgFold & (precision == 64 ? 1 : 2)) || !fused || !permuted) return 0;
  if (precision == 64 ? !tuning.fused64 : !tuning.fused32) return 0;
  // the FP64 operator at N = 15 runs on the matrix-core kernel, which keeps
  // the separate update kernel
  if (precision == 64 && mesh->N == 15 && axEngine != AX_ENGINE_VECTOR) return 0;
  return fusedS0;
}

template <typename T>
int bp_t::OperatorDotFused(const T* o_q, T* o_Aq, double* part, int* npart, const CgFold<T>* fold) {
  hipStream_t s = platform->stream;
  AxArgs<T> a;
  if (fold) a.fold = *fold;
  a.N = mesh->N;
  a.Nelements = mesh->Nelements;
  a.Nfields = 1;
  a.elementList = o_elementList.ptr();
  a.GlobalToLocal = o_G2Ls.ptr();
  a.lambda = (T)lambda;
  a.q = o_q;
  a.ldq = fusedS0;  // the first interior row
  T* AqL;
  if constexpr (std::is_same<T, double>::value) {
    a.affine = haveKappa ? 0 : mesh->opts.affine;
    a.wJ = (haveKappa && mesh->opts.affine) ? o_wJK.ptr() : o_wJ.ptr();
    a.ggeo = haveKappa ? o_ggeoK.ptr() : o_ggeo.ptr();
    a.gllw = o_gllw.ptr();
    a.D = o_D.ptr();
    a.engine = axEngine;
    AqL = o_slots.ptr();
  } else {
    if (haveKappa) return set_error(BPX_ERR_UNSUPPORTED, "bp: the FP32 operator has no variable coefficient");
    if (int st = Prepare32()) return st;
    a.affine = mesh->opts.affine;
    a.tri = tri32;
    a.wJ = o_wJ32.ptr();
    a.ggeo = o_ggeo32.ptr();
    a.gllw = o_gllw32.ptr();
    a.D = o_D32.ptr();
    AqL = o_slots32.ptr();
  }
  a.Aq = AqL;
  a.evStart = axEvStart;
  a.evStop = axEvStop;
  if (int st = launch_bp5_fused<T>(s, a, o_Aq, o_partAx.ptr())) return st;
  if (sharedOgs.NrowBlocksT > 0) {
    *npart = ogs_gather_partials(sharedOgs.NrowBlocksT);
    return launch_ogs_gather_dot_fold<T>(s, sharedOgs.NrowBlocksT, sharedOgs.o_blockRowStartsT.ptr(),
                                         sharedOgs.o_rowStartsT.ptr(), sharedOgs.o_colIdsT.ptr(), AqL, o_Aq, o_q,
                                         part, o_partAx.ptr(), naxParts, ogs.gatherMode);
  }
  *npart = 1;  // no shared rows (a single element): the Ax partials are the whole dot
  return launch_fold(s, o_partAx.ptr(), naxParts, part);
}

template int bp_t::OperatorDotFused<double>(const double*, double*, double*, int*, const CgFold<double>*);
template int bp_t::OperatorDotFused<float>(const float*, float*, double*, int*, const CgFold<float>*);

int bp_t::AxLocalKrhs(int K, const dfloat* o_q, dfloat* o_AqL, int chunk) {
  if (bp != 5 || mesh->elementType != ELEM_HEX)
    return set_error(BPX_ERR_UNSUPPORTED, "bp: k right-hand sides are implemented for BP5 on hexahedra");
  if (haveKappa) return set_error(BPX_ERR_UNSUPPORTED, "bp: k right-hand sides use the constant coefficient");
  AxArgs<double> a;
  a.N = mesh->N;
  a.Nelements = mesh->Nelements;
  a.affine = mesh->opts.affine;
  a.elementList = o_elementList.ptr();
  a.GlobalToLocal = o_GlobalToLocal.ptr();
  a.wJ = o_wJ.ptr();
  a.ggeo = o_ggeo.ptr();
  a.gllw = o_gllw.ptr();
  a.D = o_D.ptr();
  a.lambda = lambda;
  a.q = o_q;
  a.Aq = o_AqL;
  a.K = K;
  a.ldq = ogs.Ngather;
  a.ldAq = NLocal;
  a.evStart = axEvStart;
  a.evStop = axEvStop;
  return launch_bp5_ax_krhs<double>(platform->stream, a, chunk);
}

int bp_t::TimeKrhs(int K, int chunk, int launches, const dfloat* o_q, double* out) {
  if (K < 1 || launches < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "time k-RHS: K and launches must be >= 1");
  hipStream_t s = platform->stream;
  const size_t ng = (size_t)std::max<dlong>(ogs.Ngather, 1);
  deviceBuffer<double> qz, A;
  if (!o_q) {  // zeros (a kernel's power, and so its clock, depends on the data: pass q)
    if (int st = qz.alloc(ng * K)) return st;
    BPX_HIP(hipMemsetAsync(qz.ptr(), 0, ng * K * sizeof(double), s));
    o_q = qz.ptr();
  }
  if (int st = A.alloc((size_t)NLocal * K)) return st;
  // K right-hand sides at once, then one right-hand side per launch
  auto run = [&](int form) { return form == 0 ? AxLocalKrhs(K, o_q, A.ptr(), chunk) : AxLocal(o_q, A.ptr()); };
  std::vector<hipEvent_t> ev(4 * (size_t)launches, nullptr);
  int st = BPX_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      st = set_error(BPX_ERR_HIP, "time k-RHS: event create failed");
      break;
    }
  for (int form = 0; form < 2 && st == BPX_OK; ++form) {
    for (int w = 0; w < 2 && st == BPX_OK; ++w) st = run(form);
    for (int i = 0; i < launches && st == BPX_OK; ++i) {
      struct Clear {  // the hooks must not outlive their events, whatever run() does
        hipEvent_t &a, &b;
        ~Clear() { a = b = nullptr; }
      } clear{axEvStart, axEvStop};
      axEvStart = ev[4 * i + 2 * form];  // stamped by the dispatch itself
      axEvStop = ev[4 * i + 2 * form + 1];
      st = run(form);
    }
  }
  if (st == BPX_OK && hipStreamSynchronize(s) != hipSuccess) st = set_error(BPX_ERR_HIP, "time k-RHS: sync");
  double t[2] = {0, 0};
  for (int i = 0; i < launches && st == BPX_OK; ++i)
    for (int form = 0; form < 2 && st == BPX_OK; ++form) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, ev[4 * i + 2 * form], ev[4 * i + 2 * form + 1]) != hipSuccess) {
        (void)hipGetLastError();  // do not leave the failure for the next launch check
        st = set_error(BPX_ERR_HIP, "time k-RHS: elapsed time");
      }
      t[form] += ms;
    }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (st) return st;
  out[0] = t[0] * 1e3 / launches;
  out[1] = t[1] * 1e3 / launches;
  return BPX_OK;
}

int bp_t::OperatorKrhs(int K, const dfloat* o_q, dfloat* o_Aq, int chunk) {
  if (K < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "bp: K < 1");
  const size_t need = (size_t)NLocal * K;
  if (o_AqLk.length() < need) {
    if (int st = platform->finish()) return st;  // the old buffer may still be in use
    if (int st = o_AqLk.alloc(need)) return st;
  }
  if (int st = AxLocalKrhs(K, o_q, o_AqLk.ptr(), chunk)) return st;
  for (int r = 0; r < K; ++r)  // one gather per right-hand side, same maps
    if (int st = ogs.Gather<double>(*platform, o_Aq + (size_t)r * ogs.Ngather, o_AqLk.ptr() + (size_t)r * NLocal, 1))
      return st;
  return BPX_OK;
}

int bp_t::OperatorDot32(const float* o_q, float* o_Aq, double* part, int* npart, const CgFold<float>* fold) {
  if (Nfields != 1) return set_error(BPX_ERR_UNSUPPORTED, "bp: the FP32 operator covers one field (BP1/3/5)");
  if (haveKappa) return set_error(BPX_ERR_UNSUPPORTED, "bp: the FP32 operator has no variable coefficient");
  if (int st = Prepare32()) return st;
  if (fold && !(fused && tuning.fused32))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "bp: the folded CG update needs the fused FP32 operator");
  auto apply = [&]() {
    return (fused && tuning.fused32) ? OperatorDotFused<float>(o_q, o_Aq, part, npart, fold)
                                     : OperatorDotUnfused32(o_q, o_Aq, part, npart);
  };
  if (!opTimer32 || opTimer32->used + 3 > opTimer32->ev.size()) return apply();
  hipEvent_t* e = &opTimer32->ev[opTimer32->used];
  opTimer32->used += 3;
  int st;
  {
    struct Clear {  // the hooks must not outlive this application
      hipEvent_t &a, &b;
      ~Clear() { a = b = nullptr; }
    } clear{axEvStart, axEvStop};
    axEvStart = e[0];  // stamped by the Ax dispatch itself
    axEvStop = e[1];
    st = apply();
  }
  if (st) return st;
  BPX_HIP(hipEventRecord(e[2], platform->stream));
  return BPX_OK;
}

int bp_t::TimeOperatorInSolve(dfloat* o_x, const dfloat* o_b, double tol, int interval, int maxit, double* out) {
  if (maxit < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "time operator in solve: maxit < 1");
  OpTimer t;  // the first 8192 applications are timed
  t.ev.assign(3 * (size_t)std::min(maxit, 8192), nullptr);
  int st = BPX_OK;
  for (auto& e : t.ev)
    if (hipEventCreate(&e) != hipSuccess) {
      st = set_error(BPX_ERR_HIP, "time operator in solve: event create failed");
      break;
    }
  if (st == BPX_OK) {
    struct Clear {
      OpTimer*& h;
      ~Clear() { h = nullptr; }
    } clear{opTimer32};
    opTimer32 = &t;
    double stats[3];
    st = SolveMixed(o_x, o_b, tol, interval, maxit, stats);
  }
  const size_t n = t.used / 3;
  if (st == BPX_OK && n == 0) st = set_error(BPX_ERR_INVALID_ARGUMENT, "time operator in solve: no FP32 iteration ran");
  if (st == BPX_OK && hipStreamSynchronize(platform->stream) != hipSuccess)
    st = set_error(BPX_ERR_HIP, "time operator in solve: sync");
  double tax = 0, tg = 0;
  for (size_t i = 0; i < n && st == BPX_OK; ++i) {
    float a = 0, b = 0;
    if (hipEventElapsedTime(&a, t.ev[3 * i], t.ev[3 * i + 1]) != hipSuccess ||
        hipEventElapsedTime(&b, t.ev[3 * i + 1], t.ev[3 * i + 2]) != hipSuccess) {
      (void)hipGetLastError();  // do not leave the failure for the next launch check
      st = set_error(BPX_ERR_HIP, "time operator in solve: elapsed time");
    }
    tax += a;
    tg += b;
  }
  for (auto& e : t.ev)
    if (e) (void)hipEventDestroy(e);
  if (st) return st;
  out[0] = tax * 1e3 / n;
  out[1] = tg * 1e3 / n;
  out[2] = (double)n;
  out[3] = fused && tuning.fused32 ? 1.0 : 0.0;
  return BPX_OK;
}

int bp_t::OperatorDotUnfused32(const float* o_q, float* o_Aq, double* part, int* npart) {
  if (int st = ax<float>(o_q, o_AqL32.ptr(), o_wJ32.ptr(), o_ggeo32.ptr(), o_gllw32.ptr(), o_D32.ptr(),
                         o_I32.ptr(), o_MM32.ptr(), o_invV32.ptr(), SolveG2L()))
    return st;
  const ogs::ogs_t& G = SolveOgs();
  *npart = ogs_gather_partials(G.NrowBlocksT);
  return launch_ogs_gather_dot<float>(platform->stream, G.NrowBlocksT, G.o_blockRowStartsT.ptr(),
                                      G.o_rowStartsT.ptr(), G.o_colIdsT.ptr(), o_AqL32.ptr(), o_Aq, o_q, part,
                                      ogs.gatherMode);
}

namespace {
// best of three event-timed runs after one untimed run, in microseconds
template <typename F>
int time_best(hipStream_t s, F&& fn, double& us) {
  hipEvent_t e0, e1;
  BPX_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return set_error(BPX_ERR_HIP, "tune: event create failed");
  }
  int st = fn();
  float best = 1e30f;
  for (int rep = 0; rep < 3 && st == BPX_OK; ++rep) {
    float ms = 0;
    if (hipEventRecord(e0, s) != hipSuccess) {
      st = set_error(BPX_ERR_HIP, "tune: event record failed");
      break;
    }
    if ((st = fn())) break;
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      st = set_error(BPX_ERR_HIP, "tune: event timing failed");
      break;
    }
    best = std::min(best, ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st == BPX_OK) us = best * 1e3;
  return st;
}
}  // namespace

int bp_t::Tune() {
  hipStream_t s = platform->stream;
  const dlong Ng = ogs.Ngather;
  if (Ng == 0 || ogs.NrowBlocksT == 0) return BPX_OK;
  deviceBuffer<double> v, part;
  if (int st = v.alloc(2 * (size_t)Ng)) return st;
  if (int st = part.alloc(GatherPartials())) return st;
  BPX_HIP(hipMemsetAsync(o_AqL.ptr(), 0, (size_t)NLocal * sizeof(double), s));
  BPX_HIP(hipMemsetAsync(v.ptr(), 0, 2 * (size_t)Ng * sizeof(double), s));
  for (int mode = 0; mode < 2; ++mode) {
    auto run = [&] {
      return launch_ogs_gather_dot<double>(s, ogs.NrowBlocksT, ogs.o_blockRowStartsT.ptr(), ogs.o_rowStartsT.ptr(),
                                           ogs.o_colIdsT.ptr(), o_AqL.ptr(), v.ptr(), v.ptr() + Ng, part.ptr(),
                                           mode);
    };
    if (int st = time_best(s, run, tuning.gather_us[mode])) return st;
  }
  tuning.gatherMode = tuning.gather_us[1] < tuning.gather_us[0] ? 1 : 0;
  ogs.gatherMode = tuning.gatherMode;
  if (!fused) return BPX_OK;

  // the CG operator in the solve space: fused (Ax writes single-copy rows,
  // shared rows gathered) against unfused (Ax, then the full gather); same bits
  deviceBuffer<double> w;
  if (int st = w.alloc(2 * (size_t)Ns)) return st;
  BPX_HIP(hipMemsetAsync(w.ptr(), 0, 2 * (size_t)Ns * sizeof(double), s));
  int np = 0;
  if (int st = time_best(s, [&] { return OperatorDotFused<double>(w.ptr(), w.ptr() + Ns, part.ptr(), &np); },
                         tuning.fused_us))
    return st;
  if (int st = time_best(s, [&] { return OperatorDotUnfusedS(w.ptr(), w.ptr() + Ns, part.ptr(), &np); },
                         tuning.unfused_us))
    return st;
  tuning.fused64 = tuning.fused_us < tuning.unfused_us ? 1 : 0;
  return BPX_OK;
}

// The same choice for the FP32 operator, timed when it is first built
int bp_t::Tune32() {
  if (!fused) return BPX_OK;
  hipStream_t s = platform->stream;
  if (Ns == 0) return BPX_OK;
  deviceBuffer<float> v;
  deviceBuffer<double> part;
  if (int st = v.alloc(2 * (size_t)Ns)) return st;
  if (int st = part.alloc(GatherPartials())) return st;
  BPX_HIP(hipMemsetAsync(v.ptr(), 0, 2 * (size_t)Ns * sizeof(float), s));
  int np = 0;
  if (int st = time_best(s, [&] { return OperatorDotFused<float>(v.ptr(), v.ptr() + Ns, part.ptr(), &np); },
                         tuning.fused32_us))
    return st;
  if (int st = time_best(s, [&] { return OperatorDotUnfused32(v.ptr(), v.ptr() + Ns, part.ptr(), &np); },
                         tuning.unfused32_us))
    return st;
  tuning.fused32 = tuning.fused32_us < tuning.unfused32_us ? 1 : 0;
  return BPX_OK;
}

// The CG operator as the solvers apply it (solve space, tuned form), timed
// kernel by kernel with events on the library's stream: out[0] = mean Ax
// launch (us), out[1] = mean gather launch (us), out[2] = 1 if fused.
int bp_t::TimeOperator(int precision, int launches, double* out) {
  if (precision != 32 && precision != 64) return set_error(BPX_ERR_INVALID_ARGUMENT, "time operator: precision 32 or 64");
  if (launches < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "time operator: launches < 1");
  if (precision == 32)
    if (int st = Prepare32()) return st;
  hipStream_t s = platform->stream;
  const size_t n = (size_t)std::max<dlong>(Ns, 1);
  deviceBuffer<double> v64, part;
  deviceBuffer<float> v32;
  if (int st = part.alloc(std::max(GatherPartials(), 1))) return st;
  if (precision == 64) {
    if (int st = v64.alloc(2 * n)) return st;
    BPX_HIP(hipMemsetAsync(v64.ptr(), 0, 2 * n * sizeof(double), s));
  } else {
    if (int st = v32.alloc(2 * n)) return st;
    BPX_HIP(hipMemsetAsync(v32.ptr(), 0, 2 * n * sizeof(float), s));
  }
  auto run = [&]() -> int {
    int np = 0;
    return precision == 64 ? OperatorDotS(v64.ptr(), v64.ptr() + n, part.ptr(), &np)
                           : OperatorDot32(v32.ptr(), v32.ptr() + n, part.ptr(), &np);
  };
  for (int w = 0; w < 3; ++w)
    if (int st = run()) return st;
  std::vector<hipEvent_t> ev(3 * (size_t)launches, nullptr);
  int st = BPX_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      st = set_error(BPX_ERR_HIP, "time operator: event create failed");
      break;
    }
  for (int i = 0; i < launches && st == BPX_OK; ++i) {
    {
      struct Clear {  // the hooks must not outlive their events, whatever run() does
        hipEvent_t &a, &b;
        ~Clear() { a = b = nullptr; }
      } clear{axEvStart, axEvStop};
      axEvStart = ev[3 * i];  // stamped by the Ax dispatch itself
      axEvStop = ev[3 * i + 1];
      st = run();
    }
    if (st == BPX_OK && hipEventRecord(ev[3 * i + 2], s) != hipSuccess)
      st = set_error(BPX_ERR_HIP, "time operator: event record");
  }
  double tax = 0, tg = 0;
  if (st == BPX_OK && hipStreamSynchronize(s) != hipSuccess) st = set_error(BPX_ERR_HIP, "time operator: sync");
  for (int i = 0; i < launches && st == BPX_OK; ++i) {
    float a = 0, b = 0;
    if (hipEventElapsedTime(&a, ev[3 * i], ev[3 * i + 1]) != hipSuccess ||
        hipEventElapsedTime(&b, ev[3 * i + 1], ev[3 * i + 2]) != hipSuccess) {
      (void)hipGetLastError();  // do not leave the failure for the next launch check
      st = set_error(BPX_ERR_HIP, "time operator: elapsed time");
    }
    tax += a;
    tg += b;
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  if (st) return st;
  out[0] = tax * 1e3 / launches;
  out[1] = tg * 1e3 / launches;
  out[2] = fused && (precision == 64 ? tuning.fused64 : tuning.fused32) ? 1.0 : 0.0;
  return BPX_OK;
}

int bp_t::Prepare32() {
  if (have32) return BPX_OK;
  if (mesh->elementType != ELEM_HEX) return set_error(BPX_ERR_UNSUPPORTED, "bp: the FP32 operators cover hexahedra");
  hipStream_t s = platform->stream;
  auto cvt = [&](deviceBuffer<float>& dst, const std::vector<double>& src) {
    std::vector<float> f(src.begin(), src.end());
    return dst.upload(f, s);
  };
  const mesh_t& m = *mesh;
  const int affine = m.opts.affine;
  int st = BPX_OK;
  tri32 = 0;
  if (bp == 5 && !affine && fp32Geometry == 0 && mesh->NodesTrilinear()) {
    // the FP32 operator forms its geometric factors from the elements'
    // trilinear maps: 96 B per element instead of 28 B per node
    std::vector<double> tri;
    m.TrilinearCoefficients(tri);
    if ((st = cvt(o_wJ32, m.gllz)) || (st = cvt(o_ggeo32, tri)) || (st = cvt(o_gllw32, m.gllw)) ||
        (st = cvt(o_D32, m.D)))
      return st;
    tri32 = 1;
  } else if ((st = mesh->EnsureHost())) {
    return st;
  } else if (bp == 5 || bp == 6) {
    if ((st = cvt(o_wJ32, m.wJ)) || (st = cvt(o_ggeo32, m.ggeo)) || (st = cvt(o_gllw32, m.gllw)) ||
        (st = cvt(o_D32, m.D)))
      return st;
  } else if (bp == 1 || bp == 2) {
    if ((st = cvt(o_wJ32, affine ? m.wJ : m.cubwJ)) || (st = cvt(o_I32, m.cubInterp)) || (st = cvt(o_MM32, m.MM)))
      return st;
  } else {
    if ((st = cvt(o_wJ32, affine ? m.wJ : m.cubwJ)) || (st = cvt(o_ggeo32, affine ? m.ggeo : m.cubggeo)) ||
        (st = cvt(o_D32, m.cubD)) || (st = cvt(o_I32, m.cubInterp)))
      return st;
    if (affine && (st = cvt(o_invV32, m.invV))) return st;
  }
  if (bp <= 4) {
    const std::vector<double> eo = cub_even_odd_tables(m);
    if (!eo.empty() && (st = cvt(o_eo32, eo))) return st;
  }
  if ((st = o_AqL32.alloc(NLocal))) return st;
  if (fused && (st = o_slots32.alloc(o_slots.length()))) return st;
  if ((st = platform->finish())) return st;
  have32 = true;
  if ((st = Tune32())) return st;
  return platform->finish();
}

int bp_t::SetFp32Geometry(int mode) {
  if (mode != 0 && mode != 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "bp: FP32 geometry mode must be 0 or 1");
  if (mode == fp32Geometry) return BPX_OK;
  if (int st = platform->finish()) return st;  // the FP32 buffers may still be in use
  fp32Geometry = mode;
  have32 = false;  // rebuilt (and its fused form re-timed) on next use
  tri32 = 0;
  return BPX_OK;
}

int bp_t::Operator32(const float* o_q, float* o_Aq) {
  if (Nfields != 1) return set_error(BPX_ERR_UNSUPPORTED, "bp: the FP32 operator covers one field (BP1/3/5)");
  if (haveKappa) return set_error(BPX_ERR_UNSUPPORTED, "bp: the FP32 operator has no variable coefficient");
  if (int st = Prepare32()) return st;
  if (int st = ax<float>(o_q, o_AqL32.ptr(), o_wJ32.ptr(), o_ggeo32.ptr(), o_gllw32.ptr(), o_D32.ptr(),
                         o_I32.ptr(), o_MM32.ptr(), o_invV32.ptr()))
    return st;
  return ogs.Gather<float>(*platform, o_Aq, o_AqL32.ptr(), 1);
}

// device copies of the nodal coordinates and GLL weights (forcing, error norms)
int bp_t::NodeCoordinates() {
  if (o_x.ptr()) return BPX_OK;
  hipStream_t s = platform->stream;
  const mesh_t& m = *mesh;
  if (m.onDevice) {
    if (int st = o_x.copy_device(m.o_x.ptr(), m.o_x.length(), s)) return st;
    if (int st = o_y.copy_device(m.o_y.ptr(), m.o_y.length(), s)) return st;
    if (int st = o_z.copy_device(m.o_z.ptr(), m.o_z.length(), s)) return st;
  } else {
    if (int st = o_x.upload(m.x, s)) return st;
    if (int st = o_y.upload(m.y, s)) return st;
    if (m.dim == 3)
      if (int st = o_z.upload(m.z, s)) return st;
  }
  return o_gllwRhs.upload(m.gllw, s);
}

// b = gather(forcing) (BP/BP5/src/bp5Run.cpp:51-64); BP1/BP2 force with lambda = 0.
int bp_t::Forcing(dfloat* o_b) {
  hipStream_t s = platform->stream;
  const mesh_t& m = *mesh;
  if (int st = NodeCoordinates()) return st;
  if (m.elementType != ELEM_HEX) {
    deviceBuffer<double> wJ, MMd;
    if (int st = wJ.upload(m.wJ, s)) return st;
    if (int st = MMd.upload(m.MM, s)) return st;
    const double lam = (bp <= 2) ? 0.0 : lambda;
    if (int st = launch_rhs_other(s, m.elementType, m.N, m.Np, Nfields, m.Nelements, wJ.ptr(), MMd.ptr(), o_x.ptr(),
                                  o_y.ptr(), m.dim == 3 ? o_z.ptr() : o_x.ptr(), lam, o_AqL.ptr()))
      return st;
    if (int st = ogs.Gather<double>(*platform, o_b, o_AqL.ptr(), 1)) return st;
    return platform->finish();
  }
  // the forcing uses the GLL-weighted wJ of the mesh (mesh.o_wJ)
  deviceBuffer<double> wJ;
  if (int st = m.onDevice ? wJ.copy_device(m.o_wJ.ptr(), m.o_wJ.length(), s) : wJ.upload(m.wJ, s)) return st;
  const double lam = (bp <= 2) ? 0.0 : lambda;
  if (int st = launch_rhs(s, m.N, Nfields, m.opts.affine, m.Nelements, wJ.ptr(), o_gllwRhs.ptr(), o_x.ptr(),
                          o_y.ptr(), o_z.ptr(), lam, o_AqL.ptr()))
    return st;
  if (int st = ogs.Gather<double>(*platform, o_b, o_AqL.ptr(), 1)) return st;
  return platform->finish();
}

}  // namespace bpx

// extern "C" entry points of include/bpx.h. No C++ exception crosses the
// ABI: every entry point catches and returns an error code.
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "host/bpx.hpp"
#include "host/rng.hpp"

using namespace bpx;

struct bpx_platform_s {
  platform_t* p = nullptr;
};
struct bpx_mesh_s {
  mesh_t m;
};
struct bpx_ogs_s {
  ogs::ogs_t o;
  std::vector<dlong> G2L;
  std::vector<hlong> ids;
  platform_t* plat = nullptr;    // device setup: the platform it ran on
  deviceBuffer<dlong> o_G2L;     // device setup: GlobalToLocal
};
struct bpx_bp_s {
  bpx_platform_s* plat = nullptr;
  bpx_mesh_s* mesh = nullptr;
  bp_t bp;
  cg* solver = nullptr;
  deviceBuffer<double> dotPart;  // partials of bpx_bp_operator_dot
  ~bpx_bp_s() { delete solver; }
};
// solver_t whose Operator is a caller's callback (bpx_cg_solve)
struct callback_solver_t : public solver_t {
  bpx_operator_fn fn = nullptr;
  void* ctx = nullptr;
  int Operator(const dfloat* o_q, dfloat* o_Aq) override {
    const int st = fn(ctx, o_q, o_Aq);
    return st == 0 ? BPX_OK : set_error(st < 0 ? st : BPX_ERR_INVALID_ARGUMENT, "cg: the operator callback failed");
  }
};
struct bpx_cg_s {
  cg* solver = nullptr;
  ~bpx_cg_s() { delete solver; }
};
struct bpx_batch_s {
  bpx_platform_s* plat = nullptr;
  bp_batch_t b;
};

static int to_mesh_opts(const bpx_mesh_opts& o, MeshOpts& mo) {
  mo.N = o.N;
  mo.nx = o.nx;
  mo.ny = o.ny;
  mo.nz = o.nz;
  mo.affine = o.affine;
  mo.cubature = o.cubature;
  for (int c = 0; c < 3; ++c) {
    mo.lo[c] = o.lo[c];
    mo.dim[c] = o.dim[c];
  }
  mo.perturb = o.perturb;
  mo.seed = o.seed;
  mo.elementType = o.elementType == 0 ? ELEM_HEX : o.elementType;
  return BPX_OK;
}

#define BPX_GUARD_BEGIN try {
#define BPX_GUARD_END                                                        \
  }                                                                          \
  catch (const std::bad_alloc&) {                                            \
    return set_error(BPX_ERR_OUT_OF_MEMORY, "host allocation failed");       \
  }                                                                          \
  catch (const std::exception& ex) {                                         \
    return set_error(BPX_ERR_INVALID_ARGUMENT, ex.what());                   \
  }                                                                          \
  catch (...) {                                                              \
    return set_error(BPX_ERR_INVALID_ARGUMENT, "unknown C++ exception");     \
  }

static int need(const void* p, const char* what) {
  return p ? BPX_OK : set_error(BPX_ERR_INVALID_ARGUMENT, std::string(what) + " is NULL");
}

static hipStream_t S(bpx_platform p) { return p->p->stream; }

template <typename T>
static int copy_out(const std::vector<T>& v, void* dst, int64_t* count, int* dtype, int code) {
  if (count) *count = (int64_t)v.size();
  if (dtype) *dtype = code;
  if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
  return BPX_OK;
}

// a device buffer copied to host memory dst (synchronously on stream s)
template <typename T>
static int copy_out_dev(const deviceBuffer<T>& v, size_t n, void* dst, int64_t* count, int* dtype, int code,
                        hipStream_t s) {
  n = std::min(n, v.length());
  if (count) *count = (int64_t)n;
  if (dtype) *dtype = code;
  if (dst && n > 0) {
    BPX_HIP(hipMemcpyAsync(dst, v.ptr(), n * sizeof(T), hipMemcpyDeviceToHost, s));
    BPX_HIP(hipStreamSynchronize(s));
  }
  return BPX_OK;
}

extern "C" {

const char* bpx_last_error(void) { return last_error(); }
int bpx_version(void) { return 100; }

int bpx_device_count(int* count) {
  if (int st = need(count, "count")) return st;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *count = (e == hipSuccess) ? n : 0;
  return BPX_OK;
}

int bpx_random_normal(uint64_t seed, int64_t counter0, int64_t n, double* out) {
  if (n < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "random: n < 0");
  if (n > 0)
    if (int st = need(out, "out")) return st;
#pragma omp parallel for
  for (int64_t i = 0; i < n; ++i) out[i] = rng::normal(seed, (uint64_t)(counter0 + i));
  return BPX_OK;
}

// ------------------------------------------------------------ platform --
int bpx_platform_create(int device, bpx_platform* out) {
  BPX_GUARD_BEGIN
  if (int st = need(out, "out")) return st;
  platform_t* p = nullptr;
  if (int st = platform_t::Create(device, &p)) return st;
  *out = new bpx_platform_s{p};
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_platform_destroy(bpx_platform p) {
  if (!p) return BPX_OK;
  delete p->p;
  delete p;
  return BPX_OK;
}

int bpx_platform_set_stream(bpx_platform p, void* hip_stream) {
  if (int st = need(p, "platform")) return st;
  return p->p->setStream(static_cast<hipStream_t>(hip_stream));
}

void* bpx_platform_stream(bpx_platform p) { return p ? (void*)p->p->stream : nullptr; }

int bpx_platform_finish(bpx_platform p) {
  if (int st = need(p, "platform")) return st;
  return p->p->finish();
}

// ---------------------------------------------------------- Ax kernels --
#define AX_COMMON(T)           \
  AxArgs<T> a;                 \
  a.N = N;                     \
  a.Nelements = Nelements;     \
  a.elementList = elementList; \
  a.GlobalToLocal = GlobalToLocal; \
  a.q = q;                     \
  a.Aq = Aq;

static int bp5_entry(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                     const double* wJ, const double* ggeo, const double* gllw, const double* D, double lambda,
                     const double* q, double* Aq, int affine, int Nfields) {
  if (int st = need(p, "platform")) return st;
  AX_COMMON(double)
  a.Nfields = Nfields;
  a.affine = affine;
  a.wJ = wJ;
  a.ggeo = ggeo;
  a.gllw = gllw;
  a.D = D;
  a.lambda = lambda;
  return launch_bp5_ax<double>(S(p), a);
}

int bpx_bp5_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                   const double* wJ, const double* vgeo, const double* ggeo, const double* gllw, const double* D,
                   const double* S_, const double* MM, double lambda, const double* q, double* Aq) {
  (void)vgeo; (void)S_; (void)MM;
  return bp5_entry(p, N, Nelements, elementList, GlobalToLocal, wJ, ggeo, gllw, D, lambda, q, Aq, 0, 1);
}
int bpx_bp5_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                          const double* wJ, const double* vgeo, const double* ggeo, const double* gllw,
                          const double* D, const double* S_, const double* MM, double lambda, const double* q,
                          double* Aq) {
  (void)vgeo; (void)S_; (void)MM;
  return bp5_entry(p, N, Nelements, elementList, GlobalToLocal, wJ, ggeo, gllw, D, lambda, q, Aq, 1, 1);
}
}  // extern "C"
template <typename T>
static int bp5_krhs_entry(bpx_platform p, int N, int K, int chunk, int Nelements, const int* elementList,
                          const int* GlobalToLocal, const T* wJ, const T* ggeo, const T* gllw, const T* D,
                          T lambda, const T* q, int64_t ldq, T* Aq, int64_t ldAq, int affine) {
  if (int st = need(p, "platform")) return st;
  AX_COMMON(T)
  a.affine = affine;
  a.wJ = wJ;
  a.ggeo = ggeo;
  a.gllw = gllw;
  a.D = D;
  a.lambda = lambda;
  a.K = K;
  a.ldq = ldq;
  a.ldAq = ldAq;
  return launch_bp5_ax_krhs<T>(S(p), a, chunk);
}
extern "C" {

int bpx_bp5_ax_hex_krhs(bpx_platform p, int N, int K, int chunk, int Nelements, const int* elementList,
                        const int* GlobalToLocal, const double* wJ, const double* vgeo, const double* ggeo,
                        const double* gllw, const double* D, const double* S_, const double* MM, double lambda,
                        const double* q, int64_t ldq, double* Aq, int64_t ldAq) {
  (void)vgeo; (void)S_; (void)MM;
  return bp5_krhs_entry<double>(p, N, K, chunk, Nelements, elementList, GlobalToLocal, wJ, ggeo, gllw, D, lambda,
                                q, ldq, Aq, ldAq, 0);
}
int bpx_bp5_ax_affine_hex_krhs(bpx_platform p, int N, int K, int chunk, int Nelements, const int* elementList,
                               const int* GlobalToLocal, const double* wJ, const double* vgeo, const double* ggeo,
                               const double* gllw, const double* D, const double* S_, const double* MM,
                               double lambda, const double* q, int64_t ldq, double* Aq, int64_t ldAq) {
  (void)vgeo; (void)S_; (void)MM;
  return bp5_krhs_entry<double>(p, N, K, chunk, Nelements, elementList, GlobalToLocal, wJ, ggeo, gllw, D, lambda,
                                q, ldq, Aq, ldAq, 1);
}
int bpx_bp5_ax_hex_krhs_f32(bpx_platform p, int N, int K, int chunk, int Nelements, const int* elementList,
                            const int* GlobalToLocal, const float* wJ, const float* vgeo, const float* ggeo,
                            const float* gllw, const float* D, const float* S_, const float* MM, float lambda,
                            const float* q, int64_t ldq, float* Aq, int64_t ldAq) {
  (void)vgeo; (void)S_; (void)MM;
  return bp5_krhs_entry<float>(p, N, K, chunk, Nelements, elementList, GlobalToLocal, wJ, ggeo, gllw, D, lambda,
                               q, ldq, Aq, ldAq, 0);
}
int bpx_bp5_ax_affine_hex_krhs_f32(bpx_platform p, int N, int K, int chunk, int Nelements,
                                   const int* elementList, const int* GlobalToLocal, const float* wJ,
                                   const float* vgeo, const float* ggeo, const float* gllw, const float* D,
                                   const float* S_, const float* MM, float lambda, const float* q, int64_t ldq,
                                   float* Aq, int64_t ldAq) {
  (void)vgeo; (void)S_; (void)MM;
  return bp5_krhs_entry<float>(p, N, K, chunk, Nelements, elementList, GlobalToLocal, wJ, ggeo, gllw, D, lambda,
                               q, ldq, Aq, ldAq, 1);
}
int bpx_bp6_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                   const double* wJ, const double* vgeo, const double* ggeo, const double* gllw, const double* D,
                   const double* S_, const double* MM, double lambda, const double* q, double* Aq) {
  (void)vgeo; (void)S_; (void)MM;
  return bp5_entry(p, N, Nelements, elementList, GlobalToLocal, wJ, ggeo, gllw, D, lambda, q, Aq, 0, 3);
}
int bpx_bp6_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                          const double* wJ, const double* vgeo, const double* ggeo, const double* gllw,
                          const double* D, const double* S_, const double* MM, double lambda, const double* q,
                          double* Aq) {
  (void)vgeo; (void)S_; (void)MM;
  return bp5_entry(p, N, Nelements, elementList, GlobalToLocal, wJ, ggeo, gllw, D, lambda, q, Aq, 1, 3);
}

static int bp1_entry(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                     const double* cubwJ, const double* I, const double* MM, const double* q, double* Aq,
                     int affine, int Nfields) {
  if (int st = need(p, "platform")) return st;
  AX_COMMON(double)
  a.Nfields = Nfields;
  a.affine = affine;
  a.wJ = cubwJ;
  a.I = I;
  a.MM = MM;
  return launch_bp1_ax<double>(S(p), a);
}
int bpx_bp1_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                   const double* cubwJ, const double* I, const double* MM, const double* q, double* Aq) {
  return bp1_entry(p, N, Nelements, elementList, GlobalToLocal, cubwJ, I, MM, q, Aq, 0, 1);
}
int bpx_bp1_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                          const double* cubwJ, const double* I, const double* MM, const double* q, double* Aq) {
  return bp1_entry(p, N, Nelements, elementList, GlobalToLocal, cubwJ, I, MM, q, Aq, 1, 1);
}
int bpx_bp2_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                   const double* cubwJ, const double* I, const double* MM, const double* q, double* Aq) {
  return bp1_entry(p, N, Nelements, elementList, GlobalToLocal, cubwJ, I, MM, q, Aq, 0, 3);
}
int bpx_bp2_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                          const double* cubwJ, const double* I, const double* MM, const double* q, double* Aq) {
  return bp1_entry(p, N, Nelements, elementList, GlobalToLocal, cubwJ, I, MM, q, Aq, 1, 3);
}

static int bp3_entry(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                     const double* W, const double* G, const double* D, const double* I, const double* invV,
                     double lambda, const double* q, double* Aq, int affine, int Nfields) {
  if (int st = need(p, "platform")) return st;
  AX_COMMON(double)
  a.Nfields = Nfields;
  a.affine = affine;
  a.wJ = W;
  a.ggeo = G;
  a.D = D;
  a.I = I;
  a.invV = invV;
  a.lambda = lambda;
  return launch_bp3_ax<double>(S(p), a);
}
int bpx_bp3_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                   const double* cubwJ, const double* cubggeo, const double* D, const double* I,
                   const double* invV, const double* S_, const double* MM, double lambda, const double* q,
                   double* Aq) {
  (void)S_; (void)MM;
  return bp3_entry(p, N, Nelements, elementList, GlobalToLocal, cubwJ, cubggeo, D, I, invV, lambda, q, Aq, 0, 1);
}
int bpx_bp3_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                          const double* wJ, const double* ggeo, const double* D, const double* I,
                          const double* invV, const double* S_, const double* MM, double lambda, const double* q,
                          double* Aq) {
  (void)S_; (void)MM;
  return bp3_entry(p, N, Nelements, elementList, GlobalToLocal, wJ, ggeo, D, I, invV, lambda, q, Aq, 1, 1);
}
int bpx_bp4_ax_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                   const double* cubwJ, const double* cubggeo, const double* D, const double* I,
                   const double* invV, const double* S_, const double* MM, double lambda, const double* q,
                   double* Aq) {
  (void)S_; (void)MM;
  return bp3_entry(p, N, Nelements, elementList, GlobalToLocal, cubwJ, cubggeo, D, I, invV, lambda, q, Aq, 0, 3);
}
int bpx_bp4_ax_affine_hex(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                          const double* wJ, const double* ggeo, const double* D, const double* I,
                          const double* invV, const double* S_, const double* MM, double lambda, const double* q,
                          double* Aq) {
  (void)S_; (void)MM;
  return bp3_entry(p, N, Nelements, elementList, GlobalToLocal, wJ, ggeo, D, I, invV, lambda, q, Aq, 1, 3);
}

int bpx_bp5_ax_hex_f32(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                       const float* wJ, const float* vgeo, const float* ggeo, const float* gllw, const float* D,
                       const float* S_, const float* MM, float lambda, const float* q, float* Aq) {
  (void)vgeo; (void)S_; (void)MM;
  if (int st = need(p, "platform")) return st;
  AX_COMMON(float)
  a.wJ = wJ;
  a.ggeo = ggeo;
  a.gllw = gllw;
  a.D = D;
  a.lambda = lambda;
  return launch_bp5_ax<float>(S(p), a);
}

int bpx_bp3_ax_hex_f32(bpx_platform p, int N, int Nelements, const int* elementList, const int* GlobalToLocal,
                       const float* cubwJ, const float* cubggeo, const float* D, const float* I,
                       const float* invV, const float* S_, const float* MM, float lambda, const float* q,
                       float* Aq) {
  (void)S_; (void)MM;
  if (int st = need(p, "platform")) return st;
  AX_COMMON(float)
  a.wJ = cubwJ;
  a.ggeo = cubggeo;
  a.D = D;
  a.I = I;
  a.invV = invV;
  a.lambda = lambda;
  return launch_bp3_ax<float>(S(p), a);
}

int bpx_rhs_hex(bpx_platform p, int N, int Nfields, int affine, int Nelements, const double* wJ,
                const double* gllw, const double* MM, const double* x, const double* y, const double* z,
                double lambda, double* rhs) {
  (void)MM;
  if (int st = need(p, "platform")) return st;
  return launch_rhs(S(p), N, Nfields, affine, Nelements, wJ, gllw, x, y, z, lambda, rhs);
}

// ------------------------------------------------------------ ogs kernels --
int bpx_ogs_gather(bpx_platform p, int Nblocks, int K, const int* blockStarts, const int* gatherStarts,
                   const int* gatherIds, const double* q, double* gatherq) {
  if (int st = need(p, "platform")) return st;
  return launch_ogs_gather<double>(S(p), Nblocks, K, blockStarts, gatherStarts, gatherIds, q, gatherq);
}
int bpx_ogs_gather_f32(bpx_platform p, int Nblocks, int K, const int* blockStarts, const int* gatherStarts,
                       const int* gatherIds, const float* q, float* gatherq) {
  if (int st = need(p, "platform")) return st;
  return launch_ogs_gather<float>(S(p), Nblocks, K, blockStarts, gatherStarts, gatherIds, q, gatherq);
}
int bpx_ogs_scatter(bpx_platform p, int Nblocks, int K, const int* blockStarts, const int* scatterStarts,
                    const int* scatterIds, const double* gatherq, double* q) {
  if (int st = need(p, "platform")) return st;
  return launch_ogs_scatter<double>(S(p), Nblocks, K, blockStarts, scatterStarts, scatterIds, gatherq, q);
}
int bpx_ogs_gather_scatter(bpx_platform p, int Nblocks, int K, const int* blockStarts, const int* gatherStarts,
                           const int* gatherIds, double* q) {
  if (int st = need(p, "platform")) return st;
  return launch_ogs_gather_scatter<double>(S(p), Nblocks, K, blockStarts, gatherStarts, gatherIds, q);
}

// -------------------------------------------------------------- linAlg --
int bpx_linalg_set(bpx_platform p, int N, double alpha, double* x) {
  if (int st = need(p, "platform")) return st;
  return launch_set<double>(S(p), N, alpha, x);
}
int bpx_linalg_axpy(bpx_platform p, int N, double alpha, const double* x, double beta, double* y) {
  if (int st = need(p, "platform")) return st;
  return launch_axpy<double>(S(p), N, alpha, x, beta, y);
}
int bpx_linalg_inner_prod(bpx_platform p, int N, const double* x, const double* y, double* result) {
  if (int st = need(p, "platform")) return st;
  if (int st = need(result, "result")) return st;
  linAlg_t la(*p->p);
  return la.innerProd(N, x, y, result);
}
int bpx_linalg_norm2(bpx_platform p, int N, const double* x, double* result) {
  if (int st = need(p, "platform")) return st;
  if (int st = need(result, "result")) return st;
  linAlg_t la(*p->p);
  return la.norm2(N, x, result);
}
int bpx_cg_update(bpx_platform p, int N, const double* Ap, const double* o_p, double alpha, double* r, double* x,
                  double* rdotr) {
  if (int st = need(p, "platform")) return st;
  if (int st = need(rdotr, "rdotr")) return st;
  platform_t& P = *p->p;
  if (int st = launch_update_cg(P.stream, N, alpha, o_p, Ap, x, r, P.partials, P.d_scalar)) return st;
  BPX_HIP(hipMemcpyAsync(P.h_scalar, P.d_scalar, sizeof(double), hipMemcpyDeviceToHost, P.stream));
  if (int st = P.finish()) return st;
  *rdotr = *P.h_scalar;
  return BPX_OK;
}

// ------------------------------------------------------------------ cg --
int bpx_cg_create(bpx_platform p, int N, int Nhalo, bpx_cg* out) {
  BPX_GUARD_BEGIN
  if (int st = need(p, "platform")) return st;
  if (int st = need(out, "out")) return st;
  if (N < 0 || Nhalo < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "cg: N and Nhalo must be >= 0");
  bpx_cg_s* c = new bpx_cg_s;
  c->solver = new cg(*p->p, N, Nhalo);
  if (int st = c->solver->status()) {
    delete c;
    return st;
  }
  *out = c;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_cg_destroy(bpx_cg c) {
  delete c;
  return BPX_OK;
}

int bpx_cg_solve(bpx_cg c, bpx_operator_fn op, void* ctx, double* o_x, double* o_r, double tol, int maxit,
                 int verbose, int* iterations, double* rnorm_hist) {
  BPX_GUARD_BEGIN
  if (int st = need(c, "cg")) return st;
  if (!op) return set_error(BPX_ERR_INVALID_ARGUMENT, "cg: operator callback is NULL");
  callback_solver_t s;
  s.fn = op;
  s.ctx = ctx;
  int its = 0;
  const int st = c->solver->Solve(s, o_x, o_r, tol, maxit, verbose, &its);
  if (iterations) *iterations = its;
  if (rnorm_hist && st == BPX_OK) {
    std::vector<double> rr, al;
    if (int s2 = c->solver->impl.History(rr, al)) return s2;
    for (int k = 0; k <= maxit; ++k) rnorm_hist[k] = (k < (int)rr.size()) ? std::sqrt(rr[k]) : 0.0;
  }
  return st;
  BPX_GUARD_END
}

// The solve of bpx_cg_solve on the fixed work array of cg_noalloc.hip: the
// same kernels and host loop (cg_run), with no allocation of any kind.
static thread_local bool t_in_noalloc_solve = false;

int bpx_cg_solve_noalloc(void* hip_stream, int N, bpx_operator_fn op, void* ctx, double* o_x, double* o_r,
                         double tol, int maxit, int* iterations, double* rnorm_hist) {
  BPX_GUARD_BEGIN
  if (!op) return set_error(BPX_ERR_INVALID_ARGUMENT, "cg (no-alloc): operator callback is NULL");
  if (N < 0 || (long long)N > kNoAllocMaxN)
    return set_error(BPX_ERR_INVALID_ARGUMENT,
                     "cg (no-alloc): N must be in 0.." + std::to_string(kNoAllocMaxN) + " (the fixed work array)");
  if (maxit < 0 || maxit > kNoAllocMaxIter)
    return set_error(BPX_ERR_INVALID_ARGUMENT,
                     "cg (no-alloc): maxit must be in 0.." + std::to_string(kNoAllocMaxIter) + " (the fixed work array)");
  if (N > 0 && (!o_x || !o_r)) return set_error(BPX_ERR_INVALID_ARGUMENT, "cg (no-alloc): o_x or o_r is NULL");
  if (t_in_noalloc_solve)
    return set_error(BPX_ERR_INVALID_ARGUMENT, "cg (no-alloc): called from inside the operator callback of a no-alloc solve");
  hipStream_t s = static_cast<hipStream_t>(hip_stream);
  int dev = -1;
  NoAllocWork nw;
  if (int st = noalloc_work(s, &dev, &nw)) return st;
  std::lock_guard<std::mutex> lock(*nw.lock);
  t_in_noalloc_solve = true;
  struct Reset {
    ~Reset() { t_in_noalloc_solve = false; }
  } reset;
  // the array keeps the previous call's contents: clear the state, p and Ap
  // with fill kernels (no memset nodes, so the queue can also be captured)
  if (int st = launch_set<double>(s, nw.state_doubles, 0.0, reinterpret_cast<double*>(nw.state))) return st;
  if (int st = launch_set<double>(s, N, 0.0, nw.p)) return st;
  if (int st = launch_set<double>(s, N, 0.0, nw.Ap)) return st;
  CgWork<double> w;
  w.p = nw.p;
  w.Ap = nw.Ap;
  w.state = nw.state;
  w.part_rr = nw.part_rr;
  w.part_pAp = nw.part_pAp;
  w.hist_rr = nw.hist_rr;
  w.hist_alpha = nw.hist_alpha;
  struct Callback {
    bpx_operator_fn fn;
    void* ctx;
  } cb{op, ctx};
  const Callback* pcb = &cb;  // one pointer: the std::function keeps it in place, without a heap block
  const cg_t<double>::Apply apply = [pcb](const double* q, double* Aq) {
    const int st = pcb->fn(pcb->ctx, q, Aq);
    return st == 0 ? BPX_OK : set_error(st < 0 ? st : BPX_ERR_INVALID_ARGUMENT, "cg: the operator callback failed");
  };
  int its = 0;
  const int st = cg_run<double>(s, N, w, apply, cg_t<double>::ApplyDot(), o_x, o_r, tol, maxit, &its);
  if (iterations) *iterations = its;
  if (rnorm_hist && st == BPX_OK) {
    const int nh = (maxit == 0) ? 1 : its + 1;
    BPX_HIP(hipMemcpyAsync(rnorm_hist, w.hist_rr, nh * sizeof(double), hipMemcpyDeviceToHost, s));
    BPX_HIP(hipStreamSynchronize(s));
    for (int k = 0; k <= maxit; ++k) rnorm_hist[k] = (k < nh) ? std::sqrt(rnorm_hist[k]) : 0.0;
  }
  return st;
  BPX_GUARD_END
}

// ---------------------------------------------------------------- mesh --
int bpx_mesh_create(const bpx_mesh_opts* o, bpx_mesh* out) {
  BPX_GUARD_BEGIN
  if (int st = need(o, "opts")) return st;
  if (int st = need(out, "out")) return st;
  MeshOpts mo;
  to_mesh_opts(*o, mo);
  bpx_mesh_s* m = new bpx_mesh_s;
  if (int st = m->m.Setup(mo)) {
    delete m;
    return st;
  }
  *out = m;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_mesh_create_dev(bpx_platform p, const bpx_mesh_opts* o, bpx_mesh* out) {
  BPX_GUARD_BEGIN
  if (int st = need(p, "platform")) return st;
  if (int st = need(o, "opts")) return st;
  if (int st = need(out, "out")) return st;
  MeshOpts mo;
  to_mesh_opts(*o, mo);
  bpx_mesh_s* m = new bpx_mesh_s;
  if (int st = m->m.SetupDevice(*p->p, mo)) {
    delete m;
    return st;
  }
  *out = m;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_mesh_destroy(bpx_mesh m) {
  delete m;
  return BPX_OK;
}

int bpx_mesh_info(bpx_mesh m, int64_t* out, int capacity) {
  if (int st = need(m, "mesh")) return st;
  const int64_t v[] = {m->m.N,         m->m.Nq,   m->m.Np,     m->m.cubNq, m->m.cubNp,
                       m->m.Nelements, m->m.elementType, m->m.dim, m->m.Nverts};
  const int n = (int)(sizeof(v) / sizeof(v[0]));
  for (int i = 0; i < n && i < capacity; ++i) out[i] = v[i];
  return n;
}

int bpx_mesh_array(bpx_mesh m, const char* name, void* dst, int64_t* count, int* dtype) {
  BPX_GUARD_BEGIN
  if (int st = need(m, "mesh")) return st;
  if (int st = need(name, "name")) return st;
  if (int st = m->m.EnsureHost()) return st;
  const mesh_t& M = m->m;
  const std::string s(name);
#define F64(f) if (s == #f) return copy_out(M.f, dst, count, dtype, 0);
  F64(gllz) F64(gllw) F64(D) F64(MM) F64(cubr) F64(cubw) F64(cubInterp) F64(cubD) F64(invV)
  F64(x) F64(y) F64(z) F64(wJ) F64(ggeo) F64(cubwJ) F64(cubggeo) F64(EX) F64(EY) F64(EZ)
  F64(rN) F64(sN) F64(tN) F64(Dsim) F64(Ssim)
#undef F64
  if (s == "globalIds") return copy_out(M.globalIds, dst, count, dtype, 2);
  return set_error(BPX_ERR_INVALID_ARGUMENT, "mesh: unknown array " + s);
  BPX_GUARD_END
}

// ----------------------------------------------------------------- ogs --
int bpx_ogs_setup(int N, const int64_t* ids, bpx_ogs* out) {
  BPX_GUARD_BEGIN
  if (int st = need(out, "out")) return st;
  if (N > 0)
    if (int st = need(ids, "ids")) return st;
  bpx_ogs_s* o = new bpx_ogs_s;
  o->ids.assign(ids, ids + N);
  if (int st = o->o.Setup(N, o->ids.data())) {
    delete o;
    return st;
  }
  o->G2L.resize(N);
  o->o.SetupGlobalToLocalMapping(o->G2L.data());
  *out = o;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_ogs_setup_dev(bpx_platform p, int N, const int64_t* o_ids, bpx_ogs* out) {
  BPX_GUARD_BEGIN
  if (int st = need(p, "platform")) return st;
  if (int st = need(out, "out")) return st;
  if (N > 0)
    if (int st = need(o_ids, "o_ids")) return st;
  bpx_ogs_s* o = new bpx_ogs_s;
  o->plat = p->p;
  if (int st = o->o_G2L.alloc(std::max(N, 0))) {
    delete o;
    return st;
  }
  if (int st = o->o.SetupDevice(*p->p, N, (const hlong*)o_ids, o->o_G2L.ptr())) {
    delete o;
    return st;
  }
  *out = o;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_ogs_setup_masked(bpx_mesh m, int Nfields, bpx_ogs* out) {
  BPX_GUARD_BEGIN
  if (int st = need(m, "mesh")) return st;
  if (int st = need(out, "out")) return st;
  if (Nfields < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: Nfields < 1");
  if (m->m.onDevice) {  // a device mesh: ids and maps formed on its device
    platform_t& P = *m->m.devPlatform;
    const size_t NL = (size_t)m->m.Nelements * m->m.Np * Nfields;
    deviceBuffer<hlong> ids;
    if (int st = ids.alloc(NL)) return st;
    if (int st = m->m.MaskedGlobalIdsDevice(Nfields, ids.ptr())) return st;
    bpx_ogs_s* o = new bpx_ogs_s;
    o->plat = &P;
    int st = o->o_G2L.alloc(NL);
    if (!st) st = o->o.SetupDevice(P, (dlong)NL, ids.ptr(), o->o_G2L.ptr());
    if (!st) st = ids.download(o->ids, P.stream);
    if (st) {
      delete o;
      return st;
    }
    *out = o;
    return BPX_OK;
  }
  bpx_ogs_s* o = new bpx_ogs_s;
  m->m.MaskedGlobalIds(Nfields, o->ids);
  if (int st = o->o.Setup((dlong)o->ids.size(), o->ids.data())) {
    delete o;
    return st;
  }
  o->G2L.resize(o->ids.size());
  o->o.SetupGlobalToLocalMapping(o->G2L.data());
  *out = o;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_ogs_destroy(bpx_ogs o) {
  delete o;
  return BPX_OK;
}

int bpx_ogs_info(bpx_ogs o, int64_t* out, int capacity) {
  if (int st = need(o, "ogs")) return st;
  const int64_t v[] = {o->o.N, o->o.Ngather, o->o.nnzT, o->o.NrowBlocksT};
  const int n = 4;
  for (int i = 0; i < n && i < capacity; ++i) out[i] = v[i];
  return n;
}

int bpx_ogs_array(bpx_ogs o, const char* name, void* dst, int64_t* count, int* dtype) {
  BPX_GUARD_BEGIN
  if (int st = need(o, "ogs")) return st;
  if (int st = need(name, "name")) return st;
  const std::string s(name);
  if (o->o.onDevice) {  // maps formed on the device: copied down on first request
    if (int st = o->o.EnsureHost(*o->plat)) return st;
    if (o->G2L.empty() && o->o_G2L.length() > 0)
      if (int st = o->o_G2L.download(o->G2L, o->plat->stream)) return st;
  }
  if (s == "rowStartsT") return copy_out(o->o.rowStartsT, dst, count, dtype, 1);
  if (s == "colIdsT") return copy_out(o->o.colIdsT, dst, count, dtype, 1);
  if (s == "blockRowStartsT") return copy_out(o->o.blockRowStartsT, dst, count, dtype, 1);
  if (s == "GlobalToLocal") return copy_out(o->G2L, dst, count, dtype, 1);
  if (s == "maskedIds") return copy_out(o->ids, dst, count, dtype, 2);
  return set_error(BPX_ERR_INVALID_ARGUMENT, "ogs: unknown array " + s);
  BPX_GUARD_END
}

int bpx_ogs_gather_dev(bpx_platform p, bpx_ogs o, int K, const double* v, double* gv) {
  if (int st = need(p, "platform")) return st;
  if (int st = need(o, "ogs")) return st;
  if (!o->o.o_rowStartsT.ptr() && o->o.Ngather > 0)
    if (int st = o->o.Upload(*p->p)) return st;
  return o->o.Gather<double>(*p->p, gv, v, K);
}

int bpx_ogs_scatter_dev(bpx_platform p, bpx_ogs o, int K, const double* gv, double* v) {
  if (int st = need(p, "platform")) return st;
  if (int st = need(o, "ogs")) return st;
  if (!o->o.o_rowStartsT.ptr() && o->o.Ngather > 0)
    if (int st = o->o.Upload(*p->p)) return st;
  return o->o.Scatter<double>(*p->p, v, gv, K);
}

int bpx_ogs_gather_scatter_dev(bpx_platform p, bpx_ogs o, int K, double* v) {
  if (int st = need(p, "platform")) return st;
  if (int st = need(o, "ogs")) return st;
  if (!o->o.o_rowStartsT.ptr() && o->o.Ngather > 0)
    if (int st = o->o.Upload(*p->p)) return st;
  return o->o.GatherScatter(*p->p, v, K);
}

// ------------------------------------------------------------------ BP --
int bpx_bp_create(bpx_platform p, bpx_mesh m, int bp, double lambda, bpx_bp* out) {
  BPX_GUARD_BEGIN
  if (int st = need(p, "platform")) return st;
  if (int st = need(m, "mesh")) return st;
  if (int st = need(out, "out")) return st;
  bpx_bp_s* b = new bpx_bp_s;
  b->plat = p;
  b->mesh = m;
  if (int st = b->bp.Setup(*p->p, m->m, bp, lambda)) {
    delete b;
    return st;
  }
  b->solver = new cg(*p->p, b->bp.Ns, 0, b->bp.GatherPartials());
  if (int st = b->solver->status()) {
    delete b;
    return st;
  }
  *out = b;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_bp_destroy(bpx_bp b) {
  delete b;
  return BPX_OK;
}

int bpx_bp_info(bpx_bp b, int64_t* out, int capacity) {
  if (int st = need(b, "bp")) return st;
  const int64_t v[] = {b->bp.bp, b->bp.Nfields, b->bp.ogs.Ngather, b->bp.NLocal, b->bp.mesh->Nelements,
                       b->bp.mesh->N, b->bp.LatticeNodes()};
  const int n = 7;
  for (int i = 0; i < n && i < capacity; ++i) out[i] = v[i];
  return n;
}

int bpx_bp_array(bpx_bp b, const char* name, void* dst, int64_t* count, int* dtype) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  if (int st = need(name, "name")) return st;
  bp_t& B = b->bp;
  hipStream_t st = B.platform->stream;
  const std::string s(name);
  const size_t big = (size_t)-1;
  if (s == "rowStartsT") return copy_out_dev(B.ogs.o_rowStartsT, big, dst, count, dtype, 1, st);
  if (s == "colIdsT") return copy_out_dev(B.ogs.o_colIdsT, big, dst, count, dtype, 1, st);
  if (s == "blockRowStartsT") return copy_out_dev(B.ogs.o_blockRowStartsT, big, dst, count, dtype, 1, st);
  if (s == "GlobalToLocal") return copy_out_dev(B.o_GlobalToLocal, big, dst, count, dtype, 1, st);
  if (s == "wJ") return copy_out_dev(B.o_wJ, big, dst, count, dtype, 0, st);
  if (s == "ggeo") return copy_out_dev(B.o_ggeo, big, dst, count, dtype, 0, st);
  if (s == "solvePerm") return copy_out_dev(B.o_sperm, big, dst, count, dtype, 1, st);
  if (s == "solveGlobalToLocal") return copy_out_dev(B.o_G2Ls, big, dst, count, dtype, 1, st);
  if (s == "solveRowStarts") return copy_out_dev(B.solveOgs.o_rowStartsT, big, dst, count, dtype, 1, st);
  if (s == "solveColIds") return copy_out_dev(B.solveOgs.o_colIdsT, big, dst, count, dtype, 1, st);
  if (s == "solveBlocks") return copy_out_dev(B.solveOgs.o_blockRowStartsT, big, dst, count, dtype, 1, st);
  if (s == "sharedRowStarts") return copy_out_dev(B.sharedOgs.o_rowStartsT, big, dst, count, dtype, 1, st);
  if (s == "sharedColIds") return copy_out_dev(B.sharedOgs.o_colIdsT, big, dst, count, dtype, 1, st);
  if (s == "sharedBlocks") return copy_out_dev(B.sharedOgs.o_blockRowStartsT, big, dst, count, dtype, 1, st);
  return set_error(BPX_ERR_INVALID_ARGUMENT, "bp: unknown array " + s);
  BPX_GUARD_END
}

int bpx_bp_tuning(bpx_bp b, double* out, int capacity) {
  if (int st = need(b, "bp")) return st;
  const auto& t = b->bp.tuning;
  const double v[] = {(double)t.gatherMode, t.gather_us[0], t.gather_us[1],
                      b->bp.fused ? (double)t.fused64 : -1.0, t.fused_us, t.unfused_us,
                      b->bp.fused ? (double)t.fused32 : -1.0, t.fused32_us, t.unfused32_us,
                      b->bp.have32 ? (double)b->bp.tri32 : -1.0};
  const int n = 10;
  for (int i = 0; i < n && i < capacity; ++i) out[i] = v[i];
  return n;
}

int bpx_bp_operator(bpx_bp b, const double* o_q, double* o_Aq) {
  if (int st = need(b, "bp")) return st;
  return b->bp.Operator(o_q, o_Aq);
}

int bpx_bp_set_tuning(bpx_bp b, int gatherMode, int fused, int fused32) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  auto bad = [](int v) { return v < -1 || v > 1; };
  if (bad(gatherMode) || bad(fused) || bad(fused32))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "bp: tuning choice out of range");
  if ((fused == 1 || fused32 == 1) && !b->bp.fused)
    return set_error(BPX_ERR_UNSUPPORTED, "bp: no fused operator for this problem (BP5, one field)");
  auto& t = b->bp.tuning;
  if (gatherMode >= 0) t.gatherMode = b->bp.ogs.gatherMode = gatherMode;
  if (fused >= 0) t.fused64 = fused;
  if (fused32 >= 0) {
    if (int st = b->bp.Prepare32()) return st;  // so a later first use does not re-time
    t.fused32 = fused32;
  }
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_bp_set_fp32_geometry(bpx_bp b, int mode) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  return b->bp.SetFp32Geometry(mode);
  BPX_GUARD_END
}

int bpx_bp_set_cg_fold(bpx_bp b, int mask) {
  if (int st = need(b, "bp")) return st;
  if (mask < 0 || mask > 3) return set_error(BPX_ERR_INVALID_ARGUMENT, "bp: cg fold mask must be 0..3");
  b->bp.cgFold = mask;
  return BPX_OK;
}

int bpx_bp_set_ax_engine(bpx_bp b, int engine) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  if (engine != AX_ENGINE_AUTO && engine != AX_ENGINE_VECTOR && engine != AX_ENGINE_MATRIX)
    return set_error(BPX_ERR_INVALID_ARGUMENT, "bp: Ax engine must be 0 (auto), 1 (vector ALU) or 2 (matrix cores)");
  auto& bp = b->bp;
  if (engine == AX_ENGINE_MATRIX && (bp.bp != 5 || bp.Nfields != 1 || bp.mesh->N != 15))
    return set_error(BPX_ERR_UNSUPPORTED, "bp: the matrix-core Ax covers BP5 at N = 15 only");
  bp.axEngine = engine;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_bp_set_cub_form(bpx_bp b, int form) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  if (form < CUB_FORM_AUTO || form > CUB_FORM_PENCIL_ROWS)
    return set_error(BPX_ERR_INVALID_ARGUMENT,
                     "bp: cubature kernel form must be 0 (auto), 1 (pencil), 2 (wavefront) or 3 (pencil, row-by-row "
                     "contractions)");
  b->bp.cubForm = form;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_bp_time_operator(bpx_bp b, int precision, int launches, double* out) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  if (int st = need(out, "out")) return st;
  return b->bp.TimeOperator(precision, launches, out);
  BPX_GUARD_END
}

int bpx_bp_time_krhs(bpx_bp b, int K, int chunk, int launches, const double* o_q, double* out) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  if (int st = need(out, "out")) return st;
  return b->bp.TimeKrhs(K, chunk, launches, o_q, out);
  BPX_GUARD_END
}

int bpx_bp_operator_dot(bpx_bp b, const double* o_q, double* o_Aq, double* qAq) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  if (int st = need(qAq, "qAq")) return st;
  platform_t& P = *b->plat->p;
  const int n = b->bp.GatherPartials();
  if (b->dotPart.length() < (size_t)n)
    if (int st = b->dotPart.alloc(n)) return st;
  int np = 0;
  if (int st = b->bp.OperatorDot(o_q, o_Aq, b->dotPart.ptr(), &np)) return st;
  if (int st = launch_fold(P.stream, b->dotPart.ptr(), np, P.d_scalar)) return st;
  BPX_HIP(hipMemcpyAsync(P.h_scalar, P.d_scalar, sizeof(double), hipMemcpyDeviceToHost, P.stream));
  BPX_HIP(hipStreamSynchronize(P.stream));
  *qAq = *P.h_scalar;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_bp_ax_local(bpx_bp b, const double* o_q, double* o_AqL) {
  if (int st = need(b, "bp")) return st;
  return b->bp.AxLocal(o_q, o_AqL);
}

int bpx_bp_operator_f32(bpx_bp b, const float* o_q, float* o_Aq) {
  if (int st = need(b, "bp")) return st;
  return b->bp.Operator32(o_q, o_Aq);
}

int bpx_bp_operator_krhs(bpx_bp b, int K, int chunk, const double* o_q, double* o_Aq) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  return b->bp.OperatorKrhs(K, o_q, o_Aq, chunk);
  BPX_GUARD_END
}

int bpx_bp_set_kappa(bpx_bp b, const double* o_kappa) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  return b->bp.SetKappa(o_kappa);
  BPX_GUARD_END
}

int bpx_bp_kappa_gradient(bpx_bp b, const double* o_u, const double* o_lam, double* o_grad) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  if (!o_u || !o_lam || !o_grad) return set_error(BPX_ERR_INVALID_ARGUMENT, "kappa gradient: null pointer");
  return b->bp.KappaGradient(o_u, o_lam, o_grad);
  BPX_GUARD_END
}

int bpx_mesh_plot_fields(bpx_mesh m, int Nfields, const double* QL, const char* filename) {
  BPX_GUARD_BEGIN
  if (int st = need(m, "mesh")) return st;
  if (int st = need(filename, "filename")) return st;
  if (int st = m->m.EnsureHost()) return st;
  return m->m.PlotFields(QL, Nfields, filename);
  BPX_GUARD_END
}

int bpx_bp_plot_solution(bpx_bp b, const double* o_u, const char* filename) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  if (int st = need(o_u, "o_u")) return st;
  if (int st = need(filename, "filename")) return st;
  return b->bp.PlotSolution(o_u, filename);
  BPX_GUARD_END
}

int bpx_bp_error_norms(bpx_bp b, const double* o_u, double* out) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  if (int st = need(o_u, "o_u")) return st;
  if (int st = need(out, "out")) return st;
  return b->bp.ErrorNorms(o_u, out);
  BPX_GUARD_END
}

int bpx_bp_rhs_forcing(bpx_bp b, double* o_b) {
  if (int st = need(b, "bp")) return st;
  return b->bp.Forcing(o_b);
}

int bpx_bp_solve(bpx_bp b, double* o_x, double* o_r, double tol, int maxit, int verbose, int* iterations,
                 double* rnorm_hist) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  int its = 0;
  const int st = b->bp.SolveCG(*b->solver, o_x, o_r, tol, maxit, verbose, &its);
  if (iterations) *iterations = its;
  if (rnorm_hist) {
    std::vector<double> rr, al;
    if (int s2 = b->solver->impl.History(rr, al)) return s2;
    for (int i = 0; i <= maxit; ++i) rnorm_hist[i] = (i < (int)rr.size()) ? std::sqrt(rr[i]) : 0.0;
  }
  return st;
  BPX_GUARD_END
}

int bpx_bp_solve_mixed(bpx_bp b, double* o_x, const double* o_b, double tol, int interval, int maxit,
                       double* stats) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  return b->bp.SolveMixed(o_x, o_b, tol, interval, maxit, stats);
  BPX_GUARD_END
}

int bpx_bp_time_operator_in_solve(bpx_bp b, double* o_x, const double* o_b, double tol, int interval, int maxit,
                                  double* out) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "bp")) return st;
  if (int st = need(out, "out")) return st;
  return b->bp.TimeOperatorInSolve(o_x, o_b, tol, interval, maxit, out);
  BPX_GUARD_END
}

// ----------------------------------------------------------- batch ----
int bpx_batch_create(bpx_platform p, int nprob, const bpx_mesh_opts* opts, double lambda, bpx_batch* out) {
  BPX_GUARD_BEGIN
  if (int st = need(p, "platform")) return st;
  if (int st = need(opts, "opts")) return st;
  if (int st = need(out, "out")) return st;
  if (nprob < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "batch: nprob < 1");
  std::vector<MeshOpts> v(nprob);
  for (int i = 0; i < nprob; ++i)
    if (int st = to_mesh_opts(opts[i], v[i])) return st;
  bpx_batch_s* b = new bpx_batch_s;
  b->plat = p;
  if (int st = b->b.Setup(*p->p, v, lambda)) {
    delete b;
    return st;
  }
  *out = b;
  return BPX_OK;
  BPX_GUARD_END
}

int bpx_batch_destroy(bpx_batch b) {
  delete b;
  return BPX_OK;
}

int bpx_batch_info(bpx_batch b, int64_t* out, int capacity) {
  if (int st = need(b, "batch")) return st;
  const int64_t v[] = {b->b.nprob, b->b.Ntot, (int64_t)b->b.groups.size(), b->b.nchunks, b->b.nsupers,
                       b->b.NLocalTot, b->b.replays};
  const int n = 7;
  for (int i = 0; i < n && i < capacity; ++i) out[i] = v[i];
  return n;
}

int bpx_batch_offsets(bpx_batch b, int64_t* off) {
  if (int st = need(b, "batch")) return st;
  if (int st = need(off, "off")) return st;
  for (int i = 0; i <= b->b.nprob; ++i) off[i] = b->b.off[i];
  return BPX_OK;
}

int bpx_batch_tuning(bpx_batch b, double* out, int capacity) {
  if (int st = need(b, "batch")) return st;
  const int n = 6 * (int)b->b.groups.size();
  for (size_t g = 0; g < b->b.groups.size(); ++g) {
    const auto& G = b->b.groups[g];
    const double v[6] = {(double)G.N, (double)G.variant, (double)bp5_batch_epb(G.Nq, G.variant), G.us[0], G.us[1],
                         G.us[2]};
    for (int k = 0; k < 6; ++k)
      if ((int)(6 * g + k) < capacity) out[6 * g + k] = v[k];
  }
  return n;
}

int bpx_batch_solve(bpx_batch b, double* o_x, const double* o_b, double tol, int maxit, int* iterations,
                    double* rnorm) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "batch")) return st;
  return b->b.Solve(o_x, o_b, tol, maxit, iterations, rnorm, nullptr);
  BPX_GUARD_END
}

int bpx_batch_solve_mixed(bpx_batch b, double* o_x, const double* o_b, double tol, int interval, int maxit,
                          int* iterations, int* updates, double* relres) {
  BPX_GUARD_BEGIN
  if (int st = need(b, "batch")) return st;
  return b->b.SolveMixed(o_x, o_b, tol, interval, maxit, iterations, updates, relres, nullptr);
  BPX_GUARD_END
}

}  // extern "C"

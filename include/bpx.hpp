/*
 * bpx.hpp — C++ interface of libbpx with the class shapes of the
 * reference's host API, header-only over the C ABI of bpx.h, so C++ host
 * code written against benchParanumal's platform_t / mesh_t / ogs_t /
 * halo_t / linAlg_t / solver_t / linearSolver_t / cg / bpN_t reads the same
 * and needs no internal header of the library.
 *
 *  - platform_t      include/platform.hpp:61-185 (device, stream, malloc)
 *  - deviceMemory<T> include/memory.hpp:288-460 (ref-counted, offsets)
 *  - linAlg_t        include/linAlg.hpp:38-115 (set, axpy, innerProd, norm2)
 *  - mesh_t          include/mesh.hpp (hex box setup, PlotFields)
 *  - ogs::ogs_t      include/ogs.hpp:216-341 (Gather, Scatter, GatherScatter)
 *  - ogs::halo_t     include/ogs.hpp:355-379 (empty on one GPU)
 *  - solver_t        include/solver.hpp:34-50 (the Operator a solver applies)
 *  - linearSolver_t, cg  include/linearSolver.hpp:37-81
 *  - bp_t            BP/BP5/bp5.hpp:43-74 and its BP1-BP6 siblings
 *
 * Errors: as the reference's LIBP_ABORT (libs/core/exception.cpp:75-79),
 * every failure throws libbpx::exception (code = the bpx_status). Device
 * work is queued on the platform's stream; calls that return host values
 * synchronise it.
 */
#ifndef BPX_HPP
#define BPX_HPP

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "bpx.h"

namespace libbpx {

typedef double dfloat;
typedef int dlong;
typedef int64_t hlong;

class exception : public std::runtime_error {
 public:
  exception(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int status) {
  if (status != BPX_OK) throw exception(status, bpx_last_error());
}

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw exception(BPX_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------ memory ----
// deviceMemory<T>: a ref-counted device allocation, or a view of one at an
// offset (operator+, as memory.hpp:323-327), or a non-owning view of a
// caller's pointer.
template <typename T>
class deviceMemory {
 public:
  deviceMemory() = default;
  static deviceMemory view(T* p, size_t n) {
    deviceMemory m;
    m.base_ = std::shared_ptr<T>(p, [](T*) {});
    m.length_ = n;
    return m;
  }
  static deviceMemory allocate(size_t n) {
    deviceMemory m;
    void* p = nullptr;
    if (n) check_hip(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
    m.base_ = std::shared_ptr<T>(static_cast<T*>(p), [](T* q) {
      if (q) (void)hipFree(q);
    });
    m.length_ = n;
    return m;
  }
  T* ptr() const { return base_.get() + offset_; }
  size_t length() const { return length_; }
  deviceMemory operator+(size_t off) const {
    if (off > length_) throw exception(BPX_ERR_INVALID_ARGUMENT, "deviceMemory: offset past the end");
    deviceMemory m(*this);
    m.offset_ += off;
    m.length_ -= off;
    return m;
  }
  // synchronous copies (the host side may be a temporary)
  void copyFrom(const T* src, size_t n) {
    if (n > length_) throw exception(BPX_ERR_INVALID_ARGUMENT, "deviceMemory: copy larger than the buffer");
    if (n) check_hip(hipMemcpy(ptr(), src, n * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  }
  void copyFrom(const std::vector<T>& v) { copyFrom(v.data(), v.size()); }
  void copyTo(T* dst, size_t n) const {
    if (n > length_) throw exception(BPX_ERR_INVALID_ARGUMENT, "deviceMemory: copy larger than the buffer");
    if (n) check_hip(hipMemcpy(dst, ptr(), n * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy");
  }
  void copyTo(std::vector<T>& v) const {
    v.resize(length_);
    copyTo(v.data(), length_);
  }

 private:
  std::shared_ptr<T> base_;
  size_t offset_ = 0, length_ = 0;
};

// ---------------------------------------------------------- platform ----
class linAlg_t;

class platform_t {
 public:
  explicit platform_t(int device = 0) { check(bpx_platform_create(device, &h_)); }
  ~platform_t() { bpx_platform_destroy(h_); }
  platform_t(const platform_t&) = delete;
  platform_t& operator=(const platform_t&) = delete;

  template <typename T>
  deviceMemory<T> malloc(size_t n) const {
    return deviceMemory<T>::allocate(n);
  }
  template <typename T>
  deviceMemory<T> malloc(const std::vector<T>& host) const {
    deviceMemory<T> m = deviceMemory<T>::allocate(host.size());
    m.copyFrom(host);
    return m;
  }
  void finish() const { check(bpx_platform_finish(h_)); }
  hipStream_t stream() const { return static_cast<hipStream_t>(bpx_platform_stream(h_)); }
  void setStream(hipStream_t s) { check(bpx_platform_set_stream(h_, s)); }
  inline linAlg_t linAlg() const;
  bpx_platform handle() const { return h_; }

 private:
  bpx_platform h_ = nullptr;
};

// ------------------------------------------------------------ linAlg ----
class linAlg_t {
 public:
  explicit linAlg_t(bpx_platform p) : p_(p) {}
  void set(dlong N, dfloat alpha, deviceMemory<dfloat> o_x) const { check(bpx_linalg_set(p_, N, alpha, o_x.ptr())); }
  void axpy(dlong N, dfloat alpha, deviceMemory<dfloat> o_x, dfloat beta, deviceMemory<dfloat> o_y) const {
    check(bpx_linalg_axpy(p_, N, alpha, o_x.ptr(), beta, o_y.ptr()));
  }
  dfloat innerProd(dlong N, deviceMemory<dfloat> o_x, deviceMemory<dfloat> o_y) const {
    double r = 0;
    check(bpx_linalg_inner_prod(p_, N, o_x.ptr(), o_y.ptr(), &r));
    return r;
  }
  dfloat norm2(dlong N, deviceMemory<dfloat> o_x) const {
    double r = 0;
    check(bpx_linalg_norm2(p_, N, o_x.ptr(), &r));
    return r;
  }

 private:
  bpx_platform p_;
};

inline linAlg_t platform_t::linAlg() const { return linAlg_t(h_); }

// -------------------------------------------------------------- mesh ----
class mesh_t {
 public:
  // the hex box: N, nx, ny, nz, affine, cubature, lo, dim, perturb, seed
  explicit mesh_t(const bpx_mesh_opts& opts) {
    check(bpx_mesh_create(&opts, &h_));
    int64_t v[6];
    bpx_mesh_info(h_, v, 6);
    N = (int)v[0];
    Nq = (int)v[1];
    Np = (int)v[2];
    cubNq = (int)v[3];
    cubNp = (int)v[4];
    Nelements = (dlong)v[5];
  }
  ~mesh_t() { bpx_mesh_destroy(h_); }
  mesh_t(const mesh_t&) = delete;
  mesh_t& operator=(const mesh_t&) = delete;
  // host copy of a mesh array (x, y, z, wJ, ggeo, D, gllw, ...)
  std::vector<double> array(const char* name) const {
    int64_t n = 0;
    int dt = 0;
    check(bpx_mesh_array(h_, name, nullptr, &n, &dt));
    if (dt != 0) throw exception(BPX_ERR_INVALID_ARGUMENT, std::string("mesh: not a double array: ") + name);
    std::vector<double> v((size_t)n);
    check(bpx_mesh_array(h_, name, v.data(), &n, &dt));
    return v;
  }
  // mesh_t::PlotFields (libs/mesh/meshPlotFields3D.cpp): host data [e][f][n]
  void PlotFields(const std::vector<dfloat>& QL, int Nfields, const std::string& fileName) const {
    check(bpx_mesh_plot_fields(h_, Nfields, QL.data(), fileName.c_str()));
  }
  bpx_mesh handle() const { return h_; }
  int N = 0, Nq = 0, Np = 0, cubNq = 0, cubNp = 0;
  dlong Nelements = 0;

 private:
  bpx_mesh h_ = nullptr;
};

// --------------------------------------------------------------- ogs ----
namespace ogs {

class ogs_t {
 public:
  ogs_t() = default;
  ~ogs_t() { bpx_ogs_destroy(h_); }
  ogs_t(const ogs_t&) = delete;
  ogs_t& operator=(const ogs_t&) = delete;
  // ogs_t::Setup on one rank, Signed, unique (ids[n] = 0: masked)
  void Setup(dlong N_, const hlong* ids, platform_t& platform) {
    reset(platform);
    check(bpx_ogs_setup(N_, ids, &h_));
    info();
  }
  // mesh_t::MaskedGatherScatterSetup(Nfields) followed by Setup
  void SetupMasked(mesh_t& mesh, int Nfields, platform_t& platform) {
    reset(platform);
    check(bpx_ogs_setup_masked(mesh.handle(), Nfields, &h_));
    info();
  }
  // ogs_t::SetupGlobalToLocalMapping (libs/ogs/ogsSetup.cpp:833-849)
  void SetupGlobalToLocalMapping(std::vector<dlong>& GlobalToLocal) const {
    int64_t n = 0;
    int dt = 0;
    check(bpx_ogs_array(h_, "GlobalToLocal", nullptr, &n, &dt));
    GlobalToLocal.resize((size_t)n);
    check(bpx_ogs_array(h_, "GlobalToLocal", GlobalToLocal.data(), &n, &dt));
  }
  // Gather(o_gv, o_v, k, Add, Trans), Scatter(o_v, o_gv, k, NoTrans),
  // GatherScatter(o_v, k, Add, Sym)
  void Gather(deviceMemory<dfloat> o_gv, deviceMemory<dfloat> o_v, int k = 1) const {
    check(bpx_ogs_gather_dev(p_, h_, k, o_v.ptr(), o_gv.ptr()));
  }
  void Scatter(deviceMemory<dfloat> o_v, deviceMemory<dfloat> o_gv, int k = 1) const {
    check(bpx_ogs_scatter_dev(p_, h_, k, o_gv.ptr(), o_v.ptr()));
  }
  void GatherScatter(deviceMemory<dfloat> o_v, int k = 1) const {
    check(bpx_ogs_gather_scatter_dev(p_, h_, k, o_v.ptr()));
  }
  dlong N = 0, Ngather = 0;
  bpx_ogs handle() const { return h_; }

 private:
  void reset(platform_t& platform) {
    bpx_ogs_destroy(h_);
    h_ = nullptr;
    p_ = platform.handle();
  }
  void info() {
    int64_t v[4];
    bpx_ogs_info(h_, v, 4);
    N = (dlong)v[0];
    Ngather = (dlong)v[1];
  }
  bpx_ogs h_ = nullptr;
  bpx_platform p_ = nullptr;
};

// halo_t: the exchange of ghost values between ranks. One GPU owns every
// node, so the halo is empty and each exchange moves nothing (the
// reference posts no messages on a single rank, libs/ogs/ogsHalo.cpp:46-145).
class halo_t {
 public:
  void SetupFromGather(const ogs_t& o) {
    Nlocal = o.N;
    Nhalo = 0;
    NhaloP = 0;
  }
  template <typename T>
  void Exchange(deviceMemory<T> o_v, const int k) const {
    (void)o_v;
    (void)k;
  }
  template <typename T>
  void ExchangeStart(deviceMemory<T> o_v, const int k) const {
    Exchange(o_v, k);
  }
  template <typename T>
  void ExchangeFinish(deviceMemory<T> o_v, const int k) const {
    Exchange(o_v, k);
  }
  dlong Nlocal = 0, Nhalo = 0, NhaloP = 0;
};

}  // namespace ogs

// ------------------------------------------------------ solver / CG -----
class solver_t {
 public:
  virtual ~solver_t() = default;
  // o_Aq = A o_q, queued on the platform's stream
  virtual void Operator(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_Aq) = 0;
};

class linearSolver_t {
 public:
  virtual ~linearSolver_t() = default;
  // returns the iterations taken; o_rhs holds the residual on exit
  virtual int Solve(solver_t& solver, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_rhs, const dfloat tol,
                    const int MAXIT, const int verbose) = 0;
};

class cg : public linearSolver_t {
 public:
  cg(platform_t& platform, dlong N, dlong Nhalo) : N_(N + Nhalo) {
    check(bpx_cg_create(platform.handle(), N, Nhalo, &h_));
  }
  ~cg() override { bpx_cg_destroy(h_); }
  cg(const cg&) = delete;
  cg& operator=(const cg&) = delete;

  int Solve(solver_t& solver, deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_rhs, const dfloat tol,
            const int MAXIT, const int verbose) override {
    Call call{&solver, (size_t)N_, nullptr};
    history_.assign((size_t)MAXIT + 1, 0.0);
    int its = 0;
    const int st = bpx_cg_solve(h_, &cg::trampoline, &call, o_x.ptr(), o_rhs.ptr(), tol, MAXIT, verbose, &its,
                                history_.data());
    if (call.error) std::rethrow_exception(call.error);
    check(st);
    history_.resize((size_t)its + 1);
    return its;
  }
  // residual norms of the last Solve (entry 0 = initial)
  const std::vector<double>& residualHistory() const { return history_; }

 private:
  struct Call {
    solver_t* solver;
    size_t n;
    std::exception_ptr error;
  };
  // solver_t::Operator behind the C callback; an exception is kept and
  // rethrown once the solve has returned (nothing throws across the ABI)
  static int trampoline(void* ctx, const double* o_q, double* o_Aq) {
    Call* c = static_cast<Call*>(ctx);
    try {
      deviceMemory<dfloat> q = deviceMemory<dfloat>::view(const_cast<double*>(o_q), c->n);
      deviceMemory<dfloat> Aq = deviceMemory<dfloat>::view(o_Aq, c->n);
      c->solver->Operator(q, Aq);
      return 0;
    } catch (...) {
      c->error = std::current_exception();
      return BPX_ERR_INVALID_ARGUMENT;
    }
  }
  bpx_cg h_ = nullptr;
  dlong N_ = 0;
  std::vector<double> history_;
};

// ---------------------------------------------------------------- BP ----
// bpN_t for N = 1..6 on a hex box mesh: its masked ogs, GlobalToLocal and
// operator data live on the device; Operator is bpN_t::Operator.
class bp_t : public solver_t {
 public:
  bp_t(platform_t& platform, mesh_t& mesh, int bp, dfloat lambda) : platform_(platform) {
    check(bpx_bp_create(platform.handle(), mesh.handle(), bp, lambda, &h_));
    int64_t v[7];
    bpx_bp_info(h_, v, 7);
    Nfields = (int)v[1];
    Ngather = (dlong)v[2];
    NLocal = (dlong)v[3];
  }
  ~bp_t() override { bpx_bp_destroy(h_); }
  bp_t(const bp_t&) = delete;
  bp_t& operator=(const bp_t&) = delete;

  void Operator(deviceMemory<dfloat>& o_q, deviceMemory<dfloat>& o_Aq) override {
    check(bpx_bp_operator(h_, o_q.ptr(), o_Aq.ptr()));
  }
  // b = gather(forcing) (BP/BP5/src/bp5Run.cpp:51-64)
  void Forcing(deviceMemory<dfloat>& o_b) { check(bpx_bp_rhs_forcing(h_, o_b.ptr())); }
  // the library's own CG on this operator (solve-space rows, fused
  // operator + dot); rnorm gets MAXIT + 1 residual norms
  int Solve(deviceMemory<dfloat>& o_x, deviceMemory<dfloat>& o_r, dfloat tol, int MAXIT,
            std::vector<double>* rnorm = nullptr) {
    int its = 0;
    if (rnorm) rnorm->assign((size_t)MAXIT + 1, 0.0);
    check(bpx_bp_solve(h_, o_x.ptr(), o_r.ptr(), tol, MAXIT, 0, &its, rnorm ? rnorm->data() : nullptr));
    if (rnorm) rnorm->resize((size_t)its + 1);
    return its;
  }
  void PlotSolution(deviceMemory<dfloat>& o_u, const std::string& fileName) {
    check(bpx_bp_plot_solution(h_, o_u.ptr(), fileName.c_str()));
  }
  int Nfields = 1;
  dlong Ngather = 0, NLocal = 0;
  bpx_bp handle() const { return h_; }

 private:
  platform_t& platform_;
  bpx_bp h_ = nullptr;
};

}  // namespace libbpx

#endif /* BPX_HPP */

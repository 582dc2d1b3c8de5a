// C++ host layer of the BK/BP hot path. Class and method names follow the
// reference interfaces the path sits behind (platform_t, mesh_t,
// ogs::ogs_t, linAlg_t, solver_t, linearSolver_t/cg, bpN_t) so host code
// written against benchParanumal reads the same; the device work goes
// through the HIP launchers of bpx_internal.hpp.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <algorithm>
#include <deque>
#include <vector>

#include "../bpx_internal.hpp"

namespace bpx {

typedef double dfloat;

// ------------------------------------------------------------ device ----
// Owning device allocation (the reference's deviceMemory<T> from
// platform_t::malloc, include/memory.hpp:288-460, without offsets).
template <typename T>
class deviceBuffer {
 public:
  deviceBuffer() = default;
  ~deviceBuffer() { release(); }
  deviceBuffer(const deviceBuffer&) = delete;
  deviceBuffer& operator=(const deviceBuffer&) = delete;
  deviceBuffer(deviceBuffer&& o) noexcept : ptr_(o.ptr_), n_(o.n_) {
    o.ptr_ = nullptr;
    o.n_ = 0;
  }
  deviceBuffer& operator=(deviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ptr_ = o.ptr_;
      n_ = o.n_;
      o.ptr_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  int alloc(size_t n) {
    release();
    if (n == 0) return BPX_OK;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e != hipSuccess) return hip_error(e, "hipMalloc");
    ptr_ = static_cast<T*>(p);
    n_ = n;
    return BPX_OK;
  }
  // synchronous copy: the host source may be a temporary
  int upload(const T* h, size_t n, hipStream_t s) {
    (void)s;
    if (int st = alloc(n)) return st;
    if (n == 0) return BPX_OK;
    BPX_HIP(hipMemcpy(ptr_, h, n * sizeof(T), hipMemcpyHostToDevice));
    return BPX_OK;
  }
  int upload(const std::vector<T>& v, hipStream_t s) { return upload(v.data(), v.size(), s); }
  // device-to-device copy of n values, queued on s
  int copy_device(const T* d, size_t n, hipStream_t s) {
    if (int st = alloc(n)) return st;
    if (n == 0) return BPX_OK;
    BPX_HIP(hipMemcpyAsync(ptr_, d, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    return BPX_OK;
  }
  // synchronous copy of the whole buffer into h (resized to length())
  int download(std::vector<T>& h, hipStream_t s) const {
    h.resize(n_);
    if (n_ == 0) return BPX_OK;
    BPX_HIP(hipMemcpyAsync(h.data(), ptr_, n_ * sizeof(T), hipMemcpyDeviceToHost, s));
    BPX_HIP(hipStreamSynchronize(s));
    return BPX_OK;
  }
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    n_ = 0;
  }
  T* ptr() const { return ptr_; }
  size_t length() const { return n_; }

 private:
  T* ptr_ = nullptr;
  size_t n_ = 0;
};

// ---------------------------------------------------------- platform ----
// platform_t (include/platform.hpp:61-185): one device, one stream, the
// linAlg scratch (reduction partials) and a pinned host scalar.
class platform_t {
 public:
  ~platform_t();
  static int Create(int device, platform_t** out);
  int finish();
  int setStream(hipStream_t s);

  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  double* partials = nullptr;  // kRedBlocks reduction partials
  double* d_scalar = nullptr;  // device result slot
  double* h_scalar = nullptr;  // pinned host copy
};

// ------------------------------------------------------------ linAlg ----
// linAlg_t (include/linAlg.hpp:38-115; libs/core/linAlg.cpp:38-89).
class linAlg_t {
 public:
  explicit linAlg_t(platform_t& p) : platform(p) {}
  int set(dlong N, dfloat alpha, dfloat* o_x);
  int axpy(dlong N, dfloat alpha, const dfloat* o_x, dfloat beta, dfloat* o_y);
  int innerProd(dlong N, const dfloat* o_x, const dfloat* o_y, dfloat* result);  // synchronises
  int norm2(dlong N, const dfloat* o_x, dfloat* result);                           // synchronises
  platform_t& platform;
};

// -------------------------------------------------------------- mesh ----
// element families, by the reference's codes (libs/mesh/meshSetElementType.cpp)
enum { ELEM_TRI = 3, ELEM_QUAD = 4, ELEM_TET = 6, ELEM_HEX = 12 };

struct MeshOpts {
  int N = 1, nx = 1, ny = 1, nz = 1;
  int affine = 0, cubature = 0;
  double lo[3] = {-1, -1, -1}, dim[3] = {2, 2, 2};
  double perturb = 0.0;
  uint64_t seed = 0;
  int elementType = ELEM_HEX;  // 0 is taken as ELEM_HEX
};

// mesh_t for the hex box (libs/mesh/meshSetup.cpp:31-85), host data.
class mesh_t {
 public:
  int Setup(const MeshOpts& o);
  // the part of Setup that stays on the host in both setups: checks, 1D
  // operators, element vertices
  int SetupVertices(const MeshOpts& o);
  // mesh_t::MaskedGatherScatterSetup ids (meshMaskedGatherScatterSetup.cpp:31-77)
  void MaskedGlobalIds(int Nfields, std::vector<hlong>& ids) const;
  // per element, for x, y and z in turn, the coefficients of r, s, t, rs,
  // rt, st, rst of the trilinear map through its 8 vertices, and one pad
  // (24 values): the nodal coordinates are this map at the GLL nodes
  void TrilinearCoefficients(std::vector<double>& c) const;
  // whether every local node lies on its element's trilinear map of the 8
  // vertices at the GLL point (to 1e-12 of the box size): then the
  // geometric factors are the map's exact derivatives
  bool NodesAreTrilinear() const;
  // mesh_t::PlotFields (libs/mesh/meshPlotFields3D.cpp:31-121): VTU file of
  // the local nodes with Nfields values each (QL: host, [e][f][n]); each
  // element is drawn as N^3 sub-cubes of 6 tetrahedra
  int PlotFields(const double* QL, int Nfields, const std::string& fileName) const;

  // GPU-side setup: the same mesh as Setup, with the nodes, global ids and
  // geometric and cubature factors formed on the device (bit-identical to
  // Setup's host arrays) into the o_ buffers below. The vertices and the 1D
  // operators stay on the host; the per-node host arrays are left empty
  // until EnsureHost copies them down.
  int SetupDevice(platform_t& p, const MeshOpts& o);
  int EnsureHost();
  // whether every node is its element's trilinear map (NodesAreTrilinear),
  // checked on the device for a device mesh
  bool NodesTrilinear();
  // masked ids (MaskedGlobalIds) formed on the device from o_globalIds
  int MaskedGlobalIdsDevice(int Nfields, hlong* o_ids) const;
  bool onDevice = false;
  platform_t* devPlatform = nullptr;
  deviceBuffer<double> o_x, o_y, o_z, o_wJ, o_ggeo, o_cubwJ, o_cubggeo;
  deviceBuffer<hlong> o_globalIds;

  // quad, triangle and tetrahedral boxes (host/mesh_other.cpp)
  int SetupOther(const MeshOpts& o);
  int elementType = ELEM_HEX, dim = 3, Nverts = 8, Nggeo = 6;
  // triangles and tetrahedra: reference nodes, their lattice indices, the
  // packed D (dim x Np x Np) and stiffness matrices (3 or 6 x Np x Np); MM
  // holds their Np x Np mass matrix (1D mass for hex and quad)
  std::vector<double> rN, sN, tN, Dsim, Ssim;
  std::vector<int> nodeIdx;
  // non-hex meshes: 1 for a local node on the box boundary
  std::vector<char> boundary;

  MeshOpts opts;
  int N = 0, Nq = 0, Np = 0, cubNq = 0, cubNp = 0;
  dlong Nelements = 0;
  std::vector<double> gllz, gllw, D, MM, cubr, cubw, cubInterp, cubD, invV;
  std::vector<double> EX, EY, EZ, x, y, z;
  std::vector<double> wJ, ggeo, cubwJ, cubggeo;
  std::vector<hlong> globalIds;
};

// ---------------------------------------------------------------- ogs ---
namespace ogs {
// ogs::ogs_t (include/ogs.hpp:216-341) on one rank: Signed, unique, the
// T (transpose) maps only; the N maps pick a random member per group in
// the reference (ogsSetup.cpp:361-382) and are off the path.
class ogs_t {
 public:
  int Setup(dlong N, const hlong* ids);
  int SetupRows(dlong N, dlong Nrows, const dlong* rowOf);
  void SetupGlobalToLocalMapping(dlong* GlobalToLocal) const;
  int Upload(platform_t& p);  // device copies of the T maps
  // Setup and SetupRows on the device (device ids / rowOf; the same maps,
  // left in the o_ buffers; host vectors empty until EnsureHost). o_G2L
  // (device, N entries, may be null) gets GlobalToLocal.
  int SetupDevice(platform_t& p, dlong N, const hlong* o_ids, dlong* o_G2L);
  int SetupRowsDevice(platform_t& p, dlong N, dlong Nrows, const dlong* o_rowOf);
  int EnsureHost(platform_t& p);
  bool onDevice = false;
  template <typename T>
  int Gather(platform_t& p, T* o_gv, const T* o_v, int K) const;  // Add, Trans
  template <typename T>
  int Scatter(platform_t& p, T* o_v, const T* o_gv, int K) const;  // NoTrans
  int GatherScatter(platform_t& p, dfloat* o_v, int K) const;

  dlong N = 0, Ngather = 0, nnzT = 0, NrowBlocksT = 0;
  int gatherMode = 0;  // workgroup order of the gather (set by the tuner; same results either way)
  std::vector<dlong> rowStartsT, colIdsT, blockRowStartsT;
  deviceBuffer<dlong> o_rowStartsT, o_colIdsT, o_blockRowStartsT;
};

// halo_t (include/ogs.hpp:355-379): the exchange of ghost values of q
// between ranks before the elements on a rank boundary apply Ax
// (BP/BP5/src/bp5Operator.cpp:31-50). On one GPU every node is owned, so
// the halo is empty (Nhalo = 0) and each exchange moves nothing, as
// libs/ogs/ogsHalo.cpp:46-145 posts no messages on a single rank. The calls
// stay in the operator so the multi-GPU exchange slots in at the same
// points.
class halo_t {
 public:
  int SetupFromGather(const ogs_t& o) {
    Nlocal = o.N;
    Nhalo = 0;
    NhaloP = 0;
    return BPX_OK;
  }
  template <typename T>
  int Exchange(T* o_v, int k) const {
    (void)o_v;
    (void)k;
    return BPX_OK;
  }
  template <typename T>
  int ExchangeStart(T* o_v, int k) const {
    return Exchange(o_v, k);
  }
  template <typename T>
  int ExchangeFinish(T* o_v, int k) const {
    return Exchange(o_v, k);
  }
  dlong Nlocal = 0, Nhalo = 0, NhaloP = 0;
};
}  // namespace ogs

// ------------------------------------------------------ solver / CG -----
// solver_t (include/solver.hpp:34-50)
class solver_t {
 public:
  virtual ~solver_t() = default;
  virtual int Operator(const dfloat* o_q, dfloat* o_Aq) = 0;
  // Optional: Operator fused with the FP64 partials of q.Aq (used by cg).
  virtual bool HasOperatorDot() const { return false; }
  virtual int OperatorDot(const dfloat* o_q, dfloat* o_Aq, double* part, int* npart) {
    (void)o_q; (void)o_Aq; (void)part; (void)npart;
    return set_error(BPX_ERR_UNSUPPORTED, "solver: no fused operator + dot");
  }
  // Optional: OperatorDot that also does the CG direction update of the
  // rows from FoldRows() on (see CgFold); 0 rows = not offered.
  virtual dlong FoldRows() const { return 0; }
  virtual int OperatorDotFold(const CgFold<dfloat>& fold, dfloat* o_Aq, double* part, int* npart) {
    (void)fold; (void)o_Aq; (void)part; (void)npart;
    return set_error(BPX_ERR_UNSUPPORTED, "solver: no folded CG operator");
  }
};

// linearSolver_t (include/linearSolver.hpp:37-81)
class linearSolver_t {
 public:
  virtual ~linearSolver_t() = default;
  virtual int Solve(solver_t& solver, dfloat* o_x, dfloat* o_r, dfloat tol, int MAXIT, int verbose,
                    int* iterations) = 0;
};

// Device work buffers of one CG solve: p and Ap (N entries), the state,
// kRedBlocks partials of r.r, the p.Ap partials (as many as the operator
// writes, at least kRedBlocks) and MAXIT + 1 entries each of the r.r and
// alpha histories.
template <typename T>
struct CgWork {
  T* p = nullptr;
  T* Ap = nullptr;
  CgState* state = nullptr;
  double* part_rr = nullptr;
  double* part_pAp = nullptr;
  double* hist_rr = nullptr;
  double* hist_alpha = nullptr;
};

// cg (libs/core/linearSolverCG.cpp:33-156) with device-resident scalars.
// T is the vector/operator precision; reductions accumulate in FP64.
template <typename T>
class cg_t {
 public:
  // apply(o_q, o_Aq): the operator in precision T.
  typedef std::function<int(const T*, T*)> Apply;
  // apply_dot(o_q, o_Aq, part, &npart): the operator fused with the p.Ap
  // partial sums (npart FP64 partials written to part); may be empty.
  typedef std::function<int(const T*, T*, double*, int*)> ApplyDot;
  // apply_fold(fold, o_Ap, part, &npart): apply_dot on fold.p, with the
  // direction update of iteration fold.k folded in for the rows from
  // fold_rows on (k_cg_update_p then covers the rows below fold_rows)
  typedef std::function<int(const CgFold<T>&, T*, double*, int*)> ApplyFold;

  // npart_max: most partials apply_dot writes (at least kRedBlocks is kept)
  int Setup(platform_t& p, dlong N, int npart_max = 0);
  int Solve(const Apply& apply, const ApplyDot& apply_dot, T* o_x, T* o_r, double tol, int MAXIT,
            int* iterations, const ApplyFold& apply_fold = ApplyFold(), dlong fold_rows = 0);
  // squared residual norms r.r of the last Solve (entry 0 = initial)
  int History(std::vector<double>& rr, std::vector<double>& alpha);

  platform_t* platform = nullptr;
  dlong N = 0;
  deviceBuffer<T> o_p, o_Ap;
  deviceBuffer<CgState> o_state;
  deviceBuffer<double> o_part_rr, o_part_pAp;
  deviceBuffer<double> o_hist_rr, o_hist_alpha;
  int hist_len = 0;
};

// The CG iteration of cg_t::Solve on caller-provided work buffers (queued
// on stream s; allocates nothing).
template <typename T>
int cg_run(hipStream_t s, dlong N, const CgWork<T>& w, const typename cg_t<T>::Apply& apply,
           const typename cg_t<T>::ApplyDot& apply_dot, T* o_x, T* o_r, double tol, int MAXIT, int* iterations,
           const typename cg_t<T>::ApplyFold& apply_fold = typename cg_t<T>::ApplyFold(), dlong fold_rows = 0);

class cg : public linearSolver_t {
 public:
  cg(platform_t& p, dlong N, dlong Nhalo, int npart_max = 0);
  int status() const { return status_; }
  int Solve(solver_t& solver, dfloat* o_x, dfloat* o_r, dfloat tol, int MAXIT, int verbose,
            int* iterations) override;
  cg_t<double> impl;

 private:
  int status_ = BPX_OK;
};

// ------------------------------------------------------------------ BP --
// bpN_t (BP/BP5/bp5.hpp:43-74) for N = 1..6 on the hex box.
class bp_t : public solver_t {
 public:
  int Setup(platform_t& p, mesh_t& m, int bp, dfloat lambda);
  int Operator(const dfloat* o_q, dfloat* o_Aq) override;
  // Operator with the gather fused with the q.Aq partial sums
  bool HasOperatorDot() const override { return true; }
  int OperatorDot(const dfloat* o_q, dfloat* o_Aq, double* part, int* npart) override;
  int GatherPartials() const {
    return std::max({ogs_gather_partials(ogs.NrowBlocksT), ogs_gather_partials(sharedOgs.NrowBlocksT),
                     ogs_gather_partials(solveOgs.NrowBlocksT), 1});
  }

  // Solve space: the row order the CG solvers run in. For BP5 with one
  // field the rows are renumbered: rows with two or more local copies
  // first (in reference order), then, element by element, the element's
  // interior nodes (in local node order), each run padded to 32 rows (see
  // fused_interior_pad). The fused operator's Ax then writes whole cache
  // lines of Ap and the shared-row gather writes the rest; in the
  // reference order both would write parts of the same lines, and a
  // partial line costs HBM a read-modify-write. Padding rows belong to no
  // node and stay zero. Other problems keep the reference order
  // (permuted = false, Ns = Ngather).
  bool permuted = false;
  dlong Ns = 0;
  deviceBuffer<dlong> o_sperm;  // solve row -> reference row (-1 padding)
  deviceBuffer<dlong> o_G2Ls;   // local node -> solve row (-1 masked)
  ogs::ogs_t solveOgs;          // the gather onto all solve rows
  const dlong* SolveG2L() const { return permuted ? o_G2Ls.ptr() : o_GlobalToLocal.ptr(); }
  const ogs::ogs_t& SolveOgs() const { return permuted ? solveOgs : ogs; }
  // reference order <-> solve order (Ngather <-> Ns vectors)
  int ToSolve(const dfloat* o_ref, dfloat* o_s);
  int FromSolve(const dfloat* o_s, dfloat* o_ref);
  // the operators on solve-space vectors: plain (FP64), with the p.Ap
  // partials (FP64 and FP32; fused or not as tuned)
  int OperatorS(const dfloat* o_q, dfloat* o_Aq);
  int OperatorDotS(const dfloat* o_q, dfloat* o_Aq, double* part, int* npart);
  int OperatorDot32(const float* o_q, float* o_Aq, double* part, int* npart, const CgFold<float>* fold = nullptr);
  // cg::Solve on reference-order x and r, run in the solve space
  int SolveCG(cg& solver, dfloat* o_x, dfloat* o_r, dfloat tol, int maxit, int verbose, int* iterations);
  deviceBuffer<double> o_sx, o_sr, o_sb;  // solve-space copies (created on first use)
  // times the launch choices that give identical results (gather
  // workgroup order; fused or unfused CG operator) and keeps the fastest
  int Tune();
  struct Tuning {
    int gatherMode = 0;
    double gather_us[2] = {0, 0};
    int fused64 = 0;             // FP64 CG operator: 1 = fused Ax + shared-row gather, 0 = Ax + full gather
    double fused_us = -1, unfused_us = -1;
    int fused32 = 0;             // the same choice for the FP32 operator, timed on its first use
    double fused32_us = -1, unfused32_us = -1;
  } tuning;
  int Tune32();
  // BP5, one field, N = 15, FP64: where the r and s contractions run
  // (AX_ENGINE_AUTO: the matrix cores; see AxArgs::engine)
  int axEngine = AX_ENGINE_AUTO;
  // BP1-BP4 non-affine at Nq <= 8: pencil or wavefront-per-element kernels
  int cubForm = CUB_FORM_AUTO;
  // the CG operator launched back to back and timed kernel by kernel: the
  // Ax by its own dispatch stamps, the gather from the Ax's end to an event
  // after it
  int TimeOperator(int precision, int launches, double* out);
  // the FP32 operator timed inside a mixed solve (SolveMixed's arguments):
  // out = {mean Ax us, mean gather us, applications timed, fused}
  int TimeOperatorInSolve(dfloat* o_x, const dfloat* o_b, double tol, int interval, int maxit, double* out);
  // when set, each FP32 CG-operator application takes the next three
  // events: its Ax's start and end (stamped by the dispatch) and one after
  // its gather
  struct OpTimer {
    std::vector<hipEvent_t> ev;
    size_t used = 0;
  };
  OpTimer* opTimer32 = nullptr;
  // when set, the next Ax launch (BP1-BP6) stamps its start and end into
  // these (AxArgs::evStart / evStop)
  hipEvent_t axEvStart = nullptr, axEvStop = nullptr;
  // Fused operator for CG (BP5, one field): single-copy nodes written by
  // the Ax straight into Aq, shared nodes gathered (see launch_bp5_fused).
  // fused = the maps are built; tuning.fused64/fused32 = it is used.
  bool fused = false;
  int PrepareFused();
  int PrepareFusedDevice();  // PrepareFused from the device maps (device mesh)
  template <typename T>
  int OperatorDotFused(const T* o_q, T* o_Aq, double* part, int* npart, const CgFold<T>* fold = nullptr);
  // the rows a folded CG operator leaves to k_cg_update_p (the solve space's
  // rows before the first interior run), or 0 where no fold applies:
  // precision 64 or 32, the fused operator in use, and cgFold set
  dlong FoldRows(int precision) const;
  // fold the CG direction update into the fused operator (same bits): bit 0
  // for the FP64 solves, bit 1 for the FP32 iterations of the mixed solve
  int cgFold = 1;
  int OperatorDotUnfusedS(const dfloat* o_q, dfloat* o_Aq, double* part, int* npart);
  int OperatorDotUnfused32(const float* o_q, float* o_Aq, double* part, int* npart);
  // Ax into the local vector; G2L = the GlobalToLocal map to read q
  // through (default: the reference rows)
  int AxLocal(const dfloat* o_q, dfloat* o_AqL, const dlong* G2L = nullptr);
  int Operator32(const float* o_q, float* o_Aq);
  int Forcing(dfloat* o_b);
  int NodeCoordinates();  // o_x, o_y, o_z, o_gllwRhs on first use
  // the solution (reference row order) scattered to the local nodes and
  // written as a VTU file (BP/BP5/src/bp5Run.cpp:220-238)
  int PlotSolution(const dfloat* o_u, const std::string& fileName);
  // error against the forcing's manufactured solution (one field):
  // out = {max |u_h - u|, ||u_h - u||_L2, ||u_h - u||_L2 / ||u||_L2}
  int ErrorNorms(const dfloat* o_u, double* out);
  // K right-hand sides at once (BP5), one vector after another: o_q and
  // o_Aq hold K vectors of Ngather, o_AqL K of NLocal; chunk as for
  // launch_bp5_ax_krhs (0 = default)
  int AxLocalKrhs(int K, const dfloat* o_q, dfloat* o_AqL, int chunk);
  // the k-RHS Ax (K, chunk) on o_q (K vectors of Ngather; null: zeros)
  // and the single-vector Ax on its first vector, each launched `launches`
  // times after 2 warm-ups and timed by HIP events its own dispatch
  // stamps: out = {mean k-RHS launch us, mean single launch us}
  int TimeKrhs(int K, int chunk, int launches, const dfloat* o_q, double* out);
  int OperatorKrhs(int K, const dfloat* o_q, dfloat* o_Aq, int chunk);
  // Variable coefficient (BP5): kappa on the node lattice (device, Nnodes
  // values); nullptr restores the constant coefficient. The operator then
  // scales G by kappa at each node.
  int SetKappa(const dfloat* o_kappa);
  // o_grad (Nnodes) = -sum over local copies of (D lam).G(D u)
  int KappaGradient(const dfloat* o_u, const dfloat* o_lam, dfloat* o_grad);
  hlong LatticeNodes() const;
  // FP32 CG with FP64 reliable updates every `interval` iterations until
  // ||b - A x|| / ||b|| < tol; stats = {updates, FP32 iterations, rel. residual}
  int SolveMixed(dfloat* o_x, const dfloat* o_b, double tol, int interval, int maxit, double* stats);

  // work buffers of SolveMixed, kept between solves
  struct MixedWork {
    dlong N = -1;
    int maxit = 0;
    deviceBuffer<double> Ax64, part_true, part_rr, part_pAp, hist_rr, hist_alpha, norm, xs, bs;
    deviceBuffer<float> r32, p32, Ap32, c32;
    deviceBuffer<CgState> state;
  } mixed;

  platform_t* platform = nullptr;
  mesh_t* mesh = nullptr;
  int bp = 5, Nfields = 1;
  dfloat lambda = 1.0;
  ogs::ogs_t ogs;
  ogs::halo_t gHalo;  // ghost exchange of q (empty on one GPU)
  dlong NLocal = 0;
  std::vector<dlong> GlobalToLocal;
  deviceBuffer<dlong> o_GlobalToLocal, o_elementList;
  deviceBuffer<double> o_wJ, o_ggeo, o_gllw, o_D, o_I, o_MM, o_invV;
  deviceBuffer<double> o_eo;  // even-odd tables of the BP1-BP4 pencil kernel (N <= 7, non-affine)
  // triangles and tetrahedra: the Ax as one GEMM (launch_simplex_ax_gemm):
  // the stacked reference matrices, the gathered element values, the
  // product, and the rocBLAS handle
  deviceBuffer<double> o_Sall, o_Qs, o_Zs;
  std::shared_ptr<void> blas;
  deviceBuffer<double> o_AqL;
  deviceBuffer<double> o_AqLk;  // K local vectors for OperatorKrhs, grown on demand
  // variable coefficient: lattice point of each local node, the lattice
  // sum maps, kappa-scaled factors (per node) and, for affine meshes, wJ per node
  bool haveKappa = false;
  deviceBuffer<dlong> o_lattice;
  ogs::ogs_t latticeOgs;
  deviceBuffer<double> o_ggeoK, o_wJK, o_gradL;
  // fused operator (solve space): the surface-node values of each element
  // (FP64; FP32 in o_slots32), the shared rows' maps over them (rows
  // 0..Nshared-1 of the solve space), the first interior row, the Ax
  // kernel's p.Ap partials
  deviceBuffer<double> o_slots;
  deviceBuffer<float> o_slots32;
  ogs::ogs_t sharedOgs;
  dlong fusedS0 = 0;
  deviceBuffer<double> o_partAx;
  int naxParts = 0;
  // FP32 copies (created on first use)
  int Prepare32();
  bool have32 = false;
  int tri32 = 0;  // the FP32 BP5 operator forms its geometric factors from the trilinear maps
  // FP32 BP5 geometry: 0 = from the elements' trilinear maps when every
  // node is its element's trilinear map of the vertices (checked on the
  // host), else stored factors rounded to FP32; 1 = stored factors rounded
  // to FP32 always. Changing it rebuilds the FP32 operator on next use.
  int fp32Geometry = 0;
  int SetFp32Geometry(int mode);
  deviceBuffer<float> o_wJ32, o_ggeo32, o_gllw32, o_D32, o_I32, o_MM32, o_invV32, o_AqL32;
  deviceBuffer<float> o_eo32;
  // nodal coordinates for the forcing (created on first use)
  deviceBuffer<double> o_x, o_y, o_z, o_gllwRhs;

 private:
  template <typename T>
  int ax(const T* q, T* AqL, const T* wJ, const T* ggeo, const T* gllw, const T* D, const T* I,
         const T* MM, const T* invV, const dlong* G2L = nullptr);
};

// bp_t seen through the solve space (what cg::Solve applies inside SolveCG)
class bp_solve_view_t : public solver_t {
 public:
  explicit bp_solve_view_t(bp_t& b) : bp(b) {}
  int Operator(const dfloat* o_q, dfloat* o_Aq) override { return bp.OperatorS(o_q, o_Aq); }
  bool HasOperatorDot() const override { return true; }
  int OperatorDot(const dfloat* o_q, dfloat* o_Aq, double* part, int* npart) override {
    return bp.OperatorDotS(o_q, o_Aq, part, npart);
  }
  dlong FoldRows() const override { return bp.FoldRows(64); }
  int OperatorDotFold(const CgFold<dfloat>& fold, dfloat* o_Aq, double* part, int* npart) override {
    return bp.OperatorDotFused<double>(fold.p, o_Aq, part, npart, &fold);
  }
  bp_t& bp;
};

// Ragged batch of BP5 problems (box meshes of any size and degree), one CG
// per problem, all problems advanced together from a replayed hipGraph.
class bp_batch_t {
 public:
  ~bp_batch_t();
  int Setup(platform_t& p, const std::vector<MeshOpts>& opts, dfloat lambda);
  // x = A_i^{-1} b per problem (x starts from 0); per problem: iterations,
  // final ||r||, breakdown flag (host arrays of nprob, may be null)
  int Solve(double* o_x, const double* o_b, double tol, int maxit, int* iterations, double* rnorm,
            int* breakdown);
  // FP32 CG per problem with FP64 defect corrections every `interval`
  // iterations (the recurrence of bp_t::SolveMixed, every problem at once,
  // x starting from 0) until ||b - A x|| <= tol ||b||; per problem (host
  // arrays, may be null): FP32 iterations, corrections, final
  // ||b - A x|| / ||b||, breakdown flag
  int SolveMixed(double* o_x, const double* o_b, double tol, int interval, int maxit, int* iterations,
                 int* updates, double* relres, int* breakdown);

  struct Group {  // problems of one degree: one Ax launch over their elements
    int N = 0, Nq = 0;
    dlong Nel = 0;
    long long loff = 0;  // start in the concatenated local vector
    deviceBuffer<dlong> elementList, G2L, elemPid;
    deviceBuffer<double> wJ, ggeo, D;
    deviceBuffer<float> wJ32, ggeo32, D32;  // rounded to FP32 for the mixed solve
    int variant = 0;             // workgroup shape chosen by the tuner
    double us[3] = {-1, -1, -1};  // its timings (us; -1 = shape not available)
  };
  platform_t* platform = nullptr;
  dfloat lambda = 1.0;
  int nprob = 0, nchunks = 0, nsupers = 0;
  long long replays = 0;
  std::vector<dlong> off;  // global offsets, nprob+1
  dlong Ntot = 0, NLocalTot = 0;
  std::deque<Group> groups;  // deque: Group holds device buffers, which do not move

 private:
  int Tune();
  int AxGroup(const Group& G, int variant, const double* q, double* AqL) const;
  int Iteration(double* o_x, double tol);
  int Capture(double* o_x, double tol);
  int Prepare32();
  int Segment(double* o_x, const double* o_b, double tol, int interval);
  int CaptureMixed(double* o_x, const double* o_b, double tol, int interval);
  bool have32 = false;
  hipGraphExec_t graphMixed = nullptr;
  double* graphMixedX = nullptr;
  const double* graphMixedB = nullptr;
  double graphMixedTol = -1;
  int graphMixedInterval = 0;
  deviceBuffer<float> o_r32, o_p32, o_Ap32, o_c32, o_AqL32;
  deviceBuffer<double> o_Ax64, o_part_true;
  deviceBuffer<BatchState> o_inner;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int* h_count = nullptr;
  hipGraphExec_t graphExec = nullptr;
  double* graphX = nullptr;
  double graphTol = -1;
  deviceBuffer<dlong> o_rowStarts, o_colIds;
  deviceBuffer<BatchSuper> o_supers;
  deviceBuffer<BatchChunk> o_chunks;
  deviceBuffer<int2> o_pApRange, o_rrRange;
  deviceBuffer<BatchState> o_state;
  deviceBuffer<double> o_r, o_p, o_Ap, o_AqL, o_part_rr, o_part_pAp;
  deviceBuffer<int> o_count;
};

}  // namespace bpx

// Internal declarations shared by the HIP kernel files (*.hip) and the host
// code (host/*.cpp, capi.cpp). Not installed; the public ABI is include/bpx.h.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>

#include "bpx.h"

namespace bpx {

typedef int dlong;        // device index width (reference include/types.h)
typedef long long hlong;  // global id width

// ------------------------------------------------------------ errors -----
int set_error(int code, const std::string& msg);
const char* last_error();
int hip_error(hipError_t e, const char* what);

#define BPX_HIP(call)                                       \
  do {                                                      \
    hipError_t bpx_e_ = (call);                             \
    if (bpx_e_ != hipSuccess) return ::bpx::hip_error(bpx_e_, #call); \
  } while (0)

// ------------------------------------------------- kernel parameter packs --
struct CgState;

// The CG direction update folded into the fused BP5 operator (one field):
// for the element's interior rows, the only copies of their rows, the Ax
// kernel forms x <- x + alpha_{k-1} p and p <- r + beta p itself, with the
// arithmetic and the stopping logic of k_cg_update_p, which then runs on
// the rows before the first interior row only.
template <typename T>
struct CgFold {
  int on = 0;
  int k = 0;        // CG iteration
  int pending = 0;  // the x update of iteration k-1 is still to be applied
  T* p = nullptr;   // the direction (the operator's q, updated in place)
  const T* r = nullptr;
  T* x = nullptr;
  const double* hist_rr = nullptr;     // r.r before iteration k (k_cg_update_p wrote entry k)
  const double* hist_alpha = nullptr;  // alpha of iteration k-1
  const CgState* st = nullptr;
};

// Ax kernels: positional arguments of the OKL kernels, templated on the
// value type. Fields that a BP does not use stay null.
template <typename T>
struct AxArgs {
  int N = 0;           // polynomial degree
  int Nelements = 0;   // elements in elementList
  int Nfields = 1;     // 1 (BP1/3/5) or 3 (BP2/4/6)
  int affine = 0;
  const int* elementList = nullptr;
  const int* GlobalToLocal = nullptr;
  const T* wJ = nullptr;        // wJ (BP5/6), cubwJ (BP1-4 non-affine), J per element (affine)
  const T* ggeo = nullptr;      // ggeo / cubggeo, 6 per node (or per element, affine)
  const T* gllw = nullptr;      // GLL weights (BP5 affine)
  const T* D = nullptr;         // D (BP5), cubD (BP3)
  const T* I = nullptr;         // cubInterp [cubNq x Nq] (BP1-4)
  const T* MM = nullptr;        // 1D mass (BP1 affine)
  const T* invV = nullptr;      // inverse Vandermonde (BP3 affine)
  T lambda = 0;
  const T* q = nullptr;
  T* Aq = nullptr;
  // k right-hand sides (BP5 only): right-hand side r of q at q + r*ldq,
  // of Aq at Aq + r*ldAq (one vector after another)
  int K = 1;
  long long ldq = 0, ldAq = 0;
  // k_bp5_krhs_wg (chunk 0, non-affine): lane sheets per workgroup (1, 2,
  // 4, 8, 16; 0 = default) and D's rows and columns in registers where
  // they fit (>= 0, the default) or read from LDS (-1); the same bits
  int krhs_sheets = 0;
  int krhs_dreg = 0;
  // fused BP5 operator only
  CgFold<T> fold;
  // BP5, FP64, one field, Nq = 16: the r and s contractions on the matrix
  // cores (launch_bp5_ax_mfma16) or on the vector ALU; the two agree to
  // rounding, not bit for bit. AX_ENGINE_AUTO = matrix cores where they
  // apply (the faster: 814 against 1084 us at 24^3, p = 15).
  int engine = 0;
  // FP32 BP5, non-affine: ggeo holds the coefficients of each element's
  // trilinear map (24 per element) and wJ the GLL nodes; the kernel forms
  // the geometric factors itself
  int tri = 0;
  // BP1-BP4 non-affine at Nq <= 8: the pencil kernel (CUB_FORM_AUTO or
  // CUB_FORM_PENCIL) or the wavefront-per-element kernels (CUB_FORM_WAVE);
  // the two agree to rounding, not bit for bit
  int cubForm = 0;
  // BP1-BP4 non-affine pencil kernel: the even-odd tables of I, I^T, D and
  // D^T (cub_even_odd_tables), or null for the row-by-row form
  const T* eo = nullptr;
  // when set (BP1-BP6 Ax), the kernel's dispatch stamps its start and end
  // into these events (hipExtLaunchKernelGGL: no barrier packets around it)
  hipEvent_t evStart = nullptr, evStop = nullptr;
};
enum { AX_ENGINE_AUTO = 0, AX_ENGINE_VECTOR = 1, AX_ENGINE_MATRIX = 2 };
enum { CUB_FORM_AUTO = 0, CUB_FORM_PENCIL = 1, CUB_FORM_WAVE = 2, CUB_FORM_PENCIL_ROWS = 3 };
#if defined(__HIPCC__)
}  // namespace bpx
#include <hip/hip_ext.h>
namespace bpx {
// a launch whose dispatch stamps its own start and end into e0 / e1 when
// they are set (kernel time without the gaps event packets would add)
template <typename K, typename... A>
void launch_stamped(K kernel, dim3 grid, dim3 block, hipStream_t s, hipEvent_t e0, hipEvent_t e1, A... args) {
  if (e0)
    hipExtLaunchKernelGGL(kernel, grid, block, 0, s, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}
// the same with dynamic LDS
template <typename K, typename... A>
void launch_stamped_lds(K kernel, dim3 grid, dim3 block, size_t lds, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                        A... args) {
  if (e0)
    hipExtLaunchKernelGGL(kernel, grid, block, lds, s, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kernel, grid, block, lds, s, args...);
}
#endif
// whether the BP5 Ax of these arguments runs on the matrix cores
template <typename T>
inline bool bp5_uses_matrix_cores(const AxArgs<T>& a) {
  return sizeof(T) == 8 && a.N == 15 && a.Nfields == 1 && a.engine != AX_ENGINE_VECTOR;
}

// Launchers (defined in the .hip files). Return BPX_OK or an error code.
template <typename T> int launch_bp5_ax(hipStream_t s, const AxArgs<T>& a);
// one wavefront per element, r and s contractions on the matrix cores
// (ax_bp5_mfma.hip); BP5 at N = 15, FP64, one field
int launch_bp5_ax_mfma16(hipStream_t s, const AxArgs<double>& a);
int launch_bp5_fused_mfma16(hipStream_t s, const AxArgs<double>& a, double* Ap, double* part);
// a.K right-hand sides in one launch; chunk = right-hand sides held in
// registers per sweep (1, 2 or 4, dividing K; 0 picks the default)
template <typename T> int launch_bp5_ax_krhs(hipStream_t s, const AxArgs<T>& a, int chunk);
// Output layout of the fused operator (BP5, one field, Nq nodes per
// direction). Element e's interior nodes are Ap rows a.ldq + e *
// fused_interior_pad(Nq) + fused_interior_index(...); its surface nodes go
// to a.Aq + e * fused_surface_pad(Nq) + fused_surface_index(...), both in
// local node order (k, j, i); runs are padded to 32 entries (whole cache
// lines) and the padding is written with zeros. p.Ap partials of the
// interior rows go to part[bp5_workgroups(N, Nelements)].
constexpr int fused_pad(int n) { return (n + 31) / 32 * 32; }
constexpr int fused_interior(int Nq) { return Nq > 2 ? (Nq - 2) * (Nq - 2) * (Nq - 2) : 0; }
constexpr int fused_interior_pad(int Nq) { return fused_pad(fused_interior(Nq)); }
constexpr int fused_surface_pad(int Nq) { return fused_pad(Nq * Nq * Nq - fused_interior(Nq)); }
constexpr int fused_interior_index(int Nq, int i, int j, int k) {
  return ((k - 1) * (Nq - 2) + (j - 1)) * (Nq - 2) + (i - 1);
}
// -1 for an interior node; surface nodes: the whole bottom slab, then the
// ring of each interior slab, then the whole top slab, each in (j, i) order
constexpr int fused_surface_index(int Nq, int i, int j, int k) {
  const int Nq2 = Nq * Nq, ring = Nq2 - (Nq - 2) * (Nq - 2);
  if (k == 0) return j * Nq + i;
  if (k == Nq - 1) return Nq2 + (Nq - 2) * ring + j * Nq + i;
  if (i != 0 && i != Nq - 1 && j != 0 && j != Nq - 1) return -1;
  const int rp = (j == 0) ? i : (j == Nq - 1) ? Nq + 2 * (Nq - 2) + i : Nq + 2 * (j - 1) + (i == 0 ? 0 : 1);
  return Nq2 + (k - 1) * ring + rp;
}
// inverses: the element-local node (k * Nq + j) * Nq + i of interior
// index n, and of surface index n
constexpr int fused_interior_node(int Nq, int n) {
  const int m = Nq - 2;
  return ((n / (m * m) + 1) * Nq + (n / m) % m + 1) * Nq + n % m + 1;
}
constexpr int fused_surface_node(int Nq, int n) {
  const int Nq2 = Nq * Nq, ring = Nq2 - (Nq - 2) * (Nq - 2), top = Nq2 + (Nq - 2) * ring;
  if (n < Nq2) return n;
  if (n >= top) return (Nq - 1) * Nq2 + (n - top);
  const int k = 1 + (n - Nq2) / ring, rp = (n - Nq2) % ring;
  int i = 0, j = 0;
  if (rp < Nq) {
    j = 0;
    i = rp;
  } else if (rp >= Nq + 2 * (Nq - 2)) {
    j = Nq - 1;
    i = rp - Nq - 2 * (Nq - 2);
  } else {
    j = 1 + (rp - Nq) / 2;
    i = ((rp - Nq) % 2) ? Nq - 1 : 0;
  }
  return (k * Nq + j) * Nq + i;
}
template <typename T> int launch_bp5_fused(hipStream_t s, const AxArgs<T>& a, T* Ap, double* part);
int bp5_workgroups(int N, int Nelements);
template <typename T> int launch_bp1_ax(hipStream_t s, const AxArgs<T>& a);
template <typename T> int launch_bp3_ax(hipStream_t s, const AxArgs<T>& a);

// variable-coefficient BP5: kappa-scaled geometry (affine meshes expanded
// to per-node factors) and the adjoint gradient per local node
int launch_kappa_geometry(hipStream_t s, long long NLocal, int Nq, int affine, const int* lattice,
                          const double* kappa, const double* ggeo, const double* wJ, const double* gllw,
                          double* ggeoK, double* wJK);
int launch_bp5_kappa_grad(hipStream_t s, int N, int Nelements, int affine, const int* G2L, const double* ggeo,
                          const double* gllw, const double* D, const double* u, const double* lam, double* gL);

// ------------------------------------------------ ragged batch CG ------
// A chunk of one problem's slice of the concatenated vectors; first marks
// the chunk that writes the problem's state.
struct BatchChunk {
  int pid, n0, n1, first;
};
// Rows [r0, r1) of the concatenated gather maps, all of problem pid.
struct BatchSuper {
  int pid, r0, r1, pad;
};
// Device-resident CG state of one problem.
struct BatchState {
  double rr_k;     // r.r of the current iteration
  double rr_prev;  // r.r of the previous iteration
  double tol2;
  double rnorm;    // ||r|| when the problem stopped
  int iter, done, breakdown, maxit;
  int updates, pad;  // mixed solve: FP64 corrections made
};
// BP5 Ax over the elements of one degree group; variant = workgroup shape
// (bp5_batch_epb: elements per workgroup, 0 if the variant does not exist)
int launch_bp5_batch(hipStream_t s, int N, int variant, int Nelements, const int* elementList, const int* G2L,
                     const double* wJ, const double* ggeo, const double* D, double lambda, const double* q,
                     double* Aq, const int* elemPid, const BatchState* done);
int bp5_batch_epb(int Nq, int variant);
int launch_batch_init(hipStream_t s, int nchunks, const BatchChunk* chunks, const double* b, double* x, double* r,
                      double* part_rr, BatchState* st, int maxit);
int launch_batch_update_p(hipStream_t s, int nchunks, const BatchChunk* chunks, const int2* rrRange, const double* r,
                          double* p, const double* part_rr, BatchState* st, double tol);
int launch_batch_update_xr(hipStream_t s, int nchunks, const BatchChunk* chunks, const int2* pApRange,
                           const double* p, const double* Ap, double* x, double* r, const double* part_pAp,
                           double* part_rr, BatchState* st);
int launch_batch_gather(hipStream_t s, int nsupers, const BatchSuper* supers, const int* rowStarts,
                        const int* colIds, const double* q, double* gq, const double* p, double* part,
                        const BatchState* st);
int launch_batch_count_done(hipStream_t s, int nprob, const BatchState* st, int* count);
// quads, triangles and tetrahedra (ax_other.hip, FP64): bp 1..6 on the
// [element][field][node] local vectors. Quads: W = wJ (BP5/6, per node) or
// cubwJ (BP1-4, per Gauss point), G = ggeo or cubggeo (3 per point), Dm = D
// (Nq x Nq) or cubD ((Nq+1)^2), I = the Gauss interpolation. Simplices: wJ
// and ggeo per element, S the packed stiffness, MM the mass matrix.
int launch_quad_ax(hipStream_t s, int bp, int N, int Nfields, int Nelements, const int* elementList, const int* G2L,
                   const double* W, const double* G, const double* Dm, const double* I, double lambda,
                   const double* q, double* Aq);
// the simplex Ax as gather -> one rocBLAS DGEMM (Z = Sall Q) -> combine;
// Sall: (NG+1)*Np x Np (MM alone, Np x Np, for BP1/BP2), column-major;
// Q: Np x Nelements*Nfields; Z: rows(Sall) x Nelements*Nfields
int simplex_blas_create(hipStream_t s, void** handle);
void simplex_blas_destroy(void* handle);
int launch_simplex_ax_gemm(hipStream_t s, void* blas, int bp, int dim, int Np, int Nfields, int Nelements,
                           const int* elementList, const int* G2L, const double* wJ, const double* ggeo,
                           const double* Sall, double* Q, double* Z, double lambda, const double* q, double* Aq);
int launch_rhs_other(hipStream_t s, int elementType, int N, int Np, int Nfields, int Nelements, const double* wJ,
                     const double* MM, const double* x, const double* y, const double* z, double lambda,
                     double* rhs);
// FP32 Ax of the mixed batch (factors rounded to FP32), the one workgroup shape
int launch_bp5_batch_f32(hipStream_t s, int N, int Nelements, const int* elementList, const int* G2L,
                         const float* wJ, const float* ggeo, const float* D, float lambda, const float* q,
                         float* Aq, const int* elemPid, const BatchState* done);
// Mixed batch: FP32 CG per problem (inner state) with FP64 corrections
// (outer state); see batch.hip
int launch_to_f32(hipStream_t s, size_t n, const double* a, float* out);
int launch_bm_init(hipStream_t s, int nchunks, const BatchChunk* chunks, const double* b, double* x, float* r,
                   float* c, double* part_true, double* part_rr, BatchState* inner, BatchState* outer, int maxit);
int launch_bm_decide(hipStream_t s, int nprob, const int2* rrRange, const double* part_true, BatchState* inner,
                     BatchState* outer, double tol, int init);
int launch_bm_update_p(hipStream_t s, int nchunks, const BatchChunk* chunks, const int2* rrRange, const float* r,
                       float* p, const double* part_rr, BatchState* inner, const BatchState* outer);
int launch_bm_update_xr(hipStream_t s, int nchunks, const BatchChunk* chunks, const int2* pApRange, const float* p,
                        const float* Ap, float* c, float* r, const double* part_pAp, double* part_rr,
                        BatchState* inner, BatchState* outer);
int launch_bm_flush(hipStream_t s, int nchunks, const BatchChunk* chunks, float* c, double* x,
                    const BatchState* outer);
int launch_bm_true_residual(hipStream_t s, int nchunks, const BatchChunk* chunks, const double* b, const double* Ax,
                            float* r, double* part_true, double* part_rr, const BatchState* outer);
int launch_batch_gather_f32(hipStream_t s, int nsupers, const BatchSuper* supers, const int* rowStarts,
                            const int* colIds, const float* q, float* gq, const float* p, double* part,
                            const BatchState* st);
// the gather alone (no p.Ap partials)
int launch_batch_gather_only(hipStream_t s, int nsupers, const BatchSuper* supers, const int* rowStarts,
                             const int* colIds, const double* q, double* gq, const BatchState* st);

// per-workgroup partials (max |diff|, sum wJ diff^2, sum wJ (c u)^2) of a
// one-field solution u (read through G2L) against c sin(pi x) sin(pi y)
// sin(pi z) at the local nodes; part holds 3 * kRedBlocks doubles
int launch_error_norms(hipStream_t s, long long NLocal, const int* G2L, const double* u, const double* x,
                       const double* y, const double* z, const double* wJ, double c, double* part);
int launch_rhs(hipStream_t s, int N, int Nfields, int affine, int Nelements, const double* wJ,
               const double* gllw, const double* x, const double* y, const double* z,
               double lambda, double* rhs);

// ogs (libs/ogs/okl/ogsKernels.okl semantics, K interleaved vectors).
// mode: order in which the gather's workgroups take super-blocks (0 strided,
// 1 XCD-blocked); it changes speed only, never a result bit.
template <typename T>
int launch_ogs_gather(hipStream_t s, int Nblocks, int K, const int* blockStarts,
                      const int* rowStarts, const int* colIds, const T* q, T* gq, int mode = 0);
template <typename T>
int launch_ogs_scatter(hipStream_t s, int Nblocks, int K, const int* blockStarts,
                       const int* rowStarts, const int* colIds, const T* gq, T* q);
template <typename T>
int launch_ogs_gather_scatter(hipStream_t s, int Nblocks, int K, const int* blockStarts,
                              const int* rowStarts, const int* colIds, T* q);

// gather with a fused p.Ap: one FP64 partial per super-block of row
// blocks, in part[0..n) with n = ogs_gather_partials(Nblocks)
int ogs_gather_partials(int Nblocks);
// the same, where super-block sb also adds axpart[sb*nax/nsb ..
// (sb+1)*nax/nsb) to its partial (nsb = ogs_gather_partials(Nblocks))
template <typename T>
int launch_ogs_gather_dot_fold(hipStream_t s, int Nblocks, const int* blockStarts, const int* rowStarts,
                               const int* colIds, const T* q, T* gq, const T* p, double* part, const double* axpart,
                               int nax, int mode = 0);
template <typename T>
int launch_ogs_gather_dot(hipStream_t s, int Nblocks, const int* blockStarts, const int* rowStarts,
                          const int* colIds, const T* q, T* gq, const T* p, double* part, int mode = 0);

// linAlg. Reductions run on a fixed grid of kRedBlocks workgroups and write
// kRedBlocks FP64 partials; launch_fold sums partials into a device scalar.
constexpr int kRedThreads = 1024;
constexpr int kRedBlocks = 256;
template <typename T> int launch_set(hipStream_t s, dlong N, T alpha, T* x);
template <typename T> int launch_axpy(hipStream_t s, dlong N, T alpha, const T* x, T beta, T* y);
// y = alpha*x + beta*y in FP64 with x given in FP32 (defect-correction update)
int launch_axpy_f32_to_f64(hipStream_t s, dlong N, double alpha, const float* x, double beta,
                           double* y);
int launch_copy_f64_to_f32(hipStream_t s, dlong N, double scale, const double* x, float* y);
template <typename T>
int launch_dot_partials(hipStream_t s, dlong N, const T* x, const T* y, double* part);
int launch_fold(hipStream_t s, const double* part, int n, double* out);
// x.y into *out (device); part holds kRedBlocks doubles
template <typename T>
int launch_dot(hipStream_t s, dlong N, const T* x, const T* y, double* part, double* out);
// cg::UpdateCG (r -= alpha Ap, x += alpha p) with r.r into *out (device)
int launch_update_cg(hipStream_t s, dlong N, double alpha, const double* p, const double* Ap, double* x,
                     double* r, double* part, double* out);

// Device-resident CG state (one per solver).
struct CgState {
  double tol2;    // stop when r.r <= tol2
  int iter;       // iterations completed (valid once done, or after finish)
  int done;       // 1 once r.r <= tol2, the iteration cap is reached, or breakdown
  int breakdown;  // 1 when p.Ap or r.r stopped being finite/positive
  int maxit;
  // k + 1 when the stopping test of k_cg_update_p fired at iteration k, else
  // 0. That kernel records its stop here, never in `done`: its other
  // workgroups test `done` and `stop_at <= k` on entry, and a stop written
  // by its own first workgroup must not make them skip the x update of
  // iteration k - 1, which every row still needs. The kernels after it
  // treat stop_at != 0 as done, and k_cg_finish folds it into `done`.
  int stop_at;
};

// Work buffers of the allocation-free CG solve, carved from the file-scope
// device array of cg_noalloc.hip (sizes: kNoAllocMaxN for p and Ap,
// kNoAllocMaxIter + 1 for each history, kRedBlocks for each partials array).
constexpr long long kNoAllocMaxN = BPX_CG_NOALLOC_MAX_N;
constexpr int kNoAllocMaxIter = BPX_CG_NOALLOC_MAX_ITER;
struct NoAllocWork {
  double* p = nullptr;
  double* Ap = nullptr;
  double* hist_rr = nullptr;
  double* hist_alpha = nullptr;
  double* part_rr = nullptr;
  double* part_pAp = nullptr;
  CgState* state = nullptr;
  int state_doubles = 0;   // the state region in doubles (cleared per solve)
  std::mutex* lock = nullptr;  // serialises this process's solves on the device
};
// the device of stream s (the current device for the null stream) and that
// device's work buffers
int noalloc_work(hipStream_t s, int* device, NoAllocWork* w);
int noalloc_scratch_bytes(size_t* bytes);

// CG steps, templated on the vector type (double: plain CG; float: the
// inner solve of the mixed-precision path). Iteration k is passed by the
// host. hist_rr[k] = r.r before iteration k (length maxit+1), hist_alpha[k]
// = alpha of iteration k. part_rr holds kRedBlocks partials of r.r.
template <typename T>
int launch_cg_init(hipStream_t s, dlong N, const T* Ap, T* r, double* part_rr);
template <typename T>
int launch_cg_update_p(hipStream_t s, dlong N, int k, const T* r, T* p, T* x, const double* part_rr, int nrr,
                       CgState* st, double* hist_rr, const double* hist_alpha, double tol, bool pending);
template <typename T>
int launch_cg_update_r(hipStream_t s, dlong N, int k, const T* Ap, T* r, const double* part_pAp, int npAp,
                       double* part_rr, CgState* st, const double* hist_rr, double* hist_alpha);
// x <- x + alpha_k p (the deferred update of the last iteration k);
// when_running: skipped if the recurrence has stopped
template <typename T>
int launch_cg_flush_x(hipStream_t s, dlong N, int k, const T* p, T* x, const double* hist_alpha,
                      const CgState* st, bool when_running);
int launch_cg_finish(hipStream_t s, int kend, int maxit, const double* part_rr, int nrr, CgState* st,
                     double* hist_rr);

// Row permutation over N rows, perm[i] = reference row of row i (-1 =
// padding): to_ref = false: out[i] = in[perm[i]] (0 on padding);
// to_ref = true: out[perm[i]] = in[i].
int launch_permute(hipStream_t s, dlong N, const dlong* perm, const double* in, double* out, bool to_ref);

// Mixed precision reliable update: x64 += c32, c32 = 0; then from Ax (FP64)
// r32 = float(b - Ax) with kRedBlocks partials of (b-Ax).(b-Ax) in
// part_true and of r32.r32 in part_rr32.
int launch_accumulate_f32(hipStream_t s, dlong N, float* c32, double* x64);
int launch_true_residual(hipStream_t s, dlong N, const double* b, const double* Ax, float* r32, double* part_true,
                         double* part_rr32);

}  // namespace bpx

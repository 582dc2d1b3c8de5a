// Internal declarations shared by the HIP kernel files (*.hip) and the host
// code (host/*.cpp, capi.cpp). Not installed; the public ABI is include/bpx.h.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
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
};

// Launchers (defined in the .hip files). Return BPX_OK or an error code.
template <typename T> int launch_bp5_ax(hipStream_t s, const AxArgs<T>& a);
template <typename T> int launch_bp1_ax(hipStream_t s, const AxArgs<T>& a);
template <typename T> int launch_bp3_ax(hipStream_t s, const AxArgs<T>& a);

int launch_rhs(hipStream_t s, int N, int Nfields, int affine, int Nelements, const double* wJ,
               const double* gllw, const double* x, const double* y, const double* z,
               double lambda, double* rhs);

// ogs (libs/ogs/okl/ogsKernels.okl semantics, K interleaved vectors)
template <typename T>
int launch_ogs_gather(hipStream_t s, int Nblocks, int K, const int* blockStarts,
                      const int* rowStarts, const int* colIds, const T* q, T* gq);
template <typename T>
int launch_ogs_scatter(hipStream_t s, int Nblocks, int K, const int* blockStarts,
                       const int* rowStarts, const int* colIds, const T* gq, T* q);
template <typename T>
int launch_ogs_gather_scatter(hipStream_t s, int Nblocks, int K, const int* blockStarts,
                              const int* rowStarts, const int* colIds, T* q);

// linAlg. Reductions write their result to a device scalar *out (after a
// two-level deterministic tree); `scratch` holds kReduceBlocks partials
// plus a ticket counter (see reduce_scratch_bytes).
constexpr int kReduceThreads = 256;
constexpr int kReduceBlocks = 1024;
size_t reduce_scratch_bytes();
template <typename T> int launch_set(hipStream_t s, dlong N, T alpha, T* x);
template <typename T> int launch_axpy(hipStream_t s, dlong N, T alpha, const T* x, T beta, T* y);
// y = alpha*x + beta*y in FP64 with x given in FP32 (defect-correction update)
int launch_axpy_f32_to_f64(hipStream_t s, dlong N, double alpha, const float* x, double beta,
                           double* y);
int launch_copy_f64_to_f32(hipStream_t s, dlong N, double scale, const double* x, float* y);
template <typename T>
int launch_dot(hipStream_t s, dlong N, const T* x, const T* y, void* scratch, double* out);

// Device-resident CG state (one per solver).
struct CgState {
  double rdotr;       // current r.r
  double rdotr_prev;  // previous r.r (for beta)
  double alpha;
  double pAp;
  double tol2;        // stop when rdotr <= tol2
  int iter;           // iterations completed
  int done;           // 1 once rdotr <= tol2 (or breakdown)
  int breakdown;      // 1 when p.Ap or r.r stopped being finite/positive
  int maxit;
};

// CG kernels, templated on the vector type (double: plain CG; float: the
// inner solve of the mixed-precision path). hist_rr/hist_alpha: device
// arrays of length maxit+1 / maxit (may be null).
template <typename T>
int launch_cg_init(hipStream_t s, dlong N, const T* Ap, T* r, void* scratch, CgState* st,
                   double tol, int maxit, double* hist_rr);
template <typename T>
int launch_cg_update_p(hipStream_t s, dlong N, const T* r, T* p, const CgState* st);
template <typename T>
int launch_cg_pAp(hipStream_t s, dlong N, const T* p, const T* Ap, void* scratch, CgState* st);
template <typename T>
int launch_cg_update_xr(hipStream_t s, dlong N, const T* p, const T* Ap, T* x, T* r,
                        void* scratch, CgState* st, double* hist_rr, double* hist_alpha);

// Geometric factors on the device from nodal coordinates
// (libs/mesh/meshGeometricFactorsHex3D.cpp:84-139 semantics).
}  // namespace bpx

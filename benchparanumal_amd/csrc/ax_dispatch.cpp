// Run-time degree -> compiled Nq instantiation for the Ax launchers.
#include <type_traits>

#include "ax_common.hpp"

namespace bpx {
namespace {

// FP32 kernels exist for one field (the mixed-precision inner solves);
// FP64 kernels for one and three fields.
#define BPX_AX_CASE(FN, NQ)                                                   \
  case NQ:                                                                    \
    if (a.Nfields == 1) return FN<T, NQ, 1>(s, a);                            \
    if constexpr (std::is_same<T, double>::value) {                           \
      if (a.Nfields == 3) return FN<T, NQ, 3>(s, a);                          \
    }                                                                         \
    break;

#define BPX_AX_SWITCH(FN)                                                     \
  switch (a.N + 1) {                                                          \
    BPX_AX_CASE(FN, 2) BPX_AX_CASE(FN, 3) BPX_AX_CASE(FN, 4)                  \
    BPX_AX_CASE(FN, 5) BPX_AX_CASE(FN, 6) BPX_AX_CASE(FN, 7)                  \
    BPX_AX_CASE(FN, 8) BPX_AX_CASE(FN, 9) BPX_AX_CASE(FN, 10)                 \
    BPX_AX_CASE(FN, 11) BPX_AX_CASE(FN, 12) BPX_AX_CASE(FN, 13)               \
    BPX_AX_CASE(FN, 14) BPX_AX_CASE(FN, 15) BPX_AX_CASE(FN, 16)               \
    default:                                                                  \
      return set_error(BPX_ERR_UNSUPPORTED, "Ax: degree N must be in 1..15"); \
  }                                                                           \
  return set_error(BPX_ERR_UNSUPPORTED, "Ax: unsupported field count / precision");

int check(const char* who, int Nelements, const void* q, const void* Aq, const void* G2L,
          const void* list) {
  if (Nelements < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, std::string(who) + ": Nelements < 0");
  if (Nelements > 0 && (!q || !Aq || !G2L || !list))
    return set_error(BPX_ERR_INVALID_ARGUMENT, std::string(who) + ": null device pointer");
  return BPX_OK;
}

}  // namespace

template <typename T>
int launch_bp5_ax(hipStream_t s, const AxArgs<T>& a) {
  if (int st = check("bp5 Ax", a.Nelements, a.q, a.Aq, a.GlobalToLocal, a.elementList)) return st;
  if (a.Nelements == 0) return BPX_OK;
  if (!a.wJ || !a.ggeo || !a.D || (a.affine && !a.gllw))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 Ax: null geometry/operator pointer");
  if (a.engine == AX_ENGINE_MATRIX && !bp5_uses_matrix_cores(a))
    return set_error(BPX_ERR_UNSUPPORTED, "bp5 Ax: the matrix-core kernel covers FP64, one field, N = 15");
  if constexpr (std::is_same<T, double>::value) {
    if (bp5_uses_matrix_cores(a)) return launch_bp5_ax_mfma16(s, a);
  }
  BPX_AX_SWITCH(launch_bp5_nq)
}

template <typename T>
int launch_bp1_ax(hipStream_t s, const AxArgs<T>& a) {
  if (int st = check("bp1 Ax", a.Nelements, a.q, a.Aq, a.GlobalToLocal, a.elementList)) return st;
  if (a.Nelements == 0) return BPX_OK;
  if (!a.wJ || (a.affine ? !a.MM : !a.I))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "bp1 Ax: null geometry/operator pointer");
  BPX_AX_SWITCH(launch_bp1_nq)
}

template <typename T>
int launch_bp3_ax(hipStream_t s, const AxArgs<T>& a) {
  if (int st = check("bp3 Ax", a.Nelements, a.q, a.Aq, a.GlobalToLocal, a.elementList)) return st;
  if (a.Nelements == 0) return BPX_OK;
  if (!a.wJ || !a.ggeo || !a.D || (a.affine ? !a.invV : !a.I))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "bp3 Ax: null geometry/operator pointer");
  BPX_AX_SWITCH(launch_bp3_nq)
}

#define BPX_KRHS_CASE(NQ) \
  case NQ:                 \
    return launch_bp5_krhs_nq<T, NQ>(s, a, chunk);

template <typename T>
int launch_bp5_ax_krhs(hipStream_t s, const AxArgs<T>& a, int chunk) {
  if (int st = check("bp5 k-RHS Ax", a.Nelements, a.q, a.Aq, a.GlobalToLocal, a.elementList)) return st;
  if (a.K < 1) return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 k-RHS Ax: K < 1");
  if (a.K > 1 && (a.ldq < 1 || a.ldAq < 1))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 k-RHS Ax: leading dimensions must be positive");
  if (a.Nfields != 1) return set_error(BPX_ERR_UNSUPPORTED, "bp5 k-RHS Ax: one field per right-hand side");
  if (chunk > 0 && ((chunk != 1 && chunk != 2 && chunk != 4) || a.K % chunk != 0))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 k-RHS Ax: chunk must be 0, -1, -2, or 1, 2 or 4 dividing K");
  if (chunk < -2) return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 k-RHS Ax: chunk must be 0, -1, -2, 1, 2 or 4");
  if (a.Nelements == 0) return BPX_OK;
  if (!a.wJ || !a.ggeo || !a.D || (a.affine && !a.gllw))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 k-RHS Ax: null geometry/operator pointer");
  switch (a.N + 1) {
    BPX_KRHS_CASE(2) BPX_KRHS_CASE(3) BPX_KRHS_CASE(4) BPX_KRHS_CASE(5) BPX_KRHS_CASE(6)
    BPX_KRHS_CASE(7) BPX_KRHS_CASE(8) BPX_KRHS_CASE(9) BPX_KRHS_CASE(10) BPX_KRHS_CASE(11)
    BPX_KRHS_CASE(12) BPX_KRHS_CASE(13) BPX_KRHS_CASE(14) BPX_KRHS_CASE(15) BPX_KRHS_CASE(16)
    default:
      return set_error(BPX_ERR_UNSUPPORTED, "bp5 k-RHS Ax: degree N must be in 1..15");
  }
}

template int launch_bp5_ax_krhs<double>(hipStream_t, const AxArgs<double>&, int);
template int launch_bp5_ax_krhs<float>(hipStream_t, const AxArgs<float>&, int);

#define BPX_FUSED_CASE(NQ) \
  case NQ:                  \
    return launch_bp5_fused_nq<T, NQ>(s, a, Ap, part);

template <typename T>
int launch_bp5_fused(hipStream_t s, const AxArgs<T>& a, T* Ap, double* part) {
  if (int st = check("bp5 fused Ax", a.Nelements, a.q, a.Aq, a.GlobalToLocal, a.elementList)) return st;
  if (a.Nelements == 0) return BPX_OK;
  if (a.Nfields != 1 || !Ap || !part) return set_error(BPX_ERR_INVALID_ARGUMENT, "bp5 fused Ax: bad arguments");
  if (a.engine == AX_ENGINE_MATRIX && !bp5_uses_matrix_cores(a))
    return set_error(BPX_ERR_UNSUPPORTED, "bp5 fused Ax: the matrix-core kernel covers FP64, N = 15");
  if constexpr (std::is_same<T, double>::value) {
    if (bp5_uses_matrix_cores(a)) {
      if (a.fold.on)
        return set_error(BPX_ERR_UNSUPPORTED, "bp5 fused Ax: the matrix-core kernel has no folded CG update");
      return launch_bp5_fused_mfma16(s, a, Ap, part);
    }
  }
  switch (a.N + 1) {
    BPX_FUSED_CASE(2) BPX_FUSED_CASE(3) BPX_FUSED_CASE(4) BPX_FUSED_CASE(5) BPX_FUSED_CASE(6)
    BPX_FUSED_CASE(7) BPX_FUSED_CASE(8) BPX_FUSED_CASE(9) BPX_FUSED_CASE(10) BPX_FUSED_CASE(11)
    BPX_FUSED_CASE(12) BPX_FUSED_CASE(13) BPX_FUSED_CASE(14) BPX_FUSED_CASE(15) BPX_FUSED_CASE(16)
    default:
      return set_error(BPX_ERR_UNSUPPORTED, "bp5 fused Ax: degree N must be in 1..15");
  }
}

template int launch_bp5_fused<double>(hipStream_t, const AxArgs<double>&, double*, double*);
template int launch_bp5_fused<float>(hipStream_t, const AxArgs<float>&, float*, double*);

int bp5_workgroups(int N, int Nelements) {
  const int Nq = N + 1;
  const int epb = Nq >= 8 ? 1 : 256 / (Nq * Nq);
  return (Nelements + epb - 1) / epb;
}

#define BPX_BATCH_CASE(NQ) \
  case NQ:                  \
    return launch_bp5_batch_nq<NQ>(s, variant, Nelements, elementList, G2L, wJ, ggeo, D, lambda, q, Aq, elemPid, done);

int launch_bp5_batch(hipStream_t s, int N, int variant, int Nelements, const int* elementList, const int* G2L,
                     const double* wJ, const double* ggeo, const double* D, double lambda, const double* q,
                     double* Aq, const int* elemPid, const BatchState* done) {
  if (Nelements <= 0) return BPX_OK;
  switch (N + 1) {
    BPX_BATCH_CASE(2) BPX_BATCH_CASE(3) BPX_BATCH_CASE(4) BPX_BATCH_CASE(5) BPX_BATCH_CASE(6)
    BPX_BATCH_CASE(7) BPX_BATCH_CASE(8) BPX_BATCH_CASE(9) BPX_BATCH_CASE(10) BPX_BATCH_CASE(11)
    BPX_BATCH_CASE(12) BPX_BATCH_CASE(13) BPX_BATCH_CASE(14) BPX_BATCH_CASE(15) BPX_BATCH_CASE(16)
    default:
      return set_error(BPX_ERR_UNSUPPORTED, "batch Ax: degree N must be in 1..15");
  }
}

#define BPX_BATCH32_CASE(NQ) \
  case NQ:                    \
    return launch_bp5_batch_f32_nq<NQ>(s, Nelements, elementList, G2L, wJ, ggeo, D, lambda, q, Aq, elemPid, done);

int launch_bp5_batch_f32(hipStream_t s, int N, int Nelements, const int* elementList, const int* G2L,
                         const float* wJ, const float* ggeo, const float* D, float lambda, const float* q,
                         float* Aq, const int* elemPid, const BatchState* done) {
  if (Nelements <= 0) return BPX_OK;
  switch (N + 1) {
    BPX_BATCH32_CASE(2) BPX_BATCH32_CASE(3) BPX_BATCH32_CASE(4) BPX_BATCH32_CASE(5) BPX_BATCH32_CASE(6)
    BPX_BATCH32_CASE(7) BPX_BATCH32_CASE(8) BPX_BATCH32_CASE(9) BPX_BATCH32_CASE(10) BPX_BATCH32_CASE(11)
    BPX_BATCH32_CASE(12) BPX_BATCH32_CASE(13) BPX_BATCH32_CASE(14) BPX_BATCH32_CASE(15) BPX_BATCH32_CASE(16)
    default:
      return set_error(BPX_ERR_UNSUPPORTED, "batch Ax: degree N must be in 1..15");
  }
}

int bp5_batch_epb(int Nq, int variant) {
  const int epb = Nq >= 8 ? 1 : 256 / (Nq * Nq);
  if (variant == 0) return epb;
  if (variant == 1) return Nq * Nq * 2 * epb <= 1024 ? 2 * epb : 0;
  if (variant == 2) return epb >= 2 ? epb / 2 : 0;
  return 0;
}

#define BPX_KGRAD_CASE(NQ) \
  case NQ:                  \
    return launch_bp5_kappa_grad_nq<NQ>(s, Nelements, affine, G2L, ggeo, gllw, D, u, lam, gL);

int launch_bp5_kappa_grad(hipStream_t s, int N, int Nelements, int affine, const int* G2L, const double* ggeo,
                          const double* gllw, const double* D, const double* u, const double* lam, double* gL) {
  if (Nelements < 0) return set_error(BPX_ERR_INVALID_ARGUMENT, "kappa gradient: Nelements < 0");
  if (Nelements == 0) return BPX_OK;
  if (!G2L || !ggeo || !D || !u || !lam || !gL || (affine && !gllw))
    return set_error(BPX_ERR_INVALID_ARGUMENT, "kappa gradient: null device pointer");
  switch (N + 1) {
    BPX_KGRAD_CASE(2) BPX_KGRAD_CASE(3) BPX_KGRAD_CASE(4) BPX_KGRAD_CASE(5) BPX_KGRAD_CASE(6)
    BPX_KGRAD_CASE(7) BPX_KGRAD_CASE(8) BPX_KGRAD_CASE(9) BPX_KGRAD_CASE(10) BPX_KGRAD_CASE(11)
    BPX_KGRAD_CASE(12) BPX_KGRAD_CASE(13) BPX_KGRAD_CASE(14) BPX_KGRAD_CASE(15) BPX_KGRAD_CASE(16)
    default:
      return set_error(BPX_ERR_UNSUPPORTED, "kappa gradient: degree N must be in 1..15");
  }
}
template int launch_bp5_ax<double>(hipStream_t, const AxArgs<double>&);
template int launch_bp5_ax<float>(hipStream_t, const AxArgs<float>&);
template int launch_bp1_ax<double>(hipStream_t, const AxArgs<double>&);
template int launch_bp1_ax<float>(hipStream_t, const AxArgs<float>&);
template int launch_bp3_ax<double>(hipStream_t, const AxArgs<double>&);
template int launch_bp3_ax<float>(hipStream_t, const AxArgs<float>&);

}  // namespace bpx

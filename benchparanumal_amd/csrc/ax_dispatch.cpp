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

template int launch_bp5_ax<double>(hipStream_t, const AxArgs<double>&);
template int launch_bp5_ax<float>(hipStream_t, const AxArgs<float>&);
template int launch_bp1_ax<double>(hipStream_t, const AxArgs<double>&);
template int launch_bp1_ax<float>(hipStream_t, const AxArgs<float>&);
template int launch_bp3_ax<double>(hipStream_t, const AxArgs<double>&);
template int launch_bp3_ax<float>(hipStream_t, const AxArgs<float>&);

}  // namespace bpx

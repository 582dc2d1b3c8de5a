// Per-degree Ax launchers. Each ax_bp*.hip translation unit is compiled
// once per Nq (-DBPX_NQ) and instantiates its launcher for that Nq;
// ax_dispatch.cpp maps the run-time degree onto them.
#pragma once
#include "bpx_internal.hpp"

namespace bpx {

template <typename T, int NQ, int NF> int launch_bp5_nq(hipStream_t s, const AxArgs<T>& a);
template <typename T, int NQ>
int launch_bp5_fused_nq(hipStream_t s, const AxArgs<T>& a, T* Ap, double* part);
template <typename T, int NQ> int launch_bp5_krhs_nq(hipStream_t s, const AxArgs<T>& a, int chunk);
template <int NQ>
int launch_bp5_batch_nq(hipStream_t s, int variant, int Nelements, const int* elementList, const int* G2L,
                        const double* wJ, const double* ggeo, const double* D, double lambda, const double* q,
                        double* Aq, const int* elemPid, const BatchState* done);
template <int NQ>
int launch_bp5_batch_f32_nq(hipStream_t s, int Nelements, const int* elementList, const int* G2L, const float* wJ,
                            const float* ggeo, const float* D, float lambda, const float* q, float* Aq,
                            const int* elemPid, const BatchState* done);
template <int NQ>
int launch_bp5_kappa_grad_nq(hipStream_t s, int Nelements, int affine, const int* G2L, const double* ggeo,
                             const double* gllw, const double* D, const double* u, const double* lam, double* gL);
template <typename T, int NQ, int NF> int launch_bp1_nq(hipStream_t s, const AxArgs<T>& a);
template <typename T, int NQ, int NF> int launch_bp3_nq(hipStream_t s, const AxArgs<T>& a);

}  // namespace bpx

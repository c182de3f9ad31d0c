/*
 * accfft_operators_gpu.h -- drop-in replacement of the reference's
 * include/accfft_operators_gpu.h (:31-67) for one MI355X on a 1x1 grid, plus
 * the GPU inverse Laplacian / inverse biharmonic that the reference only has
 * on the CPU (include/accfft_operators.h:38-59).
 *
 * Inline wrappers over the C ABI (accfft_amd.h); blocking like the
 * reference.  Timers accumulate the GPU time in timer[4] (doc/timing.md);
 * nothing is written to timer[6] (the reference writes past a 6-element
 * array there, src/operators_gpu.txx:87,136).
 */
#ifndef ACCFFT_OPERATORS_GPU_H
#define ACCFFT_OPERATORS_GPU_H

#include <bitset>

#include "accfft_gpu.h"
#include "accfft_gpuf.h"

namespace accfft_amd_detail {
template <class P>
inline bool baked(P* plan) {
  if (!plan || !plan->r2c_plan_baked) {
    int procid = plan ? plan->procid : 0;
    PCOUT << "Error in accfft_grad! plan is not correctly made." << std::endl;
    return false;
  }
  return true;
}
inline unsigned xyz_or_all(std::bitset<3>* p) { return p ? mask(*p) : 7u; }
}  // namespace accfft_amd_detail

/* operators_gpu.txx:172-278 (one fused per-axis pass per component) */
template <typename T, typename Tp>
inline void accfft_grad_gpu_t(T* A_x, T* A_y, T* A_z, T* A, Tp* plan, std::bitset<3>* pXYZ = NULL,
                              double* timer = NULL) {
  if (!accfft_amd_detail::baked(plan)) return;
  accfft_amd_detail::finish(plan,
                            accfft_amd_grad(plan->h, A_x, A_y, A_z, A,
                                            accfft_amd_detail::xyz_or_all(pXYZ), timer, 0),
                            "grad");
}
/* operators_gpu.txx:72-170 */
template <typename T, typename Tp>
inline void accfft_grad_gpu_slow_t(T* A_x, T* A_y, T* A_z, T* A, Tp* plan,
                                   std::bitset<3>* pXYZ = NULL, double* timer = NULL) {
  if (!accfft_amd_detail::baked(plan)) return;
  accfft_amd_detail::finish(plan,
                            accfft_amd_grad(plan->h, A_x, A_y, A_z, A,
                                            accfft_amd_detail::xyz_or_all(pXYZ), timer, 1),
                            "grad_slow");
}
/* operators_gpu.txx:280-334 */
template <typename T, typename Tp>
inline void accfft_laplace_gpu_t(T* LA, T* A, Tp* plan, double* timer = NULL) {
  if (!accfft_amd_detail::baked(plan)) return;
  accfft_amd_detail::finish(plan, accfft_amd_laplace(plan->h, LA, A, timer), "laplace");
}
/* operators_gpu.txx:448-553 */
template <typename T, typename Tp>
inline void accfft_divergence_gpu_t(T* divA, T* A_x, T* A_y, T* A_z, Tp* plan,
                                    double* timer = NULL) {
  if (!accfft_amd_detail::baked(plan)) return;
  accfft_amd_detail::finish(plan, accfft_amd_divergence(plan->h, divA, A_x, A_y, A_z, timer, 0),
                            "divergence");
}
/* operators_gpu.txx:336-446 */
template <typename T, typename Tp>
inline void accfft_divergence_gpu_slow_t(T* divA, T* A_x, T* A_y, T* A_z, Tp* plan,
                                         double* timer = NULL) {
  if (!accfft_amd_detail::baked(plan)) return;
  accfft_amd_detail::finish(plan, accfft_amd_divergence(plan->h, divA, A_x, A_y, A_z, timer, 1),
                            "divergence_slow");
}
/* operators_gpu.txx:555-609 */
template <typename T, typename Tp>
inline void accfft_biharmonic_gpu_t(T* LA, T* A, Tp* plan, double* timer = NULL) {
  if (!accfft_amd_detail::baked(plan)) return;
  accfft_amd_detail::finish(plan, accfft_amd_biharmonic(plan->h, LA, A, timer), "biharmonic");
}
/* GPU counterparts of accfft_inv_laplace_t / accfft_inv_biharmonic_t
 * (src/operators.txx:1270-1376, CPU-only upstream) */
template <typename T, typename Tp>
inline void accfft_inv_laplace_gpu_t(T* invLA, T* A, Tp* plan, double* timer = NULL) {
  if (!accfft_amd_detail::baked(plan)) return;
  accfft_amd_detail::finish(plan, accfft_amd_inv_laplace(plan->h, invLA, A, timer), "inv_laplace");
}
template <typename T, typename Tp>
inline void accfft_inv_biharmonic_gpu_t(T* invBA, T* A, Tp* plan, double* timer = NULL) {
  if (!accfft_amd_detail::baked(plan)) return;
  accfft_amd_detail::finish(plan, accfft_amd_inv_biharmonic(plan->h, invBA, A, timer),
                            "inv_biharmonic");
}

/* FP64 entry points (src/operators_gpu.cpp:85-135) */
inline void accfft_grad_gpu(double* A_x, double* A_y, double* A_z, double* A,
                            accfft_plan_gpu* plan, std::bitset<3>* pXYZ = NULL,
                            double* timer = NULL) {
  accfft_grad_gpu_t<double, accfft_plan_gpu>(A_x, A_y, A_z, A, plan, pXYZ, timer);
}
inline void accfft_laplace_gpu(double* LA, double* A, accfft_plan_gpu* plan,
                               double* timer = NULL) {
  accfft_laplace_gpu_t<double, accfft_plan_gpu>(LA, A, plan, timer);
}
inline void accfft_divergence_gpu(double* divA, double* A_x, double* A_y, double* A_z,
                                  accfft_plan_gpu* plan, double* timer = NULL) {
  accfft_divergence_gpu_t<double, accfft_plan_gpu>(divA, A_x, A_y, A_z, plan, timer);
}
/* declared but never defined upstream (include/accfft_operators_gpu.h:51) */
inline void accfft_divergence_slow_gpu(double* divA, double* A_x, double* A_y, double* A_z,
                                       accfft_plan_gpu* plan, double* timer = NULL) {
  accfft_divergence_gpu_slow_t<double, accfft_plan_gpu>(divA, A_x, A_y, A_z, plan, timer);
}
inline void accfft_biharmonic_gpu(double* LA, double* A, accfft_plan_gpu* plan,
                                  double* timer = NULL) {
  accfft_biharmonic_gpu_t<double, accfft_plan_gpu>(LA, A, plan, timer);
}
inline void accfft_inv_laplace_gpu(double* invLA, double* A, accfft_plan_gpu* plan,
                                   double* timer = NULL) {
  accfft_inv_laplace_gpu_t<double, accfft_plan_gpu>(invLA, A, plan, timer);
}
inline void accfft_inv_biharmonic_gpu(double* invBA, double* A, accfft_plan_gpu* plan,
                                      double* timer = NULL) {
  accfft_inv_biharmonic_gpu_t<double, accfft_plan_gpu>(invBA, A, plan, timer);
}

/* FP32 entry points (src/operators_gpuf.cpp:93-143) */
inline void accfft_grad_gpuf(float* A_x, float* A_y, float* A_z, float* A,
                             accfft_plan_gpuf* plan, std::bitset<3>* pXYZ = NULL,
                             double* timer = NULL) {
  accfft_grad_gpu_t<float, accfft_plan_gpuf>(A_x, A_y, A_z, A, plan, pXYZ, timer);
}
inline void accfft_laplace_gpuf(float* LA, float* A, accfft_plan_gpuf* plan,
                                double* timer = NULL) {
  accfft_laplace_gpu_t<float, accfft_plan_gpuf>(LA, A, plan, timer);
}
inline void accfft_divergence_slow_gpuf(float* divA, float* A_x, float* A_y, float* A_z,
                                        accfft_plan_gpuf* plan, double* timer = NULL) {
  accfft_divergence_gpu_slow_t<float, accfft_plan_gpuf>(divA, A_x, A_y, A_z, plan, timer);
}
inline void accfft_divergence_gpuf(float* divA, float* A_x, float* A_y, float* A_z,
                                   accfft_plan_gpuf* plan, double* timer = NULL) {
  accfft_divergence_gpu_t<float, accfft_plan_gpuf>(divA, A_x, A_y, A_z, plan, timer);
}
inline void accfft_biharmonic_gpuf(float* LA, float* A, accfft_plan_gpuf* plan,
                                   double* timer = NULL) {
  accfft_biharmonic_gpu_t<float, accfft_plan_gpuf>(LA, A, plan, timer);
}
inline void accfft_inv_laplace_gpuf(float* invLA, float* A, accfft_plan_gpuf* plan,
                                    double* timer = NULL) {
  accfft_inv_laplace_gpu_t<float, accfft_plan_gpuf>(invLA, A, plan, timer);
}
inline void accfft_inv_biharmonic_gpuf(float* invBA, float* A, accfft_plan_gpuf* plan,
                                       double* timer = NULL) {
  accfft_inv_biharmonic_gpu_t<float, accfft_plan_gpuf>(invBA, A, plan, timer);
}

#endif /* ACCFFT_OPERATORS_GPU_H */

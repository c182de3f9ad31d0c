/*
 * accfft_gpuf.h -- drop-in replacement of the reference's
 * include/accfft_gpuf.h (FP32 GPU API, :97-158) for one MI355X on a 1x1
 * grid.  Inline wrappers over the C ABI (accfft_amd.h); blocking like the
 * reference.  Link with -laccfft_amd.
 */
#ifndef ACCFFT_GPUF_H
#define ACCFFT_GPUF_H

#include <bitset>
#include <cstddef>

#include "accfft_amd_dropin.hpp"
#include "accfft_common.h"

struct accfft_plan_gpuf : accfft_amd_detail::PlanFields<float, Complexf> {};

inline int dfft_get_local_size_gpuf(int N0, int N1, int N2, int* isize, int* istart,
                                    MPI_Comm c_comm) {
  return dfft_get_local_size_t<float>(N0, N1, N2, isize, istart, c_comm);
}

inline void accfft_cleanup_gpuf() {}

inline void accfft_destroy_plan_gpu(accfft_plan_gpuf* plan) { accfft_amd_detail::destroy(plan); }
inline void accfft_destroy_plan(accfft_plan_gpuf* plan) { accfft_destroy_plan_gpu(plan); }

/* src/accfft_gpuf.cpp:443-630 */
inline void accfft_execute_gpuf(accfft_plan_gpuf* plan, int direction, float* data_d = NULL,
                                float* data_out_d = NULL, double* timer = NULL,
                                std::bitset<3> xyz = 111) {
  if (!plan || !plan->r2c_plan_baked) {
    accfft_amd_detail::not_baked(plan ? plan->procid : 0);
    return;
  }
  const unsigned m = accfft_amd_detail::mask(xyz);
  const int st = direction == -1 ? accfft_amd_execute_r2c(plan->h, data_d, data_out_d, timer, m)
                                 : accfft_amd_execute_c2r(plan->h, data_d, data_out_d, timer, m);
  accfft_amd_detail::finish(plan, st, "execute");
}

/* src/accfft_gpuf.cpp:632-654 */
inline void accfft_execute_r2c_gpuf(accfft_plan_gpuf* plan, float* data = NULL,
                                    Complexf* data_out = NULL, double* timer = NULL,
                                    std::bitset<3> xyz = 111) {
  accfft_execute_gpuf(plan, -1, data, (float*)data_out, timer, xyz);
}

/* src/accfft_gpuf.cpp:656-676 */
inline void accfft_execute_c2r_gpuf(accfft_plan_gpuf* plan, Complexf* data = NULL,
                                    float* data_out = NULL, double* timer = NULL,
                                    std::bitset<3> xyz = 111) {
  accfft_execute_gpuf(plan, 1, (float*)data, data_out, timer, xyz);
}

/* src/accfft_gpuf.cpp:1137 */
inline void accfft_execute_c2c_gpuf(accfft_plan_gpuf* plan, int direction, Complexf* data_d = NULL,
                                    Complexf* data_out_d = NULL, double* timer = NULL,
                                    std::bitset<3> xyz = 111) {
  if (!plan || !plan->c2c_plan_baked) {
    accfft_amd_detail::not_baked(plan ? plan->procid : 0);
    return;
  }
  const int st = accfft_amd_execute_c2c(plan->h, direction, data_d, data_out_d, timer,
                                        accfft_amd_detail::mask(xyz));
  accfft_amd_detail::finish(plan, st, "execute_c2c");
}

/* src/accfft_gpuf.cpp:697-934 */
inline accfft_plan_gpuf* accfft_plan_dft_3d_c2c_gpuf(int* n, Complexf* data_d,
                                                     Complexf* data_out_d, MPI_Comm c_comm,
                                                     unsigned flags = ACCFFT_MEASURE) {
  return accfft_amd_detail::make_plan<accfft_plan_gpuf>(false, n, data_d, data_out_d, c_comm,
                                                        flags, ACCFFT_AMD_F32);
}

inline int accfft_local_size_dft_c2c_gpuf(int* n, int* isize, int* istart, int* osize,
                                          int* ostart, MPI_Comm c_comm) {
  return accfft_local_size_dft_c2c_t<float>(n, isize, istart, osize, ostart, c_comm);
}

/* src/accfft_gpuf.cpp:74-441 */
inline accfft_plan_gpuf* accfft_plan_dft_3d_r2c_gpuf(int* n, float* data_d, float* data_out_d,
                                                     MPI_Comm c_comm,
                                                     unsigned flags = ACCFFT_MEASURE) {
  return accfft_amd_detail::make_plan<accfft_plan_gpuf>(true, n, data_d, data_out_d, c_comm,
                                                        flags, ACCFFT_AMD_F32);
}

/* src/accfft_gpuf.cpp:59-62 */
inline int accfft_local_size_dft_r2c_gpuf(int* n, int* isize, int* istart, int* osize,
                                          int* ostart, MPI_Comm c_comm) {
  return accfft_local_size_dft_r2c_t<float>(n, isize, istart, osize, ostart, c_comm);
}

template <typename T, typename Tc>
inline void accfft_execute_r2c_gpu_t(accfft_plan_gpuf* plan, T* data, Tc* data_out,
                                     double* timer = NULL, std::bitset<3> XYZ = 111) {
  accfft_execute_gpuf(plan, -1, (float*)data, (float*)data_out, timer, XYZ);
}
template <typename Tc, typename T>
inline void accfft_execute_c2r_gpu_t(accfft_plan_gpuf* plan, Tc* data, T* data_out,
                                     double* timer = NULL, std::bitset<3> XYZ = 111) {
  accfft_execute_gpuf(plan, 1, (float*)data, (float*)data_out, timer, XYZ);
}

namespace accfft_amd_detail {
template <class P>
inline void axis_exec_f(P* plan, int axis, bool fwd, const void* in, void* out, double* timer) {
  if (!plan || !plan->r2c_plan_baked) {
    not_baked(plan ? plan->procid : 0);
    return;
  }
  const int st = fwd ? accfft_amd_execute_r2c_axis(plan->h, axis, in, out, timer)
                     : accfft_amd_execute_c2r_axis(plan->h, axis, in, out, timer);
  finish(plan, st, "execute_axis");
}
}  // namespace accfft_amd_detail

template <typename T, typename Tc>
inline void accfft_execute_r2c_z_gpu_t(accfft_plan_gpuf* plan, T* data, Tc* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec_f(plan, 2, true, data, data_out, timer);
}
template <typename Tc, typename T>
inline void accfft_execute_c2r_z_gpu_t(accfft_plan_gpuf* plan, Tc* data, T* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec_f(plan, 2, false, data, data_out, timer);
}
template <typename T, typename Tc>
inline void accfft_execute_r2c_y_gpu_t(accfft_plan_gpuf* plan, T* data, Tc* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec_f(plan, 1, true, data, data_out, timer);
}
template <typename Tc, typename T>
inline void accfft_execute_c2r_y_gpu_t(accfft_plan_gpuf* plan, Tc* data, T* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec_f(plan, 1, false, data, data_out, timer);
}
template <typename T, typename Tc>
inline void accfft_execute_r2c_x_gpu_t(accfft_plan_gpuf* plan, T* data, Tc* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec_f(plan, 0, true, data, data_out, timer);
}
template <typename Tc, typename T>
inline void accfft_execute_c2r_x_gpu_t(accfft_plan_gpuf* plan, Tc* data, T* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec_f(plan, 0, false, data, data_out, timer);
}

#endif /* ACCFFT_GPUF_H */

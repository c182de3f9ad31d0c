/*
 * accfft_gpu.h -- drop-in replacement of the reference's include/accfft_gpu.h
 * (FP64 GPU API, :97-155) for one MI355X on a 1x1 grid.
 *
 * Same struct name, function names, signatures and default arguments; each
 * function is an inline wrapper over the C ABI of libaccfft_amd.so
 * (accfft_amd.h) and, like the reference, blocks until the GPU work is done.
 * Link with -laccfft_amd.
 */
#ifndef ACCFFT_GPU_H
#define ACCFFT_GPU_H

#include <bitset>
#include <cstddef>

#include "accfft_amd_dropin.hpp"
#include "accfft_common.h"

struct accfft_plan_gpu : accfft_amd_detail::PlanFields<double, Complex> {};

inline int dfft_get_local_size_gpu(int N0, int N1, int N2, int* isize, int* istart,
                                   MPI_Comm c_comm) {
  return dfft_get_local_size_t<double>(N0, N1, N2, isize, istart, c_comm);
}

/* src/accfft_gpu.cpp:62-65 */
inline int accfft_local_size_dft_r2c_gpu(int* n, int* isize, int* istart, int* osize, int* ostart,
                                         MPI_Comm c_comm) {
  return accfft_local_size_dft_r2c_t<double>(n, isize, istart, osize, ostart, c_comm);
}

/* src/accfft_gpu.cpp:77-442 */
inline accfft_plan_gpu* accfft_plan_dft_3d_r2c_gpu(int* n, double* data_d, double* data_out_d,
                                                   MPI_Comm c_comm,
                                                   unsigned flags = ACCFFT_MEASURE) {
  return accfft_amd_detail::make_plan<accfft_plan_gpu>(true, n, data_d, data_out_d, c_comm, flags,
                                                       ACCFFT_AMD_F64);
}

/* src/accfft_gpu.cpp:689-692 */
inline int accfft_local_size_dft_c2c_gpu(int* n, int* isize, int* istart, int* osize, int* ostart,
                                         MPI_Comm c_comm) {
  return accfft_local_size_dft_c2c_t<double>(n, isize, istart, osize, ostart, c_comm);
}

/* src/accfft_gpu.cpp:704-931 */
inline accfft_plan_gpu* accfft_plan_dft_3d_c2c_gpu(int* n, Complex* data_d, Complex* data_out_d,
                                                   MPI_Comm c_comm,
                                                   unsigned flags = ACCFFT_MEASURE) {
  return accfft_amd_detail::make_plan<accfft_plan_gpu>(false, n, data_d, data_out_d, c_comm,
                                                       flags, ACCFFT_AMD_F64);
}

/* src/accfft_gpu.cpp:1150-1192 (also frees the plan struct) */
inline void accfft_destroy_plan_gpu(accfft_plan_gpu* plan) { accfft_amd_detail::destroy(plan); }
inline void accfft_destroy_plan(accfft_plan_gpu* plan) { accfft_destroy_plan_gpu(plan); }

/* src/accfft_gpu.cpp:444-623 */
inline void accfft_execute_gpu(accfft_plan_gpu* plan, int direction, double* data_d,
                               double* data_out_d, double* timer = NULL,
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

/* src/accfft_gpu.cpp:635-653 */
inline void accfft_execute_r2c_gpu(accfft_plan_gpu* plan, double* data, Complex* data_out,
                                   double* timer = NULL, std::bitset<3> xyz = 111) {
  accfft_execute_gpu(plan, -1, data, (double*)data_out, timer, xyz);
}

/* src/accfft_gpu.cpp:655-675 */
inline void accfft_execute_c2r_gpu(accfft_plan_gpu* plan, Complex* data, double* data_out,
                                   double* timer = NULL, std::bitset<3> xyz = 111) {
  accfft_execute_gpu(plan, 1, (double*)data, data_out, timer, xyz);
}

/* src/accfft_gpu.cpp:943-1136 */
inline void accfft_execute_c2c_gpu(accfft_plan_gpu* plan, int direction, Complex* data_d,
                                   Complex* data_out_d, double* timer = NULL,
                                   std::bitset<3> xyz = 111) {
  if (!plan || !plan->c2c_plan_baked) {
    accfft_amd_detail::not_baked(plan ? plan->procid : 0);
    return;
  }
  const int st = accfft_amd_execute_c2c(plan->h, direction, data_d, data_out_d, timer,
                                        accfft_amd_detail::mask(xyz));
  accfft_amd_detail::finish(plan, st, "execute_c2c");
}

/* src/accfft_gpu.cpp:45-47 */
inline void accfft_cleanup_gpu() {}

template <typename T, typename Tc>
inline void accfft_execute_r2c_gpu_t(accfft_plan_gpu* plan, T* data, Tc* data_out,
                                     double* timer = NULL, std::bitset<3> XYZ = 111) {
  accfft_execute_gpu(plan, -1, (double*)data, (double*)data_out, timer, XYZ);
}
template <typename Tc, typename T>
inline void accfft_execute_c2r_gpu_t(accfft_plan_gpu* plan, Tc* data, T* data_out,
                                     double* timer = NULL, std::bitset<3> XYZ = 111) {
  accfft_execute_gpu(plan, 1, (double*)data, (double*)data_out, timer, XYZ);
}

template <typename T>
inline int accfft_local_size_dft_r2c_gpu_t(int* n, int* isize, int* istart, int* osize,
                                           int* ostart, MPI_Comm c_comm) {
  return accfft_local_size_dft_r2c_t<T>(n, isize, istart, osize, ostart, c_comm);
}

namespace accfft_amd_detail {
template <class P>
inline void axis_exec(P* plan, int axis, bool fwd, const void* in, void* out, double* timer) {
  if (!plan || !plan->r2c_plan_baked) {
    not_baked(plan ? plan->procid : 0);
    return;
  }
  const int st = fwd ? accfft_amd_execute_r2c_axis(plan->h, axis, in, out, timer)
                     : accfft_amd_execute_c2r_axis(plan->h, axis, in, out, timer);
  finish(plan, st, "execute_axis");
}
}  // namespace accfft_amd_detail

/* Per-axis transforms, src/accfft_gpu.cpp:1212-1604: z -> [N0][N1][N2/2+1],
 * y -> [N0][N1/2+1][N2], x -> [N0/2+1][N1][N2]. */
template <typename T, typename Tc>
inline void accfft_execute_r2c_z_gpu_t(accfft_plan_gpu* plan, T* data, Tc* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec(plan, 2, true, data, data_out, timer);
}
template <typename Tc, typename T>
inline void accfft_execute_c2r_z_gpu_t(accfft_plan_gpu* plan, Tc* data, T* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec(plan, 2, false, data, data_out, timer);
}
template <typename T, typename Tc>
inline void accfft_execute_r2c_y_gpu_t(accfft_plan_gpu* plan, T* data, Tc* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec(plan, 1, true, data, data_out, timer);
}
template <typename Tc, typename T>
inline void accfft_execute_c2r_y_gpu_t(accfft_plan_gpu* plan, Tc* data, T* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec(plan, 1, false, data, data_out, timer);
}
template <typename T, typename Tc>
inline void accfft_execute_r2c_x_gpu_t(accfft_plan_gpu* plan, T* data, Tc* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec(plan, 0, true, data, data_out, timer);
}
template <typename Tc, typename T>
inline void accfft_execute_c2r_x_gpu_t(accfft_plan_gpu* plan, Tc* data, T* data_out,
                                       double* timer = NULL) {
  accfft_amd_detail::axis_exec(plan, 0, false, data, data_out, timer);
}

#endif /* ACCFFT_GPU_H */

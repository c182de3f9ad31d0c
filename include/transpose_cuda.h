/*
 * transpose_cuda.h -- drop-in replacement of the local-transpose part of the
 * reference's include/transpose_cuda.h (:133-218): the same template
 * functions, running the gfx950 kernels of libaccfft_amd.so on the null
 * stream.  (The header keeps the reference's file name so that existing
 * #include lines resolve; nothing in it uses CUDA.)  The MPI transpose plans
 * (T_Plan_gpu, Mem_Mgr_gpu) have no counterpart on one GPU: at P = 1 they
 * return at entry (src/transpose_gpu.cpp:3819-3821, :1164-1179).
 */
#ifndef ACCFFT_AMD_TRANSPOSE_CUDA_H
#define ACCFFT_AMD_TRANSPOSE_CUDA_H

#include <stddef.h>
#include <stdio.h>

#include "accfft_amd.h"

namespace accfft_amd_detail {
inline void tr_check(int st, const char* what) {
  if (st != ACCFFT_AMD_OK)
    fprintf(stderr, "accfft_amd: %s failed: %s\n", what, accfft_amd_status_string(st));
}
}  // namespace accfft_amd_detail

/* out-of-place [r][c][t] -> [c][r][t] (src/transpose_cuda.cu:242-257) */
template <typename T>
inline void local_transpose_cuda(int r, int c, int n_tuples, T* in, T* out) {
  accfft_amd_detail::tr_check(
      accfft_amd_local_transpose(r, c, n_tuples, in, out, (int)sizeof(T), nullptr),
      "local_transpose_cuda");
}

/* unpack with a ragged last row (src/transpose_cuda.cu:207-224) */
template <typename T>
inline void local_transpose_cuda(int r, int c, int n_tuples, int n_tuples2, T* in, T* out) {
  accfft_amd_detail::tr_check(accfft_amd_local_transpose_ragged(r, c, n_tuples, n_tuples2, in,
                                                                out, (int)sizeof(T), nullptr),
                              "local_transpose_cuda");
}

/* in place (src/transpose_cuda.cu:227-239) */
template <typename T>
inline void local_transpose_cuda(int r, int c, int n_tuples, T* A) {
  accfft_amd_detail::tr_check(
      accfft_amd_local_transpose_inplace(r, c, n_tuples, A, (int)sizeof(T), nullptr),
      "local_transpose_cuda");
}

/* pack with a ragged last column (src/transpose_cuda.cu:78-95) */
template <typename T>
inline void local_transpose_col_cuda(int r, int c, int n_tuples, int n_tuples2, T* in, T* out) {
  accfft_amd_detail::tr_check(accfft_amd_local_transpose_col(r, c, n_tuples, n_tuples2, in, out,
                                                             (int)sizeof(T), nullptr),
                              "local_transpose_col_cuda");
}

/* pack for the pencil exchange (src/transpose_cuda.cu:97-143) */
template <typename T>
inline void memcpy_v1_h1(int nprocs_1, int howmany, int local_n0, int n_tuples,
                         ptrdiff_t* local_n1_proc, T* send_recv_d, T* data, int idist, int N1,
                         ptrdiff_t* local_1_start_proc) {
  static_assert(sizeof(ptrdiff_t) == sizeof(int64_t), "64-bit ptrdiff_t");
  accfft_amd_detail::tr_check(
      accfft_amd_pencil_pack(nprocs_1, howmany, local_n0, n_tuples, (const int64_t*)local_n1_proc,
                             send_recv_d, data, idist, N1, (const int64_t*)local_1_start_proc,
                             (int)sizeof(T), nullptr),
      "memcpy_v1_h1");
}

/* unpack after the pencil exchange (src/transpose_cuda.cu:145-197) */
template <typename T>
inline void memcpy_v1_h2(int nprocs_0, int howmany, ptrdiff_t* local_0_start_proc,
                         ptrdiff_t* local_n0_proc, T* data, int odist, int local_n1,
                         int n_tuples, T* send_recv_cpu) {
  static_assert(sizeof(ptrdiff_t) == sizeof(int64_t), "64-bit ptrdiff_t");
  accfft_amd_detail::tr_check(
      accfft_amd_pencil_unpack(nprocs_0, howmany, (const int64_t*)local_0_start_proc,
                               (const int64_t*)local_n0_proc, data, odist, local_n1, n_tuples,
                               send_recv_cpu, (int)sizeof(T), nullptr),
      "memcpy_v1_h2");
}

#endif /* ACCFFT_AMD_TRANSPOSE_CUDA_H */

/*
 * accfft_common.h -- drop-in replacement of the reference's
 * include/accfft_common.h for the single-GPU (1x1 grid) build.
 *
 * Same macros (ACCFFT_FORWARD/BACKWARD/INVERSE, ACCFFT_ESTIMATE/MEASURE/
 * PATIENT, include/accfft_common.h:37-42), the Complex typedefs (:31-32) and
 * the functions accfft_create_comm, accfft_init, accfft_alloc, accfft_free
 * and the local-size templates (:44-58), inline over the C ABI in
 * accfft_amd.h.
 *
 * Communicator: with ACCFFT_AMD_USE_MPI defined, <mpi.h> is used and the
 * communicator must hold one rank.  Otherwise a single-rank stand-in
 * (accfft_mpi1.h) provides MPI_Comm and the few MPI calls AccFFT programs
 * make, so that a one-GPU program builds without an MPI installation.
 */
#ifndef ACCFFT_COMMON_H
#define ACCFFT_COMMON_H

#ifdef ACCFFT_AMD_USE_MPI
#include <mpi.h>
#else
#include "accfft_mpi1.h"
#endif

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <cmath>
#include <cstddef>
#include <iostream>

#include "accfft_amd.h"

#define PCOUT if (procid == 0) std::cout
typedef double Complex[2];
typedef float Complexf[2];

#define ACCFFT_FORWARD -1
#define ACCFFT_BACKWARD +1
#define ACCFFT_INVERSE +1
#define ACCFFT_ESTIMATE 1
#define ACCFFT_MEASURE 2
#define ACCFFT_PATIENT 4

namespace accfft_amd_detail {
// Returns the process grid of a communicator: this build runs on one rank.
inline bool single_rank(MPI_Comm c) {
  int size = 1;
  MPI_Comm_size(c, &size);
  if (size != 1) {
    fprintf(stderr, "accfft_amd: the single-GPU build supports one rank, got %d\n", size);
    return false;
  }
  return true;
}
// int return of the reference; -1 when the 64-bit value does not fit
// (src/accfft_common.cpp:179-184 reports -1 on overflow).
inline int to_int(int64_t v) { return (v < 0 || v > 2147483647LL) ? -1 : (int)v; }
}  // namespace accfft_amd_detail

/* accfft_create_comm (src/accfft_common.cpp:275-305): c_dims must multiply
 * to the communicator size (1 here); otherwise they are reset, as upstream. */
inline void accfft_create_comm(MPI_Comm in_comm, int* c_dims, MPI_Comm* c_comm) {
  int nprocs = 1;
  MPI_Comm_size(in_comm, &nprocs);
  if (c_dims[0] * c_dims[1] != nprocs) {
    c_dims[0] = nprocs;
    c_dims[1] = 1;
  }
  *c_comm = in_comm;
}

inline int accfft_init() { return 0; }
inline int accfft_init(int) { return 0; }

/* Host allocation (src/accfft_common.cpp:35-46 uses fftw_malloc). */
inline void* accfft_alloc(ptrdiff_t size) {
  void* p = nullptr;
  if (posix_memalign(&p, 64, (size_t)(size > 0 ? size : 1)) != 0) return nullptr;
  return p;
}
inline void accfft_free(void* ptr) { free(ptr); }

/* dfft_get_local_size_t (src/accfft_common.cpp:49-91) on a 1x1 grid. */
template <typename T>
inline int dfft_get_local_size_t(int N0, int N1, int N2, int* isize, int* istart, MPI_Comm) {
  isize[0] = N0;
  isize[1] = N1;
  isize[2] = N2;
  istart[0] = istart[1] = istart[2] = 0;
  return accfft_amd_detail::to_int((int64_t)N0 * N1 * N2 * (int64_t)sizeof(T));
}

/* accfft_local_size_dft_r2c_t<T> (src/accfft_common.cpp:98-192): the byte
 * count is the element count times sizeof(T), whatever T is (double,
 * Complex, float, Complexf), as upstream. */
template <typename T>
inline int accfft_local_size_dft_r2c_t(int* n, int* isize, int* istart, int* osize, int* ostart,
                                       MPI_Comm c_comm) {
  if (!accfft_amd_detail::single_rank(c_comm)) return -1;
  int64_t amax = 0;  // with 4-byte elements
  if (accfft_amd_local_size_dft_r2c(n, isize, istart, osize, ostart, ACCFFT_AMD_F32, 1, 1, 0, 0,
                                    &amax))
    return -1;
  return accfft_amd_detail::to_int(amax / 4 * (int64_t)sizeof(T));
}

/* accfft_local_size_dft_c2c_t<T> (src/accfft_common.cpp:203-256). */
template <typename T>
inline int accfft_local_size_dft_c2c_t(int* n, int* isize, int* istart, int* osize, int* ostart,
                                       MPI_Comm c_comm) {
  if (!accfft_amd_detail::single_rank(c_comm)) return -1;
  int64_t amax = 0;
  if (accfft_amd_local_size_dft_c2c(n, isize, istart, osize, ostart, ACCFFT_AMD_F32, 1, 1, 0, 0,
                                    &amax))
    return -1;
  return accfft_amd_detail::to_int(amax / 4 * (int64_t)sizeof(T));
}

#endif /* ACCFFT_COMMON_H */

/*
 * accfft_mpi1.h -- single-rank stand-in for the handful of MPI calls that
 * AccFFT programs make (MPI_Init, MPI_Comm_rank/size, MPI_Barrier,
 * MPI_Reduce, MPI_Wtime, MPI_Comm_free, MPI_Finalize), so a one-GPU
 * program written against the reference API (e.g. steps/step1/step1_gpu.cpp)
 * builds without an MPI installation.  Every communicator has one rank.
 * Define ACCFFT_AMD_USE_MPI before including accfft_common.h to use a real
 * <mpi.h> instead.
 */
#ifndef ACCFFT_MPI1_H
#define ACCFFT_MPI1_H

#include <string.h>
#include <time.h>

typedef int MPI_Comm;
typedef int MPI_Datatype;
typedef int MPI_Op;
typedef long long MPI_Offset;

#define MPI_COMM_WORLD 0
#define MPI_COMM_NULL (-1)
#define MPI_SUCCESS 0
#define MPI_DOUBLE 8
#define MPI_FLOAT 4
#define MPI_INT 4
#define MPI_LONG_LONG 8
#define MPI_SUM 1
#define MPI_MAX 2
#define MPI_MIN 3

inline int MPI_Init(int*, char***) { return MPI_SUCCESS; }
inline int MPI_Finalize() { return MPI_SUCCESS; }
inline int MPI_Comm_rank(MPI_Comm, int* r) {
  *r = 0;
  return MPI_SUCCESS;
}
inline int MPI_Comm_size(MPI_Comm, int* s) {
  *s = 1;
  return MPI_SUCCESS;
}
inline int MPI_Barrier(MPI_Comm) { return MPI_SUCCESS; }
inline int MPI_Comm_free(MPI_Comm* c) {
  *c = MPI_COMM_NULL;
  return MPI_SUCCESS;
}
// One rank: a reduction copies count elements of the datatype's size.
inline int MPI_Reduce(const void* s, void* r, int count, MPI_Datatype dt, MPI_Op, int,
                      MPI_Comm) {
  if (s != r) memcpy(r, s, (size_t)count * (size_t)dt);
  return MPI_SUCCESS;
}
inline int MPI_Allreduce(const void* s, void* r, int count, MPI_Datatype dt, MPI_Op, MPI_Comm) {
  if (s != r) memcpy(r, s, (size_t)count * (size_t)dt);
  return MPI_SUCCESS;
}
inline double MPI_Wtime() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

#endif /* ACCFFT_MPI1_H */

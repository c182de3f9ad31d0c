/*
 * accfft_amd_dropin.hpp -- shared pieces of the drop-in C++ headers
 * (accfft_gpu.h, accfft_gpuf.h, accfft_operators_gpu.h): the plan struct
 * layout of include/accfft_gpu.h:37-95 and the blocking call wrapper.
 */
#ifndef ACCFFT_AMD_DROPIN_HPP
#define ACCFFT_AMD_DROPIN_HPP

#include <stdio.h>

#include <bitset>
#include <iostream>

#include "accfft_amd.h"
#include "accfft_common.h"

namespace accfft_amd_detail {

/* The fields of accfft_plan_gpu(f) that user code reads (include/
 * accfft_gpu.h:37-95), filled for a 1x1 grid, plus the C-ABI handle. */
template <typename Real, typename Cplx>
struct PlanFields {
  int N[3] = {0, 0, 0};
  int alloc_max = 0;
  int64_t alloc_max64 = 0;
  int coord[2] = {0, 0}, np[2] = {1, 1}, periods[2] = {0, 0};
  MPI_Comm c_comm{}, row_comm{}, col_comm{};
  int osize_0[3] = {0, 0, 0}, ostart_0[3] = {0, 0, 0};
  int osize_1[3] = {0, 0, 0}, ostart_1[3] = {0, 0, 0};
  int osize_2[3] = {0, 0, 0}, ostart_2[3] = {0, 0, 0};
  int osize_1i[3] = {0, 0, 0}, ostart_1i[3] = {0, 0, 0};
  int osize_2i[3] = {0, 0, 0}, ostart_2i[3] = {0, 0, 0};
  int isize[3] = {0, 0, 0}, istart[3] = {0, 0, 0};
  int osize[3] = {0, 0, 0}, ostart[3] = {0, 0, 0};
  int osize_y[3] = {0, 0, 0}, osize_yi[3] = {0, 0, 0}, ostart_y[3] = {0, 0, 0};
  int osize_x[3] = {0, 0, 0}, osize_xi[3] = {0, 0, 0}, ostart_x[3] = {0, 0, 0};
  Real* data = nullptr;
  Real* data_out = nullptr;
  Cplx* data_c = nullptr;
  Cplx* data_out_c = nullptr;
  int procid = 0;
  bool inplace = false;
  bool oneD = true;
  bool r2c_plan_baked = false;
  bool c2c_plan_baked = false;
  accfft_amd_plan* h = nullptr;
};

inline unsigned mask(std::bitset<3> b) { return (unsigned)(b.to_ulong() & 7ul); }

inline bool ok(int st, const char* what) {
  if (st == ACCFFT_AMD_OK) return true;
  fprintf(stderr, "accfft_amd: %s failed: %s\n", what, accfft_amd_status_string(st));
  return false;
}

/* Fills the staged layouts of a 1x1 grid (src/accfft_gpu.cpp:104-168). */
template <class P>
inline void fill_layouts(P* p, bool r2c) {
  const int n0 = p->N[0], n1 = p->N[1], n2 = p->N[2];
  const int nz = r2c ? n2 / 2 + 1 : n2;
  int s0[3] = {n0, n1, nz}, s1[3] = {n0, n1, nz}, s2[3] = {n0, n1, nz};
  for (int i = 0; i < 3; ++i) {
    p->osize_0[i] = s0[i];
    p->osize_1[i] = s1[i];
    p->osize_2[i] = s2[i];
    p->osize_1i[i] = s1[i];
    p->osize_2i[i] = s2[i];
    p->ostart_0[i] = p->ostart_1[i] = p->ostart_2[i] = p->ostart_1i[i] = p->ostart_2i[i] = 0;
    p->ostart_y[i] = 0;
  }
  p->osize_y[0] = n0; p->osize_y[1] = n1; p->osize_y[2] = n2;
  p->osize_yi[0] = n0; p->osize_yi[1] = n1 / 2 + 1; p->osize_yi[2] = n2;
  p->osize_x[0] = n0; p->osize_x[1] = n1 * n2; p->osize_x[2] = 1;
  p->osize_xi[0] = n0 / 2 + 1; p->osize_xi[1] = n1 * n2; p->osize_xi[2] = 1;
  p->ostart_x[0] = 0; p->ostart_x[1] = -256; p->ostart_x[2] = -256;  // -1<<8 upstream
}

template <class P>
inline P* make_plan(bool r2c, int* n, void* data, void* data_out, MPI_Comm c_comm, unsigned flags,
                    int fp) {
  if (!single_rank(c_comm)) return nullptr;
  accfft_amd_plan* h = nullptr;
  const int st = r2c ? accfft_amd_plan_dft_3d_r2c(n, data, data_out, fp, flags, &h)
                     : accfft_amd_plan_dft_3d_c2c(n, data, data_out, fp, flags, &h);
  if (!ok(st, r2c ? "plan_dft_3d_r2c" : "plan_dft_3d_c2c")) return nullptr;
  P* p = new P;
  p->h = h;
  for (int i = 0; i < 3; ++i) p->N[i] = n[i];
  int kind = 0, inplace = 0, f = 0;
  accfft_amd_plan_info(h, nullptr, p->isize, p->istart, p->osize, p->ostart, &p->alloc_max64, &f,
                       &kind, &inplace);
  p->alloc_max = to_int(p->alloc_max64);
  p->inplace = inplace != 0;
  p->c_comm = p->row_comm = p->col_comm = c_comm;
  fill_layouts(p, r2c);
  if (r2c) {
    p->data = (decltype(p->data))data;
    p->data_out = (decltype(p->data_out))data_out;
    p->r2c_plan_baked = true;
  } else {
    p->data_c = (decltype(p->data_c))data;
    p->data_out_c = (decltype(p->data_out_c))data_out;
    p->c2c_plan_baked = true;
  }
  return p;
}

// Every drop-in call is blocking, as the reference (src/accfft_gpu.cpp:620).
template <class P>
inline void finish(P* p, int st, const char* what) {
  ok(st, what);
  accfft_amd_plan_sync(p->h);
}

template <class P>
inline void destroy(P* p) {
  if (!p) return;
  if (p->h) accfft_amd_destroy_plan(p->h);
  p->h = nullptr;
  delete p;
}

inline void not_baked(int procid) {
  if (procid == 0)
    std::cout << "Error. r2c plan has not been made correctly. Please first create the plan "
                 "before calling execute functions."
              << std::endl;
}

}  // namespace accfft_amd_detail

#endif /* ACCFFT_AMD_DROPIN_HPP */

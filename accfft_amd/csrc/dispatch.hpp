// dispatch.hpp -- per-precision entry points compiled in fft_f32.hip and
// fft_f64.hip and called by the C ABI layer (capi.cpp).
#pragma once
#include <stdint.h>

struct accfft_amd_plan;

#define AFFT_DECLARE(SUF)                                                                   \
  int afft_tables_##SUF(accfft_amd_plan* p);                                                \
  int afft_tune_##SUF(accfft_amd_plan* p);                                                  \
  int afft_r2c_##SUF(accfft_amd_plan* p, const void* in, void* out, unsigned xyz);          \
  int afft_c2r_##SUF(accfft_amd_plan* p, void* in, void* out, unsigned xyz);                \
  int afft_c2c_##SUF(accfft_amd_plan* p, int dir, void* in, void* out, unsigned xyz);       \
  int afft_r2c_axis_##SUF(accfft_amd_plan* p, int axis, const void* in, void* out);         \
  int afft_c2r_axis_##SUF(accfft_amd_plan* p, int axis, const void* in, void* out);         \
  int afft_grad_##SUF(accfft_amd_plan* p, void* Ax, void* Ay, void* Az, const void* A,      \
                      unsigned mask, int slow);                                             \
  int afft_spectral_##SUF(accfft_amd_plan* p, int sym, void* out, const void* A);           \
  int afft_spectral_ws_##SUF(accfft_amd_plan* p, int sym, void* out, const void* A,         \
                             void* ws);                                                     \
  int afft_div_##SUF(accfft_amd_plan* p, void* div, const void* Ax, const void* Ay,         \
                     const void* Az, int slow);                                             \
  int afft_mult_##SUF(int kind, void* wA, const void* A, const int* n, const int* osize,    \
                      const int* ostart, double scale, void* stream);                       \
  int afft_axpy_##SUF(int64_t n, double alpha, const void* x, void* y, void* stream);       \
  int afft_r2c_slab_##SUF(accfft_amd_plan* p, const void* in, void* out);                   \
  int afft_c2r_slab_##SUF(accfft_amd_plan* p, void* in, void* out);                         \
  int afft_r2c_pencil_##SUF(accfft_amd_plan* p, const void* in, void* out);                 \
  int afft_c2r_pencil_##SUF(accfft_amd_plan* p, void* in, void* out);                       \
  int afft_adj_r2c_##SUF(accfft_amd_plan* p, const void* G, void* x);                       \
  int afft_adj_c2r_##SUF(accfft_amd_plan* p, const void* y, void* G);                       \
  int afft_adj_grad_##SUF(accfft_amd_plan* p, void* out, const void* gx, const void* gy,    \
                          const void* gz);                                                  \
  int afft_adj_div_##SUF(accfft_amd_plan* p, void* ox, void* oy, void* oz, const void* v);

// transpose.hip: the inverse of accfft_amd_pencil_pack
int afft_pencil_scatter(int nprocs, int howmany, int local_n0, int64_t n_tuples,
                        const int64_t* cnt, const void* packed, void* data, int64_t idist,
                        int64_t N1, const int64_t* start, int elem_size, void* stream);

AFFT_DECLARE(f32)
AFFT_DECLARE(f64)
#undef AFFT_DECLARE

/*
 * accfft_poisson_gpu.h -- drop-in for the Poisson Fourier filter of the
 * reference's step4 program (steps/step4/kernels.cu:113-146,
 * poisson_fourier_filter_gpu, declared by step4_gpu.cpp:89-93), over the C
 * ABI of libaccfft_amd.so, plus a fused solve.
 *
 * poisson_fourier_filter_gpu(data_hat, N, isize, istart, methodNb): in place
 * on the half spectrum block isize at offset istart (the osize / ostart of
 * an r2c plan), every mode but (0,0,0) divided by the Poisson eigenvalue of
 * the unit box times N[0] N[1] N[2]; methodNb 0 = the 7-point discrete
 * Laplacian, 1 = -4 pi^2 |k|^2; the mean set to zero.  Ordered on the
 * default stream like the reference's kernel launch.
 *
 * accfft_poisson_solve_gpu(f)(phi, rho, plan, methodNb, timer): no
 * counterpart in the reference; the step4 sequence r2c -> filter -> c2r
 * with the filter fused into the x pass, rho preserved.
 */
#ifndef ACCFFT_POISSON_GPU_H
#define ACCFFT_POISSON_GPU_H

#include "accfft_gpu.h"
#include "accfft_gpuf.h"

inline void poisson_fourier_filter_gpu(Complex* data_hat, int N[3], int isize[3], int istart[3],
                                       int methodNb) {
  accfft_amd_detail::ok(accfft_amd_poisson_filter(ACCFFT_AMD_F64, data_hat, N, isize, istart,
                                                  methodNb, NULL),
                        "poisson_fourier_filter_gpu");
}
inline void poisson_fourier_filter_gpu(Complexf* data_hat, int N[3], int isize[3], int istart[3],
                                       int methodNb) {
  accfft_amd_detail::ok(accfft_amd_poisson_filter(ACCFFT_AMD_F32, data_hat, N, isize, istart,
                                                  methodNb, NULL),
                        "poisson_fourier_filter_gpu");
}

inline void accfft_poisson_solve_gpu(double* phi, double* rho, accfft_plan_gpu* plan,
                                     int methodNb, double* timer = NULL) {
  if (!plan || !plan->h) return;
  accfft_amd_detail::ok(accfft_amd_poisson_solve(plan->h, methodNb, phi, rho, timer),
                        "accfft_poisson_solve_gpu");
  accfft_amd_plan_sync(plan->h);
}
inline void accfft_poisson_solve_gpuf(float* phi, float* rho, accfft_plan_gpuf* plan,
                                      int methodNb, double* timer = NULL) {
  if (!plan || !plan->h) return;
  accfft_amd_detail::ok(accfft_amd_poisson_solve(plan->h, methodNb, phi, rho, timer),
                        "accfft_poisson_solve_gpuf");
  accfft_amd_plan_sync(plan->h);
}

#endif /* ACCFFT_POISSON_GPU_H */

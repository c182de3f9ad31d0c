// dropin_steps.cpp -- a user program written against the reference's GPU C++
// API (accfft_gpu.h, accfft_gpuf.h, accfft_operators_gpu.h,
// transpose_cuda.h), built against the drop-in headers of this repository.
//
// It follows the checks of the reference's example programs:
//   step1 (steps/step1/step1_gpu.cpp:21-209): Gaussian exp(-4|x-pi|^2),
//     r2c then c2r, relative L1 of c2r(r2c(a))/N^3 - a below 1e-10 (FP64)
//     or 1e-5 (FP32);
//   step3 (steps/step3/step3_gpu.cpp): the same for c2c;
//   step4 (steps/step4/step4_gpu.cpp:520-640, kernels.cu:24-146): Poisson
//     solve of the sine case r2c -> poisson_fourier_filter_gpu -> c2r against
//     its exact solution (method 1) and the discrete one (method 0), and the
//     fused solve;
//   step5 (steps/step5/step5_gpu.cpp:20-133, :559-699): grad, laplace,
//     biharmonic and divergence of sin x sin y sin z + exp(-32|x-pi|^2)
//     against closed forms (relative L1 < 1e-10 in FP64), plus the GPU
//     inverse Laplacian (invlap(laplace f) = f - mean f).
// The only change from a CUDA build of such a program is hipMalloc /
// hipMemcpy in place of cudaMalloc / cudaMemcpy.  Exit status 0 iff every
// check passes.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include <vector>

#include "accfft_gpu.h"
#include "accfft_gpuf.h"
#include "accfft_operators_gpu.h"
#include "accfft_poisson_gpu.h"
#include "transpose_cuda.h"

static int g_fail = 0;

static void report(const char* what, double err, double tol) {
  const bool ok = err < tol;
  printf("%-40s rel L1 %.3e  %s\n", what, err, ok ? "CORRECT" : "NOT CORRECT");
  if (!ok) g_fail = 1;
}

static double rel_l1(const std::vector<double>& a, const std::vector<double>& ref) {
  double e = 0, n = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    e += fabs(a[i] - ref[i]);
    n += fabs(ref[i]);
  }
  return e / n;
}

// ------------------------------------------------------------------ fields
static double gauss4(double X, double Y, double Z) {
  const double p = M_PI;
  return exp(-4 * ((X - p) * (X - p) + (Y - p) * (Y - p) + (Z - p) * (Z - p)));
}
static const double SIG = 32;
static double f5(double X, double Y, double Z) {
  const double p = M_PI;
  return sin(X) * sin(Y) * sin(Z) +
         exp(-SIG * ((X - p) * (X - p) + (Y - p) * (Y - p) + (Z - p) * (Z - p)));
}
static double f5_grad(double X, double Y, double Z, int d) {
  const double p = M_PI;
  const double g = exp(-SIG * ((X - p) * (X - p) + (Y - p) * (Y - p) + (Z - p) * (Z - p)));
  if (d == 0) return cos(X) * sin(Y) * sin(Z) - 2 * SIG * (X - p) * g;
  if (d == 1) return sin(X) * cos(Y) * sin(Z) - 2 * SIG * (Y - p) * g;
  return sin(X) * sin(Y) * cos(Z) - 2 * SIG * (Z - p) * g;
}
static double f5_lap(double X, double Y, double Z) {
  const double p = M_PI;
  const double r2 = (X - p) * (X - p) + (Y - p) * (Y - p) + (Z - p) * (Z - p);
  return (4 * SIG * SIG * r2 - 6 * SIG) * exp(-SIG * r2) - 3 * sin(X) * sin(Y) * sin(Z);
}
static double f5_bih(double X, double Y, double Z) {
  const double p = M_PI, s2 = SIG * SIG, s3 = s2 * SIG, s4 = s3 * SIG;
  const double x = X - p, y = Y - p, z = Z - p, r2 = x * x + y * y + z * z;
  const double f = exp(-SIG * r2);
  auto d4 = [&](double u) { return 12 * s2 * f - 48 * s3 * u * u * f + 16 * s4 * u * u * u * u * f; };
  auto d22 = [&](double u, double v) {
    return 4 * s2 * f - 8 * s3 * v * v * f - 8 * s3 * u * u * f + 16 * s4 * u * u * v * v * f;
  };
  return 9 * sin(X) * sin(Y) * sin(Z) + d4(x) + d4(y) + d4(z) + 2 * d22(x, y) + 2 * d22(x, z) +
         2 * d22(y, z);
}

template <typename F>
static std::vector<double> sample(const int* n, F f) {
  std::vector<double> v((size_t)n[0] * n[1] * n[2]);
  for (int i = 0; i < n[0]; ++i)
    for (int j = 0; j < n[1]; ++j)
      for (int k = 0; k < n[2]; ++k)
        v[((size_t)i * n[1] + j) * n[2] + k] =
            f(2 * M_PI / n[0] * i, 2 * M_PI / n[1] * j, 2 * M_PI / n[2] * k);
  return v;
}

template <typename T>
static T* to_dev(const std::vector<double>& h) {
  std::vector<T> t(h.begin(), h.end());
  T* d = nullptr;
  (void)hipMalloc((void**)&d, t.size() * sizeof(T));
  (void)hipMemcpy(d, t.data(), t.size() * sizeof(T), hipMemcpyHostToDevice);
  return d;
}
template <typename T>
static std::vector<double> to_host(const T* d, size_t n) {
  std::vector<T> t(n);
  (void)hipMemcpy(t.data(), d, n * sizeof(T), hipMemcpyDeviceToHost);
  return std::vector<double>(t.begin(), t.end());
}

// ------------------------------------------------------------------ step1 / step3
static void step1_fp64(int* n) {
  int c_dims[2] = {0}, isize[3], osize[3], istart[3], ostart[3];
  MPI_Comm c_comm;
  accfft_create_comm(MPI_COMM_WORLD, c_dims, &c_comm);
  const int alloc_max = accfft_local_size_dft_r2c_gpu(n, isize, istart, osize, ostart, c_comm);
  const size_t N = (size_t)isize[0] * isize[1] * isize[2];
  std::vector<double> a = sample(n, gauss4);
  double* data = to_dev<double>(a);
  Complex* data_hat;
  double* data2;
  (void)hipMalloc((void**)&data_hat, alloc_max);
  (void)hipMalloc((void**)&data2, N * sizeof(double));
  accfft_plan_gpu* plan = accfft_plan_dft_3d_r2c_gpu(n, data, (double*)data_hat, c_comm,
                                                     ACCFFT_MEASURE);
  double timer[5] = {0};
  accfft_execute_r2c_gpu(plan, data, data_hat, timer);
  accfft_execute_c2r_gpu(plan, data_hat, data2, timer);
  std::vector<double> b = to_host(data2, N);
  for (double& v : b) v /= (double)n[0] * n[1] * n[2];
  report("step1 r2c/c2r FP64 round trip", rel_l1(b, a), 1e-10);
  printf("  timer[4] (FFT seconds) = %.3e\n", timer[4]);
  accfft_destroy_plan_gpu(plan);
  (void)hipFree(data);
  (void)hipFree(data_hat);
  (void)hipFree(data2);
  accfft_cleanup_gpu();
}

static void step1_fp32(int* n) {
  int c_dims[2] = {1, 1}, isize[3], osize[3], istart[3], ostart[3];
  MPI_Comm c_comm;
  accfft_create_comm(MPI_COMM_WORLD, c_dims, &c_comm);
  const int alloc_max = accfft_local_size_dft_r2c_gpuf(n, isize, istart, osize, ostart, c_comm);
  const size_t N = (size_t)isize[0] * isize[1] * isize[2];
  std::vector<double> a = sample(n, gauss4);
  float* data = to_dev<float>(a);
  Complexf* data_hat;
  float* data2;
  (void)hipMalloc((void**)&data_hat, alloc_max);
  (void)hipMalloc((void**)&data2, N * sizeof(float));
  accfft_plan_gpuf* plan = accfft_plan_dft_3d_r2c_gpuf(n, data, (float*)data_hat, c_comm);
  accfft_execute_r2c_gpuf(plan, data, data_hat);
  accfft_execute_c2r_gpuf(plan, data_hat, data2);
  std::vector<double> b = to_host(data2, N);
  for (double& v : b) v /= (double)n[0] * n[1] * n[2];
  report("step1 r2c/c2r FP32 round trip", rel_l1(b, a), 1e-5);
  accfft_destroy_plan(plan);
  (void)hipFree(data);
  (void)hipFree(data_hat);
  (void)hipFree(data2);
}

// step2: in-place r2c with rows padded to 2*(N2/2+1) reals
static void step2_fp64(int* n) {
  MPI_Comm c_comm;
  int c_dims[2] = {1, 1}, isize[3], osize[3], istart[3], ostart[3];
  accfft_create_comm(MPI_COMM_WORLD, c_dims, &c_comm);
  const int alloc_max = accfft_local_size_dft_r2c_gpu(n, isize, istart, osize, ostart, c_comm);
  const int nh2 = 2 * (n[2] / 2 + 1);
  std::vector<double> a = sample(n, gauss4);
  std::vector<double> pad((size_t)n[0] * n[1] * nh2, 0.0);
  for (size_t l = 0; l < (size_t)n[0] * n[1]; ++l)
    for (int k = 0; k < n[2]; ++k) pad[l * nh2 + k] = a[l * n[2] + k];
  double* data;
  (void)hipMalloc((void**)&data, alloc_max);
  (void)hipMemcpy(data, pad.data(), pad.size() * sizeof(double), hipMemcpyHostToDevice);
  accfft_plan_gpu* plan = accfft_plan_dft_3d_r2c_gpu(n, data, data, c_comm, ACCFFT_MEASURE);
  accfft_execute_r2c_gpu(plan, data, (Complex*)data);
  accfft_execute_c2r_gpu(plan, (Complex*)data, data);
  std::vector<double> b = to_host(data, pad.size()), c(a.size());
  for (size_t l = 0; l < (size_t)n[0] * n[1]; ++l)
    for (int k = 0; k < n[2]; ++k) c[l * n[2] + k] = b[l * nh2 + k] / ((double)n[0] * n[1] * n[2]);
  report("step2 in-place r2c/c2r FP64", rel_l1(c, a), 1e-10);
  accfft_destroy_plan(plan);
  (void)hipFree(data);
}

// step3 with and without INPLACE (steps/step3/step3_gpu.cpp:52-54, 61-62,
// 83-84, 99-100): the in-place build plans and executes with data_hat = data
static void step3_fp64(int* n, bool inplace) {
  MPI_Comm c_comm;
  int c_dims[2] = {1, 1}, isize[3], osize[3], istart[3], ostart[3];
  accfft_create_comm(MPI_COMM_WORLD, c_dims, &c_comm);
  const int alloc_max = accfft_local_size_dft_c2c_gpu(n, isize, istart, osize, ostart, c_comm);
  const size_t N = (size_t)n[0] * n[1] * n[2];
  std::vector<double> a = sample(n, gauss4), z(2 * N);
  for (size_t i = 0; i < N; ++i) {
    z[2 * i] = a[i];
    z[2 * i + 1] = 0.5 * a[i];
  }
  Complex *data, *data_hat;
  (void)hipMalloc((void**)&data, alloc_max);
  if (inplace) data_hat = data;
  else (void)hipMalloc((void**)&data_hat, alloc_max);
  (void)hipMemcpy(data, z.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice);
  accfft_plan_gpu* plan = accfft_plan_dft_3d_c2c_gpu(n, data, data_hat, c_comm, ACCFFT_MEASURE);
  accfft_execute_c2c_gpu(plan, ACCFFT_FORWARD, data, data_hat);
  // spectrum of the forward transform against a direct DFT at two bins
  std::vector<double> h = to_host((double*)data_hat, 2 * N);
  double worst = 0, scale = 0;
  for (size_t l = 0; l < N; ++l) scale = fmax(scale, fabs(h[2 * l]));
  const int bins[2][3] = {{0, 0, 0}, {1, 2, 3 % n[2]}};
  for (int b = 0; b < 2; ++b) {
    double re = 0, im = 0;
    for (int i = 0; i < n[0]; ++i)
      for (int j = 0; j < n[1]; ++j)
        for (int k = 0; k < n[2]; ++k) {
          const double ph = -2 * M_PI * ((double)bins[b][0] * i / n[0] +
                                          (double)bins[b][1] * j / n[1] +
                                          (double)bins[b][2] * k / n[2]);
          const size_t l = ((size_t)i * n[1] + j) * n[2] + k;
          re += z[2 * l] * cos(ph) - z[2 * l + 1] * sin(ph);
          im += z[2 * l] * sin(ph) + z[2 * l + 1] * cos(ph);
        }
    const size_t l = ((size_t)bins[b][0] * n[1] + bins[b][1]) * n[2] + bins[b][2];
    worst = fmax(worst, fmax(fabs(h[2 * l] - re), fabs(h[2 * l + 1] - im)));
  }
  accfft_execute_c2c_gpu(plan, ACCFFT_BACKWARD, data_hat, data);
  std::vector<double> b = to_host((double*)data, 2 * N);
  for (double& v : b) v /= (double)N;
  report(inplace ? "step3 in-place c2c FP64 round trip" : "step3 c2c FP64 round trip",
         rel_l1(b, z), 1e-10);
  report(inplace ? "step3 in-place c2c FP64 DFT bins" : "step3 c2c FP64 DFT bins",
         worst / scale, 1e-10);
  accfft_destroy_plan(plan);
  (void)hipFree(data);
  if (!inplace) (void)hipFree(data_hat);
}

// the FP32 twin of the in-place build (steps/step3/step3_gpuf.cpp)
static void step3_fp32_inplace(int* n) {
  MPI_Comm c_comm;
  int c_dims[2] = {1, 1}, isize[3], osize[3], istart[3], ostart[3];
  accfft_create_comm(MPI_COMM_WORLD, c_dims, &c_comm);
  const int alloc_max = accfft_local_size_dft_c2c_gpuf(n, isize, istart, osize, ostart, c_comm);
  const size_t N = (size_t)n[0] * n[1] * n[2];
  std::vector<double> a = sample(n, gauss4), z(2 * N);
  for (size_t i = 0; i < N; ++i) {
    z[2 * i] = a[i];
    z[2 * i + 1] = 0.5 * a[i];
  }
  std::vector<float> zf(z.begin(), z.end());
  Complexf* data;
  (void)hipMalloc((void**)&data, alloc_max);
  (void)hipMemcpy(data, zf.data(), 2 * N * sizeof(float), hipMemcpyHostToDevice);
  accfft_plan_gpuf* plan = accfft_plan_dft_3d_c2c_gpuf(n, data, data, c_comm, ACCFFT_MEASURE);
  accfft_execute_c2c_gpuf(plan, ACCFFT_FORWARD, data, data);
  accfft_execute_c2c_gpuf(plan, ACCFFT_BACKWARD, data, data);
  std::vector<double> b = to_host((float*)data, 2 * N);
  for (double& v : b) v /= (double)N;
  report("step3 in-place c2c FP32 round trip", rel_l1(b, z), 1e-5);
  accfft_destroy_plan(plan);
  (void)hipFree(data);
}

// ------------------------------------------------------------------ step4
static void step4_fp64(int* n) {
  const size_t N = (size_t)n[0] * n[1] * n[2];
  std::vector<double> rho(N);
  for (int i = 0; i < n[0]; ++i)
    for (int j = 0; j < n[1]; ++j)
      for (int k = 0; k < n[2]; ++k)
        rho[((size_t)i * n[1] + j) * n[2] + k] = sin(2 * M_PI * i / n[0]) *
                                                 sin(2 * M_PI * j / n[1]) *
                                                 sin(2 * M_PI * k / n[2]);
  int isize[3], istart[3], osize[3], ostart[3];
  MPI_Comm c_comm;
  int c_dims[2] = {0, 0};
  accfft_create_comm(MPI_COMM_WORLD, c_dims, &c_comm);
  const int alloc_max = accfft_local_size_dft_r2c_gpu(n, isize, istart, osize, ostart, c_comm);
  double* d_rho = to_dev<double>(rho);
  double* d_phi = nullptr;
  Complex* phi_hat = nullptr;
  (void)hipMalloc((void**)&d_phi, N * sizeof(double));
  (void)hipMalloc((void**)&phi_hat, alloc_max);
  accfft_plan_gpu* plan = accfft_plan_dft_3d_r2c_gpu(n, d_rho, (double*)phi_hat, c_comm, ACCFFT_MEASURE);
  double ev0 = 0;
  for (int d = 0; d < 3; ++d) ev0 += 2 * (cos(2 * M_PI / n[d]) - 1) * n[d] * n[d];
  const double ev[2] = {ev0, -12 * M_PI * M_PI};
  const char* names[2] = {"step4 Poisson method 0 FP64", "step4 Poisson method 1 FP64"};
  for (int m = 0; m < 2; ++m) {
    std::vector<double> exact(N);
    for (size_t q = 0; q < N; ++q) exact[q] = rho[q] / ev[m];
    accfft_execute_r2c_gpu(plan, d_rho, phi_hat);
    poisson_fourier_filter_gpu(phi_hat, n, osize, ostart, m);
    accfft_execute_c2r_gpu(plan, phi_hat, d_phi);
    report(names[m], rel_l1(to_host(d_phi, N), exact), 1e-10);
    accfft_poisson_solve_gpu(d_phi, d_rho, plan, m);
    report(m ? "step4 fused Poisson solve method 1" : "step4 fused Poisson solve method 0",
           rel_l1(to_host(d_phi, N), exact), 1e-10);
  }
  accfft_destroy_plan_gpu(plan);
  (void)hipFree(d_rho);
  (void)hipFree(d_phi);
  (void)hipFree(phi_hat);
}

// ------------------------------------------------------------------ step5
static void step5_fp64(int* n) {
  MPI_Comm c_comm;
  int c_dims[2] = {1, 1}, isize[3], osize[3], istart[3], ostart[3];
  accfft_create_comm(MPI_COMM_WORLD, c_dims, &c_comm);
  const int alloc_max = accfft_local_size_dft_r2c_gpu(n, isize, istart, osize, ostart, c_comm);
  const size_t N = (size_t)n[0] * n[1] * n[2];
  std::vector<double> f = sample(n, f5);
  double* A = to_dev<double>(f);
  double *gx, *gy, *gz, *out, *div;
  Complex* spec;
  for (double** p : {&gx, &gy, &gz, &out, &div}) (void)hipMalloc((void**)p, N * sizeof(double));
  (void)hipMalloc((void**)&spec, alloc_max);
  accfft_plan_gpu* plan = accfft_plan_dft_3d_r2c_gpu(n, A, (double*)spec, c_comm, ACCFFT_MEASURE);
  std::bitset<3> xyz = 0;
  xyz.flip();
  accfft_grad_gpu(gx, gy, gz, A, plan, &xyz);
  const char* names[3] = {"step5 grad x FP64", "step5 grad y FP64", "step5 grad z FP64"};
  double* g[3] = {gx, gy, gz};
  for (int d = 0; d < 3; ++d)
    report(names[d], rel_l1(to_host(g[d], N),
                            sample(n, [d](double X, double Y, double Z) { return f5_grad(X, Y, Z, d); })),
           1e-10);
  accfft_laplace_gpu(out, A, plan);
  const std::vector<double> lap = sample(n, f5_lap);
  report("step5 laplace FP64", rel_l1(to_host(out, N), lap), 1e-10);
  accfft_divergence_gpu(div, gx, gy, gz, plan);
  report("step5 div(grad) = laplace FP64", rel_l1(to_host(div, N), lap), 1e-10);
  accfft_biharmonic_gpu(out, A, plan);
  report("step5 biharmonic FP64", rel_l1(to_host(out, N), sample(n, f5_bih)), 1e-9);
  // invlap(laplace(f)) = f - mean(f)
  double* L = div;
  accfft_laplace_gpu(L, A, plan);
  accfft_inv_laplace_gpu(out, L, plan);
  double mean = 0;
  for (double v : f) mean += v;
  mean /= (double)N;
  std::vector<double> fm(f);
  for (double& v : fm) v -= mean;
  report("inv_laplace(laplace f) = f - mean FP64", rel_l1(to_host(out, N), fm), 1e-10);
  accfft_destroy_plan(plan);
  for (double* p : {A, gx, gy, gz, out, div}) (void)hipFree(p);
  (void)hipFree(spec);
}

// FP32 operators: checked against the FP64 path on the same FP32-rounded
// input.  A smooth field's high-|k| content is ~0, so the FP32 rounding
// noise of any FFT (white over k, ~eps32 = 6e-8 relative) dominates there
// once multiplied by the symbol: first derivatives are bounded by
// eps32 * |k|max (~7e-6 at 128^3), second derivatives by eps32 * |k|^2max
// = 6e-8 * 3 * 64^2 ~ 7e-4.  Those bounds are the thresholds.
static void step5_fp32(int* n) {
  MPI_Comm c_comm;
  int c_dims[2] = {1, 1}, isize[3], osize[3], istart[3], ostart[3];
  accfft_create_comm(MPI_COMM_WORLD, c_dims, &c_comm);
  const int alloc_max = accfft_local_size_dft_r2c_gpuf(n, isize, istart, osize, ostart, c_comm);
  const int alloc64 = accfft_local_size_dft_r2c_gpu(n, isize, istart, osize, ostart, c_comm);
  const size_t N = (size_t)n[0] * n[1] * n[2];
  std::vector<double> f = sample(n, f5);
  for (double& v : f) v = (double)(float)v;
  float* A = to_dev<float>(f);
  double* A64 = to_dev<double>(f);
  float *gx, *gy, *gz, *out;
  double *r64, *gx64, *gy64, *gz64;
  Complexf* spec;
  Complex* spec64;
  for (float** p : {&gx, &gy, &gz, &out}) (void)hipMalloc((void**)p, N * sizeof(float));
  for (double** p : {&r64, &gx64, &gy64, &gz64}) (void)hipMalloc((void**)p, N * sizeof(double));
  (void)hipMalloc((void**)&spec, alloc_max);
  (void)hipMalloc((void**)&spec64, alloc64);
  accfft_plan_gpuf* plan = accfft_plan_dft_3d_r2c_gpuf(n, A, (float*)spec, c_comm);
  accfft_plan_gpu* plan64 = accfft_plan_dft_3d_r2c_gpu(n, A64, (double*)spec64, c_comm);
  accfft_grad_gpuf(gx, gy, gz, A, plan);
  accfft_grad_gpu(gx64, gy64, gz64, A64, plan64);
  const double bound1 = 6e-8 * 1.7320508 * (n[0] / 2);
  report("step5 grad x FP32 vs FP64 path", rel_l1(to_host(gx, N), to_host(gx64, N)), bound1);
  report("step5 grad z FP32 vs FP64 path", rel_l1(to_host(gz, N), to_host(gz64, N)), bound1);
  accfft_laplace_gpuf(out, A, plan);
  accfft_laplace_gpu(r64, A64, plan64);
  const double kmax2 = 3.0 * (n[0] / 2) * (n[0] / 2);
  const double bound2 = 6e-8 * kmax2;
  report("step5 laplace FP32 vs FP64 path", rel_l1(to_host(out, N), to_host(r64, N)), bound2);
  accfft_divergence_gpuf(out, gx, gy, gz, plan);
  report("step5 div(grad) FP32 vs FP64 laplace", rel_l1(to_host(out, N), to_host(r64, N)), bound2);
  accfft_destroy_plan(plan);
  accfft_destroy_plan(plan64);
  for (float* p : {A, gx, gy, gz, out}) (void)hipFree(p);
  for (double* p : {A64, r64, gx64, gy64, gz64}) (void)hipFree(p);
  (void)hipFree(spec);
  (void)hipFree(spec64);
}

// ------------------------------------------------------------------ transposes
static void transposes() {
  const int r = 5, c = 7, t = 3;
  std::vector<double> a(r * c * t), b(r * c * t);
  for (size_t i = 0; i < a.size(); ++i) a[i] = (double)i;
  double *d, *o;
  (void)hipMalloc((void**)&d, a.size() * sizeof(double));
  (void)hipMalloc((void**)&o, a.size() * sizeof(double));
  (void)hipMemcpy(d, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice);
  local_transpose_cuda(r, c, t, d, o);
  local_transpose_cuda(r, c, t, d);  // in place: same permutation
  (void)hipDeviceSynchronize();
  std::vector<double> x = to_host(o, a.size()), y = to_host(d, a.size());
  double e = 0;
  for (int j = 0; j < c; ++j)
    for (int i = 0; i < r; ++i)
      for (int u = 0; u < t; ++u) {
        const double want = a[(i * c + j) * t + u];
        e += fabs(x[(j * r + i) * t + u] - want) + fabs(y[(j * r + i) * t + u] - want);
      }
  report("local_transpose_cuda (out of / in place)", e, 1e-300);
  (void)hipFree(d);
  (void)hipFree(o);
}

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  int n1[3] = {64, 64, 64};
  int n5[3] = {128, 128, 128};
  if (argc == 4) {
    n1[0] = atoi(argv[1]);
    n1[1] = atoi(argv[2]);
    n1[2] = atoi(argv[3]);
  }
  step1_fp64(n1);
  step1_fp32(n1);
  int n2[3] = {32, 48, 30};
  step2_fp64(n2);
  step3_fp64(n2, false);
  step3_fp64(n2, true);
  int n3[3] = {64, 512, 32};
  step3_fp64(n3, true);
  step3_fp32_inplace(n3);
  int n4[3] = {64, 48, 40};
  step4_fp64(n4);
  step5_fp64(n5);
  step5_fp32(n5);
  transposes();
  MPI_Finalize();
  printf(g_fail ? "SOME CHECKS FAILED\n" : "ALL CHECKS PASSED\n");
  return g_fail;
}

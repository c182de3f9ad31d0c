// fft_long.hpp -- transforms of lines too long for one LDS tile (any axis
// length whose tile would exceed the 160 KiB of LDS: FP32 > 10240 points,
// FP64 > 5120 points that are not powers of two up to 4096).
//
// Each radix stage of the line's mixed-radix Stockham schedule (the dynamic
// engine's schedule: radices 16, 8, 4, 2, 9, 3, 5, 7, then any other prime)
// is one pass over HBM: butterfly j of stage (radix R, span NS) reads
// element j + r N/R of every line, multiplies by exp(-+2 pi i r (j mod NS)
// / (NS R)) from the axis's length-N twiddle table, does the R-point DFT
// and writes element (j - j mod NS) R + j mod NS + r NS, so that after the
// last stage the line is in natural order.  Radices above 16 (primes) use
// one thread per output, O(R) work each.  Stages ping-pong between the
// output and a workspace of the same layout.
//
// Lines are addressed as in the strided tiles: element n of line L sits at
// (L / C) D + L % C + n S (contiguous z lines: C = 1, D = pitch, S = 1).
// Real-to-complex z passes pack two real lines a, b into one complex line
// a + i b (one complex FFT per pair), then split
//   A[k] = (Z[k] + conj Z[N-k]) / 2,  B[k] = -i (Z[k] - conj Z[N-k]) / 2;
// complex-to-real passes do the reverse (imaginary parts of DC and, for
// even N, Nyquist dropped, as c2r does).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_device.hpp"

namespace afft {

struct GGeom {
  int64_t C, D, S;
  __device__ __host__ __forceinline__ int64_t base(int64_t line) const {
    return (line / C) * D + line % C;
  }
};

// Lines whose tile does not fit the LDS of one workgroup.
inline bool long_length(int n, size_t csz) {
  return !(n >= 16 && n <= 4096 && (n & (n - 1)) == 0) && 2 * (size_t)n * csz > 160 * 1024;
}

constexpr int GTPB = 256;

// lines_fast: consecutive threads take consecutive lines (strided axes, whose
// lines are adjacent columns); otherwise consecutive butterflies of a line.
template <typename T, int R, int SIGN>
__global__ void __launch_bounds__(GTPB)
    gstage_kernel(const cpx<T>* __restrict__ src, GGeom gs, cpx<T>* __restrict__ dst, GGeom gd,
                  int64_t nlines, int N, int NS, const cpx<T>* __restrict__ tw, int lines_fast) {
  const int64_t NB = N / R;
  const int64_t total = nlines * NB;
  const int64_t tstep = N / ((int64_t)NS * R);
  for (int64_t b = (int64_t)blockIdx.x * GTPB + threadIdx.x; b < total;
       b += (int64_t)gridDim.x * GTPB) {
    int64_t line, j;
    if (lines_fast) {
      j = b / nlines;
      line = b - j * nlines;
    } else {
      line = b / NB;
      j = b - line * NB;
    }
    const int64_t k = j % NS;
    const int64_t bs = gs.base(line), bd = gd.base(line);
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cpx<T> x = src[bs + (j + r * NB) * gs.S];
      if (r > 0 && NS > 1) {
        const cpx<T> w = tw[(r * k * tstep) % N];
        x = (SIGN < 0) ? cmul(x, w) : cmulc(x, w);
      }
      v[r] = x;
    }
    dft<T, R, SIGN>(v);
    const int64_t o = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[bd + (o + r * NS) * gd.S] = v[r];
  }
}

// Any radix R (a prime above 16): one thread per output q of a butterfly.
template <typename T, int SIGN>
__global__ void __launch_bounds__(GTPB)
    gstage_prime_kernel(const cpx<T>* __restrict__ src, GGeom gs, cpx<T>* __restrict__ dst,
                        GGeom gd, int64_t nlines, int N, int R, int NS,
                        const cpx<T>* __restrict__ tw, int lines_fast) {
  const int64_t NB = N / R;
  const int64_t total = nlines * NB * R;
  const int64_t tstep = N / ((int64_t)NS * R);
  for (int64_t b = (int64_t)blockIdx.x * GTPB + threadIdx.x; b < total;
       b += (int64_t)gridDim.x * GTPB) {
    int64_t line, j, q;
    if (lines_fast) {
      line = b % nlines;
      const int64_t u = b / nlines;
      q = u % R;
      j = u / R;
    } else {
      line = b / (NB * R);
      const int64_t u = b - line * NB * R;
      j = u / R;
      q = u - j * R;
    }
    const int64_t k = j % NS;
    const int64_t bs = gs.base(line), bd = gd.base(line);
    cpx<T> acc = {(T)0, (T)0};
    for (int r = 0; r < R; ++r) {
      // twiddle of the stage times the DFT kernel exp(-+2 pi i r q / R)
      const int64_t m = ((int64_t)r * k * tstep + ((int64_t)r * q % R) * NB) % N;
      const cpx<T> w = tw[m];
      const cpx<T> x = src[bs + (j + (int64_t)r * NB) * gs.S];
      const cpx<T> t = (SIGN < 0) ? cmul(x, w) : cmulc(x, w);
      acc = cadd(acc, t);
    }
    dst[bd + ((j - k) * R + k + q * NS) * gd.S] = acc;
  }
}

template <typename T>
__global__ void __launch_bounds__(GTPB)
    gcopy_kernel(const cpx<T>* __restrict__ src, GGeom gs, cpx<T>* __restrict__ dst, GGeom gd,
                 int64_t nlines, int N, int lines_fast) {
  const int64_t total = nlines * N;
  for (int64_t b = (int64_t)blockIdx.x * GTPB + threadIdx.x; b < total;
       b += (int64_t)gridDim.x * GTPB) {
    const int64_t line = lines_fast ? b % nlines : b / N;
    const int64_t n = lines_fast ? b / nlines : b % N;
    dst[gd.base(line) + n * gd.S] = src[gs.base(line) + n * gs.S];
  }
}

// Two real lines (2 m, 2 m + 1; pitch ip) -> one complex line m (pitch N).
template <typename T>
__global__ void __launch_bounds__(GTPB)
    gpack_r2c_kernel(const T* __restrict__ in, int64_t ip, cpx<T>* __restrict__ z, int64_t nlines,
                     int N) {
  const int64_t np = (nlines + 1) / 2;
  const int64_t total = np * N;
  for (int64_t b = (int64_t)blockIdx.x * GTPB + threadIdx.x; b < total;
       b += (int64_t)gridDim.x * GTPB) {
    const int64_t m = b / N, n = b - m * N;
    const T a = in[2 * m * ip + n];
    const T c = (2 * m + 1 < nlines) ? in[(2 * m + 1) * ip + n] : (T)0;
    z[b] = cpx<T>{a, c};
  }
}

// Complex line m (pitch N) -> half spectra of lines 2 m, 2 m + 1 (pitch op).
template <typename T>
__global__ void __launch_bounds__(GTPB)
    gsplit_r2c_kernel(const cpx<T>* __restrict__ z, cpx<T>* __restrict__ out, int64_t op,
                      int64_t nlines, int N) {
  const int64_t nh = N / 2 + 1, np = (nlines + 1) / 2;
  const int64_t total = np * nh;
  const T h = (T)0.5;
  for (int64_t b = (int64_t)blockIdx.x * GTPB + threadIdx.x; b < total;
       b += (int64_t)gridDim.x * GTPB) {
    const int64_t m = b / nh, k = b - m * nh;
    const cpx<T> z1 = z[m * N + k];
    const cpx<T> z2 = z[m * N + (k == 0 ? 0 : N - k)];
    out[2 * m * op + k] = cpx<T>{h * (z1.x + z2.x), h * (z1.y - z2.y)};
    if (2 * m + 1 < nlines) out[(2 * m + 1) * op + k] = cpx<T>{h * (z1.y + z2.y), h * (z2.x - z1.x)};
  }
}

// Half spectra of lines 2 m, 2 m + 1 (pitch ip) -> complex line m (pitch N)
// whose inverse transform is a + i b.
template <typename T>
__global__ void __launch_bounds__(GTPB)
    gpack_c2r_kernel(const cpx<T>* __restrict__ in, int64_t ip, cpx<T>* __restrict__ z,
                     int64_t nlines, int N) {
  const int64_t nh = N / 2 + 1, np = (nlines + 1) / 2;
  const int64_t total = np * nh;
  for (int64_t b = (int64_t)blockIdx.x * GTPB + threadIdx.x; b < total;
       b += (int64_t)gridDim.x * GTPB) {
    const int64_t m = b / nh, k = b - m * nh;
    cpx<T> A = in[2 * m * ip + k];
    cpx<T> B = (2 * m + 1 < nlines) ? in[(2 * m + 1) * ip + k] : cpx<T>{(T)0, (T)0};
    const bool edge = k == 0 || 2 * k == N;
    if (edge) {
      A.y = (T)0;
      B.y = (T)0;
    }
    z[m * N + k] = cpx<T>{A.x - B.y, A.y + B.x};
    if (!edge) z[m * N + N - k] = cpx<T>{A.x + B.y, B.x - A.y};
  }
}

// Complex line m (pitch N) -> real lines 2 m, 2 m + 1 (pitch op).
template <typename T>
__global__ void __launch_bounds__(GTPB)
    gunpack_c2r_kernel(const cpx<T>* __restrict__ z, T* __restrict__ out, int64_t op,
                       int64_t nlines, int N) {
  const int64_t np = (nlines + 1) / 2;
  const int64_t total = np * N;
  for (int64_t b = (int64_t)blockIdx.x * GTPB + threadIdx.x; b < total;
       b += (int64_t)gridDim.x * GTPB) {
    const int64_t m = b / N, n = b - m * N;
    const cpx<T> v = z[b];
    out[2 * m * op + n] = v.x;
    if (2 * m + 1 < nlines) out[(2 * m + 1) * op + n] = v.y;
  }
}

// Bluestein's algorithm for lines whose length N has a prime factor above
// 61 (Bluestein 1970): with nk = (n^2 + k^2 - (k - n)^2) / 2,
//   X[k] = c[k] sum_n (x[n] c[n]) h[k - n],  c[n] = exp(s i pi n^2 / N),
//   h[m] = conj c[m],
// a circular convolution of power-of-two length M >= 2N - 1: a = x c
// zero-padded to M, a = IFFT_M(FFT_M(a) H), X = c a on k < N, where H is
// the plan's FFT_M of h (circularly extended) divided by M.  chirp holds
// exp(-i pi n^2 / N); the inverse transform (s = +1) uses its conjugate.
template <typename T>
__global__ void __launch_bounds__(GTPB)
    blue_pre_kernel(const cpx<T>* __restrict__ src, GGeom gs, cpx<T>* __restrict__ a,
                    int64_t line0, int64_t nl, int N, int M, const cpx<T>* __restrict__ chirp,
                    int sign) {
  const int64_t total = nl * M;
  for (int64_t b = (int64_t)blockIdx.x * GTPB + threadIdx.x; b < total;
       b += (int64_t)gridDim.x * GTPB) {
    const int64_t l = b / M, n = b - l * M;
    cpx<T> v = {(T)0, (T)0};
    if (n < N) {
      const cpx<T> x = src[gs.base(line0 + l) + n * gs.S];
      v = sign < 0 ? cmul(x, chirp[n]) : cmulc(x, chirp[n]);
    }
    a[b] = v;
  }
}

template <typename T>
__global__ void __launch_bounds__(GTPB)
    blue_mul_kernel(cpx<T>* __restrict__ a, int64_t nl, int M, const cpx<T>* __restrict__ H) {
  const int64_t total = nl * M;
  for (int64_t b = (int64_t)blockIdx.x * GTPB + threadIdx.x; b < total;
       b += (int64_t)gridDim.x * GTPB)
    a[b] = cmul(a[b], H[b % M]);
}

template <typename T>
__global__ void __launch_bounds__(GTPB)
    blue_post_kernel(const cpx<T>* __restrict__ a, cpx<T>* __restrict__ dst, GGeom gd,
                     int64_t line0, int64_t nl, int N, int M, const cpx<T>* __restrict__ chirp,
                     int sign) {
  const int64_t total = nl * N;
  for (int64_t b = (int64_t)blockIdx.x * GTPB + threadIdx.x; b < total;
       b += (int64_t)gridDim.x * GTPB) {
    const int64_t l = b / N, k = b - l * N;
    const cpx<T> v = a[l * M + k];
    dst[gd.base(line0 + l) + k * gd.S] = sign < 0 ? cmul(v, chirp[k]) : cmulc(v, chirp[k]);
  }
}

// Whether a length runs Bluestein's algorithm: a prime factor above 61.
inline bool bluestein_length(int n) {
  int big = 1, m = n;
  for (int f = 2; (long)f * f <= m; ++f)
    while (m % f == 0) {
      big = f;
      m /= f;
    }
  if (m > 1) big = m;  // the remaining factor is a prime above every one divided out
  return big > 61;
}

inline unsigned gblocks(int64_t work) {
  int64_t b = (work + GTPB - 1) / GTPB;
  if (b > 65536) b = 65536;
  return (unsigned)(b > 0 ? b : 1);
}

template <typename T, int SIGN>
hipError_t launch_gstage(const cpx<T>* src, GGeom gs, cpx<T>* dst, GGeom gd, int64_t nlines,
                         int N, int R, int NS, const cpx<T>* tw, int lines_fast, hipStream_t st) {
  const dim3 b(GTPB);
  const dim3 g(gblocks(nlines * (N / R)));
  switch (R) {
#define AFFT_GST(RR)                                                                            \
  case RR:                                                                                      \
    hipLaunchKernelGGL((gstage_kernel<T, RR, SIGN>), g, b, 0, st, src, gs, dst, gd, nlines, N, \
                       NS, tw, lines_fast);                                                     \
    break;
    AFFT_GST(2)
    AFFT_GST(3)
    AFFT_GST(4)
    AFFT_GST(5)
    AFFT_GST(7)
    AFFT_GST(8)
    AFFT_GST(9)
    AFFT_GST(16)
#undef AFFT_GST
    default:
      hipLaunchKernelGGL((gstage_prime_kernel<T, SIGN>), dim3(gblocks(nlines * (int64_t)N)), b, 0,
                         st, src, gs, dst, gd, nlines, N, R, NS, tw, lines_fast);
  }
  return hipGetLastError();
}

// Complex transform of nlines lines of length N from (src, gs) to (dst, gd)
// through the workspace ws (laid out as gd).  src may equal dst.
template <typename T>
hipError_t long_c2c(const cpx<T>* src, GGeom gs, cpx<T>* dst, GGeom gd, cpx<T>* ws,
                    int64_t nlines, const DynSched& sc, const cpx<T>* tw, int sign, int lines_fast,
                    hipStream_t st) {
  const int S = sc.nst, N = sc.n;
  if (S == 0) return hipSuccess;
  const bool inplace = src == dst;
  const cpx<T>* cur = src;
  GGeom gc = gs;
  int NS = 1;
  for (int s = 0; s < S; ++s) {
    // the last stage writes dst (out of place); in place, even stages write ws
    cpx<T>* to = inplace ? ((s % 2 == 0) ? ws : dst) : (((S - 1 - s) % 2 == 0) ? dst : ws);
    hipError_t e = sign < 0 ? launch_gstage<T, -1>(cur, gc, to, gd, nlines, N, sc.rad[s], NS, tw,
                                                   lines_fast, st)
                            : launch_gstage<T, +1>(cur, gc, to, gd, nlines, N, sc.rad[s], NS, tw,
                                                   lines_fast, st);
    if (e != hipSuccess) return e;
    cur = to;
    gc = gd;
    NS *= sc.rad[s];
  }
  if (cur != dst) {
    hipLaunchKernelGGL((gcopy_kernel<T>), dim3(gblocks(nlines * (int64_t)N)), dim3(GTPB), 0, st,
                       cur, gd, dst, gd, nlines, N, lines_fast);
    return hipGetLastError();
  }
  return hipSuccess;
}

}  // namespace afft

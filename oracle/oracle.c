/*
 * oracle.c -- CPU FP64 restatement of AccFFT's single-rank GPU hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library in accfft_amd/: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product links it.
 *
 * What it restates (reference file:line):
 *   - sizes:       src/accfft_common.cpp:49-91   (dfft_get_local_size_t)
 *                  src/accfft_common.cpp:98-192  (accfft_local_size_dft_r2c_t)
 *                  src/accfft_common.cpp:203-256 (accfft_local_size_dft_c2c_t)
 *   - 3-D r2c/c2r: src/accfft_gpu.cpp:444-623 stage order on a 1x1 grid:
 *                  forward = z r2c (fplan_0, :477-489), y c2c (fplan_1,
 *                  :499-516), x c2c (fplan_2, :527-539); inverse = x c2c,
 *                  y c2c, z c2r (:540-603).  Unnormalised, sign -1 forward.
 *   - 3-D c2c:     src/accfft_gpu.cpp:943-1136 (same order, complex z).
 *   - per axis:    src/accfft_gpu.cpp:1212-1604 (z: [N0][N1][Nh];
 *                  y: [N0][N1/2+1][N2]; x: [N0/2+1][N1][N2]).
 *   - multipliers: src/operators_cuda.cu:15-205 (grad x/y/z, laplace,
 *                  biharmonic) and src/operators.txx:276-392 (inverse
 *                  laplace, inverse biharmonic).
 *   - operators:   src/operators_gpu.txx:72-609 (grad fast/slow, laplace,
 *                  divergence fast/slow, biharmonic) and
 *                  src/operators.txx:1270-1321 (inverse laplace).
 *   - transposes:  src/transpose.cpp:664-769 and src/transpose_cuda.cu:6-75.
 *
 * The 1-D transform is a mixed-radix Stockham DFT with radix-4/2 fast paths
 * and an O(p^2) butterfly for every other prime factor p; twiddles come from
 * a table of exp(-2*pi*i*m/n) evaluated in long double and rounded once.
 * The c2r of a half spectrum uses only Re(X[0]) and, for even n,
 * Re(X[n/2]), which is what a Hermitian-input c2r computes.
 *
 * Parity pins for this file live in tests/test_oracle.py: numpy.fft golden
 * fixtures, brute-force O(n^2) DFTs, closed forms and invariants.
 */
#include <math.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* 1-D complex DFT plan                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
  int n;
  int nf;
  int fac[64];
  double *tw; /* 2*n doubles: exp(-2 pi i m / n), m in [0, n) */
} orc_plan1d;

static void plan1d_init(orc_plan1d *p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  for (int f = 3; (long)f * f <= m; f += 2)
    while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
  if (m > 1) p->fac[p->nf++] = m;
  p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < n; ++k) {
    /* exact octant-free evaluation in extended precision; rounding once */
    long double a = twopi * (long double)k / (long double)n;
    p->tw[2 * k] = (double)cosl(a);
    p->tw[2 * k + 1] = (double)(-sinl(a));
  }
  /* exact values at the points where they are known */
  for (int k = 0; k < n; ++k) {
    if ((4L * k) % n == 0) {
      int q = (int)((4L * k) / n);
      static const double c[4] = {1, 0, -1, 0}, s[4] = {0, -1, 0, 1};
      p->tw[2 * k] = c[q];
      p->tw[2 * k + 1] = s[q];
    }
  }
}

static void plan1d_free(orc_plan1d *p) { free(p->tw); p->tw = NULL; }

/* One unnormalised DFT of length p->n.  sign = -1 forward, +1 backward.
 * x: input (complex interleaved), y: output, w: scratch of 2n doubles.
 * Stockham autosort: stage with radix R and span Ns reads x[j + r*n/R]
 * and writes y[(j/Ns)*Ns*R + j%Ns + r*Ns]. */
static void dft1d(const orc_plan1d *p, const double *xin, double *yout,
                  double *w, int sign) {
  const int n = p->n;
  if (n == 1) { yout[0] = xin[0]; yout[1] = xin[1]; return; }
  double *a = w, *b = yout;
  /* choose the buffers so that the final stage lands in yout */
  if (p->nf % 2 == 1) { a = yout; b = w; }
  memcpy(a, xin, sizeof(double) * 2 * n);
  int Ns = 1;
  for (int s = 0; s < p->nf; ++s) {
    const int R = p->fac[s];
    const int stride = n / R;
    const int tstep = n / (Ns * R); /* twiddle index step */
    for (int j = 0; j < stride; ++j) {
      const int k = j % Ns;
      double vr[64], vi[64];
      double *vrp = vr, *vip = vi;
      double *heapr = NULL;
      if (R > 64) {
        heapr = (double *)malloc(sizeof(double) * 2 * R);
        vrp = heapr; vip = heapr + R;
      }
      for (int r = 0; r < R; ++r) {
        double xr = a[2 * (j + r * stride)], xi = a[2 * (j + r * stride) + 1];
        long m = ((long)r * k * tstep) % n;
        double c = p->tw[2 * m], sn = p->tw[2 * m + 1] * (sign < 0 ? 1.0 : -1.0);
        vrp[r] = xr * c - xi * sn;
        vip[r] = xr * sn + xi * c;
      }
      const int base = (j / Ns) * Ns * R + k;
      if (R == 2) {
        double ar = vrp[0], ai = vip[0], br = vrp[1], bi = vip[1];
        b[2 * base] = ar + br; b[2 * base + 1] = ai + bi;
        b[2 * (base + Ns)] = ar - br; b[2 * (base + Ns) + 1] = ai - bi;
      } else if (R == 4) {
        double s0r = vrp[0] + vrp[2], s0i = vip[0] + vip[2];
        double d0r = vrp[0] - vrp[2], d0i = vip[0] - vip[2];
        double s1r = vrp[1] + vrp[3], s1i = vip[1] + vip[3];
        double d1r = vrp[1] - vrp[3], d1i = vip[1] - vip[3];
        /* multiply d1 by -i (forward) or +i (backward) */
        double tr = sign < 0 ? d1i : -d1i, ti = sign < 0 ? -d1r : d1r;
        b[2 * base] = s0r + s1r; b[2 * base + 1] = s0i + s1i;
        b[2 * (base + Ns)] = d0r + tr; b[2 * (base + Ns) + 1] = d0i + ti;
        b[2 * (base + 2 * Ns)] = s0r - s1r; b[2 * (base + 2 * Ns) + 1] = s0i - s1i;
        b[2 * (base + 3 * Ns)] = d0r - tr; b[2 * (base + 3 * Ns) + 1] = d0i - ti;
      } else {
        /* generic O(R^2) butterfly using the length-n table */
        const int rstep = n / R;
        for (int q = 0; q < R; ++q) {
          double sr = 0, si = 0;
          for (int r = 0; r < R; ++r) {
            long m = ((long)r * q % R) * rstep;
            double c = p->tw[2 * m], sn = p->tw[2 * m + 1] * (sign < 0 ? 1.0 : -1.0);
            sr += vrp[r] * c - vip[r] * sn;
            si += vrp[r] * sn + vip[r] * c;
          }
          b[2 * (base + q * Ns)] = sr;
          b[2 * (base + q * Ns) + 1] = si;
        }
      }
      free(heapr);
    }
    double *t = a; a = b; b = t;
    Ns *= R;
  }
  /* after nf swaps the result is in a */
  if (a != yout) memcpy(yout, a, sizeof(double) * 2 * n);
}

/* Strided batched complex transform along one axis of a complex array.
 * Element (line l, position t) lives at base(l) + t*stride, where lines are
 * enumerated as l = o*inner + c with base = o*odist + c (complex units). */
static void c2c_axis(double *data, int n, long stride, long outer, long odist,
                     long inner, int sign) {
  orc_plan1d p;
  plan1d_init(&p, n);
  long total = outer * inner;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 6 * (size_t)n);
    double *tmp = buf + 2 * n, *w = buf + 4 * n;
#pragma omp for schedule(static)
    for (long l = 0; l < total; ++l) {
      long o = l / inner, c = l % inner;
      double *base = data + 2 * (o * odist + c);
      for (int t = 0; t < n; ++t) {
        buf[2 * t] = base[2 * t * stride];
        buf[2 * t + 1] = base[2 * t * stride + 1];
      }
      dft1d(&p, buf, tmp, w, sign);
      for (int t = 0; t < n; ++t) {
        base[2 * t * stride] = tmp[2 * t];
        base[2 * t * stride + 1] = tmp[2 * t + 1];
      }
    }
    free(buf);
  }
  plan1d_free(&p);
}

/* Real-to-half-complex along one axis.  in: real, element t of line l at
 * in[ib(l) + t*istride]; out: complex (interleaved), element k (k < n/2+1)
 * at out[ob(l) + k*ostride] (complex units).  Lines: l = o*inner + c,
 * ib = o*iodist + c, ob = o*oodist + c. */
static void r2c_axis(const double *in, double *out, int n, long istride,
                     long ostride, long outer, long iodist, long oodist,
                     long inner) {
  orc_plan1d p;
  plan1d_init(&p, n);
  const int nh = n / 2 + 1;
  long total = outer * inner;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 6 * (size_t)n);
    double *tmp = buf + 2 * n, *w = buf + 4 * n;
#pragma omp for schedule(static)
    for (long l = 0; l < total; ++l) {
      long o = l / inner, c = l % inner;
      const double *ib = in + o * iodist + c;
      double *ob = out + 2 * (o * oodist + c);
      for (int t = 0; t < n; ++t) { buf[2 * t] = ib[t * istride]; buf[2 * t + 1] = 0; }
      dft1d(&p, buf, tmp, w, -1);
      for (int k = 0; k < nh; ++k) {
        ob[2 * k * ostride] = tmp[2 * k];
        ob[2 * k * ostride + 1] = tmp[2 * k + 1];
      }
    }
    free(buf);
  }
  plan1d_free(&p);
}

/* Half-complex-to-real along one axis (inverse, unnormalised). */
static void c2r_axis(const double *in, double *out, int n, long istride,
                     long ostride, long outer, long iodist, long oodist,
                     long inner) {
  orc_plan1d p;
  plan1d_init(&p, n);
  const int nh = n / 2 + 1;
  long total = outer * inner;
#pragma omp parallel
  {
    double *buf = (double *)malloc(sizeof(double) * 6 * (size_t)n);
    double *tmp = buf + 2 * n, *w = buf + 4 * n;
#pragma omp for schedule(static)
    for (long l = 0; l < total; ++l) {
      long o = l / inner, c = l % inner;
      const double *ib = in + 2 * (o * iodist + c);
      double *ob = out + o * oodist + c;
      for (int k = 0; k < nh; ++k) {
        buf[2 * k] = ib[2 * k * istride];
        buf[2 * k + 1] = ib[2 * k * istride + 1];
      }
      /* Hermitian input: DC (and Nyquist for even n) are real */
      buf[1] = 0;
      if (n % 2 == 0) buf[2 * (n / 2) + 1] = 0;
      for (int k = nh; k < n; ++k) {
        buf[2 * k] = buf[2 * (n - k)];
        buf[2 * k + 1] = -buf[2 * (n - k) + 1];
      }
      dft1d(&p, buf, tmp, w, +1);
      for (int t = 0; t < n; ++t) ob[t * ostride] = tmp[2 * t];
    }
    free(buf);
  }
  plan1d_free(&p);
}

/* ------------------------------------------------------------------ */
/* exported: threads, 1-D DFT                                          */
/* ------------------------------------------------------------------ */
ORC_API void orc_set_threads(int nt) {
#ifdef _OPENMP
  if (nt > 0) omp_set_num_threads(nt);
#else
  (void)nt;
#endif
}

ORC_API int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Unnormalised 1-D complex DFT; sign -1 forward (ACCFFT_FORWARD), +1 back. */
ORC_API void orc_dft1d(int n, const double *in, double *out, int sign) {
  orc_plan1d p;
  plan1d_init(&p, n);
  double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  dft1d(&p, in, out, w, sign);
  free(w);
  plan1d_free(&p);
}

/* Brute-force O(n^2) DFT (independent of the table and the factorisation),
 * used to pin dft1d itself. */
ORC_API void orc_dft1d_naive(int n, const double *in, double *out, int sign) {
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < n; ++k) {
    long double sr = 0, si = 0;
    for (int t = 0; t < n; ++t) {
      long m = ((long)k * t) % n;
      long double a = sign * twopi * (long double)m / (long double)n;
      long double c = cosl(a), s = sinl(a);
      sr += in[2 * t] * c - in[2 * t + 1] * s;
      si += in[2 * t] * s + in[2 * t + 1] * c;
    }
    out[2 * k] = (double)sr;
    out[2 * k + 1] = (double)si;
  }
}

/* ------------------------------------------------------------------ */
/* 3-D transforms on a 1x1 grid (src/accfft_gpu.cpp:444-623, 943-1136) */
/* ------------------------------------------------------------------ */
/* xyz bit d set <=> axis d transformed (bit 0 = x, bit 2 = z), as the
 * std::bitset<3> argument of accfft_execute_gpu. */
ORC_API void orc_r2c_3d(const int *n, const double *in, double *out, int xyz) {
  const long N0 = n[0], N1 = n[1], N2 = n[2], Nh = N2 / 2 + 1;
  if (xyz & 4) {
    r2c_axis(in, out, (int)N2, 1, 1, N0 * N1, N2, Nh, 1);
  } else {
    /* z skipped: the reference runs y and x in place on the input buffer
     * reinterpreted as complex [N0][N1][Nh] (accfft_gpu.cpp:488-489) */
    memcpy(out, in, sizeof(double) * 2 * N0 * N1 * Nh);
  }
  if (xyz & 2) c2c_axis(out, (int)N1, Nh, N0, N1 * Nh, Nh, -1);
  if (xyz & 1) c2c_axis(out, (int)N0, N1 * Nh, 1, 0, N1 * Nh, -1);
}

/* Inverse: x c2c, y c2c, z c2r.  The input is not modified (the reference
 * overwrites it; the values written to out are the same). */
ORC_API void orc_c2r_3d(const int *n, const double *in, double *out, int xyz) {
  const long N0 = n[0], N1 = n[1], N2 = n[2], Nh = N2 / 2 + 1;
  size_t csz = sizeof(double) * 2 * N0 * N1 * Nh;
  double *w = (double *)malloc(csz);
  memcpy(w, in, csz);
  if (xyz & 1) c2c_axis(w, (int)N0, N1 * Nh, 1, 0, N1 * Nh, +1);
  if (xyz & 2) c2c_axis(w, (int)N1, Nh, N0, N1 * Nh, Nh, +1);
  if (xyz & 4)
    c2r_axis(w, out, (int)N2, 1, 1, N0 * N1, Nh, N2, 1);
  else
    memcpy(out, w, csz);
  free(w);
}

ORC_API void orc_c2c_3d(const int *n, int dir, const double *in, double *out,
                        int xyz) {
  const long N0 = n[0], N1 = n[1], N2 = n[2];
  memcpy(out, in, sizeof(double) * 2 * N0 * N1 * N2);
  if (dir < 0) {
    if (xyz & 4) c2c_axis(out, (int)N2, 1, N0 * N1, N2, 1, -1);
    if (xyz & 2) c2c_axis(out, (int)N1, N2, N0, N1 * N2, N2, -1);
    if (xyz & 1) c2c_axis(out, (int)N0, N1 * N2, 1, 0, N1 * N2, -1);
  } else {
    if (xyz & 1) c2c_axis(out, (int)N0, N1 * N2, 1, 0, N1 * N2, +1);
    if (xyz & 2) c2c_axis(out, (int)N1, N2, N0, N1 * N2, N2, +1);
    if (xyz & 4) c2c_axis(out, (int)N2, 1, N0 * N1, N2, 1, +1);
  }
}

/* Per-axis transforms (src/accfft_gpu.cpp:1212-1604).
 * axis 2 (z): real [N0][N1][N2] <-> complex [N0][N1][N2/2+1]
 * axis 1 (y): real [N0][N1][N2] <-> complex [N0][N1/2+1][N2]
 * axis 0 (x): real [N0][N1][N2] <-> complex [N0/2+1][N1][N2]            */
ORC_API void orc_r2c_axis(const int *n, int axis, const double *in,
                          double *out) {
  const long N0 = n[0], N1 = n[1], N2 = n[2];
  if (axis == 2) r2c_axis(in, out, (int)N2, 1, 1, N0 * N1, N2, N2 / 2 + 1, 1);
  else if (axis == 1)
    r2c_axis(in, out, (int)N1, N2, N2, N0, N1 * N2, (N1 / 2 + 1) * N2, N2);
  else
    r2c_axis(in, out, (int)N0, N1 * N2, N1 * N2, 1, 0, 0, N1 * N2);
}

ORC_API void orc_c2r_axis(const int *n, int axis, const double *in,
                          double *out) {
  const long N0 = n[0], N1 = n[1], N2 = n[2];
  if (axis == 2) c2r_axis(in, out, (int)N2, 1, 1, N0 * N1, N2 / 2 + 1, N2, 1);
  else if (axis == 1)
    c2r_axis(in, out, (int)N1, N2, N2, N0, (N1 / 2 + 1) * N2, N1 * N2, N2);
  else
    c2r_axis(in, out, (int)N0, N1 * N2, N1 * N2, 1, 0, 0, N1 * N2);
}

/* ------------------------------------------------------------------ */
/* Spectral multipliers                                                 */
/* ------------------------------------------------------------------ */
/* Index -> wavenumber with the reference's Nyquist rule
 * (src/operators_cuda.cu:32-37): w = X; if X > N/2, w -= N; if X == N/2,
 * w = 0 (integer N/2, so for odd N the mode (N-1)/2 is zeroed). */
static long wave_nyq0(long X, long N) {
  long w = X;
  if (X > N / 2) w -= N;
  if (X == N / 2) w = 0;
  return w;
}
/* Same without the Nyquist filter (src/operators.txx:360-370,
 * mult_wave_number_inv_biharmonic). */
static long wave_plain(long X, long N) {
  long w = X;
  if (X > N / 2) w -= N;
  return w;
}

enum {
  ORC_GRAD_X = 0,
  ORC_GRAD_Y = 1,
  ORC_GRAD_Z = 2,
  ORC_LAPLACE = 3,
  ORC_BIHARMONIC = 4,
  ORC_INV_LAPLACE = 5,
  ORC_INV_BIHARMONIC = 6,
  ORC_POISSON_DISCRETE = 7,
  ORC_POISSON_CONTINUOUS = 8
};

/* out = symbol(kind) * scale * in over a complex block of osize with global
 * offsets ostart (src/operators_cuda.cu:15-205; src/operators.txx:276-392).
 * Products are formed in double as in the reference: grad stores
 * (-scale*w*Im, scale*w*Re); the others (scale*wave)*value. */
ORC_API void orc_mult(int kind, const int *N, const int *osize,
                      const int *ostart, double scale, const double *in,
                      double *out) {
  long total = (long)osize[0] * osize[1] * osize[2];
#pragma omp parallel for schedule(static)
  for (long ptr = 0; ptr < total; ++ptr) {
    long k = ptr % osize[2];
    long j = (ptr / osize[2]) % osize[1];
    long i = ptr / ((long)osize[1] * osize[2]);
    long X = i + ostart[0], Y = j + ostart[1], Z = k + ostart[2];
    double re = in[2 * ptr], im = in[2 * ptr + 1];
    if (kind <= ORC_GRAD_Z) {
      long w = kind == ORC_GRAD_X   ? wave_nyq0(X, N[0])
               : kind == ORC_GRAD_Y ? wave_nyq0(Y, N[1])
                                    : wave_nyq0(Z, N[2]);
      out[2 * ptr] = -scale * w * im;
      out[2 * ptr + 1] = scale * w * re;
      continue;
    }
    if (kind == ORC_POISSON_DISCRETE || kind == ORC_POISSON_CONTINUOUS) {
      /* Poisson Fourier filter, steps/step4/kernels.cu:24-110 (unit box,
       * dx = 1/NX): method 0 = 2 sum (cos(2 pi k/N) - 1)/dx^2, method 1 =
       * -4 pi^2 sum kk^2 with kk = k - N when k > N/2 (a double comparison);
       * times NX NY NZ (here 1/scale); the (0,0,0) mode is set to zero. */
      if (X == 0 && Y == 0 && Z == 0) {
        out[2 * ptr] = 0.0;
        out[2 * ptr + 1] = 0.0;
        continue;
      }
      double NX = N[0], NY = N[1], NZ = N[2];
      double dx = 1.0 / NX, dy = 1.0 / NY, dz = 1.0 / NZ;
      double sf;
      if (kind == ORC_POISSON_DISCRETE) {
        sf = 2 * ((cos(2 * M_PI * X / NX) - 1) / (dx * dx) +
                  (cos(2 * M_PI * Y / NY) - 1) / (dy * dy) +
                  (cos(2 * M_PI * Z / NZ) - 1) / (dz * dz));
      } else {
        double kx = X, ky = Y, kz = Z;
        if (kx > NX / 2) kx -= NX;
        if (ky > NY / 2) ky -= NY;
        if (kz > NZ / 2) kz -= NZ;
        sf = -4 * M_PI * M_PI * (kx * kx + ky * ky + kz * kz);
      }
      sf /= scale;
      out[2 * ptr] = re / sf;
      out[2 * ptr + 1] = im / sf;
      continue;
    }
    double f;
    if (kind == ORC_INV_BIHARMONIC) {
      long wx = wave_plain(X, N[0]), wy = wave_plain(Y, N[1]),
           wz = wave_plain(Z, N[2]);
      double wave = (double)(-wx * wx - wy * wy - wz * wz);
      wave *= wave;
      if (wave == 0) wave = 1;
      f = 1.0 / wave;
    } else {
      long wx = wave_nyq0(X, N[0]), wy = wave_nyq0(Y, N[1]),
           wz = wave_nyq0(Z, N[2]);
      long wave = -wx * wx - wy * wy - wz * wz;
      if (kind == ORC_LAPLACE) f = (double)wave;
      else if (kind == ORC_BIHARMONIC) f = (double)(wave * wave);
      else f = wave == 0 ? 0.0 : 1.0 / (double)wave;
    }
    out[2 * ptr] = scale * f * re;
    out[2 * ptr + 1] = scale * f * im;
  }
}

/* ------------------------------------------------------------------ */
/* Operators (src/operators_gpu.txx)                                    */
/* ------------------------------------------------------------------ */
/* accfft_grad_gpu_t (:172-278): per axis d, per-axis r2c along d,
 * multiply by i*w_d / N_d, per-axis c2r along d.  Null outputs and clear
 * mask bits are skipped. */
ORC_API void orc_grad(const int *n, const double *A, double *Ax, double *Ay,
                      double *Az, int mask) {
  const long N0 = n[0], N1 = n[1], N2 = n[2];
  double *outs[3] = {Ax, Ay, Az};
  for (int d = 0; d < 3; ++d) {
    if (!(mask & (1 << d)) || !outs[d]) continue;
    int osz[3] = {(int)N0, (int)N1, (int)N2};
    osz[d] = n[d] / 2 + 1;
    int ost[3] = {0, 0, 0};
    long csz = 2L * osz[0] * osz[1] * osz[2];
    double *h = (double *)malloc(sizeof(double) * csz);
    double *t = (double *)malloc(sizeof(double) * csz);
    orc_r2c_axis(n, d, A, h);
    orc_mult(ORC_GRAD_X + d, n, osz, ost, 1.0 / n[d], h, t);
    orc_c2r_axis(n, d, t, outs[d]);
    free(h);
    free(t);
  }
}

/* accfft_grad_gpu_slow_t (:72-170): one 3-D r2c, then per component a
 * multiply with scale 1/(N0 N1 N2) and a 3-D c2r. */
ORC_API void orc_grad_slow(const int *n, const double *A, double *Ax,
                           double *Ay, double *Az, int mask) {
  const long N0 = n[0], N1 = n[1], N2 = n[2], Nh = N2 / 2 + 1;
  int osz[3] = {(int)N0, (int)N1, (int)Nh}, ost[3] = {0, 0, 0};
  long csz = 2L * N0 * N1 * Nh;
  double *h = (double *)malloc(sizeof(double) * csz);
  double *t = (double *)malloc(sizeof(double) * csz);
  orc_r2c_3d(n, A, h, 7);
  double scale = 1.0 / ((double)N0 * N1 * N2);
  double *outs[3] = {Ax, Ay, Az};
  for (int d = 0; d < 3; ++d) {
    if (!(mask & (1 << d)) || !outs[d]) continue;
    orc_mult(ORC_GRAD_X + d, n, osz, ost, scale, h, t);
    orc_c2r_3d(n, t, outs[d], 7);
  }
  free(h);
  free(t);
}

/* laplace (:280-334), biharmonic (:555-609), inverse laplace
 * (src/operators.txx:1270-1321), inverse biharmonic (operators.txx:336-392
 * multiplier), the step4 Poisson solve (steps/step4/step4_gpu.cpp:520-640):
 * 3-D r2c, multiply with scale 1/(N0 N1 N2), 3-D c2r. */
ORC_API void orc_spectral_op(int kind, const int *n, const double *A,
                             double *out) {
  const long N0 = n[0], N1 = n[1], N2 = n[2], Nh = N2 / 2 + 1;
  int osz[3] = {(int)N0, (int)N1, (int)Nh}, ost[3] = {0, 0, 0};
  long csz = 2L * N0 * N1 * Nh;
  double *h = (double *)malloc(sizeof(double) * csz);
  double *t = (double *)malloc(sizeof(double) * csz);
  orc_r2c_3d(n, A, h, 7);
  orc_mult(kind, n, osz, ost, 1.0 / ((double)N0 * N1 * N2), h, t);
  orc_c2r_3d(n, t, out, 7);
  free(h);
  free(t);
}

/* accfft_divergence_gpu_t (:448-553): sum over d of the per-axis
 * derivative of component d (per-axis r2c, i*w_d/N_d, per-axis c2r). */
ORC_API void orc_divergence(const int *n, const double *Ax, const double *Ay,
                            const double *Az, double *div) {
  const long R = (long)n[0] * n[1] * n[2];
  const double *ins[3] = {Ax, Ay, Az};
  double *tmp = (double *)malloc(sizeof(double) * R);
  memset(div, 0, sizeof(double) * R);
  for (int d = 0; d < 3; ++d) {
    double *o[3] = {NULL, NULL, NULL};
    o[d] = tmp;
    orc_grad(n, ins[d], o[0], o[1], o[2], 1 << d);
    for (long i = 0; i < R; ++i) div[i] += tmp[i];
  }
  free(tmp);
}

/* accfft_divergence_gpu_slow_t (:336-446): for component d, a 3-D r2c over
 * axes {d, z}, multiply by i*w_d with scale 1/prod(N over {d, z}), c2r over
 * the same axes; summed. */
ORC_API void orc_divergence_slow(const int *n, const double *Ax,
                                 const double *Ay, const double *Az,
                                 double *div) {
  const long N0 = n[0], N1 = n[1], N2 = n[2], Nh = N2 / 2 + 1;
  const long R = N0 * N1 * N2;
  int osz[3] = {(int)N0, (int)N1, (int)Nh}, ost[3] = {0, 0, 0};
  long csz = 2L * N0 * N1 * Nh;
  double *h = (double *)malloc(sizeof(double) * csz);
  double *t = (double *)malloc(sizeof(double) * csz);
  double *tmp = (double *)malloc(sizeof(double) * R);
  const double *ins[3] = {Ax, Ay, Az};
  memset(div, 0, sizeof(double) * R);
  for (int d = 0; d < 3; ++d) {
    int xyz = (1 << d) | 4;
    double scale = 1.0;
    for (int e = 0; e < 3; ++e)
      if (xyz & (1 << e)) scale *= n[e];
    orc_r2c_3d(n, ins[d], h, xyz);
    orc_mult(ORC_GRAD_X + d, n, osz, ost, 1.0 / scale, h, t);
    orc_c2r_3d(n, t, tmp, xyz);
    for (long i = 0; i < R; ++i) div[i] += tmp[i];
  }
  free(h);
  free(t);
  free(tmp);
}

/* ------------------------------------------------------------------ */
/* Local transposes (src/transpose.cpp:664-769; transpose_cuda.cu:6-75) */
/* Byte-generic: elem is sizeof(T).                                     */
/* ------------------------------------------------------------------ */
/* [r][c][t] -> [c][r][t] (local_transpose / local_transpose_cuda1; the
 * in-place cycle-following form cuda3 computes the same permutation). */
ORC_API void orc_local_transpose(int r, int c, int t, const void *in,
                                 void *out, int elem) {
  const char *src = (const char *)in;
  char *dst = (char *)out;
  size_t ptr = 0, tb = (size_t)t * elem;
  for (int j = 0; j < c; ++j)
    for (int i = 0; i < r; ++i) {
      memcpy(dst + ptr, src + ((size_t)i * c + j) * tb, tb);
      ptr += tb;
    }
}

/* Unpack with a ragged last row: row r-1 holds tuples of t2 elements
 * (local_transpose(r,c,t,t2) / local_transpose_cuda2). */
ORC_API void orc_local_transpose_ragged(int r, int c, int t, int t2,
                                        const void *in, void *out, int elem) {
  const char *src = (const char *)in;
  char *dst = (char *)out;
  size_t ptr = 0;
  for (int j = 0; j < c; ++j)
    for (int i = 0; i < r; ++i) {
      if (i == r - 1) {
        memcpy(dst + ptr, src + ((size_t)i * c * t + (size_t)j * t2) * elem,
               (size_t)t2 * elem);
        ptr += (size_t)t2 * elem;
      } else {
        memcpy(dst + ptr, src + ((size_t)i * c + j) * t * elem,
               (size_t)t * elem);
        ptr += (size_t)t * elem;
      }
    }
}

/* Pack with a ragged last column (local_transpose_col /
 * local_transpose_cuda4). */
ORC_API void orc_local_transpose_col(int r, int c, int t, int t2,
                                     const void *in, void *out, int elem) {
  const char *src = (const char *)in;
  char *dst = (char *)out;
  size_t ptr = 0;
  size_t row = (size_t)(c - 1) * t + t2;
  for (int j = 0; j < c; ++j)
    for (int i = 0; i < r; ++i) {
      size_t len = (j == c - 1) ? (size_t)t2 : (size_t)t;
      memcpy(dst + ptr, src + ((size_t)i * row + (size_t)j * t) * elem,
             len * elem);
      ptr += len * elem;
    }
}

/* ------------------------------------------------------------------ */
/* Sizes (src/accfft_common.cpp:49-256), P0 x P1 grid, rank at coords.  */
/* ------------------------------------------------------------------ */
/* dfft_get_local_size_t: returns bytes as a 64-bit value (the reference
 * returns an int). */
static int64_t local_size(int N0, int N1, int tuple, int *isize, int *istart,
                          int p0, int p1, int c0, int c1, int elem) {
  isize[2] = tuple;
  isize[0] = (int)ceil(N0 / (double)p0);
  isize[1] = (int)ceil(N1 / (double)p1);
  istart[0] = isize[0] * c0;
  istart[1] = isize[1] * c1;
  istart[2] = 0;
  if ((N0 - isize[0] * c0) < isize[0]) {
    isize[0] = N0 - isize[0] * c0;
    isize[0] *= (int)(isize[0] > 0);
    istart[0] = N0 - isize[0];
  }
  if ((N1 - isize[1] * c1) < isize[1]) {
    isize[1] = N1 - isize[1] * c1;
    isize[1] *= (int)(isize[1] > 0);
    istart[1] = N1 - isize[1];
  }
  return (int64_t)isize[0] * isize[1] * isize[2] * elem;
}

static void swapi(int *a, int *b) { int t = *a; *a = *b; *b = t; }
static int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }

/* accfft_local_size_dft_r2c_t<T> with sizeof(T) = elem. Returns alloc_max
 * in bytes (64-bit). */
ORC_API int64_t orc_local_size_r2c(const int *n, int *isize, int *istart,
                                   int *osize, int *ostart, int p0, int p1,
                                   int c0, int c1, int elem) {
  int osize_0[3], ostart_0[3], osize_1[3], ostart_1[3], osize_2[3],
      ostart_2[3], osize_y[3], ostart_y[3], osize_yi[3], ostart_yi[3],
      osize_x[3], ostart_x[3], osize_xi[3], ostart_xi[3];
  int64_t a, amax = 0;
  int nt = (n[2] / 2 + 1) * 2;
  a = local_size(n[0], n[1], nt, osize_0, ostart_0, p0, p1, c0, c1, elem);
  amax = max64(amax, a);
  a = local_size(n[0], nt / 2, n[1], osize_1, ostart_1, p0, p1, c0, c1, elem);
  amax = max64(amax, 2 * a);
  a = local_size(n[1], nt / 2, n[0], osize_2, ostart_2, p0, p1, c0, c1, elem);
  amax = max64(amax, 2 * a);
  swapi(&osize_1[1], &osize_1[2]);
  swapi(&ostart_1[1], &ostart_1[2]);
  swapi(&ostart_2[1], &ostart_2[2]);
  swapi(&ostart_2[0], &ostart_2[1]);
  swapi(&osize_2[1], &osize_2[2]);
  swapi(&osize_2[0], &osize_2[1]);
  local_size(n[0], n[1], n[2], isize, istart, p0, p1, c0, c1, elem);
  for (int i = 0; i < 3; ++i) { osize[i] = osize_2[i]; ostart[i] = ostart_2[i]; }
  a = local_size(n[0], n[2], n[1], osize_y, ostart_y, p0, p1, c0, c1, elem);
  amax = max64(amax, a);
  a = local_size(n[0], n[2], n[1] / 2 + 1, osize_yi, ostart_yi, p0, p1, c0, c1,
                 elem);
  amax = max64(amax, 2 * a);
  local_size(isize[1] * n[2], n[2], n[0], osize_x, ostart_x, p0, p1, c0, c1,
             elem);
  osize_x[1] = osize_x[0];
  osize_x[0] = osize_x[2];
  osize_x[2] = 1;
  swapi(&osize_x[0], &osize_x[1]);
  amax = max64(amax, (int64_t)osize_x[0] * osize_x[1] * osize_x[2] * elem);
  local_size(isize[1] * n[2], n[2], n[0] / 2 + 1, osize_xi, ostart_xi, p0, p1,
             c0, c1, elem);
  osize_xi[1] = osize_xi[0];
  osize_xi[0] = osize_xi[2];
  osize_xi[2] = 1;
  swapi(&osize_xi[0], &osize_xi[1]);
  amax = max64(amax,
               2 * (int64_t)osize_xi[0] * osize_xi[1] * osize_xi[2] * elem);
  return amax;
}

/* accfft_local_size_dft_c2c_t<T>. */
ORC_API int64_t orc_local_size_c2c(const int *n, int *isize, int *istart,
                                   int *osize, int *ostart, int p0, int p1,
                                   int c0, int c1, int elem) {
  int osize_0[3], ostart_0[3], osize_1[3], ostart_1[3], osize_2[3],
      ostart_2[3];
  int64_t amax = 0;
  amax = max64(amax, local_size(n[0], n[1], n[2], osize_0, ostart_0, p0, p1,
                                c0, c1, elem));
  amax = max64(amax, local_size(n[0], n[2], n[1], osize_1, ostart_1, p0, p1,
                                c0, c1, elem));
  amax = max64(amax, local_size(n[1], n[2], n[0], osize_2, ostart_2, p0, p1,
                                c0, c1, elem));
  amax *= 2;
  swapi(&ostart_2[1], &ostart_2[2]);
  swapi(&ostart_2[0], &ostart_2[1]);
  swapi(&osize_2[1], &osize_2[2]);
  swapi(&osize_2[0], &osize_2[1]);
  local_size(n[0], n[1], n[2], isize, istart, p0, p1, c0, c1, elem);
  for (int i = 0; i < 3; ++i) { osize[i] = osize_2[i]; ostart[i] = ostart_2[i]; }
  return amax;
}

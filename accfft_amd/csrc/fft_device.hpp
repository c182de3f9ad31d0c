// fft_device.hpp -- device-side building blocks of the gfx950 FFT kernels.
//
// The reference delegates all FFT arithmetic to cuFFT (cufftExecD2Z/Z2Z/Z2D
// at src/accfft_gpu.cpp:480-593).  Here every 1-D transform is a Stockham
// autosort DFT run inside LDS by one workgroup:
//   stage with radix R and span NS (product of the earlier radices) reads
//   element j + r*N/R, applies exp(sign*2*pi*i*r*(j%NS)/(NS*R)), does an
//   R-point DFT and writes element (j/NS)*NS*R + j%NS + r*NS.
// Two engines share that recurrence:
//   * the static engine (power-of-two N, radix schedule fixed at compile
//     time) keeps each thread's butterflies in VGPRs across the barrier and
//     works in place in LDS, with the per-stage twiddles staged in LDS;
//   * the dynamic engine (any N, schedule passed at run time) ping-pongs
//     between two LDS buffers and reads twiddles from a length-N table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace afft {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

// Pointer into device (global) memory.  Loops that hide a running pointer
// from the compiler (asm volatile "+v", so that it does not keep every row
// address in registers) type it as global: a generic pointer compiles to
// flat loads and stores, which also count on the LDS counter, so every
// wait for an LDS result would wait for them as well.  (The host pass,
// which only parses device code, sees a plain pointer.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
#else
template <typename T>
using gptr = T*;
#endif
template <typename T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
  return (gptr<T>)p;
}
template <typename T>
__device__ __forceinline__ cpx<T> ld_global(gptr<const cpx<T>> p) {
  cpx<T> r;
  r = *p;
  return r;
}

template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {
  return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
// multiply by SIGN*i
template <int SIGN, typename T>
__device__ __forceinline__ cpx<T> mul_si(cpx<T> a) {
  if constexpr (SIGN > 0) return {-a.y, a.x};
  else return {a.y, -a.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }

// ------------------------------------------------------------------
// Small DFTs in registers, natural order in and out:
//   X[q] = sum_r v[r] * exp(SIGN * 2*pi*i * r*q / R)
// ------------------------------------------------------------------
template <typename T, int SIGN>
__device__ __forceinline__ void dft2(cpx<T>* v) {
  cpx<T> a = v[0];
  v[0] = cadd(a, v[1]);
  v[1] = csub(a, v[1]);
}

template <typename T, int SIGN>
__device__ __forceinline__ void dft4(cpx<T>* v) {
  cpx<T> s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
  cpx<T> s1 = cadd(v[1], v[3]), d1 = mul_si<SIGN>(csub(v[1], v[3]));
  v[0] = cadd(s0, s1);
  v[2] = csub(s0, s1);
  v[1] = cadd(d0, d1);
  v[3] = csub(d0, d1);
}

// 8 = 2 x 4 (decimation in time)
template <typename T, int SIGN>
__device__ __forceinline__ void dft8(cpx<T>* v) {
  const T c = (T)0.70710678118654752440;
  cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
  cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
  dft4<T, SIGN>(e);
  dft4<T, SIGN>(o);
  // o[k] *= W8^k, W8 = exp(SIGN*i*pi/4)
  cpx<T> o1 = {c * (o[1].x - SIGN * o[1].y), c * (o[1].y + SIGN * o[1].x)};
  cpx<T> o2 = mul_si<SIGN>(o[2]);
  cpx<T> o3 = {c * (-o[3].x - SIGN * o[3].y), c * (-o[3].y + SIGN * o[3].x)};
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// 16 = 2 x 8 (decimation in time)
template <typename T, int SIGN>
__device__ __forceinline__ void dft16(cpx<T>* v) {
  const T c1 = (T)0.92387953251128675613;  // cos(pi/8)
  const T s1 = (T)0.38268343236508977173;  // sin(pi/8)
  const T c2 = (T)0.70710678118654752440;
  cpx<T> e[8], o[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    e[r] = v[2 * r];
    o[r] = v[2 * r + 1];
  }
  dft8<T, SIGN>(e);
  dft8<T, SIGN>(o);
  // W16^k = cos(pi k/8) + SIGN*i*sin(pi k/8)
  const T cw[8] = {(T)1, c1, c2, s1, (T)0, -s1, -c2, -c1};
  const T sw[8] = {(T)0, s1, c2, c1, (T)1, c1, c2, s1};
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    cpx<T> t;
    if (k == 0) t = o[0];
    else if (k == 4) t = mul_si<SIGN>(o[4]);
    else {
      const T wr = cw[k], wi = SIGN * sw[k];
      t = {o[k].x * wr - o[k].y * wi, o[k].x * wi + o[k].y * wr};
    }
    v[k] = cadd(e[k], t);
    v[k + 8] = csub(e[k], t);
  }
}

// Odd prime R in {3, 5, 7} by the symmetric pair form:
//   X[q]   = v0 + sum_r a_r cos(2 pi r q/R) + SIGN*i sum_r b_r sin(2 pi r q/R)
//   X[R-q] = v0 + sum_r a_r cos(.)          - SIGN*i sum_r b_r sin(.)
// with a_r = v[r] + v[R-r], b_r = v[r] - v[R-r].
template <int R> struct OddConst;
template <> struct OddConst<3> {
  static __device__ __forceinline__ double c(int m) { return -0.5; }
  static __device__ __forceinline__ double s(int m) {
    return m == 1 ? 0.86602540378443864676 : -0.86602540378443864676;
  }
};
template <> struct OddConst<5> {
  static __device__ __forceinline__ double c(int m) {
    return (m == 1 || m == 4) ? 0.30901699437494742410 : -0.80901699437494742410;
  }
  static __device__ __forceinline__ double s(int m) {
    return m == 1 ? 0.95105651629515357212
         : m == 2 ? 0.58778525229247312917
         : m == 3 ? -0.58778525229247312917
                  : -0.95105651629515357212;
  }
};
template <> struct OddConst<7> {
  static __device__ __forceinline__ double c(int m) {
    return (m == 1 || m == 6) ? 0.62348980185873353053
         : (m == 2 || m == 5) ? -0.22252093395631440429
                              : -0.90096886790241912624;
  }
  static __device__ __forceinline__ double s(int m) {
    const double t[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702,
                         0.43388373911755812048, -0.43388373911755812048,
                         -0.97492791218182360702, -0.78183148246802980871};
    return t[m];
  }
};

template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft_odd(cpx<T>* v) {
  constexpr int H = (R - 1) / 2;
  cpx<T> a[H], b[H];
  cpx<T> v0 = v[0];
  cpx<T> sum = v[0];
#pragma unroll
  for (int r = 1; r <= H; ++r) {
    a[r - 1] = cadd(v[r], v[R - r]);
    b[r - 1] = csub(v[r], v[R - r]);
    sum = cadd(sum, a[r - 1]);
  }
  v[0] = sum;
#pragma unroll
  for (int q = 1; q <= H; ++q) {
    cpx<T> ca = v0, sb = {(T)0, (T)0};
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      const int m = (r * q) % R;
      const T cc = (T)OddConst<R>::c(m), ss = (T)OddConst<R>::s(m);
      ca.x += a[r - 1].x * cc;
      ca.y += a[r - 1].y * cc;
      sb.x += b[r - 1].x * ss;
      sb.y += b[r - 1].y * ss;
    }
    cpx<T> isb = mul_si<SIGN>(sb);
    v[q] = cadd(ca, isb);
    v[R - q] = csub(ca, isb);
  }
}

// Radix 9 as 3 x 3 (n = 3 n1 + n2, k = k1 + 3 k2): DFT-3 over n1 for each
// n2, twiddles W9^(n2 k1), DFT-3 over n2 for each k1.
template <typename T, int SIGN>
__device__ __forceinline__ void dft9(cpx<T>* v) {
  cpx<T> y[3][3];
#pragma unroll
  for (int n2 = 0; n2 < 3; ++n2) {
    cpx<T> t[3] = {v[n2], v[3 + n2], v[6 + n2]};
    dft_odd<T, 3, SIGN>(t);
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) y[n2][k1] = t[k1];
  }
  // W9^m = cos(2 pi m/9) + SIGN i sin(2 pi m/9), m = 1, 2, 4
  const T c1 = (T)0.76604444311897803520, s1 = (T)0.64278760968653932632;
  const T c2 = (T)0.17364817766693034885, s2 = (T)0.98480775301220805936;
  const T c4 = (T)-0.93969262078590838405, s4 = (T)0.34202014332566873304;
  auto rot = [](cpx<T> x, T c, T s) {
    const T si = SIGN * s;
    return cpx<T>{x.x * c - x.y * si, x.x * si + x.y * c};
  };
  y[1][1] = rot(y[1][1], c1, s1);
  y[1][2] = rot(y[1][2], c2, s2);
  y[2][1] = rot(y[2][1], c2, s2);
  y[2][2] = rot(y[2][2], c4, s4);
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    cpx<T> t[3] = {y[0][k1], y[1][k1], y[2][k1]};
    dft_odd<T, 3, SIGN>(t);
    v[k1] = t[0];
    v[k1 + 3] = t[1];
    v[k1 + 6] = t[2];
  }
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft(cpx<T>* v) {
  if constexpr (R == 2) dft2<T, SIGN>(v);
  else if constexpr (R == 4) dft4<T, SIGN>(v);
  else if constexpr (R == 8) dft8<T, SIGN>(v);
  else if constexpr (R == 9) dft9<T, SIGN>(v);
  else if constexpr (R == 16) dft16<T, SIGN>(v);
  else dft_odd<T, R, SIGN>(v);
}

// ------------------------------------------------------------------
// Static engine: compile-time radix schedules for power-of-two N.
// ------------------------------------------------------------------
template <int N> struct Sched;
template <> struct Sched<16> { static constexpr int S = 2; static constexpr int R[3] = {4, 4, 0}; };
template <> struct Sched<32> { static constexpr int S = 2; static constexpr int R[3] = {8, 4, 0}; };
template <> struct Sched<64> { static constexpr int S = 2; static constexpr int R[3] = {8, 8, 0}; };
template <> struct Sched<128> { static constexpr int S = 2; static constexpr int R[3] = {8, 16, 0}; };
template <> struct Sched<256> { static constexpr int S = 2; static constexpr int R[3] = {16, 16, 0}; };
template <> struct Sched<512> { static constexpr int S = 3; static constexpr int R[3] = {8, 8, 8}; };
template <> struct Sched<1024> { static constexpr int S = 3; static constexpr int R[3] = {16, 8, 8}; };
template <> struct Sched<2048> { static constexpr int S = 3; static constexpr int R[3] = {16, 16, 8}; };
template <> struct Sched<4096> { static constexpr int S = 3; static constexpr int R[3] = {16, 16, 16}; };

// Schedules used with the split-precision matrix-core stages: a radix-16
// first stage (on the matrix cores, no twiddles), then the smaller radices
// of the VALU schedule; 2048 and 4096 need more than one radix-16 stage.
// Each matrix-core stage costs more VALU work (operand splitting) than a
// VALU radix-16 stage, so the schedules keep them few.
template <int N> struct SchedSplit { static constexpr int S = 0; static constexpr int R[3] = {0, 0, 0}; };
template <> struct SchedSplit<256> { static constexpr int S = 3; static constexpr int R[3] = {16, 4, 4}; };
template <> struct SchedSplit<512> { static constexpr int S = 3; static constexpr int R[3] = {16, 8, 4}; };
template <> struct SchedSplit<1024> { static constexpr int S = 3; static constexpr int R[3] = {16, 8, 8}; };
template <> struct SchedSplit<2048> { static constexpr int S = 3; static constexpr int R[3] = {16, 16, 8}; };
template <> struct SchedSplit<4096> { static constexpr int S = 3; static constexpr int R[3] = {16, 16, 16}; };

// Schedules with more, smaller radices (more butterflies per stage, so more
// threads per line): the second level of the two-level strided tiles.
template <int N> struct SchedWide { static constexpr int S = 0; static constexpr int R[3] = {0, 0, 0}; };
template <> struct SchedWide<64> { static constexpr int S = 3; static constexpr int R[3] = {4, 4, 4}; };
template <> struct SchedWide<128> { static constexpr int S = 3; static constexpr int R[3] = {8, 4, 4}; };
template <> struct SchedWide<256> { static constexpr int S = 3; static constexpr int R[3] = {4, 8, 8}; };

// Schedule id of the static engine: 0 = Sched, SCHED_WIDE = SchedWide,
// any other value = a split-precision kind (SchedSplit).
constexpr int SCHED_WIDE = 8;
__host__ __device__ constexpr bool split_sched(int id) { return id != 0 && id != SCHED_WIDE; }

template <int N, int SPLIT>
struct SchedOf {
  static constexpr int S = SPLIT == SCHED_WIDE ? SchedWide<N>::S
                         : SPLIT ? SchedSplit<N>::S : Sched<N>::S;
  static constexpr int r(int i) {
    return SPLIT == SCHED_WIDE ? SchedWide<N>::R[i] : SPLIT ? SchedSplit<N>::R[i] : Sched<N>::R[i];
  }
};

template <int N, int SPLIT = 0>
__host__ __device__ constexpr int sched_ns(int s) {  // span before stage s
  int ns = 1;
  for (int i = 0; i < s; ++i) ns *= SchedOf<N, SPLIT>::r(i);
  return ns;
}
template <int N, int SPLIT = 0>
__host__ __device__ constexpr int sched_twoff(int s) {  // twiddle table offset of stage s
  int off = 0;
  for (int i = 1; i < s; ++i) off += (SchedOf<N, SPLIT>::r(i) - 1) * sched_ns<N, SPLIT>(i);
  return off;
}
template <int N, int SPLIT = 0>
__host__ __device__ constexpr int sched_twn() { return sched_twoff<N, SPLIT>(SchedOf<N, SPLIT>::S); }
template <int N, int SPLIT = 0>
__host__ __device__ constexpr int sched_minnb() {  // fewest butterflies in any stage
  int m = N;
  for (int i = 0; i < SchedOf<N, SPLIT>::S; ++i)
    if (N / SchedOf<N, SPLIT>::r(i) < m) m = N / SchedOf<N, SPLIT>::r(i);
  return m;
}

// LDS index of element e of a tile whose lines are padded by one element
// after every 16 (lines of 16k elements; line p starts at 17k p).  A
// Stockham stage writes with a lane stride of R elements (R = 8, 16): padded,
// the lanes of a 32-lane (8-byte elements) or 16-lane (16-byte elements) LDS
// access fall on distinct banks instead of 2 to 16 of them.
__device__ __forceinline__ int lds_pad16(int e) { return e + (e >> 4); }
// Element n of a line of 8-byte elements, swizzled instead of padded: the
// low five bits (a 256-byte chunk, the 64 banks once) XOR 5 x the chunk
// number.  A permutation of every aligned 32 elements, so it costs no LDS;
// the stride-R stage writes of the Stockham engine, which fall on the same
// banks in a plain line, spread over all of them (a CPU model of the
// 512-2048-point schedules' stage accesses: 1.04-1.08x the conflict-free
// cycles, against 1.7-2.0x padded by one element in 16 and 2.7-3.7x plain).
__device__ __forceinline__ int lds_swz32(int n) { return n ^ (((n >> 5) * 5) & 31); }
// The same for element e of a tile of contiguous lines of L elements.
template <int L>
__device__ __forceinline__ int lds_swz_tile(int e) {
  return e ^ ((((e & (L - 1)) >> 5) * 5) & 31);
}

// Accessor over one line in LDS with compile-time element stride ES; PAD:
// the line is stored padded (1: lds_pad16) or swizzled (2: lds_swz32), ES
// = 1 only.
template <typename T, int ES, int PAD = 0>
struct LineAcc {
  static_assert(!PAD || ES == 1, "padded and swizzled lines are contiguous");
  cpx<T>* p;
  static __device__ __forceinline__ int at(int n) {
    return PAD == 1 ? lds_pad16(n) : PAD == 2 ? lds_swz32(n) : n * ES;
  }
  __device__ __forceinline__ cpx<T> ld(int n) const { return p[at(n)]; }
  __device__ __forceinline__ void st(int n, cpx<T> v) const { p[at(n)] = v; }
};

// No element-wise operation before a transform's first stage.
struct NoPre {};

// One in-place stage.  Thread q of NT threads on the line performs
// butterflies j = q + NT*i.  Twiddles for stage s live at tw[TWOFF +
// (r-1)*NS + k] = exp(-2 pi i r k / (NS R)); conjugated for SIGN = +1.
// pre(n, x): applied to element n of the line as it is loaded (a spectral
// symbol folded into the first stage of an inverse transform).
template <typename T, int N, int R, int NS, int NT, int SIGN, int TWOFF, int ES, int PAD = 0,
          typename PRE = NoPre>
__device__ __forceinline__ void sstage(const LineAcc<T, ES, PAD>& acc, int q,
                                       const cpx<T>* __restrict__ tw, const PRE& pre = PRE{}) {
  constexpr int NB = N / R;
  constexpr int PER = NB / NT;
  static_assert(NB % NT == 0, "threads per line must divide the butterfly count");
  cpx<T> v[PER][R];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int j = q + NT * i;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v[i][r] = acc.ld(j + r * NB);
      if constexpr (!std::is_same<PRE, NoPre>::value) v[i][r] = pre(j + r * NB, v[i][r]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int j = q + NT * i;
    const int k = (NS == 1) ? 0 : (j & (NS - 1));
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const cpx<T> w = tw[TWOFF + (r - 1) * NS + k];
        v[i][r] = (SIGN < 0) ? cmul(v[i][r], w) : cmulc(v[i][r], w);
      }
    }
    dft<T, R, SIGN>(v[i]);
    const int base = (j - k) * R + k;  // (j/NS)*NS*R + k
#pragma unroll
    for (int r = 0; r < R; ++r) acc.st(base + r * NS, v[i][r]);
  }
  __syncthreads();
}

// ------------------------------------------------------------------
// Dynamic engine: any N up to the LDS budget; schedule at run time.
// ------------------------------------------------------------------
struct DynSched {
  int n;
  int nst;
  int rad[20];
};

// Flat butterfly index b -> (line, j).  Layout Z: line stride LS, element
// stride 1, threads walk j fastest.  Layout S: element stride W, line (= the
// column inside the tile) fastest.
// q = b / d with r = b % d for 0 <= b < 2^22 and d >= 1, given rcp =
// 1.0f / d: the float quotient is within one of the true one (relative
// error below 2^-23), so one correction step makes it exact.  A few VALU
// operations instead of the integer division sequence.
__device__ __forceinline__ int fdivmod(int b, int d, float rcp, int& r) {
  int q = (int)((float)b * rcp);
  r = b - q * d;
  if (r < 0) {
    --q;
    r += d;
  } else if (r >= d) {
    ++q;
    r -= d;
  }
  return q;
}

struct LayZ {
  int ls;
  __device__ __forceinline__ int addr(int line, int n) const { return line * ls + n; }
  __device__ __forceinline__ void map(int b, int nb, float rnb, int& line, int& j) const {
    line = fdivmod(b, nb, rnb, j);
  }
};
struct LayS {
  int w;
  __device__ __forceinline__ int addr(int line, int n) const { return n * w + line; }
  __device__ __forceinline__ void map(int b, int nb, float rnb, int& line, int& j) const {
    j = b >> (__ffs(w) - 1);  // w is a power of two
    line = b & (w - 1);
  }
};

template <typename T, int R, int SIGN, class Lay>
__device__ __forceinline__ void dstage(const cpx<T>* src, cpx<T>* dst, int N, int NS,
                                       int nlines, const Lay& lay,
                                       const cpx<T>* __restrict__ twg) {
  const int NB = N / R;
  const int total = nlines * NB;
  const int tstep = N / (NS * R);
  const float rnb = 1.0f / (float)NB, rns = 1.0f / (float)NS;
  for (int b = threadIdx.x; b < total; b += blockDim.x) {
    int line, j, k;
    lay.map(b, NB, rnb, line, j);
    (void)fdivmod(j, NS, rns, k);
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cpx<T> x = src[lay.addr(line, j + r * NB)];
      if (r > 0 && NS > 1) {
        const cpx<T> w = twg[r * k * tstep];  // r k tstep < R NS tstep = N
        x = (SIGN < 0) ? cmul(x, w) : cmulc(x, w);
      }
      v[r] = x;
    }
    dft<T, R, SIGN>(v);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[lay.addr(line, base + r * NS)] = v[r];
  }
}

// Generic radix (any R, O(R^2) per butterfly) straight from LDS.
template <typename T, int SIGN, class Lay>
__device__ __forceinline__ void dstage_generic(const cpx<T>* src, cpx<T>* dst, int N, int R,
                                               int NS, int nlines, const Lay& lay,
                                               const cpx<T>* __restrict__ twg) {
  const int NB = N / R;
  const int total = nlines * NB;
  const int tstep = N / (NS * R);
  const int rstep = N / R;
  const float rnb = 1.0f / (float)NB, rns = 1.0f / (float)NS;
  for (int b = threadIdx.x; b < total; b += blockDim.x) {
    int line, j, k;
    lay.map(b, NB, rnb, line, j);
    (void)fdivmod(j, NS, rns, k);
    const int base = (j - k) * R + k;
    for (int q = 0; q < R; ++q) {
      cpx<T> acc = {(T)0, (T)0};
      for (int r = 0; r < R; ++r) {
        const long m = ((long)r * k * tstep + (long)((r * q) % R) * rstep) % N;
        const cpx<T> w = twg[m];
        const cpx<T> x = src[lay.addr(line, j + r * NB)];
        const cpx<T> t = (SIGN < 0) ? cmul(x, w) : cmulc(x, w);
        acc = cadd(acc, t);
      }
      dst[lay.addr(line, base + q * NS)] = acc;
    }
  }
}

// Runs the whole transform; data starts in a, the returned pointer holds the
// result (a or b).  Caller must __syncthreads() before; returns after a
// barrier.
template <typename T, int SIGN, class Lay>
__device__ cpx<T>* dyn_fft(cpx<T>* a, cpx<T>* b, const DynSched& sc, int nlines,
                           const Lay& lay, const cpx<T>* __restrict__ twg) {
  int NS = 1;
  const int N = sc.n;
  for (int s = 0; s < sc.nst; ++s) {
    const int R = sc.rad[s];
    switch (R) {
      case 2: dstage<T, 2, SIGN>(a, b, N, NS, nlines, lay, twg); break;
      case 3: dstage<T, 3, SIGN>(a, b, N, NS, nlines, lay, twg); break;
      case 4: dstage<T, 4, SIGN>(a, b, N, NS, nlines, lay, twg); break;
      case 5: dstage<T, 5, SIGN>(a, b, N, NS, nlines, lay, twg); break;
      case 7: dstage<T, 7, SIGN>(a, b, N, NS, nlines, lay, twg); break;
      case 8: dstage<T, 8, SIGN>(a, b, N, NS, nlines, lay, twg); break;
      case 9: dstage<T, 9, SIGN>(a, b, N, NS, nlines, lay, twg); break;
      case 16: dstage<T, 16, SIGN>(a, b, N, NS, nlines, lay, twg); break;
      default: dstage_generic<T, SIGN>(a, b, N, R, NS, nlines, lay, twg); break;
    }
    __syncthreads();
    cpx<T>* t = a;
    a = b;
    b = t;
    NS *= R;
  }
  return a;
}

// ------------------------------------------------------------------
// Spectral symbols (src/operators_cuda.cu:15-205, operators.txx:276-392)
// ------------------------------------------------------------------
enum SymKind {
  SYM_GRAD_X = 0,
  SYM_GRAD_Y = 1,
  SYM_GRAD_Z = 2,
  SYM_LAPLACE = 3,
  SYM_BIHARMONIC = 4,
  SYM_INV_LAPLACE = 5,
  SYM_INV_BIHARMONIC = 6,
  // Poisson Fourier filter of the step4 solver (steps/step4/kernels.cu:24-110),
  // unit box, divided by the FFT size: method 0 = the discrete 7-point
  // Laplacian's eigenvalues, method 1 = the continuous -4 pi^2 |k|^2
  SYM_POISSON_DISCRETE = 7,
  SYM_POISSON_CONTINUOUS = 8
};

// w = X; if X > N/2: w -= N; if X == N/2: w = 0  (operators_cuda.cu:32-37)
__device__ __forceinline__ long wave_nyq0(long X, long N) {
  long w = X;
  if (X > N / 2) w -= N;
  if (X == N / 2) w = 0;
  return w;
}
// Odd-symmetric wavenumber of the per-axis r2c -> i*w -> c2r chain
// (operators_gpu.txx:172-278): modes k and N-k are both zeroed when either
// equals N/2 (integer division).
__device__ __forceinline__ long wave_sym(long k, long N) {
  if (k == N / 2 || N - k == N / 2) return 0;
  return k > N / 2 ? k - N : k;
}

// Applies symbol `kind` at global (X, Y, Z) with FP64 factor arithmetic as
// the reference, result stored in T.
template <typename T>
__device__ __forceinline__ cpx<T> apply_sym(int kind, long X, long Y, long Z, long N0, long N1,
                                            long N2, double scale, cpx<T> v) {
  if (kind <= SYM_GRAD_Z) {
    const long w = kind == SYM_GRAD_X ? wave_nyq0(X, N0)
                 : kind == SYM_GRAD_Y ? wave_nyq0(Y, N1)
                                      : wave_nyq0(Z, N2);
    const T f = (T)(scale * (double)w);
    return {-f * v.y, f * v.x};
  }
  if (kind == SYM_POISSON_DISCRETE || kind == SYM_POISSON_CONTINUOUS) {
    // every mode but (0, 0, 0) divided by its eigenvalue times N0 N1 N2
    // (1 / scale); the mean is set to zero.  Method 1 shifts k > N/2 (a
    // floating comparison: the even-N Nyquist mode stays +N/2).
    if (X == 0 && Y == 0 && Z == 0) return {(T)0, (T)0};
    double ev;
    if (kind == SYM_POISSON_DISCRETE) {
      const double tp = 6.283185307179586476925;
      const double n0 = (double)N0, n1 = (double)N1, n2 = (double)N2;
      ev = 2.0 * ((cos(tp * (double)X / n0) - 1.0) * (n0 * n0) +
                  (cos(tp * (double)Y / n1) - 1.0) * (n1 * n1) +
                  (cos(tp * (double)Z / n2) - 1.0) * (n2 * n2));
    } else {
      const double kx = (double)X > 0.5 * (double)N0 ? (double)(X - N0) : (double)X;
      const double ky = (double)Y > 0.5 * (double)N1 ? (double)(Y - N1) : (double)Y;
      const double kz = (double)Z > 0.5 * (double)N2 ? (double)(Z - N2) : (double)Z;
      ev = -39.47841760435743447534 * (kx * kx + ky * ky + kz * kz);  // -4 pi^2 |k|^2
    }
    const double den = ev / scale;
    return {(T)((double)v.x / den), (T)((double)v.y / den)};
  }
  double f;
  if (kind == SYM_INV_BIHARMONIC) {
    long wx = X > N0 / 2 ? X - N0 : X, wy = Y > N1 / 2 ? Y - N1 : Y, wz = Z > N2 / 2 ? Z - N2 : Z;
    double wave = (double)(-wx * wx - wy * wy - wz * wz);
    wave *= wave;
    if (wave == 0) wave = 1;
    f = 1.0 / wave;
  } else {
    const long wx = wave_nyq0(X, N0), wy = wave_nyq0(Y, N1), wz = wave_nyq0(Z, N2);
    const long wave = -wx * wx - wy * wy - wz * wz;
    if (kind == SYM_LAPLACE) f = (double)wave;
    else if (kind == SYM_BIHARMONIC) f = (double)(wave * wave);
    else f = wave == 0 ? 0.0 : 1.0 / (double)wave;
  }
  const T g = (T)(scale * f);
  return {g * v.x, g * v.y};
}

}  // namespace afft

// fft_mfma.hpp -- radix-16 DFT stages on the gfx950 matrix cores with
// split-precision operands and FP32 accumulation (FP32 plans only).
//
// A radix-16 DFT of 32 butterflies is one real matrix product
//   [Re Y; Im Y] (32 x 32) = [C -S; S C] (32 x 32) * [Re X; Im X] (32 x 32)
// with C = cos(2 pi q r / 16), S = SIGN sin(2 pi q r / 16).  It runs as two
// v_mfma_f32_32x32x16 K-blocks (real parts, imaginary parts) per operand
// pair.  Each FP32 operand is split into P terms of a 16-bit format,
//   x = x0 + x1 (+ x2),  x0 = fmt(x), x1 = fmt(x - x0), ...
// and the products x_i * w_j with i + j < P are accumulated in FP32:
//   SPLIT_BF16   2 bf16 terms, 3 products  (~16 significant bits)
//   SPLIT_BF16X3 3 bf16 terms, 6 products  (~24 bits)
//   SPLIT_FP16   2 fp16 terms, 3 products  (~22 bits); each butterfly's
//                inputs are scaled by a power of two into [-1, 1] first so
//                that fp16's range holds any FP32 input.
// Lane layout (32x32x16): lane l holds A[l%32][8(l/32)+v], B[8(l/32)+v][l%32]
// for v < 8, and D[8(j/4) + 4(l/32) + j%4][l%32] in accumulator j < 16.  A
// lane (n, h) = (l%32, l/32) therefore feeds inputs x_{8h..8h+7} of
// butterfly n and receives the complete outputs Y_q, q in
// {4h..4h+3, 8+4h..11+4h}: accumulator j (< 8) is Re Y_q, j + 8 is Im Y_q.
#pragma once
#include "fft_device.hpp"

namespace afft {

enum SplitKind { SPLIT_NONE = 0, SPLIT_BF16 = 1, SPLIT_BF16X3 = 2, SPLIT_FP16 = 3 };

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8_t;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16_t;

template <int SPLIT>
struct SplitFmt;
template <>
struct SplitFmt<SPLIT_BF16> {
  using V = bf16x8_t;
  using E = __bf16;
  static constexpr int P = 2;
  static __device__ __forceinline__ f32x16_t mma(V a, V b, f32x16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct SplitFmt<SPLIT_BF16X3> {
  using V = bf16x8_t;
  using E = __bf16;
  static constexpr int P = 3;
  static __device__ __forceinline__ f32x16_t mma(V a, V b, f32x16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct SplitFmt<SPLIT_FP16> {
  using V = f16x8_t;
  using E = _Float16;
  static constexpr int P = 2;
  static __device__ __forceinline__ f32x16_t mma(V a, V b, f32x16_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  }
};

// Splits 8 FP32 values into P vectors of the 16-bit format.
template <int SPLIT>
__device__ __forceinline__ void split8(const float* v,
                                       typename SplitFmt<SPLIT>::V* parts) {
  using F = SplitFmt<SPLIT>;
  using E = typename F::E;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    float r = v[i];
#pragma unroll
    for (int p = 0; p < F::P; ++p) {
      const E e = (E)r;
      parts[p][i] = e;
      r -= (float)e;
    }
  }
}

// cos(2 pi m / 16), m in [0, 16)
__device__ __forceinline__ float cos16(int m) {
  const float c1 = 0.92387953251128675613f, c2 = 0.70710678118654752440f,
              c3 = 0.38268343236508977173f;
  const float t[16] = {1.f, c1, c2, c3, 0.f, -c3, -c2, -c1, -1.f, -c1, -c2, -c3, 0.f, c3, c2, c1};
  return t[m & 15];
}

// The two K-blocks of the DFT-16 matrix as seen by this lane:
// A0[m][r] (multiplies Re x_r), A1[m][r] (multiplies Im x_r), r = 8h + v.
template <int SPLIT, int SIGN>
__device__ __forceinline__ void dft16_operands(typename SplitFmt<SPLIT>::V* A0,
                                               typename SplitFmt<SPLIT>::V* A1) {
  const int lane = threadIdx.x & 63;
  const int m = lane & 31, h = lane >> 5;
  const int q = m & 15;
  float a0[8], a1[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int r = 8 * h + v;
    const int ph = (q * r) & 15;
    const float c = cos16(ph);
    const float s = (float)SIGN * cos16(ph - 4);  // sin(x) = cos(x - pi/2)
    if (m < 16) {  // Re Y_q = sum C Re x - S Im x
      a0[v] = c;
      a1[v] = -s;
    } else {  // Im Y_q = sum S Re x + C Im x
      a0[v] = s;
      a1[v] = c;
    }
  }
  split8<SPLIT>(a0, A0);
  split8<SPLIT>(a1, A1);
}

// One lane's share of 32 radix-16 butterflies: x[8] are inputs
// x_{8h..8h+7} (already twiddled) of butterfly lane%32; returns the 8
// outputs Y_q of this lane (q as in the header comment), in y[8].
template <int SPLIT>
__device__ __forceinline__ void dft16_mfma(const cpx<float>* x, const typename SplitFmt<SPLIT>::V* A0,
                                           const typename SplitFmt<SPLIT>::V* A1, cpx<float>* y) {
  using F = SplitFmt<SPLIT>;
  using V = typename F::V;
  float re[8], im[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    re[v] = x[v].x;
    im[v] = x[v].y;
  }
  float scale = 1.f;
  if constexpr (SPLIT == SPLIT_FP16) {
    // a power of two that maps the butterfly's largest component into [0.5, 1)
    float mx = 0.f;
#pragma unroll
    for (int v = 0; v < 8; ++v) mx = fmaxf(mx, fmaxf(fabsf(re[v]), fabsf(im[v])));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    int e = 0;
    (void)frexpf(mx, &e);
    const float down = ldexpf(1.f, -e);
    scale = ldexpf(1.f, e);
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      re[v] *= down;
      im[v] *= down;
    }
  }
  V B0[F::P], B1[F::P];
  split8<SPLIT>(re, B0);
  split8<SPLIT>(im, B1);
  f32x16_t d = {};
#pragma unroll
  for (int i = 0; i < F::P; ++i)
#pragma unroll
    for (int j = 0; j < F::P - i; ++j) {
      d = F::mma(A0[i], B0[j], d);
      d = F::mma(A1[i], B1[j], d);
    }
#pragma unroll
  for (int j = 0; j < 8; ++j) y[j] = {d[j] * scale, d[j + 8] * scale};
}

// Output index q of accumulator pair j (< 8) for lane half h.
__device__ __forceinline__ int dft16_q(int j, int h) { return 8 * (j >> 2) + 4 * h + (j & 3); }

// One in-place radix-16 Stockham stage of a tile held in LDS: element n of
// line L lives at tile[L*LS + n*ES].  LINES lines of N points; the 4 waves
// of the block share the LINES*N/16 butterflies, 32 per wave per round.
template <int N, int NS, int SIGN, int TWOFF, int ES, int LINES, int SPLIT>
__device__ __forceinline__ void mstage(cpx<float>* tile, int LS, const cpx<float>* __restrict__ tw) {
  using V = typename SplitFmt<SPLIT>::V;
  constexpr int NB = N / 16;
  constexpr int TB = LINES * NB;
  constexpr int WAVES = 4;
  constexpr int IT = TB / (32 * WAVES);
  static_assert(TB % (32 * WAVES) == 0, "tile too small for the matrix-core stage");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n = lane & 31, h = lane >> 5;
  V A0[SplitFmt<SPLIT>::P], A1[SplitFmt<SPLIT>::P];
  dft16_operands<SPLIT, SIGN>(A0, A1);
  cpx<float> x[IT][8];
  int lo[IT], jb[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int B = (wave + WAVES * it) * 32 + n;
    int line, j;
    if constexpr (ES == 1) {  // contiguous lines: butterflies of a line adjacent
      line = B / NB;
      j = B % NB;
    } else {  // strided tile: columns (lines) fastest
      line = B % LINES;
      j = B / LINES;
    }
    lo[it] = line * LS;
    jb[it] = j;
#pragma unroll
    for (int v = 0; v < 8; ++v) x[it][v] = tile[lo[it] + (j + (8 * h + v) * NB) * ES];
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int j = jb[it];
    const int k = (NS == 1) ? 0 : (j & (NS - 1));
    if constexpr (NS > 1) {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int r = 8 * h + v;
        if (r > 0) {
          const cpx<float> w = tw[TWOFF + (r - 1) * NS + k];
          x[it][v] = (SIGN < 0) ? cmul(x[it][v], w) : cmulc(x[it][v], w);
        }
      }
    }
    cpx<float> y[8];
    dft16_mfma<SPLIT>(x[it], A0, A1, y);
    const int base = (j - k) * 16 + k;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) tile[lo[it] + (base + dft16_q(jj, h) * NS) * ES] = y[jj];
  }
  __syncthreads();
}

// Self-test: one wavefront computes 32 DFT-16s of x (32 x 16 complex,
// butterfly-major) into y with the matrix-core path; used by the tests to
// pin the lane layout and the split error against the FP64 oracle.
template <int SPLIT, int SIGN>
__global__ void __launch_bounds__(64) dft16_selftest_kernel(const cpx<float>* x, cpx<float>* y) {
  using V = typename SplitFmt<SPLIT>::V;
  const int lane = threadIdx.x;
  const int n = lane & 31, h = lane >> 5;
  V A0[SplitFmt<SPLIT>::P], A1[SplitFmt<SPLIT>::P];
  dft16_operands<SPLIT, SIGN>(A0, A1);
  cpx<float> in[8], out[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) in[v] = x[n * 16 + 8 * h + v];
  dft16_mfma<SPLIT>(in, A0, A1, out);
#pragma unroll
  for (int j = 0; j < 8; ++j) y[n * 16 + dft16_q(j, h)] = out[j];
}

// Full transform of the lines of a tile held in LDS (natural order in and
// out), with the radix-16 stages on the matrix cores when SPLIT != 0.
// acc/q: this thread's line and index within it (vector stages);
// tile/LS: the whole tile (matrix-core stages).  Caller must
// __syncthreads() before (data and twiddles staged).
// pre: element-wise operation on the input as the first stage loads it
// (vector first stage only).
template <typename T, int N, int NT, int SIGN, int ES, int SPLIT, int LINES, int PAD = 0,
          typename PRE = NoPre>
__device__ __forceinline__ void static_fft(const LineAcc<T, ES, PAD>& acc, int q,
                                           const cpx<T>* __restrict__ tw, cpx<T>* tile, int LS,
                                           const PRE& pre = PRE{}) {
  using S = SchedOf<N, SPLIT>;
  constexpr int R0 = S::r(0), R1 = S::r(1), R2 = S::r(2);
  static_assert(!PAD || !split_sched(SPLIT), "matrix-core stages read the unpadded tile");
  static_assert(std::is_same<PRE, NoPre>::value || !(split_sched(SPLIT) && R0 == 16),
                "an input operation needs a vector first stage");
  if constexpr (split_sched(SPLIT) && R0 == 16)
    mstage<N, 1, SIGN, 0, ES, LINES, SPLIT>(tile, LS, tw);
  else
    sstage<T, N, R0, 1, NT, SIGN, 0, ES, PAD, PRE>(acc, q, tw, pre);
  if constexpr (S::S > 1) {
    if constexpr (split_sched(SPLIT) && R1 == 16)
      mstage<N, sched_ns<N, SPLIT>(1), SIGN, sched_twoff<N, SPLIT>(1), ES, LINES, SPLIT>(tile, LS, tw);
    else
      sstage<T, N, R1, sched_ns<N, SPLIT>(1), NT, SIGN, sched_twoff<N, SPLIT>(1), ES>(acc, q, tw);
  }
  if constexpr (S::S > 2) {
    if constexpr (split_sched(SPLIT) && R2 == 16)
      mstage<N, sched_ns<N, SPLIT>(2), SIGN, sched_twoff<N, SPLIT>(2), ES, LINES, SPLIT>(tile, LS, tw);
    else
      sstage<T, N, R2, sched_ns<N, SPLIT>(2), NT, SIGN, sched_twoff<N, SPLIT>(2), ES>(acc, q, tw);
  }
}

}  // namespace afft

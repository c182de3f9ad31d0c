// fft_kernels.hpp -- batched 1-D FFT kernels for the three axis passes.
//
// zfft_*  : lines contiguous in memory (the z pass; cuFFT fplan_0/iplan_0 of
//           src/accfft_gpu.cpp:173-206 and the c2c fplan_0 of :767-789).
// sfft_*  : lines strided in memory (the y and x passes; fplan_1/fplan_2 of
//           :208-253, the per-axis fplan_y/fplan_x of :255-323).  A tile is
//           W consecutive columns x N rows, so every row segment a wavefront
//           touches is W contiguous elements.
//
// Modes:
//   M_C2C   complex in -> DFT(SIGN) -> complex out
//   M_R2C   two real lines a,b packed as a+ib -> DFT(-1) -> the two half
//           spectra A[k] = (Z[k]+conj Z[N-k])/2, B[k] = -i(Z[k]-conj Z[N-k])/2
//   M_C2R   two half spectra -> Z = A + iB (Hermitian extension, Im of DC
//           and Nyquist ignored) -> DFT(+1) -> Re to line a, Im to line b
//   M_MULT  complex in -> DFT(-1) -> spectral symbol -> DFT(+1) -> out
//           (the multiply of operators_gpu.txx fused between the forward
//           and inverse x passes)
//   M_DERIV two real lines -> DFT(-1) -> i*w_sym(k)*scale -> DFT(+1) -> real
//           (the per-axis r2c / grad_mult / c2r chain of
//           operators_gpu.txx:172-278 in one pass, optional accumulate)
#pragma once
#include <type_traits>

#include "fft_mfma.hpp"

namespace afft {

enum Mode { M_C2C = 0, M_R2C = 1, M_C2R = 2, M_MULT = 3, M_DERIV = 4 };

template <typename T>
struct FftArgs {
  const void* in;
  void* out;
  const cpx<T>* tw;  // static engine: per-stage table; dynamic: exp(-2 pi i m/n)
  DynSched sc;       // sc.n is the transform length for both engines
  int64_t nlines;    // z: lines (real lines for pair modes); s: columns G
  int64_t idist, odist;              // z: line distance (elements of in/out type)
  int64_t C, Din, Dout, Sin, Sout;   // s: column g -> (g/C)*D + g%C, row stride S
  int lpw;                           // dynamic: lines (z) or columns (s) per block
  int sym;                           // M_MULT symbol kind
  double scale;                      // M_MULT / M_DERIV scale
  long N0, N1, N2;                   // M_MULT global sizes
  int64_t Ch;                        // M_MULT: column g -> (g/Ch, g%Ch)
  int maxis;                         // M_MULT: axis of the rows (0: x, 1: y)
  int accumulate;                    // M_DERIV: out += result
  int xgroup;                        // s: tiles per XCD chunk (see xcd_tile)
  int reverse;                       // process blocks last-to-first
  int adjw;                          // z R2C: x2, z C2R: x1/2 on bins other than
                                     // DC / Nyquist (the adjoint weights)
  // strided C2C only: rows split in blocks of 2^shift, block b at b * B
  // (row r at (r % 2^shift) * S + (r >> shift) * B); shift 0 = one block.
  // This is the pack / unpack of a slab exchange fused into the y pass.
  int ishift, oshift;
  int64_t Bin, Bout;
  // z R2C store / C2R load only: the half-spectrum columns in blocks of
  // zcol (the last one cut at n/2+1), block q of every line stored together
  // at q * zcol * nlines as [nlines][width of block q]; 0 = plain rows.
  // This is the pack / unpack of a pencil share's row exchange fused into
  // the z pass.  zmag = ceil(2^32 / zcol): k / zcol = (k zmag) >> 32 exactly
  // for k, zcol < 2^16 (zcol = 1: unused).
  int64_t zcol;
  uint32_t zmag;
  // persistent strided tiles only: the plan's tile queue (QUEUE_WORDS
  // counters, zero between launches; nullptr = a fixed tile sequence per
  // workgroup).  Set by the launcher.
  unsigned* queue;
};

// Tile queue of the persistent strided tiles: one counter per XCD class of
// workgroups (blockIdx & 7) at word 32 x, the count of finished workgroups
// at word QUEUE_DONE; 128 bytes apart.
constexpr int QUEUE_DONE = 8 * 32;
constexpr int QUEUE_WORDS = QUEUE_DONE + 32;

// Element k of half-spectrum line g of a z pass (see FftArgs::zcol).
template <typename T>
__device__ __forceinline__ int64_t zcol_off(const FftArgs<T>& a, int64_t g, int64_t dist, int k,
                                            int nh) {
  if (a.zcol == 0) return g * dist + k;
  const int zc = (int)a.zcol;
  const int s = zc == 1 ? k : (int)__umulhi((unsigned)k, a.zmag) * zc;
  const int w = (nh - s) < zc ? (nh - s) : zc;
  return (int64_t)s * a.nlines + g * w + (k - s);
}

__device__ __forceinline__ int64_t row_off(int r, int64_t S, int shift, int64_t B) {
  if (shift == 0) return (int64_t)r * S;
  return (int64_t)(r & ((1 << shift) - 1)) * S + (int64_t)(r >> shift) * B;
}

// -------------------------------------------------------------- configs
// Global loads each thread of a dynamic-engine kernel keeps in flight.
constexpr int DYN_BATCH = 8;
template <int N, int SPLIT = 0, int ZT = 256>
struct ZCfg {
  static constexpr int NT = sched_minnb<N, SPLIT>();
  static constexpr int TPB = NT > ZT ? NT : ZT;
  static constexpr int LPW = TPB / NT;
};

template <typename T, int N, int SPLIT = 0, int LDSK = 64>
struct SCfg {
  static constexpr int TPB = 256;
  // keep the tile at <= LDSK KiB of LDS and at least one row per thread group
  static constexpr int W0 = (int)(LDSK * 1024 / (N * (int)sizeof(cpx<T>)));
  static constexpr int WMAX = W0 > 64 ? 64 : W0;
  static constexpr int WMIN = TPB / sched_minnb<N, SPLIT>();
  static constexpr int W = WMAX > WMIN ? WMAX : WMIN;
  static constexpr int NT = TPB / W;
};

template <typename T>
__device__ __forceinline__ cpx<T> ldc(const void* p, int64_t i) {
  return reinterpret_cast<const cpx<T>*>(p)[i];
}
template <typename T>
__device__ __forceinline__ T ldr(const void* p, int64_t i) {
  return reinterpret_cast<const T*>(p)[i];
}
// Non-temporal result stores (AFFT_NT_STORES=1) were measured slower on
// MI355X: the y pass of the r2c layout writes rows that straddle 128-byte
// lines, and partial-line non-temporal writes cannot merge in L2.
#ifndef AFFT_NT_STORES
#define AFFT_NT_STORES 0
#endif
template <typename T>
__device__ __forceinline__ void stc(void* p, int64_t i, cpx<T> v) {
#if AFFT_NT_STORES
  T* q = reinterpret_cast<T*>(p) + 2 * i;
  __builtin_nontemporal_store(v.x, q);
  __builtin_nontemporal_store(v.y, q + 1);
#else
  reinterpret_cast<cpx<T>*>(p)[i] = v;
#endif
}
template <typename T>
__device__ __forceinline__ void str(void* p, int64_t i, T v) {
#if AFFT_NT_STORES
  __builtin_nontemporal_store(v, reinterpret_cast<T*>(p) + i);
#else
  reinterpret_cast<T*>(p)[i] = v;
#endif
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  This
// maps block b to a tile so that each XCD takes chunks of g consecutive
// tiles: neighbouring tiles (which share cache lines when rows are not
// 128-byte aligned) meet in one L2, while the tiles in flight still cover a
// contiguous address range (all HBM channels busy).  g = 1 is the identity.
// A bijection on [0, nb): blocks past the last whole group of 8*g keep
// their index.
// g = 0 gives each XCD one contiguous range.
__device__ __forceinline__ int64_t xcd_tile(unsigned b, unsigned nb, unsigned g) {
  if (g == 0u) {
    const unsigned x = b & 7u, i = b >> 3, q = nb >> 3, r = nb & 7u;
    return (int64_t)x * q + (x < r ? x : r) + i;
  }
  const unsigned full = (nb / (8u * g)) * 8u * g;
  if (g == 1u || b >= full) return b;
  const unsigned x = b & 7u, i = b >> 3;
  return (int64_t)(i / g) * 8u * g + (int64_t)x * g + (i % g);
}

// Copies the static engine's twiddle table into LDS with an unrolled loop
// (issued after the caller's data loads so that all loads overlap).
template <typename T, int TWN, int TPB>
__device__ __forceinline__ void stage_twiddles(cpx<T>* dst, const cpx<T>* __restrict__ src,
                                               int t) {
  if constexpr (TWN > 0) {
    cpx<T> v[(TWN + TPB - 1) / TPB];
#pragma unroll
    for (int j = 0; j < (TWN + TPB - 1) / TPB; ++j) {
      const int i = t + j * TPB;
      if (i < TWN) v[j] = src[i];
    }
#pragma unroll
    for (int j = 0; j < (TWN + TPB - 1) / TPB; ++j) {
      const int i = t + j * TPB;
      if (i < TWN) dst[i] = v[j];
    }
  }
}

// -------------------------------------------------------------- engines
// N > 0: static engine; N == 0: dynamic engine.
template <typename T, int N, int NT, int SIGN, int SPLIT = 0, int LINES = 256 / NT, int PAD = 0,
          typename PRE = NoPre>
__device__ __forceinline__ cpx<T>* zrun(cpx<T>* A, cpx<T>* B, const FftArgs<T>& a,
                                        const cpx<T>* twl, int lpw, const PRE& pre = PRE{}) {
  if constexpr (N > 0) {
    const int line = threadIdx.x / NT, q = threadIdx.x % NT;
    LineAcc<T, 1, PAD> acc{A + line * (PAD == 1 ? N + N / 16 : N)};
    static_fft<T, N, NT, SIGN, 1, SPLIT, LINES, PAD, PRE>(acc, q, twl, A, N, pre);
    return A;
  } else {
    LayZ lay{a.sc.n};
    return dyn_fft<T, SIGN>(A, B, a.sc, lpw, lay, a.tw);
  }
}

template <typename T, int N, int W, int SIGN, int SPLIT = 0>
__device__ __forceinline__ cpx<T>* srun(cpx<T>* A, cpx<T>* B, const FftArgs<T>& a,
                                        const cpx<T>* twl, int w_dyn) {
  if constexpr (N > 0) {
    const int w = threadIdx.x % W, q = threadIdx.x / W;
    LineAcc<T, W> acc{A + w};
    static_fft<T, N, 256 / W, SIGN, W, SPLIT, W>(acc, q, twl, A, 1);
    return A;
  } else {
    LayS lay{w_dyn};
    return dyn_fft<T, SIGN>(A, B, a.sc, w_dyn, lay, a.tw);
  }
}

// ====================================================================
// Contiguous lines (z)
// ====================================================================
// ZT: threads per block (lines per block = ZT / threads per line).
template <typename T, int N, int MODE, int SIGN, int SPLIT = 0, int ZT = 256>
__global__ void __launch_bounds__(ZT > 256 ? ZT : 256) zfft_kernel(FftArgs<T> a) {
  constexpr bool STATIC = N > 0;
  constexpr int NS_ = STATIC ? N : 16;
  static_assert(ZT == 256 || SPLIT == 0, "matrix-core stages assume 256-thread blocks");
  using ZC = ZCfg<NS_, SPLIT, ZT>;
  constexpr int NT = ZC::NT;
  constexpr int TPB = STATIC ? ZC::TPB : 256;
  constexpr int TWN = STATIC ? sched_twn<NS_, SPLIT>() : 0;
  constexpr bool PAIR = (MODE == M_R2C || MODE == M_C2R || MODE == M_DERIV);
  const int n = STATIC ? N : a.sc.n;
  const int nh = n / 2 + 1;
  const int lpw = STATIC ? ZC::LPW : a.lpw;  // complex lines per block
  const int t = threadIdx.x;
  // static tiles are stored padded (lds_pad16): the stage writes spread
  // over the banks (the swizzle of lds_swz32 measured slower here: its
  // index arithmetic costs the derivative pass a wave per SIMD)
  constexpr int ZPAD = (STATIC && SPLIT == 0) ? 1 : 0;
  auto P = [](int e) { return ZPAD ? lds_pad16(e) : e; };

  cpx<T>* bufA;
  cpx<T>* bufB;
  const cpx<T>* twl;
  if constexpr (STATIC) {
    constexpr int LPW = ZC::LPW;
    __shared__ cpx<T> s_buf[LPW * NS_ + (ZPAD ? LPW * NS_ / 16 : 0)];
    __shared__ cpx<T> s_tw[TWN > 0 ? TWN : 1];
    bufA = s_buf;
    bufB = s_buf;
    twl = s_tw;
  } else {
    extern __shared__ float4 dsm[];
    bufA = reinterpret_cast<cpx<T>*>(dsm);
    bufB = bufA + lpw * n;
    twl = a.tw;
  }

  // reverse: walk the lines last-to-first, so a pass can begin with the data
  // its predecessor wrote last (still in the memory-side cache)
  const int64_t bx = xcd_tile(blockIdx.x, gridDim.x, (unsigned)a.xgroup);
  const int64_t blk = a.reverse ? (int64_t)gridDim.x - 1 - bx : bx;
  const int64_t l0 = blk * lpw;  // first complex line / pair
  const int64_t nl = a.nlines;

  // ---------------- load
  if constexpr (STATIC) {
    // every thread stages its elements in registers so that all of its
    // global loads are in flight together, then fills LDS
    constexpr int LPW = ZC::LPW;
    constexpr int EP = LPW * NS_ / TPB;  // elements per thread
    constexpr int NH = NS_ / 2 + 1;
    if constexpr (MODE == M_C2R) {
      // Z = A + iB on [0, NH) and conj(A) + i conj(B) on (NH, N): each
      // half-spectrum element is read once, coalesced along k
      constexpr int NHT = LPW * NH;
      constexpr int EH = (NHT + TPB - 1) / TPB;
      cpx<T> A[EH], B[EH];
#pragma unroll
      for (int j = 0; j < EH; ++j) {
        const int q = t + j * TPB;
        A[j] = B[j] = cpx<T>{(T)0, (T)0};
        if (q < NHT) {
          const int p = q / NH, k = q % NH;
          const int64_t ga = 2 * (l0 + p), gb = ga + 1;
          if (ga < nl) A[j] = ldc<T>(a.in, zcol_off(a, ga, a.idist, k, NH));
          if (gb < nl) B[j] = ldc<T>(a.in, zcol_off(a, gb, a.idist, k, NH));
        }
      }
      stage_twiddles<T, TWN, TPB>(const_cast<cpx<T>*>(twl), a.tw, t);
#pragma unroll
      for (int j = 0; j < EH; ++j) {
        const int q = t + j * TPB;
        if (q < NHT) {
          const int p = q / NH, k = q % NH;
          cpx<T> x = A[j], y = B[j];
          if (k == 0 || 2 * k == NS_) {
            x.y = (T)0;
            y.y = (T)0;
          } else if (a.adjw) {
            x = cscale(x, (T)0.5);
            y = cscale(y, (T)0.5);
          }
          bufA[P(p * NS_ + k)] = cpx<T>{x.x - y.y, x.y + y.x};
          if (k != 0 && 2 * k != NS_) bufA[P(p * NS_ + NS_ - k)] = cpx<T>{x.x + y.y, y.x - x.y};
        }
      }
    } else {
    cpx<T> v[EP];
#pragma unroll
    for (int j = 0; j < EP; ++j) {
      const int e = t + j * TPB;
      const int p = e / NS_, i = e % NS_;
      if constexpr (MODE == M_C2C || MODE == M_MULT) {
        const int64_t g = l0 + p;
        v[j] = g < nl ? ldc<T>(a.in, g * a.idist + i) : cpx<T>{(T)0, (T)0};
      } else if constexpr (MODE == M_R2C || MODE == M_DERIV) {
        const int64_t ra = 2 * (l0 + p), rb = ra + 1;
        v[j].x = ra < nl ? ldr<T>(a.in, ra * a.idist + i) : (T)0;
        v[j].y = rb < nl ? ldr<T>(a.in, rb * a.idist + i) : (T)0;
      }
    }
    stage_twiddles<T, TWN, TPB>(const_cast<cpx<T>*>(twl), a.tw, t);
#pragma unroll
    for (int j = 0; j < EP; ++j) bufA[P(t + j * TPB)] = v[j];
    }
  } else {
    // dynamic engine: batches of DYN_BATCH loads per thread in flight
    // before their LDS writes (a flat load cannot move past an LDS store,
    // so one load per iteration would wait out the memory latency each time)
    const int ne = lpw * n;
    for (int e0 = t; e0 < ne; e0 += DYN_BATCH * TPB) {
      cpx<T> v[DYN_BATCH];
#pragma unroll
      for (int j = 0; j < DYN_BATCH; ++j) {
        const int e = e0 + j * TPB;
        const int p = e / n, i = e - p * n;
        cpx<T> z = {(T)0, (T)0};
        if constexpr (MODE == M_C2C || MODE == M_MULT) {
          const int64_t g = l0 + p;
          if (e < ne && g < nl) z = ldc<T>(a.in, g * a.idist + i);
        } else if constexpr (MODE == M_R2C || MODE == M_DERIV) {
          const int64_t ra = 2 * (l0 + p), rb = ra + 1;
          if (e < ne && ra < nl) z.x = ldr<T>(a.in, ra * a.idist + i);
          if (e < ne && rb < nl) z.y = ldr<T>(a.in, rb * a.idist + i);
        } else {  // M_C2R
          const int64_t ga = 2 * (l0 + p), gb = ga + 1;
          const int k = i < nh ? i : n - i;
          cpx<T> A = {(T)0, (T)0}, B = {(T)0, (T)0};
          if (e < ne && ga < nl) A = ldc<T>(a.in, zcol_off(a, ga, a.idist, k, nh));
          if (e < ne && gb < nl) B = ldc<T>(a.in, zcol_off(a, gb, a.idist, k, nh));
          if (k == 0 || 2 * k == n) {
            A.y = (T)0;
            B.y = (T)0;
          } else if (a.adjw) {
            A = cscale(A, (T)0.5);
            B = cscale(B, (T)0.5);
          }
          z = i < nh ? cpx<T>{A.x - B.y, A.y + B.x} : cpx<T>{A.x + B.y, B.x - A.y};
        }
        v[j] = z;
      }
#pragma unroll
      for (int j = 0; j < DYN_BATCH; ++j) {
        const int e = e0 + j * TPB;
        if (e < ne) bufA[e] = v[j];
      }
    }
  }
  __syncthreads();

  // ---------------- transform
  cpx<T>* res;
  if constexpr (MODE == M_C2C) {
    res = zrun<T, N, NT, SIGN, SPLIT, ZC::LPW, ZPAD>(bufA, bufB, a, twl, lpw);
  } else if constexpr (MODE == M_R2C) {
    res = zrun<T, N, NT, -1, SPLIT, ZC::LPW, ZPAD>(bufA, bufB, a, twl, lpw);
  } else if constexpr (MODE == M_C2R) {
    res = zrun<T, N, NT, +1, SPLIT, ZC::LPW, ZPAD>(bufA, bufB, a, twl, lpw);
  } else {  // M_DERIV and M_MULT (z has no MULT use; treat as DERIV-style symbol)
    res = zrun<T, N, NT, -1, SPLIT, ZC::LPW, ZPAD>(bufA, bufB, a, twl, lpw);
    if constexpr (STATIC && !(split_sched(SPLIT) && SchedOf<NS_, SPLIT>::r(0) == 16)) {
      // i w(k) scale folded into the inverse transform's first-stage loads
      // (no LDS round trip and barrier of its own)
      const T sc = (T)a.scale;
      auto sym = [sc](int k, cpx<T> v) -> cpx<T> {
        const T f = sc * (T)(int)wave_sym(k, NS_);
        return {-f * v.y, f * v.x};
      };
      res = zrun<T, N, NT, +1, SPLIT, ZC::LPW, ZPAD>(res, res, a, twl, lpw, sym);
    } else {
      const int epm = STATIC ? ZC::LPW * NS_ / TPB : (lpw * n + TPB - 1) / TPB;
#pragma unroll
      for (int j = 0; j < epm; ++j) {
        const int e = t + j * TPB;
        if (!STATIC && e >= lpw * n) break;
        const int i = e % n;
        const T f = (T)(a.scale * (double)wave_sym(i, n));
        const cpx<T> v = res[P(e)];
        res[P(e)] = {-f * v.y, f * v.x};
      }
      __syncthreads();
      cpx<T>* other = (res == bufA) ? bufB : bufA;
      res = zrun<T, N, NT, +1, SPLIT, ZC::LPW, ZPAD>(res, other, a, twl, lpw);
    }
  }

  // ---------------- store
  // static: trip counts are compile-time so the loops unroll
  constexpr int EPS = STATIC ? ZC::LPW * NS_ / TPB : 1;
  constexpr int EPH = STATIC ? (2 * ZC::LPW * (NS_ / 2 + 1) + TPB - 1) / TPB : 1;
  const int ep = STATIC ? EPS : (lpw * n + TPB - 1) / TPB;
  const int eph = STATIC ? EPH : (2 * lpw * nh + TPB - 1) / TPB;
  if constexpr (MODE == M_C2C || MODE == M_MULT) {
#pragma unroll
    for (int j = 0; j < ep; ++j) {
      const int e = t + j * TPB;
      if (!STATIC && e >= lpw * n) break;
      const int p = e / n, i = e - p * n;
      const int64_t g = l0 + p;
      if (g < nl) stc<T>(a.out, g * a.odist + i, res[P(e)]);
    }
  } else if constexpr (MODE == M_R2C) {
#pragma unroll
    for (int j = 0; j < eph; ++j) {
      const int e = t + j * TPB;
      if (e >= 2 * lpw * nh) break;
      const int rl = e / nh, k = e - rl * nh;
      const int p = rl >> 1;
      const int64_t g = 2 * l0 + rl;
      if (g >= nl) continue;
      const cpx<T> z1 = res[P(p * n + k)];
      const cpx<T> z2 = res[P(p * n + (k == 0 ? 0 : n - k))];
      const T h = (T)0.5;
      cpx<T> o;
      if ((rl & 1) == 0) o = {h * (z1.x + z2.x), h * (z1.y - z2.y)};
      else o = {h * (z1.y + z2.y), h * (z2.x - z1.x)};
      if (a.adjw && k != 0 && 2 * k != n) o = cscale(o, (T)2);
      stc<T>(a.out, zcol_off(a, g, a.odist, k, n / 2 + 1), o);
    }
  } else {  // real outputs
    if constexpr (MODE == M_DERIV && STATIC) {
      if (a.accumulate) {
        // all of the thread's accumulator loads in flight together, then
        // the stores (a load-add-store per element would wait on each load)
        cpx<T> o[EPS];
#pragma unroll
        for (int j = 0; j < EPS; ++j) {
          const int e = t + j * TPB;
          const int p = e / n, i = e - p * n;
          const int64_t ra = 2 * (l0 + p), rb = ra + 1;
          o[j].x = ra < nl ? ldr<T>(a.out, ra * a.odist + i) : (T)0;
          o[j].y = rb < nl ? ldr<T>(a.out, rb * a.odist + i) : (T)0;
        }
#pragma unroll
        for (int j = 0; j < EPS; ++j) {
          const int e = t + j * TPB;
          const int p = e / n, i = e - p * n;
          const int64_t ra = 2 * (l0 + p), rb = ra + 1;
          const cpx<T> z = res[P(e)];
          if (ra < nl) str<T>(a.out, ra * a.odist + i, o[j].x + z.x);
          if (rb < nl) str<T>(a.out, rb * a.odist + i, o[j].y + z.y);
        }
        return;
      }
    }
#pragma unroll
    for (int j = 0; j < ep; ++j) {
      const int e = t + j * TPB;
      if (!STATIC && e >= lpw * n) break;
      const int p = e / n, i = e - p * n;
      const int64_t ra = 2 * (l0 + p), rb = ra + 1;
      const cpx<T> z = res[P(e)];
      if constexpr (MODE == M_DERIV && !STATIC) {
        if (a.accumulate) {
          if (ra < nl) str<T>(a.out, ra * a.odist + i, ldr<T>(a.out, ra * a.odist + i) + z.x);
          if (rb < nl) str<T>(a.out, rb * a.odist + i, ldr<T>(a.out, rb * a.odist + i) + z.y);
          continue;
        }
      }
      if (ra < nl) str<T>(a.out, ra * a.odist + i, z.x);
      if (rb < nl) str<T>(a.out, rb * a.odist + i, z.y);
    }
  }
}

// ====================================================================
// Strided lines (y, x)
// ====================================================================
// Threads per block of the dynamic-engine strided kernel: 512 in FP32 (twice
// the butterflies in flight per tile: 1000^3 round trip 28.1 -> 24.6 ms,
// 729^3 9.4 -> 8.3 ms), 256 in FP64 (512 threads: 768^3 18.6 -> 22.0 ms).
template <typename T>
constexpr int dyn_s_tpb() { return sizeof(T) == 4 ? 512 : 256; }

template <typename T, int N, int MODE, int SIGN, int SPLIT = 0, int LDSK = 64>
__global__ void __launch_bounds__(N > 0 ? 256 : dyn_s_tpb<T>()) sfft_kernel(FftArgs<T> a) {
  constexpr bool STATIC = N > 0;
  constexpr int NS_ = STATIC ? N : 16;
  constexpr int TPB = STATIC ? 256 : dyn_s_tpb<T>();
  constexpr int WS = SCfg<T, NS_, SPLIT, LDSK>::W;
  constexpr int TWN = STATIC ? sched_twn<NS_, SPLIT>() : 0;
  const int n = STATIC ? N : a.sc.n;
  const int nh = n / 2 + 1;
  const int W = STATIC ? WS : a.lpw;  // complex columns per tile
  const int t = threadIdx.x;

  cpx<T>* bufA;
  cpx<T>* bufB;
  const cpx<T>* twl;
  if constexpr (STATIC) {
    __shared__ cpx<T> s_buf[WS * NS_];
    __shared__ cpx<T> s_tw[TWN > 0 ? TWN : 1];
    bufA = s_buf;
    bufB = s_buf;
    twl = s_tw;
  } else {
    extern __shared__ float4 dsm[];
    bufA = reinterpret_cast<cpx<T>*>(dsm);
    bufB = bufA + W * n;
    twl = a.tw;
  }

  // Each thread keeps one column w for the whole block (TPB % W == 0).
  const int w = t % W;
  const int r0 = t / W, rstep = TPB / W;
  const int64_t G = a.nlines;
  constexpr bool PAIR = (MODE == M_R2C || MODE == M_C2R || MODE == M_DERIV);
  int64_t ga, gb;  // columns (real columns for pair modes)
  int64_t tile = xcd_tile(blockIdx.x, gridDim.x, (unsigned)a.xgroup);
  if (a.reverse) tile = (int64_t)gridDim.x - 1 - tile;
  if constexpr (PAIR) {
    ga = 2 * (tile * W + w);
    gb = ga + 1;
  } else {
    ga = tile * W + w;
    gb = ga;
  }
  const bool va = ga < G, vb = gb < G;
  const int64_t ia = (ga / a.C) * a.Din + (ga % a.C), ib = (gb / a.C) * a.Din + (gb % a.C);
  const int64_t oa = (ga / a.C) * a.Dout + (ga % a.C), ob = (gb / a.C) * a.Dout + (gb % a.C);

  // ---------------- load
  // rows handled by this thread: r = r0 + j*rstep, j < rp
  constexpr int RP = STATIC ? NS_ * WS / TPB : 1;
  const int rp = STATIC ? RP : (n - r0 + rstep - 1) / rstep;
  if constexpr (STATIC) {
    constexpr int RS = TPB / WS;
    constexpr int NH = NS_ / 2 + 1;
    cpx<T> v[RP];
#pragma unroll
    for (int j = 0; j < RP; ++j) {
      const int r = r0 + j * RS;
      if constexpr (MODE == M_C2C || MODE == M_MULT) {
        v[j] = va ? ldc<T>(a.in, ia + row_off(r, a.Sin, a.ishift, a.Bin)) : cpx<T>{(T)0, (T)0};
      } else if constexpr (MODE == M_R2C || MODE == M_DERIV) {
        v[j].x = va ? ldr<T>(a.in, ia + (int64_t)r * a.Sin) : (T)0;
        v[j].y = vb ? ldr<T>(a.in, ib + (int64_t)r * a.Sin) : (T)0;
      } else {
        const int k = r < NH ? r : NS_ - r;
        cpx<T> A = va ? ldc<T>(a.in, ia + (int64_t)k * a.Sin) : cpx<T>{(T)0, (T)0};
        cpx<T> B = vb ? ldc<T>(a.in, ib + (int64_t)k * a.Sin) : cpx<T>{(T)0, (T)0};
        if (k == 0 || 2 * k == NS_) { A.y = (T)0; B.y = (T)0; }
        v[j] = r < NH ? cpx<T>{A.x - B.y, A.y + B.x} : cpx<T>{A.x + B.y, B.x - A.y};
      }
    }
    stage_twiddles<T, TWN, TPB>(const_cast<cpx<T>*>(twl), a.tw, t);
#pragma unroll
    for (int j = 0; j < RP; ++j) bufA[(r0 + j * RS) * WS + w] = v[j];
  } else {
    // dynamic engine: batches of DYN_BATCH row loads in flight per thread
    // before their LDS writes (as in the z kernel)
    for (int rb0 = r0; rb0 < n; rb0 += DYN_BATCH * rstep) {
      cpx<T> v[DYN_BATCH];
#pragma unroll
      for (int j = 0; j < DYN_BATCH; ++j) {
        const int r = rb0 + j * rstep;
        const bool in = r < n;
        cpx<T> z = {(T)0, (T)0};
        if constexpr (MODE == M_C2C || MODE == M_MULT) {
          if (in && va) z = ldc<T>(a.in, ia + row_off(r, a.Sin, a.ishift, a.Bin));
        } else if constexpr (MODE == M_R2C || MODE == M_DERIV) {
          if (in && va) z.x = ldr<T>(a.in, ia + (int64_t)r * a.Sin);
          if (in && vb) z.y = ldr<T>(a.in, ib + (int64_t)r * a.Sin);
        } else {  // M_C2R
          const int k = r < nh ? r : n - r;
          cpx<T> A = {(T)0, (T)0}, B = {(T)0, (T)0};
          if (in && va) A = ldc<T>(a.in, ia + (int64_t)k * a.Sin);
          if (in && vb) B = ldc<T>(a.in, ib + (int64_t)k * a.Sin);
          if (k == 0 || 2 * k == n) { A.y = (T)0; B.y = (T)0; }
          if (r < nh) z = {A.x - B.y, A.y + B.x};
          else z = {A.x + B.y, B.x - A.y};
        }
        v[j] = z;
      }
#pragma unroll
      for (int j = 0; j < DYN_BATCH; ++j) {
        const int r = rb0 + j * rstep;
        if (r < n) bufA[r * W + w] = v[j];
      }
    }
  }
  __syncthreads();

  // ---------------- transform
  cpx<T>* res;
  if constexpr (MODE == M_C2C) {
    res = srun<T, N, WS, SIGN, SPLIT>(bufA, bufB, a, twl, W);
  } else if constexpr (MODE == M_R2C) {
    res = srun<T, N, WS, -1, SPLIT>(bufA, bufB, a, twl, W);
  } else if constexpr (MODE == M_C2R) {
    res = srun<T, N, WS, +1, SPLIT>(bufA, bufB, a, twl, W);
  } else {
    res = srun<T, N, WS, -1, SPLIT>(bufA, bufB, a, twl, W);
#pragma unroll
    for (int j = 0; j < rp; ++j) {
      const int r = r0 + j * rstep;
      const cpx<T> v = res[r * W + w];
      cpx<T> o;
      if constexpr (MODE == M_DERIV) {
        const T f = (T)(a.scale * (double)wave_sym(r, n));
        o = {-f * v.y, f * v.x};
      } else {
        const long u = (long)(ga / a.Ch), Z = (long)(ga % a.Ch);
        const long X = a.maxis == 0 ? (long)r : u, Y = a.maxis == 0 ? u : (long)r;
        o = apply_sym<T>(a.sym, X, Y, Z, a.N0, a.N1, a.N2, a.scale, v);
      }
      res[r * W + w] = o;
    }
    __syncthreads();
    cpx<T>* other = (res == bufA) ? bufB : bufA;
    res = srun<T, N, WS, +1, SPLIT>(res, other, a, twl, W);
  }

  // ---------------- store
  if constexpr (MODE == M_C2C || MODE == M_MULT) {
    if (va) {
#pragma unroll
      for (int j = 0; j < rp; ++j) {
        const int r = r0 + j * rstep;
        stc<T>(a.out, oa + row_off(r, a.Sout, a.oshift, a.Bout), res[r * W + w]);
      }
    }
  } else if constexpr (MODE == M_R2C) {
    const T h = (T)0.5;
    const int rph = STATIC ? (NS_ / 2 + 1 + TPB / WS - 1) / (TPB / WS) : (nh + rstep - 1) / rstep;
#pragma unroll
    for (int j = 0; j < rph; ++j) {
      const int k = r0 + j * rstep;
      if (k >= nh) break;
      const cpx<T> z1 = res[k * W + w];
      const cpx<T> z2 = res[(k == 0 ? 0 : n - k) * W + w];
      if (va) stc<T>(a.out, oa + (int64_t)k * a.Sout, cpx<T>{h * (z1.x + z2.x), h * (z1.y - z2.y)});
      if (vb) stc<T>(a.out, ob + (int64_t)k * a.Sout, cpx<T>{h * (z1.y + z2.y), h * (z2.x - z1.x)});
    }
  } else {
    if constexpr (MODE == M_DERIV && STATIC) {
      if (a.accumulate) {
        // the accumulator loads all in flight, then the stores
        cpx<T> o[RP];
#pragma unroll
        for (int j = 0; j < RP; ++j) {
          const int64_t off = (int64_t)(r0 + j * rstep) * a.Sout;
          o[j].x = va ? ldr<T>(a.out, oa + off) : (T)0;
          o[j].y = vb ? ldr<T>(a.out, ob + off) : (T)0;
        }
#pragma unroll
        for (int j = 0; j < RP; ++j) {
          const int r = r0 + j * rstep;
          const cpx<T> z = res[r * W + w];
          if (va) str<T>(a.out, oa + (int64_t)r * a.Sout, o[j].x + z.x);
          if (vb) str<T>(a.out, ob + (int64_t)r * a.Sout, o[j].y + z.y);
        }
        return;
      }
    }
#pragma unroll
    for (int j = 0; j < rp; ++j) {
      const int r = r0 + j * rstep;
      const cpx<T> z = res[r * W + w];
      if constexpr (MODE == M_DERIV && !STATIC) {
        if (a.accumulate) {
          if (va) str<T>(a.out, oa + (int64_t)r * a.Sout, ldr<T>(a.out, oa + (int64_t)r * a.Sout) + z.x);
          if (vb) str<T>(a.out, ob + (int64_t)r * a.Sout, ldr<T>(a.out, ob + (int64_t)r * a.Sout) + z.y);
          continue;
        }
      }
      if (va) str<T>(a.out, oa + (int64_t)r * a.Sout, z.x);
      if (vb) str<T>(a.out, ob + (int64_t)r * a.Sout, z.y);
    }
  }
}

// ====================================================================
// Strided C2C, two-level (four-step) tiles for long lines
// ====================================================================
// N = R1 * M.  A tile of W columns is held in registers after the first
// level: thread (c, n2) loads x[n2 + M n1] for n1 < R1 straight from HBM,
// does the R1-point DFT and the twiddle exp(-+2 pi i n2 k1 / N) in
// registers; the M-point DFTs over n2 (one per k1 and column) then run in
// LDS, KC values of k1 at a time, and store X[k1 + R1 k2].  Only a
// KC/R1 share of the tile is ever in LDS, so W (the row segment) is 2-4x
// wider than one-level tiles of the same LDS budget.
// TWF = 1 (R1 = 16) reads only w, w^2, w^4, w^8 of the first-level
// twiddles w^k1 = exp(-+2 pi i n2 k1 / N) from a table of N/2 entries and
// forms the other powers with at most three complex products (a few ulp):
// half the LDS table of TWF = 0 and four live twiddles instead of fifteen.  SUB is the schedule id of
// the second level (0 or SCHED_WIDE: more threads per M-point line).
// Variant V = 0 is the default of each length; the others are kept for
// measurement (selected with ACCFFT_AMD_S4V).
template <typename T, int N, int V = 0>
struct S4Cfg {
  static constexpr bool OK = false;
  static constexpr int R1 = 16, W = 1, TPB = 256, KC = 1, TWF = 0, SUB = 0, WPE = 1, PB = 0;
};
// WPE: waves per SIMD the register allocation must allow (launch bound).
// PB > 0: persistent kernel with PB workgroups per CU that loads the next
// tile while it transforms the current one.
#define AFFT_S4CFG(TT, NN, VV, R1_, W_, TPB_, KC_, TWF_, SUB_, WPE_, PB_)    \
  template <>                                                                \
  struct S4Cfg<TT, NN, VV> {                                                 \
    static constexpr bool OK = true;                                         \
    static constexpr int R1 = R1_, W = W_, TPB = TPB_, KC = KC_, TWF = TWF_, SUB = SUB_, \
                         WPE = WPE_, PB = PB_;                               \
  };
// 256: only the slab share's 256-point y stage (its plan tables force it)
AFFT_S4CFG(float, 256, 0, 16, 32, 512, 8, 0, 0, 1, 0)
AFFT_S4CFG(float, 512, 0, 16, 16, 512, 8, 0, 0, 1, 0)
AFFT_S4CFG(float, 512, 1, 16, 32, 1024, 8, 0, 0, 1, 0)
AFFT_S4CFG(float, 512, 2, 16, 16, 512, 8, 1, 0, 1, 0)
AFFT_S4CFG(float, 512, 3, 16, 16, 512, 16, 0, 0, 4, 2)
AFFT_S4CFG(float, 1024, 0, 16, 8, 512, 8, 0, 0, 1, 0)
AFFT_S4CFG(float, 1024, 1, 16, 16, 1024, 8, 0, 0, 1, 0)
// (1024 v2: power twiddles keep it within 128 VGPRs; with the full table
// two spilled registers were reloaded per tile, +6 % HBM reads)
AFFT_S4CFG(float, 1024, 2, 16, 16, 1024, 16, 1, 0, 4, 1)
AFFT_S4CFG(float, 1024, 3, 16, 8, 512, 16, 0, 0, 4, 2)
AFFT_S4CFG(float, 2048, 0, 16, 8, 1024, 16, 0, 0, 1, 0)
AFFT_S4CFG(float, 2048, 1, 16, 4, 512, 16, 0, 0, 1, 0)
AFFT_S4CFG(float, 2048, 2, 16, 8, 1024, 16, 0, 0, 4, 1)
AFFT_S4CFG(float, 2048, 3, 16, 8, 1024, 16, 1, 0, 4, 1)
AFFT_S4CFG(double, 512, 0, 16, 16, 512, 8, 0, 0, 1, 0)
AFFT_S4CFG(double, 512, 1, 16, 8, 256, 16, 0, 0, 1, 0)
#undef AFFT_S4CFG

// One tile's work, shared by the one-shot and the persistent kernel.
// SPLIT != 0 (FP32): the first level's radix-16 DFTs run on the matrix cores
// with split operands (fft_mfma.hpp); the second level stays on the VALU.
template <typename T, int N, int SIGN, bool BLK, int V, int SPLIT = 0>
struct S4Tile {
  using CF = S4Cfg<T, N, V>;
  static constexpr int R1 = CF::R1, W = CF::W, TPB = CF::TPB, KC = CF::KC;
  static constexpr int M = N / R1;
  static constexpr int D1 = W * M / TPB;      // first-level DFTs per thread
  static constexpr int RS1 = TPB / W;         // n2 step between them
  static constexpr int NT = TPB / (KC * W);   // threads per M-point line
  static constexpr int SUB = CF::SUB;
  static constexpr int TWM = sched_twn<M, SUB>();
  static constexpr int TWF = CF::TWF;
  static constexpr int NTW1 = TWF ? N / 2 : N;  // first-level twiddle entries in LDS
  static_assert(!TWF || R1 == 16, "power twiddles are written for R1 = 16");
  static_assert(D1 >= 1 && D1 * TPB == W * M, "tile must split evenly");
  static_assert(NT >= 1 && NT * KC * W == TPB, "lines must split evenly");
  static_assert(R1 % KC == 0, "k1 chunks must split evenly");
  static_assert(!SPLIT || (R1 == 16 && TWF == 0 && 32 % W == 0 && TPB % 64 == 0 &&
                           (KC == 8 || KC == 16)),
                "matrix-core first level: R1 = 16, plain twiddles, lanes l and l^32 in one column");

  // Issues the loads of tile `tile` (x[n2 + M n1] of column c) into v.
  static __device__ __forceinline__ void load(cpx<T> (&v)[D1][R1], const FftArgs<T>& a,
                                              int64_t tile, int c, int r0) {
    const int64_t col = tile * W + c;
    const bool ok = col < a.nlines;
    const int64_t ib = (col / a.C) * a.Din + (col % a.C);
    const cpx<T>* src = reinterpret_cast<const cpx<T>*>(a.in) + ib;
#pragma unroll
    for (int d = 0; d < D1; ++d) {
      if constexpr (BLK) {
        // rows r0 + d RS1 + M n1 in exchange blocks; the row counter is
        // opaque to the compiler, so the R1 offsets are formed one at a
        // time instead of all being kept in registers
        const gptr<const cpx<T>> g = to_global(src);
        int row = r0 + d * RS1;
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
          v[d][n1] = ok ? ld_global(g + row_off(row, a.Sin, a.ishift, a.Bin)) : cpx<T>{(T)0, (T)0};
          row += M;
          asm volatile("" : "+v"(row));
        }
      } else {
        // running pointer, opaque to the compiler: the R1 row offsets are
        // not precomputed into registers
        gptr<const cpx<T>> p = to_global(src + (int64_t)(r0 + d * RS1) * a.Sin);
        const int64_t step = (int64_t)M * a.Sin;
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
          v[d][n1] = ok ? ld_global(p) : cpx<T>{(T)0, (T)0};
          p += step;
          asm volatile("" : "+v"(p));
        }
      }
    }
  }

  // First level: R1-point DFTs and twiddles in registers.
  //
  // Matrix-core form: lanes l and l^32 of a wavefront hold butterflies A
  // (the lane with bit 5 clear) and B.  They swap halves (one
  // v_permlane32_swap per register pair, no LDS traffic) so that every
  // lane holds inputs 8h..8h+7 of A in v[0..7] and of B in v[8..15]
  // (h = bit 5), two 32-butterfly products give each lane outputs
  // q = dft16_q(j, h) of A (v[j]) and B (v[8 + j]), j < 8, and put() writes
  // them to their k1 = q rows directly (no swap back).
  static __device__ __forceinline__ void first(cpx<T> (&v)[D1][R1], int t, const cpx<T>* tw1) {
    if constexpr (SPLIT != 0) {
      using SF = SplitFmt<SPLIT>;
      const int h = (t >> 5) & 1;
      typename SF::V A0[SF::P], A1[SF::P];
      dft16_operands<SPLIT, SIGN>(A0, A1);
#pragma unroll
      for (int d = 0; d < D1; ++d) {
        cpx<float> x0[8], x1[8], y0[8], y1[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          // v_permlane32_swap(x, y) swaps lanes 32..63 of x with lanes 0..31
          // of y: x becomes (A_i | A_{8+i}), y becomes (B_i | B_{8+i})
          const auto re = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[d][i].x),
                                                           __float_as_uint(v[d][8 + i].x), false, false);
          const auto im = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[d][i].y),
                                                           __float_as_uint(v[d][8 + i].y), false, false);
          x0[i] = {__uint_as_float(re[0]), __uint_as_float(im[0])};
          x1[i] = {__uint_as_float(re[1]), __uint_as_float(im[1])};
        }
        dft16_mfma<SPLIT>(x0, A0, A1, y0);
        dft16_mfma<SPLIT>(x1, A0, A1, y1);
        const int n2a = (t & ~32) / W + d * RS1, n2b = (t | 32) / W + d * RS1;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int q = 8 * (j >> 2) + 4 * h + (j & 3);
          const cpx<T> wa = tw1[n2a * q], wb = tw1[n2b * q];
          v[d][j] = (SIGN < 0) ? cmul(y0[j], wa) : cmulc(y0[j], wa);
          v[d][8 + j] = (SIGN < 0) ? cmul(y1[j], wb) : cmulc(y1[j], wb);
        }
      }
      return;
    }
    const int r0 = t / W;
#pragma unroll
    for (int d = 0; d < D1; ++d) {
      dft<T, R1, SIGN>(v[d]);
      const int n2 = r0 + d * RS1;
      if constexpr (TWF) {
        // w^k1 from w^1, w^2, w^4, w^8 (indices n2 * 2^b < N/2)
        const cpx<T> p1 = tw1[n2], p2 = tw1[2 * n2], p4 = tw1[4 * n2], p8 = tw1[8 * n2];
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) {
          cpx<T> w = (k1 & 1) ? p1 : (k1 & 2) ? p2 : (k1 & 4) ? p4 : p8;
          const int lo = k1 & -k1;
          if ((k1 & 2) && lo < 2) w = cmul(w, p2);
          if ((k1 & 4) && lo < 4) w = cmul(w, p4);
          if ((k1 & 8) && lo < 8) w = cmul(w, p8);
          v[d][k1] = (SIGN < 0) ? cmul(v[d][k1], w) : cmulc(v[d][k1], w);
        }
      } else {
#pragma unroll
        for (int k1 = 1; k1 < R1; ++k1) {
          const cpx<T> w = tw1[n2 * k1];
          v[d][k1] = (SIGN < 0) ? cmul(v[d][k1], w) : cmulc(v[d][k1], w);
        }
      }
    }
  }

  // Values k1 = kc0 .. kc0+KC-1 of the first level into the tile buffer.
  static __device__ __forceinline__ void put(const cpx<T> (&v)[D1][R1], int kc0, int c, int r0,
                                             cpx<T>* buf) {
    if constexpr (SPLIT != 0) {
      // v[d][j]: output q of butterfly A (j < 8) or B; chunks of 8 hold the
      // j with (j & 7) < 4 (q < 8) first
      const int t = threadIdx.x, h = (t >> 5) & 1;
#pragma unroll
      for (int d = 0; d < D1; ++d) {
        const int n2a = (t & ~32) / W + d * RS1, n2b = (t | 32) / W + d * RS1;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int jj = j & 7;
          if (KC == 8 && ((jj < 4) != (kc0 == 0))) continue;
          const int q = 8 * (jj >> 2) + 4 * h + (jj & 3);
          buf[((q - kc0) * M + (j < 8 ? n2a : n2b)) * W + c] = v[d][j];
        }
      }
      return;
    }
#pragma unroll
    for (int d = 0; d < D1; ++d)
#pragma unroll
      for (int kk = 0; kk < KC; ++kk) buf[(kk * M + r0 + d * RS1) * W + c] = v[d][kc0 + kk];
  }

  // Second level of one chunk (M-point DFTs over n2 in LDS) and the stores
  // of X[k1 + R1 k2], k1 in the chunk.  Caller has put the chunk and
  // synchronised; ends after a barrier (buffer free again).
  static __device__ __forceinline__ void second(const FftArgs<T>& a, int64_t tile, int kc0, int c,
                                                int r0, cpx<T>* buf, const cpx<T>* tw2) {
    const int64_t col = tile * W + c;
    const bool ok = col < a.nlines;
    const int64_t ob = (col / a.C) * a.Dout + (col % a.C);
    const int q = r0 % NT, kl = r0 / NT;
    LineAcc<T, W> acc{buf + kl * M * W + c};
    static_fft<T, M, NT, SIGN, W, SUB, 1>(acc, q, tw2, buf, 0);
    if (ok) {
      cpx<T>* dst = reinterpret_cast<cpx<T>*>(a.out) + ob;
      if constexpr (BLK) {
        // rows kc0 + kk + R1 k2 in exchange blocks, one opaque row counter
        const gptr<cpx<T>> g = to_global(dst);
        int row0 = kc0 + R1 * r0;
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) {
          int row = row0;
#pragma unroll
          for (int k2 = r0; k2 < M; k2 += RS1) {
            g[row_off(row, a.Sout, a.oshift, a.Bout)] = buf[(kk * M + k2) * W + c];
            row += R1 * RS1;
            asm volatile("" : "+v"(row));
          }
          row0 += 1;
          asm volatile("" : "+v"(row0));
        }
      } else {
        // running pointers (rows kc0 + kk + R1 * k2), opaque to the
        // compiler so that the row addresses are not all kept in registers
        gptr<cpx<T>> p0 = to_global(dst + (int64_t)(kc0 + R1 * r0) * a.Sout);
        const int64_t step = (int64_t)R1 * RS1 * a.Sout;
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) {
          gptr<cpx<T>> p = p0;
#pragma unroll
          for (int k2 = r0; k2 < M; k2 += RS1) {
            *p = buf[(kk * M + k2) * W + c];
            p += step;
            asm volatile("" : "+v"(p));
          }
          p0 += a.Sout;
          asm volatile("" : "+v"(p0));
        }
      }
    }
    __syncthreads();
  }

  // The whole tile: first level, then the chunks of the second level.
  static __device__ __forceinline__ void run(cpx<T> (&v)[D1][R1], const FftArgs<T>& a,
                                             int64_t tile, int c, int r0, cpx<T>* buf,
                                             const cpx<T>* tw1, const cpx<T>* tw2) {
    first(v, (int)threadIdx.x, tw1);
#pragma unroll
    for (int kc0 = 0; kc0 < R1; kc0 += KC) {
      put(v, kc0, c, r0, buf);
      __syncthreads();
      second(a, tile, kc0, c, r0, buf, tw2);
    }
  }

  // Twiddles into LDS: the loads are issued first (tv1/tv2), the LDS writes
  // after the caller's data loads so that all loads are in flight together.
  static constexpr int E1 = (NTW1 + TPB - 1) / TPB, E2 = (TWM + TPB - 1) / TPB > 0 ? (TWM + TPB - 1) / TPB : 1;
  static __device__ __forceinline__ void tw_load(cpx<T> (&tv1)[E1], cpx<T> (&tv2)[E2],
                                                 const FftArgs<T>& a, const cpx<T>* twsub, int t) {
#pragma unroll
    for (int j = 0; j < E1; ++j)
      if (t + j * TPB < NTW1) tv1[j] = a.tw[t + j * TPB];
#pragma unroll
    for (int j = 0; j < E2; ++j)
      if (t + j * TPB < TWM) tv2[j] = twsub[t + j * TPB];
  }
  static __device__ __forceinline__ void tw_store(const cpx<T> (&tv1)[E1], const cpx<T> (&tv2)[E2],
                                                  cpx<T>* tw1, cpx<T>* tw2, int t) {
#pragma unroll
    for (int j = 0; j < E1; ++j)
      if (t + j * TPB < NTW1) tw1[t + j * TPB] = tv1[j];
#pragma unroll
    for (int j = 0; j < E2; ++j)
      if (t + j * TPB < TWM) tw2[t + j * TPB] = tv2[j];
  }

  // Virtual block vb of a launch of nb blocks -> tile (XCD chunk map, order).
  static __device__ __forceinline__ int64_t tile_of(int64_t vb, int64_t nb, const FftArgs<T>& a) {
    int64_t tile = xcd_tile((unsigned)vb, (unsigned)nb, (unsigned)a.xgroup);
    return a.reverse ? nb - 1 - tile : tile;
  }
};

template <typename T, int N, int SIGN, bool BLK, int V, int SPLIT = 0>
__global__ void __launch_bounds__((S4Cfg<T, N, V>::TPB), (S4Cfg<T, N, V>::WPE))
    sfft4_kernel(FftArgs<T> a, const cpx<T>* __restrict__ twsub) {
  using K = S4Tile<T, N, SIGN, BLK, V, SPLIT>;
  __shared__ cpx<T> buf[K::KC * K::M * K::W];
  __shared__ cpx<T> tw1[K::NTW1];
  __shared__ cpx<T> tw2[K::TWM > 0 ? K::TWM : 1];
  const int t = threadIdx.x;
  const int c = t % K::W, r0 = t / K::W;
  cpx<T> tv1[K::E1], tv2[K::E2];
  K::tw_load(tv1, tv2, a, twsub, t);
  cpx<T> v[K::D1][K::R1];
  if constexpr (S4Cfg<T, N, V>::PB == 0) {
    const int64_t tile = K::tile_of(blockIdx.x, gridDim.x, a);
    K::load(v, a, tile, c, r0);
    K::tw_store(tv1, tv2, tw1, tw2, t);
    __syncthreads();
    K::run(v, a, tile, c, r0, buf, tw1, tw2);
  } else {
    // Persistent: block b starts with virtual block b of a launch of one
    // block per tile (grid G a multiple of 8, so each keeps its XCD) and
    // takes its next ones from the plan's queue: the k-th take of class
    // x = b & 7 is virtual block G + 8 k + x, so the tiles the blocks of an
    // XCD work on at any time stay neighbours (the chunk map of tile_of
    // gives them adjacent columns, and their row segments meet in the same
    // DRAM pages), where a fixed sequence b, b + G, ... drifts apart (an
    // access-pattern copy of the 2048-row tiles: 3.6 TB/s fixed, 4.7 TB/s
    // queued).  Without a queue the fixed sequence is used.  The whole
    // first level goes to LDS in one chunk, after which the registers take
    // the next tile's loads while the second level and the stores of the
    // current tile run.
    static_assert(K::KC == K::R1, "persistent tiles need the whole first level in LDS");
    __shared__ int64_t next_vb;
    unsigned* const q = a.queue;
    const unsigned xc = blockIdx.x & 7u;
    const int64_t ntiles = (a.nlines + K::W - 1) / K::W;
    int64_t vb = blockIdx.x;
    int64_t tile = K::tile_of(vb, ntiles, a);
    K::load(v, a, tile, c, r0);
    K::tw_store(tv1, tv2, tw1, tw2, t);
    __syncthreads();
    for (;;) {
      // the take is issued before the first level so that its latency
      // overlaps it; next_vb is read after the barrier below and not
      // written again before second()'s closing barrier
      unsigned k = 0;
      if (q != nullptr && t == 0) k = atomicAdd(&q[xc * 32], 1u);
      K::first(v, t, tw1);
      K::put(v, 0, c, r0, buf);
      if (q != nullptr && t == 0) next_vb = (int64_t)gridDim.x + 8 * (int64_t)k + xc;
      __syncthreads();
      const int64_t nvb = q != nullptr ? next_vb : vb + gridDim.x;
      const bool more = nvb < ntiles;
      const int64_t ntile = more ? K::tile_of(nvb, ntiles, a) : 0;
      if (more) K::load(v, a, ntile, c, r0);
      K::second(a, tile, 0, c, r0, buf, tw2);
      if (!more) break;
      vb = nvb;
      tile = ntile;
    }
    // Every block has made its last take (its result ended the loop) before
    // it counts itself done; the last one to finish zeroes the queue for
    // the next launch.
    if (q != nullptr && t == 0) {
      if (atomicAdd(&q[QUEUE_DONE], 1u) == gridDim.x - 1) {
        for (int i = 0; i < 8; ++i) atomicExch(&q[i * 32], 0u);
        atomicExch(&q[QUEUE_DONE], 0u);
      }
    }
  }
}

// ====================================================================
// Fused two-level tiles: forward FFT, spectral symbol, inverse FFT
// ====================================================================
// M_MULT (a 3-D symbol, the x pass of laplace and friends) and M_DERIV
// (i w(k) / N along this axis, on a real field viewed as complex pairs of
// adjacent columns) with the whole tile of W columns x N rows resident:
//   forward: x[n2 + M n1] -> registers, R1-point DFTs and twiddles, LDS,
//            M-point DFTs over n2 -> X[k1 + R1 k2] at (k1, k2) in LDS;
//   symbol:  in LDS;
//   inverse: M-point inverse DFTs over k2 -> Y[k1][n2], twiddles
//            exp(+2 pi i k1 n2 / N) and R1-point inverse DFTs over k1 in
//            registers -> x'[n2 + M n1] stored straight to HBM
// (x'[n2 + M n1] = sum_k1 e^{2 pi i k1 n1 / R1} e^{2 pi i k1 n2 / N}
//  sum_k2 X[k1 + R1 k2] e^{2 pi i k2 n2 / M}).  One read and one write of
// the operand, like the one-level fused tile, but without its LDS
// round trips of the first and last levels.
// KC: values of k1 per LDS chunk in the M_DERIV kernel (the M-point levels
// and the symbol run chunk by chunk; each thread keeps its k1 values in
// registers between the levels); WPE: waves per SIMD its register
// allocation must allow.
template <typename T, int N>
struct S4FCfg {
  static constexpr bool OK = false;
  static constexpr int W = 1, TPB = 256, KC = 16, WPE = 1;
};
#define AFFT_S4FCFG(TT, NN, W_, TPB_, KC_, WPE_)        \
  template <>                                           \
  struct S4FCfg<TT, NN> {                               \
    static constexpr bool OK = true;                    \
    static constexpr int W = W_, TPB = TPB_, KC = KC_, WPE = WPE_; \
  };
AFFT_S4FCFG(float, 512, 16, 512, 8, 6)
AFFT_S4FCFG(float, 1024, 8, 512, 8, 6)
AFFT_S4FCFG(float, 2048, 4, 512, 16, 1)
#undef AFFT_S4FCFG
// M_MULT tiles where they differ from the M_DERIV ones: at 512 points
// eight-column tiles of 256 threads, four workgroups per CU instead of two
// (512^3 Poisson solve 1.102 -> 1.055 ms, laplace-family unchanged; the
// derivative passes lose 8 % with them).
template <typename T, int N>
struct S4FMult : S4FCfg<T, N> {};
template <>
struct S4FMult<float, 512> {
  static constexpr bool OK = true;
  static constexpr int W = 8, TPB = 256, KC = 16, WPE = 1;
};
template <typename T, int N, int MODE>
using S4FOf = typename std::conditional<MODE == M_MULT, S4FMult<T, N>, S4FCfg<T, N>>::type;

template <typename T, int N, int MODE>
__global__ void __launch_bounds__((S4FOf<T, N, MODE>::TPB), (MODE == M_DERIV ? S4FOf<T, N, MODE>::WPE : 1))
    sfft4f_kernel(FftArgs<T> a, const cpx<T>* __restrict__ twsub) {
  using CF = S4FOf<T, N, MODE>;
  // M_MULT keeps the whole tile in LDS (chunked, its symbol kinds need more
  // registers than two workgroups per CU allow)
  constexpr int R1 = 16, W = CF::W, TPB = CF::TPB, M = N / R1, KC = MODE == M_DERIV ? CF::KC : R1;
  constexpr int D1 = W * M / TPB, RS1 = TPB / W, NT = TPB / (KC * W);
  static_assert(R1 % KC == 0, "k1 chunks must split evenly");
  constexpr int TWM = sched_twn<M, 0>();
  static_assert(MODE == M_MULT || MODE == M_DERIV, "fused modes only");
  static_assert(D1 >= 1 && D1 * TPB == W * M, "tile must split evenly");
  static_assert(NT >= 1 && NT * KC * W == TPB, "lines must split evenly");
  __shared__ cpx<T> buf[KC * M * W];
  __shared__ cpx<T> tw1[N];
  __shared__ cpx<T> tw2[TWM > 0 ? TWM : 1];
  const int t = threadIdx.x;
  const int c = t % W, r0 = t / W;
  int64_t tile = xcd_tile(blockIdx.x, gridDim.x, (unsigned)a.xgroup);
  if (a.reverse) tile = (int64_t)gridDim.x - 1 - tile;
  const int64_t col = tile * W + c;
  const bool ok = col < a.nlines;
  const int64_t ib = (col / a.C) * a.Din + (col % a.C);
  const int64_t ob = (col / a.C) * a.Dout + (col % a.C);

  constexpr int E1 = (N + TPB - 1) / TPB, E2 = (TWM + TPB - 1) / TPB;
  cpx<T> tv1[E1], tv2[E2 > 0 ? E2 : 1];
#pragma unroll
  for (int j = 0; j < E1; ++j)
    if (t + j * TPB < N) tv1[j] = a.tw[t + j * TPB];
#pragma unroll
  for (int j = 0; j < E2; ++j)
    if (t + j * TPB < TWM) tv2[j] = twsub[t + j * TPB];
  cpx<T> v[D1][R1];
  {
    const cpx<T>* src = reinterpret_cast<const cpx<T>*>(a.in) + ib;
#pragma unroll
    for (int d = 0; d < D1; ++d) {
      gptr<const cpx<T>> p = to_global(src + (int64_t)(r0 + d * RS1) * a.Sin);
      const int64_t step = (int64_t)M * a.Sin;
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) {
        v[d][n1] = ok ? ld_global(p) : cpx<T>{(T)0, (T)0};
        p += step;
        asm volatile("" : "+v"(p));
      }
    }
  }
#pragma unroll
  for (int j = 0; j < E1; ++j)
    if (t + j * TPB < N) tw1[t + j * TPB] = tv1[j];
#pragma unroll
  for (int j = 0; j < E2; ++j)
    if (t + j * TPB < TWM) tw2[t + j * TPB] = tv2[j];
  __syncthreads();

  // forward first level in registers
#pragma unroll
  for (int d = 0; d < D1; ++d) {
    dft<T, R1, -1>(v[d]);
    const int n2 = r0 + d * RS1;
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) v[d][k1] = cmul(v[d][k1], tw1[n2 * k1]);
  }
  const int q = r0 % NT, kl = r0 / NT;  // line (k1 = kc0 + kl, column c), thread q of NT
  LineAcc<T, W> acc{buf + kl * M * W + c};
  // per chunk of KC values of k1: to LDS, forward M-point level, symbol,
  // inverse M-point level, back to the registers they came from (kc0 a
  // compile-time constant, so that v stays in registers)
  static_assert(KC == R1 || KC * 2 == R1, "one or two chunks");
  auto chunk = [&](auto kc0c) {
  constexpr int kc0 = decltype(kc0c)::value;
#pragma unroll
  for (int d = 0; d < D1; ++d)
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) buf[(kk * M + r0 + d * RS1) * W + c] = v[d][kc0 + kk];
  __syncthreads();
  static_fft<T, M, NT, -1, W, 0, 1>(acc, q, tw2, buf, 0);

  // symbol on X[k1 + R1 k2], folded into the first-stage loads of the
  // inverse M-point level (no LDS round trip and barrier of its own):
  // element n of this thread's line (k1 = kc0 + kl, column c) is
  // X[k1 + R1 n]; the column's fixed-axis wavenumbers are this thread's.
  {
    int64_t u = 0, Z = 0;
    if constexpr (MODE == M_MULT) {
      u = col / a.Ch;
      Z = col % a.Ch;
    }
    // one inverse level per symbol kind (the kind is uniform: the switch
    // stays out of the loads).  The grad and laplace-family symbols take a
    // fast form: the wavenumbers of the two fixed axes once per thread, the
    // varying one per element, 32-bit integers and T arithmetic (|k|^2 sums
    // below 2^24 are exact in FP32); the Poisson symbols use apply_sym.
    const int Nk = (int)(a.maxis == 0 ? a.N0 : a.N1), Nf = (int)(a.maxis == 0 ? a.N1 : a.N0);
    const int N2i = (int)a.N2;
    auto nyq0 = [](int x, int n) { int w = x > n / 2 ? x - n : x; return x == n / 2 ? 0 : w; };
    auto plain = [](int x, int n) { return x > n / 2 ? x - n : x; };
    const int wf0 = nyq0((int)u, Nf), wz0 = nyq0((int)Z, N2i);
    const int wf1 = plain((int)u, Nf), wz1 = plain((int)Z, N2i);
    const T sc = (T)a.scale;
    auto sym_elem = [&](auto kind, int k, cpx<T> x) -> cpx<T> {
      constexpr int KIND = decltype(kind)::value;
      if constexpr (MODE == M_DERIV) {
        const T f = sc * (T)wave_sym(k, N);
        return cpx<T>{-f * x.y, f * x.x};
      } else if constexpr (KIND <= SYM_GRAD_Z) {
        // i * w_d * scale, w_d of the varying axis or a fixed one
        const bool on_k = (KIND == SYM_GRAD_X && a.maxis == 0) || (KIND == SYM_GRAD_Y && a.maxis == 1);
        const int w = on_k ? nyq0(k, Nk) : KIND == SYM_GRAD_Z ? wz0 : wf0;
        const T f = sc * (T)w;
        return cpx<T>{-f * x.y, f * x.x};
      } else if constexpr (KIND <= SYM_INV_LAPLACE) {
        const int wk = nyq0(k, Nk);
        const T wave = -(T)(wk * wk + wf0 * wf0 + wz0 * wz0);
        T f;
        if constexpr (KIND == SYM_LAPLACE) f = sc * wave;
        else if constexpr (KIND == SYM_BIHARMONIC) f = (sc * wave) * wave;
        else f = wave == (T)0 ? (T)0 : sc / wave;
        return cpx<T>{f * x.x, f * x.y};
      } else if constexpr (KIND == SYM_INV_BIHARMONIC) {
        const int wk = plain(k, Nk);
        const T wave = (T)(wk * wk + wf1 * wf1 + wz1 * wz1);
        const T f = wave == (T)0 ? sc : (sc / wave) / wave;
        return cpx<T>{f * x.x, f * x.y};
      } else {
        const long X = a.maxis == 0 ? (long)k : (long)u, Y = a.maxis == 0 ? (long)u : (long)k;
        return apply_sym<T>(KIND, X, Y, (long)Z, a.N0, a.N1, a.N2, a.scale, x);
      }
    };
    const int k1 = kc0 + kl;
    auto inverse = [&](auto kind) {
      auto pre = [&](int n, cpx<T> x) { return sym_elem(kind, k1 + R1 * n, x); };
      static_fft<T, M, NT, +1, W, 0, 1>(acc, q, tw2, buf, 0, pre);
    };
    if constexpr (MODE == M_DERIV) {
      inverse(std::integral_constant<int, 0>{});
    } else {
      switch (a.sym) {
        case SYM_GRAD_X: inverse(std::integral_constant<int, SYM_GRAD_X>{}); break;
        case SYM_GRAD_Y: inverse(std::integral_constant<int, SYM_GRAD_Y>{}); break;
        case SYM_GRAD_Z: inverse(std::integral_constant<int, SYM_GRAD_Z>{}); break;
        case SYM_LAPLACE: inverse(std::integral_constant<int, SYM_LAPLACE>{}); break;
        case SYM_BIHARMONIC: inverse(std::integral_constant<int, SYM_BIHARMONIC>{}); break;
        case SYM_INV_LAPLACE: inverse(std::integral_constant<int, SYM_INV_LAPLACE>{}); break;
        case SYM_INV_BIHARMONIC: inverse(std::integral_constant<int, SYM_INV_BIHARMONIC>{}); break;
        case SYM_POISSON_DISCRETE: inverse(std::integral_constant<int, SYM_POISSON_DISCRETE>{}); break;
        default: inverse(std::integral_constant<int, SYM_POISSON_CONTINUOUS>{}); break;
      }
    }
  }
#pragma unroll
  for (int d = 0; d < D1; ++d)
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) v[d][kc0 + kk] = buf[(kk * M + r0 + d * RS1) * W + c];
  __syncthreads();
  };
  chunk(std::integral_constant<int, 0>{});
  if constexpr (KC < R1) chunk(std::integral_constant<int, KC>{});

  // inverse first level: twiddles, R1-point inverse DFTs, store
  cpx<T>* dst = reinterpret_cast<cpx<T>*>(a.out) + ob;
#pragma unroll
  for (int d = 0; d < D1; ++d) {
    const int n2 = r0 + d * RS1;
    cpx<T> w[R1];
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) w[k1] = v[d][k1];
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) w[k1] = cmulc(w[k1], tw1[n2 * k1]);
    dft<T, R1, +1>(w);
    if (ok) {
      gptr<cpx<T>> p = to_global(dst + (int64_t)n2 * a.Sout);
      const int64_t step = (int64_t)M * a.Sout;
      if (MODE == M_DERIV && a.accumulate) {  // uniform: one loop per case
        // the accumulator loads all in flight, then the stores
        cpx<T> o[R1];
        gptr<const cpx<T>> pl = p;
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
          o[n1] = ld_global(pl);
          pl += step;
          asm volatile("" : "+v"(pl));
        }
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
          *p = cpx<T>{o[n1].x + w[n1].x, o[n1].y + w[n1].y};
          p += step;
          asm volatile("" : "+v"(p));
        }
      } else {
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) {
          *p = w[n1];
          p += step;
          asm volatile("" : "+v"(p));
        }
      }
    }
  }
}

// ====================================================================
// z pass fused with the first level of a two-level y transform (slab share)
// ====================================================================
// LDS tiles of the fused z kernels: padded by one element in 16
// (lds_pad16) up to 512 points, swizzled (lds_swz32, no extra LDS) from 1024
// points, whose tiles fill the LDS at two or four blocks per CU: the
// Stockham stage writes of stride R, which otherwise fall on the same
// banks, spread out.  (At 512 points the swizzle measured 2 % slower than
// the padding: z C2R 206 against 202 us.)
template <int NZ>
constexpr int zy_pad() { return NZ <= 512 ? 1 : 2; }

// N1 = R * M along y.  A block owns the rows y = n2 + M n1 (n1 < R) of one
// x-plane: the R real z lines as R/2 packed complex FFTs of length NZ in
// LDS, then, per half-spectrum column k, the R row spectra, their R-point
// DFT over n1 and the twiddle exp(-2 pi i n2 k1 / N1):
//   out[x][k1 M + n2][k] = e^{-2 pi i n2 k1/N1} sum_n1 e^{-2 pi i n1 k1/R} Z_{n2 + M n1}[k].
// The y transform is then finished by M-point FFTs over the consecutive
// rows k1 M + n2 (n2 < M) of each k1, whose outputs are rows k1 + R k2.
// (X[k1 + R k2] = sum_n2 e^{-2 pi i n2 k2/M} e^{-2 pi i n2 k1/N1}
//  sum_n1 x[n2 + M n1] e^{-2 pi i n1 k1/R}.)  Compared with a z pass and a
// y pass over N1-point columns, the traffic is the same and the y pass
// reads and writes rows in wide segments.
// IL (interleaved rows): the outputs go to rows k1 + R n2 instead of
// k1 M + n2, so that the M-point stage over the rows k1 + R n2 (n2 < M)
// writes its outputs k1 + R k2 over the same rows, in place (the 1 x 1
// plans, whose y transform ends in the output array itself).
// a: in / out, idist / odist (row pitches of in and out), Din / Dout (plane
// strides), C = M, nlines = planes * M blocks, tw = the z length's static
// table; twy: the y length's exp(-2 pi i m / N1).  The tile (64 KiB at
// NZ = 2048, R = 8) and the staged z twiddles (16 KiB) fit two blocks per
// CU only unpadded (with a padded tile and the twiddles read from global
// memory instead, the twiddle loads held 40 more VGPRs and both kernels
// ran 6-11 % slower).
template <typename T, int NZ, int R, bool IL = false>
__global__ void __launch_bounds__((R / 2 * sched_minnb<NZ, 0>()))
    zy_r2c_kernel(FftArgs<T> a, const cpx<T>* __restrict__ twy) {
  constexpr int NT = sched_minnb<NZ, 0>(), LINES = R / 2, TPB = LINES * NT;
  constexpr int NH = NZ / 2 + 1, EP = LINES * NZ / TPB, TWN = sched_twn<NZ, 0>();
  constexpr int ZP = zy_pad<NZ>();
  auto P = [](int e) { return ZP == 1 ? lds_pad16(e) : lds_swz_tile<NZ>(e); };
  __shared__ cpx<T> buf[LINES * NZ + (ZP == 1 ? LINES * NZ / 16 : 0)];
  __shared__ cpx<T> s_tw[TWN];
  const int t = threadIdx.x;
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x, (unsigned)a.xgroup);
  const int64_t M = a.C, x = blk / M, n2 = blk % M;
  const T* src = reinterpret_cast<const T*>(a.in) + x * a.Din + n2 * a.idist;
  const int64_t pair = 2 * M * a.idist;  // rows n2 + M (2p) and n2 + M (2p + 1)
  cpx<T> v[EP];
#pragma unroll
  for (int j = 0; j < EP; ++j) {
    const int e = t + j * TPB, p = e / NZ, i = e % NZ;
    const T* ra = src + p * pair;
    v[j] = cpx<T>{ra[i], ra[M * a.idist + i]};
  }
  stage_twiddles<T, TWN, TPB>(s_tw, a.tw, t);
#pragma unroll
  for (int j = 0; j < EP; ++j) buf[P(t + j * TPB)] = v[j];
  __syncthreads();
  {
    const int line = t / NT, q = t % NT;
    LineAcc<T, 1, ZP> acc{buf + P(line * NZ)};
    static_fft<T, NZ, NT, -1, 1, 0, LINES>(acc, q, s_tw, buf, NZ);
  }
  cpx<T> w[R];
#pragma unroll
  for (int k1 = 0; k1 < R; ++k1) w[k1] = twy[n2 * k1];
  // row k1 + R n2 (IL) or k1 M + n2
  cpx<T>* dst = reinterpret_cast<cpx<T>*>(a.out) + x * a.Dout + n2 * (IL ? R : 1) * a.odist;
  const int64_t rstep = (IL ? 1 : M) * a.odist;
  const T h = (T)0.5;
  for (int k = t; k < NH; k += TPB) {
    cpx<T> u[R];
#pragma unroll
    for (int p = 0; p < LINES; ++p) {
      const cpx<T> z1 = buf[P(p * NZ + k)];
      const cpx<T> z2 = buf[P(p * NZ + (k == 0 ? 0 : NZ - k))];
      u[2 * p] = cpx<T>{h * (z1.x + z2.x), h * (z1.y - z2.y)};
      u[2 * p + 1] = cpx<T>{h * (z1.y + z2.y), h * (z2.x - z1.x)};
    }
    dft<T, R, -1>(u);
#pragma unroll
    for (int k1 = 0; k1 < R; ++k1) dst[k1 * rstep + k] = k1 ? cmul(u[k1], w[k1]) : u[k1];
  }
}

// The inverse: rows k1 M + n2 (IL: k1 + R n2) of the y transform's M-point stage (k1 < R)
// -> twiddles exp(+2 pi i n2 k1 / N1) and R-point inverse DFTs over k1 ->
// the half spectra of the rows n2 + M n1, packed in pairs into complex
// lines (imaginary parts of DC and Nyquist dropped, as c2r does) -> NZ-point
// inverse FFTs -> real rows.  a: in / out, idist / odist, Din / Dout, C = M,
// nlines, tw as above.
template <typename T, int NZ, int R, bool IL = false>
__global__ void __launch_bounds__((R / 2 * sched_minnb<NZ, 0>()))
    zy_c2r_kernel(FftArgs<T> a, const cpx<T>* __restrict__ twy) {
  constexpr int NT = sched_minnb<NZ, 0>(), LINES = R / 2, TPB = LINES * NT;
  constexpr int NH = NZ / 2 + 1, EP = LINES * NZ / TPB, TWN = sched_twn<NZ, 0>();
  constexpr int ZP = zy_pad<NZ>();
  auto P = [](int e) { return ZP == 1 ? lds_pad16(e) : lds_swz_tile<NZ>(e); };
  __shared__ cpx<T> buf[LINES * NZ + (ZP == 1 ? LINES * NZ / 16 : 0)];
  __shared__ cpx<T> s_tw[TWN];
  const int t = threadIdx.x;
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x, (unsigned)a.xgroup);
  const int64_t M = a.C, x = blk / M, n2 = blk % M;
  // row k1 + R n2 (IL) or k1 M + n2
  const cpx<T>* src = reinterpret_cast<const cpx<T>*>(a.in) + x * a.Din + n2 * (IL ? R : 1) * a.idist;
  const int64_t rstep = (IL ? 1 : M) * a.idist;
  cpx<T> w[R];
#pragma unroll
  for (int k1 = 0; k1 < R; ++k1) w[k1] = twy[n2 * k1];
  // all of the thread's loads in flight before the first LDS write
  constexpr int KI = (NH + TPB - 1) / TPB;
  cpx<T> uu[KI][R];
#pragma unroll
  for (int ki = 0; ki < KI; ++ki) {
    const int k = t + ki * TPB;
    if (k < NH) {
#pragma unroll
      for (int k1 = 0; k1 < R; ++k1) uu[ki][k1] = src[k1 * rstep + k];
    }
  }
#pragma unroll
  for (int ki = 0; ki < KI; ++ki) {
    const int k = t + ki * TPB;
    if (k >= NH) break;
    cpx<T>(&u)[R] = uu[ki];
#pragma unroll
    for (int k1 = 1; k1 < R; ++k1) u[k1] = cmulc(u[k1], w[k1]);
    dft<T, R, +1>(u);
#pragma unroll
    for (int p = 0; p < LINES; ++p) {
      cpx<T> A = u[2 * p], B = u[2 * p + 1];
      const bool edge = k == 0 || 2 * k == NZ;
      if (edge) {
        A.y = (T)0;
        B.y = (T)0;
      }
      buf[P(p * NZ + k)] = cpx<T>{A.x - B.y, A.y + B.x};
      if (!edge) buf[P(p * NZ + NZ - k)] = cpx<T>{A.x + B.y, B.x - A.y};
    }
  }
  stage_twiddles<T, TWN, TPB>(s_tw, a.tw, t);
  __syncthreads();
  {
    const int line = t / NT, q = t % NT;
    LineAcc<T, 1, ZP> acc{buf + P(line * NZ)};
    static_fft<T, NZ, NT, +1, 1, 0, LINES>(acc, q, s_tw, buf, NZ);
  }
  T* dst = reinterpret_cast<T*>(a.out) + x * a.Dout + n2 * a.odist;
  const int64_t pair = 2 * M * a.odist;
#pragma unroll
  for (int j = 0; j < EP; ++j) {
    const int e = t + j * TPB, p = e / NZ, i = e % NZ;
    const cpx<T> z = buf[P(e)];
    T* ra = dst + p * pair;
    ra[i] = z.x;
    ra[M * a.odist + i] = z.y;
  }
}

// ====================================================================
// Element-wise kernels
// ====================================================================
// out = symbol(kind) * scale * in over a complex block [osize] at global
// offset ostart (src/operators_cuda.cu:15-205; operators.txx:276-392).
template <typename T>
__global__ void __launch_bounds__(256)
    mult_kernel(const cpx<T>* __restrict__ in, cpx<T>* __restrict__ out, int kind, long N0,
                long N1, long N2, long os0, long os1, long os2, long st0, long st1, long st2,
                double scale) {
  const int64_t total = (int64_t)os0 * os1 * os2;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total;
       p += (int64_t)gridDim.x * blockDim.x) {
    const long k = (long)(p % os2);
    const long j = (long)((p / os2) % os1);
    const long i = (long)(p / ((int64_t)os1 * os2));
    out[p] = apply_sym<T>(kind, i + st0, j + st1, k + st2, N0, N1, N2, scale, in[p]);
  }
}

// y += alpha * x (daxpy_cu, src/operators_cuda.cu:207-215)
template <typename T>
__global__ void __launch_bounds__(256)
    axpy_kernel(int64_t n, T alpha, const T* __restrict__ x, T* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] += alpha * x[i];
}

}  // namespace afft

// fft_launch.hpp -- host-side selection of an FFT kernel instance and its
// launch shape.  Included once per precision (fft_f32.hip, fft_f64.hip).
#pragma once
#include <utility>
#include <vector>

#include "fft_kernels.hpp"
#include "fft_long.hpp"

namespace afft {

// Measurement overrides (tile widths, two-level variants, XCD chunks, pass
// order) are read from the environment only in builds made with
// ACCFFT_AMD_SWEEP defined (ACCFFT_AMD_SWEEP=1 python -m accfft_amd.build);
// the shipped library ignores the environment and always uses the built-in
// rules and the plan's MEASURE choices.
inline const char* sweep_env(const char* name) {
#ifdef ACCFFT_AMD_SWEEP
  return getenv(name);
#else
  (void)name;
  return nullptr;
#endif
}

// Twiddle tables for one transform length, resident in device memory.
template <typename T>
struct AxisTab {
  int n = 0;
  bool is_static = false;
  DynSched sc{};
  const cpx<T>* tw_static = nullptr;  // per-stage table (static engine)
  const cpx<T>* tw_full = nullptr;    // exp(-2 pi i m/n), m < n
  const cpx<T>* tw_split = nullptr;   // per-stage table of the split schedule
  const cpx<T>* tw_sub = nullptr;     // per-stage table of length n/16 (two-level tiles)
  const cpx<T>* tw_subw = nullptr;    // the same for the SCHED_WIDE schedule
  // strided C2C configuration picked by the plan's MEASURE tuning: -1 the
  // built-in rule, -2 one-level tiles, v >= 0 two-level variant S4Cfg<T, n, v>
  int s4v = -1;
  // tiles per XCD chunk picked by the tuning (xcd_tile's g); -1 the rule
  int xg = -1;
  // the plan's tile queue of the persistent strided tiles (QUEUE_WORDS
  // zeroed counters, see sfft4_kernel); nullptr = fixed tile sequences
  unsigned* queue = nullptr;
  // lines longer than one LDS tile: global radix passes (fft_long.hpp)
  bool longl = false;
  // long lines with a prime factor above 61: Bluestein's algorithm over
  // power-of-two convolutions of length bm (0: not used); bchirp =
  // exp(-i pi n^2 / n), bker[0 / 1] = the convolution kernel's spectrum / bm
  // for the forward / inverse transform, btw and bsc the bm-point tables
  int bm = 0;
  DynSched bsc{};
  const cpx<T>* bchirp = nullptr;
  const cpx<T>* bker[2] = {nullptr, nullptr};
  const cpx<T>* btw = nullptr;
};

// Lengths with a matrix-core (split-precision) schedule.
inline bool split_length(int n) { return n >= 256 && n <= 4096 && (n & (n - 1)) == 0; }
inline bool split_mode(int mode) { return mode == M_C2C || mode == M_R2C || mode == M_C2R; }

// FP32 pass with radix-16 stages on the matrix cores (fft_split.hip).
hipError_t launch_fft_split(bool strided, int mode, int sign, FftArgs<float> a,
                            const AxisTab<float>& tab, int split, hipStream_t st);

inline bool static_length(int n) {
  return n >= 16 && n <= 4096 && (n & (n - 1)) == 0;
}

inline bool pair_mode(int mode) { return mode == M_R2C || mode == M_C2R || mode == M_DERIV; }

inline int dyn_lines(int n) {
  int l = 2048 / n;
  if (l < 1) l = 1;
  if (l > 64) l = 64;
  return l;
}
// Strided dynamic tiles: the widest power-of-two column count whose buffer
// stays within 32 KiB (4096 FP32 / 2048 FP64 elements; two buffers per
// workgroup).  FP64 tiles of 4096 elements (128 KiB per workgroup, one per
// CU) ran 768^3 in 22.4 ms against 18.9 ms.
inline int dyn_cols(int n, size_t csz) {
  int w = 1;
  while (w * 2 <= 64 && (size_t)w * 2 * n * csz <= 32768) w *= 2;
  return w;
}
inline size_t dyn_z_lds(int n, int lpw, size_t csz) { return 2 * (size_t)lpw * n * csz; }
inline size_t dyn_s_lds(int n, int w, size_t csz) { return 2 * (size_t)w * n * csz; }

template <typename T, int N, int ZT = 256>
hipError_t launch_z_n(int mode, int sign, FftArgs<T>& a, int64_t grid, size_t shm,
                      hipStream_t st) {
  dim3 g((unsigned)grid), b(N > 0 ? ZCfg<(N > 0 ? N : 16), 0, ZT>::TPB : 256);
  switch (mode) {
    case M_C2C:
      if (sign < 0) hipLaunchKernelGGL((zfft_kernel<T, N, M_C2C, -1, 0, ZT>), g, b, shm, st, a);
      else hipLaunchKernelGGL((zfft_kernel<T, N, M_C2C, +1, 0, ZT>), g, b, shm, st, a);
      break;
    case M_R2C: hipLaunchKernelGGL((zfft_kernel<T, N, M_R2C, -1, 0, ZT>), g, b, shm, st, a); break;
    case M_C2R: hipLaunchKernelGGL((zfft_kernel<T, N, M_C2R, +1, 0, ZT>), g, b, shm, st, a); break;
    case M_DERIV: hipLaunchKernelGGL((zfft_kernel<T, N, M_DERIV, -1, 0, ZT>), g, b, shm, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, int N, int LDSK = 64>
hipError_t launch_s_n(int mode, int sign, FftArgs<T>& a, int64_t grid, size_t shm,
                      hipStream_t st) {
  dim3 g((unsigned)grid), b(N > 0 ? 256 : dyn_s_tpb<T>());
  switch (mode) {
    case M_C2C:
      if (sign < 0) hipLaunchKernelGGL((sfft_kernel<T, N, M_C2C, -1, 0, LDSK>), g, b, shm, st, a);
      else hipLaunchKernelGGL((sfft_kernel<T, N, M_C2C, +1, 0, LDSK>), g, b, shm, st, a);
      break;
    case M_R2C: hipLaunchKernelGGL((sfft_kernel<T, N, M_R2C, -1, 0, LDSK>), g, b, shm, st, a); break;
    case M_C2R: hipLaunchKernelGGL((sfft_kernel<T, N, M_C2R, +1, 0, LDSK>), g, b, shm, st, a); break;
    case M_DERIV: hipLaunchKernelGGL((sfft_kernel<T, N, M_DERIV, -1, 0, LDSK>), g, b, shm, st, a); break;
    case M_MULT: hipLaunchKernelGGL((sfft_kernel<T, N, M_MULT, -1, 0, LDSK>), g, b, shm, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Raise the dynamic-LDS cap of every dynamic-engine kernel once.
template <typename T>
inline void dyn_lds_attrs() {
  static bool done = false;
  if (done) return;
  done = true;
  const int cap = 160 * 1024;
  (void)hipFuncSetAttribute((const void*)zfft_kernel<T, 0, M_C2C, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  (void)hipFuncSetAttribute((const void*)zfft_kernel<T, 0, M_C2C, +1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  (void)hipFuncSetAttribute((const void*)zfft_kernel<T, 0, M_R2C, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  (void)hipFuncSetAttribute((const void*)zfft_kernel<T, 0, M_C2R, +1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  (void)hipFuncSetAttribute((const void*)zfft_kernel<T, 0, M_DERIV, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  (void)hipFuncSetAttribute((const void*)sfft_kernel<T, 0, M_C2C, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  (void)hipFuncSetAttribute((const void*)sfft_kernel<T, 0, M_C2C, +1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  (void)hipFuncSetAttribute((const void*)sfft_kernel<T, 0, M_R2C, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  (void)hipFuncSetAttribute((const void*)sfft_kernel<T, 0, M_C2R, +1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  (void)hipFuncSetAttribute((const void*)sfft_kernel<T, 0, M_DERIV, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  (void)hipFuncSetAttribute((const void*)sfft_kernel<T, 0, M_MULT, -1>, hipFuncAttributeMaxDynamicSharedMemorySize, cap);
}

// Strided C2C tiles of half the LDS (32 KiB: twice the workgroups per CU,
// half-width row segments) for the lengths whose bit log2(n) is set in
// ACCFFT_AMD_SNARROW (read once).  Default: 128 and 256, whose passes are
// too short to fill the chip twice over with 64 KiB tiles (FP32 256^3:
// strided passes 38 -> 28 us, round trip 0.206 -> 0.172 ms; 128^3 0.072 ->
// 0.064 ms; FP64 tiles are the same either way), for passes over at most
// narrow_max_bytes: on large operands the wider row segments win (the
// 256-point y passes of the 2048-point slab share: 262 -> 249 us).
constexpr size_t narrow_max_bytes = (size_t)256 << 20;
inline bool narrow_tiles(int n) {
  static const unsigned mask = [] {
    const char* e = sweep_env("ACCFFT_AMD_SNARROW");
    return e ? (unsigned)strtoul(e, nullptr, 0) : (1u << 7) | (1u << 8);
  }();
  if (n != 128 && n != 256 && n != 512 && n != 1024 && n != 2048) return false;
  int l = 0;
  while ((1 << l) < n) ++l;
  return (mask >> l) & 1u;
}

// Strided C2C tiles of 128 KiB of LDS (one workgroup per CU, twice the row
// segment) for the lengths whose bit log2(n) is set in ACCFFT_AMD_SWIDE.
inline bool wide_tiles(int n) {
  static const unsigned mask = [] {
    const char* e = sweep_env("ACCFFT_AMD_SWIDE");
    return e ? (unsigned)strtoul(e, nullptr, 0) : 0u;
  }();
  if (n != 1024 && n != 2048 && n != 4096) return false;
  int l = 0;
  while ((1 << l) < n) ++l;
  return (mask >> l) & 1u;
}

// Two-level strided C2C tiles (sfft4_kernel): on by default where they were
// measured faster on MI355X (FP32 512, 1024, 2048; FP64 512 is faster with
// the one-level tiles).  For measurement, ACCFFT_AMD_S4 (bit log2(n): use
// them) overrides that choice and ACCFFT_AMD_S4V ("n:v,n:v", e.g.
// "2048:2,1024:3") picks the configuration variant of a length.
inline unsigned env_mask(const char* name, unsigned dflt) {
  const char* e = sweep_env(name);
  return e ? (unsigned)strtoul(e, nullptr, 0) : dflt;
}
inline int ilog2i(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}
template <typename T>
inline bool four_step_default(int n) {
  static const unsigned mask =
      env_mask("ACCFFT_AMD_S4", sizeof(T) == 4 ? (1u << 9) | (1u << 10) | (1u << 11) : 0u);
  return (mask >> ilog2i(n)) & 1u;
}
inline bool four_step_env_set() {
  static const bool set = sweep_env("ACCFFT_AMD_S4") != nullptr;
  return set;
}
// Two-level tiles for this axis: ACCFFT_AMD_S4 if set, else the tuned
// choice, else the built-in rule.
template <typename T>
inline bool four_step(const AxisTab<T>& tab) {
  if (!four_step_env_set() && tab.s4v != -1) return tab.s4v >= 0;
  return four_step_default<T>(tab.n);
}
// Variant of the two-level configuration for length n (S4Cfg<T, n, V>);
// blk: rows in exchange blocks (the slab plan's y pass); tuned: the plan's
// MEASURE choice (-1 none).  ACCFFT_AMD_S4V overrides; the built-in rule is
// the persistent tiles (variant 2, tile queue) for 2048-point rows, plain
// or in exchange blocks, and 1024-point plain rows, variant 0 otherwise
// (the 2048 x pass of the 256 x 2048 x 2048 slab share: forward 5.84 ms
// against 6.47 ms with variant 0; the 2048 y pass of a 2048^3 rank on a
// 4 x 2 pencil grid: forward 6.50 against 6.81 ms; 1024^3 round trip 9.78
// against 10.11 ms).
inline int four_step_rule(int n, bool blk) { return (n == 2048 || (n == 1024 && !blk)) ? 2 : 0; }
inline int four_step_variant(int n, bool blk, int tuned = -1) {
  static const std::vector<std::pair<int, int>> tab = [] {
    std::vector<std::pair<int, int>> t;
    const char* e = sweep_env("ACCFFT_AMD_S4V");
    while (e && *e) {
      char* end = nullptr;
      const long len = strtol(e, &end, 10);
      if (end == e || *end != ':') break;
      e = end + 1;
      const long v = strtol(e, &end, 10);
      if (end == e) break;
      t.emplace_back((int)len, (int)v);
      e = *end == ',' ? end + 1 : end;
    }
    return t;
  }();
  for (const auto& kv : tab)
    if (kv.first == n) return kv.second;
  if (tuned >= 0) return tuned;
  return four_step_rule(n, blk);
}

// Whether S4Cfg<T, n, v> exists (the candidates of the MEASURE tuning).
template <typename T>
inline bool s4_config_ok(int n, int v) {
#define AFFT_S4OK(NN)                                                        \
  if (n == NN) switch (v) {                                                  \
      case 0: return S4Cfg<T, NN, 0>::OK;                                    \
      case 1: return S4Cfg<T, NN, 1>::OK;                                    \
      case 2: return S4Cfg<T, NN, 2>::OK;                                    \
      case 3: return S4Cfg<T, NN, 3>::OK;                                    \
      default: return false;                                                 \
    }
  AFFT_S4OK(256)
  AFFT_S4OK(512)
  AFFT_S4OK(1024)
  AFFT_S4OK(2048)
#undef AFFT_S4OK
  return false;
}

// Compute units of the current device (read once per process).
inline int device_cus() {
  static const int cus = [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) {
      (void)hipGetLastError();
      return 256;
    }
    return n;
  }();
  return cus;
}

template <typename T, int N, int V, int SPLIT = 0>
hipError_t launch_s4v(int sign, FftArgs<T>& a, const AxisTab<T>& tab, hipStream_t st) {
  if constexpr (!S4Cfg<T, N, V>::OK) {
    return hipErrorNotSupported;
  } else {
    using CF = S4Cfg<T, N, V>;
    const int64_t tiles = (a.nlines + CF::W - 1) / CF::W;
    a.tw = tab.tw_full;
    a.lpw = CF::W;
    int64_t grid = tiles;
    a.queue = nullptr;
    if constexpr (CF::PB > 0) {
      // persistent: PB workgroups per CU, a multiple of 8 (one per XCD);
      // the queue hands out the tiles beyond the first grid's worth (every
      // class of blockIdx & 7 must have workgroups)
      const int64_t cap = (int64_t)device_cus() * CF::PB;
      if (grid > cap) grid = cap >= 8 ? cap / 8 * 8 : cap;
      static const bool queued = [] {
        const char* e = sweep_env("ACCFFT_AMD_QUEUE");  // 0: fixed sequences (sweep builds)
        return !e || atoi(e) != 0;
      }();
      if (queued && grid < tiles && grid % 8 == 0) a.queue = tab.queue;
    }
    dim3 g((unsigned)grid), b(CF::TPB);
    const bool blk = a.ishift != 0 || a.oshift != 0;
    const cpx<T>* twsub = CF::SUB == SCHED_WIDE ? tab.tw_subw : tab.tw_sub;
#define AFFT_S4(S, B) hipLaunchKernelGGL((sfft4_kernel<T, N, S, B, V, SPLIT>), g, b, 0, st, a, twsub)
    if (blk) {
      if (sign < 0) AFFT_S4(-1, true);
      else AFFT_S4(+1, true);
    } else {
      if (sign < 0) AFFT_S4(-1, false);
      else AFFT_S4(+1, false);
    }
#undef AFFT_S4
    return hipGetLastError();
  }
}

template <typename T, int N>
hipError_t launch_s4(int sign, FftArgs<T>& a, const AxisTab<T>& tab, hipStream_t st) {
  const bool blk = a.ishift != 0 || a.oshift != 0;
  int v = four_step_variant(N, blk, tab.s4v);
  if (!s4_config_ok<T>(N, v)) v = four_step_rule(N, blk);  // an override with no such configuration
  switch (v) {
    case 1: return launch_s4v<T, N, 1>(sign, a, tab, st);
    case 2: return launch_s4v<T, N, 2>(sign, a, tab, st);
    case 3: return launch_s4v<T, N, 3>(sign, a, tab, st);
    default: return launch_s4v<T, N, 0>(sign, a, tab, st);
  }
}

// Fused two-level tiles for a strided M_MULT / M_DERIV pass (FP32 512,
// 1024, 2048); hipErrorNotSupported sends the caller to the one-level tile.
// M_DERIV runs on the complex view of adjacent real columns, which needs
// even column counts, strides and 8-byte aligned pointers.
template <typename T>
hipError_t launch_s4f(int mode, FftArgs<T> a, const AxisTab<T>& tab, hipStream_t st) {
  if (mode == M_DERIV) {
    if ((a.nlines | a.C | a.Din | a.Dout | a.Sin | a.Sout) & 1) return hipErrorNotSupported;
    if (((uintptr_t)a.in | (uintptr_t)a.out) % sizeof(cpx<T>)) return hipErrorNotSupported;
    a.nlines /= 2;
    a.C /= 2;
    a.Din /= 2;
    a.Dout /= 2;
    a.Sin /= 2;
    a.Sout /= 2;
  }
  a.tw = tab.tw_full;
#define AFFT_S4F(NN)                                                                        \
  if (tab.n == NN) {                                                                        \
    if constexpr (S4FCfg<T, NN>::OK) {                                                      \
      if (mode == M_MULT) {                                                                 \
        using CF = S4FOf<T, NN, M_MULT>;                                                    \
        a.lpw = CF::W;                                                                      \
        dim3 g((unsigned)((a.nlines + CF::W - 1) / CF::W)), b(CF::TPB);                     \
        hipLaunchKernelGGL((sfft4f_kernel<T, NN, M_MULT>), g, b, 0, st, a, tab.tw_sub);    \
      } else {                                                                              \
        using CF = S4FOf<T, NN, M_DERIV>;                                                   \
        a.lpw = CF::W;                                                                      \
        dim3 g((unsigned)((a.nlines + CF::W - 1) / CF::W)), b(CF::TPB);                     \
        hipLaunchKernelGGL((sfft4f_kernel<T, NN, M_DERIV>), g, b, 0, st, a, tab.tw_sub);   \
      }                                                                                     \
      return hipGetLastError();                                                             \
    }                                                                                       \
  }
  AFFT_S4F(512)
  AFFT_S4F(1024)
  AFFT_S4F(2048)
#undef AFFT_S4F
  return hipErrorNotSupported;
}

// Launch one axis pass.  strided = false: contiguous lines (z layout).
// Returns hipErrorNotSupported when the length does not fit one LDS tile.
template <typename T>
hipError_t launch_fft(bool strided, int mode, int sign, FftArgs<T> a, const AxisTab<T>& tab,
                      hipStream_t st) {
  const int n = tab.n;
  a.sc = tab.sc;
  a.sc.n = n;
  const int64_t units = pair_mode(mode) ? (a.nlines + 1) / 2 : a.nlines;
  if (units <= 0) return hipSuccess;
  const size_t csz = sizeof(cpx<T>);
  if (tab.is_static) {
    a.tw = tab.tw_static;
    int per;
    if (strided && (mode == M_MULT || mode == M_DERIV) && tab.tw_sub && four_step<T>(tab)) {
      const hipError_t e = launch_s4f<T>(mode, a, tab, st);
      if (e != hipErrorNotSupported) return e;
    }
    if (strided && mode == M_C2C && tab.tw_sub && four_step<T>(tab)) {
      switch (n) {
        case 256: return launch_s4<T, 256>(sign, a, tab, st);
        case 512: return launch_s4<T, 512>(sign, a, tab, st);
        case 1024: return launch_s4<T, 1024>(sign, a, tab, st);
        case 2048: return launch_s4<T, 2048>(sign, a, tab, st);
        default: break;
      }
    }
#ifdef ACCFFT_AMD_SWEEP
    // 128 KiB one-level tiles, one workgroup per CU (measured slower:
    // 1024^3 y passes 2.19 -> 2.98 ms); sweep builds only
    if (sizeof(T) == 4 && strided && mode == M_C2C && wide_tiles(n)) {
      const int w = n == 1024 ? SCfg<T, 1024, 0, 128>::W
                  : n == 2048 ? SCfg<T, 2048, 0, 128>::W : SCfg<T, 4096, 0, 128>::W;
      a.lpw = w;
      const int64_t grid = (units + w - 1) / w;
      dim3 g((unsigned)grid), b(256);
#define AFFT_WIDE(NN)                                                                     \
  if constexpr (sizeof(T) == 4)                                                           \
    if (n == NN) {                                                                        \
      if (sign < 0) hipLaunchKernelGGL((sfft_kernel<T, NN, M_C2C, -1, 0, 128>), g, b, 0, st, a); \
      else hipLaunchKernelGGL((sfft_kernel<T, NN, M_C2C, +1, 0, 128>), g, b, 0, st, a);    \
      return hipGetLastError();                                                           \
    }
      AFFT_WIDE(1024)
      AFFT_WIDE(2048)
      AFFT_WIDE(4096)
#undef AFFT_WIDE
    }
#endif
    if (strided && mode == M_C2C && n >= 512 && narrow_tiles(n)) {
      const int w = n == 512 ? SCfg<T, 512, 0, 32>::W
                  : n == 1024 ? SCfg<T, 1024, 0, 32>::W : SCfg<T, 2048, 0, 32>::W;
      a.lpw = w;
      const int64_t grid = (units + w - 1) / w;
      dim3 g((unsigned)grid), b(256);
#define AFFT_NARROW(NN)                                                                  \
  if (n == NN) {                                                                         \
    if (sign < 0) hipLaunchKernelGGL((sfft_kernel<T, NN, M_C2C, -1, 0, 32>), g, b, 0, st, a); \
    else hipLaunchKernelGGL((sfft_kernel<T, NN, M_C2C, +1, 0, 32>), g, b, 0, st, a);      \
    return hipGetLastError();                                                            \
  }
      AFFT_NARROW(512)
      AFFT_NARROW(1024)
      AFFT_NARROW(2048)
#undef AFFT_NARROW
    }
#define AFFT_CASE(NN)                                                              \
  case NN:                                                                         \
    if constexpr (NN == 128 || NN == 256) {                                        \
      if (strided && narrow_tiles(NN) && units * NN * csz <= narrow_max_bytes) {   \
        per = SCfg<T, NN, 0, 32>::W;                                               \
        a.lpw = per;                                                               \
        return launch_s_n<T, NN, 32>(mode, sign, a, (units + per - 1) / per, 0, st); \
      }                                                                            \
    }                                                                              \
    per = strided ? SCfg<T, NN>::W : ZCfg<NN>::LPW;                                \
    a.lpw = per;                                                                   \
    return strided ? launch_s_n<T, NN>(mode, sign, a, (units + per - 1) / per, 0, st) \
                   : launch_z_n<T, NN>(mode, sign, a, (units + per - 1) / per, 0, st);
    switch (n) {
      AFFT_CASE(16)
      AFFT_CASE(32)
      AFFT_CASE(64)
      AFFT_CASE(128)
      AFFT_CASE(256)
      AFFT_CASE(512)
      AFFT_CASE(1024)
      AFFT_CASE(2048)
      AFFT_CASE(4096)
      default: return hipErrorInvalidValue;
    }
#undef AFFT_CASE
  }
  a.tw = tab.tw_full;
  dyn_lds_attrs<T>();
  if (strided) {
    const int w = dyn_cols(n, csz);
    const size_t shm = dyn_s_lds(n, w, csz);
    if (shm > 160 * 1024) return hipErrorNotSupported;
    a.lpw = w;
    return launch_s_n<T, 0>(mode, sign, a, (units + w - 1) / w, shm, st);
  } else {
    const int l = dyn_lines(n);
    const size_t shm = dyn_z_lds(n, l, csz);
    if (shm > 160 * 1024) return hipErrorNotSupported;
    a.lpw = l;
    return launch_z_n<T, 0>(mode, sign, a, (units + l - 1) / l, shm, st);
  }
}

// Fused z pass and first radix-8 level of the y transform (zy_r2c_kernel /
// zy_c2r_kernel): one block per (plane, n2).  il: interleaved output rows
// k1 + R n2 (1 x 1 plans) instead of k1 M + n2 (the slab share's
// workspace).  FP32, z lengths 256 ... 2048.  (FP64 was measured slower
// fused: 512^3 round trip 2.44 ms against 2.39 ms with separate passes,
// 256^3 0.326 against 0.309 ms; its 512-point tile and twiddles take
// 40 KiB of LDS per 256 threads, four workgroups per CU.)
constexpr int ZY_R = 8;
template <typename T>
inline bool zy_length(int nz) {
  return sizeof(T) == 4 && (nz == 256 || nz == 512 || nz == 1024 || nz == 2048);
}
// 1 x 1 r2c plans whose z pass can carry the first radix-8 level of the y
// transform: FP32, a z length of the fused kernels and N1 = 8 M with a
// static M-point second level (M = 16 ... 256).  Split plans keep their
// matrix-core z stages.
template <typename T>
inline bool zy_plain_ok(int kind, int split, const int* n) {
  if (kind != 1 || split || n[1] % ZY_R) return false;
  const int m = n[1] / ZY_R;
  return zy_length<T>(n[2]) && m >= 16 && m <= 256 && (m & (m - 1)) == 0;
}
template <typename T, int NZ>
hipError_t launch_zy_n(int sign, bool il, const FftArgs<T>& a, const cpx<T>* twy, hipStream_t st) {
  dim3 g((unsigned)a.nlines), b(ZY_R / 2 * sched_minnb<NZ, 0>());
  if (il) {
    if (sign < 0) hipLaunchKernelGGL((zy_r2c_kernel<T, NZ, ZY_R, true>), g, b, 0, st, a, twy);
    else hipLaunchKernelGGL((zy_c2r_kernel<T, NZ, ZY_R, true>), g, b, 0, st, a, twy);
  } else {
    if (sign < 0) hipLaunchKernelGGL((zy_r2c_kernel<T, NZ, ZY_R, false>), g, b, 0, st, a, twy);
    else hipLaunchKernelGGL((zy_c2r_kernel<T, NZ, ZY_R, false>), g, b, 0, st, a, twy);
  }
  return hipGetLastError();
}
template <typename T>
hipError_t launch_zy(int sign, int nz, bool il, const FftArgs<T>& a, const cpx<T>* twy,
                     hipStream_t st) {
  if constexpr (sizeof(T) != 4) {
    return hipErrorNotSupported;
  } else {
    if (a.nlines <= 0) return hipSuccess;
    switch (nz) {
      case 256: return launch_zy_n<T, 256>(sign, il, a, twy, st);
      case 512: return launch_zy_n<T, 512>(sign, il, a, twy, st);
      case 1024: return launch_zy_n<T, 1024>(sign, il, a, twy, st);
      case 2048: return launch_zy_n<T, 2048>(sign, il, a, twy, st);
      default: return hipErrorNotSupported;
    }
  }
}

template <typename T>
hipError_t launch_mult(const cpx<T>* in, cpx<T>* out, int kind, const long N[3],
                       const long os[3], const long ost[3], double scale, hipStream_t st) {
  const int64_t total = (int64_t)os[0] * os[1] * os[2];
  if (total <= 0) return hipSuccess;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((mult_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, in, out, kind,
                     N[0], N[1], N[2], os[0], os[1], os[2], ost[0], ost[1], ost[2], scale);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_axpy(int64_t n, T alpha, const T* x, T* y, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((axpy_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, n, alpha, x, y);
  return hipGetLastError();
}

}  // namespace afft

// exec.hpp -- the execute paths of a plan, per precision T.
//
// Stage order follows src/accfft_gpu.cpp on a 1x1 grid:
//   r2c (:471-539):  z R2C (in -> out), y C2C in place, x C2C in place
//   c2r (:540-603):  x C2C, y C2C (in place on the input, which the
//                    reference overwrites), z C2R (-> out)
//   c2c (:943-1136): same order with a complex z pass
// with the y pass issued as one launch instead of N0 (the reference loops
// cufftExecZ2Z per x slab, :501-509).  The transposes between stages are
// no-ops at P = 1 (src/transpose_gpu.cpp:3819-3821, :1164-1179) and are
// not issued.
#pragma once
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "plan.hpp"

namespace afft {

inline int status_of(hipError_t e) {
  if (e == hipSuccess) return ACCFFT_AMD_OK;
  if (e == hipErrorNotSupported) return ACCFFT_AMD_ERR_UNSUPPORTED;
  if (e == hipErrorOutOfMemory) return ACCFFT_AMD_ERR_ALLOC;
  return ACCFFT_AMD_ERR_HIP;
}

#define AFFT_TRY(x)                         \
  do {                                      \
    hipError_t e_ = (x);                    \
    if (e_ != hipSuccess) return status_of(e_); \
  } while (0)

inline int ensure_ws(accfft_amd_plan* p, int i, size_t bytes) {
  if (p->ws_bytes[i] >= bytes) return ACCFFT_AMD_OK;
  if (p->ws[i]) {
    (void)hipStreamSynchronize(p->stream);
    (void)hipFree(p->ws[i]);
    p->ws[i] = nullptr;
    p->ws_bytes[i] = 0;
  }
  if (hipMalloc(&p->ws[i], bytes) != hipSuccess) {
    (void)hipGetLastError();
    return ACCFFT_AMD_ERR_ALLOC;
  }
  p->ws_bytes[i] = bytes;
  return ACCFFT_AMD_OK;
}

// Tiles per XCD chunk for a strided pass: rows whose stride is not a
// multiple of 128 bytes share cache lines between neighbouring tiles, so
// they go to one XCD in chunks of 32 (512^3 FP32 y passes 216/220 us with
// one contiguous range per XCD, 203/213 with chunks of 16, 201/211 with 32,
// 209/223 with 64: smaller chunks re-read the lines shared across chunk
// edges, larger ones spread the XCDs' concurrent tiles over more DRAM
// pages); aligned rows in chunks of 64.
// ACCFFT_AMD_XCD_GROUP_{MISALIGNED,ALIGNED} override (1 = round robin,
// 0 = one contiguous range per XCD).
inline int env_int(const char* name, int dflt) {
  const char* e = sweep_env(name);
  return e ? atoi(e) : dflt;
}
// Blocks per XCD chunk of a z pass: consecutive blocks read and write
// adjacent rows (2056-byte rows of the r2c half spectrum share lines at
// block edges), so chunks keep those edges in one L2.  512^3 FP32 z R2C /
// C2R 196/207 us round robin, 183/201 with chunks of 16, 180/197 with 32,
// 188/199 with 128.  ACCFFT_AMD_XCD_GROUP_Z overrides; 1 = round robin.
inline int xcd_group_z() {
  static const int g = env_int("ACCFFT_AMD_XCD_GROUP_Z", 32);
  return g;
}
inline int xcd_group_for(int64_t row_stride_bytes) {
  static const int g_mis = env_int("ACCFFT_AMD_XCD_GROUP_MISALIGNED", 32);
  static const int g_al = env_int("ACCFFT_AMD_XCD_GROUP_ALIGNED", 64);
  return (row_stride_bytes % 128) ? g_mis : g_al;
}
// Pass order: bit k of the mask reverses pass k of a 3-D transform
// (forward: 0 z, 1 y, 2 x; inverse: 3 x, 4 y, 5 z).  A pass that starts
// where the previous one ended reads what is still in the 256 MB
// memory-side cache.  Default: forward y and inverse x and y walk their
// tiles last-to-first (512^3 FP32 round trip 1.30 -> 1.28 ms; the inverse
// x pass alone 234 -> 207 us with one-level tiles).
// ACCFFT_AMD_REVERSE overrides the mask.
inline bool reverse_pass(int k) {
  static const int m = env_int("ACCFFT_AMD_REVERSE", (1 << 1) | (1 << 3) | (1 << 4));
  return (m >> k) & 1;
}

template <typename T>
inline const char* mode_name(int mode, int sign) {
  switch (mode) {
    case M_C2C: return sign < 0 ? "c2c_fwd" : "c2c_inv";
    case M_R2C: return "r2c";
    case M_C2R: return "c2r";
    case M_MULT: return "mult";
    default: return "deriv";
  }
}

template <typename F>
inline int prof_launch(accfft_amd_plan* p, const char* name, double bytes, F&& launch);

// Launches one pass; with profiling on, brackets it with events labelled
// "<z|s>fft_<mode>_n<N>_<tag>" and records its algorithmic bytes
// (elements read + elements written, each once).
template <typename T>
inline int long_pass(accfft_amd_plan* p, bool strided, int mode, int sign, const FftArgs<T>& a,
                     const AxisTab<T>& tab);

template <typename T>
inline int run_pass(accfft_amd_plan* p, const char* tag, bool strided, int mode, int sign,
                    FftArgs<T> a, const AxisTab<T>& tab, int order = -1) {
  if (tab.longl) return long_pass<T>(p, strided, mode, sign, a, tab);
  if (order >= 0) a.reverse = reverse_pass(order) ? 1 : 0;
  if (strided) {
    const bool real_in = (mode == M_R2C || mode == M_DERIV);
    const int64_t rs = a.Sin * (int64_t)(real_in ? sizeof(T) : 2 * sizeof(T));
    a.xgroup = tab.xg >= 0 ? tab.xg : xcd_group_for(rs);
  } else {
    a.xgroup = xcd_group_z();
  }
  auto launch = [&]() -> hipError_t {
    if constexpr (sizeof(T) == 4) {
      if (p->split && tab.tw_split && split_mode(mode))
        return launch_fft_split(strided, mode, sign, a, tab, p->split, p->stream);
    }
    return launch_fft<T>(strided, mode, sign, a, tab, p->stream);
  };
  if (!p->prof) return status_of(launch());
  const int64_t n = tab.n;
  const int64_t nh = n / 2 + 1;
  const double rs = sizeof(T), cs = 2.0 * sizeof(T);
  double bytes;
  switch (mode) {
    case M_R2C: bytes = a.nlines * (n * rs + nh * cs); break;
    case M_C2R: bytes = a.nlines * (nh * cs + n * rs); break;
    case M_DERIV: bytes = a.nlines * (2.0 * n * rs) * (a.accumulate ? 1.5 : 1.0); break;
    default: bytes = a.nlines * (2.0 * n * cs); break;
  }
  char name[96];
  const bool split = sizeof(T) == 4 && p->split && tab.tw_split && split_mode(mode);
  static const char* sname[4] = {"", "_bf16x2", "_bf16x3", "_fp16x2"};
  snprintf(name, sizeof(name), "%cfft_%s_n%d_%s_%s%s", strided ? 's' : 'z',
           mode_name<T>(mode, sign), (int)n, tag, sizeof(T) == 4 ? "f32" : "f64",
           split ? sname[p->split & 3] : "");
  return prof_launch(p, name, bytes, launch);
}

// Runs launch() between profiling events recorded under `name`.
template <typename F>
inline int prof_launch(accfft_amd_plan* p, const char* name, double bytes, F&& launch) {
  int kind = -1;
  for (size_t i = 0; i < p->prof_kinds.size(); ++i)
    if (p->prof_kinds[i].name == name) kind = (int)i;
  if (kind < 0) {
    accfft_amd_plan::ProfKind k;
    k.name = name;
    p->prof_kinds.push_back(k);
    kind = (int)p->prof_kinds.size() - 1;
  }
  p->prof_kinds[kind].bytes = bytes;
  hipEvent_t e0 = prof_event(p), e1 = prof_event(p);
  (void)hipEventRecord(e0, p->stream);
  const int st = status_of(launch());
  (void)hipEventRecord(e1, p->stream);
  p->prof_pending.push_back({kind, e0, e1});
  if (p->prof_pending.size() > 4096) return st ? st : prof_collect(p);
  return st;
}

// A pass over lines longer than one LDS tile (tab.longl): complex lines of
// any axis, and the real-to-complex / complex-to-real z passes (two real
// lines per complex FFT), as global radix passes (fft_long.hpp) through
// workspaces 2 and 3.  The fused modes, strided real passes, exchange-block
// maps and the adjoint weights have no long-line form.
// nlines long lines from (src, gs) to (dst, gd) (src may equal dst): the
// radix passes of long_c2c through ws, or Bluestein's algorithm in chunks
// of lines whose two convolution buffers (workspaces 4, 5) hold at most
// 256 MiB each.
template <typename T>
inline int long_lines(accfft_amd_plan* p, const AxisTab<T>& tab, const cpx<T>* src, GGeom gs,
                      cpx<T>* dst, GGeom gd, cpx<T>* ws, int64_t nlines, int sign,
                      int lines_fast) {
  using C = cpx<T>;
  if (!tab.bm)
    return status_of(long_c2c<T>(src, gs, dst, gd, ws, nlines, tab.sc, tab.tw_full, sign,
                                 lines_fast, p->stream));
  const int N = tab.n, M = tab.bm;
  int64_t chunk = ((int64_t)256 << 20) / ((int64_t)M * (int64_t)sizeof(C));
  if (chunk < 1) chunk = 1;
  if (chunk > nlines) chunk = nlines;
  int st;
  if ((st = ensure_ws(p, 4, (size_t)chunk * M * sizeof(C))) ||
      (st = ensure_ws(p, 5, (size_t)chunk * M * sizeof(C))))
    return st;
  C* a = (C*)p->ws[4];
  C* w = (C*)p->ws[5];
  const GGeom ga = {1, M, 1};
  const dim3 b(GTPB);
  for (int64_t l0 = 0; l0 < nlines; l0 += chunk) {
    const int64_t nl = nlines - l0 < chunk ? nlines - l0 : chunk;
    hipLaunchKernelGGL((blue_pre_kernel<T>), dim3(gblocks(nl * M)), b, 0, p->stream, src, gs, a,
                       l0, nl, N, M, tab.bchirp, sign);
    if ((st = status_of(hipGetLastError()))) return st;
    if ((st = status_of(long_c2c<T>(a, ga, a, ga, w, nl, tab.bsc, tab.btw, -1, 0, p->stream))))
      return st;
    hipLaunchKernelGGL((blue_mul_kernel<T>), dim3(gblocks(nl * M)), b, 0, p->stream, a, nl, M,
                       tab.bker[sign > 0 ? 1 : 0]);
    if ((st = status_of(hipGetLastError()))) return st;
    if ((st = status_of(long_c2c<T>(a, ga, a, ga, w, nl, tab.bsc, tab.btw, +1, 0, p->stream))))
      return st;
    hipLaunchKernelGGL((blue_post_kernel<T>), dim3(gblocks(nl * N)), b, 0, p->stream, a, dst, gd,
                       l0, nl, N, M, tab.bchirp, sign);
    if ((st = status_of(hipGetLastError()))) return st;
  }
  return ACCFFT_AMD_OK;
}

template <typename T>
inline int long_pass(accfft_amd_plan* p, bool strided, int mode, int sign, const FftArgs<T>& a,
                     const AxisTab<T>& tab) {
  using C = cpx<T>;
  const int N = tab.n;
  if (a.ishift || a.oshift || a.adjw || a.nlines <= 0)
    return a.nlines <= 0 ? ACCFFT_AMD_OK : ACCFFT_AMD_ERR_UNSUPPORTED;
  int st;
  if (mode == M_C2C) {
    GGeom gs, gd;
    if (strided) {
      gs = {a.C, a.Din, a.Sin};
      gd = {a.C, a.Dout, a.Sout};
    } else {
      gs = {1, a.idist, 1};
      gd = {1, a.odist, 1};
    }
    const int64_t L = a.nlines - 1;
    const int64_t extent = (L / gd.C) * gd.D + L % gd.C + (int64_t)(N - 1) * gd.S + 1;
    if (!tab.bm && (st = ensure_ws(p, 2, (size_t)extent * sizeof(C)))) return st;
    return long_lines<T>(p, tab, (const C*)a.in, gs, (C*)a.out, gd, (C*)p->ws[2], a.nlines, sign,
                         strided ? 1 : 0);
  }
  if (strided || (mode != M_R2C && mode != M_C2R)) return ACCFFT_AMD_ERR_UNSUPPORTED;
  // z passes: line pairs packed into complex lines of length N (pitch N)
  const int64_t np = (a.nlines + 1) / 2;
  const size_t wb = (size_t)np * N * sizeof(C);
  if ((!tab.bm && (st = ensure_ws(p, 2, wb))) || (st = ensure_ws(p, 3, wb))) return st;
  C* z = (C*)p->ws[3];
  const GGeom gz = {1, N, 1};
  const dim3 b(GTPB);
  if (mode == M_R2C) {
    hipLaunchKernelGGL((gpack_r2c_kernel<T>), dim3(gblocks(np * N)), b, 0, p->stream,
                       (const T*)a.in, a.idist, z, a.nlines, N);
    if ((st = status_of(hipGetLastError()))) return st;
    if ((st = long_lines<T>(p, tab, z, gz, z, gz, (C*)p->ws[2], np, -1, 0))) return st;
    hipLaunchKernelGGL((gsplit_r2c_kernel<T>), dim3(gblocks(np * (N / 2 + 1))), b, 0, p->stream,
                       z, (C*)a.out, a.odist, a.nlines, N);
    return status_of(hipGetLastError());
  }
  hipLaunchKernelGGL((gpack_c2r_kernel<T>), dim3(gblocks(np * (N / 2 + 1))), b, 0, p->stream,
                     (const C*)a.in, a.idist, z, a.nlines, N);
  if ((st = status_of(hipGetLastError()))) return st;
  if ((st = long_lines<T>(p, tab, z, gz, z, gz, (C*)p->ws[2], np, +1, 0))) return st;
  hipLaunchKernelGGL((gunpack_c2r_kernel<T>), dim3(gblocks(np * N)), b, 0, p->stream, z,
                     (T*)a.out, a.odist, a.nlines, N);
  return status_of(hipGetLastError());
}

template <typename T>
struct Ops {
  using C = cpx<T>;

  static const AxisTab<T>& tab(const accfft_amd_plan* p, int ax);

  static FftArgs<T> zargs(const void* in, void* out, int64_t nlines, int64_t idist,
                          int64_t odist) {
    FftArgs<T> a{};
    a.in = in;
    a.out = out;
    a.nlines = nlines;
    a.idist = idist;
    a.odist = odist;
    a.C = 1;
    return a;
  }
  static FftArgs<T> sargs(const void* in, void* out, int64_t G, int64_t Cc, int64_t Din,
                          int64_t Dout, int64_t Sin, int64_t Sout) {
    FftArgs<T> a{};
    a.in = in;
    a.out = out;
    a.nlines = G;
    a.C = Cc;
    a.Din = Din;
    a.Dout = Dout;
    a.Sin = Sin;
    a.Sout = Sout;
    return a;
  }

  // ---------------------------------------------------------- building blocks
  // z pass of the r2c layout: real [N0][N1][N2 (or 2Nh)] <-> [N0][N1][Nh]
  static int z_r2c(accfft_amd_plan* p, const void* in, void* out, bool padded) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    auto a = zargs(in, out, N0 * N1, padded ? 2 * Nh : N2, Nh);
    return run_pass<T>(p, "z", false, M_R2C, -1, a, tab(p, 2), 0);
  }
  static int z_c2r(accfft_amd_plan* p, const void* in, void* out, bool padded) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    auto a = zargs(in, out, N0 * N1, Nh, padded ? 2 * Nh : N2);
    return run_pass<T>(p, "z", false, M_C2R, +1, a, tab(p, 2), 5);
  }
  // y pass over a complex [N0][N1][L] array (L = Nh for r2c, N2 for c2c)
  static FftArgs<T> y_args(const accfft_amd_plan* p, const void* in, void* out, int64_t L) {
    const int64_t N0 = p->n[0], N1 = p->n[1];
    return sargs(in, out, N0 * L, L, N1 * L, N1 * L, L, L);
  }
  // x pass over a complex [N0][N1][L] array
  static FftArgs<T> x_args(const accfft_amd_plan* p, const void* in, void* out, int64_t L) {
    const int64_t N1 = p->n[1];
    return sargs(in, out, N1 * L, N1 * L, 0, 0, N1 * L, N1 * L);
  }
  static int y_c2c(accfft_amd_plan* p, const void* in, void* out, int64_t L, int sign) {
    return run_pass<T>(p, "y", true, M_C2C, sign, y_args(p, in, out, L), tab(p, 1),
                       sign < 0 ? 1 : 4);
  }
  static int x_c2c(accfft_amd_plan* p, const void* in, void* out, int64_t L, int sign) {
    return run_pass<T>(p, "x", true, M_C2C, sign, x_args(p, in, out, L), tab(p, 0),
                       sign < 0 ? 2 : 3);
  }
  // forward FFT along x (axis 0) or y (1), symbol, inverse FFT as three
  // passes: the fused pass for axes longer than one LDS tile
  static int mult_unfused(accfft_amd_plan* p, int axis, const void* in, void* out, int64_t L,
                          int sym, double scale) {
    int st;
    if ((st = axis == 0 ? x_c2c(p, in, out, L, -1) : y_c2c(p, in, out, L, -1))) return st;
    const long N[3] = {p->n[0], p->n[1], p->n[2]}, os[3] = {p->n[0], p->n[1], (long)L},
               ost[3] = {0, 0, 0};
    if ((st = status_of(launch_mult<T>((const C*)out, (C*)out, sym, N, os, ost, scale, p->stream))))
      return st;
    return axis == 0 ? x_c2c(p, out, out, L, +1) : y_c2c(p, out, out, L, +1);
  }
  // forward x, symbol, inverse x, fused (rows along x)
  static int x_mult(accfft_amd_plan* p, const void* in, void* out, int64_t L, int sym,
                    double scale) {
    if (tab(p, 0).longl) return mult_unfused(p, 0, in, out, L, sym, scale);
    auto a = x_args(p, in, out, L);
    a.sym = sym;
    a.scale = scale;
    a.N0 = p->n[0];
    a.N1 = p->n[1];
    a.N2 = p->n[2];
    a.Ch = L;
    a.maxis = 0;
    return run_pass<T>(p, "x", true, M_MULT, -1, a, tab(p, 0));
  }
  static int y_mult(accfft_amd_plan* p, const void* in, void* out, int64_t L, int sym,
                    double scale) {
    if (tab(p, 1).longl) return mult_unfused(p, 1, in, out, L, sym, scale);
    auto a = y_args(p, in, out, L);
    a.sym = sym;
    a.scale = scale;
    a.N0 = p->n[0];
    a.N1 = p->n[1];
    a.N2 = p->n[2];
    a.Ch = L;
    a.maxis = 1;
    return run_pass<T>(p, "y", true, M_MULT, -1, a, tab(p, 1));
  }
  // fused per-axis derivative of a real field (operators_gpu.txx:172-278)
  // The per-axis derivative along an axis longer than one LDS tile, which
  // has no fused form: z r2c, the axis's multiplier i w(k) (along z) or a
  // forward FFT, multiplier and inverse FFT along x / y, z c2r; out (+)=.
  // Along x / y this is the reference's per-axis r2c -> multiplier -> c2r
  // (the z transform pair cancels) when the axis length is even; for odd
  // lengths the rule that zeroes k = N/2 (integer division) keeps its
  // negative partner in a full spectrum but not in the per-axis half
  // spectrum, so odd long x / y axes are refused.
  static int long_deriv(accfft_amd_plan* p, int axis, const void* in, void* out, int accumulate,
                        double sign) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    if (axis != 2 && (p->n[axis] & 1)) return ACCFFT_AMD_ERR_UNSUPPORTED;
    int st;
    if ((st = ensure_ws(p, 0, (size_t)N0 * N1 * Nh * sizeof(C)))) return st;
    if (accumulate && (st = ensure_ws(p, 1, (size_t)N0 * N1 * N2 * sizeof(T)))) return st;
    void* w = p->ws[0];
    if ((st = z_r2c(p, in, w, false))) return st;
    if (axis == 2) {
      const long N[3] = {p->n[0], p->n[1], p->n[2]}, os[3] = {p->n[0], p->n[1], (long)Nh},
                 ost[3] = {0, 0, 0};
      if ((st = status_of(launch_mult<T>((const C*)w, (C*)w, SYM_GRAD_Z, N, os, ost,
                                         sign / (double)N2, p->stream))))
        return st;
    } else {
      const double scale = sign / ((double)p->n[axis] * N2);
      if ((st = mult_unfused(p, axis, w, w, Nh, SYM_GRAD_X + axis, scale))) return st;
    }
    void* dst = accumulate ? p->ws[1] : out;
    if ((st = z_c2r(p, w, dst, false))) return st;
    if (!accumulate) return ACCFFT_AMD_OK;
    return status_of(launch_axpy<T>(N0 * N1 * N2, (T)1, (const T*)dst, (T*)out, p->stream));
  }
  static int deriv(accfft_amd_plan* p, int axis, const void* in, void* out, int accumulate,
                   double sign = 1.0) {
    if (tab(p, axis).longl) return long_deriv(p, axis, in, out, accumulate, sign);
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2];
    FftArgs<T> a;
    bool strided = true;
    if (axis == 2) {
      a = zargs(in, out, N0 * N1, N2, N2);
      strided = false;
    } else if (axis == 1) {
      a = sargs(in, out, N0 * N2, N2, N1 * N2, N1 * N2, N2, N2);
    } else {
      a = sargs(in, out, N1 * N2, N1 * N2, 0, 0, N1 * N2, N1 * N2);
    }
    a.scale = sign / (double)p->n[axis];
    a.accumulate = accumulate;
    return run_pass<T>(p, axis == 0 ? "x" : axis == 1 ? "y" : "z", strided, M_DERIV, -1, a, tab(p, axis));
  }

  // ---------------------------------------------------------- public paths
  static int r2c(accfft_amd_plan* p, const void* in, void* out, unsigned xyz) {
    const int64_t Nh = p->n[2] / 2 + 1;
    void* o = out;
    int st;
    if ((xyz & 6) == 6 && yfused(p) && !p->inplace) {
      if ((st = zy_fwd(p, in, out))) return st;
      return (xyz & 1) ? x_c2c(p, out, out, Nh, -1) : ACCFFT_AMD_OK;
    }
    if (xyz & 4) {
      if ((st = z_r2c(p, in, out, p->inplace))) return st;
    } else {
      o = const_cast<void*>(in);  // src/accfft_gpu.cpp:488-489
    }
    if (xyz & 2)
      if ((st = y_c2c(p, o, o, Nh, -1))) return st;
    if (xyz & 1)
      if ((st = x_c2c(p, o, o, Nh, -1))) return st;
    return ACCFFT_AMD_OK;
  }

  static int c2r(accfft_amd_plan* p, void* in, void* out, unsigned xyz) {
    const int64_t N0 = p->n[0], N1 = p->n[1], Nh = p->n[2] / 2 + 1;
    const bool keep = (p->flags & ACCFFT_AMD_PRESERVE_INPUT) && (xyz & 4) && (xyz & 3);
    void* cur = in;
    int st;
    if (keep) {
      if ((st = ensure_ws(p, 0, (size_t)N0 * N1 * Nh * sizeof(C)))) return st;
    }
    if (xyz & 1) {
      void* dst = keep ? p->ws[0] : cur;
      if ((st = x_c2c(p, cur, dst, Nh, +1))) return st;
      cur = dst;
    }
    if ((xyz & 6) == 6 && yfused(p) && !p->inplace)
      return zy_inv(p, cur, (keep && cur == in) ? p->ws[0] : cur, out);
    if (xyz & 2) {
      void* dst = (keep && cur == in) ? p->ws[0] : cur;
      if ((st = y_c2c(p, cur, dst, Nh, +1))) return st;
      cur = dst;
    }
    if (xyz & 4) return z_c2r(p, cur, out, p->inplace);
    return ACCFFT_AMD_OK;  // result stays in the input (src/accfft_gpu.cpp:602-603)
  }

  static int c2c(accfft_amd_plan* p, int dir, void* in, void* out, unsigned xyz) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2];
    int st;
    if (dir < 0) {
      void* o = out;
      if (xyz & 4) {
        auto a = zargs(in, out, N0 * N1, N2, N2);
        if ((st = run_pass<T>(p, "z", false, M_C2C, -1, a, tab(p, 2)))) return st;
      } else {
        o = in;
      }
      if (xyz & 2)
        if ((st = y_c2c(p, o, o, N2, -1))) return st;
      if (xyz & 1)
        if ((st = x_c2c(p, o, o, N2, -1))) return st;
      return ACCFFT_AMD_OK;
    }
    const bool keep = (p->flags & ACCFFT_AMD_PRESERVE_INPUT) && (xyz & 4) && (xyz & 3);
    void* cur = in;
    if (keep)
      if ((st = ensure_ws(p, 0, (size_t)N0 * N1 * N2 * sizeof(C)))) return st;
    if (xyz & 1) {
      void* dst = keep ? p->ws[0] : cur;
      if ((st = x_c2c(p, cur, dst, N2, +1))) return st;
      cur = dst;
    }
    if (xyz & 2) {
      void* dst = (keep && cur == in) ? p->ws[0] : cur;
      if ((st = y_c2c(p, cur, dst, N2, +1))) return st;
      cur = dst;
    }
    if (xyz & 4) {
      auto a = zargs(cur, out, N0 * N1, N2, N2);
      return run_pass<T>(p, "z", false, M_C2C, +1, a, tab(p, 2));
    }
    return ACCFFT_AMD_OK;
  }

  // src/accfft_gpu.cpp:1212-1604
  static int r2c_axis(accfft_amd_plan* p, int axis, const void* in, void* out) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2];
    if (axis == 2) return z_r2c(p, in, out, false);
    FftArgs<T> a;
    if (axis == 1) a = sargs(in, out, N0 * N2, N2, N1 * N2, (N1 / 2 + 1) * N2, N2, N2);
    else a = sargs(in, out, N1 * N2, N1 * N2, 0, 0, N1 * N2, N1 * N2);
    return run_pass<T>(p, axis == 0 ? "x" : "y", true, M_R2C, -1, a, tab(p, axis));
  }
  static int c2r_axis(accfft_amd_plan* p, int axis, const void* in, void* out) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2];
    if (axis == 2) return z_c2r(p, in, out, false);
    FftArgs<T> a;
    if (axis == 1) a = sargs(in, out, N0 * N2, N2, (N1 / 2 + 1) * N2, N1 * N2, N2, N2);
    else a = sargs(in, out, N1 * N2, N1 * N2, 0, 0, N1 * N2, N1 * N2);
    return run_pass<T>(p, axis == 0 ? "x" : "y", true, M_C2R, +1, a, tab(p, axis));
  }

  // ---------------------------------------------------------- slab share
  // One rank's share of an r2c on a P x 1 grid (src/accfft_gpu.cpp:471-539
  // with T_plan_2 of src/transpose_gpu.cpp:3809-4021 between the y and x
  // passes).  Input [N0/P][N1][N2] real, output [N0][N1/P][Nh]; the pack
  // (local_transpose_col into [P][N0/P][N1/P][Nh], transpose_gpu.cpp:3889)
  // is the y pass's store, the exchange is the plan's hook (identity when
  // none: the packed send buffer is the receive buffer), the x pass reads
  // the received [N0][N1/P][Nh] in place.
  // Calls the plan's exchange hook; in a timed call its span is recorded
  // for timer[2] (the reference's communication time, doc/timing.md).
  static int exchange(accfft_amd_plan* p, const void* send, void* recv, int64_t block, int dir) {
    return comm_span(p, [&] {
      return p->exchange(send, recv, block, p->nprocs, dir, p->stream, p->exchange_user);
    });
  }
  // Runs one exchange hook call; in a timed call its span is bracketed by
  // the next pair of the plan's exchange events (summed into timer[2]).
  template <typename F>
  static int comm_span(accfft_amd_plan* p, F&& call) {
    const int k = p->comm_pairs;
    const bool rec = p->timing && k < accfft_amd_plan::NEVX / 2;
    if (rec) {
      for (int e = 2 * k; e < 2 * k + 2; ++e)
        if (!p->evx[e] && hipEventCreate(&p->evx[e]) != hipSuccess) return ACCFFT_AMD_ERR_HIP;
      (void)hipEventRecord(p->evx[2 * k], p->stream);
    }
    const int st = call();
    if (rec) {
      (void)hipEventRecord(p->evx[2 * k + 1], p->stream);
      p->comm_pairs = k + 1;
    }
    return st;
  }

  // Row pitch of the plan-internal z-y workspace: Nh complex rounded up to
  // 128 bytes when that costs at most 2 % more memory, so that the y pass's
  // row segments do not straddle cache lines (the user-visible layouts keep
  // the reference's pitch Nh).
  static int64_t ws_pitch(int64_t Nh) {
    const int64_t q = 128 / (int64_t)sizeof(C);
    const int64_t np = (Nh + q - 1) / q * q;
    return (np - Nh) * 50 <= Nh ? np : Nh;
  }
  // Slab shares with 2048-point y and z axes (FP32): the z pass carries the
  // first radix-8 level of the y transform (zy_r2c_kernel / zy_c2r_kernel)
  // and the y pass is ZY_R launches of 256-point FFTs over consecutive
  // rows, which read and write wide row segments instead of the 64-byte
  // segments of 2048-row tiles.  ACCFFT_AMD_ZY=0 (sweep builds) keeps the
  // separate passes.
  static bool zy_fused(const accfft_amd_plan* p) {
    if constexpr (sizeof(T) != 4) {
      return false;
    } else {
      static const bool on = env_int("ACCFFT_AMD_ZY", 1) != 0;
      const int64_t n1 = p->n[1] / p->nprocs;
      return on && p->kind == 3 && p->tf_ysub.n == 2048 / ZY_R &&
             (p->nprocs == 1 || (n1 >= 2 * ZY_R && (n1 & (n1 - 1)) == 0));
    }
  }
  static const AxisTab<T>& ysub(const accfft_amd_plan* p) {
    if constexpr (sizeof(T) == 4) return p->tf_ysub;
    else return p->td_ysub;
  }
  // il: interleaved rows k1 + 8 n2 (1 x 1 plans) or k1 M + n2 (slab)
  static int run_zy(accfft_amd_plan* p, int sign, bool il, const FftArgs<T>& a) {
    auto launch = [&]() {
      return launch_zy<T>(sign, p->n[2], il, a, tab(p, 1).tw_full, p->stream);
    };
    if (!p->prof) return status_of(launch());
    const double n = (double)p->n[2], nh = (double)(p->n[2] / 2 + 1);
    const double bytes = (double)a.nlines * ZY_R * (n * sizeof(T) + nh * 2 * sizeof(T));
    char name[64];
    snprintf(name, sizeof(name), "zyfft_%s_n%d_z_%s", sign < 0 ? "r2c" : "c2r", p->n[2],
             sizeof(T) == 4 ? "f32" : "f64");
    return prof_launch(p, name, bytes, launch);
  }
  // ---- 1 x 1 r2c plans with the y transform's first level in the z pass
  // (p->yfuse): z R2C + radix-8 level over n1 of y = n2 + M n1 into rows
  // k1 + 8 n2, then the M-point stage over the rows k1 + 8 n2 (n2 < M) of
  // each k1, whose outputs k1 + 8 k2 land on the same rows (in place).
  // Same traffic as separate z and y passes; the y stage reads and writes
  // rows in segments of 64 columns instead of 16, and the z pass writes
  // (reads, inverse) 8 consecutive rows per block.  Not for in-place plans
  // (the z pass would overwrite rows other blocks still read).
  static bool yfused(const accfft_amd_plan* p) { return p->kind == 1 && p->yfuse; }
  static FftArgs<T> y2_args(const accfft_amd_plan* p, const void* in, void* out) {
    const int64_t N0 = p->n[0], N1 = p->n[1], Nh = p->n[2] / 2 + 1, R = ZY_R;
    return sargs(in, out, N0 * R * Nh, R * Nh, N1 * Nh, N1 * Nh, R * Nh, R * Nh);
  }
  // real [N0][N1][N2] -> half spectrum [N0][N1][Nh] transformed along z and y
  static int zy_fwd(accfft_amd_plan* p, const void* in, void* out) {
    const int64_t N2 = p->n[2], Nh = N2 / 2 + 1;
    int st;
    if ((st = run_zy(p, -1, true, zy_args(p, in, out, p->n[0], N2, Nh)))) return st;
    return run_pass<T>(p, "y", true, M_C2C, -1, y2_args(p, out, out), ysub(p), 1);
  }
  // inverse y stage from in to mid (may be in), then z C2R from mid to out
  static int zy_inv(accfft_amd_plan* p, const void* in, void* mid, void* out) {
    const int64_t N2 = p->n[2], Nh = N2 / 2 + 1;
    int st;
    if ((st = run_pass<T>(p, "y", true, M_C2C, +1, y2_args(p, in, mid), ysub(p), 4))) return st;
    return run_zy(p, +1, true, zy_args(p, mid, out, p->n[0], Nh, N2));
  }
  static FftArgs<T> zy_args(const accfft_amd_plan* p, const void* in, void* out, int64_t n0,
                            int64_t idist, int64_t odist) {
    const int64_t N1 = p->n[1], M = N1 / ZY_R;
    FftArgs<T> a{};
    a.in = in;
    a.out = out;
    a.idist = idist;
    a.odist = odist;
    a.Din = N1 * idist;
    a.Dout = N1 * odist;
    a.C = M;
    a.nlines = n0 * M;
    a.tw = tab(p, 2).tw_static;
    a.xgroup = xcd_group_z();
    return a;
  }

  static int r2c_slab(accfft_amd_plan* p, const void* in, void* out) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    const int P = p->nprocs;
    const int64_t n0 = N0 / P, n1 = N1 / P, Np = ws_pitch(Nh);
    const size_t cb = (size_t)n0 * N1 * Nh * sizeof(C);
    int st;
    if ((st = ensure_ws(p, 0, (size_t)n0 * N1 * Np * sizeof(C)))) return st;
    if (p->exchange && (st = ensure_ws(p, 1, cb))) return st;
    void* zbuf = p->ws[0];
    void* send = p->exchange ? p->ws[1] : out;
    if (zy_fused(p)) {
      const int64_t M = N1 / ZY_R;
      if ((st = run_zy(p, -1, false, zy_args(p, in, zbuf, n0, N2, Np)))) return st;
      for (int k1 = 0; k1 < ZY_R; ++k1) {
        // rows k1 M + n2 -> rows k1 + ZY_R k2 of the exchange blocks
        auto ay = sargs((const C*)zbuf + k1 * M * Np, (C*)send + k1 * Nh, n0 * Nh, Nh, N1 * Np,
                        n1 * Nh, Np, ZY_R * Nh);
        ay.oshift = P == 1 ? 0 : ilog2(n1) - ilog2(ZY_R);
        ay.Bout = n0 * n1 * Nh;
        if ((st = run_pass<T>(p, "y", true, M_C2C, -1, ay, ysub(p), 1))) return st;
      }
    } else {
      auto az = zargs(in, zbuf, n0 * N1, N2, Np);
      if ((st = run_pass<T>(p, "z", false, M_R2C, -1, az, tab(p, 2), 0))) return st;
      auto ay = sargs(zbuf, send, n0 * Nh, Nh, N1 * Np, n1 * Nh, Np, Nh);
      if (P > 1) out_blocks(ay, n1, n0 * n1 * Nh);
      if ((st = run_pass<T>(p, "y", true, M_C2C, -1, ay, tab(p, 1), 1))) return st;
    }
    if (p->exchange && (st = exchange(p, send, out, (int64_t)(n0 * n1 * Nh * sizeof(C)), -1)))
      return st;
    auto ax = sargs(out, out, n1 * Nh, n1 * Nh, 0, 0, n1 * Nh, n1 * Nh);
    return run_pass<T>(p, "x", true, M_C2C, -1, ax, tab(p, 0), 2);
  }
  // The inverse: x pass in place on the input (which it overwrites, as the
  // reference's c2r does), exchange, y pass reading the received blocks
  // (the unpack local_transpose of transpose_gpu.cpp:1411 as its load), z c2r.
  static int c2r_slab(accfft_amd_plan* p, void* in, void* out) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    const int P = p->nprocs;
    const int64_t n0 = N0 / P, n1 = N1 / P, Np = ws_pitch(Nh);
    const size_t cb = (size_t)n0 * N1 * Nh * sizeof(C);
    int st;
    if ((st = ensure_ws(p, 0, (size_t)n0 * N1 * Np * sizeof(C)))) return st;
    if (p->exchange && (st = ensure_ws(p, 1, cb))) return st;
    auto ax = sargs(in, in, n1 * Nh, n1 * Nh, 0, 0, n1 * Nh, n1 * Nh);
    if ((st = run_pass<T>(p, "x", true, M_C2C, +1, ax, tab(p, 0), 3))) return st;
    const void* recv = in;
    if (p->exchange) {
      if ((st = exchange(p, in, p->ws[1], (int64_t)(n0 * n1 * Nh * sizeof(C)), +1))) return st;
      recv = p->ws[1];
    }
    void* ybuf = p->ws[0];
    if (zy_fused(p)) {
      const int64_t M = N1 / ZY_R;
      for (int k1 = 0; k1 < ZY_R; ++k1) {
        // rows k1 + ZY_R k2 of the received blocks -> rows k1 M + n2
        auto ay = sargs((const C*)recv + k1 * Nh, (C*)ybuf + k1 * M * Np, n0 * Nh, Nh, n1 * Nh,
                        N1 * Np, ZY_R * Nh, Np);
        ay.ishift = P == 1 ? 0 : ilog2(n1) - ilog2(ZY_R);
        ay.Bin = n0 * n1 * Nh;
        if ((st = run_pass<T>(p, "y", true, M_C2C, +1, ay, ysub(p), 4))) return st;
      }
      return run_zy(p, +1, false, zy_args(p, ybuf, out, n0, Np, N2));
    }
    auto ay = sargs(recv, ybuf, n0 * Nh, Nh, n1 * Nh, N1 * Np, Nh, Np);
    if (P > 1) in_blocks(ay, n1, n0 * n1 * Nh);
    if ((st = run_pass<T>(p, "y", true, M_C2C, +1, ay, tab(p, 1), 4))) return st;
    auto az = zargs(ybuf, out, n0 * N1, Np, N2);
    return run_pass<T>(p, "z", false, M_C2R, +1, az, tab(p, 2), 5);
  }
  // ---------------------------------------------------------- pencil share
  // One rank's share of an r2c on a p0 x p1 grid (src/accfft_gpu.cpp:471-623):
  // [n0][n1][N2] real -> z r2c -> T_plan_1 over the p1 ranks of the row
  // communicator -> [n0][N1][nz] -> y c2c -> T_plan_2 over the p0 ranks of
  // the column communicator -> [N0][ny][nz] -> x c2c (n0, n1 = isize0, 1;
  // ny, nz = osize1, 2).  Each transpose is the pack of memcpy_v1_h1
  // (accfft_amd_pencil_pack), the exchange (the plan's exchange_v hook, an
  // MPI_Alltoallv in the reference), and the unpack of memcpy_v1_h2
  // (accfft_amd_pencil_unpack) or, for T_plan_2, the receive straight into
  // the output rows; the inverse transposes pack the row ranges and
  // scatter the received column ranges back (afft_pencil_scatter).  A
  // communicator of one rank exchanges nothing (transpose_gpu.cpp:1814-1816).
  struct Xchg {  // per-peer byte counts and offsets of one exchange
    int64_t sb[64], so[64], rb[64], ro[64];
    int P;
    bool same() const {  // every block sent has the size of the one received
      for (int q = 0; q < P; ++q)
        if (sb[q] != rb[q]) return false;
      return true;
    }
  };
  // T_plan_1 (comm 0) and T_plan_2 (comm 1) of the forward transform; the
  // inverse exchanges are the same with send and receive swapped.
  static void pencil_tables(const accfft_amd_plan* p, Xchg& t1, Xchg& t2) {
    const int64_t cs = sizeof(C), n0 = p->isize[0], n1 = p->isize[1];
    const int64_t ny = p->osize[1], nz = p->osize[2];
    t1.P = p->p1;
    for (int q = 0; q < p->p1; ++q) {
      t1.sb[q] = n0 * n1 * p->zn[q] * cs;  // z columns of peer q, [n0][n1][zn_q]
      t1.so[q] = n0 * n1 * p->zs[q] * cs;
      t1.rb[q] = n0 * p->y1n[q] * nz * cs;  // peer q's y rows, [n0][y1n_q][nz]
      t1.ro[q] = n0 * p->y1s[q] * nz * cs;
    }
    t2.P = p->p0;
    for (int q = 0; q < p->p0; ++q) {
      t2.sb[q] = n0 * p->yn[q] * nz * cs;  // y rows of peer q, [n0][yn_q][nz]
      t2.so[q] = n0 * p->ys[q] * nz * cs;
      t2.rb[q] = p->xn[q] * ny * nz * cs;  // peer q's x rows, [xn_q][ny][nz]
      t2.ro[q] = p->xs[q] * ny * nz * cs;
    }
  }
  static void swap_sides(Xchg& t) {
    for (int q = 0; q < t.P; ++q) {
      std::swap(t.sb[q], t.rb[q]);
      std::swap(t.so[q], t.ro[q]);
    }
  }
  // Runs one exchange through the hook; returns where the received blocks
  // are (recv, or send itself for the identity exchange of a plan without
  // a hook).
  static int pencil_exchange(accfft_amd_plan* p, int comm, const Xchg& t, const void* send,
                             void* recv, int dir, const void** got) {
    if (!p->exchange_v) {
      *got = send;
      return ACCFFT_AMD_OK;
    }
    *got = recv;
    return comm_span(p, [&] {
      return p->exchange_v(comm, send, t.sb, t.so, recv, t.rb, t.ro, t.P, dir, p->stream,
                           p->exchange_user);
    });
  }
  static int pencil_ws(accfft_amd_plan* p) {
    const int64_t n0 = p->isize[0], n1 = p->isize[1], N1 = p->n[1], Nh = p->n[2] / 2 + 1;
    const int64_t nz = p->osize[2];
    const size_t b = (size_t)std::max(n0 * n1 * Nh, n0 * N1 * nz) * sizeof(C);
    const bool hook = p->exchange_v != nullptr;
    // separate copies: 0 the z output and y buffer, 1 the send blocks, 2 the
    // receive blocks of the row exchange; fused: 0 the row exchange's blocks
    // of the z side, 1 received blocks, 2 the y pass's exchange blocks
    bool need[3];
    if (pencil_fused(p)) {
      need[0] = true;
      need[1] = hook && (p->p0 > 1 || p->p1 > 1);
      need[2] = p->p1 > 1 || (p->p0 > 1 && hook);
    } else {
      need[0] = true;
      need[1] = p->p1 > 1 || (p->p0 > 1 && hook);
      need[2] = p->p1 > 1 && hook;
    }
    int st;
    for (int i = 0; i < 3; ++i)
      if (need[i] && (st = ensure_ws(p, i, b))) return st;
    return ACCFFT_AMD_OK;
  }
  // Whether the copies ride on the FFT passes: the y rows split in equal
  // power-of-two blocks over both communicators (or one rank), so that the
  // y pass loads the row exchange's received blocks and stores the column
  // exchange's send blocks as row blocks (FftArgs::ishift / oshift), and
  // the z pass stores (loads, inverse) the row exchange's column blocks
  // (FftArgs::zcol).  The forward then moves R + 5 C bytes like a 1 x 1
  // plan; other splits run the separate pack / unpack copies.
  static bool pencil_fused(const accfft_amd_plan* p) {
    auto even = [&](int P) {
      if (P == 1) return true;
      const int b = p->n[1] / P;
      return p->n[1] % P == 0 && (b & (b - 1)) == 0;
    };
    return even(p->p0) && even(p->p1);
  }
  // column blocks of width w (< 2^16: the z lengths of one LDS tile)
  static void set_zcol(FftArgs<T>& a, int64_t w) {
    a.zcol = w;
    a.zmag = w > 1 ? (uint32_t)((((uint64_t)1 << 32) + (uint64_t)w - 1) / (uint64_t)w) : 0u;
  }
  static int r2c_pencil_fused(accfft_amd_plan* p, const void* in, void* out, const Xchg& t1,
                              const Xchg& t2) {
    const int64_t N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    const int64_t n0 = p->isize[0], n1 = p->isize[1], ny = p->osize[1], nz = p->osize[2];
    const int P0 = p->p0, P1 = p->p1;
    const bool hook = p->exchange_v != nullptr;
    int st;
    // z r2c into the row exchange's send blocks [q][n0 n1][zn_q]
    auto az = zargs(in, p->ws[0], n0 * n1, N2, Nh);
    if (P1 > 1) set_zcol(az, p->zn[0]);
    if ((st = run_pass<T>(p, "z", false, M_R2C, -1, az, tab(p, 2), 0))) return st;
    const void* yin = p->ws[0];
    if (P1 > 1 && (st = pencil_exchange(p, 0, t1, p->ws[0], p->ws[1], -1, &yin))) return st;
    // y c2c from the received row blocks [q][n0][y1n_q][nz] into the column
    // exchange's send blocks [q][n0][yn_q][nz] (the output for one rank)
    void* yout = (P0 > 1 && hook) ? p->ws[2] : out;
    auto ay = sargs(yin, yout, n0 * nz, nz, (P1 > 1 ? n1 : N1) * nz, (P0 > 1 ? ny : N1) * nz,
                    nz, nz);
    if (P1 > 1) in_blocks(ay, n1, n0 * n1 * nz);
    if (P0 > 1) out_blocks(ay, ny, n0 * ny * nz);
    if ((st = run_pass<T>(p, "y", true, M_C2C, -1, ay, tab(p, 1), 1))) return st;
    const void* got;
    if (P0 > 1 && (st = pencil_exchange(p, 1, t2, yout, out, -1, &got))) return st;
    auto ax = sargs(out, out, ny * nz, ny * nz, 0, 0, ny * nz, ny * nz);
    return run_pass<T>(p, "x", true, M_C2C, -1, ax, tab(p, 0), 2);
  }
  static int c2r_pencil_fused(accfft_amd_plan* p, void* in, void* out, const Xchg& t1,
                              const Xchg& t2) {
    const int64_t N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    const int64_t n0 = p->isize[0], n1 = p->isize[1], ny = p->osize[1], nz = p->osize[2];
    const int P0 = p->p0, P1 = p->p1;
    int st;
    auto ax = sargs(in, in, ny * nz, ny * nz, 0, 0, ny * nz, ny * nz);
    if ((st = run_pass<T>(p, "x", true, M_C2C, +1, ax, tab(p, 0), 3))) return st;
    const void* yin = in;
    if (P0 > 1 && (st = pencil_exchange(p, 1, t2, in, p->ws[1], +1, &yin))) return st;
    // y c2c from the received column blocks into the row exchange's send
    // blocks (the z pass's input for one rank)
    void* yout = P1 > 1 ? p->ws[2] : p->ws[0];
    auto ay = sargs(yin, yout, n0 * nz, nz, (P0 > 1 ? ny : N1) * nz, (P1 > 1 ? n1 : N1) * nz,
                    nz, nz);
    if (P0 > 1) in_blocks(ay, ny, n0 * ny * nz);
    if (P1 > 1) out_blocks(ay, n1, n0 * n1 * nz);
    if ((st = run_pass<T>(p, "y", true, M_C2C, +1, ay, tab(p, 1), 4))) return st;
    const void* zin = yout;
    if (P1 > 1 && (st = pencil_exchange(p, 0, t1, yout, p->ws[0], +1, &zin))) return st;
    // z c2r from the received column blocks [q][n0 n1][zn_q]
    auto az = zargs(zin, out, n0 * n1, Nh, N2);
    if (P1 > 1) set_zcol(az, p->zn[0]);
    return run_pass<T>(p, "z", false, M_C2R, +1, az, tab(p, 2), 5);
  }
  static int r2c_pencil(accfft_amd_plan* p, const void* in, void* out) {
    const int64_t N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    const int64_t n0 = p->isize[0], n1 = p->isize[1], ny = p->osize[1], nz = p->osize[2];
    const int P0 = p->p0, P1 = p->p1;
    const int es = (int)sizeof(C);
    Xchg t1, t2;
    pencil_tables(p, t1, t2);
    if (!p->exchange_v && ((P1 > 1 && !t1.same()) || (P0 > 1 && !t2.same())))
      return ACCFFT_AMD_ERR_UNSUPPORTED;
    int st;
    if ((st = pencil_ws(p))) return st;
    if (pencil_fused(p)) return r2c_pencil_fused(p, in, out, t1, t2);
    void* A = p->ws[0];
    auto az = zargs(in, A, n0 * n1, N2, Nh);
    if ((st = run_pass<T>(p, "z", false, M_R2C, -1, az, tab(p, 2), 0))) return st;
    if (P1 > 1) {  // [n0][n1][Nh] -> [n0][N1][nz]
      const void* got;
      if ((st = accfft_amd_pencil_pack(P1, (int)n0, (int)n1, 1, p->zn.data(), p->ws[1], A,
                                       n1 * Nh, Nh, p->zs.data(), es, p->stream)))
        return st;
      if ((st = pencil_exchange(p, 0, t1, p->ws[1], p->ws[2], -1, &got))) return st;
      if ((st = accfft_amd_pencil_unpack(P1, (int)n0, p->y1s.data(), p->y1n.data(), A, N1 * nz,
                                         nz, 1, got, es, p->stream)))
        return st;
    }
    // y c2c over [n0][N1][nz]; with one rank along x it is already the output
    void* yout = P0 > 1 ? A : out;
    auto ay = sargs(A, yout, n0 * nz, nz, N1 * nz, N1 * nz, nz, nz);
    if ((st = run_pass<T>(p, "y", true, M_C2C, -1, ay, tab(p, 1), 1))) return st;
    if (P0 > 1) {  // [n0][N1][nz] -> [N0][ny][nz], received in place
      void* send = p->exchange_v ? p->ws[1] : out;
      const void* got;
      if ((st = accfft_amd_pencil_pack(P0, (int)n0, 1, nz, p->yn.data(), send, A, N1 * nz, N1,
                                       p->ys.data(), es, p->stream)))
        return st;
      if ((st = pencil_exchange(p, 1, t2, send, out, -1, &got))) return st;
    }
    auto ax = sargs(out, out, ny * nz, ny * nz, 0, 0, ny * nz, ny * nz);
    return run_pass<T>(p, "x", true, M_C2C, -1, ax, tab(p, 0), 2);
  }
  // The inverse: x c2c in place on the input (which it overwrites, as the
  // reference's c2r does), T_plan_2i, y c2c, T_plan_1i, z c2r.
  static int c2r_pencil(accfft_amd_plan* p, void* in, void* out) {
    const int64_t N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    const int64_t n0 = p->isize[0], n1 = p->isize[1], ny = p->osize[1], nz = p->osize[2];
    const int P0 = p->p0, P1 = p->p1;
    const int es = (int)sizeof(C);
    Xchg t1, t2;
    pencil_tables(p, t1, t2);
    swap_sides(t1);
    swap_sides(t2);
    if (!p->exchange_v && ((P1 > 1 && !t1.same()) || (P0 > 1 && !t2.same())))
      return ACCFFT_AMD_ERR_UNSUPPORTED;
    int st;
    if ((st = pencil_ws(p))) return st;
    if (pencil_fused(p)) return c2r_pencil_fused(p, in, out, t1, t2);
    void* A = p->ws[0];
    auto ax = sargs(in, in, ny * nz, ny * nz, 0, 0, ny * nz, ny * nz);
    if ((st = run_pass<T>(p, "x", true, M_C2C, +1, ax, tab(p, 0), 3))) return st;
    const void* ysrc = in;
    if (P0 > 1) {  // [N0][ny][nz] -> [n0][N1][nz]
      const void* got;
      if ((st = pencil_exchange(p, 1, t2, in, p->ws[1], +1, &got))) return st;
      if ((st = afft_pencil_scatter(P0, (int)n0, 1, nz, p->yn.data(), got, A, N1 * nz, N1,
                                    p->ys.data(), es, p->stream)))
        return st;
      ysrc = A;
    }
    auto ay = sargs(ysrc, A, n0 * nz, nz, N1 * nz, N1 * nz, nz, nz);
    if ((st = run_pass<T>(p, "y", true, M_C2C, +1, ay, tab(p, 1), 4))) return st;
    if (P1 > 1) {  // [n0][N1][nz] -> [n0][n1][Nh]
      const void* got;
      if ((st = accfft_amd_pencil_pack(P1, (int)n0, 1, nz, p->y1n.data(), p->ws[1], A, N1 * nz,
                                       N1, p->y1s.data(), es, p->stream)))
        return st;
      if ((st = pencil_exchange(p, 0, t1, p->ws[1], p->ws[2], +1, &got))) return st;
      if ((st = afft_pencil_scatter(P1, (int)n0, (int)n1, 1, p->zn.data(), got, A, n1 * Nh, Nh,
                                    p->zs.data(), es, p->stream)))
        return st;
    }
    auto az = zargs(A, out, n0 * n1, Nh, N2);
    return run_pass<T>(p, "z", false, M_C2R, +1, az, tab(p, 2), 5);
  }
  static int ilog2(int64_t v) {
    int s = 0;
    while ((int64_t(1) << s) < v) ++s;
    return s;
  }
  // Rows of a strided pass's input / output in blocks of `rows` (a power of
  // two) at distance B (FftArgs::ishift / oshift); blocks of one row are
  // shift 0 with the row stride B.
  static void in_blocks(FftArgs<T>& a, int64_t rows, int64_t B) {
    if (rows == 1) a.Sin = B;
    else a.ishift = ilog2(rows);
    a.Bin = B;
  }
  static void out_blocks(FftArgs<T>& a, int64_t rows, int64_t B) {
    if (rows == 1) a.Sout = B;
    else a.oshift = ilog2(rows);
    a.Bout = B;
  }

  // ---------------------------------------------------------- adjoints
  // The operators here are real-linear maps; with <a, b> = Re sum conj(a) b
  // on spectra and the plain dot product on real fields:
  //   r2c^T(G) = c2r(w G),  w = 1/2 on z bins other than DC / Nyquist
  //   c2r^T(y) = (1/w) r2c(y)
  // Neither overwrites its input.
  static int adjoint_r2c(accfft_amd_plan* p, const void* G, void* x) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    int st;
    if ((st = ensure_ws(p, 0, (size_t)N0 * N1 * Nh * sizeof(C)))) return st;
    void* w = p->ws[0];
    if ((st = x_c2c(p, G, w, Nh, +1))) return st;
    if ((st = y_c2c(p, w, w, Nh, +1))) return st;
    auto a = zargs(w, x, N0 * N1, Nh, N2);
    a.adjw = 1;
    return run_pass<T>(p, "z", false, M_C2R, +1, a, tab(p, 2));
  }
  static int adjoint_c2r(accfft_amd_plan* p, const void* y, void* G) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    int st;
    auto a = zargs(y, G, N0 * N1, N2, Nh);
    a.adjw = 1;
    if ((st = run_pass<T>(p, "z", false, M_R2C, -1, a, tab(p, 2)))) return st;
    if ((st = y_c2c(p, G, G, Nh, -1))) return st;
    return x_c2c(p, G, G, Nh, -1);
  }
  // grad^T(gx, gy, gz) = -div(gx, gy, gz); div^T(v) = -grad(v): the per-axis
  // derivative with the symmetric Nyquist filter is skew-adjoint.
  static int adjoint_grad(accfft_amd_plan* p, void* out, const void* gx, const void* gy,
                          const void* gz) {
    const void* g[3] = {gx, gy, gz};
    int st, first = 1;
    for (int d = 0; d < 3; ++d) {
      if (!g[d]) continue;
      if ((st = deriv(p, d, g[d], out, first ? 0 : 1, -1.0))) return st;
      first = 0;
    }
    if (first) return ACCFFT_AMD_ERR_NULL;
    return ACCFFT_AMD_OK;
  }
  static int adjoint_div(accfft_amd_plan* p, void* ox, void* oy, void* oz, const void* v) {
    void* o[3] = {ox, oy, oz};
    int st;
    for (int d = 0; d < 3; ++d)
      if (o[d])
        if ((st = deriv(p, d, v, o[d], 0, -1.0))) return st;
    return ACCFFT_AMD_OK;
  }

  // ---------------------------------------------------------- operators
  static int grad(accfft_amd_plan* p, void* Ax, void* Ay, void* Az, const void* A,
                  unsigned mask, bool slow) {
    void* outs[3] = {Ax, Ay, Az};
    int st;
    if (!slow) {
      for (int d = 0; d < 3; ++d)
        if ((mask & (1u << d)) && outs[d])
          if ((st = deriv(p, d, A, outs[d], 0))) return st;
      return ACCFFT_AMD_OK;
    }
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    const size_t cb = (size_t)N0 * N1 * Nh * sizeof(C);
    if ((st = ensure_ws(p, 0, cb)) || (st = ensure_ws(p, 1, cb))) return st;
    if (yfused(p)) {
      if ((st = zy_fwd(p, A, p->ws[0]))) return st;
    } else {
      if ((st = z_r2c(p, A, p->ws[0], false))) return st;
      if ((st = y_c2c(p, p->ws[0], p->ws[0], Nh, -1))) return st;
    }
    const double scale = 1.0 / ((double)N0 * N1 * N2);
    for (int d = 0; d < 3; ++d) {
      if (!(mask & (1u << d)) || !outs[d]) continue;
      if ((st = x_mult(p, p->ws[0], p->ws[1], Nh, SYM_GRAD_X + d, scale))) return st;
      if (yfused(p)) {
        if ((st = zy_inv(p, p->ws[1], p->ws[1], outs[d]))) return st;
        continue;
      }
      if ((st = y_c2c(p, p->ws[1], p->ws[1], Nh, +1))) return st;
      if ((st = z_c2r(p, p->ws[1], outs[d], false))) return st;
    }
    return ACCFFT_AMD_OK;
  }

  // 3-D r2c -> symbol -> 3-D c2r with the multiply fused into the x pass
  static int spectral(accfft_amd_plan* p, int sym, void* out, const void* A) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    int st;
    if ((st = ensure_ws(p, 0, (size_t)N0 * N1 * Nh * sizeof(C)))) return st;
    return spectral_in(p, sym, out, A, p->ws[0]);
  }
  // the same passes over a caller-provided spectrum w of N0 N1 (N2/2+1)
  // complex values (every element is written before it is read)
  static int spectral_in(accfft_amd_plan* p, int sym, void* out, const void* A, void* w) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    int st;
    if (yfused(p)) {
      if ((st = zy_fwd(p, A, w))) return st;
      if ((st = x_mult(p, w, w, Nh, sym, 1.0 / ((double)N0 * N1 * N2)))) return st;
      return zy_inv(p, w, w, out);
    }
    if ((st = z_r2c(p, A, w, false))) return st;
    if ((st = y_c2c(p, w, w, Nh, -1))) return st;
    if ((st = x_mult(p, w, w, Nh, sym, 1.0 / ((double)N0 * N1 * N2)))) return st;
    if ((st = y_c2c(p, w, w, Nh, +1))) return st;
    return z_c2r(p, w, out, false);
  }

  static int divergence(accfft_amd_plan* p, void* div, const void* Ax, const void* Ay,
                        const void* Az, bool slow) {
    int st;
    if (!slow) {
      if ((st = deriv(p, 0, Ax, div, 0))) return st;
      if ((st = deriv(p, 1, Ay, div, 1))) return st;
      return deriv(p, 2, Az, div, 1);
    }
    // operators_gpu.txx:336-446: component d transformed over axes {d, z}
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    const int64_t R = N0 * N1 * N2;
    if ((st = ensure_ws(p, 0, (size_t)N0 * N1 * Nh * sizeof(C))) ||
        (st = ensure_ws(p, 1, (size_t)R * sizeof(T))))
      return st;
    void* w = p->ws[0];
    if ((st = z_r2c(p, Ax, w, false))) return st;
    if ((st = x_mult(p, w, w, Nh, SYM_GRAD_X, 1.0 / ((double)N0 * N2)))) return st;
    if ((st = z_c2r(p, w, div, false))) return st;
    if ((st = z_r2c(p, Ay, w, false))) return st;
    if ((st = y_mult(p, w, w, Nh, SYM_GRAD_Y, 1.0 / ((double)N1 * N2)))) return st;
    if ((st = z_c2r(p, w, p->ws[1], false))) return st;
    if ((st = status_of(launch_axpy<T>(R, (T)1, (const T*)p->ws[1], (T*)div, p->stream))))
      return st;
    return deriv(p, 2, Az, div, 1);
  }

  // ---------------------------------------------------------- MEASURE tuning
  // The reference's ACCFFT_MEASURE tunes its local FFT and transposes at
  // plan time (doc/flags.md).  Here, every strided C2C pass of the plan
  // whose length has several tile configurations times each candidate
  // (one-level tiles, the two-level variants, each with four XCD chunk
  // sizes) on its real shape over a zeroed temporary buffer, and keeps the
  // fastest when it beats the built-in rule by more than 3 %.  It guards
  // against shapes and machines where the rule loses: with a slower build
  // of the two-level tiles (512^3 y passes 285/315 us) it picked the
  // one-level tiles (236/240 us) on its own.  If the buffer cannot be allocated the
  // rule stays.  As with FFTW's MEASURE, the choice can change results in
  // the last bits.  Split-precision plans time their own (matrix-core)
  // tiles the same way.
  static int tune(accfft_amd_plan* p) {
    const int64_t N0 = p->n[0], N1 = p->n[1], N2 = p->n[2], Nh = N2 / 2 + 1;
    struct Pass {
      int ax, order;
      FftArgs<T> a;
      int64_t in_elems, out_elems;
    };
    std::vector<Pass> passes;
    if (p->kind == 1 || p->kind == 2) {
      const int64_t L = p->kind == 1 ? Nh : N2, e = N0 * N1 * L;
      if (!yfused(p)) passes.push_back({1, 1, y_args(p, nullptr, nullptr, L), e, 0});
      passes.push_back({0, 2, x_args(p, nullptr, nullptr, L), e, 0});
    } else if (p->kind == 3) {
      const int P = p->nprocs;
      const int64_t n0 = N0 / P, n1 = N1 / P, Np = ws_pitch(Nh);
      auto ay = sargs(nullptr, nullptr, n0 * Nh, Nh, N1 * Np, n1 * Nh, Np, Nh);
      if (P > 1) out_blocks(ay, n1, n0 * n1 * Nh);
      if (!zy_fused(p)) passes.push_back({1, 1, ay, n0 * N1 * Np, N0 * n1 * Nh});
      passes.push_back({0, 2, sargs(nullptr, nullptr, n1 * Nh, n1 * Nh, 0, 0, n1 * Nh, n1 * Nh),
                        N0 * n1 * Nh, 0});
    }
    int64_t elems = 0;
    bool any = false;
    for (const Pass& q : passes) {
      const AxisTab<T>& t = tab(p, q.ax);
      if (!t.is_static || !t.tw_sub || (t.n != 512 && t.n != 1024 && t.n != 2048)) continue;
      any = true;
      if (q.in_elems + q.out_elems > elems) elems = q.in_elems + q.out_elems;
    }
    if (!any) return ACCFFT_AMD_OK;
    void* buf = nullptr;
    if (hipMalloc(&buf, (size_t)elems * sizeof(C)) != hipSuccess) {
      (void)hipGetLastError();
      return ACCFFT_AMD_OK;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = ACCFFT_AMD_OK;
    if (hipMemsetAsync(buf, 0, (size_t)elems * sizeof(C), p->stream) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
      st = ACCFFT_AMD_ERR_HIP;
    for (const Pass& q : passes) {
      if (st) break;
      AxisTab<T>& t = const_cast<AxisTab<T>&>(tab(p, q.ax));
      if (!t.is_static || !t.tw_sub || (t.n != 512 && t.n != 1024 && t.n != 2048)) continue;
      FftArgs<T> a = q.a;
      a.in = buf;
      a.out = q.out_elems ? (void*)((C*)buf + q.in_elems) : buf;
      // mean of 3 launches after one warm-up; < 0 when the candidate fails
      auto measure = [&](int cand, int xg) -> float {
        t.s4v = cand;
        t.xg = xg;
        for (int r = 0; r < 4; ++r) {
          if (r == 1) (void)hipEventRecord(e0, p->stream);
          if (run_pass<T>(p, "tune", true, M_C2C, -1, a, t, q.order)) return -1.f;
        }
        (void)hipEventRecord(e1, p->stream);
        float ms = -1.f;
        if (hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
          return -1.f;
        return ms / 3.f;
      };
      // candidates: tile configuration (the rule -1, one-level tiles -2, the
      // variants) x tiles per XCD chunk (the rule -1, 16, 32, 64, one range
      // per XCD 0); two interleaved rounds, each candidate's best kept
      // (clock ramps and neighbours' noise hit every candidate alike)
      constexpr int NG = 5;
      const int groups[NG] = {-1, 16, 32, 64, 0};
      float best_of[6][NG];
      for (auto& row : best_of)
        for (float& b : row) b = -1.f;
      for (int round = 0; round < 2; ++round)
        for (int cand = -1; cand < 4; cand = cand == -1 ? -2 : cand == -2 ? 0 : cand + 1) {
          if (cand >= 0 && !s4_config_ok<T>(t.n, cand)) continue;
          for (int gi = 0; gi < NG; ++gi) {
            const float ms = measure(cand, groups[gi]);
            float& b = best_of[cand + 2][gi];
            if (ms > 0.f && (b < 0.f || ms < b)) b = ms;
          }
        }
      const float rule = best_of[1][0];
      int best = -1, best_g = -1;
      float best_ms = rule;
      for (int cand = -2; cand < 4; ++cand)
        for (int gi = 0; gi < NG; ++gi) {
          const float ms = best_of[cand + 2][gi];
          if (ms > 0.f && rule > 0.f && ms < best_ms) {
            best_ms = ms;
            best = cand;
            best_g = groups[gi];
          }
        }
      const bool take = best_ms < 0.97f * rule;
      t.s4v = take ? best : -1;
      t.xg = take ? best_g : -1;
    }
    (void)hipGetLastError();
    (void)hipStreamSynchronize(p->stream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
    return st;
  }
};

template <>
inline const AxisTab<float>& Ops<float>::tab(const accfft_amd_plan* p, int ax) {
  return p->tf[ax];
}
template <>
inline const AxisTab<double>& Ops<double>::tab(const accfft_amd_plan* p, int ax) {
  return p->td[ax];
}

}  // namespace afft

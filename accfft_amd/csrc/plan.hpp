// plan.hpp -- the plan object behind the C ABI (include/accfft_amd.h).
//
// Mirrors struct accfft_plan_gpu(f) (include/accfft_gpu.h:37-95) on a 1x1
// grid: sizes, the data pointers recorded at plan time, the in-place flag,
// and in place of cuFFT handles the device twiddle tables of each axis.
// Scratch (the reference's Mem_Mgr_gpu, src/transpose_gpu.cpp:58-131) is a
// pair of workspaces allocated on first use.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/accfft_amd.h"
#include "fft_launch.hpp"

struct accfft_amd_plan {
  int n[3] = {0, 0, 0};
  int fp = ACCFFT_AMD_F64;
  // 1 = r2c, 2 = c2c, 3 = r2c slab share (accfft_amd_plan_dft_3d_r2c_slab),
  // 4 = r2c pencil share (accfft_amd_plan_dft_3d_r2c_pencil)
  int kind = 0;
  int nprocs = 1;  // slab share: ranks of the 1-D split (n[0] and n[1] are global)
  accfft_amd_exchange_fn exchange = nullptr;  // slab share: NULL = identity
  accfft_amd_exchange_v_fn exchange_v = nullptr;  // pencil share: NULL = identity
  void* exchange_user = nullptr;
  // pencil share: the p0 x p1 grid, this rank's coordinates, and the
  // per-peer counts and starts of the ceil-block splits
  // (dfft_get_local_size, src/accfft_common.cpp:49-91): x rows over p0
  // (xn, xs), input y rows over p1 (y1n, y1s), output y rows over p0
  // (yn, ys), output z columns of the half spectrum over p1 (zn, zs)
  int p0 = 1, p1 = 1, c0 = 0, c1 = 0;
  std::vector<int64_t> xn, xs, y1n, y1s, yn, ys, zn, zs;
  bool inplace = false;
  unsigned flags = 0;
  void* data = nullptr;
  void* data_out = nullptr;
  int64_t alloc_max = 0;
  int isize[3] = {0, 0, 0}, istart[3] = {0, 0, 0};
  int osize[3] = {0, 0, 0}, ostart[3] = {0, 0, 0};
  afft::AxisTab<float> tf[3];
  afft::AxisTab<double> td[3];
  // tables of the second level (length N1/8) of a y transform whose first
  // radix-8 level runs inside the z pass: the slab share with 2048-point y
  // and z axes (FP32), and 1 x 1 r2c plans whose shape allows it
  // (afft::zy_plain_ok); n == 0 when not built
  afft::AxisTab<float> tf_ysub;
  afft::AxisTab<double> td_ysub;
  // 1 x 1 r2c plans: the z pass carries the y transform's first radix-8
  // level (set at plan creation when the tables exist; tuning choice -3 on
  // the y axis turns it off, -1 restores it)
  bool yfuse = false;
  std::vector<void*> dev_mem;  // table allocations
  // 0, 1: spectra of the operators, adjoints and slab plans (and 2 of the
  // pencil plans); 2, 3: the ping-pong buffers of lines longer than one LDS
  // tile (fft_long.hpp); 4, 5: Bluestein's convolution buffers
  static constexpr int NWS = 6;
  void* ws[NWS] = {};
  size_t ws_bytes[NWS] = {};
  hipStream_t stream = nullptr;
  // stream of the previous call: when a call comes on another stream, that
  // stream first waits for everything already queued on the previous one
  // (the plan's workspaces are shared by all its calls)
  hipStream_t last_stream = nullptr;
  bool called = false;
  hipEvent_t ev_switch = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_ms = 0.f;
  // timer contract (timer[2] = exchange, timer[4] = local FFT): set while a
  // timed call runs; each exchange hook call of the call is bracketed by
  // the event pair evx[2 k], evx[2 k + 1] (k < comm_pairs)
  bool timing = false;
  static constexpr int NEVX = 4;
  int comm_pairs = 0;
  hipEvent_t evx[NEVX] = {nullptr, nullptr, nullptr, nullptr};
  int device = 0;
  int split = 0;  // afft::SplitKind of the radix-16 stages (FP32 plans)
  // per-launch event profiling (accfft_amd_plan_set_profiling)
  struct ProfKind {
    std::string name;
    double total_ms = 0;
    int64_t launches = 0;
    double bytes = 0;  // algorithmic bytes of the last launch
  };
  struct ProfPending {
    int kind;
    hipEvent_t a, b;
  };
  bool prof = false;
  std::vector<ProfKind> prof_kinds;
  std::vector<ProfPending> prof_pending;
  std::vector<hipEvent_t> prof_pool;
};

namespace afft {

inline hipEvent_t prof_event(accfft_amd_plan* p) {
  if (!p->prof_pool.empty()) {
    hipEvent_t e = p->prof_pool.back();
    p->prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Folds finished launch timings into the per-kind totals.
inline int prof_collect(accfft_amd_plan* p) {
  for (auto& q : p->prof_pending) {
    if (hipEventSynchronize(q.b) != hipSuccess) return ACCFFT_AMD_ERR_HIP;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, q.a, q.b);
    p->prof_kinds[q.kind].total_ms += ms;
    p->prof_kinds[q.kind].launches += 1;
    p->prof_pool.push_back(q.a);
    p->prof_pool.push_back(q.b);
  }
  p->prof_pending.clear();
  return ACCFFT_AMD_OK;
}

// exp(-2 pi i m / n) in long double, rounded once to T.
template <typename T>
inline cpx<T> host_root(long m, long n) {
  m %= n;
  if (m < 0) m += n;
  if ((4 * m) % n == 0) {
    static const T c[4] = {1, 0, -1, 0}, s[4] = {0, -1, 0, 1};
    const int q = (int)((4 * m) / n);
    return {c[q], s[q]};
  }
  const long double a = 6.283185307179586476925286766559005768L * (long double)m / (long double)n;
  return {(T)cosl(a), (T)(-sinl(a))};
}

// split: schedule id (0, SCHED_WIDE or a split-precision kind)
inline int static_sched(int n, int R[3], int split = 0) {
#define AFFT_SCHED(NN)                                              \
  case NN:                                                          \
    for (int i = 0; i < 3; ++i) R[i] = SchedOf<NN, 0>::r(i);        \
    if (split == SCHED_WIDE && SchedWide<NN>::S > 0) {              \
      for (int i = 0; i < 3; ++i) R[i] = SchedWide<NN>::R[i];       \
      return SchedWide<NN>::S;                                      \
    }                                                               \
    if (split_sched(split) && SchedSplit<NN>::S > 0) {              \
      for (int i = 0; i < 3; ++i) R[i] = SchedSplit<NN>::R[i];      \
      return SchedSplit<NN>::S;                                     \
    }                                                               \
    return Sched<NN>::S;
  switch (n) {
    AFFT_SCHED(16)
    AFFT_SCHED(32)
    AFFT_SCHED(64)
    AFFT_SCHED(128)
    AFFT_SCHED(256)
    AFFT_SCHED(512)
    AFFT_SCHED(1024)
    AFFT_SCHED(2048)
    AFFT_SCHED(4096)
    default: return 0;
  }
#undef AFFT_SCHED
}

inline DynSched dyn_sched(int n) {
  DynSched s{};
  s.n = n;
  s.nst = 0;
  int m = n;
  // 9 before 3: a power of three takes half the LDS passes
  const int pref[] = {16, 8, 4, 2, 9, 3, 5, 7};
  for (int f : pref)
    while (m % f == 0 && s.nst < 20) {
      s.rad[s.nst++] = f;
      m /= f;
    }
  for (int f = 11; (long)f * f <= m; f += 2)
    while (m % f == 0 && s.nst < 20) {
      s.rad[s.nst++] = f;
      m /= f;
    }
  if (m > 1 && s.nst < 20) s.rad[s.nst++] = m;
  if (n == 1) {
    s.nst = 0;
  }
  return s;
}

// Builds the per-stage table of the static engine: for stage s >= 1 with
// span NS and radix R, entries (r-1)*NS + k = exp(-2 pi i r k / (NS R)).
template <typename T>
inline std::vector<cpx<T>> static_table(int n, int split = 0) {
  int R[3];
  const int S = static_sched(n, R, split);
  std::vector<cpx<T>> tw;
  int ns = R[0];
  for (int s = 1; s < S; ++s) {
    for (int r = 1; r < R[s]; ++r)
      for (int k = 0; k < ns; ++k) tw.push_back(host_root<T>((long)r * k, (long)ns * R[s]));
    ns *= R[s];
  }
  if (tw.empty()) tw.push_back({(T)1, (T)0});
  return tw;
}

template <typename T>
inline std::vector<cpx<T>> full_table(int n) {
  std::vector<cpx<T>> tw(n > 0 ? n : 1);
  for (int m = 0; m < n; ++m) tw[m] = host_root<T>(m, n);
  return tw;
}

}  // namespace afft

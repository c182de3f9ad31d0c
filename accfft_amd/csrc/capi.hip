// capi.hip -- the extern "C" boundary declared in include/accfft_amd.h.
//
// Plan creation restates accfft_plan_dft_3d_{r2c,c2c}_gpu
// (src/accfft_gpu.cpp:77-442, :704-931) for a 1x1 grid: sizes come from the
// same local-size rules (src/accfft_common.cpp:49-256, kept in int64 here),
// in-place is decided by pointer equality (:99-103), and the cuFFT plans are
// replaced by device twiddle tables per axis length.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/accfft_amd.h"
#include "dispatch.hpp"
#include "plan.hpp"

namespace {

// One axis of dfft_get_local_size_t (src/accfft_common.cpp:49-91): blocks
// of ceil(N / P); the block that would run past N is cut, later ones are
// empty and start at N.
void split_block(int N, int P, int c, int* size, int* start) {
  *size = (int)std::ceil(N / (double)P);
  *start = *size * c;
  if ((N - *size * c) < *size) {
    *size = N - *size * c;
    *size *= (int)(*size > 0);
    *start = N - *size;
  }
}

// dfft_get_local_size_t (src/accfft_common.cpp:49-91); returns bytes.
int64_t local_size(int N0, int N1, int tuple, int* isize, int* istart, int p0, int p1, int c0,
                   int c1, int elem) {
  isize[2] = tuple;
  istart[2] = 0;
  split_block(N0, p0, c0, &isize[0], &istart[0]);
  split_block(N1, p1, c1, &isize[1], &istart[1]);
  return (int64_t)isize[0] * isize[1] * isize[2] * elem;
}

// Per-peer sizes and starts of one axis's split over P ranks.
void split_all(int N, int P, std::vector<int64_t>& n, std::vector<int64_t>& s) {
  n.resize(P);
  s.resize(P);
  for (int q = 0; q < P; ++q) {
    int a, b;
    split_block(N, P, q, &a, &b);
    n[q] = a;
    s[q] = b;
  }
}

bool bad_grid(const int n[3], int p0, int p1, int c0, int c1) {
  if (!n) return true;
  for (int i = 0; i < 3; ++i)
    if (n[i] <= 0) return true;
  return p0 <= 0 || p1 <= 0 || c0 < 0 || c1 < 0 || c0 >= p0 || c1 >= p1;
}

// accfft_local_size_dft_r2c_t (src/accfft_common.cpp:98-192)
int64_t size_r2c(const int* n, int* isize, int* istart, int* osize, int* ostart, int p0, int p1,
                 int c0, int c1, int elem) {
  int o0[3], s0[3], o1[3], s1[3], o2[3], s2[3], oy[3], sy[3], oyi[3], syi[3], ox[3], sx[3],
      oxi[3], sxi[3];
  const int nt = (n[2] / 2 + 1) * 2;
  int64_t amax = 0;
  amax = std::max(amax, local_size(n[0], n[1], nt, o0, s0, p0, p1, c0, c1, elem));
  amax = std::max(amax, 2 * local_size(n[0], nt / 2, n[1], o1, s1, p0, p1, c0, c1, elem));
  amax = std::max(amax, 2 * local_size(n[1], nt / 2, n[0], o2, s2, p0, p1, c0, c1, elem));
  std::swap(o2[1], o2[2]);
  std::swap(o2[0], o2[1]);
  std::swap(s2[1], s2[2]);
  std::swap(s2[0], s2[1]);
  local_size(n[0], n[1], n[2], isize, istart, p0, p1, c0, c1, elem);
  for (int i = 0; i < 3; ++i) {
    osize[i] = o2[i];
    ostart[i] = s2[i];
  }
  amax = std::max(amax, local_size(n[0], n[2], n[1], oy, sy, p0, p1, c0, c1, elem));
  amax = std::max(amax, 2 * local_size(n[0], n[2], n[1] / 2 + 1, oyi, syi, p0, p1, c0, c1, elem));
  local_size(isize[1] * n[2], n[2], n[0], ox, sx, p0, p1, c0, c1, elem);
  amax = std::max(amax, (int64_t)ox[2] * ox[0] * 1 * elem);
  local_size(isize[1] * n[2], n[2], n[0] / 2 + 1, oxi, sxi, p0, p1, c0, c1, elem);
  amax = std::max(amax, 2 * (int64_t)oxi[2] * oxi[0] * 1 * elem);
  return amax;
}

// accfft_local_size_dft_c2c_t (src/accfft_common.cpp:203-256)
int64_t size_c2c(const int* n, int* isize, int* istart, int* osize, int* ostart, int p0, int p1,
                 int c0, int c1, int elem) {
  int o0[3], s0[3], o1[3], s1[3], o2[3], s2[3];
  int64_t amax = 0;
  amax = std::max(amax, local_size(n[0], n[1], n[2], o0, s0, p0, p1, c0, c1, elem));
  amax = std::max(amax, local_size(n[0], n[2], n[1], o1, s1, p0, p1, c0, c1, elem));
  amax = std::max(amax, local_size(n[1], n[2], n[0], o2, s2, p0, p1, c0, c1, elem));
  amax *= 2;
  std::swap(s2[1], s2[2]);
  std::swap(s2[0], s2[1]);
  std::swap(o2[1], o2[2]);
  std::swap(o2[0], o2[1]);
  local_size(n[0], n[1], n[2], isize, istart, p0, p1, c0, c1, elem);
  for (int i = 0; i < 3; ++i) {
    osize[i] = o2[i];
    ostart[i] = s2[i];
  }
  return amax;
}

int elem_of(int fp) { return fp == ACCFFT_AMD_F32 ? 4 : 8; }

void free_plan(accfft_amd_plan* p) {
  if (!p) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (void* m : p->dev_mem) (void)hipFree(m);
  for (int i = 0; i < accfft_amd_plan::NWS; ++i)
    if (p->ws[i]) (void)hipFree(p->ws[i]);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  if (p->ev_switch) (void)hipEventDestroy(p->ev_switch);
  for (hipEvent_t e : p->evx)
    if (e) (void)hipEventDestroy(e);
  for (auto& q : p->prof_pending) {
    (void)hipEventDestroy(q.a);
    (void)hipEventDestroy(q.b);
  }
  for (hipEvent_t e : p->prof_pool) (void)hipEventDestroy(e);
  delete p;
}

// ACCFFT_MEASURE / ACCFFT_PATIENT: time the tile configurations of the
// strided passes on this plan's shapes (doc/flags.md); ACCFFT_ESTIMATE
// keeps the built-in rule.
int maybe_tune(accfft_amd_plan* p) {
  if (!(p->flags & (ACCFFT_AMD_MEASURE | ACCFFT_AMD_PATIENT))) return ACCFFT_AMD_OK;
  return p->fp == ACCFFT_AMD_F32 ? afft_tune_f32(p) : afft_tune_f64(p);
}

// The tile queue of the plan's persistent strided passes (sfft4_kernel):
// zeroed once here, left zeroed by every launch that uses it.
int make_queue(accfft_amd_plan* p) {
  void* q = nullptr;
  const size_t bytes = afft::QUEUE_WORDS * sizeof(unsigned);
  if (hipMalloc(&q, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return ACCFFT_AMD_ERR_ALLOC;
  }
  p->dev_mem.push_back(q);
  // a blocking copy (not a memset, which may run after the plan's first
  // launch on a non-blocking stream)
  const std::vector<unsigned> zero(afft::QUEUE_WORDS, 0u);
  if (hipMemcpy(q, zero.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return ACCFFT_AMD_ERR_HIP;
  unsigned* const u = (unsigned*)q;
  for (int ax = 0; ax < 3; ++ax) {
    p->tf[ax].queue = u;
    p->td[ax].queue = u;
  }
  p->tf_ysub.queue = u;
  p->td_ysub.queue = u;
  return ACCFFT_AMD_OK;
}

int make_plan(int kind, const int n[3], void* data, void* data_out, int fp, unsigned flags,
              accfft_amd_plan** out) {
  if (!out) return ACCFFT_AMD_ERR_NULL;
  *out = nullptr;
  if (bad_grid(n, 1, 1, 0, 0)) return ACCFFT_AMD_ERR_INVALID;
  if (fp != ACCFFT_AMD_F32 && fp != ACCFFT_AMD_F64) return ACCFFT_AMD_ERR_INVALID;
  // slab and pencil shares: every length must fit one LDS tile (lines
  // longer than a tile run as global radix passes, fft_long.hpp, on 1 x 1
  // plans only)
  for (int i = 0; i < 3 && kind >= 3; ++i)
    if (afft::long_length(n[i], fp == ACCFFT_AMD_F32 ? 8 : 16)) return ACCFFT_AMD_ERR_UNSUPPORTED;
  accfft_amd_plan* p = new (std::nothrow) accfft_amd_plan();
  if (!p) return ACCFFT_AMD_ERR_ALLOC;
  for (int i = 0; i < 3; ++i) p->n[i] = n[i];
  p->fp = fp;
  p->kind = kind;
  p->flags = flags;
  p->data = data;
  p->data_out = data_out;
  p->inplace = (data == data_out);
  if (flags & ACCFFT_AMD_SPLIT_FP16) p->split = afft::SPLIT_FP16;
  else if (flags & ACCFFT_AMD_SPLIT_BF16X3) p->split = afft::SPLIT_BF16X3;
  else if (flags & ACCFFT_AMD_SPLIT_BF16) p->split = afft::SPLIT_BF16;
  if (p->split && fp != ACCFFT_AMD_F32) {
    delete p;
    return ACCFFT_AMD_ERR_INVALID;
  }
  const int elem = elem_of(fp);
  p->alloc_max = kind == 1 ? size_r2c(n, p->isize, p->istart, p->osize, p->ostart, 1, 1, 0, 0, elem)
                           : size_c2c(n, p->isize, p->istart, p->osize, p->ostart, 1, 1, 0, 0, elem);
  (void)hipGetDevice(&p->device);
  if (hipEventCreate(&p->ev0) != hipSuccess || hipEventCreate(&p->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&p->ev_switch, hipEventDisableTiming) != hipSuccess) {
    free_plan(p);
    return ACCFFT_AMD_ERR_HIP;
  }
  int st = fp == ACCFFT_AMD_F32 ? afft_tables_f32(p) : afft_tables_f64(p);
  if (!st) st = make_queue(p);
  if (!st && kind < 3) st = maybe_tune(p);
  if (st) {
    free_plan(p);
    return st;
  }
  *out = p;
  return ACCFFT_AMD_OK;
}

// Whether a plan's real side / complex side holds no elements (a pencil
// share of a ragged split can be empty; its buffers may then be NULL).
bool real_empty(const accfft_amd_plan* p) {
  return (int64_t)p->isize[0] * p->isize[1] * p->isize[2] == 0;
}
bool cplx_empty(const accfft_amd_plan* p) {
  return (int64_t)p->osize[0] * p->osize[1] * p->osize[2] == 0;
}

// Whether the plan holds the tables of a y transform's second level.
bool ysub_built(const accfft_amd_plan* p) {
  return p->fp == ACCFFT_AMD_F32 ? p->tf_ysub.n > 0 : p->td_ysub.n > 0;
}

// Wraps one execute: events around the launches, optional timer / blocking.
// Events are recorded only when a timer or a blocking call asks for them:
// each marker costs ~10 us of queue time between kernels.
template <typename F>
int timed(accfft_amd_plan* p, double* timer, F&& body) {
  // Calls of one plan on different streams share its workspaces: a call on
  // a new stream waits for the work already queued on the previous one.
  if (p->called && p->stream != p->last_stream) {
    if (hipEventRecord(p->ev_switch, p->last_stream) != hipSuccess ||
        hipStreamWaitEvent(p->stream, p->ev_switch, 0) != hipSuccess)
      return ACCFFT_AMD_ERR_HIP;
  }
  p->called = true;
  p->last_stream = p->stream;
  const bool want = timer || (p->flags & ACCFFT_AMD_BLOCKING);
  if (!want) return body();
  p->timing = timer != nullptr;
  p->comm_pairs = 0;
  (void)hipEventRecord(p->ev0, p->stream);
  int st = body();
  (void)hipEventRecord(p->ev1, p->stream);
  p->timing = false;
  if (st) return st;
  {
    if (hipEventSynchronize(p->ev1) != hipSuccess) return ACCFFT_AMD_ERR_HIP;
    float ms = 0.f, comm = 0.f;
    (void)hipEventElapsedTime(&ms, p->ev0, p->ev1);
    for (int k = 0; k < p->comm_pairs; ++k) {
      float c = 0.f;
      (void)hipEventElapsedTime(&c, p->evx[2 * k], p->evx[2 * k + 1]);
      comm += c;
    }
    p->last_ms = ms;
    // doc/timing.md: timer[2] communication, timer[4] local FFT
    if (timer) {
      timer[2] += comm / 1000.0;
      timer[4] += (ms - comm) / 1000.0;
    }
  }
  return ACCFFT_AMD_OK;
}

}  // namespace

// Scratch of accfft_amd_inv_laplace_noalloc: one spectrum of at most
// ACCFFT_AMD_NOALLOC_MAX_BYTES, a file-scope device array that each GPU
// holds from the library's load until the process ends.  Host code uses
// the array only through the device address hipGetSymbolAddress returns
// for the current device (the name itself is a host-side stand-in).
static __device__ __align__(256) unsigned char scratch[ACCFFT_AMD_NOALLOC_MAX_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current.
int scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return ACCFFT_AMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_dev[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      return ACCFFT_AMD_ERR_HIP;
    }
    scratch_dev[device] = ptr;
  }
  *out = scratch_dev[device];
  return ACCFFT_AMD_OK;
}
}  // namespace

extern "C" {

const char* accfft_amd_status_string(int s) {
  switch (s) {
    case ACCFFT_AMD_OK: return "ok";
    case ACCFFT_AMD_ERR_INVALID: return "invalid argument";
    case ACCFFT_AMD_ERR_OVERFLOW: return "size overflow";
    case ACCFFT_AMD_ERR_UNSUPPORTED: return "unsupported transform length";
    case ACCFFT_AMD_ERR_NOT_BAKED: return "plan of the wrong kind for this call";
    case ACCFFT_AMD_ERR_HIP: return "HIP runtime error";
    case ACCFFT_AMD_ERR_SCRATCH:
      return "workspace larger than the fixed scratch of the no-allocation entry point";
    case ACCFFT_AMD_ERR_ALLOC: return "device allocation failed";
    case ACCFFT_AMD_ERR_NULL: return "NULL plan or pointer";
    default: return "unknown status";
  }
}

int accfft_amd_version(void) { return 100; }

int accfft_amd_local_size_dft_r2c(const int n[3], int isize[3], int istart[3], int osize[3],
                                  int ostart[3], int fp, int p0, int p1, int c0, int c1,
                                  int64_t* alloc_max) {
  if (!isize || !istart || !osize || !ostart) return ACCFFT_AMD_ERR_NULL;
  if (bad_grid(n, p0, p1, c0, c1)) return ACCFFT_AMD_ERR_INVALID;
  if (fp != ACCFFT_AMD_F32 && fp != ACCFFT_AMD_F64) return ACCFFT_AMD_ERR_INVALID;
  const int64_t a = size_r2c(n, isize, istart, osize, ostart, p0, p1, c0, c1, elem_of(fp));
  if (alloc_max) *alloc_max = a;
  return ACCFFT_AMD_OK;
}

int accfft_amd_local_size_dft_c2c(const int n[3], int isize[3], int istart[3], int osize[3],
                                  int ostart[3], int fp, int p0, int p1, int c0, int c1,
                                  int64_t* alloc_max) {
  if (!isize || !istart || !osize || !ostart) return ACCFFT_AMD_ERR_NULL;
  if (bad_grid(n, p0, p1, c0, c1)) return ACCFFT_AMD_ERR_INVALID;
  if (fp != ACCFFT_AMD_F32 && fp != ACCFFT_AMD_F64) return ACCFFT_AMD_ERR_INVALID;
  const int64_t a = size_c2c(n, isize, istart, osize, ostart, p0, p1, c0, c1, elem_of(fp));
  if (alloc_max) *alloc_max = a;
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_dft_3d_r2c(const int n[3], void* data, void* data_out, int fp,
                               unsigned flags, accfft_amd_plan** plan) {
  return make_plan(1, n, data, data_out, fp, flags, plan);
}

int accfft_amd_plan_dft_3d_c2c(const int n[3], void* data, void* data_out, int fp,
                               unsigned flags, accfft_amd_plan** plan) {
  return make_plan(2, n, data, data_out, fp, flags, plan);
}

int accfft_amd_plan_dft_3d_r2c_slab(const int n[3], int nprocs, void* data, void* data_out,
                                    int fp, unsigned flags, accfft_amd_plan** plan) {
  if (!plan || !n) return ACCFFT_AMD_ERR_NULL;
  *plan = nullptr;
  if (nprocs < 1 || n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return ACCFFT_AMD_ERR_INVALID;
  if (n[0] % nprocs || n[1] % nprocs) return ACCFFT_AMD_ERR_UNSUPPORTED;
  const int n1 = n[1] / nprocs;
  if (nprocs > 1 && (n1 & (n1 - 1))) return ACCFFT_AMD_ERR_UNSUPPORTED;
  if (data == data_out) return ACCFFT_AMD_ERR_UNSUPPORTED;  // out of place only
  accfft_amd_plan* p = nullptr;
  int st = make_plan(3, n, data, data_out, fp, flags, &p);
  if (st) return st;
  p->nprocs = nprocs;
  p->alloc_max = size_r2c(n, p->isize, p->istart, p->osize, p->ostart, nprocs, 1, 0, 0,
                          elem_of(fp));
  if ((st = maybe_tune(p))) {
    free_plan(p);
    return st;
  }
  *plan = p;
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_dft_3d_r2c_pencil(const int n[3], int p0, int p1, int c0, int c1,
                                      void* data, void* data_out, int fp, unsigned flags,
                                      accfft_amd_plan** plan) {
  if (!plan || !n) return ACCFFT_AMD_ERR_NULL;
  *plan = nullptr;
  if (bad_grid(n, p0, p1, c0, c1)) return ACCFFT_AMD_ERR_INVALID;
  if (p0 > 64 || p1 > 64) return ACCFFT_AMD_ERR_UNSUPPORTED;
  if (data && data == data_out) return ACCFFT_AMD_ERR_UNSUPPORTED;  // out of place only
  if (flags & (ACCFFT_AMD_SPLIT_BF16 | ACCFFT_AMD_SPLIT_BF16X3 | ACCFFT_AMD_SPLIT_FP16))
    return ACCFFT_AMD_ERR_UNSUPPORTED;  // FP32 passes only
  accfft_amd_plan* p = nullptr;
  int st = make_plan(4, n, data, data_out, fp, flags, &p);
  if (st) return st;
  p->inplace = false;  // out of place (buffers may also be given only at execute)
  p->p0 = p0;
  p->p1 = p1;
  p->c0 = c0;
  p->c1 = c1;
  p->alloc_max = size_r2c(n, p->isize, p->istart, p->osize, p->ostart, p0, p1, c0, c1,
                          elem_of(fp));
  const int Nh = n[2] / 2 + 1;
  split_all(n[0], p0, p->xn, p->xs);
  split_all(n[1], p1, p->y1n, p->y1s);
  split_all(n[1], p0, p->yn, p->ys);
  split_all(Nh, p1, p->zn, p->zs);
  *plan = p;
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_get_tuning(const accfft_amd_plan* p, int choice[3]) {
  if (!p || !choice) return ACCFFT_AMD_ERR_NULL;
  for (int ax = 0; ax < 3; ++ax)
    choice[ax] = p->fp == ACCFFT_AMD_F32 ? p->tf[ax].s4v : p->td[ax].s4v;
  if (p->kind == 1 && ysub_built(p) && !p->yfuse) choice[1] = -3;
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_set_tuning(accfft_amd_plan* p, int axis, int choice) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (p->kind == 1 && axis == 1 && ysub_built(p)) {
    if (choice == -3) {
      p->yfuse = false;
      return ACCFFT_AMD_OK;
    }
    if (choice == -1) p->yfuse = true;
  }
  if (axis < 0 || axis > 2 || choice < -2) return ACCFFT_AMD_ERR_INVALID;
  const int n = p->n[axis];
  const bool has = p->fp == ACCFFT_AMD_F32 ? p->tf[axis].tw_sub != nullptr
                                           : p->td[axis].tw_sub != nullptr;
  if (choice >= 0) {
    const bool ok = p->fp == ACCFFT_AMD_F32 ? afft::s4_config_ok<float>(n, choice)
                                            : afft::s4_config_ok<double>(n, choice);
    if (!ok || !has) return ACCFFT_AMD_ERR_INVALID;
  }
  if (p->fp == ACCFFT_AMD_F32) {
    p->tf[axis].s4v = choice;
    p->tf[axis].xg = -1;
  } else {
    p->td[axis].s4v = choice;
    p->td[axis].xg = -1;
  }
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_set_exchange(accfft_amd_plan* p, accfft_amd_exchange_fn fn, void* user) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 3) return ACCFFT_AMD_ERR_NOT_BAKED;
  p->exchange = fn;
  p->exchange_user = user;
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_set_exchange_v(accfft_amd_plan* p, accfft_amd_exchange_v_fn fn, void* user) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 4) return ACCFFT_AMD_ERR_NOT_BAKED;
  p->exchange_v = fn;
  p->exchange_user = user;
  return ACCFFT_AMD_OK;
}

int accfft_amd_destroy_plan(accfft_amd_plan* plan) {
  if (!plan) return ACCFFT_AMD_ERR_NULL;
  free_plan(plan);
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_set_stream(accfft_amd_plan* plan, void* stream) {
  if (!plan) return ACCFFT_AMD_ERR_NULL;
  plan->stream = (hipStream_t)stream;
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_info(const accfft_amd_plan* p, int n[3], int isize[3], int istart[3],
                         int osize[3], int ostart[3], int64_t* alloc_max, int* fp, int* kind,
                         int* inplace) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  for (int i = 0; i < 3; ++i) {
    if (n) n[i] = p->n[i];
    if (isize) isize[i] = p->isize[i];
    if (istart) istart[i] = p->istart[i];
    if (osize) osize[i] = p->osize[i];
    if (ostart) ostart[i] = p->ostart[i];
  }
  if (alloc_max) *alloc_max = p->alloc_max;
  if (fp) *fp = p->fp;
  if (kind) *kind = p->kind;
  if (inplace) *inplace = p->inplace ? 1 : 0;
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_data(const accfft_amd_plan* p, void** data, void** data_out) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (data) *data = p->data;
  if (data_out) *data_out = p->data_out;
  return ACCFFT_AMD_OK;
}

int accfft_amd_execute_r2c(accfft_amd_plan* p, const void* in, void* out, double* timer,
                           unsigned xyz) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (p->kind < 1 || p->kind == 2 || p->kind > 4) return ACCFFT_AMD_ERR_NOT_BAKED;
  if (!in) in = p->data;
  if (!out) out = p->data_out;
  if ((!in && !real_empty(p)) || (!out && !cplx_empty(p))) return ACCFFT_AMD_ERR_NULL;
  xyz &= 7u;
  if (p->kind == 3) {
    if (xyz != 7u) return ACCFFT_AMD_ERR_UNSUPPORTED;
    return timed(p, timer, [&] {
      return p->fp == ACCFFT_AMD_F32 ? afft_r2c_slab_f32(p, in, out) : afft_r2c_slab_f64(p, in, out);
    });
  }
  if (p->kind == 4) {
    if (xyz != 7u) return ACCFFT_AMD_ERR_UNSUPPORTED;
    if (in && in == out) return ACCFFT_AMD_ERR_UNSUPPORTED;
    return timed(p, timer, [&] {
      return p->fp == ACCFFT_AMD_F32 ? afft_r2c_pencil_f32(p, in, out)
                                     : afft_r2c_pencil_f64(p, in, out);
    });
  }
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_r2c_f32(p, in, out, xyz) : afft_r2c_f64(p, in, out, xyz);
  });
}

int accfft_amd_execute_c2r(accfft_amd_plan* p, void* in, void* out, double* timer,
                           unsigned xyz) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (p->kind < 1 || p->kind == 2 || p->kind > 4) return ACCFFT_AMD_ERR_NOT_BAKED;
  if (!in) in = p->data_out;
  if (!out) out = p->data;
  if ((!in && !cplx_empty(p)) || (!out && !real_empty(p))) return ACCFFT_AMD_ERR_NULL;
  xyz &= 7u;
  if (p->kind == 3) {
    if (xyz != 7u) return ACCFFT_AMD_ERR_UNSUPPORTED;
    return timed(p, timer, [&] {
      return p->fp == ACCFFT_AMD_F32 ? afft_c2r_slab_f32(p, in, out) : afft_c2r_slab_f64(p, in, out);
    });
  }
  if (p->kind == 4) {
    if (xyz != 7u) return ACCFFT_AMD_ERR_UNSUPPORTED;
    if (in && in == out) return ACCFFT_AMD_ERR_UNSUPPORTED;
    return timed(p, timer, [&] {
      return p->fp == ACCFFT_AMD_F32 ? afft_c2r_pencil_f32(p, in, out)
                                     : afft_c2r_pencil_f64(p, in, out);
    });
  }
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_c2r_f32(p, in, out, xyz) : afft_c2r_f64(p, in, out, xyz);
  });
}

int accfft_amd_execute_c2c(accfft_amd_plan* p, int direction, void* in, void* out,
                           double* timer, unsigned xyz) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 2) return ACCFFT_AMD_ERR_NOT_BAKED;
  if (direction != -1 && direction != 1) return ACCFFT_AMD_ERR_INVALID;
  if (!in) in = p->data;
  if (!out) out = p->data_out;
  if (!in || !out) return ACCFFT_AMD_ERR_NULL;
  xyz &= 7u;
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_c2c_f32(p, direction, in, out, xyz)
                                   : afft_c2c_f64(p, direction, in, out, xyz);
  });
}

int accfft_amd_execute_r2c_axis(accfft_amd_plan* p, int axis, const void* in, void* out,
                                double* timer) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 1) return ACCFFT_AMD_ERR_NOT_BAKED;
  if (axis < 0 || axis > 2) return ACCFFT_AMD_ERR_INVALID;
  if (!in) in = p->data;
  if (!out) out = p->data_out;
  if (!in || !out) return ACCFFT_AMD_ERR_NULL;
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_r2c_axis_f32(p, axis, in, out)
                                   : afft_r2c_axis_f64(p, axis, in, out);
  });
}

int accfft_amd_execute_c2r_axis(accfft_amd_plan* p, int axis, const void* in, void* out,
                                double* timer) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 1) return ACCFFT_AMD_ERR_NOT_BAKED;
  if (axis < 0 || axis > 2) return ACCFFT_AMD_ERR_INVALID;
  if (!in) in = p->data_out;
  if (!out) out = p->data;
  if (!in || !out) return ACCFFT_AMD_ERR_NULL;
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_c2r_axis_f32(p, axis, in, out)
                                   : afft_c2r_axis_f64(p, axis, in, out);
  });
}

int accfft_amd_grad(accfft_amd_plan* p, void* Ax, void* Ay, void* Az, const void* A,
                    unsigned mask, double* timer, int slow) {
  if (!p || !A) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 1) return ACCFFT_AMD_ERR_NOT_BAKED;
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_grad_f32(p, Ax, Ay, Az, A, mask, slow)
                                   : afft_grad_f64(p, Ax, Ay, Az, A, mask, slow);
  });
}

static int spectral(accfft_amd_plan* p, int sym, void* out, const void* A, double* timer) {
  if (!p || !A || !out) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 1) return ACCFFT_AMD_ERR_NOT_BAKED;
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_spectral_f32(p, sym, out, A)
                                   : afft_spectral_f64(p, sym, out, A);
  });
}

int accfft_amd_laplace(accfft_amd_plan* p, void* LA, const void* A, double* timer) {
  return spectral(p, afft::SYM_LAPLACE, LA, A, timer);
}
int accfft_amd_biharmonic(accfft_amd_plan* p, void* BA, const void* A, double* timer) {
  return spectral(p, afft::SYM_BIHARMONIC, BA, A, timer);
}
int accfft_amd_inv_laplace(accfft_amd_plan* p, void* out, const void* A, double* timer) {
  return spectral(p, afft::SYM_INV_LAPLACE, out, A, timer);
}
int accfft_amd_inv_laplace_noalloc(accfft_amd_plan* p, void* out, const void* A, double* timer) {
  if (!p || !A || !out) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 1) return ACCFFT_AMD_ERR_NOT_BAKED;
  const int64_t cb = (int64_t)p->n[0] * p->n[1] * (p->n[2] / 2 + 1) * 2 * elem_of(p->fp);
  if (cb > (int64_t)ACCFFT_AMD_NOALLOC_MAX_BYTES) return ACCFFT_AMD_ERR_SCRATCH;
  // lines longer than one LDS tile run through workspaces of the plan's own
  for (int ax = 0; ax < 3; ++ax)
    if (p->fp == ACCFFT_AMD_F32 ? p->tf[ax].longl : p->td[ax].longl)
      return ACCFFT_AMD_ERR_UNSUPPORTED;
  int device = -1;
  if (hipGetDevice(&device) != hipSuccess) return ACCFFT_AMD_ERR_HIP;
  if (device != p->device) return ACCFFT_AMD_ERR_INVALID;
  void* ws = nullptr;
  int st = scratch_address(device, &ws);
  if (st) return st;
  // per-launch profiling would create events; it is suspended for this call
  const bool prof = p->prof;
  p->prof = false;
  st = timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_spectral_ws_f32(p, afft::SYM_INV_LAPLACE, out, A, ws)
                                   : afft_spectral_ws_f64(p, afft::SYM_INV_LAPLACE, out, A, ws);
  });
  p->prof = prof;
  return st;
}

int accfft_amd_inv_biharmonic(accfft_amd_plan* p, void* out, const void* A, double* timer) {
  return spectral(p, afft::SYM_INV_BIHARMONIC, out, A, timer);
}
int accfft_amd_poisson_solve(accfft_amd_plan* p, int method, void* phi, const void* rho,
                             double* timer) {
  if (method != 0 && method != 1) return ACCFFT_AMD_ERR_INVALID;
  return spectral(p, method == 0 ? afft::SYM_POISSON_DISCRETE : afft::SYM_POISSON_CONTINUOUS,
                  phi, rho, timer);
}

int accfft_amd_divergence(accfft_amd_plan* p, void* divA, const void* Ax, const void* Ay,
                          const void* Az, double* timer, int slow) {
  if (!p || !divA || !Ax || !Ay || !Az) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 1) return ACCFFT_AMD_ERR_NOT_BAKED;
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_div_f32(p, divA, Ax, Ay, Az, slow)
                                   : afft_div_f64(p, divA, Ax, Ay, Az, slow);
  });
}

int accfft_amd_mult_wave_number(int kind, int fp, void* wA, const void* A, const int n[3],
                                const int osize[3], const int ostart[3], unsigned scale_xyz,
                                void* stream) {
  if (!wA || !A || !n || !osize || !ostart) return ACCFFT_AMD_ERR_NULL;
  if (kind < 0 || kind > 8) return ACCFFT_AMD_ERR_INVALID;
  if (fp != ACCFFT_AMD_F32 && fp != ACCFFT_AMD_F64) return ACCFFT_AMD_ERR_INVALID;
  for (int i = 0; i < 3; ++i)
    if (n[i] <= 0 || osize[i] < 0) return ACCFFT_AMD_ERR_INVALID;
  double scale = 1.0;
  if (kind <= 2) {
    for (int i = 0; i < 3; ++i)
      if (scale_xyz & (1u << i)) scale *= n[i];
  } else {
    scale = (double)n[0] * n[1] * n[2];
  }
  scale = 1.0 / scale;
  return fp == ACCFFT_AMD_F32 ? afft_mult_f32(kind, wA, A, n, osize, ostart, scale, stream)
                              : afft_mult_f64(kind, wA, A, n, osize, ostart, scale, stream);
}

int accfft_amd_poisson_filter(int fp, void* data_hat, const int n[3], const int osize[3],
                              const int ostart[3], int method, void* stream) {
  if (method != 0 && method != 1) return ACCFFT_AMD_ERR_INVALID;
  return accfft_amd_mult_wave_number(method == 0 ? afft::SYM_POISSON_DISCRETE
                                                 : afft::SYM_POISSON_CONTINUOUS,
                                     fp, data_hat, data_hat, n, osize, ostart, 7u, stream);
}

int accfft_amd_axpy(int fp, int64_t n, double alpha, const void* x, void* y, void* stream) {
  if (!x || !y) return ACCFFT_AMD_ERR_NULL;
  if (n < 0) return ACCFFT_AMD_ERR_INVALID;
  return fp == ACCFFT_AMD_F32 ? afft_axpy_f32(n, alpha, x, y, stream)
                              : afft_axpy_f64(n, alpha, x, y, stream);
}

int accfft_amd_plan_sync(accfft_amd_plan* p) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  return hipStreamSynchronize(p->stream) == hipSuccess ? ACCFFT_AMD_OK : ACCFFT_AMD_ERR_HIP;
}

int accfft_amd_adjoint_r2c(accfft_amd_plan* p, const void* G, void* x, double* timer) {
  if (!p || !G || !x) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 1) return ACCFFT_AMD_ERR_NOT_BAKED;
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_adj_r2c_f32(p, G, x) : afft_adj_r2c_f64(p, G, x);
  });
}

int accfft_amd_adjoint_c2r(accfft_amd_plan* p, const void* y, void* G, double* timer) {
  if (!p || !y || !G) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 1) return ACCFFT_AMD_ERR_NOT_BAKED;
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_adj_c2r_f32(p, y, G) : afft_adj_c2r_f64(p, y, G);
  });
}

int accfft_amd_adjoint_grad(accfft_amd_plan* p, void* out, const void* gx, const void* gy,
                            const void* gz, double* timer) {
  if (!p || !out) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 1) return ACCFFT_AMD_ERR_NOT_BAKED;
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_adj_grad_f32(p, out, gx, gy, gz)
                                   : afft_adj_grad_f64(p, out, gx, gy, gz);
  });
}

int accfft_amd_adjoint_divergence(accfft_amd_plan* p, void* ox, void* oy, void* oz,
                                  const void* v, double* timer) {
  if (!p || !v) return ACCFFT_AMD_ERR_NULL;
  if (p->kind != 1) return ACCFFT_AMD_ERR_NOT_BAKED;
  return timed(p, timer, [&] {
    return p->fp == ACCFFT_AMD_F32 ? afft_adj_div_f32(p, ox, oy, oz, v)
                                   : afft_adj_div_f64(p, ox, oy, oz, v);
  });
}

int accfft_amd_plan_release_workspace(accfft_amd_plan* p) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (hipStreamSynchronize(p->stream) != hipSuccess) return ACCFFT_AMD_ERR_HIP;
  for (int i = 0; i < accfft_amd_plan::NWS; ++i) {
    if (p->ws[i]) (void)hipFree(p->ws[i]);
    p->ws[i] = nullptr;
    p->ws_bytes[i] = 0;
  }
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_workspace_bytes(const accfft_amd_plan* p, int64_t* bytes) {
  if (!p || !bytes) return ACCFFT_AMD_ERR_NULL;
  int64_t total = 0;
  for (int i = 0; i < accfft_amd_plan::NWS; ++i) total += (int64_t)p->ws_bytes[i];
  *bytes = total;
  return ACCFFT_AMD_OK;
}

int accfft_amd_plan_set_profiling(accfft_amd_plan* p, int enable) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  if (!enable) {
    int st = afft::prof_collect(p);
    p->prof = false;
    return st;
  }
  p->prof = true;
  return ACCFFT_AMD_OK;
}

int accfft_amd_profile_reset(accfft_amd_plan* p) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  int st = afft::prof_collect(p);
  p->prof_kinds.clear();
  return st;
}

int accfft_amd_profile_count(accfft_amd_plan* p, int* nkinds) {
  if (!p || !nkinds) return ACCFFT_AMD_ERR_NULL;
  int st = afft::prof_collect(p);
  *nkinds = (int)p->prof_kinds.size();
  return st;
}

int accfft_amd_profile_get(accfft_amd_plan* p, int idx, char* name, int name_len,
                           double* total_ms, int64_t* launches, double* bytes_per_launch) {
  if (!p) return ACCFFT_AMD_ERR_NULL;
  int st = afft::prof_collect(p);
  if (st) return st;
  if (idx < 0 || idx >= (int)p->prof_kinds.size()) return ACCFFT_AMD_ERR_INVALID;
  const auto& k = p->prof_kinds[idx];
  if (name && name_len > 0) {
    strncpy(name, k.name.c_str(), (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  if (total_ms) *total_ms = k.total_ms;
  if (launches) *launches = k.launches;
  if (bytes_per_launch) *bytes_per_launch = k.bytes;
  return ACCFFT_AMD_OK;
}

int accfft_amd_last_elapsed_ms(const accfft_amd_plan* p, float* ms) {
  if (!p || !ms) return ACCFFT_AMD_ERR_NULL;
  *ms = p->last_ms;
  return ACCFFT_AMD_OK;
}

}  // extern "C"

// transpose.hip -- local (single-GPU) tuple transposes.
//
// Same permutations as the kernels of src/transpose_cuda.cu (and their CPU
// definitions at src/transpose.cpp:664-769):
//   plain   [r][c][t] -> [c][r][t]                          (cuda1, :6-11)
//   ragged  unpack; row r-1 holds tuples of t2 elements      (cuda2, :13-21)
//   col     pack; column c-1 holds tuples of t2 elements     (cuda4, :59-66)
//   inplace [r][c][t] -> [c][r][t] in place, by cycles       (cuda3, :23-57)
// and the batched row-segment copies of the pencil exchange
//   pencil pack    memcpy_v1_h1 (:97-143)
//   pencil unpack  memcpy_v1_h2 (:145-197)
// whose segments the reference lists in host pointer tables uploaded per
// call (memcpy_cuda, :68-75); here each workgroup computes its segment from
// the per-rank counts and starts, passed by value.
// Each workgroup moves one tuple (or a 32 KiB chunk of a long tuple) with
// the widest access the tuple's alignment allows, so long tuples stream at
// full width; short tuples are one small copy each.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/accfft_amd.h"

namespace {

constexpr int TPB = 256;
constexpr int64_t CHUNK = 32768;  // bytes of a tuple per workgroup

// Copies len bytes from s to d; every thread of the block participates.
__device__ __forceinline__ void copy_bytes(const char* __restrict__ s, char* __restrict__ d,
                                           int64_t len) {
  const uintptr_t al = ((uintptr_t)s | (uintptr_t)d | (uintptr_t)len);
  if ((al & 15) == 0) {
    const uint4* s4 = (const uint4*)s;
    uint4* d4 = (uint4*)d;
    for (int64_t i = threadIdx.x; i < len / 16; i += TPB) d4[i] = s4[i];
  } else if ((al & 7) == 0) {
    const uint2* s2 = (const uint2*)s;
    uint2* d2 = (uint2*)d;
    for (int64_t i = threadIdx.x; i < len / 8; i += TPB) d2[i] = s2[i];
  } else if ((al & 3) == 0) {
    const uint32_t* s1 = (const uint32_t*)s;
    uint32_t* d1 = (uint32_t*)d;
    for (int64_t i = threadIdx.x; i < len / 4; i += TPB) d1[i] = s1[i];
  } else {
    for (int64_t i = threadIdx.x; i < len; i += TPB) d[i] = s[i];
  }
}

// kind 0: plain, 1: ragged last row (unpack), 2: ragged last column (pack).
// Block x = tuple (i, j) with i < r, j < c; block y = chunk of the tuple.
__global__ void __launch_bounds__(TPB)
    tuple_transpose_kernel(int kind, int r, int c, int64_t t, int64_t t2,
                           const char* __restrict__ in, char* __restrict__ out, int esz) {
  const int64_t b = blockIdx.x;
  const int i = (int)(b / c), j = (int)(b % c);
  int64_t nt, src, dst;  // in elements
  if (kind == 0) {
    nt = t;
    src = ((int64_t)i * c + j) * t;
    dst = ((int64_t)j * r + i) * t;
  } else if (kind == 1) {
    nt = (i == r - 1) ? t2 : t;
    src = (int64_t)i * c * t + (int64_t)j * nt;
    dst = (int64_t)j * ((int64_t)(r - 1) * t + t2) + (int64_t)i * t;
  } else {
    nt = (j == c - 1) ? t2 : t;
    src = (int64_t)i * ((int64_t)(c - 1) * t + t2) + (int64_t)j * t;
    dst = (int64_t)j * r * t + (int64_t)i * nt;
  }
  const int64_t len = nt * esz;
  for (int64_t off = (int64_t)blockIdx.y * CHUNK; off < len; off += (int64_t)gridDim.y * CHUNK) {
    const int64_t l = (len - off) < CHUNK ? (len - off) : CHUNK;
    copy_bytes(in + src * esz + off, out + dst * esz + off, l);
  }
}

// In-place [r][c] tuple transpose by cycle following.  Input tuple p goes to
// f(p) = (p % c) * r + p / c; the block whose p is the smallest index of its
// cycle moves the whole cycle, each thread its own byte offsets, so no
// inter-thread synchronisation is needed.
__global__ void __launch_bounds__(TPB)
    tuple_transpose_inplace_kernel(int r, int c, int64_t t, char* __restrict__ A, int esz) {
  const int64_t n = (int64_t)r * c;
  const int64_t p = blockIdx.x;
  if (p >= n) return;
  auto f = [&](int64_t q) { return (q % c) * r + q / c; };
  auto finv = [&](int64_t q) { return (q % r) * c + q / r; };
  if (f(p) == p) return;
  for (int64_t q = f(p); q != p; q = f(q))
    if (q < p) return;  // not the cycle leader
  const int64_t L = t * esz;
  const uintptr_t al = (uintptr_t)A | (uintptr_t)L;
  if ((al & 15) == 0) {
    uint4* V = (uint4*)A;
    const int64_t Lv = L / 16;
    for (int64_t o = threadIdx.x; o < Lv; o += TPB) {
      const uint4 tmp = V[p * Lv + o];
      int64_t cur = p;
      for (int64_t prev = finv(cur); prev != p; prev = finv(cur)) {
        V[cur * Lv + o] = V[prev * Lv + o];
        cur = prev;
      }
      V[cur * Lv + o] = tmp;
    }
  } else if ((al & 3) == 0) {
    uint32_t* V = (uint32_t*)A;
    const int64_t Lv = L / 4;
    for (int64_t o = threadIdx.x; o < Lv; o += TPB) {
      const uint32_t tmp = V[p * Lv + o];
      int64_t cur = p;
      for (int64_t prev = finv(cur); prev != p; prev = finv(cur)) {
        V[cur * Lv + o] = V[prev * Lv + o];
        cur = prev;
      }
      V[cur * Lv + o] = tmp;
    }
  } else {
    for (int64_t o = threadIdx.x; o < L; o += TPB) {
      const char tmp = A[p * L + o];
      int64_t cur = p;
      for (int64_t prev = finv(cur); prev != p; prev = finv(cur)) {
        A[cur * L + o] = A[prev * L + o];
        cur = prev;
      }
      A[cur * L + o] = tmp;
    }
  }
}

// Per-rank counts (a), starts (b) and prefix sums (pre) of a pencil exchange.
constexpr int MAXP = 64;
struct RankTab {
  int64_t a[MAXP], b[MAXP], pre[MAXP + 1];
};

// Pack (memcpy_v1_h1): segment (proc, h, i) copies the cnt[proc] tuples of
// row i of batch h that start at column start[proc] into the send buffer
// [proc][h][i][cnt[proc] tuples].  a = cnt, b = start, pre = prefix of cnt.
// scatter = 1: the inverse copy (packed segment -> its place in the rows).
__global__ void __launch_bounds__(TPB)
    pencil_pack_kernel(int64_t howmany, int64_t local_n0, int64_t t, int64_t N1, int64_t idist,
                       RankTab rk, const char* __restrict__ data, char* __restrict__ send,
                       int esz, int scatter) {
  const int64_t per = howmany * local_n0;
  const int64_t seg = blockIdx.x;
  const int proc = (int)(seg / per);
  const int64_t r = seg - (int64_t)proc * per;
  const int64_t h = r / local_n0, i = r - h * local_n0;
  const int64_t cnt = rk.a[proc];
  const int64_t src = h * idist + (i * N1 + rk.b[proc]) * t;
  const int64_t dst = t * (rk.pre[proc] * per + r * cnt);
  const int64_t len = cnt * t * esz;
  for (int64_t off = (int64_t)blockIdx.y * CHUNK; off < len; off += (int64_t)gridDim.y * CHUNK) {
    const int64_t l = (len - off) < CHUNK ? (len - off) : CHUNK;
    if (scatter) copy_bytes(send + dst * esz + off, const_cast<char*>(data) + src * esz + off, l);
    else copy_bytes(data + src * esz + off, send + dst * esz + off, l);
  }
}

// Unpack (memcpy_v1_h2): the received buffer is a sequence of segments of
// local_n1 * t elements, (proc, h, i) for i in proc's row range
// [start0[proc], start0[proc] + n0[proc]); segment goes to row i of batch h.
// a = n0, b = start0, pre = prefix of howmany * n0.
__global__ void __launch_bounds__(TPB)
    pencil_unpack_kernel(int nprocs, int64_t howmany, int64_t local_n1, int64_t t, int64_t odist,
                         RankTab rk, const char* __restrict__ recv, char* __restrict__ data,
                         int esz) {
  const int64_t seg = blockIdx.x;
  int proc = 0;
  while (proc + 1 < nprocs && rk.pre[proc + 1] <= seg) ++proc;
  const int64_t r = seg - rk.pre[proc];
  const int64_t h = r / rk.a[proc], i = rk.b[proc] + (r - h * rk.a[proc]);
  const int64_t seglen = local_n1 * t;
  const int64_t src = seg * seglen;
  const int64_t dst = h * odist + i * seglen;
  const int64_t len = seglen * esz;
  for (int64_t off = (int64_t)blockIdx.y * CHUNK; off < len; off += (int64_t)gridDim.y * CHUNK) {
    const int64_t l = (len - off) < CHUNK ? (len - off) : CHUNK;
    copy_bytes(recv + src * esz + off, data + dst * esz + off, l);
  }
}

unsigned chunks_for(int64_t bytes) {
  int64_t c = (bytes + CHUNK - 1) / CHUNK;
  if (c < 1) c = 1;
  if (c > 65535) c = 65535;
  return (unsigned)c;
}

int launch_tuples(int kind, int r, int c, int64_t t, int64_t t2, const void* in, void* out,
                  int esz, void* stream) {
  if (r < 0 || c < 0 || t < 0 || t2 < 0 || esz <= 0) return ACCFFT_AMD_ERR_INVALID;
  if (r == 0 || c == 0) return ACCFFT_AMD_OK;  // src/transpose.cpp:667-668
  if (!in || !out) return ACCFFT_AMD_ERR_NULL;
  const int64_t tuples = (int64_t)r * c;
  if (tuples > 0x7fffffffLL) return ACCFFT_AMD_ERR_OVERFLOW;
  const int64_t tmax = (t > t2 ? t : t2) * esz;
  int64_t chunks = (tmax + CHUNK - 1) / CHUNK;
  if (chunks < 1) chunks = 1;
  if (chunks > 65535) chunks = 65535;
  hipLaunchKernelGGL(tuple_transpose_kernel, dim3((unsigned)tuples, (unsigned)chunks), dim3(TPB),
                     0, (hipStream_t)stream, kind, r, c, t, t2, (const char*)in, (char*)out, esz);
  return hipGetLastError() == hipSuccess ? ACCFFT_AMD_OK : ACCFFT_AMD_ERR_HIP;
}

// Column-range pack (memcpy_v1_h1) and, scatter = 1, its inverse.
int pack_cols(int scatter, int nprocs_1, int howmany, int local_n0, int64_t n_tuples,
              const int64_t* local_n1_proc, void* send_recv, void* data, int64_t idist,
              int64_t N1, const int64_t* local_1_start_proc, int elem_size, void* stream) {
  if (nprocs_1 < 1 || nprocs_1 > MAXP || howmany < 0 || local_n0 < 0 || n_tuples < 0 ||
      elem_size <= 0 || idist < 0 || N1 < 0)
    return ACCFFT_AMD_ERR_INVALID;
  if (!local_n1_proc || !local_1_start_proc) return ACCFFT_AMD_ERR_NULL;
  RankTab rk{};
  int64_t maxlen = 0;
  rk.pre[0] = 0;
  for (int p = 0; p < nprocs_1; ++p) {
    const int64_t cnt = local_n1_proc[p], st = local_1_start_proc[p];
    if (cnt < 0 || st < 0 || st + cnt > N1) return ACCFFT_AMD_ERR_INVALID;
    rk.a[p] = cnt;
    rk.b[p] = st;
    rk.pre[p + 1] = rk.pre[p] + cnt;
    if (cnt * n_tuples > maxlen) maxlen = cnt * n_tuples;
  }
  const int64_t segs = (int64_t)nprocs_1 * howmany * local_n0;
  if (segs == 0 || maxlen == 0) return ACCFFT_AMD_OK;
  if (!send_recv || !data) return ACCFFT_AMD_ERR_NULL;
  if (segs > 0x7fffffffLL) return ACCFFT_AMD_ERR_OVERFLOW;
  hipLaunchKernelGGL(pencil_pack_kernel, dim3((unsigned)segs, chunks_for(maxlen * elem_size)),
                     dim3(TPB), 0, (hipStream_t)stream, (int64_t)howmany, (int64_t)local_n0,
                     n_tuples, N1, idist, rk, (const char*)data, (char*)send_recv, elem_size, scatter);
  return hipGetLastError() == hipSuccess ? ACCFFT_AMD_OK : ACCFFT_AMD_ERR_HIP;
}


}  // namespace

// The inverse of accfft_amd_pencil_pack (used by the pencil share's inverse
// transpose): packed segment (proc, h, i) goes back to row i of batch h,
// columns [start[proc], start[proc] + cnt[proc]).
int afft_pencil_scatter(int nprocs, int howmany, int local_n0, int64_t n_tuples,
                        const int64_t* cnt, const void* packed, void* data, int64_t idist,
                        int64_t N1, const int64_t* start, int elem_size, void* stream) {
  return pack_cols(1, nprocs, howmany, local_n0, n_tuples, cnt, const_cast<void*>(packed), data,
                   idist, N1, start, elem_size, stream);
}

extern "C" {

int accfft_amd_local_transpose(int r, int c, int64_t t, const void* in, void* out,
                               int elem_size, void* stream) {
  return launch_tuples(0, r, c, t, t, in, out, elem_size, stream);
}

int accfft_amd_local_transpose_ragged(int r, int c, int64_t t, int64_t t2, const void* in,
                                      void* out, int elem_size, void* stream) {
  return launch_tuples(1, r, c, t, t2, in, out, elem_size, stream);
}

int accfft_amd_local_transpose_col(int r, int c, int64_t t, int64_t t2, const void* in,
                                   void* out, int elem_size, void* stream) {
  return launch_tuples(2, r, c, t, t2, in, out, elem_size, stream);
}

int accfft_amd_local_transpose_inplace(int r, int c, int64_t t, void* A, int elem_size,
                                       void* stream) {
  if (r < 0 || c < 0 || t < 0 || elem_size <= 0) return ACCFFT_AMD_ERR_INVALID;
  if (r == 0 || c == 0 || t == 0) return ACCFFT_AMD_OK;
  if (!A) return ACCFFT_AMD_ERR_NULL;
  const int64_t tuples = (int64_t)r * c;
  if (tuples > 0x7fffffffLL) return ACCFFT_AMD_ERR_OVERFLOW;
  hipLaunchKernelGGL(tuple_transpose_inplace_kernel, dim3((unsigned)tuples), dim3(TPB), 0,
                     (hipStream_t)stream, r, c, t, (char*)A, elem_size);
  return hipGetLastError() == hipSuccess ? ACCFFT_AMD_OK : ACCFFT_AMD_ERR_HIP;
}

int accfft_amd_pencil_pack(int nprocs_1, int howmany, int local_n0, int64_t n_tuples,
                           const int64_t* local_n1_proc, void* send_recv, const void* data,
                           int64_t idist, int64_t N1, const int64_t* local_1_start_proc,
                           int elem_size, void* stream) {
  return pack_cols(0, nprocs_1, howmany, local_n0, n_tuples, local_n1_proc, send_recv,
                   const_cast<void*>(data), idist, N1, local_1_start_proc, elem_size, stream);
}

int accfft_amd_pencil_unpack(int nprocs_0, int howmany, const int64_t* local_0_start_proc,
                             const int64_t* local_n0_proc, void* data, int64_t odist,
                             int64_t local_n1, int64_t n_tuples, const void* send_recv,
                             int elem_size, void* stream) {
  if (nprocs_0 < 1 || nprocs_0 > MAXP || howmany < 0 || local_n1 < 0 || n_tuples < 0 ||
      elem_size <= 0 || odist < 0)
    return ACCFFT_AMD_ERR_INVALID;
  if (!local_0_start_proc || !local_n0_proc) return ACCFFT_AMD_ERR_NULL;
  RankTab rk{};
  rk.pre[0] = 0;
  for (int p = 0; p < nprocs_0; ++p) {
    const int64_t n0 = local_n0_proc[p], st = local_0_start_proc[p];
    if (n0 < 0 || st < 0) return ACCFFT_AMD_ERR_INVALID;
    rk.a[p] = n0;
    rk.b[p] = st;
    rk.pre[p + 1] = rk.pre[p] + (int64_t)howmany * n0;
  }
  const int64_t seglen = local_n1 * n_tuples;
  const int64_t segs = rk.pre[nprocs_0];
  if (segs == 0 || seglen == 0) return ACCFFT_AMD_OK;
  if (!send_recv || !data) return ACCFFT_AMD_ERR_NULL;
  if (segs > 0x7fffffffLL) return ACCFFT_AMD_ERR_OVERFLOW;
  hipLaunchKernelGGL(pencil_unpack_kernel, dim3((unsigned)segs, chunks_for(seglen * elem_size)),
                     dim3(TPB), 0, (hipStream_t)stream, nprocs_0, (int64_t)howmany, local_n1,
                     n_tuples, odist, rk, (const char*)send_recv, (char*)data, elem_size);
  return hipGetLastError() == hipSuccess ? ACCFFT_AMD_OK : ACCFFT_AMD_ERR_HIP;
}

}  // extern "C"

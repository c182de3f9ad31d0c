// transpose.hip -- local (single-GPU) tuple transposes.
//
// Same permutations as the kernels of src/transpose_cuda.cu (and their CPU
// definitions at src/transpose.cpp:664-769):
//   plain   [r][c][t] -> [c][r][t]                          (cuda1, :6-11)
//   ragged  unpack; row r-1 holds tuples of t2 elements      (cuda2, :13-21)
//   col     pack; column c-1 holds tuples of t2 elements     (cuda4, :59-66)
//   inplace [r][c][t] -> [c][r][t] in place, by cycles       (cuda3, :23-57)
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

}  // namespace

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

}  // extern "C"

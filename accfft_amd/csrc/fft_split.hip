// fft_split.hip -- FP32 kernel instances whose radix-16 stages run on the
// matrix cores with split BF16 / FP16 operands (fft_mfma.hpp), and the
// one-wavefront DFT-16 self-test.
#include "../../include/accfft_amd.h"
#include "fft_launch.hpp"

namespace afft {

template <int N, int SPLIT>
static hipError_t launch_split_n(bool strided, int mode, int sign, FftArgs<float>& a,
                                 int64_t units, hipStream_t st) {
  const int per = strided ? SCfg<float, N, SPLIT>::W : ZCfg<N, SPLIT>::LPW;
  a.lpw = per;
  dim3 g((unsigned)((units + per - 1) / per)), b(256);
  if (strided) {
    if (mode != M_C2C) return hipErrorInvalidValue;
    if (sign < 0) hipLaunchKernelGGL((sfft_kernel<float, N, M_C2C, -1, SPLIT>), g, b, 0, st, a);
    else hipLaunchKernelGGL((sfft_kernel<float, N, M_C2C, +1, SPLIT>), g, b, 0, st, a);
  } else {
    switch (mode) {
      case M_C2C:
        if (sign < 0) hipLaunchKernelGGL((zfft_kernel<float, N, M_C2C, -1, SPLIT>), g, b, 0, st, a);
        else hipLaunchKernelGGL((zfft_kernel<float, N, M_C2C, +1, SPLIT>), g, b, 0, st, a);
        break;
      case M_R2C: hipLaunchKernelGGL((zfft_kernel<float, N, M_R2C, -1, SPLIT>), g, b, 0, st, a); break;
      case M_C2R: hipLaunchKernelGGL((zfft_kernel<float, N, M_C2R, +1, SPLIT>), g, b, 0, st, a); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

// Two-level strided tiles whose first radix-16 level runs on the matrix
// cores (configurations that keep lanes l and l^32 in one column; one-shot
// or persistent).
template <int N, int V>
constexpr bool split_s4_ok() {
  using CF = S4Cfg<float, N, V>;
  return CF::OK && CF::R1 == 16 && CF::TWF == 0 && 32 % CF::W == 0 && CF::TPB % 64 == 0 &&
         (CF::KC == 8 || CF::KC == 16);
}

template <int N, int SPLIT>
static hipError_t launch_split_s4(int sign, FftArgs<float>& a, const AxisTab<float>& tab,
                                  hipStream_t st) {
  a.sc = tab.sc;
  a.sc.n = tab.n;
  switch (four_step_variant(N, a.ishift != 0 || a.oshift != 0, tab.s4v)) {
#define AFFT_SV(VV) \
  case VV:          \
    if constexpr (split_s4_ok<N, VV>()) return launch_s4v<float, N, VV, SPLIT>(sign, a, tab, st); \
    break;
    AFFT_SV(1)
    AFFT_SV(2)
    AFFT_SV(3)
#undef AFFT_SV
    default: break;
  }
  return launch_s4v<float, N, 0, SPLIT>(sign, a, tab, st);
}

template <int SPLIT>
static hipError_t launch_split_k(bool strided, int mode, int sign, FftArgs<float>& a,
                                 int64_t units, int n, hipStream_t st) {
  switch (n) {
    case 256: return launch_split_n<256, SPLIT>(strided, mode, sign, a, units, st);
    case 512: return launch_split_n<512, SPLIT>(strided, mode, sign, a, units, st);
    case 1024: return launch_split_n<1024, SPLIT>(strided, mode, sign, a, units, st);
    case 2048: return launch_split_n<2048, SPLIT>(strided, mode, sign, a, units, st);
    case 4096: return launch_split_n<4096, SPLIT>(strided, mode, sign, a, units, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fft_split(bool strided, int mode, int sign, FftArgs<float> a,
                            const AxisTab<float>& tab, int split, hipStream_t st) {
  const int64_t units = pair_mode(mode) ? (a.nlines + 1) / 2 : a.nlines;
  if (units <= 0) return hipSuccess;
  if (strided && mode == M_C2C && tab.tw_sub && four_step<float>(tab) &&
      (tab.n == 512 || tab.n == 1024 || tab.n == 2048)) {
#define AFFT_S4S(NN)                                                                    \
  if (tab.n == NN) switch (split) {                                                     \
      case SPLIT_BF16: return launch_split_s4<NN, SPLIT_BF16>(sign, a, tab, st);        \
      case SPLIT_BF16X3: return launch_split_s4<NN, SPLIT_BF16X3>(sign, a, tab, st);    \
      case SPLIT_FP16: return launch_split_s4<NN, SPLIT_FP16>(sign, a, tab, st);        \
      default: return hipErrorInvalidValue;                                             \
    }
    AFFT_S4S(512)
    AFFT_S4S(1024)
    AFFT_S4S(2048)
#undef AFFT_S4S
  }
  a.sc = tab.sc;
  a.sc.n = tab.n;
  a.tw = tab.tw_split;
  switch (split) {
    case SPLIT_BF16: return launch_split_k<SPLIT_BF16>(strided, mode, sign, a, units, tab.n, st);
    case SPLIT_BF16X3: return launch_split_k<SPLIT_BF16X3>(strided, mode, sign, a, units, tab.n, st);
    case SPLIT_FP16: return launch_split_k<SPLIT_FP16>(strided, mode, sign, a, units, tab.n, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace afft

using afft::cpx;

extern "C" int accfft_amd_selftest_dft16(int split, int sign, const void* x, void* y,
                                         void* stream) {
  if (!x || !y) return ACCFFT_AMD_ERR_NULL;
  if (sign != -1 && sign != 1) return ACCFFT_AMD_ERR_INVALID;
  const cpx<float>* xi = (const cpx<float>*)x;
  cpx<float>* yo = (cpx<float>*)y;
  hipStream_t st = (hipStream_t)stream;
#define AFFT_ST(S)                                                                                  \
  if (sign < 0) hipLaunchKernelGGL((afft::dft16_selftest_kernel<S, -1>), dim3(1), dim3(64), 0, st, xi, yo); \
  else hipLaunchKernelGGL((afft::dft16_selftest_kernel<S, +1>), dim3(1), dim3(64), 0, st, xi, yo);
  switch (split) {
    case afft::SPLIT_BF16: AFFT_ST(afft::SPLIT_BF16) break;
    case afft::SPLIT_BF16X3: AFFT_ST(afft::SPLIT_BF16X3) break;
    case afft::SPLIT_FP16: AFFT_ST(afft::SPLIT_FP16) break;
    default: return ACCFFT_AMD_ERR_INVALID;
  }
#undef AFFT_ST
  return hipGetLastError() == hipSuccess ? ACCFFT_AMD_OK : ACCFFT_AMD_ERR_HIP;
}

// Single-precision instantiation of the FFT kernels and execute paths.
#define AFFT_T float
#define AFFT_SUF f32
#include "fft_prec.inc"

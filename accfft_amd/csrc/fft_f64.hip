// Double-precision instantiation of the FFT kernels and execute paths.
#define AFFT_T double
#define AFFT_SUF f64
#include "fft_prec.inc"

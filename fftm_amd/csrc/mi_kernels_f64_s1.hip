// F64 pass kernels, compiled 2^a 3^b 5^c 7^d lengths (default tile shape): 20, 40, 72, 112, 192, 288, 448, 768, 1000, 1400, 1792.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( double, f64, 20 )
FFTM_MI_DEFINE_LENGTH( double, f64, 40 )
FFTM_MI_DEFINE_LENGTH( double, f64, 72 )
FFTM_MI_DEFINE_LENGTH( double, f64, 112 )
FFTM_MI_DEFINE_LENGTH( double, f64, 192 )
FFTM_MI_DEFINE_LENGTH( double, f64, 288 )
FFTM_MI_DEFINE_LENGTH( double, f64, 448 )
FFTM_MI_DEFINE_LENGTH( double, f64, 768 )
FFTM_MI_DEFINE_LENGTH( double, f64, 1000 )
FFTM_MI_DEFINE_LENGTH( double, f64, 1400 )
FFTM_MI_DEFINE_LENGTH( double, f64, 1792 )

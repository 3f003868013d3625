// F64 pass kernels, compiled 2^a 3^b 5^c 7^d lengths (default tile shape): 24, 48, 80, 120, 200, 320, 480, 800, 1152, 1536, 1920.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( double, f64, 24 )
FFTM_MI_DEFINE_LENGTH( double, f64, 48 )
FFTM_MI_DEFINE_LENGTH( double, f64, 80 )
FFTM_MI_DEFINE_LENGTH( double, f64, 120 )
FFTM_MI_DEFINE_LENGTH( double, f64, 200 )
FFTM_MI_DEFINE_LENGTH( double, f64, 320 )
FFTM_MI_DEFINE_LENGTH( double, f64, 480 )
FFTM_MI_DEFINE_LENGTH( double, f64, 800 )
FFTM_MI_DEFINE_LENGTH( double, f64, 1152 )
FFTM_MI_DEFINE_LENGTH( double, f64, 1536 )
FFTM_MI_DEFINE_LENGTH( double, f64, 1920 )

// F64 pass kernels, compiled 2^a 3^b 5^c 7^d lengths (default tile shape): 12, 36, 60, 100, 160, 240, 400, 640, 960, 1344, 1680, 2100.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( double, f64, 12 )
FFTM_MI_DEFINE_LENGTH( double, f64, 36 )
FFTM_MI_DEFINE_LENGTH( double, f64, 60 )
FFTM_MI_DEFINE_LENGTH( double, f64, 100 )
FFTM_MI_DEFINE_LENGTH( double, f64, 160 )
FFTM_MI_DEFINE_LENGTH( double, f64, 240 )
FFTM_MI_DEFINE_LENGTH( double, f64, 400 )
FFTM_MI_DEFINE_LENGTH( double, f64, 640 )
FFTM_MI_DEFINE_LENGTH( double, f64, 960 )
FFTM_MI_DEFINE_LENGTH( double, f64, 1344 )
FFTM_MI_DEFINE_LENGTH( double, f64, 1680 )
FFTM_MI_DEFINE_LENGTH( double, f64, 2100 )

// F64 pass kernels, compiled 2^a 3^b 5^c 7^d lengths (default tile shape): 28, 56, 96, 144, 224, 384, 576, 896, 1280, 1600, 2000.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( double, f64, 28 )
FFTM_MI_DEFINE_LENGTH( double, f64, 56 )
FFTM_MI_DEFINE_LENGTH( double, f64, 96 )
FFTM_MI_DEFINE_LENGTH( double, f64, 144 )
FFTM_MI_DEFINE_LENGTH( double, f64, 224 )
FFTM_MI_DEFINE_LENGTH( double, f64, 384 )
FFTM_MI_DEFINE_LENGTH( double, f64, 576 )
FFTM_MI_DEFINE_LENGTH( double, f64, 896 )
FFTM_MI_DEFINE_LENGTH( double, f64, 1280 )
FFTM_MI_DEFINE_LENGTH( double, f64, 1600 )
FFTM_MI_DEFINE_LENGTH( double, f64, 2000 )

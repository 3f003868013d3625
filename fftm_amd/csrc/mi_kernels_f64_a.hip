// F64 pass kernels, lengths 2, 4, 8, 16, 32, 64, 128, 256.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( double, f64, 2 )
FFTM_MI_DEFINE_LENGTH( double, f64, 4 )
FFTM_MI_DEFINE_LENGTH( double, f64, 8 )
FFTM_MI_DEFINE_LENGTH( double, f64, 16 )
FFTM_MI_DEFINE_LENGTH( double, f64, 32 )
FFTM_MI_DEFINE_LENGTH( double, f64, 64 )
FFTM_MI_DEFINE_LENGTH( double, f64, 128 )
FFTM_MI_DEFINE_LENGTH( double, f64, 256 )

// F64 pass kernels, lengths 2048, 4096.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( double, f64, 2048 )
FFTM_MI_DEFINE_LENGTH( double, f64, 4096 )

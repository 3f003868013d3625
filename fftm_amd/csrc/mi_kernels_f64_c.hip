// F64 pass kernels, lengths 1024.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( double, f64, 1024 )

// Fused filter pass kernels (forward C2C, x H(k), inverse C2C), FP64.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_FILTER( double, f64 )

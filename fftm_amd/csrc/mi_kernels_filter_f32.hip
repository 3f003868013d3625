// Fused filter pass kernels (forward C2C, x H(k), inverse C2C), FP32.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_FILTER( float, f32 )

// F32 pass kernels, lengths 2048, 4096.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( float, f32, 2048 )
FFTM_MI_DEFINE_LENGTH( float, f32, 4096 )

// F32 pass kernels, lengths 512.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( float, f32, 512 )

// F32 pass kernels, lengths 1024.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( float, f32, 1024 )

// F32 pass kernels, lengths 2, 4, 8, 16, 32, 64, 128, 256.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( float, f32, 2 )
FFTM_MI_DEFINE_LENGTH( float, f32, 4 )
FFTM_MI_DEFINE_LENGTH( float, f32, 8 )
FFTM_MI_DEFINE_LENGTH( float, f32, 16 )
FFTM_MI_DEFINE_LENGTH( float, f32, 32 )
FFTM_MI_DEFINE_LENGTH( float, f32, 64 )
FFTM_MI_DEFINE_LENGTH( float, f32, 128 )
FFTM_MI_DEFINE_LENGTH( float, f32, 256 )

// F32 pass kernels, compiled 2^a 3^b 5^c 7^d lengths (default tile shape): 28, 56, 96, 144, 224, 384, 576, 896, 1280, 1600, 2000.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( float, f32, 28 )
FFTM_MI_DEFINE_LENGTH( float, f32, 56 )
FFTM_MI_DEFINE_LENGTH( float, f32, 96 )
FFTM_MI_DEFINE_LENGTH( float, f32, 144 )
FFTM_MI_DEFINE_LENGTH( float, f32, 224 )
FFTM_MI_DEFINE_LENGTH( float, f32, 384 )
FFTM_MI_DEFINE_LENGTH( float, f32, 576 )
FFTM_MI_DEFINE_LENGTH( float, f32, 896 )
FFTM_MI_DEFINE_LENGTH( float, f32, 1280 )
FFTM_MI_DEFINE_LENGTH( float, f32, 1600 )
FFTM_MI_DEFINE_LENGTH( float, f32, 2000 )

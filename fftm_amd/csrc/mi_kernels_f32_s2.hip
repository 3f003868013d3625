// F32 pass kernels, compiled 2^a 3^b 5^c 7^d lengths (default tile shape): 24, 48, 80, 120, 200, 320, 480, 800, 1152, 1536, 1920.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( float, f32, 24 )
FFTM_MI_DEFINE_LENGTH( float, f32, 48 )
FFTM_MI_DEFINE_LENGTH( float, f32, 80 )
FFTM_MI_DEFINE_LENGTH( float, f32, 120 )
FFTM_MI_DEFINE_LENGTH( float, f32, 200 )
FFTM_MI_DEFINE_LENGTH( float, f32, 320 )
FFTM_MI_DEFINE_LENGTH( float, f32, 480 )
FFTM_MI_DEFINE_LENGTH( float, f32, 800 )
FFTM_MI_DEFINE_LENGTH( float, f32, 1152 )
FFTM_MI_DEFINE_LENGTH( float, f32, 1536 )
FFTM_MI_DEFINE_LENGTH( float, f32, 1920 )

// F32 pass kernels, compiled 2^a 3^b 5^c 7^d lengths (default tile shape): 12, 36, 60, 100, 160, 240, 400, 640, 960, 1344, 1680, 2100.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_LENGTH( float, f32, 12 )
FFTM_MI_DEFINE_LENGTH( float, f32, 36 )
FFTM_MI_DEFINE_LENGTH( float, f32, 60 )
FFTM_MI_DEFINE_LENGTH( float, f32, 100 )
FFTM_MI_DEFINE_LENGTH( float, f32, 160 )
FFTM_MI_DEFINE_LENGTH( float, f32, 240 )
FFTM_MI_DEFINE_LENGTH( float, f32, 400 )
FFTM_MI_DEFINE_LENGTH( float, f32, 640 )
FFTM_MI_DEFINE_LENGTH( float, f32, 960 )
FFTM_MI_DEFINE_LENGTH( float, f32, 1344 )
FFTM_MI_DEFINE_LENGTH( float, f32, 1680 )
FFTM_MI_DEFINE_LENGTH( float, f32, 2100 )

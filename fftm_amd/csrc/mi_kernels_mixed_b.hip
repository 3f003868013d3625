// FP16-storage pass kernels (FP32 arithmetic), lengths 512, 1024.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_MIXED_LENGTH( 512 )
FFTM_MI_DEFINE_MIXED_LENGTH( 1024 )

// hi + lo FP16-storage pass kernels (FP32 arithmetic), lengths 128, 256.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_MIXED_SPLIT_LENGTH( 128 )
FFTM_MI_DEFINE_MIXED_SPLIT_LENGTH( 256 )

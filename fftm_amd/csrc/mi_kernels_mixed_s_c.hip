// hi + lo FP16-storage pass kernels (FP32 arithmetic), lengths 1024.
#include "mi_kernels_impl.hpp"

FFTM_MI_DEFINE_MIXED_SPLIT_LENGTH( 1024 )

// Alternative radix schedules of the strided FP32-arithmetic C2C passes, length 1024
// (8 16 8 on the vector ALUs, and with the radix-16 stage on the matrix cores).
#include "mi_kernels_impl.hpp"

namespace fftm_mi
{
hipError_t launch_rs_f32_n1024( const pass_params<float> &p, bool inv, int shape, int rs, int iom, hipStream_t s )
{
    return detail::launch_rs_n<1024>( p, inv, shape, rs, iom, s );
}
} // namespace fftm_mi

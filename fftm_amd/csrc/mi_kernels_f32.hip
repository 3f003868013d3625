// FP32 instantiations of the pass kernels.
#include "mi_kernels_impl.hpp"

namespace fftm_mi
{

hipError_t launch_pass_f32( const pass_params<float> &p, int N, int kind, bool inv, bool loadb, hipStream_t s )
{
    return detail::launch_pass_t<float>( p, N, kind, inv, loadb, s );
}

hipError_t launch_generic_f32( const pass_params<float> &p, const radix_list &rl, int N, int cols, int kind, bool inv,
                               hipStream_t s )
{
    return detail::launch_generic_t<float>( p, rl, N, cols, kind, inv, s );
}

int fixed_tile_cols_f32( int N, bool loadb )
{
    return detail::tile_cols_t<float>( N, loadb );
}

} // namespace fftm_mi

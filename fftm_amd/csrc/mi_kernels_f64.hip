// FP64 instantiations of the pass kernels.
#include "mi_kernels_impl.hpp"

namespace fftm_mi
{

hipError_t launch_pass_f64( const pass_params<double> &p, int N, int kind, bool inv, bool loadb, hipStream_t s )
{
    return detail::launch_pass_t<double>( p, N, kind, inv, loadb, s );
}

hipError_t launch_generic_f64( const pass_params<double> &p, const radix_list &rl, int N, int cols, int kind, bool inv,
                               hipStream_t s )
{
    return detail::launch_generic_t<double>( p, rl, N, cols, kind, inv, s );
}

int fixed_tile_cols_f64( int N, bool loadb )
{
    return detail::tile_cols_t<double>( N, loadb );
}

} // namespace fftm_mi

// Instantiation and dispatch of the pass kernels for one precision T and one length N.
// Included by the per-length translation units (mi_kernels_<prec>_<group>.hip).
#pragma once

#include "mi_launch.hpp"

namespace fftm_mi
{
namespace detail
{

// tile-shape variants compiled for a length (small lengths have one)
template <int N>
constexpr int variants_for()
{
    return N >= 64 ? NUM_VARIANTS : 1;
}

template <typename T, int N, int KIND, bool INV, bool LOADB, int VAR>
hipError_t launch_fixed( const pass_params<T> &p, hipStream_t s )
{
    using K              = pass_kernel<T, N, KIND, INV, LOADB, VAR>;
    const int64_t per    = (int64_t)K::TW * K::CPT;
    const int64_t ntiles = ( p.ncols + per - 1 ) / per;
    if ( ntiles <= 0 )
        return hipSuccess;
    if ( ntiles > 0x7fffffffLL )
        return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL( ( k_pass<T, N, KIND, INV, LOADB, VAR> ), dim3( (unsigned)ntiles ), dim3( K::NT ), 0, s, p );
    return hipGetLastError();
}

template <typename T, int N, int VAR>
hipError_t launch_nv( const pass_params<T> &p, int kind, bool inv, bool loadb, hipStream_t s )
{
    if ( kind == KIND_C2C )
    {
        if ( inv )
            return loadb ? launch_fixed<T, N, KIND_C2C, true, true, VAR>( p, s )
                         : launch_fixed<T, N, KIND_C2C, true, false, VAR>( p, s );
        return loadb ? launch_fixed<T, N, KIND_C2C, false, true, VAR>( p, s )
                     : launch_fixed<T, N, KIND_C2C, false, false, VAR>( p, s );
    }
    if ( kind == KIND_R2C )
        return loadb ? launch_fixed<T, N, KIND_R2C, false, true, VAR>( p, s )
                     : launch_fixed<T, N, KIND_R2C, false, false, VAR>( p, s );
    return loadb ? launch_fixed<T, N, KIND_C2R, true, true, VAR>( p, s )
                 : launch_fixed<T, N, KIND_C2R, true, false, VAR>( p, s );
}

template <typename T, int N>
hipError_t launch_n( const pass_params<T> &p, int kind, bool inv, bool loadb, int var, hipStream_t s )
{
    if constexpr ( variants_for<N>() > 1 )
    {
        switch ( var )
        {
        case 1: return launch_nv<T, N, 1>( p, kind, inv, loadb, s );
        case 2: return launch_nv<T, N, 2>( p, kind, inv, loadb, s );
        case 3: return launch_nv<T, N, 3>( p, kind, inv, loadb, s );
        case 4: return launch_nv<T, N, 4>( p, kind, inv, loadb, s );
        case 5: return launch_nv<T, N, 5>( p, kind, inv, loadb, s );
        default: break;
        }
    }
    return launch_nv<T, N, 0>( p, kind, inv, loadb, s );
}

template <typename T, int N, int VAR>
int tile_cols_nv( bool loadb )
{
    return loadb ? pass_kernel<T, N, KIND_C2C, false, true, VAR>::TW : pass_kernel<T, N, KIND_C2C, false, false, VAR>::TW;
}

template <typename T, int N>
int tile_cols_n( bool loadb, int var )
{
    if constexpr ( variants_for<N>() > 1 )
    {
        switch ( var )
        {
        case 1: return tile_cols_nv<T, N, 1>( loadb );
        case 2: return tile_cols_nv<T, N, 2>( loadb );
        case 3: return tile_cols_nv<T, N, 3>( loadb );
        case 4: return tile_cols_nv<T, N, 4>( loadb );
        case 5: return tile_cols_nv<T, N, 5>( loadb );
        default: break;
        }
    }
    return tile_cols_nv<T, N, 0>( loadb );
}

template <typename T>
hipError_t launch_generic_t( const pass_params<T> &p, const radix_list &rl, int N, int cols, int kind, bool inv,
                             hipStream_t s )
{
    const int64_t cpt    = kind == KIND_C2C ? 1 : 2;
    const int64_t per    = (int64_t)cols * cpt;
    const int64_t ntiles = ( p.ncols + per - 1 ) / per;
    if ( ntiles <= 0 )
        return hipSuccess;
    if ( ntiles > 0x7fffffffLL )
        return hipErrorInvalidConfiguration;
    const size_t lds = 2 * (size_t)N * (size_t)cols * sizeof( cpx<T> );
    const dim3   grid( (unsigned)ntiles ), block( GEN_THREADS );
    if ( kind == KIND_C2C )
    {
        if ( inv )
            hipLaunchKernelGGL( ( k_pass_generic<T, KIND_C2C, true> ), grid, block, lds, s, p, rl, N, cols );
        else
            hipLaunchKernelGGL( ( k_pass_generic<T, KIND_C2C, false> ), grid, block, lds, s, p, rl, N, cols );
    }
    else if ( kind == KIND_R2C )
        hipLaunchKernelGGL( ( k_pass_generic<T, KIND_R2C, false> ), grid, block, lds, s, p, rl, N, cols );
    else
        hipLaunchKernelGGL( ( k_pass_generic<T, KIND_C2R, true> ), grid, block, lds, s, p, rl, N, cols );
    return hipGetLastError();
}

} // namespace detail
} // namespace fftm_mi

// One length N for precision T: defines launch_<tn>_n<N> and tile_cols_<tn>_n<N>.
#define FFTM_MI_DEFINE_LENGTH( T, TN, N )                                                                              \
    namespace fftm_mi                                                                                                  \
    {                                                                                                                  \
    hipError_t launch_##TN##_n##N( const pass_params<T> &p, int kind, bool inv, bool loadb, int var, hipStream_t s ) \
    {                                                                                                                  \
        return detail::launch_n<T, N>( p, kind, inv, loadb, var, s );                                                  \
    }                                                                                                                  \
    int tile_cols_##TN##_n##N( bool loadb, int var )                                                                   \
    {                                                                                                                  \
        return detail::tile_cols_n<T, N>( loadb, var );                                                                \
    }                                                                                                                  \
    }

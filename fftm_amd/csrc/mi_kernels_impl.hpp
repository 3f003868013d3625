// Instantiation and dispatch of the pass kernels for one precision T.
// Included once per precision by mi_kernels_f32.hip / mi_kernels_f64.hip.
#pragma once

#include "mi_launch.hpp"

namespace fftm_mi
{
namespace detail
{

template <typename T, int N, int KIND, bool INV, bool LOADB>
hipError_t launch_fixed( const pass_params<T> &p, hipStream_t s )
{
    using K                = pass_kernel<T, N, KIND, INV, LOADB>;
    const int64_t per      = (int64_t)K::TW * K::CPT;
    const int64_t ntiles   = ( p.ncols + per - 1 ) / per;
    if ( ntiles <= 0 )
        return hipSuccess;
    if ( ntiles > 0x7fffffffLL )
        return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL( ( k_pass<T, N, KIND, INV, LOADB> ), dim3( (unsigned)ntiles ), dim3( K::NT ), 0, s, p );
    return hipGetLastError();
}

template <typename T, int N>
hipError_t launch_n( const pass_params<T> &p, int kind, bool inv, bool loadb, hipStream_t s )
{
    if ( kind == KIND_C2C )
    {
        if ( inv )
            return loadb ? launch_fixed<T, N, KIND_C2C, true, true>( p, s )
                         : launch_fixed<T, N, KIND_C2C, true, false>( p, s );
        return loadb ? launch_fixed<T, N, KIND_C2C, false, true>( p, s )
                     : launch_fixed<T, N, KIND_C2C, false, false>( p, s );
    }
    if ( kind == KIND_R2C )
        return loadb ? launch_fixed<T, N, KIND_R2C, false, true>( p, s )
                     : launch_fixed<T, N, KIND_R2C, false, false>( p, s );
    return loadb ? launch_fixed<T, N, KIND_C2R, true, true>( p, s ) : launch_fixed<T, N, KIND_C2R, true, false>( p, s );
}

template <typename T>
hipError_t launch_pass_t( const pass_params<T> &p, int N, int kind, bool inv, bool loadb, hipStream_t s )
{
    switch ( N )
    {
    case 2: return launch_n<T, 2>( p, kind, inv, loadb, s );
    case 4: return launch_n<T, 4>( p, kind, inv, loadb, s );
    case 8: return launch_n<T, 8>( p, kind, inv, loadb, s );
    case 16: return launch_n<T, 16>( p, kind, inv, loadb, s );
    case 32: return launch_n<T, 32>( p, kind, inv, loadb, s );
    case 64: return launch_n<T, 64>( p, kind, inv, loadb, s );
    case 128: return launch_n<T, 128>( p, kind, inv, loadb, s );
    case 256: return launch_n<T, 256>( p, kind, inv, loadb, s );
    case 512: return launch_n<T, 512>( p, kind, inv, loadb, s );
    case 1024: return launch_n<T, 1024>( p, kind, inv, loadb, s );
    case 2048: return launch_n<T, 2048>( p, kind, inv, loadb, s );
    case 4096: return launch_n<T, 4096>( p, kind, inv, loadb, s );
    default: return hipErrorInvalidValue;
    }
}

template <typename T, int N>
int tile_cols_n( bool loadb )
{
    return loadb ? pass_kernel<T, N, KIND_C2C, false, true>::TW : pass_kernel<T, N, KIND_C2C, false, false>::TW;
}

template <typename T>
int tile_cols_t( int N, bool loadb )
{
    switch ( N )
    {
    case 2: return tile_cols_n<T, 2>( loadb );
    case 4: return tile_cols_n<T, 4>( loadb );
    case 8: return tile_cols_n<T, 8>( loadb );
    case 16: return tile_cols_n<T, 16>( loadb );
    case 32: return tile_cols_n<T, 32>( loadb );
    case 64: return tile_cols_n<T, 64>( loadb );
    case 128: return tile_cols_n<T, 128>( loadb );
    case 256: return tile_cols_n<T, 256>( loadb );
    case 512: return tile_cols_n<T, 512>( loadb );
    case 1024: return tile_cols_n<T, 1024>( loadb );
    case 2048: return tile_cols_n<T, 2048>( loadb );
    case 4096: return tile_cols_n<T, 4096>( loadb );
    default: return 0;
    }
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

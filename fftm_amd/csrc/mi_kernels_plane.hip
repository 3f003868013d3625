// Plane-pass kernel instantiations (square planes 16..128, FP32 and FP64).
#include "mi_launch.hpp"
#include "mi_plane.hpp"

namespace fftm_mi
{
namespace
{

template <typename T, int N>
hipError_t launch_plane_n( const plane_params<T> &p, bool fwd, hipStream_t s )
{
    // planes of 64 x 64 and smaller: two planes per workgroup share the barriers
    constexpr int PPW = N <= 64 ? 2 : 1;
    using G           = plane_geom<N, N, (int)sizeof( cpx<T> ), PPW>;
    if ( p.nplanes <= 0 )
        return hipSuccess;
    const int64_t groups = ( p.nplanes + PPW - 1 ) / PPW;
    if ( groups > 0x7fffffffLL )
        return hipErrorInvalidConfiguration;
    const dim3 grid( (unsigned)groups ), block( G::NT );
    if ( fwd )
        hipLaunchKernelGGL( ( k_plane<T, N, N, true, PPW> ), grid, block, 0, s, p );
    else
        hipLaunchKernelGGL( ( k_plane<T, N, N, false, PPW> ), grid, block, 0, s, p );
    return hipGetLastError();
}

template <typename T, int N>
hipError_t launch_plane_rx_n( const plane_params<T> &p, bool fwd, hipStream_t s )
{
    using G = plane_rx_geom<N, (int)sizeof( cpx<T> )>;
    if ( p.nplanes <= 0 )
        return hipSuccess;
    const int64_t groups = ( p.nplanes + G::PPW - 1 ) / G::PPW;
    if ( groups > 0x7fffffffLL )
        return hipErrorInvalidConfiguration;
    const dim3 grid( (unsigned)groups ), block( G::NT );
    // real rows staged through LDS unless B is contiguous in the real array
    const bool staged = p.rb != 1;
    if ( fwd && staged )
        hipLaunchKernelGGL( ( k_plane_rx<T, N, true, true> ), grid, block, 0, s, p );
    else if ( fwd )
        hipLaunchKernelGGL( ( k_plane_rx<T, N, true, false> ), grid, block, 0, s, p );
    else if ( staged )
        hipLaunchKernelGGL( ( k_plane_rx<T, N, false, true> ), grid, block, 0, s, p );
    else
        hipLaunchKernelGGL( ( k_plane_rx<T, N, false, false> ), grid, block, 0, s, p );
    return hipGetLastError();
}

template <typename T>
hipError_t launch_plane_rx_t( const plane_params<T> &p, int n, bool fwd, hipStream_t s )
{
    switch ( n )
    {
    case 16: return launch_plane_rx_n<T, 16>( p, fwd, s );
    case 32: return launch_plane_rx_n<T, 32>( p, fwd, s );
    case 64: return launch_plane_rx_n<T, 64>( p, fwd, s );
    case 128: return launch_plane_rx_n<T, 128>( p, fwd, s );
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t launch_plane_t( const plane_params<T> &p, int na, int nb, bool fwd, hipStream_t s )
{
    if ( na != nb )
        return hipErrorInvalidValue;
    switch ( na )
    {
    case 16: return launch_plane_n<T, 16>( p, fwd, s );
    case 32: return launch_plane_n<T, 32>( p, fwd, s );
    case 64: return launch_plane_n<T, 64>( p, fwd, s );
    case 128: return launch_plane_n<T, 128>( p, fwd, s );
    default: return hipErrorInvalidValue;
    }
}

} // namespace

bool plane_supported( int64_t na, int64_t nb, int prec )
{
    (void)prec;
    return na == nb && ( na == 16 || na == 32 || na == 64 || na == 128 );
}

hipError_t launch_plane_f32( const plane_params<float> &p, int na, int nb, bool fwd, hipStream_t s )
{
    return launch_plane_t<float>( p, na, nb, fwd, s );
}

hipError_t launch_plane_f64( const plane_params<double> &p, int na, int nb, bool fwd, hipStream_t s )
{
    return launch_plane_t<double>( p, na, nb, fwd, s );
}

hipError_t launch_plane_rx_f32( const plane_params<float> &p, int n, bool fwd, hipStream_t s )
{
    return launch_plane_rx_t<float>( p, n, fwd, s );
}

hipError_t launch_plane_rx_f64( const plane_params<double> &p, int n, bool fwd, hipStream_t s )
{
    return launch_plane_rx_t<double>( p, n, fwd, s );
}

} // namespace fftm_mi

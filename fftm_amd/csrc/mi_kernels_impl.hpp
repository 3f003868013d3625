// Instantiation and dispatch of the pass kernels for one precision T and one length N.
// Included by the per-length translation units (mi_kernels_<prec>_<group>.hip).
#pragma once

#include "mi_launch.hpp"

#include <type_traits>

namespace fftm_mi
{
namespace detail
{

// tile-shape variants compiled for a length (small lengths have one)
template <int N>
constexpr int variants_for()
{
    return ( N & ( N - 1 ) ) == 0 && N >= 64 ? NUM_VARIANTS : 1;
}

template <typename T, int N, int KIND, bool INV, bool LOADB, int VAR, int IOM = IO_F32, int RS = 0>
hipError_t launch_fixed( const pass_params<T> &p, hipStream_t s )
{
    using K              = pass_kernel<T, N, KIND, INV, LOADB, VAR, IOM, false, RS>;
    const int64_t per    = (int64_t)K::TW * K::CPT;
    const int64_t ntiles = ( p.ncols + per - 1 ) / per;
    if ( ntiles <= 0 )
        return hipSuccess;
    if ( ntiles > 0x7fffffffLL )
        return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL( ( k_pass<T, N, KIND, INV, LOADB, VAR, IOM, RS> ), dim3( (unsigned)ntiles ), dim3( K::NT ), 0, s,
                        p );
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
        case 6: return launch_nv<T, N, 6>( p, kind, inv, loadb, s );
        case 7: return launch_nv<T, N, 7>( p, kind, inv, loadb, s );
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
        case 6: return tile_cols_nv<T, N, 6>( loadb );
        case 7: return tile_cols_nv<T, N, 7>( loadb );
        default: break;
        }
    }
    return tile_cols_nv<T, N, 0>( loadb );
}

// fused filter pass (forward C2C, x H(k), inverse C2C), default tile shape
template <typename T, int N, bool LB, int VAR>
hipError_t launch_filter_nv( const pass_params<T> &p, const filter_params<T> &f, hipStream_t s )
{
    using K              = pass_kernel<T, N, KIND_C2C, false, LB, VAR, IO_F32, true>;
    const int64_t ntiles = ( p.ncols + K::TW - 1 ) / K::TW;
    if ( ntiles <= 0 )
        return hipSuccess;
    if ( ntiles > 0x7fffffffLL )
        return hipErrorInvalidConfiguration;
    hipLaunchKernelGGL( ( k_pass_filter<T, N, LB, VAR> ), dim3( (unsigned)ntiles ), dim3( K::NT ), 0, s, p, f );
    return hipGetLastError();
}

// tile shapes of the filter pass: 0, 2 (twice the threads per column), 4 (twice the columns)
template <typename T, int N>
hipError_t launch_filter_n( const pass_params<T> &p, const filter_params<T> &f, bool loadb, int var, hipStream_t s )
{
    if constexpr ( variants_for<N>() > 1 )
    {
        if ( var == 2 )
            return loadb ? launch_filter_nv<T, N, true, 2>( p, f, s ) : launch_filter_nv<T, N, false, 2>( p, f, s );
        if ( var == 4 )
            return loadb ? launch_filter_nv<T, N, true, 4>( p, f, s ) : launch_filter_nv<T, N, false, 4>( p, f, s );
    }
    return loadb ? launch_filter_nv<T, N, true, 0>( p, f, s ) : launch_filter_nv<T, N, false, 0>( p, f, s );
}

// FP16-stored spectra (FP32 arithmetic): the storage modes each kind can have
template <int N, int VAR, int IOM>
hipError_t launch_mixed_io( const pass_params<float> &p, int kind, bool inv, bool loadb, hipStream_t s )
{
    if ( kind == KIND_C2C )
    {
        if ( inv )
            return loadb ? launch_fixed<float, N, KIND_C2C, true, true, VAR, IOM>( p, s )
                         : launch_fixed<float, N, KIND_C2C, true, false, VAR, IOM>( p, s );
        return loadb ? launch_fixed<float, N, KIND_C2C, false, true, VAR, IOM>( p, s )
                     : launch_fixed<float, N, KIND_C2C, false, false, VAR, IOM>( p, s );
    }
    if constexpr ( ( IOM & ~IO_SPLIT ) == IO_H16_OUT )
    {
        if ( kind == KIND_R2C )
            return loadb ? launch_fixed<float, N, KIND_R2C, false, true, VAR, IOM>( p, s )
                         : launch_fixed<float, N, KIND_R2C, false, false, VAR, IOM>( p, s );
    }
    if constexpr ( ( IOM & ~IO_SPLIT ) == IO_H16_IN )
    {
        if ( kind == KIND_C2R )
            return loadb ? launch_fixed<float, N, KIND_C2R, true, true, VAR, IOM>( p, s )
                         : launch_fixed<float, N, KIND_C2R, true, false, VAR, IOM>( p, s );
    }
    return hipErrorInvalidValue;
}

template <int N, int VAR>
hipError_t launch_mixed_v( const pass_params<float> &p, int kind, bool inv, bool loadb, int iom, hipStream_t s )
{
    switch ( iom )
    {
    case IO_H16_IN: return launch_mixed_io<N, VAR, IO_H16_IN>( p, kind, inv, loadb, s );
    case IO_H16_OUT: return launch_mixed_io<N, VAR, IO_H16_OUT>( p, kind, inv, loadb, s );
    case IO_H16_IN | IO_H16_OUT: return launch_mixed_io<N, VAR, IO_H16_IN | IO_H16_OUT>( p, kind, inv, loadb, s );
    default: return hipErrorInvalidValue;
    }
}

// tile-shape variants compiled for the FP16-storage kernels
constexpr bool mixed_variant( int var ) { return var == 0 || var == 6 || var == 7; }

// hi + lo FP16-stored spectra (IO_SPLIT): tile shapes 0, 2 and 6
template <int N>
hipError_t launch_mixed_split_n( const pass_params<float> &p, int kind, bool inv, bool loadb, int var, int iom,
                                 hipStream_t s )
{
    auto go = [&]( auto vtag ) -> hipError_t {
        constexpr int V = decltype( vtag )::value;
        switch ( iom )
        {
        case IO_H16_IN | IO_SPLIT: return launch_mixed_io<N, V, IO_H16_IN | IO_SPLIT>( p, kind, inv, loadb, s );
        case IO_H16_OUT | IO_SPLIT: return launch_mixed_io<N, V, IO_H16_OUT | IO_SPLIT>( p, kind, inv, loadb, s );
        case IO_H16_IN | IO_H16_OUT | IO_SPLIT:
            return launch_mixed_io<N, V, IO_H16_IN | IO_H16_OUT | IO_SPLIT>( p, kind, inv, loadb, s );
        default: return hipErrorInvalidValue;
        }
    };
    if ( var == 0 )
        return go( std::integral_constant<int, 0>{} );
    if ( var == 2 )
        return go( std::integral_constant<int, 2>{} );
    if ( var == 6 )
        return go( std::integral_constant<int, 6>{} );
    return hipErrorInvalidValue;
}

template <int N>
hipError_t launch_mixed_n( const pass_params<float> &p, int kind, bool inv, bool loadb, int var, int iom, hipStream_t s )
{
    switch ( var )
    {
    case 0: return launch_mixed_v<N, 0>( p, kind, inv, loadb, iom, s );
    case 2: return launch_mixed_v<N, 2>( p, kind, inv, loadb, iom, s );
    case 6: return launch_mixed_v<N, 6>( p, kind, inv, loadb, iom, s );
    case 7: return launch_mixed_v<N, 7>( p, kind, inv, loadb, iom, s );
    default: return hipErrorInvalidValue;
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

// Alternative radix schedules (RS 1: make_rplan_alt on the vector ALUs, RS 2: its radix-16
// stage on the matrix cores) of the strided C2C passes, FP32 arithmetic with FP32 or
// FP16-stored lines, tile shapes 0 and 6 (schedule 1 also in shape 2).
template <int N, int VAR, int RS>
hipError_t launch_rs_v( const pass_params<float> &p, bool inv, int iom, hipStream_t s )
{
    switch ( iom )
    {
    case IO_F32:
        return inv ? launch_fixed<float, N, KIND_C2C, true, true, VAR, IO_F32, RS>( p, s )
                   : launch_fixed<float, N, KIND_C2C, false, true, VAR, IO_F32, RS>( p, s );
    case IO_H16_IN:
        return inv ? launch_fixed<float, N, KIND_C2C, true, true, VAR, IO_H16_IN, RS>( p, s )
                   : launch_fixed<float, N, KIND_C2C, false, true, VAR, IO_H16_IN, RS>( p, s );
    case IO_H16_OUT:
        return inv ? launch_fixed<float, N, KIND_C2C, true, true, VAR, IO_H16_OUT, RS>( p, s )
                   : launch_fixed<float, N, KIND_C2C, false, true, VAR, IO_H16_OUT, RS>( p, s );
    case IO_H16_IN | IO_H16_OUT:
        return inv ? launch_fixed<float, N, KIND_C2C, true, true, VAR, IO_H16_IN | IO_H16_OUT, RS>( p, s )
                   : launch_fixed<float, N, KIND_C2C, false, true, VAR, IO_H16_IN | IO_H16_OUT, RS>( p, s );
    default: return hipErrorInvalidValue;
    }
}

template <int N>
hipError_t launch_rs_n( const pass_params<float> &p, bool inv, int shape, int rs, int iom, hipStream_t s )
{
    if ( shape == 0 )
        return rs == 1 ? launch_rs_v<N, 0, 1>( p, inv, iom, s ) : launch_rs_v<N, 0, 2>( p, inv, iom, s );
    if ( shape == 6 )
        return rs == 1 ? launch_rs_v<N, 6, 1>( p, inv, iom, s ) : launch_rs_v<N, 6, 2>( p, inv, iom, s );
    if ( shape == 2 && rs == 1 )
        return launch_rs_v<N, 2, 1>( p, inv, iom, s );
    return hipErrorInvalidValue;
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

// One length N of the FP16-storage kernels: defines launch_mixed_f32_n<N>.
#define FFTM_MI_DEFINE_MIXED_LENGTH( N )                                                                               \
    namespace fftm_mi                                                                                                  \
    {                                                                                                                  \
    hipError_t launch_mixed_f32_n##N( const pass_params<float> &p, int kind, bool inv, bool loadb, int var, int iom,   \
                                      hipStream_t s )                                                                  \
    {                                                                                                                  \
        return detail::launch_mixed_n<N>( p, kind, inv, loadb, var, iom, s );                                          \
    }                                                                                                                  \
    }

// One length N of the hi + lo FP16-storage kernels: defines launch_mixed_split_f32_n<N>.
#define FFTM_MI_DEFINE_MIXED_SPLIT_LENGTH( N )                                                                         \
    namespace fftm_mi                                                                                                  \
    {                                                                                                                  \
    hipError_t launch_mixed_split_f32_n##N( const pass_params<float> &p, int kind, bool inv, bool loadb, int var,      \
                                            int iom, hipStream_t s )                                                   \
    {                                                                                                                  \
        return detail::launch_mixed_split_n<N>( p, kind, inv, loadb, var, iom, s );                                    \
    }                                                                                                                  \
    }

// The fused filter pass for precision T (type name TN) over every compiled length.
#define FFTM_MI_DEFINE_FILTER( T, TN )                                                                                 \
    namespace fftm_mi                                                                                                  \
    {                                                                                                                  \
    hipError_t launch_filter_##TN( const pass_params<T> &p, const filter_params<T> &f, int N, bool loadb, int var,     \
                                   hipStream_t s )                                                                     \
    {                                                                                                                  \
        switch ( N )                                                                                                   \
        {                                                                                                              \
        case 2: return detail::launch_filter_n<T, 2>( p, f, loadb, var, s );                                                \
        case 4: return detail::launch_filter_n<T, 4>( p, f, loadb, var, s );                                                \
        case 8: return detail::launch_filter_n<T, 8>( p, f, loadb, var, s );                                                \
        case 16: return detail::launch_filter_n<T, 16>( p, f, loadb, var, s );                                              \
        case 32: return detail::launch_filter_n<T, 32>( p, f, loadb, var, s );                                              \
        case 64: return detail::launch_filter_n<T, 64>( p, f, loadb, var, s );                                              \
        case 128: return detail::launch_filter_n<T, 128>( p, f, loadb, var, s );                                            \
        case 256: return detail::launch_filter_n<T, 256>( p, f, loadb, var, s );                                            \
        case 512: return detail::launch_filter_n<T, 512>( p, f, loadb, var, s );                                            \
        case 1024: return detail::launch_filter_n<T, 1024>( p, f, loadb, var, s );                                          \
        case 2048: return detail::launch_filter_n<T, 2048>( p, f, loadb, var, s );                                          \
        case 4096: return detail::launch_filter_n<T, 4096>( p, f, loadb, var, s );                                          \
        default: return hipErrorInvalidValue;                                                                          \
        }                                                                                                              \
    }                                                                                                                  \
    }

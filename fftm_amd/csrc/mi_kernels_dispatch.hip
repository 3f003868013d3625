// Dispatch of the fixed-length pass kernels (one definition per length lives in the
// mi_kernels_<prec>_<group>.hip units) and the generic kernel launchers.
#include "mi_kernels_impl.hpp"

#define FFTM_MI_LENGTHS( X, T, TN )                                                                                    \
    X( T, TN, 2 ) X( T, TN, 4 ) X( T, TN, 8 ) X( T, TN, 16 ) X( T, TN, 32 ) X( T, TN, 64 ) X( T, TN, 128 )             \
        X( T, TN, 256 ) X( T, TN, 512 ) X( T, TN, 1024 ) X( T, TN, 2048 ) X( T, TN, 4096 )

// compiled lengths with factors 3, 5, 7 (must match smooth_lengths() in mi_launch.hpp)
#define FFTM_MI_SMOOTH_LENGTHS( X, T, TN )                                                                             \
    X( T, TN, 12 ) \
    X( T, TN, 20 ) \
    X( T, TN, 24 ) \
    X( T, TN, 28 ) \
    X( T, TN, 36 ) \
    X( T, TN, 40 ) \
    X( T, TN, 48 ) \
    X( T, TN, 56 ) \
    X( T, TN, 60 ) \
    X( T, TN, 72 ) \
    X( T, TN, 80 ) \
    X( T, TN, 96 ) \
    X( T, TN, 100 ) \
    X( T, TN, 112 ) \
    X( T, TN, 120 ) \
    X( T, TN, 144 ) \
    X( T, TN, 160 ) \
    X( T, TN, 192 ) \
    X( T, TN, 200 ) \
    X( T, TN, 224 ) \
    X( T, TN, 240 ) \
    X( T, TN, 288 ) \
    X( T, TN, 320 ) \
    X( T, TN, 384 ) \
    X( T, TN, 400 ) \
    X( T, TN, 448 ) \
    X( T, TN, 480 ) \
    X( T, TN, 576 ) \
    X( T, TN, 640 ) \
    X( T, TN, 768 ) \
    X( T, TN, 800 ) \
    X( T, TN, 896 ) \
    X( T, TN, 960 ) \
    X( T, TN, 1000 ) \
    X( T, TN, 1152 ) \
    X( T, TN, 1280 ) \
    X( T, TN, 1344 ) \
    X( T, TN, 1400 ) \
    X( T, TN, 1536 ) \
    X( T, TN, 1600 ) \
    X( T, TN, 1680 ) \
    X( T, TN, 1792 ) \
    X( T, TN, 1920 ) \
    X( T, TN, 2000 ) \
    X( T, TN, 2100 )

#define FFTM_MI_DECLARE( T, TN, N )                                                                                    \
    hipError_t launch_##TN##_n##N( const pass_params<T> &, int, bool, bool, int, hipStream_t );                       \
    int        tile_cols_##TN##_n##N( bool, int );

#define FFTM_MI_CASE_LAUNCH( T, TN, N )                                                                                \
    case N: return launch_##TN##_n##N( p, kind, inv, loadb, var, s );

#define FFTM_MI_CASE_COLS( T, TN, N )                                                                                  \
    case N: return tile_cols_##TN##_n##N( loadb, var );

namespace fftm_mi
{

FFTM_MI_LENGTHS( FFTM_MI_DECLARE, float, f32 )
FFTM_MI_LENGTHS( FFTM_MI_DECLARE, double, f64 )
FFTM_MI_SMOOTH_LENGTHS( FFTM_MI_DECLARE, float, f32 )
FFTM_MI_SMOOTH_LENGTHS( FFTM_MI_DECLARE, double, f64 )

hipError_t launch_pass_f32( const pass_params<float> &p, int N, int kind, bool inv, bool loadb, int var, hipStream_t s )
{
    switch ( N )
    {
        FFTM_MI_LENGTHS( FFTM_MI_CASE_LAUNCH, float, f32 )
        FFTM_MI_SMOOTH_LENGTHS( FFTM_MI_CASE_LAUNCH, float, f32 )
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_pass_f64( const pass_params<double> &p, int N, int kind, bool inv, bool loadb, int var, hipStream_t s )
{
    switch ( N )
    {
        FFTM_MI_LENGTHS( FFTM_MI_CASE_LAUNCH, double, f64 )
        FFTM_MI_SMOOTH_LENGTHS( FFTM_MI_CASE_LAUNCH, double, f64 )
    default: return hipErrorInvalidValue;
    }
}

int fixed_tile_cols_f32( int N, bool loadb, int var )
{
    switch ( N )
    {
        FFTM_MI_LENGTHS( FFTM_MI_CASE_COLS, float, f32 )
        FFTM_MI_SMOOTH_LENGTHS( FFTM_MI_CASE_COLS, float, f32 )
    default: return 0;
    }
}

int fixed_tile_cols_f64( int N, bool loadb, int var )
{
    switch ( N )
    {
        FFTM_MI_LENGTHS( FFTM_MI_CASE_COLS, double, f64 )
        FFTM_MI_SMOOTH_LENGTHS( FFTM_MI_CASE_COLS, double, f64 )
    default: return 0;
    }
}

#define FFTM_MI_MIXED_LENGTHS( X ) X( 128 ) X( 256 ) X( 512 ) X( 1024 ) X( 2048 )
#define FFTM_MI_DECLARE_MIXED( N )                                                                                     \
    hipError_t launch_mixed_f32_n##N( const pass_params<float> &, int, bool, bool, int, int, hipStream_t );
#define FFTM_MI_CASE_MIXED( N )                                                                                        \
    case N: return launch_mixed_f32_n##N( p, kind, inv, loadb, var, iom, s );

FFTM_MI_MIXED_LENGTHS( FFTM_MI_DECLARE_MIXED )

#define FFTM_MI_MIXED_SPLIT_LENGTHS( X ) X( 128 ) X( 256 ) X( 512 ) X( 1024 )
#define FFTM_MI_DECLARE_MIXED_SPLIT( N )                                                                               \
    hipError_t launch_mixed_split_f32_n##N( const pass_params<float> &, int, bool, bool, int, int, hipStream_t );
#define FFTM_MI_CASE_MIXED_SPLIT( N )                                                                                  \
    case N: return launch_mixed_split_f32_n##N( p, kind, inv, loadb, var, iom, s );

FFTM_MI_MIXED_SPLIT_LENGTHS( FFTM_MI_DECLARE_MIXED_SPLIT )

hipError_t launch_pass_mixed_f32( const pass_params<float> &p, int N, int kind, bool inv, bool loadb, int var, int iom,
                                  hipStream_t s )
{
    if ( iom & IO_SPLIT )
    {
        switch ( N )
        {
            FFTM_MI_MIXED_SPLIT_LENGTHS( FFTM_MI_CASE_MIXED_SPLIT )
        default: return hipErrorInvalidValue;
        }
    }
    switch ( N )
    {
        FFTM_MI_MIXED_LENGTHS( FFTM_MI_CASE_MIXED )
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_rs_f32_n512( const pass_params<float> &, bool, int, int, int, hipStream_t );
hipError_t launch_rs_f32_n1024( const pass_params<float> &, bool, int, int, int, hipStream_t );

bool rs_supported( int64_t N, int kind, bool loadb, int prec, int shape, int rs )
{
    return ( N == 512 || N == 1024 ) && kind == KIND_C2C && loadb && prec == 0 &&
           ( shape == 0 || shape == 6 || ( shape == 2 && rs == 1 ) );
}

hipError_t launch_pass_rs_f32( const pass_params<float> &p, int N, bool inv, int shape, int rs, int iom, hipStream_t s )
{
    if ( rs < 1 || rs > 2 )
        return hipErrorInvalidValue;
    switch ( N )
    {
    case 512: return launch_rs_f32_n512( p, inv, shape, rs, iom, s );
    case 1024: return launch_rs_f32_n1024( p, inv, shape, rs, iom, s );
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_generic_f32( const pass_params<float> &p, const radix_list &rl, int N, int cols, int kind, bool inv,
                               hipStream_t s )
{
    return detail::launch_generic_t<float>( p, rl, N, cols, kind, inv, s );
}

hipError_t launch_generic_f64( const pass_params<double> &p, const radix_list &rl, int N, int cols, int kind, bool inv,
                               hipStream_t s )
{
    return detail::launch_generic_t<double>( p, rl, N, cols, kind, inv, s );
}

} // namespace fftm_mi

// Host-side launch entry points for the pass kernels (one translation unit per
// precision instantiates the compile-time-N kernels; see mi_kernels_f32.hip /
// mi_kernels_f64.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mi_pass.hpp"
#include "mi_plane.hpp"

namespace fftm_mi
{

// compiled lengths of the fixed-N kernels: powers of two 2..4096 (all tile-shape
// variants from 64 up) and a set of 2^a 3^b 5^c 7^d lengths (default tile shape)
inline bool pow2_length( int64_t N )
{
    return N >= 2 && N <= 4096 && ( N & ( N - 1 ) ) == 0;
}
inline bool fixed_supported( int64_t N )
{
    static const int64_t smooth[] = { 12, 20, 24, 28, 36, 40, 48, 56, 60, 72, 80, 96, 100, 112, 120, 144, 160, 192, 200, 224, 240, 288, 320, 384, 400, 448, 480, 576, 640, 768, 800, 896, 960, 1000, 1152, 1280, 1344, 1400, 1536, 1600, 1680, 1792, 1920, 2000, 2100 };
    if ( pow2_length( N ) )
        return true;
    for ( int64_t v : smooth )
        if ( v == N )
            return true;
    return false;
}
// tile-shape variants compiled for a length
inline int fixed_variants( int64_t N )
{
    return pow2_length( N ) && N >= 64 ? NUM_VARIANTS : 1;
}

// columns per workgroup of a fixed-N kernel (complex lines)
int fixed_tile_cols_f32( int N, bool loadb, int var );
int fixed_tile_cols_f64( int N, bool loadb, int var );

hipError_t launch_pass_f32( const pass_params<float> &p, int N, int kind, bool inv, bool loadb, int var, hipStream_t s );
hipError_t launch_pass_f64( const pass_params<double> &p, int N, int kind, bool inv, bool loadb, int var, hipStream_t s );
// plane passes (real axis + one complex axis in LDS), square planes 16..128
bool       plane_supported( int64_t na, int64_t nb, int prec );
hipError_t launch_plane_f32( const plane_params<float> &p, int na, int nb, bool fwd, hipStream_t s );
hipError_t launch_plane_f64( const plane_params<double> &p, int na, int nb, bool fwd, hipStream_t s );
// register-resident plane kernel (k_plane_rx): square planes 16..128 whose B axis is
// contiguous in both arrays (rb = cb = 1)
hipError_t launch_plane_rx_f32( const plane_params<float> &p, int n, bool fwd, hipStream_t s );
hipError_t launch_plane_rx_f64( const plane_params<double> &p, int n, bool fwd, hipStream_t s );

// FP16-stored spectra with FP32 arithmetic (iom: bit 0 FP16 input, bit 1 FP16 output);
// lengths 128..2048, tile-shape variants 0, 2, 6, 7
inline bool mixed_supported( int64_t N )
{
    return N >= 128 && N <= 2048 && ( N & ( N - 1 ) ) == 0;
}
inline bool mixed_variant_ok( int var )
{
    const int v = variant_shape_of( var );
    return v == 0 || v == 2 || v == 6 || v == 7;
}
// a variant number: a tile shape 0..NUM_VARIANTS-1, plus 8 * the tile order (0..3), plus 32 *
// the radix schedule (0 default; 1 the alternative schedule; 2 the alternative schedule with
// its radix-16 stage on the matrix cores; 1 and 2 exist where rs_supported)
inline bool variant_number_ok( int var )
{
    return var >= 0 && variant_rs_of( var ) <= 2;
}
// alternative radix schedules are compiled for strided C2C passes of length 512 and 1024
// with FP32 arithmetic (prec 0), tile shapes 0 and 6 (and 2 with schedule 1)
bool       rs_supported( int64_t N, int kind, bool loadb, int prec, int shape, int rs );
hipError_t launch_pass_rs_f32( const pass_params<float> &p, int N, bool inv, int shape, int rs, int iom, hipStream_t s );
// the order bits of pass_params::vec for a variant number
inline int variant_vec_order( int var )
{
    static constexpr int bits[NUM_ORDERS] = { 0, VEC_RR, VEC_CHUNK8, VEC_CHUNK32 };
    return bits[variant_order_of( var )];
}
hipError_t launch_pass_mixed_f32( const pass_params<float> &p, int N, int kind, bool inv, bool loadb, int var, int iom,
                                  hipStream_t s );

// fused filter pass along a compiled length (forward C2C, x H(k), inverse C2C)
hipError_t launch_filter_f32( const pass_params<float> &p, const filter_params<float> &f, int N, bool loadb, int var,
                              hipStream_t s );
hipError_t launch_filter_f64( const pass_params<double> &p, const filter_params<double> &f, int N, bool loadb, int var,
                              hipStream_t s );

hipError_t launch_generic_f32( const pass_params<float> &p, const radix_list &rl, int N, int cols, int kind, bool inv,
                               hipStream_t s );
hipError_t launch_generic_f64( const pass_params<double> &p, const radix_list &rl, int N, int cols, int kind, bool inv,
                               hipStream_t s );

} // namespace fftm_mi

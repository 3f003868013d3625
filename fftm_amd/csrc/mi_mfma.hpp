// Radix-16 DFT of a wavefront's butterflies as dense FP16 matrix products on the CDNA4
// matrix cores (v_mfma_f32_32x32x16_f16), accumulated in FP32, with split hi + lo FP16
// operands so that the result keeps FP32-level accuracy.
//
// A 16-point DFT y = F x is a real 32 x 32 product on [Re x; Im x]:
//     [Re y]   [ Fr  -Fi ] [Re x]
//     [Im y] = [ Fi   Fr ] [Im x],   F = Fr + i Fi,  F_mk = e^{-+2 pi i m k / 16}.
// One MFMA takes a 32 x 16 block of that matrix (A) and 16 rows of 32 butterflies'
// inputs (B); the Re rows and the Im rows of x are the two K blocks.  With x = xh + xl and
// F = Fh + Fl (FP16 each), y = Fh xh + Fh xl + Fl xh + O(2^-22 |F| |x|): three products
// per K block, six MFMAs per 32 butterflies.
//
// Lane layouts of v_mfma_f32_32x32x16_f16 (wave64):
//   A: lane l holds A[l % 32][8 (l / 32) + q], q = 0..7
//   B: lane l holds B[8 (l / 32) + q][l % 32]
//   D: lane l holds D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32], v = 0..15
// Every lane starts with all 16 inputs of its own butterfly.  The two halves of the wave
// (lanes l and l ^ 32) swap inputs so that lane l (half h) holds inputs 8h..8h+7 of both
// butterflies (its own and its partner's); two sets of MFMAs transform the half-0 lanes'
// butterflies (set A) and the half-1 lanes' (set B); a second swap returns every output to
// the lane that owns its butterfly.  Each butterfly is scaled by a power of two before the
// FP16 split so that its largest component lies in [2^14, 2^15): FP16 holds the range, the
// split keeps the precision.
#pragma once

#include "mi_complex.hpp"

namespace fftm_mi
{

typedef _Float16 mf_h8 __attribute__( ( ext_vector_type( 8 ) ) );
typedef float    mf_f16 __attribute__( ( ext_vector_type( 16 ) ) );

// DFT-matrix operands of this lane: a[0], a[1] = hi, lo of the Re-input block; a[2], a[3] =
// hi, lo of the Im-input block.  tw: the pass's length-N table e^{-2 pi i k / N}.
template <bool INV>
__device__ __forceinline__ void dft16_mfma_operands( mf_h8 ( &a )[4], const cpx<float> *tw, int N )
{
    const int l = threadIdx.x & 63;
    const int i = l & 31, m = i & 15, h = l >> 5;
#pragma unroll
    for ( int q = 0; q < 8; ++q )
    {
        const int        k  = 8 * h + q;
        const cpx<float> t  = tw[( ( m * k ) & 15 ) * ( N / 16 )]; // cos - i sin of 2 pi m k / 16
        const float      fr = t.x;
        const float      fi = INV ? -t.y : t.y;                    // Im of F_mk
        const float      v0 = i < 16 ? fr : fi;                     // block multiplying Re x
        const float      v1 = i < 16 ? -fi : fr;                    // block multiplying Im x
        const _Float16   h0 = (_Float16)v0, h1 = (_Float16)v1;
        a[0][q]             = h0;
        a[1][q]             = (_Float16)( v0 - (float)h0 );
        a[2][q]             = h1;
        a[3][q]             = (_Float16)( v1 - (float)h1 );
    }
}

// power-of-two exponent placing amax in [2^14, 2^15)
__device__ __forceinline__ int mf_exponent( float amax )
{
    int e = 0;
    (void)__builtin_frexpf( amax, &e );
    e -= 15;
    return e < -126 ? -126 : ( e > 126 ? 126 : e );
}

// y = F x for the 16 inputs of one butterfly set held as 8 per lane (see header), scaled
// by 2^-e; returns the 8 outputs this lane holds (m = 4h..4h+3, 8+4h..8+4h+3), unscaled
__device__ __forceinline__ void dft16_mfma_set( const cpx<float> ( &xin )[8], int e, const mf_h8 ( &a )[4],
                                                cpx<float> ( &yout )[8] )
{
    mf_h8 reh, rel, imh, iml;
#pragma unroll
    for ( int q = 0; q < 8; ++q )
    {
        const float    xr = __builtin_ldexpf( xin[q].x, -e ), xi = __builtin_ldexpf( xin[q].y, -e );
        const _Float16 hr = (_Float16)xr, hi = (_Float16)xi;
        reh[q]            = hr;
        rel[q]            = (_Float16)( xr - (float)hr );
        imh[q]            = hi;
        iml[q]            = (_Float16)( xi - (float)hi );
    }
    mf_f16 d = {};
    d        = __builtin_amdgcn_mfma_f32_32x32x16_f16( a[0], reh, d, 0, 0, 0 );
    d        = __builtin_amdgcn_mfma_f32_32x32x16_f16( a[0], rel, d, 0, 0, 0 );
    d        = __builtin_amdgcn_mfma_f32_32x32x16_f16( a[1], reh, d, 0, 0, 0 );
    d        = __builtin_amdgcn_mfma_f32_32x32x16_f16( a[2], imh, d, 0, 0, 0 );
    d        = __builtin_amdgcn_mfma_f32_32x32x16_f16( a[2], iml, d, 0, 0, 0 );
    d        = __builtin_amdgcn_mfma_f32_32x32x16_f16( a[3], imh, d, 0, 0, 0 );
    // rows 8 (v / 4) + 4h + v % 4: v = 0..3 -> Re y[4h + r], 4..7 -> Re y[8 + 4h + r],
    // 8..11 -> Im y[4h + r], 12..15 -> Im y[8 + 4h + r]
#pragma unroll
    for ( int r = 0; r < 4; ++r )
    {
        yout[r]     = mk<float>( __builtin_ldexpf( d[r], e ), __builtin_ldexpf( d[8 + r], e ) );
        yout[4 + r] = mk<float>( __builtin_ldexpf( d[4 + r], e ), __builtin_ldexpf( d[12 + r], e ) );
    }
}

// In-place 16-point DFT of every lane's butterfly v[0..15] (natural order in and out, as
// dft_fixed<16>), for all 64 lanes of the wave together (every lane must call it).
__device__ __forceinline__ void dft16_mfma( cpx<float> ( &v )[16], const mf_h8 ( &a )[4] )
{
    const int h = ( threadIdx.x >> 5 ) & 1;
    // half 0 keeps inputs 0..7 and sends 8..15; half 1 keeps 8..15 and sends 0..7
    cpx<float> mine[8], other[8];
#pragma unroll
    for ( int q = 0; q < 8; ++q )
    {
        const cpx<float> s = h ? v[q] : v[8 + q];
        mine[q]            = h ? v[8 + q] : v[q];
        other[q]           = mk<float>( __shfl_xor( s.x, 32, 64 ), __shfl_xor( s.y, 32, 64 ) );
    }
    // scale of each butterfly: its own largest component, shared with the partner lane
    float mo = 0.f;
#pragma unroll
    for ( int q = 0; q < 16; ++q )
        mo = fmaxf( mo, fmaxf( fabsf( v[q].x ), fabsf( v[q].y ) ) );
    const float mp = __shfl_xor( mo, 32, 64 );
    const int   eo = mf_exponent( mo ), ep = mf_exponent( mp );
    // set A: the half-0 lanes' butterflies (half 1 supplies their inputs 8..15);
    // set B: the half-1 lanes' butterflies (half 0 supplies their inputs 0..7)
    cpx<float> sa[8], sb[8], ya[8], yb[8];
#pragma unroll
    for ( int q = 0; q < 8; ++q )
    {
        sa[q] = h ? other[q] : mine[q];
        sb[q] = h ? mine[q] : other[q];
    }
    dft16_mfma_set( sa, h ? ep : eo, a, ya );
    dft16_mfma_set( sb, h ? eo : ep, a, yb );
    // return the outputs: half 0 owns set A (holds m in 0..3, 8..11 of it) and needs
    // m in 4..7, 12..15, which its partner holds in ya; it sends its yb to the partner
    // slot q of o (this lane's own outputs of its half) is m = 4h + q (q < 4) or
    // 8 + 4h + (q - 4); slot q of r (from the partner) is the same with 1 - h.  Written with
    // selects so that every register index is a constant.
#pragma unroll
    for ( int q = 0; q < 8; ++q )
    {
        const cpx<float> s  = h ? ya[q] : yb[q];
        const cpx<float> r  = mk<float>( __shfl_xor( s.x, 32, 64 ), __shfl_xor( s.y, 32, 64 ) );
        const cpx<float> o  = h ? yb[q] : ya[q];
        const int        lo = q < 4 ? q : 8 + ( q - 4 ); // m of slot q for half 0
        v[lo]               = h ? r : o;
        v[lo + 4]           = h ? o : r;
    }
}

} // namespace fftm_mi

// Complex arithmetic and small-radix DFT butterflies for the CDNA4 FFT passes.
//
// All butterflies compute the unnormalised DFT  y[m] = sum_q x[q] e^{s 2 pi i q m / R}
// with s = -1 (forward, INV = false) or s = +1 (inverse, INV = true), matching the
// hipFFT/FFTW sign convention used by ffts (HIPFFT_FORWARD = -1, HIPFFT_BACKWARD = +1;
// source/external_wrap/hipfft_wrap.h:363-417).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fftm_mi
{

template <typename T>
struct alignas( 2 * sizeof( T ) ) cpx
{
    T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> mk( T a, T b )
{
    cpx<T> r;
    r.x = a;
    r.y = b;
    return r;
}

template <typename T>
__device__ __forceinline__ cpx<T> add( cpx<T> a, cpx<T> b )
{
    return mk<T>( a.x + b.x, a.y + b.y );
}

template <typename T>
__device__ __forceinline__ cpx<T> sub( cpx<T> a, cpx<T> b )
{
    return mk<T>( a.x - b.x, a.y - b.y );
}

// Fused multiply-add in the precision of T (one rounding; v_fma_f32 / v_fma_f64).
__device__ __forceinline__ float fmat( float a, float b, float c ) { return __builtin_fmaf( a, b, c ); }
__device__ __forceinline__ double fmat( double a, double b, double c ) { return __builtin_fma( a, b, c ); }

// Complex products use explicit fused multiply-adds (the library is compiled with
// -ffp-contract=off), so every kernel instantiation rounds identically whatever the
// unrolling: tile-shape variants stay bitwise equal.
template <typename T>
__device__ __forceinline__ cpx<T> mul( cpx<T> a, cpx<T> b )
{
    return mk<T>( fmat( a.x, b.x, -( a.y * b.y ) ), fmat( a.x, b.y, a.y * b.x ) );
}

// a * conj(b)
template <typename T>
__device__ __forceinline__ cpx<T> mulc( cpx<T> a, cpx<T> b )
{
    return mk<T>( fmat( a.x, b.x, a.y * b.y ), fmat( a.y, b.x, -( a.x * b.y ) ) );
}

template <typename T>
__device__ __forceinline__ cpx<T> scl( cpx<T> a, T s )
{
    return mk<T>( a.x * s, a.y * s );
}

template <typename T>
__device__ __forceinline__ cpx<T> conj( cpx<T> a )
{
    return mk<T>( a.x, -a.y );
}

// multiply by s*i  (s = -1 forward, +1 inverse)
template <bool INV, typename T>
__device__ __forceinline__ cpx<T> mul_si( cpx<T> a )
{
    return INV ? mk<T>( -a.y, a.x ) : mk<T>( a.y, -a.x );
}

// twiddle multiply: a * w (forward) or a * conj(w) (inverse), w = e^{-2 pi i k / N}
template <bool INV, typename T>
__device__ __forceinline__ cpx<T> twmul( cpx<T> a, cpx<T> w )
{
    return INV ? mulc( a, w ) : mul( a, w );
}

// ---------------------------------------------------------------------------
// Butterflies on register arrays
// ---------------------------------------------------------------------------

template <bool INV, typename T>
__device__ __forceinline__ void dft2( cpx<T> *v )
{
    const cpx<T> a = v[0], b = v[1];
    v[0]               = add( a, b );
    v[1]               = sub( a, b );
}

template <bool INV, typename T>
__device__ __forceinline__ void dft4( cpx<T> *v )
{
    const cpx<T> a = add( v[0], v[2] ), b = sub( v[0], v[2] );
    const cpx<T> c = add( v[1], v[3] ), d = mul_si<INV>( sub( v[1], v[3] ) );
    v[0]               = add( a, c );
    v[2]               = sub( a, c );
    v[1]               = add( b, d );
    v[3]               = sub( b, d );
}

// radix 8 = radix-2 DIF split followed by two radix-4 DFTs (in-place, natural order)
template <bool INV, typename T>
__device__ __forceinline__ void dft8( cpx<T> *v )
{
    const T      c = (T)0.70710678118654752440084436210484904;
    cpx<T>       a[4], b[4];
#pragma unroll
    for ( int i = 0; i < 4; ++i )
    {
        a[i] = add( v[i], v[i + 4] );
        b[i] = sub( v[i], v[i + 4] );
    }
    // b1 *= W8^{s}, b2 *= W8^{2s} = s*i, b3 *= W8^{3s}
    if ( !INV )
    {
        b[1] = mk<T>( c * ( b[1].x + b[1].y ), c * ( b[1].y - b[1].x ) );
        b[3] = mk<T>( c * ( b[3].y - b[3].x ), -c * ( b[3].x + b[3].y ) );
    }
    else
    {
        b[1] = mk<T>( c * ( b[1].x - b[1].y ), c * ( b[1].x + b[1].y ) );
        b[3] = mk<T>( -c * ( b[3].x + b[3].y ), c * ( b[3].x - b[3].y ) );
    }
    b[2] = mul_si<INV>( b[2] );
    dft4<INV>( a );
    dft4<INV>( b );
#pragma unroll
    for ( int i = 0; i < 4; ++i )
    {
        v[2 * i]     = a[i];
        v[2 * i + 1] = b[i];
    }
}

// radix 16 = 4 x 4 with internal twiddles W16^{n2 k1}
template <bool INV, typename T>
__device__ __forceinline__ void dft16( cpx<T> *v )
{
    const T c1 = (T)0.92387953251128675612818318939678829; // cos(pi/8)
    const T s1 = (T)0.38268343236508977172845998403039887; // sin(pi/8)
    const T c2 = (T)0.70710678118654752440084436210484904; // cos(pi/4)
    cpx<T>  y[4][4];
#pragma unroll
    for ( int n2 = 0; n2 < 4; ++n2 )
    {
        cpx<T> u[4] = { v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12] };
        dft4<INV>( u );
#pragma unroll
        for ( int k1 = 0; k1 < 4; ++k1 )
            y[n2][k1] = u[k1];
    }
    // W16^e for e = n2*k1 in {1,2,3,4,6,9}: (cos, -sin) of e*pi/8 for forward
    const T sg = INV ? (T)1 : (T)-1;
    const cpx<T> w1 = mk<T>( c1, sg * s1 ), w2 = mk<T>( c2, sg * c2 ), w3 = mk<T>( s1, sg * c1 );
    const cpx<T> w6 = mk<T>( -c2, sg * c2 ), w9 = mk<T>( -c1, -sg * s1 );
    y[1][1] = mul( y[1][1], w1 );
    y[1][2] = mul( y[1][2], w2 );
    y[1][3] = mul( y[1][3], w3 );
    y[2][1] = mul( y[2][1], w2 );
    y[2][2] = mul_si<INV>( y[2][2] ); // W16^4 = s*i
    y[2][3] = mul( y[2][3], w6 );
    y[3][1] = mul( y[3][1], w3 );
    y[3][2] = mul( y[3][2], w6 );
    y[3][3] = mul( y[3][3], w9 );
#pragma unroll
    for ( int k1 = 0; k1 < 4; ++k1 )
    {
        cpx<T> u[4] = { y[0][k1], y[1][k1], y[2][k1], y[3][k1] };
        dft4<INV>( u );
#pragma unroll
        for ( int k2 = 0; k2 < 4; ++k2 )
            v[k1 + 4 * k2] = u[k2];
    }
}

template <bool INV, typename T>
__device__ __forceinline__ void dft3( cpx<T> *v )
{
    const T      s  = (T)0.86602540378443864676372317075293618; // sqrt(3)/2
    const cpx<T> t1 = add( v[1], v[2] );
    const cpx<T> d  = sub( v[1], v[2] );
    const cpx<T> t2 = mk<T>( v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y );
    v[0]            = add( v[0], t1 );
    // (s_sign * i) * sqrt(3)/2 * d
    const cpx<T> e  = mul_si<INV>( scl( d, s ) );
    v[1]            = add( t2, e );
    v[2]            = sub( t2, e );
}

template <bool INV, typename T>
__device__ __forceinline__ void dft5( cpx<T> *v )
{
    const T c1 = (T)0.30901699437494742410229341718281906;  // cos(2pi/5)
    const T c2 = (T)-0.80901699437494742410229341718281906; // cos(4pi/5)
    const T s1 = (T)0.95105651629515357211643933337938214;  // sin(2pi/5)
    const T s2 = (T)0.58778525229247312916870595463907277;  // sin(4pi/5)
    const cpx<T> a1 = add( v[1], v[4] ), b1 = sub( v[1], v[4] );
    const cpx<T> a2 = add( v[2], v[3] ), b2 = sub( v[2], v[3] );
    const cpx<T> x0 = v[0];
    const cpx<T> r1 = mk<T>( x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y );
    const cpx<T> r2 = mk<T>( x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y );
    // i_s * (s1 b1 + s2 b2), i_s * (s2 b1 - s1 b2)
    const cpx<T> q1 = mul_si<INV>( mk<T>( s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y ) );
    const cpx<T> q2 = mul_si<INV>( mk<T>( s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y ) );
    v[0]            = mk<T>( x0.x + a1.x + a2.x, x0.y + a1.y + a2.y );
    v[1]            = add( r1, q1 );
    v[4]            = sub( r1, q1 );
    v[2]            = add( r2, q2 );
    v[3]            = sub( r2, q2 );
}

template <bool INV, typename T>
__device__ __forceinline__ void dft7( cpx<T> *v )
{
    const T c1 = (T)0.62348980185873353052500488400423981;  // cos(2pi/7)
    const T c2 = (T)-0.22252093395631440428890256449679476; // cos(4pi/7)
    const T c3 = (T)-0.90096886790241912623610231950744505; // cos(6pi/7)
    const T s1 = (T)0.78183148246802980870844452667405775;  // sin(2pi/7)
    const T s2 = (T)0.97492791218182360701813168299393122;  // sin(4pi/7)
    const T s3 = (T)0.43388373911755812047576833284835875;  // sin(6pi/7)
    const cpx<T> a1 = add( v[1], v[6] ), b1 = sub( v[1], v[6] );
    const cpx<T> a2 = add( v[2], v[5] ), b2 = sub( v[2], v[5] );
    const cpx<T> a3 = add( v[3], v[4] ), b3 = sub( v[3], v[4] );
    const cpx<T> x0 = v[0];
    const cpx<T> r1 = mk<T>( x0.x + c1 * a1.x + c2 * a2.x + c3 * a3.x, x0.y + c1 * a1.y + c2 * a2.y + c3 * a3.y );
    const cpx<T> r2 = mk<T>( x0.x + c2 * a1.x + c3 * a2.x + c1 * a3.x, x0.y + c2 * a1.y + c3 * a2.y + c1 * a3.y );
    const cpx<T> r3 = mk<T>( x0.x + c3 * a1.x + c1 * a2.x + c2 * a3.x, x0.y + c3 * a1.y + c1 * a2.y + c2 * a3.y );
    const cpx<T> q1 = mul_si<INV>(
        mk<T>( s1 * b1.x + s2 * b2.x + s3 * b3.x, s1 * b1.y + s2 * b2.y + s3 * b3.y ) );
    const cpx<T> q2 = mul_si<INV>(
        mk<T>( s2 * b1.x - s3 * b2.x - s1 * b3.x, s2 * b1.y - s3 * b2.y - s1 * b3.y ) );
    const cpx<T> q3 = mul_si<INV>(
        mk<T>( s3 * b1.x - s1 * b2.x + s2 * b3.x, s3 * b1.y - s1 * b2.y + s2 * b3.y ) );
    v[0] = mk<T>( x0.x + a1.x + a2.x + a3.x, x0.y + a1.y + a2.y + a3.y );
    v[1] = add( r1, q1 );
    v[6] = sub( r1, q1 );
    v[2] = add( r2, q2 );
    v[5] = sub( r2, q2 );
    v[3] = add( r3, q3 );
    v[4] = sub( r3, q3 );
}

// Generic radix R (any R <= 32): O(R^2) using the length-N twiddle table
// (W_R^m = tw[m * N / R]).  Used for primes > 7 only.
template <bool INV, typename T>
__device__ __forceinline__ void dft_generic( cpx<T> *v, int R, const cpx<T> *__restrict__ tw, int64_t N )
{
    cpx<T>        o[32];
    const int64_t step = N / R;
    for ( int m = 0; m < R; ++m )
    {
        cpx<T> s = mk<T>( 0, 0 );
        for ( int q = 0; q < R; ++q )
        {
            const cpx<T> w = tw[( ( q * m ) % R ) * step];
            s              = add( s, twmul<INV>( v[q], w ) );
        }
        o[m] = s;
    }
    for ( int m = 0; m < R; ++m )
        v[m] = o[m];
}

template <int R, bool INV, typename T>
__device__ __forceinline__ void dft_fixed( cpx<T> *v, const cpx<T> *__restrict__ tw, int64_t N )
{
    if constexpr ( R == 2 )
        dft2<INV>( v );
    else if constexpr ( R == 3 )
        dft3<INV>( v );
    else if constexpr ( R == 4 )
        dft4<INV>( v );
    else if constexpr ( R == 5 )
        dft5<INV>( v );
    else if constexpr ( R == 7 )
        dft7<INV>( v );
    else if constexpr ( R == 8 )
        dft8<INV>( v );
    else if constexpr ( R == 16 )
        dft16<INV>( v );
    else
        dft_generic<INV>( v, R, tw, N );
}

// Runtime-radix dispatch (generic kernel)
template <bool INV, typename T>
__device__ __forceinline__ void dft_runtime( cpx<T> *v, int R, const cpx<T> *__restrict__ tw, int64_t N )
{
    switch ( R )
    {
    case 2: dft2<INV>( v ); break;
    case 3: dft3<INV>( v ); break;
    case 4: dft4<INV>( v ); break;
    case 5: dft5<INV>( v ); break;
    case 7: dft7<INV>( v ); break;
    case 8: dft8<INV>( v ); break;
    case 16: dft16<INV>( v ); break;
    default: dft_generic<INV>( v, R, tw, N ); break;
    }
}

// ---------------------------------------------------------------------------
// Workgroup -> tile mapping that keeps neighbouring tiles on one XCD.
// Workgroups are dispatched round-robin over the 8 XCDs (wg i -> XCD i % 8), so
// consecutive tiles (which share cache lines at their edges) are assigned to the
// workgroups i, i+8, i+16, ... of the same XCD and its L2.
// ---------------------------------------------------------------------------
// Chunks of CH consecutive tiles per XCD, the chunks dealt to the XCDs round-robin
// (workgroup w runs on XCD w % 8).  Workgroups past the last whole round keep their own
// index, so the mapping stays a permutation of [0, nwg).
__device__ __forceinline__ int64_t xcd_chunk( int64_t wg, int64_t nwg, int64_t ch )
{
    constexpr int64_t NX   = 8;
    const int64_t     full = nwg / ( NX * ch ) * ( NX * ch );
    if ( wg >= full )
        return wg;
    const int64_t x = wg % NX, k = wg / NX;
    return ( ( k / ch ) * NX + x ) * ch + k % ch;
}

__device__ __forceinline__ int64_t xcd_tile( int64_t wg, int64_t nwg )
{
    constexpr int64_t NX = 8;
    if ( nwg < 2 * NX )
        return wg;
    const int64_t xcd  = wg % NX;
    const int64_t k    = wg / NX;
    const int64_t base = nwg / NX, rem = nwg % NX;
    const int64_t start = xcd * base + ( xcd < rem ? xcd : rem );
    return start + k;
}

} // namespace fftm_mi

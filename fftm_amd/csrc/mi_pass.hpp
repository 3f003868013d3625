// One FFT "pass": a batch of length-N transforms along one axis of an array in HBM.
//
// A pass reads every element once and writes every element once.  The batch is a
// set of "columns" (one column = one 1D line to transform), enumerated by a
// flattened index col in [0, ncols) decomposed over up to three batch dimensions
// (col = c0 + d0 * (c1 + d1 * c2)) with independent input and output strides, so
// a pass can also permute the batch layout (the 4D stage transposes of ffts are
// folded into the passes this way).
//
// Two thread mappings, chosen by the input geometry:
//   A  (element stride 1: the line is contiguous)       -> lanes walk along the line
//   B  (line strided, adjacent columns contiguous)       -> lanes walk across columns
// so that first-stage loads and last-stage stores are 128-byte coalesced either way.
// When the output geometry wants the other mapping, the last stage is re-staged
// through LDS before the store.
//
// Real axes use the two-for-one trick: real columns 2q and 2q+1 are packed into one
// complex line z = a + i b, transformed once, and separated with Hermitian symmetry
// (R2C: A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = (Z[k] - conj Z[N-k]) / 2i; C2R: the
// inverse, built at load time from the half spectra, ignoring the imaginary parts
// of the DC and Nyquist bins).  This works for every N and for both mappings.
//
// The 1D transform is a Stockham autosort FFT: stage s with radix R and span Ns
// reads x[j + r N/R], twiddles by W_{Ns R}^{r (j mod Ns)}, runs a radix-R DFT and
// writes y[(j / Ns) Ns R + j mod Ns + m Ns].  Stage 0 reads HBM, the last stage
// writes HBM (or LDS when post-processing / re-staging), the stages in between
// exchange through LDS.  Twiddles come from a length-N table e^{-2 pi i k / N}
// generated on the host in FP64 (long double) and rounded once.
#pragma once

#include "mi_complex.hpp"
#include "mi_mfma.hpp"

#include <type_traits>

namespace fftm_mi
{

enum : int
{
    KIND_C2C = 0,
    KIND_R2C = 1,
    KIND_C2R = 2
};

template <typename T>
struct pass_params
{
    const void   *in;
    void         *out;
    const cpx<T> *tw;
    int64_t       ies, oes; // element strides along the transformed axis
    int64_t       ncols;    // columns (real columns for R2C / C2R)
    int64_t       d0, d1;   // batch dims 0 and 1 (dim 2 is the rest)
    int64_t       is0, is1, is2;
    int64_t       os0, os1, os2;
    int32_t       store_b; // 1: store with the column-fastest mapping, 0: line-fastest
    int32_t       wmode;   // real-axis bin weights: 0 none, 1 R2C output x w_k, 2 C2R input / w_k
    int32_t       vec;     // bit 0: paired (8/16-byte) loads allowed, bit 1: paired stores allowed
    T             scale;   // multiplies every output (1 for the plain transforms)
    // FP16-stored spectra (mixed precision): every stored line carries a power-of-two
    // exponent; element value = FP16 value * 2^exponent.  Exponent indices are linear in
    // the coordinates, like the data offsets.
    const int8_t *xin;                  // exponents of the FP16 input lines
    int8_t       *xout;                 // exponents of the FP16 output lines (written here)
    int64_t       xs0, xs1, xs2, xsn;   // input exponent strides: batch dims, transform axis
    int64_t       ys0, ys1, ys2;        // output exponent strides: batch dims
    // division by d0 and d1 as a multiply and a shift (q = col * mg >> sh, exact for every
    // col < 2^31; mg = 0: divide)
    uint64_t      mg0, mg1;
    int32_t       sh0, sh1;
};

// n / d with the pass's multiply-and-shift constants (n < 2^31 whenever m != 0).  MG: the
// launch always carries the constants, so no 64-bit division is compiled (pass kernels of
// 32 or more points: their launches with more columns than the constants cover are refused
// on the host, and such an array would exceed the device memory anyway)
template <bool MG = false>
__device__ __forceinline__ int64_t div_col( int64_t n, int64_t d, uint64_t m, int sh )
{
    if constexpr ( MG )
        return (int64_t)( ( (uint64_t)n * m ) >> sh );
    else
        return m ? (int64_t)( ( (uint64_t)n * m ) >> sh ) : n / d;
}

typedef _Float16 h16x2 __attribute__( ( ext_vector_type( 2 ) ) );
typedef _Float16 h16x4 __attribute__( ( ext_vector_type( 4 ) ) );

// storage modes of a pass (template parameter IOM): bit 0 FP16 input, bit 1 FP16 output
enum : int
{
    IO_F32    = 0,
    IO_H16_IN = 1,
    IO_H16_OUT = 2,
    // with IO_H16_IN / IO_H16_OUT: every FP16-stored bin is a hi + lo pair (8 bytes,
    // {hi.re, hi.im, lo.re, lo.im}, value = (hi + lo 2^-11) 2^e): FP32-level precision
    IO_SPLIT = 4
};

// two consecutive elements of T (8 bytes for float, 16 for double) and two consecutive
// FP32 complex values (16 bytes)
template <typename T>
struct pair_of;
template <>
struct pair_of<float>
{
    typedef float type __attribute__( ( ext_vector_type( 2 ) ) );
};
template <>
struct pair_of<double>
{
    typedef double type __attribute__( ( ext_vector_type( 2 ) ) );
};
typedef float f32x4 __attribute__( ( ext_vector_type( 4 ) ) );

// Diagonal spectral operator applied between a forward and an inverse transform of the
// same lines (the fused filter pass): H(k) from the wavenumber vector of each bin.
enum : int
{
    OP_POISSON       = 1, // H = -scale / |k|^2, H(0) = 0   (periodic -Laplace inverse)
    OP_GAUSSIAN      = 2, // H = scale * exp(-|k|^2 / (2 sigma^2))
    OP_SHELL         = 3, // H = scale * coeffs[min(S - 1, floor(S |k| / kmax))]
    OP_DERIVATIVE    = 4, // H = i scale k_a                 (d / dx_a)
    OP_POISSON_GRAD  = 5, // H = -i scale k_a / |k|^2, H(0) = 0 (gradient of the Poisson solve)
    OP_BANDPASS      = 6  // H = scale where klo <= |k| <= khi, else 0
};

template <typename T>
struct filter_params
{
    int32_t  kind;
    int32_t  nshells;
    int32_t  nd;       // memory axes of the (dense) spectrum array, slowest first
    int32_t  dax;      // memory axis of the derivative (OP_DERIVATIVE, OP_POISSON_GRAD)
    int32_t  line_m; // memory axis of the fused filter pass's lines
    int32_t  pad_;
    int64_t  ext[4];   // memory extents
    int64_t  nlen[4];  // full logical length of each memory axis; 0 marks the halved real axis
    double   scale;
    double   inv_two_sigma2;
    int64_t  kmax2;    // sum over axes of (n_a / 2)^2 (radial shells)
    double   klo2, khi2; // OP_BANDPASS: squared band edges
    const T *coeffs;   // OP_SHELL coefficients (T[nshells])
    uint64_t mg[4];    // division by ext[m] as a multiply and a shift (mg = 0: divide)
    int32_t  sh[4];
};

// Radial shell of a bin with |k|^2 = k2: the largest s with s^2 kmax2 <= S^2 k2, i.e.
// floor(S |k| / kmax) evaluated exactly, clamped to S - 1.  Integer arithmetic: on bins
// where S |k| / kmax is an integer (|k|^2 = 3 on a 6^3 cube with S = 3) a floating
// floor(sqrt(k2) * S / kmax) can land one shell low.  S <= 1024 and k2, kmax2 < 2^25 keep
// every product below 2^46.
__device__ __forceinline__ int shell_of( int64_t k2, int S, int64_t kmax2 )
{
    const int64_t rhs = (int64_t)S * S * k2;
    int           s   = (int)sqrt( (double)rhs / (double)kmax2 ); // within one of the answer
#pragma unroll
    for ( int i = 0; i < 2; ++i )
        s = ( s > 0 && (int64_t)s * s * kmax2 > rhs ) ? s - 1 : s;
#pragma unroll
    for ( int i = 0; i < 2; ++i )
        s = ( (int64_t)( s + 1 ) * ( s + 1 ) * kmax2 <= rhs ) ? s + 1 : s;
    return s < S - 1 ? s : S - 1;
}

// wavenumbers of the bin at spectrum offset off: returns sum_a k_a^2 and sets kd to the
// wavenumber along memory axis f.dax (k = i on the halved axis, k = i <= n/2 ? i : i - n
// elsewhere)
template <typename T>
__device__ __forceinline__ float filter_k2( const filter_params<T> &f, int64_t off, float &kd )
{
    float k2 = 0.f;
    kd       = 0.f;
    for ( int m = f.nd - 1; m >= 0; --m )
    {
        const int64_t q   = div_col( off, f.ext[m], f.mg[m], f.sh[m] );
        const int64_t idx = off - q * f.ext[m];
        off               = q;
        const int64_t n = f.nlen[m];
        const int64_t k = ( n == 0 || idx <= n / 2 ) ? idx : idx - n;
        k2 += (float)( k * k );
        if ( m == f.dax )
            kd = (float)k;
    }
    return k2;
}

// H(k) as (real, imaginary) parts
template <typename T>
__device__ __forceinline__ cpx<T> filter_h( const filter_params<T> &f, float k2, float kd )
{
    switch ( f.kind )
    {
    case OP_POISSON: return mk<T>( k2 == 0.f ? (T)0 : (T)( -f.scale ) / (T)k2, (T)0 );
    case OP_GAUSSIAN:
        if constexpr ( sizeof( T ) == 4 )
            return mk<T>( (T)f.scale * __expf( -k2 * (float)f.inv_two_sigma2 ), (T)0 );
        else
            return mk<T>( (T)( f.scale * exp( -(double)k2 * f.inv_two_sigma2 ) ), (T)0 );
    case OP_DERIVATIVE: return mk<T>( (T)0, (T)f.scale * (T)kd );
    case OP_POISSON_GRAD: return mk<T>( (T)0, k2 == 0.f ? (T)0 : (T)( -f.scale ) * (T)kd / (T)k2 );
    case OP_BANDPASS:
        return mk<T>( ( (double)k2 >= f.klo2 && (double)k2 <= f.khi2 ) ? (T)f.scale : (T)0, (T)0 );
    default:
        // k2 is a sum of squared integers below 2^24, exact in FP32
        return mk<T>( (T)f.scale * f.coeffs[shell_of( (int64_t)k2, f.nshells, f.kmax2 )], (T)0 );
    }
}

// z * H with H real or purely imaginary (exact: no rounding from zero parts)
template <typename T>
__device__ __forceinline__ cpx<T> apply_h( cpx<T> z, cpx<T> h, int kind )
{
    if ( kind == OP_DERIVATIVE || kind == OP_POISSON_GRAD )
        return mk<T>( -z.y * h.y, z.x * h.y );
    return mk<T>( z.x * h.x, z.y * h.x );
}

// per-workgroup state of a pass kernel (shared by the forward and inverse halves of
// the fused filter pass)
template <typename T, typename LW>
struct pass_ctx
{
    const pass_params<T>   *p;
    const filter_params<T> *f;
    LW                     *lds;
    const cpx<T>           *twl; // stage twiddles: the workgroup's LDS copy, or p->tw
    cpx<T>                 *twb; // the LDS copy, filled from twr at the first stage exchange
    cpx<T>                  twr[4];
    bool                    twpend;
    int                     c, t;
    int64_t                 tile;
    int64_t                 ia, ib, oa, ob;
    int64_t                 xa, xb, ya, yb; // exponent indices (FP16 storage)
    bool                    va, vb;
};

// Runtime description for the generic (any-radix) kernel.
struct radix_list
{
    int32_t n;      // number of stages
    int32_t r[20];  // radices (<= 16 each)
};

// ---------------------------------------------------------------------------
// compile-time radix schedule for a given N
// ---------------------------------------------------------------------------
struct rplan
{
    int n;
    int r[24];
};

// big_last moves a leading radix 16 behind the radix-8 stages (lines read from HBM in
// pairs of consecutive elements need an even number of butterflies per thread in stage 0)
constexpr rplan make_rplan( int N, bool big_last = false )
{
    rplan p{ 0, {} };
    int   m = N, k = 0;
    while ( m % 2 == 0 )
    {
        m /= 2;
        ++k;
    }
    if ( k % 3 == 1 && k >= 4 )
    {
        p.r[p.n++] = 16;
        k -= 4;
    }
    else if ( k % 3 == 2 )
    {
        p.r[p.n++] = 4;
        k -= 2;
    }
    else if ( k % 3 == 1 )
    {
        p.r[p.n++] = 2;
        k -= 1;
    }
    while ( k >= 3 )
    {
        p.r[p.n++] = 8;
        k -= 3;
    }
    if ( big_last && p.n > 1 && p.r[0] == 16 )
    {
        for ( int i = 0; i + 1 < p.n; ++i )
            p.r[i] = p.r[i + 1];
        p.r[p.n - 1] = 16;
    }
    for ( int f = 3; f <= m; f += 2 )
        while ( m % f == 0 )
        {
            p.r[p.n++] = f;
            m /= f;
        }
    return p;
}

// The alternative radix schedule of a power-of-two length (a tuning candidate): radix 8,
// then radix 16 in the second stage, then radix 8s and a final 4 or 2, e.g. 512 = 8 16 4,
// 1024 = 8 16 8.  Its radix-16 stage is an LDS-to-LDS stage, which the matrix-core variant
// runs as FP16 MFMA products (dft16_mfma).
constexpr rplan make_rplan_alt( int N )
{
    int k = 0;
    for ( int m = N; m > 1 && m % 2 == 0; m /= 2 )
        ++k;
    if ( ( 1 << k ) != N || k < 7 )
        return make_rplan( N );
    rplan p{ 0, {} };
    p.r[p.n++] = 8;
    p.r[p.n++] = 16;
    k -= 7;
    while ( k >= 3 )
    {
        p.r[p.n++] = 8;
        k -= 3;
    }
    if ( k > 0 )
        p.r[p.n++] = 1 << k;
    return p;
}

constexpr int rplan_max( const rplan &p )
{
    int r = 1;
    for ( int i = 0; i < p.n; ++i )
        r = p.r[i] > r ? p.r[i] : r;
    return r;
}

constexpr int rplan_span( const rplan &p, int s ) // product of radices before stage s
{
    int v = 1;
    for ( int i = 0; i < s; ++i )
        v *= p.r[i];
    return v;
}

// ---------------------------------------------------------------------------
// Launch geometry (shared by the kernel instantiations and the host launcher)
// ---------------------------------------------------------------------------
struct geom
{
    int tw;  // columns (complex lines) per workgroup
    int tpc; // threads per column
    int pad; // mode A: one pad element every 2^pad elements (0 = no padding)
};

constexpr int lds_limit_bytes() { return 65536; }
constexpr int lds_limit_wide_bytes() { return 163840; }

constexpr int geom_pa( int N, int pad ) { return pad ? N + ( N >> pad ) : N; }

// var selects a tile-shape variant (0 = default); the tuner measures them per pass:
//   1: half the columns per workgroup (smaller LDS tile, more workgroups per CU)
//   2: twice the threads per column (fewer elements per thread, larger workgroup)
//   3: the tile of 4 exchanged through LDS in two rounds (real parts, then imaginary
//      parts): 256-byte row segments at variant 0's LDS per workgroup
//   4: twice the columns (256-byte row segments; LDS up to 160 KB on gfx950)
//   5: twice the columns and twice the threads per column (1024-thread workgroups)
//   6: the tile of 0, exchanged through LDS in two rounds (real parts, then imaginary
//      parts): half the LDS per workgroup, so more tiles in flight per CU
//   7: the tile of 0 exchanged in two rounds, registers capped for 4 waves per SIMD
//      (some spilling; wins on the longest contiguous R2C lines)
constexpr int NUM_VARIANTS = 8;
// A pass's variant number is shape + 8 * order.  The tile order is a runtime choice:
//   0: each XCD takes one contiguous range of tiles (xcd_tile), so neighbouring tiles
//      share one L2;
//   1: tile t runs as workgroup t, so consecutive tiles go to different XCDs;
//   2: chunks of 8 consecutive tiles per XCD, dealt round-robin (xcd_chunk);
//   3: chunks of 32.
// The host passes the order to the kernel in pass_params::vec (VEC_RR, VEC_CHUNK8,
// VEC_CHUNK32).
constexpr int NUM_ORDERS  = 4;
constexpr int VEC_RR      = 4;
constexpr int VEC_CHUNK8  = 8;
constexpr int VEC_CHUNK32 = 16;
// bit 5 of pass_params::vec: run the tiles in reverse order (last tile first), so that a
// pass reads first what the pass before it wrote last, while that is still in the
// memory-side cache
constexpr int VEC_REV     = 32;
constexpr int variant_shape_of( int var ) { return var & 7; }
constexpr int variant_order_of( int var ) { return ( var >> 3 ) & 3; }
constexpr int variant_rs_of( int var ) { return var >> 5; }

__device__ __forceinline__ int64_t ordered_tile( int order_bits )
{
    const int64_t wg = ( order_bits & VEC_REV ) ? (int64_t)gridDim.x - 1 - blockIdx.x : (int64_t)blockIdx.x;
    if ( order_bits & VEC_RR )
        return wg;
    if ( order_bits & VEC_CHUNK8 )
        return xcd_chunk( wg, gridDim.x, 8 );
    if ( order_bits & VEC_CHUNK32 )
        return xcd_chunk( wg, gridDim.x, 32 );
    return xcd_tile( wg, gridDim.x );
}

constexpr bool variant_split( int var ) { return var >= 6 || var == 3; }
constexpr int  variant_shape( int var ) { return var >= 6 ? 0 : ( var == 3 ? 4 : var ); }

constexpr geom pick_geom( int N, int csize, bool loadb, int var = 0 )
{
    const rplan p  = make_rplan( N );
    const int   rm = rplan_max( p );
    int         lo = N / 32 > 1 ? N / 32 : 1;
    int         hi = N / rm > 1 ? N / rm : 1;
    if ( lo > hi )
        lo = hi;
    geom g{ 1, 1, 3 };
    if ( loadb )
    {
        const int tw0 = 128 / csize; // one 128-byte row segment
        int       tpc = 256 / tw0;
        tpc           = tpc < lo ? lo : ( tpc > hi ? hi : tpc );
        int tw        = 256 / tpc > 0 ? 256 / tpc : 1;
        while ( tw > 1 && N * tw * csize > lds_limit_bytes() )
        {
            tw /= 2;
            if ( tpc * 2 <= hi )
                tpc *= 2;
        }
        g.tw  = tw;
        g.tpc = tpc;
        g.pad = 0;
    }
    else
    {
        int tpc = hi > 256 ? 256 : hi;
        if ( tpc < lo )
            tpc = lo;
        // two butterflies per thread when the tile still fits: the first and last stages
        // then move pairs of consecutive elements (8- or 16-byte accesses)
        if ( tpc >= 2 && tpc / 2 >= lo && ( 256 / ( tpc / 2 ) ) * geom_pa( N, 3 ) * csize <= lds_limit_bytes() )
            tpc /= 2;
        int tw  = 256 / tpc > 0 ? 256 / tpc : 1;
        int pad = 3;
        while ( tw > 1 && tw * geom_pa( N, pad ) * csize > lds_limit_bytes() )
            tw /= 2;
        if ( tw * geom_pa( N, pad ) * csize > lds_limit_bytes() )
            pad = 0;
        g.tw  = tw;
        g.tpc = tpc;
        g.pad = pad;
    }
    if ( var == 1 && g.tw > 1 )
        g.tw /= 2;
    if ( var == 2 && g.tpc * 2 <= hi && g.tpc * 2 * g.tw <= 1024 )
        g.tpc *= 2;
    if ( var == 4 || var == 5 )
    {
        const int pa    = loadb ? N : geom_pa( N, g.pad );
        const int tw2   = g.tw * 2;
        const int tpc2  = var == 5 && g.tpc * 2 <= hi ? g.tpc * 2 : g.tpc;
        if ( pa * tw2 * csize <= lds_limit_wide_bytes() && tw2 * tpc2 <= 1024 )
        {
            g.tw  = tw2;
            g.tpc = tpc2;
        }
    }
    return g;
}

// ---------------------------------------------------------------------------
// column address helpers
// ---------------------------------------------------------------------------
template <bool MG = false, typename T>
__device__ __forceinline__ void col_offsets( const pass_params<T> &p, int64_t col, int64_t &io, int64_t &oo )
{
    const int64_t r  = div_col<MG>( col, p.d0, p.mg0, p.sh0 );
    const int64_t c0 = col - r * p.d0;
    const int64_t c2 = div_col<MG>( r, p.d1, p.mg1, p.sh1 );
    const int64_t c1 = r - c2 * p.d1;
    io               = c0 * p.is0 + c1 * p.is1 + c2 * p.is2;
    oo               = c0 * p.os0 + c1 * p.os1 + c2 * p.os2;
}

// exponent indices of column col: input lines (consumer side), output line (producer side)
template <bool MG = false, typename T>
__device__ __forceinline__ void col_exponents( const pass_params<T> &p, int64_t col, int64_t &xi, int64_t &yo )
{
    const int64_t r  = div_col<MG>( col, p.d0, p.mg0, p.sh0 );
    const int64_t c0 = col - r * p.d0;
    const int64_t c2 = div_col<MG>( r, p.d1, p.mg1, p.sh1 );
    const int64_t c1 = r - c2 * p.d1;
    xi               = c0 * p.xs0 + c1 * p.xs1 + c2 * p.xs2;
    yo               = c0 * p.ys0 + c1 * p.ys1 + c2 * p.ys2;
}

// storage element of an FP16-stored bin: one FP16 pair, or a hi + lo pair of pairs
template <bool SPLIT>
struct h16_elem
{
    typedef h16x2 type;
};
template <>
struct h16_elem<true>
{
    typedef h16x4 type;
};

// FP16-stored bin * 2^e in FP32
template <bool SPLIT>
__device__ __forceinline__ cpx<float> h16_decode( typename h16_elem<SPLIT>::type h, int e )
{
    if constexpr ( SPLIT )
        return mk<float>( __builtin_ldexpf( (float)h.x + (float)h.z * ( 1.0f / 2048.0f ), e ),
                          __builtin_ldexpf( (float)h.y + (float)h.w * ( 1.0f / 2048.0f ), e ) );
    else
        return mk<float>( __builtin_ldexpf( (float)h.x, e ), __builtin_ldexpf( (float)h.y, e ) );
}

template <bool SPLIT>
__device__ __forceinline__ cpx<float> h16_load( const void *base, int64_t idx, const int8_t *xp, int64_t xidx )
{
    using E = typename h16_elem<SPLIT>::type;
    return h16_decode<SPLIT>( static_cast<const E *>( base )[idx], xp[xidx] );
}

// exponent of a line whose largest component magnitude is amax: the stored values
// amax * 2^-e then lie in [2^14, 2^15), inside the FP16 range (65504)
__device__ __forceinline__ int h16_exponent( float amax )
{
    int e = 0;
    (void)__builtin_frexpf( amax, &e );
    e -= 15;
    return e < -127 ? -127 : ( e > 127 ? 127 : e );
}

// (a, b) 2^-e rounded to FP16; with SPLIT also the rounding residuals times 2^11 (at most
// half an FP16 ulp of a value below 2^15, i.e. 8, so the scaled residual stays below 2^14)
template <bool SPLIT>
__device__ __forceinline__ typename h16_elem<SPLIT>::type h16_pack( float a, float b, int e )
{
    const float    sa = __builtin_ldexpf( a, -e ), sb = __builtin_ldexpf( b, -e );
    const _Float16 ha = (_Float16)sa, hb = (_Float16)sb;
    if constexpr ( SPLIT )
        return h16x4{ ha, hb, (_Float16)( ( sa - (float)ha ) * 2048.0f ), (_Float16)( ( sb - (float)hb ) * 2048.0f ) };
    else
        return h16x2{ ha, hb };
}

// element n of an FP16-stored input line (C2C) or of the Hermitian extension of two
// FP16-stored half spectra (C2R), as in load_line
template <int KIND, bool SPLIT>
__device__ __forceinline__ cpx<float> load_line_h16( const pass_params<float> &p, int64_t ia, int64_t ib, int64_t xa,
                                                     int64_t xb, bool va, bool vb, int n, int N )
{
    // unconditional loads (a column past the end reads element 0, which exists) and a
    // select afterwards: a load under a branch makes the compiler wait for it at the
    // join, one memory latency per element
    if constexpr ( KIND == KIND_C2C )
    {
        const cpx<float> v = h16_load<SPLIT>( p.in, va ? ia + (int64_t)n * p.ies : 0, p.xin, va ? xa + (int64_t)n * p.xsn : 0 );
        return va ? v : mk<float>( 0, 0 );
    }
    else
    {
        const int  h  = N / 2;
        const int  nn = n <= h ? n : N - n;
        cpx<float> a  = h16_load<SPLIT>( p.in, va ? ia + (int64_t)nn * p.ies : 0, p.xin, va ? xa + (int64_t)nn * p.xsn : 0 );
        cpx<float> b  = h16_load<SPLIT>( p.in, vb ? ib + (int64_t)nn * p.ies : 0, p.xin, vb ? xb + (int64_t)nn * p.xsn : 0 );
        if ( !va )
            a = mk<float>( 0, 0 );
        if ( !vb )
            b = mk<float>( 0, 0 );
        if ( n > h )
        {
            a.y = -a.y;
            b.y = -b.y;
        }
        if ( n == 0 || 2 * n == N )
        {
            a.y = 0;
            b.y = 0;
        }
        else if ( p.wmode == 2 )
        {
            a = scl( a, 0.5f );
            b = scl( b, 0.5f );
        }
        return mk<float>( a.x - b.y, a.y + b.x );
    }
}

// element n of the (possibly paired) input line, as one complex value
template <int KIND, typename T>
__device__ __forceinline__ cpx<T> load_line( const pass_params<T> &p, int64_t ia, int64_t ib, bool va, bool vb,
                                             int n, int N )
{
    if constexpr ( KIND == KIND_C2C )
    {
        if ( !va )
            return mk<T>( 0, 0 );
        return static_cast<const cpx<T> *>( p.in )[ia + (int64_t)n * p.ies];
    }
    else if constexpr ( KIND == KIND_R2C )
    {
        const T *src = static_cast<const T *>( p.in );
        const T  a   = va ? src[ia + (int64_t)n * p.ies] : (T)0;
        const T  b   = vb ? src[ib + (int64_t)n * p.ies] : (T)0;
        return mk<T>( a, b );
    }
    else
    {
        const cpx<T> *src = static_cast<const cpx<T> *>( p.in );
        const int     h   = N / 2;
        const int     nn  = n <= h ? n : N - n;
        cpx<T>        a   = va ? src[ia + (int64_t)nn * p.ies] : mk<T>( 0, 0 );
        cpx<T>        b   = vb ? src[ib + (int64_t)nn * p.ies] : mk<T>( 0, 0 );
        if ( n > h )
        {
            a.y = -a.y;
            b.y = -b.y;
        }
        if ( n == 0 || 2 * n == N )
        {
            a.y = 0;
            b.y = 0;
        }
        else if ( p.wmode == 2 )
        {
            // adjoint of R2C: interior bins of the half spectrum carry weight 1/2
            a = scl( a, (T)0.5 );
            b = scl( b, (T)0.5 );
        }
        // z = A~ + i B~
        return mk<T>( a.x - b.y, a.y + b.x );
    }
}

// store element n of a transformed line (C2C / C2R; R2C goes through r2c_store)
template <int KIND, typename T>
__device__ __forceinline__ void store_line( const pass_params<T> &p, int64_t oa, int64_t ob, bool va, bool vb, int n,
                                            cpx<T> z )
{
    if constexpr ( KIND == KIND_C2C )
    {
        if ( va )
            static_cast<cpx<T> *>( p.out )[oa + (int64_t)n * p.oes] = scl( z, p.scale );
    }
    else
    {
        T *dst = static_cast<T *>( p.out );
        if ( va )
            dst[oa + (int64_t)n * p.oes] = z.x * p.scale;
        if ( vb )
            dst[ob + (int64_t)n * p.oes] = z.y * p.scale;
    }
}

// ---------------------------------------------------------------------------
// the compile-time-N pass kernel
// ---------------------------------------------------------------------------
// RS selects the radix schedule: 0 the default (make_rplan), 1 the alternative
// (make_rplan_alt), 2 the alternative with its radix-16 stage on the matrix cores
// (FP16 hi + lo operands, FP32 accumulation; strided FP32-arithmetic C2C passes)
template <typename T, int N, int KIND, bool INV, bool LOADB, int VAR = 0, int IOM = IO_F32, bool FLT = false,
          int RS = 0>
struct pass_kernel
{
    static_assert( RS == 0 || ( KIND == KIND_C2C && LOADB && !FLT && std::is_same<T, float>::value ),
                   "alternative radix schedules are compiled for strided FP32 C2C passes" );
    static_assert( !FLT || ( KIND == KIND_C2C && !INV && IOM == IO_F32 && !variant_split( VAR ) ),
                   "the fused filter pass is a forward C2C pass with whole-value LDS exchange" );
    static_assert( IOM == IO_F32 || std::is_same<T, float>::value, "FP16 storage is for FP32 transforms" );
    static_assert( IOM != IO_SPLIT, "hi + lo storage qualifies an FP16-stored side" );
    static constexpr bool SPL = ( IOM & IO_SPLIT ) != 0;
    using HE                  = typename h16_elem<SPL>::type;
    static_assert( !( IOM & IO_H16_IN ) || KIND != KIND_R2C, "R2C reads real FP32 input" );
    static_assert( !( IOM & IO_H16_OUT ) || KIND != KIND_C2R, "C2R writes real FP32 output" );
    using C                    = cpx<T>;
    // R2C restages its output through LDS, so only its stage-0 loads want pairs
    static constexpr rplan RP  = RS ? make_rplan_alt( N ) : make_rplan( N, KIND == KIND_R2C );
    static constexpr bool  SPLIT = variant_split( VAR );
    // FP16-stored lines move 4 bytes per bin: their variant 7 doubles the columns (128-byte
    // FP16 row segments) with the split exchange, instead of capping registers
    static constexpr int   SHAPE = ( IOM != IO_F32 && VAR == 7 ) ? 4 : variant_shape( VAR );
    static constexpr geom  G   = pick_geom( N, (int)sizeof( C ), LOADB, SHAPE );
    static constexpr int   TW  = G.tw;
    static constexpr int   TPC = G.tpc;
    static constexpr int   NT  = TW * TPC;
    static constexpr int   PA  = geom_pa( N, G.pad );
    static constexpr int   PB  = TW;
    static constexpr int   LDS = LOADB ? N * PB : TW * PA;
    static constexpr int   CPT = KIND == KIND_C2C ? 1 : 2;
    // column arithmetic by the multiply-and-shift constants alone (see div_col)
    static constexpr bool  MG  = N >= 32;
    static_assert( LDS * (int)sizeof( C ) / ( SPLIT ? 2 : 1 ) <=
                       ( SHAPE >= 4 ? lds_limit_wide_bytes() : lds_limit_bytes() ),
                   "LDS tile too large" );
    // LDS words: complex values, or one component at a time with the split exchange
    using LW = typename std::conditional<SPLIT, T, C>::type;
    // C2R along contiguous lines stages its half spectra in LDS (whole-value words)
    static constexpr bool STAGE_C2R = KIND == KIND_C2R && !LOADB && !SPLIT && 2 * TW * ( N / 2 + 1 ) <= LDS;
    // the split exchange halves the LDS tile so that more workgroups share a CU: ask the
    // register allocator for the matching occupancy (waves per SIMD)
    static constexpr int WAVES_PER_EU = [] {
        if ( ( VAR != 7 && VAR != 3 ) || IOM != IO_F32 )
            return 1;
        const int lds_wg  = LDS * (int)sizeof( LW );
        const int wgs     = lds_wg > 0 ? 163840 / lds_wg : 8;
        const int waves   = wgs * ( NT / 64 ) / 4;
        return waves < 1 ? 1 : ( waves > 4 ? 4 : waves );
    }();
    // register-capped kernels form the output offsets again at the direct store instead of
    // holding them through the stages (held, they were spilled to scratch: 8 bytes written
    // per thread)
    static constexpr bool OUT_LATE = MG && WAVES_PER_EU > 1;

    __device__ static __forceinline__ int L( int n, int c )
    {
        if constexpr ( LOADB )
            return n * PB + c;
        else
            return c * PA + ( G.pad ? n + ( n >> G.pad ) : n );
    }

    // Stage S handles butterflies j in [0, N / R_S).  Thread t takes j = t + TPC b, except
    // in the first and last stages of the line-contiguous mapping, where it takes pairs
    // (2 (t + TPC (b / 2)) + b % 2) so that its HBM accesses cover consecutive elements.
    template <int S>
    static constexpr bool pair_stage()
    {
        if constexpr ( LOADB )
            return false;
        else
        {
            constexpr int NBF = N / RP.r[S];
            constexpr int NB  = ( NBF + TPC - 1 ) / TPC;
            return ( S == 0 || S == RP.n - 1 ) && NB % 2 == 0 && NBF % 2 == 0;
        }
    }
    template <int S>
    __device__ static __forceinline__ int jmap( int t, int b )
    {
        if constexpr ( pair_stage<S>() )
            return 2 * ( t + TPC * ( b >> 1 ) ) + ( b & 1 );
        else
            return t + TPC * b;
    }
    // paired HBM accesses exist for real lines (two reals) and FP32 complex lines; the
    // real side of an R2C / C2R pass stays FP32 when its spectrum side is FP16-stored
    static constexpr bool PAIR_LOAD  = !( IOM & IO_H16_IN ) && pair_stage<0>() &&
                                      ( KIND == KIND_R2C || ( KIND == KIND_C2C && sizeof( T ) == 4 ) );
    static constexpr bool PAIR_STORE = !( IOM & IO_H16_OUT ) && pair_stage<RP.n - 1>() &&
                                       ( KIND == KIND_C2R || ( KIND == KIND_C2C && sizeof( T ) == 4 ) );

    // largest component magnitude over the TPC threads that share a column: each thread
    // passes its own maximum; column-fastest thread order when cf
    __device__ static __forceinline__ float column_max( float *red, float lm, int c, bool cf )
    {
        red[threadIdx.x] = lm;
        __syncthreads();
        float m = 0.f;
#pragma unroll 8
        for ( int t = 0; t < TPC; ++t )
            m = fmaxf( m, red[cf ? t * TW + c : c * TPC + t] );
        return m;
    }
    __device__ static __forceinline__ float cmag( const C &z ) { return fmaxf( fabsf( (float)z.x ), fabsf( (float)z.y ) ); }

    using ctx = pass_ctx<T, LW>;

    // The stage twiddles are read from an LDS copy of the table: a twiddle load through
    // the vector memory path returns behind the tile loads issued before it, so under full
    // HBM traffic every stage would wait on the memory queue.  LDS loads have their own
    // counter.  Same values, same arithmetic.
    static constexpr bool TWL = RP.n > 1 && N * (int)sizeof( C ) <= 8192 &&
                                LDS * (int)sizeof( LW ) + N * (int)sizeof( C ) <= lds_limit_wide_bytes() - 8192;
    // The table is loaded into registers before the tile (at most 4 values per thread) and
    // written to LDS after the tile's loads are issued, so that no tile load waits for it.
    static constexpr int  TWPT = ( N + NT - 1 ) / NT;
    static constexpr bool TWR  = TWL && TWPT <= 4;
    // Strided FP16 lines whose exponents run along the tile's columns stage the tile's
    // exponents in LDS, loaded as dwords (XND per element index): the dwords themselves
    // (XDW), or only the N x TW exponent bytes where the dwords would take LDS that a
    // workgroup needs.  Either is used where the extra LDS leaves the number of workgroups
    // per CU unchanged, or at 3 or more (register use holds these kernels to 3 or fewer
    // workgroups per CU).  XBASE: the kernel's other LDS arrays.
    static constexpr int XND   = TW / 4 + 1;
    static constexpr int XBASE = LDS * (int)sizeof( LW ) + ( TWL ? N * (int)sizeof( C ) : 0 ) + 2 * N +
                                 ( ( IOM & IO_H16_OUT ) ? 8 * NT : 0 );
    static constexpr bool x_lds_ok( int extra )
    {
        const int wg0 = lds_limit_wide_bytes() / XBASE, wg1 = lds_limit_wide_bytes() / ( XBASE + extra );
        return wg1 == wg0 || wg1 >= 3;
    }
    static constexpr bool XDW  = x_lds_ok( N * XND * 4 );
    static constexpr bool XROW = LOADB && KIND == KIND_C2C && ( IOM & IO_H16_IN ) && TW % 4 == 0 &&
                                 ( XDW || x_lds_ok( N * TW ) );
    static constexpr int  XLDS = XDW ? N * XND * 4 : N * TW;

    // ---- LDS exchange helpers: P = 2 moves whole values, P = 0 / 1 the real /
    // imaginary components (split exchange) ------------------------------------------
    template <int P>
    __device__ static __forceinline__ void lds_st( LW *l, int idx, const C &z )
    {
        if constexpr ( P == 2 )
            l[idx] = z;
        else if constexpr ( P == 0 )
            l[idx] = z.x;
        else
            l[idx] = z.y;
    }
    template <int P>
    __device__ static __forceinline__ void lds_ld( const LW *l, int idx, C &z )
    {
        if constexpr ( P == 2 )
            z = l[idx];
        else if constexpr ( P == 0 )
            z.x = l[idx];
        else
            z.y = l[idx];
    }
    // results of stage S (radix R, span NS) into LDS at their Stockham positions
    template <int P, int S, int R, int NS, int NB>
    __device__ static __forceinline__ void put_stage( const C ( &v )[NB][R], ctx &x )
    {
        constexpr int NBF = N / R;
#pragma unroll
        for ( int b = 0; b < NB; ++b )
        {
            const int j = jmap<S>( x.t, b );
            if ( NBF % TPC == 0 || j < NBF )
            {
                const int d = ( j / NS ) * NS * R + ( j % NS );
#pragma unroll
                for ( int m = 0; m < R; ++m )
                    lds_st<P>( x.lds, L( d + m * NS, x.c ), v[b][m] );
            }
        }
    }
    // inputs of stage S (radix R) from LDS
    template <int P, int S, int R, int NB>
    __device__ static __forceinline__ void get_stage( C ( &w )[NB][R], ctx &x )
    {
        constexpr int NBF = N / R;
#pragma unroll
        for ( int b = 0; b < NB; ++b )
        {
            const int j = jmap<S>( x.t, b );
            if ( NBF % TPC == 0 || j < NBF )
            {
#pragma unroll
                for ( int r = 0; r < R; ++r )
                    lds_ld<P>( x.lds, L( j + r * NBF, x.c ), w[b][r] );
            }
        }
    }
    // restaged output elements: element t2 + i TPC of column c2 (and, for R2C, its mirror)
    template <int P, int NI>
    __device__ static __forceinline__ void get_restage( C ( &zk )[NI], C ( &zm )[NI], const LW *l, int c2, int t2 )
    {
        constexpr int LIM = KIND == KIND_R2C ? N / 2 + 1 : N;
#pragma unroll
        for ( int i = 0; i < NI; ++i )
        {
            const int k = t2 + i * TPC;
            if ( LIM % TPC == 0 || k < LIM )
            {
                lds_ld<P>( l, L( k, c2 ), zk[i] );
                if constexpr ( KIND == KIND_R2C )
                    lds_ld<P>( l, L( k == 0 ? 0 : N - k, c2 ), zm[i] );
            }
        }
    }

    // ---- final store of the last stage -------------------------------------------
    template <int R, int NS, int NB>
    __device__ static __forceinline__ void final_store( C ( &v )[NB][R], ctx &x )
    {
        const pass_params<T> &p      = *x.p;
        constexpr int         NBF    = N / R;
        if constexpr ( FLT )
        {
            // forward transform of the lines done (natural order: element d + m NS):
            // multiply by H(k), then run the inverse transform of the same lines from LDS
            // the column's offset gives the wavenumbers off the line axis; the element
            // index along the line gives the line's own
            float       kdc = 0.f;
            const float k2c = filter_k2( *x.f, x.oa, kdc );
            const bool  dline = x.f->dax == x.f->line_m;
#pragma unroll
            for ( int b = 0; b < NB; ++b )
            {
                const int j = jmap<RP.n - 1>( x.t, b );
                if ( NBF % TPC == 0 || j < NBF )
                {
                    const int d = ( j / NS ) * NS * R + ( j % NS );
#pragma unroll
                    for ( int m = 0; m < R; ++m )
                    {
                        const int n  = d + m * NS;
                        const int kk = n <= N / 2 ? n : n - N;
                        v[b][m]      = apply_h( v[b][m],
                                                filter_h( *x.f, k2c + (float)( kk * kk ), dline ? (float)kk : kdc ),
                                                x.f->kind );
                    }
                }
            }
            put_stage<2, RP.n - 1, R, NS, NB>( v, x );
            __syncthreads();
            pass_kernel<T, N, KIND_C2C, true, LOADB, VAR, IO_F32, false>::run_from_lds( x );
            return;
        }
        const bool direct = ( KIND != KIND_R2C ) && ( p.store_b == (int)LOADB );
        if constexpr ( IOM & IO_H16_OUT )
        {
            if ( direct )
            {
                // one exponent per line: the column's largest magnitude over its threads
                __shared__ float red[NT];
                float            lm = 0.f;
#pragma unroll
                for ( int b = 0; b < NB; ++b )
                {
                    const int j = jmap<RP.n - 1>( x.t, b );
                    if ( NBF % TPC == 0 || j < NBF )
                    {
#pragma unroll
                        for ( int m = 0; m < R; ++m )
                            lm = fmaxf( lm, cmag( v[b][m] ) );
                    }
                }
                const int e = h16_exponent( column_max( red, lm, x.c, LOADB ) * fabsf( (float)p.scale ) );
                HE       *dst = static_cast<HE *>( p.out );
#pragma unroll
                for ( int b = 0; b < NB; ++b )
                {
                    const int j = jmap<RP.n - 1>( x.t, b );
                    if ( ( NBF % TPC == 0 || j < NBF ) && x.va )
                    {
                        const int d = ( j / NS ) * NS * R + ( j % NS );
#pragma unroll
                        for ( int m = 0; m < R; ++m )
                            dst[x.oa + (int64_t)( d + m * NS ) * p.oes] =
                                h16_pack<SPL>( v[b][m].x * p.scale, v[b][m].y * p.scale, e );
                    }
                }
                if ( x.t == 0 && x.va )
                    p.xout[x.ya] = (int8_t)e;
                return;
            }
        }
        if ( direct )
        {
            if constexpr ( OUT_LATE )
            {
                // an opaque copy of the column index keeps the compiler from reusing the
                // offsets formed at the start
                int c = x.c;
                asm volatile( "" : "+v"( c ) );
                const int64_t col = ( x.tile * TW + c ) * CPT;
                int64_t       ia, ib;
                col_offsets<MG>( p, col, ia, x.oa );
                if ( CPT == 2 )
                    col_offsets<MG>( p, col + 1, ib, x.ob );
            }
            if constexpr ( PAIR_STORE )
            {
                if ( p.vec & 2 )
                {
                    // last stage: NS = N / R, so butterfly j writes elements j + m NS, and the
                    // pair (j, j + 1) writes two consecutive elements per m
#pragma unroll
                    for ( int b = 0; b < NB; b += 2 )
                    {
                        const int j = jmap<RP.n - 1>( x.t, b );
                        if ( NBF % TPC == 0 || j < NBF )
                        {
#pragma unroll
                            for ( int m = 0; m < R; ++m )
                            {
                                const int64_t n  = j + m * NS;
                                const C       z0 = v[b][m], z1 = v[b + 1][m];
                                if constexpr ( KIND == KIND_C2R )
                                {
                                    using V2 = typename pair_of<T>::type;
                                    T *dst   = static_cast<T *>( p.out );
                                    if ( x.va )
                                        *reinterpret_cast<V2 *>( dst + x.oa + n ) = V2{ z0.x * p.scale, z1.x * p.scale };
                                    if ( x.vb )
                                        *reinterpret_cast<V2 *>( dst + x.ob + n ) = V2{ z0.y * p.scale, z1.y * p.scale };
                                }
                                else
                                {
                                    C *dst = static_cast<C *>( p.out );
                                    if ( x.va )
                                        *reinterpret_cast<f32x4 *>( dst + x.oa + n ) =
                                            f32x4{ z0.x * p.scale, z0.y * p.scale, z1.x * p.scale, z1.y * p.scale };
                                }
                            }
                        }
                    }
                    return;
                }
            }
#pragma unroll
            for ( int b = 0; b < NB; ++b )
            {
                const int j = jmap<RP.n - 1>( x.t, b );
                if ( NBF % TPC == 0 || j < NBF )
                {
                    const int d = ( j / NS ) * NS * R + ( j % NS );
#pragma unroll
                    for ( int m = 0; m < R; ++m )
                        store_line<KIND>( p, x.oa, x.ob, x.va, x.vb, d + m * NS, v[b][m] );
                }
            }
            return;
        }
        // stage the transformed lines through LDS, then store with the output mapping
        const int     tid = threadIdx.x;
        const int     c2  = p.store_b ? tid % TW : tid / TPC;
        const int     t2  = p.store_b ? tid / TW : tid % TPC;
        constexpr int LIM = KIND == KIND_R2C ? N / 2 + 1 : N;
        constexpr int NI  = ( LIM + TPC - 1 ) / TPC;
        C             zk[NI], zm[NI];
        if constexpr ( SPLIT )
        {
            put_stage<0, RP.n - 1, R, NS, NB>( v, x );
            __syncthreads();
            get_restage<0, NI>( zk, zm, x.lds, c2, t2 );
            __syncthreads();
            put_stage<1, RP.n - 1, R, NS, NB>( v, x );
            __syncthreads();
            get_restage<1, NI>( zk, zm, x.lds, c2, t2 );
        }
        else
        {
            put_stage<2, RP.n - 1, R, NS, NB>( v, x );
            __syncthreads();
            get_restage<2, NI>( zk, zm, x.lds, c2, t2 );
        }
        const int64_t col = ( x.tile * TW + c2 ) * CPT;
        const bool    va  = col < p.ncols;
        const bool    vb  = CPT == 2 && col + 1 < p.ncols;
        int64_t       ia, oa, ib = 0, ob = 0;
        col_offsets<MG>( p, col, ia, oa );
        if ( CPT == 2 )
            col_offsets<MG>( p, col + 1, ib, ob );
        if constexpr ( KIND == KIND_R2C )
        {
            // A[k] = (Z[k] + conj Z[N-k]) / 2 ;  B[k] = (Z[k] - conj Z[N-k]) / (2i)
            C ra[NI], rb[NI];
#pragma unroll
            for ( int i = 0; i < NI; ++i )
            {
                const int k = t2 + i * TPC;
                if ( LIM % TPC == 0 || k < LIM )
                {
                    // adjoint of C2R: interior bins of the half spectrum carry weight 2
                    const T h = ( p.wmode == 1 && k != 0 && 2 * k != N ) ? p.scale : (T)0.5 * p.scale;
                    const C a = zk[i], m = zm[i];
                    ra[i]     = mk<T>( h * ( a.x + m.x ), h * ( a.y - m.y ) );
                    rb[i]     = mk<T>( h * ( a.y + m.y ), h * ( m.x - a.x ) );
                }
            }
            if constexpr ( IOM & IO_H16_OUT )
            {
                __shared__ float red[2 * NT];
                float            ma = 0.f, mb = 0.f;
#pragma unroll
                for ( int i = 0; i < NI; ++i )
                    if ( LIM % TPC == 0 || t2 + i * TPC < LIM )
                    {
                        ma = fmaxf( ma, cmag( ra[i] ) );
                        mb = fmaxf( mb, cmag( rb[i] ) );
                    }
                const bool cf = p.store_b != 0;
                red[NT + tid] = mb;
                const float cma = column_max( red, ma, c2, cf );
                float       cmb = 0.f;
#pragma unroll 8
                for ( int t = 0; t < TPC; ++t )
                    cmb = fmaxf( cmb, red[NT + ( cf ? t * TW + c2 : c2 * TPC + t )] );
                const int ea = h16_exponent( cma ), eb = h16_exponent( cmb );
                int64_t   xi, ya, yb = 0;
                col_exponents<MG>( p, col, xi, ya );
                col_exponents<MG>( p, col + 1, xi, yb );
                HE *out = static_cast<HE *>( p.out );
#pragma unroll
                for ( int i = 0; i < NI; ++i )
                {
                    const int k = t2 + i * TPC;
                    if ( LIM % TPC == 0 || k < LIM )
                    {
                        if ( va )
                            out[oa + (int64_t)k * p.oes] = h16_pack<SPL>( ra[i].x, ra[i].y, ea );
                        if ( vb )
                            out[ob + (int64_t)k * p.oes] = h16_pack<SPL>( rb[i].x, rb[i].y, eb );
                    }
                }
                if ( t2 == 0 )
                {
                    if ( va )
                        p.xout[ya] = (int8_t)ea;
                    if ( vb )
                        p.xout[yb] = (int8_t)eb;
                }
            }
            else
            {
                cpx<T> *out = static_cast<cpx<T> *>( p.out );
#pragma unroll
                for ( int i = 0; i < NI; ++i )
                {
                    const int k = t2 + i * TPC;
                    if ( LIM % TPC == 0 || k < LIM )
                    {
                        if ( va )
                            out[oa + (int64_t)k * p.oes] = ra[i];
                        if ( vb )
                            out[ob + (int64_t)k * p.oes] = rb[i];
                    }
                }
            }
        }
        else if constexpr ( IOM & IO_H16_OUT )
        {
            // C2C restaged to the other mapping, FP16 output
            __shared__ float red[NT];
            float            lm = 0.f;
#pragma unroll
            for ( int i = 0; i < NI; ++i )
                if ( LIM % TPC == 0 || t2 + i * TPC < LIM )
                    lm = fmaxf( lm, cmag( zk[i] ) );
            const int e = h16_exponent( column_max( red, lm, c2, p.store_b != 0 ) * fabsf( (float)p.scale ) );
            int64_t   xi, ya;
            col_exponents<MG>( p, col, xi, ya );
            HE *out = static_cast<HE *>( p.out );
#pragma unroll
            for ( int i = 0; i < NI; ++i )
            {
                const int n = t2 + i * TPC;
                if ( ( LIM % TPC == 0 || n < LIM ) && va )
                    out[oa + (int64_t)n * p.oes] = h16_pack<SPL>( zk[i].x * p.scale, zk[i].y * p.scale, e );
            }
            if ( t2 == 0 && va )
                p.xout[ya] = (int8_t)e;
        }
        else
        {
            (void)zm;
#pragma unroll
            for ( int i = 0; i < NI; ++i )
            {
                const int n = t2 + i * TPC;
                if ( LIM % TPC == 0 || n < LIM )
                    store_line<KIND>( p, oa, ob, va, vb, n, zk[i] );
            }
        }
    }

    // ---- stage chain ----------------------------------------------------------------
    template <int S, int NB, int R>
    __device__ static __forceinline__ void chain( C ( &v )[NB][R], ctx &x )
    {
        constexpr int NS = rplan_span( RP, S );
        if constexpr ( S == RP.n - 1 )
        {
            final_store<R, NS, NB>( v, x );
        }
        else
        {
            constexpr int R2   = RP.r[S + 1];
            constexpr int NS2  = NS * R;
            constexpr int NBF2 = N / R2;
            constexpr int NB2  = ( NBF2 + TPC - 1 ) / TPC;
            constexpr int TST  = N / ( NS2 * R2 );
            // twiddles of the next stage are fetched before the exchange, so their
            // latency hides behind the LDS writes and the barrier (the split-exchange
            // variants keep registers for occupancy instead and fetch them after it)
            constexpr bool TWPF = !SPLIT;
            C              tws[NB2][R2];
            if constexpr ( TWL && S == 0 )
            {
                if constexpr ( TWR )
                {
                    if ( x.twpend )
                    {
#pragma unroll
                        for ( int i = 0; i < TWPT; ++i )
                        {
                            const int e = (int)threadIdx.x + i * NT;
                            if ( N % NT == 0 || e < N )
                                x.twb[e] = x.twr[i];
                        }
                        x.twpend = false;
                    }
                }
                __syncthreads(); // the LDS twiddle table is complete
            }
            if constexpr ( TWPF )
            {
#pragma unroll
                for ( int b = 0; b < NB2; ++b )
                {
                    const int j = jmap<S + 1>( x.t, b );
                    if ( NBF2 % TPC == 0 || j < NBF2 )
                    {
                        const int k = j % NS2;
#pragma unroll
                        for ( int r = 1; r < R2; ++r )
                            tws[b][r] = x.twl[r * k * TST];
                    }
                }
            }
            C w[NB2][R2];
            // the matrix-core radix-16 stage: every lane of the wave holds whole butterflies
            constexpr bool MF = RS == 2 && R2 == 16;
            static_assert( !MF || NBF2 % TPC == 0, "matrix-core stage needs whole butterflies on every lane" );
            mf_h8 am[4];
            if constexpr ( MF )
                dft16_mfma_operands<INV>( am, reinterpret_cast<const cpx<float> *>( x.p->tw ), N );
            if constexpr ( SPLIT )
            {
                put_stage<0, S, R, NS, NB>( v, x );
                __syncthreads();
                get_stage<0, S + 1, R2, NB2>( w, x );
                __syncthreads();
                put_stage<1, S, R, NS, NB>( v, x );
                __syncthreads();
                get_stage<1, S + 1, R2, NB2>( w, x );
            }
            else
            {
                put_stage<2, S, R, NS, NB>( v, x );
                __syncthreads();
                get_stage<2, S + 1, R2, NB2>( w, x );
            }
#pragma unroll
            for ( int b = 0; b < NB2; ++b )
            {
                const int j = jmap<S + 1>( x.t, b );
                if ( NBF2 % TPC == 0 || j < NBF2 )
                {
                    if constexpr ( !TWPF )
                    {
                        const int k = j % NS2;
#pragma unroll
                        for ( int r = 1; r < R2; ++r )
                            tws[b][r] = x.twl[r * k * TST];
                    }
#pragma unroll
                    for ( int r = 1; r < R2; ++r )
                        w[b][r] = twmul<INV>( w[b][r], tws[b][r] );
                    if constexpr ( MF )
                        dft16_mfma( reinterpret_cast<cpx<float>( & )[16]>( w[b] ), am );
                    else
                        dft_fixed<R2, INV>( w[b], x.p->tw, N );
                }
            }
            __syncthreads();
            chain<S + 1, NB2, R2>( w, x );
        }
    }

    // stage 0 from LDS (element n of column c at L(n, c), natural order), then the
    // remaining stages and the store; used by the inverse half of the filter pass
    __device__ static __forceinline__ void run_from_lds( ctx &x )
    {
        constexpr int R0   = RP.r[0];
        constexpr int NBF0 = N / R0;
        constexpr int NB0  = ( NBF0 + TPC - 1 ) / TPC;
        C             v[NB0][R0];
        get_stage<2, 0, R0, NB0>( v, x );
#pragma unroll
        for ( int b = 0; b < NB0; ++b )
        {
            const int j = jmap<0>( x.t, b );
            if ( NBF0 % TPC == 0 || j < NBF0 )
                dft_fixed<R0, INV>( v[b], x.p->tw, N );
        }
        __syncthreads();
        chain<0, NB0, R0>( v, x );
    }

    __device__ static __forceinline__ void run( const pass_params<T> &p, const filter_params<T> *f = nullptr )
    {
        __shared__ LW lds[LDS];
        ctx           x;
        x.p          = &p;
        x.f          = f;
        x.lds        = lds;
        const int tid = threadIdx.x;
        x.twpend = false;
        if constexpr ( TWL )
        {
            __shared__ C twb[N];
            x.twl = twb;
            x.twb = twb;
            if constexpr ( TWR )
            {
#pragma unroll
                for ( int i = 0; i < TWPT; ++i )
                {
                    const int e = tid + i * NT;
                    x.twr[i]    = p.tw[N % NT == 0 || e < N ? e : 0];
                }
                x.twpend = true;
            }
            else
            {
                for ( int i = tid; i < N; i += NT )
                    twb[i] = p.tw[i];
            }
        }
        else
            x.twl = p.tw;
        x.c          = LOADB ? tid % TW : tid / TPC;
        x.t          = LOADB ? tid / TW : tid % TPC;
        x.tile       = ordered_tile( p.vec );
        const int64_t col = ( x.tile * TW + x.c ) * CPT;
        x.va         = col < p.ncols;
        x.vb         = CPT == 2 && col + 1 < p.ncols;
        x.ib = x.ob = 0;
        col_offsets<MG>( p, col, x.ia, x.oa );
        if ( CPT == 2 )
            col_offsets<MG>( p, col + 1, x.ib, x.ob );
        x.xa = x.xb = x.ya = x.yb = 0;
        if constexpr ( IOM != IO_F32 )
        {
            col_exponents<MG>( p, col, x.xa, x.ya );
            if ( CPT == 2 )
                col_exponents<MG>( p, col + 1, x.xb, x.yb );
        }

        constexpr int R0   = RP.r[0];
        constexpr int NBF0 = N / R0;
        constexpr int NB0  = ( NBF0 + TPC - 1 ) / TPC;
        C             v[NB0][R0];
        bool          loaded = false;
        if constexpr ( STAGE_C2R )
        {
            // C2R along contiguous lines: the tile's 2 TW half spectra go through LDS with
            // coalesced loads, each bin read once (the Hermitian extension reads every
            // interior bin twice); stage 0 then builds z = A~ + i B~ from LDS
            constexpr int NH = N / 2 + 1;
            __shared__ int64_t loff[2 * TW], xoff[2 * TW];
            C                 *st = reinterpret_cast<C *>( lds );
            // 2 TW lines can exceed the thread count (TPC = 1 for N <= 16): strided fill
            for ( int l = tid; l < 2 * TW; l += NT )
            {
                const int64_t c  = x.tile * TW * 2 + l;
                int64_t       io = 0, oo = 0, xi = 0, yo = 0;
                col_offsets<MG>( p, c < p.ncols ? c : 0, io, oo );
                if constexpr ( IOM & IO_H16_IN )
                    col_exponents<MG>( p, c < p.ncols ? c : 0, xi, yo );
                loff[l] = c < p.ncols ? io : -1;
                xoff[l] = xi;
            }
            __syncthreads();
            // every load is issued before the first LDS write, unconditionally (an item
            // past the end or in a column past the end reads element 0, which exists) and
            // selected afterwards, so that no load waits on a branch
            constexpr int ITEMS = 2 * TW * NH;
            constexpr int PER   = ( ITEMS + NT - 1 ) / NT;
            C             zs[PER];
            bool          xrow = false;
            if constexpr ( IOM & IO_H16_IN )
            {
                // FP16 lines whose exponent index runs along the line (xsn = 1) and does not
                // depend on the fastest batch index (xs0 = 0: the C2R pass of the 3D layout,
                // bin (x, y, z) of y-line exponent z + x xs2).  When the tile's lines lie in
                // one batch row they share one row of NH exponents, loaded once per tile as
                // XD2 dwords into LDS (up to 7 bytes past the row: the work area pads its
                // exponent arrays), instead of one byte load per bin.  Same values, same
                // arithmetic.
                constexpr int       XD2 = ( NH + 3 ) / 4 + 1;
                __shared__ uint32_t xr2[XD2];
                const int64_t       c0 = x.tile * TW * 2;
                xrow = p.xsn == 1 && p.xs0 == 0 && ( reinterpret_cast<uintptr_t>( p.xin ) & 3 ) == 0 &&
                       c0 + 2 * TW <= p.ncols && c0 % p.d0 + 2 * TW <= p.d0;
                if ( xrow )
                {
                    const int64_t   xf = xoff[0];
                    const uint32_t *xw = reinterpret_cast<const uint32_t *>( p.xin );
                    uint32_t        er = 0;
                    if ( tid < XD2 )
                        er = xw[( xf >> 2 ) + tid];
                    const HE *srch = static_cast<const HE *>( p.in );
                    HE        raw[PER];
#pragma unroll
                    for ( int i = 0; i < PER; ++i )
                    {
                        const int e  = tid + i * NT;
                        const int ec = ITEMS % NT == 0 || e < ITEMS ? e : ITEMS - 1;
                        raw[i]       = srch[loff[ec / NH] + (int64_t)( ec % NH ) * p.ies];
                    }
                    if ( tid < XD2 )
                        xr2[tid] = er;
                    __syncthreads();
                    const uint8_t *xb = reinterpret_cast<const uint8_t *>( xr2 ) + ( xf & 3 );
#pragma unroll
                    for ( int i = 0; i < PER; ++i )
                    {
                        const int e = tid + i * NT;
                        if ( ITEMS % NT == 0 || e < ITEMS )
                            st[e] = h16_decode<SPL>( raw[i], (int8_t)xb[e % NH] ); // line * NH + k == e
                    }
                }
            }
            if ( !xrow )
            {
#pragma unroll
                for ( int i = 0; i < PER; ++i )
                {
                    const int     e    = tid + i * NT;
                    const bool    in   = ITEMS % NT == 0 || e < ITEMS;
                    const int     ec   = in ? e : ITEMS - 1;
                    const int     line = ec / NH, k = ec % NH;
                    const int64_t lo   = loff[line];
                    const bool    ok   = in && lo >= 0;
                    const int64_t src  = ok ? lo + (int64_t)k * p.ies : 0;
                    C             z;
                    if constexpr ( IOM & IO_H16_IN )
                        z = h16_load<SPL>( p.in, src, p.xin, ok ? xoff[line] + (int64_t)k * p.xsn : 0 );
                    else
                        z = static_cast<const C *>( p.in )[src];
                    zs[i] = ok ? z : mk<T>( 0, 0 );
                }
#pragma unroll
                for ( int i = 0; i < PER; ++i )
                {
                    const int e = tid + i * NT;
                    if ( ITEMS % NT == 0 || e < ITEMS )
                        st[e] = zs[i]; // line * NH + k == e
                }
            }
            __syncthreads();
#pragma unroll
            for ( int b = 0; b < NB0; ++b )
            {
                const int j = jmap<0>( x.t, b );
                if ( NBF0 % TPC == 0 || j < NBF0 )
                {
#pragma unroll
                    for ( int r = 0; r < R0; ++r )
                    {
                        const int n  = j + r * NBF0;
                        const int nn = n <= N / 2 ? n : N - n;
                        C         a  = st[( 2 * x.c ) * NH + nn];
                        C         bq = st[( 2 * x.c + 1 ) * NH + nn];
                        if ( n > N / 2 )
                        {
                            a.y  = -a.y;
                            bq.y = -bq.y;
                        }
                        if ( n == 0 || 2 * n == N )
                        {
                            a.y  = 0;
                            bq.y = 0;
                        }
                        else if ( p.wmode == 2 )
                        {
                            a  = scl( a, (T)0.5 );
                            bq = scl( bq, (T)0.5 );
                        }
                        v[b][r] = mk<T>( a.x - bq.y, a.y + bq.x );
                    }
                    dft_fixed<R0, INV>( v[b], p.tw, N );
                }
            }
            __syncthreads(); // the exchange buffer aliases the staging area
            loaded = true;
        }
        if constexpr ( PAIR_LOAD )
        {
            if ( p.vec & 1 )
            {
                // butterflies j, j + 1 read elements j + r NBF0 and j + 1 + r NBF0: one
                // 8-byte (two reals) or 16-byte (two FP32 complex) load per line and r
#pragma unroll
                for ( int b = 0; b < NB0; b += 2 )
                {
                    const int j = jmap<0>( x.t, b );
                    if ( NBF0 % TPC == 0 || j < NBF0 )
                    {
#pragma unroll
                        for ( int r = 0; r < R0; ++r )
                        {
                            const int64_t n = j + r * NBF0;
                            if constexpr ( KIND == KIND_R2C )
                            {
                                using V2      = typename pair_of<T>::type;
                                const T *src  = static_cast<const T *>( p.in );
                                const V2 a    = x.va ? *reinterpret_cast<const V2 *>( src + x.ia + n ) : V2{ 0, 0 };
                                const V2 bq   = x.vb ? *reinterpret_cast<const V2 *>( src + x.ib + n ) : V2{ 0, 0 };
                                v[b][r]       = mk<T>( a.x, bq.x );
                                v[b + 1][r]   = mk<T>( a.y, bq.y );
                            }
                            else
                            {
                                const C    *src = static_cast<const C *>( p.in );
                                const f32x4 q   = x.va ? *reinterpret_cast<const f32x4 *>( src + x.ia + n )
                                                       : f32x4{ 0, 0, 0, 0 };
                                v[b][r]         = mk<T>( q.x, q.y );
                                v[b + 1][r]     = mk<T>( q.z, q.w );
                            }
                        }
                        dft_fixed<R0, INV>( v[b], p.tw, N );
                        dft_fixed<R0, INV>( v[b + 1], p.tw, N );
                    }
                }
                loaded = true;
            }
        }
        if constexpr ( LOADB && KIND == KIND_C2C && ( IOM & IO_H16_IN ) )
        {
            // strided FP16 lines whose exponents do not depend on the fastest batch index
            // (the y pass of the 3D layout: element n of every column of one x plane has
            // the exponent of z-line (x, n)).  With d0 >= TW a tile's columns span at most
            // two such planes, so the tile stages their 2 N exponents in LDS, a few byte
            // loads per thread instead of one per element.  Same values, same arithmetic.
            if ( p.xs0 == 0 && p.d0 >= TW )
            {
                __shared__ int8_t xe[2 * N];
                const int64_t     cf = x.tile * TW < p.ncols ? x.tile * TW : p.ncols - 1;
                const int64_t     cl = cf + TW - 1 < p.ncols ? cf + TW - 1 : p.ncols - 1;
                int64_t           xf = 0, xl = 0, yd = 0;
                col_exponents<MG>( p, cf, xf, yd );
                col_exponents<MG>( p, cl, xl, yd );
                // the exponent loads are issued first and the data loads after them, all
                // unconditionally (indices past the end read element 0, which exists)
                constexpr int EP = ( 2 * N + NT - 1 ) / NT;
                int8_t        er[EP];
#pragma unroll
                for ( int i = 0; i < EP; ++i )
                {
                    const int e  = tid + i * NT;
                    const int ec = ( 2 * N ) % NT == 0 || e < 2 * N ? e : 0;
                    er[i]        = p.xin[( ec < N ? xf : xl ) + (int64_t)( ec % N ) * p.xsn];
                }
                // the raw stored bin waits in the value's registers (an FP16 pair in the real
                // component's, a hi + lo pair in both)
                const HE *src = static_cast<const HE *>( p.in );
#pragma unroll
                for ( int b = 0; b < NB0; ++b )
                {
                    const int  j  = jmap<0>( x.t, b );
                    const bool ok = x.va && ( NBF0 % TPC == 0 || j < NBF0 );
#pragma unroll
                    for ( int r = 0; r < R0; ++r )
                    {
                        const HE raw = src[ok ? x.ia + (int64_t)( j + r * NBF0 ) * p.ies : 0];
                        if constexpr ( SPL )
                            v[b][r] = __builtin_bit_cast( C, raw );
                        else
                            v[b][r].x = __builtin_bit_cast( float, raw );
                    }
                }
#pragma unroll
                for ( int i = 0; i < EP; ++i )
                {
                    const int e = tid + i * NT;
                    if ( ( 2 * N ) % NT == 0 || e < 2 * N )
                        xe[e] = er[i];
                }
                __syncthreads();
                const int g = x.xa == xf ? 0 : N;
#pragma unroll
                for ( int b = 0; b < NB0; ++b )
                {
                    const int j = jmap<0>( x.t, b );
                    if ( NBF0 % TPC == 0 || j < NBF0 )
                    {
#pragma unroll
                        for ( int r = 0; r < R0; ++r )
                        {
                            const int e = xe[g + j + r * NBF0];
                            HE        h;
                            if constexpr ( SPL )
                                h = __builtin_bit_cast( HE, v[b][r] );
                            else
                                h = __builtin_bit_cast( HE, v[b][r].x );
                            v[b][r] = x.va ? h16_decode<SPL>( h, e ) : mk<float>( 0, 0 );
                        }
                        dft_fixed<R0, INV>( v[b], p.tw, N );
                    }
                }
                loaded = true;
            }
        }
        if constexpr ( XROW )
        {
            // strided FP16 lines whose exponent index runs along the columns (xs0 = 1: the x
            // pass of the 3D layout, bin (x, z) of y-line exponent z + x xsn, and the inverse y
            // pass).  When the tile's TW columns lie in one batch row, the exponents of its
            // element n are TW consecutive bytes: the tile loads them as XND aligned dwords
            // per element index and keeps the TW bytes of each in LDS, instead of one byte
            // load per element.  The loads may read up to 4 bytes past a row's last
            // exponent (the work area pads its exponent arrays).  Same values, same
            // arithmetic.
            const int64_t cf = x.tile * TW;
            if ( !loaded && p.xs0 == 1 && ( reinterpret_cast<uintptr_t>( p.xin ) & 3 ) == 0 && cf + TW <= p.ncols &&
                 cf % p.d0 + TW <= p.d0 )
            {
                __shared__ uint32_t xr32[XLDS / 4];
                uint8_t            *xr = reinterpret_cast<uint8_t *>( xr32 );
                int64_t             xf = 0, yd = 0;
                col_exponents<MG>( p, cf, xf, yd );
                const uint32_t *xw = reinterpret_cast<const uint32_t *>( p.xin );
                // the exponent loads are issued first and the data loads after them, all
                // unconditionally (indices past the end read the tile's first dword)
                constexpr int ITEMS = N * XND;
                constexpr int EP    = ( ITEMS + NT - 1 ) / NT;
                uint32_t      er[EP];
#pragma unroll
                for ( int i = 0; i < EP; ++i )
                {
                    const int     e  = tid + i * NT;
                    const int     ec = ITEMS % NT == 0 || e < ITEMS ? e : 0;
                    const int64_t a  = xf + (int64_t)( ec / XND ) * p.xsn;
                    er[i]            = xw[( a >> 2 ) + ec % XND];
                }
                const HE *src = static_cast<const HE *>( p.in );
#pragma unroll
                for ( int b = 0; b < NB0; ++b )
                {
                    const int  j  = jmap<0>( x.t, b );
                    const bool ok = NBF0 % TPC == 0 || j < NBF0;
#pragma unroll
                    for ( int r = 0; r < R0; ++r )
                    {
                        const HE raw = src[ok ? x.ia + (int64_t)( j + r * NBF0 ) * p.ies : 0];
                        if constexpr ( SPL )
                            v[b][r] = __builtin_bit_cast( C, raw );
                        else
                            v[b][r].x = __builtin_bit_cast( float, raw );
                    }
                }
#pragma unroll
                for ( int i = 0; i < EP; ++i )
                {
                    const int e = tid + i * NT;
                    if ( ITEMS % NT == 0 || e < ITEMS )
                    {
                        if constexpr ( XDW )
                            xr32[e] = er[i];
                        else
                        {
                            // byte q of dword k of row n is exponent (a & ~3) + 4 k + q of the
                            // row, column 4 k + q - (a & 3) of the tile
                            const int n  = e / XND;
                            const int c0 = 4 * ( e % XND ) - (int)( ( xf + (int64_t)n * p.xsn ) & 3 );
#pragma unroll
                            for ( int q = 0; q < 4; ++q )
                                if ( c0 + q >= 0 && c0 + q < TW )
                                    xr[n * TW + c0 + q] = (uint8_t)( er[i] >> ( 8 * q ) );
                        }
                    }
                }
                __syncthreads();
#pragma unroll
                for ( int b = 0; b < NB0; ++b )
                {
                    const int j = jmap<0>( x.t, b );
                    if ( NBF0 % TPC == 0 || j < NBF0 )
                    {
#pragma unroll
                        for ( int r = 0; r < R0; ++r )
                        {
                            const int n = j + r * NBF0;
                            int       e;
                            if constexpr ( XDW )
                                e = (int8_t)xr[n * XND * 4 + (int)( ( xf + (int64_t)n * p.xsn ) & 3 ) + x.c];
                            else
                                e = (int8_t)xr[n * TW + x.c];
                            HE            h;
                            if constexpr ( SPL )
                                h = __builtin_bit_cast( HE, v[b][r] );
                            else
                                h = __builtin_bit_cast( HE, v[b][r].x );
                            v[b][r] = h16_decode<SPL>( h, e );
                        }
                        dft_fixed<R0, INV>( v[b], p.tw, N );
                    }
                }
                loaded = true;
            }
        }
        if ( !loaded )
        {
#pragma unroll
            for ( int b = 0; b < NB0; ++b )
            {
                const int j = jmap<0>( x.t, b );
                if ( NBF0 % TPC == 0 || j < NBF0 )
                {
#pragma unroll
                    for ( int r = 0; r < R0; ++r )
                    {
                        if constexpr ( IOM & IO_H16_IN )
                            v[b][r] = load_line_h16<KIND, SPL>( p, x.ia, x.ib, x.xa, x.xb, x.va, x.vb, j + r * NBF0, N );
                        else
                            v[b][r] = load_line<KIND>( p, x.ia, x.ib, x.va, x.vb, j + r * NBF0, N );
                    }
                    dft_fixed<R0, INV>( v[b], p.tw, N );
                }
            }
        }
        chain<0, NB0, R0>( v, x );
    }
};

template <typename T, int N, int KIND, bool INV, bool LOADB, int VAR, int IOM = IO_F32, int RS = 0>
__global__ void __launch_bounds__( ( pass_kernel<T, N, KIND, INV, LOADB, VAR, IOM, false, RS>::NT ) )
    __attribute__( ( amdgpu_waves_per_eu( pass_kernel<T, N, KIND, INV, LOADB, VAR, IOM, false, RS>::WAVES_PER_EU ) ) )
    k_pass( const pass_params<T> p )
{
    pass_kernel<T, N, KIND, INV, LOADB, VAR, IOM, false, RS>::run( p );
}

// fused filter pass: forward C2C along the lines, x H(k), inverse C2C, one HBM round trip
template <typename T, int N, bool LOADB, int VAR>
__global__ void __launch_bounds__( ( pass_kernel<T, N, KIND_C2C, false, LOADB, VAR, IO_F32, true>::NT ) )
    k_pass_filter( const pass_params<T> p, const filter_params<T> f )
{
    pass_kernel<T, N, KIND_C2C, false, LOADB, VAR, IO_F32, true>::run( p, &f );
}

// ---------------------------------------------------------------------------
// Generic kernel: runtime N and radix list (radices <= 16), ping-pong LDS,
// column-fastest mapping.  Used for lengths without a compiled specialisation.
// ---------------------------------------------------------------------------
constexpr int GEN_THREADS = 256;

template <typename T, int R, bool INV>
__device__ __forceinline__ void gen_stage( const cpx<T> *src, cpx<T> *dst, int N, int Ns, int tw_cols, int c,
                                           int t, int tpc, const cpx<T> *__restrict__ tw )
{
    const int nbf = N / R;
    for ( int j = t; j < nbf; j += tpc )
    {
        cpx<T> v[R];
        for ( int r = 0; r < R; ++r )
            v[r] = src[( j + r * nbf ) * tw_cols + c];
        const int k    = j % Ns;
        const int tstp = N / ( Ns * R );
        for ( int r = 1; r < R; ++r )
            v[r] = twmul<INV>( v[r], tw[r * k * tstp] );
        dft_fixed<R, INV>( v, tw, N );
        const int d = ( j / Ns ) * Ns * R + k;
        for ( int m = 0; m < R; ++m )
            dst[( d + m * Ns ) * tw_cols + c] = v[m];
    }
}

template <typename T, bool INV>
__device__ __forceinline__ void gen_stage_rt( int R, const cpx<T> *src, cpx<T> *dst, int N, int Ns, int tw_cols,
                                              int c, int t, int tpc, const cpx<T> *__restrict__ tw )
{
    switch ( R )
    {
    case 2: gen_stage<T, 2, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 3: gen_stage<T, 3, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 4: gen_stage<T, 4, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 5: gen_stage<T, 5, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 7: gen_stage<T, 7, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 8: gen_stage<T, 8, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 11: gen_stage<T, 11, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 13: gen_stage<T, 13, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    case 16: gen_stage<T, 16, INV>( src, dst, N, Ns, tw_cols, c, t, tpc, tw ); break;
    default: break; // host never plans other radices
    }
}

template <typename T, int KIND, bool INV>
__global__ void __launch_bounds__( GEN_THREADS )
    k_pass_generic( const pass_params<T> p, const radix_list rl, int N, int tw_cols )
{
    extern __shared__ unsigned char gen_lds_raw[];
    cpx<T>       *buf0 = reinterpret_cast<cpx<T> *>( gen_lds_raw );
    cpx<T>       *buf1 = buf0 + (int64_t)N * tw_cols;
    constexpr int CPT  = KIND == KIND_C2C ? 1 : 2;
    const int     tid  = threadIdx.x;
    const int     tpc  = GEN_THREADS / tw_cols;
    const int     c    = tid % tw_cols;
    const int     t    = tid / tw_cols;
    const int64_t tile = xcd_tile( blockIdx.x, gridDim.x );
    const int64_t col  = ( tile * tw_cols + c ) * CPT;
    const bool    va   = c < tw_cols && col < p.ncols;
    const bool    vb   = CPT == 2 && c < tw_cols && col + 1 < p.ncols;
    int64_t       ia = 0, oa = 0, ib = 0, ob = 0;
    col_offsets( p, col, ia, oa );
    if ( CPT == 2 )
        col_offsets( p, col + 1, ib, ob );
    if ( t < tpc )
        for ( int n = t; n < N; n += tpc )
            buf0[n * tw_cols + c] = load_line<KIND>( p, ia, ib, va, vb, n, N );
    __syncthreads();
    cpx<T> *src = buf0, *dst = buf1;
    int     Ns  = 1;
    for ( int s = 0; s < rl.n; ++s )
    {
        if ( t < tpc )
            gen_stage_rt<T, INV>( rl.r[s], src, dst, N, Ns, tw_cols, c, t, tpc, p.tw );
        __syncthreads();
        Ns *= rl.r[s];
        cpx<T> *tmp = src;
        src         = dst;
        dst         = tmp;
    }
    if ( t >= tpc )
        return;
    if constexpr ( KIND == KIND_R2C )
    {
        cpx<T> *out = static_cast<cpx<T> *>( p.out );
        for ( int k = t; k <= N / 2; k += tpc )
        {
            const T      h  = ( p.wmode == 1 && k != 0 && 2 * k != N ) ? p.scale : (T)0.5 * p.scale;
            const cpx<T> zk = src[k * tw_cols + c];
            const cpx<T> zm = src[( k == 0 ? 0 : N - k ) * tw_cols + c];
            if ( va )
                out[oa + (int64_t)k * p.oes] = mk<T>( h * ( zk.x + zm.x ), h * ( zk.y - zm.y ) );
            if ( vb )
                out[ob + (int64_t)k * p.oes] = mk<T>( h * ( zk.y + zm.y ), h * ( zm.x - zk.x ) );
        }
    }
    else
    {
        for ( int n = t; n < N; n += tpc )
            store_line<KIND>( p, oa, ob, va, vb, n, src[n * tw_cols + c] );
    }
}

} // namespace fftm_mi

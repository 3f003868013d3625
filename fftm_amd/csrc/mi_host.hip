// Host side of the MI355X FFT backend: descriptor validation, decomposition of a
// plan into HBM passes, pass execution, the fused ffts pipelines, the 4D
// transposes, and the C ABI declared in include/fftm_mi.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fftm_mi.h"
#include "mi_fill.hpp"
#include "mi_launch.hpp"
#include "mi_spectral.hpp"
#include "mi_transpose.hpp"

namespace fftm_mi
{
namespace
{

thread_local std::string g_err;

int fail( int code, const std::string &msg )
{
    g_err = msg;
    return code;
}

int ok()
{
    return FFTM_MI_SUCCESS;
}

enum : int
{
    BUF_IN   = 0,
    BUF_OUT  = 1,
    BUF_WORK = 2
};

struct bdim
{
    int64_t n, is, os;
    int64_t xs = 0, ys = 0; // exponent strides of FP16-stored input / output lines
};

// One HBM pass: a batch of 1D transforms along one axis.
struct pass_rt
{
    int        kind   = KIND_C2C;
    bool       inv    = false;
    int64_t    N      = 1;
    int        src    = BUF_IN;
    int        dst    = BUF_OUT;
    int64_t    ies    = 1, oes = 1;
    int64_t    ncols  = 0;
    int64_t    d0 = 1, d1 = 1, is0 = 0, is1 = 0, is2 = 0, os0 = 0, os1 = 0, os2 = 0;
    bool       loadb  = false, storeb = false;
    bool       generic = false;
    radix_list rl{};
    int        gen_cols = 1;
    int        wmode    = 0; // real-axis bin weights (adjoints), see pass_params::wmode
    int        var      = 0; // tile-shape variant of the fixed-length kernel
    bool       rev      = false; // tiles in reverse order (alternate passes of a transform pair)
    double     scale    = 1.0;
    // FP16-stored spectra (mixed precision): IO_H16_IN / IO_H16_OUT bits, exponent
    // strides (input: batch dims and transform axis; output: batch dims) and the byte
    // offsets of the exponent arrays in the work area
    bool       filter   = false; // fused filter pass (forward C2C, x H(k), inverse C2C)
    int        ax       = -1;    // logical axis transformed (ffts passes)
    int        iom      = 0;
    int64_t    xsn = 0, xs0 = 0, xs1 = 0, xs2 = 0, ys0 = 0, ys1 = 0, ys2 = 0;
    int64_t    xoff = 0, yoff = 0;
    // plane pass (the real axis A of length N and axis B of length NB together)
    bool       plane    = false;
    int64_t    NB       = 0;
    int64_t    ra = 0, rb = 0, ca = 0, cb = 0;      // real / complex strides along A and B
    int64_t    nplanes = 0, pd0 = 1, ri0 = 0, ri1 = 0, ci0 = 0, ci1 = 0;
};

bool is_pow2( int64_t n )
{
    return n >= 1 && ( n & ( n - 1 ) ) == 0;
}

// radices for the generic kernel (each <= 16); false if N has a prime factor > 13
bool generic_radices( int64_t N, radix_list &rl )
{
    rl.n      = 0;
    int64_t m = N;
    int     k = 0;
    while ( m % 2 == 0 )
    {
        m /= 2;
        ++k;
    }
    while ( k >= 4 )
    {
        rl.r[rl.n++] = 16;
        k -= 4;
    }
    if ( k == 3 )
        rl.r[rl.n++] = 8;
    else if ( k == 2 )
        rl.r[rl.n++] = 4;
    else if ( k == 1 )
        rl.r[rl.n++] = 2;
    const int primes[] = { 3, 5, 7, 11, 13 };
    for ( int p : primes )
        while ( m % p == 0 )
        {
            if ( rl.n >= 20 )
                return false;
            rl.r[rl.n++] = p;
            m /= p;
        }
    return m == 1;
}

int csize_of( int prec )
{
    return prec == FFTM_MI_F32 ? 8 : 16;
}

// Decide how a pass of length N runs (fixed kernel or generic); fills N-dependent fields.
int configure_kernel( pass_rt &p, int prec )
{
    if ( fixed_supported( p.N ) )
    {
        p.generic = false;
        return ok();
    }
    if ( !generic_radices( p.N, p.rl ) )
        return fail( FFTM_MI_NOT_SUPPORTED,
                     "fftm_mi: transform length " + std::to_string( p.N ) +
                         " has a prime factor above 13 (not supported yet)" );
    const int cs   = csize_of( prec );
    int       cols = prec == FFTM_MI_F32 ? 16 : 8;
    while ( cols > 1 && 2 * p.N * cols * cs > 65536 )
        cols /= 2;
    if ( 2 * p.N * cols * cs > 65536 )
        return fail( FFTM_MI_NOT_SUPPORTED,
                     "fftm_mi: transform length " + std::to_string( p.N ) + " exceeds the single-pass limit" );
    p.generic  = true;
    p.gen_cols = cols;
    return ok();
}

// Fill the batch decomposition of a pass from its batch dims.
int set_batch( pass_rt &p, std::vector<bdim> dims )
{
    std::vector<bdim> v;
    for ( const bdim &d : dims )
        if ( d.n > 1 )
            v.push_back( d );
    p.loadb  = p.ies != 1;
    p.storeb = p.oes != 1;
    const bool by_out = !p.loadb && p.storeb;
    std::stable_sort( v.begin(), v.end(), [by_out]( const bdim &a, const bdim &b ) {
        return by_out ? a.os < b.os : a.is < b.is;
    } );
    std::vector<bdim> m;
    for ( const bdim &d : v )
    {
        if ( !m.empty() && m.back().n * m.back().is == d.is && m.back().n * m.back().os == d.os &&
             m.back().n * m.back().xs == d.xs && m.back().n * m.back().ys == d.ys )
            m.back().n *= d.n;
        else
            m.push_back( d );
    }
    if ( m.size() > 3 )
        return fail( FFTM_MI_NOT_SUPPORTED, "fftm_mi: pass needs more than 3 batch dimensions" );
    while ( m.size() < 3 )
        m.push_back( bdim{ 1, 0, 0 } );
    p.xs0 = m[0].xs;
    p.xs1 = m[1].xs;
    p.xs2 = m[2].xs;
    p.ys0 = m[0].ys;
    p.ys1 = m[1].ys;
    p.ys2 = m[2].ys;
    p.ncols = m[0].n * m[1].n * m[2].n;
    p.d0    = m[0].n;
    p.d1    = m[1].n;
    p.is0   = m[0].is;
    p.is1   = m[1].is;
    p.is2   = m[2].is;
    p.os0   = m[0].os;
    p.os1   = m[1].os;
    p.os2   = m[2].os;
    return ok();
}

// ---------------------------------------------------------------------------
// twiddle tables (one per distinct length, owned by the plan / ffts handle)
// ---------------------------------------------------------------------------
struct tables
{
    int                         prec = 0;
    std::map<int64_t, void *>   by_n;

    ~tables() { release(); }

    void release()
    {
        for ( auto &kv : by_n )
            if ( kv.second )
                (void)hipFree( kv.second );
        by_n.clear();
    }

    int ensure( int64_t N )
    {
        if ( by_n.count( N ) )
            return ok();
        const size_t cs = (size_t)csize_of( prec );
        std::vector<unsigned char> host( cs * (size_t)N );
        const long double two_pi = 6.28318530717958647692528676655900577L;
        for ( int64_t k = 0; k < N; ++k )
        {
            const long double a = -two_pi * (long double)k / (long double)N;
            const long double c = cosl( a ), s = sinl( a );
            if ( prec == FFTM_MI_F32 )
            {
                float v[2] = { (float)c, (float)s };
                std::memcpy( host.data() + cs * (size_t)k, v, sizeof( v ) );
            }
            else
            {
                double v[2] = { (double)c, (double)s };
                std::memcpy( host.data() + cs * (size_t)k, v, sizeof( v ) );
            }
        }
        void      *d = nullptr;
        hipError_t e = hipMalloc( &d, host.size() );
        if ( e != hipSuccess )
            return fail( FFTM_MI_ALLOC_FAILED, std::string( "fftm_mi: hipMalloc(twiddles): " ) + hipGetErrorString( e ) );
        e = hipMemcpy( d, host.data(), host.size(), hipMemcpyHostToDevice );
        if ( e != hipSuccess )
        {
            (void)hipFree( d );
            return fail( FFTM_MI_ALLOC_FAILED, std::string( "fftm_mi: twiddle upload: " ) + hipGetErrorString( e ) );
        }
        by_n[N] = d;
        return ok();
    }
};

// ---------------------------------------------------------------------------
// pass execution
// ---------------------------------------------------------------------------
template <typename T>
hipError_t launch_plane_one( const pass_rt &r, const tables &tb, const void *src, void *dst, hipStream_t s )
{
    plane_params<T> p;
    p.in      = src;
    p.out     = dst;
    p.twa     = static_cast<const cpx<T> *>( tb.by_n.at( r.N ) );
    p.twb     = static_cast<const cpx<T> *>( tb.by_n.at( r.NB ) );
    p.ra      = r.ra;
    p.rb      = r.rb;
    p.ca      = r.ca;
    p.cb      = r.cb;
    p.nplanes = r.nplanes;
    p.d0      = r.pd0;
    p.ri0     = r.ri0;
    p.ri1     = r.ri1;
    p.ci0     = r.ci0;
    p.ci1     = r.ci1;
    p.wmode   = r.wmode;
    p.order   = variant_vec_order( r.var );
    p.scale   = (T)r.scale;
    const bool fwd = r.kind == KIND_R2C;
    // shape 1: the register-resident kernel (square planes); its real rows 2q, 2q+1 move
    // as one pair where B is contiguous, every row start is even and the array aligned
    const void *real = fwd ? src : static_cast<const void *>( dst );
    const bool  rx   = variant_shape_of( r.var ) == 1 && r.N == r.NB;
    p.pair           = ( r.rb == 1 && r.ra % 2 == 0 && ( r.pd0 <= 1 || r.ri0 % 2 == 0 ) &&
               ( r.nplanes <= r.pd0 || r.ri1 % 2 == 0 ) &&
               ( reinterpret_cast<uintptr_t>( real ) % ( 2 * sizeof( T ) ) ) == 0 )
                         ? 1
                         : 0;
    p.pad_ = 0;
    if constexpr ( std::is_same<T, float>::value )
        return rx ? launch_plane_rx_f32( p, (int)r.N, fwd, s ) : launch_plane_f32( p, (int)r.N, (int)r.NB, fwd, s );
    else
        return rx ? launch_plane_rx_f64( p, (int)r.N, fwd, s ) : launch_plane_f64( p, (int)r.N, (int)r.NB, fwd, s );
}

// Paired accesses (two consecutive elements per load / store) need unit element stride,
// a 16-byte aligned base and an even element offset for every line start.
bool even_line_offsets( const pass_rt &r, int64_t s0, int64_t s1, int64_t s2 )
{
    const int64_t d2 = ( r.ncols + r.d0 * r.d1 - 1 ) / ( r.d0 * r.d1 );
    return ( r.d0 <= 1 || s0 % 2 == 0 ) && ( r.d1 <= 1 || s1 % 2 == 0 ) && ( d2 <= 1 || s2 % 2 == 0 );
}

// Multiply-and-shift constants for dividing any n < 2^31 by d (1 <= d < 2^31): with
// l = ceil(log2 d) and m = ceil(2^(31 + l) / d), n / d = (n m) >> (31 + l) exactly
// (Granlund and Montgomery, "Division by invariant integers using multiplication", 1994,
// theorem 4.2: 2^(31 + l) <= m d <= 2^(31 + l) + 2^l).  m <= 2^32, so n m < 2^63.
void div_magic( int64_t d, uint64_t &m, int32_t &sh )
{
    int l = 0;
    while ( ( (int64_t)1 << l ) < d )
        ++l;
    sh = 31 + l;
    m  = ( ( (uint64_t)1 << sh ) + (uint64_t)d - 1 ) / (uint64_t)d;
}

int32_t pair_access_bits( const pass_rt &r, const void *src, const void *dst )
{
    int32_t bits = 0;
    if ( !r.loadb && r.ies == 1 && ( reinterpret_cast<uintptr_t>( src ) & 15 ) == 0 &&
         even_line_offsets( r, r.is0, r.is1, r.is2 ) )
        bits |= 1;
    if ( !r.storeb && r.oes == 1 && ( reinterpret_cast<uintptr_t>( dst ) & 15 ) == 0 &&
         even_line_offsets( r, r.os0, r.os1, r.os2 ) )
        bits |= 2;
    return bits;
}

template <typename T>
hipError_t launch_one( const pass_rt &r, const tables &tb, const void *src, void *dst, hipStream_t s,
                       void *work = nullptr, const void *filter = nullptr )
{
    if ( r.plane )
        return launch_plane_one<T>( r, tb, src, dst, s );
    pass_params<T> p{};
    p.in      = src;
    p.out     = dst;
    p.tw      = static_cast<const cpx<T> *>( tb.by_n.at( r.N ) );
    p.ies     = r.ies;
    p.oes     = r.oes;
    p.ncols   = r.ncols;
    p.d0      = r.d0;
    p.d1      = r.d1;
    p.is0     = r.is0;
    p.is1     = r.is1;
    p.is2     = r.is2;
    p.os0     = r.os0;
    p.os1     = r.os1;
    p.os2     = r.os2;
    // every column index a kernel forms stays below ncols plus one tile (at most 2048
    // columns), inside the constants' exact range n < 2^31
    if ( r.ncols < ( (int64_t)1 << 31 ) - 4096 && r.d0 >= 1 && r.d1 >= 1 )
    {
        div_magic( r.d0, p.mg0, p.sh0 );
        div_magic( r.d1, p.mg1, p.sh1 );
    }
    else if ( r.N >= 32 )
        return hipErrorInvalidValue; // those kernels compile no division fallback
    p.store_b = r.storeb ? 1 : 0;
    p.wmode   = r.wmode;
    p.vec     = pair_access_bits( r, src, dst );
    p.scale   = (T)r.scale;
    // tile shape (compile-time kernel), tile order (runtime bit) and radix schedule
    const int  shape = variant_shape_of( r.var );
    const int  rr    = variant_vec_order( r.var );
    const int  rs    = variant_rs_of( r.var );
    p.vec |= rr | ( r.rev ? VEC_REV : 0 );
    if constexpr ( std::is_same<T, float>::value )
    {
        if ( r.iom )
        {
            unsigned char *w = static_cast<unsigned char *>( work );
            p.xin            = ( r.iom & IO_H16_IN ) ? reinterpret_cast<const int8_t *>( w + r.xoff ) : nullptr;
            p.xout           = ( r.iom & IO_H16_OUT ) ? reinterpret_cast<int8_t *>( w + r.yoff ) : nullptr;
            p.xsn            = r.xsn;
            p.xs0            = r.xs0;
            p.xs1            = r.xs1;
            p.xs2            = r.xs2;
            p.ys0            = r.ys0;
            p.ys1            = r.ys1;
            p.ys2            = r.ys2;
            if ( rs )
                return launch_pass_rs_f32( p, (int)r.N, r.inv, shape, rs, r.iom, s );
            return launch_pass_mixed_f32( p, (int)r.N, r.kind, r.inv, r.loadb, shape, r.iom, s );
        }
        if ( rs )
            return launch_pass_rs_f32( p, (int)r.N, r.inv, shape, rs, IO_F32, s );
        if ( r.filter )
            return launch_filter_f32( p, *static_cast<const filter_params<float> *>( filter ), (int)r.N, r.loadb, shape, s );
        return r.generic ? launch_generic_f32( p, r.rl, (int)r.N, r.gen_cols, r.kind, r.inv, s )
                         : launch_pass_f32( p, (int)r.N, r.kind, r.inv, r.loadb, shape, s );
    }
    else
    {
        if ( r.filter )
            return launch_filter_f64( p, *static_cast<const filter_params<double> *>( filter ), (int)r.N, r.loadb, shape, s );
        return r.generic ? launch_generic_f64( p, r.rl, (int)r.N, r.gen_cols, r.kind, r.inv, s )
                         : launch_pass_f64( p, (int)r.N, r.kind, r.inv, r.loadb, shape, s );
    }
}

int run_passes( const std::vector<pass_rt> &passes, const tables &tb, int prec, const void *in, void *out,
                void *work, hipStream_t s, bool flip_c2c, bool sync, const char *what, int real_wmode = 0,
                const void *filter = nullptr )
{
    for ( const pass_rt &r0 : passes )
    {
        pass_rt r = r0;
        if ( flip_c2c && r.kind == KIND_C2C )
            r.inv = !r.inv;
        if ( r.kind != KIND_C2C )
            r.wmode = real_wmode;
        const void *src = r.src == BUF_IN ? in : ( r.src == BUF_OUT ? (const void *)out : (const void *)work );
        void       *dst = r.dst == BUF_OUT ? out : ( r.dst == BUF_WORK ? work : const_cast<void *>( in ) );
        hipError_t  e   = prec == FFTM_MI_F32 ? launch_one<float>( r, tb, src, dst, s, work, filter )
                                              : launch_one<double>( r, tb, src, dst, s, work, filter );
        if ( e != hipSuccess )
            return fail( FFTM_MI_EXEC_FAILED,
                         std::string( "fftm_mi: " ) + what + ": kernel launch failed: " + hipGetErrorString( e ) );
    }
    if ( sync )
    {
        const hipError_t e = hipDeviceSynchronize();
        if ( e != hipSuccess )
            return fail( FFTM_MI_EXEC_FAILED,
                         std::string( "fftm_mi: " ) + what + ": device synchronization failed: " + hipGetErrorString( e ) );
    }
    return ok();
}

// Default tile order of a new pass: chunks of 32 tiles per XCD (order 3) where the pass
// has enough work for every XCD to run several chunks, else one contiguous range per
// XCD (order 0).  Measured with the default shapes: 512^3 pair 1517 -> 1333 us, 256^3
// 177 -> 168 us, 128^4 2653 -> 2585 us; below ~2000 tiles (64^3, 128^3) order 0 is
// 2-6 % faster.  The tuner times every order regardless.
int default_order( const pass_rt &p )
{
    const bool big = p.plane ? p.nplanes >= 2048 : p.ncols >= 32768;
    return big ? 3 : 0;
}

// a variant number this pass can run: FP16-storage passes have tile shapes 0, 2, 6, 7; the
// alternative radix schedules exist for the strided FP32-arithmetic C2C passes of lengths
// 512 and 1024 in tile shapes 0 and 6 (schedule 1 also in shape 2)
bool pass_variant_ok( const pass_rt &r, int var, int prec )
{
    if ( !variant_number_ok( var ) )
        return false;
    if ( r.iom && !mixed_variant_ok( var ) )
        return false;
    const int rs = variant_rs_of( var );
    if ( ( r.iom & IO_SPLIT ) && ( variant_shape_of( var ) == 7 || rs != 0 ) )
        return false; // hi + lo storage: tile shapes 0, 2 and 6, default radix schedule
    return rs == 0 ||
           ( !r.plane && !r.generic && !r.filter && rs_supported( r.N, r.kind, r.loadb, prec, variant_shape_of( var ), rs ) );
}

int prepare_passes( std::vector<pass_rt> &passes, int prec, tables *tb )
{
    for ( pass_rt &p : passes )
        p.var = variant_shape_of( p.var ) + 8 * default_order( p );
    for ( pass_rt &p : passes )
    {
        if ( p.iom )
        {
            if ( !mixed_supported( p.N ) || ( ( p.iom & IO_SPLIT ) && p.N > 1024 ) )
                return fail( FFTM_MI_NOT_SUPPORTED, ( p.iom & IO_SPLIT )
                                                        ? "fftm_mi: hi + lo FP16-stored spectra need power-of-two "
                                                          "lengths 128..1024, got " + std::to_string( p.N )
                                                        : "fftm_mi: FP16-stored spectra need power-of-two lengths "
                                                          "128..2048, got " + std::to_string( p.N ) );
            p.generic = false;
        }
        else if ( !p.plane )
        {
            const int rc = configure_kernel( p, prec );
            if ( rc != FFTM_MI_SUCCESS )
                return rc;
        }
        if ( tb )
        {
            int rt = tb->ensure( p.N );
            if ( rt == FFTM_MI_SUCCESS && p.plane )
                rt = tb->ensure( p.NB );
            if ( rt != FFTM_MI_SUCCESS )
                return rt;
        }
    }
    return ok();
}

// ---------------------------------------------------------------------------
// plan decomposition (hipFFT advanced layout semantics)
// ---------------------------------------------------------------------------
struct plan_geom
{
    int     rank;
    int64_t n[3];
    int64_t ext_in[3], ext_out[3];
    int64_t emb_in[3], emb_out[3];
    int64_t istride, idist, ostride, odist, batch;
    int64_t pin[3], pout[3]; // element strides of each axis
};

int64_t span_of( int rank, const int64_t *ext, const int64_t *pitch, int64_t dist, int64_t batch )
{
    int64_t off = ( batch - 1 ) * dist;
    for ( int k = 0; k < rank; ++k )
        off += ( ext[k] - 1 ) * pitch[k];
    return off + 1;
}

int geom_from_desc( const fftm_mi_plan_desc &d, plan_geom &g )
{
    if ( d.rank < 1 || d.rank > 3 )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi: rank must be 1, 2 or 3" );
    if ( d.dir < FFTM_MI_R2C || d.dir > FFTM_MI_C2CB )
        return fail( FFTM_MI_INVALID_TYPE, "fftm_mi: unknown transform direction" );
    if ( d.precision != FFTM_MI_F32 && d.precision != FFTM_MI_F64 )
        return fail( FFTM_MI_INVALID_TYPE, "fftm_mi: unknown precision" );
    if ( d.batch < 1 )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi: batch must be >= 1" );
    g.rank  = d.rank;
    g.batch = d.batch;
    for ( int k = 0; k < d.rank; ++k )
    {
        if ( d.n[k] < 1 )
            return fail( FFTM_MI_INVALID_SIZE, "fftm_mi: transform sizes must be >= 1" );
        g.n[k]       = d.n[k];
        g.ext_in[k]  = d.n[k];
        g.ext_out[k] = d.n[k];
    }
    const int     L  = d.rank - 1;
    const int64_t nh = d.n[L] / 2 + 1;
    if ( d.dir == FFTM_MI_R2C )
        g.ext_out[L] = nh;
    if ( d.dir == FFTM_MI_C2R )
        g.ext_in[L] = nh;
    if ( d.default_layout )
    {
        g.istride = g.ostride = 1;
        g.idist = g.odist = 1;
        for ( int k = 0; k < d.rank; ++k )
        {
            g.emb_in[k]  = g.ext_in[k];
            g.emb_out[k] = g.ext_out[k];
            g.idist *= g.ext_in[k];
            g.odist *= g.ext_out[k];
        }
    }
    else
    {
        g.istride = d.istride;
        g.ostride = d.ostride;
        g.idist   = d.idist;
        g.odist   = d.odist;
        if ( g.istride < 1 || g.ostride < 1 )
            return fail( FFTM_MI_INVALID_VALUE, "fftm_mi: istride/ostride must be >= 1" );
        if ( d.batch > 1 && ( g.idist < 1 || g.odist < 1 ) )
            return fail( FFTM_MI_INVALID_VALUE, "fftm_mi: idist/odist must be >= 1 for batch > 1" );
        for ( int k = 0; k < d.rank; ++k )
        {
            g.emb_in[k]  = d.inembed[k];
            g.emb_out[k] = d.onembed[k];
            if ( k > 0 && ( g.emb_in[k] < g.ext_in[k] || g.emb_out[k] < g.ext_out[k] ) )
                return fail( FFTM_MI_INVALID_VALUE, "fftm_mi: inembed/onembed smaller than the transform extent" );
        }
    }
    for ( int k = 0; k < d.rank; ++k )
    {
        int64_t pi = g.istride, po = g.ostride;
        for ( int m = k + 1; m < d.rank; ++m )
        {
            pi *= g.emb_in[m];
            po *= g.emb_out[m];
        }
        g.pin[k]  = pi;
        g.pout[k] = po;
    }
    return ok();
}

// Build the pass list of one plan.  work_elems: complex elements of work needed.
int build_plan_passes( const fftm_mi_plan_desc &d, std::vector<pass_rt> &out, int64_t &work_elems )
{
    plan_geom g;
    int       rc = geom_from_desc( d, g );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    out.clear();
    work_elems     = 0;
    const int r    = g.rank;
    const int L    = r - 1;
    auto other_dims = [&]( int ax, const int64_t *ext, const int64_t *pi, const int64_t *po, int64_t di,
                           int64_t dout ) {
        std::vector<bdim> v;
        for ( int k = 0; k < r; ++k )
            if ( k != ax )
                v.push_back( bdim{ ext[k], pi[k], po[k] } );
        v.push_back( bdim{ g.batch, di, dout } );
        return v;
    };
    if ( d.dir == FFTM_MI_C2CF || d.dir == FFTM_MI_C2CB )
    {
        const bool inv   = d.dir == FFTM_MI_C2CB;
        bool       first = true;
        for ( int ax = L; ax >= 0; --ax )
        {
            pass_rt p;
            p.kind = KIND_C2C;
            p.inv  = inv;
            p.N    = g.n[ax];
            if ( first )
            {
                p.src = BUF_IN;
                p.dst = BUF_OUT;
                p.ies = g.pin[ax];
                p.oes = g.pout[ax];
                rc    = set_batch( p, other_dims( ax, g.n, g.pin, g.pout, g.idist, g.odist ) );
            }
            else
            {
                p.src = BUF_OUT;
                p.dst = BUF_OUT;
                p.ies = p.oes = g.pout[ax];
                rc            = set_batch( p, other_dims( ax, g.n, g.pout, g.pout, g.odist, g.odist ) );
            }
            if ( rc != FFTM_MI_SUCCESS )
                return rc;
            first = false;
            out.push_back( p );
        }
        return ok();
    }
    if ( d.dir == FFTM_MI_R2C )
    {
        pass_rt p;
        p.kind = KIND_R2C;
        p.N    = g.n[L];
        p.src  = BUF_IN;
        p.dst  = BUF_OUT;
        p.ies  = g.pin[L];
        p.oes  = g.pout[L];
        rc     = set_batch( p, other_dims( L, g.n, g.pin, g.pout, g.idist, g.odist ) );
        if ( rc != FFTM_MI_SUCCESS )
            return rc;
        out.push_back( p );
        for ( int ax = L - 1; ax >= 0; --ax )
        {
            pass_rt q;
            q.kind = KIND_C2C;
            q.N    = g.n[ax];
            q.src = q.dst = BUF_OUT;
            q.ies = q.oes = g.pout[ax];
            rc            = set_batch( q, other_dims( ax, g.ext_out, g.pout, g.pout, g.odist, g.odist ) );
            if ( rc != FFTM_MI_SUCCESS )
                return rc;
            out.push_back( q );
        }
        return ok();
    }
    // C2R: inverse C2C along axes 0..L-1 (into work, or in place on the input), then C2R along L
    const bool overwrite = ( d.flags & FFTM_MI_FLAG_OVERWRITE_INPUT ) != 0;
    const int  scratch   = overwrite ? BUF_IN : BUF_WORK;
    if ( L > 0 && !overwrite )
        work_elems = span_of( r, g.ext_in, g.pin, g.idist, g.batch );
    bool first = true;
    for ( int ax = 0; ax < L; ++ax )
    {
        pass_rt q;
        q.kind = KIND_C2C;
        q.inv  = true;
        q.N    = g.n[ax];
        q.src  = first ? BUF_IN : scratch;
        q.dst  = scratch;
        q.ies = q.oes = g.pin[ax];
        rc            = set_batch( q, other_dims( ax, g.ext_in, g.pin, g.pin, g.idist, g.idist ) );
        if ( rc != FFTM_MI_SUCCESS )
            return rc;
        first = false;
        out.push_back( q );
    }
    pass_rt p;
    p.kind = KIND_C2R;
    p.inv  = true;
    p.N    = g.n[L];
    p.src  = L > 0 ? scratch : BUF_IN;
    p.dst  = BUF_OUT;
    p.ies  = g.pin[L];
    p.oes  = g.pout[L];
    rc     = set_batch( p, other_dims( L, g.n, g.pin, g.pout, g.idist, g.odist ) );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    out.push_back( p );
    return ok();
}

// ---------------------------------------------------------------------------
// fused ffts pipelines
// ---------------------------------------------------------------------------
// A strided 3-D/4-D array description: extents and element strides per logical axis.
struct arr4
{
    int64_t n[4];
    int64_t s[4];
};

// pass along logical axis `ax` between two arrays of the same logical extents
// (for the real pass the complex side's extent of axis ax is N/2+1)
pass_rt axis_pass( int kind, bool inv, int ax, int dim, const int64_t *ext, const int64_t *si, const int64_t *so,
                   int src, int dst, int64_t N, int &rc )
{
    pass_rt p;
    p.kind = kind;
    p.inv  = inv;
    p.N    = N;
    p.src  = src;
    p.dst  = dst;
    p.ax   = ax;
    p.ies  = si[ax];
    p.oes  = so[ax];
    std::vector<bdim> v;
    for ( int k = 0; k < dim; ++k )
        if ( k != ax )
            v.push_back( bdim{ ext[k], si[k], so[k] } );
    rc = set_batch( p, v );
    return p;
}

// Exponent strides of the lines a pass writes (FP16-stored spectra): line ids are dense
// over the pass's batch axes, in the order of the output data strides.
void dense_line_strides( int ax, int dim, const int64_t *ext, const int64_t *so, int64_t *ys )
{
    int order[4], n = 0;
    for ( int k = 0; k < dim; ++k )
        if ( k != ax )
            order[n++] = k;
    std::sort( order, order + n, [&]( int a, int b ) { return so[a] < so[b]; } );
    int64_t acc = 1;
    for ( int i = 0; i < n; ++i )
    {
        ys[order[i]] = acc;
        acc *= ext[order[i]];
    }
    ys[ax] = 0;
}

// axis_pass with FP16 storage: xs (per logical axis) = exponent strides of the input
// lines or null for FP32 input; ys = exponent strides of the output lines or null
pass_rt axis_pass_h16( int kind, bool inv, int ax, int dim, const int64_t *ext, const int64_t *si, const int64_t *so,
                       int src, int dst, int64_t N, const int64_t *xs, const int64_t *ys, int &rc, bool split = false )
{
    pass_rt p;
    p.kind = kind;
    p.inv  = inv;
    p.N    = N;
    p.src  = src;
    p.dst  = dst;
    p.ies  = si[ax];
    p.oes  = so[ax];
    p.ax   = ax;
    p.iom  = ( xs ? IO_H16_IN : 0 ) | ( ys ? IO_H16_OUT : 0 ) | ( split ? IO_SPLIT : 0 );
    p.xsn  = xs ? xs[ax] : 0;
    std::vector<bdim> v;
    for ( int k = 0; k < dim; ++k )
        if ( k != ax )
            v.push_back( bdim{ ext[k], si[k], so[k], xs ? xs[k] : 0, ys ? ys[k] : 0 } );
    rc = set_batch( p, v );
    // strided lines: twice the columns per workgroup (FP16 rows of 128 bytes), split exchange
    // (hi + lo storage: 8-byte bins, the split exchange at the default tile)
    p.var = p.loadb ? ( split ? 6 : 7 ) : 0;
    return p;
}

struct ffts_layout
{
    int     dim;
    int64_t n[4];    // logical extents (x, y, z[, w]); the last axis is the real one
    int64_t sr[4];   // real array strides per logical axis
    int64_t sc[4];   // spectrum strides per logical axis
};

int make_ffts_layout( int32_t layout, const int64_t *nn, ffts_layout &L )
{
    if ( layout == FFTM_MI_LAYOUT_3D )
    {
        L.dim = 3;
        for ( int k = 0; k < 3; ++k )
            L.n[k] = nn[k];
        const int64_t nx = nn[0], ny = nn[1], nz = nn[2], nzh = nz / 2 + 1;
        (void)nx;
        L.sr[0] = ny * nz;
        L.sr[1] = nz;
        L.sr[2] = 1;
        L.sc[0] = ny * nzh;
        L.sc[1] = nzh;
        L.sc[2] = 1;
        return ok();
    }
    if ( layout < FFTM_MI_LAYOUT_4D_PENCIL || layout > FFTM_MI_LAYOUT_4D_XZWY )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi: unknown ffts layout" );
    L.dim = 4;
    for ( int k = 0; k < 4; ++k )
        L.n[k] = nn[k];
    const int64_t nx = nn[0], ny = nn[1], nz = nn[2], nw = nn[3], nwh = nw / 2 + 1;
    if ( layout == FFTM_MI_LAYOUT_4D_PENCIL )
    {
        // real [w][z][y][x] (ffts.hpp:126, arranger 0123)
        L.sr[0] = 1;
        L.sr[1] = nx;
        L.sr[2] = nx * ny;
        L.sr[3] = nx * ny * nz;
    }
    else
    {
        // real [x][y][z][w] (ffts.hpp:144, arranger 3210)
        L.sr[0] = ny * nz * nw;
        L.sr[1] = nz * nw;
        L.sr[2] = nw;
        L.sr[3] = 1;
    }
    if ( layout == FFTM_MI_LAYOUT_4D_SLAB )
    {
        // spectrum [x][y][z][wh] (ffts.hpp:149-150, arranger 2103 over (y,z,w,x))
        L.sc[0] = ny * nz * nwh;
        L.sc[1] = nz * nwh;
        L.sc[2] = nwh;
        L.sc[3] = 1;
    }
    else
    {
        // spectrum [y][z][wh][x] (ffts.hpp:133-134, arranger 3210 over (y,z,w,x);
        // fftm native xzwy, mpi_transpose_4d_same_xw.h:986-988)
        L.sc[0] = 1;
        L.sc[1] = nz * nwh * nx;
        L.sc[2] = nwh * nx;
        L.sc[3] = nx;
    }
    return ok();
}

// The axis paired with the real axis in a plane pass: the one that makes the plane
// contiguous on at least one side (3D: y; 4D slab: z; 4D pencil / xzwy: x).
int plane_partner( int32_t layout )
{
    switch ( layout )
    {
    case FFTM_MI_LAYOUT_3D: return 1;
    case FFTM_MI_LAYOUT_4D_SLAB: return 2;
    default: return 0;
    }
}

// Replace the real-axis pass and the C2C pass along `B` by one plane pass when the
// (B x A) half-spectrum plane fits in LDS and there are at least 64 planes.
void fuse_plane( const ffts_layout &L, int32_t layout, int prec, std::vector<pass_rt> &fwd, std::vector<pass_rt> &bwd )
{
    const int     d  = L.dim;
    const int     A  = d - 1;
    const int     B  = plane_partner( layout );
    const int64_t na = L.n[A], nb = L.n[B];
    if ( !plane_supported( na, nb, prec ) )
        return;
    int64_t nplanes = 1;
    int     other[2], no = 0;
    for ( int k = 0; k < d; ++k )
        if ( k != A && k != B )
        {
            nplanes *= L.n[k];
            other[no++] = k;
        }
    // Fewer planes than CUs still pay off: the fused pair has four passes instead of six,
    // and problems this small are bound by their launches (128^3 FP64 round trip 62 ->
    // 50 us, 64^3 30 -> 24 us).  Below 64 planes the split passes keep more workgroups.
    if ( nplanes < 64 )
        return;
    pass_rt p;
    p.plane   = true;
    p.N       = na;
    p.NB      = nb;
    p.ra      = L.sr[A];
    p.rb      = L.sr[B];
    // default shape: the register-resident kernel (1) from 32 points up, measured faster
    // or level in every layout; the tuner times both
    p.var     = na >= 32 ? 1 : 0;
    p.ca      = L.sc[A];
    p.cb      = L.sc[B];
    p.nplanes = nplanes;
    p.pd0     = no > 0 ? L.n[other[0]] : 1;
    p.ri0     = no > 0 ? L.sr[other[0]] : 0;
    p.ci0     = no > 0 ? L.sc[other[0]] : 0;
    p.ri1     = no > 1 ? L.sr[other[1]] : 0;
    p.ci1     = no > 1 ? L.sc[other[1]] : 0;
    // forward: drop the real pass (first) and the C2C pass along B, plane first
    std::vector<pass_rt> f;
    pass_rt              pf = p;
    pf.kind                 = KIND_R2C;
    pf.src                  = BUF_IN;
    pf.dst                  = BUF_OUT;
    f.push_back( pf );
    for ( size_t i = 1; i < fwd.size(); ++i )
        if ( !( fwd[i].kind == KIND_C2C && fwd[i].N == nb && fwd[i].ies == L.sc[B] ) )
            f.push_back( fwd[i] );
    // backward: drop the C2C pass along B and the final C2R, plane last
    std::vector<pass_rt> b;
    for ( size_t i = 0; i + 1 < bwd.size(); ++i )
        if ( !( bwd[i].kind == KIND_C2C && bwd[i].N == nb && bwd[i].ies == L.sc[B] ) )
            b.push_back( bwd[i] );
    pass_rt pb = p;
    pb.kind    = KIND_C2R;
    pb.inv     = true;
    pb.dst     = BUF_OUT;
    pb.src     = b.empty() ? BUF_IN : b.back().dst;
    // the first remaining pass reads the user input
    if ( !b.empty() )
        b.front().src = BUF_IN;
    b.push_back( pb );
    fwd.swap( f );
    bwd.swap( b );
}

int build_ffts_passes( const ffts_layout &L, bool overwrite, std::vector<pass_rt> &fwd, std::vector<pass_rt> &bwd,
                       int64_t &work_elems )
{
    const int d  = L.dim;
    const int R  = d - 1; // real axis (last logical)
    int64_t   extc[4];
    for ( int k = 0; k < d; ++k )
        extc[k] = L.n[k];
    extc[R] = L.n[R] / 2 + 1;
    int rc  = 0;
    fwd.clear();
    bwd.clear();
    // forward: real axis first (real -> spectrum layout), then C2C over axes R-1 .. 0 in place
    fwd.push_back( axis_pass( KIND_R2C, false, R, d, L.n, L.sr, L.sc, BUF_IN, BUF_OUT, L.n[R], rc ) );
    if ( rc )
        return rc;
    for ( int ax = R - 1; ax >= 0; --ax )
    {
        fwd.push_back( axis_pass( KIND_C2C, false, ax, d, extc, L.sc, L.sc, BUF_OUT, BUF_OUT, L.n[ax], rc ) );
        if ( rc )
            return rc;
    }
    // backward: C2C over axes 0 .. R-1 (first reads the input), then C2R into the real layout
    const int scratch = overwrite ? BUF_IN : BUF_WORK;
    work_elems        = overwrite ? 0 : extc[0] * extc[1] * extc[2] * ( d == 4 ? extc[3] : 1 );
    for ( int ax = 0; ax < R; ++ax )
    {
        bwd.push_back(
            axis_pass( KIND_C2C, true, ax, d, extc, L.sc, L.sc, ax == 0 ? BUF_IN : scratch, scratch, L.n[ax], rc ) );
        if ( rc )
            return rc;
    }
    bwd.push_back( axis_pass( KIND_C2R, true, R, d, L.n, L.sc, L.sr, R > 0 ? scratch : BUF_IN, BUF_OUT, L.n[R], rc ) );
    return rc;
}

// 3D transform with FP16-stored intermediate spectra: the real-axis pass writes the
// half spectrum in FP16 to the work area, the middle pass keeps it there, the last
// forward pass writes the FP32 spectrum; backward mirrors it.  Work area:
// [FP16 spectrum: 4 bytes per bin][exponents E0][exponents E1], 256-byte aligned.
int build_ffts_passes_h16( const ffts_layout &L, std::vector<pass_rt> &fwd, std::vector<pass_rt> &bwd,
                           size_t &work_bytes, bool split = false )
{
    if ( L.dim != 3 )
        return fail( FFTM_MI_NOT_SUPPORTED, "fftm_mi: FP16-stored spectra are implemented for the 3D layout" );
    const int64_t n0 = L.n[0], n1 = L.n[1], n2 = L.n[2], nzh = n2 / 2 + 1;
    const int64_t extc[3] = { n0, n1, nzh };
    auto          align   = []( size_t b ) { return ( b + 255 ) & ~(size_t)255; };
    const size_t  hbytes  = align( (size_t)( n0 * n1 * nzh ) * ( split ? 8 : 4 ) );
    // exponent arrays padded by 16 bytes: a strided pass reads a tile's exponents as whole
    // dwords, up to 4 bytes past the last one
    const size_t  e0bytes = align( (size_t)std::max( n0 * n1, n1 * nzh ) + 16 );
    const size_t  e1bytes = align( (size_t)( n0 * nzh ) + 16 );
    const int64_t E0 = (int64_t)hbytes, E1 = (int64_t)( hbytes + e0bytes );
    work_bytes       = hbytes + e0bytes + e1bytes;
    int64_t yP0[3], yP1[3], yQ0[3], yQ1[3];
    dense_line_strides( 2, 3, L.n, L.sc, yP0 );
    dense_line_strides( 1, 3, extc, L.sc, yP1 );
    dense_line_strides( 0, 3, extc, L.sc, yQ0 );
    dense_line_strides( 1, 3, extc, L.sc, yQ1 );
    int rc = 0;
    fwd.clear();
    bwd.clear();
    pass_rt p0 = axis_pass_h16( KIND_R2C, false, 2, 3, L.n, L.sr, L.sc, BUF_IN, BUF_WORK, n2, nullptr, yP0, rc, split );
    p0.yoff    = E0;
    if ( rc )
        return rc;
    pass_rt p1 = axis_pass_h16( KIND_C2C, false, 1, 3, extc, L.sc, L.sc, BUF_WORK, BUF_WORK, n1, yP0, yP1, rc, split );
    p1.xoff    = E0;
    p1.yoff    = E1;
    if ( rc )
        return rc;
    pass_rt p2 = axis_pass_h16( KIND_C2C, false, 0, 3, extc, L.sc, L.sc, BUF_WORK, BUF_OUT, n0, yP1, nullptr, rc, split );
    p2.xoff    = E1;
    if ( rc )
        return rc;
    pass_rt q0 = axis_pass_h16( KIND_C2C, true, 0, 3, extc, L.sc, L.sc, BUF_IN, BUF_WORK, n0, nullptr, yQ0, rc, split );
    q0.yoff    = E0;
    if ( rc )
        return rc;
    pass_rt q1 = axis_pass_h16( KIND_C2C, true, 1, 3, extc, L.sc, L.sc, BUF_WORK, BUF_WORK, n1, yQ0, yQ1, rc, split );
    q1.xoff    = E0;
    q1.yoff    = E1;
    if ( rc )
        return rc;
    pass_rt q2 = axis_pass_h16( KIND_C2R, true, 2, 3, L.n, L.sc, L.sr, BUF_WORK, BUF_OUT, n2, yQ1, nullptr, rc, split );
    q2.xoff    = E1;
    if ( rc )
        return rc;
    fwd = { p0, p1, p2 };
    bwd = { q0, q1, q2 };
    return ok();
}

} // namespace
} // namespace fftm_mi

using namespace fftm_mi;

struct fftm_mi_plan_s
{
    fftm_mi_plan_desc    desc{};
    std::vector<pass_rt> passes;
    tables               tb;
    size_t               work_size = 0;
    void                *work      = nullptr;
    hipStream_t          stream    = nullptr;
};

struct fftm_mi_ffts_s
{
    int32_t              layout = 0;
    int32_t              prec   = 0;     // arithmetic precision (FP32 for FFTM_MI_F32_H16)
    bool                 h16    = false; // FP16-stored intermediate spectra
    bool                 h16x2  = false; // ... as hi + lo FP16 pairs (8 bytes per bin)
    bool                 fwd_work = false; // forward passes use the work area
    int32_t              flags  = 0;
    int64_t              n[4]   = { 0, 0, 0, 0 };
    std::vector<pass_rt> fwd, bwd;
    tables               tb;
    size_t               work_size = 0;
    void                *work      = nullptr;
    hipStream_t          stream    = nullptr;
    spec_geom            sg{};
    double              *partials  = nullptr; // shell-gradient partial sums (SHELL_BLOCKS x MAX_SHELLS)
    double              *scratch   = nullptr; // this device's `scratch` (shell_gradient_noalloc)
    hipError_t           scratch_err = hipSuccess;
    // filter chain real -> real: forward passes into the work area, the fused filter pass
    // (or the forward passes, an elementwise multiply and the backward passes)
    std::vector<pass_rt> flt_pre, flt_post;
    bool                 flt_ready = false, flt_fused = false;
    int                  flt_line_ax = -1; // logical axis of the fused filter pass's lines
    int                  flt_var     = -1; // tuned variant of the fused filter pass (-1: from the forward pass)

    ~fftm_mi_ffts_s()
    {
        if ( partials )
            (void)hipFree( partials );
    }
};

namespace
{
constexpr int SHELL_BLOCKS = 512;
constexpr int MAX_SHELLS   = 1024;
} // namespace

// Scratch of fftm_mi_ffts_shell_gradient_noalloc: the shell partial sums (SHELL_BLOCKS x
// nshells doubles, 4 MiB at most), one copy per GPU for the life of the process.  Host
// code never uses the array's name as an address: it uses the per-device pointer that
// hipGetSymbolAddress returned for the device that was current at handle creation.
static __device__ __align__( 256 ) double scratch[SHELL_BLOCKS * MAX_SHELLS];

namespace
{
constexpr int MAX_DEVICES = 64;
std::mutex    g_scratch_mu;
void         *g_scratch_ptr[MAX_DEVICES] = {};

// The current device's address of `scratch` (looked up once per device), or nullptr
// with the HIP error in *err.
void *scratch_for_current_device( hipError_t *err )
{
    *err    = hipSuccess;
    int dev = 0;
    if ( ( *err = hipGetDevice( &dev ) ) != hipSuccess )
        return nullptr;
    if ( dev < 0 || dev >= MAX_DEVICES )
    {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk( g_scratch_mu );
    if ( !g_scratch_ptr[dev] )
    {
        void *ptr = nullptr;
        if ( ( *err = hipGetSymbolAddress( &ptr, HIP_SYMBOL( scratch ) ) ) != hipSuccess )
            return nullptr;
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

spec_geom make_spec_geom( const ffts_layout &L )
{
    spec_geom g{};
    g.nd        = L.dim;
    g.half_axis = L.dim - 1;
    int order[4] = { 0, 1, 2, 3 };
    std::sort( order, order + L.dim, [&]( int a, int b ) { return L.sc[a] > L.sc[b]; } );
    g.total     = 1;
    int64_t kmax2 = 0;
    for ( int m = 0; m < L.dim; ++m )
    {
        const int a = order[m];
        g.ax[m]     = a;
        g.ext[m]    = a == L.dim - 1 ? L.n[a] / 2 + 1 : L.n[a];
        g.total *= g.ext[m];
        g.n[a] = L.n[a];
        kmax2 += ( L.n[a] / 2 ) * ( L.n[a] / 2 );
    }
    g.kmax2   = kmax2 > 0 ? kmax2 : 1; // all-ones sizes: every bin has k = 0
    g.nshells = 1;
    if ( g.total < ( (int64_t)1 << 31 ) )
        for ( int m = 0; m < L.dim; ++m )
            div_magic( g.ext[m], g.mg[m], g.sh[m] );
    return g;
}

// grad[s] from the partial sums in `partials` (SHELL_BLOCKS x nshells doubles)
void launch_shell_gradient( fftm_mi_ffts_s *h, const void *grad_spectrum, const void *spectrum, void *grad,
                            int nshells, double *partials )
{
    spec_geom g = h->sg;
    g.nshells   = nshells;
    const size_t lds = sizeof( double ) * (size_t)nshells;
    if ( h->prec == FFTM_MI_F32 )
    {
        hipLaunchKernelGGL( k_shell_partials<float>, dim3( SHELL_BLOCKS ), dim3( 256 ), lds, h->stream, g,
                            static_cast<const cpx<float> *>( grad_spectrum ), static_cast<const cpx<float> *>( spectrum ),
                            partials );
        hipLaunchKernelGGL( k_shell_reduce<float>, dim3( ( nshells + 255 ) / 256 ), dim3( 256 ), 0, h->stream,
                            partials, SHELL_BLOCKS, nshells, static_cast<float *>( grad ) );
    }
    else
    {
        hipLaunchKernelGGL( k_shell_partials<double>, dim3( SHELL_BLOCKS ), dim3( 256 ), lds, h->stream, g,
                            static_cast<const cpx<double> *>( grad_spectrum ),
                            static_cast<const cpx<double> *>( spectrum ), partials );
        hipLaunchKernelGGL( k_shell_reduce<double>, dim3( ( nshells + 255 ) / 256 ), dim3( 256 ), 0, h->stream,
                            partials, SHELL_BLOCKS, nshells, static_cast<double *>( grad ) );
    }
}
} // namespace

extern "C" {

const char *fftm_mi_last_error( void )
{
    return g_err.c_str();
}

int fftm_mi_version( void )
{
    return 100; // 0.1.0
}

int fftm_mi_fill_random_real( int32_t precision, int32_t rank, const int64_t *n, const int64_t *strides,
                              const int64_t *start, uint64_t seed, void *dst, void *stream, int32_t sync )
{
    if ( !n || !strides || !dst )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_fill_random_real: null argument" );
    if ( rank != 3 && rank != 4 )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_fill_random_real: rank must be 3 or 4" );
    if ( precision != FFTM_MI_F32 && precision != FFTM_MI_F64 )
        return fail( FFTM_MI_INVALID_TYPE, "fftm_mi_fill_random_real: precision must be FFTM_MI_F32 or FFTM_MI_F64" );
    fill_params p{};
    p.nd    = rank;
    p.seed  = seed;
    p.total = 1;
    int order[4];
    for ( int a = 0; a < rank; ++a )
    {
        if ( n[a] < 1 || strides[a] < 1 )
            return fail( FFTM_MI_INVALID_SIZE, "fftm_mi_fill_random_real: extents and strides must be >= 1" );
        if ( start && start[a] < 0 )
            return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_fill_random_real: start offsets must be >= 0" );
        p.start[a] = start ? (uint64_t)start[a] : 0;
        p.total *= n[a];
        order[a] = a;
    }
    // memory order: largest stride first (ties keep the logical order)
    std::stable_sort( order, order + rank, [&]( int a, int b ) { return strides[a] > strides[b]; } );
    for ( int m = 0; m < rank; ++m )
    {
        p.ax[m]  = order[m];
        p.ext[m] = n[order[m]];
        p.str[m] = strides[order[m]];
    }
    hipStream_t   s      = static_cast<hipStream_t>( stream );
    const int64_t blocks = std::min<int64_t>( ( p.total + 255 ) / 256, 16384 );
    if ( precision == FFTM_MI_F32 )
        hipLaunchKernelGGL( k_fill_random<float>, dim3( (unsigned)blocks ), dim3( 256 ), 0, s, p,
                            static_cast<float *>( dst ) );
    else
        hipLaunchKernelGGL( k_fill_random<double>, dim3( (unsigned)blocks ), dim3( 256 ), 0, s, p,
                            static_cast<double *>( dst ) );
    hipError_t e = hipGetLastError();
    if ( e == hipSuccess && sync )
        e = hipStreamSynchronize( s );
    if ( e != hipSuccess )
        return fail( FFTM_MI_EXEC_FAILED, std::string( "fftm_mi_fill_random_real: " ) + hipGetErrorString( e ) );
    return ok();
}

int fftm_mi_plan_validate( const fftm_mi_plan_desc *desc, size_t *work_size )
{
    g_err.clear();
    if ( !desc )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_plan_validate: null descriptor" );
    std::vector<pass_rt> passes;
    int64_t              we = 0;
    int                  rc = build_plan_passes( *desc, passes, we );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    rc = prepare_passes( passes, desc->precision, nullptr );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    if ( work_size )
        *work_size = (size_t)we * (size_t)csize_of( desc->precision );
    return ok();
}

int fftm_mi_plan_create( const fftm_mi_plan_desc *desc, fftm_mi_plan *plan )
{
    g_err.clear();
    if ( !desc || !plan )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_plan_create: null argument" );
    *plan = nullptr;
    std::unique_ptr<fftm_mi_plan_s> p( new ( std::nothrow ) fftm_mi_plan_s() );
    if ( !p )
        return fail( FFTM_MI_ALLOC_FAILED, "fftm_mi_plan_create: out of host memory" );
    p->desc    = *desc;
    p->tb.prec = desc->precision;
    int64_t we = 0;
    int     rc = build_plan_passes( *desc, p->passes, we );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    rc = prepare_passes( p->passes, desc->precision, &p->tb );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    p->work_size = (size_t)we * (size_t)csize_of( desc->precision );
    *plan        = p.release();
    return ok();
}

int fftm_mi_plan_get_work_size( fftm_mi_plan plan, size_t *bytes )
{
    if ( !plan || !bytes )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_plan_get_work_size: null argument" );
    *bytes = plan->work_size;
    return ok();
}

int fftm_mi_plan_set_work_area( fftm_mi_plan plan, void *work_area )
{
    if ( !plan )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_plan_set_work_area: null plan" );
    plan->work = work_area;
    return ok();
}

int fftm_mi_plan_set_stream( fftm_mi_plan plan, void *stream )
{
    if ( !plan )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_plan_set_stream: null plan" );
    plan->stream = static_cast<hipStream_t>( stream );
    return ok();
}

int fftm_mi_plan_exec( fftm_mi_plan plan, int32_t dir, const void *in, void *out, int32_t sync )
{
    if ( !plan )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_plan_exec: null plan" );
    const int32_t pd   = plan->desc.dir;
    bool          flip = false;
    if ( dir != pd )
    {
        const bool c2c = ( pd == FFTM_MI_C2CF || pd == FFTM_MI_C2CB ) && ( dir == FFTM_MI_C2CF || dir == FFTM_MI_C2CB );
        if ( !c2c )
            return fail( FFTM_MI_INVALID_TYPE, "fftm_mi_plan_exec: incompatible execution direction" );
        flip = true;
    }
    if ( !in || !out )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_plan_exec: null data pointer" );
    if ( ( pd == FFTM_MI_R2C || pd == FFTM_MI_C2R ) && in == out )
        return fail( FFTM_MI_NOT_SUPPORTED, "fftm_mi_plan_exec: in-place R2C/C2R is not supported" );
    if ( plan->work_size > 0 && !plan->work )
        return fail( FFTM_MI_NO_WORKSPACE, "fftm_mi_plan_exec: plan needs a work area (set_work_area)" );
    return run_passes( plan->passes, plan->tb, plan->desc.precision, in, out, plan->work, plan->stream, flip,
                       sync != 0, "plan_exec" );
}

int fftm_mi_plan_get_descriptor( fftm_mi_plan plan, fftm_mi_plan_desc *desc )
{
    if ( !plan || !desc )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_plan_get_descriptor: null argument" );
    *desc = plan->desc;
    return ok();
}

int fftm_mi_plan_num_passes( fftm_mi_plan plan, int32_t *passes )
{
    if ( !plan || !passes )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_plan_num_passes: null argument" );
    *passes = (int32_t)plan->passes.size();
    return ok();
}

int fftm_mi_plan_destroy( fftm_mi_plan plan )
{
    if ( !plan )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_plan_destroy: null plan" );
    delete plan;
    return ok();
}

int fftm_mi_transpose_4d( int32_t elem_bytes, const int64_t src_dims[4], const int32_t perm[4], const void *src,
                          void *dst, void *stream, int32_t sync )
{
    if ( !src_dims || !perm || !src || !dst )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_transpose_4d: null argument" );
    if ( elem_bytes != 4 && elem_bytes != 8 && elem_bytes != 16 )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_transpose_4d: elem_bytes must be 4, 8 or 16" );
    int seen = 0;
    for ( int k = 0; k < 4; ++k )
    {
        if ( perm[k] < 0 || perm[k] > 3 || ( seen & ( 1 << perm[k] ) ) )
            return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_transpose_4d: perm is not a permutation of 0..3" );
        seen |= 1 << perm[k];
        if ( src_dims[k] < 1 )
            return fail( FFTM_MI_INVALID_SIZE, "fftm_mi_transpose_4d: dims must be >= 1" );
    }
    if ( src == dst )
        return fail( FFTM_MI_NOT_SUPPORTED, "fftm_mi_transpose_4d: in-place transpose is not supported" );
    tr_params p;
    p.src = src;
    p.dst = dst;
    int64_t dd[4], dstr[4];
    for ( int k = 0; k < 4; ++k )
    {
        p.d[k] = src_dims[k];
        dd[k]  = src_dims[perm[k]];
    }
    dstr[3] = 1;
    for ( int k = 2; k >= 0; --k )
        dstr[k] = dstr[k + 1] * dd[k + 1];
    for ( int k = 0; k < 4; ++k )
        p.ds[perm[k]] = dstr[k];
    hipStream_t s = static_cast<hipStream_t>( stream );
    if ( perm[3] == 3 )
    {
        const int64_t rows = p.d[0] * p.d[1] * p.d[2];
        const int64_t gx   = std::min<int64_t>( ( p.d[3] + 255 ) / 256, 1024 );
        const int64_t gy   = std::min<int64_t>( rows, 65535 );
        const int64_t gz   = ( rows + gy - 1 ) / gy;
        const dim3    grid( (unsigned)gx, (unsigned)gy, (unsigned)gz );
        if ( elem_bytes == 4 )
            hipLaunchKernelGGL( k_tr_rows<float>, grid, dim3( 256 ), 0, s, p );
        else if ( elem_bytes == 8 )
            hipLaunchKernelGGL( k_tr_rows<float2>, grid, dim3( 256 ), 0, s, p );
        else
            hipLaunchKernelGGL( k_tr_rows<double2>, grid, dim3( 256 ), 0, s, p );
    }
    else
    {
        const int A = perm[3];
        int       U = -1, V = -1;
        for ( int k = 0; k < 3; ++k )
            if ( k != A )
            {
                if ( U < 0 )
                    U = k;
                else
                    V = k;
            }
        const int64_t gz = p.d[U] * p.d[V];
        if ( gz > 65535 || ( p.d[A] + 31 ) / 32 > 65535 )
            return fail( FFTM_MI_NOT_SUPPORTED, "fftm_mi_transpose_4d: grid too large" );
        const dim3 grid( (unsigned)( ( p.d[3] + 31 ) / 32 ), (unsigned)( ( p.d[A] + 31 ) / 32 ), (unsigned)gz );
        if ( elem_bytes == 4 )
            hipLaunchKernelGGL( k_tr_tile<float>, grid, dim3( 256 ), 0, s, p, A, U, V );
        else if ( elem_bytes == 8 )
            hipLaunchKernelGGL( k_tr_tile<float2>, grid, dim3( 256 ), 0, s, p, A, U, V );
        else
            hipLaunchKernelGGL( k_tr_tile<double2>, grid, dim3( 256 ), 0, s, p, A, U, V );
    }
    hipError_t e = hipGetLastError();
    if ( e != hipSuccess )
        return fail( FFTM_MI_EXEC_FAILED, std::string( "fftm_mi_transpose_4d: launch failed: " ) + hipGetErrorString( e ) );
    if ( sync )
    {
        e = hipDeviceSynchronize();
        if ( e != hipSuccess )
            return fail( FFTM_MI_EXEC_FAILED,
                         std::string( "fftm_mi_transpose_4d: synchronization failed: " ) + hipGetErrorString( e ) );
    }
    return ok();
}

int fftm_mi_ffts_validate( int32_t layout, const int64_t *n, int32_t precision, int32_t flags, size_t *work_size )
{
    g_err.clear();
    if ( !n )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_validate: null sizes" );
    if ( precision != FFTM_MI_F32 && precision != FFTM_MI_F64 && precision != FFTM_MI_F32_H16 &&
         precision != FFTM_MI_F32_H16X2 )
        return fail( FFTM_MI_INVALID_TYPE, "fftm_mi_ffts_validate: unknown precision" );
    const int dim = layout == FFTM_MI_LAYOUT_3D ? 3 : 4;
    for ( int k = 0; k < dim; ++k )
        if ( n[k] < 1 )
            return fail( FFTM_MI_INVALID_SIZE, "fftm_mi_ffts_validate: sizes must be >= 1" );
    ffts_layout L;
    int         rc = make_ffts_layout( layout, n, L );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    if ( precision == FFTM_MI_F32_H16 || precision == FFTM_MI_F32_H16X2 )
    {
        std::vector<pass_rt> f, b;
        size_t               wb = 0;
        rc                      = build_ffts_passes_h16( L, f, b, wb, precision == FFTM_MI_F32_H16X2 );
        if ( rc == FFTM_MI_SUCCESS )
            rc = prepare_passes( f, FFTM_MI_F32, nullptr );
        if ( rc == FFTM_MI_SUCCESS )
            rc = prepare_passes( b, FFTM_MI_F32, nullptr );
        if ( rc == FFTM_MI_SUCCESS && work_size )
            *work_size = wb;
        return rc;
    }
    std::vector<pass_rt> f, b;
    int64_t              we = 0;
    rc = build_ffts_passes( L, ( flags & FFTM_MI_FLAG_OVERWRITE_INPUT ) != 0, f, b, we );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    if ( !( flags & FFTM_MI_FLAG_NO_PLANE_FUSION ) )
        fuse_plane( L, layout, precision, f, b );
    rc = prepare_passes( f, precision, nullptr );
    if ( rc == FFTM_MI_SUCCESS )
        rc = prepare_passes( b, precision, nullptr );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    if ( work_size )
        *work_size = (size_t)we * (size_t)csize_of( precision );
    return ok();
}

int fftm_mi_ffts_create( int32_t layout, const int64_t *n, int32_t precision, int32_t flags, fftm_mi_ffts *h )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_create: null handle" );
    *h        = nullptr;
    size_t ws = 0;
    int    rc = fftm_mi_ffts_validate( layout, n, precision, flags, &ws );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    std::unique_ptr<fftm_mi_ffts_s> p( new ( std::nothrow ) fftm_mi_ffts_s() );
    if ( !p )
        return fail( FFTM_MI_ALLOC_FAILED, "fftm_mi_ffts_create: out of host memory" );
    p->layout  = layout;
    p->h16     = precision == FFTM_MI_F32_H16 || precision == FFTM_MI_F32_H16X2;
    p->h16x2   = precision == FFTM_MI_F32_H16X2;
    p->prec    = p->h16 ? FFTM_MI_F32 : precision;
    p->flags   = flags;
    p->tb.prec = p->prec;
    const int dim = layout == FFTM_MI_LAYOUT_3D ? 3 : 4;
    for ( int k = 0; k < dim; ++k )
        p->n[k] = n[k];
    ffts_layout L;
    make_ffts_layout( layout, n, L );
    if ( p->h16 )
    {
        size_t wb = 0;
        build_ffts_passes_h16( L, p->fwd, p->bwd, wb, p->h16x2 );
        p->fwd_work = true;
    }
    else
    {
        int64_t we = 0;
        build_ffts_passes( L, ( flags & FFTM_MI_FLAG_OVERWRITE_INPUT ) != 0, p->fwd, p->bwd, we );
        if ( !( flags & FFTM_MI_FLAG_NO_PLANE_FUSION ) )
            fuse_plane( L, layout, precision, p->fwd, p->bwd );
    }
    rc = prepare_passes( p->fwd, p->prec, &p->tb );
    if ( rc == FFTM_MI_SUCCESS )
        rc = prepare_passes( p->bwd, p->prec, &p->tb );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    // Every second pass of the forward + backward sequence runs its tiles last-first.  Each
    // pass walks its columns in increasing memory order and writes its output in that
    // order, so the reversed pass after it starts on the most recently written data, which
    // the memory-side cache (256 MB on MI355X) still holds.  Results are unchanged.
    for ( size_t i = 0; i < p->fwd.size(); ++i )
        p->fwd[i].rev = ( i & 1 ) != 0;
    for ( size_t i = 0; i < p->bwd.size(); ++i )
        p->bwd[i].rev = ( ( p->fwd.size() + i ) & 1 ) != 0;
    p->work_size = ws;
    p->sg        = make_spec_geom( L );
    p->scratch   = static_cast<double *>( scratch_for_current_device( &p->scratch_err ) );
    *h           = p.release();
    if ( const char *cache = std::getenv( "FFTM_MI_AUTOTUNE_CACHE" ) )
        if ( *cache )
        {
            int32_t applied = 0;
            (void)fftm_mi_ffts_load_tuning( *h, cache, &applied );
            g_err.clear();
        }
    return ok();
}

int fftm_mi_ffts_get_work_size( fftm_mi_ffts h, size_t *bytes )
{
    if ( !h || !bytes )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_get_work_size: null argument" );
    *bytes = h->work_size;
    return ok();
}

int fftm_mi_ffts_set_work_area( fftm_mi_ffts h, void *work_area )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_set_work_area: null handle" );
    h->work = work_area;
    return ok();
}

int fftm_mi_ffts_set_stream( fftm_mi_ffts h, void *stream )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_set_stream: null handle" );
    h->stream = static_cast<hipStream_t>( stream );
    return ok();
}

int fftm_mi_ffts_forward( fftm_mi_ffts h, const void *in, void *out, int32_t sync )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_forward: null handle" );
    if ( !in || !out )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_forward: null data pointer" );
    if ( in == out )
        return fail( FFTM_MI_NOT_SUPPORTED, "fftm_mi_ffts_forward: in-place transform is not supported" );
    if ( h->fwd_work && !h->work )
        return fail( FFTM_MI_NO_WORKSPACE, "fftm_mi_ffts_forward: a work area is required (set_work_area)" );
    return run_passes( h->fwd, h->tb, h->prec, in, out, h->work, h->stream, false, sync != 0, "ffts_forward" );
}

int fftm_mi_ffts_backward( fftm_mi_ffts h, const void *in, void *out, int32_t sync )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_backward: null handle" );
    if ( !in || !out )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_backward: null data pointer" );
    if ( in == out )
        return fail( FFTM_MI_NOT_SUPPORTED, "fftm_mi_ffts_backward: in-place transform is not supported" );
    if ( h->work_size > 0 && !h->work )
        return fail( FFTM_MI_NO_WORKSPACE, "fftm_mi_ffts_backward: a work area is required (set_work_area)" );
    return run_passes( h->bwd, h->tb, h->prec, in, out, h->work, h->stream, false, sync != 0, "ffts_backward" );
}

int fftm_mi_ffts_num_passes( fftm_mi_ffts h, int32_t *forward_passes, int32_t *backward_passes )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_num_passes: null handle" );
    if ( forward_passes )
        *forward_passes = (int32_t)h->fwd.size();
    if ( backward_passes )
        *backward_passes = (int32_t)h->bwd.size();
    return ok();
}

int fftm_mi_ffts_run_pass( fftm_mi_ffts h, int32_t backward, int32_t index, const void *in, void *out, int32_t sync )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_run_pass: null handle" );
    const std::vector<pass_rt> &v = backward ? h->bwd : h->fwd;
    if ( index < 0 || index >= (int32_t)v.size() )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_run_pass: pass index out of range" );
    if ( !in || !out )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_run_pass: null data pointer" );
    if ( ( backward || h->fwd_work ) && h->work_size > 0 && !h->work )
        return fail( FFTM_MI_NO_WORKSPACE, "fftm_mi_ffts_run_pass: a work area is required (set_work_area)" );
    std::vector<pass_rt> one( 1, v[(size_t)index] );
    return run_passes( one, h->tb, h->prec, in, out, h->work, h->stream, false, sync != 0, "ffts_run_pass" );
}

int fftm_mi_ffts_pass_bytes( fftm_mi_ffts h, int32_t backward, int32_t index, int64_t *bytes )
{
    if ( !h || !bytes )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_pass_bytes: null argument" );
    const std::vector<pass_rt> &v = backward ? h->bwd : h->fwd;
    if ( index < 0 || index >= (int32_t)v.size() )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_pass_bytes: pass index out of range" );
    const pass_rt &p  = v[(size_t)index];
    const int64_t  cs = csize_of( h->prec ), rs = cs / 2;
    const int64_t  nh = p.N / 2 + 1;
    if ( p.iom )
    {
        // FP16 side: 4 bytes per bin; exponents: one byte per stored line
        const int64_t lines = p.kind == KIND_C2C ? p.ncols : p.ncols;
        const int64_t hs = ( p.iom & IO_SPLIT ) ? 8 : 4; // bytes per FP16-stored bin
        const int64_t ic = ( p.iom & IO_H16_IN ) ? hs : cs, oc = ( p.iom & IO_H16_OUT ) ? hs : cs;
        const int64_t ex = ( ( p.iom & IO_H16_IN ) ? lines : 0 ) + ( ( p.iom & IO_H16_OUT ) ? lines : 0 );
        if ( p.kind == KIND_C2C )
            *bytes = p.ncols * p.N * ( ic + oc ) + ex;
        else if ( p.kind == KIND_R2C )
            *bytes = p.ncols * ( p.N * rs + nh * oc ) + ex;
        else
            *bytes = p.ncols * ( nh * ic + p.N * rs ) + ex;
        return ok();
    }
    if ( p.plane )
        *bytes = p.nplanes * p.NB * ( p.N * rs + nh * cs );
    else if ( p.kind == KIND_C2C )
        *bytes = 2 * p.ncols * p.N * cs;
    else
        *bytes = p.ncols * ( p.N * rs + nh * cs );
    return ok();
}

int fftm_mi_ffts_r2c_adjoint( fftm_mi_ffts h, const void *spectrum_in, void *real_out, int32_t sync )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_r2c_adjoint: null handle" );
    if ( !spectrum_in || !real_out || spectrum_in == real_out )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_r2c_adjoint: bad data pointers" );
    if ( h->work_size > 0 && !h->work )
        return fail( FFTM_MI_NO_WORKSPACE, "fftm_mi_ffts_r2c_adjoint: a work area is required (set_work_area)" );
    return run_passes( h->bwd, h->tb, h->prec, spectrum_in, real_out, h->work, h->stream, false, sync != 0,
                       "ffts_r2c_adjoint", 2 );
}

int fftm_mi_ffts_c2r_adjoint( fftm_mi_ffts h, const void *real_in, void *spectrum_out, int32_t sync )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_c2r_adjoint: null handle" );
    if ( !real_in || !spectrum_out || real_in == spectrum_out )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_c2r_adjoint: bad data pointers" );
    if ( h->fwd_work && !h->work )
        return fail( FFTM_MI_NO_WORKSPACE, "fftm_mi_ffts_c2r_adjoint: a work area is required (set_work_area)" );
    return run_passes( h->fwd, h->tb, h->prec, real_in, spectrum_out, h->work, h->stream, false, sync != 0,
                       "ffts_c2r_adjoint", 1 );
}

static int spectral_launch_check( fftm_mi_ffts h, bool sync, const char *what )
{
    hipError_t e = hipGetLastError();
    if ( e != hipSuccess )
        return fail( FFTM_MI_EXEC_FAILED, std::string( what ) + ": launch failed: " + hipGetErrorString( e ) );
    if ( sync )
    {
        e = hipDeviceSynchronize();
        if ( e != hipSuccess )
            return fail( FFTM_MI_EXEC_FAILED, std::string( what ) + ": synchronization failed: " + hipGetErrorString( e ) );
    }
    (void)h;
    return ok();
}

int fftm_mi_ffts_spectral_multiply( fftm_mi_ffts h, const void *in, void *out, const void *filter,
                                    const void *shell_coeffs, int32_t nshells, int32_t sync )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_spectral_multiply: null handle" );
    if ( !in || !out )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_spectral_multiply: null data pointer" );
    if ( ( filter == nullptr ) == ( shell_coeffs == nullptr ) )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_spectral_multiply: pass exactly one of filter / shell_coeffs" );
    if ( shell_coeffs && ( nshells < 1 || nshells > MAX_SHELLS ) )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_spectral_multiply: nshells must be in 1..1024" );
    spec_geom g = h->sg;
    g.nshells   = shell_coeffs ? nshells : 1;
    const int64_t blocks = std::min<int64_t>( ( g.total + 255 ) / 256, 8192 );
    if ( h->prec == FFTM_MI_F32 )
        hipLaunchKernelGGL( k_spectral_multiply<float>, dim3( (unsigned)blocks ), dim3( 256 ), 0, h->stream, g,
                            static_cast<const cpx<float> *>( in ), static_cast<cpx<float> *>( out ),
                            static_cast<const float *>( filter ), static_cast<const float *>( shell_coeffs ) );
    else
        hipLaunchKernelGGL( k_spectral_multiply<double>, dim3( (unsigned)blocks ), dim3( 256 ), 0, h->stream, g,
                            static_cast<const cpx<double> *>( in ), static_cast<cpx<double> *>( out ),
                            static_cast<const double *>( filter ), static_cast<const double *>( shell_coeffs ) );
    return spectral_launch_check( h, sync != 0, "fftm_mi_ffts_spectral_multiply" );
}

int fftm_mi_ffts_shell_gradient( fftm_mi_ffts h, const void *grad_spectrum, const void *spectrum, void *grad,
                                 int32_t nshells, int32_t sync )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_shell_gradient: null handle" );
    if ( !grad_spectrum || !spectrum || !grad )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_shell_gradient: null data pointer" );
    if ( nshells < 1 || nshells > MAX_SHELLS )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_shell_gradient: nshells must be in 1..1024" );
    if ( !h->partials )
    {
        const hipError_t e = hipMalloc( &h->partials, sizeof( double ) * SHELL_BLOCKS * MAX_SHELLS );
        if ( e != hipSuccess )
        {
            h->partials = nullptr;
            return fail( FFTM_MI_ALLOC_FAILED, std::string( "fftm_mi_ffts_shell_gradient: " ) + hipGetErrorString( e ) );
        }
    }
    launch_shell_gradient( h, grad_spectrum, spectrum, grad, nshells, h->partials );
    return spectral_launch_check( h, sync != 0, "fftm_mi_ffts_shell_gradient" );
}

int fftm_mi_ffts_shell_gradient_noalloc( fftm_mi_ffts h, const void *grad_spectrum, const void *spectrum, void *grad,
                                         int32_t nshells, int32_t sync )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_shell_gradient_noalloc: null handle" );
    if ( !grad_spectrum || !spectrum || !grad )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_shell_gradient_noalloc: null data pointer" );
    if ( nshells < 1 || nshells > MAX_SHELLS )
        return fail( FFTM_MI_INVALID_VALUE,
                     "fftm_mi_ffts_shell_gradient_noalloc: nshells must be in 1..1024 (the static scratch holds "
                     "512 x 1024 partial sums)" );
    if ( !h->scratch )
        return fail( FFTM_MI_SETUP_FAILED, std::string( "fftm_mi_ffts_shell_gradient_noalloc: no device address for "
                                                        "the static scratch: " ) +
                                               hipGetErrorString( h->scratch_err ) );
    // every partial sum the reduction reads is written by the partials kernel of this
    // call (SHELL_BLOCKS x nshells), so the scratch needs no clearing
    launch_shell_gradient( h, grad_spectrum, spectrum, grad, nshells, h->scratch );
    return spectral_launch_check( h, sync != 0, "fftm_mi_ffts_shell_gradient_noalloc" );
}

} // extern "C"

namespace
{

int check_op( const fftm_mi_ffts_s &h, const fftm_mi_spectral_op *op, const char *what )
{
    if ( !op )
        return fail( FFTM_MI_INVALID_VALUE, std::string( what ) + ": null operator" );
    if ( op->kind < FFTM_MI_OP_POISSON || op->kind > FFTM_MI_OP_BANDPASS )
        return fail( FFTM_MI_INVALID_VALUE, std::string( what ) + ": unknown operator kind" );
    if ( ( op->kind == FFTM_MI_OP_DERIVATIVE || op->kind == FFTM_MI_OP_POISSON_GRADIENT ) &&
         ( op->axis < 0 || op->axis >= h.sg.nd ) )
        return fail( FFTM_MI_INVALID_VALUE, std::string( what ) + ": derivative axis out of range" );
    if ( op->kind == FFTM_MI_OP_BANDPASS && !( op->khi >= op->klo && op->klo >= 0.0 ) )
        return fail( FFTM_MI_INVALID_VALUE, std::string( what ) + ": band-pass needs 0 <= klo <= khi" );
    if ( op->kind == FFTM_MI_OP_GAUSSIAN && !( op->sigma > 0.0 ) )
        return fail( FFTM_MI_INVALID_VALUE, std::string( what ) + ": Gaussian sigma must be > 0" );
    if ( op->kind == FFTM_MI_OP_SHELL && ( !op->coeffs || op->nshells < 1 || op->nshells > MAX_SHELLS ) )
        return fail( FFTM_MI_INVALID_VALUE, std::string( what ) + ": shell operator needs coeffs and 1..1024 shells" );
    (void)h;
    return ok();
}

template <typename T>
filter_params<T> make_filter_params( const fftm_mi_ffts_s &h, const fftm_mi_spectral_op &op, int line_ax = -1 )
{
    filter_params<T> f{};
    f.kind     = op.kind;
    f.nshells  = op.kind == FFTM_MI_OP_SHELL ? op.nshells : 1;
    f.nd       = h.sg.nd;
    f.dax      = -1;
    f.line_m = -2;
    for ( int m = 0; m < h.sg.nd; ++m )
    {
        const int a = h.sg.ax[m];
        f.ext[m]    = h.sg.ext[m];
        f.mg[m]     = h.sg.mg[m];
        f.sh[m]     = h.sg.sh[m];
        f.nlen[m]   = a == h.sg.half_axis ? 0 : h.sg.n[a];
        if ( ( op.kind == FFTM_MI_OP_DERIVATIVE || op.kind == FFTM_MI_OP_POISSON_GRADIENT ) && a == op.axis )
            f.dax = m;
        if ( a == line_ax )
            f.line_m = m;
    }
    f.klo2 = op.klo * op.klo;
    f.khi2 = op.khi * op.khi;
    f.scale          = op.scale;
    f.inv_two_sigma2 = op.kind == FFTM_MI_OP_GAUSSIAN ? 1.0 / ( 2.0 * op.sigma * op.sigma ) : 0.0;
    f.kmax2          = h.sg.kmax2;
    f.coeffs         = static_cast<const T *>( op.coeffs );
    return f;
}

bool same_lines( const pass_rt &a, const pass_rt &b )
{
    return a.N == b.N && a.ies == b.ies && a.oes == b.oes && a.ncols == b.ncols && a.d0 == b.d0 && a.d1 == b.d1 &&
           a.is0 == b.is0 && a.is1 == b.is1 && a.is2 == b.is2 && a.os0 == b.os0 && a.os1 == b.os1 && a.os2 == b.os2;
}

// The real -> real filter chain: forward passes with the spectrum in the work area, then
// either the fused filter pass (last forward axis == first backward axis) or an
// elementwise multiply, then the remaining backward passes.
int ensure_filter_chain( fftm_mi_ffts_s &h )
{
    if ( h.flt_ready )
        return ok();
    if ( h.h16 )
        return fail( FFTM_MI_NOT_SUPPORTED, "fftm_mi_ffts_filter: not implemented for FP16-stored spectra" );
    if ( h.flags & FFTM_MI_FLAG_OVERWRITE_INPUT )
        return fail( FFTM_MI_NOT_SUPPORTED,
                     "fftm_mi_ffts_filter: keeps the spectrum in the work area (create without OVERWRITE_INPUT)" );
    std::vector<pass_rt> pre = h.fwd, post = h.bwd;
    for ( pass_rt &r : pre )
    {
        if ( r.src == BUF_OUT )
            r.src = BUF_WORK;
        if ( r.dst == BUF_OUT )
            r.dst = BUF_WORK;
    }
    for ( pass_rt &r : post )
        if ( r.src == BUF_IN )
            r.src = BUF_WORK;
    const pass_rt &lf = pre.back(), &fb = post.front();
    const bool     fusable = !lf.plane && !lf.generic && lf.kind == KIND_C2C && !lf.inv && !fb.plane && !fb.generic &&
                         fb.kind == KIND_C2C && fb.inv && same_lines( lf, fb ) && pow2_length( lf.N ) &&
                         lf.src == BUF_WORK && lf.dst == BUF_WORK && fb.dst == BUF_WORK;
    if ( fusable )
    {
        // the tile shape follows the tuned forward pass it replaces (shapes 0, 2, 4)
        pre.back().filter = true;
        const int lv      = variant_shape_of( lf.var );
        pre.back().var    = h.flt_var >= 0 ? h.flt_var
                                           : ( ( lv == 2 || lv == 4 ) ? lv : ( ( lv == 3 || lv == 5 ) ? 4 : 0 ) ) |
                                                 ( lf.var & ( 3 << 3 ) );
        post.erase( post.begin() );
    }
    h.flt_pre     = pre;
    h.flt_post    = post;
    h.flt_fused   = fusable;
    h.flt_line_ax = fusable ? lf.ax : -1;
    h.flt_ready = true;
    return ok();
}

template <typename T>
int spectral_op_launch( fftm_mi_ffts_s &h, const filter_params<T> &f, const void *in, void *out )
{
    const int64_t total = h.sg.total;
    const int64_t blocks = std::min<int64_t>( ( total + 255 ) / 256, 8192 );
    hipLaunchKernelGGL( k_spectral_op<T>, dim3( (unsigned)std::max<int64_t>( blocks, 1 ) ), dim3( 256 ), 0, h.stream,
                        f, static_cast<const cpx<T> *>( in ), static_cast<cpx<T> *>( out ), total );
    const hipError_t e = hipGetLastError();
    if ( e != hipSuccess )
        return fail( FFTM_MI_EXEC_FAILED, std::string( "fftm_mi: spectral operator launch failed: " ) + hipGetErrorString( e ) );
    return ok();
}

template <typename T>
int filter_run( fftm_mi_ffts_s &h, const void *in, void *out, const fftm_mi_spectral_op &op, bool sync )
{
    const filter_params<T> f  = make_filter_params<T>( h, op, h.flt_line_ax );
    int                    rc = run_passes( h.flt_pre, h.tb, h.prec, in, out, h.work, h.stream, false, false,
                                            "ffts_filter", 0, &f );
    if ( rc == FFTM_MI_SUCCESS && !h.flt_fused )
        rc = spectral_op_launch<T>( h, f, h.work, h.work );
    if ( rc == FFTM_MI_SUCCESS )
        rc = run_passes( h.flt_post, h.tb, h.prec, in, out, h.work, h.stream, false, sync, "ffts_filter" );
    return rc;
}

} // namespace

extern "C" {

int fftm_mi_ffts_spectral_apply( fftm_mi_ffts h, const void *spectrum_in, void *spectrum_out,
                                 const fftm_mi_spectral_op *op, int32_t sync )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_spectral_apply: null handle" );
    if ( !spectrum_in || !spectrum_out )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_spectral_apply: null data pointer" );
    int rc = check_op( *h, op, "fftm_mi_ffts_spectral_apply" );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    rc = h->prec == FFTM_MI_F32 ? spectral_op_launch<float>( *h, make_filter_params<float>( *h, *op ), spectrum_in, spectrum_out )
                                : spectral_op_launch<double>( *h, make_filter_params<double>( *h, *op ), spectrum_in, spectrum_out );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    return spectral_launch_check( h, sync != 0, "fftm_mi_ffts_spectral_apply" );
}

int fftm_mi_ffts_filter( fftm_mi_ffts h, const void *real_in, void *real_out, const fftm_mi_spectral_op *op,
                         int32_t sync )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_filter: null handle" );
    if ( !real_in || !real_out )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_filter: null data pointer" );
    int rc = check_op( *h, op, "fftm_mi_ffts_filter" );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    rc = ensure_filter_chain( *h );
    if ( rc != FFTM_MI_SUCCESS )
        return rc;
    if ( !h->work )
        return fail( FFTM_MI_NO_WORKSPACE, "fftm_mi_ffts_filter: a work area is required (set_work_area)" );
    return h->prec == FFTM_MI_F32 ? filter_run<float>( *h, real_in, real_out, *op, sync != 0 )
                                  : filter_run<double>( *h, real_in, real_out, *op, sync != 0 );
}

int fftm_mi_ffts_set_pass_variant( fftm_mi_ffts h, int32_t backward, int32_t index, int32_t variant )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_set_pass_variant: null handle" );
    std::vector<pass_rt> &v = backward ? h->bwd : h->fwd;
    if ( index < 0 || index >= (int32_t)v.size() )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_set_pass_variant: pass index out of range" );
    if ( !variant_number_ok( variant ) )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_set_pass_variant: variant out of range" );
    if ( v[(size_t)index].iom && !mixed_variant_ok( variant ) )
        return fail( FFTM_MI_NOT_SUPPORTED, "fftm_mi_ffts_set_pass_variant: FP16-storage passes have tile shapes 0, 2, 6, 7" );
    if ( !pass_variant_ok( v[(size_t)index], variant, h->prec ) )
        return fail( FFTM_MI_NOT_SUPPORTED,
                     "fftm_mi_ffts_set_pass_variant: radix schedules 1 and 2 exist for strided C2C passes of length 512 "
                     "and 1024 in tile shapes 0 and 6, schedule 1 also in shape 2 (FP32 arithmetic; hi + lo FP16 "
                     "storage: tile shapes 0, 2 and 6, schedule 0)" );
    v[(size_t)index].var = variant;
    h->flt_ready         = false;
    return ok();
}

int fftm_mi_ffts_get_pass_variant( fftm_mi_ffts h, int32_t backward, int32_t index, int32_t *variant )
{
    if ( !h || !variant )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_get_pass_variant: null argument" );
    const std::vector<pass_rt> &v = backward ? h->bwd : h->fwd;
    if ( index < 0 || index >= (int32_t)v.size() )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_get_pass_variant: pass index out of range" );
    *variant = v[(size_t)index].var;
    return ok();
}

int fftm_mi_ffts_destroy( fftm_mi_ffts h )
{
    if ( !h )
        return fail( FFTM_MI_INVALID_PLAN, "fftm_mi_ffts_destroy: null handle" );
    delete h;
    return ok();
}

} // extern "C"

// ---------------------------------------------------------------------------
// Tile-shape tuning with the reference's autotune cache protocol
// (source/fftm_autotune.hpp:126-169 KEY=VALUE cache, header line, tmp + rename;
//  detail/fftm_autotune_hardware.inc:87-98, 285-337 FNV-1a-64 "v2:" signature;
//  fftm_autotune_measure.hpp:293-400 median of iterations after warmup, per-candidate
//  records, fastest valid candidate wins).  Here the candidates are the variants of
//  each pass (tile shape x tile order); validity = output bitwise equal to variant 0.
// ---------------------------------------------------------------------------
namespace
{

std::string fnv1a64_hex( const std::string &v )
{
    uint64_t h = 14695981039346656037ull;
    for ( unsigned char c : v )
    {
        h ^= (uint64_t)c;
        h *= 1099511628211ull;
    }
    char buf[17];
    std::snprintf( buf, sizeof( buf ), "%016llx", (unsigned long long)h );
    return buf;
}

std::string trim_copy( const std::string &s )
{
    size_t a = s.find_first_not_of( " \t\r\n" );
    if ( a == std::string::npos )
        return "";
    size_t b = s.find_last_not_of( " \t\r\n" );
    return s.substr( a, b - a + 1 );
}

using cache_map = std::map<std::string, std::string>;

// returns 0 ok, 1 missing file, -line for a malformed line
int cache_load( const std::string &path, cache_map &out )
{
    std::ifstream in( path.c_str() );
    if ( !in )
        return 1;
    std::string line;
    int         no = 0;
    while ( std::getline( in, line ) )
    {
        ++no;
        line = trim_copy( line );
        if ( line.empty() || line[0] == '#' )
            continue;
        const size_t eq = line.find( '=' );
        if ( eq == std::string::npos )
            return -no;
        out[trim_copy( line.substr( 0, eq ) )] = trim_copy( line.substr( eq + 1 ) );
    }
    return 0;
}

bool cache_save( const std::string &path, const cache_map &m )
{
    const std::string tmp = path + ".tmp";
    {
        std::ofstream o( tmp.c_str() );
        if ( !o )
            return false;
        o << "# FFTM C++ autotune cache\n";
        for ( const auto &kv : m )
            o << kv.first << "=" << kv.second << "\n";
    }
    if ( std::rename( tmp.c_str(), path.c_str() ) != 0 )
    {
        std::remove( tmp.c_str() );
        return false;
    }
    return true;
}

// canonical hardware string and signature of the current device
bool hardware_signature( std::string &canonical, std::string &signature, std::string &arch, std::string &model )
{
    int dev = 0;
    if ( hipGetDevice( &dev ) != hipSuccess )
        return false;
    hipDeviceProp_t pr;
    if ( hipGetDeviceProperties( &pr, dev ) != hipSuccess )
        return false;
    int rt = 0, drv = 0;
    (void)hipRuntimeGetVersion( &rt );
    (void)hipDriverGetVersion( &drv );
    arch  = pr.gcnArchName;
    model = pr.name;
    for ( char &c : model )
        if ( c == ' ' || c == ';' || c == '=' )
            c = '_';
    std::ostringstream o;
    o << "version=2;device_classes=" << arch << ";device_models=" << model << ";compute_units=" << pr.multiProcessorCount
      << ";device_total_bytes=" << (unsigned long long)pr.totalGlobalMem << ";runtime_versions=" << rt
      << ";driver_versions=" << drv;
    canonical = o.str();
    signature = "v2:" + fnv1a64_hex( canonical );
    return true;
}

std::string problem_key( const fftm_mi_ffts_s &h )
{
    static const char *names[] = { "3D", "4D_PENCIL", "4D_SLAB", "4D_XZWY" };
    std::ostringstream o;
    o << "FFTM_MI_TILE_" << names[h.layout] << "_" << h.n[0] << "x" << h.n[1] << "x" << h.n[2];
    if ( h.layout != FFTM_MI_LAYOUT_3D )
        o << "x" << h.n[3];
    o << ( h.h16x2 ? "_F32_H16X2" : ( h.h16 ? "_F32_H16" : ( h.prec == FFTM_MI_F32 ? "_F32" : "_F64" ) ) );
    // creation flags change the pass list (plane fusion, in-place backward): records are
    // per pass index, so a handle only takes records measured with the same flags
    if ( h.flags )
        o << "_FLAGS" << h.flags;
    return o.str();
}

std::string pass_key( const std::string &prob, int bw, size_t i )
{
    return prob + ( bw ? "_BWD" : "_FWD" ) + std::to_string( i );
}

__global__ void k_count_diff( const uint32_t *a, const uint32_t *b, int64_t n, unsigned long long *count )
{
    unsigned long long c = 0;
    for ( int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256 )
        c += a[i] != b[i] ? 1ull : 0ull;
    if ( c )
        atomicAdd( count, c );
}

// variants of the fused filter pass: tile shapes 0, 2 and 4, each in every tile order
constexpr int FILTER_CANDIDATES[] = { 0, 2, 4, 8, 10, 12, 16, 18, 20, 24, 26, 28 };

bool filter_variant_ok( int var )
{
    for ( int c : FILTER_CANDIDATES )
        if ( var == c )
            return true;
    return false;
}

// Bitwise comparison of n bytes (a multiple of 4) on the handle's stream: *equal is set
// when they match.  `ref` receives a copy instead when `capture` is set.
hipError_t tune_compare( const fftm_mi_ffts_s &h, void *ref, const void *buf, size_t n, bool capture,
                         unsigned long long *dcount, bool *equal )
{
    *equal        = true;
    hipError_t he = hipSuccess;
    if ( capture )
    {
        he = hipMemcpyAsync( ref, buf, n, hipMemcpyDeviceToDevice, h.stream );
        return he == hipSuccess ? hipStreamSynchronize( h.stream ) : he;
    }
    unsigned long long diff = 0;
    he                      = hipMemsetAsync( dcount, 0, sizeof( unsigned long long ), h.stream );
    if ( he == hipSuccess )
    {
        hipLaunchKernelGGL( k_count_diff, dim3( 1024 ), dim3( 256 ), 0, h.stream, static_cast<const uint32_t *>( ref ),
                            static_cast<const uint32_t *>( buf ), (int64_t)( n / 4 ), dcount );
        he = hipGetLastError();
    }
    if ( he == hipSuccess )
        he = hipMemcpyAsync( &diff, dcount, sizeof( diff ), hipMemcpyDeviceToHost, h.stream );
    if ( he == hipSuccess )
        he = hipStreamSynchronize( h.stream );
    *equal = diff == 0;
    return he;
}

// mx[0] = max |a - b| (a NaN counts as infinite), mx[1] = max |b|, as the bits of
// non-negative floats (which order like the unsigned integers)
__global__ void __launch_bounds__( 256 ) k_absdiff_max( const float *a, const float *b, int64_t n, unsigned *mx )
{
    __shared__ float sd[256], sm[256];
    float            d = 0.f, m = 0.f;
    for ( int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256 )
    {
        const float x = a[i], y = b[i];
        const float e = fabsf( x - y );
        d             = fmaxf( d, e == e ? e : __builtin_inff() );
        m             = fmaxf( m, fabsf( y ) );
    }
    sd[threadIdx.x] = d;
    sm[threadIdx.x] = m;
    __syncthreads();
    for ( int w = 128; w > 0; w >>= 1 )
    {
        if ( (int)threadIdx.x < w )
        {
            sd[threadIdx.x] = fmaxf( sd[threadIdx.x], sd[threadIdx.x + w] );
            sm[threadIdx.x] = fmaxf( sm[threadIdx.x], sm[threadIdx.x + w] );
        }
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
    {
        atomicMax( mx, __float_as_uint( sd[0] ) );
        atomicMax( mx + 1, __float_as_uint( sm[0] ) );
    }
}

// buf (n bytes of FP32 values) against the captured ref: max |buf - ref| <= tol max |ref|
hipError_t tune_compare_tol( const fftm_mi_ffts_s &h, const void *ref, const void *buf, size_t n, double tol,
                             unsigned long long *scratch8, bool *ok )
{
    *ok              = false;
    unsigned    *mx  = reinterpret_cast<unsigned *>( scratch8 );
    unsigned     hm[2] = { 0, 0 };
    hipError_t   he  = hipMemsetAsync( mx, 0, 2 * sizeof( unsigned ), h.stream );
    if ( he == hipSuccess )
    {
        hipLaunchKernelGGL( k_absdiff_max, dim3( 1024 ), dim3( 256 ), 0, h.stream, static_cast<const float *>( buf ),
                            static_cast<const float *>( ref ), (int64_t)( n / 4 ), mx );
        he = hipGetLastError();
    }
    if ( he == hipSuccess )
        he = hipMemcpyAsync( hm, mx, sizeof( hm ), hipMemcpyDeviceToHost, h.stream );
    if ( he == hipSuccess )
        he = hipStreamSynchronize( h.stream );
    float d = 0.f, m = 0.f;
    std::memcpy( &d, &hm[0], 4 );
    std::memcpy( &m, &hm[1], 4 );
    *ok = std::isfinite( d ) && d <= tol * m;
    return he;
}

// The fused filter pass (forward passes, x H(k) inside the last one, inverse passes) in
// each of its variants, with a Gaussian low-pass as the operator: the fastest one whose
// output is bitwise equal to variant 0's is kept in h.flt_var.  Layouts whose filter
// chain does not fuse have nothing to tune.
int tune_filter( fftm_mi_ffts_s &h, const void *real_in, void *real_out, void *ref, size_t rbytes,
                 unsigned long long *dcount, hipEvent_t e0, hipEvent_t e1, int warmup, int iterations,
                 const std::string &key, cache_map &rec )
{
    if ( h.h16 || ( h.flags & FFTM_MI_FLAG_OVERWRITE_INPUT ) )
        return ok();
    h.flt_var   = -1;
    h.flt_ready = false;
    int rc      = ensure_filter_chain( h );
    if ( rc != FFTM_MI_SUCCESS || !h.flt_fused )
        return rc;
    fftm_mi_spectral_op op{};
    op.kind  = FFTM_MI_OP_GAUSSIAN;
    op.scale = 1.0;
    op.sigma = 8.0;
    auto run = [&]() {
        return h.prec == FFTM_MI_F32 ? filter_run<float>( h, real_in, real_out, op, false )
                                     : filter_run<double>( h, real_in, real_out, op, false );
    };
    double best_ms = 1e300;
    int    best    = -1;
    for ( int var : FILTER_CANDIDATES )
    {
        h.flt_var   = var;
        h.flt_ready = false;
        rc          = ensure_filter_chain( h );
        if ( rc == FFTM_MI_SUCCESS )
            rc = run();
        if ( rc != FFTM_MI_SUCCESS )
            break;
        bool       valid = true;
        hipError_t he    = tune_compare( h, ref, real_out, rbytes, var == 0, dcount, &valid );
        std::vector<float> ms;
        for ( int w = 0; w < warmup && he == hipSuccess && rc == 0; ++w )
            rc = run();
        for ( int it = 0; it < iterations && he == hipSuccess && rc == 0; ++it )
        {
            float t = 0.f;
            he      = hipEventRecord( e0, h.stream );
            if ( he == hipSuccess )
                rc = run();
            if ( he == hipSuccess && rc == 0 )
                he = hipEventRecord( e1, h.stream );
            if ( he == hipSuccess && rc == 0 )
                he = hipEventSynchronize( e1 );
            if ( he == hipSuccess && rc == 0 )
                he = hipEventElapsedTime( &t, e0, e1 );
            ms.push_back( t );
        }
        if ( he != hipSuccess )
            rc = fail( FFTM_MI_EXEC_FAILED, std::string( "fftm_mi_ffts_autotune: " ) + hipGetErrorString( he ) );
        if ( rc != 0 )
            break;
        std::sort( ms.begin(), ms.end() );
        const double med = ms.size() % 2 ? ms[ms.size() / 2] : 0.5 * ( ms[ms.size() / 2 - 1] + ms[ms.size() / 2] );
        char         pfx[32];
        std::snprintf( pfx, sizeof( pfx ), "_CANDIDATE_%03d_", var );
        std::ostringstream mo;
        mo << std::setprecision( 9 ) << med;
        rec[key + pfx + "VALID"]     = valid ? "1" : "0";
        rec[key + pfx + "MEDIAN_MS"] = valid ? mo.str() : "nan";
        if ( valid && med < best_ms )
        {
            best_ms = med;
            best    = var;
        }
    }
    h.flt_var   = rc == 0 ? best : -1;
    h.flt_ready = false;
    if ( rc == 0 && best >= 0 )
    {
        rec[key] = std::to_string( best );
        std::ostringstream mo;
        mo << std::setprecision( 9 ) << best_ms;
        rec[key + "_WINNER_MEDIAN_MS"] = mo.str();
    }
    return rc;
}

} // namespace

extern "C" {

int fftm_mi_fnv1a64( const char *text, char *hex17 )
{
    if ( !text || !hex17 )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_fnv1a64: null argument" );
    const std::string h = fnv1a64_hex( text );
    std::memcpy( hex17, h.c_str(), 17 );
    return ok();
}

int fftm_mi_hardware_signature( char *buf, size_t len )
{
    std::string canon, sig, arch, model;
    if ( !hardware_signature( canon, sig, arch, model ) )
        return fail( FFTM_MI_INVALID_DEVICE, "fftm_mi_hardware_signature: no device" );
    if ( !buf || len < sig.size() + 1 )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_hardware_signature: buffer too small" );
    std::memcpy( buf, sig.c_str(), sig.size() + 1 );
    return ok();
}

int fftm_mi_ffts_load_tuning( fftm_mi_ffts h, const char *cache_path, int32_t *applied )
{
    if ( !h || !cache_path )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_load_tuning: null argument" );
    if ( applied )
        *applied = 0;
    cache_map  m;
    const int  rc = cache_load( cache_path, m );
    if ( rc == 1 )
        return ok(); // no cache yet: defaults stay
    if ( rc < 0 )
        return fail( FFTM_MI_INVALID_VALUE, std::string( "fftm_mi_ffts_load_tuning: " ) + cache_path + ":" +
                                                std::to_string( -rc ) + " is not KEY=VALUE" );
    std::string canon, sig, arch, model;
    if ( !hardware_signature( canon, sig, arch, model ) )
        return fail( FFTM_MI_INVALID_DEVICE, "fftm_mi_ffts_load_tuning: no device" );
    if ( m["FFTM_AUTOTUNE_HARDWARE_SIGNATURE"] != sig )
        return ok(); // measured on other hardware: not applied
    const std::string prob = problem_key( *h );
    int               n    = 0;
    for ( int bw = 0; bw < 2; ++bw )
    {
        std::vector<pass_rt> &v = bw ? h->bwd : h->fwd;
        for ( size_t i = 0; i < v.size(); ++i )
        {
            auto it = m.find( pass_key( prob, bw, i ) );
            if ( it == m.end() )
                continue;
            const int var = std::atoi( it->second.c_str() );
            if ( pass_variant_ok( v[i], var, h->prec ) )
            {
                v[i].var = var;
                ++n;
            }
        }
    }
    auto it = m.find( prob + "_FILTER" );
    if ( it != m.end() && filter_variant_ok( std::atoi( it->second.c_str() ) ) )
    {
        h->flt_var = std::atoi( it->second.c_str() );
        ++n;
    }
    h->flt_ready = false;
    if ( applied )
        *applied = n;
    return ok();
}

int fftm_mi_ffts_autotune( fftm_mi_ffts h, const void *real_in, void *spectrum_out, int32_t warmup,
                           int32_t iterations, const char *cache_path )
{
    if ( !h || !real_in || !spectrum_out )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_autotune: null argument" );
    if ( h->work_size > 0 && !h->work )
        return fail( FFTM_MI_NO_WORKSPACE, "fftm_mi_ffts_autotune: a work area is required" );
    if ( iterations < 1 || warmup < 0 )
        return fail( FFTM_MI_INVALID_VALUE, "fftm_mi_ffts_autotune: iterations >= 1, warmup >= 0" );
    // the real output of the backward passes needs its own buffer: same bytes as the input
    int64_t nreal = 1, ncplx = 1;
    const int dim = h->layout == FFTM_MI_LAYOUT_3D ? 3 : 4;
    for ( int k = 0; k < dim; ++k )
    {
        nreal *= h->n[k];
        ncplx *= k == dim - 1 ? h->n[k] / 2 + 1 : h->n[k];
    }
    const size_t rbytes = (size_t)nreal * ( h->prec == FFTM_MI_F32 ? 4 : 8 );
    const size_t cbytes = (size_t)ncplx * ( h->prec == FFTM_MI_F32 ? 8 : 16 );
    void               *real_out = nullptr, *ref = nullptr;
    unsigned long long *dcount   = nullptr;
    hipEvent_t          e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        if ( real_out )
            (void)hipFree( real_out );
        if ( ref )
            (void)hipFree( ref );
        if ( dcount )
            (void)hipFree( dcount );
        if ( e0 )
            (void)hipEventDestroy( e0 );
        if ( e1 )
            (void)hipEventDestroy( e1 );
    };
    // ref holds a copy of any pass's whole destination: the real output, the spectrum or the
    // work area (which is larger than the spectrum with hi + lo FP16 storage)
    const size_t refbytes = std::max( std::max( rbytes, cbytes ), h->work_size );
    if ( hipMalloc( &real_out, rbytes ) != hipSuccess || hipMalloc( &ref, refbytes ) != hipSuccess ||
         hipMalloc( &dcount, sizeof( unsigned long long ) ) != hipSuccess || hipEventCreate( &e0 ) != hipSuccess ||
         hipEventCreate( &e1 ) != hipSuccess )
    {
        cleanup();
        return fail( FFTM_MI_ALLOC_FAILED, "fftm_mi_ffts_autotune: scratch allocation failed" );
    }
    const std::string prob = problem_key( *h );
    cache_map         rec;
    int               rc = 0;
    for ( int bw = 0; bw < 2 && rc == 0; ++bw )
    {
        std::vector<pass_rt> &v   = bw ? h->bwd : h->fwd;
        const void           *in  = bw ? (const void *)spectrum_out : real_in;
        void                 *out = bw ? real_out : spectrum_out;
        for ( size_t i = 0; i < v.size() && rc == 0; ++i )
        {
            // the pass writes into: out (its dst) -- compare the whole destination buffer
            const pass_rt &cur     = v[i];
            void          *dstbuf  = cur.dst == BUF_OUT ? out : ( cur.dst == BUF_WORK ? h->work : const_cast<void *>( in ) );
            const size_t   dbytes  = cur.dst == BUF_OUT ? ( bw ? rbytes : cbytes )
                                                        : ( cur.dst == BUF_WORK ? h->work_size : cbytes );
            std::vector<pass_rt> prefix( v.begin(), v.begin() + (long)i );
            std::vector<pass_rt> one( 1, cur );
            // the pass before this one in the forward + backward sequence, with its buffers
            std::vector<pass_rt> prev;
            const void          *prev_in  = real_in;
            void                *prev_out = spectrum_out;
            if ( i > 0 )
            {
                prev.push_back( v[i - 1] );
                prev_in  = in;
                prev_out = out;
            }
            else if ( bw && !h->fwd.empty() )
                prev.push_back( h->fwd.back() );
            double               best_ms = 1e300;
            int                  best    = cur.var;
            // tile shapes 0..nvar-1, each in every tile order
            // (plane passes: shape 0 the LDS kernel, 1 the register kernel; any-radix
            // passes have one shape and their own order)
            const int            nvar    = cur.generic ? 1 : ( cur.plane ? 2 : fixed_variants( cur.N ) );
            const int            ncand   = cur.generic ? 1 : NUM_ORDERS * nvar;
            // radix schedules 1 and 2 (alternative schedule, and its radix-16 stage on the
            // matrix cores) where compiled, for passes with FP32 output (their validity is a
            // tolerance against variant 0's output, see below)
            const int            nrs     = cur.iom ? 1 : 3;
            for ( int cand = 0; cand < ncand * nrs && rc == 0; ++cand )
            {
                const int base = cand % ncand;
                const int var  = base % nvar + 8 * ( base / nvar ) + 32 * ( cand / ncand );
                if ( !pass_variant_ok( cur, var, h->prec ) )
                    continue;
                one[0].var = var;
                // regenerate the pass input, run once, check against variant 0
                if ( bw )
                    rc = run_passes( h->fwd, h->tb, h->prec, real_in, spectrum_out, h->work, h->stream, false, false, "autotune" );
                if ( rc == 0 )
                    rc = run_passes( prefix, h->tb, h->prec, in, out, h->work, h->stream, false, false, "autotune" );
                if ( rc == 0 )
                    rc = run_passes( one, h->tb, h->prec, in, out, h->work, h->stream, false, true, "autotune" );
                if ( rc != 0 )
                    break;
                // every copy, clear and comparison is ordered on the handle's stream (which
                // may be non-blocking with respect to the null stream)
                // valid: bitwise equal to variant 0 (same radix schedule: same arithmetic in
                // another tile shape or order), or, for another radix schedule, within the
                // FP32 parity tolerance of variant 0's output (relative max 1e-5, tighter than
                // the reference's 1000 eps log2 N round-trip check,
                // fftm_autotune_measure.hpp:837-845)
                bool       valid = true;
                hipError_t he    = variant_rs_of( var ) == 0
                                       ? tune_compare( *h, ref, dstbuf, dbytes, var == 0, dcount, &valid )
                                       : tune_compare_tol( *h, ref, dstbuf, dbytes, 1e-5, dcount, &valid );
                if ( he != hipSuccess )
                {
                    rc = fail( FFTM_MI_EXEC_FAILED, std::string( "fftm_mi_ffts_autotune: " ) + hipGetErrorString( he ) );
                    break;
                }
                // timing: warmup, then `iterations` launches, median.  Each timed launch
                // follows the pass that precedes it in the forward + backward sequence, as
                // in a transform pair: what that pass leaves in the memory-side cache (and
                // the order it wrote in) changes which candidate is fastest.
                std::vector<float> ms;
                auto               run_prev = [&]() {
                    if ( prev.empty() )
                        return 0;
                    return run_passes( prev, h->tb, h->prec, prev_in, prev_out, h->work, h->stream, false, false,
                                       "autotune" );
                };
                for ( int w = 0; w < warmup && rc == 0; ++w )
                {
                    rc = run_prev();
                    if ( rc == 0 )
                        rc = run_passes( one, h->tb, h->prec, in, out, h->work, h->stream, false, false, "autotune" );
                }
                for ( int it = 0; it < iterations && rc == 0; ++it )
                {
                    float t = 0.f;
                    rc      = run_prev();
                    if ( rc != 0 )
                        break;
                    he = hipEventRecord( e0, h->stream );
                    if ( he == hipSuccess )
                        rc = run_passes( one, h->tb, h->prec, in, out, h->work, h->stream, false, false, "autotune" );
                    if ( he == hipSuccess && rc == 0 )
                        he = hipEventRecord( e1, h->stream );
                    if ( he == hipSuccess && rc == 0 )
                        he = hipEventSynchronize( e1 );
                    if ( he == hipSuccess && rc == 0 )
                        he = hipEventElapsedTime( &t, e0, e1 );
                    if ( he != hipSuccess )
                        rc = fail( FFTM_MI_EXEC_FAILED, std::string( "fftm_mi_ffts_autotune: " ) + hipGetErrorString( he ) );
                    ms.push_back( t );
                }
                if ( rc != 0 )
                    break;
                std::sort( ms.begin(), ms.end() );
                const double med = ms.size() % 2 ? ms[ms.size() / 2] : 0.5 * ( ms[ms.size() / 2 - 1] + ms[ms.size() / 2] );
                char         pfx[32];
                std::snprintf( pfx, sizeof( pfx ), "_CANDIDATE_%03d_", var );
                const std::string k = pass_key( prob, bw, i ) + pfx;
                std::ostringstream mo;
                mo << std::setprecision( 9 ) << med;
                rec[k + "VALID"]     = valid ? "1" : "0";
                rec[k + "MEDIAN_MS"] = valid ? mo.str() : "nan";
                if ( valid && med < best_ms )
                {
                    best_ms = med;
                    best    = var;
                }
            }
            if ( rc == 0 )
            {
                h->flt_ready                    = false;
                v[i].var                        = best;
                rec[pass_key( prob, bw, i )]    = std::to_string( best );
                std::ostringstream mo;
                mo << std::setprecision( 9 ) << best_ms;
                rec[pass_key( prob, bw, i ) + "_WINNER_MEDIAN_MS"] = mo.str();
            }
        }
    }
    if ( rc == 0 )
        rc = tune_filter( *h, real_in, real_out, ref, rbytes, dcount, e0, e1, warmup, iterations, prob + "_FILTER", rec );
    cleanup();
    if ( rc != 0 )
        return rc;
    if ( cache_path && *cache_path )
    {
        cache_map m;
        const int lr = cache_load( cache_path, m );
        if ( lr < 0 )
            m.clear(); // a malformed cache is replaced
        std::string canon, sig, arch, model;
        if ( !hardware_signature( canon, sig, arch, model ) )
            return fail( FFTM_MI_INVALID_DEVICE, "fftm_mi_ffts_autotune: no device" );
        if ( m.count( "FFTM_AUTOTUNE_HARDWARE_SIGNATURE" ) && m["FFTM_AUTOTUNE_HARDWARE_SIGNATURE"] != sig )
            m.clear(); // measured elsewhere: start a new cache for this hardware
        m["FFTM_AUTOTUNE_SCHEMA"]                     = "2";
        m["FFTM_AUTOTUNE_SOURCE"]                     = "cpp-measured-v2";
        m["FFTM_AUTOTUNE_MEASURED"]                   = "1";
        m["FFTM_AUTOTUNE_MEASURE_WARMUP"]             = std::to_string( warmup );
        m["FFTM_AUTOTUNE_MEASURE_ITERATIONS"]         = std::to_string( iterations );
        m["FFTM_AUTOTUNE_HARDWARE_SIGNATURE_VERSION"] = "2";
        m["FFTM_AUTOTUNE_HARDWARE_SIGNATURE"]         = sig;
        m["FFTM_AUTOTUNE_HARDWARE_CANONICAL"]         = canon;
        m["FFTM_AUTOTUNE_GPU_ARCHITECTURES"]          = arch;
        m["FFTM_AUTOTUNE_GPU_MODELS"]                 = model;
        for ( const auto &kv : rec )
            m[kv.first] = kv.second;
        if ( !cache_save( cache_path, m ) )
            return fail( FFTM_MI_INVALID_VALUE, std::string( "fftm_mi_ffts_autotune: cannot write " ) + cache_path );
    }
    return ok();
}

} // extern "C"

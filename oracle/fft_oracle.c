/*
 * fft_oracle.c -- CPU FP64 restatement of the `ffts` R2C/C2R arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is built into
 * oracle/liboracle.so and is loaded by tests/, by __graft_entry__.smoke() and by
 * the cpu_baseline leg of bench.py, always as the checker.  The product library
 * (fftm_amd/) never links, loads or calls it.
 *
 * What it restates (reference files cited per function):
 *   - splitmix64 / unit_random_from_state / fill_random_real_{3d,4d}
 *       source/tests/detail/fft_benchmark_common.h:201-257
 *   - FFTW harness deterministic input  cpu_fftw_mpi/fftw_mpi_benchmark.cpp:215-219
 *   - the arithmetic of one hipFFT/cuFFT advanced-layout plan
 *       (rank 1..3, n/inembed/istride/idist/onembed/ostride/odist/batch, types
 *        R2C / C2R / C2C forward / C2C backward, unnormalised), i.e. what
 *       hipfftExecR2C/C2R/C2C compute for the descriptors ffts builds
 *       source/external_wrap/hipfft_wrap.h:303-417,1253-1275,
 *       source/ffts.hpp:659-776.
 *     C2R is defined for every input (not only Hermitian ones) the way
 *     numpy.fft.irfftn / FFTW define it: inverse C2C along all but the last axis,
 *     then a 1D complex-to-real transform along the last axis that takes the
 *     Hermitian extension of bins 0..n/2 and ignores the imaginary parts of the
 *     DC and (even n) Nyquist bins.
 *   - a brute-force long-double DFT for tiny shapes (pinning only).
 *
 * The 1D engine is a Stockham autosort FFT over any factorisation (radix 4, 2,
 * 3, 5 and a generic O(p^2) DFT for any other prime), with twiddles e^{-2 pi i k/n}
 * evaluated in long double and rounded once.  Passes over pencils run under
 * OpenMP when the library is built with it (bench.py times it that way).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define FO_R2C 1
#define FO_C2R 2
#define FO_C2CF 3
#define FO_C2CB 4

/* ------------------------------------------------------------------------ */
/* Random inputs: fft_benchmark_common.h:201-257                              */
/* ------------------------------------------------------------------------ */

uint64_t fo_splitmix64( uint64_t x )
{
    x += 0x9E3779B97F4A7C15ull;
    x = ( x ^ ( x >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
    x = ( x ^ ( x >> 27 ) ) * 0x94D049BB133111EBull;
    return x ^ ( x >> 31 );
}

/* unit_random_from_state<double>: 2*u - 1 with u = (bits >> 11) * 2^-53. */
double fo_unit_random_f64( uint64_t state )
{
    const uint64_t bits = fo_splitmix64( state );
    const double   unit = (double)( bits >> 11 ) * ( 1.0 / 9007199254740992.0 );
    return 2.0 * unit - 1.0;
}

/* unit_random_from_state<float>: static_cast<float>(2.f * static_cast<float>(unit) - 1.f). */
float fo_unit_random_f32( uint64_t state )
{
    const uint64_t bits = fo_splitmix64( state );
    const double   unit = (double)( bits >> 11 ) * ( 1.0 / 9007199254740992.0 );
    const float    u    = (float)unit;
    return 2.0f * u - 1.0f;
}

static uint64_t fo_state_3d( uint64_t seed, uint64_t gx, uint64_t gy, uint64_t gz )
{
    return seed ^ ( gx * 0xD6E8FEB86659FD93ull ) ^ ( gy * 0xA5A3564E27F886A5ull ) ^ ( gz * 0x9E3779B97F4A7C15ull );
}

static uint64_t fo_state_4d( uint64_t seed, uint64_t gx, uint64_t gy, uint64_t gz, uint64_t gw )
{
    return seed ^ ( gx * 0xD6E8FEB86659FD93ull ) ^ ( gy * 0xA5A3564E27F886A5ull ) ^ ( gz * 0x9E3779B97F4A7C15ull ) ^
           ( gw * 0x94D049BB133111EBull );
}

/*
 * fill_random_real_3d_functor (fft_benchmark_common.h:217-234) over logical
 * (x,y,z).  `sx,sy,sz` are element strides of the destination, so any
 * physical arrangement can be filled.  prec: 0 = float, 1 = double.
 */
void fo_fill_random_real_3d(
    void *dst, int prec, int64_t nx, int64_t ny, int64_t nz, int64_t sx, int64_t sy, int64_t sz, uint64_t seed
)
{
#pragma omp parallel for collapse( 2 ) schedule( static )
    for ( int64_t x = 0; x < nx; ++x )
        for ( int64_t y = 0; y < ny; ++y )
            for ( int64_t z = 0; z < nz; ++z )
            {
                const uint64_t s   = fo_state_3d( seed, (uint64_t)x, (uint64_t)y, (uint64_t)z );
                const int64_t  off = x * sx + y * sy + z * sz;
                if ( prec == 0 )
                    ( (float *)dst )[off] = fo_unit_random_f32( s );
                else
                    ( (double *)dst )[off] = fo_unit_random_f64( s );
            }
}

/* fill_random_real_4d_functor (fft_benchmark_common.h:236-257) over logical (x,y,z,w). */
void fo_fill_random_real_4d(
    void *dst, int prec, int64_t nx, int64_t ny, int64_t nz, int64_t nw, int64_t sx, int64_t sy, int64_t sz, int64_t sw,
    uint64_t seed
)
{
#pragma omp parallel for collapse( 2 ) schedule( static )
    for ( int64_t x = 0; x < nx; ++x )
        for ( int64_t y = 0; y < ny; ++y )
            for ( int64_t z = 0; z < nz; ++z )
                for ( int64_t w = 0; w < nw; ++w )
                {
                    const uint64_t s   = fo_state_4d( seed, (uint64_t)x, (uint64_t)y, (uint64_t)z, (uint64_t)w );
                    const int64_t  off = x * sx + y * sy + z * sz + w * sw;
                    if ( prec == 0 )
                        ( (float *)dst )[off] = fo_unit_random_f32( s );
                    else
                        ( (double *)dst )[off] = fo_unit_random_f64( s );
                }
}

/* input_value of the FFTW harness, cpu_fftw_mpi/fftw_mpi_benchmark.cpp:215-219. */
double fo_fftw_input_value( uint64_t global_index )
{
    const uint64_t code = ( global_index * 17ull + 13ull ) & 1023ull;
    return ( (double)code - 511.5 ) / 512.0;
}

/* ------------------------------------------------------------------------ */
/* 1D engine                                                                 */
/* ------------------------------------------------------------------------ */

typedef struct
{
    int64_t n;
    int     nfac;
    int64_t fac[64];
    double *tw; /* 2n: e^{-2 pi i k / n} */
} fo_plan1d;

static void fo_plan1d_init( fo_plan1d *p, int64_t n )
{
    p->n    = n;
    p->nfac = 0;
    int64_t m = n;
    while ( m % 4 == 0 )
    {
        p->fac[p->nfac++] = 4;
        m /= 4;
    }
    while ( m % 2 == 0 )
    {
        p->fac[p->nfac++] = 2;
        m /= 2;
    }
    for ( int64_t f = 3; f * f <= m; f += 2 )
        while ( m % f == 0 )
        {
            p->fac[p->nfac++] = f;
            m /= f;
        }
    if ( m > 1 )
        p->fac[p->nfac++] = m;
    p->tw = (double *)malloc( sizeof( double ) * 2 * (size_t)( n > 0 ? n : 1 ) );
    for ( int64_t k = 0; k < n; ++k )
    {
        /* reduce the angle to the first octant-free form k/n in long double */
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
        p->tw[2 * k]        = (double)cosl( a );
        p->tw[2 * k + 1]    = (double)sinl( a );
    }
}

static void fo_plan1d_free( fo_plan1d *p )
{
    free( p->tw );
    p->tw = NULL;
}

/*
 * Stockham autosort FFT of one contiguous complex pencil (interleaved).
 * sign = -1 forward (e^{-i}), +1 backward (e^{+i}); unnormalised.
 * `scratch` holds 2n doubles.  Result is left in `data`.
 */
static void fo_fft1d( const fo_plan1d *p, double *data, double *scratch, int sign )
{
    const int64_t n = p->n;
    if ( n <= 1 )
        return;
    double *src = data, *dst = scratch;
    int64_t ns  = 1;
    double  vr[64], vi[64];
    double *br = NULL, *bi = NULL;
    for ( int f = 0; f < p->nfac; ++f )
    {
        const int64_t r   = p->fac[f];
        const int64_t nr  = n / r;
        const int64_t span = ns * r;
        double       *ur  = vr, *ui = vi;
        if ( r > 64 )
        {
            br = (double *)malloc( sizeof( double ) * (size_t)r );
            bi = (double *)malloc( sizeof( double ) * (size_t)r );
            ur = br;
            ui = bi;
        }
        for ( int64_t j = 0; j < nr; ++j )
        {
            const int64_t k = j % ns;
            /* twiddled inputs: v[q] = src[j + q*nr] * w_span^{q k} */
            for ( int64_t q = 0; q < r; ++q )
            {
                const double  xr  = src[2 * ( j + q * nr )];
                const double  xi  = src[2 * ( j + q * nr ) + 1];
                const int64_t t   = ( ( q * k ) % span ) * ( n / span );
                const double  wr  = p->tw[2 * t];
                const double  wi  = sign < 0 ? p->tw[2 * t + 1] : -p->tw[2 * t + 1];
                ur[q]             = xr * wr - xi * wi;
                ui[q]             = xr * wi + xi * wr;
            }
            const int64_t d0 = ( j / ns ) * span + k;
            if ( r == 2 )
            {
                dst[2 * d0]                = ur[0] + ur[1];
                dst[2 * d0 + 1]            = ui[0] + ui[1];
                dst[2 * ( d0 + ns )]       = ur[0] - ur[1];
                dst[2 * ( d0 + ns ) + 1]   = ui[0] - ui[1];
                continue;
            }
            if ( r == 4 )
            {
                /* y_m = sum_q u_q e^{sign 2 pi i q m / 4}; e^{sign i pi/2} = sign * i */
                const double s  = (double)sign;
                const double ar = ur[0] + ur[2], ai = ui[0] + ui[2];
                const double br2 = ur[0] - ur[2], bi2 = ui[0] - ui[2];
                const double cr = ur[1] + ur[3], ci = ui[1] + ui[3];
                const double dr = ur[1] - ur[3], di = ui[1] - ui[3];
                /* (sign*i) * d = (-sign*di, sign*dr) */
                dst[2 * d0]                  = ar + cr;
                dst[2 * d0 + 1]              = ai + ci;
                dst[2 * ( d0 + ns )]         = br2 - s * di;
                dst[2 * ( d0 + ns ) + 1]     = bi2 + s * dr;
                dst[2 * ( d0 + 2 * ns )]     = ar - cr;
                dst[2 * ( d0 + 2 * ns ) + 1] = ai - ci;
                dst[2 * ( d0 + 3 * ns )]     = br2 + s * di;
                dst[2 * ( d0 + 3 * ns ) + 1] = bi2 - s * dr;
                continue;
            }
            for ( int64_t m2 = 0; m2 < r; ++m2 )
            {
                double sr = 0.0, si = 0.0;
                for ( int64_t q = 0; q < r; ++q )
                {
                    const int64_t t  = ( ( q * m2 ) % r ) * nr; /* e^{-2 pi i q m / r} */
                    const double  wr = p->tw[2 * t];
                    const double  wi = sign < 0 ? p->tw[2 * t + 1] : -p->tw[2 * t + 1];
                    sr += ur[q] * wr - ui[q] * wi;
                    si += ur[q] * wi + ui[q] * wr;
                }
                dst[2 * ( d0 + m2 * ns )]     = sr;
                dst[2 * ( d0 + m2 * ns ) + 1] = si;
            }
        }
        if ( r > 64 )
        {
            free( br );
            free( bi );
            br = bi = NULL;
        }
        ns *= r;
        double *tmp = src;
        src         = dst;
        dst         = tmp;
    }
    if ( src != data )
        memcpy( data, src, sizeof( double ) * 2 * (size_t)n );
}

/* Public 1D complex transform of a contiguous pencil (for tests). */
void fo_fft1d_c2c( int64_t n, const double *in, double *out, int sign )
{
    fo_plan1d p;
    fo_plan1d_init( &p, n );
    double *scratch = (double *)malloc( sizeof( double ) * 2 * (size_t)( n > 0 ? n : 1 ) );
    memcpy( out, in, sizeof( double ) * 2 * (size_t)n );
    fo_fft1d( &p, out, scratch, sign );
    free( scratch );
    fo_plan1d_free( &p );
}

/* ------------------------------------------------------------------------ */
/* Advanced-layout plans (hipfftMakePlanMany64 semantics)                     */
/* ------------------------------------------------------------------------ */

/*
 * Element offset of logical index (b, i[0..rank-1]) in a layout with embed
 * sizes `emb`, element stride `st` and batch distance `dist`
 * (hipFFT/cuFFT advanced data layout).
 */
static int64_t fo_offset( int rank, const int64_t *emb, int64_t st, int64_t dist, int64_t b, const int64_t *i )
{
    int64_t lin = 0;
    for ( int k = 0; k < rank; ++k )
        lin = lin * emb[k] + i[k];
    return b * dist + lin * st;
}

/*
 * Run one pass along axis `ax` of a rank-`rank` complex array in place
 * (layout emb/st/dist, logical extents `ext`, `batch` copies).
 */
static void fo_pass_c2c_inplace(
    double *buf, int rank, const int64_t *ext, const int64_t *emb, int64_t st, int64_t dist, int64_t batch, int ax,
    int sign
)
{
    const int64_t n = ext[ax];
    int64_t       npencil = batch;
    for ( int k = 0; k < rank; ++k )
        if ( k != ax )
            npencil *= ext[k];
    fo_plan1d p;
    fo_plan1d_init( &p, n );
#pragma omp parallel
    {
        double *line    = (double *)malloc( sizeof( double ) * 2 * (size_t)n );
        double *scratch = (double *)malloc( sizeof( double ) * 2 * (size_t)n );
#pragma omp for schedule( static )
        for ( int64_t pid = 0; pid < npencil; ++pid )
        {
            int64_t idx[3] = { 0, 0, 0 };
            int64_t rem    = pid;
            for ( int k = rank - 1; k >= 0; --k )
            {
                if ( k == ax )
                    continue;
                idx[k] = rem % ext[k];
                rem /= ext[k];
            }
            const int64_t b = rem;
            for ( int64_t t = 0; t < n; ++t )
            {
                idx[ax]           = t;
                const int64_t off = fo_offset( rank, emb, st, dist, b, idx );
                line[2 * t]       = buf[2 * off];
                line[2 * t + 1]   = buf[2 * off + 1];
            }
            fo_fft1d( &p, line, scratch, sign );
            for ( int64_t t = 0; t < n; ++t )
            {
                idx[ax]             = t;
                const int64_t off   = fo_offset( rank, emb, st, dist, b, idx );
                buf[2 * off]        = line[2 * t];
                buf[2 * off + 1]    = line[2 * t + 1];
            }
        }
        free( line );
        free( scratch );
    }
    fo_plan1d_free( &p );
}

static int64_t fo_layout_span( int rank, const int64_t *ext, const int64_t *emb, int64_t st, int64_t dist, int64_t batch )
{
    int64_t last[3] = { 0, 0, 0 };
    for ( int k = 0; k < rank; ++k )
        last[k] = ext[k] - 1;
    return fo_offset( rank, emb, st, dist, batch - 1, last ) + 1;
}

/*
 * Execute one plan on FP64 host buffers.
 *   type: FO_R2C / FO_C2R / FO_C2CF / FO_C2CB
 *   n[rank]: logical transform extents; for R2C/C2R the last one is the real length.
 *   inembed/onembed may be NULL for the default (tight, row-major) layout, in which
 *   case strides/dists are taken from the default layout as hipFFT does.
 * Real buffers hold doubles; complex buffers hold interleaved (re, im) doubles.
 * `in` is never modified.  Returns 0 on success, <0 on bad arguments.
 */
int fo_plan_exec(
    int type, int rank, const int64_t *n, const int64_t *inembed, int64_t istride, int64_t idist,
    const int64_t *onembed, int64_t ostride, int64_t odist, int64_t batch, const double *in, double *out
)
{
    if ( rank < 1 || rank > 3 || batch < 1 )
        return -1;
    for ( int k = 0; k < rank; ++k )
        if ( n[k] < 1 )
            return -1;
    const int64_t last  = n[rank - 1];
    const int64_t lasth = last / 2 + 1;

    int64_t ext_real[3], ext_cplx[3];
    for ( int k = 0; k < rank; ++k )
    {
        ext_real[k] = n[k];
        ext_cplx[k] = n[k];
    }
    if ( type == FO_R2C || type == FO_C2R )
        ext_cplx[rank - 1] = lasth;

    const int64_t *ext_in  = ( type == FO_R2C ) ? ext_real : ext_cplx;
    const int64_t *ext_out = ( type == FO_C2R ) ? ext_real : ext_cplx;

    int64_t iemb[3], oemb[3];
    if ( inembed == NULL )
    {
        for ( int k = 0; k < rank; ++k )
            iemb[k] = ext_in[k];
        istride = 1;
        idist   = 1;
        for ( int k = 0; k < rank; ++k )
            idist *= ext_in[k];
    }
    else
        for ( int k = 0; k < rank; ++k )
            iemb[k] = inembed[k];
    if ( onembed == NULL )
    {
        for ( int k = 0; k < rank; ++k )
            oemb[k] = ext_out[k];
        ostride = 1;
        odist   = 1;
        for ( int k = 0; k < rank; ++k )
            odist *= ext_out[k];
    }
    else
        for ( int k = 0; k < rank; ++k )
            oemb[k] = onembed[k];

    if ( type == FO_C2CF || type == FO_C2CB )
    {
        const int sign = type == FO_C2CF ? -1 : +1;
        /* copy in -> out with the layout change, then pass in place on out */
        int64_t npts = batch;
        for ( int k = 0; k < rank; ++k )
            npts *= n[k];
#pragma omp parallel for schedule( static )
        for ( int64_t q = 0; q < npts; ++q )
        {
            int64_t idx[3] = { 0, 0, 0 }, rem = q;
            for ( int k = rank - 1; k >= 0; --k )
            {
                idx[k] = rem % n[k];
                rem /= n[k];
            }
            const int64_t oi = fo_offset( rank, iemb, istride, idist, rem, idx );
            const int64_t oo = fo_offset( rank, oemb, ostride, odist, rem, idx );
            out[2 * oo]      = in[2 * oi];
            out[2 * oo + 1]  = in[2 * oi + 1];
        }
        for ( int ax = rank - 1; ax >= 0; --ax )
            fo_pass_c2c_inplace( out, rank, n, oemb, ostride, odist, batch, ax, sign );
        return 0;
    }

    if ( type == FO_R2C )
    {
        /* last axis: real -> half complex (full complex FFT of a real line, keep 0..n/2) */
        int64_t npencil = batch;
        for ( int k = 0; k < rank - 1; ++k )
            npencil *= n[k];
        fo_plan1d p;
        fo_plan1d_init( &p, last );
#pragma omp parallel
        {
            double *line    = (double *)malloc( sizeof( double ) * 2 * (size_t)last );
            double *scratch = (double *)malloc( sizeof( double ) * 2 * (size_t)last );
#pragma omp for schedule( static )
            for ( int64_t pid = 0; pid < npencil; ++pid )
            {
                int64_t idx[3] = { 0, 0, 0 }, rem = pid;
                for ( int k = rank - 2; k >= 0; --k )
                {
                    idx[k] = rem % n[k];
                    rem /= n[k];
                }
                for ( int64_t t = 0; t < last; ++t )
                {
                    idx[rank - 1]   = t;
                    line[2 * t]     = in[fo_offset( rank, iemb, istride, idist, rem, idx )];
                    line[2 * t + 1] = 0.0;
                }
                fo_fft1d( &p, line, scratch, -1 );
                for ( int64_t t = 0; t < lasth; ++t )
                {
                    idx[rank - 1]     = t;
                    const int64_t off = fo_offset( rank, oemb, ostride, odist, rem, idx );
                    out[2 * off]      = line[2 * t];
                    out[2 * off + 1]  = line[2 * t + 1];
                }
            }
            free( line );
            free( scratch );
        }
        fo_plan1d_free( &p );
        for ( int ax = rank - 2; ax >= 0; --ax )
            fo_pass_c2c_inplace( out, rank, ext_cplx, oemb, ostride, odist, batch, ax, -1 );
        return 0;
    }

    if ( type == FO_C2R )
    {
        /* work copy of the input (the oracle never modifies `in`) */
        const int64_t span = fo_layout_span( rank, ext_cplx, iemb, istride, idist, batch );
        double       *work = (double *)malloc( sizeof( double ) * 2 * (size_t)span );
        memcpy( work, in, sizeof( double ) * 2 * (size_t)span );
        for ( int ax = 0; ax < rank - 1; ++ax )
            fo_pass_c2c_inplace( work, rank, ext_cplx, iemb, istride, idist, batch, ax, +1 );
        int64_t npencil = batch;
        for ( int k = 0; k < rank - 1; ++k )
            npencil *= n[k];
        fo_plan1d p;
        fo_plan1d_init( &p, last );
#pragma omp parallel
        {
            double *line    = (double *)malloc( sizeof( double ) * 2 * (size_t)last );
            double *scratch = (double *)malloc( sizeof( double ) * 2 * (size_t)last );
#pragma omp for schedule( static )
            for ( int64_t pid = 0; pid < npencil; ++pid )
            {
                int64_t idx[3] = { 0, 0, 0 }, rem = pid;
                for ( int k = rank - 2; k >= 0; --k )
                {
                    idx[k] = rem % n[k];
                    rem /= n[k];
                }
                /* Hermitian extension of bins 0..last/2 */
                for ( int64_t t = 0; t < last; ++t )
                {
                    const int64_t src = t < lasth ? t : last - t;
                    idx[rank - 1]     = src;
                    const int64_t off = fo_offset( rank, iemb, istride, idist, rem, idx );
                    line[2 * t]       = work[2 * off];
                    line[2 * t + 1]   = t < lasth ? work[2 * off + 1] : -work[2 * off + 1];
                }
                line[1] = 0.0; /* imaginary part of DC is ignored */
                if ( last % 2 == 0 )
                    line[2 * ( last / 2 ) + 1] = 0.0; /* and of the Nyquist bin */
                fo_fft1d( &p, line, scratch, +1 );
                for ( int64_t t = 0; t < last; ++t )
                {
                    idx[rank - 1]                                         = t;
                    out[fo_offset( rank, oemb, ostride, odist, rem, idx )] = line[2 * t];
                }
            }
            free( line );
            free( scratch );
        }
        fo_plan1d_free( &p );
        free( work );
        return 0;
    }
    return -2;
}

/* Dense row-major helpers: numpy.fft.rfftn / irfftn (x N) semantics. */
int fo_rfftn( int rank, const int64_t *n, const double *in, double *out )
{
    return fo_plan_exec( FO_R2C, rank, n, NULL, 1, 1, NULL, 1, 1, 1, in, out );
}

int fo_irfftn( int rank, const int64_t *n, const double *in, double *out )
{
    return fo_plan_exec( FO_C2R, rank, n, NULL, 1, 1, NULL, 1, 1, 1, in, out );
}

/*
 * Dense row-major R2C of any rank 1..4 (4D is needed by the ffts 4D oracle;
 * hipFFT plans stop at rank 3, the ffts 4D pipelines compose them).
 */
int fo_rfftn_any( int rank, const int64_t *n, const double *in, double *out )
{
    if ( rank < 1 || rank > 4 )
        return -1;
    if ( rank <= 3 )
        return fo_rfftn( rank, n, in, out );
    /* 4D: R2C over the trailing 3 axes, batch over axis 0, then C2C along axis 0 */
    const int64_t lasth = n[3] / 2 + 1;
    const int64_t inner = n[1] * n[2] * n[3];
    const int64_t innc  = n[1] * n[2] * lasth;
    int           rc    = fo_plan_exec( FO_R2C, 3, n + 1, NULL, 1, inner, NULL, 1, innc, n[0], in, out );
    if ( rc != 0 )
        return rc;
    const int64_t n0[1] = { n[0] };
    /* axis-0 transform: rank-1 C2C with stride innc, batch innc, dist 1 (in place) */
    double *tmp = (double *)malloc( sizeof( double ) * 2 * (size_t)( n[0] * innc ) );
    memcpy( tmp, out, sizeof( double ) * 2 * (size_t)( n[0] * innc ) );
    rc = fo_plan_exec( FO_C2CF, 1, n0, n0, innc, 1, n0, innc, 1, innc, tmp, out );
    free( tmp );
    return rc;
}

int fo_irfftn_any( int rank, const int64_t *n, const double *in, double *out )
{
    if ( rank < 1 || rank > 4 )
        return -1;
    if ( rank <= 3 )
        return fo_irfftn( rank, n, in, out );
    const int64_t lasth = n[3] / 2 + 1;
    const int64_t inner = n[1] * n[2] * n[3];
    const int64_t innc  = n[1] * n[2] * lasth;
    const int64_t n0[1] = { n[0] };
    double       *tmp   = (double *)malloc( sizeof( double ) * 2 * (size_t)( n[0] * innc ) );
    int           rc    = fo_plan_exec( FO_C2CB, 1, n0, n0, innc, 1, n0, innc, 1, innc, in, tmp );
    if ( rc == 0 )
        rc = fo_plan_exec( FO_C2R, 3, n + 1, NULL, 1, innc, NULL, 1, inner, n[0], tmp, out );
    free( tmp );
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Brute-force long-double DFT (pinning of the engine above; tiny shapes)    */
/* ------------------------------------------------------------------------ */

/*
 * Dense row-major R2C of rank 1..4 by the definition
 *   X[k] = sum_j x[j] exp(-2 pi i sum_a k_a j_a / n_a),  k_last in [0, n_last/2].
 * O(N * N_c); only for shapes with a few thousand points.
 */
void fo_brute_r2c( int rank, const int64_t *n, const double *in, double *out )
{
    int64_t ext[4] = { 1, 1, 1, 1 };
    for ( int a = 0; a < rank; ++a )
        ext[4 - rank + a] = n[a];
    const int64_t nh = ext[3] / 2 + 1;
    const long double two_pi = 6.28318530717958647692528676655900577L;
    for ( int64_t k0 = 0; k0 < ext[0]; ++k0 )
        for ( int64_t k1 = 0; k1 < ext[1]; ++k1 )
            for ( int64_t k2 = 0; k2 < ext[2]; ++k2 )
                for ( int64_t k3 = 0; k3 < nh; ++k3 )
                {
                    long double sr = 0.0L, si = 0.0L;
                    int64_t     j  = 0;
                    for ( int64_t j0 = 0; j0 < ext[0]; ++j0 )
                        for ( int64_t j1 = 0; j1 < ext[1]; ++j1 )
                            for ( int64_t j2 = 0; j2 < ext[2]; ++j2 )
                                for ( int64_t j3 = 0; j3 < ext[3]; ++j3, ++j )
                                {
                                    /* phase as an exact rational turn count, reduced before the cosine */
                                    const long double turns =
                                        (long double)( ( k0 * j0 ) % ext[0] ) / (long double)ext[0] +
                                        (long double)( ( k1 * j1 ) % ext[1] ) / (long double)ext[1] +
                                        (long double)( ( k2 * j2 ) % ext[2] ) / (long double)ext[2] +
                                        (long double)( ( k3 * j3 ) % ext[3] ) / (long double)ext[3];
                                    const long double a = -two_pi * turns;
                                    sr += (long double)in[j] * cosl( a );
                                    si += (long double)in[j] * sinl( a );
                                }
                    const int64_t o = ( ( k0 * ext[1] + k1 ) * ext[2] + k2 ) * nh + k3;
                    out[2 * o]      = (double)sr;
                    out[2 * o + 1]  = (double)si;
                }
}

int fo_max_threads( void )
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void fo_set_threads( int n )
{
#ifdef _OPENMP
    if ( n > 0 )
        omp_set_num_threads( n );
#else
    (void)n;
#endif
}

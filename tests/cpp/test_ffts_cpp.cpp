// C++ drop-in check of the MI355X ffts frontend and plan wrappers, written like the
// reference's own drivers (source/tests/test_benchmark_ffts_3D.cu:40-145,
// source/tests/test_fft_3D.cu:24-60, versioned_ffts_3d_test_driver.h:62-67):
//   1. plan + fft_wrap_many smoke: 12 x 10 x 8 R2C through add_plan_3D / activate / exec,
//      checked against the closed-form spectrum of the versioned driver's sample_value;
//   2. benchmark-style round trips (fill_random_real_{3d,4d} splitmix64 input, forward,
//      backward, scale 1/N, RMS vs the regenerated input <= 10 eps) for 3D and each 4D
//      strategy, appending the reference CSV schema to stdout.
// Exit status 0 = every check passed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

#include "../../include/fftm_mi/ffts_mi.hpp"

namespace
{

__host__ __device__ inline unsigned long long splitmix64( unsigned long long x )
{
    x += 0x9E3779B97F4A7C15ull;
    x = ( x ^ ( x >> 30 ) ) * 0xBF58476D1CE4E5B9ull;
    x = ( x ^ ( x >> 27 ) ) * 0x94D049BB133111EBull;
    return x ^ ( x >> 31 );
}

template <class T>
__host__ __device__ inline T unit_random( unsigned long long state )
{
    const unsigned long long bits = splitmix64( state );
    const double             unit = static_cast<double>( bits >> 11 ) * ( 1.0 / 9007199254740992.0 );
    return static_cast<T>( T( 2 ) * static_cast<T>( unit ) - T( 1 ) );
}

// logical (x,y,z[,w]) with element strides sx..sw
template <class T>
__global__ void fill_random( T *a, int nx, int ny, int nz, int nw, long long sx, long long sy, long long sz,
                             long long sw, unsigned long long seed )
{
    const long long total = (long long)nx * ny * nz * nw;
    for ( long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
          i += (long long)gridDim.x * blockDim.x )
    {
        const long long w = i % nw, z = ( i / nw ) % nz, y = ( i / ( (long long)nw * nz ) ) % ny,
                        x = i / ( (long long)nw * nz * ny );
        unsigned long long st = seed ^ ( (unsigned long long)x * 0xD6E8FEB86659FD93ull ) ^
                                ( (unsigned long long)y * 0xA5A3564E27F886A5ull ) ^
                                ( (unsigned long long)z * 0x9E3779B97F4A7C15ull );
        if ( nw > 1 )
            st ^= (unsigned long long)w * 0x94D049BB133111EBull;
        a[x * sx + y * sy + z * sz + w * sw] = unit_random<T>( st );
    }
}

int failures = 0;

void expect( bool ok, const std::string &what )
{
    std::printf( "%s: %s\n", ok ? "PASS" : "FAIL", what.c_str() );
    if ( !ok )
        ++failures;
}

template <class T>
double roundtrip_rms( const std::vector<T> &y, const std::vector<T> &x, double n )
{
    double s = 0;
    for ( std::size_t i = 0; i < x.size(); ++i )
    {
        const double d = (double)y[i] / n - (double)x[i];
        s += d * d;
    }
    return std::sqrt( s / (double)x.size() );
}

// 1. plan-level smoke at the reference smoke size (test_fft_3D.cu:40)
void plan_smoke()
{
    using T            = double;
    using wrap_many    = fftm::wrap::mi::mi_wrap_many<T>;
    using C            = typename wrap_many::complex;
    const long long Nx = 12, Ny = 10, Nz = 8, Nzc = Nz / 2 + 1;
    fftm::mi::device_array<T, 3> t_in;
    fftm::mi::device_array<C, 3> t_out;
    t_in.init( Nx, Ny, Nz );
    t_out.init( Nx, Ny, Nzc );
    std::vector<T> h( Nx * Ny * Nz );
    const double   pi = 3.14159265358979323846;
    for ( long long i = 0; i < Nx; ++i )
        for ( long long j = 0; j < Ny; ++j )
            for ( long long k = 0; k < Nz; ++k )
            {
                const double x = 2 * pi * i / Nx, y = 2 * pi * j / Ny, z = 2 * pi * k / Nz;
                h[( i * Ny + j ) * Nz + k] =
                    std::sin( x ) + 0.5 * std::cos( y ) - 0.25 * std::sin( 2 * z ) + 0.125 * std::sin( x + y - z );
            }
    fftm::wrap::mi::mi_runtime_api::memcpy( t_in.raw_ptr(), h.data(), h.size() * sizeof( T ) );
    wrap_many fft;
    fft.template add_plan_3D<fftm::direction::R2C>( "3D_r2c", Nx, Ny, Nz, Nx, Ny, Nz, 1, Nx * Ny * Nz, Nx, Ny, Nzc, 1,
                                                   Nx * Ny * Nzc, 1 );
    fft.activate();
    fft.exec( "3D_r2c", t_in, t_out );
    std::vector<C> X( Nx * Ny * Nzc );
    fftm::wrap::mi::mi_runtime_api::memcpy( X.data(), t_out.raw_ptr(), X.size() * sizeof( C ) );
    const double N   = double( Nx * Ny * Nz );
    auto         at  = [&]( long long a, long long b, long long c ) { return X[( a * Ny + b ) * Nzc + c]; };
    double       err = 0;
    for ( long long a = 0; a < Nx; ++a )
        for ( long long b = 0; b < Ny; ++b )
            for ( long long c = 0; c < Nzc; ++c )
            {
                double er = 0, ei = 0;
                if ( a == 1 && b == 0 && c == 0 )
                    ei += -N / 2;
                if ( a == Nx - 1 && b == 0 && c == 0 )
                    ei += N / 2;
                if ( a == 0 && ( b == 1 || b == Ny - 1 ) && c == 0 )
                    er += N / 4;
                if ( a == 0 && b == 0 && c == 2 )
                    ei += N / 8;
                if ( a == Nx - 1 && b == Ny - 1 && c == 1 )
                    ei += N / 16;
                err = std::max( err, std::hypot( at( a, b, c ).x - er, at( a, b, c ).y - ei ) );
            }
    std::printf( "plan smoke 12x10x8: work_size=%zu max|X - exact|/N=%.3e\n", fft.get_work_size(), err / N );
    expect( err / N < 1e-14, "plan smoke closed-form spectrum" );
    auto d = fft.plan_descriptor( "3D_r2c" );
    expect( d.rank == 3 && d.n[2] == Nz && d.onembed[2] == Nzc && d.dir == fftm::direction::R2C, "descriptor" );
    bool threw = false;
    try
    {
        fft.template add_plan_1D<fftm::direction::C2CF>( "late", 8, 8, 1, 8, 8, 1, 8, 1 );
    }
    catch ( const std::logic_error & )
    {
        threw = true;
    }
    expect( threw, "add_plan after activate throws logic_error" );
}

template <class T, class S, std::size_t Dim>
struct arrays
{
    typename fftm::mi::ffts<T, S>::template real_array_t<Dim>    real;
    typename fftm::mi::ffts<T, S>::template complex_array_t<Dim> hat;
};

template <class T, class S>
void init_arrays( fftm::mi::ffts<T, S> &f, arrays<T, S, 3> &a, std::size_t nx, std::size_t ny, std::size_t nz,
                  std::size_t )
{
    f.init( nx, ny, nz );
    a.real.init( nx, ny, nz );
    a.hat.init( nx, ny, nz / 2 + 1 );
}

template <class T, class S>
void init_arrays( fftm::mi::ffts<T, S> &f, arrays<T, S, 4> &a, std::size_t nx, std::size_t ny, std::size_t nz,
                  std::size_t nw )
{
    f.init( nx, ny, nz, nw );
    const auto r = f.real_shape_4d();
    const auto c = f.complex_shape_4d();
    a.real.init( r[0], r[1], r[2], r[3] );
    a.hat.init( c[0], c[1], c[2], c[3] );
}

template <class T, class S, std::size_t Dim>
void roundtrip( std::size_t nx, std::size_t ny, std::size_t nz, std::size_t nw, int times, int warmup )
{
    fftm::mi::ffts<T, S>     f;
    arrays<T, S, Dim>        a;
    const unsigned long long seed0 = 0x1234ABCDEF987654ull;
    double                   max_l2 = 0, total_ms = 0;
    std::vector<double>      walls;
    const double             eps = 10.0 * std::numeric_limits<T>::epsilon();
    init_arrays( f, a, nx, ny, nz, nw );
    const std::size_t nr = a.real.size();
    std::vector<T>    x( nr ), y( nr );
    long long         sx, sy, sz, sw;
    const std::string sname = Dim == 3 ? "mi-3d" : fftm::mi::ffts<T, S>::strategy_name();
    if ( Dim == 3 )
    {
        sx = ny * nz, sy = nz, sz = 1, sw = 0;
    }
    else if ( sname == "pencil-pencil-fused" )
    {
        sx = 1, sy = nx, sz = nx * ny, sw = nx * ny * nz;
    }
    else
    {
        sx = ny * nz * nw, sy = nz * nw, sz = nw, sw = 1;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate( &e0 );
    (void)hipEventCreate( &e1 );
    for ( int it = -warmup; it < times; ++it )
    {
        const unsigned long long seed = seed0 + (unsigned long long)( it < 0 ? it + warmup : it );
        hipLaunchKernelGGL( fill_random<T>, dim3( 1024 ), dim3( 256 ), 0, 0, a.real.raw_ptr(), (int)nx, (int)ny,
                            (int)nz, Dim == 3 ? 1 : (int)nw, sx, sy, sz, sw, seed );
        fftm::wrap::mi::mi_runtime_api::device_synchronize();
        if ( it >= 0 )
            fftm::wrap::mi::mi_runtime_api::memcpy( x.data(), a.real.raw_ptr(), nr * sizeof( T ) );
        (void)hipEventRecord( e0, 0 );
        f.forward( a.real, a.hat );
        f.backward( a.hat, a.real );
        (void)hipEventRecord( e1, 0 );
        (void)hipEventSynchronize( e1 );
        float ms = 0;
        (void)hipEventElapsedTime( &ms, e0, e1 );
        if ( it < 0 )
            continue;
        walls.push_back( ms );
        total_ms += ms;
        fftm::wrap::mi::mi_runtime_api::memcpy( y.data(), a.real.raw_ptr(), nr * sizeof( T ) );
        max_l2 = std::max( max_l2, roundtrip_rms( y, x, (double)nr ) );
    }
    double mean = total_ms / times, var = 0;
    for ( double w : walls )
        var += ( w - mean ) * ( w - mean );
    const double sd = std::sqrt( var / times );
    // CSV row of test_benchmark_ffts_3D.cu:130-142
    std::printf( "CSV,\"ffts-%zud\",1,\"%s\",\"none\",1,1,1,%zu,%zu,%zu,%zu,%d,%d,%.3e,%.6e,%.6e,%.6e,\"stdout\"\n",
                 Dim, sname.c_str(), nx, ny, nz, Dim == 3 ? (std::size_t)0 : nw, times, warmup, eps, mean, sd, max_l2 );
    expect( max_l2 <= eps, "round trip " + std::to_string( Dim ) + "D " + sname + ( sizeof( T ) == 4 ? " f32" : " f64" ) );
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
}

// periodic Poisson solve -Lap u = 3 sin x sin y sin z on [0, 2 pi)^3 through ffts::filter
// (versioned_ffts_3d_test_driver.h:187-209, 640-655: forward, solve functor, backward and
// the 1/N scale in one call); exact solution sin x sin y sin z
__global__ void fill_poisson_rhs( float *a, int n, float h )
{
    const long long total = (long long)n * n * n;
    for ( long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
          i += (long long)gridDim.x * blockDim.x )
    {
        const long long z = i % n, y = ( i / n ) % n, x = i / ( (long long)n * n );
        a[i] = -3.0f * sinf( h * x ) * sinf( h * y ) * sinf( h * z );
    }
}

void poisson_filter()
{
    const int                 n = 64;
    fftm::mi::ffts<float>     f;
    f.init( n, n, n );
    fftm::mi::device_array<float, 3> rhs, u;
    rhs.init( n, n, n );
    u.init( n, n, n );
    const double h = 2.0 * 3.14159265358979323846 / n;
    hipLaunchKernelGGL( fill_poisson_rhs, dim3( 256 ), dim3( 256 ), 0, 0, rhs.raw_ptr(), n, (float)h );
    fftm::mi::spectral_op<float> op;
    op.kind  = FFTM_MI_OP_POISSON;
    op.scale = 1.0 / ( (double)n * n * n );
    f.filter( rhs, u, op );
    std::vector<float> hu( (std::size_t)n * n * n );
    fftm::wrap::mi::mi_runtime_api::memcpy( hu.data(), u.raw_ptr(), hu.size() * sizeof( float ) );
    double err = 0;
    for ( int x = 0; x < n; ++x )
        for ( int y = 0; y < n; ++y )
            for ( int z = 0; z < n; ++z )
            {
                const double ex = std::sin( h * x ) * std::sin( h * y ) * std::sin( h * z );
                err             = std::max( err, std::fabs( hu[( (std::size_t)x * n + y ) * n + z] - ex ) );
            }
    std::printf( "poisson filter 64^3 f32: max |u - exact| = %.3e\n", err );
    expect( err < 2e-6, "poisson filter" );
}

} // namespace

int main()
{
    try
    {
        plan_smoke();
        poisson_filter();
        using P = fftm::mi::strategy_4d_pencil_pencil;
        using S = fftm::mi::strategy_4d_slab_slab;
        using X = fftm::mi::strategy_4d_slab_native_xzwy;
        std::printf( "CSV,benchmark,num_gpus,strategy,mode,p1,p2,p3,nx,ny,nz,nw,times,warmup,epsilon,avg_wall_ms,"
                     "stddev_wall_ms,max_l2_diff,directory\n" );
        roundtrip<double, P, 3>( 128, 128, 128, 0, 10, 3 );
        roundtrip<float, P, 3>( 256, 256, 256, 0, 10, 3 );
        roundtrip<double, P, 4>( 16, 16, 16, 16, 5, 1 );
        roundtrip<float, P, 4>( 64, 64, 64, 64, 10, 3 );
        roundtrip<float, S, 4>( 64, 64, 64, 64, 10, 3 );
        roundtrip<float, X, 4>( 64, 64, 64, 64, 10, 3 );
        roundtrip<double, S, 4>( 12, 10, 8, 10, 5, 1 );
        bool threw = false;
        try
        {
            fftm::mi::ffts<float> g;
            g.init( 8, 8, 8 );
            g.init( 8, 8, 8 );
        }
        catch ( const std::logic_error & )
        {
            threw = true;
        }
        expect( threw, "double init throws logic_error" );
    }
    catch ( const std::exception &e )
    {
        std::printf( "FAIL: exception: %s\n", e.what() );
        return 2;
    }
    std::printf( "%s (%d failures)\n", failures ? "FAILED" : "ALL PASSED", failures );
    return failures ? 1 : 0;
}
